"""Benchmark: lambda's render loop on one MI355X.

Workload (the plain method at scale): a seeded procedural scene of ~1.0 M
triangles (64 instances of a 15,680-triangle noise-displaced geodesic
sphere over a ground plane, Lambertian / Blinn / thin-film materials, two
area lights), 1920x1080, 32 bounces, bounce probability 0.8, 4 hero
wavelengths on 360-830 nm, FP32, 64 samples per pixel.  One step = one
render of that frame (Renderer::render: 64 x 2,073,600 camera paths, the
linear RGB image read back to the host); K steps are K such frames with
different seeds.  value = samples/s over the timed frames, scene resident
in HBM.

Also reported:
  render_rays   rays the timed render traced (extend + shadow, counted on the
                device) per second and per sample.
  roofline      extend (closest-hit) kernel on a 2^22-ray batch (camera rays
                + first-bounce rays) of the same scene: algorithmic bytes per
                ray = 32 (ray) + 16 (hit) + 128 per BVH8 node fetched + 48
                per triangle tested, counts from the host traversal of the
                same 128-byte BVH8 records (lamh_trace_counts), checked equal
                to the kernel's counting variant; time from HIP events around
                the kernel on the library's stream.  Peak HBM 8.0 TB/s.
                "achieved" is algorithmic bytes over time; the scene (75 MB)
                stays in L2/MALL, so the HBM bytes the counters measure
                (traffic, hbm_measured_gbs: from the committed rocprofv3
                --pmc profile of the same batch) are fewer.
  shadow_roofline  the any-hit kernel on shadow rays from the batch's hit
                points to uniform points on the quad lights: 32 + 4 +
                128 per node + 48 per triangle.
  render_kernels_roofline  the render's own extend and shadow kernels in a
                profiled render of two passes of the timed render's shape
                (one lane, HIP events per launch): their node and triangle
                counts (LAM_FLAG_PROFILE_FETCH) over their time in a second
                render of the same samples.
  cpu_baseline  the CPU oracle (C restatement of the reference render loop,
                OpenMP over pixels) on a bounded row sample of the same
                frame and settings, on this host's cores.
  adjoint       the differentiation workload: box room with 16 Lambertian
                albedos and 64 thin-film spheres at 512x512; parameters 64
                thicknesses + 16 albedos x 95 grid coefficients (and the same
                gradient on the albedo CSVs' 81 samples each: 64 + 16 x 81 =
                1,360 values); target at
                parameters x (1 +- U[0,0.15]); L2 image loss; forward 64 spp
                + path-replay adjoint 64 spp, 16 bounces, p=0.8.  grads/s =
                complete gradients per second at that sample count.
  guiding       the learned proposal on the glossy interior (200 instanced
                2,352-tri rounded boxes with Blinn lobes, room shell, one
                light hidden on a rimmed cabinet top, facing the ceiling
                corner; 470,448 tris), 1024x1024, 32 bounces, p=0.8: two
                independent plain renders and two independent guided
                renders (each with a fresh network trained online on its
                first 32 passes, 128 Adam steps after each, training time
                included; the proposal, α = 0.6, stands in for the
                Lambertian and Blinn Kd lobes at every vertex) of the same
                spp.  --guide-layout plate: the light under the ceiling
                above a plate instead.
                relMSE = mean over pixels of Var(Y) / (Ȳ² + 0.01 mean(Ȳ)²),
                Y = luminance, Var estimated from the two runs (both
                estimators are unbiased); time to relMSE 1e-3 =
                run time x relMSE / 1e-3 (variance falls as 1/spp).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import lambda_amd as la  # noqa: E402
from lambda_amd import scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def scene_path(seed=2):
    d = os.path.join(tempfile.gettempdir(), f"lambda_amd_c2_seed{seed}")
    p = os.path.join(d, "icosphere_field.nls")
    if not os.path.exists(p):
        scenes.icosphere_field(d, seed=seed)
    return p


def camera_rays(desc):
    """pixel-centre pinhole rays, sample::camera (cg.hpp:1534-1542)"""
    c = desc.camera
    w, h = c.width, c.height
    j, i = np.meshgrid(np.arange(w, dtype=np.float32), np.arange(h, dtype=np.float32))
    u, v = j.ravel() + 0.5, i.ravel() + 0.5
    pos, X, Y, Z = (np.array(list(a), np.float32) for a in (c.pos, c.x, c.y, c.z))
    wij = pos + X * (-0.5 * c.w + c.delta * u)[:, None] + Y * (-0.5 * c.h + c.delta * v)[:, None] - c.D * Z
    d = wij - pos
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(len(u), dtype=la.RAY_DTYPE)
    r["o"] = pos
    r["d"] = d
    r["tmax"] = np.float32(3.4e38)
    r["skip"] = la.LAM_NONE
    return r


def _traffic(name, n):
    """HBM bytes per launch of a ray-batch kernel from the committed
    rocprofv3 --pmc runs (FETCH_SIZE and WRITE_SIZE scaled by a copy of
    known size), or None when no profile of this batch size is committed"""
    rec_path = os.path.join(ROOT, "profiles", f"{name}.traffic.json")
    if os.path.exists(rec_path):
        rec = json.load(open(rec_path))
        if rec.get("rays_per_launch") == n:
            return round(rec["traffic_bytes"]), f"profiles/{name}.traffic.json (rocprofv3 --pmc)"
    return None, None


def extend_batch(scene, ctx, n_target=1 << 22):
    """2^22 rays: the camera rays of the frame and first-bounce rays from
    their hit points in uniformly random directions (seed 0x5EED)"""
    desc = scene.desc
    cam = camera_rays(desc)
    hits = ctx.extend(cam)
    m = hits["obj"] != la.LAM_NONE
    rng = np.random.default_rng(0x5EED)
    need = n_target - len(cam)
    src = np.nonzero(m)[0]
    src = src[rng.integers(0, len(src), need)]
    b = np.zeros(need, dtype=la.RAY_DTYPE)
    b["o"] = cam["o"][src] + cam["d"][src] * hits["t"][src][:, None]
    dd = rng.normal(size=(need, 3)).astype(np.float32)
    b["d"] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    b["tmax"] = np.float32(3.4e38)
    b["skip"] = la.LAM_NONE
    return np.concatenate([cam, b])


def shadow_batch(scene, rays, hits, rng):
    """shadow rays from the hit points of `rays` to uniform points on a
    uniformly chosen quad light (tmax = distance, the light's object
    skipped), as the light samples of tracePath make them"""
    desc = scene.desc
    m = hits["obj"] != la.LAM_NONE
    p = rays["o"][m] + rays["d"][m] * hits["t"][m][:, None]
    lights = [l for l in scene.lights() if l.type == la.LIGHT_QUADDIR]
    pick = rng.integers(0, len(lights), len(p))
    x = np.zeros_like(p)
    skip = np.zeros(len(p), np.uint32)
    for k, l in enumerate(lights):
        sel = pick == k
        M = np.array(desc.objects[l.object].T.M, np.float32).reshape(3, 4)
        sx, sy = (2 * rng.random((2, sel.sum())) - 1).astype(np.float32)
        x[sel] = M[:, 3] + M[:, 0] * sx[:, None] + M[:, 1] * sy[:, None]
        skip[sel] = l.object
    r = x - p
    dist = np.linalg.norm(r, axis=1)
    out = np.zeros(len(p), dtype=la.RAY_DTYPE)
    out["o"] = p
    out["d"] = r / dist[:, None]
    out["tmax"] = dist
    out["skip"] = skip
    return out


def _batch_roofline(ctx, scene, rays, any_hit, reps, name):
    """one ray-batch kernel: per-ray nodes and triangles from the host
    traversal of the product's BVH8 (lamh_trace_counts, independent of the
    kernel), checked equal to the kernel's own counting variant; the mean
    launch time from HIP events around `reps` launches on the library's
    stream"""
    n = len(rays)
    host_out, hc = la.host_trace(scene, rays, any_hit=any_hit)
    b_rays, b_out, b_cnt = la.DeviceBuffer.from_array(rays), la.DeviceBuffer(n * 16), la.DeviceBuffer(n * 8)
    if any_hit:
        ctx.occluded_dev(b_rays.ptr, n, b_out.ptr, b_cnt.ptr)
        dev_out = b_out.download(np.uint32, n)
        same = bool(np.array_equal(dev_out, host_out))
    else:
        ctx.extend_dev(b_rays.ptr, n, b_out.ptr, b_cnt.ptr)
        dev_out = b_out.download(la.HIT_DTYPE, n)
        same = bool(np.array_equal(dev_out.view(np.uint32), host_out.view(np.uint32)))
    kc = b_cnt.download(np.uint32, 2 * n).reshape(n, 2)
    nodes, tris = hc[:, 0].astype(np.float64), hc[:, 1].astype(np.float64)
    per_ray_io = (32 + 4) if any_hit else (32 + 16)
    bytes_per_launch = float(np.sum(per_ray_io + 128 * nodes + 48 * tris))
    if any_hit:
        ctx.occluded_dev(b_rays.ptr, n, b_out.ptr)  # warm
        avg_ms = ctx.occluded_dev_timed(b_rays.ptr, n, b_out.ptr, reps)
    else:
        ctx.extend_dev(b_rays.ptr, n, b_out.ptr)  # warm
        avg_ms = ctx.extend_dev_timed(b_rays.ptr, n, b_out.ptr, reps)
    for b in (b_rays, b_out, b_cnt):
        b.free()
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, src = _traffic(name, n)
    out = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "achieved_is": "algorithmic bytes (counts of the host traversal) per mean launch time",
        "kernel": "k_trace_rays<any>" if any_hit else "k_trace_rays<closest>", "rays_per_launch": n,
        "avg_ms": round(avg_ms, 4), "rays_per_s": round(n / (avg_ms * 1e-3), 1),
        "nodes_per_ray": round(nodes.mean(), 3), "tris_per_ray": round(tris.mean(), 3),
        "bytes_per_launch": bytes_per_launch, "counts_from": "host traversal of the BVH8 records",
        "kernel_counts_equal": bool(np.array_equal(kc, hc)), "results_equal_host": same,
    }
    if traffic:
        out["traffic_source"] = src
        out["hbm_measured_gbs"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
    return out


def extend_roofline(ctx, scene, reps=10):
    rays = extend_batch(scene, ctx)
    return _batch_roofline(ctx, scene, rays, False, reps, "extend_icosphere_field_4Mrays")


def shadow_roofline(ctx, scene, reps=10):
    rays = extend_batch(scene, ctx)
    hits = ctx.extend(rays)
    sh = shadow_batch(scene, rays, hits, np.random.default_rng(0x5EED + 1))
    return _batch_roofline(ctx, scene, sh, True, reps, "shadow_icosphere_field_batch")


def render_kernel_roofline(ctx, bounces, p, hero, spp):
    """the render's own extend and shadow kernels: one render with
    LAM_FLAG_PROFILE | LAM_FLAG_PROFILE_FETCH (one lane, every traversal
    launch between HIP events, the kernels counting nodes and triangles)
    gives the algorithmic bytes; a second render of the same samples with
    LAM_FLAG_PROFILE alone gives the kernels' time without the counting.
    spp = two passes of the timed render's shape: the queue lengths, and
    with them the kernels' occupancy, follow the samples per pass"""
    base = dict(spp=spp, max_scatterings=bounces, sample_p=p, hero=hero, seed=11)
    ctx.reset_timing()
    ctx.render(la.params(flags=la.LAM_FLAG_PROFILE | la.LAM_FLAG_PROFILE_FETCH, **base))
    c = ctx.timing()
    ctx.reset_timing()
    ctx.render(la.params(flags=la.LAM_FLAG_PROFILE, **base))
    t = ctx.timing()
    out = {}
    for name, ms, nodes, tris, rays, io in (
            ("extend", t.extend_ms, c.extend_nodes, c.extend_tris, c.extend_traced, 32 + 16),
            ("shadow", t.shadow_ms, c.shadow_nodes, c.shadow_tris, c.shadow_traced, 32 + 4)):
        b = io * rays + 128 * nodes + 48 * tris
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[name] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel_ms": round(ms, 3), "rays": int(rays),
                     "nodes_per_ray": round(nodes / max(rays, 1), 3), "tris_per_ray": round(tris / max(rays, 1), 3),
                     "rays_per_s": round(rays / (ms * 1e-3), 1) if ms > 0 else 0.0}
    out["workload"] = (f"{spp}-spp profiled render (one lane, no graph): {t.passes} passes of "
                       f"{t.pass_samples} samples of every pixel")
    return out


_CPU_INFO = {}


def _cpu_info():
    """CPU model, sockets, physical cores and logical CPUs from lscpu (or
    /proc/cpuinfo), and the CPUs this process may run on"""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(ln.split(":", 1) for ln in out.splitlines() if ":" in ln)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name", "")
        sockets = int(kv.get("Socket(s)", "0") or 0)
        cores = int(kv.get("Core(s) per socket", "0") or 0)
        info["sockets"] = sockets
        info["physical_cores"] = sockets * cores
        info["logical_cpus"] = int(kv.get("CPU(s)", "0") or 0)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    if not info.get("model"):
        try:
            with open("/proc/cpuinfo") as f:
                info["model"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except OSError:
        info["usable_cpus"] = os.cpu_count()
    return info


def cpu_baseline(scene, spp, bounces, p, hero, budget_s=20.0):
    """The FP32 oracle (OpenMP over pixels, collapse(2), recursive paths,
    -O3 -mavx2) timed on this host: threads = OMP_NUM_THREADS where it is set
    (the CPU share of one GPU on a shared host), else every usable CPU;
    OMP_PROC_BIND=spread (set by main() before the OpenMP runtime loads).
    It traverses one BVH over the flattened instances, so it is faster than
    the reference's per-object linear loop on scenes of many instances."""
    from oracle.binding import Oracle
    orc = Oracle(scene)
    info = dict(_CPU_INFO) or _cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info.get("usable_cpus") or 1
    h = scene.height
    w = scene.width
    pr = la.params(spp=1, max_scatterings=bounces, sample_p=p, hero=hero, seed=7)
    # calibrate on one full frame at 1 spp (after a stripe that starts the
    # threads), then size the sample's spp to about budget_s seconds
    orc.render(pr, threads=threads, row_step=64, row_phase=32)
    t0 = time.perf_counter()
    orc.render(pr, threads=threads)
    dt = time.perf_counter() - t0
    spp_cpu = 1
    if dt < budget_s / 2:
        spp_cpu = int(max(1, min(spp, round(budget_s / max(dt, 1e-6)))))
        pr = la.params(spp=spp_cpu, max_scatterings=bounces, sample_p=p, hero=hero, seed=8)
        t0 = time.perf_counter()
        orc.render(pr, threads=threads)
        dt = time.perf_counter() - t0
    samples = h * w * spp_cpu
    return {"value": round(samples / dt, 1), "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"full frame {w}x{h}, {spp_cpu} spp, {bounces} bounces, p={p}, hero {hero}; {dt:.1f} s",
            "cpu": info.get("model", ""), "sockets": info.get("sockets"),
            "physical_cores": info.get("physical_cores"), "host_logical_cpus": info.get("logical_cpus"),
            "usable_cpus": info.get("usable_cpus"),
            "note": "one BVH over the flattened instances (the reference loops over objects linearly)",
            "omp": {k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_SCHEDULE", "OMP_NUM_THREADS")}}


def adjoint_bench(spp=64, bounces=16, p=0.8, seed=3):
    d = os.path.join(tempfile.gettempdir(), f"lambda_amd_c3_seed{seed}")
    path = os.path.join(d, "box_room_films.nls")
    if not os.path.exists(path):
        scenes.box_room_films(d, seed=seed, width=512)
    scene = la.load_nls(path)
    dp = la.DiffParams(scene, d=[f"film{i}" for i in range(64)], albedo=list(range(16)))
    th = dp.values(scene)
    rng = np.random.default_rng(seed)
    th_t = (th * (1 + rng.uniform(-0.15, 0.15, th.shape))).astype(np.float32)
    ctx = la.Context(0)
    pr = la.params(spp=spp, max_scatterings=bounces, sample_p=p, hero=scene.hero, seed=seed)
    ctx.upload(dp.apply(scene, th_t))
    target = ctx.render(pr)
    ctx.upload(scene)
    ctx.adjoint(la.params(spp=1, max_scatterings=bounces, sample_p=p, hero=scene.hero, seed=seed + 1),
                dp, np.zeros_like(target))  # warm-up
    t0 = time.perf_counter()
    rgb = ctx.render(pr)
    fwd = time.perf_counter() - t0
    dloss = 2 * (rgb - target)
    t0 = time.perf_counter()
    grad = ctx.adjoint(pr, dp, dloss)
    adj = time.perf_counter() - t0
    adj_stream = ctx.timing().render_ms
    ctx.close()
    # albedo gradients on each material's CSV sample grid (transpose of the
    # loader's box filter), beside the 64 thickness partials
    t0 = time.perf_counter()
    gn = scene.desc.grid.n
    csv_grads = []
    for m in range(16):
        name = scene.material_name(int(m))
        csv = os.path.join(d, f"albedo_{name[len('mat_'):]}.csv")
        lam_csv = np.loadtxt(csv, delimiter=",", dtype=np.float32)[:, 0]
        o = dp.mat_albedo[m]
        csv_grads.append(la.spectrum_grad_to_samples(scene.desc.grid, lam_csv, grad[o:o + gn]))
    map_s = time.perf_counter() - t0
    paths = scene.width * scene.height * spp
    return {"workload": "box room, 16 Lambertian albedos + 64 thin-film spheres (11,564 tris), 512x512, "
                        "L2 loss to a +-15% perturbed target, 16 bounces, p=0.8, hero 4",
            "n_params": int(dp.n_params),
            "n_params_csv_grid": int((dp.mat_d >= 0).sum() + sum(g.size for g in csv_grads)),
            "csv_map_s": round(map_s, 4), "spp": spp, "forward_s": round(fwd, 4), "adjoint_s": round(adj, 4),
            "adjoint_stream_ms": round(adj_stream, 2), "adjoint_paths_per_s": round(paths / adj, 1),
            "grads_per_s": round(1.0 / adj, 3), "loss": float(((rgb - target) ** 2).sum()),
            "grad_finite": bool(np.isfinite(grad).all())}


def guiding_bench(spp=1024, bounces=32, p=0.8, seed=4, width=1024, alpha=0.6, train_passes=32, max_vertex=0,
                  steps_per_pass=128, batch=16384, materials=0, lobes=0, layout="cabinet"):
    """Learned proposal on the glossy interior: two independent plain runs
    and two guided runs of `spp` samples per pixel (the guided runs train
    the network online on their first `train_passes` passes, inside the timed
    call); relMSE of one run from the difference of its two runs."""
    d = os.path.join(tempfile.gettempdir(), f"lambda_amd_c4_{layout}_seed{seed}_{width}")
    path = os.path.join(d, "glossy_interior.nls")
    if not os.path.exists(path):
        scenes.glossy_interior(d, seed=seed, width=width, layout=layout)
    scene = la.load_nls(path)
    ctx = la.Context(0)
    ctx.upload(scene)

    def lum(img):
        return img.astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])

    def relmse(a, b):
        ya, yb = lum(a), lum(b)
        mu = 0.5 * (ya + yb)
        # variance of one run from the two: (a - b)^2 / 2
        return float(np.mean(0.5 * (ya - yb) ** 2 / (mu ** 2 + 0.01 * np.mean(mu) ** 2)))

    pr = lambda s: la.params(spp=spp, max_scatterings=bounces, sample_p=p, hero=scene.hero, seed=s)
    ctx.render(la.params(spp=2, max_scatterings=bounces, sample_p=p, hero=scene.hero, seed=99))  # warm-up
    plain, tp = [], []
    for s in (1, 2):
        t0 = time.perf_counter()
        plain.append(ctx.render(pr(s)))
        tp.append(time.perf_counter() - t0)
    guided, tg, tel = [], [], []
    for s in (1, 2):
        ctx.guide_reset(s)
        gp = la.guide_params(alpha=alpha, train_passes=train_passes, steps_per_pass=steps_per_pass, batch=batch,
                             lr=5e-3, seed=s, max_vertex=max_vertex, materials=materials, lobes=lobes)
        t0 = time.perf_counter()
        guided.append(ctx.render_guided(pr(100 + s), gp))
        tg.append(time.perf_counter() - t0)
        tel.append(ctx.timing())
    rp, rg = relmse(*plain), relmse(*guided)
    t_p, t_g = float(np.mean(tp)), float(np.mean(tg))
    ctx.close()
    samples = scene.width * scene.height * spp
    hidden = {"cabinet": "light on a rimmed cabinet top facing the ceiling corner",
              "plate": "light under the ceiling above a plate"}[layout]
    return {"workload": f"glossy interior ({scene.n_tris:,} tris: room shell + 200 Blinn rounded boxes, "
                        f"{hidden}), {scene.width}x{scene.height}, {bounces} bounces, p={p}, hero 4",
            "layout": layout,
            "spp_per_run": spp, "train_passes": train_passes, "alpha": alpha, "max_vertex": max_vertex,
            "lobes": lobes or 3,
            "steps_per_pass": steps_per_pass, "batch": batch,
            "plain": {"run_s": round(t_p, 3), "samples_per_s": round(samples / t_p, 1), "relmse": round(rp, 5),
                      "time_to_relmse_1e-3_s": round(t_p * rp / 1e-3, 2)},
            "guided": {"run_s": round(t_g, 3), "samples_per_s": round(samples / t_g, 1), "relmse": round(rg, 5),
                       "time_to_relmse_1e-3_s": round(t_g * rg / 1e-3, 2),
                       "train_ms": round(float(np.mean([t.guide_train_ms for t in tel])), 1),
                       "train_records": int(np.mean([t.guide_records for t in tel])),
                       "loss_first": round(float(np.mean([t.guide_loss_first for t in tel])), 4),
                       "loss_last": round(float(np.mean([t.guide_loss_last for t in tel])), 4)},
            "speedup_to_relmse_1e-3": round((t_p * rp) / (t_g * rg), 3),
            "mean_luminance": {"plain": round(float(lum(plain[0] + plain[1]).mean() / 2), 3),
                               "guided": round(float(lum(guided[0] + guided[1]).mean() / 2), 3)}}


def stress_bench(spp=256, bounces=32, p=0.8, seed=5, adj_spp=32):
    """Combined stress: 640 noise-displaced icospheres (10,035,200 tris, 70%
    thin-film materials) at 3840x2160, guided, then an adjoint over every
    thin-film thickness at 1024x1024 (same geometry, another camera)."""
    d = os.path.join(tempfile.gettempdir(), f"lambda_amd_c5_seed{seed}")
    d1 = os.path.join(tempfile.gettempdir(), f"lambda_amd_c5_seed{seed}_1024")
    kw = dict(seed=seed, n_inst=640, material_p=(0.15, 0.15, 0.7))
    p4k = os.path.join(d, "icosphere_field.nls")
    p1k = os.path.join(d1, "icosphere_field.nls")
    if not os.path.exists(p4k):
        scenes.icosphere_field(d, width=3840, **kw)
    if not os.path.exists(p1k):
        scenes.icosphere_field(d1, width=1024, ar=1.0, **kw)
    t0 = time.perf_counter()
    scene = la.load_nls(p4k)
    t_load = time.perf_counter() - t0
    ctx = la.Context(0)
    t0 = time.perf_counter()
    ctx.upload(scene)
    t_up = time.perf_counter() - t0
    st = ctx.stats()
    log(f"stress: {scene.n_tris} tris, load {t_load:.1f}s, upload (BVH build) {t_up:.1f}s, depth {st.depth}, "
        f"stack bound {st.stack_bound}")
    # the same scene with the BVH built on the device, and the plain render on either tree
    pq = la.params(spp=16, max_scatterings=bounces, sample_p=p, hero=scene.hero, seed=1)
    t0 = time.perf_counter()
    ref16 = ctx.render(pq)
    t_host_tree = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx.upload(scene, gpu_bvh=True)
    t_up_gpu = time.perf_counter() - t0
    sg = ctx.stats()
    t0 = time.perf_counter()
    gpu16 = ctx.render(pq)
    t_gpu_tree = time.perf_counter() - t0
    same = bool(np.array_equal(ref16, gpu16))
    ctx.upload(scene)
    pr = la.params(spp=spp, max_scatterings=bounces, sample_p=p, hero=scene.hero, seed=1)
    gp = la.guide_params(alpha=0.6, train_passes=32, steps_per_pass=128, batch=16384, lr=5e-3, seed=1)
    ctx.guide_reset(1)
    t0 = time.perf_counter()
    img = ctx.render_guided(pr, gp)
    t_g = time.perf_counter() - t0
    tm = ctx.timing()
    s1k = la.load_nls(p1k)
    ctx.set_camera(s1k)
    films = [i for i, m in enumerate(s1k.materials()) if m.type == la.MAT_THINFILM]
    dp = la.DiffParams(s1k, d=films)
    pa = la.params(spp=adj_spp, max_scatterings=bounces, sample_p=p, hero=scene.hero, seed=3)
    t0 = time.perf_counter()
    fwd = ctx.render(pa)
    t_f = time.perf_counter() - t0
    # L2 loss against 0.9 x the image: dloss/drgb = 0.2 img
    t0 = time.perf_counter()
    grad = ctx.adjoint(pa, dp, (0.2 * fwd).astype(np.float32))
    t_a = time.perf_counter() - t0
    ctx.close()
    return {"workload": f"640 icospheres ({scene.n_tris:,} tris, 70% thin film), {scene.width}x{scene.height}, "
                        f"{spp} spp guided (trained on the first 32), {bounces} bounces; adjoint over "
                        f"{dp.n_params} thin-film thicknesses at 1024x1024, {adj_spp} spp",
            "load_s": round(t_load, 2), "upload_s": round(t_up, 2), "bvh_depth": st.depth,
            "stack_bound": st.stack_bound,
            "gpu_bvh": {"upload_s": round(t_up_gpu, 2), "nodes": sg.n_nodes, "depth": sg.depth,
                        "stack_bound": sg.stack_bound, "render_16spp_s": round(t_gpu_tree, 3),
                        "host_tree_render_16spp_s": round(t_host_tree, 3), "same_image": same},
            "guided_s": round(t_g, 3), "guided_samples_per_s": round(scene.width * scene.height * spp / t_g, 1),
            "train_ms": round(tm.guide_train_ms, 1), "image_finite": bool(np.isfinite(img).all()),
            "forward_s": round(t_f, 3), "adjoint_s": round(t_a, 3), "n_params": dp.n_params,
            "grad_finite": bool(np.isfinite(grad).all()), "grad_nonzero": int(np.count_nonzero(grad))}


def gpu_bvh_bench(bounces=32, p=0.8, spp=4):
    """The icosphere field's BVH8 built on the device (LAM_UPLOAD_GPU_BVH)
    next to the host builder: upload time, tree shape, render time and the
    image (equal bit for bit: the closest hit does not depend on the tree)."""
    scene = la.load_nls(scene_path())
    ctx = la.Context(0)
    pr = la.params(spp=spp, max_scatterings=bounces, sample_p=p, hero=4, seed=1)
    out = {"workload": f"icosphere field ({scene.n_tris:,} tris), {scene.width}x{scene.height}, {spp} spp"}
    imgs = []
    for name, gpu in (("host", False), ("gpu", True)):
        t0 = time.perf_counter()
        ctx.upload(scene, gpu_bvh=gpu)
        t_up = time.perf_counter() - t0
        st = ctx.stats()
        ctx.render(la.params(spp=1, max_scatterings=bounces, sample_p=p, hero=4, seed=9))
        t0 = time.perf_counter()
        imgs.append(ctx.render(pr))
        out[name] = {"upload_s": round(t_up, 3), "nodes": st.n_nodes, "depth": st.depth,
                     "stack_bound": st.stack_bound, "render_s": round(time.perf_counter() - t0, 4)}
    out["same_image"] = bool(np.array_equal(*imgs))
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed frames (each one render of --spp samples per pixel)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed frames before the timed ones")
    ap.add_argument("--spp", type=int, default=64, help="samples per pixel of one frame; 64 is the configuration's")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-guiding", action="store_true")
    ap.add_argument("--legs", default="render,roofline,adjoint,guiding,stress,cpu",
                    help="comma list of legs to run (profiling runs pick one); 'stress' is the 10 M-triangle "
                         "4K guided + adjoint run, 'gpu_bvh' (not in the default) the device BVH build of the "
                         "render scene")
    ap.add_argument("--guide-spp", type=int, default=1024)
    ap.add_argument("--guide-train-passes", type=int, default=32)
    ap.add_argument("--guide-steps-per-pass", type=int, default=128)
    ap.add_argument("--guide-max-vertex", type=int, default=0,
                    help="guide path vertices of bounce index below this (0 = every vertex)")
    ap.add_argument("--guide-lobes", type=int, default=3,
                    help="guided lobes: 1 Lambertian, 2 Blinn Kd, 4 Blinn specular (0 = 1 + 2)")
    ap.add_argument("--guide-alpha", type=float, default=0.6)
    ap.add_argument("--guide-layout", default="cabinet", choices=("cabinet", "plate"),
                    help="where the glossy interior hides its light (scenes.glossy_interior)")
    ap.add_argument("--flags", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0, help="pass lanes (0 = the library's default)")
    ap.add_argument("--max-iters", type=int, default=0,
                    help="bounce iterations of a pass before its tail goes to the finish kernel (0 = the library's)")
    ap.add_argument("--samples-per-pass", type=int, default=0,
                    help="samples of every pixel per wavefront pass (0 = the library's default)")
    ap.add_argument("--extend-only", action="store_true",
                    help="profiling helper: only the extend and shadow ray-batch microbenchmarks plus a 1 GiB "
                         "device copy (the copy calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    # the CPU baseline's OpenMP threads spread over the cores (read by the
    # OpenMP runtime when the oracle library loads)
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    # the host's CPUs before anything pins this process's threads
    _CPU_INFO.update(_cpu_info())
    if args.gpus != 1:
        print(json.dumps({"error": "this renderer runs on one GPU; --gpus > 1 is not supported "
                                   "(samples would shard by pixel tiles with one film all-reduce)"}))
        sys.exit(2)
    legs = set(args.legs.split(","))
    if args.no_cpu:
        legs.discard("cpu")
    if args.no_roofline:
        legs.discard("roofline")
    if args.no_adjoint:
        legs.discard("adjoint")
    if args.no_guiding:
        legs.discard("guiding")
    if "render" not in legs and not args.extend_only:
        out = {}
        if "stress" in legs:
            out["stress"] = stress_bench()
        if "adjoint" in legs:
            out["adjoint"] = adjoint_bench()
        if "guiding" in legs:
            out["guiding"] = guiding_bench(spp=args.guide_spp, train_passes=args.guide_train_passes,
                                           max_vertex=args.guide_max_vertex,
                                           steps_per_pass=args.guide_steps_per_pass, lobes=args.guide_lobes,
                                           alpha=args.guide_alpha, layout=args.guide_layout)
        if "gpu_bvh" in legs:
            out["gpu_bvh"] = gpu_bvh_bench()
        print(json.dumps(out), flush=True)
        return
    bounces, p, hero = 32, 0.8, 4
    t0 = time.perf_counter()
    scene = la.load_nls(scene_path())
    ctx = la.Context(0)
    ctx.upload(scene)
    if args.samples_per_pass:
        ctx.set_wave_config(samples_per_pass=args.samples_per_pass)
    if args.lanes:
        ctx.set_wave_config(lanes=args.lanes)
    if args.max_iters:
        ctx.set_wave_config(max_iters=args.max_iters)
    st = ctx.stats()
    log(f"scene: {scene.n_tris} tris, {st.n_nodes} BVH8 nodes, depth {st.depth}, stack bound "
        f"{st.stack_bound}, setup {time.perf_counter() - t0:.1f}s")
    w, h = scene.width, scene.height
    if args.extend_only:
        rl = extend_roofline(ctx, scene, reps=3)
        sl = shadow_roofline(ctx, scene, reps=3)
        nbytes = 1 << 30
        a, b = la.DeviceBuffer(nbytes), la.DeviceBuffer(nbytes)
        for _ in range(3):  # device to device: the counters' scale reference
            assert la._hip().hipMemcpy(b.ptr, a.ptr, nbytes, 3) == 0
        a.free()
        b.free()
        print(json.dumps({"extend": rl, "shadow": sl, "copy_bytes": nbytes}), flush=True)
        ctx.close()
        return
    # warm-up: whole frames of the timed shape, so that the lanes' path
    # state and graphs exist before the timed frames
    spp = args.spp
    warm = args.warmup
    for fw in range(warm):
        ctx.render(la.params(spp=spp, max_scatterings=bounces, sample_p=p, hero=hero, seed=1000 + fw,
                             flags=args.flags))
    ctx.reset_timing()
    stream_ms = 0.0
    t0 = time.perf_counter()
    for f in range(args.steps):
        img = ctx.render(la.params(spp=spp, max_scatterings=bounces, sample_p=p, hero=hero, seed=1 + f,
                                   flags=args.flags))
        stream_ms += ctx.timing().render_ms
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    samples = w * h * spp * args.steps
    log(f"render: {args.steps} frames of {spp} spp in {wall:.3f}s wall, {stream_ms:.1f} ms on stream; "
        f"last frame mean {img.mean():.4g} finite {np.isfinite(img).all()}")
    m_pass, n_pass = tm.pass_samples, tm.passes  # the timed frames' pass shape
    out = {
        "metric": "samples/s (plain method, 1 GPU)", "value": round(samples / wall, 1), "unit": "samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": warm,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "icosphere field: 64 x 15,680-tri instances (1,003,520 tris), ground plane, "
                               "2 area lights, 1920x1080, 64 spp, 32 bounces, p=0.8, hero 4 on 360-830 nm; "
                               "one step = one frame",
                   "width": w, "height": h, "spp_per_step": spp, "samples_per_pass": m_pass, "passes": n_pass,
                   "bounces": bounces, "bounce_prob": p, "triangles": scene.n_tris},
        "render_rays": {"rays_per_s": round((tm.extend_rays + tm.shadow_rays) / wall, 1),
                        "extend_rays_per_sample": round(tm.extend_rays / samples, 4),
                        "shadow_rays_per_sample": round(tm.shadow_rays / samples, 4)},
        "stream_ms_per_step": round(stream_ms / args.steps, 3),
    }
    if "roofline" in legs:
        out["roofline"] = extend_roofline(ctx, scene)
        log(f"extend: {out['roofline']}")
        out["shadow_roofline"] = shadow_roofline(ctx, scene)
        log(f"shadow: {out['shadow_roofline']}")
        out["render_kernels_roofline"] = render_kernel_roofline(ctx, bounces, p, hero, 2 * m_pass)
        log(f"render kernels: {out['render_kernels_roofline']}")
    ctx.close()
    if "adjoint" in legs:
        out["adjoint"] = adjoint_bench()
        log(f"adjoint: {out['adjoint']}")
    if "guiding" in legs:
        out["guiding"] = guiding_bench(spp=args.guide_spp, train_passes=args.guide_train_passes,
                                       max_vertex=args.guide_max_vertex, steps_per_pass=args.guide_steps_per_pass,
                                       lobes=args.guide_lobes, alpha=args.guide_alpha, layout=args.guide_layout)
        log(f"guiding: {out['guiding']}")
    if "stress" in legs:
        out["stress"] = stress_bench()
        log(f"stress: {out['stress']}")
    if "cpu" in legs:
        out["cpu_baseline"] = cpu_baseline(scene, 1024, bounces, p, hero)
        log(f"cpu: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
