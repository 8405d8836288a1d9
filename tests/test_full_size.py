"""Parity at the benchmark workloads' full sizes (the scenes bench.py
measures, built by the same generators with the same seeds):

  box room          the plain CPU-runnable workload: 10 quads + 2 blocks
                    (44 tris), one ceiling light, 256x256, 16 spp, 8 bounces
  icosphere field   64 x 15,680-tri instances (1,003,520 tris), 1920x1080,
                    32 bounces, p = 0.8, hero 4 (the headline render)
  film room         box room with 64 thin-film spheres, 512x512, 16 bounces
                    (the differentiation workload)
  glossy interior   200 Blinn rounded boxes + room shell (470,424 tris),
                    1024x1024, 32 bounces (the guiding workload)
  stress            640 x 15,680-tri instances (10,035,200 tris), 70% thin
                    film, 3840x2160, 32 bounces (the combined stress run)

Bars as in test_gpu_parity: closest hits and shadow results bit-exact
against the oracle on ray batches of the full scene; per-path spectral
radiance within 1e-4 relative for (pixel, sample) pairs spread over the
whole frame; for the headline frame, a same-sample 1-spp render of all
2,073,600 pixels against the FP32 oracle's, through both BVH builders; the
box room's full 256x256, 16-spp frame against the oracle's; and the
adjoint of the film room at its full 512x512 frame and 16 bounces against
the oracle's forward-mode tangents and central differences of lam_render.
"""
import numpy as np
import pytest

import lambda_amd as la
from lambda_amd import scenes
from oracle import binding as ob
from tests.test_gpu_parity import _assert_hits_equal, _bounce_rays, _check_paths, _check_paths_explained
from tests.conftest import random_rays

pytestmark = pytest.mark.gpu

WORKLOADS = {
    "box_room": (lambda d: scenes.box_room(d, seed=1, width=256), 8),
    "icosphere_field": (lambda d: scenes.icosphere_field(d, seed=2), 32),
    "film_room": (lambda d: scenes.box_room_films(d, seed=3, width=512), 16),
    "glossy_interior": (lambda d: scenes.glossy_interior(d, seed=4, width=1024), 32),
    "stress": (lambda d: scenes.icosphere_field(d, seed=5, n_inst=640, material_p=(0.15, 0.15, 0.7), width=3840),
               32),
}


@pytest.fixture(scope="module")
def ctx():
    c = la.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def full(tmp_path_factory):
    cache = {}

    def get(name):
        if name not in cache:
            gen, bounces = WORKLOADS[name]
            s = la.load_nls(gen(str(tmp_path_factory.mktemp(name))))
            cache[name] = (s, ob.Oracle(s), bounces)
        return cache[name]
    return get


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_full_size_hits_bit_exact(ctx, full, name):
    scene, orc, _ = full(name)
    print(f"{name}: {scene.n_tris} tris, {scene.width}x{scene.height}", flush=True)
    ctx.upload(scene)
    rng = np.random.default_rng(21)
    v = scene.tri_v().reshape(-1, 3)
    lo, hi = v.min(axis=0), v.max(axis=0)
    rays = random_rays(rng, 1 << 17, lo, hi)
    o = orc.closest(rays)
    _assert_hits_equal(ctx.extend(rays), o, f"{name} random rays")
    b = _bounce_rays(rng, o, rays, 1 << 16)
    _assert_hits_equal(ctx.extend(b), orc.closest(b), f"{name} bounce rays")
    # shadow queries with random windows and skipped objects
    rays["tmax"] = rng.uniform(0.0, np.linalg.norm(hi - lo), len(rays)).astype(np.float32)
    rays["skip"] = np.where(rng.random(len(rays)) < 0.5,
                            rng.integers(0, scene.desc.n_objects, len(rays)), la.LAM_NONE).astype(np.uint32)
    go, oo = ctx.occluded(rays), orc.occluded(rays)
    assert np.array_equal(go, oo), f"{np.sum(go != oo)} shadow results differ"
    assert 0.02 < oo.mean() < 0.98


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_full_size_paths(ctx, full, name):
    scene, orc, bounces = full(name)
    ctx.upload(scene)
    npix = scene.width * scene.height
    rng = np.random.default_rng(22)
    pix = rng.integers(0, npix, 16384).astype(np.uint32)
    smp = rng.integers(0, 1024, 16384).astype(np.uint32)
    pr = la.params(spp=1, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=9)
    g, o = ctx.trace_paths(pr, pix, smp), orc.trace_paths(pr, pix, smp)
    # (the glossy interior's light is hidden: about 3% of its path values are non-zero)
    assert np.isfinite(o).all() and (o > 0).mean() > 0.01
    _check_paths(g, o)


def test_full_size_frame_same_samples(ctx, full):
    """the headline frame at 1 spp: every pixel against the FP32 oracle's
    render of the same samples, with the host-built and the GPU-built BVH"""
    scene, orc, bounces = full("icosphere_field")
    pr = la.params(spp=1, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=5)
    o = orc.render(pr)
    for gpu_bvh in (False, True):
        ctx.upload(scene, gpu_bvh=gpu_bvh)
        g = ctx.render(pr)
        assert g.shape == o.shape == (1080, 1920, 3)
        relmse = np.mean((g - o) ** 2) / np.mean(o ** 2)
        assert relmse < 1e-8, (gpu_bvh, relmse)


def test_box_room_full_frame_16spp(ctx, full):
    """the plain CPU-runnable workload as configured: 256x256, 16 spp, 8
    bounces, p = 0.8.  All 1,048,576 paths against the oracle's, each within
    1e-4, except paths whose first disagreeing query is degenerate: the
    ceiling light is coplanar with the ceiling and shares its edges, so a
    shadow ray to a light point on an edge (or grazing the plane) crosses a
    ceiling triangle at t = tmax, where the reference's Plücker test answers
    by rounding and the two BVHs' boxes cull differently (9 paths).  The
    image against the oracle's render of the same samples: relMSE < 1e-6
    (1e-7 from those paths; < 1e-10 without them)."""
    scene, orc, bounces = full("box_room")
    assert (scene.width, scene.height, scene.n_tris) == (256, 256, 44)
    pr = la.params(spp=16, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=1)
    ctx.upload(scene)
    npix = scene.width * scene.height
    pix = np.repeat(np.arange(npix, dtype=np.uint32), 16)
    smp = np.tile(np.arange(16, dtype=np.uint32), npix)
    p1 = la.params(spp=1, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=1)
    g, o = ctx.trace_paths(p1, pix, smp), orc.trace_paths(p1, pix, smp)
    n_bad = _check_paths_explained(ctx, scene, orc, p1, pix, smp, g, o)
    print(f"box room: {n_bad} of {len(pix)} paths differ at a degenerate query", flush=True)
    o = orc.render(pr)
    g = ctx.render(pr)
    relmse = np.mean((g - o) ** 2) / np.mean(o ** 2)
    assert relmse < 1e-6, relmse
    assert np.isfinite(g).all() and g.mean() > 0


def _film_dp(scene):
    return la.DiffParams(scene, d=[f"film{i}" for i in range(64)],
                         albedo=[f"mat_{n}" for n in ("floor", "ceiling0", "ceiling1", "ceiling2", "ceiling3",
                                                      "back0", "back1", "left", "right", "front", "tall_top",
                                                      "tall_x", "tall_y", "short_top", "short_x", "short_y")])


def test_film_room_adjoint_full_size(ctx, full):
    """the differentiation workload at its full frame (512x512, 16 bounces,
    p = 0.8): lam_adjoint against the oracle's forward-mode tangents on the
    same paths (thickness and albedo directions, 2e-3) and against central
    differences of lam_render, h = 1e-3 relative (1e-2; Richardson over
    h = 2e-3, 1e-3: 2e-3)"""
    scene, orc, bounces = full("film_room")
    assert (scene.width, scene.height) == (512, 512)
    dp = _film_dp(scene)
    assert dp.n_params == 64 + 16 * scene.desc.grid.n
    th = dp.values(scene).astype(np.float64)
    rng = np.random.default_rng(31)
    W = rng.uniform(0.5, 1.5, size=(512, 512, 3)).astype(np.float32)
    fam = {"d": np.nonzero(np.arange(dp.n_params) < 64)[0], "albedo": np.arange(64, dp.n_params)}
    V = np.zeros((2, dp.n_params))
    for k, idx in enumerate(fam.values()):
        V[k, idx] = rng.normal(size=len(idx)) * np.abs(th[idx])
    ctx.upload(scene)
    p1 = la.params(spp=1, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=41)
    g = ctx.adjoint(p1, dp, W).astype(np.float64)
    J = orc.jvp(p1, dp, V, W)
    for k, name in enumerate(fam):
        assert abs(J[k]) > 0 and abs(g @ V[k] - J[k]) <= 2e-3 * abs(J[k]), (name, g @ V[k], J[k])
    p4 = la.params(spp=4, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=42)

    def loss(t):
        ctx.upload(dp.apply(scene, t))
        return float((ctx.render(p4).astype(np.float64) * W).sum())

    ctx.upload(scene)
    g4 = ctx.adjoint(p4, dp, W).astype(np.float64)
    for k, name in enumerate(fam):
        fd, ad = [], []
        for h in (2e-3, 1e-3):
            tp = (th + h * V[k]).astype(np.float32)
            tm = (th - h * V[k]).astype(np.float32)
            ve = (tp.astype(np.float64) - tm.astype(np.float64)) / (2 * h)
            fd.append((loss(tp) - loss(tm)) / (2 * h))
            ad.append(float(g4 @ ve))
        rich = (4 * fd[1] - fd[0]) / 3
        assert abs(fd[1] - ad[1]) <= 1e-2 * abs(ad[1]), (name, fd, ad)
        assert abs(rich - ad[1]) <= 2e-3 * abs(ad[1]), (name, fd, rich, ad)
    ctx.upload(scene)


def test_stress_guided_render_is_unbiased(ctx, full):
    """the combined stress frame (10,035,200 triangles, 3840x2160, 32
    bounces): a guided render (α = 0.5, the network trained online on its
    first passes, the proposal at the first 4 vertices) against a plain
    render of independent samples, tile by tile (16 x 9 tiles of 240 x 240
    pixels): every tile's mean luminance agrees within 5 standard errors,
    and the two renders' frame means within 3"""
    scene, orc, bounces = full("stress")
    assert scene.n_tris == 10035200 and (scene.width, scene.height) == (3840, 2160)
    ctx.upload(scene)
    spp = 12
    a = ctx.render(la.params(spp=spp, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=21))
    ctx.guide_reset(5)
    g = ctx.render_guided(la.params(spp=spp, max_scatterings=bounces, sample_p=0.8, hero=scene.hero, seed=22),
                          la.guide_params(alpha=0.5, train_passes=4, steps_per_pass=8, max_vertex=4, seed=5))
    assert np.isfinite(a).all() and np.isfinite(g).all()
    la_, lg = a @ np.array([0.2126, 0.7152, 0.0722]), g @ np.array([0.2126, 0.7152, 0.0722])
    diff = (lg - la_).reshape(9, 240, 16, 240).transpose(0, 2, 1, 3).reshape(9, 16, -1)
    z = diff.mean(-1) / (diff.std(-1) / np.sqrt(diff.shape[-1]) + 1e-30)
    assert np.abs(z).max() < 5, np.abs(z).max()
    dz = (lg - la_).ravel()
    assert abs(dz.mean()) < 3 * dz.std() / np.sqrt(dz.size)
