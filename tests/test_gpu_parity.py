"""GPU parity: the HIP kernels against the CPU oracle on identical inputs.

Bars:
  (a) closest-hit t / triangle / object / prim bit-exact, ties to lowest id;
  (b) per-path spectral radiance within 1e-4 relative;
  (c) images: same-sample FP32 renders agree to film reassociation; against
      the FP64 oracle at independent samples, per-pixel z-test + relMSE.
"""
import numpy as np
import pytest

import lambda_amd as la
from oracle import binding as ob
from tests.conftest import random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = la.Context(0)
    yield c
    c.close()


def _bits(h):
    return np.stack([h["t"].view(np.uint32), h["tri"], h["obj"], h["prim"]], axis=1)


def _assert_hits_equal(g, o, what):
    gb, obits = _bits(g), _bits(o)
    bad = np.nonzero(np.any(gb != obits, axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size}/{len(g)} hits differ, first {bad[:5]}: gpu {g[bad[:3]]} oracle {o[bad[:3]]}"


def _bounce_rays(rng, hits, rays, n):
    """origins at hit points, random directions (first-bounce rays)"""
    m = hits["obj"] != la.LAM_NONE
    src = np.nonzero(m)[0][:n]
    out = np.zeros(len(src), dtype=la.RAY_DTYPE)
    t = hits["t"][src].astype(np.float32)[:, None]
    out["o"] = (rays["o"][src] + rays["d"][src] * t).astype(np.float32)
    d = rng.normal(size=(len(src), 3))
    out["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    out["tmax"] = 3.4e38
    out["skip"] = la.LAM_NONE
    return out


def test_device_math_bitwise(ctx):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-60, 60, 200000), rng.uniform(-1, 1, 100000),
                        np.array([0.0, -0.0, 1e-30, 3.14159265, 1e6])]).astype(np.float32)
    for fn in (0, 1):
        assert np.array_equal(ctx.debug_math(fn, x).view(np.uint32), ob.math(fn, x).view(np.uint32)), fn
    base = rng.uniform(-1.2, 1.2, 300000).astype(np.float32)
    ex = np.concatenate([rng.uniform(0, 2000, 150000), rng.integers(0, 1100, 150000)]).astype(np.float32)
    assert np.array_equal(ctx.debug_math(2, base, ex).view(np.uint32), ob.math(2, base, ex).view(np.uint32))
    y = rng.normal(size=300000).astype(np.float32)
    assert np.array_equal(ctx.debug_math(3, y, x[:300000]).view(np.uint32),
                          ob.math(3, y, x[:300000]).view(np.uint32))
    pos = np.abs(x)
    assert np.array_equal(ctx.debug_math(4, pos), np.sqrt(pos))  # IEEE sqrt
    den = np.where(y == 0, 1.0, y).astype(np.float32)
    assert np.array_equal(ctx.debug_math(5, x[:300000], den), x[:300000] / den)  # IEEE division


@pytest.mark.parametrize("scene_name", ["box_scene", "field_scene", "mixed_scene"])
def test_extend_bit_exact(ctx, request, scene_name):
    scene = request.getfixturevalue(scene_name)
    ctx.upload(scene)
    orc = ob.Oracle(scene)
    rng = np.random.default_rng(11)
    v = scene.tri_v().reshape(-1, 3) if scene.n_tris else np.array([[-20, -20, 0], [20, 20, 24.0]])
    lo, hi = v.min(axis=0), v.max(axis=0)
    rays = random_rays(rng, 200000, lo, hi)
    g, o = ctx.extend(rays), orc.closest(rays)
    _assert_hits_equal(g, o, "random rays")
    assert (o["obj"] != la.LAM_NONE).mean() > 0.3
    b = _bounce_rays(rng, o, rays, 100000)
    _assert_hits_equal(ctx.extend(b), orc.closest(b), "bounce rays")
    # the oracle's BVH equals brute force (its own invariant)
    sub = rays[:5000]
    _assert_hits_equal(orc.closest(sub), orc.closest(sub, brute=True), "oracle bvh vs brute")


@pytest.mark.parametrize("scene_name", ["box_scene", "field_scene", "mixed_scene", "lights_scene"])
def test_shadow_bit_exact(ctx, request, scene_name):
    scene = request.getfixturevalue(scene_name)
    ctx.upload(scene)
    orc = ob.Oracle(scene)
    rng = np.random.default_rng(12)
    v = scene.tri_v().reshape(-1, 3) if scene.n_tris else np.array([[-20, -20, 0], [20, 20, 24.0]])
    lo, hi = v.min(axis=0), v.max(axis=0)
    rays = random_rays(rng, 100000, lo, hi)
    rays["tmax"] = rng.uniform(0.0, np.linalg.norm(hi - lo), len(rays)).astype(np.float32)
    objs = np.array([o for o in range(scene.desc.n_objects)], dtype=np.uint32)
    rays["skip"] = np.where(rng.random(len(rays)) < 0.5, rng.choice(objs, len(rays)), la.LAM_NONE)
    go, oo = ctx.occluded(rays), orc.occluded(rays)
    assert np.array_equal(go, oo), f"{np.sum(go != oo)} shadow results differ"
    assert 0.05 < oo.mean() < 0.95
    # closest-hit with the same window agrees with occlusion
    gh = ctx.extend(rays)
    assert np.array_equal(go, (gh["obj"] != la.LAM_NONE).astype(np.uint32))


def _check_paths(a, b, rel=1e-4, floor=1e-6):
    """|gpu - oracle| <= rel |oracle| + floor * mean|oracle| for every path
    and wavelength.  The absolute floor covers paths whose terms cancel to
    ~0 (the forward sum and the recursive oracle associate differently)."""
    scale = max(float(np.abs(b).mean()), 1e-30)
    err = np.abs(a.astype(np.float64) - b) / (rel * np.abs(b) + floor * scale)
    bad = np.argwhere(err > 1.0)
    assert bad.size == 0, f"{len(bad)} path values out of tolerance: worst {err.max():.3g} at {bad[:3]}"


def _degenerate_query(scene, ray, tri):
    """the intersection of `ray` with triangle `tri` is decided by rounding:
    the ray grazes the triangle's plane (|cos| < 1e-4), or it crosses the
    plane within float32 uncertainty of an edge — the crossing point moves
    by about 2^-24 (|o| + |v| + t) / |cos| when the ray's and the vertices'
    coordinates are rounded, and there the reference's Plücker test
    (cg.hpp:1341-1379) and a box test of the same point can disagree"""
    v = scene.tri_v()[tri].astype(np.float64)
    o = np.asarray(ray["o"], np.float64)
    d = np.asarray(ray["d"], np.float64)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    cos = abs(d @ n) / (np.linalg.norm(d) * np.linalg.norm(n))
    if cos < 1e-4:
        return True
    t = ((v[0] - o) @ n) / (d @ n)
    p = o + t * d
    tol = 8 * 2.0 ** -24 * (np.abs(o).max() + np.abs(v).max() + abs(t) * np.abs(d).max()) / cos
    dist = []
    for k in range(3):
        a, b = v[k], v[(k + 1) % 3]
        e = b - a
        dist.append(np.linalg.norm(np.cross(e, p - a)) / np.linalg.norm(e))
    return min(dist) <= tol


def _explain_divergence(ctx, scene, orc, p, pixel, sample):
    """for a path whose GPU radiance differs from the oracle's: replay the
    oracle's queries of that path (in order) through the GPU's extend and
    occlusion kernels, and return (index, kind) of the first query they
    answer differently, or None.  Every query before it is answered the
    same, so both tracers followed the same path up to there."""
    rays, kind = orc.path_rays(p, pixel, sample)
    ho, hg = orc.closest(rays), ctx.extend(rays)
    oo, og = orc.occluded(rays), ctx.occluded(rays)
    for i in range(len(rays)):
        if kind[i] == 0 and _bits(ho[i:i + 1]).tobytes() != _bits(hg[i:i + 1]).tobytes():
            tri = ho["tri"][i] if ho["tri"][i] != la.LAM_NONE else hg["tri"][i]
            return i, 0, rays[i], tri
        if kind[i] == 1 and oo[i] != og[i]:
            h = orc.closest(rays[i:i + 1]) if oo[i] else ctx.extend(rays[i:i + 1])
            return i, 1, rays[i], h["tri"][0]
    return None


def _check_paths_explained(ctx, scene, orc, p, pixel, sample, g, o, max_frac=2e-5):
    """_check_paths, except that a path may differ when the first query on
    which the GPU and the oracle disagree is degenerate (a grazing ray or a
    crossing on a triangle edge: the answer depends on which boxes a BVH
    culls, not on the kernels' arithmetic); at most max_frac of the paths"""
    scale = max(float(np.abs(o).mean()), 1e-30)
    err = np.abs(g.astype(np.float64) - o) / (1e-4 * np.abs(o) + 1e-6 * scale)
    bad = np.nonzero(err.max(axis=1) > 1.0)[0]
    assert bad.size <= max(1, max_frac * len(pixel)), f"{bad.size} of {len(pixel)} paths differ"
    for b in bad:
        e = _explain_divergence(ctx, scene, orc, p, pixel[b], sample[b])
        assert e is not None, f"path ({pixel[b]}, {sample[b]}) differs with every query answered the same"
        i, kind, ray, tri = e
        assert tri != la.LAM_NONE and _degenerate_query(scene, ray, tri), (pixel[b], sample[b], i, kind, ray, tri)
    return bad.size


@pytest.mark.parametrize("flags", [0, la.LAM_FLAG_NO_TAIL, la.LAM_FLAG_MEGAKERNEL],
                         ids=["default", "wavefront", "megakernel"])
@pytest.mark.parametrize("scene_name,hero,p,bounces", [
    ("box_scene", 4, 0.8, 8), ("field_scene", 4, 0.8, 12), ("mixed_scene", 16, 0.8, 8),
    ("mixed_scene", 4, 0.5, 4), ("lights_scene", 16, 0.8, 8), ("lights_scene", 4, 0.6, 5)])
def test_trace_paths_parity(ctx, request, scene_name, hero, p, bounces, flags):
    scene = request.getfixturevalue(scene_name)
    ctx.upload(scene)
    orc = ob.Oracle(scene)
    npix = scene.width * scene.height
    rng = np.random.default_rng(5)
    pix = rng.integers(0, npix, 8192).astype(np.uint32)
    smp = rng.integers(0, 64, 8192).astype(np.uint32)
    pr = la.params(spp=1, max_scatterings=bounces, sample_p=p, hero=hero, seed=9, flags=flags)
    g, o = ctx.trace_paths(pr, pix, smp), orc.trace_paths(pr, pix, smp)
    assert np.isfinite(o).all()
    assert (o > 0).mean() > 0.05
    _check_paths(g, o)


@pytest.mark.parametrize("flags", [0, la.LAM_FLAG_NO_GRAPH, la.LAM_FLAG_PROFILE, la.LAM_FLAG_MEGAKERNEL,
                                   la.LAM_FLAG_NO_TAIL],
                         ids=["graph", "nograph", "profile", "megakernel", "notail"])
@pytest.mark.parametrize("scene_name,hero", [("box_scene", 4), ("mixed_scene", 16), ("lights_scene", 16)])
def test_render_matches_oracle_same_samples(ctx, request, scene_name, hero, flags):
    scene = request.getfixturevalue(scene_name)
    ctx.upload(scene)
    orc = ob.Oracle(scene)
    pr = la.params(spp=4, max_scatterings=8, sample_p=0.8, hero=hero, seed=1, flags=flags)
    g, gx = ctx.render(pr, want_xyz=True)
    o, ox = orc.render(pr, want_xyz=True)
    # XYZ projection of the tent-splat film equals the spectral film then
    # projected (linear equivalence), to float reassociation
    np.testing.assert_allclose(gx, ox, rtol=2e-4, atol=2e-5 * float(np.abs(ox).max()))
    relmse = np.mean((g - o) ** 2) / np.mean(o ** 2)
    assert relmse < 1e-8, relmse


def _shade_stage(scene):
    """what the shade kernels stage in LDS for `scene` (stage_plan in
    lam_wave.h): 3 spectra and tables, 2 the tables, 1 the spectra (the
    plain shade kernel only), 0 nothing"""
    D = scene.desc
    spec = (D.n_spectra * D.grid.n + 3) // 4 * 16
    tab = ((D.n_materials * 56 + 15) // 16 + D.n_lights * 3 + D.n_objects * 7) * 16
    if spec + tab <= 52 * 1024:
        return 3
    if tab <= 52 * 1024:
        return 2
    return 1 if spec <= 48 * 1024 else 0


def test_stage_choice_of_the_test_scenes(box_scene, mixed_scene, crowd_scene, field_scene):
    """the parity scenes cover the shade kernels' LDS staging variants (the
    glossy interior's tables alone and the stress field's none are covered
    by the full-size tests)"""
    assert _shade_stage(box_scene) == 3 and _shade_stage(mixed_scene) == 3 and _shade_stage(field_scene) == 3
    assert _shade_stage(crowd_scene) == 1


@pytest.mark.parametrize("hero", [4, 16])
def test_spectra_only_staging_parity(ctx, crowd_scene, hero):
    """a large object table (530 objects) beside small spectra: the shade
    kernel stages the spectra alone; paths and images against the oracle"""
    ctx.upload(crowd_scene)
    orc = ob.Oracle(crowd_scene)
    npix = crowd_scene.width * crowd_scene.height
    rng = np.random.default_rng(17)
    pix = rng.integers(0, npix, 8192).astype(np.uint32)
    smp = rng.integers(0, 64, 8192).astype(np.uint32)
    pr = la.params(spp=1, max_scatterings=8, sample_p=0.8, hero=hero, seed=9)
    g, o = ctx.trace_paths(pr, pix, smp), orc.trace_paths(pr, pix, smp)
    assert (o > 0).mean() > 0.05
    _check_paths(g, o)
    pr = la.params(spp=4, max_scatterings=8, sample_p=0.8, hero=hero, seed=1)
    gi, oi = ctx.render(pr), orc.render(pr)
    relmse = np.mean((gi - oi) ** 2) / np.mean(oi ** 2)
    assert relmse < 1e-8, relmse


@pytest.mark.parametrize("hero", [4, 16])
def test_fine_spectral_grid_parity(ctx, tmp_path, hero):
    """a 3,000-bin grid: the colour-matching table (36 KB) is too large for
    the film kernel's LDS copy and the spectra for the shade kernels', so
    both read them from HBM; paths and images against the oracle"""
    from tests.scene_text import mixed_scene_text
    text = mixed_scene_text(str(tmp_path)).rstrip()
    scene = la.Scene.from_text(text[:-1] + ',"spectral_grid":{"n":3000,"inf":380,"sup":780}}')
    assert scene.desc.grid.n == 3000 and 3 * 3000 * 4 > 32 * 1024
    assert _shade_stage(scene) in (0, 2)
    ctx.upload(scene)
    orc = ob.Oracle(scene)
    npix = scene.width * scene.height
    rng = np.random.default_rng(23)
    pix = rng.integers(0, npix, 4096).astype(np.uint32)
    smp = rng.integers(0, 64, 4096).astype(np.uint32)
    pr = la.params(spp=1, max_scatterings=8, sample_p=0.8, hero=hero, seed=3)
    g, o = ctx.trace_paths(pr, pix, smp), orc.trace_paths(pr, pix, smp)
    assert (o > 0).mean() > 0.05
    _check_paths(g, o)
    pr = la.params(spp=6, max_scatterings=8, sample_p=0.8, hero=hero, seed=2)
    gi, oi = ctx.render(pr), orc.render(pr)
    relmse = np.mean((gi - oi) ** 2) / np.mean(oi ** 2)
    assert relmse < 1e-8, relmse


def test_tail_handoff_is_exact(ctx, tmp_path):
    """262,144 paths: the wavefront runs the first bounces and hands the
    short queue to the finish kernel; the image must equal the pure
    wavefront's bit for bit (same per-path arithmetic, same film order)."""
    from lambda_amd import scenes
    s = la.load_nls(scenes.box_room(str(tmp_path), seed=1, width=512))
    ctx.upload(s)
    a = ctx.render(la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=5))
    b = ctx.render(la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=5, flags=la.LAM_FLAG_NO_TAIL))
    c = ctx.render(la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=5, flags=la.LAM_FLAG_MEGAKERNEL))
    assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.parametrize("flags", [0, la.LAM_FLAG_NO_GRAPH, la.LAM_FLAG_NO_TAIL], ids=["graph", "nograph", "notail"])
@pytest.mark.parametrize("spp,offset", [(2, 0), (5, 3), (8, 0)])
def test_pipelined_passes_are_exact(ctx, box_scene, flags, spp, offset):
    """passes alternating between two lanes (streams) splat into the film in
    pass order: the image equals the one-lane render bit for bit"""
    ctx.upload(box_scene)
    a = ctx.render(la.params(spp=spp, max_scatterings=8, sample_p=0.8, hero=4, seed=7, spp_offset=offset,
                             flags=flags))
    b = ctx.render(la.params(spp=spp, max_scatterings=8, sample_p=0.8, hero=4, seed=7, spp_offset=offset,
                             flags=flags | la.LAM_FLAG_NO_PIPE))
    assert np.array_equal(a, b)


@pytest.fixture(scope="module")
def sheets(tmp_path_factory):
    from lambda_amd import scenes
    return la.load_nls(scenes.sheet_stack(str(tmp_path_factory.mktemp("sheets")), n_sheets=8192, width=64))


def test_deep_stacks_bit_exact(ctx, sheets):
    """8,192 parallel sheets along the view axis: a ray down the stack pushes
    7 entries per BVH level, past the 24 held in LDS into the spill columns.
    Closest hits and occlusion stay bit-exact against the oracle; the
    wavefront render (split stacks) equals the megakernel (LDS stacks)."""
    ctx.upload(sheets)
    assert ctx.stats().stack_bound > 24 and ctx.stats().depth >= 6
    orc = ob.Oracle(sheets)
    rng = np.random.default_rng(21)
    n = 60000
    rays = np.zeros(n, dtype=la.RAY_DTYPE)
    rays["o"] = np.stack([rng.uniform(-0.35, 0.35, n), np.full(n, -0.9), rng.uniform(0.65, 1.35, n)], axis=1)
    d = np.stack([rng.normal(0, 0.05, n), np.ones(n), rng.normal(0, 0.05, n)], axis=1)
    rays["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmax"] = 3.4e38
    rays["skip"] = la.LAM_NONE
    o = orc.closest(rays)
    _assert_hits_equal(ctx.extend(rays), o, "rays down the sheet stack")
    sheet_obj = np.bincount(o["obj"][o["obj"] != la.LAM_NONE]).argmax()
    assert (o["obj"] == sheet_obj).mean() > 0.9
    # shadow rays between the sheets' first and last layer
    rays["tmax"] = rng.uniform(0.0, 1.6, n).astype(np.float32)
    assert np.array_equal(ctx.occluded(rays), orc.occluded(rays))
    pr = la.params(spp=3, max_scatterings=6, sample_p=0.8, hero=4, seed=4)
    a = ctx.render(la.params(spp=3, max_scatterings=6, sample_p=0.8, hero=4, seed=4, flags=la.LAM_FLAG_NO_TAIL))
    b = ctx.render(la.params(spp=3, max_scatterings=6, sample_p=0.8, hero=4, seed=4, flags=la.LAM_FLAG_MEGAKERNEL))
    assert np.array_equal(a, b) and np.array_equal(a, ctx.render(pr))
    g, og = ctx.render(pr), orc.render(pr)
    assert np.mean((g - og) ** 2) / np.mean(og ** 2) < 1e-8


def test_deep_bvh_uploads_and_renders(ctx, tmp_path):
    """the glossy interior with a 1 m plate 0.2 m under the ceiling builds a
    BVH whose traversal stack bound is 65 entries (above the 64 an all-LDS
    stack held): it uploads, its hits are bit-exact and every traversal
    kernel (trace queue, finish, megakernel, path hook) agrees"""
    from lambda_amd import scenes
    s = la.load_nls(scenes.glossy_interior(str(tmp_path), width=48, layout="plate", plate=0.5, plate_gap=0.2))
    ctx.upload(s)
    assert ctx.stats().stack_bound > 64
    orc = ob.Oracle(s)
    rng = np.random.default_rng(8)
    rays = random_rays(rng, 50000, np.array([-4.9, -4.9, 0.1]), np.array([4.9, 4.9, 3.9]))
    _assert_hits_equal(ctx.extend(rays), orc.closest(rays), "random rays in the deep-BVH room")
    pr = la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=6)
    a = ctx.render(pr)
    b = ctx.render(la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=6, flags=la.LAM_FLAG_MEGAKERNEL))
    c = ctx.render(la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=6, flags=la.LAM_FLAG_NO_TAIL))
    assert np.array_equal(a, b) and np.array_equal(a, c)
    o = orc.render(pr)
    assert np.mean((a - o) ** 2) / np.mean(o ** 2) < 1e-8
    pix = rng.integers(0, s.width * s.height, 2048).astype(np.uint32)
    smp = rng.integers(0, 16, 2048).astype(np.uint32)
    p1 = la.params(spp=1, max_scatterings=8, sample_p=0.8, hero=4, seed=6)
    _check_paths(ctx.trace_paths(p1, pix, smp), orc.trace_paths(p1, pix, smp))


@pytest.mark.parametrize("n,bits,span", [(1, 30, 8), (5, 30, 4), (2047, 30, 1 << 30), (2048, 8, 256),
                                         (2049, 13, 1 << 13), (100_003, 30, 1 << 30), (100_003, 30, 500),
                                         (1_000_000, 32, 1 << 32), (1_000_000, 30, 1 << 12)])
def test_radix_sort_is_the_stable_sort(ctx, n, bits, span):
    """the device BVH builder's radix sort against numpy's stable argsort:
    partial and whole tiles, few and many distinct keys (equal keys keep
    their input order), keys wider than `bits` (the high bits are ignored)"""
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)[::-1].copy()
    ko, vo = ctx.debug_radix_sort(keys, vals, bits)
    sort_key = keys & np.uint32((1 << bits) - 1) if bits < 32 else keys
    o = np.argsort(sort_key, kind="stable")
    assert np.array_equal(ko, keys[o]) and np.array_equal(vo, vals[o])


@pytest.mark.parametrize("scene_name", ["box_scene", "field_scene", "mixed_scene", "sheets"])
def test_gpu_bvh_same_hits_and_images(ctx, request, scene_name):
    """the BVH8 built on the device (linear BVH, LAM_UPLOAD_GPU_BVH) gives the
    oracle's closest hits bit for bit and the host-built BVH's images"""
    scene = request.getfixturevalue(scene_name)
    pr = la.params(spp=3, max_scatterings=8, sample_p=0.8, hero=4, seed=4)
    ctx.upload(scene)
    host_img = ctx.render(pr)
    host_stats = ctx.stats()
    ctx.upload(scene, gpu_bvh=True)
    st = ctx.stats()
    if scene.n_tris:
        assert 0 < st.n_nodes <= scene.n_tris and 2 <= st.stack_bound <= 128 and st.n_leaves > 0
    assert st.n_tris == host_stats.n_tris
    assert np.array_equal(ctx.render(pr), host_img)
    orc = ob.Oracle(scene)
    rng = np.random.default_rng(13)
    v = scene.tri_v().reshape(-1, 3) if scene.n_tris else np.array([[-20, -20, 0], [20, 20, 24.0]])
    rays = random_rays(rng, 100000, v.min(axis=0), v.max(axis=0))
    g, o = ctx.extend(rays), orc.closest(rays)
    # where the oracle's BVH (the reference's slab test, not widened) culls
    # a zero-thickness sheet that a ray grazes at a triangle edge, the
    # conservative device boxes still find the hit: there the GPU must equal
    # the oracle's brute force
    diff = np.nonzero(np.any(_bits(g) != _bits(o), axis=1))[0]
    assert len(diff) <= 10, len(diff)
    if len(diff):
        _assert_hits_equal(g[diff], orc.closest(rays[diff], brute=True), "rays the oracle's BVH misses")
    rays["tmax"] = rng.uniform(0.0, 3.0, len(rays)).astype(np.float32)
    go, oo = ctx.occluded(rays), orc.occluded(rays)
    d2 = np.nonzero(go != oo)[0]
    assert len(d2) <= 10, len(d2)
    if len(d2):
        assert np.array_equal(go[d2], (orc.closest(rays[d2], brute=True)["obj"] != la.LAM_NONE).astype(np.uint32))


def test_wave_config_keeps_the_image(ctx, box_scene):
    """lanes, tail threshold, bounce cap and trace grids change the schedule,
    not the image or the hits of a ray batch"""
    ctx.upload(box_scene)
    pr = la.params(spp=7, max_scatterings=8, sample_p=0.8, hero=4, seed=11, spp_offset=2)
    base = ctx.render(pr)
    rays = random_rays(np.random.default_rng(31), 50000, [-0.9, -0.9, 0.1], [0.9, 0.9, 1.9])
    hits = ctx.extend(rays)
    default = ctx.wave_config()
    assert default == {"lanes": 3, "tail_threshold": 131072, "tail_div": 16, "max_iters": 12, "trace_per_cu": 16,
                       "batch_per_cu": 24, "finish_per_cu": 10, "shade_per_cu": 3, "samples_per_pass": 32}
    try:
        for cfg in [dict(lanes=1), dict(lanes=2, tail_div=2, max_iters=3, trace_per_cu=1, batch_per_cu=1),
                    dict(lanes=4, tail_threshold=0, trace_per_cu=32, batch_per_cu=32, finish_per_cu=1,
                         shade_per_cu=1),
                    dict(lanes=3, finish_per_cu=32, shade_per_cu=8),
                    dict(lanes=4, tail_threshold=1 << 30, tail_div=1, max_iters=1),
                    dict(samples_per_pass=1), dict(samples_per_pass=3), dict(samples_per_pass=8, lanes=2),
                    dict(samples_per_pass=4, lanes=1)]:
            ctx.set_wave_config(**{**default, **cfg})
            assert np.array_equal(ctx.render(pr), base), cfg
            assert np.array_equal(_bits(ctx.extend(rays)), _bits(hits)), cfg
        for bad in [dict(lanes=0), dict(lanes=5), dict(tail_div=0), dict(max_iters=0), dict(trace_per_cu=0),
                    dict(trace_per_cu=33), dict(batch_per_cu=0), dict(batch_per_cu=33), dict(finish_per_cu=0),
                    dict(shade_per_cu=9), dict(samples_per_pass=0), dict(samples_per_pass=33)]:
            with pytest.raises(la.LambdaError) as e:
                ctx.set_wave_config(**{**default, **bad})
            assert e.value.code == la.LAM_E_INVALID
    finally:
        ctx.set_wave_config(**default)
    assert ctx.wave_config() == default


@pytest.mark.parametrize("spp,cfg,flags,shape", [
    (64, {}, 0, (22, 3)), (64, dict(lanes=2), 0, (32, 2)), (64, {}, la.LAM_FLAG_NO_PIPE, (32, 2)),
    (128, {}, 0, (22, 6)), (256, {}, 0, (29, 9)), (7, {}, 0, (3, 3)), (2, {}, 0, (1, 2)), (1, {}, 0, (1, 1)),
    (64, dict(samples_per_pass=16), 0, (11, 6)), (64, dict(lanes=1), 0, (32, 2))])
def test_pass_shape(ctx, box_scene, spp, cfg, flags, shape):
    """a render of spp samples goes in g >= 2 passes of at most
    samples_per_pass samples, g a multiple of the lanes; the image does not
    depend on it"""
    ctx.upload(box_scene)
    default = ctx.wave_config()
    pr = la.params(spp=spp, max_scatterings=2, sample_p=0.8, hero=4, seed=5)
    try:
        ctx.set_wave_config(**{**default, **cfg})
        img = ctx.render(la.params(spp=spp, max_scatterings=2, sample_p=0.8, hero=4, seed=5, flags=flags))
        t = ctx.timing()
        assert (t.pass_samples, t.passes) == shape
    finally:
        ctx.set_wave_config(**default)
    ctx.set_wave_config(**{**default, "lanes": 1, "samples_per_pass": 1})
    try:
        assert np.array_equal(img, ctx.render(pr))
    finally:
        ctx.set_wave_config(**default)


def test_pipelined_chunks_are_exact(ctx, tmp_path):
    """a frame of more than one chunk of path slots (4,410,000 pixels > 4 Mi):
    both lanes per chunk, the same image as one lane"""
    from lambda_amd import scenes
    s = la.load_nls(scenes.box_room(str(tmp_path), seed=2, width=2100))
    assert s.width * s.height > 1 << 22
    ctx.upload(s)
    a = ctx.render(la.params(spp=3, max_scatterings=6, sample_p=0.8, hero=4, seed=5))
    b = ctx.render(la.params(spp=3, max_scatterings=6, sample_p=0.8, hero=4, seed=5, flags=la.LAM_FLAG_NO_PIPE))
    assert np.isfinite(a).all() and a.mean() > 0
    assert np.array_equal(a, b)


def test_render_against_fp64_truth(ctx, box_scene):
    """GPU at 8x8 spp vs the FP64 oracle at 64 independent spp: per-pixel
    z-test on luminance and relMSE within the noise of both estimates."""
    ctx.upload(box_scene)
    batches = []
    for k in range(8):
        pr = la.params(spp=8, max_scatterings=8, sample_p=0.8, hero=4, seed=2, spp_offset=8 * k)
        batches.append(ctx.render(pr))
    g = np.stack(batches)
    truth = ob.Oracle(box_scene, f64=True)
    tb = [truth.render(la.params(spp=8, max_scatterings=8, sample_p=0.8, hero=4, seed=77, spp_offset=8 * k))
          for k in range(8)]
    t = np.stack(tb)
    lum = np.array([0.2126, 0.7152, 0.0722], np.float32)
    gl, tl = g @ lum, t @ lum
    diff = gl.mean(0) - tl.mean(0)
    se = np.sqrt(gl.var(0, ddof=1) / 8 + tl.var(0, ddof=1) / 8) + 1e-12
    z = diff / se
    assert np.mean(np.abs(z) > 4) < 0.01, np.mean(np.abs(z) > 4)
    assert abs(np.mean(z)) < 0.1, np.mean(z)
    relmse = np.mean((g.mean(0) - t.mean(0)) ** 2) / np.mean(t.mean(0) ** 2)
    noise = (np.mean(g.var(0, ddof=1)) / 8 + np.mean(t.var(0, ddof=1)) / 8) / np.mean(t.mean(0) ** 2)
    assert relmse < 3 * noise, (relmse, noise)


@pytest.mark.parametrize("scene_name", ["box_scene", "field_scene"])
def test_render_counts_rays(ctx, request, scene_name):
    """the rays a render traces (extend queues, shadow queues and the
    finish kernel's own queries) equal the oracle's count of tracePath's
    queries for the same (pixel, sample) paths; the pass schedule does not
    change them"""
    scene = request.getfixturevalue(scene_name)
    ctx.upload(scene)
    orc = ob.Oracle(scene)
    npix = scene.width * scene.height
    counts = {}
    for flags in (0, la.LAM_FLAG_NO_TAIL, la.LAM_FLAG_NO_PIPE, la.LAM_FLAG_NO_GRAPH):
        pr = la.params(spp=3, max_scatterings=8, sample_p=0.8, hero=4, seed=3, flags=flags)
        ctx.reset_timing()
        ctx.render(pr)
        t = ctx.timing()
        counts[flags] = (t.extend_rays, t.shadow_rays)
    assert len(set(counts.values())) == 1, counts
    pix = np.tile(np.arange(npix, dtype=np.uint32), 3)
    smp = np.repeat(np.arange(3, dtype=np.uint32), npix)
    e, s = orc.count_rays(la.params(spp=3, max_scatterings=8, sample_p=0.8, hero=4, seed=3), pix, smp)
    ge, gs = counts[0]
    # a sampling decision on a float boundary may differ on a rare path
    assert abs(ge - e) <= max(2, 1e-4 * e) and abs(gs - s) <= max(2, 1e-4 * s), (ge, e, gs, s)
    assert e >= 3 * npix
    # closed form in the closed room: no roulette and no bounce after the
    # first vertex -> 2 extend rays and 1 shadow ray per path
    if scene_name == "box_scene":
        ctx.reset_timing()
        ctx.render(la.params(spp=2, max_scatterings=0, sample_p=1.0, hero=4, seed=3))
        t = ctx.timing()
        assert (t.extend_rays, t.shadow_rays) == (4 * npix, 2 * npix)


def test_every_light_type_contributes(ctx, lights_scene):
    """point, directional and ambient lights (cg.hpp:1560-1622) on the HIP
    path: each one's light estimate alone (a scene variant with only that
    light, no BSDF continuation) matches the oracle per path, shadowed
    delta lights give exact zeros where the oracle's shadow query is
    blocked, and every type lights some paths"""
    import json
    from tests.scene_text import all_lights_scene_text
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        base = json.loads(all_lights_scene_text(d))
        for name in ("pt", "sun", "sky"):
            s = dict(base)
            s["lights"] = [l for l in base["lights"] if l["name"] == name]
            sc = la.Scene.from_text(json.dumps(s))
            ctx.upload(sc)
            orc = ob.Oracle(sc)
            npix = sc.width * sc.height
            pix = np.arange(npix, dtype=np.uint32)
            smp = np.zeros(npix, np.uint32)
            pr = la.params(spp=1, max_scatterings=0, sample_p=0.0, hero=16, seed=3)
            g, o = ctx.trace_paths(pr, pix, smp), orc.trace_paths(pr, pix, smp)
            _check_paths(g, o)
            lit = (o > 0).any(axis=1)
            assert 0.05 < lit.mean(), name
            assert np.array_equal((g > 0).any(axis=1), lit), name
            if name != "sky":  # delta lights are shadowed somewhere in the closed box
                assert lit.mean() < 0.98, name


@pytest.mark.parametrize("any_hit", [False, True], ids=["closest", "any"])
def test_kernel_counts_equal_host_traversal(ctx, field_scene, any_hit):
    """the host traversal of the BVH8 records (lamh_trace_counts, the count
    behind the roofline's bytes) gives the kernels' results and their
    per-ray node and triangle counts exactly"""
    ctx.upload(field_scene)
    rng = np.random.default_rng(41)
    v = field_scene.tri_v().reshape(-1, 3)
    rays = random_rays(rng, 100000, v.min(axis=0), v.max(axis=0))
    if any_hit:
        rays["tmax"] = rng.uniform(0, 20, len(rays)).astype(np.float32)
    host, hc = la.host_trace(field_scene, rays, any_hit=any_hit)
    n = len(rays)
    d_rays, d_out, d_cnt = la.DeviceBuffer.from_array(rays), la.DeviceBuffer(n * 16), la.DeviceBuffer(n * 8)
    if any_hit:
        ctx.occluded_dev(d_rays.ptr, n, d_out.ptr, d_cnt.ptr)
        got = d_out.download(np.uint32, n)
        assert np.array_equal(got, host) and np.array_equal(got, ctx.occluded(rays))
    else:
        ctx.extend_dev(d_rays.ptr, n, d_out.ptr, d_cnt.ptr)
        got = d_out.download(la.HIT_DTYPE, n)
        assert np.array_equal(_bits(got), _bits(host))
    kc = d_cnt.download(np.uint32, 2 * n).reshape(n, 2)
    assert np.array_equal(kc, hc)
    assert hc[:, 0].mean() > 1


def test_profiled_render_counts_fetches(ctx, box_scene):
    """LAM_FLAG_PROFILE_FETCH counts the render kernels' nodes and
    triangles and changes no pixel"""
    ctx.upload(box_scene)
    base = dict(spp=3, max_scatterings=8, sample_p=0.8, hero=4, seed=2)
    a = ctx.render(la.params(**base))
    ctx.reset_timing()
    b = ctx.render(la.params(flags=la.LAM_FLAG_PROFILE | la.LAM_FLAG_PROFILE_FETCH, **base))
    t = ctx.timing()
    assert np.array_equal(a, b)
    assert t.extend_traced > 0 and t.shadow_traced > 0
    assert t.extend_nodes >= t.extend_traced and t.extend_ms > 0 and t.shadow_ms > 0
