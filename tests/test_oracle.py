"""The CPU oracle pinned against things that do not depend on it: libm,
closed-form lighting, the reference's stated formulas, and its own FP64
truth mode."""
import math

import numpy as np
import pytest

import lambda_amd as la
from oracle import binding as ob
from oracle.binding import Oracle


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, np.int64(-2 ** 31) - a, a)
    b = np.where(b < 0, np.int64(-2 ** 31) - b, b)
    return np.abs(a - b)


def test_math_sin_cos_within_one_ulp():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-60, 60, 20000), rng.uniform(-1e-3, 1e-3, 2000)]).astype(np.float32)
    for fn, ref in ((0, math.sin), (1, math.cos)):
        got = ob.math(fn, x)
        want = np.array([ref(float(v)) for v in x], dtype=np.float32)
        assert _ulps(got, want).max() <= 1, fn


def test_math_pow_within_one_ulp():
    rng = np.random.default_rng(1)
    b = rng.uniform(1e-3, 1, 20000).astype(np.float32)
    e = np.concatenate([rng.uniform(0, 2048, 10000), rng.integers(1, 1025, 10000)]).astype(np.float32)
    got = ob.math(2, b, e)
    want = np.array([math.pow(float(u), float(v)) for u, v in zip(b, e)])
    ok = want > 1e-37  # normal range
    assert _ulps(got[ok], want[ok].astype(np.float32)).max() <= 1
    # special cases (C99 pow): negative base with integer / non-integer exponent
    got = ob.math(2, np.array([-0.5, -0.5, 0.0, 0.0, 1.0, 2.0], np.float32),
                  np.array([5.0, 2.5, 3.0, -2.0, 1e30, 0.0], np.float32))
    assert got[0] == np.float32(-0.03125) and math.isnan(got[1]) and got[2] == 0 and math.isinf(got[3])
    assert got[4] == 1 and got[5] == 1


def test_math_atan2_within_one_ulp():
    rng = np.random.default_rng(2)
    y, x = rng.normal(size=(2, 20000)).astype(np.float32)
    got = ob.math(3, y, x)
    want = np.array([math.atan2(float(u), float(v)) for u, v in zip(y, x)], dtype=np.float32)
    assert _ulps(got, want).max() <= 1
    z = ob.math(3, np.array([0.0, 0.0, 1.0, -1.0], np.float32), np.array([1.0, -1.0, 0.0, 0.0], np.float32))
    np.testing.assert_allclose(z, [0, math.pi, math.pi / 2, -math.pi / 2], rtol=1e-7)


def test_rng_uniform_and_decorrelated():
    n = 200000
    pix = np.arange(n, dtype=np.uint32) % 5000
    smp = np.arange(n, dtype=np.uint32) // 5000
    u = ob.rng(7, pix, smp, np.zeros(n, np.uint32)).astype(np.float64) / 2 ** 32
    v = ob.rng(7, pix, smp, np.ones(n, np.uint32)).astype(np.float64) / 2 ** 32
    hist, _ = np.histogram(u, bins=64, range=(0, 1))
    chi2 = np.sum((hist - n / 64) ** 2 / (n / 64))
    assert chi2 < 120  # 63 dof
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.01
    w = ob.rng(8, pix, smp, np.zeros(n, np.uint32)).astype(np.float64) / 2 ** 32
    assert abs(np.corrcoef(u, w)[0, 1]) < 0.01


def _plane_scene(light, sample_p_materials='[{"name":"m","type":"lambertian","albedo":{"val":0.6}}]',
                 cam='"pos":[0,0,5],"look_at":[0,0,0],"up":[0,1,0],"fov":20,"ar":1,"width":16'):
    return ('{"camera":{%s},"lights":[%s],'
            '"objects":[{"name":"p","type":"plane","transform":{"scale":[50,50,1]},"material":"m"}],'
            '"materials":%s}' % (cam, light, sample_p_materials))


def test_directional_light_on_lambertian_closed_form():
    """SAMPLE_P = 0: one NEE estimate per path, L = I ρ cosθ / π exactly
    (dirlight, cg.hpp:1604-1621; lambertian, cg.hpp:924-925)."""
    d = np.array([0.3, -0.2, -1.0])
    s = la.Scene.from_text(_plane_scene('{"name":"d","type":"direction","dir":[0.3,-0.2,-1.0],'
                                        '"radiant_intensity":{"val":2.0}}'))
    o = Oracle(s)
    p = la.params(spp=1, max_scatterings=4, sample_p=0.0, hero=16, seed=3)
    pix = np.arange(256, dtype=np.uint32)
    L = o.trace_paths(p, pix, np.zeros_like(pix))
    cos = -d[2] / np.linalg.norm(d)  # i·n with i = -dir (unnormalised in the reference)
    want = 2.0 * 0.6 / math.pi * (-d[2])
    np.testing.assert_allclose(L, want, rtol=1e-6)
    assert cos > 0


def test_sphere_light_irradiance_converges():
    """Cone sampling of a sphere light (cg.hpp:1671-1700): a Lambertian point
    right below a sphere of radius r at height h reflects ρ L (r/h)^2."""
    s = la.Scene.from_text(_plane_scene('{"name":"s","type":"sphere","transform":{"scale":[0.5,0.5,0.5],'
                                        '"translate":[0,0,3]},"radiance":{"val":4.0}}',
                                        cam='"pos":[0,0,1],"look_at":[0,0,0],"up":[0,1,0],"fov":0.01,"ar":1,"width":4'))
    o = Oracle(s)
    p = la.params(spp=1, max_scatterings=0, sample_p=0.0, hero=4, seed=1)
    n = 40000
    L = o.trace_paths(p, np.zeros(n, np.uint32), np.arange(n, dtype=np.uint32))
    want = 0.6 * 4.0 * (0.5 / 3.0) ** 2
    assert abs(L.mean() - want) < 4 * L.std() / math.sqrt(L.size) + 1e-6


def test_emitter_seen_directly_returns_radiance():
    """A path whose camera ray hits an emitter returns Radiance(λ) of the
    hero wavelengths (src/renderer.cpp:79-80); with a linear spectrum this
    pins hero sampling on [inf, sup] (cg.hpp:1517-1531)."""
    s = la.Scene.from_text(_plane_scene('{"name":"s","type":"sphere","transform":{"scale":[2,2,2],'
                                        '"translate":[0,0,0]},"radiance":{"val":1.0}}',
                                        cam='"pos":[0,0,10],"look_at":[0,0,0],"up":[0,1,0],"fov":2,"ar":1,"width":8'))
    o = Oracle(s)
    L = o.trace_paths(la.params(spp=1, max_scatterings=2, sample_p=0.5, hero=16), np.arange(64, dtype=np.uint32),
                      np.zeros(64, np.uint32))
    assert np.all(L == 1.0)


def test_fp32_and_fp64_oracles_agree(box_scene):
    p = la.params(spp=2, max_scatterings=8, sample_p=0.8, hero=4, seed=4)
    a = Oracle(box_scene).render(p)
    b = Oracle(box_scene, f64=True).render(p)
    assert np.mean((a - b) ** 2) / np.mean(b ** 2) < 1e-5


def test_render_row_subset_matches_full(box_scene):
    o = Oracle(box_scene)
    p = la.params(spp=1, max_scatterings=4, sample_p=0.8, hero=4, seed=5)
    full = o.render(p)
    part = o.render(p, row_step=4, row_phase=1)
    h = box_scene.height
    rows = [h - 1 - i for i in range(h) if i % 4 == 1]  # image rows are flipped
    np.testing.assert_array_equal(part[rows], full[rows])
