"""Pins of the oracle's material, light, film and camera functions against
what the mathematics fixes, each chosen so that a plausible transcription
error fails: closed forms at hand-picked geometry, normalisation integrals,
chi-square histograms of sampled directions against the density of the
sampling mechanism, the laws of reflection and refraction, and the
reference's own measured values (the flat-spectrum colour and the splat
weights of a probe build of the reference).

The functions under test are the oracle's own (oracle/oracle.c, reached
through the orc_probe_* hooks), i.e. the code trace_path and orc_render
run.  Reference: lib/numli/cg.hpp (C:) and src/renderer.cpp (R:).
"""
import math
import os

import numpy as np
import pytest
from scipy import stats

import lambda_amd as la
from oracle.binding import Oracle

PI = math.pi

MATERIALS = """[
 {"name": "lam", "type": "lambertian", "albedo": {"val": 0.6}},
 {"name": "bd", "type": "blinn", "Kd": {"val": 0.6}},
 {"name": "bs", "type": "blinn", "Ks": {"val": 0.8}, "glossiness": 40},
 {"name": "bs1", "type": "blinn", "Ks": {"val": 1.0}, "glossiness": 12},
 {"name": "bt", "type": "blinn", "Kt": {"val": 0.9}, "ior": {"val": 1.5}, "glossiness": 60},
 {"name": "bt4", "type": "blinn", "Kt": {"val": 0.9}, "ior": {"val": 1.5}, "glossiness": 4},
 {"name": "bf", "type": "blinn", "Kd": {"val": 0.2}, "Ks": {"val": 0.3}, "Kt": {"val": 0.5},
  "ior": {"val": 1.33}, "glossiness": 50},
 {"name": "tf", "type": "thinfilm", "film_ior": {"val": 1.4}, "base_ior": {"val": 1.6},
  "glossiness": 64, "thickness": 500},
 {"name": "tf_eq", "type": "thinfilm", "film_ior": {"val": 1.45}, "base_ior": {"val": 1.45},
  "glossiness": 64, "thickness": 300},
 {"name": "tf_eq2", "type": "thinfilm", "film_ior": {"val": 1.45}, "base_ior": {"val": 1.45},
  "glossiness": 64, "thickness": 650},
 {"name": "tf_air", "type": "thinfilm", "film_ior": {"val": 1.0}, "base_ior": {"val": 1.7},
  "glossiness": 64, "thickness": 420},
 {"name": "tf_a", "type": "thinfilm", "film_ior": {"val": 1.4}, "base_ior": {"val": 1.6},
  "glossiness": 64, "thickness": 200},
 {"name": "tf_b", "type": "thinfilm", "film_ior": {"val": 1.4}, "base_ior": {"val": 1.6},
  "glossiness": 64, "thickness": 574.2}
]"""


def _scene(lights, materials=MATERIALS, objects=None,
           cam='"pos":[0,0,5],"look_at":[0,0,0],"up":[0,1,0],"fov":20,"ar":1,"width":16'):
    objects = objects or '[{"name":"p","type":"plane","transform":{"scale":[50,50,1]},"material":"lam"}]'
    return la.Scene.from_text('{"camera":{%s},"lights":[%s],"objects":%s,"materials":%s}'
                              % (cam, lights, objects, materials))


@pytest.fixture(scope="module")
def mat_oracle():
    s = _scene('{"name":"d","type":"direction","dir":[0,0,-1],"radiant_intensity":{"val":1.0}}')
    o = Oracle(s)
    names = {s.material_name(i): i for i in range(s.desc.n_materials)}
    return o, names


def _frame(rng, n=None):
    """a random orthonormal frame (rows x, y, z) with z = n"""
    if n is None:
        n = rng.normal(size=3)
    z = np.asarray(n, float) / np.linalg.norm(n)
    a = rng.normal(size=3)
    x = a - z * (a @ z)
    x /= np.linalg.norm(x)
    return np.stack([x, np.cross(z, x), z])


def _tilt(F, theta, phi=0.3):
    """unit direction at polar angle theta, azimuth phi in frame F"""
    return (F[0] * math.sin(theta) * math.cos(phi) + F[1] * math.sin(theta) * math.sin(phi)
            + F[2] * math.cos(theta))


def _lam(n, H=4):
    return np.tile(np.linspace(420.0, 700.0, H), (n, 1))


def _schlick(f0, c):
    return f0 + (1 - f0) * (1 - c) ** 5


def _f0(ior):
    return ((ior - 1) / (ior + 1)) ** 2


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _uniform_2d_chi2(u, phi, nu=20, nphi=16):
    """p-value of (u, phi) being uniform on [0,1) x [-pi, pi)"""
    h, _, _ = np.histogram2d(u, phi, bins=[nu, nphi], range=[[0, 1], [-PI, PI]])
    e = len(u) / (nu * nphi)
    chi2 = ((h - e) ** 2 / e).sum()
    return stats.chi2.sf(chi2, nu * nphi - 1)


def _local(F, v):
    return v @ F.T  # components along x, y, z


# ---- Blinn (C:927-993, 1799-1902) ----------------------------------------------

def test_blinn_at_normal_incidence(mat_oracle):
    """i = o = n: f = Kd/π + (Ks + F0·Kt)(α+2)/(8π), F0 = ((1-η)/(1+η))², on
    both sides (η = 1/ior in front, ior behind: F0 is the same)."""
    o, m = mat_oracle
    n = np.array([0.0, 0.0, 1.0])
    for front in (1, 0):
        v = o.bxdf(m["bf"], _lam(1), n, n, n, front=front)
        want = 0.2 / PI + (0.3 + _f0(1.33) * 0.5) * 52 / (8 * PI)
        np.testing.assert_allclose(v, want, rtol=2e-6)


def test_blinn_schlick_at_hand_picked_angles(mat_oracle):
    """i = o at angle θ: h = i, so f = (F(θ)·Kt)(α+2)/(8π)·cos^α θ with
    Schlick F(θ) = F0 + (1-F0)(1-cos θ)^5 (Kd = Ks = 0)."""
    o, m = mat_oracle
    rng = np.random.default_rng(1)
    F = _frame(rng)
    for th in (0.2, 0.7, 1.1, 1.4):
        d = _tilt(F, th)
        v = o.bxdf(m["bt4"], _lam(1), d, d, F[2])[0]
        c = math.cos(th)
        want = _schlick(_f0(1.5), c) * 0.9 * 6 / (8 * PI) * c ** 4
        np.testing.assert_allclose(v, want, rtol=1e-4)


def test_blinn_lobes_integrate_to_their_albedo(mat_oracle):
    """∫ f dω_i over the hemisphere, o = n: the Kd lobe gives Kd; the glossy
    lobe (Ks = 1, Kt = 0) gives 1 - 2^(-(α+2)/2) (half-vector angle θ/2)."""
    o, m = mat_oracle
    rng = np.random.default_rng(2)
    F = _frame(rng)
    nt, nphi = 4000, 8
    th = (np.arange(nt) + 0.5) * (PI / 2) / nt
    ph = (np.arange(nphi) + 0.5) * 2 * PI / nphi
    T, P = np.meshgrid(th, ph, indexing="ij")
    loc = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dirs = loc @ F
    w = (np.sin(T) * (PI / 2 / nt) * (2 * PI / nphi)).reshape(-1)
    N = len(dirs)
    on = np.tile(F[2], (N, 1))
    vd = o.bxdf(m["bd"], _lam(N, 1), dirs, on, on)[:, 0]
    assert abs((vd * w).sum() - 0.6) < 1e-4
    vs = o.bxdf(m["bs1"], _lam(N, 1), dirs, on, on)[:, 0]
    assert abs((vs * w).sum() - (1 - 2 ** (-(12 + 2) / 2))) < 1e-4


@pytest.mark.parametrize("front", [1, 0])
def test_blinn_transmission_peaks_on_snell_direction(mat_oracle, front):
    """For i the Snell refraction of o through the macro surface (sin θi =
    η_rel sin θo with η_rel = 1/ior entering, ior leaving), the
    transmission half vector is the normal, so f = Kt(1-F(o·n))(α+2)/(8π);
    any other i gives less."""
    o, m = mat_oracle
    rng = np.random.default_rng(3)
    F = _frame(rng)
    ior = 1.5
    eta = 1 / ior if front else ior
    for tho in (0.1, 0.4, 0.6):
        od = _tilt(F, tho, 0.9)
        st = math.sin(tho) * eta
        # refracted: opposite side, azimuth opposite to o's
        i = -(F[0] * st * math.cos(0.9) + F[1] * st * math.sin(0.9)) - F[2] * math.sqrt(1 - st * st)
        v = o.bxdf(m["bt"], _lam(1), i, od, F[2], front=front)[0]
        want = 0.9 * (1 - _schlick(_f0(ior), math.cos(tho))) * 62 / (8 * PI)
        np.testing.assert_allclose(v, want, rtol=3e-5)
        off = _unit(i + 0.05 * F[0])
        assert o.bxdf(m["bt"], _lam(1), off, od, F[2], front=front)[0, 0] < want


def test_blinn_diffuse_sampler_is_cosine_distributed(mat_oracle):
    """Kd-only Blinn: cos²θ and φ of the sampled directions are uniform
    (cosine density), prob = p·cosθ/π and mult = Kd/π·cosθ."""
    o, m = mat_oracle
    rng = np.random.default_rng(4)
    F = _frame(rng)
    N, p = 200000, 0.8
    od = _tilt(F, 0.5)
    xi = rng.random((N, 3)).astype(np.float32)
    xi[:, 0] *= p * 0.999
    ok, d, pr, mult = o.sample_material(m["bd"], _lam(N), F, od, xi, p)
    assert ok.all()
    loc = _local(F, d.astype(float))
    assert _uniform_2d_chi2(loc[:, 2] ** 2, np.arctan2(loc[:, 1], loc[:, 0])) > 1e-4
    np.testing.assert_allclose(pr, p * loc[:, 2] / PI, rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(mult[:, 0], 0.6 / PI * loc[:, 2], rtol=1e-4, atol=1e-7)
    assert not o.sample_material(m["bd"], _lam(1), F, od, [[p * 1.001, 0.3, 0.3]], p)[0][0]


def test_lambertian_sampler_is_cosine_distributed(mat_oracle):
    o, m = mat_oracle
    rng = np.random.default_rng(5)
    F = _frame(rng)
    N, p = 200000, 0.7
    od = _tilt(F, 0.3)
    xi = rng.random((N, 3)).astype(np.float32)
    xi[:, 0] *= p
    ok, d, pr, mult = o.sample_material(m["lam"], _lam(N), F, od, xi, p)
    assert ok.all()
    loc = _local(F, d.astype(float))
    assert _uniform_2d_chi2(loc[:, 2] ** 2, np.arctan2(loc[:, 1], loc[:, 0])) > 1e-4
    np.testing.assert_allclose(pr, p * loc[:, 2] / PI, rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(mult[:, 0], 0.6 / PI * loc[:, 2], rtol=1e-4, atol=1e-7)


def test_blinn_reflection_sampler(mat_oracle):
    """Ks-only Blinn (Kt = 0, so every glossy sample reflects): the half
    vector h = (i+o)/|i+o| has cos^(α+1)θh and φh uniform (density
    (α+1)/(2π)·cos^α θh, i = reflect(o, h)); prob = p·(α+1)/(2π)·cos^(α+1)θh/4
    (no 1/(o·h): the reference's pdf) and mult = Ks(α+2)/(8π)cos^α θh."""
    o, m = mat_oracle
    rng = np.random.default_rng(6)
    F = _frame(rng)
    N, p, a = 200000, 0.9, 40.0
    od = _tilt(F, 0.5)
    xi = rng.random((N, 3)).astype(np.float32)
    xi[:, 0] *= p * 0.999
    ok, d, pr, mult = o.sample_material(m["bs"], _lam(N), F, od, xi, p)
    assert ok.all()
    h = _unit(d.astype(float) + od)
    h *= np.sign(h @ F[2])[:, None]
    loc = _local(F, h)
    assert _uniform_2d_chi2(loc[:, 2] ** (a + 1), np.arctan2(loc[:, 1], loc[:, 0])) > 1e-4
    # the law of reflection: i = 2(o·h)h - o
    np.testing.assert_allclose(d, 2 * (h @ od)[:, None] * h - od, atol=2e-5)
    ch = loc[:, 2]
    np.testing.assert_allclose(pr, p * (a + 1) / (2 * PI) * ch ** (a + 1) / 4, rtol=2e-4)
    np.testing.assert_allclose(mult[:, 0], 0.8 * (a + 2) / (8 * PI) * ch ** a, rtol=2e-4)


def test_blinn_transmission_sampler_obeys_snell(mat_oracle):
    """Kt-only Blinn, transmission branch (xl past the reflection share,
    which is F(o·h) <= 0.5 here): t, o and h are coplanar, t is on the far
    side of h with |t×h| = η|o×h| (η = 1/ior); the recovered half vector
    h = -(t + η o)/|t + η o| has cos^(α+1)θh and φh uniform; prob =
    p·(1-F(o·h))·(α+1)/(2π)·cos^(α+1)θh/4; mult = Kt(1-F)(α+2)/(8π)cos^α θh.
    Reflection branch (xl < p·F0): the law of reflection."""
    o, m = mat_oracle
    rng = np.random.default_rng(7)
    F = _frame(rng)
    N, p, a, ior = 200000, 0.9, 60.0, 1.5
    eta = 1 / ior
    od = _tilt(F, 0.7, 2.0)
    xi = rng.random((N, 3)).astype(np.float32)
    xi[:, 0] = p * (0.5 + 0.499 * xi[:, 0])
    ok, t, pr, mult = o.sample_material(m["bt"], _lam(N), F, od, xi, p)
    assert ok.all()
    t = t.astype(float)
    h = -_unit(t + eta * od)
    h *= np.sign(h @ F[2])[:, None]
    assert np.all(t @ F[2] < 0.2) and np.all((t * h).sum(1) < 0)
    # t×h is antiparallel to o×h (t on the far side of h)
    np.testing.assert_allclose(-np.einsum("ij,ij->i", np.cross(t, h), np.cross(od, h)) /
                               np.linalg.norm(np.cross(od, h), axis=1) ** 2, eta, rtol=1e-4)
    assert np.abs(np.einsum("ij,ij->i", np.cross(t, np.broadcast_to(od, t.shape)), h)).max() < 1e-4
    # unit length, and cos θt = √(1 - η² sin² θo) about h
    np.testing.assert_allclose(np.linalg.norm(t, axis=1), 1.0, atol=1e-5)
    np.testing.assert_allclose(-(t * h).sum(1), np.sqrt(1 - eta ** 2 * (1 - (h @ od) ** 2)), atol=2e-5)
    loc = _local(F, h)
    assert _uniform_2d_chi2(loc[:, 2] ** (a + 1), np.arctan2(loc[:, 1], loc[:, 0])) > 1e-4
    Fr = _schlick(_f0(ior), h @ od)
    ch = loc[:, 2]
    np.testing.assert_allclose(pr, p * (1 - Fr) * (a + 1) / (2 * PI) * ch ** (a + 1) / 4, rtol=3e-4)
    np.testing.assert_allclose(mult[:, 0], 0.9 * (1 - Fr) * (a + 2) / (8 * PI) * ch ** a, rtol=3e-4)
    # reflection branch
    xi[:, 0] = p * _f0(ior) * 0.99 * rng.random(N)
    ok, r, pr, mult = o.sample_material(m["bt"], _lam(N), F, od, xi, p)
    h = _unit(r.astype(float) + od)
    np.testing.assert_allclose(r, 2 * (h @ od)[:, None] * h - od, atol=2e-5)
    Fr = _schlick(_f0(ior), h @ od)
    np.testing.assert_allclose(pr, p * Fr * (a + 1) / (2 * PI) * (h @ F[2]) ** (a + 1) / 4, rtol=3e-4)


def test_prob_material_swaps_its_arguments(mat_oracle):
    """probForMateriali passes (i, o) into the (o, i) parameters of
    probForBlinn (C:1971-1985): prob_material(wl=X, o=Y) is the sampler's
    density of Y given the outgoing X.  Exactly so for a Kd-only and a
    Kt-only Blinn material (both branches: the half vector of a transmission
    is rebuilt from the swapped pair); for the full Blinn the other lobe's
    share is added (diffuse p·p_d·cos/π).  Lambertian: no swap."""
    o, m = mat_oracle
    rng = np.random.default_rng(8)
    F = _frame(rng)
    N, p = 4000, 0.8
    X = _tilt(F, 0.6, 1.0)
    lam = _lam(N)
    for name in ("bd", "bt", "lam"):
        xi = rng.random((N, 3)).astype(np.float32)
        xi[:, 0] *= p * 0.999
        ok, Y, pr, _ = o.sample_material(m[name], lam, F, X, xi, p)
        Y = Y[ok]
        if name == "lam":
            got = o.prob_material(m[name], F, Y, X, lam[ok, 0], p)  # Lambertian: (wl, o) as named
        else:
            got = o.prob_material(m[name], F, X, Y, lam[ok, 0], p)
        np.testing.assert_allclose(got, pr[ok], rtol=5e-4, atol=1e-9)
    # the call tracePath makes, prob_material(wl=Y, o=X), is not the density of Y for Kd-only Blinn
    xi = rng.random((N, 3)).astype(np.float32)
    xi[:, 0] *= p * 0.999
    ok, Y, pr, _ = o.sample_material(m["bd"], lam, F, X, xi, p)
    got = o.prob_material(m["bd"], F, Y, X, lam[:, 0], p)
    np.testing.assert_allclose(got, p * (X @ F[2]) / PI, rtol=1e-5)
    # full Blinn, glossy-reflection samples: + the diffuse lobe's share
    mat = o.scene.materials()[m["bf"]]
    xi = rng.random((N, 3)).astype(np.float32)
    xi[:, 0] *= p * mat.p_spec * (_f0(1.33) + mat.w_r / (mat.w_r + mat.w_t)) * 0.5
    ok, Y, pr, _ = o.sample_material(m["bf"], lam, F, X, xi, p)
    up = ok & (Y @ F[2] > 0)
    got = o.prob_material(m["bf"], F, X, Y[up], lam[up, 0], p)
    np.testing.assert_allclose(got, pr[up] + p * mat.p_d * (Y[up] @ F[2]) / PI, rtol=5e-4)


# ---- thin film (C:995-1052, 1905-1949) -------------------------------------------

def _airy(n1, n2, d, lam, theta):
    """|r|² of the two-interface amplitude sum r = (r01 + r12 e^{iδ}) /
    (1 + r01 r12 e^{iδ}) with the reference's Schlick magnitudes r = √R
    and its film-angle term cos_l = √(1-1/n1²)·(1-cos²θ) (C:1007)."""
    c = math.cos(theta)
    R01 = _schlick(((1 - n1) / (1 + n1)) ** 2, c)
    R12 = _schlick(((n1 - n2) / (n1 + n2)) ** 2, c)
    cos_l = math.sqrt(1 - 1 / n1 ** 2) * (1 - c * c)
    e = np.exp(1j * 4 * PI * n1 * d * cos_l / lam)
    r = (math.sqrt(R01) + math.sqrt(R12) * e) / (1 + math.sqrt(R01 * R12) * e)
    return abs(r) ** 2


def test_thin_film_airy_reflectance(mat_oracle):
    """i and o mirrored about n at angle θ (h = n): f = (α+2)/(8π)·|r|²."""
    o, m = mat_oracle
    rng = np.random.default_rng(9)
    F = _frame(rng)
    lam = np.array([[430.0, 520.0, 610.0, 700.0]])
    for th in (0.0, 0.35, 0.8, 1.2, 1.45):
        od = _tilt(F, th, 0.4)
        i = 2 * math.cos(th) * F[2] - od
        v = o.bxdf(m["tf"], lam, i, od, F[2])[0]
        want = np.array([_airy(1.4, 1.6, 500.0, l, th) for l in lam[0]]) * 66 / (8 * PI)
        np.testing.assert_allclose(v, want, rtol=2e-4)


def test_thin_film_limits(mat_oracle):
    """Normal incidence: cos_l = 0, so R does not depend on d; n1 = n2 gives
    R = R01; n1 = 1 gives R = R12.  Grazing: R is periodic in d with period
    λ/(2·n1·cos_l), cos_l = √(1-1/n1²)(1-cos²θ); back faces and i or o
    below the surface give 0."""
    o, m = mat_oracle
    rng = np.random.default_rng(10)
    F = _frame(rng)
    n = F[2]
    lam = np.array([[450.0, 550.0, 650.0, 700.0]])
    c = 66 / (8 * PI)
    a = o.bxdf(m["tf_eq"], lam, n, n, n)[0]
    b = o.bxdf(m["tf_eq2"], lam, n, n, n)[0]
    np.testing.assert_allclose(a, ((1 - 1.45) / (1 + 1.45)) ** 2 * c, rtol=1e-5)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(o.bxdf(m["tf_air"], lam, n, n, n)[0], ((1 - 1.7) / 2.7) ** 2 * c, rtol=1e-5)
    # the angle (near 60°) at which the period at 550 nm is 374.2 nm, the
    # thickness difference of tf_b and tf_a
    cos_l = 550.0 / (2 * 1.4 * 374.2)
    th = math.asin(math.sqrt(cos_l / math.sqrt(1 - 1 / 1.96)))
    assert abs(th - PI / 3) < 1e-3
    od = _tilt(F, th)
    i = 2 * math.cos(th) * n - od
    l1 = np.array([[550.0]])
    ra = o.bxdf(m["tf_a"], l1, i, od, n)[0, 0]
    rb = o.bxdf(m["tf_b"], l1, i, od, n)[0, 0]
    assert abs(ra - rb) < 2e-4 * ra
    # half a period away the phase term flips sign: a different value
    rc = o.bxdf(m["tf"], l1, i, od, n)[0, 0]  # d = 500, 0.8 periods from tf_a
    assert abs(rc - ra) > 0.05 * ra
    assert o.bxdf(m["tf"], lam, i, od, n, front=0).max() == 0
    assert o.bxdf(m["tf"], lam, -i, od, n).max() == 0


def test_thin_film_sampler(mat_oracle):
    """Half vector cos^(α+1) and φ uniform, i = reflect(o, h); prob =
    p·(α+1)/(2π)·(n·h)^(α+1)/4 (n·h, C:1926) while prob_material uses
    (wl·h)^(α+1) (o·h, C:1948); mult = BRDFcosθ(i, o)."""
    o, m = mat_oracle
    rng = np.random.default_rng(11)
    F = _frame(rng)
    N, p, a = 200000, 0.85, 64.0
    od = _tilt(F, 0.4)
    xi = rng.random((N, 3)).astype(np.float32)
    xi[:, 0] *= p * 0.999
    lam = _lam(N)
    ok, d, pr, mult = o.sample_material(m["tf"], lam, F, od, xi, p)
    assert ok.all()
    h = _unit(d.astype(float) + od)
    loc = _local(F, h)
    assert _uniform_2d_chi2(loc[:, 2] ** (a + 1), np.arctan2(loc[:, 1], loc[:, 0])) > 1e-4
    np.testing.assert_allclose(d, 2 * (h @ od)[:, None] * h - od, atol=2e-5)
    np.testing.assert_allclose(pr, p * (a + 1) / (2 * PI) * loc[:, 2] ** (a + 1) / 4, rtol=2e-4)
    np.testing.assert_allclose(mult, o.bxdf(m["tf"], lam, d, od, F[2]), rtol=1e-6, atol=1e-12)
    got = o.prob_material(m["tf"], F, d, od, lam[:, 0], p)
    np.testing.assert_allclose(got, p * (a + 1) / (2 * PI) * (h @ od) ** (a + 1) / 4, rtol=3e-4)
    assert not o.sample_material(m["tf"], _lam(1), F, od, [[0.1, 0.3, 0.3]], p, front=0)[0][0]


# ---- lights (C:1560-1714) and MIS (R:116-145) ---------------------------------------

def _view_down(h=0.5):
    return '"pos":[0,0,%g],"look_at":[0,0,0],"up":[0,1,0],"fov":0.01,"ar":1,"width":4' % h


def _paths(o, n, sample_p=0.0, max_scat=0, hero=4, seed=1):
    p = la.params(spp=1, max_scatterings=max_scat, sample_p=sample_p, hero=hero, seed=seed)
    return o.trace_paths(p, np.zeros(n, np.uint32), np.arange(n, dtype=np.uint32))


def _mean_within(L, want, k=4.5):
    m, s = float(L.mean()), float(L.std()) / math.sqrt(L.size)
    assert abs(m - want) < k * s + 1e-6 * abs(want), (m, want, s)


def test_quad_dir_light_nee_mean():
    """Quad "dir" light (C:1625-1668) over a Lambertian point h below its
    centre: the NEE estimate is L·cos^s·ρ/π·cosθi / ((s+1)/(2π)·cos^s·inv_area)
    = 2Lρ cosθi / ((s+1) inv_area), no 1/r² (non-physical pdf), and inv_area
    = 1/(|c0||c1|) is 4× the [-1,1]² quad's 1/area.  Its mean is checked
    against midpoint quadrature of cos θi over the quad."""
    a, b, h, s_exp, L, rho = 0.25, 0.4, 1.0, 2.0, 3.0, 0.6
    light = ('{"name":"q","type":"quaddir","specular":%g,"radiance":{"val":%g},"transform":'
             '{"scale":[%g,%g,1],"rotate":{"axis":[1,0,0],"degrees":180},"translate":[0,0,%g]}}'
             % (s_exp, L, a, b, h))
    sc = _scene(light, cam=_view_down())
    lt = sc.lights()[0]
    assert abs(lt.inv_area - 1 / (a * b)) < 1e-6 / (a * b)
    o = Oracle(sc)
    Lp = _paths(o, 60000)
    g = (np.arange(400) + 0.5) / 400
    X, Y = np.meshgrid((2 * g - 1) * a, (2 * g - 1) * b)
    ecos = np.mean(h / np.sqrt(X ** 2 + Y ** 2 + h ** 2))
    want = 2 * L * rho / ((s_exp + 1) * lt.inv_area) * ecos
    _mean_within(Lp, want)
    # sample_light's pdf is prob_light's at the sampled direction
    rng = np.random.default_rng(12)
    xi = rng.random((2000, 2))
    pos = np.zeros(3)
    fr = np.eye(3)
    d, pr, _ = o.sample_light(0, _lam(2000), pos, fr, xi)
    np.testing.assert_allclose(o.prob_light(0, pos, d), pr, rtol=1e-5)


def test_point_light_closed_form():
    """Point light (C:1580-1600): L = I·ρ·cosθ/(π d²)."""
    pos = np.array([0.3, -0.4, 2.0])
    sc = _scene('{"name":"pt","type":"point","pos":[0.3,-0.4,2.0],"radiant_intensity":{"val":5.0}}',
                cam=_view_down())
    o = Oracle(sc)
    Lp = _paths(o, 64)
    d2 = float(pos @ pos)
    np.testing.assert_allclose(Lp, 5.0 * 0.6 * (2.0 / math.sqrt(d2)) / (PI * d2), rtol=1e-4)


def test_ambient_light_mean():
    """Ambient light (C:1560-1576): uniform hemisphere NEE without a shadow
    test, pdf 1/(2π): mean ρ·L."""
    sc = _scene('{"name":"am","type":"ambient","radiance":{"val":2.0}}', cam=_view_down())
    Lp = _paths(Oracle(sc), 60000)
    _mean_within(Lp, 0.6 * 2.0)


def test_mis_with_one_sphere_light():
    """NEE + BSDF sampling under the power heuristic (R:116-145) with one
    sphere light and no bounce past the first vertex.  When roulette draws
    no BSDF sample (probability 1-p) tracePath returns the light estimate
    unweighted (R:110), so the mean is the irradiance closed form ρL(r/h)²
    plus (1-p)·∫ f·w_mat dω over the light's cone, with f = ρ/π·L·cosθ and
    w_mat = p_B²/(p_B²+p_L²), p_B = p·cosθ/π, p_L = 1/(2π(1-cos α)).  The
    extra term is 1e-4 of the total at p = 0.8 and is resolved by 4e5
    paths: a weight that does not sum to one with its partner, a pdf at the
    wrong scale or a weighted light estimate on termination each fails."""
    r, h, L, rho = 0.5, 3.0, 4.0, 0.6
    sc = _scene('{"name":"s","type":"sphere","transform":{"scale":[%g,%g,%g],"translate":[0,0,%g]},'
                '"radiance":{"val":%g}}' % (r, r, r, h, L), cam=_view_down(1.0))
    o = Oracle(sc)
    ca = math.sqrt(1 - (r / h) ** 2)
    pL = 1 / (2 * PI * (1 - ca))
    th = (np.arange(200000) + 0.5) * math.acos(ca) / 200000
    dth = math.acos(ca) / 200000
    for sp in (0.8, 0.3):
        pB = sp * np.cos(th) / PI
        w_mat = pB ** 2 / (pB ** 2 + pL ** 2)
        f = rho / PI * L * np.cos(th)
        extra = (1 - sp) * np.sum(f * w_mat * 2 * PI * np.sin(th)) * dth
        want = rho * L * (r / h) ** 2 + extra
        Lp = _paths(o, 400000, sample_p=sp)
        _mean_within(Lp, want)
        assert extra > 8 * float(Lp.std()) / math.sqrt(Lp.size) or sp < 0.5


def test_sphere_light_pdf_is_the_cone_pdf():
    """Cone sampling (C:1671-1714): every sampled direction lies in the cone
    of half-angle asin(r/d) towards the centre, its pdf is
    1/(2π(1-√(1-r²/d²))), and prob_light gives the same value for it."""
    sc = _scene('{"name":"s","type":"sphere","transform":{"scale":[0.5,0.5,0.5],"translate":[0.2,0.1,3]},'
                '"radiance":{"val":1.0}}', cam=_view_down(1.0))
    o = Oracle(sc)
    rng = np.random.default_rng(13)
    xi = rng.random((4000, 2))
    pos = np.zeros(3)
    d, pr, mult = o.sample_light(0, _lam(4000), pos, np.eye(3), xi)
    c = np.array([0.2, 0.1, 3.0])
    dist = np.linalg.norm(c)
    ca = math.sqrt(1 - 0.25 / dist ** 2)
    assert np.all(d @ (c / dist) >= ca - 1e-6)
    np.testing.assert_allclose(pr, 1 / (2 * PI * (1 - ca)), rtol=1e-4)
    np.testing.assert_allclose(o.prob_light(0, pos, d), pr, rtol=1e-4)
    assert np.all(mult == 1.0)


# ---- film (C:306-341) ---------------------------------------------------------

XYZ2RGB = np.array([[3.2404542, -1.5371385, -0.4985314],
                    [-0.9692660, 1.8760108, 0.0415560],
                    [0.0556434, -0.2040259, 1.0572252]])


def _cmf_reference_grid():
    """CIE 1931 CMFs box-filtered onto the reference's 80 coefficients the
    way coefficientλ(sampledλ) does (C:252-272): each table segment adds
    its two end samples' mean times its overlap with [l0, l0+Δ], over Δ.
    Computed here from the data files."""
    delta = 320.0 / 79
    out = []
    for c in "xyz":
        t = np.loadtxt(os.path.join(la.DATA_DIR, f"xyz_{c}.csv"), delimiter=",")
        lo, hi, mean = t[:-1, 0], t[1:, 0], 0.5 * (t[:-1, 1] + t[1:, 1])
        row = []
        for i in range(80):
            l0 = 400 + i * delta
            seg = np.clip(np.minimum(hi, l0 + delta) - np.maximum(lo, l0), 0, None)
            row.append((mean * seg).sum() / delta)
        out.append(row)
    return np.array(out), delta


def test_flat_spectrum_colour_matches_the_reference():
    """coefλ2linRGB of a flat unit spectrum is (136.05, 107.27, 102.14) in
    the reference (measured on a build of it): XYZ = Δ·Σ v_i·CMF_i, not
    normalised by ∫ȳ, then the BT.709 matrix.  The scene's CMF table is the
    box filter of the CIE data files."""
    cmf, delta = _cmf_reference_grid()
    rgb = XYZ2RGB @ (delta * cmf.sum(1))
    np.testing.assert_allclose(rgb, [136.05, 107.27, 102.14], atol=0.01)
    sc = _scene('{"name":"d","type":"direction","dir":[0,0,-1],"radiant_intensity":{"val":1.0}}')
    np.testing.assert_allclose(sc.cmf(), cmf, rtol=2e-3, atol=2e-5)


def test_flat_emitter_renders_to_delta_times_flat_colour():
    """A flat unit emitter filling the view: each hero wavelength carries
    L/prob_λ = Ω and splats weight 1 over its two bins, so the film holds
    Δ per bin (Δ/2 in the two end bins, which receive from one side only);
    the image is Δ × coefλ2linRGB of that spectrum ≈ Δ × (136.05, 107.27,
    102.14)."""
    sc = _scene('{"name":"s","type":"sphere","transform":{"scale":[40,40,40],"translate":[0,0,-45]},'
                '"radiance":{"val":1.0}}',
                objects='[{"name":"p","type":"plane","transform":{"translate":[0,0,200]},"material":"lam"}]',
                cam='"pos":[0,0,5],"look_at":[0,0,0],"up":[0,1,0],"fov":10,"ar":1,"width":32')
    o = Oracle(sc)
    img = o.render(la.params(spp=64, max_scatterings=2, sample_p=0.8, hero=16, seed=2))
    cmf, delta = _cmf_reference_grid()
    c = np.ones(80)
    c[0] = c[-1] = 0.5
    want = delta * (XYZ2RGB @ (delta * (cmf * c).sum(1)))
    px = img.reshape(-1, 3)
    np.testing.assert_allclose(px.mean(0), want, rtol=2e-3)
    # the half-weight end bins (400 and 720 nm) move the colour by < 0.5%
    np.testing.assert_allclose(want / delta, [136.05, 107.27, 102.14], rtol=5e-3)


def test_splat_weights_are_swapped_as_in_the_reference():
    """addHeroλ (C:306-320): λ = 400 + 3.1Δ puts 0.9 into bin 4 and 0.1
    into bin 3 (the tent weights are swapped; measured on the reference);
    λ below 400 goes to bin 0, λ at or above 720 to bin 79."""
    sc = _scene('{"name":"d","type":"direction","dir":[0,0,-1],"radiant_intensity":{"val":1.0}}')
    o = Oracle(sc)
    delta = 320.0 / 79
    b = o.splat([400 + 3.1 * delta], [1.0])
    np.testing.assert_allclose(b[[3, 4]], [0.1, 0.9], atol=2e-5)
    assert b.sum() == pytest.approx(1.0, abs=1e-6) and np.count_nonzero(b) == 2
    b = o.splat([395.0, 725.0, 720.0], [2.0, 3.0, 5.0])
    assert b[0] == 2.0 and b[79] == 8.0 and b.sum() == 10.0


# ---- camera and hero wavelengths (C:1517-1542, 1158-1164, 2039-2063) ------------

def _cam_oracle(up):
    cam = '"pos":[1,2,6],"look_at":[1,2,0],"up":%s,"fov":30,"ar":1.5,"width":48' % up
    sc = _scene('{"name":"d","type":"direction","dir":[0,0,-1],"radiant_intensity":{"val":1.0}}', cam=cam)
    return Oracle(sc), sc


def test_camera_pixel_centre_directions():
    """Pinhole: the ray through film position (u, v) pixels makes angle
    atan(|(u-W/2, v-H/2)|·2 tan(fov/2)/H) with the view axis, H =
    ceil(W/ar) (C:2056); u grows along up×back (right), v along up."""
    o, sc = _cam_oracle("[0,1,0]")
    W, Hh = 48, 32
    assert sc.height == Hh
    rng = np.random.default_rng(14)
    uv = rng.uniform([0, 0], [W, Hh], (500, 2))
    u = np.column_stack([uv, np.full((500, 2), 0.5), np.zeros(500)])
    ro, rd, _ = o.camera(u, 4)
    np.testing.assert_allclose(ro, np.tile([1, 2, 6], (500, 1)), atol=1e-6)
    tan = np.hypot(uv[:, 0] - W / 2, uv[:, 1] - Hh / 2) * 2 * math.tan(math.radians(15)) / Hh
    np.testing.assert_allclose(np.hypot(rd[:, 0], rd[:, 1]) / -rd[:, 2], tan, rtol=2e-5, atol=1e-6)
    assert np.all(np.sign(rd[:, 0]) == np.sign(uv[:, 0] - W / 2))
    assert np.all(np.sign(rd[:, 1]) == np.sign(uv[:, 1] - Hh / 2))


def test_camera_with_oblique_up_zooms():
    """x = up × z is not normalised (C:2058) and y = z × x has the same
    length, so an up vector at angle β to the view axis scales the film by
    sin β: tan of every ray's angle to the axis shrinks by sin β."""
    beta = math.radians(50)
    up = "[0,%r,%r]" % (math.sin(beta), math.cos(beta))
    o1, _ = _cam_oracle("[0,1,0]")
    o2, _ = _cam_oracle(up)
    rng = np.random.default_rng(15)
    u = np.column_stack([rng.uniform([0, 0], [48, 32], (300, 2)), np.full((300, 2), 0.5), np.zeros(300)])
    _, d1, _ = o1.camera(u, 4)
    _, d2, _ = o2.camera(u, 4)
    t1 = np.hypot(d1[:, 0], d1[:, 1]) / -d1[:, 2]
    t2 = np.hypot(d2[:, 0], d2[:, 1]) / -d2[:, 2]
    ok = t1 > 1e-2
    np.testing.assert_allclose(t2[ok] / t1[ok], math.sin(beta), rtol=1e-4)


def test_hero_wavelengths_wrap():
    """λ_k = λ0 + kΩ/H, wrapped into [400, 720] (C:1517-1531): H distinct
    values spaced Ω/H apart cyclically, λ0 = 400 + Ωξ."""
    o, _ = _cam_oracle("[0,1,0]")
    xi = np.array([0.0, 0.3, 0.77, 0.999])
    u = np.column_stack([np.full((4, 2), 10.0), np.full((4, 2), 0.5), xi])
    _, _, lam = o.camera(u, 16)
    np.testing.assert_allclose(lam[:, 0], 400 + 320 * xi, rtol=1e-6)
    for row in lam:
        assert row.min() >= 400 and row.max() <= 720
        s = np.sort(row)
        gaps = np.diff(np.concatenate([s, [s[0] + 320]]))
        np.testing.assert_allclose(gaps, 20.0, atol=2e-3)


def test_spectrum_lookup_is_linear_interpolation(tmp_path):
    """coefficientλ::operator() (C:289-297): linear interpolation of the
    coefficient table with v[i] at inf + iΔ, clamped to the end values —
    np.interp over the grid.  A camera ray that sees an emitter returns its
    radiance at the path's hero wavelengths (R:79-80), taken here from the
    counter RNG's hero number."""
    from lambda_amd import scenes
    from oracle import binding as ob
    lam = np.arange(380.0, 781.0, 5.0)
    scenes.write_csv(str(tmp_path / "wig.csv"), lam, 1.0 + 0.5 * np.sin(lam / 23.0) + 0.002 * (lam - 380))
    sc = _scene('{"name":"s","type":"sphere","transform":{"scale":[40,40,40],"translate":[0,0,-45]},'
                '"radiance":{"val":"%s"}}' % (tmp_path / "wig.csv"),
                objects='[{"name":"p","type":"plane","transform":{"translate":[0,0,200]},"material":"lam"}]',
                cam='"pos":[0,0,5],"look_at":[0,0,0],"up":[0,1,0],"fov":10,"ar":1,"width":16')
    o = Oracle(sc)
    n, seed = 256, 9
    pix = np.arange(n, dtype=np.uint32)
    smp = np.zeros(n, np.uint32)
    Lp = o.trace_paths(la.params(spp=1, max_scatterings=1, sample_p=0.5, hero=16, seed=seed), pix, smp)
    xi = (ob.rng(seed, pix, smp, np.full(n, 4, np.uint32)) >> 8).astype(np.float32) * np.float32(2 ** -24)
    u = np.column_stack([np.full((n, 2), 8.0), np.full((n, 2), 0.5), xi])
    _, _, hero = o.camera(u, 16)
    g = sc.desc.grid
    grid = g.inf + np.arange(g.n) * ((g.sup - g.inf) / (g.n - 1))
    table = sc.spectra()[sc.materials()[sc.objects()[1].material].spec[0]]
    want = np.interp(hero.astype(float), grid, table)
    np.testing.assert_allclose(Lp, want, rtol=2e-6)
    assert np.ptp(table) > 0.5
