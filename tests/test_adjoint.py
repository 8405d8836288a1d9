"""Path-replay adjoint (lam_adjoint): gradients of an image loss with respect
to thin-film thickness, film IOR and Lambertian albedo.

Checks:
  CPU  the oracle's forward-mode tangents (orc_jvp) against central finite
       differences of the FP64 oracle render, Richardson-extrapolated, one
       parameter family at a time -- this pins the checker;
  GPU  lam_adjoint (reverse mode, replayed paths) against the oracle's
       tangents on the same paths: directional derivatives per family and
       every thin-film thickness partial;
  GPU  lam_adjoint against central finite differences of lam_render at
       h = 1e-3 relative;
  GPU  the torch.autograd binding against lam_adjoint;
  GPU  the albedo gradient on a CSV file's samples against central
       differences of lam_render over edits of that file.
"""
import numpy as np
import pytest

import lambda_amd as la
from oracle import binding as ob


@pytest.fixture(scope="module")
def film_scene(tmp_path_factory):
    from lambda_amd import scenes
    p = scenes.box_room_films(str(tmp_path_factory.mktemp("films")), seed=3, width=32)
    return la.load_nls(p)


def _dp(scene):
    return la.DiffParams(scene, d=[f"film{i}" for i in range(64)], ior=["film3", "film10", "film27"],
                         albedo=list(range(16)))


def _families(dp):
    """index sets of the three parameter families"""
    d = dp.mat_d[dp.mat_d >= 0]
    ior = np.concatenate([np.arange(o, o + dp.grid_n) for o in dp.mat_ior[dp.mat_ior >= 0]])
    alb = np.concatenate([np.arange(o, o + dp.grid_n) for o in dp.mat_albedo[dp.mat_albedo >= 0]])
    return {"d": d, "ior": ior, "albedo": alb}


def _direction(rng, dp, theta, idx):
    v = np.zeros(dp.n_params)
    v[idx] = rng.normal(size=len(idx)) * np.abs(theta[idx])
    return v


def _effective(theta, v, h):
    """float32 parameters at θ ± h v and the direction the float32 step took"""
    tp = (theta + h * v).astype(np.float32)
    tm = (theta - h * v).astype(np.float32)
    return tp, tm, (tp.astype(np.float64) - tm.astype(np.float64)) / (2 * h)


# ---- CPU: the checker ----------------------------------------------------------------

def test_diff_params_layout(film_scene):
    dp = _dp(film_scene)
    gn = film_scene.desc.grid.n
    assert dp.n_params == 64 + 3 * gn + 16 * gn
    th = dp.values(film_scene)
    mats = film_scene.materials()
    assert th[dp.mat_d[16]] == mats[16].d
    assert np.array_equal(th[dp.mat_albedo[0]:dp.mat_albedo[0] + gn], film_scene.spectra()[mats[0].spec[0]])
    v = dp.apply(film_scene, th * 1.25)
    assert np.allclose(dp.values(v), (th * 1.25).astype(np.float32))
    # every differentiated spectrum has a row of its own
    rows = [m.spec[0] for i, m in enumerate(v.materials()) if dp.mat_ior[i] >= 0 or dp.mat_albedo[i] >= 0]
    assert len(set(rows)) == len(rows) and min(rows) >= film_scene.spectra().shape[0]
    with pytest.raises(ValueError):
        la.DiffParams(film_scene, d=[0])            # Lambertian has no thickness
    with pytest.raises(ValueError):
        la.DiffParams(film_scene, albedo=["film0"])  # thin film has no albedo
    with pytest.raises(ValueError):
        la.DiffParams(film_scene, d=["film0", 16])   # listed twice


def test_oracle_tangents_match_fd_fp64(film_scene):
    """orc_jvp against (L(θ+hv) - L(θ-hv))/2h of the FP64 oracle, Richardson
    over h = 2e-3, 1e-3 (relative), along the float32 step actually taken"""
    dp = _dp(film_scene)
    th = dp.values(film_scene).astype(np.float64)
    p = la.params(spp=2, max_scatterings=4, sample_p=0.8, hero=4, seed=7)
    rng = np.random.default_rng(11)
    W = rng.normal(size=(film_scene.height, film_scene.width, 3)).astype(np.float32)
    O = ob.Oracle(film_scene, f64=True)

    def loss(t):
        return float((ob.Oracle(dp.apply(film_scene, t), f64=True).render(p).astype(np.float64) * W).sum())

    for name, idx in _families(dp).items():
        v = _direction(rng, dp, th, idx)
        fd, jv = [], []
        for h in (2e-3, 1e-3):
            tp, tm, ve = _effective(th, v, h)
            fd.append((loss(tp) - loss(tm)) / (2 * h))
            jv.append(O.jvp(p, dp, ve, W)[0])
        rich = (4 * fd[1] - fd[0]) / 3
        assert abs(jv[1]) > 1e-3, name
        assert abs(rich - jv[1]) <= 1e-3 * abs(jv[1]), (name, fd, rich, jv)


def test_oracle_tangents_linear(film_scene):
    """the tangent is linear in the direction: J(a u + b w) = a J(u) + b J(w)"""
    dp = _dp(film_scene)
    th = dp.values(film_scene).astype(np.float64)
    p = la.params(spp=1, max_scatterings=3, sample_p=0.8, hero=4, seed=2)
    rng = np.random.default_rng(5)
    W = rng.normal(size=(film_scene.height, film_scene.width, 3)).astype(np.float32)
    u = rng.normal(size=dp.n_params) * np.abs(th)
    w = rng.normal(size=dp.n_params) * np.abs(th)
    J = ob.Oracle(film_scene, f64=True).jvp(p, dp, np.stack([u, w, 2 * u - 3 * w]), W)
    assert abs(J[2] - (2 * J[0] - 3 * J[1])) <= 1e-9 * (abs(J[0]) + abs(J[1]))


# ---- GPU ----------------------------------------------------------------------------

@pytest.fixture(scope="module")
def ctx():
    c = la.Context(0)
    yield c
    c.close()


def _rel_close(a, b, rtol, atol):
    return abs(a - b) <= rtol * abs(b) + atol


@pytest.mark.gpu
@pytest.mark.parametrize("hero", [4, 16])
def test_adjoint_matches_oracle_tangents(ctx, film_scene, hero):
    """reverse mode on the GPU against forward mode on the CPU, same paths"""
    dp = _dp(film_scene)
    th = dp.values(film_scene).astype(np.float64)
    p = la.params(spp=4 if hero == 4 else 2, max_scatterings=6, sample_p=0.8, hero=hero, seed=21)
    rng = np.random.default_rng(hero)
    W = rng.normal(size=(film_scene.height, film_scene.width, 3)).astype(np.float32)
    ctx.upload(film_scene)
    g = ctx.adjoint(p, dp, W).astype(np.float64)
    O = ob.Oracle(film_scene, f64=False)
    fams = _families(dp)
    V = np.stack([_direction(rng, dp, th, idx) for idx in fams.values()])
    J = O.jvp(p, dp, V, W)
    for (name, _), v, j in zip(fams.items(), V, J):
        assert abs(j) > 0, name
        assert _rel_close(float(g @ v), j, 2e-3, 0.0), (name, float(g @ v), j)
    # every thickness partial
    d_idx = fams["d"]
    Jd = np.concatenate([O.jvp(p, dp, np.eye(dp.n_params)[d_idx[k:k + 16]], W) for k in range(0, 64, 16)])
    scale = np.abs(Jd).max()
    bad = [(i, g[d_idx[i]], Jd[i]) for i in range(64) if not _rel_close(g[d_idx[i]], Jd[i], 2e-3, 1e-4 * scale)]
    assert not bad, bad[:5]


@pytest.mark.gpu
def test_adjoint_matches_central_fd(ctx, film_scene):
    """directional derivatives against central differences of lam_render,
    h = 1e-3 relative (bar 1e-2: the O(h²) term of the thin-film phase) and
    Richardson over h = 2e-3, 1e-3 (bar 1e-3)"""
    dp = _dp(film_scene)
    th = dp.values(film_scene).astype(np.float64)
    p = la.params(spp=4, max_scatterings=6, sample_p=0.8, hero=4, seed=5)
    rng = np.random.default_rng(8)
    W = rng.uniform(0.5, 1.5, size=(film_scene.height, film_scene.width, 3)).astype(np.float32)

    def loss(t):
        ctx.upload(dp.apply(film_scene, t))
        return float((ctx.render(p).astype(np.float64) * W).sum())

    ctx.upload(film_scene)
    g = ctx.adjoint(p, dp, W).astype(np.float64)
    for name, idx in _families(dp).items():
        v = _direction(rng, dp, th, idx)
        fd, ad = [], []
        for h in (2e-3, 1e-3):
            tp, tm, ve = _effective(th, v, h)
            fd.append((loss(tp) - loss(tm)) / (2 * h))
            ad.append(float(g @ ve))
        rich = (4 * fd[1] - fd[0]) / 3
        assert abs(fd[1] - ad[1]) <= 1e-2 * abs(ad[1]), (name, fd, ad)
        assert abs(rich - ad[1]) <= 1e-3 * abs(ad[1]), (name, fd, rich, ad)


@pytest.mark.gpu
def test_adjoint_linear_in_dloss(ctx, film_scene):
    """the gradient is linear in ∂loss/∂rgb (sums of replays per pixel)"""
    dp = _dp(film_scene)
    p = la.params(spp=2, max_scatterings=4, sample_p=0.8, hero=4, seed=1)
    rng = np.random.default_rng(3)
    shape = (film_scene.height, film_scene.width, 3)
    A = rng.normal(size=shape).astype(np.float32)
    B = rng.normal(size=shape).astype(np.float32)
    ctx.upload(film_scene)
    ga, gb, gab = ctx.adjoint(p, dp, A), ctx.adjoint(p, dp, B), ctx.adjoint(p, dp, A + B)
    scale = np.abs(gab).max()
    assert np.allclose(gab, ga + gb, rtol=1e-4, atol=1e-5 * scale)
    assert np.array_equal(ctx.adjoint(p, dp, np.zeros(shape, np.float32)), np.zeros(dp.n_params, np.float32))


@pytest.mark.gpu
def test_adjoint_lanes_agree(ctx, film_scene):
    """forward passes and replays spread over the pass lanes give the
    one-lane gradient up to the order of the float atomics, also with more
    samples than lanes and an odd sample count"""
    dp = _dp(film_scene)
    rng = np.random.default_rng(4)
    W = rng.normal(size=(film_scene.height, film_scene.width, 3)).astype(np.float32)
    ctx.upload(film_scene)
    for spp in (1, 5):
        a = ctx.adjoint(la.params(spp=spp, max_scatterings=6, sample_p=0.8, hero=4, seed=2), dp, W)
        b = ctx.adjoint(la.params(spp=spp, max_scatterings=6, sample_p=0.8, hero=4, seed=2,
                                  flags=la.LAM_FLAG_NO_PIPE), dp, W)
        assert np.count_nonzero(a) > 0
        assert np.allclose(a, b, rtol=1e-4, atol=1e-5 * np.abs(b).max()), spp


@pytest.mark.gpu
def test_adjoint_argument_errors(ctx, film_scene):
    dp = _dp(film_scene)
    p = la.params(spp=1, max_scatterings=2, sample_p=0.8, hero=4)
    ctx.upload(film_scene)
    W = np.ones((film_scene.height, film_scene.width, 3), np.float32)
    bad = _dp(film_scene)
    bad.mat_d = bad.mat_d.copy()
    bad.mat_d[16] = dp.n_params  # slot out of range
    with pytest.raises(la.LambdaError) as e:
        ctx.adjoint(p, bad, W)
    assert e.value.code == la.LAM_E_INVALID
    big = _dp(film_scene)
    big.n_params = 1 << 20
    with pytest.raises(la.LambdaError) as e:
        ctx.adjoint(p, big, W)
    assert e.value.code == la.LAM_E_LIMIT


@pytest.mark.gpu
def test_autograd_binding(ctx, film_scene):
    torch = pytest.importorskip("torch")
    from lambda_amd import autograd as lag
    dp = _dp(film_scene)
    p = la.params(spp=2, max_scatterings=4, sample_p=0.8, hero=4, seed=9)
    th0 = dp.values(film_scene)
    rng = np.random.default_rng(4)
    th_target = (th0 * (1 + rng.uniform(-0.15, 0.15, th0.shape))).astype(np.float32)
    ctx.upload(dp.apply(film_scene, th_target))
    target = torch.from_numpy(ctx.render(p))
    theta = torch.tensor(th0, requires_grad=True)
    rgb = lag.render(ctx, film_scene, dp, p, theta)
    loss = ((rgb - target) ** 2).sum()
    loss.backward()
    ctx.upload(dp.apply(film_scene, th0))
    direct = ctx.adjoint(p, dp, (2 * (rgb.detach() - target)).numpy())
    # float atomics: the order of the per-block sums varies between runs
    assert np.allclose(theta.grad.numpy(), direct, rtol=1e-4, atol=1e-6 * np.abs(direct).max())
    # a small step against the gradient lowers the loss
    with torch.no_grad():
        g = theta.grad
        step = 1e-3 * torch.abs(theta) * torch.sign(g)
        l2 = ((lag.render(ctx, film_scene, dp, p, theta - step) - target) ** 2).sum()
    assert float(l2) < loss.item()


@pytest.mark.gpu
def test_adjoint_on_csv_samples_matches_fd(ctx, tmp_path):
    """the albedo gradient mapped onto the samples of the material's CSV file
    (lamh_spectrum_grad_to_samples: the loader's box filter transposed)
    against central differences of lam_render after rewriting those samples
    in the file and loading the scene again, h = 1e-3 relative"""
    from lambda_amd import scenes
    path = scenes.box_room(str(tmp_path), seed=1, width=32)
    scene = la.load_nls(path)
    m = 0
    csv = tmp_path / f"albedo_{scene.material_name(m)[len('mat_'):]}.csv"
    data = np.loadtxt(csv, delimiter=",")
    lam, val = data[:, 0].astype(np.float32), data[:, 1]
    dp = la.DiffParams(scene, albedo=[m])
    pr = la.params(spp=8, max_scatterings=6, sample_p=0.8, hero=4, seed=5)
    rng = np.random.default_rng(12)
    W = rng.uniform(0.5, 1.5, size=(scene.height, scene.width, 3)).astype(np.float32)
    ctx.upload(scene)
    g = ctx.adjoint(pr, dp, W)
    o = int(dp.mat_albedo[m])
    gs = la.spectrum_grad_to_samples(scene.desc.grid, lam, g[o:o + scene.desc.grid.n]).astype(np.float64)
    v = rng.normal(size=len(val)) * val

    def loss(vals):
        scenes.write_csv(str(csv), lam, vals)
        ctx.upload(la.load_nls(path))
        return float((ctx.render(pr).astype(np.float64) * W).sum())

    h = 1e-3
    vp, vm = val + h * v, val - h * v
    # the direction the file's float32 values actually moved
    ve = (np.float32(vp).astype(np.float64) - np.float32(vm).astype(np.float64)) / (2 * h)
    fd = (loss(vp) - loss(vm)) / (2 * h)
    ad = float(gs @ ve)
    assert abs(ad) > 0
    assert abs(fd - ad) <= 1e-2 * abs(ad), (fd, ad)


@pytest.mark.gpu
def test_adjoint_rejects_maps_of_another_scene(ctx, film_scene, box_scene):
    """parameter maps sized for another scene's materials: a ValueError in
    Python, LAM_E_INVALID through the C ABI (no read past the maps)"""
    import ctypes as C
    p = la.params(spp=1, max_scatterings=2, sample_p=0.8, hero=4)
    ctx.upload(film_scene)
    W = np.ones((film_scene.height, film_scene.width, 3), np.float32)
    other = la.DiffParams(box_scene, albedo=[0])
    with pytest.raises(ValueError):
        ctx.adjoint(p, other, W)
    cs = other.c_struct()
    grad = np.zeros(other.n_params, np.float32)
    rc = ctx._lib.lam_adjoint(ctx.handle, C.byref(p), C.byref(cs), W.ctypes.data_as(C.POINTER(C.c_float)),
                              grad.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == la.LAM_E_INVALID


@pytest.mark.gpu
def test_trace_paths_lengths_must_match(ctx, film_scene):
    ctx.upload(film_scene)
    p = la.params(spp=1, max_scatterings=2, sample_p=0.8, hero=4)
    with pytest.raises(ValueError):
        ctx.trace_paths(p, np.arange(10, dtype=np.uint32), np.zeros(9, np.uint32))
