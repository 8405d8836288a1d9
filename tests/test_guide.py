"""Learned path guiding: the network (MFMA inference, training gradient,
Adam), and the guided render (unbiased against the plain render loop).

CPU:  the numpy oracle's hand-written backward against torch float64
      autograd of the same forward; the vMF mixture integrates to 1.
GPU:  lam_guide_infer vs the oracle forward (same bf16 roundings);
      lam_guide_loss_grad vs the oracle gradient (one chunk per block,
      several, a ragged last chunk); the training forward is the inference
      bit for bit; Adam rewrites the packed network in place as a fresh pack
      would, deterministically; Adam's first step; training fits a known
      distribution; lam_render_guided with α = 0 is lam_render; the guided
      image has the plain image's and the FP64 oracle's expectation; the
      trained network lowers the variance.
"""
import numpy as np
import pytest

import lambda_amd as la
from oracle import guide as og


def _rows(rng, n, theta=None):
    inp = rng.uniform(0, 1, (n, 6))
    om = rng.normal(size=(n, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    w = rng.exponential(1.0, n)
    return np.concatenate([inp, om, w[:, None]], axis=1).astype(np.float32)


def _theta(seed):
    """parameters drawn like lam_guide_reset's (scale only; values differ)"""
    rng = np.random.default_rng(seed)
    P = {}
    for name, shape in og._SIZES:
        fan = shape[1] if len(shape) == 2 else 1
        P[name] = rng.uniform(-1, 1, shape) * (np.sqrt(6.0 / fan) if len(shape) == 2 else 0.1)
    P["W5"] *= 0.3
    return og.pack(P).astype(np.float32)


# ---- CPU ---------------------------------------------------------------------------

def test_oracle_backward_matches_autograd():
    torch = pytest.importorskip("torch")
    th = _theta(1).astype(np.float64)
    rows = _rows(np.random.default_rng(2), 300).astype(np.float64)
    loss, g = og.loss_grad(th, rows)

    T = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    parts, off = {}, 0
    for name, shape in og._SIZES:
        n = int(np.prod(shape))
        parts[name] = T[off:off + n].reshape(shape)
        off += n

    def st(x, rounded):  # straight-through rounding
        return x + (torch.as_tensor(rounded, dtype=torch.float64) - x).detach()

    h = torch.as_tensor(og.features(rows[:, :6]), dtype=torch.float64)
    for l in range(1, 6):
        W = parts[f"W{l}"]
        Wq = st(W, og.bf16(W.detach().numpy()))
        b = parts[f"b{l}"]
        bq = st(b, b.detach().numpy().astype(np.float32).astype(np.float64))
        z = h @ Wq.T + bq
        if l < 5:
            a = torch.relu(z)
            h = st(a, og.bf16(a.detach().numpy()))
    r = z.reshape(-1, 8, 5)
    mu = r[:, :, :3] / r[:, :, :3].norm(dim=2, keepdim=True)
    lk = r[:, :, 3]
    kap = torch.exp(torch.clamp(lk, og.LOGK_MIN, og.LOGK_MAX))
    logpi = torch.log_softmax(r[:, :, 4], dim=1)
    om = torch.as_tensor(rows[:, 6:9])
    md = (mu * om[:, None, :]).sum(dim=2)
    lv = torch.log(kap) - np.log(2 * np.pi) - torch.log(-torch.expm1(-2 * kap)) + kap * (md - 1)
    logp = torch.logsumexp(logpi + lv, dim=1)
    w = torch.as_tensor(rows[:, 9])
    L = -(w / w.sum() * logp).sum()
    L.backward()
    assert abs(L.item() - loss) <= 1e-12 * abs(loss)
    gt = T.grad.numpy()
    assert np.allclose(g, gt, rtol=1e-9, atol=1e-12 * np.abs(gt).max())


def test_mixture_normalised():
    """∫ p dω = 1 on a Fibonacci sphere (κ up to ~60)"""
    rng = np.random.default_rng(4)
    N = 400000
    k = np.arange(N) + 0.5
    z = 1 - 2 * k / N
    phi = np.pi * (1 + 5 ** 0.5) * k
    s = np.sqrt(1 - z * z)
    dirs = np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
    for _ in range(3):
        raw = rng.normal(size=40)
        raw[3::5] = rng.uniform(-3, 4.1, 8)
        p = og.mix_pdf(np.repeat(raw[None, :], N, axis=0), dirs)
        assert abs(p.mean() * 4 * np.pi - 1) < 2e-3


def test_features_layout():
    """features 24d + 2j + c are sin/cos(π 2^j x_d), then the raw inputs"""
    x = np.array([[0.1, 0.25, 0.5, 0.75, 0.9, 0.3]], np.float32)
    F = og.features(x)[0]
    assert F.shape == (150,)
    assert F[144:].tolist() == og.bf16(x[0]).tolist()
    assert abs(F[24 * 2 + 2 * 0] - np.sin(np.pi * 0.5)) < 1e-2       # sin(π x_2), x_2 = 0.5
    assert abs(F[24 * 1 + 2 * 1 + 1] - np.cos(2 * np.pi * 0.25)) < 1e-2  # cos(2π x_1)


def test_features_match_sin_cos():
    """every encoded feature (quarter-turn polynomial at j = 0, 4, 8, then
    double-angle steps) is sin/cos(π 2^j x_d) of the float32 input to within
    half a bf16 step"""
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(0, 1, (20000, 6)), [[0, 1, 0.5, 0.25, 1 - 2 ** -24, 2 ** -20]]]).astype(np.float32)
    F = og.features(x)
    xd = x.astype(np.float64)
    for f in range(144):
        d, j = f // 24, (f % 24) // 2
        a = np.pi * 2.0 ** j * xd[:, d]
        want = np.cos(a) if f % 2 else np.sin(a)
        # half a bf16 step in the binade [2^e, 2^(e+1)) is 2^(e-8) <= 2^-8 |v|
        assert np.all(np.abs(F[:, f] - want) <= 2.0 ** -8 * np.maximum(np.abs(want), 2.0 ** -6) + 2e-6), f


# ---- GPU ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def ctx():
    c = la.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_infer_matches_oracle(ctx):
    th = _theta(3)
    ctx.guide_set_params(th)
    assert np.array_equal(ctx.guide_get_params(), th)
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(0, 1, (4093, 6)), np.zeros((1, 6)), np.ones((1, 6))]).astype(np.float32)
    g = ctx.guide_infer(x).astype(np.float64)
    o = og.forward(th, x)
    scale = np.abs(o).max()
    err = np.abs(g - o)
    # MFMA fp32 sums in another order; a bf16 activation can round the other way
    assert np.quantile(err, 0.99) <= 1e-3 * scale, np.quantile(err, 0.99) / scale
    assert err.max() <= 3e-2 * scale, err.max() / scale


@pytest.mark.gpu
def test_loss_grad_matches_oracle(ctx):
    th = _theta(6)
    ctx.guide_set_params(th)
    rows = _rows(np.random.default_rng(7), 2048)
    loss, g = ctx.guide_loss_grad(rows)
    lo, go = og.loss_grad(th, rows, bf16_dz=True)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    assert rel <= 5e-3, rel
    # every layer, not just the largest
    off = 0
    for name, shape in og._SIZES:
        n = int(np.prod(shape))
        a, b = g[off:off + n], go[off:off + n]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-7, name
        off += n


@pytest.mark.gpu
def test_loss_grad_multi_chunk_and_ragged(ctx):
    """a batch larger than one pass of the training grid (blocks that sum
    several 64-row chunks) and a ragged one (a partial last chunk) against
    the oracle"""
    th = _theta(16)
    ctx.guide_set_params(th)
    for n in (65536, 333):
        rows = _rows(np.random.default_rng(n), n)
        loss, g = ctx.guide_loss_grad(rows)
        lo, go = og.loss_grad(th, rows, bf16_dz=True)
        assert abs(loss - lo) <= 1e-4 * abs(lo), n
        assert np.linalg.norm(g - go) <= 5e-3 * np.linalg.norm(go), n


@pytest.mark.gpu
def test_training_forward_is_the_inference(ctx):
    """the training forward is the inference network's evaluation: its raw
    outputs equal lam_guide_infer's bit for bit"""
    th = _theta(17)
    ctx.guide_set_params(th)
    rows = _rows(np.random.default_rng(18), 1000)
    _, _, raw = ctx.guide_loss_grad(rows, want_raw=True)
    assert np.array_equal(raw, ctx.guide_infer(np.ascontiguousarray(rows[:, :6])))


@pytest.mark.gpu
def test_training_updates_the_pack_in_place_and_deterministically(ctx):
    """Adam steps rewrite each parameter's packed bf16 entries in place: the
    network afterwards evaluates exactly as one packed afresh from the same
    parameters; and two identical training runs give identical parameters"""
    rows = _rows(np.random.default_rng(19), 4096)
    x = np.ascontiguousarray(_rows(np.random.default_rng(20), 512)[:, :6])
    out = []
    for _ in range(2):
        ctx.guide_reset(21)
        ctx.guide_train(rows, 5, 2e-3)
        out.append((ctx.guide_get_params(), ctx.guide_infer(x)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    ctx.guide_set_params(out[0][0])  # a fresh pack of the same parameters
    assert np.array_equal(ctx.guide_infer(x), out[0][1])


@pytest.mark.gpu
def test_adam_first_step(ctx):
    """Adam's first step moves every parameter by lr · g/(|g| + ε)"""
    th = _theta(8)
    ctx.guide_set_params(th)
    rows = _rows(np.random.default_rng(9), 1024)
    _, g = ctx.guide_loss_grad(rows)
    lr = 1e-3
    ctx.guide_train(rows, 1, lr)
    th1 = ctx.guide_get_params().astype(np.float64)
    step = th1 - th
    # where |g| >> ε the step is -lr sign(g); the gradient is recomputed
    # inside the step (float atomics in another order), so near-zero
    # components are only bounded
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert big.sum() > 1000
    assert np.allclose(step[big], -lr * np.sign(g[big]), rtol=0, atol=1e-7 + 1e-4 * lr)
    assert np.all(np.abs(step) <= lr * (1 + 1e-5) + 1e-7)


@pytest.mark.gpu
def test_training_fits_a_direction_field(ctx):
    """rows whose ω concentrates around a direction that depends on the
    position: training lowers the loss and the mixture learns the directions"""
    rng = np.random.default_rng(10)
    n = 8192
    inp = rng.uniform(0, 1, (n, 6)).astype(np.float32)
    ang = 2 * np.pi * inp[:, 0]
    mu = np.stack([np.cos(ang), np.sin(ang), 0.3 * np.ones(n)], axis=1)
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    om = mu + 0.15 * rng.normal(size=(n, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    rows = np.concatenate([inp, om, np.ones((n, 1))], axis=1).astype(np.float32)
    ctx.guide_reset(11)
    l0, _ = ctx.guide_loss_grad(rows)
    for _ in range(10):
        ctx.guide_train(rows, 40, 3e-3)
    l1, _ = ctx.guide_loss_grad(rows)
    assert l1 < l0 - 2.0, (l0, l1)
    # the fitted density is high at the true mean direction
    raw = ctx.guide_infer(inp[:512])
    p_true = og.mix_pdf(raw, mu[:512])
    p_anti = og.mix_pdf(raw, -mu[:512])
    assert np.median(p_true) > 20 * np.median(p_anti)


@pytest.fixture(scope="module")
def glossy(tmp_path_factory):
    from lambda_amd import scenes
    p = scenes.glossy_interior(str(tmp_path_factory.mktemp("glossy")), seed=4, width=48, n_boxes=24,
                               box_subdiv=3)
    return la.load_nls(p)


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, la.LAM_FLAG_NO_TAIL])
def test_guided_alpha0_is_plain(ctx, glossy, flags):
    ctx.upload(glossy)
    p = la.params(spp=4, max_scatterings=8, sample_p=0.8, hero=4, seed=3, flags=flags)
    plain = ctx.render(p)
    ctx.guide_reset(1)
    g = ctx.render_guided(p, la.guide_params(alpha=0.0, train_passes=0))
    assert np.array_equal(plain, g)
    # with α > 0 the guided vertices take other paths (the proposal is used)
    g5 = ctx.render_guided(p, la.guide_params(alpha=0.5, train_passes=0))
    assert not np.array_equal(plain, g5)
    # two pipeline lanes (each with its own network-output buffer) give the
    # one-lane image
    p1 = la.params(spp=4, max_scatterings=8, sample_p=0.8, hero=4, seed=3, flags=flags | la.LAM_FLAG_NO_PIPE)
    assert np.array_equal(g5, ctx.render_guided(p1, la.guide_params(alpha=0.5, train_passes=0)))
    # a vertex limit past every bounce changes nothing; guiding only the
    # first vertex gives a third image (deeper vertices use the plain kernels)
    assert np.array_equal(g5, ctx.render_guided(p, la.guide_params(alpha=0.5, train_passes=0, max_vertex=64)))
    g1 = ctx.render_guided(p, la.guide_params(alpha=0.5, train_passes=0, max_vertex=1))
    assert not np.array_equal(g1, g5) and not np.array_equal(g1, plain)


@pytest.mark.gpu
def test_guided_render_unbiased(ctx, glossy):
    """guided and plain renders are independent estimates of one image: the
    per-pixel difference of their means is noise (z-test), with a trained
    network and with an untrained one, with only the first two vertices
    guided, and with the Blinn specular lobe under the mixture as well"""
    ctx.upload(glossy)
    spp = 256
    p = la.params(spp=spp, max_scatterings=8, sample_p=0.8, hero=4, seed=5)
    plain = []
    for half in range(2):  # two halves for a per-pixel variance estimate
        plain.append(ctx.render(la.params(spp=spp // 2, max_scatterings=8, sample_p=0.8, hero=4, seed=5,
                                          spp_offset=half * (spp // 2))).astype(np.float64))
    every = la.GUIDE_LAMBERTIAN | la.GUIDE_BLINN_KD | la.GUIDE_BLINN_SPEC
    for trained, mv, lobes in ((False, 0, 0), (True, 0, 0), (True, 2, 0), (True, 0, every)):
        ctx.guide_reset(2)
        gp = la.guide_params(alpha=0.5, train_passes=8 if trained else 0, steps_per_pass=8, batch=4096,
                             lr=5e-3, seed=7, max_vertex=mv, lobes=lobes)
        guided = []
        for half in range(2):
            # pure wavefront: every vertex is shaded by the guided kernel
            pg = la.params(spp=spp // 2, max_scatterings=8, sample_p=0.8, hero=4, seed=99 + half,
                           spp_offset=0, flags=la.LAM_FLAG_NO_TAIL)
            guided.append(ctx.render_guided(pg, gp if half == 0 else
                                            la.guide_params(alpha=0.5, train_passes=0, seed=7,
                                                            max_vertex=mv, lobes=lobes)).astype(np.float64))
            if trained and half == 0:
                assert ctx.timing().guide_records > 1000
        mp, mg = (plain[0] + plain[1]) / 2, (guided[0] + guided[1]) / 2
        # variance of a half-mean from the two halves; of the full mean: /2
        vp = (plain[0] - plain[1]) ** 2 / 4
        vg = (guided[0] - guided[1]) ** 2 / 4
        z = (mg - mp) / np.sqrt(vp + vg + 1e-30)
        # two-sample estimates are heavy-tailed: compare image-level sums
        tot_p, tot_g = mp.sum(), mg.sum()
        sd = np.sqrt(vp.sum() + vg.sum())
        assert abs(tot_g - tot_p) < 4 * sd, (trained, tot_p, tot_g, sd)
        assert np.median(np.abs(z)) < 1.2, (trained, np.median(np.abs(z)))


@pytest.mark.gpu
def test_trained_network_lowers_variance(ctx, tmp_path):
    """What the learned proposal gains: on the hidden-light room (Lambertian
    boxes), per-pixel variance over 8 independent renders with the network
    trained on 64 passes is well below the plain sampler's (measured ratio
    0.44) and below guiding with the untrained network (2.0x plain), at
    equal samples per pixel."""
    from lambda_amd import scenes
    sc = la.load_nls(scenes.glossy_interior(str(tmp_path), seed=4, width=64, n_boxes=24, box_subdiv=3,
                                            box_ks=0.0))
    ctx.upload(sc)
    w = np.array([0.2126, 0.7152, 0.0722])

    def rel_var(runs):
        y = np.stack([r.astype(np.float64) @ w for r in runs])
        mu = y.mean(0)
        return float(np.mean(y.var(0, ddof=1) / (mu ** 2 + 0.01 * mu.mean() ** 2)))

    mk = lambda s: la.params(spp=64, max_scatterings=16, sample_p=0.8, hero=4, seed=s)
    K = 8
    v_plain = rel_var([ctx.render(mk(10 + k)) for k in range(K)])
    frozen = la.guide_params(alpha=0.5, train_passes=0, seed=3)
    ctx.guide_reset(3)
    v_untrained = rel_var([ctx.render_guided(mk(50 + k), frozen) for k in range(K)])
    ctx.guide_reset(3)
    ctx.render_guided(la.params(spp=64, max_scatterings=16, sample_p=0.8, hero=4, seed=1),
                      la.guide_params(alpha=0.5, train_passes=64, steps_per_pass=8, batch=4096, lr=5e-3, seed=3))
    assert ctx.timing().guide_records > 100000
    v_trained = rel_var([ctx.render_guided(mk(50 + k), frozen) for k in range(K)])
    print(f"relative variance: plain {v_plain:.5f}, untrained {v_untrained:.5f}, trained {v_trained:.5f}")
    assert v_trained < 0.7 * v_plain, (v_plain, v_trained)
    assert v_trained < 0.5 * v_untrained, (v_untrained, v_trained)


def test_oracle_bf16_dz_gradient_is_close_to_straight_through():
    """the oracle's bf16-dZ backward (the library's MFMA training GEMMs)
    differs from the straight-through float64 one only by the rounding of
    dZ: within 1% per layer, and identical bias gradients"""
    th = _theta(4)
    rows = _rows(np.random.default_rng(12), 512)
    l0, g0 = og.loss_grad(th, rows)
    l1, g1 = og.loss_grad(th, rows, bf16_dz=True)
    assert l0 == l1
    off = 0
    for name, shape in og._SIZES:
        n = int(np.prod(shape))
        a, b = g1[off:off + n], g0[off:off + n]
        if name.startswith("b") and name != "b1":
            pass
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-12, name
        if name == "b5":
            assert np.array_equal(a, b)
        off += n


@pytest.mark.gpu
def test_guided_and_plain_match_the_fp64_reference(ctx, tmp_path):
    """The guiding workload's scene (glossy interior, seed 4) at 64x64:
    the plain GPU render and the guided one (network trained online on its
    first 32 passes) both converge to the FP64 oracle's 4096-spp image
    (tests/golden/glossy_interior_64_fp64.npz, written by
    tests/golden/make_glossy_fp64.py): image sums within 4 standard errors,
    per-pixel |z| of median below 1.2, relMSE against the reference printed.
    (A check of bias only: at 64x64 the network learns from 256x fewer
    records than at the workload's 1024x1024, and the guided estimate is the
    noisier one here: measured relMSE 0.26 against plain 0.16.)"""
    import os
    from lambda_amd import scenes
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "glossy_interior_64_fp64.npz"))
    sc = la.load_nls(scenes.glossy_interior(str(tmp_path), seed=int(ref["scene_seed"]), width=int(ref["width"])))
    ctx.upload(sc)
    spp, b, p = 2048, int(ref["bounces"]), float(ref["sample_p"])
    mk = lambda seed, off=0: la.params(spp=spp // 2, max_scatterings=b, sample_p=p, hero=sc.hero, seed=seed,
                                       spp_offset=off)
    plain = [ctx.render(mk(3, k * spp // 2)).astype(np.float64) for k in range(2)]
    ctx.guide_reset(5)
    guided = [ctx.render_guided(mk(5 + k), la.guide_params(train_passes=32 if k == 0 else 0, steps_per_pass=8,
                                                           batch=4096, seed=5)).astype(np.float64)
              for k in range(2)]
    r, vr = ref["mean"].astype(np.float64), ref["var_of_mean"].astype(np.float64)
    w = np.array([0.2126, 0.7152, 0.0722])
    out = {}
    for name, runs in (("plain", plain), ("guided", guided)):
        m = (runs[0] + runs[1]) / 2
        v = (runs[0] - runs[1]) ** 2 / 4
        sd = np.sqrt(v.sum() + vr.sum())
        assert abs(m.sum() - r.sum()) < 4 * sd, (name, m.sum(), r.sum(), sd)
        z = (m - r) / np.sqrt(v + vr + 1e-30)
        assert np.median(np.abs(z)) < 1.2, (name, np.median(np.abs(z)))
        y, yr = m @ w, r @ w
        out[name] = float(np.mean((y - yr) ** 2 / (yr ** 2 + 0.01 * yr.mean() ** 2)))
    print(f"relMSE against the FP64 reference at {spp} spp: {out}")
