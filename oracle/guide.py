"""TEST INFRASTRUCTURE: numpy restatement of the path-guiding network (the
checker of lam_guide_infer / lam_guide_loss_grad / lam_guide_train and of
the guided render's mixture).  Imported only by tests/.

The network has no counterpart in the reference (webstek/lambda traces with
its BSDF samplers only, src/renderer.cpp:116-145); this file restates the
library's own definition:

  inputs   6 values in [0,1]
  features f < 144: sin (f even) / cos (f odd) of π 2^j x_d, f = 24 d + 2 j + c,
           evaluated as the device does: sincos_rev at j0 = 4 floor(j/4),
           then j - j0 double-angle steps; f = 144 + d: x_d;
           rounded to bf16
  layers   150 -> 64 -> 64 -> 64 -> 64 (ReLU, activations rounded to bf16)
           -> 40; weights rounded to bf16, biases fp32, sums exact (float64)
  outputs  per lobe L: raw[5L:5L+3] mean direction (normalised), raw[5L+3]
           log κ (clamped to [-4, 9]), raw[5L+4] weight logit (softmax)
  loss     -Σ_i (w_i / Σw) log p(ω_i), p = Σ_L π_L vMF(ω; μ_L, κ_L),
           vMF = κ / (2π (1 - e^{-2κ})) e^{κ (μ·ω - 1)}
  backward float64, the bf16 roundings treated as identity (straight-through);
           optionally dZ rounded to bf16 in the products, as the library's
           bf16 MFMA training GEMMs take it
"""
from __future__ import annotations

import numpy as np

N_IN, N_FEAT, N_FEAT_PAD, N_HID, N_OUT, N_LOBES = 6, 150, 160, 64, 40, 8
LOGK_MIN, LOGK_MAX = -4.0, 9.0

# parameter vector layout (lam_guide_param_count floats)
_SIZES = [("W1", (N_HID, N_FEAT)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("W4", (N_HID, N_HID)), ("b4", (N_HID,)),
          ("W5", (N_OUT, N_HID)), ("b5", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in _SIZES)


def unpack(theta):
    out, off = {}, 0
    for name, shape in _SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[off:off + n], dtype=np.float64).reshape(shape)
        off += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).ravel() for name, _ in _SIZES])


def bf16(x):
    """round to bf16 (nearest even), returned as float64"""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sincos_rev(fr):
    """sin and cos of 2π fr, fr in [0,1), in float32 with the library's
    reduction (nearest quarter turn) and Taylor polynomials, each operation
    rounded as the device rounds it (no fused multiply-adds)"""
    f = np.float32
    fr = np.asarray(fr, dtype=f)
    qf = np.floor(fr * f(4) + f(0.5))
    t = fr - qf * f(0.25)
    x = t * f(6.28318530717958647692)
    x2 = x * x
    ps = f(1) / f(362880)
    ps = f(-1) / f(5040) + x2 * ps
    ps = f(1) / f(120) + x2 * ps
    ps = f(-1) / f(6) + x2 * ps
    sp = x + x * x2 * ps
    pc = f(-1) / f(3628800)
    pc = f(1) / f(40320) + x2 * pc
    pc = f(-1) / f(720) + x2 * pc
    pc = f(1) / f(24) + x2 * pc
    pc = f(-0.5) + x2 * pc
    cp = f(1) + x2 * pc
    q = qf.astype(np.int64) & 3
    s = np.select([q == 0, q == 1, q == 2], [sp, cp, -sp], -cp)
    c = np.select([q == 0, q == 1, q == 2], [cp, -sp, -cp], sp)
    return s.astype(f), c.astype(f)


def features(inp):
    """encoded, bf16-rounded features (n x 150)"""
    x = np.asarray(inp, dtype=np.float32).reshape(-1, N_IN)
    cols = []
    f32 = np.float32
    for f in range(144):
        d, j = f // 24, (f % 24) // 2
        j0 = j & ~3
        r = x[:, d] * f32(2.0 ** j0) * f32(0.5)
        s, c = sincos_rev(r - np.floor(r))
        for _ in range(j - j0):  # double-angle steps, float32 as on the device
            s, c = (s * c) * f32(2), (c - s) * (c + s)
        cols.append((c if f % 2 else s).astype(np.float64))
    for d in range(N_IN):
        cols.append(x[:, d].astype(np.float64))
    return bf16(np.stack(cols, axis=1))


def forward(theta, inp, keep=False):
    """raw outputs (n x 40); keep=True also returns the layer inputs and
    pre-activations for the backward pass"""
    P = unpack(theta)
    X = features(inp)
    xs, zs = [X], []
    h = X
    for l in range(1, 6):
        W = bf16(P[f"W{l}"])
        z = h @ W.T + P[f"b{l}"].astype(np.float32).astype(np.float64)
        zs.append(z)
        if l < 5:
            h = bf16(np.maximum(z, 0.0))
            xs.append(h)
    return (z, xs, zs) if keep else z


def mixture(raw):
    """(μ n x 8 x 3, κ n x 8, π n x 8) from raw outputs"""
    raw = np.asarray(raw, dtype=np.float64).reshape(-1, N_LOBES, 5)
    v = raw[:, :, :3]
    ln = np.linalg.norm(v, axis=2, keepdims=True)
    mu = np.where(ln > 1e-12, v / np.maximum(ln, 1e-300), np.array([0.0, 0.0, 1.0]))
    kappa = np.exp(np.clip(raw[:, :, 3], LOGK_MIN, LOGK_MAX))
    a = raw[:, :, 4] - raw[:, :, 4].max(axis=1, keepdims=True)
    w = np.exp(a)
    return mu, kappa, w / w.sum(axis=1, keepdims=True)


def log_vmf(mu, kappa, om):
    md = np.einsum("nlk,nk->nl", mu, om)
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (md - 1.0)


def mix_pdf(raw, om):
    mu, kappa, w = mixture(raw)
    om = np.asarray(om, dtype=np.float64).reshape(-1, 3)
    return (w * np.exp(log_vmf(mu, kappa, om))).sum(axis=1)


def loss_grad(theta, rows, bf16_dz=False):
    """(loss, gradient) of -Σ (w_i/Σw) log p(ω_i | in_i) for rows (n x 10).
    bf16_dz: dZ of each layer rounded to bf16 where it enters the weight
    gradient and the back-propagated product (as the library's bf16 MFMA
    training GEMMs take it); the bias gradients sum the unrounded dZ"""
    rows = np.asarray(rows, dtype=np.float64).reshape(-1, 10)
    inp, om, wt = rows[:, :6], rows[:, 6:9], rows[:, 9]
    raw, xs, zs = forward(theta, inp, keep=True)
    n = len(rows)
    sw = wt.sum()
    wn = wt / sw if sw > 0 else np.zeros_like(wt)
    r = raw.reshape(n, N_LOBES, 5)
    v = r[:, :, :3]
    vlen = np.linalg.norm(v, axis=2)
    mu = np.where(vlen[:, :, None] > 1e-12, v / np.maximum(vlen, 1e-300)[:, :, None], np.array([0.0, 0.0, 1.0]))
    lk = r[:, :, 3]
    kappa = np.exp(np.clip(lk, LOGK_MIN, LOGK_MAX))
    a = r[:, :, 4]
    lw = a - a.max(axis=1, keepdims=True)
    lw = lw - np.log(np.exp(lw).sum(axis=1, keepdims=True))  # log π
    md = np.einsum("nlk,nk->nl", mu, om)
    lv = np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (md - 1.0)
    z = lw + lv
    zm = z.max(axis=1, keepdims=True)
    logp = (zm + np.log(np.exp(z - zm).sum(axis=1, keepdims=True)))[:, 0]
    loss = float(-(wn * logp).sum())
    resp = np.exp(z - logp[:, None])
    pi = np.exp(lw)
    dlk = 1.0 - 2.0 * kappa / np.expm1(2.0 * kappa) + kappa * (md - 1.0)
    inside = (lk > LOGK_MIN) & (lk < LOGK_MAX)
    g_k = np.where(inside, resp * dlk, 0.0)
    gv = (om[:, None, :] - md[:, :, None] * mu) * (resp * kappa / np.maximum(vlen, 1e-300))[:, :, None]
    gv = np.where(vlen[:, :, None] > 1e-12, gv, 0.0)
    d = np.zeros((n, N_LOBES, 5))
    d[:, :, :3] = gv
    d[:, :, 3] = g_k
    d[:, :, 4] = resp - pi
    dz = -(wn[:, None, None] * d).reshape(n, N_OUT)
    # backward through the layers (straight-through bf16)
    P = unpack(theta)
    grads = {}
    for l in range(5, 0, -1):
        xin = xs[l - 1]
        dzq = bf16(dz) if bf16_dz else dz
        grads[f"W{l}"] = dzq.T @ xin
        grads[f"b{l}"] = dz.sum(axis=0)
        if l > 1:
            dx = dzq @ bf16(P[f"W{l}"])
            dz = np.where(zs[l - 2] > 0, dx, 0.0)
    # W1 covers the 150 real features
    return loss, pack(grads)
