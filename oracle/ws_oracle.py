"""CPU oracle for the learned initial guess (test infrastructure only: used by
tests/ and nothing else).

The learned guess has no counterpart in the reference; it replaces the zero
initial guess xvct = 0 before DGMRES (src/main.f90:115).  This module restates,
in numpy, what the GPU kernels compute:

* ``features``: the 12 per-face features (right-hand side scaled by
  s = 1/rms(b_1..b_n), the potential and normal field of the charges within
  Gaussian windows of 5 and 10 A, area, leaf-local curvature, the two nearest
  charge distances, the normal's cosine with the nearest-charge direction and a
  smooth burial count), from tree-order geometry and leaf ranges.
* ``forward`` / ``loss_grad``: the MLP 12 -> 3 x 128 GELU -> 2 and the gradient of
  L = (1/2B) sum |out - t/ts|^2, either in FP64 throughout (``emulate=False``,
  the mathematical reference pinned by finite differences in the tests) or with
  the kernels' BF16 rounding points emulated (``emulate=True``): inputs,
  activations and deltas rounded to BF16 (round to nearest even) where the
  kernels feed them to the matrix cores, FP32-level accumulation elsewhere.

Packed parameter layout (floats): mu[12], sig[12], ts[2], 6 unused, W1[128][12],
b1[128], W2[128][128], b2[128], W3[128][128], b3[128], W4[2][128], b4[2].
"""
import numpy as np

NF, NH, NO = 12, 128, 2
O_MU, O_SIG, O_TS, O_W1 = 0, 12, 24, 32
O_B1 = O_W1 + NH * NF
O_W2 = O_B1 + NH
O_B2 = O_W2 + NH * NH
O_W3 = O_B2 + NH
O_B3 = O_W3 + NH * NH
O_W4 = O_B3 + NH
O_B4 = O_W4 + NO * NH
NP = O_B4 + NO

C_GELU = 0.7978845608028654


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(P):
    P = np.asarray(P)
    return dict(
        mu=P[O_MU:O_MU + NF], sig=P[O_SIG:O_SIG + NF], ts=P[O_TS:O_TS + NO],
        W1=P[O_W1:O_B1].reshape(NH, NF), b1=P[O_B1:O_W2],
        W2=P[O_W2:O_B2].reshape(NH, NH), b2=P[O_B2:O_W3],
        W3=P[O_W3:O_B3].reshape(NH, NH), b3=P[O_B3:O_W4],
        W4=P[O_W4:O_B4].reshape(NO, NH), b4=P[O_B4:NP])


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_d(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 0.134145 * z * z)


def _r(x, emulate):
    return bf16(x).astype(np.float64) if emulate else x


def forward(P, feat, emulate=True):
    """Outputs (ns, 2) in normalised units, and the cache for the backward pass."""
    p = {k: np.asarray(v, dtype=np.float64) for k, v in unpack(P).items()}
    f = np.asarray(feat, dtype=np.float32)
    if emulate:
        a0 = ((f - p["mu"].astype(np.float32)) / p["sig"].astype(np.float32)).astype(np.float32)
        a0 = bf16(a0).astype(np.float64)
    else:
        a0 = (f.astype(np.float64) - p["mu"]) / p["sig"]
    W = [_r(p["W1"], emulate), _r(p["W2"], emulate), _r(p["W3"], emulate), _r(p["W4"], emulate)]
    b = [p["b1"], p["b2"], p["b3"], p["b4"]]
    acts, zs = [a0], []
    a = a0
    for l in range(3):
        z = a @ W[l].T + b[l]
        if emulate:
            z = z.astype(np.float32).astype(np.float64)
        zs.append(z)
        a = _r(gelu(z), emulate)
        acts.append(a)
    out = a @ W[3].T + b[3]
    if emulate:
        out = out.astype(np.float32).astype(np.float64)
    return out, dict(W=W, acts=acts, zs=zs, p=p)


def predict(P, feat, scale, emulate=True):
    """Tree-order guess [phi; dphi/dn] = out * ts / s."""
    out, c = forward(P, feat, emulate)
    ts = c["p"]["ts"]
    return np.concatenate([out[:, 0] * ts[0] / scale, out[:, 1] * ts[1] / scale])


def loss_grad(P, feat, target, emulate=True):
    """(packed gradient, loss) of L = (1/2B) sum_f sum_o (out - t/ts)^2."""
    out, c = forward(P, feat, emulate)
    W, acts, zs, p = c["W"], c["acts"], c["zs"], c["p"]
    t = np.asarray(target, dtype=np.float64).reshape(-1, 2)
    B = out.shape[0]
    e = out - t / p["ts"]
    if emulate:
        e = (out.astype(np.float32) - (t.astype(np.float32) / p["ts"].astype(np.float32))).astype(np.float64)
    loss = float((e ** 2).sum() / (2.0 * B))
    d = e / B
    grads = {}
    names = [("W4", "b4"), ("W3", "b3"), ("W2", "b2"), ("W1", "b1")]
    for li, (wn, bn) in enumerate(names):
        l = 3 - li                      # layer index into W/acts
        db = _r(d, emulate)             # delta as the matrix cores see it
        grads[wn] = db.T @ acts[l]
        grads[bn] = d.sum(axis=0)
        if l > 0:
            d = (db @ W[l]) * gelu_d(zs[l - 1])
            if emulate:
                d = d.astype(np.float32).astype(np.float64)
    g = np.zeros(NP)
    g[O_W1:O_B1] = grads["W1"].ravel()
    g[O_B1:O_W2] = grads["b1"]
    g[O_W2:O_B2] = grads["W2"].ravel()
    g[O_B2:O_W3] = grads["b2"]
    g[O_W3:O_B3] = grads["W3"].ravel()
    g[O_B3:O_W4] = grads["b3"]
    g[O_W4:O_B4] = grads["W4"].ravel()
    g[O_B4:NP] = grads["b4"]
    return g, loss


def loss_only(P, feat, target):
    out, c = forward(P, feat, emulate=False)
    t = np.asarray(target, dtype=np.float64).reshape(-1, 2)
    e = out - t / c["p"]["ts"]
    return float((e ** 2).sum() / (2.0 * out.shape[0]))


def features(txyz, tnrm, tarea, leaf_ranges, qpos, q, eps_p, b, h=1.0, R1=5.0, R2=10.0):
    """The 12 features per face (tree order).  leaf_ranges: (ibeg, iend)
    inclusive per leaf; b the tree-order right-hand side [2n].  Charge sums use
    float32 coordinates and charges, as the kernel does, accumulated in FP64."""
    txyz = np.asarray(txyz, dtype=np.float64)
    tnrm = np.asarray(tnrm, dtype=np.float64)
    n = txyz.shape[0]
    b = np.asarray(b, dtype=np.float64)
    s = 1.0 / np.sqrt(np.mean(b[:n] ** 2))
    su = s * 0.25 / (np.pi * eps_p)
    f = np.zeros((n, 12))
    f[:, 0] = s * b[:n]
    f[:, 1] = s * b[n:]
    f[:, 6] = tarea
    # curvature: Gaussian-weighted mean of (n_j-n_i).(c_j-c_i)/|c_j-c_i|^2 in the leaf
    for ib, ie in leaf_ranges:
        c = txyz[ib:ie + 1]
        nn = tnrm[ib:ie + 1]
        dc = c[None, :, :] - c[:, None, :]
        r2 = (dc ** 2).sum(-1)
        dn = nn[None, :, :] - nn[:, None, :]
        ok = (r2 > 0) & (r2 / (h * h) <= 9.0)
        w = np.where(ok, np.exp(-r2 / (h * h)), 0.0)
        val = np.where(ok, (dn * dc).sum(-1) / np.where(ok, r2, 1.0), 0.0)
        den = w.sum(1)
        f[ib:ie + 1, 7] = np.where(den > 0, (w * val).sum(1) / np.where(den > 0, den, 1.0), 0.0)
    q32 = np.asarray(qpos, dtype=np.float32).astype(np.float64)
    qq = np.asarray(q, dtype=np.float32).astype(np.float64)
    c32 = txyz.astype(np.float32).astype(np.float64)
    n32 = tnrm.astype(np.float32).astype(np.float64)
    for i0 in range(0, n, 2048):
        c = c32[i0:i0 + 2048]
        nv = n32[i0:i0 + 2048]
        d = q32[None, :, :] - c[:, None, :]
        r2 = (d ** 2).sum(-1)
        srt = np.argsort(r2, axis=1, kind="stable")
        r1 = np.sqrt(np.take_along_axis(r2, srt[:, :1], 1))[:, 0]
        rr2 = np.sqrt(np.take_along_axis(r2, srt[:, 1:2], 1))[:, 0] if r2.shape[1] > 1 else r1
        a1 = q32[srt[:, 0]]
        ok = (r2 > 0) & (r2 < 9.0 * R2 * R2)
        rs = np.sqrt(np.where(ok, r2, 1.0))
        pot = np.where(ok, qq[None, :] / rs, 0.0)
        fld = pot * (nv[:, None, :] * d).sum(-1) / (rs * rs)
        w1 = np.where(ok, np.exp(-r2 / (R1 * R1)), 0.0)
        w2 = np.where(ok, np.exp(-r2 / (R2 * R2)), 0.0)
        sl = slice(i0, i0 + 2048)
        f[sl, 2] = su * (pot * w1).sum(1)
        f[sl, 3] = su * (fld * w1).sum(1)
        f[sl, 4] = su * (pot * w2).sum(1)
        f[sl, 5] = su * (fld * w2).sum(1)
        f[sl, 8] = r1
        f[sl, 9] = rr2
        f[sl, 10] = np.where(r1 > 0, (nv * (c - a1)).sum(1) / np.where(r1 > 0, r1, 1.0), 0.0)
        f[sl, 11] = w2.sum(1)
    return f, s
