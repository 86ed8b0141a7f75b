"""CPU statement of the reciprocal-force surrogate - TEST INFRASTRUCTURE ONLY.

This is the checker for the surrogate's HIP kernels (md_sim_amd/csrc/surrogate.hip): tests
and the smoke check call it; the product path never does.  The surrogate has no counterpart
in the reference (src/forcem.f90 always evaluates the full k-sum), so nothing here follows a
reference line; each function states the definition the kernels implement, and the tests
pin those definitions by invariances, closed forms, finite differences and torch autograd.

Model.  Per moving site i, from all sites j within RC = 6 A (minimum image in x, y, and z
when z is periodic), with taper fc(r) = (cos(pi r / RC) + 1) / 2 and r_ij = r_i - r_j:
  invariant inputs x_i (48 wide, 39 used):
    [0:8]   sum_{q_j > 0} |q_j| fc exp(-ETA_R (r - RAD_C[m])^2)
    [8:16]  the same over q_j < 0
    [16:37] dot products A_a . A_b (a <= b, a-major) of the six moment vectors
            A_{c,s} = sum_{j in class c} |q_j| fc exp(-ETA_A (r - ANG_C[s])^2) r_ij / r
            (c: q_j > 0, q_j < 0; s over ANG_C) - the angular functions
            sum_{j,k} g(r_ij) g(r_ik) cos(theta_jik) in O(neighbours) form
    [37]    q_i, [38] q_i^2, [39:48] zero
  vector features U_i,m (16 x 3), force units, over whole molecules: every site j of a
  molecule whose lead is within RG of i's lead (frozen sites: j within RG of i's lead),
  weighted by the smooth step wg(R) = 1 - s^2 (3 - 2 s), s = clip((R - RG + WG) / WG, 0, 1),
  of that lead distance R (the reference's group cut-off, src/forcem.f90:299,360-402; it
  keeps neighbour molecules neutral), plus the other sites of i's own molecule (wg = 1):
    U_0 = facq q_i sum_j q_j wg [erf(a r)/r^2 - 2a/sqrt(pi) exp(-a^2 r^2)/r] r_ij / r
          (the field of j's Gaussian-screened charge: the near part of the k-space force)
    U_m = facq q_i sum_j q_j wg T_{m-1}(2 r / RG - 1) r_ij / r,  m = 1..7 (Chebyshev
          polynomials of the first kind: a smooth radial basis the network weighs);
    U_8 .. U_15 = 0 (the network's 16 outputs fill one matrix-core tile)
Network: xn = (x - mu) / sigma, h1 = silu(W1 xn + b1), h2 = silu(W2 h1 + b2),
c = W3 h2 + b3 (16 outputs), predicted force F_i = sum_m c_m U_i,m / su_m.  The force is
a combination of vectors, so it turns with the system; the inputs do not.
Parameters are one flat FP32 vector: W1 [64][48], b1, W2 [64][64], b2, W3 [16][64], b3.
The kernels round the MFMA operands (inputs, weights, activations, back-propagated
deltas) to bfloat16 (round to nearest even) and accumulate in FP32; ``emulate=True``
applies the same roundings here.
Loss of one frame: L = sum_i |F_i - T_i|^2 / (n t2), t2 = mean_i |T_i|^2.  The scheme
trains on the change between two frames, L = sum_i |(F_i - F'_i) - (T_i - T'_i)|^2 / (n t2)
with t2 = mean |T - T'|^2 (``change_loss_and_grad``).
"""
from __future__ import annotations

import numpy as np

RC = 6.0
RAD_C = 0.75 * (np.arange(8) + 1.0)
ETA_R = 2.0
ANG_C = np.array([1.5, 3.0, 4.5])
ETA_A = 1.0
WG = 2.0
D_IN = 48
N_USED = 39
H = 64
NV = 16
NV_USED = 8
# flat parameter layout
OFF_W1, OFF_B1 = 0, 64 * 48
OFF_W2 = OFF_B1 + 64
OFF_B2 = OFF_W2 + 64 * 64
OFF_W3 = OFF_B2 + 64
OFF_B3 = OFF_W3 + 16 * 64
N_PARAMS = OFF_B3 + 16


def _erf(x):
    from scipy.special import erf
    return erf(x)


def features(pos, q, n_move, box, z_periodic, alpha, facq, lead, rg, chunk=256):
    """x [n_move, 48] and U [n_move, 16, 3] (FP64) by brute force over all site pairs.
    lead[j]: the lead site of j's molecule (j itself for single sites and frozen sites);
    rg: the group cut-off of the vector features."""
    pos = np.asarray(pos, np.float64)
    q = np.asarray(q, np.float64)
    L = np.asarray(box, np.float64)
    lead = np.asarray(lead)
    feat = np.zeros((n_move, D_IN))
    U = np.zeros((n_move, NV, 3))
    idx_all = np.arange(len(pos))
    iu = np.triu_indices(6)

    def mi(d):
        d[..., 0] -= L[0] * np.round(d[..., 0] / L[0])
        d[..., 1] -= L[1] * np.round(d[..., 1] / L[1])
        if z_periodic:
            d[..., 2] -= L[2] * np.round(d[..., 2] / L[2])
        return d

    for i0 in range(0, n_move, chunk):
        i1 = min(n_move, i0 + chunk)
        ii = np.arange(i0, i1)
        d = mi(pos[i0:i1, None, :] - pos[None, :, :])
        r2 = (d * d).sum(-1)
        notself = idx_all[None, :] != ii[:, None]
        r = np.sqrt(np.where(notself, r2, 1.0))
        rhat = d / r[..., None]
        qj = q[None, :]
        # invariant inputs: every site within RC, cosine taper
        mask = (r2 < RC * RC) & notself
        fc = np.where(mask, 0.5 * (np.cos(np.pi * r / RC) + 1.0), 0.0)
        pos_c = (qj > 0) * np.abs(qj) * fc
        neg_c = (qj < 0) * np.abs(qj) * fc
        for m, c in enumerate(RAD_C):
            g = np.exp(-ETA_R * (r - c) ** 2)
            feat[i0:i1, m] = (pos_c * g).sum(1)
            feat[i0:i1, 8 + m] = (neg_c * g).sum(1)
        A = np.zeros((i1 - i0, 6, 3))
        for s, c in enumerate(ANG_C):
            g = np.exp(-ETA_A * (r - c) ** 2)
            A[:, s] = ((pos_c * g)[..., None] * rhat).sum(1)
            A[:, 3 + s] = ((neg_c * g)[..., None] * rhat).sum(1)
        dots = np.einsum("nak,nbk->nab", A, A)
        feat[i0:i1, 16:37] = dots[:, iu[0], iu[1]]
        qi = q[i0:i1]
        feat[i0:i1, 37] = qi
        feat[i0:i1, 38] = qi * qi
        # vector features: whole molecules by lead distance, smooth step
        dl = mi(pos[lead[ii]][:, None, :] - pos[lead][None, :, :])
        R = np.sqrt((dl * dl).sum(-1))
        sstep = np.clip((R - rg + WG) / WG, 0.0, 1.0)
        wg = np.where(R < rg, 1.0 - sstep * sstep * (3.0 - 2.0 * sstep), 0.0)
        same = lead[None, :] == lead[ii][:, None]
        wg = np.where(same, 1.0, wg) * notself
        w = qj * wg * facq * qi[:, None]
        kew = _erf(alpha * r) / (r * r) - 2.0 * alpha / np.sqrt(np.pi) * np.exp(-(alpha * r) ** 2) / r
        U[i0:i1, 0] = ((w * kew)[..., None] * rhat).sum(1)
        xr = 2.0 * r / rg - 1.0
        tm1, tm = np.ones_like(r), xr
        for m in range(NV_USED - 1):
            g = tm1 if m == 0 else tm
            if m >= 1:
                tm1, tm = tm, 2.0 * xr * tm - tm1
            U[i0:i1, 1 + m] = ((w * g)[..., None] * rhat).sum(1)
    return feat, U


def normalisation(feat, U):
    """mu, sigma per input (sigma 1 where an input does not vary) and su_m = rms |U_m|
    (1 where zero), in FP64 over all sites of one frame."""
    mu = feat.mean(0)
    sd = np.sqrt(np.maximum((feat * feat).mean(0) - mu * mu, 0.0))
    sd = np.where(sd > 1e-6 * np.maximum(1.0, np.abs(mu)), sd, 1.0)
    su = np.sqrt((U * U).sum(-1).mean(0))
    su = np.where(su > 0, su, 1.0)
    return mu, sd, su


def bf16(x):
    """Round to bfloat16, nearest even (finite inputs), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    b = a.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return (b & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(64, 48), b1=p[OFF_B1:OFF_W2], W2=p[OFF_W2:OFF_B2].reshape(64, 64),
                b2=p[OFF_B2:OFF_W3], W3=p[OFF_W3:OFF_B3].reshape(16, 64), b3=p[OFF_B3:N_PARAMS])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3]).astype(np.float64)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, xn, uh, emulate=True):
    """xn [n, 48] normalised inputs, uh [n, 16, 3] = U / su.  Returns the activations and
    the predicted forces F [n, 3]."""
    P = unpack(p)
    rb = (lambda a: bf16(a).astype(np.float64)) if emulate else (lambda a: np.asarray(a, np.float64))
    x = rb(xn)
    W1, W2, W3 = rb(P["W1"]), rb(P["W2"]), rb(P["W3"])
    z1 = x @ W1.T + P["b1"]
    h1 = rb(_silu(z1))
    z2 = h1 @ W2.T + P["b2"]
    h2 = rb(_silu(z2))
    c = h2 @ W3.T + P["b3"]
    F = np.einsum("nm,nmk->nk", c, uh)
    return dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, c=c, F=F, W1=W1, W2=W2, W3=W3)


def loss_and_grad(p, xn, uh, T, emulate=True):
    """Loss of one frame and its gradient with respect to the flat parameters."""
    rb = (lambda a: bf16(a).astype(np.float64)) if emulate else (lambda a: np.asarray(a, np.float64))
    f = forward(p, xn, uh, emulate)
    n = len(xn)
    t2 = float((T * T).sum(1).mean())
    diff = f["F"] - T
    loss = float((diff * diff).sum() / (n * t2))
    gF = 2.0 * diff / (n * t2)
    return loss, _backward(f, uh, gF, emulate), f


def change_loss_and_grad(p, xa, ua, xb, ub, Ta, Tb, emulate=True):
    """Loss of the change between frames a and b and its gradient: frame a's gradient with
    the error, frame b's with its negative."""
    fa = forward(p, xa, ua, emulate)
    fb = forward(p, xb, ub, emulate)
    dT = Ta - Tb
    n = len(xa)
    t2 = float((dT * dT).sum(1).mean())
    e = (fa["F"] - fb["F"]) - dT
    loss = float((e * e).sum() / (n * t2))
    ga = _backward(fa, ua, 2.0 * e / (n * t2), emulate)
    gb = _backward(fb, ub, -2.0 * e / (n * t2), emulate)
    return loss, ga + gb


def _backward(f, uh, gF, emulate):
    rb = (lambda a: bf16(a).astype(np.float64)) if emulate else (lambda a: np.asarray(a, np.float64))
    dc = np.einsum("nk,nmk->nm", gF, uh)
    dcb = rb(dc)
    dz2 = (dcb @ f["W3"]) * _dsilu(f["z2"])
    dz2b = rb(dz2)
    dz1 = (dz2b @ f["W2"]) * _dsilu(f["z1"])
    dz1b = rb(dz1)
    return pack(dz1b.T @ f["x"], dz1.sum(0), dz2b.T @ f["h1"], dz2.sum(0), dcb.T @ f["h2"], dc.sum(0))


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected moments, Kingma & Ba 2015, Algorithm 1)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
