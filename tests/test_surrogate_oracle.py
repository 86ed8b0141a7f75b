"""The surrogate's CPU statement (oracle/surrogate.py) pinned by what the mathematics fixes:
invariance of the inputs and equivariance of the vector features under rigid motions, a
two-charge closed form, torch's bfloat16 rounding, torch autograd gradients, central
differences, and torch's Adam."""
import numpy as np
import pytest
import torch

from oracle import surrogate as sg


def _cluster(n=40, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-5.0, 5.0, size=(n, 3))
    q = rng.choice([-0.8476, 0.4238, 1.0, -1.0], size=n)
    return pos, q


def _rot(seed=1):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(3, 3))
    Q, R = np.linalg.qr(a)
    Q = Q * np.sign(np.diag(R))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def test_inputs_invariant_vectors_equivariant():
    pos, q = _cluster()
    box = (1000.0, 1000.0, 1000.0)
    lead = np.arange(len(pos))
    lead[1:3] = 0                      # sites 0-2 form one molecule
    x0, U0 = sg.features(pos, q, len(pos), box, False, 0.31, 1389.35, lead, 8.0)
    Q = _rot()
    x1, U1 = sg.features(pos @ Q.T + np.array([3.0, -2.0, 7.0]), q, len(pos), box, False, 0.31, 1389.35, lead, 8.0)
    assert np.allclose(x1, x0, rtol=1e-10, atol=1e-12)
    assert np.allclose(U1, U0 @ Q.T, rtol=1e-10, atol=1e-9)
    assert np.abs(x0[:, 16:37]).max() > 0 and np.abs(U0).max() > 0


def test_two_charges_closed_form():
    """Two sites 2.5 A apart along x: every entry has a one-term closed form."""
    pos = np.array([[0.0, 0.0, 0.0], [2.5, 0.0, 0.0]])
    q = np.array([-0.8, 0.5])
    a, facq = 0.3, 1389.35
    x, U = sg.features(pos, q, 2, (100.0, 100.0, 100.0), False, a, facq, np.arange(2), 8.0)
    r = 2.5
    fc = 0.5 * (np.cos(np.pi * r / 6.0) + 1.0)
    from scipy.special import erf
    # site 0 sees a positive charge at +x: r_01 = r_0 - r_1 points along -x
    assert np.allclose(x[0, :8], 0.5 * fc * np.exp(-2.0 * (r - sg.RAD_C) ** 2))
    assert np.allclose(x[0, 8:16], 0.0)
    A00 = 0.5 * fc * np.exp(-1.0 * (r - 1.5) ** 2)
    assert np.isclose(x[0, 16], A00 * A00)
    kew = erf(a * r) / r ** 2 - 2 * a / np.sqrt(np.pi) * np.exp(-(a * r) ** 2) / r
    # vector features: group weight 1 well inside the group cut-off
    assert np.allclose(U[0, 0], facq * (-0.8) * 0.5 * kew * np.array([-1.0, 0, 0]))
    assert np.allclose(U[1, 0], facq * 0.5 * (-0.8) * kew * np.array([1.0, 0, 0]))
    xr = 2 * r / 8.0 - 1
    for m, tm in ((1, 1.0), (2, xr), (3, 2 * xr * xr - 1), (4, 4 * xr ** 3 - 3 * xr)):
        assert np.allclose(U[0, m], facq * (-0.8) * 0.5 * tm * np.array([-1.0, 0, 0]))
    assert x[0, 37] == -0.8 and np.isclose(x[0, 38], 0.64)
    # Newton's third law for the pair's vector features
    assert np.allclose(U[0], -U[1])


def test_minimum_image():
    """A pair across the x boundary is seen at its short image; z stays open unless periodic."""
    pos = np.array([[0.5, 5.0, 5.0], [19.5, 5.0, 5.0], [0.5, 5.0, 19.0]])
    q = np.array([1.0, -1.0, -1.0])
    L3 = (20.0, 20.0, 20.0)
    x, _ = sg.features(pos, q, 3, L3, False, 0.3, 1.0, np.arange(3), 8.0)
    assert x[0, 8:16].sum() > 0          # 1 A across x
    xz, _ = sg.features(pos[[0, 2]], q[[0, 2]], 2, L3, False, 0.3, 1.0, np.arange(2), 8.0)
    assert xz[0, 8:16].sum() == 0        # 14 A apart in open z
    xp, _ = sg.features(pos[[0, 2]], q[[0, 2]], 2, L3, True, 0.3, 1.0, np.arange(2), 8.0)
    assert xp[0, 8:16].sum() == 0        # still 6 A apart in periodic z: outside RC
    pos2 = pos.copy()
    pos2[2, 2] = 23.0                    # 2.5 A across the periodic z boundary
    xp2, _ = sg.features(pos2[[0, 2]], q[[0, 2]], 2, L3, True, 0.3, 1.0, np.arange(2), 8.0)
    assert xp2[0, 8:16].sum() > 0


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8], np.float32)])
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(sg.bf16(x), ref)


def _problem(n=50, seed=5, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, size=sg.N_PARAMS)
    xn = rng.normal(size=(n, sg.D_IN))
    uh = rng.normal(size=(n, sg.NV, 3))
    T = rng.normal(size=(n, 3))
    return p, xn, uh, T


def test_gradient_matches_torch_autograd():
    p, xn, uh, T = _problem()
    loss, g, _ = sg.loss_and_grad(p, xn, uh, T, emulate=False)
    pt = torch.tensor(p, requires_grad=True)
    P = {}
    off = 0
    for name, shape in (("W1", (64, 48)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (16, 64)),
                        ("b3", (16,))):
        k = int(np.prod(shape))
        P[name] = pt[off:off + k].reshape(shape)
        off += k
    x = torch.tensor(xn)
    h1 = torch.nn.functional.silu(x @ P["W1"].T + P["b1"])
    h2 = torch.nn.functional.silu(h1 @ P["W2"].T + P["b2"])
    c = h2 @ P["W3"].T + P["b3"]
    F = torch.einsum("nm,nmk->nk", c, torch.tensor(uh))
    Tt = torch.tensor(T)
    L = ((F - Tt) ** 2).sum() / (len(xn) * (Tt * Tt).sum(1).mean())
    L.backward()
    assert abs(L.item() - loss) < 1e-12 * abs(loss)
    assert np.allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_gradient_central_differences():
    p, xn, uh, T = _problem(n=20, seed=7)
    _, g, _ = sg.loss_and_grad(p, xn, uh, T, emulate=False)
    rng = np.random.default_rng(8)
    picks = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in
                            ((sg.OFF_W1, sg.OFF_B1), (sg.OFF_B1, sg.OFF_W2), (sg.OFF_W2, sg.OFF_B2),
                             (sg.OFF_B2, sg.OFF_W3), (sg.OFF_W3, sg.OFF_B3), (sg.OFF_B3, sg.N_PARAMS))])
    h = 1e-6
    for k in picks:
        e = np.zeros_like(p)
        e[k] = h
        lp = sg.loss_and_grad(p + e, xn, uh, T, emulate=False)[0]
        lm = sg.loss_and_grad(p - e, xn, uh, T, emulate=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g[k])), (k, fd, g[k])


def test_emulated_gradient_close_to_exact():
    """bfloat16 operand rounding perturbs the gradient by about the bfloat16 step, no more."""
    p, xn, uh, T = _problem(n=200, seed=9)
    _, g0, _ = sg.loss_and_grad(p, xn, uh, T, emulate=False)
    _, g1, _ = sg.loss_and_grad(p, xn, uh, T, emulate=True)
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)


def test_adam_matches_torch():
    rng = np.random.default_rng(11)
    p = rng.normal(size=100)
    m = np.zeros(100)
    v = np.zeros(100)
    pt = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 6):
        g = rng.normal(size=100)
        p, m, v = sg.adam(p, g, m, v, t, 1e-2)
        opt.zero_grad()
        pt.grad = torch.tensor(g)
        opt.step()
        assert np.allclose(p, pt.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_group_weight_keeps_molecules_whole():
    """A neutral three-site molecule whose lead sits inside the group cut-off contributes
    with all its sites even when a site lies beyond it, and the smooth step falls from 1 to
    0 over the last WG of the cut-off."""
    rg = 8.0
    mol = np.array([[7.9, 0.0, 0.0], [8.9, 0.0, 0.0], [7.5, 0.9, 0.0]])
    pos = np.vstack([[0.0, 0.0, 0.0], mol])
    q = np.array([1.0, -0.8, 0.4, 0.4])
    lead = np.array([0, 1, 1, 1])
    _, U = sg.features(pos, q, 1, (100.0, 100.0, 100.0), False, 0.3, 1.0, lead, rg)
    s = (7.9 - rg + sg.WG) / sg.WG
    wg = 1 - s * s * (3 - 2 * s)
    d = pos[0] - mol
    r = np.linalg.norm(d, axis=1)
    from scipy.special import erf
    kew = erf(0.3 * r) / r ** 2 - 0.6 / np.sqrt(np.pi) * np.exp(-(0.3 * r) ** 2) / r
    want = (q[1:, None] * wg * kew[:, None] * d / r[:, None]).sum(0)
    assert np.allclose(U[0, 0], want, rtol=1e-12)
    _, U2 = sg.features(pos + np.array([[0, 0, 0], [0.2, 0, 0], [0.2, 0, 0], [0.2, 0, 0]]), q, 1,
                        (100.0, 100.0, 100.0), False, 0.3, 1.0, lead, rg)
    assert np.allclose(U2[0], 0.0)       # lead at 8.1 A: the whole molecule drops out


def test_change_gradient_matches_torch_autograd():
    p, xa, ua, Ta = _problem(n=40, seed=21)
    _, xb, ub, Tb = _problem(n=40, seed=22)
    loss, g = sg.change_loss_and_grad(p, xa, ua, xb, ub, Ta, Tb, emulate=False)
    pt = torch.tensor(p, requires_grad=True)

    def net(x, u):
        W1 = pt[sg.OFF_W1:sg.OFF_B1].reshape(64, 48); b1 = pt[sg.OFF_B1:sg.OFF_W2]
        W2 = pt[sg.OFF_W2:sg.OFF_B2].reshape(64, 64); b2 = pt[sg.OFF_B2:sg.OFF_W3]
        W3 = pt[sg.OFF_W3:sg.OFF_B3].reshape(16, 64); b3 = pt[sg.OFF_B3:]
        h = torch.nn.functional.silu(torch.tensor(x) @ W1.T + b1)
        h = torch.nn.functional.silu(h @ W2.T + b2)
        return torch.einsum("nm,nmk->nk", h @ W3.T + b3, torch.tensor(u))
    dT = torch.tensor(Ta - Tb)
    e = net(xa, ua) - net(xb, ub) - dT
    L = (e ** 2).sum() / (40 * (dT * dT).sum(1).mean())
    L.backward()
    assert abs(L.item() - loss) < 1e-12 * loss
    assert np.allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-13)
