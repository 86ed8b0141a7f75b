"""Reciprocal-force surrogate with multiple time steps on the GPU: its kernels against the
CPU statement (oracle/surrogate.py), the scheme's schedule, fallback and energies, and what
the trained network gains over holding the last full sum's force (the method without it)."""
import sys

import numpy as np
import pytest

from md_sim_amd import abi, systems
from helpers import libration_frames, rms_rel
from oracle import surrogate as sg

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _weights(seed=0, n_norm=True):
    rng = np.random.default_rng(seed)
    p = np.zeros(sg.N_PARAMS)
    p[sg.OFF_W1:sg.OFF_B1] = rng.normal(0, 1 / np.sqrt(48), 64 * 48)
    p[sg.OFF_B1:sg.OFF_W2] = rng.normal(0, 0.1, 64)
    p[sg.OFF_W2:sg.OFF_B2] = rng.normal(0, 1 / np.sqrt(64), 64 * 64)
    p[sg.OFF_B2:sg.OFF_W3] = rng.normal(0, 0.1, 64)
    p[sg.OFF_W3:sg.OFF_B3] = rng.normal(0, 0.3, 16 * 64)
    p[sg.OFF_B3:] = rng.normal(0, 0.3, 16)
    mu = rng.normal(0, 0.5, 48)
    sd = rng.uniform(0.5, 2.0, 48)
    su = rng.uniform(0.5, 2.0, 16)
    return np.concatenate([p, mu, sd, su]).astype(np.float32)


def _leads(s):
    """lead site of every site: the first site of each rigid molecule, else the site"""
    lead = np.arange(s.ntota)
    lead[:3 * s.nrig] = 3 * (np.arange(3 * s.nrig) // 3)
    return lead


def _oracle_inputs(w, feat, vec):
    """normalised inputs exactly as the kernels form them (FP32)"""
    w = np.asarray(w, np.float32)
    mu, sd, su = w[sg.N_PARAMS:sg.N_PARAMS + 48], w[sg.N_PARAMS + 48:sg.N_PARAMS + 96], w[sg.N_PARAMS + 96:]
    isd = (np.float32(1) / sd).astype(np.float32)
    xn = ((feat.astype(np.float32) - mu) * isd).astype(np.float32)
    uh = vec.astype(np.float64) / su.astype(np.float64)[None, :, None]
    return w[:sg.N_PARAMS].astype(np.float64), xn.astype(np.float64), uh


def test_inputs_match_oracle(oracle_lib):
    """Inputs and vector features of every moving site (FP32 kernels) against the FP64
    brute-force statement over all site pairs."""
    from md_sim_amd.engine import ForceEngine
    _say("case surrogate inputs slab512")
    s = systems.config("slab512")
    e = ForceEngine(s)
    pos = abi.positions(s)
    e.force_get(*pos)
    e.mts_enable(train_frames=0)
    feat, vec = e.mts_features()
    c = oracle_lib.Oracle(s).constants()
    q = np.asarray(s.charges(), np.float64)
    P = np.stack(pos[:3], 1)
    rf, rv = sg.features(P, q, s.nmove, s.box, False, c["alfa"], c["facq"], _leads(s), s.rcut)
    for k in range(39):
        col = rf[:, k]
        scale = max(np.sqrt((col ** 2).mean()), 1e-6)
        assert np.abs(feat[:, k] - col).max() <= 3e-4 * scale, k
    assert np.abs(feat[:, 39:]).max() == 0
    for m in range(16):
        scale = np.sqrt((rv[:, m] ** 2).sum(1).mean())
        assert np.abs(vec[:, m] - rv[:, m]).max() <= 3e-4 * scale, m
    e.close()


def test_network_and_gradient_match_oracle():
    """Forward pass, loss on the change between two frames, its gradient (bfloat16
    operands, FP32 accumulation on the matrix cores) against the CPU statement with the same
    roundings, over 300 sites (a partial last tile); and one Adam step from that gradient."""
    from md_sim_amd.engine import mts_eval
    _say("case surrogate network gradient")
    rng = np.random.default_rng(2)
    n = 300
    w = _weights(1)
    frames = []
    for _ in range(2):
        feat = rng.normal(0, 1.5, (n, 48)).astype(np.float32)
        feat[:, 39:] = 0
        vec = rng.normal(0, 1.0, (n, 16, 3)).astype(np.float32)
        frames.append((feat, vec, rng.normal(0, 1.0, (n, 3))))
    (fa, va, Ta), (fb, vb, Tb) = frames
    out = mts_eval(w, fa, va, Ta, fb, vb, Tb, lr=1e-3, adam_t=0)
    p, xa, ua = _oracle_inputs(w, fa, va)
    _, xb, ub = _oracle_inputs(w, fb, vb)
    loss, g = sg.change_loss_and_grad(p, xa, ua, xb, ub, Ta, Tb, emulate=True)
    f = sg.forward(p, xa, ua, emulate=True)
    assert rms_rel(out["force"], f["F"]) < 1e-5
    assert abs(out["loss"] - loss) < 1e-5 * loss
    gg = out["grad"].astype(np.float64)
    assert np.linalg.norm(gg - g) < 2e-3 * np.linalg.norm(g)
    for a, b in ((sg.OFF_W1, sg.OFF_B1), (sg.OFF_B1, sg.OFF_W2), (sg.OFF_W2, sg.OFF_B2), (sg.OFF_B2, sg.OFF_W3),
                 (sg.OFF_W3, sg.OFF_B3), (sg.OFF_B3, sg.N_PARAMS)):
        assert np.linalg.norm(gg[a:b] - g[a:b]) < 5e-3 * np.linalg.norm(g[a:b]) + 1e-9, (a, b)
    p1, _, _ = sg.adam(p, gg, np.zeros_like(p), np.zeros_like(p), 1, 1e-3)
    assert np.abs(out["params"] - p1).max() < 1e-6
    # forward only
    out0 = mts_eval(w, fa, va)
    assert np.array_equal(out0["force"], out["force"])


def _frames_and_exact(s, n_steps, seed=0):
    from md_sim_amd.engine import ForceEngine
    frames = libration_frames(s, n_steps, seed=seed)
    e = ForceEngine(s, precision=abi.FP32_ACC64)
    ref = [e.force_get(*f) for f in frames]
    e.close()
    return frames, ref


def _run(s, frames, **kw):
    from md_sim_amd.engine import ForceEngine
    e = ForceEngine(s, precision=abi.FP32_ACC64)
    e.mts_enable(**kw)
    outs = [e.force_get(*f) for f in frames]
    st = e.mts_stats()
    e.close()
    return outs, st


def test_held_force_on_a_static_system():
    """Positions that do not move: every step of the scheme gives the full-sum forces and
    energies; the schedule is one full sum every `interval` steps, and each judge finds 0."""
    from md_sim_amd.engine import ForceEngine
    _say("case surrogate static held")
    s = systems.config("slab512")
    pos = abi.positions(s)
    e0 = ForceEngine(s, precision=abi.FP32_ACC64)
    r = e0.force_get(*pos)
    e0.close()
    outs, st = _run(s, [pos] * 9, model="held", interval=4, train_frames=0)
    for t, o in enumerate(outs):
        assert rms_rel(o["f_site"], r["f_site"]) < 1e-12, t
        assert abs(o["e_conf"] - r["e_conf"]) < 1e-9 * abs(r["e_conf"]), t
    assert st["n_exact"] == 3 and st["n_surrogate"] == 6 and st["n_fallback"] == 0
    assert st["n_judged"] == 2 and st["err_max"] < 1e-12


def test_network_gains_over_held_force():
    """A librating, drifting water film (H speeds of water at 300 K and 2 fs steps): after
    training on the first full-sum frames, the network's steps come closer to the full sum
    than the held force (the method without the network) and no further than the untrained
    physics prior (the near-field k-space force of whole molecules).  Printed: the mean force
    RMS errors of the three on steps without the full sum, relative to the RMS site force."""
    _say("case surrogate gain slab512")
    s = systems.config("slab512")
    n_train, n_steps = 60, 140
    frames, ref = _frames_and_exact(s, n_steps)
    kw = dict(interval=4, train_frames=n_train, tol=0.2)
    held, st_h = _run(s, frames, model="held", **kw)
    prior, st_p = _run(s, frames, model="network", train_iters=0, **kw)
    net, st_n = _run(s, frames, model="network", train_iters=2, lr=1e-3, **kw)
    for k, st in (("held", st_h), ("prior", st_p), ("network", st_n)):
        _say(k, st)
    steps = [t for t in range(n_train, n_steps) if (t - n_train + 1) % 4]   # steps without the full sum
    err = {k: float(np.mean([rms_rel(o[t]["f_site"], ref[t]["f_site"]) for t in steps]))
           for k, o in (("held", held), ("prior", prior), ("network", net))}
    _say("mean force RMS error on surrogate steps: %s" % err)
    for st in (st_h, st_p, st_n):
        assert st["n_fallback"] == 0 and st["n_surrogate"] == len(steps)
    assert err["prior"] < 0.6 * err["held"]
    assert err["network"] < 0.6 * err["held"]
    assert err["network"] <= 1.02 * err["prior"]
    # the judge (4 steps after each correction) sees more than the 1-3 step errors
    assert err["network"] <= st_n["err_mean"] <= 4.0 * err["network"]


def test_fallback_on_tolerance():
    """A tolerance no prediction meets: the first judged full sum falls back, and the next
    train_frames steps are full sums (forces equal the full-sum run's)."""
    _say("case surrogate fallback")
    s = systems.config("slab512")
    frames, ref = _frames_and_exact(s, 16, seed=3)
    outs, st = _run(s, frames, model="network", interval=4, train_frames=2, train_iters=1, tol=1e-12)
    assert st["n_fallback"] >= 1
    # steps 0-1 collect, 2-4 surrogate, 5 judged full sum -> fallback (collects), 6 collects
    for t in (0, 1, 5, 6):
        assert rms_rel(outs[t]["f_site"], ref[t]["f_site"]) < 1e-9, t


def test_energy_along_the_path():
    """On steps without the full sum, e_recip is the last full sum's value plus the work of
    the scheme's k-space force along the moved path: closer to the full sum's e_recip than
    the last full sum's value is, and equal to it on full-sum steps."""
    _say("case surrogate energy path")
    s = systems.config("slab512")
    frames, ref = _frames_and_exact(s, 24, seed=5)
    outs, _ = _run(s, frames, model="held", interval=4, train_frames=0)
    err_path, err_hold = 0.0, 0.0
    for t in range(1, 24):
        if t % 4 == 0:
            assert abs(outs[t]["e_recip"] - ref[t]["e_recip"]) < 1e-8 * abs(ref[t]["e_recip"])
        else:
            t0 = 4 * (t // 4)
            err_path += abs(outs[t]["e_recip"] - ref[t]["e_recip"])
            err_hold += abs(ref[t0]["e_recip"] - ref[t]["e_recip"])
    _say("e_recip error: path integral %.3e, held value %.3e" % (err_path, err_hold))
    assert err_path < err_hold


def _relaxed_slab512(n_iter=300):
    """slab512 after rigid-body steepest descent on the engine's energy (the same
    procedure as tests/golden/make_nve.py), with 1 fs steps."""
    import dataclasses
    from golden.make_nve import _site_jacobian
    from md_sim_amd.engine import ForceEngine
    s = systems.config("slab512")
    e = ForceEngine(s)

    def energy(sys_):
        r = e.force_get(*abi.positions(sys_))
        return r["e_conf"] + r["e_wall_lr"], r
    en, r = energy(s)
    step = 0.02
    for _ in range(n_iter):
        g_com = -r["f_com"].T
        fs = r["f_site"][:, :s.nriga].T.reshape(s.nrig, 3, 3)
        g_q = -np.einsum("mksx,msx->mk", _site_jacobian(s), fs)
        h = step / max(np.abs(g_com).max(), 2.0 * np.abs(g_q).max())
        q = s.quat - h * g_q
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        trial = dataclasses.replace(s, com=s.com - h * g_com, quat=q)
        et, rt = energy(trial)
        if et < en:
            s, en, r = trial, et, rt
            step = min(step * 1.2, 0.1)
        else:
            step *= 0.5
    e.close()
    return dataclasses.replace(s, dt=1.0)


def test_nve_drift_with_multiple_time_steps(tmp_path):
    """The reference's own time loop (Gear predictor-corrector, quaternions) on the engine's
    drop-in force module, 1000 steps of 1 fs from a relaxed slab512: with the network and the
    full sum every 2nd or every 4th step the htot drift stays within twice the FP64 run's and
    below the held force's; with the 1e-3 tolerance the judge falls back to full sums and the
    drift is the full sum's."""
    import os
    from golden.make_nve import drift
    from oracle import refrun
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "md_nve_gpu")
    if not os.access(exe, os.X_OK):
        pytest.skip("NVE driver with the engine's force module not built (oracle/build_ref.sh)")
    _say("case surrogate nve slab512")
    s = _relaxed_slab512()
    n = 1000
    mts = {"MDF_PRECISION": "fp32", "MDF_MTS_TRAIN": "100"}
    runs = {"fp64": {"MDF_PRECISION": "fp64"},
            "network/2": dict(mts, MDF_MTS="2", MDF_MTS_TOL="0.05"),
            "network/4": dict(mts, MDF_MTS="4", MDF_MTS_TOL="0.05"),
            "held/4": dict(mts, MDF_MTS="4", MDF_MTS_TOL="0.05", MDF_MTS_MODEL="held"),
            "network/4 tol 1e-3": dict(mts, MDF_MTS="4", MDF_MTS_LOG=str(tmp_path / "mts.log"))}
    d = {}
    for name, env in runs.items():
        tr = refrun.run_nve(s, n, variant="gpu", env=env, timeout=300)
        assert np.all(np.isfinite(tr[:, 1])), name
        d[name] = drift(tr)
    _say("  htot drift over %d steps (kJ/mol): %s" % (n, {k: round(v, 3) for k, v in d.items()}))
    assert abs(d["network/2"]) <= 2.0 * abs(d["fp64"]) + 1.0
    assert abs(d["network/4"]) <= 2.0 * abs(d["fp64"]) + 1.0
    assert abs(d["network/4"]) < abs(d["held/4"])
    assert abs(d["network/4 tol 1e-3"]) <= 2.0 * abs(d["fp64"]) + 1.0
    log = np.loadtxt(tmp_path / "mts.log", ndmin=2)
    assert log[-1, 4] >= 1          # the judge fell back at least once


def test_training_only_on_full_sum_steps():
    """Training happens on full-sum steps only, train_iters Adam steps each: every collected
    frame from the second on, then every full sum of the scheme (online), none on the steps
    in between (whose targets would be stale)."""
    _say("case surrogate training count")
    s = systems.config("slab512")
    frames, _ = _frames_and_exact(s, 30, seed=11)
    _, st = _run(s, frames, model="network", interval=4, train_frames=10, train_iters=2, tol=0.5)
    # frames 0..9 collect (9 pairs), then full sums at steps 13, 17, 21, 25, 29 (5 pairs)
    assert st["n_exact"] == 15 and st["n_surrogate"] == 15
    assert st["adam_steps"] == 2 * (9 + 5)
    assert st["loss_last"] < 1.0
