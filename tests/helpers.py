"""Shared helpers for the parity tests."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def rel_err(a, b):
    """max |a-b| / max |b|"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max() if b.size else 0.0
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def rms_rel(a, b):
    """force RMS error relative to the RMS force"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).mean()) / np.sqrt((b ** 2).mean()))


def golden_positions(g):
    return g["x0"], g["y0"], g["z0"], g["cx0"], g["cy0"], g["cz0"]


def libration_frames(system, n_steps, theta0=0.3, amp=0.2, seed=0):
    """Positions of a synthetic trajectory with rotations: every rigid molecule librates
    about its COM (angle theta0 sin(2 pi t / P + phi) about its own random axis, P in
    30..60 steps) while its COM oscillates (amplitude amp per axis, P in 40..80 steps);
    other moving sites oscillate like COMs.  Peak H speeds are ~0.05 A per step, about
    those of SPC/E water at 300 K with 2 fs steps.  Returns a list of (x, y, z, cx, cy, cz)."""
    from md_sim_amd import abi
    x0, y0, z0, cx0, cy0, cz0 = abi.positions(system)
    nr, nm = system.nrig, system.nmove
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=(nr, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    pr = rng.uniform(30, 60, nr)
    phr = rng.uniform(0, 2 * np.pi, nr)
    pt = rng.uniform(40, 80, (nm, 3))
    pht = rng.uniform(0, 2 * np.pi, (nm, 3))
    at = rng.uniform(-amp, amp, (nm, 3))
    com0 = np.stack([cx0, cy0, cz0], 1)
    site0 = np.stack([x0, y0, z0], 1)
    frames = []
    for t in range(n_steps):
        pos = site0.copy()
        th = theta0 * np.sin(2 * np.pi * t / pr + phr) - theta0 * np.sin(phr)
        c, s = np.cos(th), np.sin(th)
        shift = at * (np.sin(2 * np.pi * t / pt + pht) - np.sin(pht))
        com = com0 + shift[:3 * nr:3] if nr else com0
        for a in range(3):
            rel = site0[a:3 * nr:3] - com0
            kx = np.cross(axis, rel)
            kd = (axis * rel).sum(1, keepdims=True)
            rot = rel * c[:, None] + kx * s[:, None] + axis * kd * (1 - c)[:, None]
            pos[a:3 * nr:3] = com + rot
        if nm > 3 * nr:
            pos[3 * nr:nm] += shift[3 * nr:nm]
        frames.append((pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(),
                       com[:, 0].copy(), com[:, 1].copy(), com[:, 2].copy()))
    return frames
