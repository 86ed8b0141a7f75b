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
    from md_sim_amd.systems import libration_frames as lf
    return lf(system, n_steps, theta0=theta0, amp=amp, seed=seed)
