"""Host-side mirror of the reference's force module on top of the HIP library.

The reference's force module (src/forcem.f90) exposes ``force_init`` (:54), ``force_zero``
(:84) and ``force_get`` (:93) and talks through module variables; ``ForceEngine`` exposes
the same three operations on explicit arrays and returns the reference's outputs (site
forces before the COM reduction, COM forces, torques, energies per type pair, reciprocal
energies, slab and wall terms, uconf, pressure from wall forces) plus a virial.  A status
other than MDF_OK raises ``MDForceError`` carrying the library's message, where the
reference would ``stop``.

There is no CPU fallback: without the built library (``libmdforce.so``) or without a HIP
device, construction fails.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdforce.so")


class MDForceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (%d): %s" % (abi.STATUS_NAMES.get(code, "?"), code, msg))
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """The force engine's shared library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmdforce.so is not built (run __graft_entry__.build() or "
                              "python -m md_sim_amd.build)")
        # The library needs the HIP runtime as "libamdhip64.so".  PyTorch-ROCm ships its own
        # runtime under that name; loading torch first makes the library bind to it, so a
        # process that also uses torch tensors has one HIP runtime, not two.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.mdf_last_error.restype = C.c_char_p
        for fn in ("mdf_create", "mdf_compute", "mdf_compute_device", "mdf_read_outputs", "mdf_get_pairs",
                   "mdf_get_kvectors", "mdf_get_structure_factor", "mdf_stage_times", "mdf_device_results",
                   "mdf_zero", "mdf_reset_stage_times", "mdf_set_stage_timing", "mdf_autotune",
                   "mdf_set_launch_shapes", "mdf_batch_create", "mdf_batch_compute_device",
                   "mdf_batch_read_outputs", "mdf_batch_autotune", "mdf_mts_enable", "mdf_mts_get_stats",
                   "mdf_mts_get_weights", "mdf_mts_set_weights", "mdf_mts_features", "mdf_mts_eval",
                   "mdf_observables_enable", "mdf_get_observables", "mdf_chempot_enable", "mdf_get_chempot",
                   "mdf_reciprocal_kernels", "mdf_mts_eval_device", "mdf_angle_histogram_enable",
                   "mdf_get_angle_histogram"):
            getattr(L, fn).restype = C.c_int
        L.mdf_destroy.restype = None
        L.mdf_batch_destroy.restype = None
        _lib = L
    return _lib


def _check(rc: int):
    if rc != abi.MDF_OK:
        raise MDForceError(rc, lib().mdf_last_error().decode())


def _producer_stream(stream):
    if stream is not None:
        return int(stream)
    import sys as _sys
    torch = _sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        return int(torch.cuda.current_stream().cuda_stream) or None
    return None


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class ForceEngine:
    """One system on one GPU.  ``ForceEngine(system)`` is force_init; ``force_get`` is
    force_get; ``force_zero`` is force_zero."""

    def __init__(self, system, positions=None, **abi_kw):
        self.system = system
        self.sa = abi.build_abi(system, **abi_kw)
        if positions is None:
            positions = abi.positions(system)
        x, y, z, cx, cy, cz = [np.ascontiguousarray(a, np.float64) for a in positions]
        self._h = C.c_void_p()
        _check(lib().mdf_create(C.byref(self.sa.top), C.byref(self.sa.par), _dp(x), _dp(y), _dp(z),
                                _dp(cx), _dp(cy), _dp(cz), C.byref(self._h)))
        self.n_move, self.n_rig, self.n_pairs = system.nmove, system.nrig, system.nijt

    # -- reference entry points --------------------------------------------------------
    def force_get(self, x, y, z, cx, cy, cz) -> abi.ForceResult:
        keep = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, cx, cy, cz)]
        o, arrs = abi.make_outputs(self.n_move, self.n_rig, self.n_pairs)
        _check(lib().mdf_compute(self._h, *[_dp(k) for k in keep], C.byref(o)))
        return abi.collect(o, arrs, self.n_move, self.n_rig)

    def force_zero(self):
        _check(lib().mdf_zero(self._h))

    # -- device-resident path ------------------------------------------------------------
    def compute_device(self, d_x: int, d_y: int, d_z: int, d_cx: int, d_cy: int, d_cz: int, stream=None):
        """Enqueue a step on device pointers (ints, e.g. torch tensor.data_ptr()); it starts
        after the work already on ``stream`` (a hipStream_t as int; default: torch's current
        stream when torch is loaded, else the null stream)."""
        vp = C.c_void_p
        _check(lib().mdf_compute_device(self._h, vp(d_x), vp(d_y), vp(d_z), vp(d_cx), vp(d_cy), vp(d_cz),
                                        vp(_producer_stream(stream))))

    def read_outputs(self) -> abi.ForceResult:
        """Wait for the last ``compute_device`` step and return its results.  When a list row
        overflowed before that step, the steps since ran on truncated rows: this raises
        MDForceError (code MDF_ERR_LIST_OVERFLOW); if the capacity could grow, the last step
        was redone and the error carries its valid results as ``.result``."""
        o, arrs = abi.make_outputs(self.n_move, self.n_rig, self.n_pairs)
        rc = lib().mdf_read_outputs(self._h, C.byref(o))
        if rc != abi.MDF_OK:
            err = MDForceError(rc, lib().mdf_last_error().decode())
            if rc == abi.MDF_ERR_LIST_OVERFLOW and o.n_list_builds > 0:   # outputs were written
                err.result = abi.collect(o, arrs, self.n_move, self.n_rig)
            raise err
        return abi.collect(o, arrs, self.n_move, self.n_rig)

    # -- inspection ------------------------------------------------------------------------
    def pairs(self):
        n1, n2 = C.c_int32(), C.c_int32()
        pi = lambda v: v.ctypes.data_as(C.POINTER(C.c_int32))
        z = np.zeros(1, np.int32)
        lib().mdf_get_pairs(self._h, pi(z), pi(z), 0, C.byref(n1), pi(z), pi(z), 0, C.byref(n2))
        a, b = np.zeros(max(n1.value, 1), np.int32), np.zeros(max(n1.value, 1), np.int32)
        c, d = np.zeros(max(n2.value, 1), np.int32), np.zeros(max(n2.value, 1), np.int32)
        _check(lib().mdf_get_pairs(self._h, pi(a), pi(b), n1.value, C.byref(n1), pi(c), pi(d), n2.value,
                                   C.byref(n2)))
        return a[:n1.value], b[:n1.value], c[:n2.value], d[:n2.value]

    def kvectors(self):
        n = C.c_int32()
        lib().mdf_get_kvectors(self._h, None, None, 0, C.byref(n))
        kv = np.zeros(3 * n.value, np.int32)
        ak = np.zeros(n.value)
        _check(lib().mdf_get_kvectors(self._h, kv.ctypes.data_as(C.POINTER(C.c_int32)), _dp(ak), n.value,
                                      C.byref(n)))
        sf = np.zeros(2 * max(n.value, 1))
        _check(lib().mdf_get_structure_factor(self._h, _dp(sf), n.value))
        return kv.reshape(-1, 3), ak, sf[0:2 * n.value:2] + 1j * sf[1:2 * n.value:2]

    def reciprocal_on_matrix_cores(self) -> bool:
        """True when both reciprocal-sum kernels run on the matrix cores (FP64 or FP16 hi/lo MFMA)."""
        a, b = C.c_int32(), C.c_int32()
        _check(lib().mdf_reciprocal_kernels(self._h, C.byref(a), C.byref(b)))
        return bool(a.value and b.value)

    # -- observables of the reference's hot loops ---------------------------------------
    def observables_enable(self, rdel: float):
        """Accumulate ngc and efipos (the reference's measure arrays) from the next step on,
        with bin width rdel; rdel <= 0 switches them off."""
        _check(lib().mdf_observables_enable(self._h, C.c_double(rdel)))

    def observables(self, reset: bool = False):
        """(ngc [nbins, n_pairs] int64, efipos [nbins, 2]) accumulated so far."""
        nb = C.c_int32()
        _check(lib().mdf_get_observables(self._h, None, None, C.byref(nb), 0))
        ngc = np.zeros((nb.value, self.n_pairs), np.int64)
        efp = np.zeros((nb.value, 2))
        _check(lib().mdf_get_observables(self._h, ngc.ctypes.data_as(C.POINTER(C.c_int64)), _dp(efp), C.byref(nb),
                                         int(bool(reset))))
        return ngc, efp

    def angle_histogram_enable(self, ang_type, n_atypes: int):
        """Accumulate force_intra's ngang from the next step on: ang_type[n_angles] is each
        angle term's 1-based group type; None switches it off."""
        if ang_type is None:
            _check(lib().mdf_angle_histogram_enable(self._h, None, 0))
            self._n_atypes = 0
            return
        self._ang_type = np.ascontiguousarray(ang_type, np.int32)
        _check(lib().mdf_angle_histogram_enable(self._h, self._ang_type.ctypes.data_as(C.POINTER(C.c_int32)),
                                                int(n_atypes)))
        self._n_atypes = int(n_atypes)

    def angle_histogram(self, reset: bool = False):
        """ngang [202, n_atypes] int64 accumulated so far."""
        out = np.zeros((202, self._n_atypes), np.int64)
        _check(lib().mdf_get_angle_histogram(self._h, out.ctypes.data_as(C.POINTER(C.c_int64)), int(bool(reset))))
        return out

    # -- Widom test-particle insertions (the reference's chempot) ------------------------
    def chempot_enable(self, n_insert: int, wall_min: float, wall_max: float, n_bins: int, bin_mid: int,
                       temperature: float = 0.0, ew3dc: bool = True, seed: float = 0.0, max_partners: int = 0):
        """Insert up to n_insert SPC/E waters after every step, as the reference's chempot
        (src/forcem.f90:1226-1507) with its tot.inp values ichmax, lwmin, lwmax, lchmax, ilmid
        (src/main.f90:148-152).  temperature <= 0 means 300 K; seed <= 0 the reference's
        start; ew3dc: the slab term with the step's coefficient (the reference reads an
        unassigned local there)."""
        p = abi.ChempotParams(n_insert=n_insert, n_bins=n_bins, bin_mid=bin_mid, ew3dc=int(bool(ew3dc)),
                              wall_min=wall_min, wall_max=wall_max, temperature=temperature, seed=seed,
                              max_partners=max_partners)
        _check(lib().mdf_chempot_enable(self._h, C.byref(p)))
        self._chem_bins = n_bins
        self._chem_temper = temperature if temperature > 0 else 300.0

    def chempot_disable(self):
        _check(lib().mdf_chempot_enable(self._h, None))

    def chempot(self, reset: bool = False):
        """dict(pave [n_bins], inave [n_bins] int64, seed, mu): the accumulated Boltzmann
        factors and orientation counts per bin, the generator state, and the excess chemical
        potential -RT ln(sum pave / sum inave) in kJ/mol (src/measure.f90:701)."""
        nb = self._chem_bins
        pave = np.zeros(nb)
        inave = np.zeros(nb, np.int64)
        seed = C.c_double()
        _check(lib().mdf_get_chempot(self._h, _dp(pave), inave.ctypes.data_as(C.POINTER(C.c_int64)),
                                     C.byref(seed), int(bool(reset))))
        n, sp = int(inave.sum()), float(pave.sum())
        mu = -np.log(sp / n) * abi.PH_R * self._chem_temper / 1000.0 if n > 0 and sp > 0 else float("nan")
        return dict(pave=pave, inave=inave, seed=seed.value, mu=mu)

    # -- reciprocal-force surrogate with multiple time steps ------------------------------
    def mts_enable(self, interval: int = 4, model: str = "network", train_frames: int = 200, train_iters: int = 2,
                   tol: float = 1e-3, lr: float = 1e-3, seed: int = 1):
        """Full k-sum every ``interval``-th step; in between the k-space force is the
        surrogate's (``model="network"``) or the last full sum's (``"held"``) plus the
        correction from the last full sum.  The first ``train_frames`` steps (and as many
        after each fallback) run the full sum and train the network."""
        p = abi.MtsParams(interval=interval, model=abi.MTS_MODELS[model], train_frames=train_frames,
                          train_iters=train_iters, tol=tol, lr=lr, seed=seed)
        _check(lib().mdf_mts_enable(self._h, C.byref(p)))

    def mts_disable(self):
        _check(lib().mdf_mts_enable(self._h, None))

    def mts_stats(self) -> dict:
        s = abi.MtsStats()
        _check(lib().mdf_mts_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in abi.MtsStats._fields_}

    def mts_weights(self) -> np.ndarray:
        w = np.zeros(abi.MTS_WEIGHTS, np.float32)
        _check(lib().mdf_mts_get_weights(self._h, w.ctypes.data_as(C.POINTER(C.c_float)), len(w)))
        return w

    def set_mts_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _check(lib().mdf_mts_set_weights(self._h, w.ctypes.data_as(C.POINTER(C.c_float)), len(w)))

    def mts_features(self):
        """The surrogate's inputs [n_move, 48] and vector features [n_move, 16, 3] at the
        positions of the last step."""
        feat = np.zeros((self.n_move, 48), np.float32)
        vec = np.zeros((self.n_move, 16, 3), np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        _check(lib().mdf_mts_features(self._h, fp(feat), fp(vec)))
        return feat, vec

    def stage_times(self):
        names = (C.c_char_p * 16)()
        ms = (C.c_float * 16)()
        n = C.c_int32()
        _check(lib().mdf_stage_times(self._h, names, ms, 16, C.byref(n)))
        return {names[k].decode(): float(ms[k]) for k in range(n.value)}

    def autotune(self, reps: int = 10) -> dict:
        """Time candidate launch shapes on the current state and keep the fastest; returns
        the choice (pair lanes per lead, pair loop form - strided: no cut-off compaction -,
        structure-factor and back-projection blocks per launch)."""
        ch = (C.c_int32 * 4)()
        _check(lib().mdf_autotune(self._h, int(reps), ch))
        return dict(pair_lanes=ch[0], pair_strided=bool(ch[1]), sf_blocks=ch[2], bp_blocks=ch[3])

    def set_launch_shapes(self, pair_lanes: int = 16, pair_strided: bool = False, sf_blocks: int = 4096,
                          bp_blocks: int = 1024):
        _check(lib().mdf_set_launch_shapes(self._h, int(pair_lanes), int(bool(pair_strided)), int(sf_blocks),
                                           int(bp_blocks)))

    def set_stage_timing(self, on, rebuild_every_step: bool = False):
        """Per-stage HIP events on (direct launches); rebuild_every_step also builds the
        pair list on every step, so that the "build" stage times full builds."""
        _check(lib().mdf_set_stage_timing(self._h, (2 if rebuild_every_step else 1) if on else 0))

    def reset_stage_times(self):
        _check(lib().mdf_reset_stage_times(self._h))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().mdf_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mts_eval(weights, feat, vec, target=None, feat_b=None, vec_b=None, target_b=None, lr: float = 0.0,
             adam_t: int = 0):
    """The surrogate network on given inputs, on the current GPU, with the scheme's own
    kernels: predicted forces [n, 3] of frame a (feat [n, 48], vec [n, 16, 3]); with targets
    [n, 3] of frame a and of a second frame b, also the gradient of the loss on their change
    (flat parameters, then the loss) and, for lr > 0, the parameters after one Adam step."""
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    w = np.ascontiguousarray(weights, np.float32)
    feat = np.ascontiguousarray(feat, np.float32)
    vec = np.ascontiguousarray(vec, np.float32)
    n = len(feat)
    f = np.zeros(3 * n, np.float32)
    grad = np.zeros(abi.MTS_WEIGHTS - 112 + 1, np.float32)
    params = np.zeros(abi.MTS_WEIGHTS - 112, np.float32)
    args = [None] * 5
    if target is not None:
        fb = np.ascontiguousarray(feat_b, np.float32)
        vb = np.ascontiguousarray(vec_b, np.float32)
        ta = np.ascontiguousarray(np.asarray(target, np.float64).T)
        tb = np.ascontiguousarray(np.asarray(target_b, np.float64).T)
        args = [ta, fb, vb, tb]
        _check(lib().mdf_mts_eval(fp(w), n, fp(feat), fp(vec), _dp(ta), fp(fb), fp(vb), _dp(tb), C.c_float(lr),
                                  int(adam_t), fp(f), fp(grad), fp(params)))
    else:
        _check(lib().mdf_mts_eval(fp(w), n, fp(feat), fp(vec), None, None, None, None, C.c_float(0.0), 0, fp(f),
                                  None, None))
    del args
    out = {"force": f.reshape(3, n).T.copy()}
    if target is not None:
        out["grad"] = grad[:-1].copy()
        out["loss"] = float(grad[-1])
        if lr > 0:
            out["params"] = params
    return out


def mts_eval_device(weights, n: int, feat, vec, f_out, target=None, feat_b=None, vec_b=None, target_b=None,
                    lr: float = 0.0, adam_t: int = 0, grad_out=None, params_out=None, stream=None):
    """mts_eval on device-resident arrays (torch tensors on the GPU, or device pointers as
    ints), allocating nothing: weights [8448] float32, feat [n, 48], vec [n, 16, 3] float32,
    targets [3, n] float64, outputs f_out [3, n], grad_out [8337], params_out [8336] float32.
    Enqueued on ``stream`` (default: torch's current stream) and not waited for."""
    ptr = lambda t: None if t is None else C.c_void_p(t if isinstance(t, int) else t.data_ptr())
    _check(lib().mdf_mts_eval_device(ptr(weights), int(n), ptr(feat), ptr(vec), ptr(target), ptr(feat_b),
                                     ptr(vec_b), ptr(target_b), C.c_float(lr), int(adam_t), ptr(f_out),
                                     ptr(grad_out), ptr(params_out), C.c_void_p(_producer_stream(stream))))


class ForceBatch:
    """Independent systems (``ForceEngine`` handles) stepped together: one launch per stage
    over all of them, one graph per step."""

    def __init__(self, engines):
        self.engines = list(engines)
        n = len(self.engines)
        arr = (C.c_void_p * n)(*[e._h.value for e in self.engines])
        self._h = C.c_void_p()
        _check(lib().mdf_batch_create(arr, n, C.byref(self._h)))

    def compute_device(self, ptrs, stream=None):
        """ptrs[i] = (x, y, z, cx, cy, cz) device pointers (ints) of system i; ordered after
        ``stream`` as ForceEngine.compute_device."""
        n = len(self.engines)
        cols = [(C.c_void_p * n)(*[int(p[c]) if p[c] else None for p in ptrs]) for c in range(6)]
        _check(lib().mdf_batch_compute_device(self._h, *cols, C.c_void_p(_producer_stream(stream))))

    def read_outputs(self):
        n = len(self.engines)
        outs = (abi.Outputs * n)()
        keep = []
        for i, e in enumerate(self.engines):
            o, arrs = abi.make_outputs(e.n_move, e.n_rig, e.n_pairs)
            outs[i] = o
            keep.append(arrs)
        rc = lib().mdf_batch_read_outputs(self._h, outs)
        if rc != abi.MDF_OK:
            err = MDForceError(rc, lib().mdf_last_error().decode())
            if rc == abi.MDF_ERR_LIST_OVERFLOW and all(outs[i].n_list_builds > 0 for i in range(n)):
                err.result = [abi.collect(outs[i], keep[i], e.n_move, e.n_rig) for i, e in enumerate(self.engines)]
            raise err
        return [abi.collect(outs[i], keep[i], e.n_move, e.n_rig) for i, e in enumerate(self.engines)]

    def autotune(self, reps: int = 5) -> dict:
        """Time whole batch steps on the last step's inputs (still valid on the device) for
        candidate launch shapes and keep the fastest: pair lanes per lead, and the batch's
        structure-factor and back-projection blocks per launch, shared out to the systems by
        their reciprocal work."""
        ch = (C.c_int32 * 4)()
        _check(lib().mdf_batch_autotune(self._h, int(reps), ch))
        return dict(pair_lanes=ch[0], sf_blocks=ch[2], bp_blocks=ch[3])

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().mdf_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
