"""mdf_mts_eval_device, the entry point that allocates nothing, against mdf_mts_eval.

Both run the same kernels on the same inputs; the device variant keeps its scratch in one
static device array shared by every call.  Two calls in a row, the second with fewer sites
and new inputs, each equal the allocating entry point's result: the forward pass bit for bit
(same bfloat16 weight copies, same kernels), the gradient and the Adam step within FP32
rounding of the loss scale, which the device sums in another order."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _weights(seed):
    from md_sim_amd import abi
    rng = np.random.default_rng(seed)
    n_par = abi.MTS_WEIGHTS - 112
    w = np.zeros(abi.MTS_WEIGHTS, np.float32)
    w[:n_par] = rng.normal(0, 0.2, n_par)
    w[n_par:n_par + 48] = rng.normal(0, 0.3, 48)         # input means
    w[n_par + 48:n_par + 96] = rng.uniform(0.5, 2.0, 48)  # input deviations
    w[n_par + 96:] = rng.uniform(0.5, 2.0, 16)           # vector scales
    return w


def _frames(rng, n):
    out = []
    for _ in range(2):
        feat = rng.normal(0, 1.5, (n, 48)).astype(np.float32)
        feat[:, 39:] = 0
        vec = rng.normal(0, 1.0, (n, 16, 3)).astype(np.float32)
        out.append((feat, vec, rng.normal(0, 1.0, (n, 3))))
    return out


def _dev(a, dtype):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda:0")


def test_no_alloc_eval_matches_allocating_entry_point():
    import torch
    from md_sim_amd import abi
    from md_sim_amd.engine import MDForceError, mts_eval, mts_eval_device
    rng = np.random.default_rng(5)
    n_par = abi.MTS_WEIGHTS - 112
    free = []
    for call, (n, seed) in enumerate([(300, 1), (137, 2), (137, 3)]):
        _say("case no-alloc eval: call %d, n = %d" % (call, n))
        w = _weights(seed)
        (fa, va, Ta), (fb, vb, Tb) = _frames(rng, n)
        ref = mts_eval(w, fa, va, Ta, fb, vb, Tb, lr=1e-3, adam_t=call)
        ref0 = mts_eval(w, fa, va)
        dw = _dev(w, torch.float32)
        dfa, dva, dfb, dvb = [_dev(a, torch.float32) for a in (fa, va, fb, vb)]
        dTa, dTb = _dev(np.asarray(Ta).T, torch.float64), _dev(np.asarray(Tb).T, torch.float64)
        f = torch.full((3, n), np.nan, dtype=torch.float32, device="cuda:0")
        g = torch.full((n_par + 1,), np.nan, dtype=torch.float32, device="cuda:0")
        p = torch.full((n_par,), np.nan, dtype=torch.float32, device="cuda:0")
        mts_eval_device(dw, n, dfa, dva, f, dTa, dfb, dvb, dTb, lr=1e-3, adam_t=call, grad_out=g, params_out=p)
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info()[0])
        force = f.cpu().numpy().T
        assert np.array_equal(force, ref["force"]), call
        gg, rg = g.cpu().numpy(), np.concatenate([ref["grad"], [ref["loss"]]])
        assert np.abs(gg - rg).max() <= 1e-6 * np.abs(rg).max(), call
        pp = p.cpu().numpy()
        assert np.abs(pp - ref["params"]).max() <= 1e-6 * max(1.0, np.abs(ref["params"]).max()), call
        # forward only: params_out and grad_out stay untouched
        f0 = torch.zeros((3, n), dtype=torch.float32, device="cuda:0")
        p0 = torch.full((n_par,), 7.0, dtype=torch.float32, device="cuda:0")
        mts_eval_device(dw, n, dfa, dva, f0, params_out=p0)
        torch.cuda.synchronize()
        assert np.array_equal(f0.cpu().numpy().T, ref0["force"]), call
        assert bool((p0 == 7.0).all()), call
    # the calls took no device memory of their own
    assert free[1] == free[2], free
    with pytest.raises(MDForceError) as ei:
        mts_eval_device(1, (1 << 24) + 1, 1, 1, 1)   # refused before any device access
    assert ei.value.code == 1
