"""wadges_diagnostics_noalloc (the diagnostics reduction with its per-element
partials in the library's fixed device array, for callers that cannot
allocate) against wadges_diagnostics with a caller-provided aux buffer: equal
totals bit for bit on two calls in a row, the second on a smaller mesh (so the
array holds the first call's partials beyond the second's K), through the C
ABI; sizes beyond the capacity are refused."""

import ctypes

import numpy as np
import pytest
import torch

from scratch_wadge_amd import _lib
from scratch_wadge_amd import problem as P

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _both(prob, Qex):
    from scratch_wadge_amd.device import DeviceSolver
    dev = DeviceSolver(prob)
    L = _lib.load()
    wq = torch.as_tensor(prob.wq, dtype=torch.float64).to(dev.device)
    qex = None if Qex is None else torch.as_tensor(Qex, dtype=torch.float64).to(dev.device)
    aux = torch.full((prob.K * 4,), np.nan, dtype=torch.float64, device=dev.device)
    tot_a = torch.full((4,), np.nan, dtype=torch.float64, device=dev.device)
    tot_b = torch.full((4,), np.nan, dtype=torch.float64, device=dev.device)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev.device).cuda_stream)
    args = (dev.plan.h, prob.K, _ptr(wq), _ptr(dev.c["Jq"]), _ptr(dev.s["Q"]), _ptr(qex))
    _lib.check(L.wadges_diagnostics(*args, _ptr(aux), _ptr(tot_a), st), "wadges_diagnostics")
    _lib.check(L.wadges_diagnostics_noalloc(*args, _ptr(tot_b), st), "wadges_diagnostics_noalloc")
    torch.cuda.synchronize(dev.device)
    return tot_a.cpu().numpy(), tot_b.cpu().numpy(), dev


def test_noalloc_diagnostics_equal_two_calls_shrinking():
    big = P.tgv3d_problem(n=6, N=3, curved=True)          # K = 1296
    small = P.tgv3d_problem(n=3, N=3, curved=True)        # K = 162
    for prob in (big, small):
        print(f"case noalloc diagnostics K={prob.K}", flush=True)
        Qex = prob.Q * (1.0 + 1e-3 * np.sin(np.arange(prob.Q.size)))
        a, b, _ = _both(prob, Qex)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), (a, b)


def test_noalloc_diagnostics_without_exact_state_and_2d():
    prob = P.vortex2d_problem(n=8, N=3)
    print(f"case noalloc diagnostics 2D K={prob.K}", flush=True)
    a, b, _ = _both(prob, None)
    assert a[2] == 0.0 and np.array_equal(a, b)


def test_noalloc_diagnostics_refuses_beyond_capacity():
    prob = P.tgv3d_problem(n=3, N=3, curved=True)
    _, _, dev = _both(prob, None)
    L = _lib.load()
    tot = torch.zeros(4, dtype=torch.float64, device=dev.device)
    big_K = 262144 + 1
    rc = L.wadges_diagnostics_noalloc(dev.plan.h, big_K, _ptr(tot), _ptr(tot), _ptr(tot), None,
                                      _ptr(tot), None)
    assert rc == -5
    assert L.wadges_error_string(rc).decode().startswith("size beyond")
