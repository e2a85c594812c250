"""GPU: clm_expm_noalloc_d -- clm_expm_d without a workspace argument, its
workspace a device array inside the library -- against clm_expm_d with a
caller workspace on the same input: the same kernels, so the same bits. Two
calls in a row, the second with a smaller input (the array keeps the last
call's contents), and orders past CLM_EXPM_NOALLOC_MAX_N rejected before any
device work."""
import ctypes

import numpy as np
import pytest
import torch

import calaman_g_amd as cg
from calaman_g_amd import _lib
from calaman_g_amd.expm import ExpmPlan, expm_cm

pytestmark = pytest.mark.gpu

L = _lib.lib()


def _noalloc(n, da, lda, de, lde, info=None):
    plan = ExpmPlan()
    st = L.clm_expm_noalloc_d(None, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
                              int(n), ctypes.c_void_p(da.data_ptr()), int(lda),
                              ctypes.c_void_p(de.data_ptr()), int(lde),
                              None if info is None else ctypes.c_void_p(info.data_ptr()),
                              ctypes.byref(plan))
    return int(st), plan.as_dict()


def _case(dev, n, norm, seed, lda, lde):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    a *= norm / np.abs(a).sum(0).max()
    abuf = np.zeros((n, lda))
    abuf[:, :n] = a.T
    da = torch.from_numpy(abuf).to(dev)
    ref = torch.full((n, lde), -7.0, dtype=torch.float64, device=dev)
    work = torch.empty(cg.expm_buffer_size(n), dtype=torch.uint8, device=dev)
    st, want_plan = expm_cm(n, da, lda, ref, lde, work)
    assert st == 0
    got = torch.full((n, lde), -7.0, dtype=torch.float64, device=dev)
    info = torch.full((2,), -9, dtype=torch.int32, device=dev)
    st, plan = _noalloc(n, da, lda, got, lde, info)
    torch.cuda.synchronize()
    assert st == 0
    assert info.cpu().tolist() == [0, 0]
    assert plan == want_plan
    # bit for bit, padding included (only the n x n block is written)
    assert torch.equal(got.view(torch.int64), ref.view(torch.int64))
    return plan


def test_noalloc_matches_workspace_call_two_calls_shrinking(dev):
    print("noalloc: n=512 |A|_1=100, then n=200 |A|_1=8, then n=33", flush=True)
    p1 = _case(dev, 512, 100.0, seed=1, lda=512, lde=520)  # the maximum order
    assert p1["s"] >= 1
    p2 = _case(dev, 200, 8.0, seed=2, lda=203, lde=200)  # smaller: stale contents beyond it
    assert p2["m"] >= 8
    _case(dev, 33, 0.5, seed=3, lda=33, lde=33)


def test_noalloc_rejects_orders_past_the_maximum(dev):
    nmax = 512  # CLM_EXPM_NOALLOC_MAX_N
    da = torch.zeros((4, 4), dtype=torch.float64, device=dev)
    st, _ = _noalloc(nmax + 1, da, nmax + 1, da, nmax + 1)
    assert st == _lib.INVALID_VALUE
    st, _ = _noalloc(0, da, 4, da, 4)
    assert st == _lib.INVALID_VALUE
    # a non-finite norm is the ordinary INVALID_VALUE of clm_expm_d
    bad = torch.full((4, 4), float("nan"), dtype=torch.float64, device=dev)
    st, _ = _noalloc(4, bad, 4, torch.empty_like(bad), 4)
    assert st == _lib.INVALID_VALUE
