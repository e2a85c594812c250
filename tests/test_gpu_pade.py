"""GPU: the reference's own route -- LU with partial pivoting (clm_getrf_* / clm_getrs_*),
the Pade approximant (clm_pade_*) and expm<T> by Pade + LU (clm_expm_pade_*) -- against
the oracle's restatement (oracle.c: orc_getrf_*, orc_getrs_*, orc_pade_*, orc_expm_*),
the reference's own tests' shape (test/expm/expm_tests.cpp: Pade at every degree, the
expm against an oracle, the singular and ill-scaled cases).

* getrf: the pivots are the oracle's (and scipy's) exactly; L and U agree to 1e-13 of
  ||A||; P L U reconstructs A; a zero column reports its index in info; block
  columns taller than the LDS holds (the in-memory leaf) at n = 8300 complex128.
* getrs: the solve against numpy.linalg.solve.
* pade: every degree against orc_pade (1e-13 FP64; FP32 within its precision).
* expm_pade: the plan equals the oracle's (the same 1-norm, bit for bit for real input),
  the result within 1e-13 of the oracle's at n = 2048 on the benchmark matrices.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

import calaman_g_amd as cg
import oracle
from calaman_g_amd import _lib

pytestmark = pytest.mark.gpu

SINGLE = {torch.float32: True, torch.float64: False, torch.complex64: True, torch.complex128: False}
CPLX = {torch.float32: False, torch.float64: False, torch.complex64: True, torch.complex128: True}


def cm(a, dtype, dev, ld=None):
    n_rows, n_cols = a.shape
    ld = n_rows if ld is None else ld
    buf = np.zeros((n_cols, ld), dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    buf[:, :n_rows] = a.T
    return torch.from_numpy(buf).to(dtype).to(dev)


def from_cm(t, n_rows):
    return t.cpu().numpy()[:, :n_rows].T.astype(np.complex128 if t.is_complex() else np.float64)


def gaussian(n, seed, cplx, scale=1.0):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    if cplx:
        a = a + 1j * rng.standard_normal((n, n))
    return a * scale


def gpu_getrf(a, dtype, dev, ld=None):
    n = a.shape[0]
    ld = n if ld is None else ld
    da = cm(a, dtype, dev, ld)
    ipiv = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
    info = torch.full((1,), -5, dtype=torch.int32, device=dev)
    st = cg.getrf_cm(n, da, ld, ipiv, info)
    assert st == 0
    torch.cuda.synchronize()
    return da, ipiv.cpu().numpy()[:n].astype(np.int64), int(info.item())


def unpack(lu):
    n = lu.shape[0]
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    return L, U


def perm_from_ipiv(ipiv, n):
    p = np.arange(n)
    for k, q in enumerate(ipiv):
        q -= 1
        p[k], p[q] = p[q], p[k]
    return p


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n", [1, 2, 7, 33, 64, 100, 257, 1000, 2100])
def test_getrf_matches_oracle(dev, dtype, n):
    a = gaussian(n, n, CPLX[dtype])
    da, ipiv, info = gpu_getrf(a, dtype, dev, ld=n + 3)
    lu = from_cm(da, n)
    olu, oipiv, oinfo = oracle.getrf(a)
    assert info == oinfo == 0
    np.testing.assert_array_equal(ipiv, oipiv)
    _, sp_piv = scipy.linalg.lu_factor(a)
    np.testing.assert_array_equal(ipiv, sp_piv + 1)
    # two correct eliminations differ by their rounding: entry (i, j) sums about n
    # products, so the bound is componentwise, n eps (|L| |U|)_ij (Higham, ASNA 9.3)
    # (below the diagonal l_ij carries that bound divided by |u_jj|)
    oL, oU = unpack(olu)
    bound = 2 * n * np.finfo(np.float64).eps * (np.abs(oL) @ np.abs(oU))
    lower = np.tril(np.ones((n, n), dtype=bool), -1)
    bound[lower] = (bound / np.abs(np.diag(oU))[None, :])[lower]
    assert np.all(np.abs(lu - olu) <= bound + 1e-300)
    scale = np.abs(a).max()
    L, U = unpack(lu)
    p = perm_from_ipiv(ipiv, n)
    assert np.abs(L @ U - a[p]).max() <= 1e-13 * n * scale
    # the padding rows of the leading dimension are untouched
    assert np.all(da.cpu().numpy()[:, n:] == 0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64])
def test_getrf_single(dev, dtype):
    n = 300
    a = gaussian(n, 5, CPLX[dtype])
    da, ipiv, info = gpu_getrf(a, dtype, dev)
    assert info == 0
    L, U = unpack(from_cm(da, n))
    p = perm_from_ipiv(ipiv, n)
    assert np.abs(L @ U - a[p]).max() <= 2e-5 * n ** 0.5 * np.abs(a).max()


def test_getrf_zero_pivot_reports_column(dev):
    n = 50
    a = gaussian(n, 3, False)
    a[:, 17] = 0.0  # column 18 (1-based) has no pivot
    _, ipiv, info = gpu_getrf(a, torch.float64, dev)
    _, oipiv, oinfo = oracle.getrf(a)
    assert info == oinfo == 18
    np.testing.assert_array_equal(ipiv, oipiv)
    _, _, info0 = gpu_getrf(np.zeros((4, 4)), torch.float64, dev)
    assert info0 == 1


def test_getrf_tall_block_columns_in_memory(dev):
    """n = 8300 complex128: the first block columns hold more rows than the LDS leaf
    (8192 complex128 entries a column), so they run the in-memory leaf"""
    n = 8300
    torch.manual_seed(0)
    a = torch.randn((n, n), dtype=torch.complex128, device=dev)
    acm = a.t().contiguous()
    ipiv = torch.zeros(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    assert cg.getrf_cm(n, acm, n, ipiv, info) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    lu = acm.t()
    L = torch.tril(lu, -1) + torch.eye(n, dtype=lu.dtype, device=dev)
    U = torch.triu(lu)
    p = torch.from_numpy(perm_from_ipiv(ipiv.cpu().numpy().astype(np.int64), n)).to(dev)
    res = (L @ U - a[p]).abs().max().item()
    assert res <= 1e-12 * a.abs().max().item() * n ** 0.5, res
    # partial pivoting by i?amax's |re| + |im|: |l| <= |re| + |im| of the entry over
    # that of the pivot times sqrt(2), so every multiplier has modulus <= sqrt(2)
    assert L.abs().max().item() <= 2 ** 0.5 * (1 + 1e-12)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128, torch.float32, torch.complex64])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (40, 3), (300, 300), (1000, 17)])
def test_getrs_solves(dev, dtype, n, nrhs):
    cplx = CPLX[dtype]
    a = gaussian(n, n + 1, cplx) + n ** 0.5 * np.eye(n)
    b = gaussian(max(n, nrhs), n + 2, cplx)[:n, :nrhs]
    da, ipiv, info = gpu_getrf(a, dtype, dev)
    db = cm(b, dtype, dev, ld=n + 1)
    dinfo = torch.full((1,), 7, dtype=torch.int32, device=dev)
    assert cg.getrs_cm(n, nrhs, da, n, torch.from_numpy(ipiv.astype(np.int32)).to(dev), db, n + 1,
                       dinfo) == 0
    torch.cuda.synchronize()
    assert int(dinfo.item()) == 0
    x = from_cm(db, n)
    want = np.linalg.solve(a, b)
    tol = 2e-4 if SINGLE[dtype] else 1e-11
    assert np.linalg.norm(x - want) <= tol * np.linalg.norm(want)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("m", [3, 5, 7, 9, 13])
@pytest.mark.parametrize("n", [5, 64, 300])
def test_pade_matches_oracle(dev, dtype, m, n):
    a = gaussian(n, m * n, CPLX[dtype])
    a *= 0.9 * oracle.pade_theta(m) / np.abs(a).sum(0).max()
    da = cm(a, dtype, dev, ld=n + 1)
    dr = torch.full((n, n + 2), 3.0, dtype=dtype, device=dev)
    work = torch.empty(cg.pade_buffer_size(m, n, dtype), dtype=torch.uint8, device=dev)
    info = torch.full((2,), -1, dtype=torch.int32, device=dev)
    assert cg.pade_cm(m, n, da, n + 1, dr, n + 2, work, info) == 0
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0, 0]
    got = from_cm(dr, n)
    want = oracle.pade(m, a)
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
    assert np.all(dr.cpu().numpy()[:, n:] == 3.0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64])
@pytest.mark.parametrize("m", [3, 7, 13])
def test_pade_single(dev, dtype, m):
    n = 64
    a = gaussian(n, m, CPLX[dtype])
    a *= 0.9 * oracle.pade_theta(m, single=True) / np.abs(a).sum(0).max()
    da = cm(a, dtype, dev)
    dr = torch.empty_like(da)
    work = torch.empty(cg.pade_buffer_size(m, n, dtype), dtype=torch.uint8, device=dev)
    assert cg.pade_cm(m, n, da, n, dr, n, work) == 0
    want = oracle.pade(m, a.astype(np.complex64 if CPLX[dtype] else np.float32))
    assert np.linalg.norm(from_cm(dr, n) - want) <= 1e-5 * np.linalg.norm(want)


def _expm_pade(a, dtype, dev):
    n = a.shape[0]
    da = cm(a, dtype, dev, ld=n + 1)
    de = torch.full((n, n + 1), -2.0, dtype=dtype, device=dev)
    work = torch.empty(cg.expm_pade_buffer_size(n, dtype), dtype=torch.uint8, device=dev)
    info = torch.full((2,), -1, dtype=torch.int32, device=dev)
    st, plan = cg.expm_pade_cm(n, da, n + 1, de, n + 1, work, info)
    assert st == 0
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0, 0]
    assert np.all(de.cpu().numpy()[:, n:] == -2.0)
    return from_cm(de, n), plan


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n,target", [(1, 0.3), (5, 0.01), (64, 2.0), (200, 30.0), (256, 900.0)])
def test_expm_pade_matches_oracle(dev, dtype, n, target):
    a = gaussian(n, n + int(target), CPLX[dtype])
    a *= target / np.abs(a).sum(0).max()
    got, plan = _expm_pade(a, dtype, dev)
    want, oplan = oracle.expm(a, return_plan=True)
    assert plan["method"] == _lib.METHOD_PADE
    assert (plan["m"], plan["s"], plan["num_gemms"]) == tuple(oplan)
    scale = np.abs(want).max()
    assert np.linalg.norm((got - want) / scale) <= 1e-13 * np.linalg.norm(want / scale)


@pytest.mark.parametrize("cplx", [False, True])
def test_expm_pade_benchmark_matrices(dev, cplx):
    """the n = 2048 benchmark inputs at ||A||_1 = 0.5, 8, 100 (seeds 0-2)"""
    import bench
    dtype = torch.complex128 if cplx else torch.float64
    for a in bench.make_inputs(2048, cplx):
        got, plan = _expm_pade(a, dtype, dev)
        want, oplan = oracle.expm(a, return_plan=True)
        assert (plan["m"], plan["s"]) == tuple(oplan)[:2]
        assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64])
def test_expm_pade_single(dev, dtype):
    n = 128
    a = gaussian(n, 9, CPLX[dtype])
    a *= 6.0 / np.abs(a).sum(0).max()
    a = a.astype(np.complex64 if CPLX[dtype] else np.float32)
    got, plan = _expm_pade(a, dtype, dev)
    want_plan = cg.pade_plan(plan["norm1"], single=True)
    assert (plan["m"], plan["s"], plan["num_gemms"]) == (want_plan["m"], want_plan["s"],
                                                         want_plan["num_gemms"])
    want = oracle.expm(a.astype(np.complex128 if CPLX[dtype] else np.float64))
    assert np.linalg.norm(got - want) <= 2e-5 * np.linalg.norm(want)


def test_pade_rejects_bad_arguments(dev):
    x = torch.zeros((8, 8), dtype=torch.float64, device=dev)
    work = torch.empty(cg.pade_buffer_size(13, 8), dtype=torch.uint8, device=dev)
    assert cg.pade_cm(4, 8, x, 8, x.clone(), 8, work) == _lib.INVALID_VALUE
    assert cg.pade_cm(13, 8, x, 7, x.clone(), 8, work) == _lib.INVALID_VALUE
    small = torch.empty(64, dtype=torch.uint8, device=dev)
    assert cg.pade_cm(13, 8, x, 8, x.clone(), 8, small) == _lib.ALLOC_FAILED
    nan = torch.full((8, 8), float("nan"), dtype=torch.float64, device=dev)
    w2 = torch.empty(cg.expm_pade_buffer_size(8), dtype=torch.uint8, device=dev)
    assert cg.expm_pade_cm(8, nan, 8, x.clone(), 8, w2)[0] == _lib.INVALID_VALUE
