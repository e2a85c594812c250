"""GPU: the MFMA matrix product (clm_gemm_*), the engine every expm product runs on.

Checked against numpy's float64/complex128 matmul (an independent BLAS) with a
tolerance of 8*eps*k*max|A||B| elementwise -- the standard bound for a k-term
dot product -- on square, ragged and padded shapes, and the epilogue terms
(alpha, beta, in-place accumulate).
"""
import numpy as np
import pytest
import torch

import calaman_g_amd as cg
from calaman_g_amd.expm import Context, gemm_cm

pytestmark = pytest.mark.gpu


def _rand(rng, shape, cplx):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return a


def _cm_tensor(a, ld, dev):
    """device buffer holding a column-major with leading dimension ld (padding = sentinel)"""
    m, n = a.shape
    dt = torch.complex128 if np.iscomplexobj(a) else torch.float64
    buf = np.full((max(n, 1), ld), -777.0, dtype=a.dtype)
    buf[:n, :m] = a.T
    return torch.from_numpy(buf).to(dev)


def _from_cm(t, m, n):
    return t.cpu().numpy()[:n, :m].T


def _tol(a, b):
    k = a.shape[1]
    return 8 * np.finfo(float).eps * max(k, 1) * (np.abs(a).max() * np.abs(b).max() + 1e-300)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 5, 5), (16, 16, 16), (37, 53, 29), (64, 64, 64),
                                   (128, 128, 128), (130, 127, 131), (256, 256, 256),
                                   (1024, 1024, 1024), (1040, 1030, 1027)])
def test_gemm_matches_numpy(dev, cplx, m, n, k):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    a, b = _rand(rng, (m, k), cplx), _rand(rng, (k, n), cplx)
    ctx = Context()
    da, db = _cm_tensor(a, m, dev), _cm_tensor(b, k, dev)
    dc = _cm_tensor(np.zeros((m, n), dtype=a.dtype), m, dev)
    assert gemm_cm(ctx, m, n, k, 1.0, da, m, db, k, 0.0, dc, m) == 0
    torch.cuda.synchronize()
    got = _from_cm(dc, m, n)
    np.testing.assert_allclose(got, a @ b, rtol=0, atol=_tol(a, b))


@pytest.mark.parametrize("cplx", [False, True])
def test_gemm_2048(dev, cplx):
    n = 2048
    rng = np.random.default_rng(2048)
    a, b = _rand(rng, (n, n), cplx), _rand(rng, (n, n), cplx)
    got = cg.gemm(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)).cpu().numpy()
    np.testing.assert_allclose(got, a @ b, rtol=0, atol=_tol(a, b))


@pytest.mark.parametrize("cplx", [False, True])
def test_gemm_alpha_beta_and_padding(dev, cplx):
    m, n, k = 70, 45, 33
    lda, ldb, ldc = m + 3, k + 5, m + 7
    rng = np.random.default_rng(5)
    a, b, c = _rand(rng, (m, k), cplx), _rand(rng, (k, n), cplx), _rand(rng, (m, n), cplx)
    alpha, beta = (1.5 - 0.25j, -0.5 + 2j) if cplx else (1.5, -0.5)
    ctx = Context()
    da, db, dc = _cm_tensor(a, lda, dev), _cm_tensor(b, ldb, dev), _cm_tensor(c, ldc, dev)
    assert gemm_cm(ctx, m, n, k, alpha, da, lda, db, ldb, beta, dc, ldc) == 0
    torch.cuda.synchronize()
    full = dc.cpu().numpy()
    got = full[:n, :m].T
    ref = alpha * (a @ b) + beta * c
    np.testing.assert_allclose(got, ref, rtol=0, atol=4 * _tol(a, b) + 1e-14 * np.abs(c).max())
    # padding rows between m and ldc untouched
    assert np.all(full[:n, m:ldc] == -777.0)


def test_gemm_rejects_bad_leading_dimension(dev):
    ctx = Context()
    x = torch.zeros(16, dtype=torch.float64, device=dev)
    assert gemm_cm(ctx, 4, 4, 4, 1.0, x, 3, x, 4, 0.0, x, 4) == cg._lib.INVALID_VALUE


@pytest.mark.parametrize("m,n,k,ldc_pad", [(2048, 2048, 512, 2), (2048, 2048, 512, 1),
                                           (2050, 2049, 300, 3)])
def test_gemm_large_tiles_alpha_beta(dev, m, n, k, ldc_pad):
    """the 128 x 128 tiles' LDS-staged epilogue: 16-byte pieces (even ldc), element
    stores (odd ldc) and ragged edges, with alpha, beta and untouched padding"""
    rng = np.random.default_rng(m + ldc_pad)
    a, b, c = _rand(rng, (m, k), False), _rand(rng, (k, n), False), _rand(rng, (m, n), False)
    lda, ldb, ldc = m + 2, k + 2, m + ldc_pad
    ctx = Context()
    da, db, dc = _cm_tensor(a, lda, dev), _cm_tensor(b, ldb, dev), _cm_tensor(c, ldc, dev)
    assert gemm_cm(ctx, m, n, k, 1.5, da, lda, db, ldb, -0.5, dc, ldc) == 0
    torch.cuda.synchronize()
    full = dc.cpu().numpy()
    ref = 1.5 * (a @ b) - 0.5 * c
    np.testing.assert_allclose(full[:n, :m].T, ref, rtol=0,
                               atol=4 * _tol(a, b) + 1e-14 * np.abs(c).max())
    assert np.all(full[:n, m:ldc] == -777.0)
