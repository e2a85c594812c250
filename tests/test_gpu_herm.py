"""GPU: exp(A) for a symmetric / Hermitian A by its eigendecomposition
(clm_expm_herm_*: the reference's expm_herm, src/expm/expm_herm.cppm:145-199) against
the oracle's restatement (oracle/herm.py: the uplo triangle, LAPACK's ?syevd/?heevd
through numpy, U diag(exp(w)) U^H), and the reference's own checks
(test/expm/expm_tests.cpp:780-920): against the general expm on the same matrix, exp(0)
= I, the argument checks.

* FP64 / complex128 within 1e-12 (relative Frobenius) of the oracle for n = 1 .. 512,
  both triangles, leading dimensions larger than n, the other triangle and the
  diagonal's imaginary part filled with NaN (never read);
* diagonal input: exp of the diagonal to 2 ulp; zero input: the identity exactly;
  degenerate spectra (c I + a rank-2 term);
* against clm_expm_* (the reference's own oracle for this route);
* FP32 / complex64 (the sweeps run in double) within 1e-6.
"""
import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle.herm as oh

pytestmark = pytest.mark.gpu

CPLX = {torch.float32: False, torch.float64: False, torch.complex64: True, torch.complex128: True}
NPC = {torch.float32: np.float32, torch.float64: np.float64, torch.complex64: np.complex64,
       torch.complex128: np.complex128}


def hermitian(n, seed, cplx, scale=1.0):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n))
    if cplx:
        g = g + 1j * rng.standard_normal((n, n))
    return (g + g.conj().T) / 2 * scale


def poison(a, uplo):
    """the triangle expm_herm must not read set to NaN (and NaN in the imaginary part
    of a complex diagonal)"""
    b = np.array(a, dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    n = b.shape[0]
    mask = np.tril(np.ones((n, n), bool), -1) if uplo == "U" else np.triu(np.ones((n, n), bool), 1)
    b[mask] = np.nan
    if np.iscomplexobj(b):
        b.imag[np.diag_indices(n)] = np.nan
    return b


def run(a, dtype, dev, uplo="U", pad=3):
    n = a.shape[0]
    ld = n + pad
    buf = np.zeros((n, ld), dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    buf[:, :n] = a.T
    da = torch.from_numpy(buf).to(dtype).to(dev)
    de = torch.full((n, ld), 7.0, dtype=dtype, device=dev)
    work = torch.empty(cg.expm_herm_buffer_size(n, dtype), dtype=torch.uint8, device=dev)
    info = torch.full((1,), -3, dtype=torch.int32, device=dev)
    print(f"expm_herm n={n} dtype={dtype} uplo={uplo}", flush=True)
    st = cg.expm_herm_cm(uplo, n, da, ld, de, ld, work, info)
    assert st == 0
    torch.cuda.synchronize()
    out = de.cpu().numpy()
    assert np.all(out[:, n:] == 7.0)  # padding rows of each column untouched
    return out[:, :n].T.astype(np.complex128 if CPLX[dtype] else np.float64), int(info.item())


def rel(e, o):
    return np.linalg.norm(e - o) / np.linalg.norm(o)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("n", [1, 2, 5, 6, 31, 32, 33, 64, 65, 100, 128, 200, 257, 512])
def test_expm_herm_matches_oracle(dev, dtype, uplo, n):
    a = hermitian(n, 1000 + n, CPLX[dtype])
    e, info = run(poison(a, uplo), dtype, dev, uplo)
    o = oh.expm_herm(a, oh.UPPER if uplo == "U" else oh.LOWER)
    assert info == 0
    err = rel(e, o)
    print(f"  rel err {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128, torch.float32, torch.complex64])
def test_expm_herm_zero_and_diagonal(dev, dtype):
    n = 70
    e, info = run(np.zeros((n, n)), dtype, dev)
    assert info == 0 and np.array_equal(e, np.eye(n))
    real = np.float32 if dtype in (torch.float32, torch.complex64) else np.float64
    d = np.linspace(-3.0, 2.0, n).astype(real).astype(np.float64)
    e, info = run(np.diag(d), dtype, dev)
    assert info == 0
    assert np.all(e[~np.eye(n, dtype=bool)] == 0)
    # the exponentials are taken in double and rounded to the type: within 2 ulp of
    # numpy's in the type's own precision
    eps = np.finfo(real).eps
    np.testing.assert_allclose(np.diag(e).real, np.exp(d.astype(real)), rtol=2 * eps, atol=0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_expm_herm_degenerate_spectrum(dev, dtype):
    n = 150
    rng = np.random.default_rng(8)
    u = rng.standard_normal((n, 2))
    if CPLX[dtype]:
        u = u + 1j * rng.standard_normal((n, 2))
    a = 0.75 * np.eye(n) + u @ np.diag([0.3, -0.2]) @ u.conj().T / n
    e, info = run(a, dtype, dev)
    assert info == 0 and rel(e, oh.expm_herm(a)) <= 1e-12


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n,target", [(6, 1.5), (5, 2.5), (128, 8.0), (300, 20.0)])
def test_expm_herm_matches_general_expm(dev, dtype, n, target):
    a = hermitian(n, 7 * n, CPLX[dtype])
    a *= target / np.abs(a).sum(0).max()
    e, _ = run(a, dtype, dev)
    g = cg.expm(torch.from_numpy(a).to(dtype).to(dev)).cpu().numpy()
    assert rel(e, g) <= 1e-12


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64])
@pytest.mark.parametrize("n", [4, 5, 128, 300])
def test_expm_herm_single(dev, dtype, n):
    a = hermitian(n, 3 * n, CPLX[dtype]).astype(NPC[dtype]).astype(
        np.complex128 if CPLX[dtype] else np.float64)
    e, info = run(a, dtype, dev, "L")
    # the sweeps run in double: the result is the exact one rounded to single
    assert info == 0 and rel(e, oh.expm_herm(a, oh.LOWER)) <= 1e-6


def test_expm_herm_convenience(dev):
    a = hermitian(90, 4, True)
    t = torch.from_numpy(a).to(dev)
    e, info = cg.expm_herm(t, "U", return_info=True)
    assert info == 0 and rel(e.cpu().numpy(), oh.expm_herm(a)) <= 1e-12


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_expm_herm_n1000(dev, dtype):
    a = hermitian(1000, 99, CPLX[dtype], scale=0.3)
    e, info = run(a, dtype, dev, "L")
    err = rel(e, oh.expm_herm(a, oh.LOWER))
    print(f"  rel err {err:.3e}")
    assert info == 0 and err <= 1e-12
