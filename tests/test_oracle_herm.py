"""CPU: the oracle's expm_herm (oracle/herm.py) pinned against what fixes the answer:
closed forms (exp(t sigma) = cosh(t) I + sinh(t) sigma for the Pauli matrices, exp of a
diagonal, exp(0) = I), scipy's expm (Pade scaling and squaring: an independent
algorithm) on random symmetric / Hermitian matrices, and the reference's reading rule
(only the uplo triangle, the diagonal's imaginary part ignored)."""
import numpy as np
import pytest
import scipy.linalg

import oracle.herm as oh

SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
SY = np.array([[0, -1j], [1j, 0]])
SZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)


@pytest.mark.parametrize("t", [0.0, 0.3, 1.0, 4.5, -2.0])
def test_pauli_closed_forms(t):
    for s in (SX, SY, SZ):
        ref = np.cosh(t) * np.eye(2) + np.sinh(t) * s
        for uplo in (oh.UPPER, oh.LOWER):
            np.testing.assert_allclose(oh.expm_herm(t * s, uplo), ref, rtol=1e-14, atol=1e-14)


def test_zero_and_diagonal():
    np.testing.assert_array_equal(oh.expm_herm(np.zeros((5, 5))), np.eye(5))
    d = np.array([-3.0, 0.5, 2.0, 0.0])
    np.testing.assert_allclose(oh.expm_herm(np.diag(d)), np.diag(np.exp(d)), rtol=1e-15)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [1, 2, 7, 40, 100])
def test_matches_scipy_expm(cplx, n):
    rng = np.random.default_rng(n + 10 * cplx)
    g = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    a = (g + g.conj().T) / 2
    ref = scipy.linalg.expm(a)
    for uplo in (oh.UPPER, oh.LOWER):
        e = oh.expm_herm(a, uplo)
        assert np.linalg.norm(e - ref) <= 1e-13 * np.linalg.norm(ref)


def test_reads_only_the_triangle():
    rng = np.random.default_rng(4)
    n = 9
    g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = (g + g.conj().T) / 2
    ref = scipy.linalg.expm(a)
    up = np.triu(a) + np.tril(np.full((n, n), 1e30), -1)
    up[np.diag_indices(n)] += 5j  # the diagonal's imaginary part is not read
    lo = np.tril(a) + np.triu(np.full((n, n), -1e30), 1)
    for b, uplo in ((up, oh.UPPER), (lo, oh.LOWER)):
        assert np.linalg.norm(oh.expm_herm(b, uplo) - ref) <= 1e-13 * np.linalg.norm(ref)
