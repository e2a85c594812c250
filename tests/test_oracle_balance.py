"""CPU: the oracle's balancing (oracle/balance.py: gebal, apply_balance_similarity) and
1-norm estimator (lacn2), pinned where the mathematics or LAPACK fixes the answer.

gebal follows LAPACK's xGEBAL through 3.4 (the reference's src/lapack/gebal/README.md
"Differences"); scipy ships LAPACK 3.12, whose scaling measures diagonal-inclusive
2-norms and whose row search continues downwards after an exchange instead of
restarting. So the pins are:

* the reference's own spec tests (test/gebal/gebal_tests.cpp): the analytic 2x2
  [[1, 2^10], [2^-10, 1]] -> [[1, 1], [1, 1]] with D = diag(2^10, 1), idempotence, a
  diagonal matrix isolating everything, n = 1, and the exact reproduction of A from
  the reported permutation and scaling;
* scipy's dgebal on scaling where the conventions coincide: matrices with one
  off-diagonal entry in every row and column and a zero diagonal (weighted
  derangements), where the off-diagonal 1-norm equals the diagonal-inclusive
  2-norm -- scale factors and B bit for bit;
* scipy's dgebal on the permutation where the conventions coincide: the same ilo and
  ihi and the same diagonal (the spectrum read off), with the structure LAPACK
  guarantees (zero below the diagonal outside the window, no further row or column
  isolating inside it);
* stationarity: no power of two improves c f + r / f by 5% at any index of the window.

lacn2 is pinned against LAPACK's dgecon, which estimates ||A^-1||_1 by dlacn2 over
the LU solves: the same estimate to rounding and the same step count; and exactly
||A||_1 (the sequence 1, 2, 1, 1) on non-negative matrices.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as la

import oracle.balance as ob


def derangement_matrix(rng, n, cplx=False):
    while True:
        p = rng.permutation(n)
        if np.all(p != np.arange(n)):
            break
    a = np.zeros((n, n), dtype=np.complex128 if cplx else np.float64)
    w = np.exp(rng.uniform(-30, 30, n))
    if cplx:  # purely imaginary: cabs1 and the modulus agree
        w = 1j * w
    a[np.arange(n), p] = w
    return a


def sparse_matrix(rng, n):
    return rng.standard_normal((n, n)) * (rng.random((n, n)) < rng.uniform(0.05, 0.6))


def test_analytic_2x2():
    a = np.array([[1.0, 2.0 ** 10], [2.0 ** -10, 1.0]])
    b, ilo, ihi, d = ob.gebal(a, ob.SCALE)
    assert (ilo, ihi) == (1, 2)
    np.testing.assert_array_equal(d, [2.0 ** 10, 1.0])
    np.testing.assert_array_equal(b, np.ones((2, 2)))
    # the same through BOTH: nothing to isolate
    b2, ilo2, ihi2, d2 = ob.gebal(a, ob.BOTH)
    assert (ilo2, ihi2) == (1, 2) and np.array_equal(b2, b) and np.array_equal(d2, d)


def test_structural_cases():
    b, ilo, ihi, d = ob.gebal(np.diag([3.0, -1.0, 2.0]), ob.BOTH)
    assert (ilo, ihi) == (1, 1)
    b, ilo, ihi, d = ob.gebal(np.array([[5.0]]), ob.BOTH)
    assert (ilo, ihi) == (1, 1) and d[0] == 1.0 and b[0, 0] == 5.0
    b, ilo, ihi, d = ob.gebal(np.zeros((0, 0)), ob.BOTH)
    assert (ilo, ihi) == (1, 0)
    a = np.arange(9.0).reshape(3, 3) + 1
    b, ilo, ihi, d = ob.gebal(a, ob.NONE)
    assert (ilo, ihi) == (1, 3) and np.array_equal(b, a) and np.all(d == 1)


@pytest.mark.parametrize("seed", range(6))
def test_scaling_matches_lapack_on_derangements(seed):
    rng = np.random.default_rng(seed)
    for _ in range(50):
        n = int(rng.integers(2, 12))
        cplx = bool(seed % 2)
        a = derangement_matrix(rng, n, cplx)
        b, ilo, ihi, d = ob.gebal(a, ob.SCALE)
        f = la.zgebal if cplx else la.dgebal
        sb, slo, shi, sd, info = f(a, permute=0, scale=1)
        assert info == 0 and (ilo, ihi) == (slo + 1, shi + 1)
        np.testing.assert_array_equal(d, sd)
        np.testing.assert_array_equal(b, sb)


@pytest.mark.parametrize("seed", range(4))
def test_permutation_matches_lapack_spectrum_and_structure(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(80):
        n = int(rng.integers(1, 14))
        a = sparse_matrix(rng, n)
        b, ilo, ihi, d = ob.gebal(a, ob.PERMUTE)
        sb, slo, shi, sd, info = la.dgebal(a, permute=1, scale=0)
        assert (ilo, ihi) == (slo + 1, shi + 1)
        np.testing.assert_array_equal(np.sort(np.diag(b)), np.sort(np.diag(sb)))
        k, l = ilo - 1, ihi - 1
        # rows below the window and columns left of it are triangular
        for i in range(l + 1, n):
            assert np.all(b[i, :i] == 0)
        for j in range(k):
            assert np.all(b[j + 1:, j] == 0)
        # inside the window no row or column isolates an eigenvalue any more
        if ilo < ihi:
            for j in range(k, l + 1):
                row = b[j, :l + 1].copy()
                row[j] = 0
                col = b[k:l + 1, j].copy()
                col[j - k] = 0
                assert np.any(row != 0) and np.any(col != 0)
        np.testing.assert_array_equal(b, ob.apply_balance_similarity(a, ilo, ihi, d))


def _badly_scaled(rng, n, spread=40, cplx=False):
    a = rng.standard_normal((n, n))
    if cplx:
        a = a + 1j * rng.standard_normal((n, n))
    e = rng.integers(-spread, spread + 1, n).astype(float)
    return a * (2.0 ** e)[:, None] / (2.0 ** e)[None, :]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("job", [ob.SCALE, ob.BOTH])
def test_exact_reproduction_stationarity_idempotence(cplx, job):
    rng = np.random.default_rng(7 + job)
    for n in [2, 3, 8, 25, 60]:
        a = _badly_scaled(rng, n, cplx=cplx)
        a[rng.random((n, n)) < 0.15] = 0
        if n > 3:  # a row and a column that isolate an eigenvalue
            a[1, :] = 0
            a[1, 1] = 2.0
            a[:, n - 2] = 0
            a[n - 2, n - 2] = -1.0
        b, ilo, ihi, d = ob.gebal(a, job)
        # exact: powers of two
        window = d[ilo - 1:ihi]
        assert np.all(np.frexp(window)[0] == 0.5)
        np.testing.assert_array_equal(b, ob.apply_balance_similarity(a, ilo, ihi, d))
        # stationarity: no power of two improves an index by the 5% LAPACK asks
        k, l = ilo - 1, ihi - 1
        mag = ob._abs1(b)
        for i in range(k, l + 1):
            c = mag[k:l + 1, i].sum() - mag[i, i]
            r = mag[i, k:l + 1].sum() - mag[i, i]
            if c == 0 or r == 0:
                continue
            for e in range(-60, 61):
                if e == 0:
                    continue
                f = 2.0 ** e
                assert c * f + r / f >= 0.95 * (c + r) * (1 - 1e-12)
        # idempotence: the same job again finds the same window and scales nothing
        b2, ilo2, ihi2, d2 = ob.gebal(b, job)
        assert (ilo2, ihi2) == (ilo, ihi)
        assert np.all(d2[ilo2 - 1:ihi2] == 1.0)
        np.testing.assert_array_equal(b2, b)


def test_balancing_lowers_the_norm_on_a_badly_scaled_matrix():
    rng = np.random.default_rng(3)
    a = _badly_scaled(rng, 40, spread=30)
    b, *_ = ob.gebal(a, ob.SCALE)
    assert np.abs(b).sum(0).max() < 1e-6 * np.abs(a).sum(0).max()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 10, 40, 100])
def test_lacn2_matches_lapack_dgecon(n):
    rng = np.random.default_rng(n)
    for _ in range(3):
        m = rng.standard_normal((n, n))
        lu, piv, info = la.dgetrf(m)
        anorm = np.abs(m).sum(0).max()
        rcond, info = la.dgecon(lu, anorm, norm="1")
        est, v, seq = ob.lacn2(np.linalg.inv(m))
        ref = 1.0 / (rcond * anorm)
        assert abs(est - ref) <= 1e-12 * ref
        assert est <= np.abs(np.linalg.inv(m)).sum(0).max() * (1 + 1e-12)
        assert seq[0] == 1 and len(seq) <= 2 * 5 + 1


@pytest.mark.parametrize("n", [1, 2, 7, 64])
def test_lacn2_exact_on_nonnegative_and_diagonal(n):
    rng = np.random.default_rng(11 + n)
    a = rng.random((n, n))
    est, v, seq = ob.lacn2(a)
    norm = a.sum(0).max()
    assert abs(est - norm) <= 1e-14 * norm
    j = int(np.argmax(a.sum(0)))
    if n > 1:
        assert seq == [1, 2, 1, 1]
        np.testing.assert_allclose(v, a[:, j], rtol=0, atol=0)
    d = np.diag(rng.standard_normal(n))
    est, v, seq = ob.lacn2(d)
    assert abs(est - np.abs(np.diag(d)).max()) <= 1e-15 * np.abs(np.diag(d)).max()
