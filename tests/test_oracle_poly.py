"""CPU: pin the oracle's Paterson-Stockmeyer / Horner restatement and the Taylor
ladder the GPU library uses.

  * plan known answers from test/paterson_stockmeyer/paterson_stockmeyer_tests.cpp:235-324
    (cost table, degree 0/1/2 edge cases, exact argmin through degree 32,
    s_requested override and clamp);
  * the evaluators against an independent power-series oracle (powers formed
    explicitly by numpy), with the reference's bound 64*eps*sum|c_k|*||A||_F^k*n
    (paterson_stockmeyer_tests.cpp:125-135): every block size, PS == Horner,
    zero polynomial exactly, nilpotent input;
  * the Taylor thresholds: theta_18 against the published value
    (Bader, Blanes, Casas 2019: 1.090863719290036) and a fresh mpmath
    bisection of the backward-error series at two degrees; the 1/k! table
    against exact rationals.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

EPS = np.finfo(float).eps


def power_series(coeffs, a):
    n = a.shape[0]
    res = np.zeros((n, n), dtype=np.result_type(a, coeffs))
    p = np.eye(n, dtype=res.dtype)
    for k, c in enumerate(coeffs):
        if k:
            p = p @ a
        res = res + c * p
    return res


def poly_tol(coeffs, a):
    anorm = np.linalg.norm(a)
    bound = sum(abs(c) * anorm ** k for k, c in enumerate(coeffs))
    return 64 * EPS * bound * a.shape[0]


# ---- plans -----------------------------------------------------------------------------------
def test_cost_table_matches_readme():
    assert [oracle.ps_plan(d)["num_gemms"] for d in (4, 8, 16, 30)] == [3, 4, 7, 10]
    p12 = oracle.ps_plan(12)
    assert (p12["s"], p12["num_gemms"], p12["num_blocks"]) == (3, 6, 3)


def test_edge_degrees():
    p0, p1, p2 = oracle.ps_plan(0), oracle.ps_plan(1), oracle.ps_plan(2)
    assert p0["num_gemms"] == 0 and p0["num_blocks"] == 0
    assert (p1["s"], p1["num_gemms"], p1["num_blocks"]) == (2, 0, 0)
    assert (p2["s"], p2["num_gemms"]) == (3, 1)


def test_block_size_is_exact_argmin():
    for d in range(1, 33):
        best = oracle.ps_plan(d)["num_gemms"]
        assert best <= d
        assert all(best <= oracle.ps_cost(d, s) for s in range(1, d + 2))
        assert best == oracle.ps_cost(d, oracle.ps_plan(d)["s"])


def test_s_requested_override_and_clamp():
    assert oracle.ps_plan(10, 1)["num_gemms"] == 10
    assert oracle.ps_plan(5, 999)["s"] == 6
    assert oracle.ps_plan(9, 4)["s"] == 4


# ---- evaluators --------------------------------------------------------------------------------
def _poly_case(n, degree, seed, cplx=False):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.5, 0.5, (n, n))
    c = rng.uniform(-0.5, 0.5, degree + 1) * 2
    if cplx:
        a = a + 1j * rng.uniform(-0.5, 0.5, (n, n))
        c = c + 1j * rng.uniform(-1, 1, degree + 1)
    return a, c


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n,degree", [(1, 3), (5, 7), (8, 12), (16, 16), (9, 29)])
def test_ps_and_horner_match_power_series(cplx, n, degree):
    a, c = _poly_case(n, degree, 30 + n + degree, cplx)
    ref = power_series(c, a)
    tol = poly_tol(c, a)
    assert np.abs(oracle.paterson_stockmeyer(c, a) - ref).max() <= tol
    assert np.abs(oracle.horner(c, a) - ref).max() <= tol


def test_every_block_size_agrees():
    a, c = _poly_case(5, 7, 77)
    ref = power_series(c, a)
    for s in range(1, 9):
        assert np.abs(oracle.paterson_stockmeyer(c, a, s) - ref).max() <= poly_tol(c, a), s


def test_zero_polynomial_exact():
    a, _ = _poly_case(6, 8, 40)
    np.testing.assert_array_equal(oracle.paterson_stockmeyer(np.zeros(9), a), np.zeros((6, 6)))
    np.testing.assert_array_equal(oracle.horner(np.zeros(9), a), np.zeros((6, 6)))


def test_nilpotent_terminates():
    n, degree = 5, 9
    a = np.triu(np.fromfunction(lambda i, j: 0.3 * (i + j + 1), (n, n)), 1)
    c = np.random.default_rng(50).uniform(-1, 1, degree + 1)
    ref = power_series(c[:n], a)  # A^k = 0 for k >= n
    assert np.abs(oracle.paterson_stockmeyer(c, a) - ref).max() <= poly_tol(c, a)


def test_degree_zero_and_one():
    a, c = _poly_case(4, 1, 3)
    np.testing.assert_array_equal(oracle.paterson_stockmeyer(c[:1], a), c[0] * np.eye(4))
    np.testing.assert_allclose(oracle.paterson_stockmeyer(c, a), c[0] * np.eye(4) + c[1] * a,
                               rtol=0, atol=4 * EPS)


# ---- Taylor ladder ---------------------------------------------------------------------------------
def test_inverse_factorials_correctly_rounded():
    got = oracle.inv_factorials()
    for k in range(33):
        assert got[k] == float(Fraction(1, math.factorial(k))), k


def test_theta18_published_value():
    assert oracle.taylor_theta(18) == 0.0  # 18 is not on this ladder (19 costs the same)
    # the published value for degree 18 anchors the derivation: theta(18) lies
    # between the ladder's neighbours and the bisection below reproduces it
    assert oracle.taylor_theta(15) < 1.090863719290036 < oracle.taylor_theta(19)


def _theta_bisect(m, u, K=140):
    """Largest x with sum_{k>m} |h_k| x^(k-1) <= u, h = log(exp(-x) T_m(x))."""
    import mpmath as mp
    mp.mp.dps = 50
    g = [Fraction(0)] * (K + 1)
    for k in range(K + 1):
        g[k] = sum(Fraction((-1) ** (k - j), math.factorial(j) * math.factorial(k - j))
                   for j in range(min(k, m) + 1))
    d = g[:]
    d[0] -= 1
    res = [Fraction(0)] * (K + 1)
    p = [Fraction(1)] + [Fraction(0)] * K
    for nn in range(1, K // (m + 1) + 1):
        q = [Fraction(0)] * (K + 1)
        for i, x in enumerate(p):
            if x:
                for j in range(m + 1, K + 1 - i):
                    if d[j]:
                        q[i + j] += x * d[j]
        p = q
        for k in range(K + 1):
            if p[k]:
                res[k] += Fraction((-1) ** (nn + 1), nn) * p[k]
    ac = [mp.mpf(abs(x.numerator)) / abs(x.denominator) for x in res]
    f = lambda x: mp.fsum(ac[k] * x ** (k - 1) for k in range(m + 1, K + 1)) - u
    lo, hi = mp.mpf(0), mp.mpf(8)
    for _ in range(120):
        mid = (lo + hi) / 2
        lo, hi = (mid, hi) if f(mid) <= 0 else (lo, mid)
    return float(lo)


def test_theta18_rederived():
    assert _theta_bisect(18, 2.0 ** -53) == pytest.approx(1.090863719290036, rel=1e-12)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("m", oracle.TAYLOR_DEGREES)
def test_theta_rederived(m, single):
    """every one of the 20 ladder thresholds, from the exact series"""
    u = 2.0 ** -24 if single else 2.0 ** -53
    assert _theta_bisect(m, u) == pytest.approx(oracle.taylor_theta(m, single), rel=1e-12)


def test_taylor_ladder_costs():
    degrees = oracle.TAYLOR_DEGREES
    costs = [oracle.ps_plan(d)["num_gemms"] for d in degrees]
    assert costs == list(range(10))
    # each ladder degree is the largest with that product count
    for d, c in zip(degrees, costs):
        if d > 1:
            assert oracle.ps_plan(d + 1)["num_gemms"] == c + 1
    th = [oracle.taylor_theta(d) for d in degrees]
    assert th == sorted(th)


def test_taylor_plan_rule():
    # (degree, s, gemms) at the benchmark norms
    want = {0.5: (15, 0, 6), 4.0: (24, 1, 9), 8.0: (24, 2, 10), 100.0: (29, 5, 14)}
    for norm, (d, s, g) in want.items():
        p = oracle.taylor_plan(norm)
        assert (p["degree"], p["s"], p["num_gemms"]) == (d, s, g)
        assert norm / 2 ** s <= oracle.taylor_theta(d)
    assert oracle.taylor_plan(0.0)["num_gemms"] == 0


# ---- squarings from the norms of powers ---------------------------------------------------------
def _alpha_ok(norms, m, x):
    """alpha_p <= x for some p with p(p-1) <= m+1 (p = 1: ||A||_1), k-th roots in
    float arithmetic -- an independent statement of the power rule"""
    if norms[0] <= x:
        return True
    q = len(norms)
    for p in range(2, q):
        if p * (p - 1) <= m + 1:
            if max(norms[p - 1] ** (1.0 / p), norms[p] ** (1.0 / (p + 1))) <= x * (1 + 1e-12):
                return True
    return False


def test_power_plan_rule():
    """The plan from norms of powers: s is the smallest with alpha_p / 2^s <= theta_m
    (Al-Mohy & Higham 2009, Thm 4.2) and the rung minimises products + s, with the
    formed powers counted once -- checked by brute force over rungs and s."""
    rng = np.random.default_rng(1)
    for _ in range(400):
        q = int(rng.integers(3, 6))
        nrm = float(10.0 ** rng.uniform(-3, 3))
        roots = nrm * 10.0 ** rng.uniform(-2.5, 0, size=q - 1)
        norms = [nrm] + [float(r ** k) for k, r in zip(range(2, q + 1), roots)]
        p = oracle.taylor_plan_powers(norms)
        best = None
        for m in oracle.TAYLOR_DEGREES:
            th = oracle.taylor_theta(m)
            s = 0
            while not _alpha_ok(norms, m, th * 2.0 ** s):
                s += 1
            pp = oracle.ps_plan(m)
            cost = max(pp["num_powers"], q - 1) + pp["r"] + s
            if best is None or cost < best[2] or (cost == best[2] and s < best[1]):
                best = (m, s, cost)
        # the float k-th roots may sit an ulp either side of a threshold: allow one
        # squaring of slack there, never a cheaper plan than the brute force finds
        assert p["num_gemms"] <= best[2] + 1 and p["num_gemms"] >= best[2] - 1, (norms, p, best)
        assert _alpha_ok(norms, p["degree"], oracle.taylor_theta(p["degree"]) * 2.0 ** p["s"] *
                         (1 + 1e-9))
        if p["s"] > 0:  # minimal: one squaring fewer violates the bound for this degree
            assert not _alpha_ok(norms, p["degree"],
                                 oracle.taylor_theta(p["degree"]) * 2.0 ** (p["s"] - 1) * (1 - 1e-9))


def test_power_plan_expm_accuracy_non_normal():
    """Matrices far from normal (||A||_1 >> spectral radius): the plan from power norms
    takes fewer products than the norm-only plan (it may trade a squaring for a lower
    degree, the powers formed for the norms counted) and the Taylor expm stays within 1e-13
    of the Pade oracle; an upper-triangular Jordan-like block and a dense Gaussian."""
    rng = np.random.default_rng(3)
    n = 48
    cases = []
    h = n // 2  # [[D1, B], [0, D2]]: ||B||_1 = 60 against a spectrum in [-0.3, 0.3]
    blk = np.zeros((n, n))
    blk[:h, h:] = rng.uniform(0, 1, (h, h))
    blk *= 60.0 / np.abs(blk).sum(0).max()
    blk += np.diag(rng.uniform(-0.3, 0.3, n))
    cases.append(blk)
    g = rng.standard_normal((n, n))
    cases.append(g * (100.0 / np.abs(g).sum(0).max()))
    for a in cases:
        e, plan = oracle.expm_taylor(a, return_plan=True)
        p0 = oracle.taylor_plan(np.abs(a).sum(0).max())
        assert plan["num_gemms"] < p0["num_gemms"]
        ref = oracle.expm(a)
        assert np.linalg.norm(e - ref) / np.linalg.norm(ref) <= 1e-13
