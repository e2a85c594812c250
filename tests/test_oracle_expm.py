"""CPU: pin the oracle (the C restatement of the reference expm) before it is
trusted as the GPU parity target.

What pins it (none of these retypes the oracle's own formula):
  * the Pade coefficient tables against the closed form
    b_j = (2m-j)! m! / ((2m)! j! (m-j)!) in exact rationals
    (src/expm/expm_bridge.h:49-50);
  * the thresholds against Higham's published table (Higham 2005, Table 2.3)
    and the reference's plan-ladder known answers (test/expm/expm_tests.cpp:389-435);
  * exp(A) against golden vectors from mpmath at 60 digits and from
    scipy.linalg.expm (an independent Pade variant);
  * closed forms and identities from test/expm/expm_tests.cpp:471-698;
  * the 1-norm against numpy, including NaN propagation and the complex modulus;
  * ?lascl's defining property for powers of two (bitwise ldexp) and its errors;
  * ?lange against numpy's norms; GEMM and LU against numpy, GEMM bitwise
    independent of the thread count.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EPS = np.finfo(float).eps


def rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


def scaled(n, target, seed, cplx=False, uniform=True):
    """The reference's scaled_random (test/expm/expm_tests.cpp:315-336)."""
    rng = np.random.default_rng(seed)
    f = (lambda s: rng.uniform(-1, 1, s)) if uniform else rng.standard_normal
    a = f((n, n))
    if cplx:
        a = a + 1j * f((n, n))
    return a * (target / np.abs(a).sum(0).max())


# ---- Pade tables, thresholds, plan -------------------------------------------------------
@pytest.mark.parametrize("m", oracle.PADE_DEGREES)
def test_pade_coefficients_closed_form(m):
    f = math.factorial
    got = oracle.pade_coeffs(m)
    for j in range(m + 1):
        exact = Fraction(f(2 * m - j) * f(m), f(2 * m) * f(j) * f(m - j))
        assert got[j] == float(exact), (m, j)


def test_pade_thresholds_match_higham_2005():
    published = {3: 1.495585217958292e-2, 5: 2.539398330063230e-1, 7: 9.504178996162932e-1,
                 9: 2.097847961257068e0, 13: 5.371920351148152e0}
    for m, v in published.items():
        assert oracle.pade_theta(m) == pytest.approx(v, rel=1e-14)
    for bad in (0, 4, 6, 11):
        assert oracle.pade_theta(bad) == 0.0
    th = [oracle.pade_theta(m) for m in oracle.PADE_DEGREES]
    assert th == sorted(th) and len(set(th)) == 5


@pytest.mark.parametrize("single", [False, True])
def test_plan_ladder(single):
    want = {3: 2, 5: 3, 7: 4, 9: 5, 13: 6}
    for m in oracle.PADE_DEGREES:
        th = oracle.pade_theta(m, single)
        pm, ps, pg = oracle.expm_plan(th * 0.99, single)
        assert ps == 0 and pm <= m and pg == want[pm]
    assert oracle.expm_plan(oracle.pade_theta(13, single), single)[1] == 0
    pm, ps, pg = oracle.expm_plan(oracle.pade_theta(13, single) * 40, single)
    assert pm == 13 and ps > 0 and pg == 6 + ps


def test_plan_at_config_norms():
    # (m, s, gemms) at the benchmark norms, from the rule in src/expm/plan.cppm:112-130
    table = {0.5: (7, 0, 4), 1: (9, 0, 5), 4: (13, 0, 6), 8: (13, 1, 7), 16: (13, 2, 8),
             100: (13, 5, 11), 256: (13, 6, 12)}
    for norm, want in table.items():
        assert oracle.expm_plan(norm) == want


# ---- golden vectors ---------------------------------------------------------------------
def _cases():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["expm"]


def golden_tol(case):
    return 64 * EPS * max(1.0, case["norm1"]) * max(1.0, math.sqrt(case["n"]))


@pytest.mark.parametrize("case", _cases(), ids=lambda c: f"{c['key']}-{c['kind']}-n{c['n']}-{c['norm1']}")
def test_oracle_matches_golden(case):
    d = np.load(os.path.join(GOLDEN, case["file"]))
    a, e = d[case["key"] + "_A"], d[case["key"] + "_E"]
    assert rel(oracle.expm(a), e) <= golden_tol(case)
    assert rel(oracle.expm_taylor(a), e) <= 4 * golden_tol(case)


def test_golden_fixture_inputs_have_stated_norms():
    for c in _cases():
        a = np.load(os.path.join(GOLDEN, c["file"]))[c["key"] + "_A"]
        assert a.shape == (c["n"], c["n"])
        assert np.abs(a).sum(0).max() == pytest.approx(c["norm1"], rel=1e-12)


# ---- closed forms and identities ----------------------------------------------------------
def elem_tol(ref, n, factor):
    return factor * EPS * max(1.0, np.linalg.norm(ref)) * n


@pytest.mark.parametrize("cplx", [False, True])
def test_zero_is_identity(cplx):
    a = np.zeros((6, 6), dtype=complex if cplx else float)
    np.testing.assert_array_equal(oracle.expm(a), np.eye(6))
    np.testing.assert_array_equal(oracle.expm_taylor(a), np.eye(6))


@pytest.mark.parametrize("cplx", [False, True])
def test_diagonal(cplx):
    rng = np.random.default_rng(2)
    d = rng.uniform(-1, 1, 6) + (1j * rng.uniform(-1, 1, 6) if cplx else 0)
    ref = np.diag(np.exp(d))
    for f in (oracle.expm, oracle.expm_taylor):
        assert np.abs(f(np.diag(d)) - ref).max() <= elem_tol(ref, 6, 128)


@pytest.mark.parametrize("cplx", [False, True])
def test_nilpotent_series_terminates(cplx):
    n = 5
    rng = np.random.default_rng(10)
    N = np.triu(rng.uniform(-0.5, 0.5, (n, n)), 1)
    if cplx:
        N = N + 1j * np.triu(rng.uniform(-0.5, 0.5, (n, n)), 1)
    ref, p = np.eye(n, dtype=N.dtype), np.eye(n, dtype=N.dtype)
    for k in range(1, n):
        p = p @ N / k
        ref = ref + p
    for f in (oracle.expm, oracle.expm_taylor):
        assert np.abs(f(N) - ref).max() <= elem_tol(ref, n, 128)


@pytest.mark.parametrize("cplx,target", [(False, 1.5), (True, 1.2), (False, 30.0)])
def test_inverse_identity(cplx, target):
    a = scaled(6, target, 21, cplx)
    for f in (oracle.expm, oracle.expm_taylor):
        prod = f(a) @ f(-a)
        cond = np.linalg.norm(f(a)) * np.linalg.norm(f(-a))
        assert np.abs(prod - np.eye(6)).max() <= elem_tol(np.eye(6), 6, 256) * cond


@pytest.mark.parametrize("cplx", [False, True])
def test_half_squared(cplx):
    a = scaled(6, 3.0, 30, cplx)
    for f in (oracle.expm, oracle.expm_taylor):
        sq = f(a / 2) @ f(a / 2)
        assert np.abs(f(a) - sq).max() <= elem_tol(sq, 6, 256)


@pytest.mark.parametrize("cplx", [False, True])
def test_each_rung_matches_pade13(cplx):
    """test/expm/expm_tests.cpp:661-698: inside each rung, expm (cheap degree,
    no scaling) agrees with r_13 on the same matrix."""
    prev = 0.0
    for i, m in enumerate(oracle.PADE_DEGREES):
        th = oracle.pade_theta(m)
        a = scaled(5, 0.5 * (prev + th), 100 + i, cplx)
        prev = th
        e, plan = oracle.expm(a, return_plan=True)
        assert plan[1] == 0 and plan[0] <= m
        r13 = oracle.pade(13, a)
        assert np.abs(e - r13).max() <= elem_tol(r13, 5, 256)


def test_pade_rejects_off_ladder_degree():
    with pytest.raises(ValueError):
        oracle.pade(6, np.eye(3))


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_norm_is_rejected(bad):
    a = scaled(6, 1.0, 3)
    a[1, 4] = bad
    assert oracle.expm_status(a) == -2


# ---- norms ----------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 3, 255, 256, 257, 600])
def test_norm1_against_numpy(n):
    a = np.random.default_rng(n).standard_normal((n, n))
    assert oracle.norm1(a) == pytest.approx(np.abs(a).sum(0).max(), rel=8 * n * EPS)
    z = a + 1j * np.random.default_rng(n + 1).standard_normal((n, n))
    assert oracle.norm1(z) == pytest.approx(np.abs(z).sum(0).max(), rel=8 * n * EPS)


def test_norm1_complex_is_true_modulus():
    z = np.array([[3 + 4j, 0], [0, 1j]])
    assert oracle.norm1(z) == 5.0  # |3+4i| = 5, not |3| + |4| = 7


def test_norm1_nan_wins():
    a = np.ones((300, 300))
    a[299, 7] = np.nan
    assert math.isnan(oracle.norm1(a))
    a[299, 7] = np.inf
    assert oracle.norm1(a) == np.inf


@pytest.mark.parametrize("which", ["max_abs", "one", "inf", "frobenius"])
@pytest.mark.parametrize("shape", [(1, 1), (7, 3), (300, 260), (5, 600)])
def test_lange_against_numpy(which, shape):
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    for a in (rng.standard_normal(shape), rng.standard_normal(shape) + 1j * rng.standard_normal(shape)):
        ref = {"max_abs": np.abs(a).max(), "one": np.abs(a).sum(0).max(),
               "inf": np.abs(a).sum(1).max(), "frobenius": np.linalg.norm(a)}[which]
        assert oracle.lange(which, a) == pytest.approx(ref, rel=16 * max(shape) * EPS)


def test_lange_empty_and_nan():
    assert oracle.lange("one", np.zeros((0, 4))) == 0.0
    a = np.ones((4, 4))
    a[2, 2] = np.nan
    for w in ("max_abs", "one", "inf", "frobenius"):
        assert math.isnan(oracle.lange(w, a))


EPS32 = float(np.finfo(np.float32).eps)


@pytest.mark.parametrize("n", [1, 3, 256, 257, 600])
def test_norm1_single_against_numpy(n):
    rng = np.random.default_rng(n + 7)
    a = rng.standard_normal((n, n)).astype(np.float32)
    assert oracle.norm1(a) == pytest.approx(np.abs(a.astype(np.float64)).sum(0).max(),
                                            rel=4 * n * EPS32)
    z = (a + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    assert oracle.norm1(z) == pytest.approx(np.abs(z.astype(np.complex128)).sum(0).max(),
                                            rel=4 * n * EPS32)


def test_norm1_single_accumulates_in_float():
    """column (1, 2^-24, 2^-24): the fold adds each 2^-24 to 1 separately
    (lane 2 joins lane 0 at the tree step s = 2, lane 1 at s = 1), and each add
    rounds back to 1 in float; a double accumulator would give 1 + 2^-23."""
    a = np.zeros((3, 3), dtype=np.float32)
    a[:, 0] = [1.0, 2.0 ** -24, 2.0 ** -24]
    assert oracle.norm1(a) == 1.0
    assert oracle.lange("one", a) == 1.0
    assert oracle.norm1(a.astype(np.float64)) == 1.0 + 2.0 ** -23


def test_lange_single_against_numpy():
    rng = np.random.default_rng(11)
    for shape in ((1, 1), (7, 3), (300, 260), (5, 600)):
        a = rng.standard_normal(shape).astype(np.float32)
        w = a.astype(np.float64)
        for which, ref in (("max_abs", np.abs(w).max()), ("one", np.abs(w).sum(0).max()),
                           ("inf", np.abs(w).sum(1).max()), ("frobenius", np.linalg.norm(w))):
            assert oracle.lange(which, a) == pytest.approx(ref, rel=8 * max(shape) * EPS32)
    nan = np.ones((4, 4), dtype=np.float32)
    nan[1, 1] = np.nan
    assert all(math.isnan(oracle.lange(w, nan)) for w in ("max_abs", "one", "inf", "frobenius"))


def test_lascl_single_guarded_chain():
    """cto/cfrom = 1e60 overflows float if formed directly; the chain (BIGNUM
    first, then the representable rest) lands on 1e25 for 1e-35."""
    a = np.full((2, 3), 1e-35, dtype=np.float32)
    st, got = oracle.lascl(1e-30, 1e30, a)
    assert st == 0 and got.dtype == np.float32
    assert np.all(np.isfinite(got))
    assert np.allclose(got, 1e25, rtol=8 * EPS32)
    st, got = oracle.lascl(8.0, 1.0, np.full((2, 2), 3.0, dtype=np.float32))
    assert st == 0 and np.all(got == np.float32(0.375))
    for cfrom, cto in ((0.0, 1.0), (np.nan, 1.0), (1.0, np.nan)):
        assert oracle.lascl(cfrom, cto, a)[0] == -1


# ---- lascl --------------------------------------------------------------------------------------
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("s", [0, 1, 5, 7, 40])
def test_lascl_power_of_two_is_exact(cplx, s):
    a = scaled(9, 8.0, s, cplx)
    st, got = oracle.lascl(2.0 ** s, 1.0, a)
    assert st == 0
    np.testing.assert_array_equal(got, np.ldexp(a.real, -s) + (1j * np.ldexp(a.imag, -s) if cplx else 0))


def test_lascl_ratio_and_guarded_extremes():
    a = scaled(5, 1.0, 1)
    st, got = oracle.lascl(3.0, 7.0, a)
    np.testing.assert_array_equal(got, a * (7.0 / 3.0))
    # a ratio of 1e-300 / 1e300 underflows if formed directly; the guarded chain
    # applies it in representable steps
    st, got = oracle.lascl(1e300, 1e-300, np.full((2, 2), 1e300))
    assert st == 0 and np.all(got == pytest.approx(1e-300, rel=1e-13))
    for cfrom, cto in ((0.0, 1.0), (np.nan, 1.0), (1.0, np.nan)):
        assert oracle.lascl(cfrom, cto, a)[0] == -1


# ---- GEMM / LU -------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 5, 64, 131])
def test_gemm_against_numpy(n):
    rng = np.random.default_rng(n)
    a, b, c = (rng.standard_normal((n, n)) for _ in range(3))
    tol = 8 * n * EPS * np.abs(a).max() * np.abs(b).max() + 4 * EPS * np.abs(c).max()
    np.testing.assert_allclose(oracle.gemm(a, b, 1.5, -0.5, c), 1.5 * a @ b - 0.5 * c, rtol=0,
                               atol=2 * tol)
    za, zb = a + 1j * c, b - 2j * a
    np.testing.assert_allclose(oracle.gemm(za, zb), za @ zb, rtol=0, atol=32 * n * EPS * 16)


def test_gemm_independent_of_thread_count():
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal((300, 300)), rng.standard_normal((300, 300))
    t0 = oracle.num_threads()
    try:
        oracle.set_threads(1)
        one = oracle.gemm(a, b)
        oracle.set_threads(max(2, t0))
        many = oracle.gemm(a, b)
    finally:
        oracle.set_threads(t0)
    np.testing.assert_array_equal(one, many)


@pytest.mark.parametrize("cplx", [False, True])
def test_lu_solve_against_numpy(cplx):
    rng = np.random.default_rng(3)
    n = 40
    a = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    b = rng.standard_normal((n, 7)) + (1j * rng.standard_normal((n, 7)) if cplx else 0)
    x, info = oracle.lu_solve(a, b)
    assert info == 0
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=0, atol=1e-10)


def test_lu_reports_singularity():
    a = np.ones((4, 4))
    _, info = oracle.lu_solve(a, np.ones((4, 1)))
    assert info > 0


# ---- Frechet derivative (the adjoint's oracle) ---------------------------------------------
def _golden_frechet():
    import json
    import os
    g = os.path.join(os.path.dirname(__file__), "golden")
    man = json.load(open(os.path.join(g, "manifest.json")))
    data = np.load(os.path.join(g, "frechet_scipy.npz"))
    return [(c, data) for c in man["frechet"]]


@pytest.mark.parametrize("case", _golden_frechet(), ids=lambda c: f"{c[0]['key']}-{c[0]['kind']}")
def test_frechet_matches_scipy_golden(case):
    c, data = case
    k = c["key"]
    got = oracle.frechet(data[f"{k}_A"], data[f"{k}_dir"])
    ref = data[f"{k}_L"]
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("cplx", [False, True])
def test_frechet_against_central_differences(cplx):
    rng = np.random.default_rng(21)
    n = 7
    a = rng.standard_normal((n, n)) * 0.4
    e = rng.standard_normal((n, n))
    if cplx:
        a = a + 1j * rng.standard_normal((n, n)) * 0.4
        e = e + 1j * rng.standard_normal((n, n))
    h = 1e-6
    fd = (oracle.expm(a + h * e) - oracle.expm(a - h * e)) / (2 * h)
    lf = oracle.frechet(a, e)
    assert np.linalg.norm(lf - fd) <= 1e-8 * np.linalg.norm(lf)
    # linear in the direction
    assert np.allclose(oracle.frechet(a, 2.5 * e), 2.5 * lf, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("cplx", [False, True])
def test_frechet_adjoint_is_the_gradient(cplx):
    """Re <grad, D> equals the central difference of L(A) = Re tr(G^H exp(A))
    along D, and the commuting case exp'(A)[E] = E exp(A) for A = a I."""
    rng = np.random.default_rng(22)
    n = 6
    a = rng.standard_normal((n, n)) * 0.5
    g = rng.standard_normal((n, n))
    d = rng.standard_normal((n, n))
    if cplx:
        a, g, d = (x + 1j * rng.standard_normal((n, n)) * s
                   for x, s in ((a, 0.5), (g, 1.0), (d, 1.0)))
    grad = oracle.frechet_adjoint(a, g)
    loss = lambda x: np.real(np.trace(g.conj().T @ oracle.expm(x)))
    h = 1e-6
    fd = (loss(a + h * d) - loss(a - h * d)) / (2 * h)
    assert np.real(np.vdot(grad, d)) == pytest.approx(fd, rel=1e-7)
    # A = 0.3 I commutes with everything: L(A, E) = e^0.3 E exactly
    assert np.allclose(oracle.frechet(0.3 * np.eye(n), d), np.exp(0.3) * d, rtol=1e-14)
