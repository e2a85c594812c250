"""GPU: the double and complex<double> forms of the stand-alone entry points --
clm_paterson_stockmeyer_{d,z}, clm_horner_{d,z}, clm_lange_{d,z} and
clm_lascl_{d,z} -- against the oracle, mirroring the reference's own suites:

  * Paterson-Stockmeyer / Horner (test/paterson_stockmeyer/paterson_stockmeyer_tests.cpp:396-586,
    test/horner/horner_tests.cpp:269-365): even and odd degrees, every block
    size 1 .. d+1 on one input, PS against Horner directly, the zero polynomial
    exactly zero, a nilpotent input, padding rows between n and ldp untouched.
    Bound: the reference's 64 eps sum_k |c_k| ||A||_F^k n
    (paterson_stockmeyer_tests.cpp:125-135), against the oracle's explicit
    power series;
  * ?lange (test/lange/lange_tests.cpp:74-145): every norm, tall / wide /
    single row or column, padding rows holding a huge sentinel that must not be
    read; bit-for-bit the oracle's fold for a real matrix (complex moduli to a
    few ulps);
  * ?lascl (test/lascl/lascl_tests.cpp:90-200): exact power-of-two ratios, a
    ratio needing the guarded multiplier chain, padding untouched; bit-for-bit
    the oracle.
"""
import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle
from calaman_g_amd.expm import Context, horner_cm, lange_cm, lascl_cm, paterson_stockmeyer_cm

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float64).eps)


def _tdt(cplx):
    return torch.complex128 if cplx else torch.float64


def _rand(rng, shape, cplx, lo=-0.5, hi=0.5):
    a = rng.uniform(lo, hi, shape)
    if cplx:
        a = a + 1j * rng.uniform(lo, hi, shape)
    return a


def _cm(a, ld, dev, sentinel=-123.5):
    """device buffer holding a column-major (m x n) with leading dimension ld; the
    rows between m and ld hold the sentinel"""
    m, n = a.shape
    buf = np.full((max(n, 1), ld), sentinel, dtype=a.dtype)
    buf[:n, :m] = a.T
    return torch.from_numpy(buf).to(dev)


def _from_cm(t, m, n):
    return t.cpu().numpy()[:n, :m].T.copy()


def poly_tol(a, c, n):
    fro = np.linalg.norm(a)
    return 64 * EPS * n * sum(abs(complex(ck)) * fro ** k for k, ck in enumerate(c))


def run_ps(dev, a, c, s_req=0, ldp=None, sentinel=-123.5):
    n = a.shape[0]
    ldp = n if ldp is None else ldp
    cplx = np.iscomplexobj(a)
    ctx = Context()
    degree = len(c) - 1
    dc = torch.from_numpy(np.asarray(c, dtype=a.dtype)).to(dev)
    da = _cm(a, n, dev)
    dp = torch.from_numpy(np.full((n, ldp), sentinel, dtype=a.dtype)).to(dev)
    work = torch.empty(max(cg.ps_buffer_size(n, degree, s_req, _tdt(cplx)), 1), dtype=torch.uint8,
                       device=dev)
    st = paterson_stockmeyer_cm(ctx, n, dc, degree, da, n, dp, ldp, work, s_req)
    torch.cuda.synchronize()
    return st, dp.cpu().numpy()


def run_horner(dev, a, c, ldp=None, sentinel=-123.5):
    n = a.shape[0]
    ldp = n if ldp is None else ldp
    cplx = np.iscomplexobj(a)
    ctx = Context()
    degree = len(c) - 1
    dc = torch.from_numpy(np.asarray(c, dtype=a.dtype)).to(dev)
    da = _cm(a, n, dev)
    dp = torch.from_numpy(np.full((n, ldp), sentinel, dtype=a.dtype)).to(dev)
    work = torch.empty(cg.horner_buffer_size(n, _tdt(cplx)), dtype=torch.uint8, device=dev)
    st = horner_cm(ctx, n, dc, degree, da, n, dp, ldp, work)
    torch.cuda.synchronize()
    return st, dp.cpu().numpy()


def power_series(a, c):
    """the reference's independent oracle: sum_k c_k A^k with explicit powers
    (paterson_stockmeyer_tests.cpp:98-117), in numpy float64 / complex128"""
    n = a.shape[0]
    out = np.zeros_like(a)
    p = np.eye(n, dtype=a.dtype)
    for k, ck in enumerate(c):
        if k:
            p = p @ a
        out = out + ck * p
    return out


# ---- Paterson-Stockmeyer and Horner ---------------------------------------------------------
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n,degree,seed", [(5, 0, 2), (6, 1, 4), (5, 2, 12), (6, 4, 13), (5, 3, 22),
                                           (6, 5, 23), (5, 6, 22), (6, 10, 23), (16, 12, 30),
                                           (16, 16, 31), (16, 4, 30), (16, 5, 31), (64, 18, 40),
                                           (200, 29, 41)])
def test_ps_and_horner_match_power_series(dev, cplx, n, degree, seed):
    rng = np.random.default_rng(seed)
    a = _rand(rng, (n, n), cplx)
    if n >= 64:
        a *= 4.0 / np.sqrt(n)
    c = _rand(rng, (degree + 1,), cplx) * 2.0
    want = power_series(a, c)
    tol = poly_tol(a, c, n)
    st, pbuf = run_ps(dev, a, c)
    assert st == 0
    got_ps = pbuf[:, :n].T
    assert np.abs(got_ps - want).max() <= tol
    # the oracle's own restatement of the reference's scheme agrees as tightly
    assert np.abs(got_ps - oracle.paterson_stockmeyer(c, a)).max() <= tol
    st, hbuf = run_horner(dev, a, c)
    assert st == 0
    got_h = hbuf[:, :n].T
    assert np.abs(got_h - want).max() <= tol
    assert np.abs(got_h - oracle.horner(c, a)).max() <= tol


@pytest.mark.parametrize("cplx", [False, True])
def test_ps_every_block_size_agrees(dev, cplx):
    """one polynomial, every block size from 1 (degenerate Horner) to d+1 (no
    Horner step) (paterson_stockmeyer_tests.cpp:432-459)"""
    n, degree = 5, 7
    rng = np.random.default_rng(77)
    a = _rand(rng, (n, n), cplx)
    c = _rand(rng, (degree + 1,), cplx) * 2.0
    want = power_series(a, c)
    tol = poly_tol(a, c, n)
    for s in range(1, degree + 2):
        st, buf = run_ps(dev, a, c, s_req=s)
        assert st == 0, s
        got = buf[:, :n].T
        assert np.abs(got - want).max() <= tol, s
        assert np.abs(got - oracle.paterson_stockmeyer(c, a, s)).max() <= tol, s


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("degree", [3, 6, 9])
def test_ps_matches_horner(dev, cplx, degree):
    n = 8
    rng = np.random.default_rng(90 + degree)
    a = _rand(rng, (n, n), cplx)
    c = _rand(rng, (degree + 1,), cplx) * 2.0
    _, ps = run_ps(dev, a, c)
    _, hn = run_horner(dev, a, c)
    assert np.abs(ps - hn).max() <= poly_tol(a, c, n)


@pytest.mark.parametrize("cplx", [False, True])
def test_zero_polynomial_is_exactly_zero(dev, cplx):
    n = 6
    rng = np.random.default_rng(40)
    a = _rand(rng, (n, n), cplx)
    for degree, runner in ((8, run_ps), (4, run_horner)):
        c = np.zeros(degree + 1, dtype=a.dtype)
        st, buf = runner(dev, a, c, sentinel=7.0)
        assert st == 0
        assert np.all(buf[:, :n] == 0.0)


@pytest.mark.parametrize("cplx", [False, True])
def test_nilpotent_series_terminates(dev, cplx):
    """strictly upper-triangular N: N^k = 0 for k >= n, so the degree-9 polynomial
    equals its degree-(n-1) truncation (paterson_stockmeyer_tests.cpp:500-530)"""
    n, degree = 5, 9
    a = np.zeros((n, n), dtype=np.complex128 if cplx else np.float64)
    for j in range(n):
        for i in range(j):
            a[i, j] = 0.3 * (i + j + 1) * ((1 + 0.5j) if cplx else 1)
    rng = np.random.default_rng(50)
    c = _rand(rng, (degree + 1,), cplx, -1.0, 1.0)
    truncated = power_series(a, c[:n])
    tol = poly_tol(a, c, n)
    for runner in (run_ps, run_horner):
        st, buf = runner(dev, a, c)
        assert st == 0
        got = buf[:, :n].T
        assert np.abs(got - truncated).max() <= tol
        assert np.abs(got - power_series(a, c)).max() <= tol


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("degree", [7, 8])
def test_padding_rows_untouched(dev, cplx, degree):
    """only the n x n block of P is written (opposite Horner parities for 7 and 8)"""
    n, ldp = 5, 8
    rng = np.random.default_rng(60 + degree)
    a = _rand(rng, (n, n), cplx)
    c = _rand(rng, (degree + 1,), cplx) * 2.0
    want = power_series(a, c)
    for runner in (run_ps, run_horner):
        st, buf = runner(dev, a, c, ldp=ldp)
        assert st == 0
        assert np.abs(buf[:, :n].T - want).max() <= poly_tol(a, c, n)
        assert np.all(buf[:, n:] == -123.5)


# ---- ?lange ------------------------------------------------------------------------------------
def _ramp(m, n, cplx):
    """the reference's mixed-sign ramp a(i) = (i % 19) - 9 (lange_tests.cpp:106-109)"""
    idx = np.arange(m * n).reshape(n, m).T  # column-major linear index
    a = (idx % 19 - 9).astype(np.float64)
    if cplx:
        a = a + 1j * ((idx % 7) - 3)
    return a


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("which", ["max_abs", "one", "inf", "frobenius"])
@pytest.mark.parametrize("shape,pad", [((1, 1), 0), ((1, 37), 2), ((37, 1), 0), ((7, 5), 3),
                                       ((300, 40), 1), ((40, 700), 5), ((1025, 513), 0)])
def test_lange_fp64(dev, cplx, which, shape, pad):
    m, n = shape
    ld = m + pad
    for a in (_ramp(m, n, cplx), _rand(np.random.default_rng(m * 7 + n), (m, n), cplx, -3, 3)):
        da = _cm(a, ld, dev, sentinel=1e6)
        out = torch.empty(1, dtype=torch.float64, device=dev)
        scratch = torch.empty(max(m, n), dtype=torch.float64, device=dev)
        code = {"max_abs": 0, "one": 1, "inf": 2, "frobenius": 3}[which]
        assert lange_cm(code, m, n, da, ld, out, scratch) == 0
        torch.cuda.synchronize()
        got, want = out.item(), oracle.lange(which, a)
        if cplx or which == "frobenius":
            assert abs(got - want) <= 8 * EPS * want
        else:
            assert got == want
        ref = {"max_abs": np.abs(a).max(), "one": np.abs(a).sum(0).max(),
               "inf": np.abs(a).sum(1).max(), "frobenius": np.linalg.norm(a)}[which]
        assert got == pytest.approx(ref, rel=1e-12)


def test_lange_fp64_empty_is_zero(dev):
    out = torch.full((1,), 5.0, dtype=torch.float64, device=dev)
    scratch = torch.empty(4, dtype=torch.float64, device=dev)
    buf = torch.zeros(4, dtype=torch.float64, device=dev)
    assert lange_cm(1, 0, 3, buf, 1, out, scratch) == 0
    torch.cuda.synchronize()
    assert out.item() == 0.0


# ---- ?lascl --------------------------------------------------------------------------------------
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("cfrom,cto", [(4.0, 1.0), (1.0, 0.125), (3.0, 7.0), (1e-300, 1e300),
                                       (1e300, 1e-300), (2.0 ** -1000, 2.0 ** 1000), (5.0, 0.0)])
def test_lascl_fp64_matches_oracle(dev, cplx, cfrom, cto):
    rng = np.random.default_rng(1)
    m, n, ld = 33, 17, 40
    a = _rand(rng, (m, n), cplx, -1, 1) * 1e-3
    da = _cm(a, ld, dev, sentinel=-777.0)
    assert lascl_cm(cfrom, cto, m, n, da, ld) == 0
    torch.cuda.synchronize()
    st, want = oracle.lascl(cfrom, cto, a)
    assert st == 0
    buf = da.cpu().numpy()
    assert np.array_equal(buf[:, :m].T, want, equal_nan=True)
    assert np.all(buf[:, m:] == -777.0)
    if cfrom in (4.0, 1.0) and cto in (1.0, 0.125):
        # a power-of-two ratio is exact: bitwise ldexp
        assert np.array_equal(buf[:, :m].T, a * (cto / cfrom))


@pytest.mark.parametrize("cplx", [False, True])
def test_lascl_fp64_rejects_bad_ratio(dev, cplx):
    a = np.ones((3, 3), dtype=np.complex128 if cplx else np.float64)
    da = _cm(a, 3, dev)
    for cfrom, cto in ((0.0, 1.0), (float("nan"), 1.0), (1.0, float("nan"))):
        assert lascl_cm(cfrom, cto, 3, 3, da, 3) == cg._lib.RUNTIME_INVALID_VALUE
