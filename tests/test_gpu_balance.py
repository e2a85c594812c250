"""GPU: balancing (clm_gebal_*), the balanced exponential (clm_expm_balanced_*) and the
1-norm estimator (clm_lacn2_*) against the oracle (oracle/balance.py), in the shape of
the reference's own suites (test/gebal/gebal_tests.cpp: exact reproduction, the
analytic 2x2, idempotence, the structural cases, all four element types;
test/lacn2: the estimate and its vector).

* gebal, FP64 and complex128: B, ilo, ihi and the scale array equal the oracle's bit
  for bit (B is A under exact power-of-two scalings and exchanges), for every job,
  with a leading dimension larger than n, on dense, sparse (isolating) and badly
  scaled matrices up to n = 1000. FP32 / complex64: the specification (exact
  reproduction, stationarity, idempotence), as the reference tests them.
* expm_balanced: D exp(B) D^-1 within 1e-13 of the oracle's exponential of the same B
  (normwise at the level of B), and fewer squarings than the unbalanced plan on a
  badly scaled matrix.
* lacn2, driven by torch products on the device: the same kase sequence and the
  same estimate (1e-13 relative) as the oracle; exact on non-negative matrices.
"""
import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle
import oracle.balance as ob

pytestmark = pytest.mark.gpu

CPLX = {torch.float32: False, torch.float64: False, torch.complex64: True, torch.complex128: True}
SINGLE = {torch.float32: True, torch.float64: False, torch.complex64: True, torch.complex128: False}
JOBS = {"N": ob.NONE, "P": ob.PERMUTE, "S": ob.SCALE, "B": ob.BOTH}


def cm(a, dtype, dev, ld):
    n_rows, n_cols = a.shape
    buf = np.zeros((n_cols, ld), dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    buf[:, :n_rows] = a.T
    return torch.from_numpy(buf).to(dtype).to(dev)


def from_cm(t, n_rows):
    return t.cpu().numpy()[:, :n_rows].T.astype(np.complex128 if t.is_complex() else np.float64)


def make(kind, n, seed, cplx):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    if cplx:
        a = a + 1j * rng.standard_normal((n, n))
    if kind in ("scaled", "sparse"):
        e = rng.integers(-40, 41, n).astype(float)
        a = a * (2.0 ** e)[:, None] / (2.0 ** e)[None, :]
    if kind == "sparse":
        a[rng.random((n, n)) < 0.9] = 0
        # a few rows and columns that isolate an eigenvalue
        for i in rng.choice(n, size=min(n, 3), replace=False):
            a[i, :] = 0
            a[i, i] = 1.5
    return a


def gpu_gebal(a, dtype, dev, job, ld=None):
    n = a.shape[0]
    ld = n + 3 if ld is None else ld
    da = cm(a, dtype, dev, max(ld, 1)) if n else torch.zeros(1, dtype=dtype, device=dev)
    real = torch.float32 if SINGLE[dtype] else torch.float64
    scale = torch.full((max(n, 1),), -7.0, dtype=real, device=dev)
    work = torch.zeros(max(cg.gebal_buffer_size(n), 1), dtype=torch.int32, device=dev)
    print(f"gebal n={n} dtype={dtype} job={job}", flush=True)
    st, ilo, ihi = cg.gebal_cm(job, n, da, max(ld, 1), scale, work)
    assert st == 0
    torch.cuda.synchronize()
    b = from_cm(da, n) if n else np.zeros((0, 0))
    pad = da.cpu().numpy()[:, n:] if n else None
    return b, ilo, ihi, scale.cpu().numpy()[:n].astype(np.float64), pad


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("job", ["N", "P", "S", "B"])
@pytest.mark.parametrize("kind,n", [("dense", 1), ("dense", 2), ("scaled", 9), ("sparse", 9),
                                    ("sparse", 64), ("scaled", 257), ("sparse", 300),
                                    ("scaled", 1000)])
def test_gebal_matches_oracle(dev, dtype, job, kind, n):
    a = make(kind, n, 17 * n + len(kind), CPLX[dtype])
    b, ilo, ihi, d, pad = gpu_gebal(a, dtype, dev, job)
    ob_b, o_ilo, o_ihi, o_d = ob.gebal(a, JOBS[job])
    assert (ilo, ihi) == (o_ilo, o_ihi)
    np.testing.assert_array_equal(d, o_d)
    np.testing.assert_array_equal(b, ob_b)
    np.testing.assert_array_equal(b, ob.apply_balance_similarity(a, ilo, ihi, d))
    assert np.all(pad == 0)  # the rows past n of each column are untouched


def test_gebal_analytic_2x2_and_idempotence(dev):
    a = np.array([[1.0, 2.0 ** 10], [2.0 ** -10, 1.0]])
    for dtype in (torch.float64, torch.float32, torch.complex128, torch.complex64):
        b, ilo, ihi, d, _ = gpu_gebal(a, dtype, dev, "S")
        assert (ilo, ihi) == (1, 2)
        np.testing.assert_array_equal(d, [2.0 ** 10, 1.0])
        np.testing.assert_array_equal(b, np.ones((2, 2)))
        b2, _, _, d2, _ = gpu_gebal(b, dtype, dev, "S")
        assert np.all(d2 == 1.0) and np.array_equal(b2, b)


def test_gebal_structural(dev):
    b, ilo, ihi, d, _ = gpu_gebal(np.diag([3.0, -1.0, 2.0, 7.0]), torch.float64, dev, "B")
    assert (ilo, ihi) == (1, 1)
    b, ilo, ihi, d, _ = gpu_gebal(np.array([[5.0]]), torch.float64, dev, "B")
    assert (ilo, ihi) == (1, 1) and d[0] == 1.0
    b, ilo, ihi, d, _ = gpu_gebal(np.zeros((0, 0)), torch.float64, dev, "B")
    assert (ilo, ihi) == (1, 0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64])
@pytest.mark.parametrize("kind,n", [("scaled", 40), ("sparse", 120)])
def test_gebal_single_specification(dev, dtype, kind, n):
    a = make(kind, n, 5 + n, CPLX[dtype])
    a = a.astype(np.complex64 if CPLX[dtype] else np.float32).astype(
        np.complex128 if CPLX[dtype] else np.float64)  # representable in single
    b, ilo, ihi, d, _ = gpu_gebal(a, dtype, dev, "B")
    assert np.all(np.frexp(d[ilo - 1:ihi])[0] == 0.5)
    np.testing.assert_array_equal(b, ob.apply_balance_similarity(a, ilo, ihi, d))
    # idempotence: the same job again finds the same window and scales nothing
    b2, ilo2, ihi2, d2, _ = gpu_gebal(b, dtype, dev, "B")
    assert (ilo2, ihi2) == (ilo, ihi)
    assert np.all(d2[ilo - 1:ihi] == 1.0)
    np.testing.assert_array_equal(b2, b)
    assert np.abs(b).sum(0).max() <= np.abs(a).sum(0).max()


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n", [1, 30, 300])
def test_expm_balanced_matches_oracle(dev, dtype, n):
    a = make("scaled", n, 3 * n, CPLX[dtype]) * 0.5
    print(f"expm_balanced n={n} dtype={dtype}", flush=True)
    e, plan = cg.expm_balanced(torch.from_numpy(a).to(dtype).to(dev), return_plan=True)
    torch.cuda.synchronize()
    e = e.cpu().numpy().astype(np.complex128 if CPLX[dtype] else np.float64)
    b, ilo, ihi, d = ob.gebal(a, ob.SCALE)
    eb = oracle.expm(b)
    dd = np.ones(n)
    dd[ilo - 1:ihi] = d[ilo - 1:ihi]
    # the error at the level of the balanced matrix: D^-1 (E - E_ref) D against exp(B)
    err = (e - dd[:, None] * eb / dd[None, :]) / dd[:, None] * dd[None, :]
    assert np.abs(err).sum(0).max() <= 1e-13 * max(np.abs(eb).sum(0).max(), 1.0)
    assert plan["norm1"] == pytest.approx(oracle.norm1(b), rel=1e-14)
    if n > 1:  # the balanced plan squares less than the unbalanced one
        assert plan["s"] < cg.expm_plan(oracle.norm1(a))["s"]


def test_expm_balanced_single(dev):
    # the balancing in FP32 is pinned by its specification above; here the recipe
    # around it: D exp(B) D^-1 with the B and D the device balancing gives
    n = 64
    a = make("scaled", n, 9, False) * 0.5
    ta = torch.from_numpy(a).float().to(dev)
    e = cg.expm_balanced(ta).cpu().numpy().astype(np.float64)
    b, ilo, ihi, d = cg.gebal(ta, "S")
    b = b.cpu().numpy().astype(np.float64)
    d = d.cpu().numpy().astype(np.float64)
    assert (ilo, ihi) == (1, n)
    eb = oracle.expm(b)
    err = (e - d[:, None] * eb / d[None, :]) / d[:, None] * d[None, :]
    assert np.abs(err).sum(0).max() <= 2e-5 * np.abs(eb).sum(0).max()


def _drive(a, dtype, dev):
    t = torch.from_numpy(a).to(dtype).to(dev)
    tt = t.t().contiguous()
    return cg.norm1_estimate(lambda x: t @ x, lambda x: tt @ x, a.shape[0], dtype, dev)


@pytest.mark.parametrize("n", [1, 2, 5, 10, 40, 100, 1000])
def test_lacn2_matches_oracle(dev, n):
    rng = np.random.default_rng(40 + n)
    for trial in range(3):
        a = rng.standard_normal((n, n))
        if trial == 2 and n > 1:
            a = np.linalg.inv(a)  # the dgecon use: an inverse
        print(f"lacn2 n={n} trial={trial}", flush=True)
        est, v, kases = _drive(a, torch.float64, dev)
        o_est, o_v, o_seq = ob.lacn2(a)
        assert kases == o_seq
        assert est == pytest.approx(o_est, rel=1e-13)
        np.testing.assert_allclose(v.cpu().numpy(), o_v, rtol=1e-12, atol=1e-12 * np.abs(o_v).max())


@pytest.mark.parametrize("n", [2, 64, 513])
def test_lacn2_exact_on_nonnegative(dev, n):
    a = np.random.default_rng(n).random((n, n))
    est, v, kases = _drive(a, torch.float64, dev)
    assert kases == [1, 2, 1, 1]
    assert est == pytest.approx(a.sum(0).max(), rel=1e-14)
    est32, _, kases32 = _drive(a, torch.float32, dev)
    assert kases32 == [1, 2, 1, 1]
    assert est32 == pytest.approx(a.sum(0).max(), rel=1e-5)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_gebal_non_finite_and_sweep_cap(dev, dtype):
    """an index whose norms are not finite is skipped (src/lapack/gebal/README.md
    "Differences"), exactly as the oracle skips it; max_sweeps caps the sweeps as the
    oracle's cap does"""
    n = 40
    a = make("scaled", n, 77, CPLX[dtype])
    a[3, 7] = np.inf
    a[11, 2] = np.nan
    real = torch.float64
    for sweeps in (1, 2, 0):
        da = cm(a, dtype, dev, n + 3)
        scale = torch.empty(n, dtype=real, device=dev)
        work = torch.zeros(cg.gebal_buffer_size(n), dtype=torch.int32, device=dev)
        st, ilo, ihi = cg.gebal_cm("B", n, da, n + 3, scale, work, sweeps)
        assert st == 0
        torch.cuda.synchronize()
        ob_b, o_ilo, o_ihi, o_d = ob.gebal(a, ob.BOTH, max_sweeps=sweeps if sweeps > 0 else 1000)
        assert (ilo, ihi) == (o_ilo, o_ihi)
        np.testing.assert_array_equal(scale.cpu().numpy(), o_d)
        np.testing.assert_array_equal(from_cm(da, n), ob_b)


def test_lacn2_single(dev):
    n = 100
    a = np.random.default_rng(5).standard_normal((n, n)).astype(np.float32).astype(np.float64)
    est, v, kases = _drive(a, torch.float32, dev)
    o_est, o_v, o_seq = ob.lacn2(a)
    assert kases == o_seq
    assert est == pytest.approx(o_est, rel=1e-5)
