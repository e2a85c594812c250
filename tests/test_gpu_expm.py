"""GPU parity: clm_expm_{d,z} (Taylor / Paterson-Stockmeyer on MFMA) against the
oracle's restatement of the reference expm (Pade ladder + LU + squaring), the
committed golden vectors, and the reference's own closed forms and identities
(test/expm/expm_tests.cpp).

Bar: relative Frobenius error <= 1e-12 against the oracle on the same input
bits (the parity target of the port); closed forms and identities use the
reference's elementwise tolerance factor*eps*max(1, ||X||_F)*n
(test/expm/expm_tests.cpp:204-228).
"""
import json
import os

import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle
from calaman_g_amd import _lib
from calaman_g_amd.expm import expm_cm, matrix_norm1_cm

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PARITY = 1e-12


def rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


def scaled_gaussian(n, target, seed, cplx=False):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    if cplx:
        a = a + 1j * rng.standard_normal((n, n))
    return a * (target / np.abs(a).sum(0).max())


def run_expm(a, dev, lda=None, lde=None, sentinel=-123.5):
    """expm through the C ABI with explicit column-major buffers; returns
    (status, result n x n, full output buffer, plan)."""
    n = a.shape[0]
    lda = n if lda is None else lda
    lde = n if lde is None else lde
    dt = np.complex128 if np.iscomplexobj(a) else np.float64
    abuf = np.zeros((n, lda), dtype=dt)
    abuf[:, :n] = a.T
    ebuf = np.full((n, lde), sentinel, dtype=dt)
    da = torch.from_numpy(abuf).to(dev)
    de = torch.from_numpy(ebuf).to(dev)
    tdt = torch.complex128 if dt == np.complex128 else torch.float64
    work = torch.empty(cg.expm_buffer_size(n, tdt), dtype=torch.uint8, device=dev)
    info = torch.full((2,), -9, dtype=torch.int32, device=dev)
    st, plan = expm_cm(n, da, lda, de, lde, work, info)
    torch.cuda.synchronize()
    out = de.cpu().numpy()
    if st == 0:
        assert info.cpu().tolist() == [0, 0]
    return st, out[:, :n].T.copy(), out, plan


# ---- parity with the oracle -------------------------------------------------------------
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [1, 2, 5, 16, 33, 64, 128, 200, 256])
@pytest.mark.parametrize("target", [0.5, 4.0, 8.0, 100.0])
def test_expm_matches_oracle(dev, cplx, n, target):
    a = scaled_gaussian(n, target, seed=n * 100 + int(target * 10), cplx=cplx)
    st, got, _, plan = run_expm(a, dev)
    assert st == 0
    ref, pade_plan = oracle.expm(a, return_plan=True)
    err = rel(got, ref)
    assert err <= PARITY, f"n={n} |A|={target} plan={plan} err={err:.3e}"
    # the executed plan is the rule applied to the norms of A .. A^q the device took
    norms = [plan["norm1"]] + plan["power_norm1"]
    tp = oracle.taylor_plan_powers(norms)
    assert (plan["m"], plan["s"], plan["num_gemms"], plan["ps_block"]) == (
        tp["degree"], tp["s"], tp["num_gemms"], tp["ps_block"])
    # the powers whose norms are taken: A^2, A^3 when the norm-only plan's bank holds
    # A^3, then one more while a zero norm for it would lower the product count
    p0 = oracle.taylor_plan(plan["norm1"])
    q0 = oracle.ps_plan(p0["degree"])["num_powers"] + 1
    q = 1 if q0 < 3 else 3
    while 3 <= q < 5 and oracle.taylor_plan_powers(norms[:q] + [0.0])["num_gemms"] < \
            oracle.taylor_plan_powers(norms[:q])["num_gemms"]:
        q += 1
    assert plan["norm_powers"] == q
    # each power formed past A^3 can waste at most the one product it took
    assert plan["num_gemms"] <= p0["num_gemms"] + max(0, q - 3)
    # and agrees with the CPU statement of the same algorithm more tightly
    tay, tplan = oracle.expm_taylor(a, return_plan=True)
    assert rel(got, tay) <= 1e-13
    if not cplx:  # the FP64 products are the oracle's bit for bit: so are the norms
        assert norms == [plan["norm1"]] + tplan["power_norm1"]
        assert (plan["m"], plan["s"]) == (tplan["degree"], tplan["s"])


@pytest.mark.parametrize("cplx", [False, True])
def test_expm_2048_matches_oracle(dev, cplx):
    n = 2048
    a = scaled_gaussian(n, 8.0, seed=2048, cplx=cplx)
    st, got, _, plan = run_expm(a, dev)
    assert st == 0
    ref = oracle.expm(a)
    assert rel(got, ref) <= PARITY


def test_norm1_bit_exact(dev):
    for n in (1, 7, 255, 256, 257, 1000):
        a = scaled_gaussian(n, 3.0, seed=n)
        da = torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
        colsum = torch.empty(n, dtype=torch.float64, device=dev)
        assert matrix_norm1_cm(n, da, n, colsum) == oracle.norm1(a)


def test_norm1_complex_true_modulus(dev):
    n = 300
    a = scaled_gaussian(n, 3.0, seed=1, cplx=True)
    da = torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
    colsum = torch.empty(n, dtype=torch.float64, device=dev)
    got = matrix_norm1_cm(n, da, n, colsum)
    assert got == pytest.approx(oracle.norm1(a), rel=1e-15)
    assert got == pytest.approx(np.abs(a).sum(0).max(), rel=1e-13)


# ---- golden vectors ----------------------------------------------------------------------
def _golden_cases():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["expm"]


@pytest.mark.parametrize("case", _golden_cases(), ids=lambda c: f"{c['key']}-{c['kind']}-n{c['n']}-{c['norm1']}")
def test_expm_golden(dev, case):
    data = np.load(os.path.join(GOLDEN, case["file"]))
    a, e = data[case["key"] + "_A"], data[case["key"] + "_E"]
    st, got, _, _ = run_expm(a, dev)
    assert st == 0
    ref = oracle.expm(a)
    # the oracle's own distance from the golden value bounds what parity can mean
    budget = max(PARITY, 4 * rel(ref, e))
    assert rel(got, e) <= budget
    assert rel(got, ref) <= PARITY


# ---- closed forms and identities (test/expm/expm_tests.cpp:471-655) ---------------------------
def ref_tol(oracle_val, n, factor):
    eps = np.finfo(float).eps
    return factor * eps * max(1.0, np.linalg.norm(oracle_val)) * n


@pytest.mark.parametrize("cplx", [False, True])
def test_zero_is_identity(dev, cplx):
    n = 6
    a = np.zeros((n, n), dtype=np.complex128 if cplx else np.float64)
    st, got, _, plan = run_expm(a, dev, sentinel=-7.0)
    assert st == 0
    np.testing.assert_array_equal(got, np.eye(n))


@pytest.mark.parametrize("cplx", [False, True])
def test_diagonal(dev, cplx):
    n = 6
    rng = np.random.default_rng(2)
    d = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if cplx else 0)
    st, got, _, _ = run_expm(np.diag(d), dev)
    assert st == 0
    ref = np.diag(np.exp(d))
    assert np.abs(got - ref).max() <= ref_tol(ref, n, 128)


@pytest.mark.parametrize("cplx", [False, True])
def test_nilpotent(dev, cplx):
    n = 5
    rng = np.random.default_rng(10)
    N = np.triu(rng.uniform(-0.5, 0.5, (n, n)), 1)
    if cplx:
        N = N + 1j * np.triu(rng.uniform(-0.5, 0.5, (n, n)), 1)
    ref = np.eye(n, dtype=N.dtype)
    p = np.eye(n, dtype=N.dtype)
    fact = 1.0
    for k in range(1, n):
        p = p @ N
        fact *= k
        ref = ref + p / fact
    st, got, _, _ = run_expm(N, dev)
    assert st == 0
    assert np.abs(got - ref).max() <= ref_tol(ref, n, 128)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("target", [1.5, 30.0])
def test_inverse_identity(dev, cplx, target):
    n = 24
    a = scaled_gaussian(n, target, seed=21, cplx=cplx)
    _, e1, _, _ = run_expm(a, dev)
    _, e2, _, _ = run_expm(-a, dev)
    prod = e1 @ e2
    cond = np.linalg.norm(e1) * np.linalg.norm(e2)
    assert np.abs(prod - np.eye(n)).max() <= ref_tol(np.eye(n), n, 256) * cond


@pytest.mark.parametrize("cplx", [False, True])
def test_half_squared(dev, cplx):
    n = 24
    a = scaled_gaussian(n, 3.0, seed=30, cplx=cplx)
    _, ea, _, _ = run_expm(a, dev)
    _, eh, _, _ = run_expm(a / 2, dev)
    sq = eh @ eh
    assert np.abs(ea - sq).max() <= ref_tol(sq, n, 256)


def test_padded_output_leading_dimension(dev):
    n, lde = 5, 8
    a = scaled_gaussian(n, 1.5, seed=60)
    st, got, full, _ = run_expm(a, dev, lde=lde)
    assert st == 0
    assert np.all(full[:, n:] == -123.5)
    assert rel(got, oracle.expm(a)) <= PARITY


def test_padded_input_leading_dimension(dev):
    n, lda = 37, 41
    a = scaled_gaussian(n, 8.0, seed=61, cplx=True)
    st, got, _, _ = run_expm(a, dev, lda=lda)
    assert st == 0
    assert rel(got, oracle.expm(a)) <= PARITY


# ---- failure paths (src/expm/expm.cppm:94-124) -------------------------------------------------
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_non_finite_norm_rejected(dev, bad):
    a = scaled_gaussian(8, 1.0, seed=3)
    a[2, 5] = bad
    st, _, _, _ = run_expm(a, dev)
    assert st == _lib.INVALID_VALUE


def test_undersized_workspace(dev):
    n = 16
    da = torch.zeros((n, n), dtype=torch.float64, device=dev)
    de = torch.zeros((n, n), dtype=torch.float64, device=dev)
    need = cg.expm_buffer_size(n)
    work = torch.empty(need - 1, dtype=torch.uint8, device=dev)
    st, _ = expm_cm(n, da, n, de, n, work)
    assert st == _lib.ALLOC_FAILED


def test_convenience_wrappers_on_a_side_stream(dev):
    """expm / expm_split / paterson_stockmeyer / horner / gemm / lange / lascl with a
    stream other than the current one: the input is made on the current stream,
    the wrapper orders its work after it and hands the result back; every result
    equals the same call on the current stream, over repeated rounds"""
    rng = np.random.default_rng(21)
    n = 300
    host = scaled_gaussian(n, 6.0, seed=21)
    coef = torch.from_numpy(rng.standard_normal(9)).to(dev)
    side = torch.cuda.Stream()
    base = torch.from_numpy(host).to(dev)
    want = {
        "expm": cg.expm(base), "split": cg.expm_split(base),
        "ps": cg.paterson_stockmeyer(coef, base), "horner": cg.horner(coef, base),
        "gemm": cg.gemm(base, base), "lange": cg.lange("one", base),
        "lascl": cg.lascl(4.0, 1.0, base)}
    for _ in range(3):
        a = base * 1.0  # made on the current stream just before the call
        got = {
            "expm": cg.expm(a, stream=side), "split": cg.expm_split(a, stream=side),
            "ps": cg.paterson_stockmeyer(coef, a, stream=side),
            "horner": cg.horner(coef, a, stream=side), "gemm": cg.gemm(a, a, stream=side),
            "lange": cg.lange("one", a, stream=side), "lascl": cg.lascl(4.0, 1.0, a, stream=side)}
        for k in want:
            assert torch.equal(got[k], want[k]), k
