"""GPU: the single-precision forms -- float and complex<float> -- of every entry
point: clm_gemm_{s,c}, clm_expm_{s,c}, clm_matrix_norm1_{s,c}, clm_lange_{s,c},
clm_lascl_{s,c}, clm_paterson_stockmeyer_{s,c}, clm_horner_{s,c}.

Bars:
  * products: 8*eps32*k*max|A||B| elementwise against the float64 product of the
    same float32 inputs (the k-term dot-product bound);
  * expm: relative Frobenius error against the FP64 oracle on the same float32
    input bits, at most 4x that of torch.linalg.matrix_exp in float32 (a plain
    PyTorch FP32 reference of the same op) or 32*eps32, whichever is larger
    (for 1 x 1, which torch evaluates with the scalar exp, the scheme's a-priori
    forward bound instead);
    closed forms with the reference's elementwise tolerance
    factor*eps*max(1, ||X||_F)*n (test/expm/expm_tests.cpp:204-228);
  * the 1-norm, ?lange and ?lascl of a real matrix: bit-for-bit the oracle's
    float restatement (same fold order, float accumulation); complex moduli to
    a few ulps (the device hypotf and the host's may round differently).
"""
import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle
from calaman_g_amd.expm import (Context, expm_cm, gemm_cm, horner_cm, lange_cm, lascl_cm,
                                matrix_norm1_cm, paterson_stockmeyer_cm)

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


def _rand32(rng, shape, cplx):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return a.astype(np.complex64 if cplx else np.float32)


def _cm(a, ld, dev, sentinel=-777.0):
    """device buffer holding a column-major with leading dimension ld"""
    m, n = a.shape
    buf = np.full((max(n, 1), ld), sentinel, dtype=a.dtype)
    buf[:n, :m] = a.T
    return torch.from_numpy(buf).to(dev)


def _from_cm(t, m, n):
    return t.cpu().numpy()[:n, :m].T.copy()


def _tdt(cplx):
    return torch.complex64 if cplx else torch.float32


# ---- products ---------------------------------------------------------------------------
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (16, 16, 16), (37, 53, 29), (64, 64, 64),
                                   (130, 127, 131), (256, 256, 256), (1024, 1024, 1024),
                                   (1040, 1030, 1027), (2048, 2048, 2048)])
def test_gemm_single_matches_float64_product(dev, cplx, m, n, k):
    rng = np.random.default_rng(m * 11 + n * 5 + k)
    a, b = _rand32(rng, (m, k), cplx), _rand32(rng, (k, n), cplx)
    ctx = Context()
    da, db = _cm(a, m, dev), _cm(b, k, dev)
    dc = _cm(np.zeros((m, n), dtype=a.dtype), m, dev)
    assert gemm_cm(ctx, m, n, k, 1.0, da, m, db, k, 0.0, dc, m) == 0
    torch.cuda.synchronize()
    got = _from_cm(dc, m, n).astype(np.complex128 if cplx else np.float64)
    exact = a.astype(np.complex128 if cplx else np.float64) @ b.astype(got.dtype)
    tol = 8 * EPS32 * k * np.abs(a).max() * np.abs(b).max()
    err = np.abs(got - exact).max()
    assert err <= tol, f"max err {err:.3e} > {tol:.3e}"


@pytest.mark.parametrize("cplx", [False, True])
def test_gemm_single_alpha_beta_padding(dev, cplx):
    rng = np.random.default_rng(5)
    m, n, k, ld = 70, 45, 33, 80
    a, b, c = _rand32(rng, (m, k), cplx), _rand32(rng, (k, n), cplx), _rand32(rng, (m, n), cplx)
    alpha, beta = (0.75 - 0.5j, 1.25 + 0.25j) if cplx else (0.75, -1.25)
    ctx = Context()
    da, db, dc = _cm(a, ld, dev), _cm(b, ld, dev), _cm(c, ld, dev)
    assert gemm_cm(ctx, m, n, k, alpha, da, ld, db, ld, beta, dc, ld) == 0
    torch.cuda.synchronize()
    full = dc.cpu().numpy()
    wide = np.complex128 if cplx else np.float64
    exact = alpha * (a.astype(wide) @ b.astype(wide)) + beta * c.astype(wide)
    got = full[:n, :m].T
    assert np.abs(got - exact).max() <= 16 * EPS32 * k * 4 * (np.abs(a).max() * np.abs(b).max() + 1)
    assert np.all(full[:n, m:] == np.float32(-777.0))  # padding rows untouched


# ---- expm -------------------------------------------------------------------------------
def _run_expm(a, dev, lde=None, sentinel=-123.5):
    n = a.shape[0]
    lde = n if lde is None else lde
    cplx = np.iscomplexobj(a)
    da = _cm(a, n, dev)
    de = torch.from_numpy(np.full((n, lde), sentinel, dtype=a.dtype)).to(dev)
    work = torch.empty(cg.expm_buffer_size(n, _tdt(cplx)), dtype=torch.uint8, device=dev)
    info = torch.full((2,), -9, dtype=torch.int32, device=dev)
    st, plan = expm_cm(n, da, n, de, lde, work, info)
    torch.cuda.synchronize()
    out = de.cpu().numpy()
    if st == 0:
        assert info.cpu().tolist() == [0, 0]
    return st, out[:, :n].T.copy(), out, plan


def _scaled32(n, target, seed, cplx):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    if cplx:
        a = a + 1j * rng.standard_normal((n, n))
    a = a * (target / np.abs(a).sum(0).max())
    return a.astype(np.complex64 if cplx else np.float32)


def _rel(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [1, 3, 16, 33, 100, 256])
@pytest.mark.parametrize("target", [0.1, 1.0, 4.0, 20.0, 100.0])
def test_expm_single_accuracy(dev, cplx, n, target):
    a = _scaled32(n, target, seed=n * 31 + int(target * 10), cplx=cplx)
    st, got, _, plan = _run_expm(a, dev)
    assert st == 0
    assert plan["dtype"] == (4 if cplx else 2)
    ref = oracle.expm(a.astype(np.complex128 if cplx else np.float64))
    err = _rel(got, ref)
    if n == 1:
        # torch evaluates a 1 x 1 matrix_exp with the scalar exp, so it is no
        # yardstick here; the a-priori forward bound of the scheme is: the
        # polynomial at z = a / 2^s errs by <= (m+1) u e^|z| against |e^z|, and
        # each of the s squarings doubles the relative error
        z = complex(a[0, 0]) / 2.0 ** plan["s"]
        bound = 2.0 ** plan["s"] * (plan["m"] + 1) * EPS32 * np.exp(abs(z) - z.real) + 4 * EPS32
        torch_err = float("nan")
    else:
        t = torch.from_numpy(a).to(dev)
        torch_err = _rel(torch.linalg.matrix_exp(t).cpu().numpy(), ref)
        bound = max(4 * torch_err, 32 * EPS32)
    assert err <= bound, f"n={n} |A|={target} plan={plan} err={err:.3e} torch32={torch_err:.3e}"
    # the plan is the single-precision ladder's for the norms of A .. A^q
    tp = cg.expm_plan(plan["norm1"], _tdt(cplx), plan["power_norm1"])
    assert (plan["m"], plan["s"], plan["num_gemms"]) == (tp["m"], tp["s"], tp["num_gemms"])


@pytest.mark.parametrize("cplx", [False, True])
def test_expm_single_2048(dev, cplx):
    """the large-tile kernels (2048 = 16 x 128) at the n = 2048 workload"""
    a = _scaled32(2048, 8.0, seed=2048, cplx=cplx)
    st, got, _, plan = _run_expm(a, dev)
    assert st == 0
    t = torch.from_numpy(a).to(dev)
    ref = torch.linalg.matrix_exp(t.to(torch.complex128 if cplx else torch.float64)).cpu().numpy()
    torch_err = _rel(torch.linalg.matrix_exp(t).cpu().numpy(), ref)
    err = _rel(got, ref)
    assert err <= max(4 * torch_err, 32 * EPS32), f"err={err:.3e} torch32={torch_err:.3e} {plan}"


@pytest.mark.parametrize("cplx", [False, True])
def test_expm_single_zero_is_identity_padded(dev, cplx):
    n, lde = 7, 11
    a = np.zeros((n, n), dtype=np.complex64 if cplx else np.float32)
    st, got, full, plan = _run_expm(a, dev, lde=lde)
    assert st == 0
    assert np.array_equal(got, np.eye(n))
    assert np.all(full[:, n:] == np.float32(-123.5))
    assert plan["s"] == 0


@pytest.mark.parametrize("cplx", [False, True])
def test_expm_single_diagonal(dev, cplx):
    n = 9
    d = np.linspace(-2.0, 1.5, n)
    if cplx:
        d = d + 1j * np.linspace(0.3, -1.1, n)
    a = np.diag(d).astype(np.complex64 if cplx else np.float32)
    st, got, _, _ = _run_expm(a, dev)
    assert st == 0
    want = np.diag(np.exp(a.diagonal().astype(np.complex128 if cplx else np.float64)))
    tol = 64 * EPS32 * max(1.0, np.linalg.norm(want)) * n
    assert np.abs(got - want).max() <= tol


@pytest.mark.parametrize("cplx", [False, True])
def test_expm_single_inverse_identity(dev, cplx):
    n = 48
    a = _scaled32(n, 3.0, seed=77, cplx=cplx)
    _, e1, _, _ = _run_expm(a, dev)
    _, e2, _, _ = _run_expm(-a, dev)
    wide = np.complex128 if cplx else np.float64
    prod = e1.astype(wide) @ e2.astype(wide)
    tol = 64 * EPS32 * max(1.0, np.linalg.norm(prod)) * n
    assert np.abs(prod - np.eye(n)).max() <= tol


@pytest.mark.parametrize("cplx", [False, True])
def test_expm_single_anti_hermitian_is_unitary(dev, cplx):
    """exp of a skew-symmetric / anti-Hermitian generator is orthogonal / unitary"""
    n = 64
    rng = np.random.default_rng(9)
    g = rng.standard_normal((n, n))
    if cplx:
        g = g + 1j * rng.standard_normal((n, n))
        a = (1j * 0.5 * (g + g.conj().T)).astype(np.complex64)
    else:
        a = (0.5 * (g - g.T)).astype(np.float32)
    _, e, _, _ = _run_expm(a, dev)
    wide = e.astype(np.complex128 if cplx else np.float64)
    dev_u = np.abs(wide.conj().T @ wide - np.eye(n)).max()
    assert dev_u <= 64 * EPS32 * n ** 0.5 * 8


def test_expm_single_rejects_non_finite(dev):
    a = np.ones((4, 4), dtype=np.float32)
    a[1, 2] = np.inf
    st, _, _, _ = _run_expm(a, dev)
    assert st == cg._lib.INVALID_VALUE


# ---- norms and scaling ---------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 7, 256, 257, 600])
def test_norm1_single_bit_exact(dev, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)).astype(np.float32)
    da = _cm(a, n, dev)
    colsum = torch.empty(n, dtype=torch.float32, device=dev)
    got = matrix_norm1_cm(n, da, n, colsum)
    assert np.float32(got) == np.float32(oracle.norm1(a))


def test_norm1_complex64_modulus(dev):
    n = 300
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    da = _cm(a, n, dev)
    colsum = torch.empty(n, dtype=torch.float32, device=dev)
    got = matrix_norm1_cm(n, da, n, colsum)
    want = oracle.norm1(a)
    assert abs(got - want) <= 8 * EPS32 * want
    assert abs(got - np.abs(a.astype(np.complex128)).sum(0).max()) <= 1e-5 * want


def test_norm1_single_nan_propagates(dev):
    a = np.ones((5, 5), dtype=np.float32)
    a[3, 1] = np.nan
    colsum = torch.empty(5, dtype=torch.float32, device=dev)
    assert np.isnan(matrix_norm1_cm(5, _cm(a, 5, dev), 5, colsum))


@pytest.mark.parametrize("which", ["max_abs", "one", "inf", "frobenius"])
@pytest.mark.parametrize("shape", [(1, 1), (5, 9), (300, 40), (40, 700)])
def test_lange_single(dev, which, shape):
    m, n = shape
    rng = np.random.default_rng(m + n)
    a = rng.standard_normal((m, n)).astype(np.float32)
    ld = m + 3
    da = _cm(a, ld, dev)
    out = torch.empty(1, dtype=torch.float32, device=dev)
    scratch = torch.empty(max(m, n), dtype=torch.float32, device=dev)
    code = {"max_abs": 0, "one": 1, "inf": 2, "frobenius": 3}[which]
    assert lange_cm(code, m, n, da, ld, out, scratch) == 0
    torch.cuda.synchronize()
    got, want = np.float32(out.item()), np.float32(oracle.lange(which, a))
    if which == "frobenius":  # x*x may be contracted differently into the sum
        assert abs(got - want) <= 4 * EPS32 * want
    else:
        assert got == want


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("cfrom,cto", [(4.0, 1.0), (3.0, 7.0), (1e-30, 1e30), (1e30, 1e-30)])
def test_lascl_single_matches_oracle(dev, cplx, cfrom, cto):
    rng = np.random.default_rng(1)
    m, n, ld = 33, 17, 40
    a = _rand32(rng, (m, n), cplx) * np.float32(1e-3)
    da = _cm(a, ld, dev)
    assert lascl_cm(np.float32(cfrom), np.float32(cto), m, n, da, ld) == 0
    torch.cuda.synchronize()
    st, want = oracle.lascl(float(np.float32(cfrom)), float(np.float32(cto)), a)
    assert st == 0
    got = _from_cm(da, m, n)
    assert np.array_equal(got, want, equal_nan=True)


# ---- polynomials ---------------------------------------------------------------------------------
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("degree,s_req", [(0, 0), (1, 0), (7, 0), (18, 0), (18, 3), (18, 19)])
def test_ps_and_horner_single(dev, cplx, degree, s_req):
    n = 40
    rng = np.random.default_rng(degree * 13 + s_req)
    a = (_rand32(rng, (n, n), cplx) * np.float32(0.6 / n ** 0.5))
    coef = _rand32(rng, (degree + 1,), cplx)
    wide = np.complex128 if cplx else np.float64
    want = oracle.paterson_stockmeyer(coef.astype(wide), a.astype(wide))
    ctx = Context()
    dcoef = torch.from_numpy(coef).to(dev)
    da = _cm(a, n, dev)
    tdt = _tdt(cplx)
    dp = torch.empty((n, n), dtype=tdt, device=dev)
    work = torch.empty(max(cg.ps_buffer_size(n, degree, s_req, tdt), 1), dtype=torch.uint8,
                       device=dev)
    assert paterson_stockmeyer_cm(ctx, n, dcoef, degree, da, n, dp, n, work, s_req) == 0
    hp = torch.empty((n, n), dtype=tdt, device=dev)
    hw = torch.empty(cg.horner_buffer_size(n, tdt), dtype=torch.uint8, device=dev)
    assert horner_cm(ctx, n, dcoef, degree, da, n, hp, n, hw) == 0
    torch.cuda.synchronize()
    # the reference's bound: 64 eps sum|c_k| ||A||_F^k n
    # (test/paterson_stockmeyer/paterson_stockmeyer_tests.cpp:125-135)
    fro = np.linalg.norm(a.astype(wide))
    bound = 64 * EPS32 * n * sum(abs(complex(c)) * fro ** k for k, c in enumerate(coef))
    for got in (_from_cm(dp, n, n), _from_cm(hp, n, n)):
        assert np.abs(got - want).max() <= bound
