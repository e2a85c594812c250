"""GPU: the FP8 product (clm_gemm_fp8_d) and the FP8-tail expm (clm_expm_fp8tail_d).

The product's bar is the oracle's quantized product oracle.gemm_fp8 (e4m3
rounding pinned against torch.float8_e4m3fn in test_oracle_fp8.py): exact where
every partial sum is an exact small integer multiple, elsewhere within
2^-13 sum_k |Q(a)||Q(b)|. That bound is not the FP32 one: the FP8 MFMA sums its
32-deep dot products at reduced precision -- measured on MI355X, up to
4.8e-5 (about 2^-14.3) of sum_k |Q(a)||Q(b)| for a single K = 32 product of
positive operands, where an FP32 sum would stay below 2^-19 -- and the bound
leaves a factor 2.5 above that. It is still far below the e4m3 rounding the
product exists to carry (2^-4 per operand).

The expm's bar is oracle.expm_fp8tail_model at the plan the device reports:
the same products quantized the same way, everything else exact. The device
differs from the model by the split products' and FP32 sums' rounding (and the
rare e4m3 rounding those move across a midpoint), which is held to a tenth of
the model's own FP8 error. The error against the FP64 oracle is printed.
"""
import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle
from calaman_g_amd import _lib

pytestmark = pytest.mark.gpu


def _cm(a, ld, dev, fill=-7.0):
    buf = np.full((a.shape[1], ld), fill)
    buf[:, :a.shape[0]] = a.T
    return torch.from_numpy(buf).to(dev)


def _gemm_fp8(dev, a, b, alpha=1.0, beta=0.0, c=None, pad=(0, 0, 0)):
    m, k = a.shape
    n = b.shape[1]
    lda, ldb, ldc = max(m, 1) + pad[0], max(k, 1) + pad[1], max(m, 1) + pad[2]
    da, db = _cm(a, lda, dev), _cm(b, ldb, dev)
    dc = _cm(c if c is not None else np.zeros((m, n)), ldc, dev)
    work = torch.empty(max(cg.gemm_fp8_buffer_size(m, n, k), 1), dtype=torch.uint8, device=dev)
    st = cg.gemm_fp8_cm(None, m, n, k, alpha, da, lda, db, ldb, beta, dc, ldc, work)
    assert st == 0
    torch.cuda.synchronize()
    out = dc.cpu().numpy()
    assert np.all(out[:, m:] == -7.0)  # padding rows untouched
    return out[:, :m].T.copy()


@pytest.mark.parametrize("m,n,k", [(128, 128, 128), (256, 384, 512), (100, 70, 300), (1, 1, 1),
                                   (129, 257, 130)])
def test_gemm_fp8_exact_on_representable_operands(dev, m, n, k):
    """integers |v| <= 15 at one power-of-two scale per row tile of A and per
    column tile of B: e4m3 holds them exactly and every sum is an exact FP32
    integer, so the product is exactly the float64 one"""
    rng = np.random.default_rng(m + n + k)
    a = rng.integers(-15, 16, (m, k)).astype(np.float64)
    b = rng.integers(-15, 16, (k, n)).astype(np.float64)
    a *= np.exp2(rng.integers(-30, 30, ((m + 127) // 128, 1))).repeat(128, 0)[:m]
    b *= np.exp2(rng.integers(-30, 30, (1, (n + 127) // 128))).repeat(128, 1)[:, :n]
    c = rng.integers(-4, 5, (m, n)).astype(np.float64)
    got = _gemm_fp8(dev, a, b, 0.5, 0.25, c, pad=(1, 2, 3))
    assert np.array_equal(got, 0.5 * (a @ b) + 0.25 * c)


@pytest.mark.parametrize("m,n,k", [(256, 256, 1024), (300, 200, 333), (128, 512, 128)])
def test_gemm_fp8_matches_quantized_product(dev, m, n, k):
    """Gaussian operands whose 128 x 128 tiles span 2^-8 .. 2^8 in magnitude"""
    rng = np.random.default_rng(7 * m + k)
    a = rng.standard_normal((m, k)) * np.exp2(rng.integers(-8, 9, (-(-m // 128), -(-k // 128)))) \
        .repeat(128, 0).repeat(128, 1)[:m, :k]
    b = rng.standard_normal((k, n)) * np.exp2(rng.integers(-8, 9, (-(-k // 128), -(-n // 128)))) \
        .repeat(128, 0).repeat(128, 1)[:k, :n]
    got = _gemm_fp8(dev, a, b)
    qa, qb = oracle.fp8_quantize(a)[0], oracle.fp8_quantize(b)[0]
    want = qa @ qb
    mag = np.abs(qa) @ np.abs(qb)
    assert np.all(np.abs(got - want) <= 2.0 ** -13 * mag)
    # and the quantization error itself is of e4m3 size, far above that bound
    assert np.median(np.abs(want - a @ b) / mag) > 10 * 2.0 ** -13


def test_gemm_fp8_zero_and_empty(dev):
    got = _gemm_fp8(dev, np.zeros((130, 40)), np.zeros((40, 5)))
    assert np.array_equal(got, np.zeros((130, 5)))
    c = np.arange(12.0).reshape(4, 3)
    got = _gemm_fp8(dev, np.zeros((4, 0)), np.zeros((0, 3)), 1.0, 2.0, c)  # k = 0: C = beta C
    assert np.array_equal(got, 2.0 * c)


def _mat(rng, n, norm):
    a = rng.standard_normal((n, n)) / np.sqrt(n)
    return a * (norm / np.abs(a).sum(0).max())


def _fp8tail(dev, a, **kw):
    n = a.shape[0]
    da = _cm(a, n + 1, dev)
    de = _cm(np.zeros_like(a), n + 2, dev)
    work = torch.empty(cg.expm_fp8tail_buffer_size(n), dtype=torch.uint8, device=dev)
    st, plan = cg.expm_fp8tail_cm(n, da, n + 1, de, n + 2, work, **kw)
    assert st == 0
    torch.cuda.synchronize()
    out = de.cpu().numpy()
    assert np.all(out[:, n:] == -7.0)
    return out[:, :n].T.copy(), plan


def _expected_fp8_products(degree, min_degree):
    p = oracle.ps_plan(degree, 0)
    b, r = p["s"], p["r"]
    max_power = b if r >= 1 else degree
    return sum(1 for k in range(2, max_power + 1) if k >= min_degree) + \
        sum(1 for j in range(r) if (j + 1) * b >= min_degree)


@pytest.mark.parametrize("n,norm", [(256, 1.0), (256, 16.0), (300, 256.0), (384, 16.0)])
def test_expm_fp8tail_against_model(dev, n, norm):
    rng = np.random.default_rng(n + int(norm))
    a = _mat(rng, n, norm)
    got, plan = _fp8tail(dev, a, ps_split=_lib.SPLIT_BF16X6)
    assert plan["fp8_min_degree"] == 4
    assert plan["fp8_gemms"] == _expected_fp8_products(plan["m"], 4) > 0
    model = oracle.expm_fp8tail_model(a, plan["m"], plan["s"])
    ref = oracle.expm(a)
    nrm = np.linalg.norm(ref)
    model_err = np.linalg.norm(model - ref) / nrm
    dev_err = np.linalg.norm(got - ref) / nrm
    gap = np.linalg.norm(got - model) / nrm
    print(f"\nfp8tail n={n} |A|_1={norm} plan=(m={plan['m']}, s={plan['s']}, "
          f"fp8 products={plan['fp8_gemms']}) err vs FP64 oracle {dev_err:.3e} "
          f"(model {model_err:.3e}, device-model {gap:.3e})")
    assert gap <= max(0.1 * model_err, 2e-6)


def test_expm_fp8tail_without_fp8_products_is_the_split_variant(dev):
    """fp8_min_degree above the degree leaves every product on the split planes:
    the same kernels on the same operands as clm_expm_split_d, the same bits"""
    rng = np.random.default_rng(11)
    n = 200
    a = _mat(rng, n, 20.0)
    got, plan = _fp8tail(dev, a, fp8_min_degree=64)
    assert plan["fp8_gemms"] == 0
    da = _cm(a, n, dev)
    de = torch.empty_like(da)
    work = torch.empty(cg.expm_split_buffer_size(n), dtype=torch.uint8, device=dev)
    # the FP8 tail plans from ||A||_1 alone: so does this split call
    st, splan = cg.expm_split_cm(n, da, n, de, n, work, plan_from=_lib.PLAN_NORM1)
    assert st == 0 and (splan["m"], splan["s"]) == (plan["m"], plan["s"])
    torch.cuda.synchronize()
    assert np.array_equal(got, de.cpu().numpy().T)


@pytest.mark.parametrize("md", [2, 3, 8])
def test_expm_fp8tail_min_degree(dev, md):
    """a lower threshold puts more products on FP8 and the error grows; every
    count matches the rule"""
    rng = np.random.default_rng(12)
    a = _mat(rng, 256, 4.0)
    got, plan = _fp8tail(dev, a, fp8_min_degree=md, ps_split=_lib.SPLIT_BF16X6)
    assert plan["fp8_gemms"] == _expected_fp8_products(plan["m"], md)
    model = oracle.expm_fp8tail_model(a, plan["m"], plan["s"], min_degree=md)
    ref = oracle.expm(a)
    model_err = np.linalg.norm(model - ref) / np.linalg.norm(ref)
    assert np.linalg.norm(got - model) / np.linalg.norm(ref) <= max(0.1 * model_err, 2e-6)


def test_expm_fp8tail_rejects_bad_arguments(dev):
    a = torch.zeros((8, 8), dtype=torch.float64, device=dev)
    work = torch.empty(cg.expm_fp8tail_buffer_size(8), dtype=torch.uint8, device=dev)
    e = torch.empty_like(a)
    assert cg.expm_fp8tail_cm(8, a, 8, e, 8, work, fp8_min_degree=1)[0] == _lib.INVALID_VALUE
    assert cg.expm_fp8tail_cm(8, a, 8, e, 8, work, ps_split=5)[0] == _lib.INVALID_VALUE
    small = torch.empty(cg.expm_split_buffer_size(8), dtype=torch.uint8, device=dev)
    assert cg.expm_fp8tail_cm(8, a, 8, e, 8, small)[0] == _lib.ALLOC_FAILED
    ok, plan = cg.expm_fp8tail_cm(8, a, 8, e, 8, work)
    assert ok == 0
    torch.cuda.synchronize()
    assert torch.equal(e, torch.eye(8, dtype=torch.float64, device=dev))
