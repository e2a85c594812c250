"""GPU: the BF16-split precision variant (clm_gemm_split_d, clm_expm_split_d).

No reference counterpart: the reference runs every product in the working
precision. The oracle here is the FP64 oracle (the C restatement of the
reference expm) and numpy's FP64 product, with the accuracy target the
precision variant states: relative Frobenius error <= 1e-6 for expm.
"""
import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle
from calaman_g_amd import _lib
from calaman_g_amd.expm import Context, expm_split_cm, gemm_cm

pytestmark = pytest.mark.gpu


def rel(x, y):
    """relative Frobenius error, scaled first (exp of a large norm reaches 1e159,
    whose squares overflow FP64)"""
    scale = np.abs(y).max()
    return np.linalg.norm((x - y) / scale) / np.linalg.norm(y / scale)


def cm(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)


@pytest.mark.parametrize("split,bound", [(3, 2e-5), (6, 5e-7)])
@pytest.mark.parametrize("m,n,k", [(128, 128, 128), (256, 384, 512), (1024, 1024, 1024),
                                   (37, 53, 29), (130, 127, 131)])
def test_split_gemm_accuracy(dev, split, bound, m, n, k):
    """Relative error of each entry against the exact product, measured in units
    of (|A||B|)_ij: the split keeps ~16 (3 products) or ~24 (6 products) bits."""
    rng = np.random.default_rng(m + n + k + split)
    a, b = rng.standard_normal((m, k)), rng.standard_normal((k, n))
    ctx = Context()
    dc = torch.zeros((n, m), dtype=torch.float64, device=dev)
    assert gemm_cm(ctx, m, n, k, 1.0, cm(a, dev), m, cm(b, dev), k, 0.0, dc, m, split=split) == 0
    got = dc.cpu().numpy().T
    scale = np.abs(a) @ np.abs(b)
    err = np.abs(got - a @ b) / scale
    assert err.max() <= bound, err.max()


def test_split_gemm_epilogue(dev):
    m = n = k = 256
    rng = np.random.default_rng(9)
    a, b, c = (rng.standard_normal((m, m)) for _ in range(3))
    ctx = Context()
    dc = cm(c, dev)
    assert gemm_cm(ctx, m, n, k, 2.0, cm(a, dev), m, cm(b, dev), k, -1.0, dc, m, split=6) == 0
    got = dc.cpu().numpy().T
    np.testing.assert_allclose(got, 2 * a @ b - c, rtol=0, atol=1e-5 * np.abs(a).max() * np.abs(b).max() * k ** 0.5)


def test_split_gemm_rejects_bad_split(dev):
    x = torch.zeros(16, dtype=torch.float64, device=dev)
    assert _lib.lib().clm_gemm_split_d(None, 4, 4, 4, 4, 1.0, x.data_ptr(), 4, x.data_ptr(), 4,
                                       0.0, x.data_ptr(), 4) == _lib.INVALID_VALUE


def _run_split(a, dev, ps_split=0, sq=-1):
    n = a.shape[0]
    da = cm(a, dev)
    de = torch.empty_like(da)
    work = torch.empty(cg.expm_split_buffer_size(n), dtype=torch.uint8, device=dev)
    info = torch.full((2,), -1, dtype=torch.int32, device=dev)
    st, plan = expm_split_cm(n, da, n, de, n, work, info, ps_split, sq)
    torch.cuda.synchronize()
    assert st == 0
    assert info.cpu().tolist() == [0, 0]
    return de.cpu().numpy().T, plan


# Sizes and norms where 1e-6 is attainable in single-precision-level arithmetic:
# the split products round their inputs at ~2^-24 relative, and exp's relative
# condition grows with ||A||, so for small matrices of large norm the error is set
# by that rounding, not by the schedule (measured: n = 16 at ||A||_1 = 1000
# 1.5e-6, n = 256 at 3000 1.7e-6; n = 2048 at 3000 5e-7).
_SPLIT_CASES = [(n, t) for n in (64, 256, 1024)
                for t in (0.5, 3.0, 8.0, 12.0, 40.0, 100.0, 400.0, 1000.0)] + \
               [(1024, 3000.0)] + \
               [(n, t) for n in (16, 32) for t in (0.5, 3.0, 8.0, 12.0, 40.0, 100.0)]


@pytest.mark.parametrize("n,target", _SPLIT_CASES)
def test_expm_split_meets_1e6(dev, n, target):
    rng = np.random.default_rng(n + int(target))
    a = rng.standard_normal((n, n))
    a *= target / np.abs(a).sum(0).max()
    got, plan = _run_split(a, dev)
    err = rel(got, oracle.expm(a))
    assert err <= 1e-6, (plan, err)
    s = plan["s"]
    small = plan["norm1"] / 2.0 ** s <= 1.0 + 2.0 ** -20
    half = 6 if n <= 64 else 13  # orders up to 64 take the 6-way polynomial
    assert plan["ps_split"] == ((3 if small else half) if s <= 5 else 6)
    # the last squarings split (2 for s = 4, 5 with half planes, else 1); with s > 8
    # every squaring stays FP64
    want_t = 0 if (s == 0 or s > 8) else (2 if (s in (4, 5) and plan["ps_split"] == 13) else 1)
    assert plan["split_squarings"] == want_t
    # the plan is the one the host predicts from the norms of A .. A^q (taken only
    # up to ||A||_1 = 256)
    assert plan == {**cg.expm_split_plan(plan["norm1"], power_norms=plan["power_norm1"], n=n),
                    "dtype": 1}
    assert plan["norm_powers"] in ((1,) if plan["norm1"] > 256 else (1, 3, 4, 5))


def test_expm_split_2048(dev):
    n = 2048
    rng = np.random.default_rng(2048)
    a = rng.standard_normal((n, n))
    a *= 8.0 / np.abs(a).sum(0).max()
    got, plan = _run_split(a, dev)
    assert rel(got, oracle.expm(a)) <= 1e-6


def test_expm_split_schedule_knobs(dev):
    """All-FP64 squarings and 6-way polynomial products only tighten the error."""
    n = 256
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, n))
    a *= 100.0 / np.abs(a).sum(0).max()
    ref = oracle.expm(a)
    e_auto, p_auto = _run_split(a, dev)
    e_tight, p_tight = _run_split(a, dev, ps_split=6, sq=0)
    assert p_tight["split_squarings"] == 0 and p_tight["ps_split"] == 6
    assert rel(e_tight, ref) <= rel(e_auto, ref) * 1.5
    assert rel(e_tight, ref) <= 1e-6


def test_expm_split_rejects_bad_options(dev):
    n = 8
    da = torch.zeros((n, n), dtype=torch.float64, device=dev)
    work = torch.empty(cg.expm_split_buffer_size(n), dtype=torch.uint8, device=dev)
    st, _ = expm_split_cm(n, da, n, torch.empty_like(da), n, work, None, ps_split=4)
    assert st == _lib.INVALID_VALUE


def test_expm_split_large_result_stays_finite(dev):
    """a power whose entries pass the FP32 range: split squarings scale rows and
    columns by powers of two first, so nothing overflows (BF16 planes of such a
    power -- FP32 first -- would be infinite)"""
    n = 32
    rng = np.random.default_rng(77)
    a = rng.standard_normal((n, n))
    a *= 500.0 / np.abs(a).sum(0).max()
    ref = oracle.expm(a)
    assert np.abs(ref).max() > 1e39  # beyond FP32
    got, plan = _run_split(a, dev)
    assert 1 <= plan["s"] <= 8 and plan["split_squarings"] >= 1
    assert np.all(np.isfinite(got))
    assert rel(got, ref) <= 1e-6


def test_expm_split_reused_workspace_matches_fresh(dev):
    """The half format's row / column exponents are tagged with the generation of the
    split that wrote them and cleared once per call: a workspace that held a larger
    matrix's exponents (same generations, larger exponents: without the clear they would
    win the atomic maximum) or positive garbage (0x5A5A5A5A: a huge generation that no
    split would ever overwrite) gives the same bits as a zeroed one, zero rows and
    columns included"""
    n = 96
    rng = np.random.default_rng(31)
    big = rng.standard_normal((n, n))
    big *= 300.0 / np.abs(big).sum(0).max()
    a = rng.standard_normal((n, n))
    a[::3, :] = 0.0
    a[:, 1::4] = 0.0
    a *= 20.0 / np.abs(a).sum(0).max()
    outs = []
    for fill in ("reuse", 0, 0x5A):
        work = torch.empty(cg.expm_split_buffer_size(n), dtype=torch.uint8, device=dev)
        if fill == "reuse":
            d_big = cm(big, dev)
            st, plan_big = expm_split_cm(n, d_big, n, torch.empty_like(d_big), n, work)
            assert st == 0 and plan_big["split_squarings"] > 0
        else:
            work.fill_(fill)
        da = cm(a, dev)
        de = torch.empty_like(da)
        st, plan = expm_split_cm(n, da, n, de, n, work)
        torch.cuda.synchronize()
        assert st == 0 and plan["ps_split"] == 13
        outs.append(de.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    assert rel(outs[0].T, oracle.expm(a)) <= 1e-6


@pytest.mark.parametrize("n,target", [(130, 8.0), (200, 100.0), (64, 0.5)])
def test_expm_split_padded_leading_dims_same_bits(dev, n, target):
    """odd leading dimensions (lda = n + 1, lde = n + 3) take the edge paths of the
    element-wise split kernels (unaligned, clamped loads) and of the products' LDS
    epilogue (element stores): the same bits as the contiguous call, padding untouched"""
    rng = np.random.default_rng(n + 5)
    a = rng.standard_normal((n, n))
    a *= target / np.abs(a).sum(0).max()
    want, plan = _run_split(a, dev)
    lda, lde = n + 1, n + 3
    bufa = np.full((n, lda), 3.0)
    bufa[:, :n] = a.T
    da = torch.from_numpy(bufa).to(dev)
    de = torch.full((n, lde), -5.0, dtype=torch.float64, device=dev)
    work = torch.empty(cg.expm_split_buffer_size(n), dtype=torch.uint8, device=dev)
    st, plan2 = expm_split_cm(n, da, lda, de, lde, work)
    torch.cuda.synchronize()
    assert st == 0 and plan2 == plan
    out = de.cpu().numpy()
    assert np.all(out[:, n:] == -5.0)
    assert np.array_equal(out[:, :n].T, want)
