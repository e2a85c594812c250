"""CPU: the C-ABI library loads, exports every symbol include/calaman_expm.h
declares, and its host-only logic (plans, thresholds, workspace sizes,
argument checks that return before any device call) matches the reference's
rules -- no GPU needed, as in the reference's host-only suites
(test/expm/expm_tests.cpp:389-465,
test/paterson_stockmeyer/paterson_stockmeyer_tests.cpp:231-383).
"""
import ctypes

import numpy as np
import pytest

import oracle
from calaman_g_amd import _lib

L = _lib.lib()
ST = _lib


def test_every_declared_symbol_is_exported():
    names = _lib.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(_lib._SIGNATURES), "binding table out of step with the header"


def test_version_and_status_strings():
    assert b"gfx950" in L.clm_version()
    assert L.clm_status_string(ST.INVALID_VALUE) == b"CLM_STATUS_INVALID_VALUE"
    assert L.clm_status_string(12345) == b"CLM_STATUS_UNKNOWN"
    assert ST.INVALID_VALUE & 0x7FFF0000 == ST.CLM_DOMAIN_BLAS
    assert ST.RUNTIME_INVALID_VALUE & 0x7FFF0000 == ST.CLM_DOMAIN_RUNTIME


@pytest.mark.parametrize("norm", [0.0, 1e-20, 1e-9, 2e-5, 0.003, 0.05, 0.2, 0.5, 0.64, 1.0,
                                  1.26, 2.2, 3.3, 4.0, 8.0, 16.0, 100.0, 256.0, 1e4, 1e12])
def test_expm_plan_matches_oracle(norm):
    p = L.clm_expm_plan_d(norm)
    t = oracle.taylor_plan(norm)
    assert (p.m, p.s, p.num_gemms, p.ps_block) == (t["degree"], t["s"], t["num_gemms"],
                                                   t["ps_block"])
    assert p.method == ST.METHOD_TAYLOR_PS and p.norm1 == norm
    # the reference's own ladder, restated on the host
    q = L.clm_pade_plan_d(norm)
    assert (q.m, q.s, q.num_gemms) == oracle.expm_plan(norm)


@pytest.mark.parametrize("norm", [0.0, 1e-9, 1e-4, 0.1, 0.4258730034897931, 0.5, 1.0, 1.2, 4.0,
                                  6.0, 11.25, 20.0, 100.0, 1e6])
def test_single_plans_match_oracle(norm):
    p = L.clm_expm_plan_s(norm)
    t = oracle.taylor_plan(norm, single=True)
    assert (p.m, p.s, p.num_gemms, p.ps_block) == (t["degree"], t["s"], t["num_gemms"],
                                                   t["ps_block"])
    # expm_plan<float>: comparisons in float (src/expm/plan.cppm:112-130)
    q = L.clm_pade_plan_s(norm)
    assert (q.m, q.s, q.num_gemms) == oracle.expm_plan(norm, single=True)


def _plan_powers(norms, single=False):
    arr = (ctypes.c_double * len(norms))(*norms)
    f = L.clm_expm_plan_powers_s if single else L.clm_expm_plan_powers_d
    return f(arr, len(norms))


# norms of A .. A^5 of the n = 2048 N(0,1) matrices at ||A||_1 = 0.5 / 8 / 100 (their
# k-th roots: 0.5 .082 .045 .033; 8 1.32 .71 .52 .44; 100 16.4 8.98 6.60 5.50)
_CFG2_NORMS = [
    [0.5, 0.082 ** 2, 0.045 ** 3, 0.033 ** 4],
    [8.0, 1.319 ** 2, 0.710 ** 3, 0.524 ** 4, 0.436 ** 5],
    [100.0, 16.389 ** 2, 8.978 ** 3, 6.604 ** 4, 5.504 ** 5],
]


@pytest.mark.parametrize("single", [False, True])
def test_power_plan_matches_oracle(single):
    rng = np.random.default_rng(5)
    cases = list(_CFG2_NORMS)
    for _ in range(300):
        q = int(rng.integers(1, 6))
        nrm = float(10.0 ** rng.uniform(-6, 3.5))
        # ||A^k|| <= ||A||^k; the k-th roots anywhere from tiny to ||A||
        roots = nrm * 10.0 ** rng.uniform(-3, 0, size=q - 1)
        cases.append([nrm] + [float(r ** k) for k, r in zip(range(2, q + 1), roots)])
    for norms in cases:
        p = _plan_powers(norms, single)
        t = oracle.taylor_plan_powers(norms, single)
        assert (p.m, p.s, p.num_gemms, p.ps_block) == (t["degree"], t["s"], t["num_gemms"],
                                                       t["ps_block"]), norms
        assert p.norm_powers == len(norms) and p.norm1 == norms[0]
        assert list(p.power_norm1)[:len(norms) - 1] == norms[1:]


def test_power_plan_known_answers():
    # the n = 2048 benchmark matrices: (m, s, products) drop from (15,0,6) / (24,2,10) /
    # (29,5,14) with the 1-norm alone
    got = [(p.m, p.s, p.num_gemms) for p in map(_plan_powers, _CFG2_NORMS)]
    assert got == [(8, 0, 5), (15, 0, 7), (29, 1, 10)]
    # norms of a normal matrix (||A^k|| = ||A||^k): the plan from the 1-norm alone
    for nrm in (0.3, 4.0, 100.0, 1e4):
        q0 = min(oracle.ps_plan(oracle.taylor_plan(nrm)["degree"])["num_powers"] + 1, 5)
        p = _plan_powers([nrm ** k for k in range(1, q0 + 1)])
        t = oracle.taylor_plan(nrm)
        assert (p.m, p.s, p.num_gemms) == (t["degree"], t["s"], t["num_gemms"])
    # A^2 = 0 (nilpotent of index 2): exp(A) = I + A, degree 1, no squaring, the
    # formed powers still counted
    p = _plan_powers([50.0, 0.0, 0.0, 0.0, 0.0])
    assert (p.m, p.s, p.num_gemms) == (1, 0, 4)
    # q = 1 or 2 cannot refine (alpha_2 needs ||A^3||)
    for norms in ([100.0], [100.0, 1e-6]):
        p = _plan_powers(norms)
        t = oracle.taylor_plan(100.0)
        assert (p.m, p.s) == (t["degree"], t["s"])


def test_plan_huge_norms_defined():
    """A norm whose ratio to theta_1 overflows: no undefined cast, the rung is skipped
    and the plan squares (the reference's ladder squares too)."""
    for nrm in (1e293, 1e300, 1.7e308):
        p = L.clm_expm_plan_d(nrm)
        t = oracle.taylor_plan(nrm)
        assert p.m > 1 and p.s > 900 and (p.m, p.s) == (t["degree"], t["s"])
        q = L.clm_expm_plan_s(nrm)
        assert q.m > 1 and q.s > 900


def test_thresholds_match_oracle():
    for m in list(range(0, 35)):
        assert L.clm_taylor_theta_d(m) == oracle.taylor_theta(m)
        assert L.clm_taylor_theta_s(m) == oracle.taylor_theta(m, single=True)
    for m in range(0, 15):
        assert L.clm_pade_theta_d(m) == oracle.pade_theta(m)
        assert L.clm_pade_theta_s(m) == oracle.pade_theta(m, single=True)


def test_ps_plan_functions_match_oracle():
    for d in range(0, 40):
        for s in (0, 1, 2, 3, 5, 50):
            p = L.clm_paterson_stockmeyer_plan(d, s).as_dict()
            assert p == oracle.ps_plan(d, s)
        for s in range(1, d + 2):
            assert L.clm_paterson_stockmeyer_cost(d, s) == oracle.ps_cost(d, s)


def test_ps_buffer_size_is_plan_blocks_aligned():
    """test/paterson_stockmeyer/paterson_stockmeyer_tests.cpp:297-330"""
    for n in (1, 2, 5, 7, 16):
        for degree in (2, 4, 8, 16):
            plan = oracle.ps_plan(degree)
            block = ((n * n * 8 + 255) // 256) * 256
            assert L.clm_paterson_stockmeyer_buffer_size_d(n, degree, 0) == plan["num_blocks"] * block
            zblock = ((n * n * 16 + 255) // 256) * 256
            assert L.clm_paterson_stockmeyer_buffer_size_z(n, degree, 0) == plan["num_blocks"] * zblock
    assert L.clm_paterson_stockmeyer_buffer_size_d(10, 0, 0) == 0
    assert L.clm_paterson_stockmeyer_buffer_size_d(10, 1, 0) == 0
    assert L.clm_paterson_stockmeyer_power_stride_d(5) == 32  # 200 B -> 256 B


def test_expm_and_horner_buffer_sizes():
    for n in (1, 7, 64, 2048):
        blk = ((n * n * 8 + 255) // 256) * 256
        assert L.clm_horner_buffer_size_d(n) == blk
        out = ctypes.c_size_t()
        assert L.clm_expm_buffer_size_d(n, ctypes.byref(out)) == 0
        colsum = ((n * 8 + 255) // 256) * 256
        # colsum + 32 coefficients + As + a 4-power bank + one scratch block: six n x n blocks
        assert out.value == colsum + 256 + 6 * blk
        assert L.clm_expm_buffer_size_z(n, ctypes.byref(out)) == 0
        zblk = ((n * n * 16 + 255) // 256) * 256
        assert out.value == colsum + 512 + 6 * zblk
    assert L.clm_expm_buffer_size_d(0, ctypes.byref(out)) == ST.INVALID_VALUE
    for n in (1, 7, 64, 2048):
        colsum = ((n * 4 + 255) // 256) * 256
        for f, es in ((L.clm_expm_buffer_size_s, 4), (L.clm_expm_buffer_size_c, 8)):
            assert f(n, ctypes.byref(out)) == 0
            assert out.value == colsum + 256 + 6 * (((n * n * es + 255) // 256) * 256)
        assert L.clm_horner_buffer_size_s(n) == ((n * n * 4 + 255) // 256) * 256
        assert L.clm_horner_buffer_size_c(n) == ((n * n * 8 + 255) // 256) * 256
    assert L.clm_paterson_stockmeyer_power_stride_s(5) == 64  # 100 B -> 256 B
    assert L.clm_paterson_stockmeyer_power_stride_c(5) == 32


def _p(x):
    return ctypes.c_void_p(x)


def test_expm_rejects_bad_arguments_before_device():
    """test/expm/expm_tests.cpp:441-465 with host dummy pointers."""
    a = np.ones(16)
    e = np.zeros(16)
    w = np.zeros(64)
    pa, pe, pw = (_p(x.ctypes.data) for x in (a, e, w))
    for f in (L.clm_expm_d, L.clm_expm_z, L.clm_expm_s, L.clm_expm_c):
        call = lambda n, lda, lde, work: f(None, None, n, pa, lda, pe, lde, work, 512, None, None)
        assert call(0, 4, 4, pw) == ST.INVALID_VALUE
        assert call(4, 3, 4, pw) == ST.INVALID_VALUE
        assert call(4, 4, 3, pw) == ST.INVALID_VALUE
        assert call(4, 4, 4, None) == ST.INVALID_VALUE
        assert call(46341, 46341, 46341, pw) == ST.INVALID_VALUE
        # 512 bytes is too small for n = 4: ALLOC_FAILED, still before any device call
        assert call(4, 4, 4, pw) == ST.ALLOC_FAILED


def test_ps_horner_reject_with_not_initialized():
    a = np.ones(16)
    pa = _p(a.ctypes.data)
    for f in (L.clm_paterson_stockmeyer_d, L.clm_paterson_stockmeyer_z,
              L.clm_paterson_stockmeyer_s, L.clm_paterson_stockmeyer_c):
        assert f(None, 0, pa, 3, pa, 4, pa, 4, pa, 10 ** 6, 0) == ST.NOT_INITIALIZED  # n < 1
        assert f(None, 4, pa, -1, pa, 4, pa, 4, pa, 10 ** 6, 0) == ST.NOT_INITIALIZED  # degree
        assert f(None, 4, pa, 3, pa, 3, pa, 4, pa, 10 ** 6, 0) == ST.NOT_INITIALIZED  # lda
        assert f(None, 4, pa, 3, pa, 4, pa, 3, pa, 10 ** 6, 0) == ST.NOT_INITIALIZED  # ldp
        assert f(None, 4, pa, 3, pa, 4, pa, 4, pa, 10 ** 6, -1) == ST.NOT_INITIALIZED  # s
        assert f(None, 4, None, 3, pa, 4, pa, 4, pa, 10 ** 6, 0) == ST.NOT_INITIALIZED  # coeffs
        assert f(None, 4, pa, 3, pa, 4, pa, 4, None, 0, 0) == ST.NOT_INITIALIZED  # workspace
        assert f(None, 4, pa, 3, pa, 4, pa, 4, pa, 8, 0) == ST.NOT_INITIALIZED  # small workspace
    for f in (L.clm_horner_d, L.clm_horner_z, L.clm_horner_s, L.clm_horner_c):
        assert f(None, 0, pa, 3, pa, 4, pa, 4, pa, 10 ** 6) == ST.NOT_INITIALIZED
        assert f(None, 4, pa, 3, pa, 4, pa, 4, None, 0) == ST.NOT_INITIALIZED
        assert f(None, 4, pa, 3, None, 4, pa, 4, pa, 10 ** 6) == ST.NOT_INITIALIZED


def test_lascl_lange_gemm_argument_checks():
    a = np.ones(16)
    pa = _p(a.ctypes.data)
    for f in (L.clm_lascl_d, L.clm_lascl_z, L.clm_lascl_s, L.clm_lascl_c):
        assert f(None, 0.0, 1.0, 4, 4, pa, 4) == ST.RUNTIME_INVALID_VALUE
        assert f(None, float("nan"), 1.0, 4, 4, pa, 4) == ST.RUNTIME_INVALID_VALUE
        assert f(None, 1.0, float("nan"), 4, 4, pa, 4) == ST.RUNTIME_INVALID_VALUE
        assert f(None, 1.0, 2.0, 0, 4, pa, 4) == ST.SUCCESS  # empty: nothing enqueued
        assert f(None, 1.0, 2.0, 4, 4, pa, 3) == ST.RUNTIME_INVALID_VALUE
    for f in (L.clm_lange_d, L.clm_lange_z, L.clm_lange_s, L.clm_lange_c):
        assert f(None, 7, 4, 4, pa, 4, pa, pa) == ST.RUNTIME_INVALID_VALUE
        assert f(None, 1, 4, 4, pa, 3, pa, pa) == ST.RUNTIME_INVALID_VALUE
        assert f(None, 1, 4, 4, pa, 4, None, pa) == ST.RUNTIME_INVALID_VALUE
    assert L.clm_gemm_d(None, 4, 4, 4, 1.0, pa, 3, pa, 4, 0.0, pa, 4) == ST.INVALID_VALUE
    assert L.clm_gemm_d(None, -1, 4, 4, 1.0, pa, 4, pa, 4, 0.0, pa, 4) == ST.INVALID_VALUE
    assert L.clm_gemm_d(None, 0, 4, 4, 1.0, pa, 4, pa, 4, 0.0, pa, 4) == ST.SUCCESS
    assert L.clm_matrix_norm1_d(None, 0, pa, 4, pa, ctypes.byref(ctypes.c_double())) == ST.INVALID_VALUE
    assert L.clm_gemm_s(None, 4, 4, 4, 1.0, pa, 3, pa, 4, 0.0, pa, 4) == ST.INVALID_VALUE
    assert L.clm_matrix_norm1_s(None, 0, pa, 4, pa, ctypes.byref(ctypes.c_float())) == ST.INVALID_VALUE
    one = _lib.FloatComplex(1.0, 0.0)
    assert L.clm_gemm_c(None, 4, 4, 4, one, pa, 4, pa, 3, one, pa, 4) == ST.INVALID_VALUE


def test_batched_plan_rejects_bad_entries_before_device():
    h = ctypes.c_void_p()
    e = (_lib.BatchEntry * 2)()
    assert L.clm_expm_batched_plan_create(ctypes.byref(h), 0, e) == ST.INVALID_VALUE
    assert L.clm_expm_batched_plan_create(None, 1, e) == ST.INVALID_VALUE
    good = _lib.BatchEntry(4, ST.DTYPE_F32, 4096, 4, 8192, 4)
    for bad in (_lib.BatchEntry(0, ST.DTYPE_F32, 4096, 4, 8192, 4),      # n < 1
                _lib.BatchEntry(4097, ST.DTYPE_F32, 4096, 4097, 8192, 4097),  # n too large
                _lib.BatchEntry(4, ST.DTYPE_F32, 4096, 3, 8192, 4),      # lda < n
                _lib.BatchEntry(4, ST.DTYPE_F32, 4096, 4, 8192, 3),      # lde < n
                _lib.BatchEntry(4, ST.DTYPE_F32, None, 4, 8192, 4),      # no A
                _lib.BatchEntry(4, 9, 4096, 4, 8192, 4)):               # no such dtype
        e[0], e[1] = good, bad
        assert L.clm_expm_batched_plan_create(ctypes.byref(h), 2, e) == ST.INVALID_VALUE
    size = ctypes.c_size_t()
    assert L.clm_expm_batched_buffer_size(None, ctypes.byref(size)) == ST.INVALID_VALUE
    assert L.clm_expm_batched(None, None, None, None, 0, None) == ST.INVALID_VALUE


def test_batched_tuning_table_round_trip():
    cfg = (ctypes.c_int * _lib.BATCH_BUCKETS)()
    assert L.clm_expm_batched_get_tuning(ST.DTYPE_C64, cfg) == 0
    saved = list(cfg)
    assert all(0 <= c < _lib.BATCH_TILES for c in saved)
    new = (ctypes.c_int * _lib.BATCH_BUCKETS)(*([1] * _lib.BATCH_BUCKETS))
    assert L.clm_expm_batched_set_tuning(ST.DTYPE_C64, new) == 0
    assert L.clm_expm_batched_get_tuning(ST.DTYPE_C64, cfg) == 0 and list(cfg) == [1] * 8
    bad = (ctypes.c_int * _lib.BATCH_BUCKETS)(*([_lib.BATCH_TILES] * _lib.BATCH_BUCKETS))
    assert L.clm_expm_batched_set_tuning(ST.DTYPE_C64, bad) == ST.INVALID_VALUE
    assert L.clm_expm_batched_set_tuning(7, new) == ST.INVALID_VALUE
    assert L.clm_expm_batched_set_tuning(ST.DTYPE_C64, (ctypes.c_int * 8)(*saved)) == 0


def test_adjoint_tuning_table_is_separate():
    fwd, ad = (ctypes.c_int * _lib.BATCH_BUCKETS)(), (ctypes.c_int * _lib.BATCH_BUCKETS)()
    assert L.clm_expm_batched_get_tuning(ST.DTYPE_F64, fwd) == 0
    assert L.clm_expm_batched_get_tuning_ad(ST.DTYPE_F64, ad) == 0
    saved_f, saved_a = list(fwd), list(ad)
    new = (ctypes.c_int * _lib.BATCH_BUCKETS)(*([_lib.BATCH_TILE_128W16] * _lib.BATCH_BUCKETS))
    assert L.clm_expm_batched_set_tuning_ad(ST.DTYPE_F64, new) == 0
    assert L.clm_expm_batched_get_tuning_ad(ST.DTYPE_F64, ad) == 0
    assert list(ad) == [_lib.BATCH_TILE_128W16] * 8
    assert L.clm_expm_batched_get_tuning(ST.DTYPE_F64, fwd) == 0 and list(fwd) == saved_f
    bad = (ctypes.c_int * _lib.BATCH_BUCKETS)(*([-1] * _lib.BATCH_BUCKETS))
    assert L.clm_expm_batched_set_tuning_ad(ST.DTYPE_F64, bad) == ST.INVALID_VALUE
    assert L.clm_expm_batched_autotune_ad(None, None, 9, 64, None, None) == ST.INVALID_VALUE
    assert L.clm_expm_batched_set_tuning_ad(ST.DTYPE_F64, (ctypes.c_int * 8)(*saved_a)) == 0


def test_fp8_buffer_sizes_and_argument_checks():
    """the FP8 workspaces: the split workspace plus three n x n byte planes and their
    tile exponents; the product's two operand planes and exponents; bad options
    and sizes rejected before any device call"""
    out, split = ctypes.c_size_t(), ctypes.c_size_t()
    for n in (1, 100, 128, 129, 4096):
        assert L.clm_expm_fp8tail_buffer_size_d(n, ctypes.byref(out)) == 0
        assert L.clm_expm_split_buffer_size_d(n, ctypes.byref(split)) == 0
        tiles = -(-n // 128)
        plane = ((n * n + 255) // 256) * 256 + ((tiles * tiles * 4 + 255) // 256) * 256
        assert out.value == split.value + 3 * plane
    assert L.clm_expm_fp8tail_buffer_size_d(0, ctypes.byref(out)) == ST.INVALID_VALUE
    al = lambda v: ((v + 255) // 256) * 256
    assert L.clm_gemm_fp8_buffer_size_d(300, 200, 130, ctypes.byref(out)) == 0
    assert out.value == al(300 * 130) + al(3 * 2 * 4) + al(130 * 200) + al(2 * 2 * 4)
    assert L.clm_gemm_fp8_buffer_size_d(-1, 2, 2, ctypes.byref(out)) == ST.INVALID_VALUE
    a = np.ones(64)
    pa = _p(a.ctypes.data)
    assert L.clm_gemm_fp8_d(None, 4, 4, 4, 1.0, pa, 3, pa, 4, 0.0, pa, 4, pa, 512) == ST.INVALID_VALUE
    assert L.clm_gemm_fp8_d(None, 0, 4, 4, 1.0, pa, 4, pa, 4, 0.0, pa, 4, None, 0) == ST.SUCCESS
    assert L.clm_gemm_fp8_d(None, 4, 4, 4, 1.0, pa, 4, pa, 4, 0.0, pa, 4, pa, 8) == ST.ALLOC_FAILED
    for opts in ((0, -1, 1), (5, -1, 0), (0, -2, 0), (0, -1, -3)):
        o = _lib.Fp8TailOpts(*opts)
        assert L.clm_expm_fp8tail_d(None, None, 4, pa, 4, pa, 4, pa, 10 ** 6, None,
                                    ctypes.byref(o), None) == ST.INVALID_VALUE
    assert L.clm_expm_fp8tail_d(None, None, 4, pa, 4, pa, 4, pa, 512, None, None,
                                None) == ST.ALLOC_FAILED


@pytest.mark.parametrize("norm,want", [
    (0.5, (0, 3, 0)), (1.0, (0, 3, 0)), (1.0 + 2.0 ** -52, (0, 3, 0)), (1.01, (0, 13, 0)),
    (8.0, (1, 13, 1)), (100.0, (5, 13, 2)), (1000.0, (8, 6, 1)), (1e6, (18, 6, 0))])
def test_split_schedule(norm, want):
    """(s, polynomial format, split squarings) of the precision variant: BF16 3-way
    while ||A / 2^s||_1 <= 1 (a matrix scaled to 1-norm 1 whose column sum rounds an
    ulp above stays there), half planes to s = 5 with two split squarings at s = 4, 5,
    6-way beyond, every squaring FP64 past s = 8"""
    from calaman_g_amd.expm import expm_split_plan
    p = expm_split_plan(norm)
    assert (p["s"], p["ps_split"], p["split_squarings"]) == want
    assert p["num_gemms"] == oracle.ps_plan(p["m"], 0)["num_gemms"] + p["s"]


def _split_plan_powers(n, norms):
    arr = (ctypes.c_double * len(norms))(*norms)
    opts = _lib.SplitOpts(0, -1, _lib.PLAN_POWERS)
    return L.clm_expm_split_plan_powers_d(n, arr, len(norms), ctypes.byref(opts))


def test_split_power_plan_rules():
    # the n = 2048 benchmark matrices: no squaring at 0.5 and 8, one at 100, all split
    got = [(p.m, p.s, p.ps_split, p.split_squarings)
           for p in (_split_plan_powers(2048, nm) for nm in _CFG2_NORMS)]
    assert got == [(5, 0, 3, 0), (11, 0, 13, 0), (24, 1, 13, 1)]
    # above ||A||_1 = 256 the norms of powers are not used: the 1-norm's plan
    for nrm in (300.0, 1000.0, 3000.0):
        norms = [nrm, (0.1 * nrm) ** 2, (0.05 * nrm) ** 3, (0.04 * nrm) ** 4, (0.03 * nrm) ** 5]
        p = _split_plan_powers(1024, norms)
        q = L.clm_expm_split_plan_d(nrm, None)
        assert (p.m, p.s, p.ps_split, p.split_squarings, p.norm_powers) == (
            q.m, q.s, q.ps_split, q.split_squarings, 1)
    # orders up to 64 take the 6-way polynomial where larger ones take half planes
    norms = [100.0, 26.7 ** 2, 17.3 ** 3, 13.8 ** 4, 12.1 ** 5]
    small, large = _split_plan_powers(64, norms), _split_plan_powers(65, norms)
    assert small.ps_split == _lib.SPLIT_BF16X6 and large.ps_split == _lib.SPLIT_FP16X3
    assert small.s <= 5 and large.s <= 5


def test_pade_route_sizes_and_argument_checks():
    """clm_pade_buffer_size_*: the powers the degree needs (1, 2, 3, 4, and 3 plus the
    nested X at m = 13) and V, W; expm's covers the worst degree plus A / 2^s;
    argument checks return before any device work"""
    out = ctypes.c_size_t()
    for sfx, size in (("d", 8), ("z", 16), ("s", 4), ("c", 8)):
        f = getattr(L, f"clm_pade_buffer_size_{sfx}")
        n = 100
        blocks = {}
        for m, nb in ((3, 3), (5, 4), (7, 5), (9, 6), (13, 6)):
            assert f(m, n, ctypes.byref(out)) == 0
            blocks[m] = out.value
            assert out.value >= nb * n * n * size
        assert blocks[3] < blocks[5] < blocks[7] < blocks[9]
        assert f(4, n, ctypes.byref(out)) == ST.INVALID_VALUE
        assert f(13, 0, ctypes.byref(out)) == ST.INVALID_VALUE
        assert getattr(L, f"clm_expm_pade_buffer_size_{sfx}")(n, ctypes.byref(out)) == 0
        assert out.value >= max(blocks.values()) + n * n * size
    buf = (ctypes.c_double * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    assert L.clm_pade_d(None, None, 11, 4, p, 4, p, 4, p, 1 << 20, None) == ST.INVALID_VALUE
    assert L.clm_pade_d(None, None, 3, 4, p, 3, p, 4, p, 1 << 20, None) == ST.INVALID_VALUE
    assert L.clm_pade_d(None, None, 3, 4, p, 4, p, 4, p, 16, None) == ST.ALLOC_FAILED
    assert L.clm_expm_pade_d(None, None, 0, p, 1, p, 1, p, 1 << 20, None, None) == ST.INVALID_VALUE
    assert L.clm_expm_pade_d(None, None, 4, p, 4, p, 4, p, 16, None, None) == ST.ALLOC_FAILED
    assert L.clm_getrf_d(None, None, -1, p, 1, p, p) == ST.INVALID_VALUE
    assert L.clm_getrf_d(None, None, 4, p, 3, p, p) == ST.INVALID_VALUE
    assert L.clm_getrf_d(None, None, 4, p, 4, p, None) == ST.INVALID_VALUE
    assert L.clm_getrs_d(None, None, 4, 2, p, 4, p, p, 3, None) == ST.INVALID_VALUE


def test_balancing_and_lacn2_sizes_and_argument_checks():
    """gebal_bufferSize is 2n + 1 ints; the balanced expm's workspace holds B, D, the
    gebal ints and clm_expm_*'s own; lacn2's holds the sign record, a flag and the
    reductions. Every check below returns before any device work."""
    lw = ctypes.c_int()
    assert L.clm_gebal_buffer_size(5, ctypes.byref(lw)) == 0 and lw.value == 11
    assert L.clm_gebal_buffer_size(-1, ctypes.byref(lw)) == ST.RUNTIME_INVALID_VALUE
    out, inner = ctypes.c_size_t(), ctypes.c_size_t()
    for sfx, size in (("d", 8), ("z", 16), ("s", 4), ("c", 8)):
        assert getattr(L, f"clm_expm_balanced_buffer_size_{sfx}")(100, ctypes.byref(out)) == 0
        assert getattr(L, f"clm_expm_buffer_size_{sfx}")(100, ctypes.byref(inner)) == 0
        assert out.value >= inner.value + 100 * 100 * size + 100 * (size // (1 + (sfx in "zc")))
        assert getattr(L, f"clm_expm_balanced_buffer_size_{sfx}")(0, ctypes.byref(out)) == \
            ST.INVALID_VALUE
    assert L.clm_lacn2_buffer_size_d(1000) >= 1000 * 4 + 4 + 32
    buf = (ctypes.c_double * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    ilo, ihi = ctypes.c_int(), ctypes.c_int()
    for job in (-1, 4):
        assert L.clm_gebal_d(None, job, 4, p, 4, ctypes.byref(ilo), ctypes.byref(ihi), p, p, 0) == \
            ST.RUNTIME_INVALID_VALUE
    assert L.clm_gebal_d(None, 3, 4, p, 3, ctypes.byref(ilo), ctypes.byref(ihi), p, p, 0) == \
        ST.RUNTIME_INVALID_VALUE
    assert L.clm_gebal_z(None, 3, 4, p, 4, ctypes.byref(ilo), ctypes.byref(ihi), p, None, 0) == \
        ST.RUNTIME_INVALID_VALUE
    # n = 0: ilo = 1, ihi = 0 without a launch
    assert L.clm_gebal_d(None, 3, 0, p, 1, ctypes.byref(ilo), ctypes.byref(ihi), p, p, 0) == 0
    assert (ilo.value, ihi.value) == (1, 0)
    assert L.clm_expm_balanced_d(None, None, 4, p, 3, p, 4, p, 1 << 20, None, None) == \
        ST.INVALID_VALUE
    assert L.clm_expm_balanced_d(None, None, 4, p, 4, p, 4, p, 16, None, None) == ST.ALLOC_FAILED
    est, kase, isave = ctypes.c_double(), ctypes.c_int(0), (ctypes.c_int * 3)(0, 0, 0)
    assert L.clm_lacn2_d(None, 0, p, p, p, 1 << 10, ctypes.byref(est), ctypes.byref(kase), isave) == \
        ST.INVALID_VALUE
    assert L.clm_lacn2_d(None, 8, p, p, p, 8, ctypes.byref(est), ctypes.byref(kase), isave) == \
        ST.INVALID_VALUE
    kase.value, isave[0] = 1, 9  # a state no step leaves behind
    assert L.clm_lacn2_d(None, 8, p, p, p, 1 << 10, ctypes.byref(est), ctypes.byref(kase), isave) == \
        ST.INVALID_VALUE
    kase.value, isave[0], isave[1] = 2, 4, 9  # a column index past n
    assert L.clm_lacn2_d(None, 8, p, p, p, 1 << 10, ctypes.byref(est), ctypes.byref(kase), isave) == \
        ST.INVALID_VALUE


def test_expm_herm_sizes_and_argument_checks():
    """test/expm/expm_tests.cpp:780-800: n < 1, lda < n, lde < n and a null workspace are
    INVALID_VALUE, so is a size query at n < 1; a bad fill mode too; a short workspace
    ALLOC_FAILED. The workspace holds three blocks of order n rounded up to 64."""
    out = ctypes.c_size_t()
    for sfx, size in (("d", 8), ("z", 16), ("s", 4), ("c", 8)):
        f = getattr(L, f"clm_expm_herm_buffer_size_{sfx}")
        assert f(0, ctypes.byref(out)) == ST.INVALID_VALUE
        assert f(100, ctypes.byref(out)) == 0
        assert out.value >= 3 * 128 * 128 * size
        assert f(1, ctypes.byref(out)) == 0 and out.value >= 3 * 64 * 64 * size
    buf = (ctypes.c_double * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    call = L.clm_expm_herm_d
    assert call(None, None, 1, 0, p, 4, p, 4, p, 1 << 30, None) == ST.INVALID_VALUE
    assert call(None, None, 1, 4, p, 3, p, 4, p, 1 << 30, None) == ST.INVALID_VALUE
    assert call(None, None, 1, 4, p, 4, p, 3, p, 1 << 30, None) == ST.INVALID_VALUE
    assert call(None, None, 1, 4, p, 4, p, 4, None, 1 << 30, None) == ST.INVALID_VALUE
    assert call(None, None, 2, 4, p, 4, p, 4, p, 1 << 30, None) == ST.INVALID_VALUE
    assert call(None, None, 0, 4, p, 4, p, 4, p, 64, None) == ST.ALLOC_FAILED
