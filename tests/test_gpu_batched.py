"""GPU: the batched persistent expm (clm_expm_batched / clm_expm_batched_graph).

Every matrix of a batch runs the single-matrix algorithm with the same fold
order, plan rule and product accumulation order, so the bar is BIT-FOR-BIT
equality with clm_expm_{s,d,c,z} on the same input -- which in turn is checked
against the oracle by test_gpu_expm.py / test_gpu_single.py; a few batch
members are also checked against the oracle directly here. Batches mix orders
1..300 (the 1-norm fold past 256 rows included), all four element types, padded
leading dimensions, and a non-finite member.
"""
import ctypes

import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle
from calaman_g_amd import _lib
from calaman_g_amd.expm import BatchPlan, expm_cm

pytestmark = pytest.mark.gpu

NP = {torch.float32: np.float32, torch.float64: np.float64, torch.complex64: np.complex64,
      torch.complex128: np.complex128}


def _matrix(rng, n, dtype, target):
    a = rng.standard_normal((n, n))
    if dtype.is_complex:
        a = a + 1j * rng.standard_normal((n, n))
    nrm = np.abs(a).sum(0).max()
    return (a * (target / nrm if nrm > 0 else 1.0)).astype(NP[dtype])


def _cm_buffer(a, ld, dev, fill=-5.0):
    """column-major storage with leading dimension ld, padding rows = fill"""
    n = a.shape[0]
    buf = np.full((n, ld), fill, dtype=a.dtype)
    buf[:, :n] = a.T
    return torch.from_numpy(buf).to(dev)


def _single(a, dev):
    """the single-matrix entry point on the same input bits"""
    n = a.shape[0]
    da = _cm_buffer(a, n, dev)
    de = torch.empty_like(da)
    tdt = da.dtype
    work = torch.empty(cg.expm_buffer_size(n, tdt), dtype=torch.uint8, device=dev)
    st, plan = expm_cm(n, da, n, de, n, work)
    assert st == 0
    torch.cuda.synchronize()
    return de.cpu().numpy().T.copy(), plan


def _build(rng, dev, specs):
    """specs: (n, dtype, target, lda_pad, lde_pad) -> (entries, host inputs)"""
    entries, inputs = [], []
    for n, dtype, target, pa, pe in specs:
        a = _matrix(rng, n, dtype, target)
        da = _cm_buffer(a, n + pa, dev)
        de = torch.from_numpy(np.full((n, n + pe), -5.0, dtype=NP[dtype])).to(dev)
        entries.append((n, da, n + pa, de, n + pe))
        inputs.append(a)
    return entries, inputs


def _mixed_specs(rng, count, nmax=300):
    types = [torch.float32, torch.complex64, torch.float64, torch.complex128]
    specs = []
    for i in range(count):
        n = int(rng.integers(1, nmax + 1)) if i > 3 else [1, 2, 256, 300][i]
        specs.append((n, types[i % 4], float(np.exp(rng.uniform(np.log(0.05), np.log(60.0)))),
                      int(rng.integers(0, 3)), int(rng.integers(0, 3))))
    return specs


def _check_batch(entries, inputs, results, dev):
    for (n, da, lda, de, lde), a, r in zip(entries, inputs, results):
        want, plan = _single(a, dev)
        full = de.cpu().numpy()
        got = full[:, :n].T
        assert np.array_equal(got, want), f"n={n} {a.dtype} plan={plan}"
        assert np.all(full[:, n:] == -5.0)  # padding rows untouched
        assert r["status"] == 0
        assert (r["m"], r["s"], r["num_gemms"]) == (plan["m"], plan["s"], plan["num_gemms"])
        assert r["norm1"] == pytest.approx(plan["norm1"], rel=0, abs=0)


def test_batched_mixed_types_bitwise_equals_single(dev):
    rng = np.random.default_rng(1)
    entries, inputs = _build(rng, dev, _mixed_specs(rng, 48))
    plan = BatchPlan(entries)
    work = plan.workspace()
    res = plan.results_buffer()
    assert plan.launch(work, res) == 0
    torch.cuda.synchronize()
    _check_batch(entries, inputs, BatchPlan.decode_results(res), dev)


@pytest.mark.parametrize("cfg", list(range(_lib.BATCH_TILES)))
def test_batched_every_tile_gives_the_same_bits(dev, cfg):
    """the tile shape changes which lanes compute an element, not its arithmetic"""
    rng = np.random.default_rng(10 + cfg)
    saved = {t: cg.batched_tuning(t) for t in NP}
    try:
        for t in NP:
            cg.set_batched_tuning(t, [cfg] * _lib.BATCH_BUCKETS)
        entries, inputs = _build(rng, dev, _mixed_specs(rng, 16, nmax=200))
        plan = BatchPlan(entries)
        work, res = plan.workspace(), plan.results_buffer()
        assert plan.launch(work, res) == 0
        torch.cuda.synchronize()
        _check_batch(entries, inputs, BatchPlan.decode_results(res), dev)
    finally:
        for t, c in saved.items():
            cg.set_batched_tuning(t, c)


def test_batched_against_oracle(dev):
    rng = np.random.default_rng(3)
    specs = [(64, torch.float64, 8.0, 0, 0), (100, torch.complex128, 30.0, 0, 0),
             (37, torch.float32, 4.0, 0, 0), (128, torch.complex64, 2.0, 0, 0)]
    entries, inputs = _build(rng, dev, specs)
    plan = BatchPlan(entries)
    work = plan.workspace()
    assert plan.launch(work) == 0
    torch.cuda.synchronize()
    for (n, _, _, de, _), a in zip(entries, inputs):
        wide = a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)
        ref = oracle.expm(wide)
        got = de.cpu().numpy()[:, :n].T
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        bound = 1e-12 if a.dtype in (np.float64, np.complex128) else 2e-5
        assert err <= bound, (n, a.dtype, err)


def test_batched_non_finite_member(dev):
    rng = np.random.default_rng(4)
    entries, inputs = _build(rng, dev, [(20, torch.float32, 1.0, 0, 0),
                                        (30, torch.float32, 1.0, 0, 0),
                                        (25, torch.complex64, 1.0, 0, 0)])
    bad = inputs[1].copy()
    bad[3, 4] = np.nan
    entries[1][1].copy_(_cm_buffer(bad, 30, dev))
    plan = BatchPlan(entries)
    work, res = plan.workspace(), plan.results_buffer()
    assert plan.launch(work, res) == 0
    torch.cuda.synchronize()
    r = BatchPlan.decode_results(res)
    assert r[1]["status"] == _lib.INVALID_VALUE
    out = entries[1][3].cpu().numpy()
    assert np.all(np.isnan(out[:, :30]))  # no exponential: NaN, never stale memory
    assert np.all(out[:, 30:] == -5.0)  # the padding rows untouched
    for i in (0, 2):
        want, _ = _single(inputs[i], dev)
        assert np.array_equal(entries[i][3].cpu().numpy()[:, :entries[i][0]].T, want)


def test_batched_graph_replays_match_eager(dev):
    """captured once, replayed with eager work and new input values in between;
    every replay is checked against the single-matrix results of its inputs"""
    rng = np.random.default_rng(5)
    specs = _mixed_specs(rng, 12, nmax=160)
    entries, inputs = _build(rng, dev, specs)
    plan = BatchPlan(entries)
    work, res = plan.workspace(), plan.results_buffer()
    stream = torch.cuda.Stream()
    for rep in range(3):
        if rep > 0:  # new values in the same buffers
            inputs = [_matrix(rng, n, t, g) for (n, t, g, _, _) in specs]
            for (n, da, lda, _, _), a in zip(entries, inputs):
                da.copy_(_cm_buffer(a, lda, dev))
        torch.cuda.synchronize()
        _ = torch.ones(1000, device=dev).sum().item()  # unrelated eager work
        assert plan.launch(work, res, stream=stream, graph=True) == 0
        stream.synchronize()
        _check_batch(entries, inputs, BatchPlan.decode_results(res), dev)


def test_batched_rejects_bad_plans(dev):
    a = torch.zeros((4, 4), dtype=torch.float32, device=dev)
    with pytest.raises(cg.CalamanError):
        BatchPlan([(4, a, 3, a, 4)])  # lda < n
    with pytest.raises(cg.CalamanError):
        BatchPlan([(0, a, 4, a, 4)])
    plan = BatchPlan([(4, a, 4, torch.empty_like(a), 4)])
    small = torch.empty(16, dtype=torch.uint8, device=dev)
    assert plan.launch(small) == _lib.ALLOC_FAILED
    assert plan.launch(plan.workspace(), graph=True) == _lib.INVALID_VALUE  # null stream


def test_batched_autotune_small(dev):
    """the tuner times every tile for the buckets up to n = 64 and installs its
    choice; a plan made afterwards still reproduces the single-matrix bits"""
    saved = cg.batched_tuning(torch.float32)
    try:
        table, ms = cg.batched_autotune(torch.float32, nmax=64)
        assert all(0 <= c < _lib.BATCH_TILES for c in table)
        # real FP32 has every tile distinct: all were timed
        assert all(t > 0 for t in ms[0]) and all(t > 0 for t in ms[1])
        assert cg.batched_tuning(torch.float32) == table
        for b in (0, 1):
            assert ms[b][table[b]] == min(ms[b])
        rng = np.random.default_rng(6)
        entries, inputs = _build(rng, dev, [(n, torch.float32, 2.0, 0, 0) for n in (5, 33, 64)])
        plan = BatchPlan(entries)
        work, res = plan.workspace(), plan.results_buffer()
        assert plan.launch(work, res) == 0
        torch.cuda.synchronize()
        _check_batch(entries, inputs, BatchPlan.decode_results(res), dev)
    finally:
        cg.set_batched_tuning(torch.float32, saved)


def test_expm_batched_convenience(dev):
    rng = np.random.default_rng(7)
    mats = [torch.from_numpy(_matrix(rng, n, t, 3.0)).to(dev)
            for n, t in ((10, torch.float64), (50, torch.complex64), (3, torch.float32))]
    outs = cg.expm_batched(mats)
    for a, e in zip(mats, outs):
        want = cg.expm(a)
        assert torch.equal(e, want)


def test_expm_batched_convenience_raises_on_non_finite_member(dev):
    rng = np.random.default_rng(8)
    mats = [torch.from_numpy(_matrix(rng, n, torch.float32, 2.0)).to(dev) for n in (9, 17, 5)]
    mats[1][2, 3] = float("inf")
    with pytest.raises(cg.CalamanError, match=r"members \[1\]"):
        cg.expm_batched(mats)


def test_expm_batched_convenience_on_a_side_stream(dev):
    """inputs made on the current stream, the launch on another: the wrapper orders
    the two and hands the results back to the current stream"""
    rng = np.random.default_rng(9)
    host = [_matrix(rng, n, torch.complex64, 4.0) for n in (64, 130)]
    want = [cg.expm(torch.from_numpy(h).to(dev)) for h in host]
    side = torch.cuda.Stream()
    for _ in range(3):
        mats = [torch.from_numpy(h).to(dev) * 1.0 for h in host]  # made on the current stream
        outs = cg.expm_batched(mats, stream=side)
        for o, w in zip(outs, want):
            assert torch.equal(o, w)
