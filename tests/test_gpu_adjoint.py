"""GPU: the hand-written Frechet adjoint (clm_expm_batched_backward) and the
torch.autograd.Function over it.

Oracles:
  * the block-triangular identity: the gradient of Re tr(G^H exp(A)) is
    L(A^H, G) = [exp([[A^H, G], [0, A^H]])]_12 (oracle.frechet_adjoint, itself
    pinned against scipy.linalg.expm_frechet golden values and central
    differences in test_oracle_expm.py);
  * FP64 central differences (h = 1e-6) of the library's own forward;
  * torch.linalg.matrix_exp's autograd (a plain PyTorch reference of the op).
The adjoint differentiates the computed chain exactly, so in FP64 it meets the
true derivative to the forward's accuracy (bound 1e-10 relative); in FP32 to a
small multiple of torch's FP32 gradient error.
"""
import numpy as np
import pytest
import torch

import calaman_g_amd as cg
import oracle
from calaman_g_amd import _lib
from calaman_g_amd.expm import BatchPlan, expm_batched_autograd

pytestmark = pytest.mark.gpu

NP = {torch.float32: np.float32, torch.float64: np.float64, torch.complex64: np.complex64,
      torch.complex128: np.complex128}
EPS32 = float(np.finfo(np.float32).eps)


def _mat(rng, n, dtype, target):
    a = rng.standard_normal((n, n))
    if dtype.is_complex:
        a = a + 1j * rng.standard_normal((n, n))
    return (a * (target / np.abs(a).sum(0).max())).astype(NP[dtype])


def _cm(a, ld, dev):
    buf = np.full((a.shape[0], ld), -3.0, dtype=a.dtype)
    buf[:, :a.shape[0]] = a.T
    return torch.from_numpy(buf).to(dev)


def _run(dev, mats, grads, checkpoints=4, pad=1):
    """forward + backward of an adjoint plan; returns (E list, dA list, plan results)"""
    ents, adj, outs = [], [], []
    for a, g in zip(mats, grads):
        n = a.shape[0]
        da, de = _cm(a, n + pad, dev), _cm(np.zeros_like(a), n + pad, dev)
        dg, dd = _cm(g, n + pad, dev), _cm(np.zeros_like(a), n + pad, dev)
        ents.append((n, da, n + pad, de, n + pad))
        adj.append((dg, n + pad, dd, n + pad))
        outs.append((de, dd))
    plan = BatchPlan(ents, adjoint=adj, checkpoints=checkpoints)
    work, res = plan.workspace(), plan.results_buffer()
    assert plan.launch(work, res) == 0
    assert plan.backward(work) == 0
    torch.cuda.synchronize()
    es = [de.cpu().numpy()[:, :a.shape[0]].T.copy() for (de, _), a in zip(outs, mats)]
    ds = [dd.cpu().numpy()[:, :a.shape[0]].T.copy() for (_, dd), a in zip(outs, mats)]
    pads_ok = all(np.all(dd.cpu().numpy()[:, a.shape[0]:] == -3.0) for (_, dd), a in zip(outs, mats))
    assert pads_ok
    return es, ds, BatchPlan.decode_results(res)


def _rel(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_adjoint_fp64_matches_block_triangular(dev, dtype):
    rng = np.random.default_rng(1)
    specs = [(5, 0.3), (17, 2.0), (40, 9.0), (64, 100.0), (100, 1.0)]
    mats = [_mat(rng, n, dtype, t) for n, t in specs]
    grads = [_mat(rng, n, dtype, float(n)) for n, _ in specs]
    _, ds, res = _run(dev, mats, grads)
    assert max(r["s"] for r in res) >= 2  # the squaring sweep is exercised
    for a, g, d, r in zip(mats, grads, ds, res):
        ref = oracle.frechet_adjoint(a, g)
        assert _rel(d, ref) <= 1e-10, (a.shape[0], r, _rel(d, ref))


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64])
def test_adjoint_fp32_against_block_triangular_and_torch(dev, dtype):
    rng = np.random.default_rng(2)
    specs = [(8, 0.5), (33, 4.0), (128, 20.0)]
    mats = [_mat(rng, n, dtype, t) for n, t in specs]
    grads = [_mat(rng, n, dtype, float(n)) for n, _ in specs]
    _, ds, res = _run(dev, mats, grads)
    for a, g, d in zip(mats, grads, ds):
        wide = np.complex128 if dtype.is_complex else np.float64
        ref = oracle.frechet_adjoint(a.astype(wide), g.astype(wide))
        ta = torch.from_numpy(a).to(dev).requires_grad_(True)
        torch.linalg.matrix_exp(ta).backward(torch.from_numpy(g).to(dev))
        torch_err = _rel(ta.grad.cpu().numpy(), ref)
        err = _rel(d, ref)
        assert err <= max(4 * torch_err, 64 * EPS32), (a.shape[0], err, torch_err)


def test_adjoint_central_differences(dev):
    """Re <dA, D> against FP64 central differences (h = 1e-6) of the library's
    own forward, L(A) = Re tr(G^H exp(A))"""
    rng = np.random.default_rng(3)
    n = 24
    a = _mat(rng, n, torch.complex128, 6.0)
    g = _mat(rng, n, torch.complex128, 10.0)
    d = _mat(rng, n, torch.complex128, 1.0)
    _, ds, _ = _run(dev, [a], [g])
    h = 1e-6

    def loss(x):
        e = cg.expm(torch.from_numpy(x).to(dev)).cpu().numpy()
        return float(np.real(np.trace(g.conj().T @ e)))

    fd = (loss(a + h * d) - loss(a - h * d)) / (2 * h)
    assert float(np.real(np.vdot(ds[0], d))) == pytest.approx(fd, rel=1e-7)


def test_adjoint_checkpoints_and_forward_bits(dev):
    """the checkpoint count changes what is stored, not the arithmetic: every
    count gives the same gradient bits, and the forward of an adjoint plan the
    same bits as the plain forward"""
    rng = np.random.default_rng(4)
    mats = [_mat(rng, n, t, 300.0) for n, t in ((30, torch.complex64), (50, torch.float64))]
    grads = [_mat(rng, a.shape[0], torch.complex64 if np.iscomplexobj(a) else torch.float64, 5.0)
             for a in mats]
    runs = [_run(dev, mats, grads, checkpoints=k) for k in (1, 2, 3, 64)]
    assert min(r["s"] for r in runs[0][2]) >= 4
    for es, ds, _ in runs[1:]:
        for x, y in zip(ds, runs[0][1]):
            assert np.array_equal(x, y)
    for a, e in zip(mats, runs[0][0]):
        want = cg.expm(torch.from_numpy(a).to(dev)).cpu().numpy()
        assert np.array_equal(e, want)


def test_adjoint_non_finite_member_gives_nan(dev):
    rng = np.random.default_rng(5)
    good = _mat(rng, 12, torch.float32, 1.0)
    bad = good.copy()
    bad[2, 3] = np.inf
    _, ds, res = _run(dev, [good, bad], [good, good])
    assert res[1]["status"] != 0
    assert np.all(np.isnan(ds[1])) and np.all(np.isfinite(ds[0]))


@pytest.mark.parametrize("dtype", [torch.complex128, torch.float64])
def test_autograd_function_matches_torch(dev, dtype):
    rng = np.random.default_rng(6)
    b, n = 6, 20
    a = torch.from_numpy(np.stack([_mat(rng, n, dtype, 3.0) for _ in range(b)])).to(dev)
    c = torch.from_numpy(np.stack([_mat(rng, n, dtype, 4.0) for _ in range(b)])).to(dev)
    x = a.clone().requires_grad_(True)
    loss = (c.conj() * expm_batched_autograd(x, checkpoints=2)).sum().real
    loss.backward()
    y = a.clone().requires_grad_(True)
    ((c.conj() * torch.linalg.matrix_exp(y)).sum().real).backward()
    assert torch.allclose(x.grad, y.grad, rtol=1e-10, atol=1e-12)


def test_autograd_two_forwards_of_one_shape_before_backward(dev):
    """each forward keeps its own tape: two same-shape forwards (two layers, an
    unrolled loop) and then one backward through both match torch's autograd"""
    rng = np.random.default_rng(12)
    b, n, dtype = 3, 24, torch.float64
    a1 = torch.from_numpy(np.stack([_mat(rng, n, dtype, 5.0) for _ in range(b)])).to(dev)
    a2 = torch.from_numpy(np.stack([_mat(rng, n, dtype, 0.7) for _ in range(b)])).to(dev)
    c = torch.from_numpy(np.stack([_mat(rng, n, dtype, 2.0) for _ in range(b)])).to(dev)
    x1, x2 = a1.clone().requires_grad_(True), a2.clone().requires_grad_(True)
    e1 = expm_batched_autograd(x1)
    e2 = expm_batched_autograd(x2)  # the same shape, before e1's backward
    (c * (e1 @ e2)).sum().backward()
    y1, y2 = a1.clone().requires_grad_(True), a2.clone().requires_grad_(True)
    (c * (torch.linalg.matrix_exp(y1) @ torch.linalg.matrix_exp(y2))).sum().backward()
    assert torch.allclose(x1.grad, y1.grad, rtol=1e-10, atol=1e-12)
    assert torch.allclose(x2.grad, y2.grad, rtol=1e-10, atol=1e-12)
    # and a later forward of the shape reuses a freed state without changing results
    del e1, e2
    x3 = a1.clone().requires_grad_(True)
    (c * expm_batched_autograd(x3)).sum().backward()
    y3 = a1.clone().requires_grad_(True)
    (c * torch.linalg.matrix_exp(y3)).sum().backward()
    assert torch.allclose(x3.grad, y3.grad, rtol=1e-10, atol=1e-12)


def test_autograd_non_finite_member_forward_is_nan(dev):
    rng = np.random.default_rng(13)
    a = np.stack([_mat(rng, 10, torch.float32, 1.0) for _ in range(3)])
    a[1, 0, 0] = np.nan
    e = expm_batched_autograd(torch.from_numpy(a).to(dev))
    e = e.detach().cpu().numpy()
    assert np.all(np.isnan(e[1])) and np.all(np.isfinite(e[[0, 2]]))


def test_autograd_unitary_generator_batch(dev):
    """the differentiated workload's shape at small size: A = -i t H, H = (G + G^H)/2,
    L = Re tr(C^H exp(A)) with C unitary, complex64"""
    rng = np.random.default_rng(7)
    b, n = 8, 32
    gs = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
    h = (gs + np.conj(np.transpose(gs, (0, 2, 1)))) / 2
    a = (-1j * 0.5 * h).astype(np.complex64)
    q, _ = np.linalg.qr(rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n)))
    c = torch.from_numpy(q.astype(np.complex64)).to(dev)
    x = torch.from_numpy(a).to(dev).requires_grad_(True)
    (c.conj() * expm_batched_autograd(x)).sum().real.backward()
    for i in range(b):
        ref = oracle.frechet_adjoint(a[i].astype(np.complex128), q[i])
        assert _rel(x.grad[i].cpu().numpy(), ref) <= 1e-4


@pytest.mark.parametrize("dtype", [torch.complex64, torch.float64])
def test_adjoint_every_tile_gives_the_same_bits(dev, dtype):
    """every backward product is the NN workgroup product after an exact
    conjugate transpose: the tile decides which lanes compute an element, not its
    arithmetic, so every tile gives the same gradient bits"""
    rng = np.random.default_rng(8)
    specs = [(7, 0.5), (45, 6.0), (130, 12.0)]
    mats = [_mat(rng, n, dtype, t) for n, t in specs]
    grads = [_mat(rng, n, dtype, 3.0) for n, _ in specs]
    saved = cg.batched_tuning(dtype, adjoint=True)
    try:
        ref = None
        for cfg in range(_lib.BATCH_TILES):
            cg.set_batched_tuning(dtype, [cfg] * 8, adjoint=True)
            _, ds, _ = _run(dev, mats, grads)
            if ref is None:
                ref = ds
            for x, y in zip(ds, ref):
                assert np.array_equal(x, y), cfg
    finally:
        cg.set_batched_tuning(dtype, saved, adjoint=True)


def test_adjoint_autotune_small(dev):
    saved = cg.batched_tuning(torch.complex64, adjoint=True)
    fwd = cg.batched_tuning(torch.complex64)
    try:
        table, ms = cg.batched_autotune(torch.complex64, nmax=40, adjoint=True)
        assert all(0 <= c < _lib.BATCH_TILES for c in table)
        assert cg.batched_tuning(torch.complex64, adjoint=True) == table
        assert cg.batched_tuning(torch.complex64) == fwd  # the forward table is untouched
        for b in (0, 1):
            assert min(t for t in ms[b] if t > 0) == ms[b][table[b]]
        rng = np.random.default_rng(9)
        a, g = _mat(rng, 33, torch.complex64, 2.0), _mat(rng, 33, torch.complex64, 1.0)
        _, ds, _ = _run(dev, [a], [g])
        ref = oracle.frechet_adjoint(a.astype(np.complex128), g.astype(np.complex128))
        assert _rel(ds[0], ref) <= 1e-4
    finally:
        cg.set_batched_tuning(torch.complex64, saved, adjoint=True)
