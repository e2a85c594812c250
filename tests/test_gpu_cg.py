"""GPU: the Riemannian conjugate gradient on U(n) (clm_cg_*), the caller of the
matrix exponential, against the oracle's restatement (oracle/cg.py).

* The geodesic step W <- exp(t H) W against oracle.expm(t H) @ W, one step and a
  chain of eight, with W^H W = I held (test/cg_unitary/cg_unitary_tests.cpp:309-318).
* A line search's derivative samples, taken at the interval the device chose,
  against the oracle's at the same points; the oracle's Table 1 / Table 2 stages
  on the device's samples give the device's step. With the 1-norm frequency bound
  (bit-exact on both sides) the interval itself is the oracle's.
* The first iterations of the solver against the oracle's, iterate by iterate.
* The solver maximises the Brockett criterion to R's eigenvalues in ascending
  order for all four element types and both searches -- the reference's own suite
  (test/cg_unitary/cg_unitary_tests.cpp:349-375).
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import cg as ocg
from calaman_g_amd import _lib
from calaman_g_amd import cg

pytestmark = pytest.mark.gpu

CPLX = {torch.float32: False, torch.float64: False, torch.complex64: True, torch.complex128: True}
SINGLE = {torch.float32: True, torch.float64: False, torch.complex64: True, torch.complex128: False}


def rand_herm(n, seed, cplx):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (n, n))
    if cplx:
        a = a + 1j * rng.uniform(-1, 1, (n, n))
    return (a + a.conj().T) / 2


def rand_unitary(n, seed, cplx):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))[None, :]


def rand_skew(n, seed, cplx, norm=1.0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    h = x - x.conj().T
    return h * (norm / np.abs(h).sum(0).max())


def dev_t(a, dtype, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).to(dev)


def rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128, torch.float32, torch.complex64])
@pytest.mark.parametrize("n", [4, 33, 256])
def test_geodesic_step_matches_oracle(dev, dtype, n):
    cplx = CPLX[dtype]
    H = rand_skew(n, n, cplx, 3.0)
    W = rand_unitary(n, n + 1, cplx)
    t = -0.37
    got = cg.geodesic_step(dev_t(H, dtype, dev), t, dev_t(W, dtype, dev)).cpu().numpy()
    want = oracle.expm(t * H) @ W
    tol, utol = (2e-5, 5e-5) if SINGLE[dtype] else (1e-12, 1e-12)
    assert rel(got, want) <= tol
    assert np.abs(got.conj().T @ got - np.eye(n)).max() <= utol


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_geodesic_step_chain_stays_unitary(dev, dtype):
    """eight steps along different directions, each against the oracle's step from
    the device's previous iterate, and the accumulated W against the oracle chain"""
    n = 96
    cplx = CPLX[dtype]
    W = rand_unitary(n, 5, cplx)
    Wo = W.copy()
    w = dev_t(W, dtype, dev)
    for k in range(8):
        H = rand_skew(n, 100 + k, cplx, 0.5 + k)
        t = 0.3 * (-1) ** k
        prev = w.cpu().numpy()
        w = cg.geodesic_step(dev_t(H, dtype, dev), t, w)
        got = w.cpu().numpy()
        assert rel(got, oracle.expm(t * H) @ prev) <= 1e-12
        Wo = oracle.expm(t * H) @ Wo
        assert np.abs(got.conj().T @ got - np.eye(n)).max() <= 1e-12
    assert rel(w.cpu().numpy(), Wo) <= 1e-11


def _brockett(n, seed, dtype, dev):
    R = rand_herm(n, seed, CPLX[dtype])
    nd = np.arange(1, n + 1).astype(R.dtype)
    return R, nd, cg.BrockettCost(dev_t(R, dtype, dev), dev_t(nd, dtype, dev))


def _search_inputs(n, dtype, dev, seed=3):
    cplx = CPLX[dtype]
    R, nd, cost = _brockett(n, seed, dtype, dev)
    ocost = ocg.Brockett(R, nd)
    W = rand_unitary(n, seed + 1, cplx)
    Psi = ocost.gradient(W)
    G = ocg.riemannian_gradient(Psi, W)
    return W, G, Psi, ocost, cost


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
@pytest.mark.parametrize("n", [6, 64])
@pytest.mark.parametrize("power", [False, True])
def test_poly_search_matches_oracle(dev, dtype, n, power):
    W, G, Psi, ocost, cost = _search_inputs(n, dtype, dev)
    opts = cg.search_options(use_spectral_radius=power)
    mu, info = cg.geodesic_search(dev_t(W, dtype, dev), dev_t(G, dtype, dev),
                                  dev_t(Psi, dtype, dev), cost, cg.MAXIMIZE, cg.POLYNOMIAL, opts)
    P = 5
    om_o, _ = ocg.frequency_bound(G, power)
    if power:  # same rule, both stop within 1e-4 of |omega_max|^2
        assert abs(info["omega_max"] - om_o) <= 2e-4 * om_o
    elif CPLX[dtype]:  # the 1-norm: the modulus is a hypot on each side
        assert info["omega_max"] == pytest.approx(om_o, rel=1e-15)
    else:  # the 1-norm, bit for bit
        assert info["omega_max"] == om_o
    t_mu = info["interval"]
    assert t_mu == pytest.approx(2 * np.pi / (2 * info["omega_max"]), rel=1e-15)
    h = t_mu / P
    dots, _ = ocg.sample_derivative(W, G, Psi, ocost, h, P + 1)
    got = np.array(info["samples"])
    assert np.abs(got - dots).max() <= 1e-10 * np.abs(dots).max()
    c = ocg.table1_coeffs(got, P, h, 2.0)
    mu_o, found = ocg.poly_first_root(c, t_mu)
    assert info["found"] == found
    if power:  # |omega_max| itself: the interval holds the first extremum
        assert found == 1
    if found:
        assert abs(mu - mu_o) <= 1e-12 * t_mu
        # the step improves the cost
        W1 = oracle.expm(mu * G) @ W
        assert ocost.value(W1, ocost.gradient(W1)) > ocost.value(W, Psi)
    else:  # the 1-norm over-estimates |omega_max|: T_mu can end before the extremum
        assert mu == 0.0


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64])
def test_poly_search_single_precision(dev, dtype):
    n = 16
    W, G, Psi, ocost, cost = _search_inputs(n, dtype, dev, seed=9)
    mu, info = cg.geodesic_search(dev_t(W, dtype, dev), dev_t(G, dtype, dev),
                                  dev_t(Psi, dtype, dev), cost, cg.MAXIMIZE, cg.POLYNOMIAL,
                                  cg.search_options(use_spectral_radius=False))
    r = ocg.search_poly(W, G, Psi, ocost, ocg.MAXIMIZE, use_spectral_radius=False)
    assert abs(info["omega_max"] - r["omega"]) <= 1e-6 * r["omega"]
    got = np.array(info["samples"])
    assert np.abs(got - r["dots"]).max() <= 1e-4 * np.abs(r["dots"]).max()
    assert info["found"] == 1 and abs(mu - r["mu"]) <= 1e-3 * r["mu"]


@pytest.mark.parametrize("dtype", [torch.complex128, torch.float64])
def test_dft_search_matches_oracle(dev, dtype):
    n = 8
    W, G, Psi, ocost, cost = _search_inputs(n, dtype, dev, seed=21)
    mu, info = cg.geodesic_search(dev_t(W, dtype, dev), dev_t(G, dtype, dev),
                                  dev_t(Psi, dtype, dev), cost, cg.MAXIMIZE, cg.DFT,
                                  cg.search_options(use_spectral_radius=False))
    r = ocg.search_dft(W, G, Psi, ocost, ocg.MAXIMIZE, use_spectral_radius=False)
    assert info["method_used"] == cg.DFT and r["num_samples"] == info["num_samples"] == 31
    assert info["interval"] == pytest.approx(r["interval"], rel=1e-15)
    got = np.array(info["samples"])
    assert np.abs(got - r["dots"]).max() <= 1e-10 * np.abs(r["dots"]).max()
    assert mu > 0 and abs(mu - r["mu"]) <= 1e-10 * r["interval"]


def test_search_zero_direction_and_bad_options(dev):
    n = 5
    dtype = torch.complex128
    W, G, Psi, ocost, cost = _search_inputs(n, dtype, dev)
    z = torch.zeros((n, n), dtype=dtype, device=dev)
    mu, info = cg.geodesic_search(dev_t(W, dtype, dev), z, dev_t(Psi, dtype, dev), cost)
    assert mu == 0.0 and info["found"] == 0 and info["omega_max"] == 0.0
    with pytest.raises(_lib.CalamanError):
        cg.geodesic_search(dev_t(W, dtype, dev), dev_t(G, dtype, dev), dev_t(Psi, dtype, dev),
                           cost, opts=cg.search_options(poly_order=6))


@pytest.mark.parametrize("dtype", [torch.complex128, torch.float64])
@pytest.mark.parametrize("method", [cg.POLYNOMIAL, cg.DFT])
def test_solver_iterates_match_oracle(dev, dtype, method):
    """k = 1, 2, 3 iterations from the same start: each iterate W_k against the
    oracle solver's (the 1-norm frequency bound, so every interval is exact)"""
    n = 6
    R, nd, cost = _brockett(n, 31, dtype, dev)
    W0 = rand_unitary(n, 32, CPLX[dtype])
    for k in (1, 2, 3):
        opts = cg.default_options(dtype, max_iterations=k, method=method, use_spectral_radius=0)
        got, info = cg.cg_unitary(dev_t(W0, dtype, dev), cost, cg.MAXIMIZE, opts)
        want, oinfo = ocg.cg_unitary(W0, ocg.Brockett(R, nd), ocg.MAXIMIZE, max_iterations=k,
                                     method=method, use_spectral_radius=False)
        assert info["iterations"] == oinfo["iterations"] == k
        assert info["resets"] == oinfo["resets"]
        assert abs(info["last_step"] - oinfo["last_step"]) <= 1e-10 * oinfo["last_step"]
        assert rel(got.cpu().numpy(), want) <= 1e-10
        assert info["gradient_norm_sq"] == pytest.approx(oinfo["gradient_norm_sq"], rel=1e-8)


@pytest.mark.parametrize("dtype,n,seed,method", [
    (torch.float64, 4, 1, cg.POLYNOMIAL), (torch.float64, 6, 7, cg.POLYNOMIAL),
    (torch.complex128, 4, 2, cg.POLYNOMIAL), (torch.complex128, 6, 11, cg.POLYNOMIAL),
    (torch.float32, 4, 3, cg.POLYNOMIAL), (torch.complex64, 4, 5, cg.POLYNOMIAL),
    (torch.complex128, 5, 13, cg.DFT), (torch.complex128, 16, 17, cg.POLYNOMIAL)])
def test_brockett_maximised(dev, dtype, n, seed, method):
    R, nd, cost = _brockett(n, seed, dtype, dev)
    W0 = torch.eye(n, dtype=dtype, device=dev)
    W, info = cg.cg_unitary(W0, cost, cg.MAXIMIZE, cg.default_options(dtype, method=method))
    print(f"\nn={n} {dtype} method={method} info={info}")
    single = SINGLE[dtype]
    assert info["reason"] == cg.CONVERGED or (single and info["reason"] == cg.MAX_ITERATIONS)
    W = W.cpu().numpy().astype(np.complex128 if CPLX[dtype] else np.float64)
    M = W.conj().T @ R @ W
    ev = np.linalg.eigvalsh(R)
    tol = (2e-2 if single else 1e-5) * max(1.0, np.abs(ev).max())
    assert np.abs(W.conj().T @ W - np.eye(n)).max() <= (1e-3 if single else 1e-9)
    assert np.abs(M - np.diag(np.diag(M))).max() <= tol
    d = np.diag(M).real
    assert np.all(np.diff(d) >= -tol)
    np.testing.assert_allclose(np.sort(d), ev, atol=tol)


def test_solver_rejects_bad_arguments(dev):
    n = 4
    dtype = torch.complex128
    R, nd, cost = _brockett(n, 1, dtype, dev)
    W = torch.eye(n, dtype=dtype, device=dev)
    work = torch.empty(cg.cg_buffer_size(n, cost.buffer_size(n), dtype), dtype=torch.uint8,
                       device=dev)
    small = torch.empty(1024, dtype=torch.uint8, device=dev)
    assert cg.cg_unitary_cm(n, W, cost, cg.MAXIMIZE, small)[0] == _lib.ALLOC_FAILED
    assert cg.cg_unitary_cm(n, W, cost, 7, work)[0] == _lib.INVALID_VALUE
    assert cg.cg_unitary_cm(0, W, cost, cg.MAXIMIZE, work)[0] == _lib.INVALID_VALUE
    st, info = cg.cg_unitary_cm(n, W, cost, cg.MAXIMIZE, work,
                                cg.default_options(dtype, max_iterations=0))
    assert st == 0 and info["reason"] == cg.MAX_ITERATIONS and info["iterations"] == 0


def test_reorthonormalization_keeps_the_result(dev):
    """a Newton-Schulz step every iteration leaves the optimum where it was"""
    n = 6
    dtype = torch.complex128
    R, nd, cost = _brockett(n, 11, dtype, dev)
    W0 = torch.eye(n, dtype=dtype, device=dev)
    a, ia = cg.cg_unitary(W0, cost, cg.MAXIMIZE, cg.default_options(dtype))
    b, ib = cg.cg_unitary(W0, cost, cg.MAXIMIZE,
                          cg.default_options(dtype, reorthogonalize_period=1))
    assert ia["reason"] == ib["reason"] == cg.CONVERGED
    A, B = a.cpu().numpy(), b.cpu().numpy()
    assert np.abs(B.conj().T @ B - np.eye(n)).max() <= 1e-13
    Ma, Mb = A.conj().T @ R @ A, B.conj().T @ R @ B
    np.testing.assert_allclose(np.diag(Ma).real, np.diag(Mb).real, atol=1e-8)
