"""CPU: the oracle's restatement of cg_unitary (oracle/cg.py) pinned by what the
mathematics fixes, before it checks the HIP path.

* Table 1's Vandermonde solve recovers an exact polynomial from its samples and
  agrees with numpy.polyfit; its root bracket agrees with numpy.roots.
* Table 2's centred DFT agrees with numpy.fft; the trig-root scan finds the known
  zeros of cos(3 theta) and sin(2 theta) + 1/2.
* A derivative sample is 1/2 of the central difference of J(exp(t H) W) -- a
  dropped factor 2, a sign or a conjugate in the sample fails it.
* The power iteration never exceeds ||H||_2 and lands within its tolerance of it.
* The solver maximises the Brockett criterion to R's eigenvalues in ascending order
  (the reference's own oracle, test/cg_unitary/cg_unitary_tests.cpp:309-345), for
  real and complex R and both searches.
"""
import numpy as np
import pytest

import oracle
from oracle import cg


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


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_table1_recovers_polynomial(order):
    rng = np.random.default_rng(order)
    c_true = rng.standard_normal(order + 1)
    h, scale = 0.37, -2.0
    x = np.arange(order + 1) * h
    dots = np.polyval(c_true[::-1], x) / scale
    c = cg.table1_coeffs(dots, order, h, scale)
    np.testing.assert_allclose(c, c_true, rtol=1e-10, atol=1e-12)
    fit = np.polyfit(x, dots * scale, order)[::-1]
    np.testing.assert_allclose(c, fit, rtol=1e-8, atol=1e-10)


def test_poly_first_root_matches_numpy_roots():
    c = np.polynomial.polynomial.polyfromroots([0.3, 0.7, -1.0]) * -2.5
    mu, found = cg.poly_first_root(c, 1.0)
    assert found == 1 and abs(mu - 0.3) <= 1e-12
    roots = np.roots(c[::-1])
    want = min(r.real for r in roots if abs(r.imag) < 1e-12 and 0 < r.real <= 1.0)
    assert abs(mu - want) <= 1e-12
    assert cg.poly_first_root(c, 0.25) == (0.0, 0)  # no root in (0, 0.25]
    assert cg.poly_first_root(np.array([1.0, 0.0]), 1.0) == (0.0, 0)  # constant
    # the root at upper itself is bracketed by the last grid cell
    mu, found = cg.poly_first_root(np.array([-0.5, 1.0]), 0.5)
    assert found == 1 and abs(mu - 0.5) <= 1e-12


@pytest.mark.parametrize("num", [3, 7, 31])
def test_hann_dft_matches_numpy_fft(num):
    rng = np.random.default_rng(num)
    dots = rng.standard_normal(num)
    c = cg.hann_dft(dots, num, 2.0)
    i = np.arange(num)
    x = (0.5 - 0.5 * np.cos(2 * np.pi * (i + 1) / (num + 1))) * 2.0 * dots
    X = np.fft.fft(x)
    half = (num - 1) // 2
    want = np.array([X[(k - half) % num] for k in range(num)])
    np.testing.assert_allclose(c, want, rtol=1e-12, atol=1e-12)
    # real samples: c_{-k} = conj(c_k)
    np.testing.assert_allclose(c[::-1], c.conj(), rtol=1e-12, atol=1e-12)


def test_dft_root_args_known_trig_polynomials():
    # cos(3 theta): c_{+-3} = 1/2, zeros at pi/6 + j pi/3
    c = np.zeros(7, complex)
    c[0] = c[6] = 0.5
    got = cg.dft_root_args(c)
    want = [np.pi / 6 + j * np.pi / 3 for j in range(6)]
    np.testing.assert_allclose(got, want, atol=1e-12)
    # sin(2 theta) + 1/2 = -i/2 e^{2i theta} + i/2 e^{-2i theta} + 1/2
    c = np.zeros(5, complex)
    c[2] = 0.5
    c[4] = -0.5j
    c[0] = 0.5j
    got = cg.dft_root_args(c)
    base = [7 * np.pi / 12, 11 * np.pi / 12]  # sin(2t) = -1/2 at 2t = 7pi/6, 11pi/6
    want = sorted(base + [b + np.pi for b in base])
    np.testing.assert_allclose(got, want, atol=1e-12)
    for t in got:
        assert abs(np.sin(2 * t) + 0.5) < 1e-12


def test_select_dft_step_takes_nearest_even_crossing():
    t_dft = 2 * np.pi
    args = [0.5, 1.0, 2.0, 2.5, 4.0]  # mu = args here (t_dft = 2 pi)
    costs = [0.0, 1.0, 3.0, 2.0]  # best at i = 2: mu* = 2 * 2pi / 4 = pi
    # even-indexed crossings 0.5, 2.0, 4.0: 4.0 is nearest pi (the odd 2.5 is a maximum)
    assert cg.select_dft_step(args, costs, t_dft, True) == 4.0
    assert cg.select_dft_step(args, costs, t_dft, False) == 0.5  # best at 0: nearest even is 0.5
    assert cg.select_dft_step([], costs, t_dft, True) == 0.0


@pytest.mark.parametrize("cplx", [False, True])
def test_derivative_sample_is_half_the_cost_slope(cplx):
    n = 6
    R = rand_herm(n, 1, cplx)
    cost = cg.Brockett(R, np.arange(1, n + 1).astype(R.dtype))
    W = rand_unitary(n, 2, cplx)
    Psi = cost.gradient(W)
    H = cg.riemannian_gradient(Psi, W)
    step, num = 0.05, 4
    dots, costs = cg.sample_derivative(W, H, Psi, cost, step, num, want_cost=True)

    def J(t):
        Wt = oracle.expm(t * H) @ W
        return cost.value(Wt, cost.gradient(Wt))

    eps = 1e-5
    for i in range(num):
        t = i * step
        slope = (J(t + eps) - J(t - eps)) / (2 * eps)
        assert abs(2 * dots[i] - slope) <= 1e-6 * max(1.0, abs(slope)), (i, dots[i], slope)
        assert abs(costs[i] - J(t)) <= 1e-12 * max(1.0, abs(J(t)))


@pytest.mark.parametrize("cplx", [False, True])
def test_power_iteration_bounds_spectral_norm(cplx):
    rng = np.random.default_rng(4)
    for n in (5, 40):
        X = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
        H = X - X.conj().T
        om, v, it = cg.spectral_radius(H)
        s2 = np.linalg.norm(H, 2)
        assert om <= s2 * (1 + 1e-12)
        assert om >= s2 * (1 - 2e-3), (om, s2, it)
        assert abs(np.linalg.norm(v) - 1) < 1e-12
    assert cg.spectral_radius(np.zeros((3, 3)))[0] == 0.0


@pytest.mark.parametrize("method", [cg.POLYNOMIAL, cg.DFT])
def test_search_step_improves_cost(method):
    n = 6
    R = rand_herm(n, 3, True)
    cost = cg.Brockett(R, np.arange(1, n + 1).astype(complex))
    W = rand_unitary(n, 5, True)
    Psi = cost.gradient(W)
    G = cg.riemannian_gradient(Psi, W)
    find = cg.search_dft if method == cg.DFT else cg.search_poly
    r = find(W, G, Psi, cost, cg.MAXIMIZE)
    assert r["found"] == 1 and r["mu"] > 0
    W1 = oracle.expm(r["mu"] * G) @ W
    assert cost.value(W1, cost.gradient(W1)) > cost.value(W, Psi)
    if method == cg.POLYNOMIAL:  # mu is a zero of the fitted derivative
        assert abs(np.polynomial.polynomial.polyval(r["mu"], r["coeffs"])) <= \
            1e-9 * np.abs(r["coeffs"]).max()


@pytest.mark.parametrize("n,seed,cplx,method", [(4, 1, False, cg.POLYNOMIAL),
                                                (6, 7, False, cg.POLYNOMIAL),
                                                (4, 2, True, cg.POLYNOMIAL),
                                                (6, 11, True, cg.POLYNOMIAL),
                                                (5, 13, True, cg.DFT)])
def test_brockett_maximised_to_ascending_eigenvalues(n, seed, cplx, method):
    R = rand_herm(n, seed, cplx)
    cost = cg.Brockett(R, np.arange(1, n + 1).astype(R.dtype))
    W, info = cg.cg_unitary(np.eye(n, dtype=R.dtype), cost, cg.MAXIMIZE, method=method)
    assert info["reason"] == cg.CONVERGED
    M = W.conj().T @ R @ W
    ev = np.linalg.eigvalsh(R)
    tol = 1e-5 * max(1.0, np.abs(ev).max())
    assert np.abs(W.conj().T @ W - np.eye(n)).max() <= 1e-9
    assert np.abs(M - np.diag(np.diag(M))).max() <= tol
    d = np.diag(M).real
    assert np.all(np.diff(d) >= -tol)
    np.testing.assert_allclose(np.sort(d), ev, atol=tol)
