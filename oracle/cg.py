"""CPU restatement of calaman_G's cg_unitary line searches and solver -- TEST INFRASTRUCTURE.

Only ``tests/`` import this module, as the checker of the HIP path
(calaman_g_amd/csrc/cg_unitary.hip). Everything is numpy in FP64 (complex128
for complex inputs); the one matrix exponential per search is the oracle's
Pade expm (``oracle.expm``, the C restatement of src/expm/pade.cppm).

Each function cites the reference code it restates. Where the reference's
numerics are a published method (the Abrudan-Eriksson-Koivunen line searches,
Signal Processing 89 (2009) 1704-1714, Tables 1-3) the restatement follows the
reference's own choices: the integer Vandermonde solve with partial pivoting,
the 1024-point bracket and 60 bisections, the Hann window over N + 1, the
2048-point trig-polynomial bracket and 50 bisections, and the cost-anchored step
choice.
"""
from __future__ import annotations

import numpy as np

from . import expm as _expm
from . import norm1

MAX_SAMPLES = 32  # src/cg_unitary/cg_unitary_bridge.h: kCgMaxSamples
TWO_PI = 6.283185307179586476925286766559

MINIMIZE, MAXIMIZE = 0, 1
POLYNOMIAL, DFT = 0, 1


def sign(direction: int) -> float:
    """cg_sign (src/cg_unitary/cost_function.cppm:31-34): -1 minimising, +1 maximising"""
    return 1.0 if direction == MAXIMIZE else -1.0


def _adj(x):
    return x.conj().T


# ---- the Brockett criterion (src/cg_unitary/brockett.cppm:44-130) --------------------------
class Brockett:
    """J(W) = Re trace{W^H R W N}, N = diag(nd); Psi = R W N; q = 2"""
    order = 2

    def __init__(self, R, nd):
        self.R = np.asarray(R)
        self.nd = np.asarray(nd)

    def gradient(self, W):
        return (self.R @ W) * self.nd[None, :]

    def value(self, W, Psi):
        # Re dotc(W, Psi) = Re sum conj(W) Psi (brockett.cppm:102-130)
        return float(np.real(np.vdot(W.reshape(-1, order="F"), Psi.reshape(-1, order="F"))))


def riemannian_gradient(Psi, W):
    """G = Psi W^H - W Psi^H (src/cg_unitary/cg_solver.cppm:197-222)"""
    X = Psi @ _adj(W)
    return X - _adj(X)


def dot_re(A, B) -> float:
    """Re trace{A^H B}: the real part of dotc over the packed elements
    (geodesic_search.cppm:184-230)"""
    return float(np.real(np.vdot(A.reshape(-1, order="F"), B.reshape(-1, order="F"))))


# ---- frequency bound (geodesic_search.cppm:232-336) -----------------------------------------
def spectral_radius(H, v=None):
    """|omega_max| of a skew-Hermitian H by power iteration on H^H H: at most 100
    steps, stop when the Rayleigh quotient moves by <= 1e-4 of itself; v is the
    warm start (ones when absent or zero). Returns (omega, v, iterations)."""
    n = H.shape[0]
    dt = np.complex128 if np.iscomplexobj(H) else np.float64
    v = np.zeros(n, dt) if v is None else np.array(v, dtype=dt)
    nv = np.linalg.norm(v)
    if not nv > 0:
        v = np.ones(n, dt)
        nv = np.linalg.norm(v)
    if not nv > 0:
        return 0.0, v, 0
    v = v / nv
    omega2 = 0.0
    it = 0
    for it in range(1, 101):
        y = H @ v
        rho = float(np.linalg.norm(y)) ** 2
        z = _adj(H) @ y
        nz = float(np.linalg.norm(z))
        converged = omega2 > 0 and abs(rho - omega2) <= 1e-4 * rho
        omega2 = rho
        if not nz > 0:
            break
        v = z / nz
        if converged:
            break
    return float(np.sqrt(max(omega2, 0.0))), v, it


def frequency_bound(H, use_spectral_radius=True, v=None):
    if use_spectral_radius:
        om, v, _ = spectral_radius(H, v)
        return om, v
    # the 1-norm surrogate is matrix_norm1 (geodesic_search.cppm:334): the oracle's
    # restatement with the device's fold order
    return float(norm1(H)), v


# ---- the sampling sweep (geodesic_search.cppm:338-446) --------------------------------------
def sample_derivative(W, H, Psi, cost, step: float, num: int, want_cost: bool = False):
    """dots[i] = Re dotc(H W_i, Psi(W_i)), W_i = R_i W, R_i = R_{i-1} R_1,
    R_1 = expm(step H); sample 0 is W itself with the caller's Psi. Returns the
    derivative dots and, with want_cost, the cost values."""
    R1 = _expm(step * H)
    dots = np.zeros(num)
    costs = np.zeros(num)
    Ri = None
    for i in range(num):
        if i == 0:
            Wi, Pi = W, Psi
        else:
            Ri = R1.copy() if i == 1 else Ri @ R1
            Wi = Ri @ W
            Pi = cost.gradient(Wi)
        dots[i] = dot_re(H @ Wi, Pi)
        if want_cost:
            costs[i] = cost.value(Wi, Pi)
    return dots, costs


# ---- Table 1 (src/cg_unitary/cg_unitary.cu:90-220) -------------------------------------------
def table1_coeffs(dots, order: int, h: float, scale: float):
    """the degree-`order` polynomial through scale * dots[i] at mu = i h: the
    integer Vandermonde system M[i-1][j-1] = i^j for the coefficients times h^j,
    by Gaussian elimination with partial pivoting, then unscaled by h^j"""
    jp = [scale * float(d) for d in dots[:order + 1]]
    c = np.zeros(order + 1)
    c[0] = jp[0]
    if order < 1:
        return c
    m = np.zeros((order, order + 1))
    for i in range(1, order + 1):
        p = 1.0
        for j in range(1, order + 1):
            p *= i
            m[i - 1, j - 1] = p
        m[i - 1, order] = jp[i] - jp[0]
    for col in range(order):
        piv = col + int(np.argmax(np.abs(m[col:order, col])))
        if not abs(m[piv, col]) > 0:
            return c
        if piv != col:
            m[[col, piv], col:] = m[[piv, col], col:]
        for r in range(col + 1, order):
            f = m[r, col] / m[col, col]
            m[r, col:] -= f * m[col, col:]
    b = np.zeros(order)
    for i in range(order - 1, -1, -1):
        acc = m[i, order]
        for j in range(i + 1, order):
            acc -= m[i, j] * b[j]
        b[i] = acc / m[i, i]
    hp = 1.0
    for j in range(1, order + 1):
        hp *= h
        c[j] = b[j - 1] / hp if hp > 0 else 0.0
    return c


def _horner(c, x):
    acc = c[-1]
    for j in range(len(c) - 2, -1, -1):
        acc = acc * x + c[j]
    return acc


def _straddles(fa, fb):
    return (fa < 0 and fb >= 0) or (fa > 0 and fb <= 0)


def poly_first_root(c, upper: float, grid: int = 1024, bisections: int = 60):
    """the first sign change of the polynomial over (0, upper] on a 1024-point
    grid, refined by 60 bisections; (mu, found)"""
    if not upper > 0 or len(c) < 2:
        return 0.0, 0
    step = upper / grid
    x0, f0 = 0.0, _horner(c, 0.0)
    for i in range(1, grid + 1):
        x1 = upper if i == grid else i * step
        f1 = _horner(c, x1)
        if _straddles(f0, f1):
            lo, hi, flo = x0, x1, f0
            for _ in range(bisections):
                mid = 0.5 * (lo + hi)
                fm = _horner(c, mid)
                if (flo < 0 and fm < 0) or (flo > 0 and fm > 0):
                    lo, flo = mid, fm
                else:
                    hi = mid
            root = 0.5 * (lo + hi)
            if root > 0:
                return root, 1
        x0, f0 = x1, f1
    return 0.0, 0


# ---- Table 2 (src/cg_unitary/cg_unitary.cu:222-370) ------------------------------------------
def hann_dft(dots, num: int, scale: float):
    """centred DFT c_k, k = -(N-1)/2 .. (N-1)/2, of the Hann-windowed samples
    w_i = 1/2 - 1/2 cos(2 pi (i + 1) / (N + 1))"""
    i = np.arange(num)
    w = 0.5 - 0.5 * np.cos(TWO_PI * (i + 1) / (num + 1))
    x = w * scale * np.asarray(dots[:num], dtype=np.float64)
    half = (num - 1) // 2
    k = np.arange(num) - half
    ang = -TWO_PI * np.outer(k, i) / num
    return (x[None, :] * np.cos(ang)).sum(1) + 1j * (x[None, :] * np.sin(ang)).sum(1)


def trig_eval(c, theta):
    half = (len(c) - 1) // 2
    k = np.arange(len(c)) - half
    return float(np.sum(c.real * np.cos(k * theta) - c.imag * np.sin(k * theta)))


def dft_root_args(c, grid: int = 2048, bisections: int = 50):
    """zero crossings over (0, 2 pi) of D(theta) = sum_k c_k e^{ik theta}, bracketed
    on a 2048-point grid and refined by 50 bisections, ascending, at most 32"""
    if len(c) < 2:
        return []
    step = TWO_PI / grid
    out = []
    x0, f0 = 0.0, trig_eval(c, 0.0)
    for i in range(1, grid + 1):
        if len(out) >= MAX_SAMPLES:
            break
        x1 = i * step
        f1 = trig_eval(c, x1)
        if _straddles(f0, f1) and x1 > 0:
            lo, hi, flo = x0, x1, f0
            for _ in range(bisections):
                mid = 0.5 * (lo + hi)
                fm = trig_eval(c, mid)
                if (flo < 0) == (fm < 0):
                    lo, flo = mid, fm
                else:
                    hi = mid
            root = 0.5 * (lo + hi)
            if 0 < root < TWO_PI:
                out.append(root)
        x0, f0 = x1, f1
    return out


def select_dft_step(args, costs, t_dft: float, maximizing: bool) -> float:
    """the even-indexed crossing (a minimum of the descent) nearest the best sampled
    cost's mu (cg_unitary.cu:322-370)"""
    n = len(costs)
    if not args or n < 1:
        return 0.0
    best_i, best = 0, costs[0]
    for i in range(1, n):
        if (costs[i] > best) if maximizing else (costs[i] < best):
            best, best_i = costs[i], i
    mu_star = best_i * t_dft / n
    chosen, closest = 0.0, -1.0
    for a in args[0::2]:
        mu = a * t_dft / TWO_PI
        if not mu > 0:
            continue
        d = abs(mu - mu_star)
        if closest < 0 or d < closest:
            closest, chosen = d, mu
    return chosen


# ---- the searches (geodesic_search.cppm:448-660) ----------------------------------------------
def search_poly(W, H, Psi, cost, direction, poly_order=5, use_spectral_radius=True, v=None):
    """Table 1: returns dict(mu, omega, interval, num_samples, dots, coeffs, found, v)"""
    omega, v = frequency_bound(H, use_spectral_radius, v)
    out = dict(mu=0.0, omega=omega, interval=0.0, num_samples=0, dots=None, coeffs=None,
               found=0, v=v)
    if not omega > 0:
        return out
    t_mu = TWO_PI / (cost.order * omega)
    h = t_mu / poly_order
    sg = sign(direction)
    dots, _ = sample_derivative(W, H, Psi, cost, sg * h, poly_order + 1)
    c = table1_coeffs(dots, poly_order, h, 2.0 * sg)
    mu, found = poly_first_root(c, t_mu)
    usable = found != 0 and 0 < mu <= t_mu
    out.update(mu=mu if usable else 0.0, interval=t_mu, num_samples=poly_order + 1, dots=dots,
               coeffs=c, found=int(usable))
    return out


def search_dft(W, H, Psi, cost, direction, dft_factor=5, dft_periods=10, poly_order=5,
               use_spectral_radius=True, v=None):
    """Table 2 with the polynomial search as its fallback"""
    omega, v2 = frequency_bound(H, use_spectral_radius, v)
    out = dict(mu=0.0, omega=omega, interval=0.0, num_samples=0, dots=None, found=0, v=v2)
    if not omega > 0:
        return out
    t_mu = TWO_PI / (cost.order * omega)
    num = 2 * ((dft_factor * dft_periods) // 2) + 1
    if num > MAX_SAMPLES - 1:
        num = ((MAX_SAMPLES - 1) // 2) * 2 + 1
    num = max(num, 3)
    periods = num // dft_factor
    t_dft = max(periods, 1) * t_mu
    h = t_dft / num
    sg = sign(direction)
    dots, costs = sample_derivative(W, H, Psi, cost, sg * h, num, want_cost=True)
    c = hann_dft(dots, num, 2.0 * sg)
    args = dft_root_args(c)
    mu = select_dft_step(args, costs, t_dft, direction == MAXIMIZE)
    out.update(interval=t_dft, num_samples=num, dots=dots, costs=costs, args=args)
    if mu > 0:
        out.update(mu=mu, found=1)
        return out
    return search_poly(W, H, Psi, cost, direction, poly_order, use_spectral_radius, v2)


# ---- the solver (src/cg_unitary/cg_solver.cppm:178-432) ---------------------------------------
CONVERGED, MAX_ITERATIONS, LINE_SEARCH_FAILED = 0, 1, 2


def cg_unitary(W, cost, direction, max_iterations=300, tol=1e-20, method=POLYNOMIAL,
               reset_period=0, **search):
    """Table 3: returns (W, info dict)"""
    W = np.array(W, dtype=np.complex128 if np.iscomplexobj(W) else np.float64)
    n = W.shape[0]
    sg = sign(direction)
    period = reset_period if reset_period > 0 else n * n
    info = dict(reason=MAX_ITERATIONS, iterations=0, resets=0, line_search_failures=0)
    Psi = cost.gradient(W)
    G = riemannian_gradient(Psi, W)
    D = G.copy()
    info["resets"] += 1
    gg = 0.5 * np.linalg.norm(G) ** 2
    v = None
    find = search_dft if method == DFT else search_poly
    for k in range(max_iterations):
        if k > 0 and k % period == 0:
            D = G.copy()
            info["resets"] += 1
        if gg < tol:
            info["reason"] = CONVERGED
            break
        r = find(W, D, Psi, cost, direction, v=v, **search)
        v = r["v"]
        mu = r["mu"]
        if not mu > 0:
            info["line_search_failures"] += 1
            D = G.copy()
            info["resets"] += 1
            r = find(W, D, Psi, cost, direction, v=v, **search)
            v = r["v"]
            mu = r["mu"]
            if not mu > 0:
                info["iterations"] = k
                info["reason"] = LINE_SEARCH_FAILED
                break
        info["last_step"] = mu
        W = _expm(sg * mu * D) @ W
        Psi = cost.gradient(W)
        Gn = riemannian_gradient(Psi, W)
        ggn = 0.5 * np.linalg.norm(Gn) ** 2
        num = dot_re(Gn - G, Gn)
        gamma = num / (2.0 * gg) if gg > 0 else 0.0
        D = Gn + gamma * D
        G, gg = Gn, ggn
        if not dot_re(D, G) > 0:
            D = G.copy()
            info["resets"] += 1
        info["iterations"] = k + 1
    info["gradient_norm_sq"] = gg
    return W, info
