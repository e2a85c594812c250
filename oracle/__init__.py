"""CPU oracle for calaman_G's matrix-exponential path -- TEST INFRASTRUCTURE.

This package is the checker, never the thing measured or shipped: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
import it. It wraps ``liboracle.so`` (``oracle.c``, a C restatement of the
reference's Pade ladder, norms, ?lascl, Paterson-Stockmeyer and Horner, plus
the build's Taylor/Paterson-Stockmeyer algorithm) with numpy-facing helpers.

Matrices are passed as numpy arrays in Fortran (column-major) order; complex
matrices as complex128, whose memory layout is the interleaved (re, im) pairs
the C code expects.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc, no GPU)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_pade_coeffs.restype = _dp
        L.orc_pade_theta_d.restype = ctypes.c_double
        L.orc_pade_theta_s.restype = ctypes.c_float
        L.orc_pade_theta_s.argtypes = [ctypes.c_int]
        L.orc_expm_plan_d.argtypes = [ctypes.c_double, _ip]
        L.orc_expm_plan_s.argtypes = [ctypes.c_float, _ip]
        for f in ("orc_norm1_d", "orc_norm1_z"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        for f in ("orc_lange_d", "orc_lange_z"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, _dp,
                                      ctypes.c_size_t]
        for f in ("orc_lascl_d", "orc_lascl_z"):
            getattr(L, f).argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_size_t,
                                      ctypes.c_size_t, _dp, ctypes.c_size_t]
        for f in ("orc_norm1_s", "orc_norm1_c"):
            getattr(L, f).restype = ctypes.c_float
            getattr(L, f).argtypes = [ctypes.c_int, _fp, ctypes.c_int]
        for f in ("orc_lange_s", "orc_lange_c"):
            getattr(L, f).restype = ctypes.c_float
            getattr(L, f).argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, _fp,
                                      ctypes.c_size_t]
        for f in ("orc_lascl_s", "orc_lascl_c"):
            getattr(L, f).argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_size_t,
                                      ctypes.c_size_t, _fp, ctypes.c_size_t]
        L.orc_gemm_d.argtypes = [ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, _dp,
                                 ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int]
        L.orc_gemm_z.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                 _dp, ctypes.c_int]
        for f in ("orc_getrf_d", "orc_getrf_z"):
            getattr(L, f).argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip]
        for f in ("orc_getrs_d", "orc_getrs_z"):
            getattr(L, f).argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp,
                                      ctypes.c_int]
        for f in ("orc_pade_d", "orc_pade_z"):
            getattr(L, f).argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                      ctypes.c_int, _ip]
        for f in ("orc_expm_d", "orc_expm_z"):
            getattr(L, f).argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip]
        for f in ("orc_expm_taylor_d", "orc_expm_taylor_z"):
            getattr(L, f).argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip,
                                      _dp]
        L.orc_ps_plan.argtypes = [ctypes.c_int, ctypes.c_int, _ip]
        for f in ("orc_ps_d", "orc_ps_z"):
            getattr(L, f).argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                      ctypes.c_int, ctypes.c_int]
        for f in ("orc_horner_d", "orc_horner_z"):
            getattr(L, f).argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                      ctypes.c_int]
        for f in ("orc_taylor_theta_d", "orc_taylor_theta_s"):
            getattr(L, f).restype = ctypes.c_double
        L.orc_inv_factorials.restype = _dp
        L.orc_taylor_plan_d.argtypes = [ctypes.c_double, _ip]
        L.orc_taylor_plan_s.argtypes = [ctypes.c_double, _ip]
        L.orc_taylor_plan_powers_d.argtypes = [_dp, ctypes.c_int, _ip]
        L.orc_taylor_plan_powers_s.argtypes = [_dp, ctypes.c_int, _ip]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def _is_single(a) -> bool:
    return np.asarray(a).dtype in (np.float32, np.complex64)


def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _is_complex(a) -> bool:
    return np.iscomplexobj(a)


def set_threads(t: int) -> None:
    lib().orc_set_num_threads(int(t))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ---- Pade ladder -------------------------------------------------------------
PADE_DEGREES = (3, 5, 7, 9, 13)


def pade_coeffs(m: int):
    p = lib().orc_pade_coeffs(m)
    return None if not p else [p[j] for j in range(m + 1)]


def pade_theta(m: int, single: bool = False) -> float:
    return float(lib().orc_pade_theta_s(m) if single else lib().orc_pade_theta_d(m))


def pade_num_gemms(m: int) -> int:
    return int(lib().orc_pade_num_gemms(m))


def expm_plan(norm1: float, single: bool = False):
    out = (ctypes.c_int * 3)()
    if single:
        lib().orc_expm_plan_s(ctypes.c_float(norm1), out)
    else:
        lib().orc_expm_plan_d(norm1, out)
    return tuple(out)


# ---- norms -----------------------------------------------------------------
def norm1(A) -> float:
    """max_j sum_i |A_ij|; float32/complex64 inputs reduce in float."""
    c = _is_complex(A)
    if _is_single(A):
        A = _fortran(A, np.complex64 if c else np.float32)
        f = lib().orc_norm1_c if c else lib().orc_norm1_s
        return float(f(A.shape[0], _fptr(A), A.shape[0]))
    A = _fortran(A, np.complex128 if c else np.float64)
    n = A.shape[0]
    f = lib().orc_norm1_z if c else lib().orc_norm1_d
    return float(f(n, _ptr(A), n))


LANGE = {"max_abs": 0, "one": 1, "inf": 2, "frobenius": 3}


def lange(which: str, A) -> float:
    c = _is_complex(A)
    if _is_single(A):
        A = _fortran(A, np.complex64 if c else np.float32)
        m, n = A.shape
        f = lib().orc_lange_c if c else lib().orc_lange_s
        return float(f(LANGE[which], m, n, _fptr(A), max(m, 1)))
    A = _fortran(A, np.complex128 if c else np.float64)
    m, n = A.shape
    f = lib().orc_lange_z if c else lib().orc_lange_d
    return float(f(LANGE[which], m, n, _ptr(A), max(m, 1)))


def lascl(cfrom: float, cto: float, A):
    """Returns (status, scaled copy)."""
    c = _is_complex(A)
    if _is_single(A):
        A = _fortran(A, np.complex64 if c else np.float32).copy(order="F")
        m, n = A.shape
        f = lib().orc_lascl_c if c else lib().orc_lascl_s
        st = f(cfrom, cto, m, n, _fptr(A), max(m, 1))
        return int(st), A
    A = _fortran(A, np.complex128 if c else np.float64).copy(order="F")
    m, n = A.shape
    f = lib().orc_lascl_z if c else lib().orc_lascl_d
    st = f(cfrom, cto, m, n, _ptr(A), max(m, 1))
    return int(st), A


# ---- BLAS / LAPACK pieces -------------------------------------------------------
def gemm(A, B, alpha=1.0, beta=0.0, C=None):
    c = _is_complex(A) or _is_complex(B) or (C is not None and _is_complex(C))
    dt = np.complex128 if c else np.float64
    A = _fortran(A, dt)
    B = _fortran(B, dt)
    n = A.shape[0]
    C = np.zeros((n, n), dtype=dt, order="F") if C is None else _fortran(C, dt).copy(order="F")
    if c:
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.array([complex(beta).real, complex(beta).imag])
        lib().orc_gemm_z(n, _ptr(al), _ptr(A), n, _ptr(B), n, _ptr(be), _ptr(C), n)
    else:
        lib().orc_gemm_d(n, float(alpha), _ptr(A), n, _ptr(B), n, float(beta), _ptr(C), n)
    return C


def lu_solve(A, B):
    """Solve A X = B with the oracle's getrf/getrs. Returns (X, info)."""
    c = _is_complex(A) or _is_complex(B)
    dt = np.complex128 if c else np.float64
    LU = _fortran(A, dt).copy(order="F")
    X = _fortran(B, dt).copy(order="F")
    n = LU.shape[0]
    nrhs = X.shape[1]
    ipiv = (ctypes.c_int * n)()
    if c:
        info = lib().orc_getrf_z(n, _ptr(LU), n, ipiv)
        lib().orc_getrs_z(n, nrhs, _ptr(LU), n, ipiv, _ptr(X), n)
    else:
        info = lib().orc_getrf_d(n, _ptr(LU), n, ipiv)
        lib().orc_getrs_d(n, nrhs, _ptr(LU), n, ipiv, _ptr(X), n)
    return X, int(info)


def getrf(A):
    """The oracle's partial-pivot LU (?getrf semantics): (LU, ipiv 1-based, info)"""
    c = _is_complex(A)
    dt = np.complex128 if c else np.float64
    LU = _fortran(A, dt).copy(order="F")
    n = LU.shape[0]
    ipiv = (ctypes.c_int * max(n, 1))()
    f = lib().orc_getrf_z if c else lib().orc_getrf_d
    info = f(n, _ptr(LU), n, ipiv)
    return LU, np.array(ipiv[:n], dtype=np.int64), int(info)


# ---- Pade / expm ------------------------------------------------------------------
def pade(m: int, A):
    c = _is_complex(A)
    dt = np.complex128 if c else np.float64
    A = _fortran(A, dt)
    n = A.shape[0]
    R = np.zeros((n, n), dtype=dt, order="F")
    info = (ctypes.c_int * 2)()
    f = lib().orc_pade_z if c else lib().orc_pade_d
    st = f(m, n, _ptr(A), n, _ptr(R), n, info)
    if st != 0:
        raise ValueError(f"orc_pade status {st}")
    return R


def expm(A, return_plan: bool = False):
    """The reference expm restated: Pade ladder + scaling and squaring."""
    c = _is_complex(A)
    dt = np.complex128 if c else np.float64
    A = _fortran(A, dt)
    n = A.shape[0]
    E = np.zeros((n, n), dtype=dt, order="F")
    plan = (ctypes.c_int * 3)()
    f = lib().orc_expm_z if c else lib().orc_expm_d
    st = f(n, _ptr(A), n, _ptr(E), n, plan)
    if st != 0:
        raise ValueError(f"orc_expm status {st}")
    return (E, tuple(plan)) if return_plan else E


def expm_status(A) -> int:
    c = _is_complex(A)
    dt = np.complex128 if c else np.float64
    A = _fortran(A, dt)
    n = A.shape[0]
    E = np.zeros((n, n), dtype=dt, order="F")
    f = lib().orc_expm_z if c else lib().orc_expm_d
    return int(f(n, _ptr(A), n, _ptr(E), n, None))


def expm_taylor(A, return_plan: bool = False):
    """The build's algorithm on the CPU: the norm-only plan's power bank formed
    from A, the plan from the exact 1-norms of A .. A^q (taylor_plan_powers),
    Taylor / Paterson-Stockmeyer at A / 2^s, s squarings. The plan, when
    returned, is a dict: degree, s, num_gemms, ps_block, norm_powers (q) and
    power_norm1 (||A^k||_1 for k = 2 .. q)."""
    c = _is_complex(A)
    dt = np.complex128 if c else np.float64
    A = _fortran(A, dt)
    n = A.shape[0]
    E = np.zeros((n, n), dtype=dt, order="F")
    plan = (ctypes.c_int * 5)()
    pnorm = np.zeros(4)
    f = lib().orc_expm_taylor_z if c else lib().orc_expm_taylor_d
    st = f(n, _ptr(A), n, _ptr(E), n, plan, _ptr(pnorm))
    if st != 0:
        raise ValueError(f"orc_expm_taylor status {st}")
    if not return_plan:
        return E
    q = int(plan[4])
    info = dict(zip(("degree", "s", "num_gemms", "ps_block", "norm_powers"), tuple(plan)))
    info["power_norm1"] = [float(v) for v in pnorm[:max(q - 1, 0)]]
    return E, info


# ---- Frechet derivative -----------------------------------------------------------
def frechet(A, E):
    """L(A, E), the Frechet derivative of exp at A in direction E, as the (1, 2)
    block of exp([[A, E], [0, A]]) (Mathias 1996; Higham 2008, Thm 3.6), with
    the oracle's expm of the 2n x 2n block matrix."""
    c = _is_complex(A) or _is_complex(E)
    dt = np.complex128 if c else np.float64
    A = np.asarray(A, dtype=dt)
    E = np.asarray(E, dtype=dt)
    n = A.shape[0]
    big = np.zeros((2 * n, 2 * n), dtype=dt)
    big[:n, :n] = A
    big[n:, n:] = A
    big[:n, n:] = E
    return np.asarray(expm(big))[:n, n:].copy()


def frechet_adjoint(A, G):
    """The gradient of Re tr(G^H exp(A)) with respect to A, in the convention
    grad = dL/dRe + i dL/dIm: L(A^H, G), the adjoint of E -> L(A, E) under the
    Frobenius inner product."""
    A = np.asarray(A)
    return frechet(A.conj().T, G)


# ---- polynomials ------------------------------------------------------------------
def ps_plan(degree: int, s_requested: int = 0):
    out = (ctypes.c_int * 5)()
    lib().orc_ps_plan(degree, s_requested, out)
    return dict(zip(("s", "r", "num_powers", "num_blocks", "num_gemms"), tuple(out)))


def ps_cost(degree: int, s: int) -> int:
    return int(lib().orc_ps_cost(degree, s))


def paterson_stockmeyer(coeffs, A, s_requested: int = 0):
    c = _is_complex(A) or _is_complex(coeffs)
    dt = np.complex128 if c else np.float64
    A = _fortran(A, dt)
    co = np.ascontiguousarray(np.asarray(coeffs, dtype=dt))
    n = A.shape[0]
    P = np.zeros((n, n), dtype=dt, order="F")
    f = lib().orc_ps_z if c else lib().orc_ps_d
    st = f(n, _ptr(co), len(co) - 1, _ptr(A), n, _ptr(P), n, s_requested)
    if st != 0:
        raise ValueError(f"orc_ps status {st}")
    return P


def horner(coeffs, A):
    c = _is_complex(A) or _is_complex(coeffs)
    dt = np.complex128 if c else np.float64
    A = _fortran(A, dt)
    co = np.ascontiguousarray(np.asarray(coeffs, dtype=dt))
    n = A.shape[0]
    P = np.zeros((n, n), dtype=dt, order="F")
    f = lib().orc_horner_z if c else lib().orc_horner_d
    st = f(n, _ptr(co), len(co) - 1, _ptr(A), n, _ptr(P), n)
    if st != 0:
        raise ValueError(f"orc_horner status {st}")
    return P


# ---- Taylor ladder -------------------------------------------------------------------
TAYLOR_DEGREES = (1, 2, 3, 5, 8, 11, 15, 19, 24, 29)


def taylor_theta(degree: int, single: bool = False) -> float:
    f = lib().orc_taylor_theta_s if single else lib().orc_taylor_theta_d
    return float(f(degree))


def taylor_plan(norm1: float, single: bool = False):
    """The plan from the 1-norm alone."""
    out = (ctypes.c_int * 4)()
    (lib().orc_taylor_plan_s if single else lib().orc_taylor_plan_d)(norm1, out)
    return dict(zip(("degree", "s", "num_gemms", "ps_block"), tuple(out)))


def taylor_plan_powers(norms, single: bool = False):
    """The plan from norms[k-1] = ||A^k||_1, k = 1 .. q: s from
    alpha_p = max(||A^p||^(1/p), ||A^(p+1)||^(1/(p+1))) (Al-Mohy & Higham 2009,
    Thm 4.2), the q - 1 powers already formed counted once."""
    pn = np.ascontiguousarray(norms, dtype=np.float64)
    out = (ctypes.c_int * 4)()
    f = lib().orc_taylor_plan_powers_s if single else lib().orc_taylor_plan_powers_d
    f(_ptr(pn), len(pn), out)
    return dict(zip(("degree", "s", "num_gemms", "ps_block"), tuple(out)))


def inv_factorials():
    p = lib().orc_inv_factorials()
    return [p[k] for k in range(33)]


# ---- FP8 (OCP e4m3) tile quantization and the FP8-tail evaluation ---------------------
# No reference counterpart: the precision variant's own arithmetic, restated in
# numpy. e4m3 (OCP 8-bit floating point, E4M3 encoding): 1 sign, 4 exponent bits
# (bias 7), 3 fraction bits; normal magnitudes 2^-6 .. 448, subnormal quantum
# 2^-9, no infinities. Pinned against torch.float8_e4m3fn conversion and the
# format's code table in tests/test_oracle_fp8.py.
E4M3_MAX = 448.0
FP8_TILE = 128


def e4m3_round(x):
    """Round-to-nearest-even onto e4m3 of float64 values with |x| <= 448, after a
    first rounding to float32 (the device converts FP64 -> FP32 -> e4m3)."""
    x = np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)
    if np.any(np.abs(x) > E4M3_MAX):
        raise ValueError("e4m3_round: magnitude beyond 448")
    _, ex = np.frexp(x)                        # |x| = f 2^ex, f in [0.5, 1)
    e = np.maximum(ex - 1, -6)                 # binade exponent, subnormals at 2^-6
    q = np.ldexp(1.0, e - 3)                   # quantum: 3 fraction bits
    return np.round(x / q) * q                 # exact division; np.round is half-to-even


def fp8_tile_exponent(mx: float) -> int:
    """The largest e with mx * 2^e <= 448 (0 for a zero tile)."""
    if not mx > 0.0 or not np.isfinite(mx):
        return 0
    f, k = np.frexp(mx)                        # mx = f 2^k; 448 = 0.875 2^9
    return int(9 - k) if f <= 0.875 else int(8 - k)


def fp8_quantize(X, tile: int = FP8_TILE):
    """X with every tile x tile block rounded to e4m3 at its own power-of-two
    scale and scaled back: Q(X) = 2^-e e4m3(X 2^e) per tile. Returns (Q(X), exps)
    with exps[ti, tj] the tile exponents."""
    X = np.asarray(X, dtype=np.float64)
    rows, cols = X.shape
    tr, tc = -(-rows // tile), -(-cols // tile)
    out = np.empty_like(X)
    exps = np.zeros((tr, tc), dtype=np.int64)
    for ti in range(tr):
        for tj in range(tc):
            blk = X[ti * tile:(ti + 1) * tile, tj * tile:(tj + 1) * tile]
            e = fp8_tile_exponent(float(np.abs(blk).max()) if blk.size else 0.0)
            exps[ti, tj] = e
            out[ti * tile:(ti + 1) * tile, tj * tile:(tj + 1) * tile] = \
                np.ldexp(e4m3_round(np.ldexp(blk, e)), -e)
    return out, exps


def gemm_fp8(A, B):
    """Q(A) Q(B) in float64: the product the FP8 MFMA path forms, up to its FP32
    summation."""
    return fp8_quantize(A)[0] @ fp8_quantize(B)[0]


def expm_fp8tail_model(A, degree: int, s: int, min_degree: int = 4):
    """The FP8-tail evaluation at a given plan (Taylor degree, squarings s) in
    float64: Paterson-Stockmeyer over A / 2^s with the block size of ps_plan;
    every product whose result holds only Taylor terms of degree >= min_degree
    (the power A^k for k >= min_degree, the Horner product R_(j+1) X with
    (j+1) b >= min_degree) takes e4m3-quantized operands (gemm_fp8), every
    other product and the squarings are exact float64 products."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    As = np.ldexp(A, -s)
    c = inv_factorials()
    p = ps_plan(degree, 0)
    b, r = p["s"], p["r"]
    max_power = b if r >= 1 else degree
    pw = {0: np.eye(n), 1: As}
    for k in range(2, max_power + 1):
        pw[k] = gemm_fp8(pw[k - 1], As) if k >= min_degree else pw[k - 1] @ As

    def block(j):
        terms = min(degree - j * b + 1, b)
        return sum(c[j * b + t] * pw[t] for t in range(terms))

    X = pw[b] if (r >= 1 and b >= 2) else As
    R = block(r)
    for j in range(r - 1, -1, -1):
        prod = gemm_fp8(R, X) if (j + 1) * b >= min_degree else R @ X
        R = prod + block(j)
    for _ in range(s):
        R = R @ R
    return R
