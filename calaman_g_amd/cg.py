"""Python host mirror of calaman_G's cg_unitary (src/cg_unitary/) over the C ABI of
include/calaman_cg.h.

Riemannian conjugate gradient on U(n): minimise or maximise a real J(W) subject
to W^H W = I along geodesics exp(mu H) W, the step from the almost-periodicity
line searches of Abrudan, Eriksson & Koivunen (Signal Processing 89 (2009),
Tables 1-3). The arithmetic -- the MFMA products, the matrix exponentials, the
reductions, the power iteration and the searches' scalar stages -- runs in
libcalaman_expm.so; this module moves torch tensors in and out.

As in ``expm``, ``*_cm`` functions take tensors as column-major device buffers
(the C calling convention) and the convenience functions take ordinary matrices.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from ._lib import CgBrockett, CgCost, CgInfo, CgOptions, CgSearchInfo, CgSearchOptions, check, lib
from .expm import _is_single, _on_stream, _ptr, _require_device, _stream_ptr, _suffix, from_cm, \
    to_cm

MINIMIZE, MAXIMIZE = _lib.CG_MINIMIZE, _lib.CG_MAXIMIZE
POLYNOMIAL, DFT = _lib.CG_SEARCH_POLYNOMIAL, _lib.CG_SEARCH_DFT
CONVERGED, MAX_ITERATIONS, LINE_SEARCH_FAILED, NUMERICAL_FAILURE = (
    _lib.CG_CONVERGED, _lib.CG_MAX_ITERATIONS, _lib.CG_LINE_SEARCH_FAILED,
    _lib.CG_NUMERICAL_FAILURE)

__all__ = ["MINIMIZE", "MAXIMIZE", "POLYNOMIAL", "DFT", "CONVERGED", "MAX_ITERATIONS",
           "LINE_SEARCH_FAILED", "NUMERICAL_FAILURE", "BrockettCost", "default_options",
           "search_options", "cg_buffer_size", "cg_unitary", "cg_unitary_cm", "geodesic_search",
           "geodesic_search_cm", "geodesic_step", "geodesic_step_cm"]


def default_options(dtype: torch.dtype = torch.float64, **overrides) -> CgOptions:
    """CgOptions<R> (src/cg_unitary/cg_solver.cppm:64-103): 300 iterations, tol 1e-20
    (1e-10 for float32/complex64), the polynomial search with P = 5, power iteration,
    resets every n^2 iterations, no re-orthonormalization. Keyword overrides name
    CgOptions fields, or CgSearchOptions fields for the search."""
    o = CgOptions()
    getattr(lib(), "clm_cg_default_options_s" if _is_single(dtype) else
            "clm_cg_default_options_d")(ctypes.byref(o))
    for k, v in overrides.items():
        if k in dict(CgSearchOptions._fields_):
            setattr(o.search, k, int(v))
        else:
            setattr(o, k, v)
    return o


def search_options(poly_order: int = 5, dft_factor: int = 5, dft_periods: int = 10,
                   use_spectral_radius: bool = True) -> CgSearchOptions:
    """GeodesicSearchOptions (src/cg_unitary/geodesic_search.cppm:68-98)"""
    return CgSearchOptions(int(poly_order), int(dft_factor), int(dft_periods),
                           int(bool(use_spectral_radius)))


class BrockettCost:
    """brockett_cost<T> (src/cg_unitary/brockett.cppm:29-130): J(W) = trace{W^H R W N}
    with R Hermitian and N = diag(nd). Holds R (column-major) and nd on the device
    for as long as the cost is used."""

    def __init__(self, R: torch.Tensor, nd: torch.Tensor):
        _require_device(R, "R")
        n = R.shape[0]
        if R.shape != (n, n) or nd.numel() != n:
            raise ValueError("R must be n x n and nd hold n entries")
        self.dtype = R.dtype
        self.n = n
        self.Rcm = to_cm(R)
        self.nd = nd.to(device=R.device, dtype=R.dtype).contiguous()
        self.spec = CgBrockett(self.Rcm.data_ptr(), n, self.nd.data_ptr())
        self.cost = CgCost()
        check(getattr(lib(), f"clm_cg_brockett_cost_{_suffix(R.dtype)}")(
            ctypes.byref(self.spec), ctypes.byref(self.cost)), "clm_cg_brockett_cost")

    def buffer_size(self, n: int) -> int:
        return int(self.cost.buffer_size(self.cost.user, int(n)))


def cg_buffer_size(n: int, cost_bytes: int, dtype: torch.dtype = torch.complex128) -> int:
    """cg_unitary_bufferSize<T> (src/cg_unitary/buffer_size.cppm:120-146)"""
    out = ctypes.c_size_t()
    check(getattr(lib(), f"clm_cg_unitary_buffer_size_{_suffix(dtype)}")(
        int(n), int(cost_bytes), ctypes.byref(out)), "clm_cg_unitary_buffer_size")
    return int(out.value)


def _work_for(n: int, cost: BrockettCost, dtype, device, work):
    need = cg_buffer_size(n, cost.buffer_size(n), dtype)
    if work is None:
        work = torch.empty(need, dtype=torch.uint8, device=device)
    return work


# ---- column-major layer ------------------------------------------------------------------------
def cg_unitary_cm(n: int, dW: torch.Tensor, cost: BrockettCost, direction: int,
                  dWork: torch.Tensor, opts: Optional[CgOptions] = None,
                  stream: Optional[torch.cuda.Stream] = None, ctx=None) -> tuple:
    """clm_cg_unitary_{s,d,c,z}; returns (status, info dict)"""
    info = CgInfo()
    f = getattr(lib(), f"clm_cg_unitary_{_suffix(dW.dtype)}")
    st = f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(n), _ptr(dW),
           ctypes.byref(cost.cost), int(direction), _ptr(dWork),
           dWork.numel() * dWork.element_size(), ctypes.byref(opts) if opts is not None else None,
           ctypes.byref(info))
    return int(st), info.as_dict()


def geodesic_search_cm(n: int, dW: torch.Tensor, dH: torch.Tensor, ldh: int, dPsi: torch.Tensor,
                       cost: BrockettCost, direction: int, method: int, dWork: torch.Tensor,
                       opts: Optional[CgSearchOptions] = None,
                       stream: Optional[torch.cuda.Stream] = None, ctx=None) -> tuple:
    """clm_cg_geodesic_search_{s,d,c,z}; returns (status, mu, info dict)"""
    info = CgSearchInfo()
    mu = ctypes.c_double()
    f = getattr(lib(), f"clm_cg_geodesic_search_{_suffix(dW.dtype)}")
    st = f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(n), _ptr(dW),
           _ptr(dH), int(ldh), _ptr(dPsi), ctypes.byref(cost.cost), int(direction), int(method),
           _ptr(dWork), dWork.numel() * dWork.element_size(),
           ctypes.byref(opts) if opts is not None else None, ctypes.byref(mu), ctypes.byref(info))
    return int(st), float(mu.value), info.as_dict()


def geodesic_step_cm(n: int, dH: torch.Tensor, ldh: int, t: float, dW: torch.Tensor,
                     dWork: torch.Tensor, stream: Optional[torch.cuda.Stream] = None,
                     ctx=None) -> int:
    """clm_cg_geodesic_step_{s,d,c,z}: W <- exp(t H) W"""
    f = getattr(lib(), f"clm_cg_geodesic_step_{_suffix(dW.dtype)}")
    return int(f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(n),
                 _ptr(dH), int(ldh), float(t), _ptr(dW), _ptr(dWork),
                 dWork.numel() * dWork.element_size()))


# ---- convenience layer (ordinary matrices) ------------------------------------------------------
def cg_unitary(W: torch.Tensor, cost: BrockettCost, direction: int = MAXIMIZE,
               opts: Optional[CgOptions] = None, stream: Optional[torch.cuda.Stream] = None):
    """cg_unitary<T, CostF> (src/cg_unitary/cg_solver.cppm:178-432) from the unitary
    start W; returns (W_opt, info dict). A status other than success raises."""
    _require_device(W, "W")
    n = W.shape[0]
    with _on_stream(stream) as keep:
        wcm = to_cm(W)
        work = _work_for(n, cost, W.dtype, W.device, None)
        st, info = cg_unitary_cm(n, wcm, cost, direction, work, opts, stream)
        check(st, "clm_cg_unitary")
        return keep(from_cm(wcm, n, n).contiguous()), info


def geodesic_search(W: torch.Tensor, H: torch.Tensor, Psi: torch.Tensor, cost: BrockettCost,
                    direction: int = MAXIMIZE, method: int = POLYNOMIAL,
                    opts: Optional[CgSearchOptions] = None,
                    stream: Optional[torch.cuda.Stream] = None):
    """geodesic_search<T, CostF> (src/cg_unitary/geodesic_search.cppm:679-697): the
    step mu along exp(sgn mu H) W; returns (mu, info dict)"""
    n = W.shape[0]
    with _on_stream(stream):
        wcm, hcm, pcm = to_cm(W), to_cm(H), to_cm(Psi)
        work = _work_for(n, cost, W.dtype, W.device, None)
        st, mu, info = geodesic_search_cm(n, wcm, hcm, n, pcm, cost, direction, method, work,
                                          opts, stream)
        check(st, "clm_cg_geodesic_search")
    return mu, info


def geodesic_step(H: torch.Tensor, t: float, W: torch.Tensor,
                  stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
    """exp(t H) W (src/cg_unitary/cg_solver.cppm:298-333)"""
    n = W.shape[0]
    with _on_stream(stream) as keep:
        wcm, hcm = to_cm(W), to_cm(H)
        work = torch.empty(cg_buffer_size(n, 0, W.dtype), dtype=torch.uint8, device=W.device)
        check(geodesic_step_cm(n, hcm, n, t, wcm, work, stream), "clm_cg_geodesic_step")
        return keep(from_cm(wcm, n, n).contiguous())
