"""Python host mirror of calaman_G's matrix-exponential interface over the C ABI.

Names and argument meaning follow the reference's modules
(src/expm/interface.cppm, src/paterson_stockmeyer/interface.cppm,
src/horner/interface.cppm, src/lapack/lange, src/lapack/lascl); device memory
and streams come from PyTorch (ROCm), the arithmetic from libcalaman_expm.so.

Two layers:

* ``*_cm`` functions take torch tensors purely as device buffers holding a
  COLUMN-MAJOR matrix with a leading dimension -- the reference's calling
  convention, one-to-one with the C entry points. A column-major n x n matrix
  with leading dimension ld lives in a tensor of shape (n, ld) (row j holds
  column j), or any tensor with at least ld*(n-1)+n elements.
* convenience functions (``expm``, ``matrix_norm1``, ...) take ordinary
  (row-major) torch matrices and return ordinary matrices; they transpose into
  column-major storage so the plan is chosen from the 1-norm of the matrix
  itself, as the reference does.

Errors are raised as ``CalamanError`` carrying the library status.
"""
from __future__ import annotations

import contextlib
import ctypes
import threading
import weakref
from typing import Optional

import torch

from . import _lib
from ._lib import CalamanError, ExpmPlan, check, lib

_F64 = torch.float64
_C128 = torch.complex128
_F32 = torch.float32
_C64 = torch.complex64
_SUFFIX = {_F32: "s", _F64: "d", _C64: "c", _C128: "z"}


def _suffix(dtype: torch.dtype) -> str:
    try:
        return _SUFFIX[dtype]
    except KeyError:
        raise TypeError(f"unsupported dtype {dtype}; float32, float64, complex64 and complex128 "
                        "are supported") from None


def _real_dtype(dtype: torch.dtype) -> torch.dtype:
    return _F32 if dtype in (_F32, _C64) else _F64


def _is_single(dtype: torch.dtype) -> bool:
    return dtype in (_F32, _C64)


def _stream_ptr(stream: Optional[torch.cuda.Stream]) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


@contextlib.contextmanager
def _on_stream(stream: Optional[torch.cuda.Stream]):
    """Runs a convenience wrapper's device work on `stream`: the stream first waits
    for the work already queued on the current stream (which made the input), the
    temporaries come from `stream`'s allocator pool, and afterwards the current
    stream waits for `stream`, so the result is ready wherever it is used next.
    Yields a function that marks a returned tensor as used on the current stream."""
    cur = torch.cuda.current_stream()
    if stream is None or stream == cur:
        yield lambda t: t
        return
    stream.wait_stream(cur)
    with torch.cuda.stream(stream):
        yield lambda t: (t.record_stream(cur), t)[1]
    cur.wait_stream(stream)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _require_device(t: torch.Tensor, name: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device (cuda/HIP) tensor")


def to_cm(a: torch.Tensor) -> torch.Tensor:
    """Column-major storage of a (row-major) matrix: returns a tensor whose memory
    is a in column-major order (shape (ncols, nrows), row j = column j)."""
    return a.t().contiguous()


def from_cm(buf: torch.Tensor, n_rows: int, n_cols: int, ld: Optional[int] = None) -> torch.Tensor:
    """The n_rows x n_cols matrix held column-major in buf (leading dimension ld)."""
    ld = n_rows if ld is None else ld
    flat = buf.reshape(-1)[: ld * (n_cols - 1) + n_rows] if n_cols > 0 else buf.reshape(-1)[:0]
    return flat.as_strided((n_rows, n_cols), (1, ld))


# ---- plans and sizes ------------------------------------------------------------------
def expm_plan(norm1: float, dtype: torch.dtype = _F64, power_norms=None) -> dict:
    """The plan for a 1-norm alone (the role of expm_plan<T>); with power_norms
    (||A^k||_1 for k = 2 .. q, as expm() reports them in plan["power_norm1"]) the
    plan expm() executes from the norms of A .. A^q. float32/complex64 use the
    thresholds at u = 2^-24."""
    single = _is_single(dtype)
    if power_norms:
        norms = [float(norm1)] + [float(v) for v in power_norms]
        arr = (ctypes.c_double * len(norms))(*norms)
        f = lib().clm_expm_plan_powers_s if single else lib().clm_expm_plan_powers_d
        return f(arr, len(norms)).as_dict()
    f = lib().clm_expm_plan_s if single else lib().clm_expm_plan_d
    return f(float(norm1)).as_dict()


def taylor_theta(m: int, single: bool = False) -> float:
    f = lib().clm_taylor_theta_s if single else lib().clm_taylor_theta_d
    return float(f(int(m)))


def pade_theta(m: int, single: bool = False) -> float:
    """pade_theta<double> / pade_theta<float> (src/expm/plan.cppm:51-85)."""
    f = lib().clm_pade_theta_s if single else lib().clm_pade_theta_d
    return float(f(int(m)))


def pade_plan(norm1: float, single: bool = False) -> dict:
    """The reference's own expm_plan<double> / <float> (src/expm/plan.cppm:112-130)."""
    if single:
        return lib().clm_pade_plan_s(float(norm1)).as_dict()
    return lib().clm_pade_plan_d(float(norm1)).as_dict()


def expm_buffer_size(n: int, dtype: torch.dtype = _F64) -> int:
    out = ctypes.c_size_t(0)
    check(getattr(lib(), f"clm_expm_buffer_size_{_suffix(dtype)}")(int(n), ctypes.byref(out)),
          "expm_buffer_size")
    return int(out.value)


def ps_plan(degree: int, s_requested: int = 0) -> dict:
    return lib().clm_paterson_stockmeyer_plan(int(degree), int(s_requested)).as_dict()


def ps_cost(degree: int, s: int) -> int:
    return int(lib().clm_paterson_stockmeyer_cost(int(degree), int(s)))


def ps_block_size(degree: int) -> int:
    return int(lib().clm_paterson_stockmeyer_block_size(int(degree)))


def ps_buffer_size(n: int, degree: int, s_requested: int = 0, dtype: torch.dtype = _F64) -> int:
    f = getattr(lib(), f"clm_paterson_stockmeyer_buffer_size_{_suffix(dtype)}")
    return int(f(int(n), int(degree), int(s_requested)))


def ps_power_stride(n: int, dtype: torch.dtype = _F64) -> int:
    f = getattr(lib(), f"clm_paterson_stockmeyer_power_stride_{_suffix(dtype)}")
    return int(f(int(n)))


def horner_buffer_size(n: int, dtype: torch.dtype = _F64) -> int:
    return int(getattr(lib(), f"clm_horner_buffer_size_{_suffix(dtype)}")(int(n)))


# ---- context -------------------------------------------------------------------------------
class Context:
    """clm_ctx: the stand-in for the reference's BLAS/solver handles; carries the
    stream Paterson-Stockmeyer, Horner and gemm enqueue on."""

    def __init__(self, stream: Optional[torch.cuda.Stream] = None):
        h = ctypes.c_void_p()
        check(lib().clm_create(ctypes.byref(h)), "clm_create")
        self.handle = h
        self.set_stream(stream)

    def set_stream(self, stream: Optional[torch.cuda.Stream]) -> None:
        check(lib().clm_set_stream(self.handle, ctypes.c_void_p(_stream_ptr(stream))),
              "clm_set_stream")

    def close(self) -> None:
        if self.handle:
            lib().clm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts: dict = {}


def _ctx_for(stream: Optional[torch.cuda.Stream]) -> Context:
    key = (torch.cuda.current_device(), _stream_ptr(stream))
    c = _contexts.get(key)
    if c is None:
        c = Context(stream)
        _contexts[key] = c
    return c


# ---- column-major layer (one-to-one with the C entry points) ---------------------------------
def expm_cm(n: int, dA: torch.Tensor, lda: int, dE: torch.Tensor, lde: int, dWork: torch.Tensor,
            dInfo: Optional[torch.Tensor] = None, stream: Optional[torch.cuda.Stream] = None,
            ctx: Optional[Context] = None) -> tuple:
    """clm_expm_{s,d,c,z}; returns (status, plan dict)."""
    plan = ExpmPlan()
    f = getattr(lib(), f"clm_expm_{_suffix(dA.dtype)}")
    st = f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(n), _ptr(dA),
           int(lda), _ptr(dE), int(lde), _ptr(dWork), dWork.numel() * dWork.element_size(),
           _ptr(dInfo), ctypes.byref(plan))
    return int(st), plan.as_dict()


def expm_split_cm(n: int, dA: torch.Tensor, lda: int, dE: torch.Tensor, lde: int,
                  dWork: torch.Tensor, dInfo: Optional[torch.Tensor] = None, ps_split: int = 0,
                  split_squarings: int = -1, stream: Optional[torch.cuda.Stream] = None,
                  ctx: Optional[Context] = None, plan_from: int = _lib.PLAN_POWERS) -> tuple:
    """clm_expm_split_d: the BF16-split precision variant; returns (status, plan dict).
    plan_from: PLAN_POWERS (s from the norms of powers) or PLAN_NORM1 (||A||_1 alone)."""
    plan = ExpmPlan()
    opts = _lib.SplitOpts(int(ps_split), int(split_squarings), int(plan_from))
    st = lib().clm_expm_split_d(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)),
                                int(n), _ptr(dA), int(lda), _ptr(dE), int(lde), _ptr(dWork),
                                dWork.numel() * dWork.element_size(), _ptr(dInfo),
                                ctypes.byref(opts), ctypes.byref(plan))
    return int(st), plan.as_dict()


def expm_split_buffer_size(n: int) -> int:
    out = ctypes.c_size_t(0)
    check(lib().clm_expm_split_buffer_size_d(int(n), ctypes.byref(out)), "expm_split_buffer_size")
    return int(out.value)


def expm_split_plan(norm1: float, ps_split: int = 0, split_squarings: int = -1,
                    power_norms=None, n: int = 65) -> dict:
    """The split plan expm_split executes for an order-n matrix from the norms of
    A .. A^q: norm1 and power_norms (||A^k||_1 for k = 2 .. q, as the plan reports
    them in "power_norm1"; none = the 1-norm alone)."""
    opts = _lib.SplitOpts(int(ps_split), int(split_squarings), _lib.PLAN_POWERS)
    norms = [float(norm1)] + [float(v) for v in (power_norms or [])]
    arr = (ctypes.c_double * len(norms))(*norms)
    return lib().clm_expm_split_plan_powers_d(int(n), arr, len(norms),
                                              ctypes.byref(opts)).as_dict()


def expm_fp8tail_cm(n: int, dA: torch.Tensor, lda: int, dE: torch.Tensor, lde: int,
                    dWork: torch.Tensor, dInfo: Optional[torch.Tensor] = None, ps_split: int = 0,
                    split_squarings: int = -1, fp8_min_degree: int = 0,
                    stream: Optional[torch.cuda.Stream] = None,
                    ctx: Optional[Context] = None) -> tuple:
    """clm_expm_fp8tail_d: the split variant with the products that carry only Taylor
    terms of degree >= fp8_min_degree (0 = 4) on FP8; returns (status, plan dict)."""
    plan = ExpmPlan()
    opts = _lib.Fp8TailOpts(int(ps_split), int(split_squarings), int(fp8_min_degree))
    st = lib().clm_expm_fp8tail_d(ctx.handle if ctx else None,
                                  ctypes.c_void_p(_stream_ptr(stream)), int(n), _ptr(dA),
                                  int(lda), _ptr(dE), int(lde), _ptr(dWork),
                                  dWork.numel() * dWork.element_size(), _ptr(dInfo),
                                  ctypes.byref(opts), ctypes.byref(plan))
    return int(st), plan.as_dict()


def expm_fp8tail_buffer_size(n: int) -> int:
    out = ctypes.c_size_t(0)
    check(lib().clm_expm_fp8tail_buffer_size_d(int(n), ctypes.byref(out)),
          "expm_fp8tail_buffer_size")
    return int(out.value)


def matrix_norm1_cm(n: int, dA: torch.Tensor, lda: int, dColsum: torch.Tensor,
                    stream: Optional[torch.cuda.Stream] = None) -> float:
    out = ctypes.c_float(0.0) if _is_single(dA.dtype) else ctypes.c_double(0.0)
    f = getattr(lib(), f"clm_matrix_norm1_{_suffix(dA.dtype)}")
    check(f(ctypes.c_void_p(_stream_ptr(stream)), int(n), _ptr(dA), int(lda), _ptr(dColsum),
            ctypes.byref(out)), "matrix_norm1")
    return float(out.value)


def lange_cm(which: int, m: int, n: int, dA: torch.Tensor, lda: int, dResult: torch.Tensor,
             dScratch: torch.Tensor, stream: Optional[torch.cuda.Stream] = None) -> int:
    f = getattr(lib(), f"clm_lange_{_suffix(dA.dtype)}")
    return int(f(ctypes.c_void_p(_stream_ptr(stream)), int(which), int(m), int(n), _ptr(dA),
                 int(lda), _ptr(dResult), _ptr(dScratch)))


def lascl_cm(cfrom: float, cto: float, m: int, n: int, dA: torch.Tensor, lda: int,
             stream: Optional[torch.cuda.Stream] = None) -> int:
    f = getattr(lib(), f"clm_lascl_{_suffix(dA.dtype)}")
    return int(f(ctypes.c_void_p(_stream_ptr(stream)), float(cfrom), float(cto), int(m), int(n),
                 _ptr(dA), int(lda)))


def paterson_stockmeyer_cm(ctx: Context, n: int, dCoeffs: torch.Tensor, degree: int,
                           dA: torch.Tensor, lda: int, dP: torch.Tensor, ldp: int,
                           dWork: Optional[torch.Tensor], s_requested: int = 0) -> int:
    f = getattr(lib(), f"clm_paterson_stockmeyer_{_suffix(dA.dtype)}")
    wb = 0 if dWork is None else dWork.numel() * dWork.element_size()
    return int(f(ctx.handle, int(n), _ptr(dCoeffs), int(degree), _ptr(dA), int(lda), _ptr(dP),
                 int(ldp), _ptr(dWork), wb, int(s_requested)))


def horner_cm(ctx: Context, n: int, dCoeffs: torch.Tensor, degree: int, dA: torch.Tensor,
              lda: int, dP: torch.Tensor, ldp: int, dWork: Optional[torch.Tensor]) -> int:
    f = getattr(lib(), f"clm_horner_{_suffix(dA.dtype)}")
    wb = 0 if dWork is None else dWork.numel() * dWork.element_size()
    return int(f(ctx.handle, int(n), _ptr(dCoeffs), int(degree), _ptr(dA), int(lda), _ptr(dP),
                 int(ldp), _ptr(dWork), wb))


def gemm_cm(ctx: Context, m: int, n: int, k: int, alpha, dA: torch.Tensor, lda: int,
            dB: torch.Tensor, ldb: int, beta, dC: torch.Tensor, ldc: int, split: int = 0) -> int:
    if split:
        return int(lib().clm_gemm_split_d(ctx.handle, int(split), m, n, k, float(alpha), _ptr(dA),
                                          lda, _ptr(dB), ldb, float(beta), _ptr(dC), ldc))
    sfx = _suffix(dA.dtype)
    if dA.dtype in (_C128, _C64):
        ct = _lib.DoubleComplex if dA.dtype == _C128 else _lib.FloatComplex
        a = ct(complex(alpha).real, complex(alpha).imag)
        b = ct(complex(beta).real, complex(beta).imag)
    else:
        a, b = float(alpha), float(beta)
    return int(getattr(lib(), f"clm_gemm_{sfx}")(ctx.handle, m, n, k, a, _ptr(dA), lda, _ptr(dB),
                                                  ldb, b, _ptr(dC), ldc))


def gemm_fp8_buffer_size(m: int, n: int, k: int) -> int:
    out = ctypes.c_size_t(0)
    check(lib().clm_gemm_fp8_buffer_size_d(int(m), int(n), int(k), ctypes.byref(out)),
          "gemm_fp8_buffer_size")
    return int(out.value)


def gemm_fp8_cm(ctx: Optional[Context], m: int, n: int, k: int, alpha: float, dA: torch.Tensor,
                lda: int, dB: torch.Tensor, ldb: int, beta: float, dC: torch.Tensor, ldc: int,
                dWork: torch.Tensor) -> int:
    """clm_gemm_fp8_d on column-major float64 buffers: C = alpha * Q(A) Q(B) + beta * C
    with Q the e4m3 rounding at one power-of-two scale per 128 x 128 tile."""
    return int(lib().clm_gemm_fp8_d(ctx.handle if ctx else None, int(m), int(n), int(k),
                                    float(alpha), _ptr(dA), int(lda), _ptr(dB), int(ldb),
                                    float(beta), _ptr(dC), int(ldc), _ptr(dWork),
                                    dWork.numel() * dWork.element_size()))


# ---- batched expm -----------------------------------------------------------------------------
_DTYPE_CODE = {_F32: _lib.DTYPE_F32, _F64: _lib.DTYPE_F64, _C64: _lib.DTYPE_C64,
               _C128: _lib.DTYPE_C128}


class BatchPlan:
    """clm_batch_plan over fixed column-major device buffers: ``entries`` is a list
    of (n, A, lda, E, lde) with A and E torch tensors (storage of the column-major
    matrices; element type from A.dtype). The plan keeps references to the
    tensors, so they stay alive while it can launch."""

    def __init__(self, entries, adjoint=None, checkpoints: int = 4):
        """adjoint: optional list of (G, ldg, dA, ldda) per entry -- an adjoint plan
        whose forward records the tape clm_expm_batched_backward reads."""
        self._keep = []
        self.adjoint = adjoint is not None
        ent = _lib.BatchAdEntry if self.adjoint else _lib.BatchEntry
        arr = (ent * len(entries))()
        for i, (n, a, lda, e, lde) in enumerate(entries):
            _require_device(a, "A")
            if e.dtype != a.dtype:
                raise TypeError("A and E differ in dtype")
            fields = [int(n), _DTYPE_CODE[a.dtype], a.data_ptr(), int(lda), e.data_ptr(), int(lde)]
            self._keep += [a, e]
            if self.adjoint:
                g, ldg, da, ldda = adjoint[i]
                fields += [g.data_ptr(), int(ldg), da.data_ptr(), int(ldda)]
                self._keep += [g, da]
            arr[i] = ent(*fields)
        self.count = len(entries)
        h = ctypes.c_void_p()
        if self.adjoint:
            check(lib().clm_expm_batched_ad_plan_create(ctypes.byref(h), len(entries), arr,
                                                        int(checkpoints)),
                  "clm_expm_batched_ad_plan_create")
        else:
            check(lib().clm_expm_batched_plan_create(ctypes.byref(h), len(entries), arr),
                  "clm_expm_batched_plan_create")
        self.handle = h
        nbytes = ctypes.c_size_t(0)
        check(lib().clm_expm_batched_buffer_size(h, ctypes.byref(nbytes)), "buffer_size")
        self.buffer_size = int(nbytes.value)
        self.device = entries[0][1].device if entries else None

    def workspace(self) -> torch.Tensor:
        return torch.empty(self.buffer_size, dtype=torch.uint8, device=self.device)

    def results_buffer(self) -> torch.Tensor:
        """device storage for count clm_batch_result records (24 bytes each)"""
        return torch.zeros(self.count * ctypes.sizeof(_lib.BatchResult), dtype=torch.uint8,
                           device=self.device)

    @staticmethod
    def decode_results(buf: torch.Tensor) -> list:
        raw = buf.cpu().numpy().tobytes()
        size = ctypes.sizeof(_lib.BatchResult)
        out = []
        for i in range(len(raw) // size):
            r = _lib.BatchResult.from_buffer_copy(raw[i * size:(i + 1) * size])
            out.append({f: getattr(r, f) for f, _ in r._fields_})
        return out

    def launch(self, work: torch.Tensor, results: Optional[torch.Tensor] = None,
               stream: Optional[torch.cuda.Stream] = None, ctx: Optional[Context] = None,
               graph: bool = False) -> int:
        f = lib().clm_expm_batched_graph if graph else lib().clm_expm_batched
        return int(f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), self.handle,
                     _ptr(work), work.numel(), _ptr(results)))

    def backward(self, work: torch.Tensor, stream: Optional[torch.cuda.Stream] = None,
                 ctx: Optional[Context] = None) -> int:
        """clm_expm_batched_backward: dA from G, after a forward launch on `work`"""
        return int(lib().clm_expm_batched_backward(ctx.handle if ctx else None,
                                                   ctypes.c_void_p(_stream_ptr(stream)),
                                                   self.handle, _ptr(work), work.numel()))

    def close(self):
        if getattr(self, "handle", None):
            lib().clm_expm_batched_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _BatchedAdjointState:
    """Device buffers and the adjoint plan for one (batch, n, dtype, device): the
    column-major inputs, outputs, cotangents and gradients of a (B, n, n) batch."""

    def __init__(self, b, n, dtype, device, checkpoints):
        mk = lambda: torch.empty((b, n, n), dtype=dtype, device=device)
        self.A, self.E, self.G, self.dA = mk(), mk(), mk(), mk()
        ents = [(n, self.A[i], n, self.E[i], n) for i in range(b)]
        adj = [(self.G[i], n, self.dA[i], n) for i in range(b)]
        self.plan = BatchPlan(ents, adjoint=adj, checkpoints=checkpoints)
        self.work = self.plan.workspace()


# Free adjoint states per (batch, n, dtype, device, checkpoints). Every forward takes
# a state of its own (its tape lives in the state's workspace until its backward has
# run); the state returns to the pool when the forward's autograd node is freed --
# after backward, or when the graph is dropped without one.
_ad_free: dict = {}
_ad_lock = threading.Lock()


def _ad_acquire(key):
    with _ad_lock:
        free = _ad_free.get(key)
        if free:
            return free.pop()
    return _BatchedAdjointState(*key)


def _ad_release(key, st):
    with _ad_lock:
        _ad_free.setdefault(key, []).append(st)


class ExpmBatchedFunction(torch.autograd.Function):
    """exp of a (B, n, n) batch on the batched persistent kernel, with the
    hand-written adjoint (clm_expm_batched_backward) as its backward. Each forward
    holds its own buffers, plan and tape until its autograd node is freed, so
    several forwards of one shape may precede their backwards."""

    @staticmethod
    def forward(ctx, a, checkpoints=4):
        b, n, _ = a.shape
        key = (b, n, a.dtype, a.device, int(checkpoints))
        st = _ad_acquire(key)
        ctx.state = st
        ctx.lease = weakref.finalize(ctx, _ad_release, key, st)
        # column-major storage of each A_i is A_i^T in row-major order
        st.A.copy_(a.detach().transpose(1, 2))
        check(st.plan.launch(st.work), "expm_batched forward")
        return st.E.transpose(1, 2).clone()

    @staticmethod
    def backward(ctx, grad):
        st = ctx.state
        st.G.copy_(grad.transpose(1, 2))
        check(st.plan.backward(st.work), "expm_batched backward")
        return st.dA.transpose(1, 2).clone(), None


def expm_batched_autograd(a: torch.Tensor, checkpoints: int = 4) -> torch.Tensor:
    """exp of each matrix of a (B, n, n) device batch (float32/64, complex64/128),
    differentiable through the hand-written Frechet adjoint."""
    if a.dim() != 3 or a.shape[1] != a.shape[2]:
        raise ValueError("expects a (B, n, n) batch")
    return ExpmBatchedFunction.apply(a, checkpoints)


def batched_tuning(dtype: torch.dtype, adjoint: bool = False) -> list:
    """The tile table of forward plans (adjoint=False) or of adjoint plans."""
    arr = (ctypes.c_int * _lib.BATCH_BUCKETS)()
    f = lib().clm_expm_batched_get_tuning_ad if adjoint else lib().clm_expm_batched_get_tuning
    check(f(_DTYPE_CODE[dtype], arr), "get_tuning")
    return list(arr)


def set_batched_tuning(dtype: torch.dtype, cfg, adjoint: bool = False) -> None:
    arr = (ctypes.c_int * _lib.BATCH_BUCKETS)(*cfg)
    f = lib().clm_expm_batched_set_tuning_ad if adjoint else lib().clm_expm_batched_set_tuning
    check(f(_DTYPE_CODE[dtype], arr), "set_tuning")


def batched_autotune(dtype: torch.dtype, nmax: int = 256,
                     stream: Optional[torch.cuda.Stream] = None, adjoint: bool = False) -> tuple:
    """Times every workgroup tile per order bucket up to nmax (forward launches, or
    forward + backward launches for adjoint=True) and installs the fastest;
    returns (table, ms[bucket][tile])."""
    cfg = (ctypes.c_int * _lib.BATCH_BUCKETS)()
    nt = _lib.BATCH_TILES
    ms = (ctypes.c_float * (nt * _lib.BATCH_BUCKETS))()
    f = lib().clm_expm_batched_autotune_ad if adjoint else lib().clm_expm_batched_autotune
    check(f(None, ctypes.c_void_p(_stream_ptr(stream)), _DTYPE_CODE[dtype], int(nmax), cfg, ms),
          "autotune")
    return list(cfg), [list(ms[nt * b:nt * b + nt]) for b in range(_lib.BATCH_BUCKETS)]


def expm_batched(mats, stream: Optional[torch.cuda.Stream] = None):
    """exp of every square device matrix in ``mats`` (mixed orders and element types)
    in one batched launch; returns the list of results (row-major, like expm()).
    Raises CalamanError naming the members whose 1-norm is not finite."""
    with _on_stream(stream) as hand_over:
        cms = [to_cm(a) for a in mats]
        outs = [torch.empty_like(c) for c in cms]
        plan = BatchPlan([(a.shape[0], c, a.shape[0], e, a.shape[0])
                          for a, c, e in zip(mats, cms, outs)])
        work = plan.workspace()
        res = plan.results_buffer()
        check(plan.launch(work, res, stream=stream), "expm_batched")
        bad = [i for i, r in enumerate(BatchPlan.decode_results(res)) if r["status"] != 0]
        if bad:
            raise CalamanError(_lib.INVALID_VALUE, f"expm_batched: members {bad[:8]} "
                               f"({len(bad)} in all) have a non-finite 1-norm")
        return [hand_over(e).t() for e in outs]


# ---- the reference's Pade route (clm_pade_*, clm_expm_pade_*, clm_getrf_*, clm_getrs_*) ----------
def pade_buffer_size(m: int, n: int, dtype: torch.dtype = _F64) -> int:
    out = ctypes.c_size_t()
    check(getattr(lib(), f"clm_pade_buffer_size_{_suffix(dtype)}")(int(m), int(n), ctypes.byref(out)),
          "clm_pade_buffer_size")
    return int(out.value)


def expm_pade_buffer_size(n: int, dtype: torch.dtype = _F64) -> int:
    out = ctypes.c_size_t()
    check(getattr(lib(), f"clm_expm_pade_buffer_size_{_suffix(dtype)}")(int(n), ctypes.byref(out)),
          "clm_expm_pade_buffer_size")
    return int(out.value)


def pade_cm(m: int, n: int, dA: torch.Tensor, lda: int, dR: torch.Tensor, ldr: int,
            dWork: torch.Tensor, dInfo: Optional[torch.Tensor] = None,
            stream: Optional[torch.cuda.Stream] = None, ctx: Optional[Context] = None) -> int:
    """clm_pade_{s,d,c,z}: R = r_m(A) (pade<T>, src/expm/pade.cppm:77-191)"""
    f = getattr(lib(), f"clm_pade_{_suffix(dA.dtype)}")
    return int(f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(m), int(n),
                 _ptr(dA), int(lda), _ptr(dR), int(ldr), _ptr(dWork),
                 dWork.numel() * dWork.element_size(), _ptr(dInfo)))


def expm_pade_cm(n: int, dA: torch.Tensor, lda: int, dE: torch.Tensor, lde: int,
                 dWork: torch.Tensor, dInfo: Optional[torch.Tensor] = None,
                 stream: Optional[torch.cuda.Stream] = None, ctx: Optional[Context] = None) -> tuple:
    """clm_expm_pade_{s,d,c,z}: the reference's expm<T> (Pade + LU); (status, plan dict)"""
    plan = ExpmPlan()
    f = getattr(lib(), f"clm_expm_pade_{_suffix(dA.dtype)}")
    st = f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(n), _ptr(dA),
           int(lda), _ptr(dE), int(lde), _ptr(dWork), dWork.numel() * dWork.element_size(),
           _ptr(dInfo), ctypes.byref(plan))
    return int(st), plan.as_dict()


def getrf_cm(n: int, dA: torch.Tensor, lda: int, dIpiv: torch.Tensor, dInfo: torch.Tensor,
             stream: Optional[torch.cuda.Stream] = None, ctx: Optional[Context] = None) -> int:
    """clm_getrf_{s,d,c,z}: A = P L U in place, ipiv 1-based"""
    f = getattr(lib(), f"clm_getrf_{_suffix(dA.dtype)}")
    return int(f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(n),
                 _ptr(dA), int(lda), _ptr(dIpiv), _ptr(dInfo)))


def getrs_cm(n: int, nrhs: int, dLU: torch.Tensor, lda: int, dIpiv: torch.Tensor,
             dB: torch.Tensor, ldb: int, dInfo: Optional[torch.Tensor] = None,
             stream: Optional[torch.cuda.Stream] = None, ctx: Optional[Context] = None) -> int:
    """clm_getrs_{s,d,c,z}: B <- A^-1 B from getrf's factors"""
    f = getattr(lib(), f"clm_getrs_{_suffix(dLU.dtype)}")
    return int(f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(n),
                 int(nrhs), _ptr(dLU), int(lda), _ptr(dIpiv), _ptr(dB), int(ldb), _ptr(dInfo)))


def expm_pade(a: torch.Tensor, stream: Optional[torch.cuda.Stream] = None,
              return_plan: bool = False):
    """exp(a) by the reference's own algorithm: the Pade ladder, LU solve and squarings"""
    _require_device(a, "a")
    n = a.shape[-1]
    if a.dim() != 2 or a.shape[0] != n:
        raise ValueError("expm_pade expects a square matrix")
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        ecm = torch.empty_like(acm)
        work = torch.empty(expm_pade_buffer_size(n, a.dtype), dtype=torch.uint8, device=a.device)
        info = torch.empty(2, dtype=torch.int32, device=a.device)
        st, plan = expm_pade_cm(n, acm, n, ecm, n, work, info, stream)
        check(st, "expm_pade")
        e = hand_over(ecm).t()
    return (e, plan) if return_plan else e


# ---- exp(A) for a symmetric / Hermitian A by its eigendecomposition ----------------------------
FILL_MODES = {"L": 0, "U": 1}


def _fill_mode(uplo) -> int:
    if isinstance(uplo, str):
        if uplo.upper() not in FILL_MODES:
            raise ValueError(f"uplo {uplo!r}: 'L' or 'U'")
        return FILL_MODES[uplo.upper()]
    return int(uplo)


def expm_herm_buffer_size(n: int, dtype: torch.dtype = _F64) -> int:
    out = ctypes.c_size_t()
    check(getattr(lib(), f"clm_expm_herm_buffer_size_{_suffix(dtype)}")(int(n), ctypes.byref(out)),
          "clm_expm_herm_buffer_size")
    return int(out.value)


def expm_herm_cm(uplo, n: int, dA: torch.Tensor, lda: int, dE: torch.Tensor, lde: int,
                 dWork: torch.Tensor, dInfo: Optional[torch.Tensor] = None,
                 stream: Optional[torch.cuda.Stream] = None, ctx: Optional[Context] = None) -> int:
    """clm_expm_herm_{s,d,c,z} (src/expm/expm_herm.cppm:145): exp(A) from the uplo triangle"""
    f = getattr(lib(), f"clm_expm_herm_{_suffix(dA.dtype)}")
    return int(f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), _fill_mode(uplo),
                 int(n), _ptr(dA), int(lda), _ptr(dE), int(lde), _ptr(dWork),
                 dWork.numel() * dWork.element_size(), _ptr(dInfo)))


def expm_herm(a: torch.Tensor, uplo="U", stream: Optional[torch.cuda.Stream] = None,
              return_info: bool = False):
    """exp(a) for a symmetric / Hermitian device matrix given by its uplo triangle"""
    _require_device(a, "a")
    n = a.shape[-1]
    if a.dim() != 2 or a.shape[0] != n:
        raise ValueError("expm_herm expects a square matrix")
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        ecm = torch.empty_like(acm)
        work = torch.empty(expm_herm_buffer_size(n, a.dtype), dtype=torch.uint8, device=a.device)
        info = torch.zeros(1, dtype=torch.int32, device=a.device)
        check(expm_herm_cm(uplo, n, acm, n, ecm, n, work, info, stream), "expm_herm")
        e = hand_over(ecm).t()
        info = hand_over(info)
    return (e, int(info.item())) if return_info else e


# ---- balancing (?gebal), the balanced exponential, the 1-norm estimator (?lacn2) ---------------
GEBAL_JOBS = {"N": 0, "P": 1, "S": 2, "B": 3}


def _gebal_job(job) -> int:
    if isinstance(job, str):
        if job.upper() not in GEBAL_JOBS:
            raise ValueError(f"gebal job {job!r}: one of N, P, S, B")
        return GEBAL_JOBS[job.upper()]
    return int(job)


def gebal_buffer_size(n: int) -> int:
    """clm_gebal_buffer_size: the int workspace gebal needs (gebal_bufferSize)"""
    out = ctypes.c_int()
    check(lib().clm_gebal_buffer_size(int(n), ctypes.byref(out)), "clm_gebal_buffer_size")
    return int(out.value)


def gebal_cm(job, n: int, dA: torch.Tensor, lda: int, dScale: torch.Tensor, dWork: torch.Tensor,
             max_sweeps: int = 0, stream: Optional[torch.cuda.Stream] = None) -> tuple:
    """clm_gebal_{s,d,c,z} (src/lapack/gebal/interface.cppm:128): A <- D^-1 P^T A P D in place;
    (status, ilo, ihi), both 1-based"""
    ilo, ihi = ctypes.c_int(), ctypes.c_int()
    f = getattr(lib(), f"clm_gebal_{_suffix(dA.dtype)}")
    st = f(ctypes.c_void_p(_stream_ptr(stream)), _gebal_job(job), int(n), _ptr(dA), int(lda),
           ctypes.byref(ilo), ctypes.byref(ihi), _ptr(dScale), _ptr(dWork), int(max_sweeps))
    return int(st), int(ilo.value), int(ihi.value)


def gebal(a: torch.Tensor, job="B", max_sweeps: int = 0,
          stream: Optional[torch.cuda.Stream] = None) -> tuple:
    """(B, ilo, ihi, scale) with B = D^-1 P^T a P D for a square device matrix"""
    _require_device(a, "a")
    n = a.shape[-1]
    if a.dim() != 2 or a.shape[0] != n:
        raise ValueError("gebal expects a square matrix")
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        scale = torch.empty(max(n, 1), dtype=_real_dtype(a.dtype), device=a.device)
        work = torch.empty(max(gebal_buffer_size(n), 1), dtype=torch.int32, device=a.device)
        st, ilo, ihi = gebal_cm(job, n, acm, n, scale, work, max_sweeps, stream)
        check(st, "gebal")
        b = hand_over(acm).t()
        scale = hand_over(scale)
    return b, ilo, ihi, scale[:n]


def expm_balanced_buffer_size(n: int, dtype: torch.dtype = _F64) -> int:
    out = ctypes.c_size_t()
    check(getattr(lib(), f"clm_expm_balanced_buffer_size_{_suffix(dtype)}")(int(n), ctypes.byref(out)),
          "clm_expm_balanced_buffer_size")
    return int(out.value)


def expm_balanced_cm(n: int, dA: torch.Tensor, lda: int, dE: torch.Tensor, lde: int,
                     dWork: torch.Tensor, dInfo: Optional[torch.Tensor] = None,
                     stream: Optional[torch.cuda.Stream] = None, ctx: Optional[Context] = None) -> tuple:
    """clm_expm_balanced_{s,d,c,z}: D exp(D^-1 A D) D^-1 (src/expm/README.md:100-122);
    (status, plan dict of the balanced matrix)"""
    plan = ExpmPlan()
    f = getattr(lib(), f"clm_expm_balanced_{_suffix(dA.dtype)}")
    st = f(ctx.handle if ctx else None, ctypes.c_void_p(_stream_ptr(stream)), int(n), _ptr(dA),
           int(lda), _ptr(dE), int(lde), _ptr(dWork), dWork.numel() * dWork.element_size(),
           _ptr(dInfo), ctypes.byref(plan))
    return int(st), plan.as_dict()


def expm_balanced(a: torch.Tensor, stream: Optional[torch.cuda.Stream] = None,
                  return_plan: bool = False):
    """exp(a) with gebal's scaling around the exponential: one squaring fewer for every
    factor of two the balancing takes off the 1-norm"""
    _require_device(a, "a")
    n = a.shape[-1]
    if a.dim() != 2 or a.shape[0] != n:
        raise ValueError("expm_balanced expects a square matrix")
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        ecm = torch.empty_like(acm)
        work = torch.empty(expm_balanced_buffer_size(n, a.dtype), dtype=torch.uint8, device=a.device)
        info = torch.empty(2, dtype=torch.int32, device=a.device)
        st, plan = expm_balanced_cm(n, acm, n, ecm, n, work, info, stream)
        check(st, "expm_balanced")
        e = hand_over(ecm).t()
    return (e, plan) if return_plan else e


def lacn2_buffer_size(n: int, dtype: torch.dtype = _F64) -> int:
    return int(getattr(lib(), f"clm_lacn2_buffer_size_{_suffix(dtype)}")(int(n)))


class Lacn2State:
    """The host state xLACN2 carries between calls: est, kase, isave[3]"""

    def __init__(self, dtype: torch.dtype = _F64):
        self.est = (ctypes.c_float if _is_single(dtype) else ctypes.c_double)(0.0)
        self.kase = ctypes.c_int(0)
        self.isave = (ctypes.c_int * 3)(0, 0, 0)


def lacn2_cm(n: int, dV: torch.Tensor, dX: torch.Tensor, dWork: torch.Tensor, state: Lacn2State,
             stream: Optional[torch.cuda.Stream] = None) -> int:
    """clm_lacn2_{s,d}: one reverse-communication step (src/lapack/lacn2/interface.cppm:95)"""
    f = getattr(lib(), f"clm_lacn2_{_suffix(dX.dtype)}")
    return int(f(ctypes.c_void_p(_stream_ptr(stream)), int(n), _ptr(dV), _ptr(dX), _ptr(dWork),
                 dWork.numel() * dWork.element_size(), ctypes.byref(state.est),
                 ctypes.byref(state.kase), state.isave))


def norm1_estimate(matvec, rmatvec, n: int, dtype: torch.dtype = _F64,
                   device: Optional[torch.device] = None) -> tuple:
    """||A||_1 estimated by xLACN2 from x -> A x and x -> A^T x on device vectors (the
    callables return the product, on the current stream); (est, v with ||v||_1 = est ||w||_1,
    the sequence of kase values)"""
    device = device or torch.device("cuda", torch.cuda.current_device())
    v = torch.empty(n, dtype=dtype, device=device)
    x = torch.empty(n, dtype=dtype, device=device)
    work = torch.empty(lacn2_buffer_size(n, dtype), dtype=torch.uint8, device=device)
    state = Lacn2State(dtype)
    kases = []
    while True:
        check(lacn2_cm(n, v, x, work, state), "lacn2")
        if state.kase.value == 0:
            break
        kases.append(state.kase.value)
        x.copy_(matvec(x) if state.kase.value == 1 else rmatvec(x))
    return float(state.est.value), v, kases


# ---- convenience layer (row-major torch matrices) ----------------------------------------------
def expm(a: torch.Tensor, stream: Optional[torch.cuda.Stream] = None,
         return_plan: bool = False):
    """exp(a) for a square float32/float64/complex64/complex128 device matrix."""
    _require_device(a, "a")
    n = a.shape[-1]
    if a.dim() != 2 or a.shape[0] != n:
        raise ValueError("expm expects a square matrix")
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        ecm = torch.empty_like(acm)
        work = torch.empty(expm_buffer_size(n, a.dtype), dtype=torch.uint8, device=a.device)
        info = torch.empty(2, dtype=torch.int32, device=a.device)
        st, plan = expm_cm(n, acm, n, ecm, n, work, info, stream)
        check(st, "expm")
        e = hand_over(ecm).t()
    return (e, plan) if return_plan else e


def expm_split(a: torch.Tensor, ps_split: int = 0, split_squarings: int = -1,
               stream: Optional[torch.cuda.Stream] = None, return_plan: bool = False):
    """exp(a) for a float64 device matrix on BF16-split MFMA products (~1e-6 relative)."""
    _require_device(a, "a")
    if a.dtype != _F64:
        raise TypeError("expm_split takes float64")
    n = a.shape[-1]
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        ecm = torch.empty_like(acm)
        work = torch.empty(expm_split_buffer_size(n), dtype=torch.uint8, device=a.device)
        st, plan = expm_split_cm(n, acm, n, ecm, n, work, None, ps_split, split_squarings,
                                 stream)
        check(st, "expm_split")
        e = hand_over(ecm).t()
    return (e, plan) if return_plan else e


def expm_fp8tail(a: torch.Tensor, ps_split: int = 0, split_squarings: int = -1,
                 fp8_min_degree: int = 0, stream: Optional[torch.cuda.Stream] = None,
                 return_plan: bool = False):
    """exp(a) for a float64 device matrix with the high-degree Taylor products on FP8
    (e4m3, one scale per 128 x 128 tile); the error is reported, not bounded."""
    _require_device(a, "a")
    if a.dtype != _F64:
        raise TypeError("expm_fp8tail takes float64")
    n = a.shape[-1]
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        ecm = torch.empty_like(acm)
        work = torch.empty(expm_fp8tail_buffer_size(n), dtype=torch.uint8, device=a.device)
        st, plan = expm_fp8tail_cm(n, acm, n, ecm, n, work, None, ps_split, split_squarings,
                                   fp8_min_degree, stream)
        check(st, "expm_fp8tail")
        e = hand_over(ecm).t()
    return (e, plan) if return_plan else e


def matrix_norm1(a: torch.Tensor, stream: Optional[torch.cuda.Stream] = None) -> float:
    _require_device(a, "a")
    n = a.shape[0]
    with _on_stream(stream):
        acm = to_cm(a)
        colsum = torch.empty(n, dtype=_real_dtype(a.dtype), device=a.device)
        return matrix_norm1_cm(n, acm, n, colsum, stream)


def lange(which: str, a: torch.Tensor, stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
    """The 'max_abs' | 'one' | 'inf' | 'frobenius' norm of a, as a device scalar."""
    code = {"max_abs": 0, "one": 1, "inf": 2, "frobenius": 3}[which]
    m, n = a.shape
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        out = torch.empty(1, dtype=_real_dtype(a.dtype), device=a.device)
        scratch = torch.empty(max(m, n, 1), dtype=_real_dtype(a.dtype), device=a.device)
        check(lange_cm(code, m, n, acm, max(m, 1), out, scratch, stream), "lange")
        return hand_over(out)[0]


def lascl(cfrom: float, cto: float, a: torch.Tensor,
          stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
    """a * (cto / cfrom) through ?lascl's guarded multipliers (returns a new matrix)."""
    m, n = a.shape
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        check(lascl_cm(cfrom, cto, m, n, acm, max(m, 1), stream), "lascl")
        return hand_over(acm).t()


def paterson_stockmeyer(coeffs: torch.Tensor, a: torch.Tensor, s_requested: int = 0,
                        stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
    n = a.shape[0]
    degree = coeffs.numel() - 1
    ctx = _ctx_for(stream)
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        cf = coeffs.contiguous()
        pcm = torch.empty_like(acm)
        nbytes = ps_buffer_size(n, degree, s_requested, a.dtype)
        work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=a.device)
        check(paterson_stockmeyer_cm(ctx, n, cf, degree, acm, n, pcm, n, work, s_requested),
              "paterson_stockmeyer")
        return hand_over(pcm).t()


def horner(coeffs: torch.Tensor, a: torch.Tensor,
           stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
    n = a.shape[0]
    degree = coeffs.numel() - 1
    ctx = _ctx_for(stream)
    with _on_stream(stream) as hand_over:
        acm = to_cm(a)
        cf = coeffs.contiguous()
        pcm = torch.empty_like(acm)
        work = torch.empty(horner_buffer_size(n, a.dtype), dtype=torch.uint8, device=a.device)
        check(horner_cm(ctx, n, cf, degree, acm, n, pcm, n, work), "horner")
        return hand_over(pcm).t()


def gemm(a: torch.Tensor, b: torch.Tensor, alpha=1.0, beta=0.0, c: Optional[torch.Tensor] = None,
         stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
    m, k = a.shape
    k2, n = b.shape
    if k != k2:
        raise ValueError("inner dimensions differ")
    ctx = _ctx_for(stream)
    with _on_stream(stream) as hand_over:
        acm, bcm = to_cm(a), to_cm(b)
        ccm = to_cm(c) if c is not None else torch.zeros((n, m), dtype=a.dtype, device=a.device)
        check(gemm_cm(ctx, m, n, k, alpha, acm, max(m, 1), bcm, max(k, 1), beta, ccm,
                      max(m, 1)), "gemm")
        return hand_over(ccm).t()


__all__ = [
    "CalamanError", "Context", "expm", "expm_cm", "expm_split", "expm_split_cm", "expm_split_plan", "expm_split_buffer_size",
    "expm_fp8tail", "expm_fp8tail_cm", "expm_fp8tail_buffer_size", "gemm_fp8_cm",
    "gemm_fp8_buffer_size", "expm_plan", "expm_buffer_size", "taylor_theta",
    "pade_theta", "pade_plan", "matrix_norm1", "matrix_norm1_cm", "lange", "lange_cm", "lascl",
    "lascl_cm", "paterson_stockmeyer", "paterson_stockmeyer_cm", "horner", "horner_cm", "gemm",
    "gemm_cm", "ps_plan", "ps_cost", "ps_block_size", "ps_buffer_size", "ps_power_stride",
    "horner_buffer_size", "to_cm", "from_cm", "BatchPlan", "expm_batched", "batched_tuning",
    "set_batched_tuning", "batched_autotune", "ExpmBatchedFunction", "expm_batched_autograd",
    "pade_buffer_size", "pade_cm", "expm_pade_buffer_size", "expm_pade_cm", "expm_pade",
    "getrf_cm", "getrs_cm", "GEBAL_JOBS", "gebal_buffer_size", "gebal_cm", "gebal",
    "expm_balanced_buffer_size", "expm_balanced_cm", "expm_balanced", "lacn2_buffer_size",
    "Lacn2State", "lacn2_cm", "norm1_estimate", "FILL_MODES", "expm_herm_buffer_size",
    "expm_herm_cm", "expm_herm",
]
