"""ctypes binding of libcalaman_expm.so (include/calaman_expm.h).

The library is the product: there is no CPU fallback. Loading fails loudly
when the shared object has not been built.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcalaman_expm.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "calaman_expm.h")
CG_HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "calaman_cg.h")

# status codes (calaman_expm.h)
CLM_DOMAIN_BLAS = 0x00010000
CLM_DOMAIN_SOLVER = 0x00020000
CLM_DOMAIN_RUNTIME = 0x00030000
SUCCESS = 0
NOT_INITIALIZED = CLM_DOMAIN_BLAS | 1
ALLOC_FAILED = CLM_DOMAIN_BLAS | 3
INVALID_VALUE = CLM_DOMAIN_BLAS | 7
EXECUTION_FAILED = CLM_DOMAIN_BLAS | 13
NOT_SUPPORTED = CLM_DOMAIN_BLAS | 15
RUNTIME_INVALID_VALUE = CLM_DOMAIN_RUNTIME | 1
RUNTIME_ERROR = CLM_DOMAIN_RUNTIME | 2

METHOD_TAYLOR_PS = 0
METHOD_PADE = 1
DTYPE_F64 = 1
DTYPE_F32 = 2
DTYPE_C128 = 3
DTYPE_C64 = 4

NORM_MAX_ABS, NORM_ONE, NORM_INF, NORM_FROBENIUS = 0, 1, 2, 3


class ExpmPlan(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("s", ctypes.c_int), ("num_gemms", ctypes.c_int),
                ("method", ctypes.c_int), ("ps_block", ctypes.c_int), ("dtype", ctypes.c_int),
                ("ps_split", ctypes.c_int), ("split_squarings", ctypes.c_int),
                ("norm1", ctypes.c_double), ("fp8_gemms", ctypes.c_int),
                ("fp8_min_degree", ctypes.c_int), ("norm_powers", ctypes.c_int),
                ("power_norm1", ctypes.c_double * 4)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["power_norm1"] = list(self.power_norm1)[:max(self.norm_powers - 1, 0)]
        return d


SPLIT_BF16X3 = 3
SPLIT_BF16X6 = 6
SPLIT_FP16X3 = 13


class SplitOpts(ctypes.Structure):
    _fields_ = [("ps_split", ctypes.c_int), ("split_squarings", ctypes.c_int),
                ("plan_from", ctypes.c_int)]


PLAN_POWERS, PLAN_NORM1 = 0, 1


class Fp8TailOpts(ctypes.Structure):
    _fields_ = [("ps_split", ctypes.c_int), ("split_squarings", ctypes.c_int),
                ("fp8_min_degree", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PsPlan(ctypes.Structure):
    _fields_ = [("s", ctypes.c_int), ("r", ctypes.c_int), ("num_powers", ctypes.c_int),
                ("num_blocks", ctypes.c_int), ("num_gemms", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class DoubleComplex(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class FloatComplex(ctypes.Structure):
    _fields_ = [("re", ctypes.c_float), ("im", ctypes.c_float)]


BATCH_BUCKETS = 8
BATCH_TILE_32, BATCH_TILE_64, BATCH_TILE_128x64, BATCH_TILE_128, BATCH_TILE_128W16 = 0, 1, 2, 3, 4
BATCH_TILES = 5


class BatchEntry(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("dtype", ctypes.c_int), ("A", ctypes.c_void_p),
                ("lda", ctypes.c_int), ("E", ctypes.c_void_p), ("lde", ctypes.c_int)]


class BatchAdEntry(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("dtype", ctypes.c_int), ("A", ctypes.c_void_p),
                ("lda", ctypes.c_int), ("E", ctypes.c_void_p), ("lde", ctypes.c_int),
                ("G", ctypes.c_void_p), ("ldg", ctypes.c_int), ("dA", ctypes.c_void_p),
                ("ldda", ctypes.c_int)]


class BatchResult(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("s", ctypes.c_int), ("num_gemms", ctypes.c_int),
                ("status", ctypes.c_int), ("norm1", ctypes.c_double)]


# ---- cg_unitary (include/calaman_cg.h) ----
CG_MINIMIZE, CG_MAXIMIZE = 0, 1
CG_SEARCH_POLYNOMIAL, CG_SEARCH_DFT = 0, 1
CG_CONVERGED, CG_MAX_ITERATIONS, CG_LINE_SEARCH_FAILED, CG_NUMERICAL_FAILURE = 0, 1, 2, 3
CG_MAX_SAMPLES = 32


class CgSearchOptions(ctypes.Structure):
    _fields_ = [("poly_order", ctypes.c_int), ("dft_factor", ctypes.c_int),
                ("dft_periods", ctypes.c_int), ("use_spectral_radius", ctypes.c_int)]


class CgOptions(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int), ("tol", ctypes.c_double),
                ("method", ctypes.c_int), ("search", CgSearchOptions),
                ("reset_period", ctypes.c_int), ("reorthogonalize_period", ctypes.c_int)]


class CgInfo(ctypes.Structure):
    _fields_ = [("reason", ctypes.c_int), ("iterations", ctypes.c_int), ("resets", ctypes.c_int),
                ("line_search_failures", ctypes.c_int), ("gradient_norm_sq", ctypes.c_double),
                ("last_step", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class CgSearchInfo(ctypes.Structure):
    _fields_ = [("omega_max", ctypes.c_double), ("interval", ctypes.c_double),
                ("num_samples", ctypes.c_int), ("num_roots", ctypes.c_int),
                ("found", ctypes.c_int), ("method_used", ctypes.c_int),
                ("samples", ctypes.c_double * CG_MAX_SAMPLES)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["samples"] = list(self.samples)[:self.num_samples]
        return d


CG_BUFFER_SIZE_FN = ctypes.CFUNCTYPE(ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int)
CG_GRADIENT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_size_t)
CG_VALUE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)


class CgCost(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("buffer_size", CG_BUFFER_SIZE_FN),
                ("gradient", CG_GRADIENT_FN), ("value", CG_VALUE_FN), ("user", ctypes.c_void_p)]


class CgBrockett(ctypes.Structure):
    _fields_ = [("dR", ctypes.c_void_p), ("ldr", ctypes.c_int), ("dN", ctypes.c_void_p)]


class CalamanError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        name = lib().clm_status_string(status).decode()
        super().__init__(f"{what}: {name} ({status:#x})" if what else f"{name} ({status:#x})")


_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t
_d = ctypes.c_double
_f = ctypes.c_float

# name -> (restype, argtypes)
_SIGNATURES = {
    "clm_version": (ctypes.c_char_p, []),
    "clm_status_string": (ctypes.c_char_p, [_i]),
    "clm_create": (_i, [ctypes.POINTER(_vp)]),
    "clm_destroy": (_i, [_vp]),
    "clm_set_stream": (_i, [_vp, _vp]),
    "clm_get_stream": (_i, [_vp, ctypes.POINTER(_vp)]),
    "clm_expm_plan_d": (ExpmPlan, [_d]),
    "clm_expm_plan_powers_d": (ExpmPlan, [ctypes.POINTER(_d), _i]),
    "clm_expm_plan_powers_s": (ExpmPlan, [ctypes.POINTER(_d), _i]),
    "clm_expm_split_plan_powers_d": (ExpmPlan, [_i, ctypes.POINTER(_d), _i,
                                                 ctypes.POINTER(SplitOpts)]),
    "clm_taylor_theta_d": (_d, [_i]),
    "clm_pade_theta_d": (_d, [_i]),
    "clm_pade_plan_d": (ExpmPlan, [_d]),
    "clm_expm_split_plan_d": (ExpmPlan, [_d, ctypes.POINTER(SplitOpts)]),
    "clm_taylor_theta_s": (_d, [_i]),
    "clm_expm_split_buffer_size_d": (_i, [_i, ctypes.POINTER(_sz)]),
    "clm_expm_split_d": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _vp,
                              ctypes.POINTER(SplitOpts), ctypes.POINTER(ExpmPlan)]),
    "clm_expm_fp8tail_buffer_size_d": (_i, [_i, ctypes.POINTER(_sz)]),
    "clm_expm_fp8tail_d": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _vp,
                                ctypes.POINTER(Fp8TailOpts), ctypes.POINTER(ExpmPlan)]),
    "clm_gemm_fp8_buffer_size_d": (_i, [_i, _i, _i, ctypes.POINTER(_sz)]),
    "clm_gemm_fp8_d": (_i, [_vp, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i, _vp, _sz]),
    "clm_gemm_split_d": (_i, [_vp, _i, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i]),
    "clm_expm_buffer_size_d": (_i, [_i, ctypes.POINTER(_sz)]),
    "clm_expm_buffer_size_z": (_i, [_i, ctypes.POINTER(_sz)]),
    "clm_expm_d": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _vp, ctypes.POINTER(ExpmPlan)]),
    "clm_expm_noalloc_d": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, ctypes.POINTER(ExpmPlan)]),
    "clm_expm_z": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _vp, ctypes.POINTER(ExpmPlan)]),
    "clm_matrix_norm1_d": (_i, [_vp, _i, _vp, _i, _vp, ctypes.POINTER(_d)]),
    "clm_matrix_norm1_z": (_i, [_vp, _i, _vp, _i, _vp, ctypes.POINTER(_d)]),
    "clm_lange_d": (_i, [_vp, _i, _sz, _sz, _vp, _sz, _vp, _vp]),
    "clm_lange_z": (_i, [_vp, _i, _sz, _sz, _vp, _sz, _vp, _vp]),
    "clm_lascl_d": (_i, [_vp, _d, _d, _sz, _sz, _vp, _sz]),
    "clm_lascl_z": (_i, [_vp, _d, _d, _sz, _sz, _vp, _sz]),
    "clm_paterson_stockmeyer_cost": (_i, [_i, _i]),
    "clm_paterson_stockmeyer_block_size": (_i, [_i]),
    "clm_paterson_stockmeyer_plan": (PsPlan, [_i, _i]),
    "clm_paterson_stockmeyer_power_stride_d": (_sz, [_i]),
    "clm_paterson_stockmeyer_power_stride_z": (_sz, [_i]),
    "clm_paterson_stockmeyer_buffer_size_d": (_sz, [_i, _i, _i]),
    "clm_paterson_stockmeyer_buffer_size_z": (_sz, [_i, _i, _i]),
    "clm_paterson_stockmeyer_d": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _i]),
    "clm_paterson_stockmeyer_z": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _i]),
    "clm_horner_buffer_size_d": (_sz, [_i]),
    "clm_horner_buffer_size_z": (_sz, [_i]),
    "clm_horner_d": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "clm_horner_z": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "clm_gemm_d": (_i, [_vp, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i]),
    "clm_gemm_z": (_i, [_vp, _i, _i, _i, DoubleComplex, _vp, _i, _vp, _i, DoubleComplex, _vp,
                        _i]),
    # single precision
    "clm_expm_plan_s": (ExpmPlan, [_d]),
    "clm_pade_theta_s": (_d, [_i]),
    "clm_pade_plan_s": (ExpmPlan, [_f]),
    "clm_expm_buffer_size_s": (_i, [_i, ctypes.POINTER(_sz)]),
    "clm_expm_buffer_size_c": (_i, [_i, ctypes.POINTER(_sz)]),
    "clm_expm_s": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _vp, ctypes.POINTER(ExpmPlan)]),
    "clm_expm_c": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _vp, ctypes.POINTER(ExpmPlan)]),
    "clm_matrix_norm1_s": (_i, [_vp, _i, _vp, _i, _vp, ctypes.POINTER(_f)]),
    "clm_matrix_norm1_c": (_i, [_vp, _i, _vp, _i, _vp, ctypes.POINTER(_f)]),
    "clm_lange_s": (_i, [_vp, _i, _sz, _sz, _vp, _sz, _vp, _vp]),
    "clm_lange_c": (_i, [_vp, _i, _sz, _sz, _vp, _sz, _vp, _vp]),
    "clm_lascl_s": (_i, [_vp, _f, _f, _sz, _sz, _vp, _sz]),
    "clm_lascl_c": (_i, [_vp, _f, _f, _sz, _sz, _vp, _sz]),
    "clm_paterson_stockmeyer_power_stride_s": (_sz, [_i]),
    "clm_paterson_stockmeyer_power_stride_c": (_sz, [_i]),
    "clm_paterson_stockmeyer_buffer_size_s": (_sz, [_i, _i, _i]),
    "clm_paterson_stockmeyer_buffer_size_c": (_sz, [_i, _i, _i]),
    "clm_paterson_stockmeyer_s": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _i]),
    "clm_paterson_stockmeyer_c": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _i]),
    "clm_horner_buffer_size_s": (_sz, [_i]),
    "clm_horner_buffer_size_c": (_sz, [_i]),
    "clm_horner_s": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "clm_horner_c": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "clm_gemm_s": (_i, [_vp, _i, _i, _i, _f, _vp, _i, _vp, _i, _f, _vp, _i]),
    "clm_gemm_c": (_i, [_vp, _i, _i, _i, FloatComplex, _vp, _i, _vp, _i, FloatComplex, _vp, _i]),
    # batched
    "clm_expm_batched_plan_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(BatchEntry)]),
    "clm_expm_batched_plan_destroy": (_i, [_vp]),
    "clm_expm_batched_buffer_size": (_i, [_vp, ctypes.POINTER(_sz)]),
    "clm_expm_batched": (_i, [_vp, _vp, _vp, _vp, _sz, _vp]),
    "clm_expm_batched_graph": (_i, [_vp, _vp, _vp, _vp, _sz, _vp]),
    "clm_expm_batched_get_tuning": (_i, [_i, ctypes.POINTER(_i)]),
    "clm_expm_batched_set_tuning": (_i, [_i, ctypes.POINTER(_i)]),
    "clm_expm_batched_autotune": (_i, [_vp, _vp, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_f)]),
    "clm_expm_batched_get_tuning_ad": (_i, [_i, ctypes.POINTER(_i)]),
    "clm_expm_batched_set_tuning_ad": (_i, [_i, ctypes.POINTER(_i)]),
    "clm_expm_batched_autotune_ad": (_i, [_vp, _vp, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_f)]),
    "clm_expm_batched_ad_plan_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(BatchAdEntry),
                                             _i]),
    "clm_expm_batched_backward": (_i, [_vp, _vp, _vp, _vp, _sz]),
    # cg_unitary
    "clm_cg_default_options_s": (None, [ctypes.POINTER(CgOptions)]),
    "clm_cg_default_options_d": (None, [ctypes.POINTER(CgOptions)]),
}
for _t in "sdcz":
    _SIGNATURES[f"clm_pade_buffer_size_{_t}"] = (_i, [_i, _i, ctypes.POINTER(_sz)])
    _SIGNATURES[f"clm_pade_{_t}"] = (_i, [_vp, _vp, _i, _i, _vp, _i, _vp, _i, _vp, _sz, _vp])
    _SIGNATURES[f"clm_expm_pade_buffer_size_{_t}"] = (_i, [_i, ctypes.POINTER(_sz)])
    _SIGNATURES[f"clm_expm_pade_{_t}"] = (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _vp,
                                               ctypes.POINTER(ExpmPlan)])
    _SIGNATURES[f"clm_getrf_{_t}"] = (_i, [_vp, _vp, _i, _vp, _i, _vp, _vp])
    _SIGNATURES[f"clm_getrs_{_t}"] = (_i, [_vp, _vp, _i, _i, _vp, _i, _vp, _vp, _i, _vp])
    _SIGNATURES[f"clm_cg_brockett_cost_{_t}"] = (_i, [ctypes.POINTER(CgBrockett),
                                                      ctypes.POINTER(CgCost)])
    _SIGNATURES[f"clm_cg_unitary_buffer_size_{_t}"] = (_i, [_i, _sz, ctypes.POINTER(_sz)])
    _SIGNATURES[f"clm_cg_unitary_{_t}"] = (_i, [_vp, _vp, _i, _vp, ctypes.POINTER(CgCost), _i, _vp,
                                                _sz, ctypes.POINTER(CgOptions),
                                                ctypes.POINTER(CgInfo)])
    _SIGNATURES[f"clm_cg_geodesic_search_{_t}"] = (_i, [_vp, _vp, _i, _vp, _vp, _i, _vp,
                                                        ctypes.POINTER(CgCost), _i, _i, _vp, _sz,
                                                        ctypes.POINTER(CgSearchOptions),
                                                        ctypes.POINTER(_d),
                                                        ctypes.POINTER(CgSearchInfo)])
    _SIGNATURES[f"clm_cg_geodesic_step_{_t}"] = (_i, [_vp, _vp, _i, _vp, _i, _d, _vp, _vp, _sz])
    # balancing, the balanced exponential
    _SIGNATURES[f"clm_gebal_{_t}"] = (_i, [_vp, _i, _i, _vp, _i, ctypes.POINTER(_i),
                                           ctypes.POINTER(_i), _vp, _vp, _i])
    _SIGNATURES[f"clm_expm_balanced_buffer_size_{_t}"] = (_i, [_i, ctypes.POINTER(_sz)])
    _SIGNATURES[f"clm_expm_balanced_{_t}"] = (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _vp, _sz, _vp,
                                                   ctypes.POINTER(ExpmPlan)])
    # the Hermitian route
    _SIGNATURES[f"clm_expm_herm_buffer_size_{_t}"] = (_i, [_i, ctypes.POINTER(_sz)])
    _SIGNATURES[f"clm_expm_herm_{_t}"] = (_i, [_vp, _vp, _i, _i, _vp, _i, _vp, _i, _vp, _sz, _vp])
_SIGNATURES["clm_gebal_buffer_size"] = (_i, [_i, ctypes.POINTER(_i)])
_SIGNATURES["clm_lacn2_buffer_size_s"] = (_sz, [_i])
_SIGNATURES["clm_lacn2_buffer_size_d"] = (_sz, [_i])
_SIGNATURES["clm_lacn2_s"] = (_i, [_vp, _i, _vp, _vp, _vp, _sz, ctypes.POINTER(_f),
                                   ctypes.POINTER(_i), ctypes.POINTER(_i)])
_SIGNATURES["clm_lacn2_d"] = (_i, [_vp, _i, _vp, _vp, _vp, _sz, ctypes.POINTER(_d),
                                   ctypes.POINTER(_i), ctypes.POINTER(_i)])


def header_symbols(path: str | None = None):
    """Every function the public C headers declare."""
    paths = [path] if path else [HEADER_PATH, CG_HEADER_PATH]
    text = "".join(open(p).read() for p in paths)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # a type name before a function-pointer declarator, "clm_status_t (*f)(...)", is no symbol
    return sorted(set(n for n in re.findall(r"\b(clm_[a-z0-9_]+)\s*\(", text)
                      if not n.endswith("_t")))


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the HIP library is required; there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status != SUCCESS:
        raise CalamanError(status, what)
