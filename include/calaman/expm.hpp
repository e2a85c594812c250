// calaman/expm.hpp -- C++ drop-in for calaman_G's matrix-exponential modules
// over the C ABI of libcalaman_expm.so (include/calaman_expm.h).
//
// A caller of the reference writes
//     import calaman.expm;  calaman::expm<double>(blas, solver, stream, n, dA, lda, ...);
// and after switching writes
//     #include <calaman/expm.hpp>  calaman::expm<double>(ctx, ctx, stream, n, dA, lda, ...);
// with the same argument order and meaning (src/expm/interface.cppm:72-89,
// src/paterson_stockmeyer/interface.cppm:329-332, src/horner/interface.cppm:156-159,
// src/lapack/lange/interface.cppm:82-85, src/lapack/lascl/interface.cppm:79-82).
// The two vendor handles become clm_ctx* (nullptr = the library's default
// context for the current device). Element types: float, double,
// std::complex<float>, std::complex<double>, hipFloatComplex and
// hipDoubleComplex (the complex types share one interleaved layout).
#pragma once

#include <hip/hip_complex.h>
#include <hip/hip_runtime.h>

#include <complex>
#include <cstddef>
#include <type_traits>

#include "../calaman_expm.h"

namespace calaman {

// calaman::Status (src/error_handling/error_handling.cppm:88-180): the domain
// and code of the first failure, or success.
struct Status {
  clm_status_t value = CLM_STATUS_SUCCESS;
  Status() = default;
  Status(clm_status_t v) : value(v) {}  // NOLINT: implicit, as the reference's converting ctors
  bool ok() const { return value == CLM_STATUS_SUCCESS; }
  int domain() const { return value & CLM_DOMAIN_MASK; }
  int code() const { return value & 0xffff; }
  const char *name() const { return clm_status_string(value); }
  explicit operator bool() const { return ok(); }
  bool operator==(clm_status_t v) const { return value == v; }
};

using Handle = clm_ctx *;

// ExpmPlan (src/expm/plan.cppm:94-98); m is the Taylor degree here.
struct ExpmPlan {
  int m = 0;
  int s = 0;
  int num_gemms = 0;
};

enum class MatrixNorm { max_abs = CLM_NORM_MAX_ABS, one = CLM_NORM_ONE, inf = CLM_NORM_INF,
                        frobenius = CLM_NORM_FROBENIUS };

struct PatersonStockmeyerPlan {
  int s, r, num_powers, num_blocks, num_gemms;
};

namespace detail {
// Per element type: the C entry points of its precision and its real type.
#define CALAMAN_ABI(T, E, R, SFX, CPLX)                                                           \
  template <>                                                                                    \
  struct abi<T> {                                                                                \
    using elem = E;                                                                              \
    using real = R;                                                                              \
    static constexpr bool complex = CPLX;                                                        \
    static clm_status_t buffer_size(int n, std::size_t *b) { return clm_expm_buffer_size_##SFX(n, b); } \
    static clm_status_t expm(clm_ctx *h, hipStream_t st, int n, const E *a, int lda, E *e,       \
                             int lde, void *w, std::size_t wb, int *info, clm_expm_plan *p) {    \
      return clm_expm_##SFX(h, st, n, a, lda, e, lde, w, wb, info, p);                           \
    }                                                                                            \
    static clm_status_t norm1(hipStream_t st, int n, const E *a, int lda, R *cs, R *out) {      \
      return clm_matrix_norm1_##SFX(st, n, a, lda, cs, out);                                     \
    }                                                                                            \
    static clm_status_t lange(hipStream_t st, int w, std::size_t m, std::size_t n, const E *a,  \
                              std::size_t lda, R *res, R *scr) {                                 \
      return clm_lange_##SFX(st, w, m, n, a, lda, res, scr);                                     \
    }                                                                                            \
    static clm_status_t lascl(hipStream_t st, R cf, R ct, std::size_t m, std::size_t n, E *a,   \
                              std::size_t lda) {                                                 \
      return clm_lascl_##SFX(st, cf, ct, m, n, a, lda);                                          \
    }                                                                                            \
    static std::size_t ps_stride(int n) { return clm_paterson_stockmeyer_power_stride_##SFX(n); } \
    static std::size_t ps_size(int n, int d, int s) {                                            \
      return clm_paterson_stockmeyer_buffer_size_##SFX(n, d, s);                                 \
    }                                                                                            \
    static clm_status_t ps(clm_ctx *h, int n, const E *c, int d, const E *a, int lda, E *p,     \
                           int ldp, void *w, std::size_t wb, int s) {                            \
      return clm_paterson_stockmeyer_##SFX(h, n, c, d, a, lda, p, ldp, w, wb, s);                \
    }                                                                                            \
    static std::size_t horner_size(int n) { return clm_horner_buffer_size_##SFX(n); }           \
    static clm_status_t horner(clm_ctx *h, int n, const E *c, int d, const E *a, int lda, E *p, \
                               int ldp, void *w, std::size_t wb) {                               \
      return clm_horner_##SFX(h, n, c, d, a, lda, p, ldp, w, wb);                                \
    }                                                                                            \
  };

template <typename T>
struct abi;
CALAMAN_ABI(float, float, float, s, false)
CALAMAN_ABI(double, double, double, d, false)
CALAMAN_ABI(std::complex<float>, clm_float_complex, float, c, true)
CALAMAN_ABI(std::complex<double>, clm_double_complex, double, z, true)
CALAMAN_ABI(hipFloatComplex, clm_float_complex, float, c, true)
CALAMAN_ABI(hipDoubleComplex, clm_double_complex, double, z, true)
#undef CALAMAN_ABI

template <typename T>
using elem_t = typename abi<T>::elem;
template <typename T>
inline const elem_t<T> *c(const T *p) {
  return reinterpret_cast<const elem_t<T> *>(p);
}
template <typename T>
inline elem_t<T> *c(T *p) {
  return reinterpret_cast<elem_t<T> *>(p);
}
template <typename T>
constexpr bool single = std::is_same_v<typename abi<T>::real, float>;
static_assert(sizeof(std::complex<double>) == sizeof(clm_double_complex), "complex layout");
static_assert(sizeof(std::complex<float>) == sizeof(clm_float_complex), "complex layout");
}  // namespace detail

// ComplexToRealType<T> (the reference's real type of an element type)
template <typename T>
using RealType = typename detail::abi<T>::real;

// ---- plan / thresholds -------------------------------------------------------------
template <typename T>
inline ExpmPlan expm_plan(double norm1) {
  const clm_expm_plan p = detail::single<T> ? clm_expm_plan_s(norm1) : clm_expm_plan_d(norm1);
  return ExpmPlan{p.m, p.s, p.num_gemms};
}

// backward-error threshold of the degree-m Taylor rung (the role of pade_theta<T>)
template <typename T>
inline double taylor_theta(int m) {
  return detail::single<T> ? clm_taylor_theta_s(m) : clm_taylor_theta_d(m);
}

// the reference's own Pade thresholds, for callers that predict its cost
template <typename T>
inline double pade_theta(int m) {
  return detail::single<T> ? clm_pade_theta_s(m) : clm_pade_theta_d(m);
}

// ---- expm ----------------------------------------------------------------------------------
template <typename T>
inline Status expm_bufferSize(Handle /*solver*/, int n, std::size_t *lwork_bytes) {
  return detail::abi<T>::buffer_size(n, lwork_bytes);
}

template <typename T>
inline Status expm(Handle blas, Handle /*solver*/, hipStream_t stream, int n, const T *d_A,
                   int lda, T *d_expA, int lde, void *d_work, std::size_t lwork_bytes, int *d_info,
                   ExpmPlan *plan = nullptr) {
  clm_expm_plan p{};
  const clm_status_t st = detail::abi<T>::expm(blas, stream, n, detail::c(d_A), lda,
                                               detail::c(d_expA), lde, d_work, lwork_bytes,
                                               d_info, &p);
  if (plan && st == CLM_STATUS_SUCCESS) *plan = ExpmPlan{p.m, p.s, p.num_gemms};
  return st;
}

// matrix_norm1<T> (src/expm/norm1.cppm:45-56): synchronizes, returns on the host.
template <typename T>
inline RealType<T> matrix_norm1(hipStream_t stream, int n, const T *d_A, int lda,
                                RealType<T> *d_colsum) {
  RealType<T> out = 0;
  static_cast<void>(detail::abi<T>::norm1(stream, n, detail::c(d_A), lda, d_colsum, &out));
  return out;
}

// lange<T>: like the reference, allocates its per-column scratch on the stream.
template <typename T>
inline Status lange(hipStream_t stream, MatrixNorm which, std::size_t m, std::size_t n,
                    const T *d_A, std::size_t lda, RealType<T> *d_result) {
  RealType<T> *scratch = nullptr;
  const std::size_t len = (m > n ? m : n) + 1;
  if (hipMallocAsync(reinterpret_cast<void **>(&scratch), len * sizeof(RealType<T>), stream) !=
      hipSuccess)
    return CLM_STATUS_RUNTIME_ERROR;
  const clm_status_t st = detail::abi<T>::lange(stream, static_cast<int>(which), m, n,
                                                detail::c(d_A), lda, d_result, scratch);
  static_cast<void>(hipFreeAsync(scratch, stream));
  return st;
}

template <typename T>
inline Status lascl(hipStream_t stream, RealType<T> cfrom, RealType<T> cto, std::size_t m,
                    std::size_t n, T *d_a, std::size_t lda) {
  return detail::abi<T>::lascl(stream, cfrom, cto, m, n, detail::c(d_a), lda);
}

// ---- Paterson-Stockmeyer / Horner ------------------------------------------------------------------
inline int paterson_stockmeyer_cost(int degree, int s) { return clm_paterson_stockmeyer_cost(degree, s); }
inline int paterson_stockmeyer_block_size(int degree) {
  return clm_paterson_stockmeyer_block_size(degree);
}
inline PatersonStockmeyerPlan paterson_stockmeyer_plan(int degree, int s_requested = 0) {
  const clm_ps_plan p = clm_paterson_stockmeyer_plan(degree, s_requested);
  return PatersonStockmeyerPlan{p.s, p.r, p.num_powers, p.num_blocks, p.num_gemms};
}
template <typename T>
inline std::size_t paterson_stockmeyer_power_stride(int n) {
  return detail::abi<T>::ps_stride(n);
}
template <typename T>
inline std::size_t paterson_stockmeyer_bufferSize(int n, int degree, int s_requested = 0) {
  return detail::abi<T>::ps_size(n, degree, s_requested);
}
template <typename T>
inline Status paterson_stockmeyer(Handle handle, int n, const T *d_coeffs, int degree,
                                  const T *d_A, int lda, T *d_P, int ldp, void *d_work,
                                  std::size_t work_bytes, int s_requested = 0) {
  return detail::abi<T>::ps(handle, n, detail::c(d_coeffs), degree, detail::c(d_A), lda,
                            detail::c(d_P), ldp, d_work, work_bytes, s_requested);
}

template <typename T>
inline std::size_t horner_bufferSize(int n) {
  return detail::abi<T>::horner_size(n);
}
template <typename T>
inline Status horner(Handle handle, int n, const T *d_coeffs, int degree, const T *d_A, int lda,
                     T *d_P, int ldp, void *d_work, std::size_t work_bytes) {
  return detail::abi<T>::horner(handle, n, detail::c(d_coeffs), degree, detail::c(d_A), lda,
                                detail::c(d_P), ldp, d_work, work_bytes);
}

}  // namespace calaman
