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
// context for the current device). Element types: double and
// std::complex<double> (hipDoubleComplex converts implicitly through the same
// layout).
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

// ExpmPlan (src/expm/plan.cppm:290-294); m is the Taylor degree here.
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
template <typename T>
struct abi;
template <>
struct abi<double> {
  using elem = double;
  static constexpr bool complex = false;
};
template <>
struct abi<std::complex<double>> {
  using elem = clm_double_complex;
  static constexpr bool complex = true;
};
template <>
struct abi<hipDoubleComplex> {
  using elem = clm_double_complex;
  static constexpr bool complex = true;
};
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
static_assert(sizeof(std::complex<double>) == sizeof(clm_double_complex), "complex layout");
}  // namespace detail

// ---- plan / thresholds -------------------------------------------------------------
template <typename T>
inline ExpmPlan expm_plan(double norm1) {
  const clm_expm_plan p = clm_expm_plan_d(norm1);
  return ExpmPlan{p.m, p.s, p.num_gemms};
}

// backward-error threshold of the degree-m Taylor rung (the role of pade_theta<T>)
template <typename T>
inline double taylor_theta(int m) {
  return clm_taylor_theta_d(m);
}

// the reference's own Pade thresholds, for callers that predict its cost
template <typename T>
inline double pade_theta(int m) {
  return clm_pade_theta_d(m);
}

// ---- expm ----------------------------------------------------------------------------------
template <typename T>
inline Status expm_bufferSize(Handle /*solver*/, int n, std::size_t *lwork_bytes) {
  return detail::abi<T>::complex ? clm_expm_buffer_size_z(n, lwork_bytes)
                                 : clm_expm_buffer_size_d(n, lwork_bytes);
}

template <typename T>
inline Status expm(Handle blas, Handle /*solver*/, hipStream_t stream, int n, const T *d_A,
                   int lda, T *d_expA, int lde, void *d_work, std::size_t lwork_bytes, int *d_info,
                   ExpmPlan *plan = nullptr) {
  clm_expm_plan p{};
  clm_status_t st;
  if constexpr (detail::abi<T>::complex)
    st = clm_expm_z(blas, stream, n, detail::c(d_A), lda, detail::c(d_expA), lde, d_work,
                    lwork_bytes, d_info, &p);
  else
    st = clm_expm_d(blas, stream, n, d_A, lda, d_expA, lde, d_work, lwork_bytes, d_info, &p);
  if (plan && st == CLM_STATUS_SUCCESS) *plan = ExpmPlan{p.m, p.s, p.num_gemms};
  return st;
}

// matrix_norm1<T> (src/expm/norm1.cppm:417-428): synchronizes, returns on the host.
template <typename T>
inline double matrix_norm1(hipStream_t stream, int n, const T *d_A, int lda, double *d_colsum) {
  double out = 0.0;
  if constexpr (detail::abi<T>::complex)
    clm_matrix_norm1_z(stream, n, detail::c(d_A), lda, d_colsum, &out);
  else
    clm_matrix_norm1_d(stream, n, d_A, lda, d_colsum, &out);
  return out;
}

// lange<T>: like the reference, allocates its per-column scratch on the stream.
template <typename T>
inline Status lange(hipStream_t stream, MatrixNorm which, std::size_t m, std::size_t n,
                    const T *d_A, std::size_t lda, double *d_result) {
  double *scratch = nullptr;
  const std::size_t len = (m > n ? m : n) + 1;
  if (hipMallocAsync(reinterpret_cast<void **>(&scratch), len * sizeof(double), stream) !=
      hipSuccess)
    return CLM_STATUS_RUNTIME_ERROR;
  clm_status_t st;
  if constexpr (detail::abi<T>::complex)
    st = clm_lange_z(stream, static_cast<int>(which), m, n, detail::c(d_A), lda, d_result, scratch);
  else
    st = clm_lange_d(stream, static_cast<int>(which), m, n, d_A, lda, d_result, scratch);
  static_cast<void>(hipFreeAsync(scratch, stream));
  return st;
}

template <typename T>
inline Status lascl(hipStream_t stream, double cfrom, double cto, std::size_t m, std::size_t n,
                    T *d_a, std::size_t lda) {
  if constexpr (detail::abi<T>::complex)
    return clm_lascl_z(stream, cfrom, cto, m, n, detail::c(d_a), lda);
  else
    return clm_lascl_d(stream, cfrom, cto, m, n, d_a, lda);
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
  return detail::abi<T>::complex ? clm_paterson_stockmeyer_power_stride_z(n)
                                 : clm_paterson_stockmeyer_power_stride_d(n);
}
template <typename T>
inline std::size_t paterson_stockmeyer_bufferSize(int n, int degree, int s_requested = 0) {
  return detail::abi<T>::complex ? clm_paterson_stockmeyer_buffer_size_z(n, degree, s_requested)
                                 : clm_paterson_stockmeyer_buffer_size_d(n, degree, s_requested);
}
template <typename T>
inline Status paterson_stockmeyer(Handle handle, int n, const T *d_coeffs, int degree,
                                  const T *d_A, int lda, T *d_P, int ldp, void *d_work,
                                  std::size_t work_bytes, int s_requested = 0) {
  if constexpr (detail::abi<T>::complex)
    return clm_paterson_stockmeyer_z(handle, n, detail::c(d_coeffs), degree, detail::c(d_A), lda,
                                     detail::c(d_P), ldp, d_work, work_bytes, s_requested);
  else
    return clm_paterson_stockmeyer_d(handle, n, d_coeffs, degree, d_A, lda, d_P, ldp, d_work,
                                     work_bytes, s_requested);
}

template <typename T>
inline std::size_t horner_bufferSize(int n) {
  return detail::abi<T>::complex ? clm_horner_buffer_size_z(n) : clm_horner_buffer_size_d(n);
}
template <typename T>
inline Status horner(Handle handle, int n, const T *d_coeffs, int degree, const T *d_A, int lda,
                     T *d_P, int ldp, void *d_work, std::size_t work_bytes) {
  if constexpr (detail::abi<T>::complex)
    return clm_horner_z(handle, n, detail::c(d_coeffs), degree, detail::c(d_A), lda,
                        detail::c(d_P), ldp, d_work, work_bytes);
  else
    return clm_horner_d(handle, n, d_coeffs, degree, d_A, lda, d_P, ldp, d_work, work_bytes);
}

}  // namespace calaman
