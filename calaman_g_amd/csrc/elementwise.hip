// elementwise.hip -- the HBM-bound kernels of the matrix-exponential path.
//
//   * the 1-norm: per-column |.| sums and a NaN-propagating max
//     (src/expm/expm.cu:188-245 over src/reduce_columns/reduce_columns.cuh:62-108);
//   * ?lange's four norms (src/lapack/lange/lange.cu:42-156);
//   * the 2^-s scaling (src/expm/expm.cppm:138-144) and ?lascl's multiply
//     (src/lapack/lascl/lascl.cu:45-84);
//   * the Paterson-Stockmeyer block build
//     (src/paterson_stockmeyer/paterson_stockmeyer.cu:51-110) and Horner's
//     P = alpha I (src/horner/horner.cu:39-67).
//
// Every kernel walks a column-major matrix with the row index on the lanes and
// one column per grid row (blockIdx.y), so each wave reads or writes 64
// consecutive elements of one column -- the coalesced direction -- and the
// column index needs no division. Reductions keep the reference's fold order
// (a 256-lane block, each lane folding rows tid, tid+256, ... seeded by its
// first element, then a halving LDS tree), so the 1-norm of a real matrix is
// bit-for-bit the CPU restatement's.
#include "clm_internal.hpp"

#include <mutex>

namespace clm {
namespace {

constexpr int kRB = 256;  // reduction block: 4 wave64
constexpr int kEW = 256;  // element-wise block

__device__ __forceinline__ double modulus(double x) { return fabs(x); }
__device__ __forceinline__ double modulus(zd x) { return hypot(x.re, x.im); }
__device__ __forceinline__ float modulus(float x) { return fabsf(x); }
__device__ __forceinline__ float modulus(zc x) { return hypotf(x.re, x.im); }
__device__ __forceinline__ double square(double x) { return x * x; }
__device__ __forceinline__ double square(zd x) { return x.re * x.re + x.im * x.im; }
__device__ __forceinline__ float square(float x) { return x * x; }
__device__ __forceinline__ float square(zc x) { return x.re * x.re + x.im * x.im; }

template <typename R>
__device__ __forceinline__ R nan_max(R a, R b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? b : a;
}

// pre: 0 |x|, 1 |x|^2; op: 0 plus, 1 nan_max. One block per column.
// A grid layer y > 0 reduces the matrix at A + y * mstride into out + y * gridDim.x.
template <typename T, int PRE, int OP>
__global__ void __launch_bounds__(kRB)
    reduce_columns_kernel(const T *__restrict__ A, real_t<T> *__restrict__ out, size_t m,
                          size_t lda, long mstride = 0) {
  using R = real_t<T>;
  __shared__ R sdata[kRB];
  const size_t col = blockIdx.x;
  const unsigned tid = threadIdx.x;
  A += static_cast<long>(blockIdx.y) * mstride;
  out += static_cast<size_t>(blockIdx.y) * gridDim.x;
  const T *column = A + col * lda;
  const unsigned nactive = (m < kRB) ? static_cast<unsigned>(m) : kRB;
  if (tid < nactive) {
    R acc = PRE == 0 ? modulus(column[tid]) : square(column[tid]);
    for (size_t i = tid + kRB; i < m; i += kRB) {
      const R x = PRE == 0 ? modulus(column[i]) : square(column[i]);
      acc = OP == 0 ? acc + x : nan_max(acc, x);
    }
    sdata[tid] = acc;
  }
  __syncthreads();
  for (unsigned s = kRB / 2; s > 0; s >>= 1) {
    if (tid < s && tid + s < nactive)
      sdata[tid] = OP == 0 ? sdata[tid] + sdata[tid + s] : nan_max(sdata[tid], sdata[tid + s]);
    __syncthreads();
  }
  if (tid == 0) out[col] = sdata[0];
}

// max_reduce (src/expm/expm.cu:212-245): best seeded at 0, a NaN wins and stays;
// the result goes to out[0] (values[0] when out is values). Block b reduces
// values + b * n into out[b].
template <typename R>
__global__ void __launch_bounds__(kRB) max_reduce_kernel(int n, R *values, R *out) {
  __shared__ R sdata[kRB];
  const int tid = threadIdx.x;
  values += static_cast<long>(blockIdx.x) * n;
  out += blockIdx.x;
  R best = 0;
  for (int i = tid; i < n; i += kRB) {
    const R v = values[i];
    if (v > best || isnan(v)) best = v;
  }
  sdata[tid] = best;
  __syncthreads();
  for (int s = kRB / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const R other = sdata[tid + s];
      if (other > sdata[tid] || isnan(other)) sdata[tid] = other;
    }
    __syncthreads();
  }
  if (tid == 0) out[0] = sdata[0];
}

// single-column identity-free fold of n values into out[0] (lange stage 2)
template <typename R, int OP>
__global__ void __launch_bounds__(kRB) fold_vector_kernel(const R *v, size_t n, R *out,
                                                          int take_sqrt) {
  __shared__ R sdata[kRB];
  const unsigned tid = threadIdx.x;
  const unsigned nactive = (n < kRB) ? static_cast<unsigned>(n) : kRB;
  if (tid < nactive) {
    R acc = v[tid];
    for (size_t i = tid + kRB; i < n; i += kRB) acc = OP == 0 ? acc + v[i] : nan_max(acc, v[i]);
    sdata[tid] = acc;
  }
  __syncthreads();
  for (unsigned s = kRB / 2; s > 0; s >>= 1) {
    if (tid < s && tid + s < nactive)
      sdata[tid] = OP == 0 ? sdata[tid] + sdata[tid + s] : nan_max(sdata[tid], sdata[tid + s]);
    __syncthreads();
  }
  if (tid == 0) out[0] = take_sqrt ? sqrt(sdata[0]) : sdata[0];
}

// lange 'I': one lane per row, j ascending, seed 0 (src/lapack/lange/lange.cu:108-119)
template <typename T>
__global__ void __launch_bounds__(kEW)
    row_abssum_kernel(const T *__restrict__ A, real_t<T> *__restrict__ rows, size_t m, size_t n,
                      size_t lda) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride) {
    real_t<T> acc = 0;
    for (size_t j = 0; j < n; ++j) acc += modulus(A[i + j * lda]);
    rows[i] = acc;
  }
}

template <typename R>
__global__ void zero_scalar_kernel(R *d) { d[0] = 0; }

template <typename T>
__global__ void __launch_bounds__(kEW)
    scale_copy_kernel(int n, double f, const T *__restrict__ in, long ldi, T *__restrict__ out,
                      long ldo) {
  const int i = blockIdx.x * kEW + threadIdx.x;
  const long j = blockIdx.y;
  if (i < n) out[i + j * ldo] = ops<T>::scale(in[i + j * ldi], f);
}

// bank[k-2] = 2^(-k s) bank[k-2] in place for the powers k = 2 .. count + 1 (one
// power per grid layer): the bank formed from A becomes the bank of A / 2^s.
// ldexp is exact where the result is normal and rounds once where it is not.
template <typename T>
__global__ void __launch_bounds__(kEW)
    scale_powers_kernel(int n, T *bank, long stride, long ldb, int s) {
  const int i = blockIdx.x * kEW + threadIdx.x;
  const long j = blockIdx.y;
  const int k = static_cast<int>(blockIdx.z) + 2;
  if (i >= n) return;
  T *x = bank + static_cast<long>(blockIdx.z) * stride + i + j * ldb;
  *x = ops<T>::ldexp(*x, -k * s);
}

// out[k] = 2^(-(k mod b) s) c[k], k < count: the Paterson-Stockmeyer coefficients
// with the 2^-s scaling of A folded in (block j multiplies c[jb + t] by A^t).
template <typename T>
__global__ void scale_coef_kernel(const T *__restrict__ c, T *__restrict__ out, int count, int b,
                                  int s) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) out[k] = ops<T>::ldexp(c[k], -(k % b) * s);
}

template <typename T>
__global__ void __launch_bounds__(kEW) lascl_kernel(size_t m, double mul, T *A, size_t lda) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kEW + threadIdx.x;
  const size_t j = blockIdx.y;
  if (i < m) A[i + j * lda] = ops<T>::scale(A[i + j * lda], mul);
}

// B = c0 I + c1 X1 + sum_{t>=2} c_t bank[t-2], the reference's per-element order
template <typename T>
__global__ void __launch_bounds__(kEW)
    build_block_kernel(int n, T *__restrict__ Bout, long ldb, const T *__restrict__ coef,
                       int nterms, const T *__restrict__ X1, long ldx1, const T *__restrict__ bank,
                       long stride, long ldbank) {
  const int i = blockIdx.x * kEW + threadIdx.x;
  const long j = blockIdx.y;
  if (i >= n) return;
  T acc = (i == j) ? coef[0] : ops<T>::zero();
  if (nterms > 1) acc = ops<T>::mad(coef[1], X1[i + j * ldx1], acc);
  for (int t = 2; t < nterms; ++t)
    acc = ops<T>::mad(coef[t], bank[(long)(t - 2) * stride + i + j * ldbank], acc);
  Bout[i + j * ldb] = acc;
}

template <typename T>
__global__ void __launch_bounds__(kEW)
    set_scaled_identity_kernel(int n, T *P, long ldp, const T *alpha) {
  const int i = blockIdx.x * kEW + threadIdx.x;
  const long j = blockIdx.y;
  if (i < n) P[i + j * ldp] = (i == j) ? alpha[0] : ops<T>::zero();
}

template <typename T>
__global__ void __launch_bounds__(kEW)
    copy_kernel(int n, const T *__restrict__ src, long lds, T *__restrict__ dst, long ldd) {
  const int i = blockIdx.x * kEW + threadIdx.x;
  const long j = blockIdx.y;
  if (i < n) dst[i + j * ldd] = src[i + j * lds];
}

// Copies count values into a host-mapped readback slot and then publishes seq in
// the slot's flag, after a system-scope fence: a host that sees seq sees the values.
template <typename R>
__global__ void publish_kernel(const R *__restrict__ src, int count, double *vals, unsigned *flag,
                               unsigned seq) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < count; ++i) vals[i] = static_cast<double>(src[i]);
  __threadfence_system();
  __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void fill_int_kernel(int *d, int count, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) d[i] = v;
}

inline dim3 col_grid(size_t rows, size_t cols) {
  return dim3(static_cast<unsigned>((rows + kEW - 1) / kEW), static_cast<unsigned>(cols));
}

}  // namespace

template <typename T>
hipError_t launch_norm1(hipStream_t stream, int n, const T *A, long lda, real_t<T> *colsum,
                        real_t<T> *out) {
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL((reduce_columns_kernel<T, 0, 0>), dim3(n), dim3(kRB), 0, stream, A, colsum,
                     static_cast<size_t>(n), static_cast<size_t>(lda));
  hipLaunchKernelGGL(max_reduce_kernel<real_t<T>>, dim3(1), dim3(kRB), 0, stream, n, colsum,
                     out ? out : colsum);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_norm1_powers(hipStream_t stream, int n, const T *bank, long stride, int count,
                               real_t<T> *colsums, real_t<T> *out) {
  if (n < 1 || count < 1) return hipSuccess;
  hipLaunchKernelGGL((reduce_columns_kernel<T, 0, 0>), dim3(n, count), dim3(kRB), 0, stream, bank,
                     colsums, static_cast<size_t>(n), static_cast<size_t>(n), stride);
  hipLaunchKernelGGL(max_reduce_kernel<real_t<T>>, dim3(count), dim3(kRB), 0, stream, n, colsums,
                     out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_powers(hipStream_t stream, int n, T *bank, long stride, int count, int s) {
  if (n < 1 || count < 1 || s == 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((n + kEW - 1) / kEW), static_cast<unsigned>(n),
                  static_cast<unsigned>(count));
  hipLaunchKernelGGL(scale_powers_kernel<T>, grid, dim3(kEW), 0, stream, n, bank, stride,
                     static_cast<long>(n), s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_coef(hipStream_t stream, const T *c, T *out, int count, int b, int s) {
  if (count < 1 || b < 1) return hipSuccess;
  hipLaunchKernelGGL(scale_coef_kernel<T>, dim3((count + 63) / 64), dim3(64), 0, stream, c, out,
                     count, b, s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lange(hipStream_t stream, int which, size_t m, size_t n, const T *A, size_t lda,
                        real_t<T> *d_result, real_t<T> *d_scratch) {
  using R = real_t<T>;
  if (m == 0 || n == 0) {
    hipLaunchKernelGGL(zero_scalar_kernel<R>, dim3(1), dim3(1), 0, stream, d_result);
    return hipGetLastError();
  }
  switch (which) {
  case CLM_NORM_MAX_ABS:
    hipLaunchKernelGGL((reduce_columns_kernel<T, 0, 1>), dim3(n), dim3(kRB), 0, stream, A,
                       d_scratch, m, lda);
    hipLaunchKernelGGL((fold_vector_kernel<R, 1>), dim3(1), dim3(kRB), 0, stream, d_scratch, n, d_result,
                       0);
    break;
  case CLM_NORM_ONE:
    hipLaunchKernelGGL((reduce_columns_kernel<T, 0, 0>), dim3(n), dim3(kRB), 0, stream, A,
                       d_scratch, m, lda);
    hipLaunchKernelGGL((fold_vector_kernel<R, 1>), dim3(1), dim3(kRB), 0, stream, d_scratch, n, d_result,
                       0);
    break;
  case CLM_NORM_INF: {
    const unsigned blocks = static_cast<unsigned>((m + kEW - 1) / kEW);
    hipLaunchKernelGGL(row_abssum_kernel<T>, dim3(blocks), dim3(kEW), 0, stream, A, d_scratch, m,
                       n, lda);
    hipLaunchKernelGGL((fold_vector_kernel<R, 1>), dim3(1), dim3(kRB), 0, stream, d_scratch, m, d_result,
                       0);
    break;
  }
  default:  // frobenius
    hipLaunchKernelGGL((reduce_columns_kernel<T, 1, 0>), dim3(n), dim3(kRB), 0, stream, A,
                       d_scratch, m, lda);
    hipLaunchKernelGGL((fold_vector_kernel<R, 0>), dim3(1), dim3(kRB), 0, stream, d_scratch, n, d_result,
                       1);
    break;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_copy(hipStream_t stream, int n, double f, const T *in, long ldi, T *out,
                             long ldo) {
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL(scale_copy_kernel<T>, col_grid(n, n), dim3(kEW), 0, stream, n, f, in, ldi, out,
                     ldo);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lascl(hipStream_t stream, size_t m, size_t n, double mul, T *A, size_t lda) {
  if (m == 0 || n == 0) return hipSuccess;
  hipLaunchKernelGGL(lascl_kernel<T>, col_grid(m, n), dim3(kEW), 0, stream, m, mul, A, lda);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_build_block(hipStream_t stream, int n, T *Bout, long ldb, const T *coef,
                              int nterms, const T *X1, long ldx1, const T *bank, long stride,
                              long ldbank) {
  if (n < 1 || nterms < 1) return hipSuccess;
  hipLaunchKernelGGL(build_block_kernel<T>, col_grid(n, n), dim3(kEW), 0, stream, n, Bout, ldb,
                     coef, nterms, X1, ldx1, bank, stride, ldbank);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_set_scaled_identity(hipStream_t stream, int n, T *P, long ldp, const T *alpha) {
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL(set_scaled_identity_kernel<T>, col_grid(n, n), dim3(kEW), 0, stream, n, P, ldp,
                     alpha);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy(hipStream_t stream, int n, const T *src, long lds, T *dst, long ldd) {
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL(copy_kernel<T>, col_grid(n, n), dim3(kEW), 0, stream, n, src, lds, dst, ldd);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_publish(hipStream_t stream, const R *src, int count, double *vals,
                          unsigned *flag, unsigned seq) {
  hipLaunchKernelGGL(publish_kernel<R>, dim3(1), dim3(64), 0, stream, src, count, vals, flag, seq);
  return hipGetLastError();
}
template hipError_t launch_publish<double>(hipStream_t, const double *, int, double *, unsigned *,
                                           unsigned);
template hipError_t launch_publish<float>(hipStream_t, const float *, int, double *, unsigned *,
                                          unsigned);

hipError_t launch_fill_int(hipStream_t stream, int *d, int count, int v) {
  if (count < 1) return hipSuccess;
  hipLaunchKernelGGL(fill_int_kernel, dim3((count + 63) / 64), dim3(64), 0, stream, d, count, v);
  return hipGetLastError();
}

#define CLM_INSTANTIATE(T)                                                                         \
  template hipError_t launch_norm1<T>(hipStream_t, int, const T *, long, real_t<T> *, real_t<T> *); \
  template hipError_t launch_scale_coef<T>(hipStream_t, const T *, T *, int, int, int);            \
  template hipError_t launch_scale_powers<T>(hipStream_t, int, T *, long, int, int);               \
  template hipError_t launch_norm1_powers<T>(hipStream_t, int, const T *, long, int, real_t<T> *,  \
                                             real_t<T> *);                                         \
  template hipError_t launch_lange<T>(hipStream_t, int, size_t, size_t, const T *, size_t,          \
                                      real_t<T> *, real_t<T> *);                                   \
  template hipError_t launch_scale_copy<T>(hipStream_t, int, double, const T *, long, T *, long);  \
  template hipError_t launch_lascl<T>(hipStream_t, size_t, size_t, double, T *, size_t);           \
  template hipError_t launch_build_block<T>(hipStream_t, int, T *, long, const T *, int, const T *, \
                                            long, const T *, long, long);                          \
  template hipError_t launch_set_scaled_identity<T>(hipStream_t, int, T *, long, const T *);       \
  template hipError_t launch_copy<T>(hipStream_t, int, const T *, long, T *, long);

CLM_INSTANTIATE(double)
CLM_INSTANTIATE(zd)
CLM_INSTANTIATE(float)
CLM_INSTANTIATE(zc)
#undef CLM_INSTANTIATE

// ---- 1/k! tables ------------------------------------------------------------------------
// float(Fraction(1, k!)) for k = 0..32: correctly rounded, so a Taylor
// coefficient carries one rounding at most.
#define CLM_INV_FACT(X)                                                                   \
  X(0x1.0000000000000p+0), X(0x1.0000000000000p+0), X(0x1.0000000000000p-1),                    \
      X(0x1.5555555555555p-3), X(0x1.5555555555555p-5), X(0x1.1111111111111p-7),                \
      X(0x1.6c16c16c16c17p-10), X(0x1.a01a01a01a01ap-13), X(0x1.a01a01a01a01ap-16),             \
      X(0x1.71de3a556c734p-19), X(0x1.27e4fb7789f5cp-22), X(0x1.ae64567f544e4p-26),             \
      X(0x1.1eed8eff8d898p-29), X(0x1.6124613a86d09p-33), X(0x1.93974a8c07c9dp-37),             \
      X(0x1.ae7f3e733b81fp-41), X(0x1.ae7f3e733b81fp-45), X(0x1.952c77030ad4ap-49),             \
      X(0x1.6827863b97d97p-53), X(0x1.2f49b46814157p-57), X(0x1.e542ba4020225p-62),             \
      X(0x1.71b8ef6dcf572p-66), X(0x1.0ce396db7f853p-70), X(0x1.761b41316381ap-75),             \
      X(0x1.f2cf01972f578p-80), X(0x1.3f3ccdd165fa9p-84), X(0x1.88e85fc6a4e5ap-89),             \
      X(0x1.d1ab1c2dccea3p-94), X(0x1.0a18a2635085dp-98), X(0x1.259f98b4358adp-103),            \
      X(0x1.3932c5047d60ep-108), X(0x1.434d2e783f5bcp-113), X(0x1.434d2e783f5bcp-118)
#define CLM_AS_REAL(v) v
#define CLM_AS_COMPLEX(v) zd{v, 0.0}
// the same values correctly rounded to FP32 (float(Fraction(1, k!)))
#define CLM_INV_FACT_F(X)                                                                  \
  X(0x1.000000p+0f), X(0x1.000000p+0f), X(0x1.000000p-1f), X(0x1.555556p-3f),                   \
      X(0x1.555556p-5f), X(0x1.111112p-7f), X(0x1.6c16c2p-10f), X(0x1.a01a02p-13f),             \
      X(0x1.a01a02p-16f), X(0x1.71de3ap-19f), X(0x1.27e4fcp-22f), X(0x1.ae6456p-26f),           \
      X(0x1.1eed8ep-29f), X(0x1.612462p-33f), X(0x1.93974ap-37f), X(0x1.ae7f3ep-41f),           \
      X(0x1.ae7f3ep-45f), X(0x1.952c78p-49f), X(0x1.682786p-53f), X(0x1.2f49b4p-57f),           \
      X(0x1.e542bap-62f), X(0x1.71b8f0p-66f), X(0x1.0ce396p-70f), X(0x1.761b42p-75f),           \
      X(0x1.f2cf02p-80f), X(0x1.3f3ccep-84f), X(0x1.88e860p-89f), X(0x1.d1ab1cp-94f),           \
      X(0x1.0a18a2p-98f), X(0x1.259f98p-103f), X(0x1.3932c6p-108f), X(0x1.434d2ep-113f),        \
      X(0x1.434d2ep-118f)
#define CLM_AS_COMPLEX_F(v) zc{v, 0.f}
// extern linkage: hipGetSymbolAddress resolves a __device__ variable by its
// exported name, which a namespace-scope const (internal linkage) does not have.
extern __device__ const double kInvFactD[33];
extern __device__ const zd kInvFactZ[33];
extern __device__ const float kInvFactS[33];
extern __device__ const zc kInvFactC[33];
__device__ const double kInvFactD[33] = {CLM_INV_FACT(CLM_AS_REAL)};
__device__ const zd kInvFactZ[33] = {CLM_INV_FACT(CLM_AS_COMPLEX)};
__device__ const float kInvFactS[33] = {CLM_INV_FACT_F(CLM_AS_REAL)};
__device__ const zc kInvFactC[33] = {CLM_INV_FACT_F(CLM_AS_COMPLEX_F)};
#undef CLM_AS_REAL
#undef CLM_AS_COMPLEX
#undef CLM_AS_COMPLEX_F
#undef CLM_INV_FACT
#undef CLM_INV_FACT_F

namespace {
struct DeviceTables {
  const double *d = nullptr;
  const zd *z = nullptr;
  const float *s = nullptr;
  const zc *c = nullptr;
};
DeviceTables g_tables[64];
std::mutex g_tables_mutex;

// The device address of each table, looked up once per device (each device
// holds its own copy of a __device__ array).
const DeviceTables &tables(int device) {
  std::lock_guard<std::mutex> lock(g_tables_mutex);
  DeviceTables &t = g_tables[device & 63];
  if (!t.d) {
    void *pd = nullptr, *pz = nullptr, *ps = nullptr, *pc = nullptr;
    if (hipGetSymbolAddress(&pd, HIP_SYMBOL(kInvFactD)) != hipSuccess) return t;
    if (hipGetSymbolAddress(&pz, HIP_SYMBOL(kInvFactZ)) != hipSuccess) return t;
    if (hipGetSymbolAddress(&ps, HIP_SYMBOL(kInvFactS)) != hipSuccess) return t;
    if (hipGetSymbolAddress(&pc, HIP_SYMBOL(kInvFactC)) != hipSuccess) return t;
    t.z = static_cast<const zd *>(pz);
    t.s = static_cast<const float *>(ps);
    t.c = static_cast<const zc *>(pc);
    t.d = static_cast<const double *>(pd);  // set last: it marks the entry complete
  }
  return t;
}
}  // namespace

const double *inv_factorials_d(int device) { return tables(device).d; }
const zd *inv_factorials_z(int device) { return tables(device).z; }
const float *inv_factorials_s(int device) { return tables(device).s; }
const zc *inv_factorials_c(int device) { return tables(device).c; }

template <>
const double *inv_factorials<double>(int device) {
  return inv_factorials_d(device);
}
template <>
const zd *inv_factorials<zd>(int device) {
  return inv_factorials_z(device);
}
template <>
const float *inv_factorials<float>(int device) {
  return inv_factorials_s(device);
}
template <>
const zc *inv_factorials<zc>(int device) {
  return inv_factorials_c(device);
}

}  // namespace clm
