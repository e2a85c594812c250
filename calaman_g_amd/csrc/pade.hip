// pade.hip -- the element-wise passes of the reference's Pade route
// (src/expm/pade.cppm:77-191 over src/expm/expm.cu:66-161):
//
//   * the even and odd halves of the Pade numerator in one pass over the powers,
//     V = cv[0] I + sum_k cv[k+1] P_k and W = cw[0] I + sum_k cw[k+1] P_k, the
//     products accumulated k ascending as fused multiply-adds, the identity term
//     last -- the reference's order, so the oracle restatement agrees to rounding;
//   * the split of the numerator and the denominator, R = V + U (output, leading
//     dimension ldr) and Q = V - U over V (q_m(A) = p_m(-A)).
//
// Both walk column-major n x n blocks with the row on the lanes and one column
// a grid row, as the library's other element-wise kernels.
#include "clm_internal.hpp"

namespace clm {
namespace {

constexpr int kBlk = 256;

template <typename R>
struct Coefs {
  R cv[5], cw[5];
};

__device__ inline double fma_real(double c, double x, double acc) { return fma(c, x, acc); }
__device__ inline float fma_real(float c, float x, float acc) { return fmaf(c, x, acc); }
__device__ inline zd fma_real(double c, zd x, zd acc) { return zd{fma(c, x.re, acc.re), fma(c, x.im, acc.im)}; }
__device__ inline zc fma_real(float c, zc x, zc acc) {
  return zc{fmaf(c, x.re, acc.re), fmaf(c, x.im, acc.im)};
}

template <typename T, int NP>
__global__ void __launch_bounds__(kBlk)
    even_odd_kernel(int n, const T *__restrict__ P0, const T *__restrict__ P1, const T *__restrict__ P2,
                    const T *__restrict__ P3, Coefs<real_t<T>> c, T *__restrict__ V, T *__restrict__ W) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const long idx = i + static_cast<long>(j) * n;
  const T *P[4] = {P0, P1, P2, P3};
  T v = ops<T>::zero(), w = ops<T>::zero();
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const T x = P[k][idx];
    v = fma_real(c.cv[k + 1], x, v);
    w = fma_real(c.cw[k + 1], x, w);
  }
  if (i == j) {
    v = ops<T>::add(v, ops<T>::from_real(c.cv[0]));
    w = ops<T>::add(w, ops<T>::from_real(c.cw[0]));
  }
  V[idx] = v;
  W[idx] = w;
}

__device__ inline double sub(double a, double b) { return a - b; }
__device__ inline float sub(float a, float b) { return a - b; }
__device__ inline zd sub(zd a, zd b) { return zd{a.re - b.re, a.im - b.im}; }
__device__ inline zc sub(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }

// R = V + U (leading dimension ldr), Q = V - U (leading dimension n; Q may be V:
// each element is read before it is written, by the same lane)
template <typename T>
__global__ void __launch_bounds__(kBlk)
    split_kernel(int n, const T *U, const T *V, T *R, long ldr, T *Q) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const long idx = i + static_cast<long>(j) * n;
  const T u = U[idx], v = V[idx];
  R[i + j * ldr] = ops<T>::add(v, u);
  Q[idx] = sub(v, u);
}

}  // namespace

template <typename T>
hipError_t launch_pade_even_odd(hipStream_t st, int n, int np, const T *const *P,
                                const double *cv, const double *cw, T *V, T *W) {
  if (n < 1) return hipSuccess;
  Coefs<real_t<T>> c{};
  for (int k = 0; k <= np && k < 5; ++k) {
    c.cv[k] = static_cast<real_t<T>>(cv[k]);
    c.cw[k] = static_cast<real_t<T>>(cw[k]);
  }
  const dim3 grid((n + kBlk - 1) / kBlk, n);
  const T *p0 = P[0], *p1 = np > 1 ? P[1] : nullptr, *p2 = np > 2 ? P[2] : nullptr,
          *p3 = np > 3 ? P[3] : nullptr;
  switch (np) {
  case 1: even_odd_kernel<T, 1><<<grid, kBlk, 0, st>>>(n, p0, p1, p2, p3, c, V, W); break;
  case 2: even_odd_kernel<T, 2><<<grid, kBlk, 0, st>>>(n, p0, p1, p2, p3, c, V, W); break;
  case 3: even_odd_kernel<T, 3><<<grid, kBlk, 0, st>>>(n, p0, p1, p2, p3, c, V, W); break;
  case 4: even_odd_kernel<T, 4><<<grid, kBlk, 0, st>>>(n, p0, p1, p2, p3, c, V, W); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_pade_split(hipStream_t st, int n, const T *U, const T *V, T *R, long ldr, T *Q) {
  if (n < 1) return hipSuccess;
  split_kernel<T><<<dim3((n + kBlk - 1) / kBlk, n), kBlk, 0, st>>>(n, U, V, R, ldr, Q);
  return hipGetLastError();
}

#define CLM_PADE_INST(T)                                                                        \
  template hipError_t launch_pade_even_odd<T>(hipStream_t, int, int, const T *const *,          \
                                              const double *, const double *, T *, T *);         \
  template hipError_t launch_pade_split<T>(hipStream_t, int, const T *, const T *, T *, long, T *);
CLM_PADE_INST(double)
CLM_PADE_INST(float)
CLM_PADE_INST(zd)
CLM_PADE_INST(zc)
#undef CLM_PADE_INST

}  // namespace clm
