// lu.hip -- LU factorization with partial pivoting and the triangular solves of
// the Pade route (?getrf / ?getrs, which the reference takes from the vendor
// solver: src/expm/pade.cppm:186-188).
//
// Recursive LU (the column split of LAPACK's ?getrf2): factor the left half of
// a block column, interchange and solve its right half against the unit-lower
// triangle, subtract the product from the trailing rows, factor the right half.
// The large updates are the library's MFMA products; the recursion stops at
// leaf block columns of up to 32 columns whose whole height fits the LDS
// (128 KiB: 2048 FP64 rows at 8 columns, 4096 at 4, ...), factored by one
// 1024-lane workgroup in LDS -- pivot search by a (value, index) reduction over
// every lane, the interchange, the scaling and the rank-1 update of the leaf's
// remaining columns -- so a leaf costs one load and one store of its rows. Taller
// leaves (past 16384 FP64 rows) run the same steps on the matrix in memory.
// The leaf's interchanges are applied at once to every other column of the
// matrix, which gives the same matrix as LAPACK's deferred application: the
// columns to the right are not touched in between, those to the left are final.
//
// Pivoting follows ?getf2: the pivot is the first entry of largest |.| (|re| +
// |im| for complex, as i?amax), a zero pivot leaves the column unscaled and
// sets info to its (1-based) column once, and the column is scaled by the
// reciprocal of the pivot unless the pivot is below the smallest normal number,
// when it is divided. ipiv is 1-based, as LAPACK's.
#include "clm_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>

namespace clm {
namespace {

constexpr int kLeafThreads = 1024;
constexpr size_t kLeafLdsBytes = 131072;
constexpr int kLeafMaxCols = 32;
constexpr int kTrsmNb = 64;  // triangle order solved in LDS

// i?amax's measure: |x| for real, |re| + |im| for complex
__device__ inline double cabs1(double x) { return fabs(x); }
__device__ inline double cabs1(float x) { return fabsf(x); }
__device__ inline double cabs1(zd x) { return fabs(x.re) + fabs(x.im); }
__device__ inline double cabs1(zc x) { return double(fabsf(x.re) + fabsf(x.im)); }

__device__ inline double sub(double a, double b) { return a - b; }
__device__ inline float sub(float a, float b) { return a - b; }
__device__ inline zd sub(zd a, zd b) { return zd{a.re - b.re, a.im - b.im}; }
__device__ inline zc sub(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }

__device__ inline double cdiv(double a, double b) { return a / b; }
__device__ inline float cdiv(float a, float b) { return a / b; }
// complex division, Smith's method (as the reference LAPACK's ?ladiv family guards
// the ratio of the larger component)
__device__ inline zd cdiv(zd a, zd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return zd{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.im + b.re * r;
  return zd{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ inline zc cdiv(zc a, zc b) {
  if (fabsf(b.re) >= fabsf(b.im)) {
    const float r = b.im / b.re, d = b.re + b.im * r;
    return zc{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const float r = b.re / b.im, d = b.im + b.re * r;
  return zc{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
template <typename T>
__device__ inline double min_normal() {
  return std::is_same_v<real_t<T>, float> ? 1.17549435e-38 : 2.2250738585072014e-308;
}

// ---- the leaf: a block column of m rows and w <= 32 columns -------------------------------
// P: the block (LDS copy, leading dimension ldp = m, or the matrix itself). row0 is
// the global row (and column) of the block's first row; ipiv and info get global
// 1-based indices.
template <typename T>
__device__ void leaf_factor(T *P, long ldp, int m, int w, int row0, int *ipiv, int *info,
                            double *red_v, int *red_i, T *pivot) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  for (int j = 0; j < w; ++j) {
    // pivot: the first row of largest measure among rows j .. m-1 (a NaN measure
    // ranks below every number, so it is taken only from an all-NaN column)
    double best = -1.0;
    int bi = INT_MAX;
    for (int i = j + tid; i < m; i += kLeafThreads) {
      double v = cabs1(P[i + j * ldp]);
      if (!(v == v)) v = -0.5;
      if (v > best) {  // rows ascend within a lane: the first maximum stays
        best = v;
        bi = i;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      red_v[wave] = best;
      red_i[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double b = red_v[0];
      int idx = red_i[0];
      for (int q = 1; q < kLeafThreads / 64; ++q) {
        const double v = red_v[q];
        const int i = red_i[q];
        if (v > b || (v == b && i < idx)) {
          b = v;
          idx = i;
        }
      }
      if (idx == INT_MAX) idx = j;  // row j always competes: not reached
      red_i[kLeafThreads / 64] = idx;
      *pivot = P[idx + j * ldp];  // published before the barrier: lane j exchanges it below
      ipiv[row0 + j] = row0 + idx + 1;
    }
    __syncthreads();
    const int p = red_i[kLeafThreads / 64];
    const T a = *pivot;
    const bool nonzero = !ops<T>::is_zero(a);
    if (nonzero) {
      if (p != j && tid < w) {  // interchange rows j and p of the block
        const T t = P[j + tid * ldp];
        P[j + tid * ldp] = P[p + tid * ldp];
        P[p + tid * ldp] = t;
      }
    } else if (tid == 0) {
      atomicMin(info, row0 + j + 1);
    }
    __syncthreads();
    if (nonzero) {
      // scale the column below the pivot (after the interchange the pivot is at row j)
      if (cabs1(a) >= min_normal<T>()) {
        const T r = cdiv(ops<T>::from_real(1.0), a);
        for (int i = j + 1 + tid; i < m; i += kLeafThreads) P[i + j * ldp] = ops<T>::mul(P[i + j * ldp], r);
      } else {
        for (int i = j + 1 + tid; i < m; i += kLeafThreads) P[i + j * ldp] = cdiv(P[i + j * ldp], a);
      }
    }
    __syncthreads();
    // rank-1 update of the block's remaining columns: a lane a row, across the columns
    if (j + 1 < w) {
      for (int i = j + 1 + tid; i < m; i += kLeafThreads) {
        const T lij = P[i + j * ldp];
        for (int c = j + 1; c < w; ++c)
          P[i + c * ldp] = sub(P[i + c * ldp], ops<T>::mul(lij, P[j + c * ldp]));
      }
    }
    __syncthreads();
  }
}

// the block A[row0 : n, row0 : row0 + w] copied into LDS, factored, copied back
template <typename T>
__global__ void __launch_bounds__(kLeafThreads)
    leaf_lds_kernel(int n, T *A, long lda, int row0, int w, int *ipiv, int *info) {
  __shared__ __attribute__((aligned(16))) char buf[kLeafLdsBytes];
  __shared__ double red_v[kLeafThreads / 64];
  __shared__ int red_i[kLeafThreads / 64 + 1];
  __shared__ T pivot;
  T *P = reinterpret_cast<T *>(buf);
  const int m = n - row0;
  T *blk = A + row0 + static_cast<long>(row0) * lda;
  for (int c = 0; c < w; ++c)
    for (int i = threadIdx.x; i < m; i += kLeafThreads) P[i + c * m] = blk[i + c * lda];
  __syncthreads();
  leaf_factor<T>(P, m, m, w, row0, ipiv, info, red_v, red_i, &pivot);
  for (int c = 0; c < w; ++c)
    for (int i = threadIdx.x; i < m; i += kLeafThreads) blk[i + c * lda] = P[i + c * m];
}

// the same steps on the matrix in memory, for blocks taller than the LDS holds
template <typename T>
__global__ void __launch_bounds__(kLeafThreads)
    leaf_mem_kernel(int n, T *A, long lda, int row0, int w, int *ipiv, int *info) {
  __shared__ double red_v[kLeafThreads / 64];
  __shared__ int red_i[kLeafThreads / 64 + 1];
  __shared__ T pivot;
  leaf_factor<T>(A + row0 + static_cast<long>(row0) * lda, lda, n - row0, w, row0, ipiv, info,
                 red_v, red_i, &pivot);
}

// rows k and ipiv[k] - 1 interchanged for k = k0 .. k1-1 in order, in every column
// c < ncols outside [skip0, skip1): one lane a column
template <typename T>
__global__ void __launch_bounds__(256)
    laswp_kernel(int ncols, T *A, long lda, const int *ipiv, int k0, int k1, int skip0, int skip1) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncols || (c >= skip0 && c < skip1)) return;
  T *col = A + static_cast<long>(c) * lda;
  for (int k = k0; k < k1; ++k) {
    const int p = ipiv[k] - 1;
    if (p != k) {
      const T t = col[k];
      col[k] = col[p];
      col[p] = t;
    }
  }
}

__global__ void fill_info_kernel(int *info, int v) {
  if (threadIdx.x == 0) *info = v;
}
// a first zero pivot was recorded as its column; none leaves INT_MAX, reported as 0
__global__ void finish_info_kernel(int *info) {
  if (threadIdx.x == 0 && *info == INT_MAX) *info = 0;
}

// ---- triangular solves: X = T^-1 B, T of order nb <= 64 in LDS ----------------------------------
// lower = true: T unit lower (forward substitution); false: T upper with its diagonal
// (backward). One lane a right-hand side, 64 columns a workgroup.
template <typename T, bool kLower>
__global__ void __launch_bounds__(64)
    trsm_small_kernel(int nb, int ncols, const T *Tm, long ldt, T *B, long ldb) {
  __shared__ T tri[kTrsmNb][kTrsmNb + 1];
  __shared__ T xs[kTrsmNb][kTrsmNb + 1];  // xs[col][row]
  const int lane = threadIdx.x;
  const int c0 = blockIdx.x * 64;
  for (int j = 0; j < nb; ++j)
    if (lane < nb) tri[lane][j] = Tm[lane + static_cast<long>(j) * ldt];  // tri[row][col]
  const int nc = min(64, ncols - c0);
  for (int c = 0; c < nc; ++c)
    if (lane < nb) xs[c][lane] = B[lane + static_cast<long>(c0 + c) * ldb];
  __syncthreads();
  if (lane < nc) {
    if (kLower) {
      for (int i = 1; i < nb; ++i) {
        T acc = xs[lane][i];
        for (int j = 0; j < i; ++j) acc = sub(acc, ops<T>::mul(tri[i][j], xs[lane][j]));
        xs[lane][i] = acc;
      }
    } else {
      for (int i = nb - 1; i >= 0; --i) {
        T acc = xs[lane][i];
        for (int j = i + 1; j < nb; ++j) acc = sub(acc, ops<T>::mul(tri[i][j], xs[lane][j]));
        xs[lane][i] = cdiv(acc, tri[i][i]);
      }
    }
  }
  __syncthreads();
  for (int c = 0; c < nc; ++c)
    if (lane < nb) B[lane + static_cast<long>(c0 + c) * ldb] = xs[c][lane];
}

template <typename T>
hipError_t gemm_sub(hipStream_t st, int M, int N, int K, const T *A, long lda, const T *B, long ldb,
                    T *C, long ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  GemmArgs<T> g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.alpha = ops<T>::from_real(-1.0);
  g.beta = ops<T>::from_real(1.0);
  g.Cin = C;
  g.ldcin = ldc;
  return launch_gemm<T>(st, g);
}

// X = L^-1 B (L unit lower of order nb) or X = U^-1 B (U upper), recursively:
// halves of the triangle with one product between them
template <typename T, bool kLower>
hipError_t trsm(hipStream_t st, int nb, int ncols, const T *Tm, long ldt, T *B, long ldb) {
  if (nb <= 0 || ncols <= 0) return hipSuccess;
  if (nb <= kTrsmNb) {
    trsm_small_kernel<T, kLower><<<(ncols + 63) / 64, 64, 0, st>>>(nb, ncols, Tm, ldt, B, ldb);
    return hipGetLastError();
  }
  const int n1 = nb / 2, n2 = nb - n1;
  const T *T11 = Tm, *T22 = Tm + n1 + static_cast<long>(n1) * ldt;
  T *B1 = B, *B2 = B + n1;
  hipError_t e;
  if (kLower) {
    if ((e = trsm<T, true>(st, n1, ncols, T11, ldt, B1, ldb)) != hipSuccess) return e;
    if ((e = gemm_sub<T>(st, n2, ncols, n1, Tm + n1, ldt, B1, ldb, B2, ldb)) != hipSuccess) return e;
    return trsm<T, true>(st, n2, ncols, T22, ldt, B2, ldb);
  }
  if ((e = trsm<T, false>(st, n2, ncols, T22, ldt, B2, ldb)) != hipSuccess) return e;
  if ((e = gemm_sub<T>(st, n1, ncols, n2, Tm + static_cast<long>(n1) * ldt, ldt, B2, ldb, B1, ldb)) !=
      hipSuccess)
    return e;
  return trsm<T, false>(st, n1, ncols, T11, ldt, B1, ldb);
}

// leaf width for a block column of m rows: the widest power of two up to 32 whose
// block fits the LDS (0: taller than the LDS holds even one column)
template <typename T>
int leaf_cols(int m) {
  const long cap = static_cast<long>(kLeafLdsBytes / sizeof(T)) / std::max(m, 1);
  int w = 0;
  for (int c = 1; c <= kLeafMaxCols && c <= cap; c <<= 1) w = c;
  return w;
}

template <typename T>
hipError_t factor_block(hipStream_t st, int n, T *A, long lda, int k, int w, int *ipiv, int *info) {
  const int m = n - k;
  const int lw = leaf_cols<T>(m);
  if (w <= std::max(lw, lw == 0 ? 8 : 0)) {
    if (lw > 0) leaf_lds_kernel<T><<<1, kLeafThreads, 0, st>>>(n, A, lda, k, w, ipiv, info);
    else leaf_mem_kernel<T><<<1, kLeafThreads, 0, st>>>(n, A, lda, k, w, ipiv, info);
    laswp_kernel<T><<<(n + 255) / 256, 256, 0, st>>>(n, A, lda, ipiv, k, k + w, k, k + w);
    return hipGetLastError();
  }
  const int w1 = w / 2, w2 = w - w1;
  hipError_t e;
  if ((e = factor_block<T>(st, n, A, lda, k, w1, ipiv, info)) != hipSuccess) return e;
  T *A11 = A + k + static_cast<long>(k) * lda;
  T *A12 = A + k + static_cast<long>(k + w1) * lda;
  T *A21 = A + k + w1 + static_cast<long>(k) * lda;
  T *A22 = A + k + w1 + static_cast<long>(k + w1) * lda;
  if ((e = trsm<T, true>(st, w1, w2, A11, lda, A12, lda)) != hipSuccess) return e;
  if ((e = gemm_sub<T>(st, m - w1, w2, w1, A21, lda, A12, lda, A22, lda)) != hipSuccess) return e;
  return factor_block<T>(st, n, A, lda, k + w1, w2, ipiv, info);
}

}  // namespace

template <typename T>
hipError_t launch_getrf(hipStream_t st, int n, T *A, long lda, int *ipiv, int *info) {
  fill_info_kernel<<<1, 64, 0, st>>>(info, INT_MAX);
  const hipError_t e = factor_block<T>(st, n, A, lda, 0, n, ipiv, info);
  if (e != hipSuccess) return e;
  finish_info_kernel<<<1, 64, 0, st>>>(info);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_getrs(hipStream_t st, int n, int nrhs, const T *LU, long lda, const int *ipiv,
                        T *B, long ldb) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  laswp_kernel<T><<<(nrhs + 255) / 256, 256, 0, st>>>(nrhs, B, ldb, ipiv, 0, n, 0, 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = trsm<T, true>(st, n, nrhs, LU, lda, B, ldb)) != hipSuccess) return e;
  return trsm<T, false>(st, n, nrhs, LU, lda, B, ldb);
}

#define CLM_LU_INST(T)                                                                        \
  template hipError_t launch_getrf<T>(hipStream_t, int, T *, long, int *, int *);             \
  template hipError_t launch_getrs<T>(hipStream_t, int, int, const T *, long, const int *, T *, \
                                      long);
CLM_LU_INST(double)
CLM_LU_INST(float)
CLM_LU_INST(zd)
CLM_LU_INST(zc)
#undef CLM_LU_INST

}  // namespace clm
