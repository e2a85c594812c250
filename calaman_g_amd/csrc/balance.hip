// balance.hip -- balancing (?gebal), the balanced matrix exponential and the 1-norm
// estimator (?lacn2): the reference's calaman.gebal (src/lapack/gebal/) and
// calaman.lacn2 (src/lapack/lacn2/), and the balancing recipe of
// src/expm/README.md:100-122 as one entry point.
//
// gebal (LAPACK's xGEBAL through 3.4, as the reference): first a symmetric
// permutation that isolates eigenvalues -- rows whose off-diagonal part within
// the window vanishes pushed to the bottom, then such columns to the left -- and
// then Parlett-Reinsch sweeps choosing a power-of-two diagonal D that brings each
// off-diagonal row 1-norm and column 1-norm of the window together (|re| + |im|
// for complex, radix 2, a 5% improvement to accept, the SFMIN/SFMAX guards, an
// index with non-finite norms skipped). Each permutation step marks the rows or
// columns with an off-diagonal non-zero in one full-grid pass (a lane an entry of
// a column, coalesced), picks the candidate in one workgroup and reads it back;
// the exchange is two more launches. A sweep is Gauss-Seidel in the index, so it
// runs in one 1024-lane workgroup: per index the four norms by a reduction over
// every lane, the uniform decision, then row and column scaled by exact powers
// of two; one read-back per sweep says whether anything moved.
//
// lacn2 (LAPACK's xLACN2, Hager's method with Higham's refinements, by reverse
// communication): the state machine on the host, the vectors and the sign record
// on the device, the sums and the first index of the largest magnitude reduced on
// the device and read back (one read a step).
#include "clm_internal.hpp"
#include "calaman_expm.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

namespace clm {
namespace {

constexpr int kBlk = 256;
constexpr int kSweepThreads = 1024;
constexpr int kDefaultMaxSweeps = 1000;

__device__ inline double abs1(double x) { return fabs(x); }
__device__ inline float abs1(float x) { return fabsf(x); }
__device__ inline double abs1(zd x) { return fabs(x.re) + fabs(x.im); }
__device__ inline float abs1(zc x) { return fabsf(x.re) + fabsf(x.im); }

template <typename R>
struct Limits {
  R sfmin1, sfmax1, sfmin2, sfmax2;
};
template <typename R>
Limits<R> limits() {
  // DLAMCH('S') / DLAMCH('P'): the smallest normal over the machine epsilon
  const R sfmin1 = std::is_same_v<R, float> ? R(FLT_MIN / FLT_EPSILON) : R(DBL_MIN / DBL_EPSILON);
  const R sfmin2 = sfmin1 * R(2);
  return Limits<R>{sfmin1, R(1) / sfmin1, sfmin2, R(1) / sfmin2};
}

// ---- permutation stage ---------------------------------------------------------------------
// rows: flag[j] = 1 for each row j <= l with A(j, i) != 0 for some i <= l, i != j;
// columns: flag[j] = 1 for each column j in [k, l] with A(i, j) != 0 for some i in
// [k, l], i != j. A lane an entry, a grid row a column.
template <typename T>
__global__ void __launch_bounds__(kBlk)
    mark_rows_kernel(int l, const T *__restrict__ A, long lda, int *flag) {
  const int j = blockIdx.x * kBlk + threadIdx.x;
  const int i = blockIdx.y;
  if (j <= l && j != i && !ops<T>::is_zero(A[j + static_cast<long>(i) * lda])) flag[j] = 1;
}
template <typename T>
__global__ void __launch_bounds__(kBlk)
    mark_cols_kernel(int k, int l, const T *__restrict__ A, long lda, int *flag) {
  const int i = k + blockIdx.x * kBlk + threadIdx.x;
  const int j = k + blockIdx.y;
  if (i <= l && i != j && !ops<T>::is_zero(A[i + static_cast<long>(j) * lda])) flag[j] = 1;
}
__global__ void __launch_bounds__(kBlk) clear_kernel(int *flag, int count) {
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < count; i += gridDim.x * kBlk) flag[i] = 0;
}
// the largest (want_max) or smallest unflagged index in [lo, hi] into *out, or -1
__global__ void __launch_bounds__(kBlk) pick_kernel(const int *flag, int lo, int hi, int want_max, int *out) {
  __shared__ int red[kBlk];
  int best = want_max ? -1 : INT_MAX;
  for (int j = lo + threadIdx.x; j <= hi; j += kBlk)
    if (flag[j] == 0) best = want_max ? max(best, j) : min(best, j);
  red[threadIdx.x] = best;
  __syncthreads();
  for (int s = kBlk / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s)
      red[threadIdx.x] = want_max ? max(red[threadIdx.x], red[threadIdx.x + s])
                                  : min(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = (red[0] == INT_MAX) ? -1 : red[0];
}
// columns a and b exchanged over rows [0, rows)
template <typename T>
__global__ void __launch_bounds__(kBlk) exchange_cols_kernel(T *A, long lda, int a, int b, int rows) {
  const int t = blockIdx.x * kBlk + threadIdx.x;
  if (t < rows) {
    T *ca = A + static_cast<long>(a) * lda, *cb = A + static_cast<long>(b) * lda;
    const T x = ca[t];
    ca[t] = cb[t];
    cb[t] = x;
  }
}
// rows a and b exchanged over columns [c0, n)
template <typename T>
__global__ void __launch_bounds__(kBlk)
    exchange_rows_kernel(int n, T *A, long lda, int a, int b, int c0) {
  const int c = c0 + blockIdx.x * kBlk + threadIdx.x;
  if (c < n) {
    T *col = A + static_cast<long>(c) * lda;
    const T x = col[a];
    col[a] = col[b];
    col[b] = x;
  }
}
template <typename R>
__global__ void set_real_kernel(R *p, int i, R v) {
  if (threadIdx.x == 0) p[i] = v;
}
template <typename R>
__global__ void __launch_bounds__(kBlk) fill_real_kernel(R *p, int count, R v) {
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < count; i += gridDim.x * kBlk) p[i] = v;
}

// ---- scaling stage: one Parlett-Reinsch sweep over the window [k, l] ------------------------
template <typename R>
__device__ inline void reduce4(R &c, R &r, R &ca, R &ra, R *red) {
  for (int off = 32; off > 0; off >>= 1) {
    c += __shfl_xor(c, off);
    r += __shfl_xor(r, off);
    ca = fmax(ca, __shfl_xor(ca, off));
    ra = fmax(ra, __shfl_xor(ra, off));
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int kWaves = kSweepThreads / 64;
  if (lane == 0) {
    red[wave] = c;
    red[kWaves + wave] = r;
    red[2 * kWaves + wave] = ca;
    red[3 * kWaves + wave] = ra;
  }
  __syncthreads();
  c = r = ca = ra = R(0);
  for (int q = 0; q < kWaves; ++q) {  // every lane folds the waves in the same order
    c += red[q];
    r += red[kWaves + q];
    ca = fmax(ca, red[2 * kWaves + q]);
    ra = fmax(ra, red[3 * kWaves + q]);
  }
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(kSweepThreads)
    sweep_kernel(int n, T *A, long lda, int k, int l, real_t<T> *scale, int *noconv,
                 Limits<real_t<T>> lim) {
  using R = real_t<T>;
  __shared__ R red[4 * (kSweepThreads / 64)];
  const int t = threadIdx.x;
  for (int i = k; i <= l; ++i) {
    R c = 0, r = 0, ca = 0, ra = 0;
    const T *col = A + static_cast<long>(i) * lda;
    for (int j = t; j < n; j += kSweepThreads) {
      const R a_col = j <= l ? abs1(col[j]) : R(0);
      const R a_row = j >= k ? abs1(A[i + static_cast<long>(j) * lda]) : R(0);
      if (j >= k && j <= l && j != i) {
        c += a_col;
        r += a_row;
      }
      if (j <= l) ca = fmax(ca, a_col);
      if (j >= k) ra = fmax(ra, a_row);
    }
    reduce4<R>(c, r, ca, ra, red);
    // the decision: every lane from the same values, so it is uniform
    R f = R(1);
    bool accept = false;
    if (c != R(0) && r != R(0) && isfinite(c + r + ca + ra)) {
      const R s0 = c + r;
      R g = r / R(2);
      while (!(c >= g || fmax(fmax(f, c), ca) >= lim.sfmax2 || fmin(fmin(r, g), ra) <= lim.sfmin2)) {
        f *= R(2);
        c *= R(2);
        ca *= R(2);
        r /= R(2);
        g /= R(2);
        ra /= R(2);
      }
      g = c / R(2);
      while (!(g < r || fmax(r, ra) >= lim.sfmax2 || fmin(fmin(fmin(f, c), g), ca) <= lim.sfmin2)) {
        f /= R(2);
        c /= R(2);
        g /= R(2);
        ca /= R(2);
        r *= R(2);
        ra *= R(2);
      }
      if (c + r < R(0.95) * s0) {
        const R sc = scale[i];
        accept = !((f < R(1) && sc < R(1) && f * sc <= lim.sfmin1) ||
                   (f > R(1) && sc > R(1) && sc >= lim.sfmax1 / f));
      }
    }
    if (accept) {
      const R g = R(1) / f;
      for (int j = k + t; j < n; j += kSweepThreads) {
        T &x = A[i + static_cast<long>(j) * lda];
        x = ops<T>::scale(x, g);
      }
      __syncthreads();  // A(i, i) belongs to both the row and the column
      for (int j = t; j <= l; j += kSweepThreads) {
        T &x = A[j + static_cast<long>(i) * lda];
        x = ops<T>::scale(x, f);
      }
      if (t == 0) {
        scale[i] *= f;
        *noconv = 1;
      }
    }
    __syncthreads();
  }
}

// exp(A) = D exp(D^-1 A D) D^-1 from the balanced result: the rows scaled by d_i,
// then the columns by 1 / d_j, in the element type (the reference's two diagonal
// scalings; d = 1 outside the scaled window; powers of two, so exact short of
// overflow)
template <typename T>
__global__ void __launch_bounds__(kBlk)
    rebalance_kernel(int n, T *E, long lde, const real_t<T> *d, int lo, int hi) {
  using R = real_t<T>;
  const int i = blockIdx.x * kBlk + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const R di = (i >= lo && i <= hi) ? d[i] : R(1);
  const R dj = (j >= lo && j <= hi) ? d[j] : R(1);
  if (di == dj) return;
  T x = E[i + static_cast<long>(j) * lde];
  if (di != R(1)) x = ops<T>::scale(x, di);
  if (dj != R(1)) x = ops<T>::scale(x, R(1) / dj);
  E[i + static_cast<long>(j) * lde] = x;
}

}  // namespace

template <typename T>
clm_status_t gebal_run(hipStream_t st, int job, int n, T *A, long lda, int *ilo, int *ihi,
                       real_t<T> *scale, int *work, int max_sweeps) {
  using R = real_t<T>;
  const bool permute = job == CLM_GEBAL_PERMUTE || job == CLM_GEBAL_BOTH;
  const bool scaling = job == CLM_GEBAL_SCALE || job == CLM_GEBAL_BOTH;
  *ilo = 1;
  *ihi = n;
  if (n == 0) return CLM_STATUS_SUCCESS;
  fill_real_kernel<R><<<(n + kBlk - 1) / kBlk, kBlk, 0, st>>>(scale, n, R(1));
  CLM_TRY_HIP(hipGetLastError());
  if (!permute && !scaling) return CLM_STATUS_SUCCESS;
  int *flag = work, *pick = work + n;
  const int sweeps = max_sweeps > 0 ? max_sweeps : kDefaultMaxSweeps;
  int k = 0, l = n - 1;
  if (permute) {
    for (;;) {  // rows isolating an eigenvalue, pushed to the bottom
      if (l == 0) {
        *ilo = 1;
        *ihi = 1;
        return CLM_STATUS_SUCCESS;
      }
      clear_kernel<<<(n + kBlk - 1) / kBlk, kBlk, 0, st>>>(flag, n);
      mark_rows_kernel<T><<<dim3((l + kBlk) / kBlk, l + 1), kBlk, 0, st>>>(l, A, lda, flag);
      pick_kernel<<<1, kBlk, 0, st>>>(flag, 0, l, 1, pick);
      CLM_TRY_HIP(hipGetLastError());
      int j = -1;
      CLM_TRY_HIP(hipMemcpyAsync(&j, pick, sizeof(int), hipMemcpyDeviceToHost, st));
      CLM_TRY_HIP(hipStreamSynchronize(st));
      if (j < 0) break;
      set_real_kernel<R><<<1, 64, 0, st>>>(scale, l, R(j + 1));
      if (j != l) {
        exchange_cols_kernel<T><<<(l + kBlk) / kBlk, kBlk, 0, st>>>(A, lda, j, l, l + 1);
        exchange_rows_kernel<T><<<(n - k + kBlk - 1) / kBlk, kBlk, 0, st>>>(n, A, lda, j, l, k);
      }
      CLM_TRY_HIP(hipGetLastError());
      --l;
    }
    while (k <= l) {  // columns isolating an eigenvalue, pushed to the left
      clear_kernel<<<(n + kBlk - 1) / kBlk, kBlk, 0, st>>>(flag, n);
      mark_cols_kernel<T><<<dim3((l - k + kBlk) / kBlk, l - k + 1), kBlk, 0, st>>>(k, l, A, lda, flag);
      pick_kernel<<<1, kBlk, 0, st>>>(flag, k, l, 0, pick);
      CLM_TRY_HIP(hipGetLastError());
      int j = -1;
      CLM_TRY_HIP(hipMemcpyAsync(&j, pick, sizeof(int), hipMemcpyDeviceToHost, st));
      CLM_TRY_HIP(hipStreamSynchronize(st));
      if (j < 0) break;
      set_real_kernel<R><<<1, 64, 0, st>>>(scale, k, R(j + 1));
      if (j != k) {
        exchange_cols_kernel<T><<<(l + kBlk) / kBlk, kBlk, 0, st>>>(A, lda, j, k, l + 1);
        exchange_rows_kernel<T><<<(n - k + kBlk - 1) / kBlk, kBlk, 0, st>>>(n, A, lda, j, k, k);
      }
      CLM_TRY_HIP(hipGetLastError());
      ++k;
    }
  }
  *ilo = k + 1;
  *ihi = l + 1;
  if (scaling && k <= l) {
    const Limits<R> lim = limits<R>();
    for (int s = 0; s < sweeps; ++s) {
      clear_kernel<<<1, kBlk, 0, st>>>(pick, 1);
      sweep_kernel<T><<<1, kSweepThreads, 0, st>>>(n, A, lda, k, l, scale, pick, lim);
      CLM_TRY_HIP(hipGetLastError());
      int noconv = 0;
      CLM_TRY_HIP(hipMemcpyAsync(&noconv, pick, sizeof(int), hipMemcpyDeviceToHost, st));
      CLM_TRY_HIP(hipStreamSynchronize(st));
      if (noconv == 0) break;
    }
  }
  return CLM_STATUS_SUCCESS;
}

template <typename T>
hipError_t launch_rebalance(hipStream_t st, int n, T *E, long lde, const real_t<T> *d, int lo,
                            int hi) {
  rebalance_kernel<T><<<dim3((n + kBlk - 1) / kBlk, n), kBlk, 0, st>>>(n, E, lde, d, lo, hi);
  return hipGetLastError();
}

#define CLM_BAL_INST(T)                                                                         \
  template clm_status_t gebal_run<T>(hipStream_t, int, int, T *, long, int *, int *, real_t<T> *, \
                                     int *, int);                                              \
  template hipError_t launch_rebalance<T>(hipStream_t, int, T *, long, const real_t<T> *, int, int);
CLM_BAL_INST(double)
CLM_BAL_INST(float)
CLM_BAL_INST(zd)
CLM_BAL_INST(zc)
#undef CLM_BAL_INST

// ---- lacn2 ------------------------------------------------------------------------------------
namespace {
template <typename T>
__global__ void __launch_bounds__(kBlk) lacn2_fill_kernel(int n, T v, T *x) {
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) x[i] = v;
}
template <typename T>
__global__ void __launch_bounds__(kBlk) lacn2_unit_kernel(int n, int j, T *x) {
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) x[i] = T(i == j ? 1 : 0);
}
// x(i) = altsgn (1 + i / (n - 1)), altsgn = +1, -1, ...
template <typename T>
__global__ void __launch_bounds__(kBlk) lacn2_altsgn_kernel(int n, T *x) {
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
    const T v = T(1) + static_cast<T>(i) / static_cast<T>(n - 1);
    x[i] = (i % 2 == 0) ? v : -v;
  }
}
// x(i) = sign(x(i)) (+1 for x >= 0, so -0 and NaN map as the reference's X(I) .GE. ZERO)
template <typename T>
__global__ void __launch_bounds__(kBlk) lacn2_sign_kernel(int n, T *x, int *isgn) {
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
    const int s = x[i] >= T(0) ? 1 : -1;
    x[i] = T(s);
    isgn[i] = s;
  }
}
template <typename T>
__global__ void __launch_bounds__(kBlk) lacn2_changed_kernel(int n, const T *x, const int *isgn, int *flag) {
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk)
    if ((x[i] >= T(0) ? 1 : -1) != isgn[i]) *flag = 1;
}
// out[0] = sum |x|, out[1] = the first index (1-based) of the largest |x|
// (i?amax: the first maximum; a NaN never wins a comparison), out[2] = flag
template <typename T>
__global__ void __launch_bounds__(kBlk) lacn2_reduce_kernel(int n, const T *x, const int *flag, double *out) {
  __shared__ double s_sum[kBlk];
  __shared__ double s_max[kBlk];
  __shared__ int s_idx[kBlk];
  double sum = 0.0, best = -1.0;
  int idx = INT_MAX;
  for (int i = threadIdx.x; i < n; i += kBlk) {
    const double a = fabs(double(x[i]));
    sum += a;
    if (a > best) {
      best = a;
      idx = i;
    }
  }
  s_sum[threadIdx.x] = sum;
  s_max[threadIdx.x] = best;
  s_idx[threadIdx.x] = idx;
  __syncthreads();
  for (int s = kBlk / 2; s > 0; s >>= 1) {
    const int t = threadIdx.x;
    if (t < s) {
      s_sum[t] += s_sum[t + s];
      if (s_max[t + s] > s_max[t] || (s_max[t + s] == s_max[t] && s_idx[t + s] < s_idx[t])) {
        s_max[t] = s_max[t + s];
        s_idx[t] = s_idx[t + s];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s_sum[0];
    out[1] = s_idx[0] == INT_MAX ? 1.0 : double(s_idx[0] + 1);
    out[2] = flag ? double(*flag) : 0.0;
  }
}
__global__ void lacn2_clear_flag_kernel(int *flag) {
  if (threadIdx.x == 0) *flag = 0;
}

// [isgn | n ints] [flag | 1 int] [reduction | 4 doubles]
struct Lacn2Ws {
  int *isgn = nullptr, *flag = nullptr;
  double *red = nullptr;
  size_t carve(void *base, int n) {
    Layout L(base);
    isgn = L.take<int>(static_cast<size_t>(n < 1 ? 1 : n));
    flag = L.take<int>(1);
    red = L.take<double>(4);
    return L.total();
  }
};
}  // namespace

size_t lacn2_buffer_size(int n) {
  Lacn2Ws w;
  return w.carve(nullptr, n);
}

template <typename T>
clm_status_t lacn2_step(hipStream_t st, int n, T *v, T *x, void *work, size_t bytes, T *est,
                        int *kase, int *isave) {
  constexpr int kItmax = 5;
  if (n < 1 || v == nullptr || x == nullptr || work == nullptr || est == nullptr ||
      kase == nullptr || isave == nullptr || (*kase != 0 && (isave[0] < 1 || isave[0] > 5)))
    return CLM_STATUS_INVALID_VALUE;
  // the column index the later steps carry (1-based) is read on the device: a state
  // that was not left by a previous step is refused before it can address past x
  if (*kase != 0 && isave[0] >= 3 && (isave[1] < 1 || isave[1] > n)) return CLM_STATUS_INVALID_VALUE;
  Lacn2Ws ws;
  if (bytes < ws.carve(work, n)) return CLM_STATUS_INVALID_VALUE;
  const int blocks = std::min((n + kBlk - 1) / kBlk, 1024);
  auto reduce = [&](const int *flag, double *out) -> clm_status_t {
    int device = 0;
    CLM_TRY_HIP(hipGetDevice(&device));
    lacn2_reduce_kernel<T><<<1, kBlk, 0, st>>>(n, x, flag, ws.red);
    CLM_TRY_HIP(hipGetLastError());
    return read_device_values<double>(st, device, ws.red, 3, out);
  };
  auto to_unit = [&]() -> clm_status_t {
    lacn2_unit_kernel<T><<<blocks, kBlk, 0, st>>>(n, isave[1] - 1, x);
    CLM_TRY_HIP(hipGetLastError());
    *kase = 1;
    isave[0] = 3;
    return CLM_STATUS_SUCCESS;
  };
  auto to_final = [&]() -> clm_status_t {
    lacn2_altsgn_kernel<T><<<blocks, kBlk, 0, st>>>(n, x);
    CLM_TRY_HIP(hipGetLastError());
    *kase = 1;
    isave[0] = 5;
    return CLM_STATUS_SUCCESS;
  };
  auto to_sign = [&](int next) -> clm_status_t {
    lacn2_sign_kernel<T><<<blocks, kBlk, 0, st>>>(n, x, ws.isgn);
    CLM_TRY_HIP(hipGetLastError());
    *kase = 2;
    isave[0] = next;
    return CLM_STATUS_SUCCESS;
  };
  if (*kase == 0) {
    lacn2_fill_kernel<T><<<blocks, kBlk, 0, st>>>(n, T(1) / static_cast<T>(n), x);
    CLM_TRY_HIP(hipGetLastError());
    *kase = 1;
    isave[0] = 1;
    return CLM_STATUS_SUCCESS;
  }
  double r[3] = {0, 0, 0};
  switch (isave[0]) {
  case 1: {  // x = A x for x = e / n
    if (n == 1) CLM_TRY_HIP(hipMemcpyAsync(v, x, sizeof(T), hipMemcpyDeviceToDevice, st));
    CLM_TRY(reduce(nullptr, r));
    *est = static_cast<T>(r[0]);  // n = 1: |x(1)|
    if (n == 1) {
      *kase = 0;
      return CLM_STATUS_SUCCESS;
    }
    return to_sign(2);
  }
  case 2: {  // x = A^T sign: the first column guess
    CLM_TRY(reduce(nullptr, r));
    isave[1] = static_cast<int>(r[1]);
    isave[2] = 2;
    return to_unit();
  }
  case 3: {  // x = A e_j
    CLM_TRY_HIP(hipMemcpyAsync(v, x, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    const T estold = *est;
    lacn2_clear_flag_kernel<<<1, 64, 0, st>>>(ws.flag);
    lacn2_changed_kernel<T><<<blocks, kBlk, 0, st>>>(n, x, ws.isgn, ws.flag);
    CLM_TRY(reduce(ws.flag, r));
    *est = static_cast<T>(r[0]);
    if (r[2] == 0.0 || *est <= estold) return to_final();  // repeated signs, or cycling
    return to_sign(4);
  }
  case 4: {  // x = A^T sign: the next column guess
    const int jlast = isave[1];
    CLM_TRY(reduce(nullptr, r));
    isave[1] = static_cast<int>(r[1]);
    T xs[2] = {T(0), T(0)};
    CLM_TRY_HIP(hipMemcpyAsync(&xs[0], x + (jlast - 1), sizeof(T), hipMemcpyDeviceToHost, st));
    CLM_TRY_HIP(hipMemcpyAsync(&xs[1], x + (isave[1] - 1), sizeof(T), hipMemcpyDeviceToHost, st));
    CLM_TRY_HIP(hipStreamSynchronize(st));
    if (xs[0] != std::fabs(xs[1]) && isave[2] < kItmax) {
      ++isave[2];
      return to_unit();
    }
    return to_final();
  }
  default: {  // x = A altsgn
    CLM_TRY(reduce(nullptr, r));
    const T temp = T(2) * (static_cast<T>(r[0]) / static_cast<T>(3 * n));
    if (temp > *est) {
      CLM_TRY_HIP(hipMemcpyAsync(v, x, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
      *est = temp;
    }
    *kase = 0;
    return CLM_STATUS_SUCCESS;
  }
  }
}
template clm_status_t lacn2_step<double>(hipStream_t, int, double *, double *, void *, size_t,
                                         double *, int *, int *);
template clm_status_t lacn2_step<float>(hipStream_t, int, float *, float *, void *, size_t, float *,
                                        int *, int *);

}  // namespace clm

// ---- the balanced exponential (src/expm/README.md:100-122): B = D^-1 A D by gebal's
// scaling half, exp(B), then exp(A) = D exp(B) D^-1 ----------------------------------------------
namespace clm {
namespace {
template <typename T>
struct Api;
template <>
struct Api<float> {
  static clm_status_t bytes(int n, size_t *b) { return clm_expm_buffer_size_s(n, b); }
  static clm_status_t expm(clm_ctx *c, hipStream_t st, int n, const float *a, float *e, long lde,
                           void *w, size_t b, int *info, clm_expm_plan *p) {
    return clm_expm_s(c, st, n, a, n, e, static_cast<int>(lde), w, b, info, p);
  }
};
template <>
struct Api<double> {
  static clm_status_t bytes(int n, size_t *b) { return clm_expm_buffer_size_d(n, b); }
  static clm_status_t expm(clm_ctx *c, hipStream_t st, int n, const double *a, double *e, long lde,
                           void *w, size_t b, int *info, clm_expm_plan *p) {
    return clm_expm_d(c, st, n, a, n, e, static_cast<int>(lde), w, b, info, p);
  }
};
template <>
struct Api<zc> {
  static clm_status_t bytes(int n, size_t *b) { return clm_expm_buffer_size_c(n, b); }
  static clm_status_t expm(clm_ctx *c, hipStream_t st, int n, const zc *a, zc *e, long lde, void *w,
                           size_t b, int *info, clm_expm_plan *p) {
    return clm_expm_c(c, st, n, reinterpret_cast<const clm_float_complex *>(a), n,
                      reinterpret_cast<clm_float_complex *>(e), static_cast<int>(lde), w, b, info, p);
  }
};
template <>
struct Api<zd> {
  static clm_status_t bytes(int n, size_t *b) { return clm_expm_buffer_size_z(n, b); }
  static clm_status_t expm(clm_ctx *c, hipStream_t st, int n, const zd *a, zd *e, long lde, void *w,
                           size_t b, int *info, clm_expm_plan *p) {
    return clm_expm_z(c, st, n, reinterpret_cast<const clm_double_complex *>(a), n,
                      reinterpret_cast<clm_double_complex *>(e), static_cast<int>(lde), w, b, info, p);
  }
};

// [B | n*n T] [d | n reals] [gebal ints | 2n + 1] [the expm workspace]
template <typename T>
size_t balanced_head(int n) {
  Layout L(nullptr);
  (void)L.take<T>(static_cast<size_t>(n) * n);
  (void)L.take<real_t<T>>(static_cast<size_t>(n));
  (void)L.take<int>(2 * static_cast<size_t>(n) + 1);
  return L.total();
}
template <typename T>
clm_status_t balanced_bytes(int n, size_t *bytes) {
  if (bytes == nullptr || n < 1 || n > kMaxDim) return CLM_STATUS_INVALID_VALUE;
  size_t eb = 0;
  CLM_TRY(Api<T>::bytes(n, &eb));
  *bytes = balanced_head<T>(n) + eb;
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t expm_balanced_impl(clm_ctx *ctx, hipStream_t st, int n, const T *A, int lda, T *E,
                                int lde, void *work, size_t bytes, int *info, clm_expm_plan *plan) {
  if (n < 1 || n > kMaxDim || lda < n || lde < n || A == nullptr || E == nullptr || work == nullptr)
    return CLM_STATUS_INVALID_VALUE;
  size_t need = 0;
  CLM_TRY(balanced_bytes<T>(n, &need));
  if (bytes < need) return CLM_STATUS_ALLOC_FAILED;
  Layout L(work);
  T *B = L.take<T>(static_cast<size_t>(n) * n);
  real_t<T> *d = L.take<real_t<T>>(static_cast<size_t>(n));
  int *iw = L.take<int>(2 * static_cast<size_t>(n) + 1);
  void *ework = static_cast<char *>(work) + L.total();
  CLM_TRY_HIP(launch_copy<T>(st, n, A, lda, B, n));
  int ilo = 1, ihi = n;
  CLM_TRY(gebal_run<T>(st, CLM_GEBAL_SCALE, n, B, n, &ilo, &ihi, d, iw, 0));
  CLM_TRY(Api<T>::expm(ctx, st, n, B, E, lde, ework, bytes - L.total(), info, plan));
  CLM_TRY_HIP(launch_rebalance<T>(st, n, E, lde, d, ilo - 1, ihi - 1));
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t gebal_impl(clm_stream_t stream, int job, int n, T *A, int lda, int *ilo, int *ihi,
                        real_t<T> *scale, int *work, int max_sweeps) {
  if (n < 0 || lda < n || A == nullptr || scale == nullptr || ilo == nullptr || ihi == nullptr ||
      job < CLM_GEBAL_NONE || job > CLM_GEBAL_BOTH)
    return CLM_STATUS_RUNTIME_INVALID_VALUE;
  if (n > 0 && work == nullptr) return CLM_STATUS_RUNTIME_INVALID_VALUE;
  const clm_status_t s = gebal_run<T>(static_cast<hipStream_t>(stream), job, n, A, lda, ilo, ihi,
                                      scale, work, max_sweeps);
  return s == CLM_STATUS_SUCCESS ? s : CLM_STATUS_RUNTIME_ERROR;
}

template <typename T>
clm_status_t lacn2_impl(clm_stream_t stream, int n, T *v, T *x, void *work, size_t bytes, T *est,
                        int *kase, int *isave) {
  return lacn2_step<T>(static_cast<hipStream_t>(stream), n, v, x, work, bytes, est, kase, isave);
}
}  // namespace
}  // namespace clm

using clm::zc;
using clm::zd;

extern "C" {

clm_status_t clm_gebal_buffer_size(int n, int *lwork) {
  if (lwork == nullptr || n < 0) return CLM_STATUS_RUNTIME_INVALID_VALUE;
  *lwork = 2 * n + 1;
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_gebal_s(clm_stream_t stream, int job, int n, float *dA, int lda, int *ilo,
                         int *ihi, float *dScale, int *dWork, int max_sweeps) {
  return clm::gebal_impl<float>(stream, job, n, dA, lda, ilo, ihi, dScale, dWork, max_sweeps);
}
clm_status_t clm_gebal_d(clm_stream_t stream, int job, int n, double *dA, int lda, int *ilo,
                         int *ihi, double *dScale, int *dWork, int max_sweeps) {
  return clm::gebal_impl<double>(stream, job, n, dA, lda, ilo, ihi, dScale, dWork, max_sweeps);
}
clm_status_t clm_gebal_c(clm_stream_t stream, int job, int n, clm_float_complex *dA, int lda,
                         int *ilo, int *ihi, float *dScale, int *dWork, int max_sweeps) {
  return clm::gebal_impl<zc>(stream, job, n, reinterpret_cast<zc *>(dA), lda, ilo, ihi, dScale,
                             dWork, max_sweeps);
}
clm_status_t clm_gebal_z(clm_stream_t stream, int job, int n, clm_double_complex *dA, int lda,
                         int *ilo, int *ihi, double *dScale, int *dWork, int max_sweeps) {
  return clm::gebal_impl<zd>(stream, job, n, reinterpret_cast<zd *>(dA), lda, ilo, ihi, dScale,
                             dWork, max_sweeps);
}

clm_status_t clm_expm_balanced_buffer_size_s(int n, size_t *b) { return clm::balanced_bytes<float>(n, b); }
clm_status_t clm_expm_balanced_buffer_size_d(int n, size_t *b) { return clm::balanced_bytes<double>(n, b); }
clm_status_t clm_expm_balanced_buffer_size_c(int n, size_t *b) { return clm::balanced_bytes<zc>(n, b); }
clm_status_t clm_expm_balanced_buffer_size_z(int n, size_t *b) { return clm::balanced_bytes<zd>(n, b); }

clm_status_t clm_expm_balanced_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dA, int lda,
                                 float *dE, int lde, void *dWork, size_t lwork, int *dInfo,
                                 clm_expm_plan *plan) {
  return clm::expm_balanced_impl<float>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dE, lde,
                                        dWork, lwork, dInfo, plan);
}
clm_status_t clm_expm_balanced_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA, int lda,
                                 double *dE, int lde, void *dWork, size_t lwork, int *dInfo,
                                 clm_expm_plan *plan) {
  return clm::expm_balanced_impl<double>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dE, lde,
                                         dWork, lwork, dInfo, plan);
}
clm_status_t clm_expm_balanced_c(clm_ctx *ctx, clm_stream_t stream, int n,
                                 const clm_float_complex *dA, int lda, clm_float_complex *dE,
                                 int lde, void *dWork, size_t lwork, int *dInfo,
                                 clm_expm_plan *plan) {
  return clm::expm_balanced_impl<zc>(ctx, static_cast<hipStream_t>(stream), n,
                                     reinterpret_cast<const zc *>(dA), lda,
                                     reinterpret_cast<zc *>(dE), lde, dWork, lwork, dInfo, plan);
}
clm_status_t clm_expm_balanced_z(clm_ctx *ctx, clm_stream_t stream, int n,
                                 const clm_double_complex *dA, int lda, clm_double_complex *dE,
                                 int lde, void *dWork, size_t lwork, int *dInfo,
                                 clm_expm_plan *plan) {
  return clm::expm_balanced_impl<zd>(ctx, static_cast<hipStream_t>(stream), n,
                                     reinterpret_cast<const zd *>(dA), lda,
                                     reinterpret_cast<zd *>(dE), lde, dWork, lwork, dInfo, plan);
}

size_t clm_lacn2_buffer_size_s(int n) { return clm::lacn2_buffer_size(n); }
size_t clm_lacn2_buffer_size_d(int n) { return clm::lacn2_buffer_size(n); }
clm_status_t clm_lacn2_s(clm_stream_t stream, int n, float *dV, float *dX, void *dWork,
                         size_t work_bytes, float *est, int *kase, int *isave) {
  return clm::lacn2_impl<float>(stream, n, dV, dX, dWork, work_bytes, est, kase, isave);
}
clm_status_t clm_lacn2_d(clm_stream_t stream, int n, double *dV, double *dX, void *dWork,
                         size_t work_bytes, double *est, int *kase, int *isave) {
  return clm::lacn2_impl<double>(stream, n, dV, dX, dWork, work_bytes, est, kase, isave);
}

}  // extern "C"
