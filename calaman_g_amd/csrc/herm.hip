// herm.hip -- exp(A) for a symmetric (real) or Hermitian (complex) A by its
// eigendecomposition: the reference's expm_herm (src/expm/expm_herm.cppm:145-199,
// herm_exp_scale in src/expm/expm.cu:170-184). The reference takes the
// eigendecomposition from the vendor solver (syevd / heevd); here it is a blocked
// two-sided Jacobi method of this library's own:
//
//   * the uplo triangle is mirrored into a full matrix H of order np = n rounded up
//     to 64 (zero padding adds eigenvalues 0 whose exponentials land outside the
//     n x n block the result keeps, so the block is exp(A) exactly as before);
//   * H is cut into 2k blocks of 32; a sweep pairs every block with every other once
//     over 2k - 1 rounds (the circle method). A round is two launches: one workgroup
//     per pair runs one cyclic Jacobi sweep over its 64 x 64 subproblem in LDS
//     (32 disjoint rotations at a time, circle order again), keeping the unitary Q;
//     then one launch applies Q_t^H H_tu Q_u to every off-diagonal tile (t < u, the
//     mirror written as its adjoint) and V <- V Q_u to the eigenvector tiles, skipping
//     pairs whose Q stayed the identity;
//   * an entry is rotated when its modulus exceeds delta = u ||H||_F / np, so when a
//     whole sweep rotates nothing every off-diagonal entry is at most delta and
//     ||off(H)||_F <= u ||H||_F (u = 2^-53 or 2^-24); one read-back per sweep;
//   * M = V diag(exp(w)) over the first n rows, then exp(A) = M V^H on the library's
//     MFMA product (V^H staged once, the product is NN).
#include "clm_internal.hpp"
#include "calaman_expm.h"

#include <algorithm>
#include <cmath>

namespace clm {
namespace {

constexpr int kB = 32;             // block width
constexpr int kP = 2 * kB;         // a pair of blocks: the subproblem order
constexpr int kLd = kP + 1;        // LDS row pitch
constexpr int kThreads = 256;      // one workgroup per tile (a lane a 4 x 4 block)
constexpr int kSolveThreads = 1024;  // one workgroup per subproblem (32 lanes a rotation)
constexpr int kInnerSweeps = 1;    // Jacobi sweeps of one subproblem per round: more
                                   // sweeps save no outer sweep at n >= 128 and every
                                   // rotation adds its rounding to V
constexpr int kMaxSweeps = 30;     // cap on the sweeps over the whole matrix

__device__ inline double conj_(double x) { return x; }
__device__ inline float conj_(float x) { return x; }
__device__ inline zd conj_(zd x) { return zd{x.re, -x.im}; }
__device__ inline zc conj_(zc x) { return zc{x.re, -x.im}; }
__device__ inline double re_(double x) { return x; }
__device__ inline float re_(float x) { return x; }
__device__ inline double re_(zd x) { return x.re; }
__device__ inline float re_(zc x) { return x.re; }
__device__ inline double abs_(double x) { return fabs(x); }
__device__ inline double abs_(zd x) { return hypot(x.re, x.im); }
__device__ inline double abs2_(double x) { return x * x; }
__device__ inline double abs2_(zd x) { return x.re * x.re + x.im * x.im; }
__device__ inline double exp_(double x) { return exp(x); }
// The working type: the Jacobi sweeps of the single-precision types run in double
// (their rotations' rounding would otherwise reach 1e-5 of exp(A) at n = 128)
template <typename T>
struct work_type {
  using type = T;
};
template <>
struct work_type<float> {
  using type = double;
};
template <>
struct work_type<zc> {
  using type = zd;
};
template <typename T>
using work_t = typename work_type<T>::type;
__device__ inline double widen(double x) { return x; }
__device__ inline double widen(float x) { return x; }
__device__ inline zd widen(zd x) { return x; }
__device__ inline zd widen(zc x) { return zd{x.re, x.im}; }
template <typename T>
__device__ inline T narrow(work_t<T> x);  // defined for the single-precision types
template <>
__device__ inline float narrow<float>(double x) { return static_cast<float>(x); }
template <>
__device__ inline zc narrow<zc>(zd x) { return zc{static_cast<float>(x.re), static_cast<float>(x.im)}; }

template <typename T>
__device__ inline T sub_(T a, T b) {
  return ops<T>::add(a, ops<T>::scale(b, -1.0));
}
template <typename T>
__device__ inline T mac_(T acc, T a, T b) {
  return ops<T>::add(acc, ops<T>::mul(a, b));
}

// the circle method over np2 positions (np2 even): pair t of round r
__device__ inline void pair_of(int r, int t, int np2, int &p, int &q) {
  const int m = np2 - 1;
  if (t == 0) {
    p = r;
    q = m;
  } else {
    p = (r + t) % m;
    q = (r - t + m) % m;
  }
}

// H = the uplo triangle of A mirrored (real diagonal), zero-padded to np x np, in the
// working type W
template <typename T, typename W>
__global__ void __launch_bounds__(kThreads)
    mirror_kernel(int n, int np, int upper, const T *__restrict__ A, long lda, W *__restrict__ H) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= np) return;
  W v = ops<W>::zero();
  if (i < n && j < n) {
    if (i == j) v = ops<W>::from_real(re_(A[i + static_cast<long>(i) * lda]));
    else if ((i < j) == (upper != 0)) v = widen(A[i + static_cast<long>(j) * lda]);
    else v = widen(conj_(A[j + static_cast<long>(i) * lda]));
  }
  H[i + static_cast<long>(j) * np] = v;
}
// E = the n x n result rounded to the caller's type
template <typename T, typename W>
__global__ void __launch_bounds__(kThreads)
    narrow_kernel(int n, const W *__restrict__ X, T *__restrict__ E, long lde) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) E[i + static_cast<long>(j) * lde] = narrow<T>(X[i + static_cast<long>(j) * n]);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) identity_kernel(int np, T *V) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  const int j = blockIdx.y;
  if (i < np) V[i + static_cast<long>(j) * np] = ops<T>::from_real(i == j ? 1.0 : 0.0);
}

// colsq[j] = sum_i |H(i, j)|^2
template <typename T>
__global__ void __launch_bounds__(kThreads) colsq_kernel(int np, const T *H, real_t<T> *colsq) {
  using R = real_t<T>;
  __shared__ R red[kThreads];
  const int j = blockIdx.x;
  R s = R(0);
  for (int i = threadIdx.x; i < np; i += kThreads) s += abs2_(H[i + static_cast<long>(j) * np]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) colsq[j] = red[0];
}
// delta = u sqrt(sum colsq) / np
template <typename T>
__global__ void __launch_bounds__(kThreads)
    delta_kernel(int np, const real_t<T> *colsq, real_t<T> u, real_t<T> *delta) {
  using R = real_t<T>;
  __shared__ R red[kThreads];
  R s = R(0);
  for (int j = threadIdx.x; j < np; j += kThreads) s += colsq[j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *delta = u * sqrt(red[0]) / static_cast<R>(np);
}

__global__ void clear_count_kernel(unsigned *c) {
  if (threadIdx.x == 0) *c = 0u;
}

// local index l of a 64-wide pair (p, q) -> global index
__device__ inline int glob(int l, int p, int q) { return l < kB ? p * kB + l : q * kB + (l - kB); }

// The subproblem of pair t in round r: S = H[{p,q},{p,q}] in LDS, one cyclic sweep of
// rotations J = [[c, e s], [-conj(e) s, c]] (e = S_ij / |S_ij|, t = sign(tau) /
// (|tau| + sqrt(1 + tau^2)), tau = (S_jj - S_ii) / (2 |S_ij|), c = 1 / sqrt(1 + t^2),
// s = t c: the real symmetric Jacobi rotation conjugated by the phase of S_ij),
// S <- J^H S J and Q <- Q J. 32 lanes serve one rotation of a round: each computes
// the rotation itself from the same three entries (no broadcast), updates two of
// the pair's 64 columns of rows i, j, then -- after every row update -- two rows of
// columns i, j of S and of Q. Two barriers a round. Writes S back and Q to Qbuf[t]
// when anything rotated.
template <typename W>
__global__ void __launch_bounds__(kSolveThreads)
    solve_kernel(int np, W *H, const real_t<W> *delta, int round, W *Qbuf, int *flags,
                 unsigned *count) {
  using R = real_t<W>;
  __shared__ W S[kP][kLd];
  __shared__ W Q[kP][kLd];
  __shared__ int sweep_rot;
  const int np2 = np / kB;
  const int t = blockIdx.x;
  int p, q;
  pair_of(round, t, np2, p, q);
  const long ld = np;
  for (int e = threadIdx.x; e < kP * kP; e += kSolveThreads) {
    const int r = e % kP, c = e / kP;
    S[r][c] = H[glob(r, p, q) + glob(c, p, q) * ld];
    Q[r][c] = ops<W>::from_real(r == c ? 1.0 : 0.0);
  }
  const int m = threadIdx.x >> 5;  // the rotation this lane serves (32 lanes each)
  const int sl = threadIdx.x & 31;
  const R d = *delta;
  bool any = false;
  for (int sw = 0; sw < kInnerSweeps; ++sw) {
    if (threadIdx.x == 0) sweep_rot = 0;
    __syncthreads();
    for (int r = 0; r < kP - 1; ++r) {
      int a, b;
      pair_of(r, m, kP, a, b);
      const int i = min(a, b), j = max(a, b);
      // rows i and j are this pair's alone until the barrier: the three entries are
      // read by all 32 lanes before any of them writes
      const W c = S[i][j];
      const R absc = abs_(c);
      const bool on = absc > d;  // false for NaN
      R cs = R(1), di = R(0), dj = R(0);
      W es = ops<W>::zero();
      if (on) {
        const R aii = re_(S[i][i]), ajj = re_(S[j][j]);
        const R tau = (ajj - aii) / (R(2) * absc);
        const R at = fabs(tau);
        const R tt = at > R(1e150) ? R(0.5) / tau
                                   : (tau >= R(0) ? R(1) : R(-1)) / (at + sqrt(R(1) + tau * tau));
        cs = R(1) / sqrt(R(1) + tt * tt);
        es = ops<W>::scale(c, tt * cs / absc);
        di = aii - tt * absc;
        dj = ajj + tt * absc;
#pragma unroll
        for (int h = 0; h < 2; ++h) {  // rows i, j: S <- J^H S
          const int col = sl + 32 * h;
          const W xi = S[i][col], xj = S[j][col];
          S[i][col] = sub_(ops<W>::scale(xi, cs), ops<W>::mul(es, xj));
          S[j][col] = mac_(ops<W>::scale(xj, cs), conj_(es), xi);
        }
        if (sl == 0) atomicAdd(&sweep_rot, 1);
      }
      __syncthreads();
      if (on) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {  // columns i, j: S <- S J, Q <- Q J
          const int row = sl + 32 * h;
          const W xi = S[row][i], xj = S[row][j];
          W ni = sub_(ops<W>::scale(xi, cs), ops<W>::mul(conj_(es), xj));
          W nj = mac_(ops<W>::scale(xj, cs), es, xi);
          if (row == i) {
            ni = ops<W>::from_real(di);
            nj = ops<W>::zero();
          } else if (row == j) {
            ni = ops<W>::zero();
            nj = ops<W>::from_real(dj);
          }
          S[row][i] = ni;
          S[row][j] = nj;
          const W qi = Q[row][i], qj = Q[row][j];
          Q[row][i] = sub_(ops<W>::scale(qi, cs), ops<W>::mul(conj_(es), qj));
          Q[row][j] = mac_(ops<W>::scale(qj, cs), es, qi);
        }
      }
      __syncthreads();
    }
    const bool quiet = sweep_rot == 0;
    __syncthreads();  // every lane has read the count before the next sweep clears it
    if (quiet) break;
    any = true;
  }
  if (any) {
    W *Qt = Qbuf + static_cast<long>(t) * kP * kP;
    for (int e = threadIdx.x; e < kP * kP; e += kSolveThreads) {
      const int r = e % kP, c = e / kP;
      H[glob(r, p, q) + glob(c, p, q) * ld] = S[r][c];
      Qt[r + c * kP] = Q[r][c];
    }
  }
  if (threadIdx.x == 0) {
    flags[t] = any ? 1 : 0;
    if (any) atomicAdd(count, 1u);
  }
}

// acc = X * Y over 64 x 64 tiles in LDS (adj: X^H * Y); a lane a 4 x 4 block of the
// result: rows lane % 16 + 16 a (16 consecutive rows per k across a lane group, so
// the X reads fall on distinct banks), columns 4 (lane / 16) + b (one Y element a
// lane group: a broadcast)
template <typename T, bool kAdjX>
__device__ inline void tile_product(const T (*X)[kLd], const T (*Y)[kLd], T (&acc)[4][4]) {
  const int r0 = threadIdx.x % 16, c0 = 4 * (threadIdx.x / 16);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = ops<T>::zero();
  for (int k = 0; k < kP; ++k) {
    T x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = kAdjX ? conj_(X[k][r0 + 16 * a]) : X[r0 + 16 * a][k];
#pragma unroll
    for (int b = 0; b < 4; ++b) y[b] = Y[k][c0 + b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = mac_(acc[a][b], x[a], y[b]);
  }
}
template <typename T>
__device__ inline void tile_store(T (*X)[kLd], const T (&acc)[4][4]) {
  const int r0 = threadIdx.x % 16, c0 = 4 * (threadIdx.x / 16);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) X[r0 + 16 * a][c0 + b] = acc[a][b];
}

// One launch per round: jobs [0, k(k-1)/2) are the tiles (t, u), t < u, of H:
// H_tu <- Q_t^H H_tu Q_u and H_ut = H_tu^H; jobs past them are V's tiles (64 rows,
// pair u): V <- V Q_u. Tiles whose Q are both the identity are skipped.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    update_kernel(int np, T *H, T *V, int round, const T *Qbuf, const int *flags) {
  __shared__ T X[kP][kLd];
  __shared__ T Y[kP][kLd];
  const int np2 = np / kB, k = np2 / 2;
  const int nA = k * (k - 1) / 2;
  const long ld = np;
  int job = blockIdx.x;
  T acc[4][4];
  if (job < nA) {
    int t = 0, rem = job;
    while (rem >= k - 1 - t) {
      rem -= k - 1 - t;
      ++t;
    }
    const int u = t + 1 + rem;
    const bool ft = flags[t] != 0, fu = flags[u] != 0;
    if (!ft && !fu) return;
    int pt, qt, pu, qu;
    pair_of(round, t, np2, pt, qt);
    pair_of(round, u, np2, pu, qu);
    for (int e = threadIdx.x; e < kP * kP; e += kThreads) {
      const int r = e % kP, c = e / kP;
      X[r][c] = H[glob(r, pt, qt) + glob(c, pu, qu) * ld];
    }
    if (fu) {
      const T *Qu = Qbuf + static_cast<long>(u) * kP * kP;
      for (int e = threadIdx.x; e < kP * kP; e += kThreads) Y[e % kP][e / kP] = Qu[e];
      __syncthreads();
      tile_product<T, false>(X, Y, acc);
      __syncthreads();
      tile_store<T>(X, acc);
    }
    if (ft) {
      const T *Qt = Qbuf + static_cast<long>(t) * kP * kP;
      for (int e = threadIdx.x; e < kP * kP; e += kThreads) Y[e % kP][e / kP] = Qt[e];
      __syncthreads();
      tile_product<T, true>(Y, X, acc);
      __syncthreads();
      tile_store<T>(X, acc);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kP * kP; e += kThreads) {
      const int r = e % kP, c = e / kP;
      H[glob(r, pt, qt) + glob(c, pu, qu) * ld] = X[r][c];
    }
    for (int e = threadIdx.x; e < kP * kP; e += kThreads) {  // the adjoint, rows of it coalesced
      const int c = e % kP, r = e / kP;
      H[glob(c, pu, qu) + glob(r, pt, qt) * ld] = conj_(X[r][c]);
    }
    return;
  }
  job -= nA;
  const int u = job % k, rb = job / k;
  if (flags[u] == 0) return;
  int pu, qu;
  pair_of(round, u, np2, pu, qu);
  const T *Qu = Qbuf + static_cast<long>(u) * kP * kP;
  for (int e = threadIdx.x; e < kP * kP; e += kThreads) {
    const int r = e % kP, c = e / kP;
    X[r][c] = V[rb * kP + r + glob(c, pu, qu) * ld];
    Y[r][c] = Qu[e];
  }
  __syncthreads();
  tile_product<T, false>(X, Y, acc);
  __syncthreads();
  tile_store<T>(X, acc);
  __syncthreads();
  for (int e = threadIdx.x; e < kP * kP; e += kThreads) {
    const int r = e % kP, c = e / kP;
    V[rb * kP + r + glob(c, pu, qu) * ld] = X[r][c];
  }
}

// w = the real diagonal of H; M = V diag(exp(w)) over rows < n (leading dimension n)
template <typename T>
__global__ void __launch_bounds__(kThreads)
    exp_scale_kernel(int n, int np, const T *H, const T *V, T *M) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const real_t<T> w = re_(H[j + static_cast<long>(j) * np]);
  M[i + static_cast<long>(j) * n] = ops<T>::scale(V[i + static_cast<long>(j) * np], exp_(w));
}
// W = V(0:n, :)^H (np x n, leading dimension np)
template <typename T>
__global__ void __launch_bounds__(kThreads) adjoint_rows_kernel(int n, int np, const T *V, T *W) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  const int j = blockIdx.y;
  if (k < np && j < n) W[k + static_cast<long>(j) * np] = conj_(V[j + static_cast<long>(k) * np]);
}

__global__ void set_info_kernel(int *info, int v) {
  if (threadIdx.x == 0) *info = v;
}

template <typename T>
struct HermWs {
  T *H = nullptr, *V = nullptr, *W = nullptr, *Q = nullptr;
  int *flags = nullptr;
  unsigned *count = nullptr;
  real_t<T> *delta = nullptr, *colsq = nullptr;
  size_t carve(void *base, int n) {
    const size_t np = static_cast<size_t>((n + kP - 1) / kP) * kP;
    const size_t k = np / kP;
    Layout L(base);
    H = L.take<T>(np * np);
    V = L.take<T>(np * np);
    W = L.take<T>(np * np);
    Q = L.take<T>(k * kP * kP);
    flags = L.take<int>(k);
    count = L.take<unsigned>(1);
    delta = L.take<real_t<T>>(1);
    colsq = L.take<real_t<T>>(np);
    return L.total();
  }
};

template <typename T>
clm_status_t herm_bytes(int n, size_t *bytes) {
  if (bytes == nullptr || n < 1 || n > kMaxDim) return CLM_STATUS_INVALID_VALUE;
  HermWs<work_t<T>> ws;
  *bytes = ws.carve(nullptr, n);
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t expm_herm_impl(clm_ctx *ctx, hipStream_t st, int uplo, int n, const T *A, int lda,
                            T *E, int lde, void *work, size_t bytes, int *info) {
  using W = work_t<T>;
  using R = real_t<W>;
  if (n < 1 || n > kMaxDim || lda < n || lde < n || A == nullptr || E == nullptr ||
      work == nullptr || (uplo != CLM_FILL_MODE_LOWER && uplo != CLM_FILL_MODE_UPPER))
    return CLM_STATUS_INVALID_VALUE;
  HermWs<W> ws;
  if (bytes < ws.carve(nullptr, n)) return CLM_STATUS_ALLOC_FAILED;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  ws.carve(work, n);
  const int np = (n + kP - 1) / kP * kP;
  const int np2 = np / kB, k = np2 / 2;
  const dim3 grid_sq((np + kThreads - 1) / kThreads, np);
  mirror_kernel<T, W><<<grid_sq, kThreads, 0, st>>>(n, np, uplo == CLM_FILL_MODE_UPPER, A, lda,
                                                    ws.H);
  identity_kernel<W><<<grid_sq, kThreads, 0, st>>>(np, ws.V);
  colsq_kernel<W><<<np, kThreads, 0, st>>>(np, ws.H, ws.colsq);
  // the unit roundoff of the working precision, 2^-53
  delta_kernel<W><<<1, kThreads, 0, st>>>(np, ws.colsq, R(1.1102230246251565e-16), ws.delta);
  CLM_TRY_HIP(hipGetLastError());
  bool converged = false;
  const int jobs = k * (k - 1) / 2 + k * k;
  for (int sweep = 0; sweep < kMaxSweeps && !converged; ++sweep) {
    clear_count_kernel<<<1, 64, 0, st>>>(ws.count);
    for (int r = 0; r < np2 - 1; ++r) {
      solve_kernel<W><<<k, kSolveThreads, 0, st>>>(np, ws.H, ws.delta, r, ws.Q, ws.flags, ws.count);
      update_kernel<W><<<jobs, kThreads, 0, st>>>(np, ws.H, ws.V, r, ws.Q, ws.flags);
    }
    CLM_TRY_HIP(hipGetLastError());
    unsigned rotated = 0;
    CLM_TRY_HIP(hipMemcpyAsync(&rotated, ws.count, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    CLM_TRY_HIP(hipStreamSynchronize(st));
    converged = rotated == 0;
  }
  if (info) {
    set_info_kernel<<<1, 64, 0, st>>>(info, converged ? 0 : 1);
    CLM_TRY_HIP(hipGetLastError());
  }
  // M = V diag(exp(w)) (rows < n) over W's storage first, then V(0:n,:)^H over H's (the
  // eigenvalues have been read by then), then exp(A) = M V^H: into E directly, or, for
  // the single-precision types, into V's storage and rounded into E
  W *M = ws.W;
  exp_scale_kernel<W><<<dim3((n + kThreads - 1) / kThreads, np), kThreads, 0, st>>>(n, np, ws.H,
                                                                                   ws.V, M);
  CLM_TRY_HIP(hipGetLastError());
  W *Vh = ws.H;
  adjoint_rows_kernel<W><<<dim3((np + kThreads - 1) / kThreads, n), kThreads, 0, st>>>(n, np, ws.V,
                                                                                       Vh);
  CLM_TRY_HIP(hipGetLastError());
  constexpr bool kSame = std::is_same_v<T, W>;
  GemmArgs<W> g;
  g.M = n;
  g.N = n;
  g.K = np;
  g.A = M;
  g.lda = n;
  g.B = Vh;
  g.ldb = np;
  if constexpr (kSame) {
    g.C = E;
    g.ldc = lde;
  } else {
    g.C = ws.V;
    g.ldc = n;
  }
  g.alpha = ops<W>::from_real(1.0);
  g.beta = ops<W>::zero();
  CLM_TRY_HIP(launch_gemm<W>(st, g));
  if constexpr (!kSame) {
    narrow_kernel<T, W><<<dim3((n + kThreads - 1) / kThreads, n), kThreads, 0, st>>>(n, ws.V, E,
                                                                                     lde);
    CLM_TRY_HIP(hipGetLastError());
  }
  return CLM_STATUS_SUCCESS;
}

}  // namespace
}  // namespace clm

using clm::zc;
using clm::zd;

extern "C" {

clm_status_t clm_expm_herm_buffer_size_s(int n, size_t *b) { return clm::herm_bytes<float>(n, b); }
clm_status_t clm_expm_herm_buffer_size_d(int n, size_t *b) { return clm::herm_bytes<double>(n, b); }
clm_status_t clm_expm_herm_buffer_size_c(int n, size_t *b) { return clm::herm_bytes<zc>(n, b); }
clm_status_t clm_expm_herm_buffer_size_z(int n, size_t *b) { return clm::herm_bytes<zd>(n, b); }

clm_status_t clm_expm_herm_s(clm_ctx *ctx, clm_stream_t stream, int uplo, int n, const float *dA,
                             int lda, float *dE, int lde, void *dWork, size_t lwork_bytes,
                             int *dInfo) {
  return clm::expm_herm_impl<float>(ctx, static_cast<hipStream_t>(stream), uplo, n, dA, lda, dE,
                                    lde, dWork, lwork_bytes, dInfo);
}
clm_status_t clm_expm_herm_d(clm_ctx *ctx, clm_stream_t stream, int uplo, int n, const double *dA,
                             int lda, double *dE, int lde, void *dWork, size_t lwork_bytes,
                             int *dInfo) {
  return clm::expm_herm_impl<double>(ctx, static_cast<hipStream_t>(stream), uplo, n, dA, lda, dE,
                                     lde, dWork, lwork_bytes, dInfo);
}
clm_status_t clm_expm_herm_c(clm_ctx *ctx, clm_stream_t stream, int uplo, int n,
                             const clm_float_complex *dA, int lda, clm_float_complex *dE, int lde,
                             void *dWork, size_t lwork_bytes, int *dInfo) {
  return clm::expm_herm_impl<zc>(ctx, static_cast<hipStream_t>(stream), uplo, n,
                                 reinterpret_cast<const zc *>(dA), lda, reinterpret_cast<zc *>(dE),
                                 lde, dWork, lwork_bytes, dInfo);
}
clm_status_t clm_expm_herm_z(clm_ctx *ctx, clm_stream_t stream, int uplo, int n,
                             const clm_double_complex *dA, int lda, clm_double_complex *dE, int lde,
                             void *dWork, size_t lwork_bytes, int *dInfo) {
  return clm::expm_herm_impl<zd>(ctx, static_cast<hipStream_t>(stream), uplo, n,
                                 reinterpret_cast<const zd *>(dA), lda, reinterpret_cast<zd *>(dE),
                                 lde, dWork, lwork_bytes, dInfo);
}

}  // extern "C"
