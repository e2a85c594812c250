// cg_unitary.hip -- Riemannian conjugate gradient on U(n), the caller of the
// matrix exponential (src/cg_unitary/): Abrudan, Eriksson & Koivunen, Signal
// Processing 89 (2009) 1704-1714, Tables 1-3.
//
// One iteration: Psi = dJ/dW(W), G = Psi W^H - W Psi^H; a geodesic line search
// along exp(sgn mu H) W; W <- exp(sgn mu H) W; the Polak-Ribiere update of H.
// The products run on the library's MFMA GEMMs and its expm; everything else is
// O(n^2) or scalar and stays on the device:
//
//   * Frobenius inner products and norms: a two-stage deterministic reduction
//     (fixed block count and order, FP64 sums for every element type);
//   * the power iteration for |omega_max| (geodesic_search.cppm:232-322): every
//     step's gemv, norms and convergence test are kernels that read a device
//     flag and return at once after convergence, enqueued a few steps at a time,
//     so the host reads one flag per chunk instead of two norms per step;
//   * the scalar stages of the searches (cg_unitary.cu): one 256-lane workgroup
//     each, the root brackets evaluated on all lanes -- the 1024-point grid of the
//     Table 1 polynomial and the 2048-point grid of the Table 2 trig polynomial,
//     31 terms a point -- and the brackets bisected one lane each;
//   * the Polak-Ribiere factor, the direction update and the downhill reset
//     (cg_solver.cppm:356-428) as device scalars: one read-back per iteration
//     carries <G, G> and the reset flag to the host's control flow.
//
// The sampling sweep advances the trial points as W_i = R_1 W_{i-1} with
// R_1 = exp(step H): the same points as the reference's R_i = R_{i-1} R_1,
// W_i = R_i W (geodesic_search.cppm:382-446) -- exp(i step H) W both ways -- in
// one product a sample instead of two.
#include "clm_internal.hpp"
#include "calaman_cg.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>

namespace clm {
namespace {

constexpr int kMaxSamples = CLM_CG_MAX_SAMPLES;
constexpr int kBlk = 256;
constexpr int kMaxPartials = 1024;
constexpr int kPowerMaxIter = 100;
constexpr int kPowerChunk = 6;
constexpr int kGridPoly = 1024;  // Table 1 bracket grid (cg_unitary.cu:172-220)
constexpr int kGridTrig = 2048;  // Table 2 bracket grid (cg_unitary.cu:256-320)
constexpr double kTwoPi = 6.283185307179586476925286766559;

// The device scalars of one solve. Fields read back together sit together.
struct CgScalars {
  double dots[kMaxSamples];   // derivative samples Re dotc(H W_i, Psi(W_i))
  double costs[kMaxSamples];  // sampled cost values (Table 2)
  double coeffs[8];           // Table 1 polynomial, ascending powers
  double dft_re[kMaxSamples], dft_im[kMaxSamples];
  double args[kMaxSamples];   // Table 2 root arguments, ascending
  double mu, found, num_args, pad0;
  double omega2, power_done, power_iters, power_update;
  double power_inv, power_rho, power_nz2, pad1;
  double gg, reset, gg_next, numer;
  double hg, gamma, pad2, pad3;
};

// ---- element helpers --------------------------------------------------------------------
__host__ __device__ inline double conj_of(double x) { return x; }
__host__ __device__ inline float conj_of(float x) { return x; }
__host__ __device__ inline zd conj_of(zd x) { return zd{x.re, -x.im}; }
__host__ __device__ inline zc conj_of(zc x) { return zc{x.re, -x.im}; }
__device__ inline double re_dot(double a, double b) { return a * b; }
__device__ inline double re_dot(float a, float b) { return double(a) * double(b); }
__device__ inline double re_dot(zd a, zd b) { return a.re * b.re + a.im * b.im; }
__device__ inline double re_dot(zc a, zc b) { return double(a.re) * b.re + double(a.im) * b.im; }
__device__ inline double sub(double a, double b) { return a - b; }
__device__ inline float sub(float a, float b) { return a - b; }
__device__ inline zd sub(zd a, zd b) { return zd{a.re - b.re, a.im - b.im}; }
__device__ inline zc sub(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }
__device__ inline double re_of(double a) { return a; }
__device__ inline double re_of(float a) { return a; }
__device__ inline double re_of(zd a) { return a.re; }
__device__ inline double re_of(zc a) { return a.re; }
__device__ inline void store_real(double *p, double v) { *p = v; }
__device__ inline void store_real(float *p, double v) { *p = static_cast<float>(v); }
__device__ inline void store_real(zd *p, double v) { *p = zd{v, 0.0}; }
__device__ inline void store_real(zc *p, double v) { *p = zc{static_cast<float>(v), 0.f}; }

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ inline float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ inline zd wave_sum(zd v) { return zd{wave_sum(v.re), wave_sum(v.im)}; }
__device__ inline zc wave_sum(zc v) { return zc{wave_sum(v.re), wave_sum(v.im)}; }

__device__ inline bool gated(const double *gate) { return gate != nullptr && *gate != 0.0; }

// ---- Frobenius reductions ------------------------------------------------------------------
// partials[b] = sum over block b's elements of Re conj(x - w) y (w may be null)
template <typename T>
__global__ void __launch_bounds__(kBlk)
    dot_partial_kernel(long count, const T *__restrict__ x, const T *__restrict__ w,
                       const T *__restrict__ y, double *__restrict__ partials, const double *gate) {
  if (gated(gate)) return;
  __shared__ double red[kBlk / 64];
  double acc = 0.0;
  const long stride = static_cast<long>(gridDim.x) * kBlk;
  for (long i = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x; i < count; i += stride) {
    const T a = w ? sub(x[i], w[i]) : x[i];
    acc += re_dot(a, y[i]);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// *out = scale * sum(partials[0 .. np-1]), in a fixed order
template <typename O>
__global__ void __launch_bounds__(kBlk)
    dot_final_kernel(int np, const double *__restrict__ partials, double scale, O *out,
                     const double *gate) {
  if (gated(gate)) return;
  __shared__ double red[kBlk];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kBlk) acc += partials[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlk / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) store_real(out, scale * red[0]);
}

template <typename T, typename O>
hipError_t launch_dot(hipStream_t st, long count, const T *x, const T *w, const T *y,
                      double *partials, int max_partials, double scale, O *out,
                      const double *gate = nullptr) {
  long blocks = (count + 4L * kBlk - 1) / (4L * kBlk);
  blocks = std::max(1L, std::min<long>(blocks, max_partials));
  dot_partial_kernel<T><<<static_cast<int>(blocks), kBlk, 0, st>>>(count, x, w, y, partials, gate);
  dot_final_kernel<O><<<1, kBlk, 0, st>>>(static_cast<int>(blocks), partials, scale, out, gate);
  return hipGetLastError();
}

// ---- O(n^2) matrix kernels -------------------------------------------------------------------
// Y = X^H through a 32 x 33 LDS tile (coalesced reads and writes)
template <typename T>
__global__ void __launch_bounds__(kBlk)
    adjoint_kernel(int n, const T *__restrict__ X, long ldx, T *__restrict__ Y, long ldy) {
  __shared__ T tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int r = bx + tx, c = by + j;
    if (r < n && c < n) tile[j][tx] = X[r + static_cast<long>(c) * ldx];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = by + tx, c = bx + j;
    if (r < n && c < n) Y[r + static_cast<long>(c) * ldy] = conj_of(tile[tx][j]);
  }
}

// G = X - X^H (the projection onto u(n), cg_solver.cppm:221-222)
template <typename T>
__global__ void __launch_bounds__(kBlk)
    skew_kernel(int n, const T *__restrict__ X, long ldx, T *__restrict__ G, long ldg) {
  __shared__ T tb[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {  // tb[j][tx] = X(by + tx, bx + j)
    const int r = by + tx, c = bx + j;
    if (r < n && c < n) tb[j][tx] = X[r + static_cast<long>(c) * ldx];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = bx + tx, c = by + j;
    if (r < n && c < n)
      G[r + static_cast<long>(c) * ldg] = sub(X[r + static_cast<long>(c) * ldx], conj_of(tb[tx][j]));
  }
}

// Y(i, j) = X(i, j) * d[j] (dgmm from the right, brockett.cppm:96-98)
template <typename T>
__global__ void __launch_bounds__(kBlk)
    col_scale_kernel(int n, const T *__restrict__ X, long ldx, const T *__restrict__ d,
                     T *__restrict__ Y, long ldy) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) Y[i + static_cast<long>(j) * ldy] = ops<T>::mul(X[i + static_cast<long>(j) * ldx], d[j]);
}

// Y = 3/2 I - X / 2 (one Newton-Schulz step's right factor)
template <typename T>
__global__ void __launch_bounds__(kBlk) polar_factor_kernel(int n, T *X, long ldx) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  T v = ops<T>::scale(X[i + static_cast<long>(j) * ldx], -0.5);
  if (i == j) v = ops<T>::add(v, ops<T>::from_real(1.5));
  X[i + static_cast<long>(j) * ldx] = v;
}

// D = G + gamma D
template <typename T>
__global__ void __launch_bounds__(kBlk)
    direction_kernel(long count, const T *__restrict__ G, T *__restrict__ D, const CgScalars *sc) {
  const double gamma = sc->gamma;
  const long stride = static_cast<long>(gridDim.x) * kBlk;
  for (long i = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x; i < count; i += stride)
    D[i] = ops<T>::add(G[i], ops<T>::scale(D[i], gamma));
}

// Remark 2: D = G when <D, G> <= 0; the flag is read back by the host
template <typename T>
__global__ void __launch_bounds__(kBlk)
    reset_kernel(long count, const T *__restrict__ G, T *__restrict__ D, CgScalars *sc) {
  const bool reset = !(sc->hg > 0.0);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->reset = reset ? 1.0 : 0.0;
  if (!reset) return;
  const long stride = static_cast<long>(gridDim.x) * kBlk;
  for (long i = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x; i < count; i += stride)
    D[i] = G[i];
}

// the Polak-Ribiere factor gamma = <G' - G, G'> / <G, G> (eq. 10), then <G, G> <- <G', G'>
__global__ void polak_ribiere_kernel(CgScalars *sc) {
  if (threadIdx.x != 0) return;
  const double gg = sc->gg;
  sc->gamma = gg > 0.0 ? sc->numer / (2.0 * gg) : 0.0;
  sc->gg = sc->gg_next;
}

__global__ void init_scalars_kernel(CgScalars *sc) {
  if (threadIdx.x != 0) return;
  sc->reset = 0.0;
  sc->gamma = 0.0;
}

template <typename T>
__global__ void real_parts_kernel(int count, const T *__restrict__ v, double *out) {
  const int t = threadIdx.x;
  if (t < count) out[t] = re_of(v[t]);
}

// ---- power iteration (geodesic_search.cppm:232-322) -----------------------------------------
// v <- v / ||v||, seeded with ones when v is zero (or when `fresh`); resets the state
template <typename T>
__global__ void __launch_bounds__(kBlk) power_init_kernel(int n, T *v, CgScalars *sc, int fresh) {
  __shared__ double red[kBlk];
  __shared__ double nv2;
  if (fresh)
    for (int i = threadIdx.x; i < n; i += kBlk) v[i] = ops<T>::from_real(1.0);
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlk) acc += re_dot(v[i], v[i]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlk / 2; s > 0; s >>= 1) {
      if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) nv2 = red[0];
    __syncthreads();
    if (nv2 > 0.0 || pass == 1) break;
    for (int i = threadIdx.x; i < n; i += kBlk) v[i] = ops<T>::from_real(1.0);
    __syncthreads();
  }
  const double inv = 1.0 / sqrt(nv2);
  for (int i = threadIdx.x; i < n; i += kBlk) v[i] = ops<T>::scale(v[i], inv);
  if (threadIdx.x == 0) {
    sc->omega2 = 0.0;
    sc->power_done = 0.0;
    sc->power_iters = 0.0;
    sc->power_update = -1.0;
  }
}

// y = H v: 64 rows a workgroup, four column phases summed through LDS
template <typename T>
__global__ void __launch_bounds__(kBlk)
    gemv_n_kernel(int n, const T *__restrict__ H, long ldh, const T *__restrict__ v,
                  T *__restrict__ y, const double *gate) {
  if (gated(gate)) return;
  __shared__ T part[4][64];
  const int l = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + l;
  T acc = ops<T>::zero();
  if (r < n)
    for (int j = ph; j < n; j += 4)
      acc = ops<T>::add(acc, ops<T>::mul(H[r + static_cast<long>(j) * ldh], v[j]));
  part[ph][l] = acc;
  __syncthreads();
  if (ph == 0 && r < n)
    y[r] = ops<T>::add(ops<T>::add(part[0][l], part[1][l]), ops<T>::add(part[2][l], part[3][l]));
}

// z = H^H y: one wave a column
template <typename T>
__global__ void __launch_bounds__(kBlk)
    gemv_c_kernel(int n, const T *__restrict__ H, long ldh, const T *__restrict__ y,
                  T *__restrict__ z, const double *gate) {
  if (gated(gate)) return;
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (col >= n) return;
  T acc = ops<T>::zero();
  const T *hc = H + static_cast<long>(col) * ldh;
  for (int i = lane; i < n; i += 64) acc = ops<T>::add(acc, ops<T>::mul(conj_of(hc[i]), y[i]));
  acc = wave_sum(acc);
  if (lane == 0) z[col] = acc;
}

// one step's scalar logic: rho = ||H v||^2, convergence at |rho - omega2| <= 1e-4 rho
__global__ void power_step_kernel(CgScalars *sc, int it) {
  if (threadIdx.x != 0 || sc->power_done != 0.0) return;
  const double rho = sc->power_rho;
  const double nz = sqrt(sc->power_nz2);
  const double omega2 = sc->omega2;
  const bool converged = omega2 > 0.0 && fabs(rho - omega2) <= 1e-4 * rho;
  sc->omega2 = rho;
  sc->power_iters = it + 1;
  if (!(nz > 0.0)) {
    sc->power_done = 1.0;
    return;
  }
  sc->power_inv = 1.0 / nz;
  sc->power_update = it;
  if (converged || it + 1 >= kPowerMaxIter) sc->power_done = 1.0;
}

// v = z / ||z|| when step `it` asked for it
template <typename T>
__global__ void __launch_bounds__(kBlk)
    power_update_kernel(int n, T *__restrict__ v, const T *__restrict__ z, const CgScalars *sc,
                        int it) {
  if (sc->power_update != static_cast<double>(it)) return;
  const double inv = sc->power_inv;
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) v[i] = ops<T>::scale(z[i], inv);
}

// ---- Table 1 (cg_unitary.cu:90-220) ------------------------------------------------------------
__device__ inline double poly_eval(const double *c, int order, double x) {
  double acc = c[order];
  for (int j = order - 1; j >= 0; --j) acc = acc * x + c[j];
  return acc;
}
__device__ inline bool straddles(double a, double b) {
  return (a < 0.0 && b >= 0.0) || (a > 0.0 && b <= 0.0);
}

// The degree-`order` polynomial through scale * dots[i] at mu = i h (the integer
// Vandermonde system M[i-1][j-1] = i^j by Gaussian elimination with partial
// pivoting, unscaled by h^j), then its first sign change over (0, upper] on the
// 1024-point grid -- every lane evaluating four points -- refined by 60 bisections.
__global__ void __launch_bounds__(kBlk)
    table1_kernel(CgScalars *sc, int order, double h, double scale, double upper) {
  __shared__ double c[8];
  __shared__ double f[kGridPoly + 1];
  __shared__ int first;
  if (threadIdx.x == 0) {
    first = kGridPoly + 1;
    double jp[8];
    for (int i = 0; i <= order; ++i) jp[i] = scale * sc->dots[i];
    for (int j = 0; j < 8; ++j) c[j] = 0.0;
    c[0] = jp[0];
    bool singular = false;
    if (order >= 1) {
      double m[6][7];
      for (int i = 1; i <= order; ++i) {
        double p = 1.0;
        for (int j = 1; j <= order; ++j) {
          p *= static_cast<double>(i);
          m[i - 1][j - 1] = p;
        }
        m[i - 1][order] = jp[i] - jp[0];
      }
      for (int col = 0; col < order && !singular; ++col) {
        int piv = col;
        double best = fabs(m[col][col]);
        for (int r = col + 1; r < order; ++r)
          if (fabs(m[r][col]) > best) {
            best = fabs(m[r][col]);
            piv = r;
          }
        if (!(best > 0.0)) {
          singular = true;
          break;
        }
        if (piv != col)
          for (int j = col; j <= order; ++j) {
            const double t = m[col][j];
            m[col][j] = m[piv][j];
            m[piv][j] = t;
          }
        for (int r = col + 1; r < order; ++r) {
          const double fct = m[r][col] / m[col][col];
          for (int j = col; j <= order; ++j) m[r][j] -= fct * m[col][j];
        }
      }
      if (!singular) {
        double b[6];
        for (int i = order - 1; i >= 0; --i) {
          double acc = m[i][order];
          for (int j = i + 1; j < order; ++j) acc -= m[i][j] * b[j];
          b[i] = acc / m[i][i];
        }
        double hp = 1.0;
        for (int j = 1; j <= order; ++j) {
          hp *= h;
          c[j] = hp > 0.0 ? b[j - 1] / hp : 0.0;
        }
      }
    }
    for (int j = 0; j < 8; ++j) sc->coeffs[j] = c[j];
  }
  __syncthreads();
  const bool usable = upper > 0.0 && order >= 1;
  const double step = upper / kGridPoly;
  if (usable) {
    for (int i = threadIdx.x; i <= kGridPoly; i += kBlk)
      f[i] = poly_eval(c, order, i == kGridPoly ? upper : static_cast<double>(i) * step);
    __syncthreads();
    for (int i = threadIdx.x + 1; i <= kGridPoly; i += kBlk)
      if (straddles(f[i - 1], f[i])) atomicMin(&first, i);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double mu = 0.0;
  int found = 0;
  if (usable && first <= kGridPoly) {
    double lo = static_cast<double>(first - 1) * step;
    double hi = first == kGridPoly ? upper : static_cast<double>(first) * step;
    double flo = f[first - 1];
    for (int it = 0; it < 60; ++it) {
      const double mid = 0.5 * (lo + hi);
      const double fm = poly_eval(c, order, mid);
      if ((flo < 0.0 && fm < 0.0) || (flo > 0.0 && fm > 0.0)) {
        lo = mid;
        flo = fm;
      } else {
        hi = mid;
      }
    }
    const double root = 0.5 * (lo + hi);
    if (root > 0.0) {
      mu = root;
      found = 1;
    }
  }
  sc->mu = mu;
  sc->found = found;
}

// ---- Table 2 (cg_unitary.cu:222-370) -------------------------------------------------------------
// The Hann-windowed centred DFT of scale * dots (window 1/2 - 1/2 cos(2 pi (i+1) / (N+1))),
// then the zero crossings over (0, 2 pi) of D(theta) = sum_k c_k e^{ik theta}: the
// 2049 grid values on all lanes, the first 32 brackets in ascending order, each
// refined by 50 bisections on a lane of its own.
__global__ void __launch_bounds__(kBlk) dft_kernel(CgScalars *sc, int num, double scale) {
  __shared__ double win[kMaxSamples];
  __shared__ double cre[kMaxSamples], cim[kMaxSamples];
  __shared__ double f[kGridTrig + 1];
  __shared__ int bracket[kMaxSamples];
  __shared__ int count;
  const int t = threadIdx.x;
  const int half = (num - 1) / 2;
  if (t < num) {
    const double w = 0.5 - 0.5 * cos(kTwoPi * static_cast<double>(t + 1) / static_cast<double>(num + 1));
    win[t] = w * (scale * sc->dots[t]);
  }
  __syncthreads();
  if (t < num) {
    const int k = t - half;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < num; ++i) {
      const double ang = -kTwoPi * static_cast<double>(i) * static_cast<double>(k) / static_cast<double>(num);
      re += win[i] * cos(ang);
      im += win[i] * sin(ang);
    }
    cre[t] = re;
    cim[t] = im;
    sc->dft_re[t] = re;
    sc->dft_im[t] = im;
  }
  __syncthreads();
  const double step = kTwoPi / kGridTrig;
  auto D = [&](double theta) {
    double acc = 0.0;
    for (int j = 0; j < num; ++j) {
      const double k = static_cast<double>(j - half);
      acc += cre[j] * cos(k * theta) - cim[j] * sin(k * theta);
    }
    return acc;
  };
  for (int i = t; i <= kGridTrig; i += kBlk) f[i] = D(static_cast<double>(i) * step);
  __syncthreads();
  if (t == 0) {
    int c = 0;
    for (int i = 1; i <= kGridTrig && c < kMaxSamples; ++i)
      if (straddles(f[i - 1], f[i])) bracket[c++] = i;
    count = c;
  }
  __syncthreads();
  if (t < count) {
    const int i = bracket[t];
    double lo = static_cast<double>(i - 1) * step, hi = static_cast<double>(i) * step;
    double flo = f[i - 1];
    for (int it = 0; it < 50; ++it) {
      const double mid = 0.5 * (lo + hi);
      const double fm = D(mid);
      if ((flo < 0.0) == (fm < 0.0)) {
        lo = mid;
        flo = fm;
      } else {
        hi = mid;
      }
    }
    sc->args[t] = 0.5 * (lo + hi);
  }
  if (t == 0) sc->num_args = count;
}

// steps 12-13: the even-indexed crossing (a minimum of the descent) nearest the
// best sampled cost's mu
__global__ void select_dft_kernel(CgScalars *sc, int num, double t_dft, int maximizing) {
  if (threadIdx.x != 0) return;
  sc->mu = 0.0;
  const int nargs = static_cast<int>(sc->num_args);
  if (nargs < 1 || num < 1) return;
  int best_i = 0;
  double best = sc->costs[0];
  for (int i = 1; i < num; ++i) {
    const bool better = maximizing ? (sc->costs[i] > best) : (sc->costs[i] < best);
    if (better) {
      best = sc->costs[i];
      best_i = i;
    }
  }
  const double mu_star = static_cast<double>(best_i) * t_dft / static_cast<double>(num);
  double chosen = 0.0, closest = -1.0;
  for (int l = 0; l < nargs; l += 2) {
    const double mu = sc->args[l] * t_dft / kTwoPi;
    if (!(mu > 0.0)) continue;
    const double d = fabs(mu - mu_star);
    if (closest < 0.0 || d < closest) {
      closest = d;
      chosen = mu;
    }
  }
  sc->mu = chosen;
}

// ---- element-type bindings -------------------------------------------------------------------
template <typename T>
struct Api;
template <>
struct Api<float> {
  static clm_status_t expm(clm_ctx *c, hipStream_t st, int n, const float *a, float *e, void *w,
                           size_t b) {
    return clm_expm_s(c, st, n, a, n, e, n, w, b, nullptr, nullptr);
  }
  static clm_status_t expm_bytes(int n, size_t *b) { return clm_expm_buffer_size_s(n, b); }
  static constexpr double tol = 1e-10;
};
template <>
struct Api<double> {
  static clm_status_t expm(clm_ctx *c, hipStream_t st, int n, const double *a, double *e, void *w,
                           size_t b) {
    return clm_expm_d(c, st, n, a, n, e, n, w, b, nullptr, nullptr);
  }
  static clm_status_t expm_bytes(int n, size_t *b) { return clm_expm_buffer_size_d(n, b); }
  static constexpr double tol = 1e-20;
};
template <>
struct Api<zc> {
  static clm_status_t expm(clm_ctx *c, hipStream_t st, int n, const zc *a, zc *e, void *w,
                           size_t b) {
    return clm_expm_c(c, st, n, reinterpret_cast<const clm_float_complex *>(a), n,
                      reinterpret_cast<clm_float_complex *>(e), n, w, b, nullptr, nullptr);
  }
  static clm_status_t expm_bytes(int n, size_t *b) { return clm_expm_buffer_size_c(n, b); }
  static constexpr double tol = 1e-10;
};
template <>
struct Api<zd> {
  static clm_status_t expm(clm_ctx *c, hipStream_t st, int n, const zd *a, zd *e, void *w,
                           size_t b) {
    return clm_expm_z(c, st, n, reinterpret_cast<const clm_double_complex *>(a), n,
                      reinterpret_cast<clm_double_complex *>(e), n, w, b, nullptr, nullptr);
  }
  static clm_status_t expm_bytes(int n, size_t *b) { return clm_expm_buffer_size_z(n, b); }
  static constexpr double tol = 1e-20;
};

template <typename T>
T one_of() {
  return ops<T>::from_real(1.0);
}

template <typename T>
hipError_t gemm_nn(hipStream_t st, int n, const T *A, long lda, const T *B, long ldb, T *C,
                   long ldc) {
  GemmArgs<T> g;
  g.M = g.N = g.K = n;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.alpha = one_of<T>();
  g.beta = ops<T>::zero();
  return launch_gemm<T>(st, g);
}

dim3 tiles32(int n) { return dim3((n + 31) / 32, (n + 31) / 32); }
dim3 columns(int n) { return dim3((n + kBlk - 1) / kBlk, n); }
int flat_blocks(long count) {
  return static_cast<int>(std::max(1L, std::min<long>((count + kBlk - 1) / kBlk, 4096)));
}

// ---- workspace (the role of CgSlices, src/cg_unitary/buffer_size.cppm:38-118) ----------------
template <typename T>
struct CgWs {
  T *psi = nullptr, *grad = nullptr, *grad_next = nullptr, *dir = nullptr, *tmp = nullptr;
  T *w_new = nullptr, *w_prev = nullptr, *rot = nullptr, *psi_trial = nullptr, *wadj = nullptr;
  T *pv = nullptr, *py = nullptr, *pz = nullptr;
  T *cost_vals = nullptr;
  double *colsum = nullptr;
  double *partials = nullptr;
  CgScalars *sc = nullptr;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;

  size_t carve(void *base, int n, size_t expm_bytes, size_t cost_bytes) {
    Layout L(base);
    const size_t nn = static_cast<size_t>(n) * n;
    psi = L.take<T>(nn);
    grad = L.take<T>(nn);
    grad_next = L.take<T>(nn);
    dir = L.take<T>(nn);
    tmp = L.take<T>(nn);
    w_new = L.take<T>(nn);
    w_prev = L.take<T>(nn);
    rot = L.take<T>(nn);
    psi_trial = L.take<T>(nn);
    wadj = L.take<T>(nn);
    pv = L.take<T>(n);
    py = L.take<T>(n);
    pz = L.take<T>(n);
    cost_vals = L.take<T>(kMaxSamples);
    colsum = L.take<double>(n);
    partials = L.take<double>(kMaxPartials);
    sc = L.take<CgScalars>(1);
    scratch_bytes = std::max(expm_bytes, cost_bytes);
    scratch = L.take<unsigned char>(scratch_bytes);
    return L.total();
  }
};

template <typename T>
clm_status_t cg_buffer_size(int n, size_t cost_bytes, size_t *bytes) {
  if (n < 1 || n > kMaxDim || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  size_t eb = 0;
  CLM_TRY(Api<T>::expm_bytes(n, &eb));
  CgWs<T> ws;
  *bytes = ws.carve(nullptr, n, eb, cost_bytes);
  return CLM_STATUS_SUCCESS;
}

clm_cg_search_options default_search() { return clm_cg_search_options{5, 5, 10, 1}; }

bool valid_cost(const clm_cg_cost *cost) {
  return cost != nullptr && cost->order >= 1 && cost->gradient != nullptr &&
         cost->value != nullptr && cost->buffer_size != nullptr;
}

// ---- the solver's pieces ----------------------------------------------------------------------
template <typename T>
struct Cg {
  clm_ctx *ctx = nullptr;
  hipStream_t st = nullptr;
  int n = 0;
  const clm_cg_cost *cost = nullptr;
  CgWs<T> ws;
  long count() const { return static_cast<long>(n) * n; }

  clm_status_t read(const double *src, int k, double *out) {
    return read_device_values<double>(st, ctx->device, src, k, out);
  }
  clm_status_t dot(const T *x, const T *w, const T *y, double scale, double *out,
                   const double *gate = nullptr) {
    CLM_TRY_HIP((launch_dot<T, double>(st, count(), x, w, y, ws.partials, kMaxPartials, scale, out,
                                      gate)));
    return CLM_STATUS_SUCCESS;
  }
  clm_status_t copy(const T *src, T *dst) {
    CLM_TRY_HIP(launch_copy<T>(st, n, src, n, dst, n));
    return CLM_STATUS_SUCCESS;
  }
  clm_status_t grad(const T *W, T *Psi) {
    return cost->gradient(cost->user, st, n, W, n, Psi, n, ws.scratch, ws.scratch_bytes);
  }

  // |omega_max| of H: power iteration on H^H H, or ||H||_1 (geodesic_search.cppm:324-336)
  clm_status_t frequency(const T *H, int ldh, bool power, bool fresh, double *omega) {
    *omega = 0.0;
    if (!power) {
      double v = 0.0;
      CLM_TRY_HIP(launch_norm1<T>(st, n, H, ldh, reinterpret_cast<real_t<T> *>(ws.colsum)));
      CLM_TRY(read_device_values<real_t<T>>(st, ctx->device,
                                            reinterpret_cast<real_t<T> *>(ws.colsum), 1, &v));
      if (!std::isfinite(v)) return CLM_STATUS_EXECUTION_FAILED;
      *omega = v;
      return CLM_STATUS_SUCCESS;
    }
    power_init_kernel<T><<<1, kBlk, 0, st>>>(n, ws.pv, ws.sc, fresh ? 1 : 0);
    CLM_TRY_HIP(hipGetLastError());
    const double *gate = &ws.sc->power_done;
    const int rows = (n + 63) / 64, cols = (n + 3) / 4, flat = (n + kBlk - 1) / kBlk;
    double state[3] = {0, 0, 0};
    for (int it0 = 0; it0 < kPowerMaxIter; it0 += kPowerChunk) {
      for (int it = it0; it < std::min(it0 + kPowerChunk, kPowerMaxIter); ++it) {
        gemv_n_kernel<T><<<rows, kBlk, 0, st>>>(n, H, ldh, ws.pv, ws.py, gate);
        CLM_TRY_HIP((launch_dot<T, double>(st, n, ws.py, nullptr, ws.py, ws.partials, kMaxPartials,
                                          1.0, &ws.sc->power_rho, gate)));
        gemv_c_kernel<T><<<cols, kBlk, 0, st>>>(n, H, ldh, ws.py, ws.pz, gate);
        CLM_TRY_HIP((launch_dot<T, double>(st, n, ws.pz, nullptr, ws.pz, ws.partials, kMaxPartials,
                                          1.0, &ws.sc->power_nz2, gate)));
        power_step_kernel<<<1, 64, 0, st>>>(ws.sc, it);
        power_update_kernel<T><<<flat, kBlk, 0, st>>>(n, ws.pv, ws.pz, ws.sc, it);
      }
      CLM_TRY_HIP(hipGetLastError());
      CLM_TRY(read(&ws.sc->omega2, 3, state));
      if (state[1] != 0.0) break;
    }
    *omega = std::sqrt(state[0] > 0.0 ? state[0] : 0.0);
    return CLM_STATUS_SUCCESS;
  }

  // the sampling sweep (geodesic_search.cppm:338-446): dots[i] = Re dotc(H W_i, Psi(W_i))
  // at W_i = exp(i step H) W, sample 0 the caller's W and Psi
  clm_status_t sample(const T *W, const T *H, int ldh, const T *Psi, double step, int num,
                      bool want_cost) {
    CLM_TRY_HIP(launch_scale_copy<T>(st, n, step, H, ldh, ws.tmp, n));
    CLM_TRY(Api<T>::expm(ctx, st, n, ws.tmp, ws.rot, ws.scratch, ws.scratch_bytes));
    const T *wi = W;
    T *next = ws.w_new, *spare = ws.w_prev;
    for (int i = 0; i < num; ++i) {
      const T *pi = Psi;
      if (i > 0) {
        CLM_TRY_HIP(gemm_nn<T>(st, n, ws.rot, n, wi, n, next, n));
        wi = next;
        std::swap(next, spare);
        CLM_TRY(grad(wi, ws.psi_trial));
        pi = ws.psi_trial;
      }
      CLM_TRY_HIP(gemm_nn<T>(st, n, H, ldh, wi, n, ws.tmp, n));
      CLM_TRY(dot(ws.tmp, nullptr, pi, 1.0, &ws.sc->dots[i]));
      if (want_cost)
        CLM_TRY(cost->value(cost->user, st, n, wi, n, pi, n, ws.cost_vals + i, ws.scratch,
                            ws.scratch_bytes));
    }
    return CLM_STATUS_SUCCESS;
  }

  void fill_samples(clm_cg_search_info *info, int num) {
    if (!info || num < 1) return;
    double host[kMaxSamples] = {};
    if (hipMemcpyAsync(host, ws.sc->dots, sizeof(double) * num, hipMemcpyDeviceToHost, st) ==
            hipSuccess &&
        hipStreamSynchronize(st) == hipSuccess)
      for (int i = 0; i < num; ++i) info->samples[i] = host[i];
  }

  // Table 1 (geodesic_search.cppm:448-545)
  clm_status_t search_poly(const T *W, const T *H, int ldh, const T *Psi, int direction,
                           const clm_cg_search_options &o, bool fresh, double *mu,
                           clm_cg_search_info *info) {
    *mu = 0.0;
    if (info) *info = clm_cg_search_info{};
    const int P = o.poly_order;
    if (P < 1 || P > 5) return CLM_STATUS_INVALID_VALUE;
    double omega = 0.0;
    CLM_TRY(frequency(H, ldh, o.use_spectral_radius != 0, fresh, &omega));
    if (info) {
      info->omega_max = omega;
      info->method_used = CLM_CG_SEARCH_POLYNOMIAL;
    }
    if (!(omega > 0.0)) return CLM_STATUS_SUCCESS;
    const double t_mu = kTwoPi / (static_cast<double>(cost->order) * omega);
    const double h = t_mu / P;
    const double sg = direction == CLM_CG_MAXIMIZE ? 1.0 : -1.0;
    CLM_TRY(sample(W, H, ldh, Psi, sg * h, P + 1, false));
    table1_kernel<<<1, kBlk, 0, st>>>(ws.sc, P, h, 2.0 * sg, t_mu);
    CLM_TRY_HIP(hipGetLastError());
    double r[2] = {0, 0};
    CLM_TRY(read(&ws.sc->mu, 2, r));
    const bool usable = r[1] != 0.0 && r[0] > 0.0 && r[0] <= t_mu;
    if (info) {
      info->interval = t_mu;
      info->num_samples = P + 1;
      info->num_roots = static_cast<int>(r[1]);
      info->found = usable ? 1 : 0;
      fill_samples(info, P + 1);
    }
    *mu = usable ? r[0] : 0.0;
    return CLM_STATUS_SUCCESS;
  }

  // Table 2, with Table 1 as its fallback (geodesic_search.cppm:547-678)
  clm_status_t search_dft(const T *W, const T *H, int ldh, const T *Psi, int direction,
                          const clm_cg_search_options &o, bool fresh, double *mu,
                          clm_cg_search_info *info) {
    *mu = 0.0;
    if (info) *info = clm_cg_search_info{};
    if (o.dft_factor < 2 || o.dft_periods < 1) return CLM_STATUS_INVALID_VALUE;
    double omega = 0.0;
    CLM_TRY(frequency(H, ldh, o.use_spectral_radius != 0, fresh, &omega));
    if (info) {
      info->omega_max = omega;
      info->method_used = CLM_CG_SEARCH_DFT;
    }
    if (!(omega > 0.0)) return CLM_STATUS_SUCCESS;
    const double t_mu = kTwoPi / (static_cast<double>(cost->order) * omega);
    int num = 2 * ((o.dft_factor * o.dft_periods) / 2) + 1;
    if (num > kMaxSamples - 1) num = ((kMaxSamples - 1) / 2) * 2 + 1;
    if (num < 3) num = 3;
    const int periods = num / o.dft_factor;
    const double t_dft = static_cast<double>(periods > 0 ? periods : 1) * t_mu;
    const double h = t_dft / num;
    const double sg = direction == CLM_CG_MAXIMIZE ? 1.0 : -1.0;
    CLM_TRY(sample(W, H, ldh, Psi, sg * h, num, true));
    dft_kernel<<<1, kBlk, 0, st>>>(ws.sc, num, 2.0 * sg);
    real_parts_kernel<T><<<1, 64, 0, st>>>(num, ws.cost_vals, ws.sc->costs);
    select_dft_kernel<<<1, 64, 0, st>>>(ws.sc, num, t_dft, direction == CLM_CG_MAXIMIZE ? 1 : 0);
    CLM_TRY_HIP(hipGetLastError());
    double r[3] = {0, 0, 0};
    CLM_TRY(read(&ws.sc->mu, 3, r));
    if (info) {
      info->interval = t_dft;
      info->num_samples = num;
      fill_samples(info, num);
    }
    if (r[0] > 0.0) {
      if (info) {
        info->num_roots = static_cast<int>(r[2]);
        info->found = 1;
      }
      *mu = r[0];
      return CLM_STATUS_SUCCESS;
    }
    // no candidate survived: the polynomial search, warm-started from this one's vector
    return search_poly(W, H, ldh, Psi, direction, o, false, mu, info);
  }

  clm_status_t search(const T *W, const T *H, int ldh, const T *Psi, int direction, int method,
                      const clm_cg_search_options &o, bool fresh, double *mu,
                      clm_cg_search_info *info) {
    if (method == CLM_CG_SEARCH_DFT) return search_dft(W, H, ldh, Psi, direction, o, fresh, mu, info);
    return search_poly(W, H, ldh, Psi, direction, o, fresh, mu, info);
  }

  // W <- exp(t H) W (cg_solver.cppm:298-333)
  clm_status_t step(const T *H, int ldh, double t, T *W) {
    CLM_TRY_HIP(launch_scale_copy<T>(st, n, t, H, ldh, ws.tmp, n));
    CLM_TRY(Api<T>::expm(ctx, st, n, ws.tmp, ws.rot, ws.scratch, ws.scratch_bytes));
    CLM_TRY_HIP(gemm_nn<T>(st, n, ws.rot, n, W, n, ws.w_new, n));
    return copy(ws.w_new, W);
  }

  // W <- W (3 I - W^H W) / 2: one Newton-Schulz step toward the polar factor
  clm_status_t reorthonormalize(T *W) {
    adjoint_kernel<T><<<tiles32(n), kBlk, 0, st>>>(n, W, n, ws.wadj, n);
    CLM_TRY_HIP(gemm_nn<T>(st, n, ws.wadj, n, W, n, ws.tmp, n));
    polar_factor_kernel<T><<<columns(n), kBlk, 0, st>>>(n, ws.tmp, n);
    CLM_TRY_HIP(gemm_nn<T>(st, n, W, n, ws.tmp, n, ws.w_new, n));
    return copy(ws.w_new, W);
  }

  // Psi = dJ/dW(W) into ws.psi, G = Psi W^H - W Psi^H (cg_solver.cppm:197-222)
  clm_status_t riemannian_gradient(const T *W, T *G) {
    CLM_TRY(grad(W, ws.psi));
    adjoint_kernel<T><<<tiles32(n), kBlk, 0, st>>>(n, W, n, ws.wadj, n);
    CLM_TRY_HIP(gemm_nn<T>(st, n, ws.psi, n, ws.wadj, n, ws.tmp, n));
    skew_kernel<T><<<tiles32(n), kBlk, 0, st>>>(n, ws.tmp, n, G, n);
    CLM_TRY_HIP(hipGetLastError());
    return CLM_STATUS_SUCCESS;
  }

  // Table 3 (cg_solver.cppm:178-432)
  clm_status_t solve(T *W, int direction, const clm_cg_options &o, clm_cg_info *local) {
    const double sg = direction == CLM_CG_MAXIMIZE ? 1.0 : -1.0;
    const int period = o.reset_period > 0 ? o.reset_period : n * n;
    const long cnt = count();
    const int fb = flat_blocks(cnt);
    T *G = ws.grad, *Gn = ws.grad_next;
    init_scalars_kernel<<<1, 64, 0, st>>>(ws.sc);
    CLM_TRY(riemannian_gradient(W, G));
    CLM_TRY(copy(G, ws.dir));
    ++local->resets;
    CLM_TRY(dot(G, nullptr, G, 0.5, &ws.sc->gg));
    double r[2] = {0, 0};
    CLM_TRY(read(&ws.sc->gg, 2, r));
    double gg = r[0];
    local->gradient_norm_sq = gg;
    bool fresh = true;  // the first search seeds the power iteration with ones
    for (int k = 0; k < o.max_iterations; ++k) {
      if (k > 0 && k % period == 0) {
        CLM_TRY(copy(G, ws.dir));
        ++local->resets;
      }
      if (gg < o.tol) {
        local->reason = CLM_CG_CONVERGED;
        return CLM_STATUS_SUCCESS;
      }
      double mu = 0.0;
      CLM_TRY(search(W, ws.dir, n, ws.psi, direction, o.method, o.search, fresh, &mu, nullptr));
      fresh = false;
      if (!(mu > 0.0)) {
        ++local->line_search_failures;
        CLM_TRY(copy(G, ws.dir));
        ++local->resets;
        CLM_TRY(search(W, ws.dir, n, ws.psi, direction, o.method, o.search, false, &mu, nullptr));
        if (!(mu > 0.0)) {
          local->iterations = k;
          local->reason = CLM_CG_LINE_SEARCH_FAILED;
          return CLM_STATUS_SUCCESS;
        }
      }
      local->last_step = mu;
      CLM_TRY(step(ws.dir, n, sg * mu, W));
      if (o.reorthogonalize_period > 0 && (k + 1) % o.reorthogonalize_period == 0)
        CLM_TRY(reorthonormalize(W));
      CLM_TRY(riemannian_gradient(W, Gn));
      CLM_TRY(dot(Gn, nullptr, Gn, 0.5, &ws.sc->gg_next));
      CLM_TRY(dot(Gn, G, Gn, 1.0, &ws.sc->numer));
      polak_ribiere_kernel<<<1, 64, 0, st>>>(ws.sc);
      direction_kernel<T><<<fb, kBlk, 0, st>>>(cnt, Gn, ws.dir, ws.sc);
      std::swap(G, Gn);
      CLM_TRY(dot(ws.dir, nullptr, G, 1.0, &ws.sc->hg));
      reset_kernel<T><<<fb, kBlk, 0, st>>>(cnt, G, ws.dir, ws.sc);
      CLM_TRY_HIP(hipGetLastError());
      CLM_TRY(read(&ws.sc->gg, 2, r));
      gg = r[0];
      local->gradient_norm_sq = gg;
      if (r[1] != 0.0) ++local->resets;
      local->iterations = k + 1;
    }
    local->reason = CLM_CG_MAX_ITERATIONS;
    return CLM_STATUS_SUCCESS;
  }
};

template <typename T>
clm_status_t setup(Cg<T> &cg, clm_ctx *ctx, clm_stream_t stream, int n, const clm_cg_cost *cost,
                   void *work, size_t lwork) {
  if (n < 1 || n > kMaxDim || work == nullptr || !valid_cost(cost)) return CLM_STATUS_INVALID_VALUE;
  const size_t cost_bytes = cost->buffer_size(cost->user, n);
  size_t need = 0;
  CLM_TRY(cg_buffer_size<T>(n, cost_bytes, &need));
  if (lwork < need) return CLM_STATUS_ALLOC_FAILED;
  size_t eb = 0;
  CLM_TRY(Api<T>::expm_bytes(n, &eb));
  CLM_TRY(resolve_ctx(ctx, &cg.ctx));
  cg.st = static_cast<hipStream_t>(stream);
  cg.n = n;
  cg.cost = cost;
  cg.ws.carve(work, n, eb, cost_bytes);
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t cg_unitary_impl(clm_ctx *ctx, clm_stream_t stream, int n, T *W,
                             const clm_cg_cost *cost, int direction, void *work, size_t lwork,
                             const clm_cg_options *opts, clm_cg_info *info) {
  clm_cg_info local{};
  local.reason = CLM_CG_NUMERICAL_FAILURE;
  if (info) *info = local;
  if (W == nullptr || (direction != CLM_CG_MINIMIZE && direction != CLM_CG_MAXIMIZE))
    return CLM_STATUS_INVALID_VALUE;
  clm_cg_options o{300, Api<T>::tol, CLM_CG_SEARCH_POLYNOMIAL, default_search(), 0, 0};
  if (opts) o = *opts;
  if (o.method != CLM_CG_SEARCH_POLYNOMIAL && o.method != CLM_CG_SEARCH_DFT)
    return CLM_STATUS_INVALID_VALUE;
  Cg<T> cg;
  CLM_TRY(setup(cg, ctx, stream, n, cost, work, lwork));
  const clm_status_t s = cg.solve(W, direction, o, &local);
  if (s != CLM_STATUS_SUCCESS) local.reason = CLM_CG_NUMERICAL_FAILURE;
  if (info) *info = local;
  return s;
}

template <typename T>
clm_status_t search_impl(clm_ctx *ctx, clm_stream_t stream, int n, const T *W, const T *H,
                         int ldh, const T *Psi, const clm_cg_cost *cost, int direction, int method,
                         void *work, size_t lwork, const clm_cg_search_options *opts, double *mu,
                         clm_cg_search_info *info) {
  if (mu == nullptr) return CLM_STATUS_INVALID_VALUE;
  *mu = 0.0;
  if (W == nullptr || H == nullptr || Psi == nullptr || ldh < n ||
      (direction != CLM_CG_MINIMIZE && direction != CLM_CG_MAXIMIZE) ||
      (method != CLM_CG_SEARCH_POLYNOMIAL && method != CLM_CG_SEARCH_DFT))
    return CLM_STATUS_INVALID_VALUE;
  const clm_cg_search_options o = opts ? *opts : default_search();
  Cg<T> cg;
  CLM_TRY(setup(cg, ctx, stream, n, cost, work, lwork));
  return cg.search(W, H, ldh, Psi, direction, method, o, true, mu, info);
}

// The step's cost table is not read: any order-1 stand-in sizes the workspace.
size_t no_cost_bytes(void *, int) { return 0; }
clm_status_t no_gradient(void *, clm_stream_t, int, const void *, int, void *, int, void *, size_t) {
  return CLM_STATUS_NOT_SUPPORTED;
}
clm_status_t no_value(void *, clm_stream_t, int, const void *, int, const void *, int, void *,
                      void *, size_t) {
  return CLM_STATUS_NOT_SUPPORTED;
}

template <typename T>
clm_status_t step_impl(clm_ctx *ctx, clm_stream_t stream, int n, const T *H, int ldh, double t,
                       T *W, void *work, size_t lwork) {
  if (H == nullptr || W == nullptr || ldh < n || !std::isfinite(t)) return CLM_STATUS_INVALID_VALUE;
  static const clm_cg_cost none{1, no_cost_bytes, no_gradient, no_value, nullptr};
  Cg<T> cg;
  CLM_TRY(setup(cg, ctx, stream, n, &none, work, lwork));
  return cg.step(H, ldh, t, W);
}

// ---- the Brockett criterion (src/cg_unitary/brockett.cppm:29-130) ------------------------------
template <typename T>
size_t brockett_bytes(void *, int n) {
  return Layout::align(static_cast<size_t>(n) * n * sizeof(T)) + kMaxPartials * sizeof(double);
}

// Psi = R W N: one product and a column scaling
template <typename T>
clm_status_t brockett_gradient(void *user, clm_stream_t stream, int n, const void *dW, int ldw,
                               void *dPsi, int ldp, void *work, size_t lwork) {
  const auto *b = static_cast<const clm_cg_brockett *>(user);
  if (b == nullptr || b->dR == nullptr || b->dN == nullptr || b->ldr < n)
    return CLM_STATUS_INVALID_VALUE;
  if (lwork < static_cast<size_t>(n) * n * sizeof(T)) return CLM_STATUS_ALLOC_FAILED;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  T *tmp = static_cast<T *>(work);
  CLM_TRY_HIP(gemm_nn<T>(st, n, static_cast<const T *>(b->dR), b->ldr, static_cast<const T *>(dW),
                         ldw, tmp, n));
  col_scale_kernel<T><<<columns(n), kBlk, 0, st>>>(n, tmp, n, static_cast<const T *>(b->dN),
                                                   static_cast<T *>(dPsi), ldp);
  CLM_TRY_HIP(hipGetLastError());
  return CLM_STATUS_SUCCESS;
}

// J(W) = Re dotc(W, Psi) into *dOut (Psi computed first when not given)
template <typename T>
clm_status_t brockett_value(void *user, clm_stream_t stream, int n, const void *dW, int ldw,
                            const void *dPsi, int ldp, void *dOut, void *work, size_t lwork) {
  if (ldw != n || (dPsi != nullptr && ldp != n)) return CLM_STATUS_INVALID_VALUE;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t block = Layout::align(static_cast<size_t>(n) * n * sizeof(T));
  unsigned char *w = static_cast<unsigned char *>(work);
  const T *psi = static_cast<const T *>(dPsi);
  size_t used = 0;
  if (psi == nullptr) {
    if (lwork < 2 * block + kMaxPartials * sizeof(double)) return CLM_STATUS_ALLOC_FAILED;
    CLM_TRY(brockett_gradient<T>(user, stream, n, dW, ldw, w, n, w + block, lwork - block));
    psi = reinterpret_cast<const T *>(w);
    used = 2 * block;
  } else if (lwork < block + kMaxPartials * sizeof(double)) {
    return CLM_STATUS_ALLOC_FAILED;
  } else {
    used = block;
  }
  double *partials = reinterpret_cast<double *>(w + used);
  CLM_TRY_HIP((launch_dot<T, T>(st, static_cast<long>(n) * n, static_cast<const T *>(dW), nullptr,
                               psi, partials, kMaxPartials, 1.0, static_cast<T *>(dOut))));
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t brockett_cost(clm_cg_brockett *b, clm_cg_cost *cost) {
  if (b == nullptr || cost == nullptr) return CLM_STATUS_INVALID_VALUE;
  *cost = clm_cg_cost{2, brockett_bytes<T>, brockett_gradient<T>, brockett_value<T>, b};
  return CLM_STATUS_SUCCESS;
}

void default_options(clm_cg_options *o, double tol) {
  if (o) *o = clm_cg_options{300, tol, CLM_CG_SEARCH_POLYNOMIAL, default_search(), 0, 0};
}

}  // namespace
}  // namespace clm

using clm::zc;
using clm::zd;

extern "C" {

clm_status_t clm_cg_brockett_cost_s(clm_cg_brockett *b, clm_cg_cost *c) {
  return clm::brockett_cost<float>(b, c);
}
clm_status_t clm_cg_brockett_cost_d(clm_cg_brockett *b, clm_cg_cost *c) {
  return clm::brockett_cost<double>(b, c);
}
clm_status_t clm_cg_brockett_cost_c(clm_cg_brockett *b, clm_cg_cost *c) {
  return clm::brockett_cost<zc>(b, c);
}
clm_status_t clm_cg_brockett_cost_z(clm_cg_brockett *b, clm_cg_cost *c) {
  return clm::brockett_cost<zd>(b, c);
}

void clm_cg_default_options_s(clm_cg_options *o) { clm::default_options(o, 1e-10); }
void clm_cg_default_options_d(clm_cg_options *o) { clm::default_options(o, 1e-20); }

clm_status_t clm_cg_unitary_buffer_size_s(int n, size_t cb, size_t *b) {
  return clm::cg_buffer_size<float>(n, cb, b);
}
clm_status_t clm_cg_unitary_buffer_size_d(int n, size_t cb, size_t *b) {
  return clm::cg_buffer_size<double>(n, cb, b);
}
clm_status_t clm_cg_unitary_buffer_size_c(int n, size_t cb, size_t *b) {
  return clm::cg_buffer_size<zc>(n, cb, b);
}
clm_status_t clm_cg_unitary_buffer_size_z(int n, size_t cb, size_t *b) {
  return clm::cg_buffer_size<zd>(n, cb, b);
}

clm_status_t clm_cg_unitary_s(clm_ctx *ctx, clm_stream_t stream, int n, float *dW,
                              const clm_cg_cost *cost, int direction, void *dWork, size_t lwork,
                              const clm_cg_options *opts, clm_cg_info *info) {
  return clm::cg_unitary_impl<float>(ctx, stream, n, dW, cost, direction, dWork, lwork, opts, info);
}
clm_status_t clm_cg_unitary_d(clm_ctx *ctx, clm_stream_t stream, int n, double *dW,
                              const clm_cg_cost *cost, int direction, void *dWork, size_t lwork,
                              const clm_cg_options *opts, clm_cg_info *info) {
  return clm::cg_unitary_impl<double>(ctx, stream, n, dW, cost, direction, dWork, lwork, opts, info);
}
clm_status_t clm_cg_unitary_c(clm_ctx *ctx, clm_stream_t stream, int n, clm_float_complex *dW,
                              const clm_cg_cost *cost, int direction, void *dWork, size_t lwork,
                              const clm_cg_options *opts, clm_cg_info *info) {
  return clm::cg_unitary_impl<zc>(ctx, stream, n, reinterpret_cast<zc *>(dW), cost, direction,
                                  dWork, lwork, opts, info);
}
clm_status_t clm_cg_unitary_z(clm_ctx *ctx, clm_stream_t stream, int n, clm_double_complex *dW,
                              const clm_cg_cost *cost, int direction, void *dWork, size_t lwork,
                              const clm_cg_options *opts, clm_cg_info *info) {
  return clm::cg_unitary_impl<zd>(ctx, stream, n, reinterpret_cast<zd *>(dW), cost, direction,
                                  dWork, lwork, opts, info);
}

clm_status_t clm_cg_geodesic_search_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dW,
                                      const float *dH, int ldh, const float *dPsi,
                                      const clm_cg_cost *cost, int direction, int method,
                                      void *dWork, size_t lwork, const clm_cg_search_options *opts,
                                      double *mu, clm_cg_search_info *info) {
  return clm::search_impl<float>(ctx, stream, n, dW, dH, ldh, dPsi, cost, direction, method, dWork,
                                 lwork, opts, mu, info);
}
clm_status_t clm_cg_geodesic_search_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dW,
                                      const double *dH, int ldh, const double *dPsi,
                                      const clm_cg_cost *cost, int direction, int method,
                                      void *dWork, size_t lwork, const clm_cg_search_options *opts,
                                      double *mu, clm_cg_search_info *info) {
  return clm::search_impl<double>(ctx, stream, n, dW, dH, ldh, dPsi, cost, direction, method,
                                  dWork, lwork, opts, mu, info);
}
clm_status_t clm_cg_geodesic_search_c(clm_ctx *ctx, clm_stream_t stream, int n,
                                      const clm_float_complex *dW, const clm_float_complex *dH,
                                      int ldh, const clm_float_complex *dPsi,
                                      const clm_cg_cost *cost, int direction, int method,
                                      void *dWork, size_t lwork, const clm_cg_search_options *opts,
                                      double *mu, clm_cg_search_info *info) {
  return clm::search_impl<zc>(ctx, stream, n, reinterpret_cast<const zc *>(dW),
                              reinterpret_cast<const zc *>(dH), ldh,
                              reinterpret_cast<const zc *>(dPsi), cost, direction, method, dWork,
                              lwork, opts, mu, info);
}
clm_status_t clm_cg_geodesic_search_z(clm_ctx *ctx, clm_stream_t stream, int n,
                                      const clm_double_complex *dW, const clm_double_complex *dH,
                                      int ldh, const clm_double_complex *dPsi,
                                      const clm_cg_cost *cost, int direction, int method,
                                      void *dWork, size_t lwork, const clm_cg_search_options *opts,
                                      double *mu, clm_cg_search_info *info) {
  return clm::search_impl<zd>(ctx, stream, n, reinterpret_cast<const zd *>(dW),
                              reinterpret_cast<const zd *>(dH), ldh,
                              reinterpret_cast<const zd *>(dPsi), cost, direction, method, dWork,
                              lwork, opts, mu, info);
}

clm_status_t clm_cg_geodesic_step_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dH,
                                    int ldh, double t, float *dW, void *dWork, size_t lwork) {
  return clm::step_impl<float>(ctx, stream, n, dH, ldh, t, dW, dWork, lwork);
}
clm_status_t clm_cg_geodesic_step_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dH,
                                    int ldh, double t, double *dW, void *dWork, size_t lwork) {
  return clm::step_impl<double>(ctx, stream, n, dH, ldh, t, dW, dWork, lwork);
}
clm_status_t clm_cg_geodesic_step_c(clm_ctx *ctx, clm_stream_t stream, int n,
                                    const clm_float_complex *dH, int ldh, double t,
                                    clm_float_complex *dW, void *dWork, size_t lwork) {
  return clm::step_impl<zc>(ctx, stream, n, reinterpret_cast<const zc *>(dH), ldh, t,
                            reinterpret_cast<zc *>(dW), dWork, lwork);
}
clm_status_t clm_cg_geodesic_step_z(clm_ctx *ctx, clm_stream_t stream, int n,
                                    const clm_double_complex *dH, int ldh, double t,
                                    clm_double_complex *dW, void *dWork, size_t lwork) {
  return clm::step_impl<zd>(ctx, stream, n, reinterpret_cast<const zd *>(dH), ldh, t,
                            reinterpret_cast<zd *>(dW), dWork, lwork);
}

}  // extern "C"
