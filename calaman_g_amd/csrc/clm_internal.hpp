// clm_internal.hpp -- types, workspace layout and kernel launchers shared by the
// HIP translation units of the MI355X matrix-exponential library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "calaman_expm.h"

// A context: the device it was created on and the stream it enqueues on.
struct clm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
};

#define CLM_TRY(expr)                                                  \
  do {                                                                 \
    const clm_status_t clm_try_status_ = (expr);                       \
    if (clm_try_status_ != CLM_STATUS_SUCCESS) return clm_try_status_; \
  } while (0)
#define CLM_TRY_HIP(expr)                                         \
  do {                                                            \
    if ((expr) != hipSuccess) return CLM_STATUS_EXECUTION_FAILED; \
  } while (0)

namespace clm {

// ctx, or the library's default context of the current device when ctx is null
clm_status_t resolve_ctx(clm_ctx *ctx, clm_ctx **out);

// n*n must fit an int in the reference's flat-index kernels (src/expm/detail.cppm:25-28);
// the bound is kept so the two libraries accept the same orders.
constexpr int kMaxDim = 46340;

// Interleaved double complex, 16-byte aligned so one global/LDS access moves it.
struct alignas(16) zd {
  double re, im;
};

__host__ __device__ inline zd make_zd(double r, double i) { return zd{r, i}; }

// Interleaved float complex (complex64), one 8-byte access.
struct alignas(8) zc {
  float re, im;
};

// ---- element arithmetic (real and complex share the kernels' templates) ----
template <typename T>
struct ops;

template <>
struct ops<double> {
  using real = double;
  __host__ __device__ static double zero() { return 0.0; }
  __host__ __device__ static double from_real(double r) { return r; }
  __host__ __device__ static double add(double a, double b) { return a + b; }
  __host__ __device__ static double mul(double a, double b) { return a * b; }
  // c * x + acc, the order the reference block build uses
  __host__ __device__ static double mad(double c, double x, double acc) { return c * x + acc; }
  __host__ __device__ static double scale(double x, double r) { return x * r; }
  __host__ __device__ static double ldexp(double x, int e) { return ::ldexp(x, e); }
  __host__ __device__ static bool is_zero(double a) { return a == 0.0; }
};

template <>
struct ops<float> {
  using real = float;
  __host__ __device__ static float zero() { return 0.f; }
  __host__ __device__ static float from_real(double r) { return static_cast<float>(r); }
  __host__ __device__ static float add(float a, float b) { return a + b; }
  __host__ __device__ static float mul(float a, float b) { return a * b; }
  __host__ __device__ static float mad(float c, float x, float acc) { return c * x + acc; }
  __host__ __device__ static float scale(float x, double r) { return x * static_cast<float>(r); }
  __host__ __device__ static float ldexp(float x, int e) { return ::ldexpf(x, e); }
  __host__ __device__ static bool is_zero(float a) { return a == 0.f; }
};

template <>
struct ops<zc> {
  using real = float;
  __host__ __device__ static zc zero() { return zc{0.f, 0.f}; }
  __host__ __device__ static zc from_real(double r) { return zc{static_cast<float>(r), 0.f}; }
  __host__ __device__ static zc add(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }
  __host__ __device__ static zc mul(zc a, zc b) {
    return zc{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  }
  __host__ __device__ static zc mad(zc c, zc x, zc acc) { return add(mul(c, x), acc); }
  __host__ __device__ static zc scale(zc x, double r) {
    const float f = static_cast<float>(r);
    return zc{x.re * f, x.im * f};
  }
  __host__ __device__ static zc ldexp(zc x, int e) { return zc{::ldexpf(x.re, e), ::ldexpf(x.im, e)}; }
  __host__ __device__ static bool is_zero(zc a) { return a.re == 0.f && a.im == 0.f; }
};

template <>
struct ops<zd> {
  using real = double;
  __host__ __device__ static zd zero() { return zd{0.0, 0.0}; }
  __host__ __device__ static zd from_real(double r) { return zd{r, 0.0}; }
  __host__ __device__ static zd add(zd a, zd b) { return zd{a.re + b.re, a.im + b.im}; }
  __host__ __device__ static zd mul(zd a, zd b) {
    return zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  }
  __host__ __device__ static zd mad(zd c, zd x, zd acc) { return add(mul(c, x), acc); }
  __host__ __device__ static zd scale(zd x, double r) { return zd{x.re * r, x.im * r}; }
  __host__ __device__ static zd ldexp(zd x, int e) { return zd{::ldexp(x.re, e), ::ldexp(x.im, e)}; }
  __host__ __device__ static bool is_zero(zd a) { return a.re == 0.0 && a.im == 0.0; }
};

template <typename T>
using real_t = typename ops<T>::real;

// ---- workspace carve -----------------------------------------------------------
// One caller-owned buffer carved into 256-byte aligned regions; a null base
// sizes without carving, so a size query and the routine share one layout
// (the reference's WorkspaceLayout, src/common/workspace.cppm:69-159).
class Layout {
 public:
  explicit Layout(void *base) : base_(static_cast<char *>(base)) {}
  template <typename U>
  U *take(size_t count) {
    offset_ = align(offset_);
    U *p = base_ ? reinterpret_cast<U *>(base_ + offset_) : nullptr;
    offset_ += count * sizeof(U);
    return p;
  }
  size_t total() const { return align(offset_); }
  static size_t align(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

 private:
  char *base_;
  size_t offset_ = 0;
};

// ---- GEMM ---------------------------------------------------------------------------
// C = alpha * A * B [+ beta * Cin] [+ B_blk], column-major NN, where the optional
// block term is the Paterson-Stockmeyer block
//   B_blk = coef[0] I + coef[1] X1 + sum_{t>=2} coef[t] bank[t-2]
// built in the epilogue (coef on the device). With nterms == 1 it is the Horner
// step's coef[0] I.
template <typename T>
struct GemmArgs {
  int M = 0, N = 0, K = 0;
  const T *A = nullptr;
  long lda = 0;
  const T *B = nullptr;
  long ldb = 0;
  T *C = nullptr;
  long ldc = 0;
  T alpha{};
  T beta{};
  const T *Cin = nullptr;  // read when non-null (may alias C)
  long ldcin = 0;
  const T *coef = nullptr;  // block coefficients, device
  int nterms = 0;
  const T *X1 = nullptr;
  long ldx1 = 0;
  const T *bank = nullptr;
  long bank_stride = 0;
  long ldbank = 0;
  // the split products only: when set, the epilogue also leaves each output row's
  // largest-magnitude exponent in row_e_out, tagged with generation e_gen_out (the
  // half-format split's own tags), so the next split of C needs no exponent pass
  int *row_e_out = nullptr;
  int e_gen_out = 0;
};

hipError_t launch_gemm_d(hipStream_t stream, const GemmArgs<double> &g);
// the same product on BF16 MFMA with operands split into 2 (split = 3 products)
// or 3 (split = 6 products) bfloat16 terms, FP32 accumulation, FP64 in and out
hipError_t launch_gemm_split_bf16(hipStream_t stream, const GemmArgs<double> &g, int split);
// bfloat16 planes of an FP64 rows x cols matrix (2 planes for split 3, 3 for split 6):
// pc column-major (ld rows), pr row-major (ld cols); either may be null
hipError_t launch_split_planes(hipStream_t stream, int split, int rows, int cols, const double *X,
                               long ldx, unsigned short *pc, unsigned short *pr, long pstride);
// the split product from pre-split planes: PA = row-major planes of A (M x K),
// PB = column-major planes of B (K x N); g.A / g.B are not read
// (row_e / col_e: the half format's tagged exponents of A's rows / B's columns,
// written by generations row_gen / col_gen; unused by the bfloat16 formats)
hipError_t launch_gemm_split_planes(hipStream_t stream, const GemmArgs<double> &g, int split,
                                    const unsigned short *PA, const unsigned short *PB,
                                    long pstride, const int *row_e = nullptr,
                                    const int *col_e = nullptr, int row_gen = 0, int col_gen = 0);
// IEEE-half planes (split = CLM_SPLIT_FP16X3). Exponent arrays hold values tagged
// with the generation (>= 1) of the split that wrote them, so they are cleared
// (to generation 0) once per run of splits, not before each: ...
hipError_t launch_clear_exponents(hipStream_t stream, int n, int *a, int *b, int *c,
                                  int *d = nullptr);
// ... the largest-magnitude exponent of every row (row_e) and column (col_e) of
// X, as generation gen, either array may be null ...
hipError_t launch_half_exponents(hipStream_t stream, int rows, int cols, const double *X, long ldx,
                                 int *row_e, int *col_e, int gen);
// ... and the two half planes of X scaled by them: pc column-major by col_e
// (right operand), pr row-major by row_e (left operand); either may be null
hipError_t launch_split_planes_half(hipStream_t stream, int rows, int cols, const double *X,
                                    long ldx, unsigned short *pc, unsigned short *pr, long pstride,
                                    const int *row_e, const int *col_e, int row_gen, int col_gen);
// e4m3 planes with one power-of-two exponent per 128 x 128 tile: tile (ti, tj)
// of X is scaled by 2^e, e = the tile's entry [ti + tj * ceil(rows / 128)] of
// ec / er, so its largest magnitude lands in (224, 448]; pc column-major
// (ld rows) with exponents ec, pr row-major (ld cols) with exponents er; any
// of the four may be null
hipError_t launch_quantize_fp8(hipStream_t stream, int rows, int cols, const double *X, long ldx,
                               unsigned char *pc, unsigned char *pr, int *ec, int *er);
// the product on FP8 MFMA from e4m3 planes: PA row-major planes of A (M x K),
// PB column-major planes of B (K x N) with their tile exponents; FP32 sums per
// 128-deep k-tile, FP64 epilogue; g.A / g.B are not read
hipError_t launch_gemm_fp8(hipStream_t stream, const GemmArgs<double> &g, const unsigned char *PA,
                           const unsigned char *PB, const int *eA, const int *eB);
hipError_t launch_gemm_z(hipStream_t stream, const GemmArgs<zd> &g);
template <typename T>
hipError_t launch_gemm(hipStream_t stream, const GemmArgs<T> &g);

// ---- element-wise and reduction kernels ------------------------------------------------
// 1-norm in two launches (colsums, NaN-propagating max into dColsum[0]).
template <typename T>
hipError_t launch_norm1(hipStream_t stream, int n, const T *A, long lda, real_t<T> *colsum,
                        real_t<T> *out = nullptr);
// the 1-norms of count n x n blocks (leading dimension n) `stride` elements apart:
// column sums into colsums (count * n values), maxima into out[0 .. count-1];
// each bit for bit launch_norm1's
template <typename T>
hipError_t launch_norm1_powers(hipStream_t stream, int n, const T *bank, long stride, int count,
                               real_t<T> *colsums, real_t<T> *out);
// bank[k-2] *= 2^(-k s) for k = 2 .. count + 1 (n x n blocks, leading dimension n)
template <typename T>
hipError_t launch_scale_powers(hipStream_t stream, int n, T *bank, long stride, int count, int s);
// vals[0 .. count-1] = src (device -> host-mapped memory), then *flag = seq (system
// scope, after the values)
template <typename R>
hipError_t launch_publish(hipStream_t stream, const R *src, int count, double *vals,
                          unsigned *flag, unsigned seq);
// out[k] = 2^(-(k mod b) s) c[k] for k < count (device arrays)
template <typename T>
hipError_t launch_scale_coef(hipStream_t stream, const T *c, T *out, int count, int b, int s);
// lange: which in CLM_NORM_*; result in *d_result.
template <typename T>
hipError_t launch_lange(hipStream_t stream, int which, size_t m, size_t n, const T *A, size_t lda,
                        real_t<T> *d_result, real_t<T> *d_scratch);
// out = f * in (n x n), f a real factor
template <typename T>
hipError_t launch_scale_copy(hipStream_t stream, int n, double f, const T *in, long ldi, T *out,
                             long ldo);
// A(i,j) *= mul, m x n in place
template <typename T>
hipError_t launch_lascl(hipStream_t stream, size_t m, size_t n, double mul, T *A, size_t lda);
// B = coef[0] I + coef[1] X1 + sum_t coef[t] bank[t-2]
// launch_build_block for FP64 on the 64 x 64 element-wise tiling, with each row's
// tagged exponent (generation gen, the half-format splits' encoding) into row_e
hipError_t launch_build_block_rows(hipStream_t stream, int n, double *Bout, long ldb,
                                   const double *coef, int nterms, const double *X1, long ldx1,
                                   const double *bank, long stride, long ldbank, int *row_e,
                                   int gen);
template <typename T>
hipError_t launch_build_block(hipStream_t stream, int n, T *Bout, long ldb, const T *coef,
                              int nterms, const T *X1, long ldx1, const T *bank, long stride,
                              long ldbank);
// P = (*alpha) I
template <typename T>
hipError_t launch_set_scaled_identity(hipStream_t stream, int n, T *P, long ldp, const T *alpha);
// dst = src (n x n)
template <typename T>
hipError_t launch_copy(hipStream_t stream, int n, const T *src, long lds, T *dst, long ldd);
// d[0..count) = v
hipError_t launch_fill_int(hipStream_t stream, int *d, int count, int v);

// ---- the Pade route: LU with partial pivoting, triangular solves, numerator passes ----------
// A = P L U in place (unit-lower L below the diagonal, U on and above), ipiv 1-based
// row interchanges; *info = first zero pivot's column (1-based) or 0
template <typename T>
hipError_t launch_getrf(hipStream_t st, int n, T *A, long lda, int *ipiv, int *info);
// B <- A^-1 B from getrf's factors (nrhs columns)
template <typename T>
hipError_t launch_getrs(hipStream_t st, int n, int nrhs, const T *LU, long lda, const int *ipiv,
                        T *B, long ldb);
// V = cv[0] I + sum_k cv[k+1] P[k], W = cw[0] I + sum_k cw[k+1] P[k], k < np <= 4 (n x n
// blocks, leading dimension n)
template <typename T>
hipError_t launch_pade_even_odd(hipStream_t st, int n, int np, const T *const *P,
                                const double *cv, const double *cw, T *V, T *W);
// R = V + U (leading dimension ldr), Q = V - U (Q may be V)
template <typename T>
hipError_t launch_pade_split(hipStream_t st, int n, const T *U, const T *V, T *R, long ldr, T *Q);

// src[0 .. count-1] (count <= 5) once the stream reaches them: a flag polled in
// host-mapped memory, then a blocking synchronization after 2 ms
template <typename R>
clm_status_t read_device_values(hipStream_t st, int device, const R *src, int count, double *out);

// Device tables of 1/k!, k = 0..32 (correctly rounded), for the Taylor coefficients.
const double *inv_factorials_d(int device);
const zd *inv_factorials_z(int device);
const float *inv_factorials_s(int device);
const zc *inv_factorials_c(int device);
template <typename T>
const T *inv_factorials(int device);

}  // namespace clm
