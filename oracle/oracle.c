/*
 * oracle.c -- CPU restatement of calaman_G's matrix-exponential path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the GPU library is compared against.
 * Nothing in the shipped library links or calls it; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * What is restated, and from where (paths relative to the calaman_G root):
 *   - Pade coefficient tables, degree ladder, powers/GEMM counts:
 *       src/expm/expm_bridge.h:58-130
 *   - backward-error thresholds and the (m, s) plan: src/expm/plan.cppm:51-130
 *   - the 1-norm with the device fold order: src/expm/norm1.cppm:45-56,
 *       src/reduce_columns/reduce_columns.cuh:62-108 (4 wave64 = 256 lanes),
 *       src/expm/expm.cu:210-245 (max_reduce, NaN wins)
 *   - ?lange norms: src/lapack/lange/lange.cu:42-156
 *   - ?lascl guarded multiplier loop: src/lapack/lascl/interface.cppm:79-143
 *   - pade(): src/expm/pade.cppm:77-191 and its kernels src/expm/expm.cu:73-162
 *   - expm(): src/expm/expm.cppm:87-174
 *   - Paterson-Stockmeyer: src/paterson_stockmeyer/interface.cppm:117-422,
 *       src/paterson_stockmeyer/paterson_stockmeyer.cu:51-84
 *   - Horner: src/horner/interface.cppm:156-194, src/horner/horner.cu:39-91
 * The GEMM and LU the reference takes from rocBLAS/rocSOLVER are restated as a
 * deterministic blocked GEMM (every element sums k ascending, whatever the
 * thread count) and LAPACK's blocked right-looking partial-pivot LU (?getrf
 * over ?getf2 panels, with ?getf2's pivot rule) and blocked ?getrs.
 *
 * The Taylor section at the end is the build's own algorithm (truncated
 * Taylor series evaluated by Paterson-Stockmeyer, then squared), stated on
 * the CPU with the same plan rule the GPU library uses.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

/* The reference's int-indexed kernels bound n so n*n fits an int
 * (src/expm/detail.cppm:25-28). */
#define ORC_MAX_DIM 46340

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

/* =========================================================================
 * Pade tables (src/expm/expm_bridge.h:58-113): b_j = (2m-j)! m! / ((2m)! j! (m-j)!)
 * ========================================================================= */
static const double kPade3[4] = {1.0, 1.0 / 2.0, 1.0 / 10.0, 1.0 / 120.0};
static const double kPade5[6] = {1.0, 1.0 / 2.0, 1.0 / 9.0, 1.0 / 72.0, 1.0 / 1008.0, 1.0 / 30240.0};
static const double kPade7[8] = {1.0,          1.0 / 2.0,     3.0 / 26.0,     5.0 / 312.0,
                                 5.0 / 3432.0, 1.0 / 11440.0, 1.0 / 308880.0, 1.0 / 17297280.0};
static const double kPade9[10] = {1.0,          1.0 / 2.0,      2.0 / 17.0,      7.0 / 408.0,
                                  7.0 / 4080.0, 1.0 / 8160.0,   1.0 / 159120.0,  1.0 / 4455360.0,
                                  1.0 / 196035840.0, 1.0 / 17643225600.0};
static const double kPade13[14] = {1.0,
                                   1.0 / 2.0,
                                   3.0 / 25.0,
                                   11.0 / 600.0,
                                   11.0 / 5520.0,
                                   3.0 / 18400.0,
                                   1.0 / 96600.0,
                                   1.0 / 1932000.0,
                                   1.0 / 48944000.0,
                                   1.0 / 1585785600.0,
                                   1.0 / 67395888000.0,
                                   1.0 / 3953892096000.0,
                                   1.0 / 355850288640000.0,
                                   1.0 / 64764752532480000.0};
static const int kPadeDegrees[5] = {3, 5, 7, 9, 13};

const double *orc_pade_coeffs(int m) {
  switch (m) {
  case 3: return kPade3;
  case 5: return kPade5;
  case 7: return kPade7;
  case 9: return kPade9;
  case 13: return kPade13;
  default: return NULL;
  }
}

/* src/expm/expm_bridge.h:128-130 */
int orc_pade_num_powers(int m) { return (m == 13) ? 3 : (m - 1) / 2; }
int orc_pade_num_gemms(int m) { return orc_pade_num_powers(m) + 1 + ((m == 13) ? 2 : 0); }

/* src/expm/plan.cppm:65-96 */
double orc_pade_theta_d(int m) {
  switch (m) {
  case 3: return 1.495585217958292e-2;
  case 5: return 2.539398330063232e-1;
  case 7: return 9.504178996162932e-1;
  case 9: return 2.097847961257067;
  case 13: return 5.371920351148152;
  default: return 0.0;
  }
}

float orc_pade_theta_s(int m) {
  switch (m) {
  case 3: return 4.258730034897931e-1f;
  case 5: return 1.880152698533769f;
  case 7: return 3.925724846433284f;
  case 9: return 6.249156334514102f;
  case 13: return 1.124873763647540e1f;
  default: return 0.0f;
  }
}

/* src/expm/plan.cppm:112-130: the cheapest rung whose theta covers the norm,
 * else m = 13 with s = max(0, ceil(log2(norm / theta_13))). */
void orc_expm_plan_d(double norm1, int *out) {
  for (int i = 0; i < 5; ++i) {
    const int m = kPadeDegrees[i];
    if (norm1 <= orc_pade_theta_d(m)) {
      out[0] = m; out[1] = 0; out[2] = orc_pade_num_gemms(m);
      return;
    }
  }
  /* a ratio that overflows (or a NaN norm) gets a bounded s, not a cast of an
   * infinite log2: 2^20 marks it unusable, NaN plans s = 0 (the reference rejects
   * both before it plans) */
  const double ratio = norm1 / orc_pade_theta_d(13);
  int s;
  if (!(ratio <= DBL_MAX)) s = isnan(ratio) ? 0 : (1 << 20);
  else s = (int)ceil(log2(ratio));
  if (s < 0) s = 0;
  out[0] = 13; out[1] = s; out[2] = orc_pade_num_gemms(13) + s;
}

void orc_expm_plan_s(float norm1, int *out) {
  for (int i = 0; i < 5; ++i) {
    const int m = kPadeDegrees[i];
    if (norm1 <= orc_pade_theta_s(m)) {
      out[0] = m; out[1] = 0; out[2] = orc_pade_num_gemms(m);
      return;
    }
  }
  const float ratio = norm1 / orc_pade_theta_s(13);
  int s;
  if (!(ratio <= FLT_MAX)) s = isnan(ratio) ? 0 : (1 << 20);
  else s = (int)ceilf(log2f(ratio));
  if (s < 0) s = 0;
  out[0] = 13; out[1] = s; out[2] = orc_pade_num_gemms(13) + s;
}

/* =========================================================================
 * Norms
 * ========================================================================= */

/* Lanes of the reduce_columns block on a wave64 device: 4 * 64
 * (src/reduce_columns/reduce_columns.cuh:72), and max_reduce's kMaxThreads
 * (src/expm/expm.cu:210). */
#define RC_BLOCK 256

static inline double zabs(const double *z) { return hypot(z[0], z[1]); }

/* reduce_columns_kernel with pre = |.| and op = + for one column
 * (src/reduce_columns/reduce_columns.cuh:74-107): identity-free strided fold
 * seeded by each lane's first element, then a halving tree over the active
 * lanes. */
static double colsum_fold(const double *col, int n, int is_complex) {
  double sdata[RC_BLOCK];
  const int nactive = MIN(n, RC_BLOCK);
  for (int tid = 0; tid < nactive; ++tid) {
    double acc = is_complex ? zabs(col + 2 * tid) : fabs(col[tid]);
    for (int i = tid + RC_BLOCK; i < n; i += RC_BLOCK) {
      acc = acc + (is_complex ? zabs(col + 2 * (size_t)i) : fabs(col[i]));
    }
    sdata[tid] = acc;
  }
  for (int s = RC_BLOCK / 2; s > 0; s >>= 1) {
    for (int tid = 0; tid < s; ++tid) {
      if (tid + s < nactive) sdata[tid] = sdata[tid] + sdata[tid + s];
    }
  }
  return sdata[0];
}

/* max_reduce_kernel (src/expm/expm.cu:212-245): best seeded at 0, NaN wins. */
static double max_reduce(const double *v, int n) {
  double sdata[RC_BLOCK];
  for (int tid = 0; tid < RC_BLOCK; ++tid) {
    double best = 0.0;
    for (int i = tid; i < n; i += RC_BLOCK) {
      const double x = v[i];
      if (x > best || isnan(x)) best = x;
    }
    sdata[tid] = best;
  }
  for (int s = RC_BLOCK / 2; s > 0; s >>= 1) {
    for (int tid = 0; tid < s; ++tid) {
      const double other = sdata[tid + s];
      if (other > sdata[tid] || isnan(other)) sdata[tid] = other;
    }
  }
  return sdata[0];
}

static double norm1_impl(int n, const double *A, int lda, int cplx) {
  if (n < 1) return 0.0;
  double *colsum = (double *)malloc(sizeof(double) * (size_t)n);
  const size_t es = cplx ? 2 : 1;
  for (int j = 0; j < n; ++j) colsum[j] = colsum_fold(A + es * (size_t)j * lda, n, cplx);
  const double r = max_reduce(colsum, n);
  free(colsum);
  return r;
}

/* matrix_norm1 (src/expm/norm1.cppm:45-56) */
double orc_norm1_d(int n, const double *A, int lda) { return norm1_impl(n, A, lda, 0); }
double orc_norm1_z(int n, const double *A, int lda) { return norm1_impl(n, A, lda, 1); }

/* lange's MaxFunctor (src/lapack/lange/lange.cu:70-81) */
static inline double nan_max(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? b : a;
}

/* generic reduce_columns over a column of length m, identity-free, with op */
static double rc_fold(const double *col, size_t m, int cplx, int pre, int op) {
  /* pre: 0 |.|, 1 re^2+im^2, 2 identity; op: 0 plus, 1 nan_max */
  double sdata[RC_BLOCK];
  const size_t nactive = MIN(m, (size_t)RC_BLOCK);
  for (size_t tid = 0; tid < nactive; ++tid) {
    double acc = 0.0;
    for (size_t i = tid; i < m; i += RC_BLOCK) {
      double x;
      if (pre == 0) {
        x = cplx ? zabs(col + 2 * i) : fabs(col[i]);
      } else if (pre == 1) {
        const double re = cplx ? col[2 * i] : col[i];
        const double im = cplx ? col[2 * i + 1] : 0.0;
        x = re * re + im * im;
      } else {
        x = col[i];
      }
      if (i == tid) acc = x;
      else acc = (op == 0) ? acc + x : nan_max(acc, x);
    }
    sdata[tid] = acc;
  }
  for (size_t s = RC_BLOCK / 2; s > 0; s >>= 1) {
    for (size_t tid = 0; tid < s; ++tid) {
      if (tid + s < nactive)
        sdata[tid] = (op == 0) ? sdata[tid] + sdata[tid + s] : nan_max(sdata[tid], sdata[tid + s]);
    }
  }
  return sdata[0];
}

/* lange (src/lapack/lange/lange.cu:123-156); 0 for an empty matrix
 * (src/lapack/lange/interface.cppm:87-90). */
static double lange_impl(int which, size_t m, size_t n, const double *A, size_t lda, int cplx) {
  if (m == 0 || n == 0) return 0.0;
  const size_t es = cplx ? 2 : 1;
  double r = 0.0;
  if (which == 2) { /* inf: row sums seeded 0, j ascending, then max */
    double *rows = (double *)malloc(sizeof(double) * m);
    for (size_t i = 0; i < m; ++i) {
      double acc = 0.0;
      for (size_t j = 0; j < n; ++j) {
        const double *e = A + es * (i + j * lda);
        acc += cplx ? zabs(e) : fabs(e[0]);
      }
      rows[i] = acc;
    }
    r = rc_fold(rows, m, 0, 2, 1);
    free(rows);
    return r;
  }
  double *cols = (double *)malloc(sizeof(double) * n);
  for (size_t j = 0; j < n; ++j) {
    const double *col = A + es * j * lda;
    if (which == 0) cols[j] = rc_fold(col, m, cplx, 0, 1);
    else if (which == 1) cols[j] = rc_fold(col, m, cplx, 0, 0);
    else cols[j] = rc_fold(col, m, cplx, 1, 0);
  }
  if (which == 3) r = sqrt(rc_fold(cols, n, 0, 2, 0));
  else r = rc_fold(cols, n, 0, 2, 1);
  free(cols);
  return r;
}

double orc_lange_d(int which, size_t m, size_t n, const double *A, size_t lda) {
  return lange_impl(which, m, n, A, lda, 0);
}
double orc_lange_z(int which, size_t m, size_t n, const double *A, size_t lda) {
  return lange_impl(which, m, n, A, lda, 1);
}

/* =========================================================================
 * lascl TYPE='G' (src/lapack/lascl/interface.cppm:79-143, kernel
 * src/lapack/lascl/lascl.cu:45-59)
 * ========================================================================= */
static int lascl_impl(double cfrom, double cto, size_t m, size_t n, double *A, size_t lda,
                      int cplx) {
  if (cfrom == 0.0 || isnan(cfrom) || isnan(cto)) return -1;
  if (m == 0 || n == 0) return 0;
  const double smlnum = DBL_MIN;
  const double bignum = 1.0 / smlnum;
  double cfromc = cfrom, ctoc = cto;
  int done = 0;
  const size_t es = cplx ? 2 : 1;
  do {
    const double cfrom1 = cfromc * smlnum;
    double mul;
    if (cfrom1 == cfromc) {
      mul = ctoc / cfromc;
      done = 1;
    } else {
      const double cto1 = ctoc / bignum;
      if (cto1 == ctoc) {
        mul = ctoc;
        done = 1;
        cfromc = 1.0;
      } else if (fabs(cfrom1) > fabs(ctoc) && ctoc != 0.0) {
        mul = smlnum;
        done = 0;
        cfromc = cfrom1;
      } else if (fabs(cto1) > fabs(cfromc)) {
        mul = bignum;
        done = 0;
        ctoc = cto1;
      } else {
        mul = ctoc / cfromc;
        done = 1;
      }
    }
    for (size_t j = 0; j < n; ++j) {
      double *col = A + es * j * lda;
      for (size_t i = 0; i < es * m; ++i) col[i] = col[i] * mul;
    }
  } while (!done);
  return 0;
}

int orc_lascl_d(double cfrom, double cto, size_t m, size_t n, double *A, size_t lda) {
  return lascl_impl(cfrom, cto, m, n, A, lda, 0);
}
int orc_lascl_z(double cfrom, double cto, size_t m, size_t n, double *A, size_t lda) {
  return lascl_impl(cfrom, cto, m, n, A, lda, 1);
}

/* =========================================================================
 * Single precision (float / complex<float>): the same reductions and the same
 * lascl chain in R = float, the reference's ComplexToRealType<T> for those
 * element types (src/expm/norm1.cppm:45-56 accumulates in R; lascl's
 * SMLNUM/BIGNUM are R's, src/lapack/lascl/interface.cppm:101-105).
 * ========================================================================= */
static inline float cabs_f(const float *z) { return hypotf(z[0], z[1]); }

static inline float nan_max_f(float a, float b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? b : a;
}

static float rc_fold_f(const float *col, size_t m, int cplx, int pre, int op) {
  float sdata[RC_BLOCK];
  const size_t nactive = MIN(m, (size_t)RC_BLOCK);
  for (size_t tid = 0; tid < nactive; ++tid) {
    float acc = 0.0f;
    for (size_t i = tid; i < m; i += RC_BLOCK) {
      float x;
      if (pre == 0) {
        x = cplx ? cabs_f(col + 2 * i) : fabsf(col[i]);
      } else if (pre == 1) {
        const float re = cplx ? col[2 * i] : col[i];
        const float im = cplx ? col[2 * i + 1] : 0.0f;
        x = re * re + im * im;
      } else {
        x = col[i];
      }
      if (i == tid) acc = x;
      else acc = (op == 0) ? acc + x : nan_max_f(acc, x);
    }
    sdata[tid] = acc;
  }
  for (size_t s = RC_BLOCK / 2; s > 0; s >>= 1) {
    for (size_t tid = 0; tid < s; ++tid) {
      if (tid + s < nactive)
        sdata[tid] = (op == 0) ? sdata[tid] + sdata[tid + s] : nan_max_f(sdata[tid], sdata[tid + s]);
    }
  }
  return sdata[0];
}

static float max_reduce_f(const float *v, int n) {
  float sdata[RC_BLOCK];
  for (int tid = 0; tid < RC_BLOCK; ++tid) {
    float best = 0.0f;
    for (int i = tid; i < n; i += RC_BLOCK) {
      const float x = v[i];
      if (x > best || isnan(x)) best = x;
    }
    sdata[tid] = best;
  }
  for (int s = RC_BLOCK / 2; s > 0; s >>= 1) {
    for (int tid = 0; tid < s; ++tid) {
      const float other = sdata[tid + s];
      if (other > sdata[tid] || isnan(other)) sdata[tid] = other;
    }
  }
  return sdata[0];
}

static float norm1_impl_f(int n, const float *A, int lda, int cplx) {
  if (n < 1) return 0.0f;
  float *colsum = (float *)malloc(sizeof(float) * (size_t)n);
  const size_t es = cplx ? 2 : 1;
  for (int j = 0; j < n; ++j) colsum[j] = rc_fold_f(A + es * (size_t)j * lda, (size_t)n, cplx, 0, 0);
  const float r = max_reduce_f(colsum, n);
  free(colsum);
  return r;
}

float orc_norm1_s(int n, const float *A, int lda) { return norm1_impl_f(n, A, lda, 0); }
float orc_norm1_c(int n, const float *A, int lda) { return norm1_impl_f(n, A, lda, 1); }

static float lange_impl_f(int which, size_t m, size_t n, const float *A, size_t lda, int cplx) {
  if (m == 0 || n == 0) return 0.0f;
  const size_t es = cplx ? 2 : 1;
  float r = 0.0f;
  if (which == 2) {
    float *rows = (float *)malloc(sizeof(float) * m);
    for (size_t i = 0; i < m; ++i) {
      float acc = 0.0f;
      for (size_t j = 0; j < n; ++j) {
        const float *e = A + es * (i + j * lda);
        acc += cplx ? cabs_f(e) : fabsf(e[0]);
      }
      rows[i] = acc;
    }
    r = rc_fold_f(rows, m, 0, 2, 1);
    free(rows);
    return r;
  }
  float *cols = (float *)malloc(sizeof(float) * n);
  for (size_t j = 0; j < n; ++j) {
    const float *col = A + es * j * lda;
    if (which == 0) cols[j] = rc_fold_f(col, m, cplx, 0, 1);
    else if (which == 1) cols[j] = rc_fold_f(col, m, cplx, 0, 0);
    else cols[j] = rc_fold_f(col, m, cplx, 1, 0);
  }
  if (which == 3) r = sqrtf(rc_fold_f(cols, n, 0, 2, 0));
  else r = rc_fold_f(cols, n, 0, 2, 1);
  free(cols);
  return r;
}

float orc_lange_s(int which, size_t m, size_t n, const float *A, size_t lda) {
  return lange_impl_f(which, m, n, A, lda, 0);
}
float orc_lange_c(int which, size_t m, size_t n, const float *A, size_t lda) {
  return lange_impl_f(which, m, n, A, lda, 1);
}

static int lascl_impl_f(float cfrom, float cto, size_t m, size_t n, float *A, size_t lda,
                        int cplx) {
  if (cfrom == 0.0f || isnan(cfrom) || isnan(cto)) return -1;
  if (m == 0 || n == 0) return 0;
  const float smlnum = FLT_MIN;
  const float bignum = 1.0f / smlnum;
  float cfromc = cfrom, ctoc = cto;
  int done = 0;
  const size_t es = cplx ? 2 : 1;
  do {
    const float cfrom1 = cfromc * smlnum;
    float mul;
    if (cfrom1 == cfromc) {
      mul = ctoc / cfromc;
      done = 1;
    } else {
      const float cto1 = ctoc / bignum;
      if (cto1 == ctoc) {
        mul = ctoc;
        done = 1;
        cfromc = 1.0f;
      } else if (fabsf(cfrom1) > fabsf(ctoc) && ctoc != 0.0f) {
        mul = smlnum;
        done = 0;
        cfromc = cfrom1;
      } else if (fabsf(cto1) > fabsf(cfromc)) {
        mul = bignum;
        done = 0;
        ctoc = cto1;
      } else {
        mul = ctoc / cfromc;
        done = 1;
      }
    }
    for (size_t j = 0; j < n; ++j) {
      float *col = A + es * j * lda;
      for (size_t i = 0; i < es * m; ++i) col[i] = col[i] * mul;
    }
  } while (!done);
  return 0;
}

int orc_lascl_s(float cfrom, float cto, size_t m, size_t n, float *A, size_t lda) {
  return lascl_impl_f(cfrom, cto, m, n, A, lda, 0);
}
int orc_lascl_c(float cfrom, float cto, size_t m, size_t n, float *A, size_t lda) {
  return lascl_impl_f(cfrom, cto, m, n, A, lda, 1);
}

/* =========================================================================
 * Deterministic GEMM. Every element of C is one running sum over k in
 * ascending order, started at zero and advanced by one fused multiply-add per
 * k, then alpha * sum (+ beta * C): the result does not depend on the thread
 * count or the blocking. Operands are packed once into k-contiguous strips
 * (8 rows of A, 6 columns of B) and multiplied by an AVX2 FMA micro-kernel
 * holding an 8 x 6 block of sums in twelve vector registers; the k range is
 * walked in blocks of GK_KC whose partial sums are kept, in double, in a
 * per-thread tile buffer, which leaves every element's FMA sequence unchanged.
 * ========================================================================= */
#include <immintrin.h>

/* aligned_alloc wants a size that is a multiple of the alignment (C11 7.22.3.1) */
static double *alloc64(size_t count) {
  const size_t bytes = (sizeof(double) * count + 63) & ~(size_t)63;
  return (double *)aligned_alloc(64, bytes ? bytes : 64);
}

#define GK_MR 8     /* rows per micro-tile (two 4-double vectors) */
#define GK_NR 6     /* columns per micro-tile */
#define GK_KC 256   /* k block */
#define GK_MS 8     /* micro row strips per thread tile (64 rows) */
#define GK_NS 16    /* micro column strips per thread tile (96 columns) */

/* acc[c*8 + r] += sum_k a[k*8 + r] * b[k*6 + c], k ascending, one FMA per k */
static void kernel_d_8x6(int kc, const double *a, const double *b, double *acc, int first) {
  __m256d c0l, c0h, c1l, c1h, c2l, c2h, c3l, c3h, c4l, c4h, c5l, c5h;
  if (first) {
    c0l = c0h = c1l = c1h = c2l = c2h = c3l = c3h = c4l = c4h = c5l = c5h = _mm256_setzero_pd();
  } else {
    c0l = _mm256_loadu_pd(acc + 0); c0h = _mm256_loadu_pd(acc + 4);
    c1l = _mm256_loadu_pd(acc + 8); c1h = _mm256_loadu_pd(acc + 12);
    c2l = _mm256_loadu_pd(acc + 16); c2h = _mm256_loadu_pd(acc + 20);
    c3l = _mm256_loadu_pd(acc + 24); c3h = _mm256_loadu_pd(acc + 28);
    c4l = _mm256_loadu_pd(acc + 32); c4h = _mm256_loadu_pd(acc + 36);
    c5l = _mm256_loadu_pd(acc + 40); c5h = _mm256_loadu_pd(acc + 44);
  }
  for (int k = 0; k < kc; ++k) {
    const __m256d al = _mm256_loadu_pd(a + 8 * k), ah = _mm256_loadu_pd(a + 8 * k + 4);
    const double *bk = b + 6 * k;
    __m256d bb = _mm256_broadcast_sd(bk + 0);
    c0l = _mm256_fmadd_pd(al, bb, c0l); c0h = _mm256_fmadd_pd(ah, bb, c0h);
    bb = _mm256_broadcast_sd(bk + 1);
    c1l = _mm256_fmadd_pd(al, bb, c1l); c1h = _mm256_fmadd_pd(ah, bb, c1h);
    bb = _mm256_broadcast_sd(bk + 2);
    c2l = _mm256_fmadd_pd(al, bb, c2l); c2h = _mm256_fmadd_pd(ah, bb, c2h);
    bb = _mm256_broadcast_sd(bk + 3);
    c3l = _mm256_fmadd_pd(al, bb, c3l); c3h = _mm256_fmadd_pd(ah, bb, c3h);
    bb = _mm256_broadcast_sd(bk + 4);
    c4l = _mm256_fmadd_pd(al, bb, c4l); c4h = _mm256_fmadd_pd(ah, bb, c4h);
    bb = _mm256_broadcast_sd(bk + 5);
    c5l = _mm256_fmadd_pd(al, bb, c5l); c5h = _mm256_fmadd_pd(ah, bb, c5h);
  }
  _mm256_storeu_pd(acc + 0, c0l); _mm256_storeu_pd(acc + 4, c0h);
  _mm256_storeu_pd(acc + 8, c1l); _mm256_storeu_pd(acc + 12, c1h);
  _mm256_storeu_pd(acc + 16, c2l); _mm256_storeu_pd(acc + 20, c2h);
  _mm256_storeu_pd(acc + 24, c3l); _mm256_storeu_pd(acc + 28, c3h);
  _mm256_storeu_pd(acc + 32, c4l); _mm256_storeu_pd(acc + 36, c4h);
  _mm256_storeu_pd(acc + 40, c5l); _mm256_storeu_pd(acc + 44, c5h);
}

/* C (m x n) = alpha * A (m x k) * B (k x n) + beta * C, column-major */
static void gemm_d(int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                   int ldb, double beta, double *C, int ldc) {
  if (m < 1 || n < 1) return;
  const int ms = (m + GK_MR - 1) / GK_MR, ns = (n + GK_NR - 1) / GK_NR;
  const int kk = k < 1 ? 1 : k;
  double *Ap = alloc64((size_t)ms * GK_MR * kk);
  double *Bp = alloc64((size_t)ns * GK_NR * kk);
#pragma omp parallel
  {
#pragma omp for schedule(static)
    for (int si = 0; si < ms; ++si) {
      double *dst = Ap + (size_t)si * GK_MR * kk;
      const int i0 = si * GK_MR;
      for (int p = 0; p < k; ++p) {
        const double *src = A + (size_t)p * lda + i0;
        for (int r = 0; r < GK_MR; ++r) dst[(size_t)p * GK_MR + r] = (i0 + r < m) ? src[r] : 0.0;
      }
    }
#pragma omp for schedule(static)
    for (int sj = 0; sj < ns; ++sj) {
      double *dst = Bp + (size_t)sj * GK_NR * kk;
      const int j0 = sj * GK_NR;
      for (int c = 0; c < GK_NR; ++c) {
        const double *src = B + (size_t)(j0 + c) * ldb;
        for (int p = 0; p < k; ++p) dst[(size_t)p * GK_NR + c] = (j0 + c < n) ? src[p] : 0.0;
      }
    }
    double *tile = alloc64(GK_MS * GK_NS * GK_MR * GK_NR);
    const int tm = (ms + GK_MS - 1) / GK_MS, tn = (ns + GK_NS - 1) / GK_NS;
#pragma omp for collapse(2) schedule(dynamic, 1)
    for (int ti = 0; ti < tm; ++ti) {
      for (int tj = 0; tj < tn; ++tj) {
        const int si0 = ti * GK_MS, sj0 = tj * GK_NS;
        const int nsi = MIN(GK_MS, ms - si0), nsj = MIN(GK_NS, ns - sj0);
        if (k < 1) {
          memset(tile, 0, sizeof(double) * GK_MS * GK_NS * GK_MR * GK_NR);
        }
        for (int p0 = 0; p0 < k; p0 += GK_KC) {
          const int kc = MIN(GK_KC, k - p0);
          for (int b = 0; b < nsj; ++b) {
            const double *bp = Bp + (size_t)(sj0 + b) * GK_NR * kk + (size_t)p0 * GK_NR;
            for (int a = 0; a < nsi; ++a) {
              const double *ap = Ap + (size_t)(si0 + a) * GK_MR * kk + (size_t)p0 * GK_MR;
              kernel_d_8x6(kc, ap, bp, tile + (size_t)(b * GK_MS + a) * GK_MR * GK_NR, p0 == 0);
            }
          }
        }
        for (int b = 0; b < nsj; ++b) {
          for (int c = 0; c < GK_NR; ++c) {
            const int j = (sj0 + b) * GK_NR + c;
            if (j >= n) break;
            for (int a = 0; a < nsi; ++a) {
              const double *acc = tile + (size_t)(b * GK_MS + a) * GK_MR * GK_NR + c * GK_MR;
              const int i0 = (si0 + a) * GK_MR;
              double *cc = C + (size_t)j * ldc + i0;
              const int rr = MIN(GK_MR, m - i0);
              for (int r = 0; r < rr; ++r) {
                const double v = alpha * acc[r];
                cc[r] = (beta == 0.0) ? v : fma(beta, cc[r], v);
              }
            }
          }
        }
      }
    }
    free(tile);
  }
  free(Ap);
  free(Bp);
}

void orc_gemm_d(int n, double alpha, const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc) {
  gemm_d(n, n, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* complex, interleaved: per k, re += ar*br then re -= ai*bi, im += ar*bi then
 * im += ai*br (four fused multiply-adds), k ascending from zero; strips of 4
 * rows (re and im planes) and 6 columns */
#define GZ_MR 4
#define GZ_NR 6
#define GZ_MS 16 /* 64 rows per thread tile */
#define GZ_NS 8  /* 48 columns per thread tile */

/* a: per k {re[4], im[4]}; b: per k {re[6], im[6]}; acc: [c][re4, im4] */
static void kernel_z_4x6(int kc, const double *a, const double *b, double *acc, int first) {
  __m256d r[6], i[6];
  for (int c = 0; c < 6; ++c) {
    r[c] = first ? _mm256_setzero_pd() : _mm256_loadu_pd(acc + 8 * c);
    i[c] = first ? _mm256_setzero_pd() : _mm256_loadu_pd(acc + 8 * c + 4);
  }
  for (int k = 0; k < kc; ++k) {
    const __m256d ar = _mm256_loadu_pd(a + 8 * k), ai = _mm256_loadu_pd(a + 8 * k + 4);
    const double *bk = b + 12 * k;
#pragma GCC unroll 6
    for (int c = 0; c < 6; ++c) {
      const __m256d br = _mm256_broadcast_sd(bk + c), bi = _mm256_broadcast_sd(bk + 6 + c);
      r[c] = _mm256_fmadd_pd(ar, br, r[c]);
      r[c] = _mm256_fnmadd_pd(ai, bi, r[c]);
      i[c] = _mm256_fmadd_pd(ar, bi, i[c]);
      i[c] = _mm256_fmadd_pd(ai, br, i[c]);
    }
  }
  for (int c = 0; c < 6; ++c) {
    _mm256_storeu_pd(acc + 8 * c, r[c]);
    _mm256_storeu_pd(acc + 8 * c + 4, i[c]);
  }
}

static void gemm_z(int m, int n, int k, const double *alpha, const double *A, int lda,
                   const double *B, int ldb, const double *beta, double *C, int ldc) {
  if (m < 1 || n < 1) return;
  const int ms = (m + GZ_MR - 1) / GZ_MR, ns = (n + GZ_NR - 1) / GZ_NR;
  const int kk = k < 1 ? 1 : k;
  const int beta_zero = (beta[0] == 0.0 && beta[1] == 0.0);
  double *Ap = alloc64((size_t)ms * 2 * GZ_MR * kk);
  double *Bp = alloc64((size_t)ns * 2 * GZ_NR * kk);
#pragma omp parallel
  {
#pragma omp for schedule(static)
    for (int si = 0; si < ms; ++si) {
      double *dst = Ap + (size_t)si * 2 * GZ_MR * kk;
      const int i0 = si * GZ_MR;
      for (int p = 0; p < k; ++p) {
        const double *src = A + 2 * ((size_t)p * lda + i0);
        for (int q = 0; q < GZ_MR; ++q) {
          const int ok = i0 + q < m;
          dst[(size_t)p * 8 + q] = ok ? src[2 * q] : 0.0;
          dst[(size_t)p * 8 + 4 + q] = ok ? src[2 * q + 1] : 0.0;
        }
      }
    }
#pragma omp for schedule(static)
    for (int sj = 0; sj < ns; ++sj) {
      double *dst = Bp + (size_t)sj * 2 * GZ_NR * kk;
      const int j0 = sj * GZ_NR;
      for (int c = 0; c < GZ_NR; ++c) {
        const double *src = B + 2 * (size_t)(j0 + c) * ldb;
        const int ok = j0 + c < n;
        for (int p = 0; p < k; ++p) {
          dst[(size_t)p * 12 + c] = ok ? src[2 * p] : 0.0;
          dst[(size_t)p * 12 + 6 + c] = ok ? src[2 * p + 1] : 0.0;
        }
      }
    }
    const size_t tile_elems = (size_t)GZ_MS * GZ_NS * 2 * GZ_MR * GZ_NR;
    double *tile = alloc64(tile_elems);
    const int tm = (ms + GZ_MS - 1) / GZ_MS, tn = (ns + GZ_NS - 1) / GZ_NS;
#pragma omp for collapse(2) schedule(dynamic, 1)
    for (int ti = 0; ti < tm; ++ti) {
      for (int tj = 0; tj < tn; ++tj) {
        const int si0 = ti * GZ_MS, sj0 = tj * GZ_NS;
        const int nsi = MIN(GZ_MS, ms - si0), nsj = MIN(GZ_NS, ns - sj0);
        if (k < 1) memset(tile, 0, sizeof(double) * tile_elems);
        for (int p0 = 0; p0 < k; p0 += GK_KC) {
          const int kc = MIN(GK_KC, k - p0);
          for (int b = 0; b < nsj; ++b) {
            const double *bp = Bp + (size_t)(sj0 + b) * 2 * GZ_NR * kk + (size_t)p0 * 12;
            for (int a = 0; a < nsi; ++a) {
              const double *ap = Ap + (size_t)(si0 + a) * 2 * GZ_MR * kk + (size_t)p0 * 8;
              kernel_z_4x6(kc, ap, bp, tile + (size_t)(b * GZ_MS + a) * 2 * GZ_MR * GZ_NR,
                           p0 == 0);
            }
          }
        }
        for (int b = 0; b < nsj; ++b) {
          for (int c = 0; c < GZ_NR; ++c) {
            const int j = (sj0 + b) * GZ_NR + c;
            if (j >= n) break;
            for (int a = 0; a < nsi; ++a) {
              const double *acc = tile + (size_t)(b * GZ_MS + a) * 2 * GZ_MR * GZ_NR + 8 * c;
              const int i0 = (si0 + a) * GZ_MR;
              double *cc = C + 2 * ((size_t)j * ldc + i0);
              const int rr = MIN(GZ_MR, m - i0);
              for (int q = 0; q < rr; ++q) {
                const double sr = acc[q], si = acc[4 + q];
                double vr = alpha[0] * sr - alpha[1] * si;
                double vi = alpha[0] * si + alpha[1] * sr;
                if (!beta_zero) {
                  const double cr = cc[2 * q], ci = cc[2 * q + 1];
                  vr = vr + (beta[0] * cr - beta[1] * ci);
                  vi = vi + (beta[0] * ci + beta[1] * cr);
                }
                cc[2 * q] = vr;
                cc[2 * q + 1] = vi;
              }
            }
          }
        }
      }
    }
    free(tile);
  }
  free(Ap);
  free(Bp);
}

void orc_gemm_z(int n, const double *alpha, const double *A, int lda, const double *B, int ldb,
                const double *beta, double *C, int ldc) {
  gemm_z(n, n, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* =========================================================================
 * LU: LAPACK ?getrf's right-looking blocked algorithm (panels of LU_NB
 * columns factored by ?getf2, row interchanges applied across the whole
 * matrix, U12 = L11^-1 A12, A22 -= L21 U12 on the GEMM above). ?getf2's
 * rules: pivot = first index of the largest |.| (|re|+|im| for complex, as
 * izamax), the column below the pivot scaled by its reciprocal (divided when
 * the pivot is below the safe minimum), rank-1 update of the panel.
 * ========================================================================= */
#define LU_NB 64

/* the j-th step of ?getf2 on rows j..m-1 of columns c0..c1-1 (all inside one
 * panel); returns 1 for an exactly zero pivot */
static int getf2_panel_d(int m, int c0, int c1, double *A, int lda, int *ipiv) {
  int zero = 0;
  const double sfmin = DBL_MIN;
  for (int j = c0; j < c1; ++j) {
    double *cj = A + (size_t)j * lda;
    int jp = j;
    double best = fabs(cj[j]);
    for (int i = j + 1; i < m; ++i) {
      const double v = fabs(cj[i]);
      if (v > best) { best = v; jp = i; }
    }
    ipiv[j] = jp + 1;
    if (cj[jp] != 0.0) {
      if (jp != j) {
        for (int c = c0; c < c1; ++c) {
          double *col = A + (size_t)c * lda;
          const double t = col[j]; col[j] = col[jp]; col[jp] = t;
        }
      }
      if (fabs(cj[j]) >= sfmin) {
        const double r = 1.0 / cj[j];
        for (int i = j + 1; i < m; ++i) cj[i] *= r;
      } else {
        for (int i = j + 1; i < m; ++i) cj[i] /= cj[j];
      }
    } else if (!zero) {
      zero = j + 1;
    }
#pragma omp parallel for schedule(static) if ((m - j) * (c1 - j) > 65536)
    for (int c = j + 1; c < c1; ++c) {
      double *col = A + (size_t)c * lda;
      const double u = col[j];
      if (u != 0.0) {
        for (int i = j + 1; i < m; ++i) col[i] -= cj[i] * u;
      }
    }
  }
  return zero;
}

/* rows j0..j1-1 of ipiv applied to columns [c0, c1) */
static void swap_rows_d(double *A, int lda, int c0, int c1, int j0, int j1, const int *ipiv) {
#pragma omp parallel for schedule(static) if ((c1 - c0) > 64)
  for (int c = c0; c < c1; ++c) {
    double *col = A + (size_t)c * lda;
    for (int j = j0; j < j1; ++j) {
      const int p = ipiv[j] - 1;
      if (p != j) { const double t = col[j]; col[j] = col[p]; col[p] = t; }
    }
  }
}

int orc_getrf_d(int n, double *A, int lda, int *ipiv) {
  int info = 0;
  for (int j0 = 0; j0 < n; j0 += LU_NB) {
    const int jb = MIN(LU_NB, n - j0);
    const int z = getf2_panel_d(n, j0, j0 + jb, A, lda, ipiv);
    if (z && !info) info = z;
    swap_rows_d(A, lda, 0, j0, j0, j0 + jb, ipiv);
    if (j0 + jb < n) {
      swap_rows_d(A, lda, j0 + jb, n, j0, j0 + jb, ipiv);
      /* U12 = L11^-1 A12 (unit lower), column by column */
#pragma omp parallel for schedule(static)
      for (int c = j0 + jb; c < n; ++c) {
        double *col = A + (size_t)c * lda;
        for (int p = j0; p < j0 + jb; ++p) {
          const double u = col[p];
          if (u != 0.0) {
            const double *l = A + (size_t)p * lda;
            for (int i = p + 1; i < j0 + jb; ++i) col[i] -= l[i] * u;
          }
        }
      }
      const int rest = n - j0 - jb;
      gemm_d(rest, rest, jb, -1.0, A + (size_t)j0 * lda + j0 + jb, lda,
             A + (size_t)(j0 + jb) * lda + j0, lda, 1.0, A + (size_t)(j0 + jb) * lda + j0 + jb,
             lda);
    }
  }
  return info;
}

static int getf2_panel_z(int m, int c0, int c1, double *A, int lda, int *ipiv) {
  int zero = 0;
  const double sfmin = DBL_MIN;
  for (int j = c0; j < c1; ++j) {
    double *cj = A + 2 * (size_t)j * lda;
    int jp = j;
    double best = fabs(cj[2 * j]) + fabs(cj[2 * j + 1]);
    for (int i = j + 1; i < m; ++i) {
      const double v = fabs(cj[2 * i]) + fabs(cj[2 * i + 1]);
      if (v > best) { best = v; jp = i; }
    }
    ipiv[j] = jp + 1;
    if (cj[2 * jp] != 0.0 || cj[2 * jp + 1] != 0.0) {
      if (jp != j) {
        for (int c = c0; c < c1; ++c) {
          double *col = A + 2 * (size_t)c * lda;
          for (int q = 0; q < 2; ++q) {
            const double t = col[2 * j + q]; col[2 * j + q] = col[2 * jp + q]; col[2 * jp + q] = t;
          }
        }
      }
      const double pr = cj[2 * j], pi = cj[2 * j + 1];
      if (hypot(pr, pi) >= sfmin) {
        /* reciprocal 1/(pr + i pi) by Smith's method */
        double rr, ri;
        if (fabs(pr) >= fabs(pi)) {
          const double t = pi / pr, d = pr + pi * t;
          rr = 1.0 / d; ri = -t / d;
        } else {
          const double t = pr / pi, d = pi + pr * t;
          rr = t / d; ri = -1.0 / d;
        }
        for (int i = j + 1; i < m; ++i) {
          const double xr = cj[2 * i], xi = cj[2 * i + 1];
          cj[2 * i] = xr * rr - xi * ri;
          cj[2 * i + 1] = xr * ri + xi * rr;
        }
      } else {
        for (int i = j + 1; i < m; ++i) {
          const double xr = cj[2 * i], xi = cj[2 * i + 1];
          const double d = pr * pr + pi * pi;
          cj[2 * i] = (xr * pr + xi * pi) / d;
          cj[2 * i + 1] = (xi * pr - xr * pi) / d;
        }
      }
    } else if (!zero) {
      zero = j + 1;
    }
#pragma omp parallel for schedule(static) if ((m - j) * (c1 - j) > 32768)
    for (int c = j + 1; c < c1; ++c) {
      double *col = A + 2 * (size_t)c * lda;
      const double ur = col[2 * j], ui = col[2 * j + 1];
      if (ur != 0.0 || ui != 0.0) {
        for (int i = j + 1; i < m; ++i) {
          const double lr = cj[2 * i], li = cj[2 * i + 1];
          col[2 * i] -= lr * ur - li * ui;
          col[2 * i + 1] -= lr * ui + li * ur;
        }
      }
    }
  }
  return zero;
}

static void swap_rows_z(double *A, int lda, int c0, int c1, int j0, int j1, const int *ipiv) {
#pragma omp parallel for schedule(static) if ((c1 - c0) > 64)
  for (int c = c0; c < c1; ++c) {
    double *col = A + 2 * (size_t)c * lda;
    for (int j = j0; j < j1; ++j) {
      const int p = ipiv[j] - 1;
      if (p != j) {
        for (int q = 0; q < 2; ++q) {
          const double t = col[2 * j + q]; col[2 * j + q] = col[2 * p + q]; col[2 * p + q] = t;
        }
      }
    }
  }
}

int orc_getrf_z(int n, double *A, int lda, int *ipiv) {
  int info = 0;
  const double mone[2] = {-1.0, 0.0}, one[2] = {1.0, 0.0};
  for (int j0 = 0; j0 < n; j0 += LU_NB) {
    const int jb = MIN(LU_NB, n - j0);
    const int z = getf2_panel_z(n, j0, j0 + jb, A, lda, ipiv);
    if (z && !info) info = z;
    swap_rows_z(A, lda, 0, j0, j0, j0 + jb, ipiv);
    if (j0 + jb < n) {
      swap_rows_z(A, lda, j0 + jb, n, j0, j0 + jb, ipiv);
#pragma omp parallel for schedule(static)
      for (int c = j0 + jb; c < n; ++c) {
        double *col = A + 2 * (size_t)c * lda;
        for (int p = j0; p < j0 + jb; ++p) {
          const double ur = col[2 * p], ui = col[2 * p + 1];
          if (ur != 0.0 || ui != 0.0) {
            const double *l = A + 2 * (size_t)p * lda;
            for (int i = p + 1; i < j0 + jb; ++i) {
              const double lr = l[2 * i], li = l[2 * i + 1];
              col[2 * i] -= lr * ur - li * ui;
              col[2 * i + 1] -= lr * ui + li * ur;
            }
          }
        }
      }
      const int rest = n - j0 - jb;
      gemm_z(rest, rest, jb, mone, A + 2 * ((size_t)j0 * lda + j0 + jb), lda,
             A + 2 * ((size_t)(j0 + jb) * lda + j0), lda, one,
             A + 2 * ((size_t)(j0 + jb) * lda + j0 + jb), lda);
    }
  }
  return info;
}

/* ?getrs NoTrans: row interchanges, then the unit-lower and the upper solve
 * blocked as ?trsm (Left, NoTrans) is: a diagonal block solved column by
 * column, the rest of the right-hand sides updated by one GEMM per block. */
int orc_getrs_d(int n, int nrhs, const double *LU, int lda, const int *ipiv, double *B, int ldb) {
  swap_rows_d(B, ldb, 0, nrhs, 0, n, ipiv);
  for (int k0 = 0; k0 < n; k0 += LU_NB) {
    const int kb = MIN(LU_NB, n - k0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nrhs; ++c) {
      double *b = B + (size_t)c * ldb;
      for (int k = k0; k < k0 + kb; ++k) {
        const double bk = b[k];
        if (bk != 0.0) {
          const double *l = LU + (size_t)k * lda;
          for (int i = k + 1; i < k0 + kb; ++i) b[i] -= bk * l[i];
        }
      }
    }
    if (k0 + kb < n)
      gemm_d(n - k0 - kb, nrhs, kb, -1.0, LU + (size_t)k0 * lda + k0 + kb, lda, B + k0, ldb, 1.0,
             B + k0 + kb, ldb);
  }
  const int nblk = (n + LU_NB - 1) / LU_NB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int k0 = bi * LU_NB, kb = MIN(LU_NB, n - k0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nrhs; ++c) {
      double *b = B + (size_t)c * ldb;
      for (int k = k0 + kb - 1; k >= k0; --k) {
        if (b[k] != 0.0) {
          const double *u = LU + (size_t)k * lda;
          b[k] = b[k] / u[k];
          const double bk = b[k];
          for (int i = k0; i < k; ++i) b[i] -= bk * u[i];
        }
      }
    }
    if (k0 > 0)
      gemm_d(k0, nrhs, kb, -1.0, LU + (size_t)k0 * lda, lda, B + k0, ldb, 1.0, B, ldb);
  }
  return 0;
}

int orc_getrs_z(int n, int nrhs, const double *LU, int lda, const int *ipiv, double *B, int ldb) {
  const double mone[2] = {-1.0, 0.0}, one[2] = {1.0, 0.0};
  swap_rows_z(B, ldb, 0, nrhs, 0, n, ipiv);
  for (int k0 = 0; k0 < n; k0 += LU_NB) {
    const int kb = MIN(LU_NB, n - k0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nrhs; ++c) {
      double *b = B + 2 * (size_t)c * ldb;
      for (int k = k0; k < k0 + kb; ++k) {
        const double br = b[2 * k], bi = b[2 * k + 1];
        if (br != 0.0 || bi != 0.0) {
          const double *l = LU + 2 * (size_t)k * lda;
          for (int i = k + 1; i < k0 + kb; ++i) {
            b[2 * i] -= br * l[2 * i] - bi * l[2 * i + 1];
            b[2 * i + 1] -= br * l[2 * i + 1] + bi * l[2 * i];
          }
        }
      }
    }
    if (k0 + kb < n)
      gemm_z(n - k0 - kb, nrhs, kb, mone, LU + 2 * ((size_t)k0 * lda + k0 + kb), lda,
             B + 2 * (size_t)k0, ldb, one, B + 2 * (size_t)(k0 + kb), ldb);
  }
  const int nblk = (n + LU_NB - 1) / LU_NB;
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int k0 = blk * LU_NB, kb = MIN(LU_NB, n - k0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nrhs; ++c) {
      double *b = B + 2 * (size_t)c * ldb;
      for (int k = k0 + kb - 1; k >= k0; --k) {
        if (b[2 * k] != 0.0 || b[2 * k + 1] != 0.0) {
          const double *u = LU + 2 * (size_t)k * lda;
          /* b[k] /= u[k] by Smith's method */
          const double ur = u[2 * k], ui = u[2 * k + 1], xr = b[2 * k], xi = b[2 * k + 1];
          double qr, qi;
          if (fabs(ur) >= fabs(ui)) {
            const double t = ui / ur, d = ur + ui * t;
            qr = (xr + xi * t) / d; qi = (xi - xr * t) / d;
          } else {
            const double t = ur / ui, d = ui + ur * t;
            qr = (xr * t + xi) / d; qi = (xi * t - xr) / d;
          }
          b[2 * k] = qr; b[2 * k + 1] = qi;
          for (int i = k0; i < k; ++i) {
            b[2 * i] -= qr * u[2 * i] - qi * u[2 * i + 1];
            b[2 * i + 1] -= qr * u[2 * i + 1] + qi * u[2 * i];
          }
        }
      }
    }
    if (k0 > 0)
      gemm_z(k0, nrhs, kb, mone, LU + 2 * (size_t)k0 * lda, lda, B + 2 * (size_t)k0, ldb, one, B,
             ldb);
  }
  return 0;
}

/* =========================================================================
 * Pade approximant (src/expm/pade.cppm:77-191)
 * ========================================================================= */

/* pade_even_odd (src/expm/expm.cu:91-109): v = sum_k cv[k+1]*P_k (k ascending,
 * acc + c*x), then + cv[0] on the diagonal; packed n x n. */
static void even_odd(int n, int np, double *const *P, const double *cv, const double *cw,
                     double *V, double *W, int cplx) {
  const size_t nn = (size_t)n * n;
  const size_t es = cplx ? 2 : 1;
#pragma omp parallel for schedule(static)
  for (size_t idx = 0; idx < nn; ++idx) {
    for (size_t q = 0; q < es; ++q) {
      double v = 0.0, w = 0.0;
      for (int k = 0; k < np; ++k) {
        const double x = P[k][es * idx + q];
        v = v + x * cv[k + 1];
        w = w + x * cw[k + 1];
      }
      if (q == 0 && idx % (size_t)(n + 1) == 0) {
        v = v + cv[0];
        w = w + cw[0];
      }
      V[es * idx + q] = v;
      W[es * idx + q] = w;
    }
  }
}

static int pade_impl(int m, int n, const double *A, int lda, double *R, int ldr, int *info,
                     int cplx) {
  const double *b = orc_pade_coeffs(m);
  if (b == NULL || n < 1 || n > ORC_MAX_DIM || lda < n || ldr < n) return -1;
  const size_t es = cplx ? 2 : 1;
  const size_t nn = (size_t)n * n;
  const int np = orc_pade_num_powers(m);
  double *blk[6];
  for (int i = 0; i < 6; ++i) blk[i] = (double *)malloc(sizeof(double) * es * nn);
  double *P[4] = {blk[0], blk[1], blk[2], blk[3]};
  double *X = blk[3], *V = blk[4], *W = blk[5];
  int *ipiv = (int *)malloc(sizeof(int) * (size_t)n);
  const double one_z[2] = {1.0, 0.0}, zero_z[2] = {0.0, 0.0};

#define MM(X_, ldx, Y_, ldy, Z_, ldz, acc)                                                      \
  do {                                                                                          \
    if (cplx) orc_gemm_z(n, one_z, X_, ldx, Y_, ldy, (acc) ? one_z : zero_z, Z_, ldz);          \
    else orc_gemm_d(n, 1.0, X_, ldx, Y_, ldy, (acc) ? 1.0 : 0.0, Z_, ldz);                       \
  } while (0)

  MM(A, lda, A, lda, P[0], n, 0);
  if (np >= 2) MM(P[0], n, P[0], n, P[1], n, 0);
  if (np >= 3) MM(P[1], n, P[0], n, P[2], n, 0);
  if (np >= 4) MM(P[1], n, P[1], n, P[3], n, 0);

  double cv[5] = {0}, cw[5] = {0};
  if (m != 13) {
    for (int i = 0; i <= np; ++i) { cv[i] = b[2 * i]; cw[i] = b[2 * i + 1]; }
    even_odd(n, np, P, cv, cw, V, W, cplx);
  } else {
    cv[0] = 0.0; cv[1] = b[9]; cv[2] = b[11]; cv[3] = b[13];
    cw[0] = b[1]; cw[1] = b[3]; cw[2] = b[5]; cw[3] = b[7];
    even_odd(n, 3, P, cv, cw, X, W, cplx);
    MM(P[2], n, X, n, W, n, 1);
    cv[0] = 0.0; cv[1] = b[8]; cv[2] = b[10]; cv[3] = b[12];
    cw[0] = b[0]; cw[1] = b[2]; cw[2] = b[4]; cw[3] = b[6];
    even_odd(n, 3, P, cv, cw, X, V, cplx);
    MM(P[2], n, X, n, V, n, 1);
  }
  double *U = P[0];
  MM(A, lda, W, n, U, n, 0);

  /* pade_split (src/expm/expm.cu:150-161): R = V + U (ldr), Q = V - U over V */
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) {
      const size_t idx = (size_t)i + (size_t)j * n;
      for (size_t q = 0; q < es; ++q) {
        const double u = U[es * idx + q], v = V[es * idx + q];
        R[es * ((size_t)i + (size_t)j * ldr) + q] = v + u;
        V[es * idx + q] = v - u;
      }
    }
  }
  if (cplx) {
    info[0] = orc_getrf_z(n, V, n, ipiv);
    info[1] = orc_getrs_z(n, n, V, n, ipiv, R, ldr);
  } else {
    info[0] = orc_getrf_d(n, V, n, ipiv);
    info[1] = orc_getrs_d(n, n, V, n, ipiv, R, ldr);
  }
#undef MM
  for (int i = 0; i < 6; ++i) free(blk[i]);
  free(ipiv);
  return 0;
}

int orc_pade_d(int m, int n, const double *A, int lda, double *R, int ldr, int *info) {
  return pade_impl(m, n, A, lda, R, ldr, info, 0);
}
int orc_pade_z(int m, int n, const double *A, int lda, double *R, int ldr, int *info) {
  return pade_impl(m, n, A, lda, R, ldr, info, 1);
}

/* squaring loop and odd-s copy (src/expm/expm.cppm:155-171) */
static void square_s_times(int n, int s, double *E, int lde, int cplx) {
  const size_t es = cplx ? 2 : 1;
  double *sq = (double *)malloc(sizeof(double) * es * (size_t)n * n);
  const double one_z[2] = {1.0, 0.0}, zero_z[2] = {0.0, 0.0};
  double *cur = E, *other = sq;
  int ld_cur = lde, ld_other = n;
  for (int k = 0; k < s; ++k) {
    if (cplx) orc_gemm_z(n, one_z, cur, ld_cur, cur, ld_cur, zero_z, other, ld_other);
    else orc_gemm_d(n, 1.0, cur, ld_cur, cur, ld_cur, 0.0, other, ld_other);
    double *t = cur; cur = other; other = t;
    const int tl = ld_cur; ld_cur = ld_other; ld_other = tl;
  }
  if (cur != E) {
    for (int j = 0; j < n; ++j)
      memcpy(E + es * (size_t)j * lde, cur + es * (size_t)j * ld_cur, sizeof(double) * es * n);
  }
  free(sq);
}

/* scaled copy As = 2^-s A (the geam of src/expm/expm.cppm:138-144; a power of
 * two scale is exact), packed with leading dimension n */
static double *scaled_copy(int n, const double *A, int lda, int s, int cplx) {
  const size_t es = cplx ? 2 : 1;
  double *As = (double *)malloc(sizeof(double) * es * (size_t)n * n);
  const double f = ldexp(1.0, -s);
  for (int j = 0; j < n; ++j)
    for (size_t i = 0; i < es * (size_t)n; ++i)
      As[es * (size_t)j * n + i] = f * A[es * (size_t)j * lda + i];
  return As;
}

static int expm_impl(int n, const double *A, int lda, double *E, int lde, int *plan, int cplx) {
  if (n < 1 || n > ORC_MAX_DIM || lda < n || lde < n || A == NULL || E == NULL) return -1;
  const double norm = cplx ? orc_norm1_z(n, A, lda) : orc_norm1_d(n, A, lda);
  if (!isfinite(norm)) return -2;
  int p[3];
  orc_expm_plan_d(norm, p);
  if (plan) { plan[0] = p[0]; plan[1] = p[1]; plan[2] = p[2]; }
  const double *src = A;
  int ld_src = lda;
  double *As = NULL;
  if (p[1] > 0) {
    As = scaled_copy(n, A, lda, p[1], cplx);
    src = As;
    ld_src = n;
  }
  int info[2];
  pade_impl(p[0], n, src, ld_src, E, lde, info, cplx);
  free(As);
  square_s_times(n, p[1], E, lde, cplx);
  return 0;
}

int orc_expm_d(int n, const double *A, int lda, double *E, int lde, int *plan) {
  return expm_impl(n, A, lda, E, lde, plan, 0);
}
int orc_expm_z(int n, const double *A, int lda, double *E, int lde, int *plan) {
  return expm_impl(n, A, lda, E, lde, plan, 1);
}

/* =========================================================================
 * Paterson-Stockmeyer (src/paterson_stockmeyer/interface.cppm:117-210, 329-422)
 * ========================================================================= */
int orc_ps_cost(int degree, int s) {
  if (s < 1 || degree < 1) return 0;
  const int r = degree / s;
  const int max_power = (r >= 1) ? s : degree;
  const int num_powers = (max_power >= 2) ? (max_power - 1) : 0;
  return num_powers + r;
}

int orc_ps_block_size(int degree) {
  if (degree < 1) return 1;
  int best_s = 1, best_cost = orc_ps_cost(degree, 1);
  for (int s = 2; s <= degree + 1; ++s) {
    const int c = orc_ps_cost(degree, s);
    if (c < best_cost) { best_cost = c; best_s = s; }
  }
  return best_s;
}

void orc_ps_plan(int degree, int s_requested, int *out) {
  int s = 1, r = 0, num_powers = 0, num_blocks = 0, num_gemms = 0;
  if (degree >= 0) {
    s = (s_requested > 0) ? s_requested : orc_ps_block_size(degree);
    if (s > degree + 1) s = degree + 1;
    if (s < 1) s = 1;
    r = degree / s;
    const int max_power = (r >= 1) ? s : degree;
    num_powers = (max_power >= 2) ? (max_power - 1) : 0;
    num_blocks = num_powers + ((r >= 1) ? 1 : 0);
    num_gemms = num_powers + r;
  }
  out[0] = s; out[1] = r; out[2] = num_powers; out[3] = num_blocks; out[4] = num_gemms;
}

/* build_block (src/paterson_stockmeyer/paterson_stockmeyer.cu:64-84):
 * acc = c0 on the diagonal; acc = c1*A + acc; acc = c_k*A^k + acc. */
static void build_block(int n, double *Bk, int ldb, const double *A, int lda,
                        double *const *pw, const double *c, int terms, int cplx) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) {
      if (!cplx) {
        double acc = (i == j) ? c[0] : 0.0;
        if (terms > 1) acc = c[1] * A[i + (size_t)j * lda] + acc;
        for (int k = 2; k < terms; ++k) acc = c[k] * pw[k - 2][i + (size_t)j * n] + acc;
        Bk[i + (size_t)j * ldb] = acc;
      } else {
        double ar = (i == j) ? c[0] : 0.0, ai = (i == j) ? c[1] : 0.0;
        for (int k = 1; k < terms; ++k) {
          const double *x = (k == 1) ? A + 2 * (i + (size_t)j * lda)
                                     : pw[k - 2] + 2 * (i + (size_t)j * n);
          const double cr = c[2 * k], ci = c[2 * k + 1];
          ar = (cr * x[0] - ci * x[1]) + ar;
          ai = (cr * x[1] + ci * x[0]) + ai;
        }
        Bk[2 * (i + (size_t)j * ldb)] = ar;
        Bk[2 * (i + (size_t)j * ldb) + 1] = ai;
      }
    }
  }
}

/* Paterson-Stockmeyer; the first `have` powers (A^2 .. A^(have+1)) may be
 * given in pre (n x n, leading dimension n), the rest are formed here. */
static int ps_impl_have(int n, const double *coeffs, int degree, const double *A, int lda,
                        double *P, int ldp, int s_requested, int cplx, double *const *pre,
                        int have) {
  if (n < 1 || degree < 0 || lda < n || ldp < n || s_requested < 0) return -1;
  if (coeffs == NULL || A == NULL || P == NULL) return -1;
  int pl[5];
  orc_ps_plan(degree, s_requested, pl);
  const int s = pl[0], r = pl[1], num_powers = pl[2];
  const size_t es = cplx ? 2 : 1;
  const size_t nn = (size_t)n * n;
  const int npw = num_powers > have ? num_powers : have;
  double **pw = (double **)calloc((size_t)(npw > 0 ? npw : 1), sizeof(double *));
  int *owned = (int *)calloc((size_t)(npw > 0 ? npw : 1), sizeof(int));
  for (int k = 0; k < npw; ++k) {
    if (k < have && pre) {
      pw[k] = pre[k];
    } else {
      pw[k] = (double *)malloc(sizeof(double) * es * nn);
      owned[k] = 1;
    }
  }
  double *scratch = (r >= 1) ? (double *)malloc(sizeof(double) * es * nn) : NULL;
  const double one_z[2] = {1.0, 0.0}, zero_z[2] = {0.0, 0.0};

  for (int k = 2 + (pre ? have : 0); k <= num_powers + 1; ++k) {
    const double *left = (k == 2) ? A : pw[k - 3];
    const int ldl = (k == 2) ? lda : n;
    if (cplx) orc_gemm_z(n, one_z, left, ldl, A, lda, zero_z, pw[k - 2], n);
    else orc_gemm_d(n, 1.0, left, ldl, A, lda, 0.0, pw[k - 2], n);
  }
  const double *X = (r < 1) ? NULL : (s >= 2) ? pw[s - 2] : A;
  const int ldx = (s >= 2) ? n : lda;
  double *buf[2] = {P, scratch};
  const int ld[2] = {ldp, n};
  int cur = (r % 2 == 0) ? 0 : 1;
#define TERMS(j) (MIN(s, degree - (j) * s + 1))
  build_block(n, buf[cur], ld[cur], A, lda, pw, coeffs + es * (size_t)r * s, TERMS(r), cplx);
  for (int j = r - 1; j >= 0; --j) {
    const int next = 1 - cur;
    build_block(n, buf[next], ld[next], A, lda, pw, coeffs + es * (size_t)j * s, TERMS(j), cplx);
    if (cplx) orc_gemm_z(n, one_z, buf[cur], ld[cur], X, ldx, one_z, buf[next], ld[next]);
    else orc_gemm_d(n, 1.0, buf[cur], ld[cur], X, ldx, 1.0, buf[next], ld[next]);
    cur = next;
  }
#undef TERMS
  for (int k = 0; k < npw; ++k)
    if (owned[k]) free(pw[k]);
  free(pw);
  free(owned);
  free(scratch);
  return 0;
}

static int ps_impl(int n, const double *coeffs, int degree, const double *A, int lda, double *P,
                   int ldp, int s_requested, int cplx) {
  return ps_impl_have(n, coeffs, degree, A, lda, P, ldp, s_requested, cplx, NULL, 0);
}

int orc_ps_d(int n, const double *coeffs, int degree, const double *A, int lda, double *P, int ldp,
             int s_requested) {
  return ps_impl(n, coeffs, degree, A, lda, P, ldp, s_requested, 0);
}
int orc_ps_z(int n, const double *coeffs, int degree, const double *A, int lda, double *P, int ldp,
             int s_requested) {
  return ps_impl(n, coeffs, degree, A, lda, P, ldp, s_requested, 1);
}

/* Horner (src/horner/interface.cppm:156-194; diagonal kernels src/horner/horner.cu:39-91) */
static int horner_impl(int n, const double *c, int degree, const double *A, int lda, double *P,
                       int ldp, int cplx) {
  if (n < 1 || degree < 0 || lda < n || ldp < n || c == NULL || A == NULL || P == NULL) return -1;
  const size_t es = cplx ? 2 : 1;
  double *scratch = (double *)malloc(sizeof(double) * es * (size_t)n * n);
  double *buf[2] = {P, scratch};
  const int ld[2] = {ldp, n};
  const double one_z[2] = {1.0, 0.0}, zero_z[2] = {0.0, 0.0};
  int cur = (degree % 2 == 0) ? 0 : 1;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      for (size_t q = 0; q < es; ++q)
        buf[cur][es * (i + (size_t)j * ld[cur]) + q] = (i == j) ? c[es * degree + q] : 0.0;
  for (int k = degree - 1; k >= 0; --k) {
    const int next = 1 - cur;
    if (cplx) orc_gemm_z(n, one_z, buf[cur], ld[cur], A, lda, zero_z, buf[next], ld[next]);
    else orc_gemm_d(n, 1.0, buf[cur], ld[cur], A, lda, 0.0, buf[next], ld[next]);
    cur = next;
    for (int i = 0; i < n; ++i)
      for (size_t q = 0; q < es; ++q) buf[cur][es * (i + (size_t)i * ld[cur]) + q] += c[es * k + q];
  }
  free(scratch);
  return 0;
}

int orc_horner_d(int n, const double *coeffs, int degree, const double *A, int lda, double *P,
                 int ldp) {
  return horner_impl(n, coeffs, degree, A, lda, P, ldp, 0);
}
int orc_horner_z(int n, const double *coeffs, int degree, const double *A, int lda, double *P,
                 int ldp) {
  return horner_impl(n, coeffs, degree, A, lda, P, ldp, 1);
}

/* =========================================================================
 * The build's algorithm: Taylor T_d(A / 2^s) by Paterson-Stockmeyer, squared
 * s times. theta_d is the largest x with sum_{k>d} |h_k| x^(k-1) <= u for
 * h(x) = log(exp(-x) T_d(x)), the backward-error criterion the reference
 * applies to its Pade rungs (src/expm/plan.cppm:29-41), at u = 2^-53 and 2^-24.
 * The ladder keeps the degrees whose Paterson-Stockmeyer cost
 * (orc_ps_plan) rises by one product per rung.
 * ========================================================================= */
static const int kTaylorDegrees[10] = {1, 2, 3, 5, 8, 11, 15, 19, 24, 29};
static const double kTaylorThetaD[10] = {
    2.220446049250313e-16, 2.580956802971767e-8, 1.386347866119121e-5, 2.400876357887274e-3,
    4.991228871115323e-2,  2.142358068451711e-1, 6.410835233041199e-1, 1.260381060642639,
    2.21904886936509,      3.310172839890271};
static const double kTaylorThetaS[10] = {
    1.192092800768788e-7, 5.978858893805233e-4, 1.123386473528671e-2, 1.30848716459947e-1,
    5.800524627688768e-1, 1.223479542424143,    2.21704439497472,     3.279561212635997,
    4.649782224100757,    6.041358758192571};

static const double kInvFact[33] = {
    0x1.0000000000000p+0,  0x1.0000000000000p+0,  0x1.0000000000000p-1,  0x1.5555555555555p-3,
    0x1.5555555555555p-5,  0x1.1111111111111p-7,  0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
    0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
    0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41,
    0x1.ae7f3e733b81fp-45, 0x1.952c77030ad4ap-49, 0x1.6827863b97d97p-53, 0x1.2f49b46814157p-57,
    0x1.e542ba4020225p-62, 0x1.71b8ef6dcf572p-66, 0x1.0ce396db7f853p-70, 0x1.761b41316381ap-75,
    0x1.f2cf01972f578p-80, 0x1.3f3ccdd165fa9p-84, 0x1.88e85fc6a4e5ap-89, 0x1.d1ab1c2dccea3p-94,
    0x1.0a18a2635085dp-98, 0x1.259f98b4358adp-103, 0x1.3932c5047d60ep-108,
    0x1.434d2e783f5bcp-113, 0x1.434d2e783f5bcp-118};

const double *orc_inv_factorials(void) { return kInvFact; }

static double taylor_theta(int degree, const double *tab) {
  for (int i = 0; i < 10; ++i)
    if (kTaylorDegrees[i] == degree) return tab[i];
  return 0.0;
}
double orc_taylor_theta_d(int degree) { return taylor_theta(degree, kTaylorThetaD); }
double orc_taylor_theta_s(int degree) { return taylor_theta(degree, kTaylorThetaS); }

/* Squarings rung i needs from the 1-norm alone; a ratio that overflows marks
 * the rung unusable (a huge norm against a tiny theta). */
#define ORC_UNUSABLE_SQUARINGS (1 << 20)
static int norm_squarings(double norm1, double theta) {
  if (!(norm1 > theta)) return 0;
  const double ratio = norm1 / theta;
  if (!(ratio <= DBL_MAX)) return ORC_UNUSABLE_SQUARINGS;
  const int s = (int)ceil(log2(ratio));
  return s < 0 ? 0 : s;
}

/* With pn[k-1] = ||A^k||_1 for k = 1..q: the smallest s with
 * alpha_p / 2^s <= theta_m for some p with p(p-1) <= m+1, where
 * alpha_p = max(||A^p||^(1/p), ||A^(p+1)||^(1/(p+1))) (Al-Mohy & Higham 2009,
 * Thm 4.2(a)); alpha_1 = ||A||_1. d_k <= x is tested as ||A^k|| <= x^k. */
static int power_squarings(const double *pn, int q, int m, double theta) {
  const int s_norm = norm_squarings(pn[0], theta);
  if (q < 3 || s_norm == 0 || s_norm >= ORC_UNUSABLE_SQUARINGS) return s_norm;
  for (int s = 0; s < s_norm; ++s) {
    const double x = ldexp(theta, s);
    double xp = x * x;
    for (int p = 2; p + 1 <= q && p * (p - 1) <= m + 1; ++p) {
      const double xp1 = xp * x;
      if (pn[p - 1] <= xp && pn[p] <= xp1) return s;
      xp = xp1;
    }
  }
  return s_norm;
}

/* Minimize max(ps_powers(d), q-1) + ps_horner(d) + s over the ladder (the q-1
 * powers formed to take their norms count once); ties go to the smaller s.
 * q = 1: the plan from the 1-norm alone. */
static void taylor_plan_powers(const double *pn, int q, const double *tab, int *out) {
  const int have = q > 1 ? q - 1 : 0;
  int best = -1, best_cost = 0, best_s = 0;
  for (int i = 0; i < 10; ++i) {
    int pl[5];
    orc_ps_plan(kTaylorDegrees[i], 0, pl);
    const int s = power_squarings(pn, q, kTaylorDegrees[i], tab[i]);
    const int cost = (pl[2] > have ? pl[2] : have) + pl[1] + s;
    if (best < 0 || cost < best_cost || (cost == best_cost && s < best_s)) {
      best = i; best_cost = cost; best_s = s;
    }
  }
  int pl[5];
  orc_ps_plan(kTaylorDegrees[best], 0, pl);
  out[0] = kTaylorDegrees[best]; out[1] = best_s; out[2] = best_cost; out[3] = pl[0];
}
void orc_taylor_plan_d(double norm1, int *out) { taylor_plan_powers(&norm1, 1, kTaylorThetaD, out); }
void orc_taylor_plan_s(double norm1, int *out) { taylor_plan_powers(&norm1, 1, kTaylorThetaS, out); }
void orc_taylor_plan_powers_d(const double *pn, int q, int *out) {
  taylor_plan_powers(pn, q, kTaylorThetaD, out);
}
void orc_taylor_plan_powers_s(const double *pn, int q, int *out) {
  taylor_plan_powers(pn, q, kTaylorThetaS, out);
}

/* The powers formed first: A^2 and A^3 when the norm-only plan's bank holds A^3
 * and ||A||_1 <= 2^200 (3), else none (1). */
static int norm_power_count(double norm1, int degree) {
  if (!(norm1 <= 0x1p200)) return 1;
  int pl[5];
  orc_ps_plan(degree, 0, pl);
  return pl[2] + 1 >= 3 ? 3 : 1;
}

/* Whether forming A^(q+1) can lower the plan's cost: the best plan with
 * ||A^(q+1)|| taken as 0 against the best plan without it (q < 5). */
static int worth_next_power(const double *pn, int q, const double *tab) {
  if (q < 3 || q >= 5) return 0;
  double trial[5] = {0, 0, 0, 0, 0};
  for (int k = 0; k < q; ++k) trial[k] = pn[k];
  int with_next[4], now[4];
  taylor_plan_powers(trial, q + 1, tab, with_next);
  taylor_plan_powers(pn, q, tab, now);
  return with_next[2] < now[2];
}

/* The build's Taylor expm: A^2, A^3 (then A^4, A^5 while worth_next_power) formed
 * from A itself, their exact 1-norms, the plan of taylor_plan_powers, then A / 2^s and the bank
 * scaled in place by 2^(-k s) (ldexp, exact for normal results), the polynomial
 * continued from the powers already formed, s squarings. plan[4] = q and
 * pnorm[k-2] = ||A^k||_1 (k = 2..q) when given. */
static int expm_taylor_impl(int n, const double *A, int lda, double *E, int lde, int *plan,
                            double *pnorm, int cplx) {
  if (n < 1 || n > ORC_MAX_DIM || lda < n || lde < n || A == NULL || E == NULL) return -1;
  const double norm = cplx ? orc_norm1_z(n, A, lda) : orc_norm1_d(n, A, lda);
  if (!isfinite(norm)) return -2;
  const size_t es = cplx ? 2 : 1;
  const size_t nn = (size_t)n * n;
  int p0[4];
  orc_taylor_plan_d(norm, p0);
  int q = norm_power_count(norm, p0[0]);
  double pn[5] = {norm, 0, 0, 0, 0};
  double *pw[4] = {NULL, NULL, NULL, NULL};
  const double one_z[2] = {1.0, 0.0}, zero_z[2] = {0.0, 0.0};
  /* A^2, A^3, then A^4, A^5 one at a time while worth_next_power */
  for (int k = 2; k <= q; ++k) {
    pw[k - 2] = (double *)malloc(sizeof(double) * es * nn);
    const double *left = (k == 2) ? A : pw[k - 3];
    const int ldl = (k == 2) ? lda : n;
    if (cplx) orc_gemm_z(n, one_z, left, ldl, A, lda, zero_z, pw[k - 2], n);
    else orc_gemm_d(n, 1.0, left, ldl, A, lda, 0.0, pw[k - 2], n);
    pn[k - 1] = cplx ? orc_norm1_z(n, pw[k - 2], n) : orc_norm1_d(n, pw[k - 2], n);
    if (k == q && worth_next_power(pn, q, kTaylorThetaD)) ++q;
  }
  int p[4];
  taylor_plan_powers(pn, q, kTaylorThetaD, p);
  if (plan) { plan[0] = p[0]; plan[1] = p[1]; plan[2] = p[2]; plan[3] = p[3]; plan[4] = q; }
  if (pnorm)
    for (int k = 2; k <= 5; ++k) pnorm[k - 2] = (k <= q) ? pn[k - 1] : 0.0;
  const double *src = A;
  int ld_src = lda;
  double *As = NULL;
  if (p[1] > 0) {
    As = scaled_copy(n, A, lda, p[1], cplx);
    src = As;
    ld_src = n;
    for (int k = 2; k <= q; ++k)
      for (size_t e = 0; e < es * nn; ++e) pw[k - 2][e] = ldexp(pw[k - 2][e], -k * p[1]);
  }
  double coef[66];
  for (int k = 0; k <= p[0]; ++k) {
    if (cplx) { coef[2 * k] = kInvFact[k]; coef[2 * k + 1] = 0.0; }
    else coef[k] = kInvFact[k];
  }
  ps_impl_have(n, coef, p[0], src, ld_src, E, lde, p[3], cplx, pw, q - 1);
  free(As);
  for (int k = 0; k < 4; ++k) free(pw[k]);
  square_s_times(n, p[1], E, lde, cplx);
  return 0;
}

int orc_expm_taylor_d(int n, const double *A, int lda, double *E, int lde, int *plan,
                      double *pnorm) {
  return expm_taylor_impl(n, A, lda, E, lde, plan, pnorm, 0);
}
int orc_expm_taylor_z(int n, const double *A, int lda, double *E, int lde, int *plan,
                      double *pnorm) {
  return expm_taylor_impl(n, A, lda, E, lde, plan, pnorm, 1);
}
