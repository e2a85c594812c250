/*
 * oracle.h -- CPU restatement of calaman_G's matrix-exponential path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the GPU product is
 * compared against; it is linked by nothing the product ships. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Matrices are column-major with a leading dimension. Complex matrices are
 * interleaved (re, im) doubles -- the layout of std::complex<double>,
 * hipDoubleComplex and numpy complex128 -- so a complex n x n matrix with
 * leading dimension ld is 2*ld*n doubles and element (i, j) sits at
 * 2*(i + j*ld).
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the calaman_G source root).
 */
#ifndef CALAMAN_ORACLE_H
#define CALAMAN_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Pade ladder: tables, thresholds, plan (src/expm/expm_bridge.h:58-130,
 *      src/expm/plan.cppm:51-130) ---------------------------------------- */
const double *orc_pade_coeffs(int m);
int orc_pade_num_powers(int m);
int orc_pade_num_gemms(int m);
double orc_pade_theta_d(int m);
float orc_pade_theta_s(int m);
/* out[3] = {m, s, num_gemms} */
void orc_expm_plan_d(double norm1, int *out);
void orc_expm_plan_s(float norm1, int *out);

/* ---- Norms (src/expm/norm1.cppm:45-56, src/expm/expm.cu:188-245,
 *      src/reduce_columns/reduce_columns.cuh:62-108,
 *      src/lapack/lange/lange.cu:42-156) ----------------------------------- */
double orc_norm1_d(int n, const double *A, int lda);
double orc_norm1_z(int n, const double *A, int lda);
/* which: 0 max_abs, 1 one, 2 inf, 3 frobenius (src/common/enums.h:154-159) */
double orc_lange_d(int which, size_t m, size_t n, const double *A, size_t lda);
double orc_lange_z(int which, size_t m, size_t n, const double *A, size_t lda);

/* ---- lascl TYPE='G' guarded multiplier loop
 *      (src/lapack/lascl/interface.cppm:79-143) --------------------------- */
/* returns 0, or -1 for a zero/NaN cfrom or NaN cto */
int orc_lascl_d(double cfrom, double cto, size_t m, size_t n, double *A, size_t lda);
int orc_lascl_z(double cfrom, double cto, size_t m, size_t n, double *A, size_t lda);

/* The same three in single precision (accumulation and multipliers in float);
 * complex<float> as interleaved float pairs. */
float orc_norm1_s(int n, const float *A, int lda);
float orc_norm1_c(int n, const float *A, int lda);
float orc_lange_s(int which, size_t m, size_t n, const float *A, size_t lda);
float orc_lange_c(int which, size_t m, size_t n, const float *A, size_t lda);
int orc_lascl_s(float cfrom, float cto, size_t m, size_t n, float *A, size_t lda);
int orc_lascl_c(float cfrom, float cto, size_t m, size_t n, float *A, size_t lda);

/* ---- BLAS-3 / LAPACK pieces the reference takes from the vendor ---------- */
/* C = alpha*A*B + beta*C, n x n, deterministic: each element sums k ascending,
 * independent of the thread count. beta == 0 never reads C. */
void orc_gemm_d(int n, double alpha, const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc);
void orc_gemm_z(int n, const double *alpha, const double *A, int lda, const double *B, int ldb,
                const double *beta, double *C, int ldc);
/* partial-pivot LU (LAPACK ?getrf semantics, 1-based ipiv, info > 0 singular) */
int orc_getrf_d(int n, double *A, int lda, int *ipiv);
int orc_getrf_z(int n, double *A, int lda, int *ipiv);
/* solve A X = B with the factors, nrhs right-hand sides */
int orc_getrs_d(int n, int nrhs, const double *LU, int lda, const int *ipiv, double *B, int ldb);
int orc_getrs_z(int n, int nrhs, const double *LU, int lda, const int *ipiv, double *B, int ldb);

/* ---- Pade approximant and the reference expm (src/expm/pade.cppm:77-191,
 *      src/expm/expm.cppm:87-174) ----------------------------------------- */
/* info[2] = getrf, getrs info. Returns 0 or a negative argument error. */
int orc_pade_d(int m, int n, const double *A, int lda, double *R, int ldr, int *info);
int orc_pade_z(int m, int n, const double *A, int lda, double *R, int ldr, int *info);
/* plan[3] = {m, s, num_gemms} (may be NULL). Returns 0, -1 bad argument,
 * -2 non-finite 1-norm. */
int orc_expm_d(int n, const double *A, int lda, double *E, int lde, int *plan);
int orc_expm_z(int n, const double *A, int lda, double *E, int lde, int *plan);

/* ---- Paterson-Stockmeyer and Horner (src/paterson_stockmeyer/interface.cppm:117-422,
 *      src/horner/interface.cppm:156-194) --------------------------------- */
int orc_ps_cost(int degree, int s);
int orc_ps_block_size(int degree);
/* out[5] = {s, r, num_powers, num_blocks, num_gemms} */
void orc_ps_plan(int degree, int s_requested, int *out);
int orc_ps_d(int n, const double *coeffs, int degree, const double *A, int lda, double *P, int ldp,
             int s_requested);
int orc_ps_z(int n, const double *coeffs, int degree, const double *A, int lda, double *P, int ldp,
             int s_requested);
int orc_horner_d(int n, const double *coeffs, int degree, const double *A, int lda, double *P,
                 int ldp);
int orc_horner_z(int n, const double *coeffs, int degree, const double *A, int lda, double *P,
                 int ldp);

/* ---- The build's algorithm: truncated Taylor series by Paterson-Stockmeyer,
 *      scaling and squaring. Same plan rule as the GPU library. -------------- */
double orc_taylor_theta_d(int degree);
double orc_taylor_theta_s(int degree);
const double *orc_inv_factorials(void); /* 1/k!, k = 0..32, correctly rounded */
/* out[4] = {degree, s, num_gemms, ps_block} */
void orc_taylor_plan_d(double norm1, int *out);
void orc_taylor_plan_s(double norm1, int *out);
/* the plan from pn[k-1] = ||A^k||_1, k = 1..q (out: degree, s, products, block) */
void orc_taylor_plan_powers_d(const double *pn, int q, int *out);
void orc_taylor_plan_powers_s(const double *pn, int q, int *out);
/* plan: {degree, s, products, block, q}; pnorm (4 doubles, may be NULL):
 * ||A^k||_1 for k = 2..q of the unscaled powers */
int orc_expm_taylor_d(int n, const double *A, int lda, double *E, int lde, int *plan,
                      double *pnorm);
int orc_expm_taylor_z(int n, const double *A, int lda, double *E, int lde, int *plan,
                      double *pnorm);

/* threads the oracle uses (OpenMP) */
int orc_num_threads(void);
void orc_set_num_threads(int t);

#ifdef __cplusplus
}
#endif
#endif
