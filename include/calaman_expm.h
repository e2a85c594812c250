/*
 * calaman_expm.h -- C ABI of the MI355X (gfx950) matrix-exponential library.
 *
 * Drop-in boundary for calaman_G's matrix-exponential path. Each entry point
 * names the reference interface it replaces (paths relative to the calaman_G
 * source root). The reference's C++ templates over {float, double,
 * complex<float>, complex<double>} become type-suffixed C functions here
 * (s = float, d = double, c = complex<float>, z = complex<double>); the C++
 * header calaman/expm.hpp restores the template spelling on top of these.
 *
 * Conventions kept from the reference:
 *   - matrices are column-major with a leading dimension >= n; only the n x n
 *     block of an output is written, padding rows are left untouched;
 *   - the caller owns every device buffer; the workspace is sized by the
 *     *_buffer_size query and passed in, nothing on the expm path allocates;
 *   - all work is enqueued on one stream; clm_expm_* and clm_matrix_norm1_*
 *     synchronize that stream once (to read the 1-norm that picks the plan);
 *     every other entry point is asynchronous;
 *   - a bad argument or a non-finite 1-norm is CLM_STATUS_INVALID_VALUE, an
 *     undersized workspace CLM_STATUS_ALLOC_FAILED (src/expm/expm.cppm:77-80);
 *     Paterson-Stockmeyer and Horner report every bad argument as
 *     CLM_STATUS_NOT_INITIALIZED (src/paterson_stockmeyer/interface.cppm:333-345,
 *     src/horner/interface.cppm:160-164); lascl reports a bad ratio in the
 *     runtime domain (src/lapack/lascl/interface.cppm:90-95).
 *
 * No argument-checking path touches the device: a rejected call returns before
 * any HIP call, so it can be exercised on a host without a GPU.
 */
#ifndef CALAMAN_EXPM_H
#define CALAMAN_EXPM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- element types ---------------------------------------------------------
 * Interleaved complex, layout-compatible with std::complex<T>,
 * hipDoubleComplex/hipFloatComplex and numpy complex128/complex64. */
typedef struct clm_double_complex {
  double re, im;
} clm_double_complex;
typedef struct clm_float_complex {
  float re, im;
} clm_float_complex;

/* A HIP stream (hipStream_t) as an opaque pointer; NULL is the null stream. */
typedef void *clm_stream_t;

/* ---- status ------------------------------------------------------------------
 * calaman::Status{domain, code} (src/error_handling/error_handling.cppm:88-180)
 * flattened into one int: 0 is success in every domain, otherwise
 * domain | code. */
typedef int clm_status_t;
#define CLM_DOMAIN_MASK 0x7fff0000
#define CLM_DOMAIN_BLAS 0x00010000
#define CLM_DOMAIN_SOLVER 0x00020000
#define CLM_DOMAIN_RUNTIME 0x00030000
#define CLM_STATUS_SUCCESS 0
#define CLM_STATUS_NOT_INITIALIZED (CLM_DOMAIN_BLAS | 1)
#define CLM_STATUS_ALLOC_FAILED (CLM_DOMAIN_BLAS | 3)
#define CLM_STATUS_INVALID_VALUE (CLM_DOMAIN_BLAS | 7)
#define CLM_STATUS_EXECUTION_FAILED (CLM_DOMAIN_BLAS | 13)
#define CLM_STATUS_NOT_SUPPORTED (CLM_DOMAIN_BLAS | 15)
#define CLM_STATUS_RUNTIME_INVALID_VALUE (CLM_DOMAIN_RUNTIME | 1)
#define CLM_STATUS_RUNTIME_ERROR (CLM_DOMAIN_RUNTIME | 2)

/* Status name, e.g. "CLM_STATUS_INVALID_VALUE"; never NULL. */
const char *clm_status_string(clm_status_t status);

/* ---- context -----------------------------------------------------------------
 * Replaces the reference's (wwrblasHandle_t, wwrsolverDnHandle_t) pair. A
 * context is bound to the HIP device current at creation and carries the
 * stream that clm_paterson_stockmeyer_* / clm_horner_* / clm_gemm_* enqueue on
 * (the reference takes that stream from the BLAS handle,
 * src/paterson_stockmeyer/interface.cppm:349-350). Everywhere a clm_ctx* is
 * taken, NULL means the library's default context for the current device
 * (null stream). One context per host thread, as with vendor handles. */
typedef struct clm_ctx clm_ctx;
clm_status_t clm_create(clm_ctx **ctx);
clm_status_t clm_destroy(clm_ctx *ctx);
clm_status_t clm_set_stream(clm_ctx *ctx, clm_stream_t stream);
clm_status_t clm_get_stream(clm_ctx *ctx, clm_stream_t *stream);

/* ---- plans -------------------------------------------------------------------
 * ExpmPlan{m, s, num_gemms} (src/expm/plan.cppm:94-98), extended with the
 * method, the Paterson-Stockmeyer block size, the element type and the 1-norm
 * the plan was chosen from. */
#define CLM_METHOD_TAYLOR_PS 0 /* Taylor series by Paterson-Stockmeyer (this library) */
#define CLM_METHOD_PADE 1      /* diagonal Pade + LU (the reference's route; plan only) */
#define CLM_DTYPE_F64 1
#define CLM_DTYPE_F32 2
#define CLM_DTYPE_C128 3
#define CLM_DTYPE_C64 4
typedef struct clm_expm_plan {
  int m;         /* degree: Taylor degree, or Pade degree for CLM_METHOD_PADE */
  int s;         /* squarings: the matrix is evaluated at A / 2^s */
  int num_gemms; /* matrix products the evaluation issues */
  int method;    /* CLM_METHOD_* */
  int ps_block;  /* Paterson-Stockmeyer block size (0 for Pade) */
  int dtype;     /* CLM_DTYPE_* (0 when only predicted from a norm) */
  int ps_split;  /* polynomial products: 0 = FP64 MFMA, 3 / 6 = BF16 split (clm_expm_split_d) */
  int split_squarings; /* how many of the last squarings ran split (FP16X3 planes) */
  double norm1;  /* the 1-norm the plan was chosen from */
  int fp8_gemms;      /* polynomial products that ran on FP8 (clm_expm_fp8tail_d) */
  int fp8_min_degree; /* the lowest Taylor degree an FP8 product may carry (0 otherwise) */
  int norm_powers;    /* q: the plan used ||A^k||_1 for k = 1 .. q (1: the 1-norm alone) */
  double power_norm1[4]; /* ||A^k||_1 of the unscaled powers k = 2 .. q (0 beyond q) */
} clm_expm_plan;

/* The plan for a given 1-norm alone (the role of expm_plan<T>,
 * src/expm/plan.cppm:112-130): over the Taylor degrees
 * {1,2,3,5,8,11,15,19,24,29}, minimise Paterson-Stockmeyer products + s with
 * s = max(0, ceil(log2(norm1 / theta_m))), ties to the smaller s.
 * clm_expm_* start from it; when its power bank holds A^3 (degree >= 8) and
 * ||A||_1 <= 2^200 (2^24 for the single-precision types and the split / FP8
 * variants) they form that bank from the unscaled matrix first, take the exact
 * 1-norms of the powers and execute clm_expm_plan_powers_* of them. */
clm_expm_plan clm_expm_plan_d(double norm1);
/* The plan from norms[k-1] = ||A^k||_1, k = 1 .. q (1 <= q <= 5): s is the smallest
 * with alpha_p / 2^s <= theta_m for some p with p(p-1) <= m+1, where
 * alpha_p = max(||A^p||^(1/p), ||A^(p+1)||^(1/(p+1))) (Al-Mohy & Higham 2009,
 * Thm 4.2; alpha_1 = ||A||_1), and the q - 1 powers already formed count as
 * products whether or not the chosen degree's bank uses them. No reference
 * counterpart: the reference plans from ||A||_1 alone. */
clm_expm_plan clm_expm_plan_powers_d(const double *norms, int q);
clm_expm_plan clm_expm_plan_powers_s(const double *norms, int q);
/* Backward-error threshold theta_m of the degree-m Taylor approximant at
 * u = 2^-53 (0 for a degree off the ladder); the counterpart of pade_theta<T>
 * (src/expm/plan.cppm:51-85). */
double clm_taylor_theta_d(int m);
/* The reference's own Pade ladder, restated for callers that predict its cost:
 * pade_theta<double> and expm_plan<double> (src/expm/plan.cppm:51-130). */
double clm_pade_theta_d(int m);
clm_expm_plan clm_pade_plan_d(double norm1);
/* Single precision (clm_expm_s/c): the same Taylor ladder with the thresholds at
 * u = 2^-24 (clm_taylor_theta_s); the reference's float Pade ladder
 * (pade_theta<float>, expm_plan<float>, src/expm/plan.cppm:51-68,112-130). */
clm_expm_plan clm_expm_plan_s(double norm1);
double clm_pade_theta_s(int m);
clm_expm_plan clm_pade_plan_s(float norm1);

/* ---- expm --------------------------------------------------------------------
 * expm_bufferSize<T> (src/expm/buffer_size.cppm:77-97): bytes of device
 * workspace clm_expm_* needs for order n (n >= 1). */
clm_status_t clm_expm_buffer_size_s(int n, size_t *bytes);
clm_status_t clm_expm_buffer_size_d(int n, size_t *bytes);
clm_status_t clm_expm_buffer_size_c(int n, size_t *bytes);
clm_status_t clm_expm_buffer_size_z(int n, size_t *bytes);

/* expm<T> (src/expm/expm.cppm:87-174): E = exp(A) by scaling and squaring.
 *   ctx      NULL or a context (stands in for both reference handles)
 *   stream   the stream all work is enqueued on; synchronized once
 *   dA       n x n input, leading dimension lda >= n; must not overlap dE
 *   dE       n x n output, leading dimension lde >= n
 *   dWork    >= clm_expm_buffer_size_* bytes of device memory, 256-B aligned
 *   dInfo    device int[2]; the reference's getrf/getrs info. The Taylor
 *            route needs no LU, so both are written 0.
 *   plan     optional: the executed plan
 * n must lie in [1, 46340] (src/expm/detail.cppm:25-28). The float and
 * complex<float> forms compute in FP32 (FP32 MFMA products) with the degree
 * ladder at u = 2^-24. */
clm_status_t clm_expm_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dA, int lda,
                        float *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                        clm_expm_plan *plan);
clm_status_t clm_expm_c(clm_ctx *ctx, clm_stream_t stream, int n, const clm_float_complex *dA,
                        int lda, clm_float_complex *dE, int lde, void *dWork, size_t lwork_bytes,
                        int *dInfo, clm_expm_plan *plan);
clm_status_t clm_expm_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA, int lda,
                        double *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                        clm_expm_plan *plan);
clm_status_t clm_expm_z(clm_ctx *ctx, clm_stream_t stream, int n, const clm_double_complex *dA,
                        int lda, clm_double_complex *dE, int lde, void *dWork, size_t lwork_bytes,
                        int *dInfo, clm_expm_plan *plan);

/* ---- the reference's own route: Pade approximant + LU --------------------------
 * pade<T> (src/expm/pade.cppm:77-191): R = r_m(A) = q_m(A)^-1 p_m(A) for m in
 * {3, 5, 7, 9, 13}, unscaled (accurate while ||A||_1 <= clm_pade_theta_*(m)):
 * powers A^2 .. A^8 (A^6 and the nested form at m = 13), the even and odd halves
 * of p_m in one pass, U = A W, then (V - U) R = (V + U) by clm_getrf_* /
 * clm_getrs_*. Never synchronizes. dInfo: device int[2] = getrf info, getrs info
 * (may be NULL). pade_bufferSize<T> (src/expm/buffer_size.cppm:36-48):
 * clm_pade_buffer_size_*. */
clm_status_t clm_pade_buffer_size_s(int m, int n, size_t *bytes);
clm_status_t clm_pade_buffer_size_d(int m, int n, size_t *bytes);
clm_status_t clm_pade_buffer_size_c(int m, int n, size_t *bytes);
clm_status_t clm_pade_buffer_size_z(int m, int n, size_t *bytes);
clm_status_t clm_pade_s(clm_ctx *ctx, clm_stream_t stream, int m, int n, const float *dA, int lda,
                        float *dR, int ldr, void *dWork, size_t lwork_bytes, int *dInfo);
clm_status_t clm_pade_d(clm_ctx *ctx, clm_stream_t stream, int m, int n, const double *dA, int lda,
                        double *dR, int ldr, void *dWork, size_t lwork_bytes, int *dInfo);
clm_status_t clm_pade_c(clm_ctx *ctx, clm_stream_t stream, int m, int n,
                        const clm_float_complex *dA, int lda, clm_float_complex *dR, int ldr,
                        void *dWork, size_t lwork_bytes, int *dInfo);
clm_status_t clm_pade_z(clm_ctx *ctx, clm_stream_t stream, int m, int n,
                        const clm_double_complex *dA, int lda, clm_double_complex *dR, int ldr,
                        void *dWork, size_t lwork_bytes, int *dInfo);
/* expm<T> exactly as the reference evaluates it (src/expm/expm.cppm:87-174): the
 * 1-norm (one synchronization), the Pade plan (clm_pade_plan_*), 2^-s A, r_m,
 * s squarings; dInfo (may be NULL) as clm_pade_*; plan->method =
 * CLM_METHOD_PADE. expm_bufferSize<T>: clm_expm_pade_buffer_size_*. */
clm_status_t clm_expm_pade_buffer_size_s(int n, size_t *bytes);
clm_status_t clm_expm_pade_buffer_size_d(int n, size_t *bytes);
clm_status_t clm_expm_pade_buffer_size_c(int n, size_t *bytes);
clm_status_t clm_expm_pade_buffer_size_z(int n, size_t *bytes);
clm_status_t clm_expm_pade_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dA, int lda,
                             float *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                             clm_expm_plan *plan);
clm_status_t clm_expm_pade_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA, int lda,
                             double *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                             clm_expm_plan *plan);
clm_status_t clm_expm_pade_c(clm_ctx *ctx, clm_stream_t stream, int n,
                             const clm_float_complex *dA, int lda, clm_float_complex *dE, int lde,
                             void *dWork, size_t lwork_bytes, int *dInfo, clm_expm_plan *plan);
clm_status_t clm_expm_pade_z(clm_ctx *ctx, clm_stream_t stream, int n,
                             const clm_double_complex *dA, int lda, clm_double_complex *dE, int lde,
                             void *dWork, size_t lwork_bytes, int *dInfo, clm_expm_plan *plan);
/* The LU the reference takes from the vendor solver (getrf/getrs,
 * src/expm/pade.cppm:186-188): A = P L U in place with partial pivoting (ipiv
 * 1-based; *dInfo = the first zero pivot's column, 1-based, or 0), then
 * B <- A^-1 B for nrhs right-hand sides (dInfo may be NULL; written 0). Square
 * matrices, n <= 46340. No workspace. */
clm_status_t clm_getrf_s(clm_ctx *ctx, clm_stream_t stream, int n, float *dA, int lda, int *dIpiv,
                         int *dInfo);
clm_status_t clm_getrf_d(clm_ctx *ctx, clm_stream_t stream, int n, double *dA, int lda, int *dIpiv,
                         int *dInfo);
clm_status_t clm_getrf_c(clm_ctx *ctx, clm_stream_t stream, int n, clm_float_complex *dA, int lda,
                         int *dIpiv, int *dInfo);
clm_status_t clm_getrf_z(clm_ctx *ctx, clm_stream_t stream, int n, clm_double_complex *dA, int lda,
                         int *dIpiv, int *dInfo);
clm_status_t clm_getrs_s(clm_ctx *ctx, clm_stream_t stream, int n, int nrhs, const float *dLU,
                         int lda, const int *dIpiv, float *dB, int ldb, int *dInfo);
clm_status_t clm_getrs_d(clm_ctx *ctx, clm_stream_t stream, int n, int nrhs, const double *dLU,
                         int lda, const int *dIpiv, double *dB, int ldb, int *dInfo);
clm_status_t clm_getrs_c(clm_ctx *ctx, clm_stream_t stream, int n, int nrhs,
                         const clm_float_complex *dLU, int lda, const int *dIpiv,
                         clm_float_complex *dB, int ldb, int *dInfo);
clm_status_t clm_getrs_z(clm_ctx *ctx, clm_stream_t stream, int n, int nrhs,
                         const clm_double_complex *dLU, int lda, const int *dIpiv,
                         clm_double_complex *dB, int ldb, int *dInfo);

/* ---- exp(A) for a symmetric / Hermitian A by its eigendecomposition --------------------
 * calaman::expm_herm<T> (src/expm/expm_herm.cppm:145-199): only the uplo triangle of A
 * is read (the other assumed to mirror it; the diagonal's imaginary part ignored),
 * A = U diag(w) U^H, exp(A) = U diag(exp(w)) U^H. The reference takes the
 * eigendecomposition from the vendor syevd / heevd; this library's is a blocked
 * two-sided Jacobi method on the device (off-diagonal Frobenius norm driven below
 * u ||A||_F). uplo: CLM_FILL_MODE_LOWER / CLM_FILL_MODE_UPPER (the BLAS fill-mode
 * values). dInfo (device int, may be NULL): 0, or 1 when the Jacobi sweeps did not
 * converge within 30 (the result is still written). Synchronizes once per sweep.
 * Bad arguments: CLM_STATUS_INVALID_VALUE; a short workspace: CLM_STATUS_ALLOC_FAILED.
 * expm_herm_bufferSize<T> (:97-109): clm_expm_herm_buffer_size_* (three blocks of
 * order n rounded up to 64, where the reference holds two and the solver's scratch). */
#define CLM_FILL_MODE_LOWER 0
#define CLM_FILL_MODE_UPPER 1
clm_status_t clm_expm_herm_buffer_size_s(int n, size_t *bytes);
clm_status_t clm_expm_herm_buffer_size_d(int n, size_t *bytes);
clm_status_t clm_expm_herm_buffer_size_c(int n, size_t *bytes);
clm_status_t clm_expm_herm_buffer_size_z(int n, size_t *bytes);
clm_status_t clm_expm_herm_s(clm_ctx *ctx, clm_stream_t stream, int uplo, int n, const float *dA,
                             int lda, float *dE, int lde, void *dWork, size_t lwork_bytes,
                             int *dInfo);
clm_status_t clm_expm_herm_d(clm_ctx *ctx, clm_stream_t stream, int uplo, int n, const double *dA,
                             int lda, double *dE, int lde, void *dWork, size_t lwork_bytes,
                             int *dInfo);
clm_status_t clm_expm_herm_c(clm_ctx *ctx, clm_stream_t stream, int uplo, int n,
                             const clm_float_complex *dA, int lda, clm_float_complex *dE, int lde,
                             void *dWork, size_t lwork_bytes, int *dInfo);
clm_status_t clm_expm_herm_z(clm_ctx *ctx, clm_stream_t stream, int uplo, int n,
                             const clm_double_complex *dA, int lda, clm_double_complex *dE, int lde,
                             void *dWork, size_t lwork_bytes, int *dInfo);

/* ---- balancing (?gebal), the balanced exponential, the 1-norm estimator (?lacn2) -----
 * calaman::gebal<T> (src/lapack/gebal/interface.cppm:128-262): A overwritten with
 * B = D^-1 P^T A P D as LAPACK's xGEBAL through 3.4 does it -- the permutation
 * isolating eigenvalues, then Parlett-Reinsch sweeps of powers of two (exact, so
 * undoing it gives A back bit for bit). job: CLM_GEBAL_* (GebalJob, :76). ilo,
 * ihi: host outputs, 1-based. dScale: device, n reals (real also for complex
 * T): the diagonal of D inside [ilo, ihi], the 1-based exchange index outside.
 * dWork: device ints, clm_gebal_buffer_size (gebal_bufferSize, :89-92). max_sweeps
 * <= 0: the reference's default, 1000. Synchronizes the stream (every step branches
 * on device data) unless job is CLM_GEBAL_NONE. Bad arguments:
 * CLM_STATUS_RUNTIME_INVALID_VALUE; a failed launch or copy: CLM_STATUS_RUNTIME_ERROR. */
#define CLM_GEBAL_NONE 0
#define CLM_GEBAL_PERMUTE 1
#define CLM_GEBAL_SCALE 2
#define CLM_GEBAL_BOTH 3
clm_status_t clm_gebal_buffer_size(int n, int *lwork);
clm_status_t clm_gebal_s(clm_stream_t stream, int job, int n, float *dA, int lda, int *ilo,
                         int *ihi, float *dScale, int *dWork, int max_sweeps);
clm_status_t clm_gebal_d(clm_stream_t stream, int job, int n, double *dA, int lda, int *ilo,
                         int *ihi, double *dScale, int *dWork, int max_sweeps);
clm_status_t clm_gebal_c(clm_stream_t stream, int job, int n, clm_float_complex *dA, int lda,
                         int *ilo, int *ihi, float *dScale, int *dWork, int max_sweeps);
clm_status_t clm_gebal_z(clm_stream_t stream, int job, int n, clm_double_complex *dA, int lda,
                         int *ilo, int *ihi, double *dScale, int *dWork, int max_sweeps);
/* The balancing recipe of src/expm/README.md:100-122 as one call: B = D^-1 A D by
 * gebal's scaling half (no permutation), exp(B) by clm_expm_*, then
 * exp(A) = D exp(B) D^-1 (D of powers of two: the last step is exact). A is not
 * modified. Workspace: clm_expm_balanced_buffer_size_* (B, D, the gebal ints and
 * clm_expm_*'s own). Synchronizes (the sweeps, the 1-norm, the norms of powers). */
clm_status_t clm_expm_balanced_buffer_size_s(int n, size_t *bytes);
clm_status_t clm_expm_balanced_buffer_size_d(int n, size_t *bytes);
clm_status_t clm_expm_balanced_buffer_size_c(int n, size_t *bytes);
clm_status_t clm_expm_balanced_buffer_size_z(int n, size_t *bytes);
clm_status_t clm_expm_balanced_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dA, int lda,
                                 float *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                                 clm_expm_plan *plan);
clm_status_t clm_expm_balanced_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA, int lda,
                                 double *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                                 clm_expm_plan *plan);
clm_status_t clm_expm_balanced_c(clm_ctx *ctx, clm_stream_t stream, int n,
                                 const clm_float_complex *dA, int lda, clm_float_complex *dE,
                                 int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                                 clm_expm_plan *plan);
clm_status_t clm_expm_balanced_z(clm_ctx *ctx, clm_stream_t stream, int n,
                                 const clm_double_complex *dA, int lda, clm_double_complex *dE,
                                 int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                                 clm_expm_plan *plan);
/* calaman::lacn2<T> (src/lapack/lacn2/interface.cppm:73-206): one reverse-
 * communication step of LAPACK's xLACN2 (Hager's method, Higham's refinements).
 * Start with *kase = 0; on a return with *kase 1 (2), overwrite dX with A x (A^T x)
 * and call again with everything else unchanged; *kase 0 on return: done, *est <=
 * ||A||_1 and dV = A w with est = ||v||_1 / ||w||_1. est, kase, isave[3]: host state
 * carried between calls (isave 1-based, as the reference's). dWork: device, at least
 * clm_lacn2_buffer_size_* bytes, kept across calls. The reference binds a BLAS handle
 * for its stream; this takes the stream. Bad arguments or state:
 * CLM_STATUS_INVALID_VALUE. Synchronizes on every step but the first. */
size_t clm_lacn2_buffer_size_s(int n);
size_t clm_lacn2_buffer_size_d(int n);
clm_status_t clm_lacn2_s(clm_stream_t stream, int n, float *dV, float *dX, void *dWork,
                         size_t work_bytes, float *est, int *kase, int *isave);
clm_status_t clm_lacn2_d(clm_stream_t stream, int n, double *dV, double *dX, void *dWork,
                         size_t work_bytes, double *est, int *kase, int *isave);

/* clm_expm_d without a workspace argument, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. The workspace is a file-scope
 * device array inside the library (about 12 MiB per GPU, held until the process
 * ends), so n is at most CLM_EXPM_NOALLOC_MAX_N (larger n: CLM_STATUS_INVALID_VALUE)
 * and the call allocates nothing -- no hipMalloc, hipFree or allocating library
 * routine. Every call on one GPU shares that GPU's array: calls that overlap on a
 * GPU (two host threads, or two streams without ordering) race and must be
 * serialized by the caller. ctx's device (or the current one for NULL) must be
 * the current device. Otherwise clm_expm_d's contract: same kernels, same bits,
 * one synchronization for the 1-norm and one for the norms of powers. */
#define CLM_EXPM_NOALLOC_MAX_N 512
clm_status_t clm_expm_noalloc_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA,
                                int lda, double *dE, int lde, int *dInfo, clm_expm_plan *plan);

/* ---- precision variant ---------------------------------------------------------------
 * No reference counterpart (the build's precision addition). The same algorithm
 * as clm_expm_d with its matrix products moved onto BF16 MFMA: each FP64 operand is
 * split into bfloat16 terms and the products accumulate in FP32
 * (CLM_SPLIT_BF16X3: 2 terms, 3 products; CLM_SPLIT_BF16X6: 3 terms, 6 products).
 * The 1-norm, the plan and the 2^-s scaling stay FP64; the degree comes from the
 * Taylor thresholds at u = 2^-24; the squarings run FP64 except the last
 * `split_squarings`, whose errors the remaining squarings amplify least (none
 * when s > 8), in IEEE-half planes. Target:
 * relative error <= 1e-6 against FP64. */
#define CLM_SPLIT_BF16X3 3
#define CLM_SPLIT_BF16X6 6
/* IEEE half planes, 2 terms, 3 products, every row of a left operand and column
 * of a right operand scaled by a power of two into the half range first (~22
 * bits per product, between X3 and X6 at the cost of X3) */
#define CLM_SPLIT_FP16X3 13
typedef struct clm_split_opts {
  int ps_split;        /* 0 = auto (for s <= 5: BF16X3 while ||A/2^s||_1 <= 1, else FP16X3,
                          BF16X6 for n <= 64; BF16X6 above), or CLM_SPLIT_BF16X3 / BF16X6 /
                          FP16X3 */
  int split_squarings; /* -1 = auto: 0 when s = 0 or s > 8; 2 with an FP16X3 polynomial
                          and s = 4, 5; otherwise 1. Else 0..s. Split squarings use
                          FP16X3 planes whatever the polynomial's format */
  int plan_from;       /* CLM_PLAN_POWERS (0): s from the norms of powers when the bank
                          holds A^3 and ||A||_1 <= 256; CLM_PLAN_NORM1 (1): from ||A||_1
                          alone */
} clm_split_opts;
#define CLM_PLAN_POWERS 0
#define CLM_PLAN_NORM1 1
/* the split plan for a 1-norm alone under opts (NULL = auto) for an order above
 * 64, and the plan clm_expm_split_d executes for order n from norms[k-1] =
 * ||A^k||_1, k = 1 .. q (see clm_expm_plan_powers_d: the powers are formed in the
 * norm-only plan's formats when its bank holds A^3; the split variant uses their
 * norms only while ||A||_1 <= 256, above which its polynomial's 16-bit error at
 * the larger ||A / 2^s||_1 approaches 1e-6). Orders n <= 64 take the BF16 6-way
 * polynomial where larger ones take IEEE-half planes. */
clm_expm_plan clm_expm_split_plan_d(double norm1, const clm_split_opts *opts);
clm_expm_plan clm_expm_split_plan_powers_d(int n, const double *norms, int q,
                                           const clm_split_opts *opts);
double clm_taylor_theta_s(int m); /* theta_m at u = 2^-24 */
/* workspace: the FP64 layout of clm_expm_buffer_size_d plus nine n x n bfloat16
 * planes (each operand is split into planes once, then multiplied) */
clm_status_t clm_expm_split_buffer_size_d(int n, size_t *bytes);
clm_status_t clm_expm_split_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA, int lda,
                              double *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                              const clm_split_opts *opts, clm_expm_plan *plan);

/* ---- FP8 tail: the split variant with its high-degree products on FP8 -----------------
 * No reference counterpart. The plan, the scaling and the squarings are those of
 * clm_expm_split_d; in the Paterson-Stockmeyer polynomial every product whose
 * result holds only Taylor terms of degree >= fp8_min_degree (the power A^k for
 * k >= that degree, the Horner product R_(j+1) X whose lowest term has degree
 * (j+1) b) runs on FP8 MFMA (OCP e4m3, one power-of-two scale per 128 x 128
 * tile, FP32 sums per 128-deep k-tile, FP64 epilogue). An e4m3 operand keeps 4
 * significant bits (relative rounding 2^-4), so the result's error is set by the
 * size of the degree-fp8_min_degree terms at A / 2^s and the squarings that
 * follow: it is reported, not bounded (plan->fp8_gemms counts the FP8 products). */
typedef struct clm_fp8tail_opts {
  int ps_split;        /* products below fp8_min_degree: as clm_split_opts.ps_split */
  int split_squarings; /* as clm_split_opts.split_squarings */
  int fp8_min_degree;  /* 0 = 4; otherwise >= 2 */
} clm_fp8tail_opts;
/* workspace: that of clm_expm_split_buffer_size_d plus three n x n e4m3 planes and
 * their tile exponents */
clm_status_t clm_expm_fp8tail_buffer_size_d(int n, size_t *bytes);
clm_status_t clm_expm_fp8tail_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA,
                                int lda, double *dE, int lde, void *dWork, size_t lwork_bytes,
                                int *dInfo, const clm_fp8tail_opts *opts, clm_expm_plan *plan);

/* ---- batched expm: many orders and types in one persistent launch ------------------
 * No reference counterpart (the reference runs one expm per call). Every matrix
 * of a batch goes through the algorithm of clm_expm_{s,d,c,z} -- the same
 * 1-norm, plan, scaling, Paterson-Stockmeyer products and squarings -- and its
 * result agrees bit for bit with the single-matrix call on the same input. One
 * workgroup runs one matrix's whole chain; a persistent grid takes the matrices
 * largest first. The plan is chosen on the device, nothing is read back, so the
 * launch can be captured in a hipGraph (clm_expm_batched_graph does so).
 *
 * Usage: create a plan from a HOST array of entries (device pointers inside),
 * size and allocate the workspace, then launch as often as wanted; the plan
 * holds the entries, so changing a pointer or an order means a new plan. */
typedef struct clm_batch_entry {
  int n;         /* order, 1..4096 */
  int dtype;     /* CLM_DTYPE_F32 / F64 / C64 / C128 */
  const void *A; /* device, column-major, leading dimension lda >= n */
  int lda;
  void *E;       /* device output, leading dimension lde >= n; must not overlap A */
  int lde;
} clm_batch_entry;
/* per-matrix report, indexed like the entries: the executed plan, or status
 * CLM_STATUS_INVALID_VALUE (and E untouched) for a non-finite 1-norm */
typedef struct clm_batch_result {
  int m, s, num_gemms;
  int status;
  double norm1;
} clm_batch_result;
typedef struct clm_batch_plan clm_batch_plan;

/* workgroup tile of the per-matrix products, chosen per order bucket
 * (bucket b: n in (32b, 32(b+1)], the last bucket everything above) */
#define CLM_BATCH_BUCKETS 8
#define CLM_BATCH_TILE_32 0     /* 32 x 32 */
#define CLM_BATCH_TILE_64 1     /* 64 x 64 */
#define CLM_BATCH_TILE_128x64 2 /* 128 x 64 (complex FP64: 64 x 64) */
#define CLM_BATCH_TILE_128 3    /* 128 x 128 of 4 waves (real FP32 only; others: their largest) */
#define CLM_BATCH_TILE_128W16 4 /* 128 x 128 of 16 waves (every type) */
#define CLM_BATCH_TILES 5

/* Validates the entries (CLM_STATUS_INVALID_VALUE before any device call),
 * orders them by cost, takes each one's tile from the tuning table, and copies
 * the descriptors to the current device. */
clm_status_t clm_expm_batched_plan_create(clm_batch_plan **plan, int count,
                                          const clm_batch_entry *entries);
clm_status_t clm_expm_batched_plan_destroy(clm_batch_plan *plan);
/* device workspace bytes for one launch of the plan */
clm_status_t clm_expm_batched_buffer_size(const clm_batch_plan *plan, size_t *bytes);
/* One launch per precision present (FP32 pair, FP64 pair) on `stream`, plus a
 * one-lane queue reset; asynchronous; dResults optional (device array of count).
 * A member whose 1-norm is not finite gets CLM_STATUS_INVALID_VALUE in its result
 * and NaN in its n x n output. A plan may have ONE launch (or graph replay, or
 * backward) in flight: its side streams and fork/join events are the plan's own,
 * so two launches of one plan on different streams would wait on each other's
 * events; give concurrent launches plans of their own. */
clm_status_t clm_expm_batched(clm_ctx *ctx, clm_stream_t stream, const clm_batch_plan *plan,
                              void *dWork, size_t lwork_bytes, clm_batch_result *dResults);
/* The same launch captured once into a hipGraph (per stream, workspace and
 * results pointer) and replayed; stream must not be the null stream. */
clm_status_t clm_expm_batched_graph(clm_ctx *ctx, clm_stream_t stream, clm_batch_plan *plan,
                                    void *dWork, size_t lwork_bytes, clm_batch_result *dResults);
/* The tuning table (CLM_BATCH_BUCKETS tile ids per element type) and the tuner
 * that fills it: for each bucket up to nmax it times a batch of 1024 matrices
 * of the bucket's middle order under every tile and keeps the fastest. ms_out
 * (optional, CLM_BATCH_TILES * CLM_BATCH_BUCKETS floats) receives
 * ms[bucket * CLM_BATCH_TILES + tile] (0 for a tile that is not distinct for the type).
 * Plans created afterwards use the new table. */
clm_status_t clm_expm_batched_get_tuning(int dtype, int *cfg);
clm_status_t clm_expm_batched_set_tuning(int dtype, const int *cfg);
clm_status_t clm_expm_batched_autotune(clm_ctx *ctx, clm_stream_t stream, int dtype, int nmax,
                                       int *cfg_out, float *ms_out);
/* Adjoint plans (clm_expm_batched_ad_plan_create) take their tiles from a table
 * of their own, since their launches are dominated by the backward sweep; its
 * tuner times forward + backward launches of 512 matrices per bucket. */
clm_status_t clm_expm_batched_get_tuning_ad(int dtype, int *cfg);
clm_status_t clm_expm_batched_set_tuning_ad(int dtype, const int *cfg);
clm_status_t clm_expm_batched_autotune_ad(clm_ctx *ctx, clm_stream_t stream, int dtype, int nmax,
                                          int *cfg_out, float *ms_out);

/* ---- batched adjoint (Frechet derivative backward) -----------------------------------
 * No reference counterpart. For L(E) with cotangent G = dL/dE (the conjugate
 * Wirtinger convention of PyTorch for complex types), the backward writes
 * dA = dL/dA by reverse mode through the executed chain: the squarings (from
 * `checkpoints` saved powers, squared forward again between them), the
 * Paterson-Stockmeyer Horner steps and blocks, the power bank and the scaling.
 * It equals the Frechet adjoint L(A^H, G) to within the forward's accuracy.
 *
 * An adjoint plan's forward launch (clm_expm_batched / _graph) also records a
 * tape in the workspace; clm_expm_batched_backward reads it, so the backward
 * runs after the forward of the same plan on the same workspace, with A
 * unchanged. G and dA are device matrices per entry. */
typedef struct clm_batch_ad_entry {
  int n;
  int dtype;
  const void *A;
  int lda;
  void *E;
  int lde;
  const void *G; /* cotangent of E, leading dimension ldg >= n */
  int ldg;
  void *dA;      /* gradient output, leading dimension ldda >= n */
  int ldda;
} clm_batch_ad_entry;
/* checkpoints: 1..64 saved squaring powers per matrix (the spacing adapts to
 * each matrix's s: c = ceil(s / checkpoints), recomputation costs up to c - 1
 * squarings per backward step) */
clm_status_t clm_expm_batched_ad_plan_create(clm_batch_plan **plan, int count,
                                             const clm_batch_ad_entry *entries, int checkpoints);
clm_status_t clm_expm_batched_backward(clm_ctx *ctx, clm_stream_t stream,
                                       const clm_batch_plan *plan, void *dWork,
                                       size_t lwork_bytes);

/* ---- 1-norm --------------------------------------------------------------------
 * matrix_norm1<T> (src/expm/norm1.cppm:45-56): max_j sum_i |A(i,j)| with the
 * true complex modulus; NaN propagates. dColsum is device scratch of n
 * values of the real type (float for s/c, double for d/z), which the sums
 * accumulate in. Synchronizes the stream and writes the norm to *norm (host). */
clm_status_t clm_matrix_norm1_s(clm_stream_t stream, int n, const float *dA, int lda,
                                float *dColsum, float *norm);
clm_status_t clm_matrix_norm1_c(clm_stream_t stream, int n, const clm_float_complex *dA, int lda,
                                float *dColsum, float *norm);
clm_status_t clm_matrix_norm1_d(clm_stream_t stream, int n, const double *dA, int lda,
                                double *dColsum, double *norm);
clm_status_t clm_matrix_norm1_z(clm_stream_t stream, int n, const clm_double_complex *dA, int lda,
                                double *dColsum, double *norm);

/* lange<T> (src/lapack/lange/interface.cppm:82-107): the `which` norm of an
 * m x n matrix into the DEVICE scalar dResult, asynchronously; 0 for an empty
 * matrix. which: CLM_NORM_MAX_ABS, CLM_NORM_ONE, CLM_NORM_INF,
 * CLM_NORM_FROBENIUS (src/common/enums.h:154-159). The reference allocates its
 * per-column scratch on the stream; here the caller passes dScratch of
 * max(m, n) values of the real type. */
#define CLM_NORM_MAX_ABS 0
#define CLM_NORM_ONE 1
#define CLM_NORM_INF 2
#define CLM_NORM_FROBENIUS 3
clm_status_t clm_lange_s(clm_stream_t stream, int which, size_t m, size_t n, const float *dA,
                         size_t lda, float *dResult, float *dScratch);
clm_status_t clm_lange_c(clm_stream_t stream, int which, size_t m, size_t n,
                         const clm_float_complex *dA, size_t lda, float *dResult, float *dScratch);
clm_status_t clm_lange_d(clm_stream_t stream, int which, size_t m, size_t n, const double *dA,
                         size_t lda, double *dResult, double *dScratch);
clm_status_t clm_lange_z(clm_stream_t stream, int which, size_t m, size_t n,
                         const clm_double_complex *dA, size_t lda, double *dResult,
                         double *dScratch);

/* lascl<T> TYPE='G' (src/lapack/lascl/interface.cppm:79-143): A *= cto/cfrom
 * through netlib's over/underflow-guarded chain of multipliers, in place,
 * asynchronously. A zero or NaN cfrom, or a NaN cto, is
 * CLM_STATUS_RUNTIME_INVALID_VALUE. The multipliers are formed in the real
 * type of the element (float for s/c). */
clm_status_t clm_lascl_s(clm_stream_t stream, float cfrom, float cto, size_t m, size_t n,
                         float *dA, size_t lda);
clm_status_t clm_lascl_c(clm_stream_t stream, float cfrom, float cto, size_t m, size_t n,
                         clm_float_complex *dA, size_t lda);
clm_status_t clm_lascl_d(clm_stream_t stream, double cfrom, double cto, size_t m, size_t n,
                         double *dA, size_t lda);
clm_status_t clm_lascl_z(clm_stream_t stream, double cfrom, double cto, size_t m, size_t n,
                         clm_double_complex *dA, size_t lda);

/* ---- Paterson-Stockmeyer -----------------------------------------------------------
 * PatersonStockmeyerPlan and its helpers
 * (src/paterson_stockmeyer/interface.cppm:117-228). */
typedef struct clm_ps_plan {
  int s;          /* block size */
  int r;          /* Horner steps over X = A^s */
  int num_powers; /* powers held: A^2 .. A^(num_powers+1) */
  int num_blocks; /* n x n workspace blocks */
  int num_gemms;  /* matrix products */
} clm_ps_plan;
int clm_paterson_stockmeyer_cost(int degree, int s);
int clm_paterson_stockmeyer_block_size(int degree);
clm_ps_plan clm_paterson_stockmeyer_plan(int degree, int s_requested);
size_t clm_paterson_stockmeyer_power_stride_s(int n);
size_t clm_paterson_stockmeyer_power_stride_d(int n);
size_t clm_paterson_stockmeyer_power_stride_c(int n);
size_t clm_paterson_stockmeyer_power_stride_z(int n);
size_t clm_paterson_stockmeyer_buffer_size_s(int n, int degree, int s_requested);
size_t clm_paterson_stockmeyer_buffer_size_d(int n, int degree, int s_requested);
size_t clm_paterson_stockmeyer_buffer_size_c(int n, int degree, int s_requested);
size_t clm_paterson_stockmeyer_buffer_size_z(int n, int degree, int s_requested);

/* paterson_stockmeyer<T> (src/paterson_stockmeyer/interface.cppm:329-422):
 * P = sum_{k=0}^{degree} c_k A^k, coefficients ascending in DEVICE memory,
 * asynchronous on the context's stream. */
clm_status_t clm_paterson_stockmeyer_s(clm_ctx *ctx, int n, const float *dCoeffs, int degree,
                                       const float *dA, int lda, float *dP, int ldp, void *dWork,
                                       size_t work_bytes, int s_requested);
clm_status_t clm_paterson_stockmeyer_c(clm_ctx *ctx, int n, const clm_float_complex *dCoeffs,
                                       int degree, const clm_float_complex *dA, int lda,
                                       clm_float_complex *dP, int ldp, void *dWork,
                                       size_t work_bytes, int s_requested);
clm_status_t clm_paterson_stockmeyer_d(clm_ctx *ctx, int n, const double *dCoeffs, int degree,
                                       const double *dA, int lda, double *dP, int ldp,
                                       void *dWork, size_t work_bytes, int s_requested);
clm_status_t clm_paterson_stockmeyer_z(clm_ctx *ctx, int n, const clm_double_complex *dCoeffs,
                                       int degree, const clm_double_complex *dA, int lda,
                                       clm_double_complex *dP, int ldp, void *dWork,
                                       size_t work_bytes, int s_requested);

/* horner_bufferSize<T> / horner<T> (src/horner/interface.cppm:103-194). */
size_t clm_horner_buffer_size_s(int n);
size_t clm_horner_buffer_size_d(int n);
size_t clm_horner_buffer_size_c(int n);
size_t clm_horner_buffer_size_z(int n);
clm_status_t clm_horner_s(clm_ctx *ctx, int n, const float *dCoeffs, int degree, const float *dA,
                          int lda, float *dP, int ldp, void *dWork, size_t work_bytes);
clm_status_t clm_horner_c(clm_ctx *ctx, int n, const clm_float_complex *dCoeffs, int degree,
                          const clm_float_complex *dA, int lda, clm_float_complex *dP, int ldp,
                          void *dWork, size_t work_bytes);
clm_status_t clm_horner_d(clm_ctx *ctx, int n, const double *dCoeffs, int degree,
                          const double *dA, int lda, double *dP, int ldp, void *dWork,
                          size_t work_bytes);
clm_status_t clm_horner_z(clm_ctx *ctx, int n, const clm_double_complex *dCoeffs, int degree,
                          const clm_double_complex *dA, int lda, clm_double_complex *dP, int ldp,
                          void *dWork, size_t work_bytes);

/* ---- matrix product -----------------------------------------------------------------
 * The NN product the reference takes from wwr::gemm (src/expm/pade.cppm:114-124,
 * src/paterson_stockmeyer/interface.cppm:372-373): C = alpha*A*B + beta*C for
 * an m x k A and a k x n B, on the MFMA kernels of this library, asynchronous
 * on the context's stream. beta == 0 never reads C. */
clm_status_t clm_gemm_s(clm_ctx *ctx, int m, int n, int k, float alpha, const float *dA, int lda,
                        const float *dB, int ldb, float beta, float *dC, int ldc);
clm_status_t clm_gemm_c(clm_ctx *ctx, int m, int n, int k, clm_float_complex alpha,
                        const clm_float_complex *dA, int lda, const clm_float_complex *dB,
                        int ldb, clm_float_complex beta, clm_float_complex *dC, int ldc);
clm_status_t clm_gemm_d(clm_ctx *ctx, int m, int n, int k, double alpha, const double *dA, int lda,
                        const double *dB, int ldb, double beta, double *dC, int ldc);
/* the product on BF16 MFMA with split operands (split = CLM_SPLIT_BF16X3 / X6), FP64 in/out */
clm_status_t clm_gemm_split_d(clm_ctx *ctx, int split, int m, int n, int k, double alpha,
                              const double *dA, int lda, const double *dB, int ldb, double beta,
                              double *dC, int ldc);
/* the product on FP8 MFMA (no reference counterpart): A and B are quantized to OCP
 * e4m3 with one power-of-two scale per 128 x 128 tile (the largest magnitude of a
 * tile lands in (224, 448]), multiplied with FP32 sums per 128-deep k-tile, and
 * alpha * AB + beta * C is formed in FP64. The workspace holds the e4m3 copies. */
clm_status_t clm_gemm_fp8_buffer_size_d(int m, int n, int k, size_t *bytes);
clm_status_t clm_gemm_fp8_d(clm_ctx *ctx, int m, int n, int k, double alpha, const double *dA,
                            int lda, const double *dB, int ldb, double beta, double *dC, int ldc,
                            void *dWork, size_t lwork_bytes);
clm_status_t clm_gemm_z(clm_ctx *ctx, int m, int n, int k, clm_double_complex alpha,
                        const clm_double_complex *dA, int lda, const clm_double_complex *dB,
                        int ldb, clm_double_complex beta, clm_double_complex *dC, int ldc);

/* Library version string. */
const char *clm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CALAMAN_EXPM_H */
