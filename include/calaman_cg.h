/*
 * calaman_cg.h -- C ABI of the Riemannian conjugate gradient on U(n): the
 * caller of the matrix exponential in calaman_G (src/cg_unitary/).
 *
 * Minimises or maximises a real J(W) subject to W^H W = I by conjugate
 * gradient along geodesics W(mu) = exp(mu H) W of the unitary group, with the
 * step size from one of the two almost-periodicity line searches of Abrudan,
 * Eriksson & Koivunen, Signal Processing 89 (2009) 1704-1714 (Tables 1-3).
 * The reference's C++ templates over {float, double, complex<float>,
 * complex<double>} become type-suffixed functions (s, d, c, z) as in
 * calaman_expm.h; the real types optimise over the orthogonal group.
 *
 * Conventions kept from the reference (src/cg_unitary/cg_solver.cppm:135-176):
 *   - W and every matrix the solver reads as a flat n*n vector is packed
 *     (leading dimension exactly n);
 *   - one caller-owned workspace, sized by clm_cg_unitary_buffer_size_*;
 *   - a status of CLM_STATUS_SUCCESS whenever the iteration reached one of its
 *     own stopping conditions, MaxIterations and LineSearchFailed included
 *     (read clm_cg_info.reason); a failing status means a device call failed;
 *   - the stream is synchronized a few times per iteration (the step size and
 *     the convergence test are host control flow).
 */
#ifndef CALAMAN_CG_H
#define CALAMAN_CG_H

#include <stddef.h>

#include "calaman_expm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* CgDirection (src/cg_unitary/cost_function.cppm:21-34) */
#define CLM_CG_MINIMIZE 0
#define CLM_CG_MAXIMIZE 1

/* LineSearchMethod (src/cg_unitary/geodesic_search.cppm:57-66) */
#define CLM_CG_SEARCH_POLYNOMIAL 0 /* Table 1: degree-P fit of the derivative, first root */
#define CLM_CG_SEARCH_DFT 1        /* Table 2: Hann-windowed DFT, best of several minima */

/* CgStopReason (src/cg_unitary/cg_solver.cppm:55-61) */
#define CLM_CG_CONVERGED 0
#define CLM_CG_MAX_ITERATIONS 1
#define CLM_CG_LINE_SEARCH_FAILED 2
#define CLM_CG_NUMERICAL_FAILURE 3

/* Largest number of derivative samples one search takes
 * (src/cg_unitary/cg_unitary_bridge.h: kCgMaxSamples). */
#define CLM_CG_MAX_SAMPLES 32

/* GeodesicSearchOptions (src/cg_unitary/geodesic_search.cppm:68-98) */
typedef struct clm_cg_search_options {
  int poly_order;          /* P of Table 1, 1 .. 5 (default 5) */
  int dft_factor;          /* K of Table 2, samples per period, >= 2 (default 5) */
  int dft_periods;         /* N_T of Table 2, >= 1 (default 10) */
  int use_spectral_radius; /* |omega_max| by power iteration (1, default) or ||H||_1 (0) */
} clm_cg_search_options;

/* CgOptions (src/cg_unitary/cg_solver.cppm:64-103) */
typedef struct clm_cg_options {
  int max_iterations; /* default 300 */
  double tol;         /* stop when <G, G> = ||G||_F^2 / 2 < tol (1e-20; 1e-10 for s, c) */
  int method;         /* CLM_CG_SEARCH_* */
  clm_cg_search_options search;
  int reset_period;           /* iterations between resets of H to G; 0 = n^2 */
  int reorthogonalize_period; /* iterations between re-orthonormalizations of W; 0 = never */
} clm_cg_options;

/* CgInfo + IterationInfo (src/cg_unitary/cg_solver.cppm:105-112,
 * src/iterative/iterative.cppm:40-43) */
typedef struct clm_cg_info {
  int reason;     /* CLM_CG_* stop reason */
  int iterations; /* iterations completed */
  int resets;     /* times H was reset to the steepest direction */
  int line_search_failures;
  double gradient_norm_sq; /* final <G, G> */
  double last_step;        /* final mu */
} clm_cg_info;

/* GeodesicSearchInfo (src/cg_unitary/geodesic_search.cppm:100-109), plus the
 * derivative samples Re trace{Psi(W_i) (H W_i)^H} the search took */
typedef struct clm_cg_search_info {
  double omega_max; /* the frequency estimate that set the interval */
  double interval;  /* T_mu, or T_DFT for the DFT search */
  int num_samples;  /* derivative evaluations taken */
  int num_roots;    /* candidates the root finder returned */
  int found;        /* whether a usable step came out */
  int method_used;  /* the search that produced mu (the DFT search falls back to Table 1) */
  double samples[CLM_CG_MAX_SAMPLES];
} clm_cg_search_info;

/* ---- cost functions -----------------------------------------------------------
 * unitary_cost_function<F, T> (src/cg_unitary/cost_function.cppm:36-79) as a
 * table of callbacks on device buffers of the solver's element type (pointers
 * are void* and hold that type). Every callback enqueues on `stream` and
 * returns CLM_STATUS_SUCCESS or a failure that stops the solver. */
typedef struct clm_cg_cost {
  /* q: the degree of t in J(W + tZ); sets T_mu = 2 pi / (q |omega_max|) */
  int order;
  /* bytes of device workspace the callbacks need at order n */
  size_t (*buffer_size)(void *user, int n);
  /* Psi = dJ/dW at W (n x n, leading dimensions ldw, ldp) */
  clm_status_t (*gradient)(void *user, clm_stream_t stream, int n, const void *dW, int ldw,
                           void *dPsi, int ldp, void *dWork, size_t lwork);
  /* J(W) into one device element at dOut (only its real part is read); dPsi is
   * the gradient at W, which the callback may reuse */
  clm_status_t (*value)(void *user, clm_stream_t stream, int n, const void *dW, int ldw,
                        const void *dPsi, int ldp, void *dOut, void *dWork, size_t lwork);
  void *user;
} clm_cg_cost;

/* brockett_cost<T> (src/cg_unitary/brockett.cppm:29-130): J(W) = trace{W^H R W N},
 * R Hermitian (n x n, leading dimension ldr), N = diag(dN[0 .. n-1]); order 2,
 * Psi = R W N. The library fills *cost with its own callbacks and cost->user =
 * b; b and the device arrays it names must outlive every use of *cost. */
typedef struct clm_cg_brockett {
  const void *dR;
  int ldr;
  const void *dN;
} clm_cg_brockett;
clm_status_t clm_cg_brockett_cost_s(clm_cg_brockett *b, clm_cg_cost *cost);
clm_status_t clm_cg_brockett_cost_d(clm_cg_brockett *b, clm_cg_cost *cost);
clm_status_t clm_cg_brockett_cost_c(clm_cg_brockett *b, clm_cg_cost *cost);
clm_status_t clm_cg_brockett_cost_z(clm_cg_brockett *b, clm_cg_cost *cost);

/* the defaults of CgOptions / GeodesicSearchOptions for the element type */
void clm_cg_default_options_s(clm_cg_options *opts);
void clm_cg_default_options_d(clm_cg_options *opts);

/* ---- workspace (src/cg_unitary/buffer_size.cppm:120-146) -----------------------
 * Ten n x n blocks, O(n) vectors and scalars, and one scratch region for
 * whichever of the matrix exponential or the cost function needs more. The
 * same workspace serves clm_cg_unitary_*, clm_cg_geodesic_search_* and
 * clm_cg_geodesic_step_*. */
clm_status_t clm_cg_unitary_buffer_size_s(int n, size_t cost_bytes, size_t *bytes);
clm_status_t clm_cg_unitary_buffer_size_d(int n, size_t cost_bytes, size_t *bytes);
clm_status_t clm_cg_unitary_buffer_size_c(int n, size_t cost_bytes, size_t *bytes);
clm_status_t clm_cg_unitary_buffer_size_z(int n, size_t cost_bytes, size_t *bytes);

/* ---- the solver: cg_unitary<T, CostF> (src/cg_unitary/cg_solver.cppm:178-432) ---
 * dW: in/out n x n unitary, packed; opts NULL = the defaults; info may be NULL. */
clm_status_t clm_cg_unitary_s(clm_ctx *ctx, clm_stream_t stream, int n, float *dW,
                              const clm_cg_cost *cost, int direction, void *dWork, size_t lwork,
                              const clm_cg_options *opts, clm_cg_info *info);
clm_status_t clm_cg_unitary_d(clm_ctx *ctx, clm_stream_t stream, int n, double *dW,
                              const clm_cg_cost *cost, int direction, void *dWork, size_t lwork,
                              const clm_cg_options *opts, clm_cg_info *info);
clm_status_t clm_cg_unitary_c(clm_ctx *ctx, clm_stream_t stream, int n, clm_float_complex *dW,
                              const clm_cg_cost *cost, int direction, void *dWork, size_t lwork,
                              const clm_cg_options *opts, clm_cg_info *info);
clm_status_t clm_cg_unitary_z(clm_ctx *ctx, clm_stream_t stream, int n, clm_double_complex *dW,
                              const clm_cg_cost *cost, int direction, void *dWork, size_t lwork,
                              const clm_cg_options *opts, clm_cg_info *info);

/* ---- one line search: geodesic_search<T, CostF>
 * (src/cg_unitary/geodesic_search.cppm:448-697) -----------------------------------
 * The step mu along exp(sgn mu H) W, H in u(n) (leading dimension ldh), Psi the
 * Euclidean gradient at W (packed); *mu = 0 when no usable step exists. Each
 * call starts the power iteration from the ones vector (the solver carries its
 * warm start from search to search). opts NULL = the defaults; info may be NULL. */
clm_status_t clm_cg_geodesic_search_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dW,
                                      const float *dH, int ldh, const float *dPsi,
                                      const clm_cg_cost *cost, int direction, int method,
                                      void *dWork, size_t lwork, const clm_cg_search_options *opts,
                                      double *mu, clm_cg_search_info *info);
clm_status_t clm_cg_geodesic_search_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dW,
                                      const double *dH, int ldh, const double *dPsi,
                                      const clm_cg_cost *cost, int direction, int method,
                                      void *dWork, size_t lwork, const clm_cg_search_options *opts,
                                      double *mu, clm_cg_search_info *info);
clm_status_t clm_cg_geodesic_search_c(clm_ctx *ctx, clm_stream_t stream, int n,
                                      const clm_float_complex *dW, const clm_float_complex *dH,
                                      int ldh, const clm_float_complex *dPsi,
                                      const clm_cg_cost *cost, int direction, int method,
                                      void *dWork, size_t lwork, const clm_cg_search_options *opts,
                                      double *mu, clm_cg_search_info *info);
clm_status_t clm_cg_geodesic_search_z(clm_ctx *ctx, clm_stream_t stream, int n,
                                      const clm_double_complex *dW, const clm_double_complex *dH,
                                      int ldh, const clm_double_complex *dPsi,
                                      const clm_cg_cost *cost, int direction, int method,
                                      void *dWork, size_t lwork, const clm_cg_search_options *opts,
                                      double *mu, clm_cg_search_info *info);

/* ---- the geodesic step (src/cg_unitary/cg_solver.cppm:298-333) --------------------
 * W <- exp(t H) W for H (leading dimension ldh) and a packed W: step 5 of Table 3
 * with t = sgn mu. */
clm_status_t clm_cg_geodesic_step_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dH,
                                    int ldh, double t, float *dW, void *dWork, size_t lwork);
clm_status_t clm_cg_geodesic_step_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dH,
                                    int ldh, double t, double *dW, void *dWork, size_t lwork);
clm_status_t clm_cg_geodesic_step_c(clm_ctx *ctx, clm_stream_t stream, int n,
                                    const clm_float_complex *dH, int ldh, double t,
                                    clm_float_complex *dW, void *dWork, size_t lwork);
clm_status_t clm_cg_geodesic_step_z(clm_ctx *ctx, clm_stream_t stream, int n,
                                    const clm_double_complex *dH, int ldh, double t,
                                    clm_double_complex *dW, void *dWork, size_t lwork);

#ifdef __cplusplus
}
#endif

#endif /* CALAMAN_CG_H */
