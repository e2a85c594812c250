/* The CPU oracle (oracle/oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: built by `make -C oracle sanitize` and run by
 * tests/test_sanitizers.py. Every entry point on small inputs -- odd orders,
 * padded leading dimensions, zero, huge and non-finite norms, singular and
 * nilpotent matrices -- so an out-of-bounds index, a bad cast or an overflow in
 * the checker aborts the run with a nonzero status. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int g_fail = 0;
static void expect(int ok, const char *what) {
  if (!ok) {
    printf("FAILED: %s\n", what);
    ++g_fail;
  }
}

/* a deterministic n x n matrix with leading dimension ld (complex: 2 doubles an element) */
static double *matrix(int n, int ld, int cplx, double scale, unsigned seed) {
  const int w = cplx ? 2 : 1;
  double *a = calloc((size_t)ld * n * w, sizeof(double));
  unsigned x = seed * 2654435761u + 1u;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < ld * w; ++i) {
      x = x * 1664525u + 1013904223u;
      a[(size_t)j * ld * w + i] = i < n * w ? scale * ((double)(x >> 8) / 16777216.0 - 0.5) : 7.0;
    }
  return a;
}

int main(void) {
  const double inf = INFINITY, nan = NAN;
  const double norms[] = {0.0, 5e-324, 1e-9, 0.5, 8.0, 100.0, 1e100, 1e300, 1.7976931348623157e308,
                          inf, nan};
  int out[8];
  for (size_t i = 0; i < sizeof norms / sizeof norms[0]; ++i) {
    orc_expm_plan_d(norms[i], out);
    orc_expm_plan_s((float)norms[i], out);
    orc_taylor_plan_d(norms[i], out);
    orc_taylor_plan_s(norms[i], out);
    double pn[5] = {norms[i], norms[i] * 0.1, norms[i] * 0.01, norms[i] * 1e-3, norms[i] * 1e-4};
    for (int q = 1; q <= 5; ++q) {
      orc_taylor_plan_powers_d(pn, q, out);
      orc_taylor_plan_powers_s(pn, q, out);
    }
  }
  for (int m = -1; m <= 40; ++m) {
    (void)orc_pade_theta_d(m);
    (void)orc_pade_theta_s(m);
    (void)orc_taylor_theta_d(m);
    (void)orc_taylor_theta_s(m);
    (void)orc_pade_num_gemms(m);
  }
  for (int d = -1; d <= 35; ++d) {
    (void)orc_ps_block_size(d);
    for (int s = -1; s <= 40; ++s) {
      (void)orc_ps_cost(d, s);
      orc_ps_plan(d, s, out);
    }
  }
  (void)orc_inv_factorials();

  const int orders[] = {1, 2, 5, 17, 40};
  for (size_t oi = 0; oi < sizeof orders / sizeof orders[0]; ++oi) {
    const int n = orders[oi], ld = n + 3;
    for (int cplx = 0; cplx <= 1; ++cplx) {
      const int w = cplx ? 2 : 1;
      double *A = matrix(n, ld, cplx, 1.0, (unsigned)(n + cplx));
      double *B = matrix(n, ld, cplx, 1.0, (unsigned)(n + 7));
      double *C = matrix(n, ld, cplx, 1.0, (unsigned)(n + 9));
      double *E = calloc((size_t)ld * n * w, sizeof(double));
      int plan[8], info[2];
      double pnorm[4];
      for (int t = 0; t < 4; ++t) {
        const double sc = t == 0 ? 1e-3 : t == 1 ? 1.0 : t == 2 ? 40.0 : 900.0;
        double *As = matrix(n, ld, cplx, sc, (unsigned)(n * 13 + t));
        if (cplx) {
          expect(orc_expm_z(n, As, ld, E, ld, plan) == 0, "expm_z");
          expect(orc_expm_taylor_z(n, As, ld, E, ld, plan, pnorm) == 0, "expm_taylor_z");
          (void)orc_norm1_z(n, As, ld);
          for (int which = 0; which < 4; ++which) (void)orc_lange_z(which, n, n, As, ld);
        } else {
          expect(orc_expm_d(n, As, ld, E, ld, plan) == 0, "expm_d");
          expect(orc_expm_taylor_d(n, As, ld, E, ld, plan, pnorm) == 0, "expm_taylor_d");
          (void)orc_norm1_d(n, As, ld);
          for (int which = 0; which < 4; ++which) (void)orc_lange_d(which, n, n, As, ld);
        }
        free(As);
      }
      for (int m = 3; m <= 13; m += 2) {
        if (m == 11) continue;
        if (cplx) (void)orc_pade_z(m, n, A, ld, E, ld, info);
        else (void)orc_pade_d(m, n, A, ld, E, ld, info);
      }
      /* products, LU, polynomials */
      const double one[2] = {1.0, 0.0}, half[2] = {0.5, 0.25};
      if (cplx) orc_gemm_z(n, one, A, ld, B, ld, half, C, ld);
      else orc_gemm_d(n, 1.0, A, ld, B, ld, 0.5, C, ld);
      int *ipiv = calloc((size_t)n, sizeof(int));
      double *LU = matrix(n, ld, cplx, 1.0, (unsigned)(n + 11));
      double *X = matrix(n, ld, cplx, 1.0, (unsigned)(n + 12));
      if (cplx) {
        (void)orc_getrf_z(n, LU, ld, ipiv);
        (void)orc_getrs_z(n, n, LU, ld, ipiv, X, ld);
      } else {
        (void)orc_getrf_d(n, LU, ld, ipiv);
        (void)orc_getrs_d(n, n, LU, ld, ipiv, X, ld);
      }
      /* a singular matrix: info > 0, no crash */
      memset(LU, 0, sizeof(double) * (size_t)ld * n * w);
      if (cplx) expect(orc_getrf_z(n, LU, ld, ipiv) > 0, "getrf_z singular");
      else expect(orc_getrf_d(n, LU, ld, ipiv) > 0, "getrf_d singular");
      double coef[2 * 31];
      for (int k = 0; k < 2 * 31; ++k) coef[k] = 1.0 / (k + 1);
      for (int degree = 0; degree <= 30; degree += 3)
        for (int s = 0; s <= degree + 2; s += 2) {
          if (cplx) {
            (void)orc_ps_z(n, coef, degree, A, ld, E, ld, s);
            (void)orc_horner_z(n, coef, degree, A, ld, E, ld);
          } else {
            (void)orc_ps_d(n, coef, degree, A, ld, E, ld, s);
            (void)orc_horner_d(n, coef, degree, A, ld, E, ld);
          }
        }
      /* scaling with the guarded multiplier loop, extreme ratios */
      const double pairs[][2] = {{1.0, 2.0}, {1e-300, 1e300}, {1e300, 1e-300}, {3.0, 0.0}};
      for (int p = 0; p < 4; ++p) {
        if (cplx) (void)orc_lascl_z(pairs[p][0], pairs[p][1], n, n, X, ld);
        else (void)orc_lascl_d(pairs[p][0], pairs[p][1], n, n, X, ld);
      }
      expect((cplx ? orc_lascl_z(0.0, 1.0, n, n, X, ld) : orc_lascl_d(0.0, 1.0, n, n, X, ld)) != 0,
             "lascl cfrom = 0");
      /* non-finite input: rejected, not crashed */
      A[0] = nan;
      expect((cplx ? orc_expm_z(n, A, ld, E, ld, plan) : orc_expm_d(n, A, ld, E, ld, plan)) == -2,
             "expm NaN input");
      A[0] = inf;
      expect((cplx ? orc_expm_z(n, A, ld, E, ld, plan) : orc_expm_d(n, A, ld, E, ld, plan)) == -2,
             "expm inf input");
      free(A);
      free(B);
      free(C);
      free(E);
      free(ipiv);
      free(LU);
      free(X);
    }
  }
  expect(orc_expm_d(0, NULL, 1, NULL, 1, NULL) == -1, "expm n = 0");
  printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "ok", g_fail);
  return g_fail ? 1 : 0;
}
