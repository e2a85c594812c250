// Host layer of libcalaman_expm under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Built from the library's own sources with the sanitizers on their host half (the
// device code is never launched; no GPU needed) by
// `make -C calaman_g_amd/csrc host-sanitize`, and run by
// tests/test_sanitizers.py. It drives every host path that returns before a
// device call: the plan rules (Taylor ladder from the 1-norm and from the norms
// of powers, the split schedule, the Pade ladder), the workspace sizes, the
// Paterson-Stockmeyer plan arithmetic, and the argument checks of the entry
// points -- with ordinary, extreme (0, subnormal, huge, inf, NaN) and malformed
// inputs. Any sanitizer report aborts the run with a nonzero status.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <vector>

#include "calaman_cg.h"
#include "calaman_expm.h"

namespace {
int g_fail = 0;
void expect(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++g_fail;
  }
}
}  // namespace

int main() {
  const double inf = std::numeric_limits<double>::infinity();
  const double nan = std::numeric_limits<double>::quiet_NaN();
  const double norms[] = {0.0, 5e-324, 1e-300, 1e-9, 0.5, 1.0, 8.0, 100.0, 1e4, 1e100, 1e293,
                          1e300, 1.7976931348623157e308, inf, nan, -1.0};

  // plans from the 1-norm and from the norms of powers, both precisions
  for (double a : norms) {
    const clm_expm_plan p = clm_expm_plan_d(a);
    const clm_expm_plan q = clm_expm_plan_s(a);
    expect(p.s >= 0 && q.s >= 0, "plan s >= 0");
    double pn[5] = {a, a * a * 0.01, a * a * a * 1e-4, a * 1e-3, a * 1e-5};
    for (int k = 0; k <= 6; ++k) {
      (void)clm_expm_plan_powers_d(pn, k);
      (void)clm_expm_plan_powers_s(pn, k);
      for (int n : {1, 16, 64, 65, 2048}) {
        clm_split_opts o{0, -1, CLM_PLAN_POWERS};
        (void)clm_expm_split_plan_powers_d(n, pn, k, &o);
        (void)clm_expm_split_plan_powers_d(n, pn, k, nullptr);
      }
    }
    clm_split_opts o{CLM_SPLIT_BF16X6, 3, CLM_PLAN_NORM1};
    (void)clm_expm_split_plan_d(a, &o);
    (void)clm_expm_split_plan_d(a, nullptr);
    (void)clm_pade_plan_d(a);
    (void)clm_pade_plan_s(static_cast<float>(a));
  }
  for (int m = -2; m <= 40; ++m) {
    (void)clm_taylor_theta_d(m);
    (void)clm_taylor_theta_s(m);
    (void)clm_pade_theta_d(m);
    (void)clm_pade_theta_s(m);
  }

  // Paterson-Stockmeyer plan arithmetic and sizes
  for (int d = -1; d <= 40; ++d)
    for (int s = -1; s <= 45; ++s) {
      (void)clm_paterson_stockmeyer_cost(d, s);
      (void)clm_paterson_stockmeyer_plan(d, s);
      for (int n : {0, 1, 7, 300}) {
        (void)clm_paterson_stockmeyer_buffer_size_d(n, d, s);
        (void)clm_paterson_stockmeyer_buffer_size_z(n, d, s);
        (void)clm_paterson_stockmeyer_buffer_size_s(n, d, s);
        (void)clm_paterson_stockmeyer_buffer_size_c(n, d, s);
      }
    }
  for (int d = -1; d <= 40; ++d) (void)clm_paterson_stockmeyer_block_size(d);

  // workspace sizes, including the limits
  size_t bytes = 0;
  for (int n : {-1, 0, 1, 2, 63, 64, 65, 2048, 46340, 46341}) {
    (void)clm_expm_buffer_size_d(n, &bytes);
    (void)clm_expm_buffer_size_z(n, &bytes);
    (void)clm_expm_buffer_size_s(n, &bytes);
    (void)clm_expm_buffer_size_c(n, &bytes);
    (void)clm_expm_split_buffer_size_d(n, &bytes);
    (void)clm_expm_fp8tail_buffer_size_d(n, &bytes);
    (void)clm_horner_buffer_size_d(n);
    (void)clm_gemm_fp8_buffer_size_d(n, n, n, &bytes);
    for (size_t cb : {size_t(0), size_t(1) << 20})
      (void)clm_cg_unitary_buffer_size_z(n, cb, &bytes);
  }
  for (int m = 0; m <= 14; ++m)
    for (int n : {-1, 0, 1, 7, 2048, 46341}) {
      (void)clm_pade_buffer_size_d(m, n, &bytes);
      (void)clm_pade_buffer_size_z(m, n, &bytes);
      (void)clm_pade_buffer_size_s(m, n, &bytes);
      (void)clm_pade_buffer_size_c(m, n, &bytes);
      (void)clm_expm_pade_buffer_size_d(n, &bytes);
      (void)clm_expm_pade_buffer_size_c(n, &bytes);
    }
  expect(clm_expm_buffer_size_d(8, nullptr) != CLM_STATUS_SUCCESS, "null size pointer");

  // argument checks that return before any device call
  std::vector<double> host(64, 0.0);
  double *p = host.data();
  expect(clm_expm_d(nullptr, nullptr, 0, p, 1, p, 1, p, 1 << 20, nullptr, nullptr) ==
             CLM_STATUS_INVALID_VALUE,
         "expm n = 0");
  expect(clm_expm_d(nullptr, nullptr, 4, p, 3, p, 4, p, 1 << 20, nullptr, nullptr) ==
             CLM_STATUS_INVALID_VALUE,
         "expm lda < n");
  expect(clm_expm_d(nullptr, nullptr, 4, p, 4, p, 4, nullptr, 1 << 20, nullptr, nullptr) ==
             CLM_STATUS_INVALID_VALUE,
         "expm null workspace");
  expect(clm_expm_d(nullptr, nullptr, 4, p, 4, p, 4, p, 16, nullptr, nullptr) ==
             CLM_STATUS_ALLOC_FAILED,
         "expm short workspace");
  expect(clm_expm_noalloc_d(nullptr, nullptr, CLM_EXPM_NOALLOC_MAX_N + 1, p, 600, p, 600, nullptr,
                            nullptr) == CLM_STATUS_INVALID_VALUE,
         "noalloc n too large");
  clm_split_opts bad{5, -1, CLM_PLAN_POWERS};
  expect(clm_expm_split_d(nullptr, nullptr, 4, p, 4, p, 4, p, 1 << 20, nullptr, &bad, nullptr) ==
             CLM_STATUS_INVALID_VALUE,
         "split bad ps_split");
  expect(clm_paterson_stockmeyer_d(nullptr, 0, p, 3, p, 4, p, 4, p, 1 << 20, 0) ==
             CLM_STATUS_NOT_INITIALIZED,
         "ps n = 0");
  expect(clm_horner_d(nullptr, 4, p, -1, p, 4, p, 4, p, 1 << 20) == CLM_STATUS_NOT_INITIALIZED,
         "horner degree < 0");
  expect(clm_lascl_d(nullptr, 0.0, 1.0, 4, 4, p, 4) != CLM_STATUS_SUCCESS, "lascl cfrom = 0");
  expect(clm_lascl_d(nullptr, 1.0, nan, 4, 4, p, 4) != CLM_STATUS_SUCCESS, "lascl cto NaN");
  expect(clm_lange_d(nullptr, 9, 4, 4, p, 4, p, p) != CLM_STATUS_SUCCESS, "lange bad norm");
  expect(clm_gemm_split_d(nullptr, 4, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4) ==
             CLM_STATUS_INVALID_VALUE,
         "split gemm bad split");

  expect(clm_pade_d(nullptr, nullptr, 11, 4, p, 4, p, 4, p, 1 << 20, nullptr) ==
             CLM_STATUS_INVALID_VALUE,
         "pade bad degree");
  expect(clm_pade_d(nullptr, nullptr, 13, 4, p, 4, p, 4, p, 64, nullptr) == CLM_STATUS_ALLOC_FAILED,
         "pade short workspace");
  expect(clm_expm_pade_d(nullptr, nullptr, 4, p, 2, p, 4, p, 1 << 20, nullptr, nullptr) ==
             CLM_STATUS_INVALID_VALUE,
         "expm_pade lda < n");
  expect(clm_getrf_d(nullptr, nullptr, 4, p, 4, nullptr, nullptr) == CLM_STATUS_INVALID_VALUE,
         "getrf null info");
  expect(clm_getrs_d(nullptr, nullptr, 4, 1, p, 4, nullptr, p, 4, nullptr) ==
             CLM_STATUS_INVALID_VALUE,
         "getrs null ipiv");

  for (int n : {-1, 0, 1, 63, 64, 65, 2048, 46341}) {
    (void)clm_expm_herm_buffer_size_d(n, &bytes);
    (void)clm_expm_herm_buffer_size_c(n, &bytes);
  }
  expect(clm_expm_herm_d(nullptr, nullptr, 5, 4, p, 4, p, 4, p, 1 << 20, nullptr) ==
             CLM_STATUS_INVALID_VALUE,
         "expm_herm bad fill mode");
  expect(clm_expm_herm_z(nullptr, nullptr, CLM_FILL_MODE_UPPER, 4,
                         reinterpret_cast<clm_double_complex *>(p), 4,
                         reinterpret_cast<clm_double_complex *>(p), 4, p, 64,
                         nullptr) == CLM_STATUS_ALLOC_FAILED,
         "expm_herm short workspace");

  int lw = 0, ilo = 0, ihi = 0;
  expect(clm_gebal_buffer_size(7, &lw) == CLM_STATUS_SUCCESS && lw == 15, "gebal lwork");
  expect(clm_gebal_d(nullptr, 9, 4, p, 4, &ilo, &ihi, p, reinterpret_cast<int *>(p), 0) ==
             CLM_STATUS_RUNTIME_INVALID_VALUE,
         "gebal bad job");
  expect(clm_gebal_d(nullptr, CLM_GEBAL_BOTH, 0, p, 1, &ilo, &ihi, p, nullptr, 0) ==
                 CLM_STATUS_SUCCESS &&
             ilo == 1 && ihi == 0,
         "gebal n = 0");
  for (int n : {-1, 0, 1, 2048, 46341}) {
    (void)clm_expm_balanced_buffer_size_d(n, &bytes);
    (void)clm_expm_balanced_buffer_size_c(n, &bytes);
    (void)clm_lacn2_buffer_size_d(n);
  }
  expect(clm_expm_balanced_z(nullptr, nullptr, 4, reinterpret_cast<clm_double_complex *>(p), 4,
                             reinterpret_cast<clm_double_complex *>(p), 4, p, 16, nullptr,
                             nullptr) == CLM_STATUS_ALLOC_FAILED,
         "expm_balanced short workspace");
  double est = 0.0;
  int kase = 1, isave[3] = {7, 0, 0};
  expect(clm_lacn2_d(nullptr, 8, p, p, p, 1 << 10, &est, &kase, isave) == CLM_STATUS_INVALID_VALUE,
         "lacn2 bad state");

  clm_batch_plan *bp = nullptr;
  clm_batch_entry e{0, CLM_DTYPE_F64, p, 1, p, 1};
  expect(clm_expm_batched_plan_create(&bp, 1, &e) == CLM_STATUS_INVALID_VALUE, "batch n = 0");
  expect(clm_expm_batched_plan_create(&bp, 0, &e) == CLM_STATUS_INVALID_VALUE, "batch count 0");
  int cfg[CLM_BATCH_BUCKETS] = {};
  expect(clm_expm_batched_get_tuning(CLM_DTYPE_F32, cfg) == CLM_STATUS_SUCCESS, "get tuning");
  cfg[0] = 99;
  expect(clm_expm_batched_set_tuning(CLM_DTYPE_F32, cfg) == CLM_STATUS_INVALID_VALUE,
         "set tuning bad tile");

  // cg_unitary: options, Brockett table and argument checks
  clm_cg_options o{};
  clm_cg_default_options_d(&o);
  expect(o.max_iterations == 300 && o.search.poly_order == 5, "cg defaults");
  clm_cg_brockett b{p, 4, p};
  clm_cg_cost cost{};
  expect(clm_cg_brockett_cost_z(&b, &cost) == CLM_STATUS_SUCCESS && cost.order == 2,
         "brockett table");
  expect(cost.buffer_size(cost.user, 4) > 0, "brockett bytes");
  expect(clm_cg_brockett_cost_d(nullptr, &cost) == CLM_STATUS_INVALID_VALUE, "brockett null");
  clm_cg_info info{};
  expect(clm_cg_unitary_d(nullptr, nullptr, 4, p, &cost, 7, p, 1 << 20, nullptr, &info) ==
             CLM_STATUS_INVALID_VALUE,
         "cg bad direction");
  expect(clm_cg_unitary_d(nullptr, nullptr, 4, p, nullptr, CLM_CG_MAXIMIZE, p, 1 << 20, nullptr,
                          &info) == CLM_STATUS_INVALID_VALUE,
         "cg null cost");
  expect(clm_cg_unitary_d(nullptr, nullptr, 4, p, &cost, CLM_CG_MAXIMIZE, p, 64, nullptr, &info) ==
             CLM_STATUS_ALLOC_FAILED,
         "cg short workspace");
  double mu = 1.0;
  expect(clm_cg_geodesic_search_d(nullptr, nullptr, 4, p, p, 3, p, &cost, CLM_CG_MAXIMIZE,
                                  CLM_CG_SEARCH_POLYNOMIAL, p, 1 << 20, nullptr, &mu, nullptr) ==
                 CLM_STATUS_INVALID_VALUE &&
             mu == 0.0,
         "search ldh < n");
  expect(clm_cg_geodesic_step_d(nullptr, nullptr, 4, p, 4, nan, p, p, 1 << 20) ==
             CLM_STATUS_INVALID_VALUE,
         "step t NaN");

  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "ok", g_fail);
  return g_fail ? 1 : 0;
}
