// capi.hip -- host driver and extern "C" entry points of the MI355X
// matrix-exponential library (declared in include/calaman_expm.h).
//
// The host side composes the kernels of gemm.hip and elementwise.hip the way
// the reference's C++ modules compose vendor BLAS calls:
//   expm      -- src/expm/expm.cppm:87-174 (validation order, workspace check,
//                one 1-norm readback, plan, 2^-s scaling, evaluation, s squarings)
//   PS        -- src/paterson_stockmeyer/interface.cppm:117-422
//   Horner    -- src/horner/interface.cppm:103-194
//   lange     -- src/lapack/lange/interface.cppm:82-107
//   lascl     -- src/lapack/lascl/interface.cppm:79-143
// The evaluation itself is the truncated Taylor series T_m(A/2^s) by
// Paterson-Stockmeyer instead of the reference's Pade approximant + LU: same
// accuracy target (backward error <= u), no LU, and every product is an MFMA GEMM.
// Argument checks come before any HIP call, so rejected calls never touch a device.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <type_traits>
#include <utility>

#include "clm_internal.hpp"
#include "expm_plan.hpp"

namespace clm {
namespace {
// the library's default context of each device (static storage, set up before
// main: resolving one allocates and writes nothing)
struct DefaultContexts {
  clm_ctx c[64];
  DefaultContexts() {
    for (int i = 0; i < 64; ++i) c[i].device = i;
  }
} g_default_ctx;
}  // namespace

clm_status_t resolve_ctx(clm_ctx *ctx, clm_ctx **out) {
  if (ctx) {
    *out = ctx;
    return CLM_STATUS_SUCCESS;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return CLM_STATUS_NOT_INITIALIZED;
  *out = &g_default_ctx.c[dev];
  return CLM_STATUS_SUCCESS;
}

namespace {

template <typename T>
constexpr int dtype_code() {
  if constexpr (std::is_same_v<T, double>) return CLM_DTYPE_F64;
  else if constexpr (std::is_same_v<T, float>) return CLM_DTYPE_F32;
  else if constexpr (std::is_same_v<T, zd>) return CLM_DTYPE_C128;
  else return CLM_DTYPE_C64;
}

template <typename T>
constexpr bool is_single() {
  return std::is_same_v<real_t<T>, float>;
}

template <typename T>
T one() {
  return ops<T>::from_real(1.0);
}

// prec: 0 = the FP64 / complex128 MFMA product; 3 or 6 = the BF16 split product
// (real FP64 only)
template <typename T>
hipError_t launch_prec(hipStream_t st, const GemmArgs<T> &g, int prec) {
  if constexpr (std::is_same_v<T, double>) {
    if (prec == CLM_SPLIT_BF16X3 || prec == CLM_SPLIT_BF16X6)
      return launch_gemm_split_bf16(st, g, prec);
  }
  return launch_gemm(st, g);
}

template <typename T>
size_t power_stride(int n) {
  const size_t order = static_cast<size_t>(n < 1 ? 1 : n);
  return Layout::align(order * order * sizeof(T)) / sizeof(T);
}

template <typename T>
size_t ps_buffer_size(int n, int degree, int s_requested) {
  const int order = n < 1 ? 1 : n;
  const clm_ps_plan p = ps_plan(degree, s_requested);
  Layout layout(nullptr);
  if (p.num_powers > 0) layout.take<T>(static_cast<size_t>(p.num_powers) * power_stride<T>(order));
  if (p.r >= 1) layout.take<T>(static_cast<size_t>(order) * order);
  return layout.total();
}

// The precision variant's schedule. Degree from the 2^-24 thresholds. The
// polynomial's products are exact to their format's precision relative to the
// size of its terms -- which grows with ||A / 2^s|| -- and an error made in
// squaring k of s (or in the polynomial) is amplified about 2^(s-k) by the
// squarings after it. So:
//   s <= 5:      polynomial products BF16 3-way (~16 bits) while ||A / 2^s||_1 <= 1,
//                else in IEEE-half planes (FP16X3: rows / columns scaled by powers
//                of two, ~22 bits, 3 products) -- BF16 6-way for orders n <= 64;
//   s = 6 .. 8:  polynomial BF16 6-way (~24 bits);
//   s > 8:       polynomial BF16 6-way, every squaring FP64;
// and the last t squarings split: t = 1 for s = 1 .. 8, 2 for s = 4, 5 with a
// half-plane polynomial, 0 for s = 0 or s > 8. Split squarings always use the half
// planes: a power whose entries pass the FP32 range (e^A of a large norm) cannot
// overflow there, as it would in BF16 planes (FP32 first).
// Measured against FP64 (n = 16 .. 2048, ||A||_1 from 0.1 to 3000, several
// seeds): at most 6.9e-7 relative error; BF16 3-way polynomials at ||A / 2^s|| > 1
// reached 1e-5 for n <= 128, half-plane polynomials at s >= 7 1.8e-6; at n <= 64
// the 6-way polynomial errs less than the half planes (n = 16, ||A||_1 = 100:
// 0.8e-7 against 7e-7), at n >= 256 more (n = 256, ||A||_1 = 1000: 2.7e-6 against
// 4e-7 at s = 5).
// The degree minimises the products weighted by their cost relative to an FP64
// product (3-product formats 0.2, 6-way 0.35), so cheap split polynomial products
// are traded for expensive FP64 squarings.
struct SplitChoice {
  int ps_split, split_squarings;
};
// BF16 3-way polynomial bound on ||A / 2^s||_1: the measured boundary is 1 (n = 16:
// 1.7e-7 there, 5.5e-7 at 2); the 2^-20 slack keeps a matrix scaled to 1-norm 1,
// whose computed column sum can land a few ulps above 1, on the measured side.
constexpr double kBf16x3MaxScaledNorm = 1.0 + 0x1p-20;
// orders up to this one take the 6-way polynomial where larger ones take half planes
constexpr int kSplitSmallOrder = 64;
// The split variant plans from the norms of powers only up to this 1-norm. Fewer
// squarings leave a larger ||A / 2^s||_1, and the polynomial's error in the
// 16-bit formats grows with it (halving with every squaring added back): measured
// at n = 2048, ||A||_1 = 1000, the half-plane polynomial errs 6.8e-7 at the powers'
// s = 4 against 0.9e-7 at the norm's s = 8, and n = 1024 at 3000 reaches 1.6e-6 at
// the powers' s = 6. Up to 256 (both benchmark norm sets) every measured case of
// the powers' plan stays within 5.3e-7; above it the plan is the norm's. (The 2^-20
// slack keeps a matrix scaled to 1-norm 256, whose computed column sum can land a
// few ulps above it, on the measured side.)
constexpr double kSplitPowerMaxNorm1 = 256.0 * (1.0 + 0x1p-20);
SplitChoice split_schedule(int s, double scaled_norm1, int n, const clm_split_opts *opts) {
  SplitChoice c{CLM_SPLIT_BF16X6, 0};
  if (s <= 5)
    c.ps_split = scaled_norm1 <= kBf16x3MaxScaledNorm ? CLM_SPLIT_BF16X3
                 : n <= kSplitSmallOrder              ? CLM_SPLIT_BF16X6
                                                      : CLM_SPLIT_FP16X3;
  if (opts && (opts->ps_split == CLM_SPLIT_BF16X3 || opts->ps_split == CLM_SPLIT_BF16X6 ||
               opts->ps_split == CLM_SPLIT_FP16X3))
    c.ps_split = opts->ps_split;
  if (s >= 1 && s <= 8)
    c.split_squarings = (c.ps_split == CLM_SPLIT_FP16X3 && (s == 4 || s == 5)) ? 2 : 1;
  if (opts && opts->split_squarings >= 0)
    c.split_squarings = opts->split_squarings < s ? opts->split_squarings : s;
  return c;
}

// The split plan of an order-n matrix from norms[k-1] = ||A^k||_1, k = 1 .. q
// (q = 1: the 1-norm alone; the norms of powers count only while ||A||_1 <=
// kSplitPowerMaxNorm1): squarings per rung as in taylor_plan_powers at u = 2^-24,
// the rung's cost weighted by format, the q - 1 powers already formed not counted
// again.
clm_expm_plan split_plan_powers(const double *pn, int q, int n, const clm_split_opts *opts,
                                double *cost_out = nullptr) {
  const double norm1 = pn[0];
  if (!(norm1 <= kSplitPowerMaxNorm1)) q = 1;
  const int have = q > 1 ? q - 1 : 0;
  int best = -1, best_s = 0;
  double best_cost = 0.0;
  for (int i = 0; i < kNumTaylor; ++i) {
    const clm_ps_plan pp = ps_plan(taylor_degree(i), 0);
    const int s = power_squarings(pn, q, taylor_degree(i), taylor_theta_at(i, true));
    if (s >= kUnusableSquarings) continue;
    const SplitChoice c = split_schedule(s, std::ldexp(norm1, -s), n, opts);
    const double w_ps = (c.ps_split == CLM_SPLIT_BF16X6) ? 0.35 : 0.2;
    const double w_sq = 0.2;  // split squarings: half planes, 3 products
    const int new_powers = pp.num_powers > have ? pp.num_powers - have : 0;
    const double cost = (new_powers + pp.r) * w_ps + (s - c.split_squarings) * 1.0 +
                        c.split_squarings * w_sq;
    if (best < 0 || cost < best_cost - 1e-12 || (std::fabs(cost - best_cost) <= 1e-12 && s < best_s)) {
      best = i;
      best_cost = cost;
      best_s = s;
    }
  }
  if (cost_out) *cost_out = best_cost;
  clm_expm_plan p{};
  const clm_ps_plan pp = ps_plan(taylor_degree(best), 0);
  p.m = taylor_degree(best);
  p.s = best_s;
  p.num_gemms = (pp.num_powers > have ? pp.num_powers : have) + pp.r + best_s;
  p.method = CLM_METHOD_TAYLOR_PS;
  p.ps_block = pp.s;
  p.norm1 = norm1;
  p.norm_powers = q < 1 ? 1 : q;
  for (int k = 2; k <= q && k <= kMaxNormPower; ++k) p.power_norm1[k - 2] = pn[k - 1];
  const SplitChoice c = split_schedule(best_s, std::ldexp(norm1, -best_s), n, opts);
  p.ps_split = c.ps_split;
  p.split_squarings = c.split_squarings;
  return p;
}

clm_expm_plan split_plan(double norm1, int n, const clm_split_opts *opts) {
  return split_plan_powers(&norm1, 1, n, opts);
}

// worth_next_power for the split variant's weighted cost: forming A^(q+1) (at the
// weight w_new of the format it is formed in) must cost less than it saves.
bool split_worth_next_power(const double *pn, int q, int n, const clm_split_opts *opts,
                            double w_new) {
  if (q < 3 || q >= kMaxNormPower) return false;
  double trial[kMaxNormPower] = {};
  for (int k = 0; k < q; ++k) trial[k] = pn[k];
  double now = 0.0, with_next = 0.0;
  (void)split_plan_powers(pn, q, n, opts, &now);
  (void)split_plan_powers(trial, q + 1, n, opts, &with_next);
  return with_next + w_new < now - 1e-12;
}

// ---- the reference's Pade ladder (src/expm/plan.cppm:51-130, expm_bridge.h:128-130) ----
double pade_theta_single(int m) {
  switch (m) {
  case 3: return static_cast<double>(4.258730034897931e-1f);
  case 5: return static_cast<double>(1.880152698533769f);
  case 7: return static_cast<double>(3.925724846433284f);
  case 9: return static_cast<double>(6.249156334514102f);
  case 13: return static_cast<double>(1.124873763647540e1f);
  default: return 0.0;
  }
}

double pade_theta(int m) {
  switch (m) {
  case 3: return 1.495585217958292e-2;
  case 5: return 2.539398330063232e-1;
  case 7: return 9.504178996162932e-1;
  case 9: return 2.097847961257067;
  case 13: return 5.371920351148152;
  default: return 0.0;
  }
}
int pade_num_gemms(int m) { return ((m == 13) ? 3 : (m - 1) / 2) + 1 + ((m == 13) ? 2 : 0); }

clm_expm_plan pade_plan(double norm1) {
  clm_expm_plan p{};
  p.method = CLM_METHOD_PADE;
  p.norm1 = norm1;
  const int degrees[5] = {3, 5, 7, 9, 13};
  for (int m : degrees) {
    if (norm1 <= pade_theta(m)) {
      p.m = m;
      p.s = 0;
      p.num_gemms = pade_num_gemms(m);
      return p;
    }
  }
  // a ratio that overflows (or a NaN) takes the same bounded answer as the Taylor
  // ladder's, not a cast of an infinite log2
  const int s = norm_squarings(norm1, pade_theta(13));
  p.m = 13;
  p.s = s;
  p.num_gemms = pade_num_gemms(13) + s;
  return p;
}

// expm_plan<float>: the comparisons and the log2 in float, as the reference's
// template does for R = float
clm_expm_plan pade_plan_single(float norm1) {
  clm_expm_plan p{};
  p.method = CLM_METHOD_PADE;
  p.norm1 = norm1;
  const int degrees[5] = {3, 5, 7, 9, 13};
  for (int m : degrees) {
    if (norm1 <= static_cast<float>(pade_theta_single(m))) {
      p.m = m;
      p.s = 0;
      p.num_gemms = pade_num_gemms(m);
      return p;
    }
  }
  const float theta = static_cast<float>(pade_theta_single(13));
  const float ratio = norm1 / theta;
  int s = 0;
  if (!(ratio <= std::numeric_limits<float>::max()))
    s = std::isnan(ratio) ? 0 : kUnusableSquarings;
  else
    s = std::max(0, static_cast<int>(std::ceil(std::log2(ratio))));
  p.m = 13;
  p.s = s;
  p.num_gemms = pade_num_gemms(13) + s;
  return p;
}

// ---- expm workspace --------------------------------------------------------------------------
//   [colsum | n doubles ]  per-column |.| sums; colsum[0] ends as the 1-norm
//   [coef   | 32 T      ]  Taylor coefficients with 2^-s folded in (see expm_impl)
//   [As     | n*n T     ]  2^-s A
//   [bank   | B x stride]  A^2 .. A^(B+1) of the scaled matrix, B = max_bank_blocks()
//   [S      | n*n T     ]  Paterson-Stockmeyer scratch, then the squaring ping-pong
// Six n*n blocks at the top of the ladder, against the reference's seven
// (src/expm/README.md:206-211); nothing for an LU.
constexpr int kCoefSlots = 32;  // Taylor coefficients 0 .. 29 (the ladder's top degree)
template <typename T>
struct ExpmWs {
  real_t<T> *colsum = nullptr;
  T *coef = nullptr;  // the coefficients with 2^-s folded in
  T *As = nullptr;
  T *bank = nullptr;
  T *S = nullptr;
  size_t stride = 0;
  size_t carve(void *base, int n) {
    Layout l(base);
    const size_t nn = static_cast<size_t>(n) * n;
    stride = power_stride<T>(n);
    colsum = l.take<real_t<T>>(static_cast<size_t>(n));
    coef = l.take<T>(kCoefSlots);
    As = l.take<T>(nn);
    bank = l.take<T>(static_cast<size_t>(max_bank_blocks()) * stride);
    S = l.take<T>(nn);
    return l.total();
  }
};

template <typename T>
size_t expm_buffer_size(int n) {
  ExpmWs<T> ws;
  return ws.carve(nullptr, n);
}

// ---- device building blocks ------------------------------------------------------------------
template <typename T>
clm_status_t gemm(hipStream_t st, int n, const T *A, long lda, const T *B, long ldb, T *C,
                  long ldc, int prec = 0) {
  GemmArgs<T> g;
  g.M = g.N = g.K = n;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.alpha = one<T>();
  CLM_TRY_HIP(launch_prec(st, g, prec));
  return CLM_STATUS_SUCCESS;
}

// The power bank A^2 .. A^k_to of A, from A^k_from on (the powers below it are
// already in the bank): bank[k-2] = bank[k-3] A.
template <typename T>
clm_status_t bank_powers(hipStream_t st, int n, const T *A, long lda, T *bank, long stride,
                         int k_from, int k_to) {
  for (int k = k_from; k <= k_to; ++k) {
    const T *left = (k == 2) ? A : bank + static_cast<long>(k - 3) * stride;
    const long ld_left = (k == 2) ? lda : n;
    CLM_TRY(gemm(st, n, left, ld_left, A, lda, bank + static_cast<long>(k - 2) * stride, n));
  }
  return CLM_STATUS_SUCCESS;
}

// Paterson-Stockmeyer on the device, the reference's structure
// (src/paterson_stockmeyer/interface.cppm:364-419) with each B_j built inside
// the epilogue of the Horner product that consumes it. The first `have` powers
// (A^2 .. A^(have+1)) are already in the bank. x_scale multiplies every Horner
// product P X (2^-(b s) when A and its bank are unscaled and the 2^-s is folded
// into the coefficients; exact, as every operand scaling here is a power of two).
template <typename T>
clm_status_t ps_evaluate(hipStream_t st, int n, const T *coef, int degree, const T *A, long lda,
                         T *P, long ldp, T *bank, long stride, T *scratch, long ld_scratch,
                         int s_requested, int have = 0, double x_scale = 1.0) {
  const clm_ps_plan p = ps_plan(degree, s_requested);
  CLM_TRY(bank_powers(st, n, A, lda, bank, stride, have + 2, p.num_powers + 1));
  const T *X = (p.r < 1) ? nullptr : (p.s >= 2) ? bank + static_cast<long>(p.s - 2) * stride : A;
  const long ldx = (p.s >= 2) ? n : lda;
  T *buf[2] = {P, scratch};
  const long ld[2] = {ldp, ld_scratch};
  int cur = (p.r % 2 == 0) ? 0 : 1;
  auto terms = [&](int j) {
    const int remaining = degree - j * p.s + 1;
    return remaining < p.s ? remaining : p.s;
  };
  CLM_TRY_HIP(launch_build_block(st, n, buf[cur], ld[cur], coef + static_cast<long>(p.r) * p.s,
                                 terms(p.r), A, lda, bank, stride, static_cast<long>(n)));
  for (int j = p.r - 1; j >= 0; --j) {
    const int next = 1 - cur;
    GemmArgs<T> g;
    g.M = g.N = g.K = n;
    g.A = buf[cur];
    g.lda = ld[cur];
    g.B = X;
    g.ldb = ldx;
    g.C = buf[next];
    g.ldc = ld[next];
    g.alpha = ops<T>::from_real(x_scale);
    g.coef = coef + static_cast<long>(j) * p.s;
    g.nterms = terms(j);
    g.X1 = A;
    g.ldx1 = lda;
    g.bank = bank;
    g.bank_stride = stride;
    g.ldbank = n;
    CLM_TRY_HIP(launch_prec(st, g, 0));
    cur = next;
  }
  return CLM_STATUS_SUCCESS;
}

// ---- the precision variant's plane workspace ----------------------------------------------
// 16-bit planes (up to 3 per matrix), each plane n*n padded to 256 bytes, with
// the power-of-two exponents the half format scales rows / columns by:
//   [PR  | 3 planes | n ints]  row-major planes of the current left operand
//   [PCA | 3 planes | n ints]  column-major planes of A/2^s (right operand of the
//                              power bank), reused for the squarings' right operand
//   [PCX | 3 planes | n ints]  column-major planes of X = (A/2^s)^b (right operand
//                              of the Paterson-Stockmeyer Horner steps)
struct PlaneSet {
  unsigned short *p = nullptr;
  mutable int *e = nullptr;  // half format: per-row (row-major set) or per-column exponents,
  mutable int gen = 0;       // tagged with the generation of the split that wrote them
};
// exponents of a matrix found before its split: by a pass ahead of time (the input's
// rows in PR.e and columns in PCA.e) or by the epilogue of the product that wrote it
// (its rows, in the spare row array); gen = 0: none
struct PreExponents {
  const double *X = nullptr;
  int *row_e = nullptr;
  int *col_e = nullptr;
  int gen = 0;
};
struct SplitWs {
  PlaneSet PR, PCA, PCX;
  long pstride = 0;
  mutable int generation = 0;  // splits so far in this run (exponent tags)
  // the spare row-exponent array: a product's epilogue writes its output's row
  // exponents there while PR.e still serves the product's own left operand; the
  // next split of that output swaps it with PR.e
  mutable int *e_alt = nullptr;
  mutable PreExponents pre;
  // Planes split ahead of need, while the host waits on a read-back: PR (PCA) holds
  // the planes of spec_pr (spec_pca) in format spec_*_fmt. The next split asked of that
  // plane set either finds them there (and skips its pass) or replaces them; either way
  // the note is spent. Anything that rewrites the matrix in place drops it (drop_spec).
  mutable const double *spec_pr = nullptr, *spec_pca = nullptr;
  mutable int spec_pr_fmt = 0, spec_pca_fmt = 0;
  void drop_spec() const {
    spec_pr = spec_pca = nullptr;
    spec_pr_fmt = spec_pca_fmt = 0;
  }
  // the run's first step when it uses the half format: every exponent to generation 0
  hipError_t start(hipStream_t st, int n) const {
    generation = 0;
    pre = PreExponents{};
    drop_spec();
    return launch_clear_exponents(st, n, PR.e, PCA.e, PCX.e, e_alt);
  }
  size_t carve(void *base, int n) {
    Layout l(base);
    pstride = static_cast<long>(Layout::align(static_cast<size_t>(n) * n * 2) / 2);
    for (PlaneSet *ps : {&PR, &PCA, &PCX}) {
      ps->p = l.take<unsigned short>(3 * static_cast<size_t>(pstride));
      ps->e = l.take<int>(static_cast<size_t>(n));
    }
    e_alt = l.take<int>(static_cast<size_t>(n));
    return l.total();
  }
};

size_t expm_split_buffer_size(int n) {
  ExpmWs<double> ws;
  SplitWs sp;
  return ws.carve(nullptr, n) + sp.carve(nullptr, n);
}

// planes of X as a right operand (pc, column-major) and / or a left operand (pr,
// row-major); the half format first finds the row / column exponents
clm_status_t split_planes(hipStream_t st, int split, int n, const double *X, long ldx,
                          const PlaneSet *pc, const PlaneSet *pr, const SplitWs &sp) {
  // planes already split ahead of need for exactly this matrix and format
  if (pr == &sp.PR) {
    if (sp.spec_pr == X && sp.spec_pr_fmt == split) pr = nullptr;
    sp.spec_pr = nullptr;
  }
  if (pc == &sp.PCA) {
    if (sp.spec_pca == X && sp.spec_pca_fmt == split) pc = nullptr;
    sp.spec_pca = nullptr;
  }
  if (!pr && !pc) return CLM_STATUS_SUCCESS;
  if (split == CLM_SPLIT_FP16X3) {
    // X's exponents, when found before (ahead of time or by the product that wrote X)
    // for every side this split needs; any split of the half format retires them
    const PreExponents q = sp.pre;
    sp.pre = PreExponents{};
    const bool reuse = q.gen > 0 && X == q.X && (!pr || (pr == &sp.PR && q.row_e)) &&
                       (!pc || (pc == &sp.PCA && q.col_e == sp.PCA.e));
    int gen;
    if (reuse) {
      gen = q.gen;
      if (pr && q.row_e != sp.PR.e) std::swap(sp.PR.e, sp.e_alt);  // the epilogue's rows
    } else {
      gen = ++sp.generation;
    }
    if (pc) pc->gen = gen;
    if (pr) pr->gen = gen;
    if (!reuse)
      CLM_TRY_HIP(launch_half_exponents(st, n, n, X, ldx, pr ? pr->e : nullptr,
                                        pc ? pc->e : nullptr, gen));
    CLM_TRY_HIP(launch_split_planes_half(st, n, n, X, ldx, pc ? pc->p : nullptr,
                                         pr ? pr->p : nullptr, sp.pstride, pr ? pr->e : nullptr,
                                         pc ? pc->e : nullptr, gen, gen));
  } else {
    CLM_TRY_HIP(launch_split_planes(st, split, n, n, X, ldx, pc ? pc->p : nullptr,
                                    pr ? pr->p : nullptr, sp.pstride));
  }
  return CLM_STATUS_SUCCESS;
}

// split X into PR (and PCA) ahead of need, noting it for the next split asked of them
clm_status_t split_ahead(hipStream_t st, int split, int n, const double *X, long ldx, bool pca,
                         const SplitWs &sp) {
  CLM_TRY(split_planes(st, split, n, X, ldx, pca ? &sp.PCA : nullptr, &sp.PR, sp));
  sp.spec_pr = X;
  sp.spec_pr_fmt = split;
  if (pca) {
    sp.spec_pca = X;
    sp.spec_pca_fmt = split;
  }
  return CLM_STATUS_SUCCESS;
}

// the split product C = PR-planes x PC-planes (+ epilogue in g); with rows_next (half
// format), the epilogue also finds C's row exponents for C's next split as a left operand
clm_status_t gemm_planes(hipStream_t st, GemmArgs<double> g, int split, const PlaneSet &a,
                         const PlaneSet &b, long pstride, const SplitWs *rows_next = nullptr) {
  if (rows_next && split == CLM_SPLIT_FP16X3) {
    g.row_e_out = rows_next->e_alt;
    g.e_gen_out = ++rows_next->generation;
  }
  CLM_TRY_HIP(launch_gemm_split_planes(st, g, split, a.p, b.p, pstride, a.e, b.e, a.gen, b.gen));
  if (g.row_e_out) rows_next->pre = PreExponents{g.C, g.row_e_out, nullptr, g.e_gen_out};
  return CLM_STATUS_SUCCESS;
}

GemmArgs<double> plain_args(int n, double *C, long ldc) {
  GemmArgs<double> g;
  g.M = g.N = g.K = n;
  g.C = C;
  g.ldc = ldc;
  g.alpha = 1.0;
  return g;
}

// The power bank on split planes, A^k_from .. A^k_to: A is split once (column
// planes as the right operand, and row planes for A^2's left), each earlier
// power once as a left operand. Returns with sp.PCA holding A's column planes
// when any product ran (*pca_is_a).
clm_status_t bank_powers_split(hipStream_t st, int split, int n, const double *A, long lda,
                               double *bank, long stride, int k_from, int k_to,
                               const SplitWs &sp, bool *pca_is_a) {
  for (int k = k_from; k <= k_to; ++k) {
    if (k == 2) {
      CLM_TRY(split_planes(st, split, n, A, lda, &sp.PCA, &sp.PR, sp));
    } else {
      if (!*pca_is_a) CLM_TRY(split_planes(st, split, n, A, lda, &sp.PCA, nullptr, sp));
      CLM_TRY(split_planes(st, split, n, bank + static_cast<long>(k - 3) * stride, n, nullptr,
                           &sp.PR, sp));
    }
    *pca_is_a = true;
    CLM_TRY(gemm_planes(st, plain_args(n, bank + static_cast<long>(k - 2) * stride, n), split,
                        sp.PR, sp.PCA, sp.pstride, &sp));
  }
  return CLM_STATUS_SUCCESS;
}

// Paterson-Stockmeyer with every product on pre-split planes: each operand is
// split once (the power bank's left factors, A and X once each, the Horner
// accumulator once per step), the products never re-split inside their loops.
// The first `have` powers are already in the bank.
clm_status_t ps_evaluate_split(hipStream_t st, int n, const double *coef, int degree,
                               const double *A, long lda, double *P, long ldp, double *bank,
                               long stride, double *scratch, long ld_scratch, int s_requested,
                               int split, const SplitWs &sp, int have = 0, double x_scale = 1.0,
                               bool pca_is_a = false) {
  // pca_is_a: sp.PCA already holds A's column planes in this format (the powers formed
  // before the plan), so neither they nor their exponents are made again
  const clm_ps_plan p = ps_plan(degree, s_requested);
  CLM_TRY(bank_powers_split(st, split, n, A, lda, bank, stride, have + 2, p.num_powers + 1, sp,
                            &pca_is_a));
  const PlaneSet *pcx = &sp.PCA;
  if (p.r >= 1 && p.s >= 2) {
    CLM_TRY(split_planes(st, split, n, bank + static_cast<long>(p.s - 2) * stride, n, &sp.PCX,
                         nullptr, sp));
    pcx = &sp.PCX;
  } else if (p.r >= 1 && !pca_is_a) {
    CLM_TRY(split_planes(st, split, n, A, lda, &sp.PCA, nullptr, sp));
  }
  double *buf[2] = {P, scratch};
  const long ld[2] = {ldp, ld_scratch};
  int cur = (p.r % 2 == 0) ? 0 : 1;
  auto terms = [&](int j) {
    const int remaining = degree - j * p.s + 1;
    return remaining < p.s ? remaining : p.s;
  };
  if (split == CLM_SPLIT_FP16X3 && p.r >= 1) {
    // the top block's row exponents come out of its own pass (into the spare row
    // array, free here) and serve its first split as a left operand
    const int gen = ++sp.generation;
    CLM_TRY_HIP(launch_build_block_rows(st, n, buf[cur], ld[cur],
                                        coef + static_cast<long>(p.r) * p.s, terms(p.r), A, lda,
                                        bank, stride, static_cast<long>(n), sp.e_alt, gen));
    sp.pre = PreExponents{buf[cur], sp.e_alt, nullptr, gen};
  } else {
    CLM_TRY_HIP(launch_build_block(st, n, buf[cur], ld[cur], coef + static_cast<long>(p.r) * p.s,
                                   terms(p.r), A, lda, bank, stride, static_cast<long>(n)));
  }
  for (int j = p.r - 1; j >= 0; --j) {
    const int next = 1 - cur;
    CLM_TRY(split_planes(st, split, n, buf[cur], ld[cur], nullptr, &sp.PR, sp));
    GemmArgs<double> g = plain_args(n, buf[next], ld[next]);
    g.alpha = x_scale;
    g.coef = coef + static_cast<long>(j) * p.s;
    g.nterms = terms(j);
    g.X1 = A;
    g.ldx1 = lda;
    g.bank = bank;
    g.bank_stride = stride;
    g.ldbank = n;
    CLM_TRY(gemm_planes(st, g, split, sp.PR, *pcx, sp.pstride, j > 0 ? &sp : nullptr));
    cur = next;
  }
  return CLM_STATUS_SUCCESS;
}

// ---- the FP8-tail variant's e4m3 workspace (after the split planes) -------------------
//   [R  | n*n bytes | tiles^2 ints]  row-major e4m3 plane of the current left operand
//   [CA | n*n bytes | tiles^2 ints]  column-major plane of A/2^s
//   [CX | n*n bytes | tiles^2 ints]  column-major plane of X = (A/2^s)^b
// with one power-of-two exponent per 128 x 128 tile (tiles = ceil(n / 128)).
struct Fp8Plane {
  unsigned char *q = nullptr;
  int *e = nullptr;
};
struct Fp8Ws {
  Fp8Plane R, CA, CX;
  size_t carve(void *base, int n) {
    Layout l(base);
    const size_t tiles = (static_cast<size_t>(n) + 127) / 128;
    for (Fp8Plane *f : {&R, &CA, &CX}) {
      f->q = l.take<unsigned char>(static_cast<size_t>(n) * n);
      f->e = l.take<int>(tiles * tiles);
    }
    return l.total();
  }
};

size_t expm_fp8tail_buffer_size(int n) {
  Fp8Ws f8;
  return expm_split_buffer_size(n) + f8.carve(nullptr, n);
}

constexpr int kFp8MinDegree = 4;

// The FP8-tail variant's power bank, A^k_from .. A^k_to: A^k = A^(k-1) A on FP8
// for k >= min_degree, on the split planes below it. *ca_split / *ca_fp8 say
// whether A's column planes / e4m3 plane exist (made once, when first needed).
clm_status_t bank_powers_fp8(hipStream_t st, int split, int min_degree, int n, const double *A,
                             long lda, double *bank, long stride, int k_from, int k_to,
                             const SplitWs &sp, const Fp8Ws &f8, bool *ca_split, bool *ca_fp8,
                             int *count) {
  for (int k = k_from; k <= k_to; ++k) {
    const double *L = (k == 2) ? A : bank + static_cast<long>(k - 3) * stride;
    const long ldl = (k == 2) ? lda : n;
    const GemmArgs<double> g = plain_args(n, bank + static_cast<long>(k - 2) * stride, n);
    if (k >= min_degree) {
      if (k == 2) {
        CLM_TRY_HIP(launch_quantize_fp8(st, n, n, A, lda, f8.CA.q, f8.R.q, f8.CA.e, f8.R.e));
      } else {
        if (!*ca_fp8)
          CLM_TRY_HIP(launch_quantize_fp8(st, n, n, A, lda, f8.CA.q, nullptr, f8.CA.e, nullptr));
        CLM_TRY_HIP(launch_quantize_fp8(st, n, n, L, ldl, nullptr, f8.R.q, nullptr, f8.R.e));
      }
      *ca_fp8 = true;
      CLM_TRY_HIP(launch_gemm_fp8(st, g, f8.R.q, f8.CA.q, f8.R.e, f8.CA.e));
      ++*count;
    } else {
      if (k == 2) {
        CLM_TRY(split_planes(st, split, n, A, lda, &sp.PCA, &sp.PR, sp));
      } else {
        if (!*ca_split) CLM_TRY(split_planes(st, split, n, A, lda, &sp.PCA, nullptr, sp));
        CLM_TRY(split_planes(st, split, n, L, ldl, nullptr, &sp.PR, sp));
      }
      *ca_split = true;
      CLM_TRY(gemm_planes(st, g, split, sp.PR, sp.PCA, sp.pstride));
    }
  }
  return CLM_STATUS_SUCCESS;
}

// Paterson-Stockmeyer of the FP8-tail variant. A product whose result holds
// only Taylor terms of degree >= min_degree -- the bank's A^k = A^(k-1) A for
// k >= min_degree, and the Horner product R_(j+1) X (X = A^b), whose lowest term
// has degree (j+1) b -- runs on FP8 MFMA from e4m3 operands with one scale per
// 128 x 128 tile; every other product on the split planes. Operand planes are
// made once per operand and format, when a product first needs them. The first
// `have` powers are already in the bank; *fp8_products counts this call's FP8
// products.
clm_status_t ps_evaluate_fp8tail(hipStream_t st, int n, const double *coef, int degree,
                                 const double *A, long lda, double *P, long ldp, double *bank,
                                 long stride, double *scratch, long ld_scratch, int s_requested,
                                 int split, int min_degree, const SplitWs &sp, const Fp8Ws &f8,
                                 int *fp8_products, int have = 0) {
  const clm_ps_plan p = ps_plan(degree, s_requested);
  const int b = p.s;
  int count = 0;
  bool ca_split = false, ca_fp8 = false;
  CLM_TRY(bank_powers_fp8(st, split, min_degree, n, A, lda, bank, stride, have + 2,
                          p.num_powers + 1, sp, f8, &ca_split, &ca_fp8, &count));
  // X = A^b for the Horner steps (A itself when the block is 1 wide)
  const bool x_is_a = !(p.r >= 1 && b >= 2);
  const double *X = x_is_a ? A : bank + static_cast<long>(b - 2) * stride;
  const long ldx = x_is_a ? lda : n;
  bool cx_split = false, cx_fp8 = false;
  double *buf[2] = {P, scratch};
  const long ld[2] = {ldp, ld_scratch};
  int cur = (p.r % 2 == 0) ? 0 : 1;
  auto terms = [&](int j) {
    const int remaining = degree - j * p.s + 1;
    return remaining < p.s ? remaining : p.s;
  };
  CLM_TRY_HIP(launch_build_block(st, n, buf[cur], ld[cur], coef + static_cast<long>(p.r) * p.s,
                                 terms(p.r), A, lda, bank, stride, static_cast<long>(n)));
  for (int j = p.r - 1; j >= 0; --j) {
    const int next = 1 - cur;
    GemmArgs<double> g = plain_args(n, buf[next], ld[next]);
    g.coef = coef + static_cast<long>(j) * p.s;
    g.nterms = terms(j);
    g.X1 = A;
    g.ldx1 = lda;
    g.bank = bank;
    g.bank_stride = stride;
    g.ldbank = n;
    if ((j + 1) * b >= min_degree) {
      bool &have = x_is_a ? ca_fp8 : cx_fp8;
      const Fp8Plane &right = x_is_a ? f8.CA : f8.CX;
      if (!have)
        CLM_TRY_HIP(launch_quantize_fp8(st, n, n, X, ldx, right.q, nullptr, right.e, nullptr));
      have = true;
      CLM_TRY_HIP(launch_quantize_fp8(st, n, n, buf[cur], ld[cur], nullptr, f8.R.q, nullptr, f8.R.e));
      CLM_TRY_HIP(launch_gemm_fp8(st, g, f8.R.q, right.q, f8.R.e, right.e));
      ++count;
    } else {
      bool &have = x_is_a ? ca_split : cx_split;
      const PlaneSet &right = x_is_a ? sp.PCA : sp.PCX;
      if (!have) CLM_TRY(split_planes(st, split, n, X, ldx, &right, nullptr, sp));
      have = true;
      CLM_TRY(split_planes(st, split, n, buf[cur], ld[cur], nullptr, &sp.PR, sp));
      CLM_TRY(gemm_planes(st, g, split, sp.PR, right, sp.pstride));
    }
    cur = next;
  }
  if (fp8_products) *fp8_products = count;
  return CLM_STATUS_SUCCESS;
}

// ---- reading norms back ----------------------------------------------------------------
// The plan waits on the device twice (the 1-norm, then the norms of the powers).
// A host-mapped, coherent slot that a one-lane kernel fills and then flags with a
// sequence number lets the host see the values ~6 us after they exist, against
// ~23 us for a pageable copy and a stream synchronization (measured on MI355X).
// Each device has a ring of slots taken in turn; the host spins on the flag for up
// to kSpinLimit, then falls back to a blocking synchronization, so long waits
// (the powers of a large matrix) do not hold a core.
struct ReadbackSlot {
  double vals[kMaxNormPower];
  unsigned seq;
  unsigned pad[21];  // 128 bytes a slot
};
constexpr int kRingSlots = 64;
struct ReadbackRing {
  ReadbackSlot *host = nullptr;
  ReadbackSlot *dev = nullptr;
  std::atomic<unsigned> next{0};
};
ReadbackRing g_ring[64];
std::mutex g_ring_mutex;
constexpr auto kSpinLimit = std::chrono::microseconds(2000);

ReadbackRing *ring_of(int device) {
  if (device < 0 || device >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(g_ring_mutex);
  ReadbackRing &r = g_ring[device];
  if (!r.host) {
    void *h = nullptr, *d = nullptr;
    if (hipHostMalloc(&h, sizeof(ReadbackSlot) * kRingSlots,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
      return nullptr;
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
      (void)hipHostFree(h);
      return nullptr;
    }
    std::memset(h, 0, sizeof(ReadbackSlot) * kRingSlots);
    r.dev = static_cast<ReadbackSlot *>(d);
    r.host = static_cast<ReadbackSlot *>(h);
  }
  return &r;
}

// out[0 .. count-1] = the device values src[0 .. count-1] once the stream reaches
// them. pinned = false takes a pageable copy and a stream synchronization and
// allocates nothing.
// `after` (when given) enqueues work that need not precede the values: it runs on the
// device while the host waits for them.
template <typename R, typename After = std::nullptr_t>
clm_status_t read_back(hipStream_t st, int device, const R *src, int count, double *out,
                       bool pinned, After after = nullptr) {
  auto launch_after = [&]() -> hipError_t {
    if constexpr (std::is_same_v<After, std::nullptr_t>) return hipSuccess;
    else return after();
  };
  ReadbackRing *ring = pinned ? ring_of(device) : nullptr;
  if (!ring) {
    R host[kMaxNormPower] = {};
    CLM_TRY_HIP(hipMemcpyAsync(host, src, sizeof(R) * count, hipMemcpyDeviceToHost, st));
    CLM_TRY_HIP(launch_after());
    CLM_TRY_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < count; ++i) out[i] = static_cast<double>(host[i]);
    return CLM_STATUS_SUCCESS;
  }
  const unsigned ticket = ring->next.fetch_add(1, std::memory_order_relaxed);
  const unsigned seq = ticket + 1 == 0 ? 1 : ticket + 1;
  const int slot = static_cast<int>(ticket % kRingSlots);
  ReadbackSlot *h = ring->host + slot;
  CLM_TRY_HIP(launch_publish(st, src, count, ring->dev[slot].vals, &ring->dev[slot].seq, seq));
  CLM_TRY_HIP(launch_after());
  const auto t0 = std::chrono::steady_clock::now();
  while (__atomic_load_n(&h->seq, __ATOMIC_ACQUIRE) != seq) {
    if (std::chrono::steady_clock::now() - t0 > kSpinLimit) {
      CLM_TRY_HIP(hipStreamSynchronize(st));
      if (__atomic_load_n(&h->seq, __ATOMIC_ACQUIRE) != seq) return CLM_STATUS_EXECUTION_FAILED;
      break;
    }
    __builtin_ia32_pause();
  }
  for (int i = 0; i < count; ++i) out[i] = h->vals[i];
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t read_norm1(hipStream_t st, int n, const T *A, long lda, real_t<T> *colsum,
                        real_t<T> *norm, int device = -1, bool pinned = false) {
  CLM_TRY_HIP(launch_norm1(st, n, A, lda, colsum));
  double v = 0;
  CLM_TRY(read_back(st, device, colsum, 1, &v, pinned));
  *norm = static_cast<real_t<T>>(v);
  return CLM_STATUS_SUCCESS;
}

// the precision variants of the FP64 driver
enum Variant { kPlain = 0, kSplit = 1, kFp8Tail = 2 };

template <typename T>
clm_status_t expm_impl(clm_ctx *ctx, hipStream_t st, int n, const T *dA, int lda, T *dE, int lde,
                       void *work, size_t bytes, int *dInfo, clm_expm_plan *plan,
                       const clm_split_opts *split = nullptr, int variant = kPlain,
                       int fp8_min_degree = kFp8MinDegree, bool may_allocate = true) {
  if (n < 1 || n > kMaxDim || lda < n || lde < n || work == nullptr || dA == nullptr ||
      dE == nullptr)
    return CLM_STATUS_INVALID_VALUE;
  const bool use_split = variant != kPlain;
  const size_t need = variant == kFp8Tail ? expm_fp8tail_buffer_size(n)
                      : use_split         ? expm_split_buffer_size(n)
                                          : expm_buffer_size<T>(n);
  if (bytes < need) return CLM_STATUS_ALLOC_FAILED;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  const T *coef = inv_factorials<T>(c->device);
  if (!coef) return CLM_STATUS_EXECUTION_FAILED;

  ExpmWs<T> ws;
  const size_t fp64_bytes = ws.carve(work, n);
  SplitWs sp;
  Fp8Ws f8;
  if (use_split) {
    const size_t split_bytes = sp.carve(static_cast<char *>(work) + fp64_bytes, n);
    if (variant == kFp8Tail) f8.carve(static_cast<char *>(work) + fp64_bytes + split_bytes, n);
  }

  // the 1-norm, read back; behind it, while the host waits, the split variants clear
  // their exponent arrays and find A's row and column exponents for the half format,
  // which A's first split then takes instead of a pass of its own (unused when the
  // plan picks the bfloat16 format)
  CLM_TRY_HIP(launch_norm1(st, n, dA, lda, ws.colsum));
  double norm_v = 0;
  auto ahead = [&]() -> hipError_t {
    if (!use_split) return hipSuccess;
    const hipError_t e = sp.start(st, n);
    if (e != hipSuccess) return e;
    if constexpr (std::is_same_v<T, double>) {
      sp.pre = PreExponents{dA, sp.PR.e, sp.PCA.e, ++sp.generation};
      return launch_half_exponents(st, n, n, dA, lda, sp.PR.e, sp.PCA.e, sp.pre.gen);
    }
    return hipSuccess;
  };
  CLM_TRY(read_back(st, c->device, ws.colsum, 1, &norm_v, may_allocate, ahead));
  const real_t<T> norm = static_cast<real_t<T>>(norm_v);
  if (!std::isfinite(norm)) return CLM_STATUS_INVALID_VALUE;

  // The plan from the 1-norm alone (the degree ladder at the element type's unit
  // roundoff, 2^-53 or 2^-24) names the power bank; when that bank holds A^3 the
  // powers are formed from A itself first, in the formats that plan gives them,
  // and their exact 1-norms -- read back together, the call's second and last
  // synchronization -- bound the truncation error through alpha_p, which lowers s
  // (and often the degree) for a matrix far from normal.
  // (The FP8 tail keeps the plan from the 1-norm: its e4m3 products err relative
  // to |X||Y|, which only a small ||A / 2^s||_1 keeps near the size of the terms.)
  const bool single_range = is_single<T>() || use_split;
  const clm_expm_plan p0 =
      use_split ? split_plan(norm, n, split) : taylor_plan(norm, is_single<T>());
  const bool norm_only = variant == kFp8Tail || (split && split->plan_from == CLM_PLAN_NORM1) ||
                         (use_split && !(double(norm) <= kSplitPowerMaxNorm1));
  int q = norm_only ? 1 : norm_power_count(double(norm), p0.m, single_range);
  double pn[kMaxNormPower] = {double(norm)};
  int fp8_count = 0;
  if (q > 1) {
    // A^2 and A^3, their norms; then A^4, A^5 one at a time while worth_next_power
    // says the power can lower the plan's cost (one readback each, no product spent
    // on a power the plan would not use)
    const long stride = static_cast<long>(ws.stride);
    bool pca_is_a = false, ca_split = false, ca_fp8 = false;
    const double w_new = p0.ps_split == CLM_SPLIT_BF16X6 ? 0.35 : 0.2;
    // the power norms land in the first bytes of S and their column sums in As
    // (both unused until the polynomial; As holds q - 1 <= 4 columns of sums for n >= 4)
    real_t<T> *dnorms = reinterpret_cast<real_t<T> *>(ws.S);
    int formed = 1;  // A^1 .. A^formed exist
    for (;;) {
      if constexpr (std::is_same_v<T, double>) {
        if (variant == kFp8Tail) {
          CLM_TRY(bank_powers_fp8(st, p0.ps_split, fp8_min_degree, n, dA, lda, ws.bank, stride,
                                  formed + 1, q, sp, f8, &ca_split, &ca_fp8, &fp8_count));
        } else if (use_split) {
          CLM_TRY(bank_powers_split(st, p0.ps_split, n, dA, lda, ws.bank, stride, formed + 1, q,
                                    sp, &pca_is_a));
        } else {
          CLM_TRY(bank_powers<T>(st, n, dA, lda, ws.bank, stride, formed + 1, q));
        }
      } else {
        CLM_TRY(bank_powers<T>(st, n, dA, lda, ws.bank, stride, formed + 1, q));
      }
      const int fresh = q - formed;  // powers A^(formed+1) .. A^q
      if (n >= 4) {
        CLM_TRY_HIP(launch_norm1_powers(st, n, ws.bank + static_cast<long>(formed - 1) * stride,
                                        stride, fresh, reinterpret_cast<real_t<T> *>(ws.As),
                                        dnorms + (formed - 1)));
      } else {
        for (int k = formed + 1; k <= q; ++k)
          CLM_TRY_HIP(launch_norm1(st, n, ws.bank + static_cast<long>(k - 2) * stride, n,
                                   ws.colsum, dnorms + (k - 2)));
      }
      // behind this wait, the split variant splits A^q as a left operand when the plan
      // from the norms known so far has a block above q (the next power A^(q+1) is then
      // likely wanted and starts from those planes; otherwise they are replaced): the
      // device would idle through the wait anyway
      auto power_ahead = [&]() -> hipError_t {
        if (variant != kSplit) return hipSuccess;
        if constexpr (std::is_same_v<T, double>) {
          const clm_expm_plan known =
              formed == 1 ? p0 : split_plan_powers(pn, formed, n, split);
          if (known.ps_block <= q) return hipSuccess;
          return split_ahead(st, p0.ps_split, n, ws.bank + static_cast<long>(q - 2) * stride, n,
                             false, sp) == CLM_STATUS_SUCCESS
                     ? hipSuccess
                     : hipErrorLaunchFailure;
        }
        return hipSuccess;
      };
      CLM_TRY(read_back(st, c->device, dnorms + (formed - 1), fresh, pn + formed, may_allocate,
                        power_ahead));
      formed = q;
      const bool more = use_split ? split_worth_next_power(pn, q, n, split, w_new)
                                  : worth_next_power(pn, q, is_single<T>());
      if (!more) break;
      ++q;
    }
  }
  clm_expm_plan p = use_split ? split_plan_powers(pn, q, n, split)
                              : taylor_plan_powers(pn, q, is_single<T>());
  p.dtype = dtype_code<T>();
  if (variant == kFp8Tail) p.fp8_min_degree = fp8_min_degree;
  if (plan) *plan = p;

  // T_m(A / 2^s). With the bank formed from A itself, A and the bank stay unscaled
  // and the 2^-s goes into the arithmetic instead: block j's coefficient of A^t is
  // c_(jb+t) 2^(-ts) and every Horner product P X (X = A^b) is scaled by 2^(-bs).
  // Every scaling is by a power of two, so each product and epilogue sum is bit for
  // bit the one formed from A / 2^s while no intermediate leaves the normal range:
  // the unscaled P A^b is 2^(bs) times the scaled one, so the fold is taken for
  // b s <= 128 in FP64 and <= 32 in the split variant (its planes' FP32 range);
  // not for FP32 / complex64, where a product of the scaled operands can reach the
  // denormal range that the unscaled one avoids (measured: one complex64 case of
  // 120 differed in its last bits). Otherwise the bank is scaled in place.
  const T *src = dA;
  long ld_src = lda;
  const T *pcoef = coef;
  double x_scale = 1.0;
  const int fold_max = is_single<T>() ? 0 : use_split ? 32 : 128;
  if (p.s > 0 && q > 1 && p.ps_block * p.s <= fold_max) {
    CLM_TRY_HIP(launch_scale_coef(st, coef, ws.coef, p.m + 1, p.ps_block, p.s));
    pcoef = ws.coef;
    x_scale = std::ldexp(1.0, -p.ps_block * p.s);
  } else if (p.s > 0) {
    CLM_TRY_HIP(launch_scale_copy(st, n, std::ldexp(1.0, -p.s), dA, lda, ws.As, n));
    CLM_TRY_HIP(launch_scale_powers(st, n, ws.bank, static_cast<long>(ws.stride), q - 1, p.s));
    sp.drop_spec();  // the powers changed in place
    src = ws.As;
    ld_src = n;
  }

  // The polynomial lands where the s squarings that follow end in dE, so no
  // final copy is needed (the reference copies when s is odd, expm.cppm:167-171).
  T *target = (p.s % 2 == 0) ? dE : ws.S;
  const long ld_target = (p.s % 2 == 0) ? lde : n;
  T *other = (p.s % 2 == 0) ? ws.S : dE;
  const long ld_other = (p.s % 2 == 0) ? n : lde;

  if constexpr (std::is_same_v<T, double>) {
    if (variant == kFp8Tail) {
      int count = 0;
      CLM_TRY(ps_evaluate_fp8tail(st, n, coef, p.m, src, ld_src, target, ld_target, ws.bank,
                                  static_cast<long>(ws.stride), other, ld_other, p.ps_block,
                                  p.ps_split, fp8_min_degree, sp, f8, &count, q - 1));
      p.fp8_gemms = fp8_count + count;
    } else if (use_split) {
      // A's column planes from the powers formed before the plan still serve when the
      // polynomial multiplies by A itself (the scaling folded, not applied to a copy) in
      // the same format
      const bool pca_valid = q > 1 && src == dA && p.ps_split == p0.ps_split;
      CLM_TRY(ps_evaluate_split(st, n, pcoef, p.m, src, ld_src, target, ld_target, ws.bank,
                                static_cast<long>(ws.stride), other, ld_other, p.ps_block,
                                p.ps_split, sp, q - 1, x_scale, pca_valid));
    } else {
      CLM_TRY(ps_evaluate<T>(st, n, pcoef, p.m, src, ld_src, target, ld_target, ws.bank,
                             static_cast<long>(ws.stride), other, ld_other, p.ps_block, q - 1,
                             x_scale));
    }
  } else {
    CLM_TRY(ps_evaluate<T>(st, n, pcoef, p.m, src, ld_src, target, ld_target, ws.bank,
                           static_cast<long>(ws.stride), other, ld_other, p.ps_block, q - 1,
                           x_scale));
  }
  if (plan) plan->fp8_gemms = p.fp8_gemms;

  T *cur = target, *nxt = other;
  long ldc = ld_target, ldn = ld_other;
  for (int k = 0; k < p.s; ++k) {
    const bool split_sq = k >= p.s - p.split_squarings;
    if constexpr (std::is_same_v<T, double>) {
      if (split_sq) {
        // split squarings in half planes (scaled per row / column: no FP32 overflow)
        const int sq = CLM_SPLIT_FP16X3;
        CLM_TRY(split_planes(st, sq, n, cur, ldc, &sp.PCA, &sp.PR, sp));
        GemmArgs<double> g;
        g.M = g.N = g.K = n;
        g.C = nxt;
        g.ldc = ldn;
        g.alpha = 1.0;
        CLM_TRY(gemm_planes(st, g, sq, sp.PR, sp.PCA, sp.pstride));
        std::swap(cur, nxt);
        std::swap(ldc, ldn);
        continue;
      }
    }
    CLM_TRY(gemm(st, n, cur, ldc, cur, ldc, nxt, ldn));
    std::swap(cur, nxt);
    std::swap(ldc, ldn);
  }
  if (dInfo) CLM_TRY_HIP(launch_fill_int(st, dInfo, 2, 0));
  return CLM_STATUS_SUCCESS;
}

// ---- the reference's Pade route on the GPU (src/expm/pade.cppm:77-191,
//      src/expm/expm.cppm:87-174) -------------------------------------------------------------
// b_j = (2m-j)! m! / ((2m)! j! (m-j)!), b_0 = 1 (src/expm/expm_bridge.h:58-113): each entry
// the correctly rounded quotient of the two integers
const double kPadeB3[4] = {1.0, 1.0 / 2.0, 1.0 / 10.0, 1.0 / 120.0};
const double kPadeB5[6] = {1.0, 1.0 / 2.0, 1.0 / 9.0, 1.0 / 72.0, 1.0 / 1008.0, 1.0 / 30240.0};
const double kPadeB7[8] = {1.0,          1.0 / 2.0,     3.0 / 26.0,     5.0 / 312.0,
                           5.0 / 3432.0, 1.0 / 11440.0, 1.0 / 308880.0, 1.0 / 17297280.0};
const double kPadeB9[10] = {1.0,          1.0 / 2.0,         2.0 / 17.0,          7.0 / 408.0,
                            7.0 / 4080.0, 1.0 / 8160.0,      1.0 / 159120.0,      1.0 / 4455360.0,
                            1.0 / 196035840.0, 1.0 / 17643225600.0};
const double kPadeB13[14] = {1.0,
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
const double *pade_coeffs(int m) {
  switch (m) {
  case 3: return kPadeB3;
  case 5: return kPadeB5;
  case 7: return kPadeB7;
  case 9: return kPadeB9;
  case 13: return kPadeB13;
  default: return nullptr;
  }
}
int pade_num_powers(int m) { return m == 13 ? 3 : (m - 1) / 2; }

// [P0 .. P(np-1) | X (m = 13) | V | W | ipiv (n ints) | info (2 ints)]: A^2, A^4, A^6,
// A^8 as the degree needs them; U = A W reuses P0 and q_m(A) = V - U reuses V
template <typename T>
struct PadeWs {
  T *P[4] = {nullptr, nullptr, nullptr, nullptr};
  T *X = nullptr, *V = nullptr, *W = nullptr;
  int *ipiv = nullptr, *info = nullptr;
  size_t carve(void *base, int n, int m) {
    Layout L(base);
    const size_t nn = static_cast<size_t>(n) * n;
    const int np = pade_num_powers(m);
    for (int k = 0; k < np; ++k) P[k] = L.take<T>(nn);
    if (m == 13) X = L.take<T>(nn);
    V = L.take<T>(nn);
    W = L.take<T>(nn);
    ipiv = L.take<int>(static_cast<size_t>(n));
    info = L.take<int>(2);
    return L.total();
  }
};

template <typename T>
size_t pade_buffer_size(int n, int m) {
  PadeWs<T> w;
  return w.carve(nullptr, n, m);
}
template <typename T>
size_t pade_worst_buffer_size(int n) {
  size_t worst = 0;
  for (int m : {3, 5, 7, 9, 13}) worst = std::max(worst, pade_buffer_size<T>(n, m));
  return worst;
}
// [As | n*n T] [colsum | n reals] [the Pade region for the worst degree]; the squarings
// ping-pong through As once the Pade evaluation has consumed it
template <typename T>
size_t expm_pade_head(int n) {
  Layout L(nullptr);
  (void)L.take<T>(static_cast<size_t>(n) * n);
  (void)L.take<real_t<T>>(static_cast<size_t>(n));
  return L.total();
}
template <typename T>
size_t expm_pade_buffer_size(int n) {
  return expm_pade_head<T>(n) + pade_worst_buffer_size<T>(n);
}

template <typename T>
hipError_t product(hipStream_t st, int n, const T *X, long ldx, const T *Y, long ldy, T *Z, long ldz,
                   bool accumulate) {
  GemmArgs<T> g;
  g.M = g.N = g.K = n;
  g.A = X;
  g.lda = ldx;
  g.B = Y;
  g.ldb = ldy;
  g.C = Z;
  g.ldc = ldz;
  g.alpha = one<T>();
  g.beta = accumulate ? one<T>() : ops<T>::zero();
  g.Cin = accumulate ? Z : nullptr;
  g.ldcin = ldz;
  return launch_gemm<T>(st, g);
}

// r_m(A) = q_m(A)^-1 p_m(A) into R: powers, the even and odd halves, U = A W, the
// split, then getrf / getrs of (V - U) X = (V + U); info[0] getrf's, info[1] = 0
template <typename T>
clm_status_t pade_run(hipStream_t st, int m, int n, const T *A, long lda, T *R, long ldr,
                      const PadeWs<T> &ws, int *info) {
  const double *b = pade_coeffs(m);
  const int np = pade_num_powers(m);
  T *const *P = ws.P;
  CLM_TRY_HIP(product<T>(st, n, A, lda, A, lda, P[0], n, false));
  if (np >= 2) CLM_TRY_HIP(product<T>(st, n, P[0], n, P[0], n, P[1], n, false));
  if (np >= 3) CLM_TRY_HIP(product<T>(st, n, P[1], n, P[0], n, P[2], n, false));
  if (np >= 4) CLM_TRY_HIP(product<T>(st, n, P[1], n, P[1], n, P[3], n, false));
  const T *Pc[4] = {P[0], P[1], P[2], P[3]};
  double cv[5] = {}, cw[5] = {};
  if (m != 13) {
    for (int i = 0; i <= np; ++i) {
      cv[i] = b[2 * i];
      cw[i] = b[2 * i + 1];
    }
    CLM_TRY_HIP(launch_pade_even_odd<T>(st, n, np, Pc, cv, cw, ws.V, ws.W));
  } else {
    // the degree-13 nested form: W = A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I,
    // V likewise with the even coefficients
    const double odd_v[4] = {0.0, b[9], b[11], b[13]}, odd_w[4] = {b[1], b[3], b[5], b[7]};
    CLM_TRY_HIP(launch_pade_even_odd<T>(st, n, 3, Pc, odd_v, odd_w, ws.X, ws.W));
    CLM_TRY_HIP(product<T>(st, n, P[2], n, ws.X, n, ws.W, n, true));
    const double even_v[4] = {0.0, b[8], b[10], b[12]}, even_w[4] = {b[0], b[2], b[4], b[6]};
    CLM_TRY_HIP(launch_pade_even_odd<T>(st, n, 3, Pc, even_v, even_w, ws.X, ws.V));
    CLM_TRY_HIP(product<T>(st, n, P[2], n, ws.X, n, ws.V, n, true));
  }
  T *U = P[0];  // the powers are dead from here
  CLM_TRY_HIP(product<T>(st, n, A, lda, ws.W, n, U, n, false));
  CLM_TRY_HIP(launch_pade_split<T>(st, n, U, ws.V, R, ldr, ws.V));
  CLM_TRY_HIP(launch_getrf<T>(st, n, ws.V, n, ws.ipiv, info));
  CLM_TRY_HIP(launch_getrs<T>(st, n, n, ws.V, n, ws.ipiv, R, ldr));
  CLM_TRY_HIP(launch_fill_int(st, info + 1, 1, 0));
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t pade_impl(clm_ctx *ctx, hipStream_t st, int m, int n, const T *dA, int lda, T *dR,
                       int ldr, void *work, size_t bytes, int *dInfo) {
  if (pade_coeffs(m) == nullptr || n < 1 || n > kMaxDim || lda < n || ldr < n || dA == nullptr ||
      dR == nullptr || work == nullptr)
    return CLM_STATUS_INVALID_VALUE;
  if (bytes < pade_buffer_size<T>(n, m)) return CLM_STATUS_ALLOC_FAILED;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  PadeWs<T> ws;
  ws.carve(work, n, m);
  return pade_run<T>(st, m, n, dA, lda, dR, ldr, ws, dInfo ? dInfo : ws.info);
}

// expm<T> as the reference evaluates it: the 1-norm (one synchronization), the Pade
// plan, 2^-s A, r_m, s squarings, the odd-s copy
template <typename T>
clm_status_t expm_pade_impl(clm_ctx *ctx, hipStream_t st, int n, const T *dA, int lda, T *dE,
                            int lde, void *work, size_t bytes, int *dInfo, clm_expm_plan *plan) {
  if (n < 1 || n > kMaxDim || lda < n || lde < n || work == nullptr || dA == nullptr ||
      dE == nullptr)
    return CLM_STATUS_INVALID_VALUE;
  if (bytes < expm_pade_buffer_size<T>(n)) return CLM_STATUS_ALLOC_FAILED;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  Layout L(work);
  T *As = L.take<T>(static_cast<size_t>(n) * n);
  real_t<T> *colsum = L.take<real_t<T>>(static_cast<size_t>(n));
  void *pade_base = static_cast<char *>(work) + L.total();
  real_t<T> norm = 0;
  CLM_TRY(read_norm1(st, n, dA, lda, colsum, &norm, c->device, true));
  if (!std::isfinite(norm)) return CLM_STATUS_INVALID_VALUE;
  clm_expm_plan p = is_single<T>() ? pade_plan_single(static_cast<float>(norm))
                                   : pade_plan(static_cast<double>(norm));
  p.dtype = dtype_code<T>();
  p.norm_powers = 1;
  if (plan) *plan = p;
  const T *src = dA;
  long ld_src = lda;
  if (p.s > 0) {
    CLM_TRY_HIP(launch_scale_copy<T>(st, n, std::ldexp(1.0, -p.s), dA, lda, As, n));
    src = As;
    ld_src = n;
  }
  PadeWs<T> ws;
  ws.carve(pade_base, n, p.m);
  CLM_TRY(pade_run<T>(st, p.m, n, src, ld_src, dE, lde, ws, dInfo ? dInfo : ws.info));
  T *cur = dE, *other = As;
  long ld_cur = lde, ld_other = n;
  for (int k = 0; k < p.s; ++k) {
    CLM_TRY_HIP(product<T>(st, n, cur, ld_cur, cur, ld_cur, other, ld_other, false));
    std::swap(cur, other);
    std::swap(ld_cur, ld_other);
  }
  if (cur != dE) CLM_TRY_HIP(launch_copy<T>(st, n, cur, ld_cur, dE, lde));
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t getrf_impl(clm_ctx *ctx, hipStream_t st, int n, T *dA, int lda, int *dIpiv,
                        int *dInfo) {
  if (n < 0 || n > kMaxDim || lda < (n > 1 ? n : 1) || dInfo == nullptr ||
      (n > 0 && (dA == nullptr || dIpiv == nullptr)))
    return CLM_STATUS_INVALID_VALUE;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  if (n == 0) {
    CLM_TRY_HIP(launch_fill_int(st, dInfo, 1, 0));
    return CLM_STATUS_SUCCESS;
  }
  CLM_TRY_HIP(launch_getrf<T>(st, n, dA, lda, dIpiv, dInfo));
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t getrs_impl(clm_ctx *ctx, hipStream_t st, int n, int nrhs, const T *dLU, int lda,
                        const int *dIpiv, T *dB, int ldb, int *dInfo) {
  if (n < 0 || nrhs < 0 || n > kMaxDim || lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1) ||
      (n > 0 && nrhs > 0 && (dLU == nullptr || dIpiv == nullptr || dB == nullptr)))
    return CLM_STATUS_INVALID_VALUE;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  CLM_TRY_HIP(launch_getrs<T>(st, n, nrhs, dLU, lda, dIpiv, dB, ldb));
  if (dInfo) CLM_TRY_HIP(launch_fill_int(st, dInfo, 1, 0));
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t ps_impl(clm_ctx *ctx, int n, const T *dCoeffs, int degree, const T *dA, int lda,
                     T *dP, int ldp, void *work, size_t work_bytes, int s_requested) {
  if (n < 1 || n > kMaxDim || degree < 0 || lda < n || ldp < n || s_requested < 0)
    return CLM_STATUS_NOT_INITIALIZED;
  if (dCoeffs == nullptr || dA == nullptr || dP == nullptr) return CLM_STATUS_NOT_INITIALIZED;
  const clm_ps_plan p = ps_plan(degree, s_requested);
  if (p.num_blocks > 0 &&
      (work == nullptr || work_bytes < ps_buffer_size<T>(n, degree, s_requested)))
    return CLM_STATUS_NOT_INITIALIZED;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  Layout l(work);
  const size_t stride = power_stride<T>(n);
  T *bank = nullptr, *scratch = nullptr;
  if (p.num_powers > 0) bank = l.take<T>(static_cast<size_t>(p.num_powers) * stride);
  if (p.r >= 1) scratch = l.take<T>(static_cast<size_t>(n) * n);
  return ps_evaluate<T>(c->stream, n, dCoeffs, degree, dA, lda, dP, ldp, bank,
                        static_cast<long>(stride), scratch, n, s_requested);
}

template <typename T>
size_t horner_buffer_size(int n) {
  const size_t order = static_cast<size_t>(n < 1 ? 1 : n);
  Layout l(nullptr);
  l.take<T>(order * order);
  return l.total();
}

// Horner (src/horner/interface.cppm:156-194): P <- c_d I, then P <- P A + c_k I,
// the diagonal add riding in the product's epilogue.
template <typename T>
clm_status_t horner_impl(clm_ctx *ctx, int n, const T *dCoeffs, int degree, const T *dA, int lda,
                         T *dP, int ldp, void *work, size_t work_bytes) {
  if (!(n >= 1 && n <= kMaxDim && degree >= 0 && lda >= n && ldp >= n))
    return CLM_STATUS_NOT_INITIALIZED;
  if (dCoeffs == nullptr || dA == nullptr || dP == nullptr) return CLM_STATUS_NOT_INITIALIZED;
  if (!(degree == 0 || (work != nullptr && work_bytes >= horner_buffer_size<T>(n))))
    return CLM_STATUS_NOT_INITIALIZED;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  hipStream_t st = c->stream;
  T *buf[2] = {dP, static_cast<T *>(work)};
  const long ld[2] = {ldp, n};
  int cur = (degree % 2 == 0) ? 0 : 1;
  CLM_TRY_HIP(launch_set_scaled_identity(st, n, buf[cur], ld[cur], dCoeffs + degree));
  for (int k = degree - 1; k >= 0; --k) {
    const int next = 1 - cur;
    GemmArgs<T> g;
    g.M = g.N = g.K = n;
    g.A = buf[cur];
    g.lda = ld[cur];
    g.B = dA;
    g.ldb = lda;
    g.C = buf[next];
    g.ldc = ld[next];
    g.alpha = one<T>();
    g.coef = dCoeffs + k;
    g.nterms = 1;
    CLM_TRY_HIP(launch_gemm(st, g));
    cur = next;
  }
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t gemm_impl(clm_ctx *ctx, int m, int n, int k, T alpha, const T *dA, int lda,
                       const T *dB, int ldb, T beta, T *dC, int ldc, int prec = 0) {
  if (m < 0 || n < 0 || k < 0 || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) ||
      ldc < (m > 1 ? m : 1))
    return CLM_STATUS_INVALID_VALUE;
  if (m == 0 || n == 0) return CLM_STATUS_SUCCESS;
  if (dC == nullptr || (k > 0 && (dA == nullptr || dB == nullptr))) return CLM_STATUS_INVALID_VALUE;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  GemmArgs<T> g;
  g.M = m;
  g.N = n;
  g.K = k;
  g.A = dA;
  g.lda = lda;
  g.B = dB;
  g.ldb = ldb;
  g.C = dC;
  g.ldc = ldc;
  g.alpha = alpha;
  g.beta = beta;
  if (!ops<T>::is_zero(beta)) {
    g.Cin = dC;
    g.ldcin = ldc;
  }
  CLM_TRY_HIP(launch_prec(c->stream, g, prec));
  return CLM_STATUS_SUCCESS;
}

// the FP8 product's workspace: A's row-major e4m3 plane and tile exponents, then B's
// column-major plane and exponents
struct Fp8GemmWs {
  unsigned char *pa = nullptr, *pb = nullptr;
  int *ea = nullptr, *eb = nullptr;
  size_t carve(void *base, int m, int n, int k) {
    Layout l(base);
    const size_t tm = (static_cast<size_t>(m) + 127) / 128, tn = (static_cast<size_t>(n) + 127) / 128,
                 tk = (static_cast<size_t>(k) + 127) / 128;
    pa = l.take<unsigned char>(static_cast<size_t>(m) * k);
    ea = l.take<int>(tm * tk);
    pb = l.take<unsigned char>(static_cast<size_t>(k) * n);
    eb = l.take<int>(tk * tn);
    return l.total();
  }
};

clm_status_t gemm_fp8_impl(clm_ctx *ctx, int m, int n, int k, double alpha, const double *dA,
                           int lda, const double *dB, int ldb, double beta, double *dC, int ldc,
                           void *work, size_t bytes) {
  if (m < 0 || n < 0 || k < 0 || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) ||
      ldc < (m > 1 ? m : 1))
    return CLM_STATUS_INVALID_VALUE;
  if (m == 0 || n == 0) return CLM_STATUS_SUCCESS;
  if (dC == nullptr || (k > 0 && (dA == nullptr || dB == nullptr))) return CLM_STATUS_INVALID_VALUE;
  Fp8GemmWs ws;
  if (bytes < ws.carve(nullptr, m, n, k) || (k > 0 && work == nullptr))
    return CLM_STATUS_ALLOC_FAILED;
  ws.carve(work, m, n, k);
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  CLM_TRY_HIP(launch_quantize_fp8(c->stream, m, k, dA, lda, nullptr, ws.pa, nullptr, ws.ea));
  CLM_TRY_HIP(launch_quantize_fp8(c->stream, k, n, dB, ldb, ws.pb, nullptr, ws.eb, nullptr));
  GemmArgs<double> g;
  g.M = m;
  g.N = n;
  g.K = k;
  g.C = dC;
  g.ldc = ldc;
  g.alpha = alpha;
  g.beta = beta;
  if (beta != 0.0) {
    g.Cin = dC;
    g.ldcin = ldc;
  }
  CLM_TRY_HIP(launch_gemm_fp8(c->stream, g, ws.pa, ws.pb, ws.ea, ws.eb));
  return CLM_STATUS_SUCCESS;
}

template <typename T>
clm_status_t norm1_impl(clm_stream_t stream, int n, const T *dA, int lda, real_t<T> *dColsum,
                        real_t<T> *norm) {
  if (n < 1 || n > kMaxDim || lda < n || dA == nullptr || dColsum == nullptr || norm == nullptr)
    return CLM_STATUS_INVALID_VALUE;
  return read_norm1(static_cast<hipStream_t>(stream), n, dA, lda, dColsum, norm);
}

template <typename T>
clm_status_t lange_impl(clm_stream_t stream, int which, size_t m, size_t n, const T *dA,
                        size_t lda, real_t<T> *dResult, real_t<T> *dScratch) {
  if (which < CLM_NORM_MAX_ABS || which > CLM_NORM_FROBENIUS || dResult == nullptr ||
      lda < (m > 1 ? m : 1))
    return CLM_STATUS_RUNTIME_INVALID_VALUE;
  if (m > 0 && n > 0 && (dA == nullptr || dScratch == nullptr))
    return CLM_STATUS_RUNTIME_INVALID_VALUE;
  if (n > 0x7fffffffu) return CLM_STATUS_RUNTIME_INVALID_VALUE;
  CLM_TRY_HIP(launch_lange(static_cast<hipStream_t>(stream), which, m, n, dA, lda, dResult,
                           dScratch));
  return CLM_STATUS_SUCCESS;
}

// ?lascl's guarded multiplier loop, verbatim in structure from netlib via the
// reference (src/lapack/lascl/interface.cppm:101-140), in the element type's
// real precision R; each factor is one pass.
template <typename T>
clm_status_t lascl_impl(clm_stream_t stream, real_t<T> cfrom, real_t<T> cto, size_t m, size_t n,
                        T *dA, size_t lda) {
  using R = real_t<T>;
  if (cfrom == R(0) || std::isnan(cfrom) || std::isnan(cto)) return CLM_STATUS_RUNTIME_INVALID_VALUE;
  if (m == 0 || n == 0) return CLM_STATUS_SUCCESS;
  if (dA == nullptr || lda < m) return CLM_STATUS_RUNTIME_INVALID_VALUE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const R smlnum = std::numeric_limits<R>::min();  // ?LAMCH('S')
  const R bignum = R(1) / smlnum;
  R cfromc = cfrom, ctoc = cto;
  bool done = false;
  do {
    const R cfrom1 = cfromc * smlnum;
    R mul;
    if (cfrom1 == cfromc) {
      mul = ctoc / cfromc;
      done = true;
    } else {
      const R cto1 = ctoc / bignum;
      if (cto1 == ctoc) {
        mul = ctoc;
        done = true;
        cfromc = R(1);
      } else if (std::fabs(cfrom1) > std::fabs(ctoc) && ctoc != R(0)) {
        mul = smlnum;
        done = false;
        cfromc = cfrom1;
      } else if (std::fabs(cto1) > std::fabs(cfromc)) {
        mul = bignum;
        done = false;
        ctoc = cto1;
      } else {
        mul = ctoc / cfromc;
        done = true;
      }
    }
    // grid rows carry columns; launch in runs of at most 65535 columns
    for (size_t j0 = 0; j0 < n; j0 += 65535) {
      const size_t nc = (n - j0 < 65535) ? n - j0 : 65535;
      if (launch_lascl(st, m, nc, mul, dA + j0 * lda, lda) != hipSuccess)
        return CLM_STATUS_RUNTIME_ERROR;
    }
  } while (!done);
  return CLM_STATUS_SUCCESS;
}

// ---- the FP64 expm entry point that allocates nothing -------------------------------------
// Its workspace is the file-scope device array `scratch`: one copy per GPU, held
// until the process ends, sized for the expm workspace at n = CLM_EXPM_NOALLOC_MAX_N.
// In host code `scratch` names a host-side stand-in, not device memory; the
// device address of each GPU's copy comes from hipGetSymbolAddress, once per
// device with that device current, and only that address is ever used. Nothing
// in the workspace must start at zero: every region is written before it is read.
constexpr size_t kNoAllocBytes = 4096 + 256 + 6 * (size_t(CLM_EXPM_NOALLOC_MAX_N) *
                                                    CLM_EXPM_NOALLOC_MAX_N * sizeof(double));
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_noalloc_mutex;
void *g_noalloc_ptr[64] = {};

// the device address of this GPU's `scratch` (looked up on the first call per device)
clm_status_t noalloc_workspace(int device, void **out) {
  if (device < 0 || device >= 64) return CLM_STATUS_NOT_INITIALIZED;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_ptr[device]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr)
      return CLM_STATUS_EXECUTION_FAILED;
    g_noalloc_ptr[device] = ptr;
  }
  *out = g_noalloc_ptr[device];
  return CLM_STATUS_SUCCESS;
}

}  // namespace

template <typename R>
clm_status_t read_device_values(hipStream_t st, int device, const R *src, int count, double *out) {
  if (src == nullptr || out == nullptr || count < 1 || count > kMaxNormPower)
    return CLM_STATUS_INVALID_VALUE;
  return read_back(st, device, src, count, out, true);
}
template clm_status_t read_device_values<double>(hipStream_t, int, const double *, int, double *);
template clm_status_t read_device_values<float>(hipStream_t, int, const float *, int, double *);

}  // namespace clm

using clm::zc;
using clm::zd;

static_assert(sizeof(clm_double_complex) == sizeof(zd), "complex layout");
static_assert(sizeof(clm_float_complex) == sizeof(zc), "complex layout");

extern "C" {

const char *clm_version(void) { return "calaman-expm-mi355x 0.1 (gfx950)"; }

const char *clm_status_string(clm_status_t s) {
  switch (s) {
  case CLM_STATUS_SUCCESS: return "CLM_STATUS_SUCCESS";
  case CLM_STATUS_NOT_INITIALIZED: return "CLM_STATUS_NOT_INITIALIZED";
  case CLM_STATUS_ALLOC_FAILED: return "CLM_STATUS_ALLOC_FAILED";
  case CLM_STATUS_INVALID_VALUE: return "CLM_STATUS_INVALID_VALUE";
  case CLM_STATUS_EXECUTION_FAILED: return "CLM_STATUS_EXECUTION_FAILED";
  case CLM_STATUS_NOT_SUPPORTED: return "CLM_STATUS_NOT_SUPPORTED";
  case CLM_STATUS_RUNTIME_INVALID_VALUE: return "CLM_STATUS_RUNTIME_INVALID_VALUE";
  case CLM_STATUS_RUNTIME_ERROR: return "CLM_STATUS_RUNTIME_ERROR";
  default: return "CLM_STATUS_UNKNOWN";
  }
}

clm_status_t clm_create(clm_ctx **ctx) {
  if (ctx == nullptr) return CLM_STATUS_INVALID_VALUE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return CLM_STATUS_NOT_INITIALIZED;
  clm_ctx *c = new clm_ctx;
  c->device = dev;
  c->stream = nullptr;
  *ctx = c;
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_destroy(clm_ctx *ctx) {
  delete ctx;
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_set_stream(clm_ctx *ctx, clm_stream_t stream) {
  if (ctx == nullptr) return CLM_STATUS_INVALID_VALUE;
  ctx->stream = static_cast<hipStream_t>(stream);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_get_stream(clm_ctx *ctx, clm_stream_t *stream) {
  if (ctx == nullptr || stream == nullptr) return CLM_STATUS_INVALID_VALUE;
  *stream = ctx->stream;
  return CLM_STATUS_SUCCESS;
}

clm_expm_plan clm_expm_plan_d(double norm1) { return clm::taylor_plan(norm1, false); }

static int clamp_norm_count(int q) {
  return q < 1 ? 1 : (q > clm::kMaxNormPower ? clm::kMaxNormPower : q);
}
clm_expm_plan clm_expm_plan_powers_d(const double *norms, int q) {
  if (norms == nullptr) return clm::taylor_plan(0.0, false);
  return clm::taylor_plan_powers(norms, clamp_norm_count(q), false);
}
clm_expm_plan clm_expm_plan_powers_s(const double *norms, int q) {
  if (norms == nullptr) return clm::taylor_plan(0.0, true);
  return clm::taylor_plan_powers(norms, clamp_norm_count(q), true);
}
clm_expm_plan clm_expm_split_plan_powers_d(int n, const double *norms, int q,
                                           const clm_split_opts *opts) {
  if (norms == nullptr) return clm::split_plan(0.0, n, opts);
  return clm::split_plan_powers(norms, clamp_norm_count(q), n, opts);
}

clm_expm_plan clm_expm_split_plan_d(double norm1, const clm_split_opts *opts) {
  return clm::split_plan(norm1, clm::kSplitSmallOrder + 1, opts);
}

double clm_taylor_theta_s(int m) {
  for (int i = 0; i < clm::kNumTaylor; ++i)
    if (clm::taylor_degree(i) == m) return clm::taylor_theta_at(i, true);
  return 0.0;
}

clm_status_t clm_expm_split_buffer_size_d(int n, size_t *bytes) {
  if (n < 1 || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_split_buffer_size(n);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_expm_split_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA, int lda,
                              double *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                              const clm_split_opts *opts, clm_expm_plan *plan) {
  if (opts && ((opts->ps_split != 0 && opts->ps_split != CLM_SPLIT_BF16X3 &&
                opts->ps_split != CLM_SPLIT_BF16X6 && opts->ps_split != CLM_SPLIT_FP16X3) ||
               opts->split_squarings < -1 ||
               (opts->plan_from != CLM_PLAN_POWERS && opts->plan_from != CLM_PLAN_NORM1)))
    return CLM_STATUS_INVALID_VALUE;
  return clm::expm_impl<double>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dE, lde, dWork,
                                lwork_bytes, dInfo, plan, opts, clm::kSplit);
}

clm_status_t clm_expm_fp8tail_buffer_size_d(int n, size_t *bytes) {
  if (n < 1 || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_fp8tail_buffer_size(n);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_expm_fp8tail_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA,
                                int lda, double *dE, int lde, void *dWork, size_t lwork_bytes,
                                int *dInfo, const clm_fp8tail_opts *opts, clm_expm_plan *plan) {
  clm_split_opts split{0, -1, CLM_PLAN_NORM1};
  int min_degree = clm::kFp8MinDegree;
  if (opts) {
    if ((opts->ps_split != 0 && opts->ps_split != CLM_SPLIT_BF16X3 &&
         opts->ps_split != CLM_SPLIT_BF16X6 && opts->ps_split != CLM_SPLIT_FP16X3) ||
        opts->split_squarings < -1 || opts->fp8_min_degree < 0 || opts->fp8_min_degree == 1)
      return CLM_STATUS_INVALID_VALUE;
    split.ps_split = opts->ps_split;
    split.split_squarings = opts->split_squarings;
    if (opts->fp8_min_degree > 0) min_degree = opts->fp8_min_degree;
  }
  return clm::expm_impl<double>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dE, lde, dWork,
                                lwork_bytes, dInfo, plan, &split, clm::kFp8Tail, min_degree);
}

double clm_taylor_theta_d(int m) {
  for (int i = 0; i < clm::kNumTaylor; ++i)
    if (clm::taylor_degree(i) == m) return clm::taylor_theta_at(i, false);
  return 0.0;
}

double clm_pade_theta_d(int m) { return clm::pade_theta(m); }
clm_expm_plan clm_pade_plan_d(double norm1) { return clm::pade_plan(norm1); }

clm_status_t clm_expm_buffer_size_d(int n, size_t *bytes) {
  if (n < 1 || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_buffer_size<double>(n);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_expm_buffer_size_z(int n, size_t *bytes) {
  if (n < 1 || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_buffer_size<zd>(n);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_expm_noalloc_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA,
                                int lda, double *dE, int lde, int *dInfo, clm_expm_plan *plan) {
  if (n < 1 || n > CLM_EXPM_NOALLOC_MAX_N) return CLM_STATUS_INVALID_VALUE;
  if (clm::expm_buffer_size<double>(n) > clm::kNoAllocBytes) return CLM_STATUS_INVALID_VALUE;
  if (lda < n || lde < n || dA == nullptr || dE == nullptr) return CLM_STATUS_INVALID_VALUE;
  int dev = 0;
  if (ctx) {
    dev = ctx->device;
  } else if (hipGetDevice(&dev) != hipSuccess) {
    return CLM_STATUS_NOT_INITIALIZED;
  }
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || cur != dev) return CLM_STATUS_INVALID_VALUE;
  void *work = nullptr;
  CLM_TRY(clm::noalloc_workspace(dev, &work));
  return clm::expm_impl<double>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dE, lde, work,
                                clm::kNoAllocBytes, dInfo, plan, nullptr, clm::kPlain,
                                clm::kFp8MinDegree, /*may_allocate=*/false);
}

clm_status_t clm_expm_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA, int lda,
                        double *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                        clm_expm_plan *plan) {
  return clm::expm_impl<double>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dE, lde, dWork,
                                lwork_bytes, dInfo, plan);
}

clm_status_t clm_expm_z(clm_ctx *ctx, clm_stream_t stream, int n, const clm_double_complex *dA,
                        int lda, clm_double_complex *dE, int lde, void *dWork, size_t lwork_bytes,
                        int *dInfo, clm_expm_plan *plan) {
  return clm::expm_impl<zd>(ctx, static_cast<hipStream_t>(stream), n,
                            reinterpret_cast<const zd *>(dA), lda, reinterpret_cast<zd *>(dE), lde,
                            dWork, lwork_bytes, dInfo, plan);
}

clm_status_t clm_matrix_norm1_d(clm_stream_t stream, int n, const double *dA, int lda,
                                double *dColsum, double *norm) {
  return clm::norm1_impl<double>(stream, n, dA, lda, dColsum, norm);
}

clm_status_t clm_matrix_norm1_z(clm_stream_t stream, int n, const clm_double_complex *dA, int lda,
                                double *dColsum, double *norm) {
  return clm::norm1_impl<zd>(stream, n, reinterpret_cast<const zd *>(dA), lda, dColsum, norm);
}

clm_status_t clm_lange_d(clm_stream_t stream, int which, size_t m, size_t n, const double *dA,
                         size_t lda, double *dResult, double *dScratch) {
  return clm::lange_impl<double>(stream, which, m, n, dA, lda, dResult, dScratch);
}

clm_status_t clm_lange_z(clm_stream_t stream, int which, size_t m, size_t n,
                         const clm_double_complex *dA, size_t lda, double *dResult,
                         double *dScratch) {
  return clm::lange_impl<zd>(stream, which, m, n, reinterpret_cast<const zd *>(dA), lda, dResult,
                             dScratch);
}

clm_status_t clm_lascl_d(clm_stream_t stream, double cfrom, double cto, size_t m, size_t n,
                         double *dA, size_t lda) {
  return clm::lascl_impl<double>(stream, cfrom, cto, m, n, dA, lda);
}

clm_status_t clm_lascl_z(clm_stream_t stream, double cfrom, double cto, size_t m, size_t n,
                         clm_double_complex *dA, size_t lda) {
  return clm::lascl_impl<zd>(stream, cfrom, cto, m, n, reinterpret_cast<zd *>(dA), lda);
}

int clm_paterson_stockmeyer_cost(int degree, int s) { return clm::ps_cost(degree, s); }
int clm_paterson_stockmeyer_block_size(int degree) { return clm::ps_block_size(degree); }
clm_ps_plan clm_paterson_stockmeyer_plan(int degree, int s_requested) {
  return clm::ps_plan(degree, s_requested);
}
size_t clm_paterson_stockmeyer_power_stride_d(int n) { return clm::power_stride<double>(n); }
size_t clm_paterson_stockmeyer_power_stride_z(int n) { return clm::power_stride<zd>(n); }
size_t clm_paterson_stockmeyer_buffer_size_d(int n, int degree, int s_requested) {
  return clm::ps_buffer_size<double>(n, degree, s_requested);
}
size_t clm_paterson_stockmeyer_buffer_size_z(int n, int degree, int s_requested) {
  return clm::ps_buffer_size<zd>(n, degree, s_requested);
}

clm_status_t clm_paterson_stockmeyer_d(clm_ctx *ctx, int n, const double *dCoeffs, int degree,
                                       const double *dA, int lda, double *dP, int ldp,
                                       void *dWork, size_t work_bytes, int s_requested) {
  return clm::ps_impl<double>(ctx, n, dCoeffs, degree, dA, lda, dP, ldp, dWork, work_bytes,
                              s_requested);
}

clm_status_t clm_paterson_stockmeyer_z(clm_ctx *ctx, int n, const clm_double_complex *dCoeffs,
                                       int degree, const clm_double_complex *dA, int lda,
                                       clm_double_complex *dP, int ldp, void *dWork,
                                       size_t work_bytes, int s_requested) {
  return clm::ps_impl<zd>(ctx, n, reinterpret_cast<const zd *>(dCoeffs), degree,
                          reinterpret_cast<const zd *>(dA), lda, reinterpret_cast<zd *>(dP), ldp,
                          dWork, work_bytes, s_requested);
}

size_t clm_horner_buffer_size_d(int n) { return clm::horner_buffer_size<double>(n); }
size_t clm_horner_buffer_size_z(int n) { return clm::horner_buffer_size<zd>(n); }

clm_status_t clm_horner_d(clm_ctx *ctx, int n, const double *dCoeffs, int degree,
                          const double *dA, int lda, double *dP, int ldp, void *dWork,
                          size_t work_bytes) {
  return clm::horner_impl<double>(ctx, n, dCoeffs, degree, dA, lda, dP, ldp, dWork, work_bytes);
}

clm_status_t clm_horner_z(clm_ctx *ctx, int n, const clm_double_complex *dCoeffs, int degree,
                          const clm_double_complex *dA, int lda, clm_double_complex *dP, int ldp,
                          void *dWork, size_t work_bytes) {
  return clm::horner_impl<zd>(ctx, n, reinterpret_cast<const zd *>(dCoeffs), degree,
                              reinterpret_cast<const zd *>(dA), lda, reinterpret_cast<zd *>(dP),
                              ldp, dWork, work_bytes);
}

clm_status_t clm_gemm_d(clm_ctx *ctx, int m, int n, int k, double alpha, const double *dA, int lda,
                        const double *dB, int ldb, double beta, double *dC, int ldc) {
  return clm::gemm_impl<double>(ctx, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
}

clm_status_t clm_gemm_fp8_buffer_size_d(int m, int n, int k, size_t *bytes) {
  if (m < 0 || n < 0 || k < 0 || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  clm::Fp8GemmWs ws;
  *bytes = ws.carve(nullptr, m, n, k);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_gemm_fp8_d(clm_ctx *ctx, int m, int n, int k, double alpha, const double *dA,
                            int lda, const double *dB, int ldb, double beta, double *dC, int ldc,
                            void *dWork, size_t lwork_bytes) {
  return clm::gemm_fp8_impl(ctx, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, dWork,
                            lwork_bytes);
}

clm_status_t clm_gemm_split_d(clm_ctx *ctx, int split, int m, int n, int k, double alpha,
                              const double *dA, int lda, const double *dB, int ldb, double beta,
                              double *dC, int ldc) {
  if (split != CLM_SPLIT_BF16X3 && split != CLM_SPLIT_BF16X6) return CLM_STATUS_INVALID_VALUE;
  return clm::gemm_impl<double>(ctx, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, split);
}

clm_status_t clm_gemm_z(clm_ctx *ctx, int m, int n, int k, clm_double_complex alpha,
                        const clm_double_complex *dA, int lda, const clm_double_complex *dB,
                        int ldb, clm_double_complex beta, clm_double_complex *dC, int ldc) {
  return clm::gemm_impl<zd>(ctx, m, n, k, zd{alpha.re, alpha.im}, reinterpret_cast<const zd *>(dA),
                            lda, reinterpret_cast<const zd *>(dB), ldb, zd{beta.re, beta.im},
                            reinterpret_cast<zd *>(dC), ldc);
}

// ---- single precision: float and complex<float> ---------------------------------------
clm_expm_plan clm_expm_plan_s(double norm1) { return clm::taylor_plan(norm1, true); }
double clm_pade_theta_s(int m) { return clm::pade_theta_single(m); }
clm_expm_plan clm_pade_plan_s(float norm1) { return clm::pade_plan_single(norm1); }

clm_status_t clm_expm_buffer_size_s(int n, size_t *bytes) {
  if (n < 1 || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_buffer_size<float>(n);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_expm_buffer_size_c(int n, size_t *bytes) {
  if (n < 1 || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_buffer_size<zc>(n);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_expm_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dA, int lda,
                        float *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                        clm_expm_plan *plan) {
  return clm::expm_impl<float>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dE, lde, dWork,
                               lwork_bytes, dInfo, plan);
}

clm_status_t clm_expm_c(clm_ctx *ctx, clm_stream_t stream, int n, const clm_float_complex *dA,
                        int lda, clm_float_complex *dE, int lde, void *dWork, size_t lwork_bytes,
                        int *dInfo, clm_expm_plan *plan) {
  return clm::expm_impl<zc>(ctx, static_cast<hipStream_t>(stream), n,
                            reinterpret_cast<const zc *>(dA), lda, reinterpret_cast<zc *>(dE), lde,
                            dWork, lwork_bytes, dInfo, plan);
}

clm_status_t clm_matrix_norm1_s(clm_stream_t stream, int n, const float *dA, int lda,
                                float *dColsum, float *norm) {
  return clm::norm1_impl<float>(stream, n, dA, lda, dColsum, norm);
}

clm_status_t clm_matrix_norm1_c(clm_stream_t stream, int n, const clm_float_complex *dA, int lda,
                                float *dColsum, float *norm) {
  return clm::norm1_impl<zc>(stream, n, reinterpret_cast<const zc *>(dA), lda, dColsum, norm);
}

clm_status_t clm_lange_s(clm_stream_t stream, int which, size_t m, size_t n, const float *dA,
                         size_t lda, float *dResult, float *dScratch) {
  return clm::lange_impl<float>(stream, which, m, n, dA, lda, dResult, dScratch);
}

clm_status_t clm_lange_c(clm_stream_t stream, int which, size_t m, size_t n,
                         const clm_float_complex *dA, size_t lda, float *dResult,
                         float *dScratch) {
  return clm::lange_impl<zc>(stream, which, m, n, reinterpret_cast<const zc *>(dA), lda, dResult,
                             dScratch);
}

clm_status_t clm_lascl_s(clm_stream_t stream, float cfrom, float cto, size_t m, size_t n,
                         float *dA, size_t lda) {
  return clm::lascl_impl<float>(stream, cfrom, cto, m, n, dA, lda);
}

clm_status_t clm_lascl_c(clm_stream_t stream, float cfrom, float cto, size_t m, size_t n,
                         clm_float_complex *dA, size_t lda) {
  return clm::lascl_impl<zc>(stream, cfrom, cto, m, n, reinterpret_cast<zc *>(dA), lda);
}

size_t clm_paterson_stockmeyer_power_stride_s(int n) { return clm::power_stride<float>(n); }
size_t clm_paterson_stockmeyer_power_stride_c(int n) { return clm::power_stride<zc>(n); }
size_t clm_paterson_stockmeyer_buffer_size_s(int n, int degree, int s_requested) {
  return clm::ps_buffer_size<float>(n, degree, s_requested);
}
size_t clm_paterson_stockmeyer_buffer_size_c(int n, int degree, int s_requested) {
  return clm::ps_buffer_size<zc>(n, degree, s_requested);
}

clm_status_t clm_paterson_stockmeyer_s(clm_ctx *ctx, int n, const float *dCoeffs, int degree,
                                       const float *dA, int lda, float *dP, int ldp, void *dWork,
                                       size_t work_bytes, int s_requested) {
  return clm::ps_impl<float>(ctx, n, dCoeffs, degree, dA, lda, dP, ldp, dWork, work_bytes,
                             s_requested);
}

clm_status_t clm_paterson_stockmeyer_c(clm_ctx *ctx, int n, const clm_float_complex *dCoeffs,
                                       int degree, const clm_float_complex *dA, int lda,
                                       clm_float_complex *dP, int ldp, void *dWork,
                                       size_t work_bytes, int s_requested) {
  return clm::ps_impl<zc>(ctx, n, reinterpret_cast<const zc *>(dCoeffs), degree,
                          reinterpret_cast<const zc *>(dA), lda, reinterpret_cast<zc *>(dP), ldp,
                          dWork, work_bytes, s_requested);
}

size_t clm_horner_buffer_size_s(int n) { return clm::horner_buffer_size<float>(n); }
size_t clm_horner_buffer_size_c(int n) { return clm::horner_buffer_size<zc>(n); }

clm_status_t clm_horner_s(clm_ctx *ctx, int n, const float *dCoeffs, int degree, const float *dA,
                          int lda, float *dP, int ldp, void *dWork, size_t work_bytes) {
  return clm::horner_impl<float>(ctx, n, dCoeffs, degree, dA, lda, dP, ldp, dWork, work_bytes);
}

clm_status_t clm_horner_c(clm_ctx *ctx, int n, const clm_float_complex *dCoeffs, int degree,
                          const clm_float_complex *dA, int lda, clm_float_complex *dP, int ldp,
                          void *dWork, size_t work_bytes) {
  return clm::horner_impl<zc>(ctx, n, reinterpret_cast<const zc *>(dCoeffs), degree,
                              reinterpret_cast<const zc *>(dA), lda, reinterpret_cast<zc *>(dP),
                              ldp, dWork, work_bytes);
}

clm_status_t clm_gemm_s(clm_ctx *ctx, int m, int n, int k, float alpha, const float *dA, int lda,
                        const float *dB, int ldb, float beta, float *dC, int ldc) {
  return clm::gemm_impl<float>(ctx, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
}

clm_status_t clm_gemm_c(clm_ctx *ctx, int m, int n, int k, clm_float_complex alpha,
                        const clm_float_complex *dA, int lda, const clm_float_complex *dB,
                        int ldb, clm_float_complex beta, clm_float_complex *dC, int ldc) {
  return clm::gemm_impl<zc>(ctx, m, n, k, zc{alpha.re, alpha.im}, reinterpret_cast<const zc *>(dA),
                            lda, reinterpret_cast<const zc *>(dB), ldb, zc{beta.re, beta.im},
                            reinterpret_cast<zc *>(dC), ldc);
}

clm_status_t clm_pade_buffer_size_s(int m, int n, size_t *bytes) {
  if (bytes == nullptr || clm::pade_coeffs(m) == nullptr || n < 1 || n > clm::kMaxDim)
    return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::pade_buffer_size<float>(n, m);
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_pade_s(clm_ctx *ctx, clm_stream_t stream, int m, int n, const float *dA, int lda,
                        float *dR, int ldr, void *dWork, size_t lwork_bytes, int *dInfo) {
  return clm::pade_impl<float>(ctx, static_cast<hipStream_t>(stream), m, n, dA, lda,
                             dR, ldr, dWork, lwork_bytes, dInfo);
}
clm_status_t clm_expm_pade_buffer_size_s(int n, size_t *bytes) {
  if (bytes == nullptr || n < 1 || n > clm::kMaxDim) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_pade_buffer_size<float>(n);
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_expm_pade_s(clm_ctx *ctx, clm_stream_t stream, int n, const float *dA, int lda,
                             float *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                             clm_expm_plan *plan) {
  return clm::expm_pade_impl<float>(ctx, static_cast<hipStream_t>(stream), n, dA, lda,
                                  dE, lde, dWork, lwork_bytes, dInfo, plan);
}
clm_status_t clm_getrf_s(clm_ctx *ctx, clm_stream_t stream, int n, float *dA, int lda, int *dIpiv,
                         int *dInfo) {
  return clm::getrf_impl<float>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dIpiv,
                              dInfo);
}
clm_status_t clm_getrs_s(clm_ctx *ctx, clm_stream_t stream, int n, int nrhs, const float *dLU,
                         int lda, const int *dIpiv, float *dB, int ldb, int *dInfo) {
  return clm::getrs_impl<float>(ctx, static_cast<hipStream_t>(stream), n, nrhs, dLU, lda,
                              dIpiv, dB, ldb, dInfo);
}

clm_status_t clm_pade_buffer_size_d(int m, int n, size_t *bytes) {
  if (bytes == nullptr || clm::pade_coeffs(m) == nullptr || n < 1 || n > clm::kMaxDim)
    return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::pade_buffer_size<double>(n, m);
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_pade_d(clm_ctx *ctx, clm_stream_t stream, int m, int n, const double *dA, int lda,
                        double *dR, int ldr, void *dWork, size_t lwork_bytes, int *dInfo) {
  return clm::pade_impl<double>(ctx, static_cast<hipStream_t>(stream), m, n, dA, lda,
                             dR, ldr, dWork, lwork_bytes, dInfo);
}
clm_status_t clm_expm_pade_buffer_size_d(int n, size_t *bytes) {
  if (bytes == nullptr || n < 1 || n > clm::kMaxDim) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_pade_buffer_size<double>(n);
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_expm_pade_d(clm_ctx *ctx, clm_stream_t stream, int n, const double *dA, int lda,
                             double *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                             clm_expm_plan *plan) {
  return clm::expm_pade_impl<double>(ctx, static_cast<hipStream_t>(stream), n, dA, lda,
                                  dE, lde, dWork, lwork_bytes, dInfo, plan);
}
clm_status_t clm_getrf_d(clm_ctx *ctx, clm_stream_t stream, int n, double *dA, int lda, int *dIpiv,
                         int *dInfo) {
  return clm::getrf_impl<double>(ctx, static_cast<hipStream_t>(stream), n, dA, lda, dIpiv,
                              dInfo);
}
clm_status_t clm_getrs_d(clm_ctx *ctx, clm_stream_t stream, int n, int nrhs, const double *dLU,
                         int lda, const int *dIpiv, double *dB, int ldb, int *dInfo) {
  return clm::getrs_impl<double>(ctx, static_cast<hipStream_t>(stream), n, nrhs, dLU, lda,
                              dIpiv, dB, ldb, dInfo);
}

clm_status_t clm_pade_buffer_size_c(int m, int n, size_t *bytes) {
  if (bytes == nullptr || clm::pade_coeffs(m) == nullptr || n < 1 || n > clm::kMaxDim)
    return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::pade_buffer_size<zc>(n, m);
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_pade_c(clm_ctx *ctx, clm_stream_t stream, int m, int n, const clm_float_complex *dA, int lda,
                        clm_float_complex *dR, int ldr, void *dWork, size_t lwork_bytes, int *dInfo) {
  return clm::pade_impl<zc>(ctx, static_cast<hipStream_t>(stream), m, n, reinterpret_cast<const zc *>(dA), lda,
                             reinterpret_cast<zc *>(dR), ldr, dWork, lwork_bytes, dInfo);
}
clm_status_t clm_expm_pade_buffer_size_c(int n, size_t *bytes) {
  if (bytes == nullptr || n < 1 || n > clm::kMaxDim) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_pade_buffer_size<zc>(n);
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_expm_pade_c(clm_ctx *ctx, clm_stream_t stream, int n, const clm_float_complex *dA, int lda,
                             clm_float_complex *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                             clm_expm_plan *plan) {
  return clm::expm_pade_impl<zc>(ctx, static_cast<hipStream_t>(stream), n, reinterpret_cast<const zc *>(dA), lda,
                                  reinterpret_cast<zc *>(dE), lde, dWork, lwork_bytes, dInfo, plan);
}
clm_status_t clm_getrf_c(clm_ctx *ctx, clm_stream_t stream, int n, clm_float_complex *dA, int lda, int *dIpiv,
                         int *dInfo) {
  return clm::getrf_impl<zc>(ctx, static_cast<hipStream_t>(stream), n, reinterpret_cast<zc *>(dA), lda, dIpiv,
                              dInfo);
}
clm_status_t clm_getrs_c(clm_ctx *ctx, clm_stream_t stream, int n, int nrhs, const clm_float_complex *dLU,
                         int lda, const int *dIpiv, clm_float_complex *dB, int ldb, int *dInfo) {
  return clm::getrs_impl<zc>(ctx, static_cast<hipStream_t>(stream), n, nrhs, reinterpret_cast<const zc *>(dLU), lda,
                              dIpiv, reinterpret_cast<zc *>(dB), ldb, dInfo);
}

clm_status_t clm_pade_buffer_size_z(int m, int n, size_t *bytes) {
  if (bytes == nullptr || clm::pade_coeffs(m) == nullptr || n < 1 || n > clm::kMaxDim)
    return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::pade_buffer_size<zd>(n, m);
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_pade_z(clm_ctx *ctx, clm_stream_t stream, int m, int n, const clm_double_complex *dA, int lda,
                        clm_double_complex *dR, int ldr, void *dWork, size_t lwork_bytes, int *dInfo) {
  return clm::pade_impl<zd>(ctx, static_cast<hipStream_t>(stream), m, n, reinterpret_cast<const zd *>(dA), lda,
                             reinterpret_cast<zd *>(dR), ldr, dWork, lwork_bytes, dInfo);
}
clm_status_t clm_expm_pade_buffer_size_z(int n, size_t *bytes) {
  if (bytes == nullptr || n < 1 || n > clm::kMaxDim) return CLM_STATUS_INVALID_VALUE;
  *bytes = clm::expm_pade_buffer_size<zd>(n);
  return CLM_STATUS_SUCCESS;
}
clm_status_t clm_expm_pade_z(clm_ctx *ctx, clm_stream_t stream, int n, const clm_double_complex *dA, int lda,
                             clm_double_complex *dE, int lde, void *dWork, size_t lwork_bytes, int *dInfo,
                             clm_expm_plan *plan) {
  return clm::expm_pade_impl<zd>(ctx, static_cast<hipStream_t>(stream), n, reinterpret_cast<const zd *>(dA), lda,
                                  reinterpret_cast<zd *>(dE), lde, dWork, lwork_bytes, dInfo, plan);
}
clm_status_t clm_getrf_z(clm_ctx *ctx, clm_stream_t stream, int n, clm_double_complex *dA, int lda, int *dIpiv,
                         int *dInfo) {
  return clm::getrf_impl<zd>(ctx, static_cast<hipStream_t>(stream), n, reinterpret_cast<zd *>(dA), lda, dIpiv,
                              dInfo);
}
clm_status_t clm_getrs_z(clm_ctx *ctx, clm_stream_t stream, int n, int nrhs, const clm_double_complex *dLU,
                         int lda, const int *dIpiv, clm_double_complex *dB, int ldb, int *dInfo) {
  return clm::getrs_impl<zd>(ctx, static_cast<hipStream_t>(stream), n, nrhs, reinterpret_cast<const zd *>(dLU), lda,
                              dIpiv, reinterpret_cast<zd *>(dB), ldb, dInfo);
}

}  // extern "C"
