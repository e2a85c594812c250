// expm_plan.hpp -- the plan rules of the matrix-exponential path as
// __host__ __device__ code, so the host driver (capi.hip) and the batched
// persistent kernel (batched.hip) choose the same degree, block size and
// number of squarings from the same 1-norm.
//
//   * Paterson-Stockmeyer plan, cost and block size
//     (src/paterson_stockmeyer/interface.cppm:117-210);
//   * the Taylor ladder: theta_m is the largest x with
//     sum_{k>m} |h_k| x^(k-1) <= u for h(x) = log(exp(-x) T_m(x)) -- the
//     backward-error criterion the reference applies to its Pade rungs
//     (src/expm/plan.cppm:29-41) -- evaluated by an exact rational series and
//     bisection, at u = 2^-53 and u = 2^-24. The degrees are those where the
//     Paterson-Stockmeyer product count rises by one. The plan minimises
//     products + s with s = max(0, ceil(log2(norm1 / theta_m))), ties to the
//     smaller s (the role of expm_plan<T>, src/expm/plan.cppm:112-130).
#pragma once

#include <hip/hip_runtime.h>

#include "calaman_expm.h"

namespace clm {

__host__ __device__ constexpr int ps_cost(int degree, int s) {
  if (s < 1 || degree < 1) return 0;
  const int r = degree / s;
  const int max_power = (r >= 1) ? s : degree;
  const int num_powers = (max_power >= 2) ? (max_power - 1) : 0;
  return num_powers + r;
}

__host__ __device__ constexpr int ps_block_size(int degree) {
  if (degree < 1) return 1;
  int best_s = 1, best = ps_cost(degree, 1);
  for (int s = 2; s <= degree + 1; ++s) {
    const int c = ps_cost(degree, s);
    if (c < best) {
      best = c;
      best_s = s;
    }
  }
  return best_s;
}

__host__ __device__ constexpr clm_ps_plan ps_plan(int degree, int s_requested) {
  clm_ps_plan p{1, 0, 0, 0, 0};
  if (degree < 0) return p;
  int s = (s_requested > 0) ? s_requested : ps_block_size(degree);
  if (s > degree + 1) s = degree + 1;
  if (s < 1) s = 1;
  p.s = s;
  p.r = degree / s;
  const int max_power = (p.r >= 1) ? s : degree;
  p.num_powers = (max_power >= 2) ? (max_power - 1) : 0;
  p.num_blocks = p.num_powers + ((p.r >= 1) ? 1 : 0);
  p.num_gemms = p.num_powers + p.r;
  return p;
}

constexpr int kNumTaylor = 10;

__host__ __device__ constexpr int taylor_degree(int i) {
  switch (i) {
  case 0: return 1;
  case 1: return 2;
  case 2: return 3;
  case 3: return 5;
  case 4: return 8;
  case 5: return 11;
  case 6: return 15;
  case 7: return 19;
  case 8: return 24;
  default: return 29;
  }
}

// theta_m at u = 2^-53 (single = false) and u = 2^-24 (single = true)
__host__ __device__ inline double taylor_theta_at(int i, bool single) {
  if (single) {
    switch (i) {
    case 0: return 1.192092800768788e-7;
    case 1: return 5.978858893805233e-4;
    case 2: return 1.123386473528671e-2;
    case 3: return 1.30848716459947e-1;
    case 4: return 5.800524627688768e-1;
    case 5: return 1.223479542424143;
    case 6: return 2.21704439497472;
    case 7: return 3.279561212635997;
    case 8: return 4.649782224100757;
    default: return 6.041358758192571;
    }
  }
  switch (i) {
  case 0: return 2.220446049250313e-16;
  case 1: return 2.580956802971767e-8;
  case 2: return 1.386347866119121e-5;
  case 3: return 2.400876357887274e-3;
  case 4: return 4.991228871115323e-2;
  case 5: return 2.142358068451711e-1;
  case 6: return 6.410835233041199e-1;
  case 7: return 1.260381060642639;
  case 8: return 2.21904886936509;
  default: return 3.310172839890271;
  }
}

// the largest power bank any ladder degree needs
__host__ __device__ constexpr int max_bank_blocks() {
  int mx = 0;
  for (int i = 0; i < kNumTaylor; ++i) {
    const int np = ps_plan(taylor_degree(i), 0).num_powers;
    if (np > mx) mx = np;
  }
  return mx;
}

// Squarings that rung i needs from the 1-norm alone: the smallest s >= 0 with
// norm1 / 2^s <= theta. A ratio that overflows (a rung whose theta is tiny next
// to a huge norm) marks the rung unusable instead of casting an infinite log2.
constexpr int kUnusableSquarings = 1 << 20;
__host__ __device__ inline int norm_squarings(double norm1, double theta) {
  if (!(norm1 > theta)) return 0;
  const double ratio = norm1 / theta;
  if (!(ratio <= 1.7976931348623157e308)) return kUnusableSquarings;
  const int s = static_cast<int>(ceil(log2(ratio)));
  return s < 0 ? 0 : s;
}

// ---- squarings from the norms of powers ----------------------------------------------------
// With pn[k-1] = ||A^k||_1 known for k = 1 .. q, the backward error of T_m(A / 2^s)
// is bounded through alpha_p = max(||A^p||^(1/p), ||A^(p+1)||^(1/(p+1))) for every p
// with p(p-1) <= m + 1 (Al-Mohy & Higham, SIAM J. Matrix Anal. Appl. 31 (2009)
// 970-989, Thm 4.2(a)): every k >= m + 1 is a sum of p's and (p+1)'s, so
//   || sum_{k>m} h_k (A/2^s)^k || <= sum_{k>m} |h_k| (alpha_p / 2^s)^k,
// and alpha_p / 2^s <= theta_m keeps the relative backward error <= u exactly as
// ||A||_1 / 2^s <= theta_m does (alpha_p <= ||A||_1). For a matrix far from normal
// alpha_p is far below ||A||_1 -- a dense N(0, 1) matrix of order 2048 has
// alpha_3 ~ ||A||_1 / 11 -- so s drops by several squarings. The comparison
// d_k <= x is made as ||A^k|| <= x^k (x^k by repeated multiplication), so the
// host and the device agree bit for bit without k-th roots.
constexpr int kMaxNormPower = 5;  // the largest power bank (degrees 24, 29) holds A^2 .. A^5
__host__ __device__ inline int power_squarings(const double *pn, int q, int m, double theta) {
  const int s_norm = norm_squarings(pn[0], theta);  // p = 1: alpha_1 = ||A||_1
  if (q < 3 || s_norm == 0 || s_norm >= kUnusableSquarings) return s_norm;
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

// The plan from the norms of A .. A^q: the rung minimising products + s, ties to
// the smaller s, where the q - 1 powers already formed count whether or not the
// rung's bank uses them. q = 1 is the plan from the 1-norm alone.
__host__ __device__ inline clm_expm_plan taylor_plan_powers(const double *pn, int q, bool single) {
  const int have = q > 1 ? q - 1 : 0;
  int best = -1, best_cost = 0, best_s = 0;
  for (int i = 0; i < kNumTaylor; ++i) {
    const clm_ps_plan pp = ps_plan(taylor_degree(i), 0);
    const int s = power_squarings(pn, q, taylor_degree(i), taylor_theta_at(i, single));
    const int cost = (pp.num_powers > have ? pp.num_powers : have) + pp.r + s;
    if (best < 0 || cost < best_cost || (cost == best_cost && s < best_s)) {
      best = i;
      best_cost = cost;
      best_s = s;
    }
  }
  clm_expm_plan p{};
  p.m = taylor_degree(best);
  p.s = best_s;
  p.num_gemms = best_cost;
  p.method = CLM_METHOD_TAYLOR_PS;
  p.ps_block = ps_plan(p.m, 0).s;
  p.dtype = 0;
  p.norm1 = pn[0];
  p.norm_powers = q < 1 ? 1 : q;
  for (int k = 2; k <= q && k <= kMaxNormPower; ++k) p.power_norm1[k - 2] = pn[k - 1];
  return p;
}

__host__ __device__ inline clm_expm_plan taylor_plan(double norm1, bool single) {
  return taylor_plan_powers(&norm1, 1, single);
}

// Whether the norms of powers are taken: powers of the unscaled matrix must stay
// finite (||A^5|| <= ||A||^5; FP32 also bounds the 16-bit split planes), and at
// least A^3 must be in the norm-only plan's bank for an alpha_p to exist. Returns
// the powers formed first: A^2 and A^3 (3), or none (1).
__host__ __device__ inline int norm_power_count(double norm1, int degree, bool single_range) {
  if (!(norm1 <= (single_range ? 0x1p24 : 0x1p200))) return 1;
  return ps_plan(degree, 0).num_powers + 1 >= 3 ? 3 : 1;
}

// Whether forming A^(q+1) can lower the cost of the plan from the norms of
// A .. A^q: the best plan with ||A^(q+1)|| taken as 0 (the most that power's norm
// can help) against the best plan without it, the new power counted as a product.
// A power the final plan's bank would need anyway costs nothing extra, so a bank
// that keeps growing keeps being formed; one that would be wasted is not.
__host__ __device__ inline bool worth_next_power(const double *pn, int q, bool single) {
  if (q < 3 || q >= kMaxNormPower) return false;
  double trial[kMaxNormPower] = {};
  for (int k = 0; k < q; ++k) trial[k] = pn[k];
  return taylor_plan_powers(trial, q + 1, single).num_gemms <
         taylor_plan_powers(pn, q, single).num_gemms;
}

}  // namespace clm
