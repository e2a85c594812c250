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

__host__ __device__ inline clm_expm_plan taylor_plan(double norm1, bool single) {
  int best = -1, best_cost = 0, best_s = 0;
  for (int i = 0; i < kNumTaylor; ++i) {
    const double theta = taylor_theta_at(i, single);
    int s = 0;
    if (norm1 > theta) {
      s = static_cast<int>(ceil(log2(norm1 / theta)));
      if (s < 0) s = 0;
    }
    const int cost = ps_plan(taylor_degree(i), 0).num_gemms + s;
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
  p.norm1 = norm1;
  return p;
}

}  // namespace clm
