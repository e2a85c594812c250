// batched.hip -- many matrix exponentials of uneven orders in ONE persistent
// launch (no reference counterpart: the reference runs one expm per call).
//
// Design:
//   * one workgroup of 256 lanes (4 wave64) owns one matrix at a time and runs
//     its whole chain -- 1-norm, plan, 2^-s scaling, power bank, the
//     Paterson-Stockmeyer Horner steps, the s squarings -- with workgroup
//     barriers between the steps; every step reads and writes the matrix's own
//     workspace, so workgroups never wait on one another;
//   * the grid is persistent (as many workgroups as fit on the device at once)
//     and takes matrices from a queue in decreasing cost order (n^3, x4 for
//     complex): the largest chains start first, the small ones fill the tail;
//   * the plan is chosen on the device from the device-computed 1-norm, with the
//     same rules (expm_plan.hpp) and the same fold order as the single-matrix
//     path, so nothing returns to the host and the launch can be captured in a
//     hipGraph;
//   * the products use the single-matrix kernels' tile helpers (gemm_tile.hpp)
//     over a per-matrix choice of workgroup tile (32x32 .. 128x128) picked from a
//     tuning table that clm_expm_batched_autotune measures;
//   * each matrix therefore goes through the same arithmetic, in the same
//     order, as clm_expm_{s,d,c,z} on its own: the results agree bit for bit.
//
// The kernel is instantiated per (element type, workgroup tile): one
// instantiation holding every tile would carry the register demand of the
// largest (measured: 256 VGPRs and spills, against 67 for the 64 x 64 FP32 tile
// alone), so a plan groups its matrices by (type, tile) and launches one
// persistent grid per group, back to back on the stream (one graph when
// captured).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <vector>

#include "expm_plan.hpp"
#include "gemm_tile.hpp"

namespace clm {

struct BatchDesc {
  const void *A;
  void *E;
  long long ws_off;  // byte offset of the matrix's workspace inside the caller's buffer
  int n, lda, lde, dtype;
  int index;  // position in the caller's entry array
  int cfg;    // workgroup tile configuration (CLM_BATCH_TILE_*)
  // adjoint plans only (tape_off < 0 otherwise)
  const void *G;  // cotangent of E
  void *dA;       // gradient with respect to A
  long long tape_off;
  int ldg, ldda;
  int ckpt;  // checkpoint slots for the squaring chain
  int pad_;
};

namespace {

using namespace tile;

constexpr int kBuckets = CLM_BATCH_BUCKETS;
constexpr int kMaxBatchN = 4096;
constexpr int kTiles = CLM_BATCH_TILES;
constexpr int kGroups = 4 * kTiles;  // (element type, tile) pairs
constexpr size_t kHeadBytes = 256;  // the groups' queue counters at the head of the workspace

template <typename T>
constexpr bool is_cplx() {
  return std::is_same_v<T, zd> || std::is_same_v<T, zc>;
}

template <typename T>
constexpr int dtype_of() {
  if constexpr (std::is_same_v<T, double>) return CLM_DTYPE_F64;
  else if constexpr (std::is_same_v<T, float>) return CLM_DTYPE_F32;
  else if constexpr (std::is_same_v<T, zd>) return CLM_DTYPE_C128;
  else return CLM_DTYPE_C64;
}

__device__ __forceinline__ double modulus(double x) { return fabs(x); }
__device__ __forceinline__ double modulus(zd x) { return hypot(x.re, x.im); }
__device__ __forceinline__ float modulus(float x) { return fabsf(x); }
__device__ __forceinline__ float modulus(zc x) { return hypotf(x.re, x.im); }

template <typename T>
__host__ __device__ inline long stride_of(int n) {
  return static_cast<long>(((static_cast<size_t>(n) * n * sizeof(T) + 255) & ~size_t(255)) /
                           sizeof(T));
}

// The adjoint's tape, per matrix, after a 256-byte head {m, s, c, status}:
//   P[1..r_max]      the Horner iterates P_r (= B_r) .. P_1
//   CK[0..K)         squaring checkpoints R_0, R_c, R_2c, ... (c = ceil(s / K))
//   U, V             cotangent ping-pong of the backward sweep
//   W0, W1           squarings recomputed from a checkpoint
//   Xb               cotangent of the scaled matrix X0 = A / 2^s
//   Yb[0..nbank)     cotangents of the power bank X0^2 ..
// (the backward keeps every cotangent conjugate-transposed)
__host__ __device__ constexpr int max_horner_steps() {
  int mx = 0;
  for (int i = 0; i < kNumTaylor; ++i) {
    const int r = ps_plan(taylor_degree(i), 0).r;
    if (r > mx) mx = r;
  }
  return mx;
}
// compile-time constants: the tape and workspace offsets below are computed per
// element in the adjoint's inner loops
constexpr int kMaxHornerSteps = max_horner_steps();
constexpr int kMaxBankBlocks = max_bank_blocks();
// per-matrix workspace: As | bank (kMaxBankBlocks) | S, each an aligned n x n block
template <typename T>
size_t matrix_ws_bytes(int n) {
  return static_cast<size_t>(2 + kMaxBankBlocks) * stride_of<T>(n) * sizeof(T);
}
__host__ __device__ inline int tape_blocks(int ckpt) {
  return kMaxHornerSteps + ckpt + 5 + kMaxBankBlocks;
}
template <typename T>
size_t tape_bytes(int n, int ckpt) {
  return 256 + static_cast<size_t>(tape_blocks(ckpt)) * stride_of<T>(n) * sizeof(T);
}

// ---- LDS budget per tile configuration -------------------------------------------
template <typename S, int BM, int BN, int BK, int NT = 256>
constexpr size_t real_lds() {
  using TL = RTile<S, BM, BN, NT, BK>;
  return 2 * (TL::A_STAGE + TL::B_STAGE) * sizeof(S);
}
template <typename C, int BM, int BN, int NT = 256>
constexpr size_t cplx_lds() {
  using TL = CTile<C, BM, BN, NT, kBK>;
  return 2 * TL::STAGE * sizeof(typename CxReal<C>::type);
}

// K slice of the 32 x 32 tile: 16 for FP64 (2 doubles per 16-byte vector),
// 32 for FP32 (4 floats), so that each of the 256 lanes stages one vector
template <typename S>
constexpr int small_bk() {
  return std::is_same_v<S, double> ? 16 : 32;
}

// the tile a configuration id stands for per element type: complex FP64 stops
// at 64 x 64, 128 x 128 exists for real FP32 only (others take their largest)
template <typename T>
constexpr int effective_cfg(int cfg) {
  if (cfg == CLM_BATCH_TILE_128W16) return cfg;
  if constexpr (std::is_same_v<T, zd>) return cfg > CLM_BATCH_TILE_64 ? CLM_BATCH_TILE_64 : cfg;
  else if constexpr (std::is_same_v<T, float>) return cfg;
  else return cfg > CLM_BATCH_TILE_128x64 ? CLM_BATCH_TILE_128x64 : cfg;
}

// lanes per workgroup: 16 waves for the 128 x 128 tile of 32 x 32 wave tiles
// (a CU's worth of waves on one matrix), 4 for the others
constexpr int threads_of(int cfg) { return cfg == CLM_BATCH_TILE_128W16 ? 1024 : 256; }

template <typename T, int CFG>
constexpr size_t lds_bytes() {
  constexpr int c = effective_cfg<T>(CFG);
  if constexpr (c == CLM_BATCH_TILE_128W16) {
    if constexpr (is_cplx<T>()) return cplx_lds<T, 128, 128, 1024>();
    else return real_lds<T, 128, 128, 32, 1024>();
  } else if constexpr (is_cplx<T>()) {
    if constexpr (c == CLM_BATCH_TILE_32) return cplx_lds<T, 32, 32>();
    else if constexpr (c == CLM_BATCH_TILE_64) return cplx_lds<T, 64, 64>();
    else return cplx_lds<T, 128, 64>();
  } else {
    if constexpr (c == CLM_BATCH_TILE_32) return real_lds<T, 32, 32, small_bk<T>()>();
    else if constexpr (c == CLM_BATCH_TILE_64) return real_lds<T, 64, 64, kBK>();
    else if constexpr (c == CLM_BATCH_TILE_128x64) return real_lds<T, 128, 64, kBK>();
    else return real_lds<T, 128, 128, kBK>();
  }
}

// A value every lane of the wave holds, re-declared uniform: the compiler then
// keeps it (and what is computed from it) in scalar registers.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// ---- workgroup-wide building blocks ---------------------------------------------------
template <typename S, int BM, int BN, int WM, int WN, int BK, int NT = 256>
__device__ void wg_gemm_real(const GemmArgs<S> &g, char *lds) {
  using TL = RTile<S, BM, BN, NT, BK>;
  using V = typename Mfma<S>::vec;
  constexpr int WAVES_M = BM / WM;
  static_assert((BM / WM) * (BN / WN) == NT / 64, "one wave per WM x WN sub-tile");
  constexpr int TM = WM / 16, TN = WN / 16;
  S *As = reinterpret_cast<S *>(lds);
  S *Bs = As + 2 * TL::A_STAGE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int tiles_m = (g.M + BM - 1) / BM, tiles_n = (g.N + BN - 1) / BN;
  const int KT = (g.K + BK - 1) / BK;
  for (int t = 0; t < tiles_m * tiles_n; ++t) {
    const int m0 = (t % tiles_m) * BM, n0 = (t / tiles_m) * BN;
    typename Mfma<S>::acc acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<S>::zero();
    V ra[TL::A_V], rb[TL::B_V];
    __syncthreads();  // the previous tile's last slice is consumed before restaging
    if (KT > 0) {
      rgemm_gload<S, BM, BN, NT, BK, false>(g, m0, n0, 0, tid, ra, rb);
      rgemm_sstore<S, BM, BN, NT, BK>(As, Bs, tid, ra, rb);
      __syncthreads();
      for (int kt = 0; kt < KT - 1; ++kt) {
        const int cur = kt & 1;
        rgemm_gload<S, BM, BN, NT, BK, false>(g, m0, n0, (kt + 1) * BK, tid, ra, rb);
        rgemm_compute<S, BM, BN, WM, WN, NT, BK>(As + cur * TL::A_STAGE, Bs + cur * TL::B_STAGE,
                                                  lane, wm0, wn0, acc);
        rgemm_sstore<S, BM, BN, NT, BK>(As + (cur ^ 1) * TL::A_STAGE,
                                         Bs + (cur ^ 1) * TL::B_STAGE, tid, ra, rb);
        __syncthreads();
      }
      const int last = (KT - 1) & 1;
      rgemm_compute<S, BM, BN, WM, WN, NT, BK>(As + last * TL::A_STAGE, Bs + last * TL::B_STAGE,
                                                lane, wm0, wn0, acc);
    }
    constexpr int PW = epilogue_pass_cols<S, BM, BN, WN>(2 * (TL::A_STAGE + TL::B_STAGE) * sizeof(S));
    static_assert(BN % PW == 0 && PW % WN == 0 &&
                      static_cast<size_t>(BM + 1) * PW * sizeof(S) <=
                          2 * (TL::A_STAGE + TL::B_STAGE) * sizeof(S),
                  "epilogue passes fit the staging LDS");
    S *sumt = reinterpret_cast<S *>(lds);
#pragma unroll
    for (int p0 = 0; p0 < BN; p0 += PW) {
      __syncthreads();  // the stages (or the previous pass) are no longer read
      if (wn0 >= p0 && wn0 < p0 + PW)
        epilogue_put<S, S, BM, TM, TN>(sumt, p0, lane, wm0, wn0,
                                       [&](int i, int j, int r) { return S(acc[i][j][r]); });
      __syncthreads();
      epilogue_drain<S, BM, PW, NT>(g, sumt, m0, n0 + p0);
    }
  }
  __syncthreads();
}

template <typename C, int BM, int BN, int WM, int WN, int NT = 256>
__device__ void wg_gemm_cplx(const GemmArgs<C> &g, char *lds) {
  constexpr int BK = kBK;
  using TL = CTile<C, BM, BN, NT, BK>;
  using S = typename CxReal<C>::type;
  constexpr int WAVES_M = BM / WM;
  static_assert((BM / WM) * (BN / WN) == NT / 64, "one wave per WM x WN sub-tile");
  constexpr int TM = WM / 16, TN = WN / 16;
  S *L = reinterpret_cast<S *>(lds);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int tiles_m = (g.M + BM - 1) / BM, tiles_n = (g.N + BN - 1) / BN;
  const int KT = (g.K + BK - 1) / BK;
  for (int t = 0; t < tiles_m * tiles_n; ++t) {
    const int m0 = (t % tiles_m) * BM, n0 = (t / tiles_m) * BN;
    typename Mfma<S>::acc cr[TM][TN], ci[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        cr[i][j] = Mfma<S>::zero();
        ci[i][j] = Mfma<S>::zero();
      }
    C ra[TL::A_E], rb[TL::B_E];
    __syncthreads();
    if (KT > 0) {
      cgemm_gload<C, BM, BN, NT, BK, false>(g, m0, n0, 0, tid, ra, rb);
      cgemm_sstore<C, BM, BN, NT, BK>(L, tid, ra, rb);
      __syncthreads();
      for (int kt = 0; kt < KT - 1; ++kt) {
        const int cur = kt & 1;
        cgemm_gload<C, BM, BN, NT, BK, false>(g, m0, n0, (kt + 1) * BK, tid, ra, rb);
        cgemm_compute<C, BM, BN, WM, WN, NT, BK>(L + cur * TL::STAGE, lane, wm0, wn0, cr, ci);
        cgemm_sstore<C, BM, BN, NT, BK>(L + (cur ^ 1) * TL::STAGE, tid, ra, rb);
        __syncthreads();
      }
      cgemm_compute<C, BM, BN, WM, WN, NT, BK>(L + ((KT - 1) & 1) * TL::STAGE, lane, wm0, wn0,
                                                cr, ci);
    }
    constexpr int PW = epilogue_pass_cols<C, BM, BN, WN>(2 * TL::STAGE * sizeof(S));
    static_assert(BN % PW == 0 && PW % WN == 0 &&
                      static_cast<size_t>(BM + 1) * PW * sizeof(C) <= 2 * TL::STAGE * sizeof(S),
                  "epilogue passes fit the staging LDS");
    C *sumt = reinterpret_cast<C *>(lds);
#pragma unroll
    for (int p0 = 0; p0 < BN; p0 += PW) {
      __syncthreads();  // the stages (or the previous pass) are no longer read
      if (wn0 >= p0 && wn0 < p0 + PW)
        epilogue_put<C, S, BM, TM, TN>(sumt, p0, lane, wm0, wn0, [&](int i, int j, int r) {
          return C{S(cr[i][j][r]), S(ci[i][j][r])};
        });
      __syncthreads();
      epilogue_drain<C, BM, PW, NT>(g, sumt, m0, n0 + p0);
    }
  }
  __syncthreads();
}

template <typename T, int CFG>
__device__ void wg_gemm(const GemmArgs<T> &g, char *lds) {
  constexpr int c = effective_cfg<T>(CFG);
  if constexpr (c == CLM_BATCH_TILE_128W16) {
    if constexpr (is_cplx<T>()) wg_gemm_cplx<T, 128, 128, 32, 32, 1024>(g, lds);
    else wg_gemm_real<T, 128, 128, 32, 32, 32, 1024>(g, lds);
  } else if constexpr (is_cplx<T>()) {
    if constexpr (c == CLM_BATCH_TILE_32) wg_gemm_cplx<T, 32, 32, 16, 16>(g, lds);
    else if constexpr (c == CLM_BATCH_TILE_64) wg_gemm_cplx<T, 64, 64, 32, 32>(g, lds);
    else wg_gemm_cplx<T, 128, 64, 64, 32>(g, lds);
  } else {
    if constexpr (c == CLM_BATCH_TILE_32) wg_gemm_real<T, 32, 32, 16, 16, small_bk<T>()>(g, lds);
    else if constexpr (c == CLM_BATCH_TILE_64) wg_gemm_real<T, 64, 64, 32, 32, kBK>(g, lds);
    else if constexpr (c == CLM_BATCH_TILE_128x64) wg_gemm_real<T, 128, 64, 64, 32, kBK>(g, lds);
    else wg_gemm_real<T, 128, 128, 64, 64, kBK>(g, lds);
  }
}

// complex conjugate for the adjoint (identity for real types)
__device__ __forceinline__ float conj_if(float x, bool) { return x; }
__device__ __forceinline__ double conj_if(double x, bool) { return x; }
__device__ __forceinline__ zc conj_if(zc x, bool c) { return c ? zc{x.re, -x.im} : x; }
__device__ __forceinline__ zd conj_if(zd x, bool c) { return c ? zd{x.re, -x.im} : x; }

// The 1-norm in the fold order of the single-matrix reduction (256 virtual
// lanes, each folding rows t, t+256, ... seeded by its first, then a halving
// tree): virtual lane t + 64q sits in register q of lane t of a wave; the
// tree's first two levels add registers, the remaining six are shuffles. A
// virtual lane past n holds +0, which leaves every partial sum unchanged, so
// the result is bit for bit the single-matrix kernel's. Each wave folds every
// fourth column; the NaN-propagating max runs over the columns.
template <typename T>
__device__ real_t<T> wg_norm1(int n, const T *A, long lda, char *lds) {
  using R = real_t<T>;
  R *red = reinterpret_cast<R *>(lds);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  R best = 0;
  for (int j = wave; j < n; j += nwv) {
    const T *col = A + static_cast<long>(j) * lda;
    R v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = lane + 64 * q;
      R acc = 0;
      if (t < n) {
        acc = modulus(col[t]);
        for (int i = t + 256; i < n; i += 256) acc = acc + modulus(col[i]);
      }
      v[q] = acc;
    }
    v[0] = v[0] + v[2];
    v[1] = v[1] + v[3];
    v[0] = v[0] + v[1];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
      const R o = __shfl_down(v[0], s, 64);
      if (lane < s) v[0] = v[0] + o;
    }
    const R cs = __shfl(v[0], 0, 64);
    if (cs > best || isnan(cs)) best = cs;
  }
  __syncthreads();  // the LDS arena may still be read by a previous step
  if (lane == 0) red[wave] = best;
  __syncthreads();
  R r = 0;
  for (int w = 0; w < nwv; ++w) {
    const R v = red[w];
    if (v > r || isnan(v)) r = v;
  }
  __syncthreads();
  return r;
}

// Element-wise passes over an n x n matrix: the workgroup walks the elements in
// column-major order, each lane taking EW of them per round (consecutive lanes
// on consecutive rows), and every round issues all of its loads before any of
// its stores -- a lane has EW loads in flight instead of one, which is what a
// single workgroup's pass over a matrix is bound by (load latency, not
// bandwidth). Each element's arithmetic is unchanged.
constexpr int kEw = 8;
// the round for a configuration: kEw for 4-wave workgroups; 2 for the 16-wave tile,
// which has four times the lanes (the same loads in flight per CU) and a quarter
// of the registers per lane (8 would spill)
constexpr int ew_of(int cfg) { return threads_of(cfg) >= 1024 ? 2 : kEw; }

// (row, column) of linear element e of an n x n matrix
__device__ __forceinline__ void ew_rc(int e, int n, int &i, int &j) {
  j = e / n;
  i = e - j * n;
}

// out = f * in (f a power of two: exact)
template <int EW = kEw, typename T>
__device__ void wg_scale_copy(int n, double f, const T *in, long ldi, T *out, long ldo) {
  const int total = n * n, nt = blockDim.x;
  for (int base = threadIdx.x; base < total; base += nt * EW) {
    T v[EW];
#pragma unroll
    for (int u = 0; u < EW; ++u) {
      const int e = base + u * nt;
      int i, j;
      ew_rc(e, n, i, j);
      if (e < total) v[u] = in[i + static_cast<long>(j) * ldi];
    }
#pragma unroll
    for (int u = 0; u < EW; ++u) {
      const int e = base + u * nt;
      int i, j;
      ew_rc(e, n, i, j);
      if (e < total) out[i + static_cast<long>(j) * ldo] = ops<T>::scale(v[u], f);
    }
  }
  __syncthreads();
}

// bank[k-2] = 2^(-k s) bank[k-2] in place for k = 2 .. count + 1 (ldexp: the
// single-matrix scale_powers_kernel's arithmetic)
template <int EW = kEw, typename T>
__device__ void wg_scale_powers(int n, T *bank, long stride, int count, int s) {
  const int total = n * n, nt = blockDim.x;
  for (int c = 0; c < count; ++c) {
    T *x = bank + static_cast<long>(c) * stride;
    const int e2 = -(c + 2) * s;
    for (int base = threadIdx.x; base < total; base += nt * EW) {
      T v[EW];
#pragma unroll
      for (int u = 0; u < EW; ++u) {
        const int e = base + u * nt;
        if (e < total) v[u] = x[e];
      }
#pragma unroll
      for (int u = 0; u < EW; ++u) {
        const int e = base + u * nt;
        if (e < total) x[e] = ops<T>::ldexp(v[u], e2);
      }
    }
  }
  __syncthreads();
}

// B = c0 I + c1 X1 + sum_t c_t bank[t-2], the per-element order of the
// single-matrix block build
template <int EW = kEw, typename T>
__device__ void wg_build_block(int n, T *B, long ldb, const T *coef, int nterms, const T *X1,
                               long ldx1, const T *bank, long stride, long ldbank) {
  const int total = n * n, nt = blockDim.x;
  for (int base = threadIdx.x; base < total; base += nt * EW) {
    T acc[EW];
#pragma unroll
    for (int u = 0; u < EW; ++u) {
      const int e = base + u * nt;
      int i, j;
      ew_rc(e, n, i, j);
      acc[u] = (i == j) ? coef[0] : ops<T>::zero();
      if (e < total && nterms > 1)
        acc[u] = ops<T>::mad(coef[1], X1[i + static_cast<long>(j) * ldx1], acc[u]);
    }
    for (int t = 2; t < nterms; ++t) {
      const T *bt = bank + static_cast<long>(t - 2) * stride;
#pragma unroll
      for (int u = 0; u < EW; ++u) {
        const int e = base + u * nt;
        int i, j;
        ew_rc(e, n, i, j);
        if (e < total) acc[u] = ops<T>::mad(coef[t], bt[i + j * ldbank], acc[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < EW; ++u) {
      const int e = base + u * nt;
      int i, j;
      ew_rc(e, n, i, j);
      if (e < total) B[i + static_cast<long>(j) * ldb] = acc[u];
    }
  }
  __syncthreads();
}

// dst = src
template <int EW = kEw, typename T>
__device__ void wg_copy(int n, const T *src, long lds_, T *dst, long ldd) {
  wg_scale_copy<EW>(n, 1.0, src, lds_, dst, ldd);
}

// dst = f src^H (n x n, f a power of two) through 32 x 32 LDS tiles (row 33), so
// the reads and the writes both run down columns; NT tiles per pair of barriers
// (as many as the kernel's LDS holds), every lane's loads of a batch in flight
// together; ends with a barrier
template <typename T>
constexpr size_t transpose_lds() {
  return 32 * 33 * sizeof(T);
}
template <typename T, int NTILE>
__device__ void wg_conj_transpose(int n, const T *src, long lds_, T *dst, long ldd, char *lds,
                                  double f = 1.0) {
  T *tile = reinterpret_cast<T *>(lds);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tiles = (n + 31) / 32, count = tiles * tiles;
  for (int t0 = 0; t0 < count; t0 += NTILE) {
    __syncthreads();
    for (int e = tid; e < 1024 * NTILE; e += nt) {
      const int t = t0 + (e >> 10);
      const int i = e & 31, j = (e >> 5) & 31;
      const int ti = t % tiles, tj = t / tiles;  // source tile: rows 32 ti.., columns 32 tj..
      const int gi = ti * 32 + i, gj = tj * 32 + j;
      if (t < count && gi < n && gj < n)
        tile[(e >> 10) * 1056 + j * 33 + i] = src[gi + static_cast<long>(gj) * lds_];
    }
    __syncthreads();
    for (int e = tid; e < 1024 * NTILE; e += nt) {
      const int t = t0 + (e >> 10);
      const int i = e & 31, j = (e >> 5) & 31;  // dst(32 tj + i, 32 ti + j) = conj(src(32 ti + j, 32 tj + i))
      const int ti = t % tiles, tj = t / tiles;
      const int gi = tj * 32 + i, gj = ti * 32 + j;
      if (t < count && gi < n && gj < n)
        dst[gi + static_cast<long>(gj) * ldd] =
            ops<T>::scale(conj_if(tile[(e >> 10) * 1056 + i * 33 + j], true), f);
    }
  }
  __syncthreads();
}

// dst = v on the n x n block
template <typename T>
__device__ void wg_fill(int n, T *dst, long ld, T v) {
  const int total = n * n, nt = blockDim.x;
  for (int e = threadIdx.x; e < total; e += nt) {
    int i, j;
    ew_rc(e, n, i, j);
    dst[i + static_cast<long>(j) * ld] = v;
  }
  __syncthreads();
}

// the adjoint kernel's LDS: the product's stages, at least one transpose tile
template <typename T, int CFG>
constexpr size_t kAdjointLds() {
  return lds_bytes<T, CFG>() > transpose_lds<T>() ? lds_bytes<T, CFG>() : transpose_lds<T>();
}

// the adjoint tape of one matrix (layout above tape_blocks)
template <typename T>
struct Tape {
  int *head;  // {m, s, c, status}
  T *base;    // first block
  long stride;
  int ckpt;
  __device__ T *horner(int j) const { return base + static_cast<long>(j - 1) * stride; }  // P_j, j >= 1
  __device__ T *checkpoint(int i) const {
    return base + static_cast<long>(kMaxHornerSteps + i) * stride;
  }
  __device__ T *scratch(int i) const {  // U, V, W0, W1, Xb
    return base + static_cast<long>(kMaxHornerSteps + ckpt + i) * stride;
  }
  __device__ T *bank_bar(int i) const {  // cotangent of X0^(i+2)
    return base + static_cast<long>(kMaxHornerSteps + ckpt + 5 + i) * stride;
  }
};

template <typename T>
__device__ Tape<T> tape_of(const BatchDesc &d, char *work) {
  Tape<T> t;
  t.head = reinterpret_cast<int *>(work + d.tape_off);
  t.base = reinterpret_cast<T *>(work + d.tape_off + 256);
  t.stride = stride_of<T>(d.n);
  t.ckpt = d.ckpt;
  return t;
}

template <typename T>
__device__ GemmArgs<T> square_args(int n, const T *A, long lda, const T *B, long ldb, T *C,
                                   long ldc) {
  GemmArgs<T> g;
  g.M = g.N = g.K = n;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.alpha = ops<T>::from_real(1.0);
  return g;
}

// One matrix's chain. First A^2, A^3 of A itself (none when the norm-only plan's
// degree is below 8), then A^4, A^5 while worth_next_power, whose 1-norms give
// the executed plan; the bank and A are then scaled by 2^-(k s) (exact) to the bank
// of A / 2^s. Then the chain as one sequence of products, so the product code
// of the long loop is instantiated once per tile configuration:
//   steps [0, nbank)             power bank A^2 .. A^(nbank+1) of the scaled A
//                                (from the first power not yet formed),
//   (block B_r built here)
//   steps [nbank, nbank + r)     Paterson-Stockmeyer Horner steps P <- P X + B_j,
//                                B_j formed in the product's epilogue,
//   steps [nbank + r, + s)       squarings.
// This is the single-matrix evaluator's order of operations exactly; the
// polynomial lands where the squarings end in E (the parity of s picks it).
template <typename T, int CFG>
__device__ void expm_one(const BatchDesc &d, char *work, const T *coef, char *lds,
                         clm_batch_result *results) {
  using R = real_t<T>;
  constexpr bool kSingle = std::is_same_v<R, float>;
  const int n = d.n;
  const T *A = static_cast<const T *>(d.A);
  T *E = static_cast<T *>(d.E);
  const long stride = stride_of<T>(n);
  T *As = reinterpret_cast<T *>(work + d.ws_off);
  T *bank = As + stride;
  T *S = bank + static_cast<long>(kMaxBankBlocks) * stride;

  const bool ad = d.tape_off >= 0;
  const Tape<T> tape = tape_of<T>(d, work);

  const R norm = uniform(wg_norm1(n, A, d.lda, lds));
  if (!isfinite(norm)) {
    // no exponential: E is NaN (never stale memory) and the status says why
    if (threadIdx.x == 0 && results)
      results[d.index] = clm_batch_result{0, 0, 0, CLM_STATUS_INVALID_VALUE, double(norm)};
    if (threadIdx.x == 0 && ad) tape.head[3] = CLM_STATUS_INVALID_VALUE;
    wg_fill(n, E, long(d.lde), ops<T>::from_real(__builtin_nan("")));
    return;  // uniform: every lane holds the same norm
  }
  // A^2, A^3 of A itself and their 1-norms, then A^4, A^5 while worth_next_power
  int npow =
      uniform(norm_power_count(double(norm), taylor_plan(double(norm), kSingle).m, kSingle));
  double pn[kMaxNormPower] = {double(norm)};
  if (npow > 1) {
    int formed = 1;
    for (;;) {
      for (int k = formed + 1; k <= npow; ++k) {
        const T *left = (k == 2) ? A : bank + static_cast<long>(k - 3) * stride;
        wg_gemm<T, CFG>(square_args(n, left, (k == 2) ? long(d.lda) : long(n), A, long(d.lda),
                                    bank + static_cast<long>(k - 2) * stride, long(n)),
                        lds);
      }
      for (int k = formed + 1; k <= npow; ++k)
        pn[k - 1] =
            double(uniform(wg_norm1(n, bank + static_cast<long>(k - 2) * stride, long(n), lds)));
      formed = npow;
      if (!worth_next_power(pn, npow, kSingle)) break;
      npow = uniform(npow + 1);
    }
  }
  const clm_expm_plan p = taylor_plan_powers(pn, npow, kSingle);
  const int degree = uniform(p.m), nsq = uniform(p.s);
  const int npre = npow > 1 ? npow - 1 : 0;
  if (threadIdx.x == 0 && results)
    results[d.index] =
        clm_batch_result{degree, nsq, uniform(p.num_gemms), CLM_STATUS_SUCCESS, double(norm)};
  // checkpoint spacing: R_0, R_c, R_2c, ... fill at most ckpt slots
  const int cstep = ad ? max(1, (nsq + d.ckpt - 1) / d.ckpt) : 1;
  if (threadIdx.x == 0 && ad) {
    tape.head[0] = degree;
    tape.head[1] = nsq;
    tape.head[2] = cstep;
    tape.head[3] = CLM_STATUS_SUCCESS;
  }

  const T *src = A;
  long ld_src = d.lda;
  if (nsq > 0) {
    wg_scale_copy<ew_of(CFG)>(n, ldexp(1.0, -nsq), A, d.lda, As, n);
    if (npre > 0) wg_scale_powers<ew_of(CFG)>(n, bank, stride, npre, nsq);
    src = As;
    ld_src = n;
  }
  const clm_ps_plan pp = ps_plan(degree, 0);
  const int bs = uniform(pp.s), nhorner = uniform(pp.r), nbank = uniform(pp.num_powers);
  const T *X = (bs >= 2) ? bank + static_cast<long>(bs - 2) * stride : src;
  const long ldx = (bs >= 2) ? long(n) : ld_src;
  T *target = (nsq % 2 == 0) ? E : S;
  const long ld_target = (nsq % 2 == 0) ? long(d.lde) : long(n);
  T *other = (nsq % 2 == 0) ? S : E;
  const long ld_other = (nsq % 2 == 0) ? long(n) : long(d.lde);
  // Horner ping-pong: starts where r steps end in target
  int cur = (nhorner % 2 == 0) ? 0 : 1;
  const int total = nbank + nhorner + nsq;
  T *res = target, *spare = other;  // squaring operands once the polynomial is done
  long ld_res = ld_target, ld_spare = ld_other;
  // the bank powers formed before the plan are not formed again
  for (int q = min(npre, nbank); q <= total; ++q) {
    if (q == nbank) {
      const int terms_r = min(degree - nhorner * bs + 1, bs);
      // with a tape, the Horner iterates P_r .. P_1 stay in their own blocks
      T *blk = (ad && nhorner >= 1) ? tape.horner(nhorner) : (cur == 0 ? target : other);
      const long ldb = (ad && nhorner >= 1) ? long(n) : (cur == 0 ? ld_target : ld_other);
      wg_build_block<ew_of(CFG)>(n, blk, ldb, coef + static_cast<long>(nhorner) * bs, terms_r,
                                 src, ld_src, bank, stride, static_cast<long>(n));
    }
    if (ad && nsq > 0 && q == nbank + nhorner)
      wg_copy<ew_of(CFG)>(n, res, ld_res, tape.checkpoint(0), long(n));
    if (q == total) break;
    GemmArgs<T> g;
    if (q < nbank) {
      const int k = q + 2;
      const T *left = (k == 2) ? src : bank + static_cast<long>(k - 3) * stride;
      g = square_args(n, left, (k == 2) ? ld_src : long(n), src, ld_src,
                      bank + static_cast<long>(k - 2) * stride, long(n));
    } else if (q < nbank + nhorner) {
      const int j = nhorner - 1 - (q - nbank);
      T *from = cur == 0 ? target : other;
      T *to = cur == 0 ? other : target;
      long ld_from = cur == 0 ? ld_target : ld_other, ld_to = cur == 0 ? ld_other : ld_target;
      if (ad) {
        from = tape.horner(j + 1);
        ld_from = n;
        to = (j == 0) ? target : tape.horner(j);
        ld_to = (j == 0) ? ld_target : long(n);
      }
      g = square_args(n, from, ld_from, X, ldx, to, ld_to);
      g.coef = coef + static_cast<long>(j) * bs;
      g.nterms = min(degree - j * bs + 1, bs);
      g.X1 = src;
      g.ldx1 = ld_src;
      g.bank = bank;
      g.bank_stride = stride;
      g.ldbank = n;
      cur = 1 - cur;
    } else {
      g = square_args(n, res, ld_res, res, ld_res, spare, ld_spare);
      T *t = res;
      res = spare;
      spare = t;
      const long tl = ld_res;
      ld_res = ld_spare;
      ld_spare = tl;
    }
    wg_gemm<T, CFG>(g, lds);
    if (ad && q >= nbank + nhorner) {
      const int k1 = q - (nbank + nhorner) + 1;  // this product made R_k1
      if (k1 < nsq && k1 % cstep == 0)
        wg_copy<ew_of(CFG)>(n, res, ld_res, tape.checkpoint(k1 / cstep), long(n));
    }
  }
}

// ---- the adjoint -------------------------------------------------------------------
// Reverse mode through the executed chain (the gradient of the computed
// function, which is the Frechet adjoint L(A^H, G) to within the truncation
// and rounding of the forward): with G the cotangent of E = R_s,
//   squarings  R_{k+1} = R_k^2          R_k^- = R_{k+1}^- R_k^H + R_k^H R_{k+1}^-
//   Horner     P_j = P_{j+1} X + B_j    P_{j+1}^- = P_j^- X^H,  X^- += P_{j+1}^H P_j^-,
//                                       B_j^- = P_j^-
//   blocks     B_j = sum_t c_t X0^t     (X0^t)^- += conj(c_t) B_j^-
//   bank       Y_k = Y_{k-1} X0         Y_{k-1}^- += Y_k^- X0^H,  X0^- += Y_{k-1}^H Y_k^-
//   scaling    A^- = 2^-s X0^-
// The R_k come from the checkpoints (squared forward again between them), the
// Horner iterates and the power bank from the forward's tape and workspace.

// The cotangents are kept conjugate-transposed (see adjoint_one): for the block
// terms t = 1 .. terms-1, dst_t^H += c_t bar^H, i.e. dst_t += c_t * bt with bt
// the transposed cotangent held (t = 1 into Xb, t >= 2 into the bank cotangents)
template <int EW = kEw, typename T>
__device__ void wg_block_adjoint(int n, const T *bt, const T *c, int terms, T *Xb,
                                 const Tape<T> &tape) {
  const int total = n * n, nt = blockDim.x;
  for (int base = threadIdx.x; base < total; base += nt * EW) {
    T b[EW];
#pragma unroll
    for (int u = 0; u < EW; ++u) {
      const int e = base + u * nt;
      if (e < total) b[u] = bt[e];
    }
    for (int t = 1; t < terms; ++t) {
      T *y = (t == 1) ? Xb : tape.bank_bar(t - 2);
      T v[EW];
#pragma unroll
      for (int u = 0; u < EW; ++u) {
        const int e = base + u * nt;
        if (e < total) v[u] = y[e];
      }
#pragma unroll
      for (int u = 0; u < EW; ++u) {
        const int e = base + u * nt;
        if (e < total) y[e] = ops<T>::mad(c[t], b[u], v[u]);
      }
    }
  }
  __syncthreads();
}

template <typename T>
__device__ GemmArgs<T> accumulate(GemmArgs<T> g) {
  g.Cin = g.C;
  g.ldcin = g.ldc;
  g.beta = ops<T>::from_real(1.0);
  return g;
}

template <typename T, int CFG>
__device__ void adjoint_one(const BatchDesc &d, char *work, const T *coef, char *lds) {
  using R = real_t<T>;
  const int n = d.n;
  const Tape<T> tape = tape_of<T>(d, work);
  T *dA = static_cast<T *>(d.dA);
  const T *A = static_cast<const T *>(d.A);
  if (uniform(tape.head[3]) != CLM_STATUS_SUCCESS) {
    const R nan = __builtin_nan("");
    wg_fill(n, dA, d.ldda, ops<T>::from_real(nan));
    return;
  }
  const int degree = uniform(tape.head[0]), nsq = uniform(tape.head[1]);
  const int cstep = uniform(tape.head[2]);
  const long stride = stride_of<T>(n);
  T *As = reinterpret_cast<T *>(work + d.ws_off);
  const T *bank = As + stride;
  const T *X0 = nsq > 0 ? As : A;
  const long ldx0 = nsq > 0 ? long(n) : long(d.lda);
  const clm_ps_plan pp = ps_plan(degree, 0);
  const int bs = uniform(pp.s), nhorner = uniform(pp.r), nbank = uniform(pp.num_powers);
  T *U = tape.scratch(0), *V = tape.scratch(1), *W0 = tape.scratch(2), *W1 = tape.scratch(3);
  T *Xb = tape.scratch(4);
  // The sweep runs on the conjugate transposes of the cotangents, B = bar^H:
  // every rule above turns into products of stored operands, with no operand
  // transposed, e.g. bar R^H + R^H bar = (R B + B R)^H and X0bar += Y^H Ybar
  // becomes X0bar^H += Ybar^H Y. G is transposed once on the way in and X0bar
  // once on the way out (into dA, with the 2^-s). The three phases run as one
  // sequence of steps with a single product call site (as the forward does), so
  // the product code is instantiated once:
  //   squarings, last first: for k = s-1 .. 0, R_k squared forward again from
  //     its checkpoint, then B <- R_k B + B R_k;
  //   Horner steps j = 0 .. r-1, each after the block adjoint of B_j:
  //     Xbar^H += B P_(j+1), B <- X B; then the block adjoint of B_r;
  //   power bank, highest power first: Y_(k-1)bar^H += X0 Y_k bar^H,
  //     X0bar^H += Y_k bar^H Y_(k-1).
  constexpr int kTiles = kAdjointLds<T, CFG>() / transpose_lds<T>() < 8
                             ? int(kAdjointLds<T, CFG>() / transpose_lds<T>())
                             : 8;
  wg_conj_transpose<T, kTiles>(n, static_cast<const T *>(d.G), d.ldg, U, long(n), lds);
  const T *bar = U;
  const T *X = (bs >= 2) ? bank + static_cast<long>(bs - 2) * stride : X0;
  const long ldx = (bs >= 2) ? long(n) : ldx0;
  T *Xbar = (bs >= 2) ? tape.bank_bar(bs - 2) : Xb;
  int phase = nsq > 0 ? 0 : 1;
  int k = nsq - 1, have = -1, w = 0, sub = 0;  // squarings
  int j = 0;                                   // Horner steps
  int kb = nbank + 1;                          // power bank
  bool started = false;
  const T *Rk = nullptr;
  T *out = nullptr;
  for (;;) {
    GemmArgs<T> g;
    if (phase == 0) {
      if (have < 0) {
        const int base = k / cstep;
        Rk = tape.checkpoint(base);
        have = base * cstep;
        w = 0;
        sub = 0;
      }
      if (have < k) {
        T *to = w ? W1 : W0;
        g = square_args(n, Rk, long(n), Rk, long(n), to, long(n));
        Rk = to;
        ++have;
        w ^= 1;
      } else if (sub == 0) {
        out = (bar == U) ? V : U;
        g = square_args(n, Rk, long(n), bar, long(n), out, long(n));
        sub = 1;
      } else {
        g = accumulate(square_args(n, bar, long(n), Rk, long(n), out, long(n)));
        bar = out;
        --k;
        have = -1;
        if (k < 0) phase = 1;
      }
    } else if (phase == 1) {
      if (!started) {
        wg_fill(n, Xb, long(n), ops<T>::zero());
        for (int t = 0; t < nbank; ++t) wg_fill(n, tape.bank_bar(t), long(n), ops<T>::zero());
        started = true;
        sub = 0;
      }
      if (sub == 0) {
        wg_block_adjoint<ew_of(CFG)>(n, bar, coef + static_cast<long>(j) * bs,
                                     min(degree - j * bs + 1, bs),
                         Xb, tape);
        if (j == nhorner) {
          phase = 2;
          continue;
        }
        g = accumulate(square_args(n, bar, long(n), tape.horner(j + 1), long(n), Xbar, long(n)));
        sub = 1;
      } else {
        out = (bar == U) ? V : U;
        g = square_args(n, X, ldx, bar, long(n), out, long(n));
        bar = out;
        ++j;
        sub = 0;
      }
    } else {
      if (kb < 2) break;
      T *ykb = tape.bank_bar(kb - 2);
      if (sub == 0) {
        T *ykm1b = (kb - 1 >= 2) ? tape.bank_bar(kb - 3) : Xb;
        g = accumulate(square_args(n, X0, ldx0, ykb, long(n), ykm1b, long(n)));
        sub = 1;
      } else {
        const T *ykm1 = (kb - 1 >= 2) ? bank + static_cast<long>(kb - 3) * stride : X0;
        const long ldy = (kb - 1 >= 2) ? long(n) : ldx0;
        g = accumulate(square_args(n, ykb, long(n), ykm1, ldy, Xb, long(n)));
        sub = 0;
        --kb;
      }
    }
    wg_gemm<T, CFG>(g, lds);
  }
  wg_conj_transpose<T, kTiles>(n, Xb, long(n), dA, long(d.ldda), lds, ldexp(1.0, -nsq));
}

template <typename T, int CFG>
__global__ void __launch_bounds__(threads_of(CFG))
    expm_adjoint_kernel(const BatchDesc *__restrict__ desc, int count, int *queue, char *work,
                        const T *coef) {
  __shared__ __attribute__((aligned(16))) char lds[kAdjointLds<T, CFG>()];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= count) break;
    const BatchDesc d = desc[item];
    adjoint_one<T, CFG>(d, work, coef, lds);
  }
}

// The persistent kernel of one (element type, tile) group: each workgroup
// takes the next matrix from the group's queue (one atomic per matrix, by lane
// 0, broadcast through LDS) until the queue is drained; every workgroup reaches
// the exit once the counter passes `count`.
template <typename T, int CFG>
__global__ void __launch_bounds__(threads_of(CFG))
    expm_batched_kernel(const BatchDesc *__restrict__ desc, int count, int *queue, char *work,
                        clm_batch_result *results, const T *coef) {
  __shared__ __attribute__((aligned(16))) char lds[lds_bytes<T, CFG>()];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    // uniform across the workgroup: readfirstlane tells the compiler so, which
    // keeps the descriptor and every size and pointer derived from it in SGPRs
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= count) break;
    const BatchDesc d = desc[item];
    expm_one<T, CFG>(d, work, coef, lds, results);
  }
}

// zeroes the kGroups queue counters at q (a kernel, not a memset: it is captured
// into graphs)
__global__ void reset_queue_kernel(int *q) {
  if (threadIdx.x < kGroups) q[threadIdx.x] = 0;
}

// deterministic test matrices for the tuner: entries in [-0.5, 0.5) / n
template <typename T>
__global__ void fill_tuning_matrices(T *A, int n, int count, unsigned seed) {
  const long total = static_cast<long>(n) * n * count;
  for (long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    unsigned h = static_cast<unsigned>(i) * 2654435761u ^ seed;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    const double u = (h & 0xffffff) / double(1 << 24) - 0.5;
    if constexpr (is_cplx<T>()) {
      unsigned g = h * 3266489917u;
      g ^= g >> 16;
      const double w = (g & 0xffffff) / double(1 << 24) - 0.5;
      A[i] = T{static_cast<decltype(T::re)>(u / n), static_cast<decltype(T::re)>(w / n)};
    } else {
      A[i] = static_cast<T>(u / n);
    }
  }
}

// ---- tuning table ---------------------------------------------------------------------
// cfg[path: 0 forward plans, 1 adjoint plans][precision: 0 FP32, 1 FP64]
// [0 real, 1 complex][bucket], bucket b covers n in (32 b, 32 (b + 1)], the
// last one everything above. Adjoint plans have a table of their own: their
// launches are dominated by the backward sweep, whose best tile differs.
std::mutex g_tune_mutex;
int g_tune[2][2][2][kBuckets];
bool g_tune_init = false;

void tune_defaults() {
  for (int a = 0; a < 2; ++a)
    for (int p = 0; p < 2; ++p)
      for (int c = 0; c < 2; ++c)
        for (int b = 0; b < kBuckets; ++b) {
          const int nmax = 32 * (b + 1);
          int cfg = nmax <= 32 ? CLM_BATCH_TILE_32 : nmax <= 96 ? CLM_BATCH_TILE_64
                                                                 : CLM_BATCH_TILE_128x64;
          g_tune[a][p][c][b] = cfg;
        }
  g_tune_init = true;
}

int bucket_of(int n) {
  const int b = (n - 1) / 32;
  return b < kBuckets ? b : kBuckets - 1;
}

bool is_complex_code(int dtype) { return dtype == CLM_DTYPE_C64 || dtype == CLM_DTYPE_C128; }
int precision_of(int dtype) { return (dtype == CLM_DTYPE_F64 || dtype == CLM_DTYPE_C128) ? 1 : 0; }
bool valid_dtype(int dtype) {
  return dtype == CLM_DTYPE_F32 || dtype == CLM_DTYPE_F64 || dtype == CLM_DTYPE_C64 ||
         dtype == CLM_DTYPE_C128;
}

size_t tape_bytes_of(int dtype, int n, int ckpt) {
  switch (dtype) {
  case CLM_DTYPE_F32: return tape_bytes<float>(n, ckpt);
  case CLM_DTYPE_F64: return tape_bytes<double>(n, ckpt);
  case CLM_DTYPE_C64: return tape_bytes<zc>(n, ckpt);
  default: return tape_bytes<zd>(n, ckpt);
  }
}

size_t ws_bytes_of(int dtype, int n) {
  switch (dtype) {
  case CLM_DTYPE_F32: return matrix_ws_bytes<float>(n);
  case CLM_DTYPE_F64: return matrix_ws_bytes<double>(n);
  case CLM_DTYPE_C64: return matrix_ws_bytes<zc>(n);
  default: return matrix_ws_bytes<zd>(n);
  }
}

// group id of an (element type, tile) pair; dtype index F32 0, F64 1, C64 2, C128 3
int group_of(int dtype, int cfg) {
  const int t = dtype == CLM_DTYPE_F32 ? 0 : dtype == CLM_DTYPE_F64 ? 1 : dtype == CLM_DTYPE_C64 ? 2 : 3;
  return t * kTiles + cfg;
}

// workgroups resident at once on the device for one kernel instantiation
template <typename T, int CFG>
int resident_workgroups(int device) {
  static int cached[64] = {};
  int &r = cached[device & 63];
  if (r == 0) {
    int blocks = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, expm_batched_kernel<T, CFG>, threads_of(CFG), 0) !=
            hipSuccess ||
        blocks < 1)
      blocks = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus < 1)
      cus = 256;
    r = blocks * cus;
  }
  return r;
}

template <typename T, int CFG>
hipError_t launch_group(hipStream_t st, int device, const BatchDesc *desc, int count, int *queue,
                        char *work, clm_batch_result *results) {
  const T *coef = inv_factorials<T>(device);
  if (!coef) return hipErrorInvalidValue;
  const int grid = std::min(count, resident_workgroups<T, CFG>(device));
  hipLaunchKernelGGL((expm_batched_kernel<T, CFG>), dim3(grid), dim3(threads_of(CFG)), 0, st, desc,
                     count,
                     queue, work, results, coef);
  return hipGetLastError();
}

// the occupancy and coefficient-table lookups a launch needs, done once ahead
// of it (so none of them runs inside a stream capture)
template <typename T>
bool prepare_type(int cfg, int device) {
  if (!inv_factorials<T>(device)) return false;
  switch (cfg) {
  case CLM_BATCH_TILE_32: return resident_workgroups<T, CLM_BATCH_TILE_32>(device) > 0;
  case CLM_BATCH_TILE_64: return resident_workgroups<T, CLM_BATCH_TILE_64>(device) > 0;
  case CLM_BATCH_TILE_128x64: return resident_workgroups<T, CLM_BATCH_TILE_128x64>(device) > 0;
  case CLM_BATCH_TILE_128: return resident_workgroups<T, CLM_BATCH_TILE_128>(device) > 0;
  default: return resident_workgroups<T, CLM_BATCH_TILE_128W16>(device) > 0;
  }
}

bool prepare_group(int g, int device) {
  switch (g / kTiles) {
  case 0: return prepare_type<float>(g % kTiles, device);
  case 1: return prepare_type<double>(g % kTiles, device);
  case 2: return prepare_type<zc>(g % kTiles, device);
  default: return prepare_type<zd>(g % kTiles, device);
  }
}

template <typename T, int CFG>
hipError_t launch_adjoint_group(hipStream_t st, int device, const BatchDesc *desc, int count,
                                int *queue, char *work) {
  const T *coef = inv_factorials<T>(device);
  if (!coef) return hipErrorInvalidValue;
  static int cached[64] = {};
  int &resident = cached[device & 63];
  if (resident == 0) {
    int blocks = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, expm_adjoint_kernel<T, CFG>, threads_of(CFG), 0) !=
            hipSuccess ||
        blocks < 1)
      blocks = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus < 1)
      cus = 256;
    resident = blocks * cus;
  }
  hipLaunchKernelGGL((expm_adjoint_kernel<T, CFG>), dim3(std::min(count, resident)),
                     dim3(threads_of(CFG)), 0,
                     st, desc, count, queue, work, coef);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_adjoint_type(int cfg, hipStream_t st, int device, const BatchDesc *desc,
                               int count, int *queue, char *work) {
  switch (cfg) {
  case CLM_BATCH_TILE_32:
    return launch_adjoint_group<T, CLM_BATCH_TILE_32>(st, device, desc, count, queue, work);
  case CLM_BATCH_TILE_64:
    return launch_adjoint_group<T, CLM_BATCH_TILE_64>(st, device, desc, count, queue, work);
  case CLM_BATCH_TILE_128x64:
    return launch_adjoint_group<T, CLM_BATCH_TILE_128x64>(st, device, desc, count, queue, work);
  case CLM_BATCH_TILE_128:
    return launch_adjoint_group<T, CLM_BATCH_TILE_128>(st, device, desc, count, queue, work);
  default:
    return launch_adjoint_group<T, CLM_BATCH_TILE_128W16>(st, device, desc, count, queue, work);
  }
}

template <typename T>
hipError_t launch_type(int cfg, hipStream_t st, int device, const BatchDesc *desc, int count,
                       int *queue, char *work, clm_batch_result *results) {
  switch (cfg) {
  case CLM_BATCH_TILE_32:
    return launch_group<T, CLM_BATCH_TILE_32>(st, device, desc, count, queue, work, results);
  case CLM_BATCH_TILE_64:
    return launch_group<T, CLM_BATCH_TILE_64>(st, device, desc, count, queue, work, results);
  case CLM_BATCH_TILE_128x64:
    return launch_group<T, CLM_BATCH_TILE_128x64>(st, device, desc, count, queue, work, results);
  case CLM_BATCH_TILE_128:
    return launch_group<T, CLM_BATCH_TILE_128>(st, device, desc, count, queue, work, results);
  default:
    return launch_group<T, CLM_BATCH_TILE_128W16>(st, device, desc, count, queue, work, results);
  }
}

}  // namespace
}  // namespace clm

// ---- plan object --------------------------------------------------------------------------
struct clm_batch_plan {
  int device = 0;
  int count = 0;
  int group_count[clm::kGroups] = {};  // matrices per (element type, tile) group
  int group_start[clm::kGroups] = {};
  double group_cost[clm::kGroups] = {};  // sum of n^3 (x4 complex, x2 FP64) of the members
  int launch_order[clm::kGroups] = {};   // non-empty groups, costliest first
  int launch_lane[clm::kGroups] = {};    // 0: the caller's stream, i > 0: side[i - 1]
  int num_launch = 0;
  clm::BatchDesc *d_desc = nullptr;    // grouped, each group in decreasing cost
  bool ad = false;                     // records the adjoint tape; backward allowed
  int ckpt = 0;
  size_t ws_bytes = 0;
  std::vector<clm::BatchDesc> host_desc;
  // the (type, tile) groups run concurrently: group 0 on the caller's stream, the
  // others on these side streams (default priority), forked from and joined
  // back to the caller's stream with events -- capturable into a graph
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSide] = {};
  // the captured launch (clm_expm_batched_graph)
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t graph_stream = nullptr;
  void *graph_work = nullptr;
  void *graph_results = nullptr;
};

namespace clm {
namespace {

void destroy_plan(clm_batch_plan *plan) {
  if (!plan) return;
  if (plan->graph_exec) (void)hipGraphExecDestroy(plan->graph_exec);
  if (plan->d_desc) (void)hipFree(plan->d_desc);
  for (int i = 0; i < clm_batch_plan::kSide; ++i) {
    if (plan->side[i]) (void)hipStreamDestroy(plan->side[i]);
    if (plan->join[i]) (void)hipEventDestroy(plan->join[i]);
  }
  if (plan->fork) (void)hipEventDestroy(plan->fork);
  delete plan;
}

// Enqueues launch(g, stream) for every non-empty group: the first on st, the
// others spread over the plan's side streams, which fork from st before and
// join it after.
template <typename F>
clm_status_t for_each_group(const clm_batch_plan *plan, hipStream_t st, F launch) {
  const int ng = plan->num_launch;
  if (ng == 0) return CLM_STATUS_SUCCESS;
  const int nside = std::min(ng - 1, clm_batch_plan::kSide);
  if (nside > 0) {
    CLM_TRY_HIP(hipEventRecord(plan->fork, st));
    for (int i = 0; i < nside; ++i) CLM_TRY_HIP(hipStreamWaitEvent(plan->side[i], plan->fork, 0));
  }
  for (int k = 0; k < ng; ++k) {
    const int lane = plan->launch_lane[k];
    hipStream_t s = (lane == 0) ? st : plan->side[lane - 1];
    CLM_TRY_HIP(launch(plan->launch_order[k], s));
  }
  for (int i = 0; i < nside; ++i) {
    CLM_TRY_HIP(hipEventRecord(plan->join[i], plan->side[i]));
    CLM_TRY_HIP(hipStreamWaitEvent(st, plan->join[i], 0));
  }
  return CLM_STATUS_SUCCESS;
}

clm_status_t build_plan(int count, const clm_batch_entry *entries, const int *forced_cfg,
                        clm_batch_plan **out, const clm_batch_ad_entry *ad = nullptr,
                        int ckpt = 0) {
  if (out == nullptr || count < 1 || entries == nullptr) return CLM_STATUS_INVALID_VALUE;
  for (int i = 0; i < count; ++i) {
    const clm_batch_entry &e = entries[i];
    if (e.n < 1 || e.n > kMaxBatchN || e.lda < e.n || e.lde < e.n || e.A == nullptr ||
        e.E == nullptr || !valid_dtype(e.dtype))
      return CLM_STATUS_INVALID_VALUE;
    if (ad && (ad[i].G == nullptr || ad[i].dA == nullptr || ad[i].ldg < e.n || ad[i].ldda < e.n))
      return CLM_STATUS_INVALID_VALUE;
  }
  if (ad && (ckpt < 1 || ckpt > 64)) return CLM_STATUS_INVALID_VALUE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return CLM_STATUS_NOT_INITIALIZED;
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (!g_tune_init) tune_defaults();
  }
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int i) {
    const double n = entries[i].n;
    return n * n * n * (is_complex_code(entries[i].dtype) ? 4.0 : 1.0);
  };
  std::vector<int> cfg(count);
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    for (int i = 0; i < count; ++i) {
      const clm_batch_entry &e = entries[i];
      cfg[i] = forced_cfg ? *forced_cfg
                          : g_tune[ad ? 1 : 0][precision_of(e.dtype)][is_complex_code(e.dtype)]
                                  [bucket_of(e.n)];
    }
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int ga = group_of(entries[a].dtype, cfg[a]), gb = group_of(entries[b].dtype, cfg[b]);
    if (ga != gb) return ga < gb;
    return cost(a) > cost(b);
  });
  auto *plan = new clm_batch_plan;
  plan->device = dev;
  plan->count = count;
  plan->host_desc.resize(count);
  size_t off = kHeadBytes;
  for (int k = 0; k < count; ++k) {
    const int i = order[k];
    const clm_batch_entry &e = entries[i];
    BatchDesc &d = plan->host_desc[k];
    d.A = e.A;
    d.E = e.E;
    d.ws_off = static_cast<long long>(off);
    d.n = e.n;
    d.lda = e.lda;
    d.lde = e.lde;
    d.dtype = e.dtype;
    d.index = i;
    d.cfg = cfg[i];
    off += Layout::align(ws_bytes_of(e.dtype, e.n));
    d.G = nullptr;
    d.dA = nullptr;
    d.tape_off = -1;
    d.ldg = d.ldda = 0;
    d.ckpt = 0;
    d.pad_ = 0;
    if (ad) {
      d.G = ad[i].G;
      d.dA = ad[i].dA;
      d.ldg = ad[i].ldg;
      d.ldda = ad[i].ldda;
      d.ckpt = ckpt;
      d.tape_off = static_cast<long long>(off);
      off += Layout::align(tape_bytes_of(e.dtype, e.n, ckpt));
    }
    plan->group_count[group_of(e.dtype, cfg[i])] += 1;
    plan->group_cost[group_of(e.dtype, cfg[i])] +=
        cost(i) * ((e.dtype == CLM_DTYPE_F64 || e.dtype == CLM_DTYPE_C128) ? 2.0 : 1.0);
  }
  for (int g = 1; g < kGroups; ++g)
    plan->group_start[g] = plan->group_start[g - 1] + plan->group_count[g - 1];
  // Longest-processing-time lanes: the groups in decreasing cost, each on the
  // stream whose groups so far cost least, so the heaviest groups start together
  // at the launch instead of queueing behind each other on one stream.
  {
    int ng = 0;
    for (int g = 0; g < kGroups; ++g)
      if (plan->group_count[g] > 0) plan->launch_order[ng++] = g;
    std::stable_sort(plan->launch_order, plan->launch_order + ng, [&](int a, int b) {
      return plan->group_cost[a] > plan->group_cost[b];
    });
    double load[1 + clm_batch_plan::kSide] = {};
    for (int k = 0; k < ng; ++k) {
      int best = 0;
      for (int l = 1; l < 1 + clm_batch_plan::kSide && l < ng; ++l)
        if (load[l] < load[best]) best = l;
      plan->launch_lane[k] = best;
      load[best] += plan->group_cost[plan->launch_order[k]];
    }
    plan->num_launch = ng;
  }
  for (int g = 0; g < kGroups; ++g)
    if (plan->group_count[g] > 0 && !prepare_group(g, dev)) {
      delete plan;
      return CLM_STATUS_EXECUTION_FAILED;
    }
  bool streams_ok = hipEventCreateWithFlags(&plan->fork, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < clm_batch_plan::kSide && streams_ok; ++i)
    streams_ok = hipStreamCreateWithFlags(&plan->side[i], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&plan->join[i], hipEventDisableTiming) == hipSuccess;
  if (!streams_ok) {
    destroy_plan(plan);
    return CLM_STATUS_EXECUTION_FAILED;
  }
  plan->ws_bytes = off;
  plan->ad = ad != nullptr;
  plan->ckpt = ckpt;
  if (hipMalloc(&plan->d_desc, sizeof(BatchDesc) * count) != hipSuccess) {
    delete plan;
    return CLM_STATUS_ALLOC_FAILED;
  }
  if (hipMemcpy(plan->d_desc, plan->host_desc.data(), sizeof(BatchDesc) * count,
                hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(plan->d_desc);
    delete plan;
    return CLM_STATUS_EXECUTION_FAILED;
  }
  *out = plan;
  return CLM_STATUS_SUCCESS;
}

clm_status_t launch_plan(clm_ctx *ctx, hipStream_t st, const clm_batch_plan *plan, void *work,
                         size_t bytes, clm_batch_result *results) {
  if (plan == nullptr || work == nullptr) return CLM_STATUS_INVALID_VALUE;
  if (bytes < plan->ws_bytes) return CLM_STATUS_ALLOC_FAILED;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  int *queue = static_cast<int *>(work);
  char *base = static_cast<char *>(work);
  hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, st, queue);
  CLM_TRY_HIP(hipGetLastError());
  return for_each_group(plan, st, [&](int g, hipStream_t s) -> hipError_t {
    const int count = plan->group_count[g];
    const BatchDesc *desc = plan->d_desc + plan->group_start[g];
    const int cfg = g % kTiles;
    switch (g / kTiles) {
    case 0: return launch_type<float>(cfg, s, plan->device, desc, count, queue + g, base, results);
    case 1: return launch_type<double>(cfg, s, plan->device, desc, count, queue + g, base, results);
    case 2: return launch_type<zc>(cfg, s, plan->device, desc, count, queue + g, base, results);
    default: return launch_type<zd>(cfg, s, plan->device, desc, count, queue + g, base, results);
    }
  });
}

// the adjoint sweep of an adjoint plan; reads the tape the last forward launch of
// the same plan on the same workspace recorded
clm_status_t launch_backward(clm_ctx *ctx, hipStream_t st, const clm_batch_plan *plan, void *work,
                             size_t bytes) {
  if (plan == nullptr || work == nullptr || !plan->ad) return CLM_STATUS_INVALID_VALUE;
  if (bytes < plan->ws_bytes) return CLM_STATUS_ALLOC_FAILED;
  clm_ctx *c = nullptr;
  CLM_TRY(resolve_ctx(ctx, &c));
  int *queue = static_cast<int *>(work) + kGroups;  // the second set of counters
  char *base = static_cast<char *>(work);
  hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, st, queue);
  CLM_TRY_HIP(hipGetLastError());
  return for_each_group(plan, st, [&](int g, hipStream_t s) -> hipError_t {
    const int count = plan->group_count[g];
    const BatchDesc *desc = plan->d_desc + plan->group_start[g];
    const int cfg = g % kTiles;
    switch (g / kTiles) {
    case 0: return launch_adjoint_type<float>(cfg, s, plan->device, desc, count, queue + g, base);
    case 1: return launch_adjoint_type<double>(cfg, s, plan->device, desc, count, queue + g, base);
    case 2: return launch_adjoint_type<zc>(cfg, s, plan->device, desc, count, queue + g, base);
    default: return launch_adjoint_type<zd>(cfg, s, plan->device, desc, count, queue + g, base);
    }
  });
}


// one homogeneous tuning batch: `count` matrices of order n and type dtype, tile
// configuration cfg; returns the median of `reps` timed launches in ms
template <typename T>
clm_status_t time_config(clm_ctx *ctx, hipStream_t st, int dtype, int n, int count, int cfg,
                         int reps, float *ms, bool adjoint = false) {
  const size_t mat = static_cast<size_t>(n) * n * sizeof(T);
  const int nmat = adjoint ? 4 : 2;  // A, E (+ G, dA)
  char *buf = nullptr;
  if (hipMalloc(&buf, nmat * mat * count) != hipSuccess) return CLM_STATUS_ALLOC_FAILED;
  std::vector<clm_batch_entry> entries(count);
  std::vector<clm_batch_ad_entry> ad(adjoint ? count : 0);
  for (int i = 0; i < count; ++i) {
    entries[i] = clm_batch_entry{n, dtype, buf + mat * i, n, buf + mat * (count + i), n};
    if (adjoint)
      ad[i] = clm_batch_ad_entry{n, dtype, entries[i].A, n, entries[i].E, n,
                                 buf + mat * (2 * count + i), n, buf + mat * (3 * count + i), n};
  }
  // A and (for the adjoint) the cotangents G: the first 2 count matrices are A
  // and E, so fill A, then G as a second batch of tuning matrices
  hipLaunchKernelGGL(fill_tuning_matrices<T>, dim3(1024), dim3(256), 0, st,
                     reinterpret_cast<T *>(buf), n, count, 12345u);
  if (adjoint)
    hipLaunchKernelGGL(fill_tuning_matrices<T>, dim3(1024), dim3(256), 0, st,
                       reinterpret_cast<T *>(buf + mat * 2 * count), n, count, 777u);
  clm_batch_plan *plan = nullptr;
  clm_status_t status = build_plan(count, entries.data(), &cfg, &plan,
                                   adjoint ? ad.data() : nullptr, adjoint ? 4 : 0);
  void *work = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<float> times;
  if (status == CLM_STATUS_SUCCESS && hipMalloc(&work, plan->ws_bytes) != hipSuccess)
    status = CLM_STATUS_ALLOC_FAILED;
  if (status == CLM_STATUS_SUCCESS &&
      (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    status = CLM_STATUS_EXECUTION_FAILED;
  for (int r = -1; status == CLM_STATUS_SUCCESS && r < reps; ++r) {
    if (hipEventRecord(e0, st) != hipSuccess) status = CLM_STATUS_EXECUTION_FAILED;
    if (status == CLM_STATUS_SUCCESS) status = launch_plan(ctx, st, plan, work, plan->ws_bytes, nullptr);
    if (status == CLM_STATUS_SUCCESS && adjoint)
      status = launch_backward(ctx, st, plan, work, plan->ws_bytes);
    if (status == CLM_STATUS_SUCCESS &&
        (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
      status = CLM_STATUS_EXECUTION_FAILED;
    float t = 0.f;
    if (status == CLM_STATUS_SUCCESS && r >= 0 && hipEventElapsedTime(&t, e0, e1) == hipSuccess)
      times.push_back(t);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (work) (void)hipFree(work);
  destroy_plan(plan);
  (void)hipFree(buf);
  if (status != CLM_STATUS_SUCCESS) return status;
  std::sort(times.begin(), times.end());
  *ms = times.empty() ? 0.f : times[times.size() / 2];
  return CLM_STATUS_SUCCESS;
}

}  // namespace
}  // namespace clm

extern "C" {

clm_status_t clm_expm_batched_plan_create(clm_batch_plan **plan, int count,
                                          const clm_batch_entry *entries) {
  return clm::build_plan(count, entries, nullptr, plan);
}

clm_status_t clm_expm_batched_ad_plan_create(clm_batch_plan **plan, int count,
                                             const clm_batch_ad_entry *entries, int checkpoints) {
  if (count < 1 || entries == nullptr) return CLM_STATUS_INVALID_VALUE;
  std::vector<clm_batch_entry> base(count);
  for (int i = 0; i < count; ++i)
    base[i] = clm_batch_entry{entries[i].n, entries[i].dtype, entries[i].A, entries[i].lda,
                              entries[i].E, entries[i].lde};
  return clm::build_plan(count, base.data(), nullptr, plan, entries, checkpoints);
}

clm_status_t clm_expm_batched_backward(clm_ctx *ctx, clm_stream_t stream,
                                       const clm_batch_plan *plan, void *dWork,
                                       size_t lwork_bytes) {
  return clm::launch_backward(ctx, static_cast<hipStream_t>(stream), plan, dWork, lwork_bytes);
}

clm_status_t clm_expm_batched_plan_destroy(clm_batch_plan *plan) {
  clm::destroy_plan(plan);
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_expm_batched_buffer_size(const clm_batch_plan *plan, size_t *bytes) {
  if (plan == nullptr || bytes == nullptr) return CLM_STATUS_INVALID_VALUE;
  *bytes = plan->ws_bytes;
  return CLM_STATUS_SUCCESS;
}

clm_status_t clm_expm_batched(clm_ctx *ctx, clm_stream_t stream, const clm_batch_plan *plan,
                              void *dWork, size_t lwork_bytes, clm_batch_result *dResults) {
  return clm::launch_plan(ctx, static_cast<hipStream_t>(stream), plan, dWork, lwork_bytes,
                          dResults);
}

clm_status_t clm_expm_batched_graph(clm_ctx *ctx, clm_stream_t stream, clm_batch_plan *plan,
                                    void *dWork, size_t lwork_bytes, clm_batch_result *dResults) {
  if (plan == nullptr || dWork == nullptr || stream == nullptr) return CLM_STATUS_INVALID_VALUE;
  if (lwork_bytes < plan->ws_bytes) return CLM_STATUS_ALLOC_FAILED;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool stale = plan->graph_exec == nullptr || plan->graph_stream != st ||
                     plan->graph_work != dWork || plan->graph_results != dResults;
  if (stale) {
    if (plan->graph_exec) {
      (void)hipGraphExecDestroy(plan->graph_exec);
      plan->graph_exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    CLM_TRY_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const clm_status_t status = clm::launch_plan(ctx, st, plan, dWork, lwork_bytes, dResults);
    const hipError_t end = hipStreamEndCapture(st, &graph);
    if (status != CLM_STATUS_SUCCESS) {
      if (graph) (void)hipGraphDestroy(graph);
      return status;
    }
    if (end != hipSuccess || graph == nullptr) return CLM_STATUS_EXECUTION_FAILED;
    const hipError_t inst = hipGraphInstantiate(&plan->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (inst != hipSuccess) {
      plan->graph_exec = nullptr;
      return CLM_STATUS_EXECUTION_FAILED;
    }
    plan->graph_stream = st;
    plan->graph_work = dWork;
    plan->graph_results = dResults;
  }
  CLM_TRY_HIP(hipGraphLaunch(plan->graph_exec, st));
  return CLM_STATUS_SUCCESS;
}

}  // extern "C"

namespace clm {
namespace {
clm_status_t get_tuning(int path, int dtype, int *cfg) {
  if (!valid_dtype(dtype) || cfg == nullptr) return CLM_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  if (!g_tune_init) tune_defaults();
  for (int b = 0; b < kBuckets; ++b) cfg[b] = g_tune[path][precision_of(dtype)][is_complex_code(dtype)][b];
  return CLM_STATUS_SUCCESS;
}

clm_status_t set_tuning(int path, int dtype, const int *cfg) {
  if (!valid_dtype(dtype) || cfg == nullptr) return CLM_STATUS_INVALID_VALUE;
  for (int b = 0; b < kBuckets; ++b)
    if (cfg[b] < 0 || cfg[b] >= CLM_BATCH_TILES) return CLM_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  if (!g_tune_init) tune_defaults();
  for (int b = 0; b < kBuckets; ++b) g_tune[path][precision_of(dtype)][is_complex_code(dtype)][b] = cfg[b];
  return CLM_STATUS_SUCCESS;
}

// times every distinct tile on one homogeneous batch per bucket up to nmax
// (n = 32 b + 16) and installs the fastest: forward launches of 1024 matrices,
// or (adjoint) forward + backward launches of 512
clm_status_t autotune(clm_ctx *ctx, hipStream_t st, int path, int dtype, int nmax, int *cfg_out,
                      float *ms_out) {
  if (!valid_dtype(dtype) || nmax < 1 || nmax > kMaxBatchN) return CLM_STATUS_INVALID_VALUE;
  int chosen[kBuckets];
  CLM_TRY(get_tuning(path, dtype, chosen));
  const bool single = precision_of(dtype) == 0;
  const bool cplx = is_complex_code(dtype);
  // candidate tiles: those that are distinct for this type (128 x 128 of 4 waves
  // exists for real FP32 only, 128 x 64 for all but complex FP64)
  auto distinct = [&](int cfg) {
    if (cfg == CLM_BATCH_TILE_128W16) return true;
    if (cfg == CLM_BATCH_TILE_128) return single && !cplx;
    if (cfg == CLM_BATCH_TILE_128x64) return single || !cplx;
    return true;
  };
  const bool adj = path == 1;
  for (int b = 0; b < kBuckets && 32 * b < nmax; ++b) {
    const int n = std::min(32 * b + 16, nmax);
    // enough matrices to keep every workgroup slot busy
    const int count = adj ? 512 : 1024;
    float best = 0.f;
    for (int cfg = CLM_BATCH_TILE_32; cfg < CLM_BATCH_TILES; ++cfg) {
      if (!distinct(cfg)) continue;
      float ms = 0.f;
      clm_status_t s;
      switch (dtype) {
      case CLM_DTYPE_F32: s = time_config<float>(ctx, st, dtype, n, count, cfg, 3, &ms, adj); break;
      case CLM_DTYPE_F64: s = time_config<double>(ctx, st, dtype, n, count, cfg, 3, &ms, adj); break;
      case CLM_DTYPE_C64: s = time_config<zc>(ctx, st, dtype, n, count, cfg, 3, &ms, adj); break;
      default: s = time_config<zd>(ctx, st, dtype, n, count, cfg, 3, &ms, adj); break;
      }
      if (s != CLM_STATUS_SUCCESS) return s;
      if (ms_out) ms_out[b * CLM_BATCH_TILES + cfg] = ms;
      if (cfg == CLM_BATCH_TILE_32 || ms < best) {
        best = ms;
        chosen[b] = cfg;
      }
    }
  }
  if (cfg_out)
    for (int b = 0; b < kBuckets; ++b) cfg_out[b] = chosen[b];
  return set_tuning(path, dtype, chosen);
}
}  // namespace
}  // namespace clm

extern "C" {

clm_status_t clm_expm_batched_get_tuning(int dtype, int *cfg) { return clm::get_tuning(0, dtype, cfg); }
clm_status_t clm_expm_batched_set_tuning(int dtype, const int *cfg) {
  return clm::set_tuning(0, dtype, cfg);
}
clm_status_t clm_expm_batched_get_tuning_ad(int dtype, int *cfg) {
  return clm::get_tuning(1, dtype, cfg);
}
clm_status_t clm_expm_batched_set_tuning_ad(int dtype, const int *cfg) {
  return clm::set_tuning(1, dtype, cfg);
}

clm_status_t clm_expm_batched_autotune(clm_ctx *ctx, clm_stream_t stream, int dtype, int nmax,
                                       int *cfg_out, float *ms_out) {
  return clm::autotune(ctx, static_cast<hipStream_t>(stream), 0, dtype, nmax, cfg_out, ms_out);
}
clm_status_t clm_expm_batched_autotune_ad(clm_ctx *ctx, clm_stream_t stream, int dtype, int nmax,
                                          int *cfg_out, float *ms_out) {
  return clm::autotune(ctx, static_cast<hipStream_t>(stream), 1, dtype, nmax, cfg_out, ms_out);
}

}  // extern "C"
