// gemm.hip -- matrix products on gfx950 MFMA for the four element types:
// FP64 and FP32 real, complex128 and complex64.
//
// C = alpha*A*B [+ beta*Cin] [+ B_blk], column-major NN, where B_blk is the
// Paterson-Stockmeyer block c0 I + c1 X1 + sum_t c_t bank[t-2] (or the Horner
// step's c I) formed in the epilogue. Fusing the block into the product's
// store is what removes the reference's separate build_block launch and its
// n^2 round trip through HBM before every Horner step
// (src/paterson_stockmeyer/interface.cppm:405-419 builds B_j, then runs the
// gemm with beta = 1 over it).
//
// Kernel shape:
//   * a BM x BN tile of C per workgroup of NT lanes (NT/64 wave64), each wave
//     owning a WM x WN sub-tile built from 16 x 16 MFMA tiles
//     (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32, 4 accumulator values
//     per lane);
//   * K advances in slices of BK, double-buffered in LDS: the next slice is
//     fetched global -> VGPR while the current one is multiplied, then written
//     to the other LDS stage; one barrier per slice. The loop body has no
//     branch (the last slice is peeled), which keeps the accumulators in
//     registers across iterations;
//   * A is staged as As[k][m] (m contiguous, as it arrives from column-major
//     memory), B as Bs[n][k] with a row of BK + one 16-byte vector, so the 16
//     lanes that read one k of B hit 16 distinct bank groups;
//   * workgroups are renumbered so each of the 8 XCDs (dispatched round robin
//     on the workgroup id) owns a contiguous run of tiles, walked in groups of
//     4 tile-rows: the A panels an XCD touches stay in its own L2.
//
// MFMA operand layout (16x16x4): A lane l holds A[l%16][l/16], B lane l holds
// B[l/16][l%16]; D lane l holds column l%16 at rows l/16 + 4r for the FP64
// instruction and rows 4(l/16) + r for the FP32 one (r = 0..3). The FP64 MFMA
// accumulates the four k terms of a step in order, as a sequential FMA chain.
#include "clm_internal.hpp"

namespace clm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kBK = 16;

template <typename S>
struct Mfma;

template <>
struct Mfma<double> {
  using acc = d4;
  using vec = double2;  // one 16-byte access
  static constexpr int EV = 2;
  __device__ static acc zero() { return d4{0.0, 0.0, 0.0, 0.0}; }
  __device__ static acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <>
struct Mfma<float> {
  using acc = f4;
  using vec = float4;
  static constexpr int EV = 4;
  __device__ static acc zero() { return f4{0.f, 0.f, 0.f, 0.f}; }
  __device__ static acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int &tm, int &tn) {
  const int total = tiles_m * tiles_n;
  int t = bid;
  if ((total & 7) == 0) t = (bid & 7) * (total >> 3) + (bid >> 3);
  constexpr int G = 4;
  const int per_group = G * tiles_n;
  const int group = t / per_group;
  const int first_m = group * G;
  const int gsize = min(tiles_m - first_m, G);
  const int r = t - group * per_group;
  tm = first_m + r % gsize;
  tn = r / gsize;
}

template <typename T>
__device__ __forceinline__ void epilogue_store(const GemmArgs<T> &g, int row, int col, T acc) {
  T v = ops<T>::mul(g.alpha, acc);
  if (g.Cin) v = ops<T>::add(v, ops<T>::mul(g.beta, g.Cin[row + (long)col * g.ldcin]));
  if (g.nterms > 0) {
    T e = (row == col) ? g.coef[0] : ops<T>::zero();
    if (g.nterms > 1) e = ops<T>::mad(g.coef[1], g.X1[row + (long)col * g.ldx1], e);
    for (int t = 2; t < g.nterms; ++t)
      e = ops<T>::mad(g.coef[t], g.bank[(long)(t - 2) * g.bank_stride + row + (long)col * g.ldbank],
                      e);
    v = ops<T>::add(v, e);
  }
  g.C[row + (long)col * g.ldc] = v;
}

// ------------------------------------------------------------------------------------
// real (S = double or float)
// ------------------------------------------------------------------------------------
template <typename S, int BM, int BN, int NT, int BK>
struct RTile {
  static constexpr int EV = Mfma<S>::EV;
  static constexpr int LDA_S = BM + 16;  // As[k][m]
  static constexpr int LDB_S = BK + EV;  // Bs[n][k]
  static constexpr int A_V = BM * BK / EV / NT;  // 16-byte vectors per lane per slice
  static constexpr int B_V = BN * BK / EV / NT;
  static constexpr int A_STAGE = BK * LDA_S;
  static constexpr int B_STAGE = BN * LDB_S;
  static_assert(A_V >= 1 && B_V >= 1 && (BM * BK / EV) % NT == 0 && (BN * BK / EV) % NT == 0,
                "tile staging must split evenly over the lanes");
};

template <typename S, int BM, int BN, int NT, int BK, bool kFast>
__device__ __forceinline__ void rgemm_gload(
    const GemmArgs<S> &g, int m0, int n0, int k0, int tid,
    typename Mfma<S>::vec (&ra)[RTile<S, BM, BN, NT, BK>::A_V],
    typename Mfma<S>::vec (&rb)[RTile<S, BM, BN, NT, BK>::B_V]) {
  using TL = RTile<S, BM, BN, NT, BK>;
  using V = typename Mfma<S>::vec;
  constexpr int EV = TL::EV;
#pragma unroll
  for (int it = 0; it < TL::A_V; ++it) {
    const int idx = tid + it * NT;
    const int mm = (idx % (BM / EV)) * EV, k = idx / (BM / EV);
    if constexpr (kFast) {
      ra[it] = *reinterpret_cast<const V *>(g.A + (m0 + mm) + (long)(k0 + k) * g.lda);
    } else {
      const int gm = m0 + mm, gk = k0 + k;
      S v[EV];
#pragma unroll
      for (int e = 0; e < EV; ++e)
        v[e] = (gk < g.K && gm + e < g.M) ? g.A[gm + e + (long)gk * g.lda] : S(0);
      ra[it] = *reinterpret_cast<const V *>(v);
    }
  }
#pragma unroll
  for (int it = 0; it < TL::B_V; ++it) {
    const int idx = tid + it * NT;
    const int kk = (idx % (BK / EV)) * EV, nn = idx / (BK / EV);
    if constexpr (kFast) {
      rb[it] = *reinterpret_cast<const V *>(g.B + (k0 + kk) + (long)(n0 + nn) * g.ldb);
    } else {
      const int gk = k0 + kk, gn = n0 + nn;
      S v[EV];
#pragma unroll
      for (int e = 0; e < EV; ++e)
        v[e] = (gn < g.N && gk + e < g.K) ? g.B[gk + e + (long)gn * g.ldb] : S(0);
      rb[it] = *reinterpret_cast<const V *>(v);
    }
  }
}

template <typename S, int BM, int BN, int NT, int BK>
__device__ __forceinline__ void rgemm_sstore(
    S *As, S *Bs, int tid, const typename Mfma<S>::vec (&ra)[RTile<S, BM, BN, NT, BK>::A_V],
    const typename Mfma<S>::vec (&rb)[RTile<S, BM, BN, NT, BK>::B_V]) {
  using TL = RTile<S, BM, BN, NT, BK>;
  using V = typename Mfma<S>::vec;
  constexpr int EV = TL::EV;
#pragma unroll
  for (int it = 0; it < TL::A_V; ++it) {
    const int idx = tid + it * NT;
    const int mm = (idx % (BM / EV)) * EV, k = idx / (BM / EV);
    *reinterpret_cast<V *>(&As[k * TL::LDA_S + mm]) = ra[it];
  }
#pragma unroll
  for (int it = 0; it < TL::B_V; ++it) {
    const int idx = tid + it * NT;
    const int kk = (idx % (BK / EV)) * EV, nn = idx / (BK / EV);
    *reinterpret_cast<V *>(&Bs[nn * TL::LDB_S + kk]) = rb[it];
  }
}

template <typename S, int BM, int BN, int WM, int WN, int NT, int BK>
__device__ __forceinline__ void rgemm_compute(const S *As, const S *Bs, int lane, int wm0, int wn0,
                                              typename Mfma<S>::acc (&acc)[WM / 16][WN / 16]) {
  using TL = RTile<S, BM, BN, NT, BK>;
  constexpr int TM = WM / 16, TN = WN / 16;
#pragma unroll
  for (int kk = 0; kk < BK; kk += 4) {
    const int kr = kk + (lane >> 4);
    S a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) a[i] = As[kr * TL::LDA_S + wm0 + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < TN; ++j) b[j] = Bs[(wn0 + j * 16 + (lane & 15)) * TL::LDB_S + kr];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<S>::mma(a[i], b[j], acc[i][j]);
  }
}

template <typename S, int BM, int BN, int WM, int WN, int NT, bool kFast, int BK = kBK>
__global__ void __launch_bounds__(NT)
    rgemm_mfma(const GemmArgs<S> g, const int tiles_m, const int tiles_n) {
  using TL = RTile<S, BM, BN, NT, BK>;
  using V = typename Mfma<S>::vec;
  constexpr int WAVES_M = BM / WM;
  static_assert((BM / WM) * (BN / WN) == NT / 64, "one wave per WM x WN sub-tile");
  constexpr int TM = WM / 16, TN = WN / 16;

  __shared__ __attribute__((aligned(16))) S As[2 * TL::A_STAGE];
  __shared__ __attribute__((aligned(16))) S Bs[2 * TL::B_STAGE];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int tm, tn;
  tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int wm0 = (wave % WAVES_M) * WM;
  const int wn0 = (wave / WAVES_M) * WN;

  typename Mfma<S>::acc acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<S>::zero();

  V ra[TL::A_V], rb[TL::B_V];
  const int KT = (g.K + BK - 1) / BK;
  if (KT > 0) {
    rgemm_gload<S, BM, BN, NT, BK, kFast>(g, m0, n0, 0, tid, ra, rb);
    rgemm_sstore<S, BM, BN, NT, BK>(As, Bs, tid, ra, rb);
    __syncthreads();
    for (int kt = 0; kt < KT - 1; ++kt) {
      const int cur = kt & 1;
      rgemm_gload<S, BM, BN, NT, BK, kFast>(g, m0, n0, (kt + 1) * BK, tid, ra, rb);
      rgemm_compute<S, BM, BN, WM, WN, NT, BK>(As + cur * TL::A_STAGE, Bs + cur * TL::B_STAGE,
                                               lane, wm0, wn0, acc);
      rgemm_sstore<S, BM, BN, NT, BK>(As + (cur ^ 1) * TL::A_STAGE, Bs + (cur ^ 1) * TL::B_STAGE,
                                      tid, ra, rb);
      __syncthreads();
    }
    const int last = (KT - 1) & 1;
    rgemm_compute<S, BM, BN, WM, WN, NT, BK>(As + last * TL::A_STAGE, Bs + last * TL::B_STAGE,
                                             lane, wm0, wn0, acc);
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + Mfma<S>::row(lane, r);
        const int col = n0 + wn0 + j * 16 + (lane & 15);
        if (kFast || (row < g.M && col < g.N)) epilogue_store(g, row, col, S(acc[i][j][r]));
      }
}

// ------------------------------------------------------------------------------------
// complex (C = zd or zc): four real MFMA products per 16x16 tile and k-step,
//   Cr += Ar Br - Ai Bi,  Ci += Ar Bi + Ai Br,
// with real and imaginary planes staged separately in LDS, double-buffered.
// ------------------------------------------------------------------------------------
template <typename C>
struct CxReal;
template <>
struct CxReal<zd> {
  using type = double;
};
template <>
struct CxReal<zc> {
  using type = float;
};

template <typename C, int BM, int BN, int NT, int BK>
struct CTile {
  using S = typename CxReal<C>::type;
  static constexpr int LDA_S = BM + 16;
  static constexpr int LDB_S = BK + Mfma<S>::EV;
  static constexpr int A_E = BM * BK / NT;  // complex elements per lane per slice
  static constexpr int B_E = BN * BK / NT;
  static constexpr int A_PLANE = BK * LDA_S;
  static constexpr int B_PLANE = BN * LDB_S;
  static constexpr int STAGE = 2 * (A_PLANE + B_PLANE);  // re/im planes of A and B
  static_assert(A_E >= 1 && B_E >= 1 && (BM * BK) % NT == 0 && (BN * BK) % NT == 0,
                "tile staging must split evenly over the lanes");
};

template <typename C, int BM, int BN, int NT, int BK, bool kFast>
__device__ __forceinline__ void cgemm_gload(const GemmArgs<C> &g, int m0, int n0, int k0, int tid,
                                            C (&ra)[CTile<C, BM, BN, NT, BK>::A_E],
                                            C (&rb)[CTile<C, BM, BN, NT, BK>::B_E]) {
  using TL = CTile<C, BM, BN, NT, BK>;
#pragma unroll
  for (int it = 0; it < TL::A_E; ++it) {
    const int e = tid + it * NT;
    const int gm = m0 + e % BM, gk = k0 + e / BM;
    if (kFast || (gm < g.M && gk < g.K)) ra[it] = g.A[gm + (long)gk * g.lda];
    else ra[it] = ops<C>::zero();
  }
#pragma unroll
  for (int it = 0; it < TL::B_E; ++it) {
    const int e = tid + it * NT;
    const int gk = k0 + e % BK, gn = n0 + e / BK;
    if (kFast || (gk < g.K && gn < g.N)) rb[it] = g.B[gk + (long)gn * g.ldb];
    else rb[it] = ops<C>::zero();
  }
}

template <typename C, int BM, int BN, int NT, int BK>
__device__ __forceinline__ void cgemm_sstore(typename CxReal<C>::type *L, int tid,
                                             const C (&ra)[CTile<C, BM, BN, NT, BK>::A_E],
                                             const C (&rb)[CTile<C, BM, BN, NT, BK>::B_E]) {
  using TL = CTile<C, BM, BN, NT, BK>;
#pragma unroll
  for (int it = 0; it < TL::A_E; ++it) {
    const int e = tid + it * NT;
    const int s = (e / BM) * TL::LDA_S + (e % BM);
    L[s] = ra[it].re;
    L[TL::A_PLANE + s] = ra[it].im;
  }
#pragma unroll
  for (int it = 0; it < TL::B_E; ++it) {
    const int e = tid + it * NT;
    const int s = (e / BK) * TL::LDB_S + (e % BK);
    L[2 * TL::A_PLANE + s] = rb[it].re;
    L[2 * TL::A_PLANE + TL::B_PLANE + s] = rb[it].im;
  }
}

template <typename C, int BM, int BN, int WM, int WN, int NT, int BK>
__device__ __forceinline__ void cgemm_compute(
    const typename CxReal<C>::type *L, int lane, int wm0, int wn0,
    typename Mfma<typename CxReal<C>::type>::acc (&cr)[WM / 16][WN / 16],
    typename Mfma<typename CxReal<C>::type>::acc (&ci)[WM / 16][WN / 16]) {
  using TL = CTile<C, BM, BN, NT, BK>;
  using S = typename CxReal<C>::type;
  constexpr int TM = WM / 16, TN = WN / 16;
#pragma unroll
  for (int kk = 0; kk < BK; kk += 4) {
    const int kr = kk + (lane >> 4);
    S ar[TM], ai[TM], nai[TM], br[TN], bi[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int s = kr * TL::LDA_S + wm0 + i * 16 + (lane & 15);
      ar[i] = L[s];
      ai[i] = L[TL::A_PLANE + s];
      nai[i] = -ai[i];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int s = (wn0 + j * 16 + (lane & 15)) * TL::LDB_S + kr;
      br[j] = L[2 * TL::A_PLANE + s];
      bi[j] = L[2 * TL::A_PLANE + TL::B_PLANE + s];
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        cr[i][j] = Mfma<S>::mma(ar[i], br[j], cr[i][j]);
        ci[i][j] = Mfma<S>::mma(ar[i], bi[j], ci[i][j]);
        cr[i][j] = Mfma<S>::mma(nai[i], bi[j], cr[i][j]);
        ci[i][j] = Mfma<S>::mma(ai[i], br[j], ci[i][j]);
      }
  }
}

template <typename C, int BM, int BN, int WM, int WN, int NT, bool kFast, int BK = kBK>
__global__ void __launch_bounds__(NT)
    cgemm_mfma(const GemmArgs<C> g, const int tiles_m, const int tiles_n) {
  using TL = CTile<C, BM, BN, NT, BK>;
  using S = typename CxReal<C>::type;
  constexpr int WAVES_M = BM / WM;
  static_assert((BM / WM) * (BN / WN) == NT / 64, "one wave per WM x WN sub-tile");
  constexpr int TM = WM / 16, TN = WN / 16;

  __shared__ __attribute__((aligned(16))) S L[2 * TL::STAGE];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int tm, tn;
  tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int wm0 = (wave % WAVES_M) * WM;
  const int wn0 = (wave / WAVES_M) * WN;

  typename Mfma<S>::acc cr[TM][TN], ci[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      cr[i][j] = Mfma<S>::zero();
      ci[i][j] = Mfma<S>::zero();
    }

  C ra[TL::A_E], rb[TL::B_E];
  const int KT = (g.K + BK - 1) / BK;
  if (KT > 0) {
    cgemm_gload<C, BM, BN, NT, BK, kFast>(g, m0, n0, 0, tid, ra, rb);
    cgemm_sstore<C, BM, BN, NT, BK>(L, tid, ra, rb);
    __syncthreads();
    for (int kt = 0; kt < KT - 1; ++kt) {
      const int cur = kt & 1;
      cgemm_gload<C, BM, BN, NT, BK, kFast>(g, m0, n0, (kt + 1) * BK, tid, ra, rb);
      cgemm_compute<C, BM, BN, WM, WN, NT, BK>(L + cur * TL::STAGE, lane, wm0, wn0, cr, ci);
      cgemm_sstore<C, BM, BN, NT, BK>(L + (cur ^ 1) * TL::STAGE, tid, ra, rb);
      __syncthreads();
    }
    cgemm_compute<C, BM, BN, WM, WN, NT, BK>(L + ((KT - 1) & 1) * TL::STAGE, lane, wm0, wn0, cr,
                                             ci);
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + Mfma<S>::row(lane, r);
        const int col = n0 + wn0 + j * 16 + (lane & 15);
        if (kFast || (row < g.M && col < g.N))
          epilogue_store(g, row, col, C{S(cr[i][j][r]), S(ci[i][j][r])});
      }
}

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <int BM, int BN, int NT, int BK, typename Kfast, typename Kgen, typename T>
hipError_t dispatch(hipStream_t stream, const GemmArgs<T> &g, bool fast_layout, Kfast kfast,
                    Kgen kgen) {
  const int tiles_m = (g.M + BM - 1) / BM;
  const int tiles_n = (g.N + BN - 1) / BN;
  const bool fast = fast_layout && g.M % BM == 0 && g.N % BN == 0 && g.K % BK == 0;
  const dim3 grid(tiles_m * tiles_n), block(NT);
  if (fast) hipLaunchKernelGGL(kfast, grid, block, 0, stream, g, tiles_m, tiles_n);
  else hipLaunchKernelGGL(kgen, grid, block, 0, stream, g, tiles_m, tiles_n);
  return hipGetLastError();
}

// Real products: 128 x 128 tiles of 16 waves (32 x 32 each, K slices of 32)
// once they fill the 256 CUs -- the fastest of the swept FP64 shapes at n = 2048
// and 4096, 4 waves per SIMD hiding the per-slice barrier; 64 x 64 tiles of
// 4 waves below that.
template <typename S>
hipError_t launch_real(hipStream_t stream, const GemmArgs<S> &g) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  constexpr int EV = Mfma<S>::EV;
  const bool fl = (g.lda % EV == 0) && (g.ldb % EV == 0) && aligned16(g.A) && aligned16(g.B);
  const long big_tiles = static_cast<long>((g.M + 127) / 128) * ((g.N + 127) / 128);
  if (big_tiles >= 256)
    return dispatch<128, 128, 1024, 32>(stream, g, fl,
                                        rgemm_mfma<S, 128, 128, 32, 32, 1024, true, 32>,
                                        rgemm_mfma<S, 128, 128, 32, 32, 1024, false, 32>);
  return dispatch<64, 64, 256, kBK>(stream, g, fl, rgemm_mfma<S, 64, 64, 32, 32, 256, true>,
                                    rgemm_mfma<S, 64, 64, 32, 32, 256, false>);
}

// Complex products: 128 x 64 tiles of 8 waves (32 x 32 each) once they fill the
// CUs, 64 x 64 tiles of 4 waves below that; K slices of 16, double-buffered.
template <typename C>
hipError_t launch_complex(hipStream_t stream, const GemmArgs<C> &g) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  const long big_tiles = static_cast<long>((g.M + 127) / 128) * ((g.N + 63) / 64);
  if (big_tiles >= 256)
    return dispatch<128, 64, 512, kBK>(stream, g, true, cgemm_mfma<C, 128, 64, 32, 32, 512, true>,
                                       cgemm_mfma<C, 128, 64, 32, 32, 512, false>);
  return dispatch<64, 64, 256, kBK>(stream, g, true, cgemm_mfma<C, 64, 64, 32, 32, 256, true>,
                                    cgemm_mfma<C, 64, 64, 32, 32, 256, false>);
}

}  // namespace

hipError_t launch_gemm_d(hipStream_t stream, const GemmArgs<double> &g) {
  return launch_real<double>(stream, g);
}
hipError_t launch_gemm_z(hipStream_t stream, const GemmArgs<zd> &g) {
  return launch_complex<zd>(stream, g);
}

template <>
hipError_t launch_gemm<double>(hipStream_t stream, const GemmArgs<double> &g) {
  return launch_real<double>(stream, g);
}
template <>
hipError_t launch_gemm<float>(hipStream_t stream, const GemmArgs<float> &g) {
  return launch_real<float>(stream, g);
}
template <>
hipError_t launch_gemm<zd>(hipStream_t stream, const GemmArgs<zd> &g) {
  return launch_complex<zd>(stream, g);
}
template <>
hipError_t launch_gemm<zc>(hipStream_t stream, const GemmArgs<zc> &g) {
  return launch_complex<zc>(stream, g);
}

}  // namespace clm
