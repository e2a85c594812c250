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
#include "gemm_tile.hpp"

namespace clm {
namespace {

using namespace tile;

template <typename S, int BM, int BN, int WM, int WN, int NT, bool kFast, int BK = kBK>
__global__ void __launch_bounds__(NT)
    rgemm_mfma(const GemmArgs<S> g, const int tiles_m, const int tiles_n) {
  using TL = RTile<S, BM, BN, NT, BK>;
  using V = typename Mfma<S>::vec;
  constexpr int WAVES_M = BM / WM;
  static_assert((BM / WM) * (BN / WN) == NT / 64, "one wave per WM x WN sub-tile");
  constexpr int TM = WM / 16, TN = WN / 16;

  __shared__ __attribute__((aligned(16))) S L[2 * (TL::A_STAGE + TL::B_STAGE)];
  S *As = L;
  S *Bs = L + 2 * TL::A_STAGE;

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
      // the next slice goes into the other stage half-way through this one, so its
      // LDS writes overlap the other waves' MFMAs instead of idling the CU before
      // the barrier (the k-steps still run in ascending order)
      rgemm_compute<S, BM, BN, WM, WN, NT, BK, 0, BK / 2>(As + cur * TL::A_STAGE,
                                                         Bs + cur * TL::B_STAGE, lane, wm0, wn0, acc);
      rgemm_sstore<S, BM, BN, NT, BK>(As + (cur ^ 1) * TL::A_STAGE, Bs + (cur ^ 1) * TL::B_STAGE,
                                      tid, ra, rb);
      rgemm_compute<S, BM, BN, WM, WN, NT, BK, BK / 2, BK>(As + cur * TL::A_STAGE,
                                                          Bs + cur * TL::B_STAGE, lane, wm0, wn0, acc);
      __syncthreads();
    }
    const int last = (KT - 1) & 1;
    rgemm_compute<S, BM, BN, WM, WN, NT, BK>(As + last * TL::A_STAGE, Bs + last * TL::B_STAGE,
                                             lane, wm0, wn0, acc);
  }

  // 128 x 128 tiles: the sums go column-major into the free LDS stages and the
  // whole workgroup writes them out in coalesced column pieces
  constexpr int LDT = 128 + 16 / sizeof(S);
  if constexpr (BM == 128 && BN == 128 && 128 * LDT <= 2 * (TL::A_STAGE + TL::B_STAGE)) {
    __syncthreads();  // every wave is done with the stages
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          L[(wn0 + j * 16 + (lane & 15)) * LDT + wm0 + i * 16 + Mfma<S>::row(lane, r)] =
              S(acc[i][j][r]);
    __syncthreads();
    lds_tile_epilogue<S, NT / 64, kFast>(g, m0, n0, [&](int c, int r) { return L[c * LDT + r]; });
  } else {
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
