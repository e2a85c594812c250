// gemm.hip -- FP64 and complex128 matrix products on gfx950 MFMA.
//
// C = alpha*A*B [+ beta*Cin] [+ B_blk], column-major NN, where B_blk is the
// Paterson-Stockmeyer block c0 I + c1 X1 + sum_t c_t bank[t-2] (or the Horner
// step's c I) formed in the epilogue. Fusing the block into the product's
// store is what removes the reference's separate build_block launch and its
// n^2 round trip through HBM before every Horner step
// (src/paterson_stockmeyer/interface.cppm:405-419 builds B_j, then runs the
// gemm with beta = 1 over it).
//
// Kernel shape (256 lanes = 4 wave64 per workgroup):
//   * a BM x BN tile of C per workgroup, 2 x 2 waves each owning WM x WN,
//     built from 16 x 16 MFMA tiles (v_mfma_f64_16x16x4_f64, 4 doubles of D
//     per lane);
//   * K advances in slices of 16: the next slice is fetched global -> VGPR
//     while the current one is multiplied out of LDS, then written to LDS
//     between two barriers;
//   * A is staged as As[k][m] (m contiguous, as it arrives from column-major
//     memory), B as Bs[n][k] with a row of 18 doubles, so the 16 lanes that read
//     one k of B hit 16 distinct bank pairs;
//   * workgroups are renumbered so each of the 8 XCDs (dispatched round robin
//     on the workgroup id) owns a contiguous run of tiles, walked in groups of
//     4 tile-rows: the A panels an XCD touches stay in its own L2.
//
// MFMA operand layout (16x16x4 f64): A lane l holds A[l%16][l/16], B lane l
// holds B[l/16][l%16], D lane l holds rows l/16 + 4r (r = 0..3) of column l%16.
#include "clm_internal.hpp"

namespace clm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kBK = 16;

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
// FP64
//
// NT lanes (NT/64 waves) per workgroup; K slices of 16 are double-buffered in
// LDS, so one barrier per slice separates "everyone has finished reading stage
// s" from "stage s is refilled". The loop body has no branch: the last slice is
// peeled, which keeps the accumulators resident in AGPRs across iterations.
// ------------------------------------------------------------------------------------
template <int BM, int BN, int NT, int BK>
struct DTile {
  static constexpr int LDA_S = BM + 16;  // As[k][m]
  static constexpr int LDB_S = BK + 2;   // Bs[n][k]
  static constexpr int A_V2 = BM * BK / 2 / NT;  // double2 per lane per slice
  static constexpr int B_V2 = BN * BK / 2 / NT;
  static constexpr int A_STAGE = BK * LDA_S;
  static constexpr int B_STAGE = BN * LDB_S;
  static_assert(A_V2 >= 1 && B_V2 >= 1 && (BM * BK / 2) % NT == 0 && (BN * BK / 2) % NT == 0,
                "tile staging must split evenly over the lanes");
};

template <int BM, int BN, int NT, int BK, bool kFast>
__device__ __forceinline__ void dgemm_gload(const GemmArgs<double> &g, int m0, int n0, int k0,
                                            int tid, double2 (&ra)[DTile<BM, BN, NT, BK>::A_V2],
                                            double2 (&rb)[DTile<BM, BN, NT, BK>::B_V2]) {
  using TL = DTile<BM, BN, NT, BK>;
#pragma unroll
  for (int it = 0; it < TL::A_V2; ++it) {
    const int idx = tid + it * NT;
    const int mm = (idx % (BM / 2)) * 2, k = idx / (BM / 2);
    if constexpr (kFast) {
      ra[it] = *reinterpret_cast<const double2 *>(g.A + (m0 + mm) + (long)(k0 + k) * g.lda);
    } else {
      const int gm = m0 + mm, gk = k0 + k;
      const bool kin = gk < g.K;
      ra[it].x = (kin && gm < g.M) ? g.A[gm + (long)gk * g.lda] : 0.0;
      ra[it].y = (kin && gm + 1 < g.M) ? g.A[gm + 1 + (long)gk * g.lda] : 0.0;
    }
  }
#pragma unroll
  for (int it = 0; it < TL::B_V2; ++it) {
    const int idx = tid + it * NT;
    const int kk = (idx % (BK / 2)) * 2, nn = idx / (BK / 2);
    if constexpr (kFast) {
      rb[it] = *reinterpret_cast<const double2 *>(g.B + (k0 + kk) + (long)(n0 + nn) * g.ldb);
    } else {
      const int gk = k0 + kk, gn = n0 + nn;
      const bool nin = gn < g.N;
      rb[it].x = (nin && gk < g.K) ? g.B[gk + (long)gn * g.ldb] : 0.0;
      rb[it].y = (nin && gk + 1 < g.K) ? g.B[gk + 1 + (long)gn * g.ldb] : 0.0;
    }
  }
}

template <int BM, int BN, int NT, int BK>
__device__ __forceinline__ void dgemm_sstore(double *As, double *Bs, int tid,
                                             const double2 (&ra)[DTile<BM, BN, NT, BK>::A_V2],
                                             const double2 (&rb)[DTile<BM, BN, NT, BK>::B_V2]) {
  using TL = DTile<BM, BN, NT, BK>;
#pragma unroll
  for (int it = 0; it < TL::A_V2; ++it) {
    const int idx = tid + it * NT;
    const int mm = (idx % (BM / 2)) * 2, k = idx / (BM / 2);
    *reinterpret_cast<double2 *>(&As[k * TL::LDA_S + mm]) = ra[it];
  }
#pragma unroll
  for (int it = 0; it < TL::B_V2; ++it) {
    const int idx = tid + it * NT;
    const int kk = (idx % (BK / 2)) * 2, nn = idx / (BK / 2);
    *reinterpret_cast<double2 *>(&Bs[nn * TL::LDB_S + kk]) = rb[it];
  }
}

template <int BM, int BN, int WM, int WN, int NT, int BK>
__device__ __forceinline__ void dgemm_compute(const double *As, const double *Bs, int lane, int wm0,
                                              int wn0, d4 (&acc)[WM / 16][WN / 16]) {
  using TL = DTile<BM, BN, NT, BK>;
  constexpr int TM = WM / 16, TN = WN / 16;
#pragma unroll
  for (int kk = 0; kk < BK; kk += 4) {
    const int kr = kk + (lane >> 4);
    double a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) a[i] = As[kr * TL::LDA_S + wm0 + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < TN; ++j) b[j] = Bs[(wn0 + j * 16 + (lane & 15)) * TL::LDB_S + kr];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
  }
}

template <int BM, int BN, int WM, int WN, int NT, bool kFast, int BK = kBK>
__global__ void __launch_bounds__(NT)
    dgemm_mfma(const GemmArgs<double> g, const int tiles_m, const int tiles_n) {
  using TL = DTile<BM, BN, NT, BK>;
  constexpr int WAVES_M = BM / WM;
  static_assert((BM / WM) * (BN / WN) == NT / 64, "one wave per WM x WN sub-tile");
  constexpr int TM = WM / 16, TN = WN / 16;

  __shared__ double As[2 * TL::A_STAGE];
  __shared__ double Bs[2 * TL::B_STAGE];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int tm, tn;
  tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int wm0 = (wave % WAVES_M) * WM;
  const int wn0 = (wave / WAVES_M) * WN;

  d4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  double2 ra[TL::A_V2], rb[TL::B_V2];
  const int KT = (g.K + BK - 1) / BK;
  if (KT > 0) {
    dgemm_gload<BM, BN, NT, BK, kFast>(g, m0, n0, 0, tid, ra, rb);
    dgemm_sstore<BM, BN, NT, BK>(As, Bs, tid, ra, rb);
    __syncthreads();
    for (int kt = 0; kt < KT - 1; ++kt) {
      const int cur = kt & 1;
      dgemm_gload<BM, BN, NT, BK, kFast>(g, m0, n0, (kt + 1) * BK, tid, ra, rb);
      dgemm_compute<BM, BN, WM, WN, NT, BK>(As + cur * TL::A_STAGE, Bs + cur * TL::B_STAGE, lane,
                                            wm0, wn0, acc);
      dgemm_sstore<BM, BN, NT, BK>(As + (cur ^ 1) * TL::A_STAGE, Bs + (cur ^ 1) * TL::B_STAGE, tid,
                                   ra, rb);
      __syncthreads();
    }
    const int last = (KT - 1) & 1;
    dgemm_compute<BM, BN, WM, WN, NT, BK>(As + last * TL::A_STAGE, Bs + last * TL::B_STAGE, lane,
                                          wm0, wn0, acc);
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + (lane >> 4) + 4 * r;
        const int col = n0 + wn0 + j * 16 + (lane & 15);
        if (kFast || (row < g.M && col < g.N)) epilogue_store(g, row, col, acc[i][j][r]);
      }
}

// ------------------------------------------------------------------------------------
// complex128: four real MFMA products per 16x16 tile and k-step,
//   Cr += Ar Br - Ai Bi,  Ci += Ar Bi + Ai Br,
// with real and imaginary planes staged separately in LDS.
// ------------------------------------------------------------------------------------
template <int BM, int BN, int WM, int WN, bool kFast>
__global__ void __launch_bounds__(kThreads)
    zgemm_mfma(const GemmArgs<zd> g, const int tiles_m, const int tiles_n) {
  constexpr int WAVES_M = BM / WM;
  static_assert((BM / WM) * (BN / WN) == 4, "4 waves per workgroup");
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int LDA_S = BM + 16;
  constexpr int LDB_S = kBK + 2;
  constexpr int A_ELEMS = BM * kBK / kThreads;
  constexpr int B_ELEMS = BN * kBK / kThreads;

  __shared__ double Ar[kBK * LDA_S];
  __shared__ double Ai[kBK * LDA_S];
  __shared__ double Br[BN * LDB_S];
  __shared__ double Bi[BN * LDB_S];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int tm, tn;
  tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int wm0 = (wave % WAVES_M) * WM;
  const int wn0 = (wave / WAVES_M) * WN;

  d4 cr[TM][TN], ci[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      cr[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      ci[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }

  zd ra[A_ELEMS], rb[B_ELEMS];

  auto gload = [&](const int k0) {
#pragma unroll
    for (int it = 0; it < A_ELEMS; ++it) {
      const int e = tid + it * kThreads;
      const int mm = e % BM, k = e / BM;
      const int gm = m0 + mm, gk = k0 + k;
      if (kFast || (gm < g.M && gk < g.K)) ra[it] = g.A[gm + (long)gk * g.lda];
      else ra[it] = zd{0.0, 0.0};
    }
#pragma unroll
    for (int it = 0; it < B_ELEMS; ++it) {
      const int e = tid + it * kThreads;
      const int k = e % kBK, nn = e / kBK;
      const int gk = k0 + k, gn = n0 + nn;
      if (kFast || (gk < g.K && gn < g.N)) rb[it] = g.B[gk + (long)gn * g.ldb];
      else rb[it] = zd{0.0, 0.0};
    }
  };

  auto sstore = [&]() {
#pragma unroll
    for (int it = 0; it < A_ELEMS; ++it) {
      const int e = tid + it * kThreads;
      const int s = (e / BM) * LDA_S + (e % BM);
      Ar[s] = ra[it].re;
      Ai[s] = ra[it].im;
    }
#pragma unroll
    for (int it = 0; it < B_ELEMS; ++it) {
      const int e = tid + it * kThreads;
      const int s = (e / kBK) * LDB_S + (e % kBK);
      Br[s] = rb[it].re;
      Bi[s] = rb[it].im;
    }
  };

  auto compute = [&]() {
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double ar[TM], ai[TM], nai[TM], br[TN], bi[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int s = kr * LDA_S + wm0 + i * 16 + (lane & 15);
        ar[i] = Ar[s];
        ai[i] = Ai[s];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int s = (wn0 + j * 16 + (lane & 15)) * LDB_S + kr;
        br[j] = Br[s];
        bi[j] = Bi[s];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], ci[i][j], 0, 0, 0);
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], ci[i][j], 0, 0, 0);
        }
    }
  };

  const int KT = (g.K + kBK - 1) / kBK;
  if (KT > 0) {
    gload(0);
    sstore();
  }
  __syncthreads();
  for (int kt = 0; kt < KT; ++kt) {
    const bool more = kt + 1 < KT;
    if (more) gload((kt + 1) * kBK);
    compute();
    __syncthreads();
    if (more) {
      sstore();
      __syncthreads();
    }
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + (lane >> 4) + 4 * r;
        const int col = n0 + wn0 + j * 16 + (lane & 15);
        if (kFast || (row < g.M && col < g.N))
          epilogue_store(g, row, col, zd{cr[i][j][r], ci[i][j][r]});
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

}  // namespace

hipError_t launch_gemm_d(hipStream_t stream, const GemmArgs<double> &g) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  const bool fl = (g.lda % 2 == 0) && (g.ldb % 2 == 0) && aligned16(g.A) && aligned16(g.B);
  // 128 x 128 tiles once they fill the 256 CUs; 64 x 64 below that
  const long big_tiles = static_cast<long>((g.M + 127) / 128) * ((g.N + 127) / 128);
  // 16 waves of 32 x 32 with K slices of 32: the fastest of the swept shapes at
  // n = 2048 and 4096 (4 waves per SIMD hide the per-slice barrier)
  if (big_tiles >= 256)
    return dispatch<128, 128, 1024, 32>(stream, g, fl,
                                        dgemm_mfma<128, 128, 32, 32, 1024, true, 32>,
                                        dgemm_mfma<128, 128, 32, 32, 1024, false, 32>);
  return dispatch<64, 64, 256, kBK>(stream, g, fl, dgemm_mfma<64, 64, 32, 32, 256, true>,
                                    dgemm_mfma<64, 64, 32, 32, 256, false>);
}

hipError_t launch_gemm_z(hipStream_t stream, const GemmArgs<zd> &g) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (g.M >= 1024 && g.N >= 1024)
    return dispatch<128, 64, kThreads, kBK>(stream, g, true, zgemm_mfma<128, 64, 64, 32, true>,
                                     zgemm_mfma<128, 64, 64, 32, false>);
  return dispatch<64, 64, kThreads, kBK>(stream, g, true, zgemm_mfma<64, 64, 32, 32, true>,
                                  zgemm_mfma<64, 64, 32, 32, false>);
}

template <>
hipError_t launch_gemm<double>(hipStream_t stream, const GemmArgs<double> &g) {
  return launch_gemm_d(stream, g);
}
template <>
hipError_t launch_gemm<zd>(hipStream_t stream, const GemmArgs<zd> &g) {
  return launch_gemm_z(stream, g);
}

}  // namespace clm
