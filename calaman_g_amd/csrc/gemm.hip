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
// ------------------------------------------------------------------------------------
template <int BM, int BN, int WM, int WN, bool kFast>
__global__ void __launch_bounds__(kThreads)
    dgemm_mfma(const GemmArgs<double> g, const int tiles_m, const int tiles_n) {
  constexpr int WAVES_M = BM / WM;
  static_assert((BM / WM) * (BN / WN) == 4, "4 waves per workgroup");
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int LDA_S = BM + 16;
  constexpr int LDB_S = kBK + 2;
  constexpr int A_ELEMS = BM * kBK / kThreads;
  constexpr int B_ELEMS = BN * kBK / kThreads;
  static_assert(A_ELEMS % 2 == 0 && B_ELEMS % 2 == 0, "double2 staging");

  __shared__ double As[kBK * LDA_S];
  __shared__ double Bs[BN * LDB_S];

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

  double ra[A_ELEMS], rb[B_ELEMS];

  auto gload = [&](const int k0) {
    if constexpr (kFast) {
      constexpr int ACOLS = kThreads / (BM / 2);
      const int mm = (tid % (BM / 2)) * 2;
      const int kq = tid / (BM / 2);
#pragma unroll
      for (int it = 0; it < A_ELEMS / 2; ++it) {
        const double2 v = *reinterpret_cast<const double2 *>(
            g.A + (m0 + mm) + (long)(k0 + kq + it * ACOLS) * g.lda);
        ra[2 * it] = v.x;
        ra[2 * it + 1] = v.y;
      }
      const int kk = (tid % (kBK / 2)) * 2;
      const int nq = tid / (kBK / 2);
      constexpr int BCOLS = kThreads / (kBK / 2);
#pragma unroll
      for (int it = 0; it < B_ELEMS / 2; ++it) {
        const double2 v = *reinterpret_cast<const double2 *>(
            g.B + (k0 + kk) + (long)(n0 + nq + it * BCOLS) * g.ldb);
        rb[2 * it] = v.x;
        rb[2 * it + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int it = 0; it < A_ELEMS; ++it) {
        const int e = tid + it * kThreads;
        const int mm = e % BM, k = e / BM;
        const int gm = m0 + mm, gk = k0 + k;
        ra[it] = (gm < g.M && gk < g.K) ? g.A[gm + (long)gk * g.lda] : 0.0;
      }
#pragma unroll
      for (int it = 0; it < B_ELEMS; ++it) {
        const int e = tid + it * kThreads;
        const int k = e % kBK, nn = e / kBK;
        const int gk = k0 + k, gn = n0 + nn;
        rb[it] = (gk < g.K && gn < g.N) ? g.B[gk + (long)gn * g.ldb] : 0.0;
      }
    }
  };

  auto sstore = [&]() {
    if constexpr (kFast) {
      constexpr int ACOLS = kThreads / (BM / 2);
      const int mm = (tid % (BM / 2)) * 2;
      const int kq = tid / (BM / 2);
#pragma unroll
      for (int it = 0; it < A_ELEMS / 2; ++it)
        *reinterpret_cast<double2 *>(&As[(kq + it * ACOLS) * LDA_S + mm]) =
            double2{ra[2 * it], ra[2 * it + 1]};
      const int kk = (tid % (kBK / 2)) * 2;
      const int nq = tid / (kBK / 2);
      constexpr int BCOLS = kThreads / (kBK / 2);
#pragma unroll
      for (int it = 0; it < B_ELEMS / 2; ++it)
        *reinterpret_cast<double2 *>(&Bs[(nq + it * BCOLS) * LDB_S + kk]) =
            double2{rb[2 * it], rb[2 * it + 1]};
    } else {
#pragma unroll
      for (int it = 0; it < A_ELEMS; ++it) {
        const int e = tid + it * kThreads;
        As[(e / BM) * LDA_S + (e % BM)] = ra[it];
      }
#pragma unroll
      for (int it = 0; it < B_ELEMS; ++it) {
        const int e = tid + it * kThreads;
        Bs[(e / kBK) * LDB_S + (e % kBK)] = rb[it];
      }
    }
  };

  auto compute = [&]() {
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[kr * LDA_S + wm0 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[(wn0 + j * 16 + (lane & 15)) * LDB_S + kr];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
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

template <int BM, int BN, int WM, int WN, typename Kfast, typename Kgen, typename T>
hipError_t dispatch(hipStream_t stream, const GemmArgs<T> &g, bool fast_layout, Kfast kfast,
                    Kgen kgen) {
  const int tiles_m = (g.M + BM - 1) / BM;
  const int tiles_n = (g.N + BN - 1) / BN;
  const bool fast = fast_layout && g.M % BM == 0 && g.N % BN == 0 && g.K % kBK == 0;
  const dim3 grid(tiles_m * tiles_n), block(kThreads);
  if (fast) hipLaunchKernelGGL(kfast, grid, block, 0, stream, g, tiles_m, tiles_n);
  else hipLaunchKernelGGL(kgen, grid, block, 0, stream, g, tiles_m, tiles_n);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm_d(hipStream_t stream, const GemmArgs<double> &g) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  const bool fl = (g.lda % 2 == 0) && (g.ldb % 2 == 0) && aligned16(g.A) && aligned16(g.B);
  if (g.M >= 1024 && g.N >= 1024)
    return dispatch<128, 128, 64, 64>(stream, g, fl, dgemm_mfma<128, 128, 64, 64, true>,
                                      dgemm_mfma<128, 128, 64, 64, false>);
  return dispatch<64, 64, 32, 32>(stream, g, fl, dgemm_mfma<64, 64, 32, 32, true>,
                                  dgemm_mfma<64, 64, 32, 32, false>);
}

hipError_t launch_gemm_z(hipStream_t stream, const GemmArgs<zd> &g) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (g.M >= 1024 && g.N >= 1024)
    return dispatch<128, 64, 64, 32>(stream, g, true, zgemm_mfma<128, 64, 64, 32, true>,
                                     zgemm_mfma<128, 64, 64, 32, false>);
  return dispatch<64, 64, 32, 32>(stream, g, true, zgemm_mfma<64, 64, 32, 32, true>,
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
