// gemm_split.hip -- FP64-in / FP64-out matrix products on BF16 MFMA with split
// operands and FP32 accumulation (the precision variant of the expm path).
//
// Each FP64 element x is rounded once to FP32, f = fl32(x), and split into
// bfloat16 terms f = h + m (+ l) + tiny: h = bf16(f), m = bf16(f - h),
// l = bf16(f - h - m). Each difference is exact in FP32 (h and m are f and
// f - h rounded to fewer bits), so two terms carry ~16 and three ~24 bits of f.
// The product keeps the cross terms of total order below the split:
//   SPLIT = 3:  Ah*Bh + Ah*Bm + Am*Bh                      (3 bf16 products)
//   SPLIT = 6:  the above + Ah*Bl + Am*Bm + Al*Bh           (6 bf16 products)
// all accumulated by v_mfma_f32_16x16x32_bf16 in FP32. bf16 * bf16 is exact in
// FP32, so the error is the FP32 accumulation plus the dropped terms: about
// 2^-16 per product for SPLIT = 3 and the FP32 level (~2^-24 sqrt(k)) for 6.
//
// Two ways in. The expm path splits each operand once into k-contiguous planes
// in HBM (split_planes_kernel: BF16; max_exponent_kernel + split_planes_half_kernel:
// IEEE half with per-row / per-column power-of-two scaling) and multiplies the
// planes (dgemm_split_planes_kg for the 3-product formats, dgemm_split_planes2 for
// the 6-product one). The stand-alone product (clm_gemm_split_d, dgemm_split_bf16)
// reads FP64 and splits in registers on the way into LDS. Every epilogue is the
// FP64 path's (alpha, beta*Cin, fused Paterson-Stockmeyer block), in FP64.
//
// dgemm_split_bf16 layout: a 128 x 128 tile of C per workgroup of 4 wave64, each
// wave 64 x 64 as 4 x 4 MFMA tiles of 16 x 16; K advances in slices of 32 (one
// MFMA k-step). LDS holds every plane k-contiguous -- As[m][k], Bs[n][k], rows of
// 40 bf16 (80 B) -- because the MFMA wants 8 consecutive k of one row in a lane;
// the 80-byte row makes the 16 lanes of a quarter-wave hit 16 disjoint 4-bank
// groups. A arrives m-contiguous from column-major memory and is transposed on
// the LDS write.
//
// MFMA operand layout (16x16x32 bf16 / f16): A lane l holds A[l%16][8(l/16) .. +7],
// B lane l holds B[8(l/16) .. +7][l%16], D lane l holds rows 4 (l/16) + r.
#include "clm_internal.hpp"
#include "gemm_tile.hpp"

namespace clm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
// 16 bytes of plane data in registers: a native vector, not HIP's uint4 struct,
// whose global -> register -> LDS copy the compiler turns into one memcpy at the
// LDS store -- issuing the next slice's loads only after this slice's MFMAs
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kSplitBK = 32;
constexpr int kSplitRow = 40;  // bf16 per LDS row (32 + 8 pad)

// row of D element r held by lane l (16x16 f32 output)
__device__ __forceinline__ int d_row(int lane, int r) { return 4 * (lane >> 4) + r; }

__device__ __forceinline__ unsigned short bf16_bits_rne(float f) {
  const unsigned u = __float_as_uint(f);
  const unsigned rounded = u + 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(rounded >> 16);
}
__device__ __forceinline__ float bf16_to_float(unsigned short b) {
  return __uint_as_float(static_cast<unsigned>(b) << 16);
}

// split x into NPL bf16 planes (bit patterns)
template <int NPL>
__device__ __forceinline__ void split_bf16(double x, unsigned short (&p)[NPL]) {
  float r = static_cast<float>(x);
#pragma unroll
  for (int t = 0; t < NPL; ++t) {
    p[t] = bf16_bits_rne(r);
    r = r - bf16_to_float(p[t]);
  }
}

__device__ __forceinline__ void tile_coords_s(int bid, int tiles_m, int tiles_n, int &tm, int &tn) {
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

__device__ __forceinline__ void epilogue_store_s(const GemmArgs<double> &g, int row, int col,
                                                 double acc) {
  double v = g.alpha * acc;
  if (g.Cin) v = v + g.beta * g.Cin[row + (long)col * g.ldcin];
  if (g.nterms > 0) {
    double e = (row == col) ? g.coef[0] : 0.0;
    if (g.nterms > 1) e = g.coef[1] * g.X1[row + (long)col * g.ldx1] + e;
    for (int t = 2; t < g.nterms; ++t)
      e = g.coef[t] * g.bank[(long)(t - 2) * g.bank_stride + row + (long)col * g.ldbank] + e;
    v = v + e;
  }
  g.C[row + (long)col * g.ldc] = v;
}

// planes needed per operand
template <int SPLIT>
struct SplitPlanes {
  static constexpr int NPL = (SPLIT == 3) ? 2 : 3;
};

template <int SPLIT, bool kFast>
__global__ void __launch_bounds__(256)
    dgemm_split_bf16(const GemmArgs<double> g, const int tiles_m, const int tiles_n) {
  constexpr int BM = 128, BN = 128, WM = 64, WN = 64, NT = 256;
  constexpr int NPL = SplitPlanes<SPLIT>::NPL;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int PLANE_A = BM * kSplitRow;  // bf16 per A plane
  constexpr int PLANE_B = BN * kSplitRow;
  // A: each lane stages 2 rows (m, m+1) x 8 k; B: 8 consecutive k of one column
  constexpr int A_ITERS = (BM * kSplitBK) / (NT * 16);  // = 1
  constexpr int B_ITERS = (BN * kSplitBK) / (NT * 8);   // = 2
  static_assert(A_ITERS == 1 && B_ITERS == 2, "staging shape");

  __shared__ __attribute__((aligned(16))) unsigned short As[NPL * PLANE_A];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[NPL * PLANE_B];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int tm, tn;
  tile_coords_s(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;

  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  // staging coordinates
  const int a_m = (tid % (BM / 2)) * 2;       // rows a_m, a_m + 1
  const int a_k = (tid / (BM / 2)) * 8;       // k a_k .. a_k + 7
  const int b_k = (tid % (kSplitBK / 8)) * 8;  // k b_k .. b_k + 7
  const int b_n = tid / (kSplitBK / 8);        // columns b_n, b_n + 64

  double ra[2][8];
  double rb[2][8];

  // global -> registers for the slice at k0
  auto gload = [&](const int k0) {
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const int gk = k0 + a_k + kk;
      if (kFast) {
        const double2 v = *reinterpret_cast<const double2 *>(g.A + (m0 + a_m) + (long)gk * g.lda);
        ra[0][kk] = v.x;
        ra[1][kk] = v.y;
      } else {
        const int gm = m0 + a_m;
        const bool kin = gk < g.K;
        ra[0][kk] = (kin && gm < g.M) ? g.A[gm + (long)gk * g.lda] : 0.0;
        ra[1][kk] = (kin && gm + 1 < g.M) ? g.A[gm + 1 + (long)gk * g.lda] : 0.0;
      }
    }
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int gn = n0 + b_n + it * 64;
#pragma unroll
      for (int kk = 0; kk < 8; kk += 2) {
        const int gk = k0 + b_k + kk;
        if (kFast) {
          const double2 v = *reinterpret_cast<const double2 *>(g.B + gk + (long)gn * g.ldb);
          rb[it][kk] = v.x;
          rb[it][kk + 1] = v.y;
        } else {
          const bool nin = gn < g.N;
          rb[it][kk] = (nin && gk < g.K) ? g.B[gk + (long)gn * g.ldb] : 0.0;
          rb[it][kk + 1] = (nin && gk + 1 < g.K) ? g.B[gk + 1 + (long)gn * g.ldb] : 0.0;
        }
      }
    }
  };

  const int KT = (g.K + kSplitBK - 1) / kSplitBK;
  if (KT > 0) gload(0);
  for (int kt = 0; kt < KT; ++kt) {
    __syncthreads();  // everyone is done reading the previous slice
    // ---- split and store: 8 consecutive k of one row = one 16-byte LDS write per plane ----
#pragma unroll
    for (int r2 = 0; r2 < 2; ++r2) {
      unsigned short pl[NPL][8];
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        unsigned short p[NPL];
        split_bf16<NPL>(ra[r2][kk], p);
#pragma unroll
        for (int t = 0; t < NPL; ++t) pl[t][kk] = p[t];
      }
#pragma unroll
      for (int t = 0; t < NPL; ++t) {
        uint4 w;
        w.x = pl[t][0] | (static_cast<unsigned>(pl[t][1]) << 16);
        w.y = pl[t][2] | (static_cast<unsigned>(pl[t][3]) << 16);
        w.z = pl[t][4] | (static_cast<unsigned>(pl[t][5]) << 16);
        w.w = pl[t][6] | (static_cast<unsigned>(pl[t][7]) << 16);
        *reinterpret_cast<uint4 *>(&As[t * PLANE_A + (a_m + r2) * kSplitRow + a_k]) = w;
      }
    }
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      unsigned short pl[NPL][8];
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        unsigned short p[NPL];
        split_bf16<NPL>(rb[it][kk], p);
#pragma unroll
        for (int t = 0; t < NPL; ++t) pl[t][kk] = p[t];
      }
#pragma unroll
      for (int t = 0; t < NPL; ++t) {
        uint4 w;
        w.x = pl[t][0] | (static_cast<unsigned>(pl[t][1]) << 16);
        w.y = pl[t][2] | (static_cast<unsigned>(pl[t][3]) << 16);
        w.z = pl[t][4] | (static_cast<unsigned>(pl[t][5]) << 16);
        w.w = pl[t][6] | (static_cast<unsigned>(pl[t][7]) << 16);
        *reinterpret_cast<uint4 *>(&Bs[t * PLANE_B + (b_n + it * 64) * kSplitRow + b_k]) = w;
      }
    }
    __syncthreads();
    // the next slice's loads fly while this slice is multiplied (the last
    // iteration re-reads its own slice instead of branching)
    gload((kt + 1 < KT ? kt + 1 : kt) * kSplitBK);
    // ---- MFMA: one k-step of 32 ----
    const int koff = 8 * (lane >> 4);
    bf16x8 a[NPL][TM], b[NPL][TN];
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
        a[t][i] = *reinterpret_cast<const bf16x8 *>(
            &As[t * PLANE_A + (wm0 + i * 16 + (lane & 15)) * kSplitRow + koff]);
#pragma unroll
      for (int j = 0; j < TN; ++j)
        b[t][j] = *reinterpret_cast<const bf16x8 *>(
            &Bs[t * PLANE_B + (wn0 + j * 16 + (lane & 15)) * kSplitRow + koff]);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        // smallest terms first, so the FP32 accumulator adds them before the big one
        if constexpr (SPLIT == 6) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[2][j], acc[i][j], 0, 0, 0);
        }
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[0][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[1][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
      }
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + d_row(lane, r);
        const int col = n0 + wn0 + j * 16 + (lane & 15);
        if (kFast || (row < g.M && col < g.N))
          epilogue_store_s(g, row, col, static_cast<double>(acc[i][j][r]));
      }
}

// ------------------------------------------------------------------------------------
// Split once, multiply many times. The plane kernels write the 16-bit planes of
// an FP64 matrix X (rows x cols, column-major, ldx) in the layouts the MFMA
// operands want -- k contiguous:
//   pc: column-major planes, pc[t*pstride + i + j*rows]  (X as the right operand)
//   pr: row-major planes,    pr[t*pstride + i*cols + j]  (X as the left operand)
// The row-major write goes through an LDS transpose so both the FP64 read and the
// plane writes are coalesced. dgemm_split_planes then stages the planes straight
// into LDS with 16-byte loads: its loop is loads, LDS and MFMA only.
//
// Tiling of the element-wise kernels: a 64 x 64 tile per 256-lane workgroup;
// lane t holds rows 4 (t % 16) .. +3 of the columns t / 16 + 16 q (q = 0..3), all
// 16 values loaded before any is used (two 16-byte reads per column when the rows
// allow), so a lane keeps its loads in flight together -- these passes are HBM
// bound, and a load-use-load chain per element would leave them latency bound.
// ------------------------------------------------------------------------------------
struct Tile64 {
  int i0, j0, r0, c0;  // the tile's origin; the lane's first row and first column
  bool full;           // the whole tile lies inside X, 16-byte aligned rows
  __device__ Tile64(int rows, int cols, const double *X, long ldx)
      : i0(blockIdx.x * 64), j0(blockIdx.y * 64), r0(i0 + 4 * (threadIdx.x & 15)),
        c0(j0 + (threadIdx.x >> 4)),
        full(i0 + 64 <= rows && j0 + 64 <= cols && (ldx & 1) == 0 &&
             (reinterpret_cast<uintptr_t>(X) & 15) == 0) {}
  __device__ int col(int q) const { return c0 + 16 * q; }
};

// The lane's 16 values, straight-line (edge tiles clamp the address and zero the
// value instead of branching, so no load waits on another).
template <bool kFull>
__device__ __forceinline__ void tile64_load(const Tile64 &tl, int rows, int cols,
                                            const double *__restrict__ X, long ldx,
                                            double (&v)[4][4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if constexpr (kFull) {
      const double *p = X + tl.r0 + static_cast<long>(tl.col(q)) * ldx;
      const double2 a = *reinterpret_cast<const double2 *>(p);
      const double2 b = *reinterpret_cast<const double2 *>(p + 2);
      v[q][0] = a.x;
      v[q][1] = a.y;
      v[q][2] = b.x;
      v[q][3] = b.y;
    } else {
      const int j = min(tl.col(q), cols - 1);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double x = X[min(tl.r0 + r, rows - 1) + static_cast<long>(j) * ldx];
        v[q][r] = (tl.col(q) < cols && tl.r0 + r < rows) ? x : 0.0;
      }
    }
  }
}

// whether a tile's column-major plane words go out as 8-byte stores
__device__ __forceinline__ bool col_vec_ok(int rows, long pstride, const unsigned short *pc) {
  return (rows & 3) == 0 && (pstride & 3) == 0 && (reinterpret_cast<uintptr_t>(pc) & 7) == 0;
}

// the four 16-bit words w[0..3] (rows r0 .. r0+3 of one column) into a column-major
// plane: one 8-byte store, or one store per row inside the matrix
__device__ __forceinline__ void store_col4(unsigned short *__restrict__ dst, bool vec, int nrows,
                                           const unsigned short (&w)[4]) {
  if (vec) {
    *reinterpret_cast<uint2 *>(dst) = make_uint2(w[0] | (static_cast<unsigned>(w[1]) << 16),
                                                 w[2] | (static_cast<unsigned>(w[3]) << 16));
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (r < nrows) dst[r] = w[r];
  }
}

// Row-major planes through LDS: every lane deposits its 16 values' plane words at
// [row][col] of the tile (planes 0, 1 packed in one 32-bit word, plane 2 in a
// second), then lane t writes row t / 4, columns 16 (t % 4) .. +15 of each plane:
// two 16-byte stores per plane when the rows are 16-byte aligned.
template <int NPL>
struct RowStage {
  static constexpr int W = (NPL + 1) / 2;  // 32-bit words per element
  unsigned int w[W][64][65];
};

template <int NPL>
__device__ __forceinline__ void row_stage_put(RowStage<NPL> &st, int rr, int cc,
                                              const unsigned short (&p)[NPL]) {
  unsigned int w0 = p[0];
  if constexpr (NPL > 1) w0 |= static_cast<unsigned>(p[1]) << 16;
  st.w[0][rr][cc] = w0;
  if constexpr (NPL > 2) st.w[1][rr][cc] = p[2];
}

template <int NPL>
__device__ __forceinline__ void row_stage_write(const RowStage<NPL> &st, const Tile64 &tl, int rows,
                                                int cols, unsigned short *__restrict__ pr,
                                                long pstride) {
  const int t = threadIdx.x, rr = t >> 2, cc = 16 * (t & 3);
  const int i = tl.i0 + rr, j = tl.j0 + cc;
  if (i >= rows || j >= cols) return;
  const bool vec = tl.full && (cols & 7) == 0 && (pstride & 7) == 0 &&
                   (reinterpret_cast<uintptr_t>(pr) & 15) == 0;
#pragma unroll
  for (int t2 = 0; t2 < NPL; ++t2) {
    unsigned short h[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const unsigned word = st.w[t2 / 2][rr][cc + c];
      h[c] = static_cast<unsigned short>((t2 & 1) ? word >> 16 : word);
    }
    unsigned short *dst = pr + t2 * pstride + static_cast<long>(i) * cols + j;
    if (vec) {
      uint4 lo, hi;
      lo.x = h[0] | (static_cast<unsigned>(h[1]) << 16);
      lo.y = h[2] | (static_cast<unsigned>(h[3]) << 16);
      lo.z = h[4] | (static_cast<unsigned>(h[5]) << 16);
      lo.w = h[6] | (static_cast<unsigned>(h[7]) << 16);
      hi.x = h[8] | (static_cast<unsigned>(h[9]) << 16);
      hi.y = h[10] | (static_cast<unsigned>(h[11]) << 16);
      hi.z = h[12] | (static_cast<unsigned>(h[13]) << 16);
      hi.w = h[14] | (static_cast<unsigned>(h[15]) << 16);
      reinterpret_cast<uint4 *>(dst)[0] = lo;
      reinterpret_cast<uint4 *>(dst)[1] = hi;
    } else {
#pragma unroll
      for (int c = 0; c < 16; ++c)
        if (j + c < cols) dst[c] = h[c];
    }
  }
}

template <int NPL, bool kFull>
__device__ __forceinline__ void split_planes_body(const Tile64 &tl, int rows, int cols,
                                                  const double *__restrict__ X, long ldx,
                                                  unsigned short *__restrict__ pc,
                                                  unsigned short *__restrict__ pr, long pstride,
                                                  RowStage<NPL> &st) {
  double v[4][4];
  tile64_load<kFull>(tl, rows, cols, X, ldx, v);
  const bool cvec = kFull && col_vec_ok(rows, pstride, pc);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned short p[4][NPL];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      split_bf16<NPL>(v[q][r], p[r]);
      if (pr) row_stage_put<NPL>(st, 4 * (threadIdx.x & 15) + r, (threadIdx.x >> 4) + 16 * q, p[r]);
    }
    const int j = tl.col(q);
    if (pc && j < cols) {
#pragma unroll
      for (int t = 0; t < NPL; ++t) {
        const unsigned short w[4] = {p[0][t], p[1][t], p[2][t], p[3][t]};
        store_col4(pc + t * pstride + tl.r0 + static_cast<long>(j) * rows, cvec, rows - tl.r0, w);
      }
    }
  }
  if (!pr) return;
  __syncthreads();
  row_stage_write<NPL>(st, tl, rows, cols, pr, pstride);
}

template <int NPL>
__global__ void __launch_bounds__(256)
    split_planes_kernel(int rows, int cols, const double *__restrict__ X, long ldx,
                        unsigned short *__restrict__ pc, unsigned short *__restrict__ pr,
                        long pstride) {
  __shared__ RowStage<NPL> st;
  const Tile64 tl(rows, cols, X, ldx);
  if (tl.full) split_planes_body<NPL, true>(tl, rows, cols, X, ldx, pc, pr, pstride, st);
  else split_planes_body<NPL, false>(tl, rows, cols, X, ldx, pc, pr, pstride, st);
}

// ---- IEEE-half planes with power-of-two scales -----------------------------------------
// The half format has 11 significant bits (bf16: 8) but a narrow exponent range,
// so each row of a left operand and each column of a right operand is scaled by
// a power of two that brings its largest magnitude into [2^13, 2^14) before the
// split: x' = 2^e x, h = half(fl32(x')), m = half(fl32(x') - h). Two terms carry
// ~22 bits of x, and the three products hh + hm + mh keep ~22 bits of the
// product (the dropped mm term is 2^-22 relative) -- close to the 6-product
// bfloat16 split at half its MFMA work. Products of scaled halves stay far below
// the FP32 range (2^28 per term); the epilogue divides the accumulator by
// 2^(e_row + e_col), exactly.
constexpr int kNoExponent = -(1 << 30);  // an all-zero row / column

// Exponent arrays hold tagged values, (gen << 12) | (e + 2048), gen = the
// generation of the split that wrote them (gen >= 1, counted per run of splits)
// and e clamped to [-2047, 2047] (an infinite maximum clamps to 2047, whose scale
// 2^-2034 still sends every finite entry to 0). atomicMax over tags is the max
// over e within a generation, and an entry an earlier generation left reads as
// an all-zero row's or column's -- so the arrays need one clearing per run of
// splits, not one before every split.
__device__ __forceinline__ int exp_of(int tagged, int gen) {
  return (tagged >> 12) == gen ? (tagged & 4095) - 2048 : kNoExponent;
}

// exponent of the largest |x| of each row (row_e) and column (col_e) of X, on the
// element-wise tiling: FP64 maxima (NaN ignored), reduced across the lanes that
// share a column (4 values halved to 1 over lane bits 0, 1, then bits 2, 3) or a
// row (bits 4, 5, then the four waves through LDS); one ilogb and one atomicMax
// per row and column and tile
template <bool kFull>
__device__ __forceinline__ void max_exponent_body(const Tile64 &tl, int rows, int cols,
                                                  const double *__restrict__ X, long ldx,
                                                  int *row_e, int *col_e, int gen,
                                                  double (&red)[4][64]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double v[4][4];
  tile64_load<kFull>(tl, rows, cols, X, ldx, v);
  double cm[4], rm[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    cm[q] = fmax(fmax(fabs(v[q][0]), fabs(v[q][1])), fmax(fabs(v[q][2]), fabs(v[q][3])));
#pragma unroll
  for (int r = 0; r < 4; ++r)
    rm[r] = fmax(fmax(fabs(v[0][r]), fabs(v[1][r])), fmax(fabs(v[2][r]), fabs(v[3][r])));
  // columns: after lane bit 0 a lane keeps columns q = 2 b0 + {0, 1}, after bit 1
  // the one column q = 2 b0 + b1; bits 2, 3 complete the 16-lane maximum
  {
    const bool b = lane & 1;
    const double s0 = b ? cm[0] : cm[2], s1 = b ? cm[1] : cm[3];
    const double k0 = b ? cm[2] : cm[0], k1 = b ? cm[3] : cm[1];
    cm[0] = fmax(k0, __shfl_xor(s0, 1, 64));
    cm[1] = fmax(k1, __shfl_xor(s1, 1, 64));
  }
  {
    const bool b = lane & 2;
    const double s = b ? cm[0] : cm[1], k = b ? cm[1] : cm[0];
    cm[0] = fmax(k, __shfl_xor(s, 2, 64));
  }
  cm[0] = fmax(cm[0], __shfl_xor(cm[0], 4, 64));
  cm[0] = fmax(cm[0], __shfl_xor(cm[0], 8, 64));
  if (col_e && (lane & 12) == 0) {
    const int j = tl.col(2 * (lane & 1) + ((lane >> 1) & 1));
    if (j < cols && cm[0] > 0.0) atomicMax(&col_e[j], tile::exp_tag(gen, ilogb(cm[0])));
  }
  // rows: lanes l, l ^ 16, l ^ 32, l ^ 48 share rows; after bits 4, 5 a lane keeps
  // row 2 b4 + b5 of its four
  {
    const bool b = lane & 16;
    const double s0 = b ? rm[0] : rm[2], s1 = b ? rm[1] : rm[3];
    const double k0 = b ? rm[2] : rm[0], k1 = b ? rm[3] : rm[1];
    rm[0] = fmax(k0, __shfl_xor(s0, 16, 64));
    rm[1] = fmax(k1, __shfl_xor(s1, 16, 64));
  }
  {
    const bool b = lane & 32;
    const double s = b ? rm[0] : rm[1], k = b ? rm[1] : rm[0];
    rm[0] = fmax(k, __shfl_xor(s, 32, 64));
  }
  red[wave][4 * (lane & 15) + 2 * ((lane >> 4) & 1) + ((lane >> 5) & 1)] = rm[0];
  __syncthreads();
  if (row_e && wave == 0) {
    const double m = fmax(fmax(red[0][lane], red[1][lane]), fmax(red[2][lane], red[3][lane]));
    const int i = tl.i0 + lane;
    if (i < rows && m > 0.0) atomicMax(&row_e[i], tile::exp_tag(gen, ilogb(m)));
  }
}

// The Paterson-Stockmeyer block B = c0 I + c1 X1 + sum_t c_t bank[t-2] (each entry as
// build_block_kernel forms it, the same multiply-adds in the same order) on the
// element-wise tiling, with the exponent of each row's largest |B| found on the way
// (max_exponent_body's row half): the block's first split as a left operand then needs
// no exponent pass of its own.
template <bool kFull>
__device__ __forceinline__ void build_block_rows_body(const Tile64 &tl, int n, double *Bout,
                                                      long ldb, const double *coef, int nterms,
                                                      const double *X1, long ldx1,
                                                      const double *bank, long stride,
                                                      long ldbank, int *row_e, int gen,
                                                      double (&red)[4][64]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double v[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) v[q][r] = (tl.r0 + r == tl.col(q)) ? coef[0] : 0.0;
  if (nterms > 1) {
    double x[4][4];
    tile64_load<kFull>(tl, n, n, X1, ldx1, x);
    const double c = coef[1];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) v[q][r] = ops<double>::mad(c, x[q][r], v[q][r]);
  }
  for (int t = 2; t < nterms; ++t) {
    double x[4][4];
    tile64_load<kFull>(tl, n, n, bank + static_cast<long>(t - 2) * stride, ldbank, x);
    const double c = coef[t];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) v[q][r] = ops<double>::mad(c, x[q][r], v[q][r]);
  }
  double rm[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = tl.col(q);
    if constexpr (kFull) {
      double *p = Bout + tl.r0 + static_cast<long>(j) * ldb;
      *reinterpret_cast<double2 *>(p) = double2{v[q][0], v[q][1]};
      *reinterpret_cast<double2 *>(p + 2) = double2{v[q][2], v[q][3]};
#pragma unroll
      for (int r = 0; r < 4; ++r) rm[r] = fmax(rm[r], fabs(v[q][r]));
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (j < n && tl.r0 + r < n) {
          Bout[tl.r0 + r + static_cast<long>(j) * ldb] = v[q][r];
          rm[r] = fmax(rm[r], fabs(v[q][r]));
        }
    }
  }
  // lanes l, l ^ 16, l ^ 32, l ^ 48 share rows; after bits 4, 5 a lane keeps row
  // 2 b4 + b5 of its four, then the four waves meet in LDS
  {
    const bool b = lane & 16;
    const double s0 = b ? rm[0] : rm[2], s1 = b ? rm[1] : rm[3];
    const double k0 = b ? rm[2] : rm[0], k1 = b ? rm[3] : rm[1];
    rm[0] = fmax(k0, __shfl_xor(s0, 16, 64));
    rm[1] = fmax(k1, __shfl_xor(s1, 16, 64));
  }
  {
    const bool b = lane & 32;
    const double s = b ? rm[0] : rm[1], k = b ? rm[1] : rm[0];
    rm[0] = fmax(k, __shfl_xor(s, 32, 64));
  }
  red[wave][4 * (lane & 15) + 2 * ((lane >> 4) & 1) + ((lane >> 5) & 1)] = rm[0];
  __syncthreads();
  if (wave == 0) {
    const double m = fmax(fmax(red[0][lane], red[1][lane]), fmax(red[2][lane], red[3][lane]));
    const int i = tl.i0 + lane;
    if (i < n && m > 0.0) atomicMax(&row_e[i], tile::exp_tag(gen, ilogb(m)));
  }
}

__global__ void __launch_bounds__(256)
    build_block_rows_kernel(int n, double *__restrict__ Bout, long ldb,
                            const double *__restrict__ coef, int nterms,
                            const double *__restrict__ X1, long ldx1,
                            const double *__restrict__ bank, long stride, long ldbank,
                            int *row_e, int gen, bool sources_aligned) {
  __shared__ double red[4][64];
  const Tile64 tl(n, n, Bout, ldb);
  if (tl.full && sources_aligned)
    build_block_rows_body<true>(tl, n, Bout, ldb, coef, nterms, X1, ldx1, bank, stride, ldbank,
                                row_e, gen, red);
  else
    build_block_rows_body<false>(tl, n, Bout, ldb, coef, nterms, X1, ldx1, bank, stride, ldbank,
                                 row_e, gen, red);
}

__global__ void __launch_bounds__(256)
    max_exponent_kernel(int rows, int cols, const double *__restrict__ X, long ldx, int *row_e,
                        int *col_e, int gen) {
  __shared__ double red[4][64];
  const Tile64 tl(rows, cols, X, ldx);
  if (tl.full) max_exponent_body<true>(tl, rows, cols, X, ldx, row_e, col_e, gen, red);
  else max_exponent_body<false>(tl, rows, cols, X, ldx, row_e, col_e, gen, red);
}

// the scale exponent of a row / column whose largest |x| has exponent e
__device__ __forceinline__ int half_scale(int e) { return e == kNoExponent ? 0 : 13 - e; }

__device__ __forceinline__ unsigned short half_bits(float f) {
  const _Float16 h = static_cast<_Float16>(f);
  return __builtin_bit_cast(unsigned short, h);
}
__device__ __forceinline__ float half_value(unsigned short b) {
  return static_cast<float>(__builtin_bit_cast(_Float16, b));
}

__device__ __forceinline__ void split_half(double x, int scale, unsigned short (&p)[2]) {
  const float f = static_cast<float>(ldexp(x, scale));
  p[0] = half_bits(f);
  p[1] = half_bits(f - half_value(p[0]));
}

// half planes of X: pc column-major scaled by col_e (right operand), pr row-major
// scaled by row_e (left operand); the element-wise tiling and row staging above
template <bool kFull>
__device__ __forceinline__ void split_half_body(const Tile64 &tl, int rows, int cols,
                                                const double *__restrict__ X, long ldx,
                                                unsigned short *__restrict__ pc,
                                                unsigned short *__restrict__ pr, long pstride,
                                                const int *__restrict__ row_e,
                                                const int *__restrict__ col_e, int row_gen,
                                                int col_gen, RowStage<2> &st) {
  double v[4][4];
  tile64_load<kFull>(tl, rows, cols, X, ldx, v);
  // the exponents, straight-line like the values, before the first store
  int rs[4] = {0, 0, 0, 0}, cs[4] = {0, 0, 0, 0};
  if (pr) {
#pragma unroll
    for (int r = 0; r < 4; ++r) rs[r] = half_scale(exp_of(row_e[min(tl.r0 + r, rows - 1)], row_gen));
  }
  if (pc) {
#pragma unroll
    for (int q = 0; q < 4; ++q) cs[q] = half_scale(exp_of(col_e[min(tl.col(q), cols - 1)], col_gen));
  }
  const bool cvec = kFull && col_vec_ok(rows, pstride, pc);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = tl.col(q);
    if (pc && j < cols) {
      unsigned short w0[4], w1[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        unsigned short p[2];
        split_half(v[q][r], cs[q], p);
        w0[r] = p[0];
        w1[r] = p[1];
      }
      unsigned short *dst = pc + tl.r0 + static_cast<long>(j) * rows;
      store_col4(dst, cvec, rows - tl.r0, w0);
      store_col4(dst + pstride, cvec, rows - tl.r0, w1);
    }
    if (pr) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        unsigned short p[2];
        split_half(v[q][r], rs[r], p);
        row_stage_put<2>(st, 4 * (threadIdx.x & 15) + r, (threadIdx.x >> 4) + 16 * q, p);
      }
    }
  }
  if (!pr) return;
  __syncthreads();
  row_stage_write<2>(st, tl, rows, cols, pr, pstride);
}

__global__ void __launch_bounds__(256)
    split_planes_half_kernel(int rows, int cols, const double *__restrict__ X, long ldx,
                             unsigned short *__restrict__ pc, unsigned short *__restrict__ pr,
                             long pstride, const int *__restrict__ row_e,
                             const int *__restrict__ col_e, int row_gen, int col_gen) {
  __shared__ RowStage<2> st;
  const Tile64 tl(rows, cols, X, ldx);
  if (tl.full)
    split_half_body<true>(tl, rows, cols, X, ldx, pc, pr, pstride, row_e, col_e, row_gen, col_gen,
                          st);
  else
    split_half_body<false>(tl, rows, cols, X, ldx, pc, pr, pstride, row_e, col_e, row_gen, col_gen,
                           st);
}

__global__ void clear_exponents_kernel(int *a, int *b, int *c, int *d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    if (a) a[i] = 0;
    if (b) b[i] = 0;
    if (c) c[i] = 0;
    if (d) d[i] = 0;
  }
}
// the 16x16x32 MFMA of a plane format: bf16 or (FMT = 1) IEEE half
template <int FMT>
struct PlaneMfma;
template <>
struct PlaneMfma<0> {
  using vec = bf16x8;
  __device__ static f4 mma(vec a, vec b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct PlaneMfma<1> {
  using vec = f16x8;
  __device__ static f4 mma(vec a, vec b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

// ------------------------------------------------------------------------------------
// The plane product with more waves and deeper K slices. A 128 x 128 tile of C
// is shared by NW waves (NW = 8: a 2 x 4 grid of 64 x 32 wave tiles; NW = 4:
// 2 x 2 of 64 x 64) and each LDS stage holds KS MFMA k-steps (32 KS values of
// k). At n = 2048 there are exactly as many tiles as CUs, so the waves of one
// workgroup are all a CU has to hide the load, LDS and MFMA latencies with:
// eight waves give every SIMD two, and a 64-deep slice halves the barriers
// (measured at n = 2048 against a 4-wave, 32-deep version: split 3 86 -> 68 us,
// split 6 124 -> 115 us; a second register set for a two-slice prefetch
// spilled to scratch and ran 2.7x slower). The k-steps run in ascending order and the products of a
// k-step in a fixed order (smallest terms first).
// ------------------------------------------------------------------------------------
// LDS layout of a plane tile (128 rows of BK k-values, 16-bit). 64-deep slices
// (KS = 2) use unpadded 128-byte rows with the row's eight 16-byte chunks permuted
// by an XOR with the row index mod 8: the b128 fragment reads (16 lanes on 16
// rows, one chunk each) and the staging writes (8 lanes on one row's 8 chunks) are
// both free of bank conflicts, measured with SQ_LDS_BANK_CONFLICT on MI355X -- the
// 144-byte padded rows took 1.1 extra cycles per fragment read (5.1 LDS cycles per
// b128 read against 4.0, the 256 B/clk rate). 32-deep slices keep a 16-byte pad.
template <int NPL, int KS>
struct PlaneTile {
  static constexpr int BK = 32 * KS;
  static constexpr bool kSwizzle = (KS == 2);
  static constexpr int ROW = kSwizzle ? BK : BK + 8;
  static constexpr int PLANE = 128 * ROW;
  static constexpr int STAGE = 2 * NPL * PLANE;  // A and B planes
  // element offset of (row, k) in a plane, k a multiple of 8 (a 16-byte chunk)
  __device__ static constexpr int off(int row, int k) {
    if constexpr (kSwizzle) return row * ROW + (((k >> 3) ^ (row & 7)) << 3);
    else return row * ROW + k;
  }
};

template <int NPL, int KS, int NT, bool kFast>
__device__ __forceinline__ void planes2_gload(const GemmArgs<double> &g, const unsigned short *PA,
                                              const unsigned short *PB, long pstride, int m0,
                                              int n0, int k0, int tid,
                                              u32x4 (&ra)[NPL][128 * 4 * KS / NT],
                                              u32x4 (&rb)[NPL][128 * 4 * KS / NT]) {
  constexpr int CHUNKS = 4 * KS;  // 16-byte chunks per row per slice
  constexpr int IT = 128 * CHUNKS / NT;
  const long K = g.K;
#pragma unroll
  for (int t = 0; t < NPL; ++t)
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int c = tid + it * NT;
      const int r = c / CHUNKS, kq = (c % CHUNKS) * 8;
      const int gk = k0 + kq;
      if (kFast) {
        ra[t][it] = *reinterpret_cast<const u32x4 *>(PA + t * pstride + (long)(m0 + r) * K + gk);
        rb[t][it] = *reinterpret_cast<const u32x4 *>(PB + t * pstride + (long)(n0 + r) * K + gk);
      } else {
        unsigned va[8], vb[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const bool kin = gk + e < K;
          va[e] = (kin && m0 + r < g.M) ? PA[t * pstride + (long)(m0 + r) * K + gk + e] : 0u;
          vb[e] = (kin && n0 + r < g.N) ? PB[t * pstride + (long)(n0 + r) * K + gk + e] : 0u;
        }
        ra[t][it] = u32x4{va[0] | (va[1] << 16), va[2] | (va[3] << 16), va[4] | (va[5] << 16),
                          va[6] | (va[7] << 16)};
        rb[t][it] = u32x4{vb[0] | (vb[1] << 16), vb[2] | (vb[3] << 16), vb[4] | (vb[5] << 16),
                          vb[6] | (vb[7] << 16)};
      }
    }
}

template <int NPL, int KS, int NT>
__device__ __forceinline__ void planes2_sstore(unsigned short *S, int tid,
                                               const u32x4 (&ra)[NPL][128 * 4 * KS / NT],
                                               const u32x4 (&rb)[NPL][128 * 4 * KS / NT]) {
  using PT = PlaneTile<NPL, KS>;
  constexpr int CHUNKS = 4 * KS;
  constexpr int IT = 128 * CHUNKS / NT;
#pragma unroll
  for (int t = 0; t < NPL; ++t)
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int c = tid + it * NT;
      const int r = c / CHUNKS, kq = (c % CHUNKS) * 8;
      *reinterpret_cast<u32x4 *>(&S[t * PT::PLANE + PT::off(r, kq)]) = ra[t][it];
      *reinterpret_cast<u32x4 *>(&S[(NPL + t) * PT::PLANE + PT::off(r, kq)]) = rb[t][it];
    }
}

template <int SPLIT, int FMT, int KS, int TM, int TN>
__device__ __forceinline__ void planes2_compute(const unsigned short *S, int lane, int wm0, int wn0,
                                                f4 (&acc)[TM][TN]) {
  constexpr int NPL = SplitPlanes<SPLIT>::NPL;
  using PT = PlaneTile<NPL, KS>;
  using V = typename PlaneMfma<FMT>::vec;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int koff = ks * 32 + 8 * (lane >> 4);
    V a[NPL][TM], b[NPL][TN];
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
        a[t][i] = *reinterpret_cast<const V *>(
            &S[t * PT::PLANE + PT::off(wm0 + i * 16 + (lane & 15), koff)]);
#pragma unroll
      for (int j = 0; j < TN; ++j)
        b[t][j] = *reinterpret_cast<const V *>(
            &S[(NPL + t) * PT::PLANE + PT::off(wn0 + j * 16 + (lane & 15), koff)]);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        if constexpr (SPLIT == 6) {
          acc[i][j] = PlaneMfma<FMT>::mma(a[2][i], b[0][j], acc[i][j]);
          acc[i][j] = PlaneMfma<FMT>::mma(a[1][i], b[1][j], acc[i][j]);
          acc[i][j] = PlaneMfma<FMT>::mma(a[0][i], b[2][j], acc[i][j]);
        }
        acc[i][j] = PlaneMfma<FMT>::mma(a[1][i], b[0][j], acc[i][j]);
        acc[i][j] = PlaneMfma<FMT>::mma(a[0][i], b[1][j], acc[i][j]);
        acc[i][j] = PlaneMfma<FMT>::mma(a[0][i], b[0][j], acc[i][j]);
      }
  }
}

// Epilogue of a 128 x 128 plane-product tile: the FP32 sums go down column-major
// in LDS (sum[c * kSumLd + r], one 16-byte write per lane and MFMA tile), then
// tile::lds_tile_epilogue walks them in coalesced column pieces.
constexpr int kSumLd = 132;  // FP32 per LDS column of the sum tile (128 + 4 pad)

template <int TM, int TN>
__device__ __forceinline__ void sum_tile_put(float *sum, int lane, int wm0, int wn0,
                                             const f4 (&acc)[TM][TN]) {
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
      *reinterpret_cast<f4 *>(&sum[(wn0 + j * 16 + (lane & 15)) * kSumLd + wm0 + i * 16 +
                                   d_row(lane, 0)]) = acc[i][j];
}

// the tile's FP32 sums from LDS, the half format's 2^-(e_row + e_col) undone
template <int FMT, int NW, bool kFast>
__device__ __forceinline__ void tile_epilogue(const GemmArgs<double> &g, const float *sum, int m0,
                                              int n0, const int *__restrict__ row_e,
                                              const int *__restrict__ col_e, int row_gen,
                                              int col_gen, int *row_tags = nullptr) {
  tile::lds_tile_epilogue<double, NW, kFast>(
      g, m0, n0,
      [&](int c, int r) {
        double a = static_cast<double>(sum[c * kSumLd + r]);
        if constexpr (FMT == 1)
          a = ldexp(a, -(half_scale(exp_of(row_e[m0 + r], row_gen)) +
                         half_scale(exp_of(col_e[n0 + c], col_gen))));
        return a;
      },
      row_tags);
}

template <int SPLIT, int FMT, int NW, int KS, bool kFast>
__global__ void __launch_bounds__(NW * 64)
    dgemm_split_planes2(const GemmArgs<double> g, const unsigned short *__restrict__ PA,
                        const unsigned short *__restrict__ PB, const long pstride,
                        const int tiles_m, const int tiles_n, const int *__restrict__ row_e,
                        const int *__restrict__ col_e, const int row_gen, const int col_gen) {
  constexpr int NT = NW * 64;
  constexpr int NPL = SplitPlanes<SPLIT>::NPL;
  using PT = PlaneTile<NPL, KS>;
  constexpr int WAVES_N = NW == 8 ? 4 : 2;
  constexpr int WM = 64, WN = 128 / WAVES_N;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int IT = 128 * 4 * KS / NT;
  static_assert(IT >= 1 && (128 * 4 * KS) % NT == 0, "staging shape");

  __shared__ __attribute__((aligned(16))) unsigned short L[2 * PT::STAGE];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int tm, tn;
  tile_coords_s(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * 128, n0 = tn * 128;
  const int wm0 = (wave / WAVES_N) * WM;
  const int wn0 = (wave % WAVES_N) * WN;

  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  u32x4 ra[NPL][IT], rb[NPL][IT];
  const int KT = (g.K + PT::BK - 1) / PT::BK;
  if (KT > 0) {
    planes2_gload<NPL, KS, NT, kFast>(g, PA, PB, pstride, m0, n0, 0, tid, ra, rb);
    planes2_sstore<NPL, KS, NT>(L, tid, ra, rb);
    __syncthreads();
    for (int kt = 0; kt < KT - 1; ++kt) {
      const int cur = kt & 1;
      planes2_gload<NPL, KS, NT, kFast>(g, PA, PB, pstride, m0, n0, (kt + 1) * PT::BK, tid, ra, rb);
      // the next slice's loads go out before this slice's MFMAs, whose time covers
      // their latency (the scheduler would otherwise sink them next to the LDS stores)
      __builtin_amdgcn_sched_barrier(0);
      planes2_compute<SPLIT, FMT, KS, TM, TN>(L + cur * PT::STAGE, lane, wm0, wn0, acc);
      planes2_sstore<NPL, KS, NT>(L + (cur ^ 1) * PT::STAGE, tid, ra, rb);
      __syncthreads();
    }
    planes2_compute<SPLIT, FMT, KS, TM, TN>(L + ((KT - 1) & 1) * PT::STAGE, lane, wm0, wn0, acc);
  }

  static_assert(128 * kSumLd * 4 <= 2 * PT::STAGE * 2, "the sum tile fits in the stages");
  float *sum = reinterpret_cast<float *>(L);
  __syncthreads();  // every wave is done with the stages
  sum_tile_put<TM, TN>(sum, lane, wm0, wn0, acc);
  __syncthreads();
  tile_epilogue<FMT, NW, kFast>(g, sum, m0, n0, row_e, col_e, row_gen, col_gen);
}

// ------------------------------------------------------------------------------------
// The plane product with k-grouped waves (the 3-product formats). Eight waves
// share a 128 x 128 tile as two groups of four; each group covers the whole tile
// as 2 x 2 wave tiles of 64 x 64 and takes one of the two 32-deep k-steps of
// every 64-deep LDS slice (group 0 the even steps, group 1 the odd). A 64 x 64
// wave tile reads 64 + 64 rows of each plane per k-step for 16 MFMA tiles,
// against 64 + 32 rows for 8 tiles at 64 x 32 -- a third less LDS reading per
// MFMA -- and a CU still runs two waves per SIMD. At the end the groups' FP32
// sums meet in LDS: group 1 writes its accumulators, group 0 adds them (even-step
// sum + odd-step sum) and runs the epilogue.
// ------------------------------------------------------------------------------------
// a wave's fragments of one 32-deep k-step: rows wm0.. of the A planes, wn0.. of the B planes
template <int NPL, int FMT>
__device__ __forceinline__ void planes_kg_read(const unsigned short *S, int lane, int wm0, int wn0,
                                               int ks, typename PlaneMfma<FMT>::vec (&a)[NPL][4],
                                               typename PlaneMfma<FMT>::vec (&b)[NPL][4]) {
  using PT = PlaneTile<NPL, 2>;
  using V = typename PlaneMfma<FMT>::vec;
  const int koff = ks * 32 + 8 * (lane >> 4);
#pragma unroll
  for (int t = 0; t < NPL; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[t][i] = *reinterpret_cast<const V *>(
          &S[t * PT::PLANE + PT::off(wm0 + i * 16 + (lane & 15), koff)]);
      b[t][i] = *reinterpret_cast<const V *>(
          &S[(NPL + t) * PT::PLANE + PT::off(wn0 + i * 16 + (lane & 15), koff)]);
    }
}

// the first two products of a k-step (the small terms a1 b0 and a0 b1), every tile
template <int FMT>
__device__ __forceinline__ void planes_kg_mma_small(const typename PlaneMfma<FMT>::vec (&a)[2][4],
                                                    const typename PlaneMfma<FMT>::vec (&b)[2][4],
                                                    f4 (&acc)[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[i][j] = PlaneMfma<FMT>::mma(a[1][i], b[0][j], acc[i][j]);
      acc[i][j] = PlaneMfma<FMT>::mma(a[0][i], b[1][j], acc[i][j]);
    }
}

// the last product of a k-step (a0 b0), every tile
template <int FMT>
__device__ __forceinline__ void planes_kg_mma_lead(const typename PlaneMfma<FMT>::vec (&a)[2][4],
                                                   const typename PlaneMfma<FMT>::vec (&b)[2][4],
                                                   f4 (&acc)[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = PlaneMfma<FMT>::mma(a[0][i], b[0][j], acc[i][j]);
}

template <int SPLIT, int FMT, bool kFast>
__global__ void __launch_bounds__(512)
    dgemm_split_planes_kg(const GemmArgs<double> g, const unsigned short *__restrict__ PA,
                          const unsigned short *__restrict__ PB, const long pstride,
                          const int tiles_m, const int tiles_n, const int *__restrict__ row_e,
                          const int *__restrict__ col_e, const int row_gen, const int col_gen) {
  static_assert(SPLIT == 3, "the k-grouped kernel runs the 3-product formats");
  constexpr int NT = 512, KS = 2;
  constexpr int NPL = SplitPlanes<SPLIT>::NPL;
  using PT = PlaneTile<NPL, KS>;
  constexpr int IT = 128 * 4 * KS / NT;
  static_assert(IT >= 1 && (128 * 4 * KS) % NT == 0, "staging shape");
  // the reduction buffer: 4 waves x 16 tiles x 4 rows x 64 lanes of FP32
  // the stages, or the reduction buffer and the column-major sum tile if larger
  constexpr int kRedFloats = 4 * 16 * 4 * 64 + 128 * kSumLd;
  constexpr int kLdsShorts = (2 * PT::STAGE > 2 * kRedFloats) ? 2 * PT::STAGE : 2 * kRedFloats;

  __shared__ __attribute__((aligned(16))) unsigned short L[kLdsShorts];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int grp = wave >> 2, wpos = wave & 3;
  int tm, tn;
  tile_coords_s(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * 128, n0 = tn * 128;
  const int wm0 = (wpos >> 1) * 64, wn0 = (wpos & 1) * 64;

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  // Software pipeline over the 64-deep slices: a wave holds slice kt's fragments in
  // registers; it runs the two small products of every tile, stages slice kt + 1
  // (loaded from memory one slice earlier) and issues the loads of slice kt + 2,
  // meets the barrier, then issues its LDS reads of slice kt + 1 and runs the lead
  // products of slice kt behind them, so the read latency hides under MFMAs. The
  // stage written at slice kt held slice kt - 1, whose fragments every wave had in
  // registers before the previous barrier (its small products read all of them).
  // Every accumulator still takes a1 b0, a0 b1, a0 b0 of each k-step in k order.
  using V = typename PlaneMfma<FMT>::vec;
  u32x4 ra[NPL][IT], rb[NPL][IT];
  const int KT = (g.K + PT::BK - 1) / PT::BK;
  if (KT > 0) {
    planes2_gload<NPL, KS, NT, kFast>(g, PA, PB, pstride, m0, n0, 0, tid, ra, rb);
    planes2_sstore<NPL, KS, NT>(L, tid, ra, rb);
    if (KT > 1) planes2_gload<NPL, KS, NT, kFast>(g, PA, PB, pstride, m0, n0, PT::BK, tid, ra, rb);
    __syncthreads();
    V a[NPL][4], b[NPL][4];
    planes_kg_read<NPL, FMT>(L, lane, wm0, wn0, grp, a, b);
    for (int kt = 0; kt < KT - 1; ++kt) {
      const int nxt = (kt + 1) & 1;
      planes_kg_mma_small<FMT>(a, b, acc);
      planes2_sstore<NPL, KS, NT>(L + nxt * PT::STAGE, tid, ra, rb);
      // the LDS stores spread over the small products (four MFMAs, one store)
#pragma unroll
      for (int q = 0; q < 2 * NPL * IT; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
        __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
      }
      if (kt + 2 < KT)
        planes2_gload<NPL, KS, NT, kFast>(g, PA, PB, pstride, m0, n0, (kt + 2) * PT::BK, tid, ra, rb);
      __syncthreads();
      V na[NPL][4], nb[NPL][4];
      planes_kg_read<NPL, FMT>(L + nxt * PT::STAGE, lane, wm0, wn0, grp, na, nb);
      __builtin_amdgcn_sched_barrier(0);
      planes_kg_mma_lead<FMT>(a, b, acc);
#pragma unroll
      for (int t = 0; t < NPL; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[t][i] = na[t][i];
          b[t][i] = nb[t][i];
        }
    }
    planes_kg_mma_small<FMT>(a, b, acc);
    planes_kg_mma_lead<FMT>(a, b, acc);
  }

  __syncthreads();  // every wave is done with the stages
  // group 1's sums, then (group 0) the whole tile's, column-major for the epilogue
  float *red = reinterpret_cast<float *>(L);
  float *sum = red + 4 * 16 * 4 * 64;
  static_assert(kRedFloats * 4 <= kLdsShorts * 2, "the reduction buffer and the sum tile fit");
  if (grp == 1) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[((wpos * 16 + i * 4 + j) * 4 + r) * 64 + lane] = acc[i][j][r];
  }
  __syncthreads();
  if (grp == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][r] += red[((wpos * 16 + i * 4 + j) * 4 + r) * 64 + lane];
    sum_tile_put<4, 4>(sum, lane, wm0, wn0, acc);
  }
  __syncthreads();
  // the output rows' exponents, when asked for, gather in the (now free) reduction buffer
  int *row_tags = nullptr;
  if (g.row_e_out) {
    row_tags = reinterpret_cast<int *>(red);
    if (tid < 128) row_tags[tid] = 0;
    __syncthreads();
  }
  tile_epilogue<FMT, 8, kFast>(g, sum, m0, n0, row_e, col_e, row_gen, col_gen, row_tags);
}

inline bool aligned16s(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <int SPLIT>
hipError_t launch_split(hipStream_t stream, const GemmArgs<double> &g) {
  const int tiles_m = (g.M + 127) / 128, tiles_n = (g.N + 127) / 128;
  const bool fast = g.M % 128 == 0 && g.N % 128 == 0 && g.K % kSplitBK == 0 && g.lda % 2 == 0 &&
                    g.ldb % 2 == 0 && aligned16s(g.A) && aligned16s(g.B);
  const dim3 grid(tiles_m * tiles_n), block(256);
  if (fast) hipLaunchKernelGGL((dgemm_split_bf16<SPLIT, true>), grid, block, 0, stream, g, tiles_m, tiles_n);
  else hipLaunchKernelGGL((dgemm_split_bf16<SPLIT, false>), grid, block, 0, stream, g, tiles_m, tiles_n);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm_split_bf16(hipStream_t stream, const GemmArgs<double> &g, int split) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  return split == 6 ? launch_split<6>(stream, g) : launch_split<3>(stream, g);
}

hipError_t launch_split_planes(hipStream_t stream, int split, int rows, int cols, const double *X,
                               long ldx, unsigned short *pc, unsigned short *pr, long pstride) {
  if (rows <= 0 || cols <= 0 || (!pc && !pr)) return hipSuccess;
  const dim3 grid((rows + 63) / 64, (cols + 63) / 64), block(256);
  if (split == 6)
    hipLaunchKernelGGL(split_planes_kernel<3>, grid, block, 0, stream, rows, cols, X, ldx, pc, pr,
                       pstride);
  else
    hipLaunchKernelGGL(split_planes_kernel<2>, grid, block, 0, stream, rows, cols, X, ldx, pc, pr,
                       pstride);
  return hipGetLastError();
}

hipError_t launch_gemm_split_planes(hipStream_t stream, const GemmArgs<double> &g, int split,
                                    const unsigned short *PA, const unsigned short *PB,
                                    long pstride, const int *row_e, const int *col_e, int row_gen,
                                    int col_gen) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  // the output rows' exponents come from the 3-product kernel's epilogue only
  if (g.row_e_out && split == CLM_SPLIT_BF16X6) return hipErrorInvalidValue;
  const int tiles_m = (g.M + 127) / 128, tiles_n = (g.N + 127) / 128;
  // 3-product formats: 8 waves, 64-deep slices (two LDS stages of two planes);
  // 6 products: 8 waves, 32-deep slices (three planes)
  const int bk = (split == CLM_SPLIT_BF16X6) ? 32 : 64;
  const bool fast = g.M % 128 == 0 && g.N % 128 == 0 && g.K % bk == 0 && aligned16s(PA) &&
                    aligned16s(PB) && (pstride % 8) == 0;
  const dim3 grid(tiles_m * tiles_n), block(512);
#define CLM_SPLIT_LAUNCH(S, FMT, KS, F)                                                       \
  hipLaunchKernelGGL((dgemm_split_planes2<S, FMT, 8, KS, F>), grid, block, 0, stream, g, PA, PB, \
                     pstride, tiles_m, tiles_n, row_e, col_e, row_gen, col_gen)
#define CLM_SPLIT_LAUNCH_KG(FMT, F)                                                          \
  hipLaunchKernelGGL((dgemm_split_planes_kg<3, FMT, F>), grid, block, 0, stream, g, PA, PB,    \
                     pstride, tiles_m, tiles_n, row_e, col_e, row_gen, col_gen)
  if (split == CLM_SPLIT_FP16X3) {
    if (fast) CLM_SPLIT_LAUNCH_KG(1, true);
    else CLM_SPLIT_LAUNCH_KG(1, false);
  } else if (split == CLM_SPLIT_BF16X6) {
    if (fast) CLM_SPLIT_LAUNCH(6, 0, 1, true);
    else CLM_SPLIT_LAUNCH(6, 0, 1, false);
  } else {
    if (fast) CLM_SPLIT_LAUNCH_KG(0, true);
    else CLM_SPLIT_LAUNCH_KG(0, false);
  }
#undef CLM_SPLIT_LAUNCH
#undef CLM_SPLIT_LAUNCH_KG
  return hipGetLastError();
}

hipError_t launch_clear_exponents(hipStream_t stream, int n, int *a, int *b, int *c, int *d) {
  if (n <= 0 || (!a && !b && !c && !d)) return hipSuccess;
  hipLaunchKernelGGL(clear_exponents_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, a, b, c, d,
                     n);
  return hipGetLastError();
}

hipError_t launch_build_block_rows(hipStream_t stream, int n, double *Bout, long ldb,
                                   const double *coef, int nterms, const double *X1, long ldx1,
                                   const double *bank, long stride, long ldbank, int *row_e,
                                   int gen) {
  if (n < 1) return hipSuccess;
  auto a16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool aligned = (nterms < 2 || (a16(X1) && (ldx1 & 1) == 0)) &&
                       (nterms < 3 || (a16(bank) && (ldbank & 1) == 0 && (stride & 1) == 0));
  const dim3 grid((n + 63) / 64, (n + 63) / 64), block(256);
  hipLaunchKernelGGL(build_block_rows_kernel, grid, block, 0, stream, n, Bout, ldb, coef, nterms,
                     X1, ldx1, bank, stride, ldbank, row_e, gen, aligned);
  return hipGetLastError();
}

hipError_t launch_half_exponents(hipStream_t stream, int rows, int cols, const double *X, long ldx,
                                 int *row_e, int *col_e, int gen) {
  if (rows <= 0 || cols <= 0 || (!row_e && !col_e)) return hipSuccess;
  hipLaunchKernelGGL(max_exponent_kernel, dim3((rows + 63) / 64, (cols + 63) / 64), dim3(256), 0,
                     stream, rows, cols, X, ldx, row_e, col_e, gen);
  return hipGetLastError();
}

hipError_t launch_split_planes_half(hipStream_t stream, int rows, int cols, const double *X,
                                    long ldx, unsigned short *pc, unsigned short *pr, long pstride,
                                    const int *row_e, const int *col_e, int row_gen, int col_gen) {
  if (rows <= 0 || cols <= 0 || (!pc && !pr)) return hipSuccess;
  const dim3 grid((rows + 63) / 64, (cols + 63) / 64), block(256);
  hipLaunchKernelGGL(split_planes_half_kernel, grid, block, 0, stream, rows, cols, X, ldx, pc, pr,
                     pstride, row_e, col_e, row_gen, col_gen);
  return hipGetLastError();
}

}  // namespace clm
