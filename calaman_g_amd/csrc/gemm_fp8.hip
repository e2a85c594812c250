// gemm_fp8.hip -- FP64-in / FP64-out matrix products on FP8 MFMA with a power-of-two
// scale per 128 x 128 tile (the high-degree-term products of the FP8-tail variant).
//
// Quantization: every 128 x 128 tile of an FP64 matrix X gets the exponent
// e = floor(log2(448 / max|tile|)) (448 is the largest OCP e4m3 value; a value
// beyond it would convert to NaN), and q = e4m3(x 2^e) with round-to-nearest-
// even. Left operands are stored row-major (k contiguous), right operands
// column-major (k contiguous): the MFMA wants 32 consecutive k of one row /
// column in each lane.
//
// Product: C(I, J) = sum over k-tiles K of [Aq(I, K) Bq(K, J)] 2^-(e_A(I, K) + e_B(K, J)),
// the scales applied by the MFMA itself (see dgemm_fp8_w), FP32 accumulation, and
// the FP64 path's epilogue (alpha, beta*Cin, the fused Paterson-Stockmeyer block).
#include "clm_internal.hpp"
#include "gemm_tile.hpp"

namespace clm {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kT = 128;   // quantization tile edge = k-tile depth
constexpr int kRowB = kT;  // LDS row (bytes): one k-tile of a row, unpadded
// byte offset of (row, k) in an LDS plane, k a multiple of 16: the row's eight
// 16-byte chunks permuted by an XOR with the row index mod 8, which makes the
// fragment reads (16 lanes on 16 rows, one chunk each) and the staging writes
// (8 lanes on one row's 8 chunks) free of bank conflicts (measured with
// SQ_LDS_BANK_CONFLICT on MI355X; 144-byte padded rows conflicted)
__device__ __forceinline__ int fp8_off(int row, int k) {
  return row * kRowB + (((k >> 4) ^ (row & 7)) << 4);
}

__device__ __forceinline__ unsigned fp8_pair(float a, float b) {
  return static_cast<unsigned>(__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false)) & 0xffffu;
}

// One 128 x 128 tile per workgroup of 512 lanes, read from HBM once: lane t holds
// rows 2 (t % 64) and 2 (t % 64) + 1 of the tile's columns t / 64 + 8 q, q < 16, in
// registers; the tile's max (wave shuffles, then LDS) gives its exponent (into ec
// and / or er); the pairs are converted to e4m3 two at a time and written as
// 16-bit words down the columns of the column-major plane pc (ld rows), and
// through an LDS byte tile along the rows of the row-major plane pr (ld cols).
template <bool kFast>
__global__ void __launch_bounds__(512)
    quantize_fp8_kernel(int rows, int cols, const double *__restrict__ X, long ldx,
                        unsigned char *__restrict__ pc, unsigned char *__restrict__ pr,
                        int *__restrict__ ec, int *__restrict__ er) {
  __shared__ double red[8];
  __shared__ unsigned char tile[kT][kT + 4];  // [column][row]
  __shared__ int s_e;
  const int ti = blockIdx.x, tj = blockIdx.y;
  const int i0 = ti * kT, j0 = tj * kT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ii = 2 * lane;  // the lane's row pair within the tile
  double2 v[16];
  double mx = 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int jj = wave + 8 * q;
    const int i = i0 + ii, j = j0 + jj;
    if (kFast) {
      v[q] = *reinterpret_cast<const double2 *>(X + i + (long)j * ldx);
    } else {
      v[q].x = (i < rows && j < cols) ? X[i + (long)j * ldx] : 0.0;
      v[q].y = (i + 1 < rows && j < cols) ? X[i + 1 + (long)j * ldx] : 0.0;
    }
    mx = fmax(mx, fmax(fabs(v[q].x), fabs(v[q].y)));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if (lane == 0) red[wave] = mx;
  __syncthreads();
  if (tid == 0) {
    double m = red[0];
    for (int w = 1; w < 8; ++w) m = fmax(m, red[w]);
    int e = 0;
    if (m > 0.0 && isfinite(m)) {
      e = ilogb(448.0 / m);
      if (ldexp(m, e) > 448.0) --e;  // ilogb floors; keep max * 2^e within range
    }
    s_e = e;
    if (ec) ec[ti + tj * gridDim.x] = e;
    if (er) er[ti + tj * gridDim.x] = e;
  }
  __syncthreads();
  const int e = s_e;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int jj = wave + 8 * q;
    const int i = i0 + ii, j = j0 + jj;
    const unsigned w = fp8_pair(static_cast<float>(ldexp(v[q].x, e)),
                                static_cast<float>(ldexp(v[q].y, e)));
    if (pc) {
      if (kFast) {
        *reinterpret_cast<unsigned short *>(pc + i + (long)j * rows) = static_cast<unsigned short>(w);
      } else if (j < cols) {
        if (i < rows) pc[i + (long)j * rows] = static_cast<unsigned char>(w & 0xff);
        if (i + 1 < rows) pc[i + 1 + (long)j * rows] = static_cast<unsigned char>(w >> 8);
      }
    }
    *reinterpret_cast<unsigned short *>(&tile[jj][ii]) = static_cast<unsigned short>(w);
  }
  if (!pr) return;
  __syncthreads();
  // rows of pr: lane t writes 4 consecutive columns of row t / 32 + 16 p
  for (int idx = tid; idx < kT * kT / 4; idx += 512) {
    const int ir = idx >> 5, jq = (idx & 31) * 4;
    const int i = i0 + ir, j = j0 + jq;
    if (kFast) {
      const unsigned word = tile[jq][ir] | (tile[jq + 1][ir] << 8) | (tile[jq + 2][ir] << 16) |
                            (static_cast<unsigned>(tile[jq + 3][ir]) << 24);
      *reinterpret_cast<unsigned *>(pr + (long)i * cols + j) = word;
    } else if (i < rows) {
      for (int c = 0; c < 4; ++c)
        if (j + c < cols) pr[(long)i * cols + j + c] = tile[jq + c][ir];
    }
  }
}

// ------------------------------------------------------------------------------------
// The product kernel: a 256 x 128 tile of C per workgroup of 8 wave64 (4 x 2 waves of
// 64 x 64), one 128-deep k-tile per LDS stage (A 256 x 128 and B 128 x 128 bytes, rows
// padded to 144 bytes), double-buffered. Each wave issues one scaled
// v_mfma_scale_f32_16x16x128_f8f6f4 per 16 x 16 sub-tile and k-tile: the operands'
// tile exponents enter as the instruction's E8M0 block scales (2^(127 - e) each), so
// every product lands in the FP32 accumulator already scaled. A wave's 64 rows lie in
// one 128-row quantization tile of A and the workgroup's 128 columns in one of B, so
// the scales are uniform over a wave.
// Operand layout (measured on the part with an in-bounds probe): lane l holds row
// (A) / column (B) l % 16, k = 32 (l / 16) + b in byte b of its 32 bytes; format code
// 0 is e4m3; D rows 4 (l / 16) + r of column l % 16.
// ------------------------------------------------------------------------------------
typedef int i8v __attribute__((ext_vector_type(8)));
// staging registers as a native vector: HIP's uint4 (a struct with a union) kept the
// arrays of it from being split into registers, and the compiler moved them to LDS
typedef unsigned u4v __attribute__((ext_vector_type(4)));
constexpr int kBM = 256, kBN = 128;
constexpr int kStageA = kBM * kRowB, kStageB = kBN * kRowB;
constexpr int kStageW = kStageA + kStageB;  // bytes per stage

template <bool kFast>
__device__ __forceinline__ void fp8w_gload(const GemmArgs<double> &g, const unsigned char *PA,
                                           const unsigned char *PB, int m0, int n0, int k0,
                                           int tid, u4v (&ra)[4], u4v (&rb)[2]) {
  const long K = g.K;
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const int c = tid + it * 512;
    const int r = c >> 3, kq = (c & 7) * 16;
    if (kFast) {
      ra[it] = *reinterpret_cast<const u4v *>(PA + (long)(m0 + r) * K + k0 + kq);
    } else {
      u4v v = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int e = 0; e < 16; ++e)
        if (k0 + kq + e < K && m0 + r < g.M)
          v[e >> 2] |= static_cast<unsigned>(PA[(long)(m0 + r) * K + k0 + kq + e]) << (8 * (e & 3));
      ra[it] = v;
    }
  }
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int c = tid + it * 512;
    const int r = c >> 3, kq = (c & 7) * 16;
    if (kFast) {
      rb[it] = *reinterpret_cast<const u4v *>(PB + (long)(n0 + r) * K + k0 + kq);
    } else {
      u4v v = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int e = 0; e < 16; ++e)
        if (k0 + kq + e < K && n0 + r < g.N)
          v[e >> 2] |= static_cast<unsigned>(PB[(long)(n0 + r) * K + k0 + kq + e]) << (8 * (e & 3));
      rb[it] = v;
    }
  }
}

__device__ __forceinline__ void fp8w_sstore(unsigned char *S, int tid, const u4v (&ra)[4],
                                            const u4v (&rb)[2]) {
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const int c = tid + it * 512;
    *reinterpret_cast<u4v *>(&S[fp8_off(c >> 3, (c & 7) * 16)]) = ra[it];
  }
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int c = tid + it * 512;
    *reinterpret_cast<u4v *>(&S[kStageA + fp8_off(c >> 3, (c & 7) * 16)]) = rb[it];
  }
}

// a lane's 32 fragment bytes of one row: k .. k + 31 (two chunks, swizzled apart)
__device__ __forceinline__ i8v lds_i8v(const unsigned char *S, int row, int k) {
  const u4v lo = *reinterpret_cast<const u4v *>(S + fp8_off(row, k));
  const u4v hi = *reinterpret_cast<const u4v *>(S + fp8_off(row, k + 16));
  i8v v;
  v[0] = lo[0]; v[1] = lo[1]; v[2] = lo[2]; v[3] = lo[3];
  v[4] = hi[0]; v[5] = hi[1]; v[6] = hi[2]; v[7] = hi[3];
  return v;
}

// a wave's fragments of one 128-deep k-tile
__device__ __forceinline__ void fp8w_read(const unsigned char *S, int lane, int wm0, int wn0,
                                          i8v (&a)[4], i8v (&b)[4]) {
  const int koff = 32 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = lds_i8v(S, wm0 + i * 16 + (lane & 15), koff);
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = lds_i8v(S + kStageA, wn0 + j * 16 + (lane & 15), koff);
}

// the scaled MFMAs of rows I0 .. I1 - 1 (16-row sub-tiles) of a wave's tile, every column
template <int I0, int I1>
__device__ __forceinline__ void fp8w_mma(const i8v (&a)[4], const i8v (&b)[4], int sa, int sb,
                                         f4 (&acc)[4][4]) {
#pragma unroll
  for (int i = I0; i < I1; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[i], b[j], acc[i][j], 0, 0, 0,
                                                                   sa, 0, sb);
}

// E8M0 scale of a tile exponent e: 2^-e, clamped to the format's range
__device__ __forceinline__ int e8m0_of(int e) {
  const int v = 127 - e;
  return v < 0 ? 0 : (v > 254 ? 254 : v);
}

template <bool kFast>
__global__ void __launch_bounds__(512)
    dgemm_fp8_w(const GemmArgs<double> g, const unsigned char *__restrict__ PA,
                const unsigned char *__restrict__ PB, const int *__restrict__ eA,
                const int *__restrict__ eB, const int tiles_m, const int tiles_n,
                const int qtiles_m, const int tiles_k) {
  __shared__ __attribute__((aligned(16))) unsigned char L[2 * kStageW];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
  const int m0 = tm * kBM, n0 = tn * kBN;
  const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 64;
  // the quantization tiles this wave's rows and the workgroup's columns lie in
  // (a row tile past the matrix -- rows that are all padding -- reads the last one)
  const int qm = min((m0 + wm0) / 128, qtiles_m - 1);
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // Software pipeline over the k-tiles: a wave holds k-tile kt's fragments in
  // registers; it runs the MFMAs of its first two row sub-tiles, stages k-tile kt + 1
  // (loaded one k-tile earlier, the stores spread over those MFMAs) and loads
  // k-tile kt + 2, meets the barrier, then issues its LDS reads of k-tile kt + 1 and
  // runs the last two row sub-tiles behind them. The stage written at k-tile kt held
  // k-tile kt - 1, read into registers before the previous barrier (which waits for
  // every LDS access of the wave). Each accumulator takes the k-tiles in order.
  u4v ra[4], rb[2];
  const int KT = tiles_k;
  if (KT > 0) {
    fp8w_gload<kFast>(g, PA, PB, m0, n0, 0, tid, ra, rb);
    fp8w_sstore(L, tid, ra, rb);
    if (KT > 1) fp8w_gload<kFast>(g, PA, PB, m0, n0, kT, tid, ra, rb);
    __syncthreads();
    i8v a[4], b[4];
    fp8w_read(L, lane, wm0, wn0, a, b);
    for (int kt = 0; kt < KT - 1; ++kt) {
      const int nxt = (kt + 1) & 1;
      const int sa = e8m0_of(eA[qm + kt * qtiles_m]), sb = e8m0_of(eB[kt + tn * tiles_k]);
      fp8w_mma<0, 2>(a, b, sa, sb, acc);
      fp8w_sstore(L + nxt * kStageW, tid, ra, rb);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
      }
      // the first half's results exist here: its MFMAs stay ahead of the barrier
      asm volatile("" : "+v"(acc[0][0]), "+v"(acc[0][1]), "+v"(acc[0][2]), "+v"(acc[0][3]),
                   "+v"(acc[1][0]), "+v"(acc[1][1]), "+v"(acc[1][2]), "+v"(acc[1][3]));
      if (kt + 2 < KT) fp8w_gload<kFast>(g, PA, PB, m0, n0, (kt + 2) * kT, tid, ra, rb);
      __syncthreads();
      i8v na[4], nb[4];
      fp8w_read(L + nxt * kStageW, lane, wm0, wn0, na, nb);
      __builtin_amdgcn_sched_barrier(0);
      fp8w_mma<2, 4>(a, b, sa, sb, acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = na[i];
        b[i] = nb[i];
      }
    }
    const int sa = e8m0_of(eA[qm + (KT - 1) * qtiles_m]), sb = e8m0_of(eB[KT - 1 + tn * tiles_k]);
    fp8w_mma<0, 4>(a, b, sa, sb, acc);
  }
  // the epilogue through LDS, one 128-row half of the tile at a time: the four waves
  // holding the half lay their sums down column-major, then every wave writes C (and
  // reads Cin / the power blocks) in whole column pieces (tile::lds_tile_epilogue)
  constexpr int kSumLd = 132;
  static_assert(128 * kSumLd * 4 <= 2 * kStageW, "a half tile's sums fit in the stages");
  float *sum = reinterpret_cast<float *>(L);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __syncthreads();  // the stages (or the previous half) are no longer read
    if ((wm0 >> 7) == h) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          *reinterpret_cast<f4 *>(&sum[(wn0 + j * 16 + (lane & 15)) * kSumLd + (wm0 & 127) + i * 16 +
                                       4 * (lane >> 4)]) = acc[i][j];
    }
    __syncthreads();
    // (four columns in flight per lane: the other half's accumulators are still live)
    tile::lds_tile_epilogue<double, 8, kFast, 4>(g, m0 + 128 * h, n0, [&](int c, int r) {
      return static_cast<double>(sum[c * kSumLd + r]);
    });
  }
}

inline bool aligned16_8(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace

hipError_t launch_quantize_fp8(hipStream_t stream, int rows, int cols, const double *X, long ldx,
                               unsigned char *pc, unsigned char *pr, int *ec, int *er) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const dim3 grid((rows + kT - 1) / kT, (cols + kT - 1) / kT);
  const bool fast = rows % kT == 0 && cols % kT == 0 && ldx % 2 == 0 &&
                    (reinterpret_cast<uintptr_t>(X) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(pc) & 1) == 0 &&
                    (reinterpret_cast<uintptr_t>(pr) & 3) == 0;
  if (fast)
    hipLaunchKernelGGL(quantize_fp8_kernel<true>, grid, dim3(512), 0, stream, rows, cols, X, ldx,
                       pc, pr, ec, er);
  else
    hipLaunchKernelGGL(quantize_fp8_kernel<false>, grid, dim3(512), 0, stream, rows, cols, X, ldx,
                       pc, pr, ec, er);
  return hipGetLastError();
}

hipError_t launch_gemm_fp8(hipStream_t stream, const GemmArgs<double> &g, const unsigned char *PA,
                           const unsigned char *PB, const int *eA, const int *eB) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  const int tiles_m = (g.M + kBM - 1) / kBM, tiles_n = (g.N + kBN - 1) / kBN;
  const int qtiles_m = (g.M + kT - 1) / kT;
  const int tiles_k = (g.K + kT - 1) / kT;
  const bool fast = g.M % kBM == 0 && g.N % kBN == 0 && g.K % kT == 0 && aligned16_8(PA) &&
                    aligned16_8(PB);
  const dim3 grid(tiles_m * tiles_n), block(512);
  if (fast)
    hipLaunchKernelGGL(dgemm_fp8_w<true>, grid, block, 0, stream, g, PA, PB, eA, eB, tiles_m,
                       tiles_n, qtiles_m, tiles_k);
  else
    hipLaunchKernelGGL(dgemm_fp8_w<false>, grid, block, 0, stream, g, PA, PB, eA, eB, tiles_m,
                       tiles_n, qtiles_m, tiles_k);
  return hipGetLastError();
}

}  // namespace clm
