// gemm_fp8.hip -- FP64-in / FP64-out matrix products on FP8 MFMA with a power-of-two
// scale per 128 x 128 tile (the low-order-term products of the FP8-tail variant).
//
// Quantization: every 128 x 128 tile of an FP64 matrix X gets the exponent
// e = floor(log2(448 / max|tile|)) (448 is the largest OCP e4m3 value; a value
// beyond it would convert to NaN), and q = e4m3(x 2^e) with round-to-nearest-
// even. Left operands are stored row-major (k contiguous), right operands
// column-major (k contiguous): the 16x16x32 MFMA wants 8 consecutive k of one
// row / column in each lane.
//
// Product: C(I, J) = sum over k-tiles K of [Aq(I, K) Bq(K, J)] 2^-(e_A(I, K) + e_B(K, J)).
// A 128 x 128 tile of C per workgroup of 8 wave64 (2 x 4 waves of 64 x 32); each
// LDS stage holds one 128-deep k-tile of both operands, so one slice has one
// scale pair: its four v_mfma_f32_16x16x32_fp8_fp8 k-steps accumulate a partial
// sum that is then added, scaled, to the running FP32 sum. The epilogue is the
// FP64 path's (alpha, beta*Cin, the fused Paterson-Stockmeyer block).
//
// MFMA operand layout (16x16x32 fp8, measured on this part): lane l holds A row
// l%16, k = 8(l/16) .. +7 as 8 bytes; B column l%16 likewise; D lane l holds
// rows 4(l/16) + r of column l%16.
#include "clm_internal.hpp"
#include "gemm_tile.hpp"

namespace clm {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kT = 128;        // tile edge: quantization tile = product tile = k-slice
constexpr int kRowB = kT + 16;  // LDS row (bytes): 16-byte pad
constexpr int kStage = 2 * kT * kRowB;

__device__ __forceinline__ unsigned fp8_pair(float a, float b) {
  return static_cast<unsigned>(__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false)) & 0xffffu;
}

// one 128 x 128 tile per workgroup: the tile's exponent (into ec and / or er),
// then its bytes in the column-major plane pc (ld rows) and / or the row-major
// plane pr (ld cols)
__global__ void __launch_bounds__(256)
    quantize_fp8_kernel(int rows, int cols, const double *__restrict__ X, long ldx,
                        unsigned char *__restrict__ pc, unsigned char *__restrict__ pr,
                        int *__restrict__ ec, int *__restrict__ er) {
  __shared__ double red[256];
  __shared__ unsigned char tile[kT][kT + 4];
  __shared__ int s_e;
  const int ti = blockIdx.x, tj = blockIdx.y;
  const int i0 = ti * kT, j0 = tj * kT;
  const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
  double mx = 0.0;
  for (int jj = ty; jj < kT; jj += 4)
    for (int ii = tx; ii < kT; ii += 64) {
      const int i = i0 + ii, j = j0 + jj;
      if (i < rows && j < cols) mx = fmax(mx, fabs(X[i + (long)j * ldx]));
    }
  red[tid] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
    __syncthreads();
  }
  if (tid == 0) {
    const double m = red[0];
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
  for (int jj = ty; jj < kT; jj += 4)
    for (int ii = 2 * tx; ii < kT; ii += 128) {
      const int i = i0 + ii, j = j0 + jj;
      float a = 0.f, b = 0.f;
      if (i < rows && j < cols) a = static_cast<float>(ldexp(X[i + (long)j * ldx], e));
      if (i + 1 < rows && j < cols) b = static_cast<float>(ldexp(X[i + 1 + (long)j * ldx], e));
      const unsigned q = fp8_pair(a, b);
      if (pc && j < cols) {
        if (i < rows) pc[i + (long)j * rows] = static_cast<unsigned char>(q & 0xff);
        if (i + 1 < rows) pc[i + 1 + (long)j * rows] = static_cast<unsigned char>(q >> 8);
      }
      tile[jj][ii] = static_cast<unsigned char>(q & 0xff);
      tile[jj][ii + 1] = static_cast<unsigned char>(q >> 8);
    }
  if (!pr) return;
  __syncthreads();
  for (int ii = ty; ii < kT; ii += 4)
    for (int jj = tx; jj < kT; jj += 64) {
      const int i = i0 + ii, j = j0 + jj;
      if (i < rows && j < cols) pr[(long)i * cols + j] = tile[jj][ii];
    }
}

// global -> registers: one 128 x 128 byte slice of each operand, 2 x 16 bytes per lane
template <bool kFast>
__device__ __forceinline__ void fp8_gload(const GemmArgs<double> &g, const unsigned char *PA,
                                          const unsigned char *PB, int m0, int n0, int k0, int tid,
                                          uint4 (&ra)[2], uint4 (&rb)[2]) {
  const long K = g.K;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int c = tid + it * 512;
    const int r = c >> 3, kq = (c & 7) * 16;
    if (kFast) {
      ra[it] = *reinterpret_cast<const uint4 *>(PA + (long)(m0 + r) * K + k0 + kq);
      rb[it] = *reinterpret_cast<const uint4 *>(PB + (long)(n0 + r) * K + k0 + kq);
    } else {
      unsigned char va[16], vb[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const bool kin = k0 + kq + e < K;
        va[e] = (kin && m0 + r < g.M) ? PA[(long)(m0 + r) * K + k0 + kq + e] : 0;
        vb[e] = (kin && n0 + r < g.N) ? PB[(long)(n0 + r) * K + k0 + kq + e] : 0;
      }
      ra[it] = *reinterpret_cast<const uint4 *>(va);
      rb[it] = *reinterpret_cast<const uint4 *>(vb);
    }
  }
}

__device__ __forceinline__ void fp8_sstore(unsigned char *S, int tid, const uint4 (&ra)[2],
                                           const uint4 (&rb)[2]) {
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int c = tid + it * 512;
    const int r = c >> 3, kq = (c & 7) * 16;
    *reinterpret_cast<uint4 *>(&S[r * kRowB + kq]) = ra[it];
    *reinterpret_cast<uint4 *>(&S[kT * kRowB + r * kRowB + kq]) = rb[it];
  }
}

// the 128-deep slice: four k-steps into a zeroed partial sum, then added to acc
// with the slice's scale
__device__ __forceinline__ void fp8_compute(const unsigned char *S, int lane, int wm0, int wn0,
                                            float scale, f4 (&acc)[4][2]) {
  f4 part[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const int koff = ks * 32 + 8 * (lane >> 4);
    long a[4], b[2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      a[i] = *reinterpret_cast<const long *>(&S[(wm0 + i * 16 + (lane & 15)) * kRowB + koff]);
#pragma unroll
    for (int j = 0; j < 2; ++j)
      b[j] = *reinterpret_cast<const long *>(
          &S[kT * kRowB + (wn0 + j * 16 + (lane & 15)) * kRowB + koff]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a[i], b[j], part[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = fmaf(part[i][j][r], scale, acc[i][j][r]);
}

template <bool kFast>
__global__ void __launch_bounds__(512)
    dgemm_fp8_tiles(const GemmArgs<double> g, const unsigned char *__restrict__ PA,
                    const unsigned char *__restrict__ PB, const int *__restrict__ eA,
                    const int *__restrict__ eB, const int tiles_m, const int tiles_n,
                    const int tiles_k) {
  __shared__ __attribute__((aligned(16))) unsigned char L[2 * kStage];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
  const int m0 = tm * kT, n0 = tn * kT;
  const int wm0 = (wave >> 2) * 64, wn0 = (wave & 3) * 32;
  f4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // exponents: A tiles (tm, tk) of an tiles_m x tiles_k grid, B tiles (tk, tn)
  auto scale_of = [&](int tk) {
    return ldexpf(1.0f, -(eA[tm + tk * tiles_m] + eB[tk + tn * tiles_k]));
  };
  uint4 ra[2], rb[2];
  const int KT = tiles_k;
  if (KT > 0) {
    fp8_gload<kFast>(g, PA, PB, m0, n0, 0, tid, ra, rb);
    fp8_sstore(L, tid, ra, rb);
    __syncthreads();
    for (int kt = 0; kt < KT - 1; ++kt) {
      const int cur = kt & 1;
      fp8_gload<kFast>(g, PA, PB, m0, n0, (kt + 1) * kT, tid, ra, rb);
      fp8_compute(L + cur * kStage, lane, wm0, wn0, scale_of(kt), acc);
      fp8_sstore(L + (cur ^ 1) * kStage, tid, ra, rb);
      __syncthreads();
    }
    fp8_compute(L + ((KT - 1) & 1) * kStage, lane, wm0, wn0, scale_of(KT - 1), acc);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn0 + j * 16 + (lane & 15);
        if (kFast || (row < g.M && col < g.N))
          tile::epilogue_store(g, row, col, static_cast<double>(acc[i][j][r]));
      }
}

inline bool aligned16_8(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace

hipError_t launch_quantize_fp8(hipStream_t stream, int rows, int cols, const double *X, long ldx,
                               unsigned char *pc, unsigned char *pr, int *ec, int *er) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const dim3 grid((rows + kT - 1) / kT, (cols + kT - 1) / kT);
  hipLaunchKernelGGL(quantize_fp8_kernel, grid, dim3(256), 0, stream, rows, cols, X, ldx, pc, pr,
                     ec, er);
  return hipGetLastError();
}

hipError_t launch_gemm_fp8(hipStream_t stream, const GemmArgs<double> &g, const unsigned char *PA,
                           const unsigned char *PB, const int *eA, const int *eB) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  const int tiles_m = (g.M + kT - 1) / kT, tiles_n = (g.N + kT - 1) / kT;
  const int tiles_k = (g.K + kT - 1) / kT;
  const bool fast = g.M % kT == 0 && g.N % kT == 0 && g.K % kT == 0 && aligned16_8(PA) &&
                    aligned16_8(PB);
  const dim3 grid(tiles_m * tiles_n), block(512);
  if (fast)
    hipLaunchKernelGGL(dgemm_fp8_tiles<true>, grid, block, 0, stream, g, PA, PB, eA, eB, tiles_m,
                       tiles_n, tiles_k);
  else
    hipLaunchKernelGGL(dgemm_fp8_tiles<false>, grid, block, 0, stream, g, PA, PB, eA, eB, tiles_m,
                       tiles_n, tiles_k);
  return hipGetLastError();
}

}  // namespace clm
