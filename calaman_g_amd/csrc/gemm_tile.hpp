// gemm_tile.hpp -- the MFMA tile engine shared by the stand-alone product
// kernels (gemm.hip) and the batched persistent expm kernel (batched.hip):
// MFMA traits per element type, LDS staging layouts, the global -> register
// fetch of a K slice, its register -> LDS store, and the MFMA compute of one
// slice. A kernel composes them into its own loop; both kernels use the same
// helpers, so one output element accumulates its k terms in the same order
// (ascending, four per MFMA step) whichever kernel computes it.
//
// MFMA operand layout (16x16x4): A lane l holds A[l%16][l/16], B lane l holds
// B[l/16][l%16]; D lane l holds column l%16 at rows l/16 + 4r for the FP64
// instruction and rows 4(l/16) + r for the FP32 one (r = 0..3).
#pragma once

#include "clm_internal.hpp"

namespace clm {
namespace tile {

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

// tagged exponent of the half-format splits: (gen << 12) | (e + 2048), e clamped to
// [-2047, 2047]; atomicMax over tags is the max over e within a generation
__device__ __forceinline__ int exp_tag(int gen, int e) {
  return (gen << 12) | (min(max(e, -2047), 2047) + 2048);
}

// the stored value is returned (the split products' exponent tracking reads it)
template <typename T>
__device__ __forceinline__ T epilogue_store(const GemmArgs<T> &g, int row, int col, T acc) {
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
  return v;
}

// ------------------------------------------------------------------------------------
// real (S = double or float)
// ------------------------------------------------------------------------------------
template <typename S, int BM, int BN, int NT, int BK>
struct RTile {
  static constexpr int EV = Mfma<S>::EV;
  static constexpr int LDA_S = BM + 16;  // As[k][m]
  // Bs[n][k]: a row pitch of 2 (mod 32) LDS banks, BK + 1 doubles or BK + 2 floats, so
  // the 16 rows one k-step reads (lane % 16) land in 16 distinct bank pairs; a pitch
  // of BK + one 16-byte vector put rows n and n + 8 on the same banks (the hardware
  // counted 17M conflict cycles per n = 2048 FP64 product, 3.2 per LDS instruction).
  // The rows are no longer 16-byte aligned, so B is staged element by element.
  static constexpr int LDB_S = BK + (sizeof(S) == 8 ? 1 : 2);
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
#pragma unroll
    for (int e = 0; e < EV; ++e) Bs[nn * TL::LDB_S + kk + e] = reinterpret_cast<const S *>(&rb[it])[e];
  }
}

// the k-steps [K0, K1) of a slice (default: all of it), in ascending order
template <typename S, int BM, int BN, int WM, int WN, int NT, int BK, int K0 = 0, int K1 = BK>
__device__ __forceinline__ void rgemm_compute(const S *As, const S *Bs, int lane, int wm0, int wn0,
                                              typename Mfma<S>::acc (&acc)[WM / 16][WN / 16]) {
  using TL = RTile<S, BM, BN, NT, BK>;
  constexpr int TM = WM / 16, TN = WN / 16;
#pragma unroll
  for (int kk = K0; kk < K1; kk += 4) {
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
  static constexpr int LDB_S = BK + (sizeof(S) == 8 ? 1 : 2);  // as RTile's: 2 banks (mod 32)
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

// ------------------------------------------------------------------------------------
// Epilogue of a 128 x 128 real tile through LDS. The caller lays the tile's sums
// down in LDS (after the K loop the stages are free); then the whole workgroup
// (NW waves) walks the tile in column segments: wave w takes columns w, w + NW,
// ..., a lane two consecutive rows. Every read of Cin, X1 and the power bank
// and every write of C is one contiguous column piece per wave (1 KiB in FP64)
// in 2-element vectors, with up to eight columns' reads in flight per lane --
// where the MFMA layout gives short pieces of 16 columns per instruction, one
// column at a time. The arithmetic is epilogue_store's, in its order.
// fetch(c, r) returns the tile's value at row r of tile column c.
// ------------------------------------------------------------------------------------
template <typename S>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

template <typename S>
__device__ __forceinline__ bool epilogue_vec2_ok(const GemmArgs<S> &g) {
  auto al = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & (2 * sizeof(S) - 1)) == 0; };
  return (g.ldc & 1) == 0 && al(g.C) && (!g.Cin || ((g.ldcin & 1) == 0 && al(g.Cin))) &&
         (g.nterms < 2 || ((g.ldx1 & 1) == 0 && al(g.X1))) &&
         (g.nterms < 3 || ((g.ldbank & 1) == 0 && (g.bank_stride & 1) == 0 && al(g.bank)));
}

// row_tags (FP64 only; LDS, 128 ints zeroed by the caller before a barrier): when given
// and g.row_e_out is set, the largest |value| stored in each of the tile's rows is
// folded into its tagged exponent (an LDS max over the waves, then one global max per
// row) -- the split's exponent pass, done on the values as they are written.
template <typename S, int NW, bool kFast, int MAXCH = 8, typename Fetch>
__device__ __forceinline__ void lds_tile_epilogue(const GemmArgs<S> &g, int m0, int n0,
                                                  Fetch fetch, int *row_tags = nullptr) {
  using V2 = typename Vec2<S>::type;
  using O = ops<S>;
  const bool track = row_tags != nullptr && g.row_e_out != nullptr;
  S rmax0 = S(0), rmax1 = S(0);  // this lane's rows, over its columns
  constexpr int CPW = 128 / NW;                // columns per wave
  constexpr int CH = CPW < MAXCH ? CPW : MAXCH;  // columns per chunk of reads in flight
  static_assert(128 % NW == 0 && CPW % CH == 0, "column chunks");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rl = 2 * lane, row = m0 + rl;
  auto at = [&](const S *base, long ld, int k0, int k) {
    return reinterpret_cast<const V2 *>(&base[row + static_cast<long>(n0 + wave + NW * (k0 + k)) * ld]);
  };
  if (kFast && epilogue_vec2_ok(g)) {
#pragma unroll
    for (int c0 = 0; c0 < CPW; c0 += CH) {
      V2 v[CH];
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int c = wave + NW * (c0 + k);
        v[k].x = O::mul(g.alpha, fetch(c, rl));
        v[k].y = O::mul(g.alpha, fetch(c, rl + 1));
      }
      if (g.Cin) {
        V2 cin[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) cin[k] = *at(g.Cin, g.ldcin, c0, k);
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          v[k].x = O::add(v[k].x, O::mul(g.beta, cin[k].x));
          v[k].y = O::add(v[k].y, O::mul(g.beta, cin[k].y));
        }
      }
      if (g.nterms > 0) {
        V2 e[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int col = n0 + wave + NW * (c0 + k);
          e[k].x = row == col ? g.coef[0] : O::zero();
          e[k].y = row + 1 == col ? g.coef[0] : O::zero();
        }
        if (g.nterms > 1) {
          V2 x[CH];
#pragma unroll
          for (int k = 0; k < CH; ++k) x[k] = *at(g.X1, g.ldx1, c0, k);
          const S c1 = g.coef[1];
#pragma unroll
          for (int k = 0; k < CH; ++k) {
            e[k].x = O::mad(c1, x[k].x, e[k].x);
            e[k].y = O::mad(c1, x[k].y, e[k].y);
          }
        }
        for (int t = 2; t < g.nterms; ++t) {
          const S *bk = g.bank + static_cast<long>(t - 2) * g.bank_stride;
          V2 b[CH];
#pragma unroll
          for (int k = 0; k < CH; ++k) b[k] = *at(bk, g.ldbank, c0, k);
          const S ct = g.coef[t];
#pragma unroll
          for (int k = 0; k < CH; ++k) {
            e[k].x = O::mad(ct, b[k].x, e[k].x);
            e[k].y = O::mad(ct, b[k].y, e[k].y);
          }
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          v[k].x = O::add(v[k].x, e[k].x);
          v[k].y = O::add(v[k].y, e[k].y);
        }
      }
#pragma unroll
      for (int k = 0; k < CH; ++k)
        *const_cast<V2 *>(at(g.C, g.ldc, c0, k)) = v[k];
      if constexpr (std::is_same_v<S, double>) {
        if (track) {
#pragma unroll
          for (int k = 0; k < CH; ++k) {
            rmax0 = fmax(rmax0, fabs(v[k].x));
            rmax1 = fmax(rmax1, fabs(v[k].y));
          }
        }
      }
    }
  } else {
    for (int k = 0; k < CPW; ++k) {
      const int c = wave + NW * k, col = n0 + c;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (kFast || (row + h < g.M && col < g.N)) {
          const S v = epilogue_store(g, row + h, col, fetch(c, rl + h));
          if constexpr (std::is_same_v<S, double>) {
            if (track) {
              if (h == 0) rmax0 = fmax(rmax0, fabs(v));
              else rmax1 = fmax(rmax1, fabs(v));
            }
          }
        }
    }
  }
  if constexpr (std::is_same_v<S, double>) {
    if (track) {  // uniform: every lane reaches both barriers
      if (rmax0 > 0.0) atomicMax(&row_tags[rl], exp_tag(g.e_gen_out, ilogb(rmax0)));
      if (rmax1 > 0.0) atomicMax(&row_tags[rl + 1], exp_tag(g.e_gen_out, ilogb(rmax1)));
      __syncthreads();
      const int t = threadIdx.x;
      if (t < 128 && m0 + t < g.M && row_tags[t] > 0) atomicMax(&g.row_e_out[m0 + t], row_tags[t]);
    }
  }
}

// ------------------------------------------------------------------------------------
// Epilogue of any tile through LDS in column passes (the batched kernels' products and
// the complex products): after the K loop a tile's sums are laid
// down column-major (pitch BM + 1, so the 16 columns an MFMA tile's lanes hold fall
// in distinct banks) in the free staging LDS, PW columns at a time -- the widest of
// BN, BN / 2, ... (a multiple of the wave tile's width WN) that fits -- and every
// lane then stores elements in column order. Consecutive lanes hold consecutive
// rows, so C, Cin and the power blocks are touched in whole column pieces instead
// of the MFMA layout's 4-row pieces of 16 columns. The arithmetic is
// epilogue_store's.
template <typename E, int BM, int BN, int WN>
constexpr int epilogue_pass_cols(size_t lds_bytes_) {
  int pw = BN;
  while (pw > WN && static_cast<size_t>(BM + 1) * pw * sizeof(E) > lds_bytes_) pw /= 2;
  return pw;
}

// lanes write the sums of the wave tiles whose columns lie in [p0, p0 + PW); value(i,
// j, r) is the wave's element r of MFMA tile (i, j)
template <typename E, typename S, int BM, int TM, int TN, typename Value>
__device__ __forceinline__ void epilogue_put(E *sumt, int p0, int lane, int wm0, int wn0,
                                             Value value) {
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sumt[(wn0 - p0 + j * 16 + (lane & 15)) * (BM + 1) + wm0 + i * 16 + Mfma<S>::row(lane, r)] =
            value(i, j, r);
}

template <typename E, int BM, int PW, int NT>
__device__ __forceinline__ void epilogue_drain(const GemmArgs<E> &g, const E *sumt, int m0,
                                               int c0) {
  for (int e = threadIdx.x; e < BM * PW; e += NT) {
    const int r = e % BM, c = e / BM;
    const int row = m0 + r, col = c0 + c;
    if (row < g.M && col < g.N) epilogue_store(g, row, col, sumt[c * (BM + 1) + r]);
  }
}

// C = A*B (+ epilogue of g) over the whole n x n output, tile by tile, by all
// 256 lanes; ends with a barrier, so C is complete for every lane afterwards.
}  // namespace tile
}  // namespace clm
