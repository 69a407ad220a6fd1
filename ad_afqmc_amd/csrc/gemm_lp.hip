// Reduced-precision GEMM for the walker step (FP32 and split-BF16 x3 on
// matrix cores, FP32 accumulation).
//
// Same tiling as the FP64 GEMM (WM x WN waves of TM x TN 16x16 blocks, two LDS
// buffers with register prefetch), with operands staged in LDS as FP32 and a
// 32-deep K slab.  Global operands may be double or float; they are rounded
// to FP32 when staged.
//   F32    : v_mfma_f32_16x16x4f32, eight k-steps of 4 per slab.
//   BF16X3 : each FP32 operand x is split into hi = bf16(x), lo = bf16(x - hi);
//            a product is hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16
//            (lane l holds 8 consecutive k), i.e. ~16 significant bits per
//            operand and FP32 accumulation.
// Output C may be float2 (complex64) or double2; D (double2) += alpha*result, or
// D = E + alpha*result with a separate complex-double addend E.
//
// MFMA f32 16x16 lane layout (gfx950): accumulator element v of lane l is
// D[4*(l/16) + v][l%16]; f32 16x16x4 inputs A[l%16][l/16], B[l/16][l%16];
// bf16 16x16x32 inputs A[l%16][8*(l/16)+t], B[8*(l/16)+t][l%16], t = 0..7.
#include <type_traits>

#include "afq_internal.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace afq {
namespace {

constexpr int BKL = 32;

__device__ inline double2 cmul(double2 a, double2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// LDS row strides (floats).  F32 fragments read 4 rows k..k+3: stride = 16 (mod 64)
// puts them in disjoint bank quarters.  BF16 fragments read rows 8 apart: 8*stride = 16 (mod 64).
constexpr int pad_f32(int w) { return w + ((16 - (w % 64)) % 64 + 64) % 64; }
constexpr int pad_bf16(int w) { return w + (((2 - (w % 8)) % 8) + 8) % 8; }

template <typename T>
__device__ inline float2 ld2(const T* p, long off, bool cplx) {
  if (cplx) {
    if constexpr (sizeof(T) == 8) {
      const double2 v = reinterpret_cast<const double2*>(p)[off];
      return float2{(float)v.x, (float)v.y};
    } else {
      return reinterpret_cast<const float2*>(p)[off];
    }
  }
  return float2{(float)p[off], 0.f};
}

template <bool BF, bool AC, bool BC, bool AT, bool BT, typename TA, typename TB, int WM, int WN, int TM, int TN>
struct CfgL {
  static constexpr int NT = 64 * WM * WN;
  static constexpr int BM = 16 * TM * WM, BN = 16 * TN * WN;
  static constexpr int LDA = BF ? pad_bf16(BM) : pad_f32(BM);
  static constexpr int LDB = BF ? pad_bf16(BN) : pad_f32(BN);
  static constexpr int PA = AC ? 2 : 1, PB = BC ? 2 : 1;
  static constexpr int QA = BM * BKL / 4, QB = BN * BKL / 4;
  static constexpr int RA = (QA + NT - 1) / NT, RB = (QB + NT - 1) / NT;
};

template <class C, bool AC, bool AT, typename TA>
__device__ inline void fetch_a(float (&r)[C::RA][4][2], const TA* A, long lda, int M, int K, int m0, int k0, int tid,
                               int asym) {
#pragma unroll
  for (int i = 0; i < C::RA; ++i) {
    const int q = tid + i * C::NT;
#pragma unroll
    for (int t = 0; t < 4; ++t) r[i][t][0] = r[i][t][1] = 0.f;
    if (q >= C::QA) continue;
    const int m = AT ? (q % (C::BM / 4)) * 4 : q % C::BM;
    const int k = AT ? q / (C::BM / 4) : (q / C::BM) * 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int gm = m0 + m + (AT ? t : 0), gk = k0 + k + (AT ? 0 : t);
      if (gm < M && gk < K) {
        const float2 v = ld2(A, AT ? (long)gk * lda + gm : asym ? sym_off(gm, gk) : (long)gm * lda + gk, AC);
        r[i][t][0] = v.x;
        r[i][t][1] = v.y;
      }
    }
  }
}

template <class C, bool AC, bool AT>
__device__ inline void store_a(float* As, const float (&r)[C::RA][4][2], int tid) {
#pragma unroll
  for (int i = 0; i < C::RA; ++i) {
    const int q = tid + i * C::NT;
    if (q >= C::QA) continue;
    const int m = AT ? (q % (C::BM / 4)) * 4 : q % C::BM;
    const int k = AT ? q / (C::BM / 4) : (q / C::BM) * 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int mm = m + (AT ? t : 0), kk = k + (AT ? 0 : t);
      As[kk * C::LDA + mm] = r[i][t][0];
      if (AC) As[BKL * C::LDA + kk * C::LDA + mm] = r[i][t][1];
    }
  }
}

template <class C, bool BC, bool BT, typename TB>
__device__ inline void fetch_b(float (&r)[C::RB][4][2], const TB* B, long ldb, int N, int K, int n0, int k0, int tid) {
#pragma unroll
  for (int i = 0; i < C::RB; ++i) {
    const int q = tid + i * C::NT;
#pragma unroll
    for (int t = 0; t < 4; ++t) r[i][t][0] = r[i][t][1] = 0.f;
    if (q >= C::QB) continue;
    const int n = BT ? q % C::BN : (q % (C::BN / 4)) * 4;
    const int k = BT ? (q / C::BN) * 4 : q / (C::BN / 4);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int gn = n0 + n + (BT ? 0 : t), gk = k0 + k + (BT ? t : 0);
      if (gn < N && gk < K) {
        const float2 v = ld2(B, BT ? (long)gn * ldb + gk : (long)gk * ldb + gn, BC);
        r[i][t][0] = v.x;
        r[i][t][1] = v.y;
      }
    }
  }
}

template <class C, bool BC, bool BT>
__device__ inline void store_b(float* Bs, const float (&r)[C::RB][4][2], int tid) {
#pragma unroll
  for (int i = 0; i < C::RB; ++i) {
    const int q = tid + i * C::NT;
    if (q >= C::QB) continue;
    const int n = BT ? q % C::BN : (q % (C::BN / 4)) * 4;
    const int k = BT ? (q / C::BN) * 4 : q / (C::BN / 4);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int nn = n + (BT ? 0 : t), kk = k + (BT ? t : 0);
      Bs[kk * C::LDB + nn] = r[i][t][0];
      if (BC) Bs[BKL * C::LDB + kk * C::LDB + nn] = r[i][t][1];
    }
  }
}

// Raw slab fetches for operands stored in FP64 (the VHS and force-bias
// products): each element is read at a clamped, always valid address and kept
// in its storage type; store_a_raw / store_b_raw zero what lies outside the
// operand and round to FP32 when they write the slab to LDS after the MFMAs.
// Rounding (or masking) at load time consumed each loaded value inside a
// divergent branch, so the compiler waited for the loads one at a time before
// the MFMAs.  (Measured: VHS and force bias 8-12% faster; FP32 operands, whose
// loads need no conversion, gained nothing and keep the fetches above.)
template <typename T>
__device__ inline void ld_raw(const T* p, long off, bool cplx, T& re, T& im) {
  if (cplx) {
    if constexpr (sizeof(T) == 8) {
      const double2 v = reinterpret_cast<const double2*>(p)[off];
      re = v.x;
      im = v.y;
    } else {
      const float2 v = reinterpret_cast<const float2*>(p)[off];
      re = v.x;
      im = v.y;
    }
  } else {
    re = p[off];
    im = T(0);
  }
}

template <class C, bool AC, bool AT, typename TA>
__device__ inline void fetch_a_raw(TA (&r)[C::RA][4][2], const TA* A, long lda, int M, int K, int m0, int k0, int tid,
                                   int asym) {
#pragma unroll
  for (int i = 0; i < C::RA; ++i) {
    const int q = min(tid + i * C::NT, C::QA - 1);
    const int m = AT ? (q % (C::BM / 4)) * 4 : q % C::BM;
    const int k = AT ? q / (C::BM / 4) : (q / C::BM) * 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int gm = min(m0 + m + (AT ? t : 0), M - 1), gk = min(k0 + k + (AT ? 0 : t), K - 1);
      const long off = AT ? (long)gk * lda + gm : asym ? sym_off(gm, gk) : (long)gm * lda + gk;
      ld_raw(A, off, AC, r[i][t][0], r[i][t][1]);
    }
  }
}

template <class C, bool AC, bool AT, typename TA>
__device__ inline void store_a_raw(float* As, const TA (&r)[C::RA][4][2], int tid, int M, int K, int m0, int k0) {
#pragma unroll
  for (int i = 0; i < C::RA; ++i) {
    const int q = tid + i * C::NT;
    if (q >= C::QA) continue;
    const int m = AT ? (q % (C::BM / 4)) * 4 : q % C::BM;
    const int k = AT ? q / (C::BM / 4) : (q / C::BM) * 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int mm = m + (AT ? t : 0), kk = k + (AT ? 0 : t);
      const bool ok = m0 + mm < M && k0 + kk < K;
      As[kk * C::LDA + mm] = ok ? (float)r[i][t][0] : 0.f;
      if (AC) As[BKL * C::LDA + kk * C::LDA + mm] = ok ? (float)r[i][t][1] : 0.f;
    }
  }
}

template <class C, bool BC, bool BT, typename TB>
__device__ inline void fetch_b_raw(TB (&r)[C::RB][4][2], const TB* B, long ldb, int N, int K, int n0, int k0, int tid) {
#pragma unroll
  for (int i = 0; i < C::RB; ++i) {
    const int q = min(tid + i * C::NT, C::QB - 1);
    const int n = BT ? q % C::BN : (q % (C::BN / 4)) * 4;
    const int k = BT ? (q / C::BN) * 4 : q / (C::BN / 4);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int gn = min(n0 + n + (BT ? 0 : t), N - 1), gk = min(k0 + k + (BT ? t : 0), K - 1);
      ld_raw(B, BT ? (long)gn * ldb + gk : (long)gk * ldb + gn, BC, r[i][t][0], r[i][t][1]);
    }
  }
}

template <class C, bool BC, bool BT, typename TB>
__device__ inline void store_b_raw(float* Bs, const TB (&r)[C::RB][4][2], int tid, int N, int K, int n0, int k0) {
#pragma unroll
  for (int i = 0; i < C::RB; ++i) {
    const int q = tid + i * C::NT;
    if (q >= C::QB) continue;
    const int n = BT ? q % C::BN : (q % (C::BN / 4)) * 4;
    const int k = BT ? (q / C::BN) * 4 : q / (C::BN / 4);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int nn = n + (BT ? 0 : t), kk = k + (BT ? t : 0);
      const bool ok = n0 + nn < N && k0 + kk < K;
      Bs[kk * C::LDB + nn] = ok ? (float)r[i][t][0] : 0.f;
      if (BC) Bs[BKL * C::LDB + kk * C::LDB + nn] = ok ? (float)r[i][t][1] : 0.f;
    }
  }
}

struct BfPair {
  v8bf hi, lo;
};
// 8 consecutive k (rows of the [k][m] LDS layout) of column `col`, split hi/lo.
__device__ inline BfPair frag_bf(const float* s, int ld, int k0, int col, bool neg) {
  BfPair p;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    float x = s[(k0 + t) * ld + col];
    if (neg) x = -x;
    const __bf16 h = (__bf16)x;
    p.hi[t] = h;
    p.lo[t] = (__bf16)(x - (float)h);
  }
  return p;
}

__device__ inline v4f mma3(const BfPair& a, const BfPair& b, v4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
  return c;
}

// grid: x = tiles (XCD-remapped when xcd_remap == 1), y = batch, z = K split;
// xcd_remap == 2: x = tiles x batch rounded up to 8 entries, y = z = 1.
template <bool BF, bool AC, bool BC, bool AT, bool BT, typename TA, typename TB, typename TCo, int WM, int WN, int TM,
          int TN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_lp_kernel(Gemm g, int tiles_m, int tiles_n, int xcd_remap,
                                                                long T8, int kchunk, double* ws) {
  using C = CfgL<BF, AC, BC, AT, BT, TA, TB, WM, WN, TM, TN>;
  constexpr bool CC = AC || BC;
  __shared__ float As[2][C::PA * BKL * C::LDA];
  __shared__ float Bs[2][C::PB * BKL * C::LDB];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  long tile = blockIdx.x;
  long b = blockIdx.y;
  if (xcd_remap == 1) {
    tile = (tile & 7) * T8 + (tile >> 3);
  } else if (xcd_remap == 2) {
    // batched, 1-D grid: work-group L runs on XCD L % 8; that XCD takes the
    // batch entries b = L % 8 (mod 8), all T8 tiles of one entry in a row, so
    // an operand shared by the tiles of an entry is fetched into one L2
    const long p = tile >> 3;
    b = (tile & 7) + 8 * (p / T8);
    tile = p % T8;
    if (b >= g.batch) return;  // uniform per work-group
  }
  const int mt = (int)(tile % tiles_m);
  const long ntl = tile / tiles_m;
  if (ntl >= tiles_n) return;
  const int m0 = mt * C::BM, n0 = (int)ntl * C::BN;
  const int kz = blockIdx.z;
  const int kbeg = kz * kchunk;
  const int kend = min(g.K, kbeg + kchunk);
  const TA* A = static_cast<const TA*>(g.A) + b * g.sA * (AC ? 2 : 1);
  const TB* B = static_cast<const TB*>(g.B) + b * g.sB * (BC ? 2 : 1);

  v4f accR[TM][TN], accI[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      accR[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
      accI[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    }

  // FP64-stored A (VHS, force bias): raw fetches, rounded when stored to LDS
  constexpr bool RAW = sizeof(TA) == 8;
  using RAt = std::conditional_t<RAW, TA, float>;
  using RBt = std::conditional_t<RAW, TB, float>;
  RAt ra[C::RA][4][2];
  RBt rb[C::RB][4][2];
  auto fetch = [&](int k0) {
    if constexpr (RAW) {
      fetch_a_raw<C, AC, AT, TA>(ra, A, g.lda, g.M, kend, m0, k0, tid, g.a_sym);
      fetch_b_raw<C, BC, BT, TB>(rb, B, g.ldb, g.N, kend, n0, k0, tid);
    } else {
      fetch_a<C, AC, AT, TA>(ra, A, g.lda, g.M, kend, m0, k0, tid, g.a_sym);
      fetch_b<C, BC, BT, TB>(rb, B, g.ldb, g.N, kend, n0, k0, tid);
    }
  };
  auto stash = [&](int b2, int k0) {
    if constexpr (RAW) {
      store_a_raw<C, AC, AT, TA>(As[b2], ra, tid, g.M, kend, m0, k0);
      store_b_raw<C, BC, BT, TB>(Bs[b2], rb, tid, g.N, kend, n0, k0);
    } else {
      store_a<C, AC, AT>(As[b2], ra, tid);
      store_b<C, BC, BT>(Bs[b2], rb, tid);
    }
  };
  fetch(kbeg);
  stash(0, kbeg);
  __syncthreads();
  int buf = 0;
  const int c16 = lane & 15, q4 = lane >> 4;
  for (int k0 = kbeg; k0 < kend; k0 += BKL) {
    const bool more = k0 + BKL < kend;
    if (more) fetch(k0 + BKL);
    const float* as = As[buf];
    const float* bs = Bs[buf];
    if constexpr (!BF) {
#pragma unroll
      for (int kk = 0; kk < BKL; kk += 4) {
        const int kr = kk + q4;
        float ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const int col = (wm * TM + i) * 16 + c16;
          ar[i] = as[kr * C::LDA + col];
          ai[i] = AC ? as[BKL * C::LDA + kr * C::LDA + col] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int col = (wn * TN + j) * 16 + c16;
          br[j] = bs[kr * C::LDB + col];
          bi[j] = BC ? bs[BKL * C::LDB + kr * C::LDB + col] : 0.f;
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], br[j], accR[i][j], 0, 0, 0);
            if (AC && BC) {
              accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai[i], bi[j], accR[i][j], 0, 0, 0);
              accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], bi[j], accI[i][j], 0, 0, 0);
              accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], br[j], accI[i][j], 0, 0, 0);
            } else if (AC) {
              accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], br[j], accI[i][j], 0, 0, 0);
            } else if (BC) {
              accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], bi[j], accI[i][j], 0, 0, 0);
            }
          }
      }
    } else {
      const int kq = 8 * q4;
      BfPair fbr[TN], fbi[TN];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = (wn * TN + j) * 16 + c16;
        fbr[j] = frag_bf(bs, C::LDB, kq, col, false);
        if (BC) fbi[j] = frag_bf(bs + BKL * C::LDB, C::LDB, kq, col, false);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int col = (wm * TM + i) * 16 + c16;
        const BfPair far_ = frag_bf(as, C::LDA, kq, col, false);
        if (AC && BC) {
          const BfPair fai = frag_bf(as + BKL * C::LDA, C::LDA, kq, col, false);
          const BfPair fain = frag_bf(as + BKL * C::LDA, C::LDA, kq, col, true);
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            accR[i][j] = mma3(far_, fbr[j], accR[i][j]);
            accR[i][j] = mma3(fain, fbi[j], accR[i][j]);
            accI[i][j] = mma3(far_, fbi[j], accI[i][j]);
            accI[i][j] = mma3(fai, fbr[j], accI[i][j]);
          }
        } else if (AC) {
          const BfPair fai = frag_bf(as + BKL * C::LDA, C::LDA, kq, col, false);
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            accR[i][j] = mma3(far_, fbr[j], accR[i][j]);
            accI[i][j] = mma3(fai, fbr[j], accI[i][j]);
          }
        } else {
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            accR[i][j] = mma3(far_, fbr[j], accR[i][j]);
            if (BC) accI[i][j] = mma3(far_, fbi[j], accI[i][j]);
          }
        }
      }
    }
    if (more) stash(buf ^ 1, k0 + BKL);
    __syncthreads();
    buf ^= 1;
  }

  if (ws) {  // split-K partial (double) to ws[kz][b][m][n]
    const long plane = (long)g.M * g.N;
    double* Wp = ws + ((long)kz * g.batch + b) * plane * (CC ? 2 : 1);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = m0 + (wm * TM + i) * 16 + 4 * q4 + v;
          const int n = n0 + (wn * TN + j) * 16 + c16;
          if (m >= g.M || n >= g.N) continue;
          const long off = (long)m * g.N + n;
          if (CC)
            reinterpret_cast<double2*>(Wp)[off] = double2{(double)accR[i][j][v], (double)accI[i][j][v]};
          else
            Wp[off] = accR[i][j][v];
        }
    return;
  }
  const bool has_beta = (g.beta.x != 0.0 || g.beta.y != 0.0);
  TCo* Cp = g.C ? static_cast<TCo*>(g.C) + b * g.sC : nullptr;
  double2* Dp = g.D ? static_cast<double2*>(g.D) + b * g.sD : nullptr;
  const double2* Ep = g.E ? static_cast<const double2*>(g.E) + b * g.sE : nullptr;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + (wm * TM + i) * 16 + 4 * q4 + v;
        const int n = n0 + (wn * TN + j) * 16 + c16;
        if (m >= g.M || n >= g.N) continue;
        const long off = (long)m * g.ldc + n;
        double2 r = cmul(g.alpha, double2{(double)accR[i][j][v], (double)accI[i][j][v]});
        if (Cp) {
          if (has_beta) {
            const double2 o = cmul(g.beta, double2{(double)Cp[off].x, (double)Cp[off].y});
            r.x += o.x;
            r.y += o.y;
          }
          if constexpr (sizeof(TCo) == 8)
            Cp[off] = TCo{(float)r.x, (float)r.y};
          else
            Cp[off] = TCo{r.x, r.y};
        }
        if (Dp) {
          double2* Dc = Dp + (long)m * g.ldd + n;
          double2 d = Ep ? Ep[(long)m * g.lde + n] : *Dc;
          d.x += r.x;
          d.y += r.y;
          *Dc = d;
        }
      }
}

template <bool BF, bool AC, bool BC, bool AT, bool BT, typename TA, typename TB, typename TCo, int WM, int WN, int TM,
          int TN>
int launch(const Gemm& g, hipStream_t s) {
  using C = CfgL<BF, AC, BC, AT, BT, TA, TB, WM, WN, TM, TN>;
  const int tm = (g.M + C::BM - 1) / C::BM, tn = (g.N + C::BN - 1) / C::BN;
  const long tiles = (long)tm * tn;
  const bool CC = AC || BC;
  int ksplit = 1;
  const long wgs = tiles * g.batch;
  // split-K partials are reduced into a double2 / double C only
  if (sizeof(TCo) == 16 && g.ws && !g.E && g.batch == 1 && wgs < 1024 && g.K >= 8 * BKL) {
    ksplit = (int)std::min<long>((2048 + wgs - 1) / wgs, (long)g.K / (4 * BKL));
    while (ksplit > 1 && (long)ksplit * g.M * g.N * (CC ? 2 : 1) > g.ws_elems) --ksplit;
  }
  int kchunk = g.K;
  if (ksplit > 1) {
    kchunk = ((g.K + ksplit - 1) / ksplit + BKL - 1) / BKL * BKL;
    ksplit = (g.K + kchunk - 1) / kchunk;
  }
  const int remap = (g.batch == 1 && tiles >= 64) ? 1 : (g.batch > 1 && tiles > 1 && ksplit == 1) ? 2 : 0;
  const long T8 = remap == 1 ? (tiles + 7) / 8 : remap == 2 ? tiles : 0;
  const long gx = remap == 1 ? T8 * 8 : remap == 2 ? tiles * ((g.batch + 7) / 8 * 8) : tiles;
  const unsigned gy = remap == 2 ? 1u : (unsigned)g.batch;
  if (gx > 0x7fffffffL || gy > 65535u) return fail(AFQ_E_UNSUPPORTED, "gemm_lp: grid too large");
  double* ws = ksplit > 1 ? static_cast<double*>(g.ws) : nullptr;
  hipLaunchKernelGGL((gemm_lp_kernel<BF, AC, BC, AT, BT, TA, TB, TCo, WM, WN, TM, TN>),
                     dim3((unsigned)gx, gy, ksplit), dim3(C::NT), 0, s, g, tm, tn, remap, T8, kchunk, ws);
  AFQ_TRY(launch_status("gemm_lp_kernel"));
  if (ksplit > 1) return gemm_splitk_reduce(g, ws, ksplit, CC, s);
  return AFQ_OK;
}

// Split-BF16 x3 walker-panel Taylor product: complex float A (N x N per
// walker, or its packed lower block triangle) times a complex panel B (N x c,
// complex double or float, c <= 16 TN), complex float C = alpha A.B and
// complex double D += C (or D = E + C).  Each of the NW waves owns 16 rows of A
// and all TN column blocks; a lane loads the 8 consecutive k of its row that
// the 16x16x32 bf16 MFMA takes from it straight from global memory and splits
// them in registers, so A never passes through LDS.  The B slab (32 k) is
// staged once per work-group as four bf16 planes (re hi, re lo, im hi, im lo)
// with k contiguous, so a B fragment is one 16-byte LDS read.
constexpr int PROW = 40;  // bf16 per LDS row of a B plane: 32 k + 8 pad (80 bytes: a fragment's 16 rows hit distinct banks)

template <int NW, int TN, typename TB, bool ASYM>
__global__ __launch_bounds__(64 * NW) void panel_bf3_kernel(Gemm g, int tiles_m, int tiles_n) {
  constexpr int BN = 16 * TN, PLANE = BN * PROW;
  constexpr int NE = 32 * BN, RE = (NE + 64 * NW - 1) / (64 * NW);  // B elements per slab
  __shared__ __align__(16) __bf16 Bs[2][4 * PLANE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long T = (long)tiles_m * tiles_n;
  const long L = blockIdx.x;
  const long p = L >> 3;
  const long b = (L & 7) + 8 * (p / T);
  if (b >= g.batch) return;  // uniform per work-group
  const long tile = p % T;
  const int mt = (int)(tile % tiles_m), nt = (int)(tile / tiles_m);
  const int m0 = mt * 16 * NW, n0 = nt * BN;
  const int r = lane & 15, q = lane >> 4;
  const int row = m0 + wave * 16 + r;
  const bool rok = row < g.M;
  const bool active = m0 + wave * 16 < g.M;
  const float2* A = static_cast<const float2*>(g.A) + b * g.sA;
  const TB* B = static_cast<const TB*>(g.B) + b * g.sB * 2;
  const int rowc = min(row, g.M - 1);

  // A is read at clamped (always valid) addresses by every wave and masked by
  // mask_a only when the slab is about to be used, so that no loaded value is
  // consumed inside a divergent branch (which makes the compiler wait for each
  // load there and serialises the slab's loads ahead of the MFMAs).
  auto load_a = [&](float2 (&x)[8], int k0) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = min(k0 + 8 * q + t, g.K - 1);
      x[t] = A[ASYM ? sym_off(rowc, k) : (long)rowc * g.lda + k];
    }
  };
  auto mask_a = [&](float2 (&x)[8], int k0) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const bool ok = rok && k0 + 8 * q + t < g.K;
      x[t].x = ok ? x[t].x : 0.f;
      x[t].y = ok ? x[t].y : 0.f;
    }
  };
  // hi / lo bf16 of the real (im = false) or imaginary parts, negated on request
  auto split8 = [](const float2 (&x)[8], bool im, v8bf& hi, v8bf& lo, bool neg) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const float v = im ? x[t].y : x[t].x;
      const float y = neg ? -v : v;
      const __bf16 h = (__bf16)y;
      hi[t] = h;
      lo[t] = (__bf16)(y - (float)h);
    }
  };
  // B slab rows k0 .. k0+31: element e = (k, n), n fastest; read raw at
  // clamped addresses, masked, rounded and split when stashed
  auto load_b = [&](TB (&x)[RE][2], int k0) {
#pragma unroll
    for (int i = 0; i < RE; ++i) {
      const int e = min(tid + i * 64 * NW, NE - 1);
      const int gk = min(k0 + e / BN, g.K - 1), gn = min(n0 + e % BN, g.N - 1);
      const long off = ((long)gk * g.ldb + gn) * 2;
      x[i][0] = B[off];
      x[i][1] = B[off + 1];
    }
  };
  auto stash_b = [&](__bf16* bs, const TB (&x)[RE][2], int k0) {
#pragma unroll
    for (int i = 0; i < RE; ++i) {
      const int e = tid + i * 64 * NW;
      if (e >= NE) continue;
      const int kk = e / BN, nn = e % BN;
      const bool ok = k0 + kk < g.K && n0 + nn < g.N;
      const int o = nn * PROW + kk;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float y = ok ? (float)x[i][c] : 0.f;
        const __bf16 h = (__bf16)y;
        bs[(2 * c) * PLANE + o] = h;
        bs[(2 * c + 1) * PLANE + o] = (__bf16)(y - (float)h);
      }
    }
  };

  v4f accR[TN], accI[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    accR[j] = v4f{0.f, 0.f, 0.f, 0.f};
    accI[j] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  float2 ac[8], an[8];
  TB rb[RE][2];
  load_a(ac, 0);
  mask_a(ac, 0);
  load_b(rb, 0);
  stash_b(Bs[0], rb, 0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < g.K; k0 += 32) {
    const bool more = k0 + 32 < g.K;
    if (more) {
      load_a(an, k0 + 32);
      load_b(rb, k0 + 32);
    }
    // every wave multiplies (a wave past M has all-zero A and stores nothing):
    // a divergent branch here would make the compiler settle the prefetched
    // A registers, and so wait for their loads, before the MFMAs
    {
      v8bf rh, rl, ih, il, nih, nil;
      split8(ac, false, rh, rl, false);
      split8(ac, true, ih, il, false);
      split8(ac, true, nih, nil, true);
      const __bf16* bs = Bs[buf];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int o = (j * 16 + r) * PROW + 8 * q;
        const v8bf brh = *reinterpret_cast<const v8bf*>(bs + o);
        const v8bf brl = *reinterpret_cast<const v8bf*>(bs + PLANE + o);
        const v8bf bih = *reinterpret_cast<const v8bf*>(bs + 2 * PLANE + o);
        const v8bf bil = *reinterpret_cast<const v8bf*>(bs + 3 * PLANE + o);
        // re += ar br - ai bi ; im += ar bi + ai br, each as hi*hi + lo*hi + hi*lo
        accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh, brh, accR[j], 0, 0, 0);
        accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rl, brh, accR[j], 0, 0, 0);
        accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh, brl, accR[j], 0, 0, 0);
        accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, bih, accR[j], 0, 0, 0);
        accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nil, bih, accR[j], 0, 0, 0);
        accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, bil, accR[j], 0, 0, 0);
        accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh, bih, accI[j], 0, 0, 0);
        accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rl, bih, accI[j], 0, 0, 0);
        accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh, bil, accI[j], 0, 0, 0);
        accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ih, brh, accI[j], 0, 0, 0);
        accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(il, brh, accI[j], 0, 0, 0);
        accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ih, brl, accI[j], 0, 0, 0);
      }
    }
    if (more) {
      stash_b(Bs[buf ^ 1], rb, k0 + 32);
#pragma unroll
      for (int t = 0; t < 8; ++t) ac[t] = an[t];
      mask_a(ac, k0 + 32);
    }
    __syncthreads();
    buf ^= 1;
  }
  if (!active) return;

  // accumulator element v of lane l is (4(l/16) + v, l%16)
  float2* Cp = g.C ? static_cast<float2*>(g.C) + b * g.sC : nullptr;
  double2* Dp = g.D ? static_cast<double2*>(g.D) + b * g.sD : nullptr;
  const double2* Ep = g.E ? static_cast<const double2*>(g.E) + b * g.sE : nullptr;
#pragma unroll
  for (int j = 0; j < TN; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wave * 16 + 4 * q + v;
      const int n = n0 + j * 16 + r;
      if (m >= g.M || n >= g.N) continue;
      const double2 x = cmul(g.alpha, double2{(double)accR[j][v], (double)accI[j][v]});
      if (Cp) Cp[(long)m * g.ldc + n] = float2{(float)x.x, (float)x.y};
      if (Dp) {
        double2* Dc = Dp + (long)m * g.ldd + n;
        double2 d = Ep ? Ep[(long)m * g.lde + n] : *Dc;
        d.x += x.x;
        d.y += x.y;
        *Dc = d;
      }
    }
}

template <int NW, int TN, typename TB>
int launch_panel_bf3(const Gemm& g, hipStream_t s) {
  const int tm = (g.M + 16 * NW - 1) / (16 * NW), tn = (g.N + 16 * TN - 1) / (16 * TN);
  const long gx = (long)tm * tn * ((g.batch + 7) / 8 * 8);
  if (gx > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "gemm_lp: grid too large");
  if (g.a_sym)
    hipLaunchKernelGGL((panel_bf3_kernel<NW, TN, TB, true>), dim3((unsigned)gx), dim3(64 * NW), 0, s, g, tm, tn);
  else
    hipLaunchKernelGGL((panel_bf3_kernel<NW, TN, TB, false>), dim3((unsigned)gx), dim3(64 * NW), 0, s, g, tm, tn);
  return launch_status("panel_bf3_kernel");
}

// FP32 walker-panel product with a real A (the one-body propagator's E - I,
// N x N, one per batch entry or shared with sA = 0) and a complex double panel
// B (N x c, c <= 16 TN): complex double D = E + alpha A.B (D += alpha A.B
// without E), complex float C = alpha A.B when given.  Each of the NW waves
// owns 16 rows of A and all TN column blocks.  In a 32-deep slab, step s of
// the 16x16x4 f32 MFMA takes k = 8q + s from lane (r, q), so a lane reads the 8
// consecutive k of its row straight from global memory; the B slab is staged
// once per work-group as two FP32 planes (re, im) with k contiguous, and a
// lane's 8 values per plane are two 16-byte LDS reads.  As in
// panel_bf3_kernel, loads go to clamped addresses and are masked when used.
constexpr int FROW = 36;  // floats per LDS row of a B plane: 32 k + 4 pad (144 bytes)

template <int NW, int TN>
__global__ __launch_bounds__(64 * NW) void panel_f32_kernel(Gemm g, int tiles_m, int tiles_n) {
  constexpr int BN = 16 * TN, PLANE = BN * FROW;
  constexpr int NE = 32 * BN, RE = (NE + 64 * NW - 1) / (64 * NW);  // B elements per slab
  __shared__ __align__(16) float Bs[2][2 * PLANE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long T = (long)tiles_m * tiles_n;
  const long L = blockIdx.x;
  const long p = L >> 3;
  const long b = (L & 7) + 8 * (p / T);
  if (b >= g.batch) return;  // uniform per work-group
  const long tile = p % T;
  const int mt = (int)(tile % tiles_m), nt = (int)(tile / tiles_m);
  const int m0 = mt * 16 * NW, n0 = nt * BN;
  const int r = lane & 15, q = lane >> 4;
  const int row = m0 + wave * 16 + r;
  const bool rok = row < g.M;
  const bool active = m0 + wave * 16 < g.M;
  const float* A = static_cast<const float*>(g.A) + b * g.sA;
  const double* B = static_cast<const double*>(g.B) + b * g.sB * 2;
  const long arow = (long)min(row, g.M - 1) * g.lda;

  auto load_a = [&](float (&x)[8], int k0) {
#pragma unroll
    for (int t = 0; t < 8; ++t) x[t] = A[arow + min(k0 + 8 * q + t, g.K - 1)];
  };
  auto mask_a = [&](float (&x)[8], int k0) {
#pragma unroll
    for (int t = 0; t < 8; ++t) x[t] = (rok && k0 + 8 * q + t < g.K) ? x[t] : 0.f;
  };
  // B slab rows k0 .. k0+31: element e = (k, n), n fastest
  auto load_b = [&](double (&x)[RE][2], int k0) {
#pragma unroll
    for (int i = 0; i < RE; ++i) {
      const int e = min(tid + i * 64 * NW, NE - 1);
      const int gk = min(k0 + e / BN, g.K - 1), gn = min(n0 + e % BN, g.N - 1);
      const long off = ((long)gk * g.ldb + gn) * 2;
      x[i][0] = B[off];
      x[i][1] = B[off + 1];
    }
  };
  auto stash_b = [&](float* bs, const double (&x)[RE][2], int k0) {
#pragma unroll
    for (int i = 0; i < RE; ++i) {
      const int e = tid + i * 64 * NW;
      if (e >= NE) continue;
      const int kk = e / BN, nn = e % BN;
      const bool ok = k0 + kk < g.K && n0 + nn < g.N;
      bs[nn * FROW + kk] = ok ? (float)x[i][0] : 0.f;
      bs[PLANE + nn * FROW + kk] = ok ? (float)x[i][1] : 0.f;
    }
  };

  v4f accR[TN], accI[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    accR[j] = v4f{0.f, 0.f, 0.f, 0.f};
    accI[j] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  float ac[8], an[8];
  double rb[RE][2];
  load_a(ac, 0);
  mask_a(ac, 0);
  load_b(rb, 0);
  stash_b(Bs[0], rb, 0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < g.K; k0 += 32) {
    const bool more = k0 + 32 < g.K;
    if (more) {
      load_a(an, k0 + 32);
      load_b(rb, k0 + 32);
    }
    // every wave multiplies (a wave past M has all-zero A and stores nothing)
    const float* bs = Bs[buf];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const float* pr = bs + (j * 16 + r) * FROW + 8 * q;
      const float4 r0 = *reinterpret_cast<const float4*>(pr), r1 = *reinterpret_cast<const float4*>(pr + 4);
      const float4 i0 = *reinterpret_cast<const float4*>(pr + PLANE), i1 = *reinterpret_cast<const float4*>(pr + PLANE + 4);
      const float br[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
      const float bi[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        accR[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t], br[t], accR[j], 0, 0, 0);
        accI[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[t], bi[t], accI[j], 0, 0, 0);
      }
    }
    if (more) {
      stash_b(Bs[buf ^ 1], rb, k0 + 32);
#pragma unroll
      for (int t = 0; t < 8; ++t) ac[t] = an[t];
      mask_a(ac, k0 + 32);
    }
    __syncthreads();
    buf ^= 1;
  }
  if (!active) return;

  // accumulator element v of lane l is (4(l/16) + v, l%16)
  float2* Cp = g.C ? static_cast<float2*>(g.C) + b * g.sC : nullptr;
  double2* Dp = g.D ? static_cast<double2*>(g.D) + b * g.sD : nullptr;
  const double2* Ep = g.E ? static_cast<const double2*>(g.E) + b * g.sE : nullptr;
#pragma unroll
  for (int j = 0; j < TN; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wave * 16 + 4 * q + v;
      const int n = n0 + j * 16 + r;
      if (m >= g.M || n >= g.N) continue;
      const double2 x = cmul(g.alpha, double2{(double)accR[j][v], (double)accI[j][v]});
      if (Cp) Cp[(long)m * g.ldc + n] = float2{(float)x.x, (float)x.y};
      if (Dp) {
        double2* Dc = Dp + (long)m * g.ldd + n;
        double2 d = Ep ? Ep[(long)m * g.lde + n] : *Dc;
        d.x += x.x;
        d.y += x.y;
        *Dc = d;
      }
    }
}

template <int NW, int TN>
int launch_panel_f32(const Gemm& g, hipStream_t s) {
  const int tm = (g.M + 16 * NW - 1) / (16 * NW), tn = (g.N + 16 * TN - 1) / (16 * TN);
  const long gx = (long)tm * tn * ((g.batch + 7) / 8 * 8);
  if (gx > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "gemm_lp: grid too large");
  hipLaunchKernelGGL((panel_f32_kernel<NW, TN>), dim3((unsigned)gx), dim3(64 * NW), 0, s, g, tm, tn);
  return launch_status("panel_f32_kernel");
}

template <bool BF, bool AC, bool BC, bool AT, bool BT, typename TA, typename TB, typename TCo>
int dispatch_shape(const Gemm& g, hipStream_t s) {
  if constexpr (!BC && !BT)  // wide real-B tiles (the VHS product), as in the FP64 GEMM
    if (g.batch == 1 && g.N >= 4096 && g.M >= 64 && g.K >= 64)
      return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 2, 8, 2, 2>(g, s);
  if (g.N <= 16) return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 1, 1, 1>(g, s);
  if (g.N <= 32) return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 1, 1, 2>(g, s);
  // walker-panel products (the Taylor terms, N = occupied orbitals in 33..64):
  // more waves per work-group, each with fewer blocks, measured fastest at
  // N = 40 and 60 (bf16x3: 4 x 2 waves of 1 x 2 blocks; FP32: 4 x 3 or 4 x 4
  // waves of one block)
  if (g.M > 64 && g.N <= 64) {
    if constexpr (BF) return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 2, 1, 2>(g, s);
    if (g.N <= 48) return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 3, 1, 1>(g, s);
    return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 4, 1, 1>(g, s);
  }
  // both spins' columns of unrestricted walkers (65..128 columns)
  if (g.M > 64 && g.N <= 128 && g.batch > 1) {
    if constexpr (BF) return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 4, 1, 2>(g, s);
    return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 4, 1, 1>(g, s);
  }
  if (g.N <= 48) return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 1, 1, 3>(g, s);
  if (g.N <= 64 && g.M > 64) return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 4, 1, 1, 4>(g, s);
  return launch<BF, AC, BC, AT, BT, TA, TB, TCo, 2, 2, 2, 2>(g, s);
}

template <bool BF>
int dispatch(LpKind kind, const Gemm& g, hipStream_t s) {
  switch (kind) {
    case LpKind::VHS:  // complex double x real float -> complex float
      return dispatch_shape<BF, true, false, false, false, double, float, float2>(g, s);
    case LpKind::TAYLOR_D:  // complex float x complex double -> complex float (+ D)
      // bf16x3: the register-panel kernel at 33..48 columns (0.081 vs 0.149 ms per
      // first term at N=200, 40 columns, 512 walkers, packed A) and at 65..128
      // columns (0.224 vs 0.329 ms at N=300, 118 columns, 256 walkers)
      if constexpr (BF)
        if (g.batch > 1 && g.M > 64 && !g.ws && !g.a_trans && !g.b_trans) {
          if (g.N > 32 && g.N <= 48) return launch_panel_bf3<16, 3, double>(g, s);
          if (g.N > 64 && g.N <= 128) return launch_panel_bf3<8, 8, double>(g, s);
        }
      return dispatch_shape<BF, true, true, false, false, float, double, float2>(g, s);
    case LpKind::TAYLOR_F:  // complex float x complex float -> complex float (+ D)
      // bf16x3 at 33..48 columns: the register-panel kernel with 16 waves (0.120 vs
      // 0.143 ms per term at N=200, 40 columns, 512 walkers); at 65..128 columns
      // (both spins of unrestricted walkers) with 8 waves of 8 column blocks
      // (0.231 vs 0.273 ms per term at N=300, 118 columns, 256 walkers, packed A)
      if constexpr (BF)
        if (g.batch > 1 && g.M > 64 && !g.ws && !g.a_trans && !g.b_trans) {
          if (g.N > 32 && g.N <= 48) return launch_panel_bf3<16, 3, float>(g, s);
          if (g.N > 64 && g.N <= 128) return launch_panel_bf3<8, 8, float>(g, s);
        }
      return dispatch_shape<BF, true, true, false, false, float, float, float2>(g, s);
    case LpKind::FORCE_BIAS:  // complex double x real float (transposed) -> complex double
      return dispatch_shape<BF, true, false, false, true, double, float, double2>(g, s);
    case LpKind::ONE_BODY:  // real float x complex double -> D = E + result (complex double)
      // 33..64 columns (restricted walkers at O = 40, one spin of unrestricted
      // walkers at O = 60): the FP32 register-panel kernel
      if (g.batch > 1 && g.M > 64 && g.N > 32 && g.N <= 64 && !g.ws && !g.a_trans && !g.b_trans && !g.a_sym)
        return g.N <= 48 ? launch_panel_f32<8, 3>(g, s) : launch_panel_f32<8, 4>(g, s);
      return dispatch_shape<false, false, true, false, false, float, double, double2>(g, s);
  }
  return fail(AFQ_E_INVALID, "gemm_lp: unknown kind");
}

}  // namespace

int gemm_lp(LpKind kind, bool bf16x3, const Gemm& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return AFQ_OK;
  if (g.K <= 0) return fail(AFQ_E_INVALID, "gemm_lp: K must be positive");
  if (g.E && !g.D) return fail(AFQ_E_INVALID, "gemm_lp: an addend E needs an output D");
  if (g.a_sym && (g.a_trans || g.a_sym != g.M || g.a_sym != g.K))
    return fail(AFQ_E_INVALID, "gemm_lp: a packed symmetric A is square and not transposed");
  if (!g.C && !g.D) return fail(AFQ_E_INVALID, "gemm_lp: no output");
  return bf16x3 ? dispatch<true>(kind, g, s) : dispatch<false>(kind, g, s);
}

}  // namespace afq
