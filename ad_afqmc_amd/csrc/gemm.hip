// Batched FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// A work-group of WM x WN waves computes a BM x BN tile of C, each wave a
// TM x TN grid of 16x16 MFMA blocks (BM = 16 TM WM, BN = 16 TN WN).  K advances
// in slabs of 16 through two LDS buffers: the next slab is fetched from global
// memory into registers while the current one is multiplied, then stored to
// the other buffer (one barrier per slab).  Complex operands are split into
// real/imaginary LDS planes; complex x real costs two real MFMAs per block,
// complex x complex four, all accumulated in FP64.
//
// LDS rows are padded so their stride is 16 (mod 32) doubles: the four k rows
// read by one MFMA fragment then fall in disjoint bank halves.
//
// Two tile shapes: 64x64 (2x2 waves of 2x2 blocks) for large products, and
// "panel" tiles 64 x 16*TN (4x1 waves of 1 x TN blocks) for products whose N is
// a walker's occupied-orbital count, so at most 15 padded columns are wasted.
// Large single products are ordered XCD-aware (work-groups sharing a B panel
// land on the same XCD) and may be split along K into a workspace that a
// second kernel reduces in a fixed order.
//
// MFMA f64 16x16x4 lane layout (gfx950): lane l feeds A[l%16][l/16] and
// B[l/16][l%16]; accumulator element v of lane l is D[l/16 + 4v][l%16].
#include "afq_internal.h"
#include "group.h"

namespace afq {
namespace {

constexpr int BK = 16;

__device__ inline double2 cmul(double2 a, double2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

constexpr int padded(int w) { return ((w % 32) == 16) ? w : w + ((48 - (w % 32)) % 32); }

template <bool AC, bool BC, bool AT, bool BT, int WM, int WN, int TM, int TN>
struct Cfg {
  static constexpr int NT = 64 * WM * WN;
  static constexpr int BM = 16 * TM * WM, BN = 16 * TN * WN;
  static constexpr int LDA = padded(BM), LDB = padded(BN);
  static constexpr int PA = AC ? 2 : 1, PB = BC ? 2 : 1;
  // quads (4 consecutive elements along the contiguous dimension) per tile
  static constexpr int QA = BM * BK / 4, QB = BN * BK / 4;
  static constexpr int RA = (QA + NT - 1) / NT, RB = (QB + NT - 1) / NT;
};

// Loads quad q of the A tile (rows m0.., k0..) into registers.
template <class C, bool AC, bool AT>
__device__ inline void fetch_a(double (&r)[C::RA][4][2], const double* A, long lda, int M, int K, int m0, int k0,
                               int tid, int asym) {
#pragma unroll
  for (int i = 0; i < C::RA; ++i) {
    const int q = tid + i * C::NT;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      r[i][t][0] = 0.0;
      r[i][t][1] = 0.0;
    }
    if (q >= C::QA) continue;
    int m, k;
    if (!AT) {  // k contiguous: quad = 4 consecutive k of row m
      m = q % C::BM;
      k = (q / C::BM) * 4;
    } else {  // m contiguous: quad = 4 consecutive m of column k
      m = (q % (C::BM / 4)) * 4;
      k = q / (C::BM / 4);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int gm = m0 + m + (AT ? t : 0), gk = k0 + k + (AT ? 0 : t);
      if (gm < M && gk < K) {
        const long off = AT ? (long)gk * lda + gm : asym ? sym_off(gm, gk) : (long)gm * lda + gk;
        if (AC) {
          const double2 v = reinterpret_cast<const double2*>(A)[off];
          r[i][t][0] = v.x;
          r[i][t][1] = v.y;
        } else {
          r[i][t][0] = A[off];
        }
      }
    }
  }
}

template <class C, bool AC, bool AT>
__device__ inline void store_a(double* As, const double (&r)[C::RA][4][2], int tid) {
#pragma unroll
  for (int i = 0; i < C::RA; ++i) {
    const int q = tid + i * C::NT;
    if (q >= C::QA) continue;
    int m, k;
    if (!AT) {
      m = q % C::BM;
      k = (q / C::BM) * 4;
    } else {
      m = (q % (C::BM / 4)) * 4;
      k = q / (C::BM / 4);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int mm = m + (AT ? t : 0), kk = k + (AT ? 0 : t);
      As[kk * C::LDA + mm] = r[i][t][0];
      if (AC) As[BK * C::LDA + kk * C::LDA + mm] = r[i][t][1];
    }
  }
}

template <class C, bool BC, bool BT>
__device__ inline void fetch_b(double (&r)[C::RB][4][2], const double* B, long ldb, int N, int K, int n0, int k0,
                               int tid) {
#pragma unroll
  for (int i = 0; i < C::RB; ++i) {
    const int q = tid + i * C::NT;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      r[i][t][0] = 0.0;
      r[i][t][1] = 0.0;
    }
    if (q >= C::QB) continue;
    int n, k;
    if (!BT) {  // n contiguous: quad = 4 consecutive n of row k
      n = (q % (C::BN / 4)) * 4;
      k = q / (C::BN / 4);
    } else {  // k contiguous: quad = 4 consecutive k of column n
      n = q % C::BN;
      k = (q / C::BN) * 4;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int gn = n0 + n + (BT ? 0 : t), gk = k0 + k + (BT ? t : 0);
      if (gn < N && gk < K) {
        const long off = BT ? (long)gn * ldb + gk : (long)gk * ldb + gn;
        if (BC) {
          const double2 v = reinterpret_cast<const double2*>(B)[off];
          r[i][t][0] = v.x;
          r[i][t][1] = v.y;
        } else {
          r[i][t][0] = B[off];
        }
      }
    }
  }
}

template <class C, bool BC, bool BT>
__device__ inline void store_b(double* Bs, const double (&r)[C::RB][4][2], int tid) {
#pragma unroll
  for (int i = 0; i < C::RB; ++i) {
    const int q = tid + i * C::NT;
    if (q >= C::QB) continue;
    int n, k;
    if (!BT) {
      n = (q % (C::BN / 4)) * 4;
      k = q / (C::BN / 4);
    } else {
      n = q % C::BN;
      k = (q / C::BN) * 4;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int nn = n + (BT ? 0 : t), kk = k + (BT ? t : 0);
      Bs[kk * C::LDB + nn] = r[i][t][0];
      if (BC) Bs[BK * C::LDB + kk * C::LDB + nn] = r[i][t][1];
    }
  }
}

// One work-group's tile `tile` of batch entry b, K slice kz (the whole
// product when kchunk >= K and ws is null).
// PART (complex A and B only): 0 complex C; 1 real C from Re(A.B); 2 real C from Im(A.B)
template <bool AC, bool BC, bool AT, bool BT, int WM, int WN, int TM, int TN, int PART = 0>
__device__ __forceinline__ void gemm_f64_body(const Gemm& g, int tiles_m, int tiles_n, long tile, long b, int kz,
                                              int kchunk, double* ws) {
  static_assert(PART == 0 || (AC && BC), "a real part of a complex product needs complex A and B");
  using C = Cfg<AC, BC, AT, BT, WM, WN, TM, TN>;
  constexpr bool CC = AC || BC;
  __shared__ double As[2][C::PA * BK * C::LDA];
  __shared__ double Bs[2][C::PB * BK * C::LDB];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int mt = (int)(tile % tiles_m);
  const long ntl = tile / tiles_m;
  if (ntl >= tiles_n) return;  // uniform per work-group
  const int m0 = mt * C::BM, n0 = (int)ntl * C::BN;
  const int kbeg = kz * kchunk;
  const int kend = min(g.K, kbeg + kchunk);
  const double* A = static_cast<const double*>(g.A) + b * g.sA * (AC ? 2 : 1);
  const double* B = static_cast<const double*>(g.B) + b * g.sB * (BC ? 2 : 1);

  v4d accR[TM][TN], accI[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      accR[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    }
  const double asg = g.a_conj ? -1.0 : 1.0;  // sign of A's imaginary part

  double ra[C::RA][4][2], rb[C::RB][4][2];
  // K is processed as [kbeg, kend); the fetch helpers clip against kend.
  fetch_a<C, AC, AT>(ra, A, g.lda, g.M, kend, m0, kbeg, tid, g.a_sym);
  fetch_b<C, BC, BT>(rb, B, g.ldb, g.N, kend, n0, kbeg, tid);
  store_a<C, AC, AT>(As[0], ra, tid);
  store_b<C, BC, BT>(Bs[0], rb, tid);
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) {
      fetch_a<C, AC, AT>(ra, A, g.lda, g.M, kend, m0, k0 + BK, tid, g.a_sym);
      fetch_b<C, BC, BT>(rb, B, g.ldb, g.N, kend, n0, k0 + BK, tid);
    }
    const double* as = As[buf];
    const double* bs = Bs[buf];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      double ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int col = (wm * TM + i) * 16 + c;
        ar[i] = as[kr * C::LDA + col];
        ai[i] = AC ? asg * as[BK * C::LDA + kr * C::LDA + col] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = (wn * TN + j) * 16 + c;
        br[j] = bs[kr * C::LDB + col];
        bi[j] = BC ? bs[BK * C::LDB + kr * C::LDB + col] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          if (PART != 2) accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accR[i][j], 0, 0, 0);
          if (AC && BC) {
            if (PART != 2) accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[i], bi[j], accR[i][j], 0, 0, 0);
            if (PART != 1) {
              accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], accI[i][j], 0, 0, 0);
              accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], accI[i][j], 0, 0, 0);
            }
          } else if (AC) {
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], accI[i][j], 0, 0, 0);
          } else if (BC) {
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], accI[i][j], 0, 0, 0);
          }
        }
    }
    if (more) {
      store_a<C, AC, AT>(As[buf ^ 1], ra, tid);
      store_b<C, BC, BT>(Bs[buf ^ 1], rb, tid);
    }
    __syncthreads();
    buf ^= 1;
  }

  // ---- epilogue
  if (ws) {  // split-K partial: raw accumulators to ws[kz][b][m][n]
    const long plane = (long)g.M * g.N;
    double* W = ws + ((long)kz * g.batch + b) * plane * (CC ? 2 : 1);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = m0 + (wm * TM + i) * 16 + (lane >> 4) + 4 * v;
          const int n = n0 + (wn * TN + j) * 16 + (lane & 15);
          if (m >= g.M || n >= g.N) continue;
          const long off = (long)m * g.N + n;
          if (CC)
            reinterpret_cast<double2*>(W)[off] = double2{accR[i][j][v], accI[i][j][v]};
          else
            W[off] = accR[i][j][v];
        }
    return;
  }
  const bool has_beta = (g.beta.x != 0.0 || g.beta.y != 0.0);
  if constexpr (PART != 0) {  // real C (no D, no split K)
    double* Cr = static_cast<double*>(g.C) + b * g.sC;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = m0 + (wm * TM + i) * 16 + (lane >> 4) + 4 * v;
          const int n = n0 + (wn * TN + j) * 16 + (lane & 15);
          if (m >= g.M || n >= g.N) continue;
          const long off = (long)m * g.ldc + n;
          double r = PART == 1 ? g.alpha.x * accR[i][j][v] : -g.alpha.y * accI[i][j][v];
          if (has_beta) r += g.beta.x * Cr[off];
          Cr[off] = r;
        }
    return;
  }
  double* Cp = static_cast<double*>(g.C) + b * g.sC * (CC ? 2 : 1);
  double* Dp = g.D ? static_cast<double*>(g.D) + b * g.sD * (CC ? 2 : 1) : nullptr;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + (wm * TM + i) * 16 + (lane >> 4) + 4 * v;
        const int n = n0 + (wn * TN + j) * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const long off = (long)m * g.ldc + n;
        if (CC) {
          double2 r = cmul(g.alpha, double2{accR[i][j][v], accI[i][j][v]});
          double2* Cc = reinterpret_cast<double2*>(Cp);
          if (has_beta) {
            const double2 o = cmul(g.beta, Cc[off]);
            r.x += o.x;
            r.y += o.y;
          }
          Cc[off] = r;
          if (Dp) {
            double2* Dc = reinterpret_cast<double2*>(Dp) + (long)m * g.ldd + n;
            double2 d = *Dc;
            d.x += r.x;
            d.y += r.y;
            *Dc = d;
          }
        } else {
          double r = g.alpha.x * accR[i][j][v];
          if (has_beta) r += g.beta.x * Cp[off];
          Cp[off] = r;
          if (Dp) Dp[(long)m * g.ldd + n] += r;
        }
      }
}

// grid: x = tiles (XCD-remapped when xcd_remap == 1), y = batch, z = K split;
// xcd_remap == 2: x = tiles x batch rounded up to 8 entries, y = z = 1.
template <bool AC, bool BC, bool AT, bool BT, int WM, int WN, int TM, int TN, int PART = 0>
__global__ __launch_bounds__(64 * WM * WN) void gemm_f64_kernel(Gemm g, int tiles_m, int tiles_n, int xcd_remap,
                                                                 long T8, int kchunk, double* ws) {
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
  gemm_f64_body<AC, BC, AT, BT, WM, WN, TM, TN, PART>(g, tiles_m, tiles_n, tile, b, (int)blockIdx.z, kchunk, ws);
}

// Two batched products with the same operand form and tile configuration in
// one launch (both spins of a step): work-groups [0, items0) run g0 and the
// rest g1, each in gemm_f64_kernel's batched XCD-aware order (items0 is a
// multiple of 8, so the order holds for g1 too).  No K split.
template <bool AC, bool BC, bool AT, bool BT, int WM, int WN, int TM, int TN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_f64_pair_kernel(Gemm g0, Gemm g1, int tm0, int tn0, int tm1,
                                                                      int tn1, long items0) {
  const long item = blockIdx.x;
  const bool second = item >= items0;  // uniform per work-group
  const Gemm g = second ? g1 : g0;
  const int tm = second ? tm1 : tm0, tn = second ? tn1 : tn0;
  const long local = second ? item - items0 : item;
  const long tiles = (long)tm * tn;
  const long p = local >> 3;
  const long b = (local & 7) + 8 * (p / tiles);
  if (b >= g.batch) return;  // uniform per work-group
  gemm_f64_body<AC, BC, AT, BT, WM, WN, TM, TN>(g, tm, tn, p % tiles, b, 0, g.K, nullptr);
}

// Grouped batched products: work-group i belongs to the job j with
// prefix[j] <= i < prefix[j + 1] (a multiple of 8); within it the local index
// is mapped as gemm_f64_kernel's xcd_remap == 2: XCD local % 8 takes the batch
// entries local % 8 (mod 8), all tiles of one entry in a row.  No K split.
template <bool AC, bool BC, bool AT, bool BT>
__global__ __launch_bounds__(256) void gemm_f64_grouped_kernel(const GemmJob* __restrict__ jobs,
                                                               const int* __restrict__ prefix, int njobs) {
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const Gemm g = jobs[j].g;
  const int tm = jobs[j].tiles_m, tn = jobs[j].tiles_n;
  const int local = item - prefix[j];
  const int tiles = tm * tn;
  const int p = local >> 3;
  const int b = (local & 7) + 8 * (p / tiles);
  if (b >= g.batch) return;  // uniform per work-group
  gemm_f64_body<AC, BC, AT, BT, 2, 2, 2, 2>(g, tm, tn, p % tiles, b, 0, g.K, nullptr);
}

// C = alpha * sum_kz ws[kz] + beta * C  (fixed summation order)
template <bool CC>
__global__ void splitk_reduce_kernel(Gemm g, const double* ws, int ksplit) {
  const long plane = (long)g.M * g.N;
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long b = blockIdx.y;
  if (idx >= plane) return;
  const int m = (int)(idx / g.N), n = (int)(idx % g.N);
  const bool has_beta = (g.beta.x != 0.0 || g.beta.y != 0.0);
  if (CC) {
    double2 s = {0.0, 0.0};
    for (int z = 0; z < ksplit; ++z) {
      const double2 v = reinterpret_cast<const double2*>(ws)[((long)z * gridDim.y + b) * plane + idx];
      s.x += v.x;
      s.y += v.y;
    }
    double2 r = cmul(g.alpha, s);
    double2* Cc = reinterpret_cast<double2*>(static_cast<double*>(g.C) + b * g.sC * 2) + (long)m * g.ldc + n;
    if (has_beta) {
      const double2 o = cmul(g.beta, *Cc);
      r.x += o.x;
      r.y += o.y;
    }
    *Cc = r;
    if (g.D) {
      double2* Dc = reinterpret_cast<double2*>(static_cast<double*>(g.D) + b * g.sD * 2) + (long)m * g.ldd + n;
      Dc->x += r.x;
      Dc->y += r.y;
    }
  } else {
    double s = 0.0;
    for (int z = 0; z < ksplit; ++z) s += ws[((long)z * gridDim.y + b) * plane + idx];
    double r = g.alpha.x * s;
    double* Cr = static_cast<double*>(g.C) + b * g.sC + (long)m * g.ldc + n;
    if (has_beta) r += g.beta.x * *Cr;
    *Cr = r;
    if (g.D) static_cast<double*>(g.D)[b * g.sD + (long)m * g.ldd + n] += r;
  }
}

template <bool AC, bool BC, bool AT, bool BT, int WM, int WN, int TM, int TN, int PART = 0>
int launch(const Gemm& g, hipStream_t s) {
  using C = Cfg<AC, BC, AT, BT, WM, WN, TM, TN>;
  const int tm = (g.M + C::BM - 1) / C::BM, tn = (g.N + C::BN - 1) / C::BN;
  const long tiles = (long)tm * tn;
  int ksplit = 1;
  const long wgs = tiles * g.batch;
  if (PART == 0 && g.ws && g.batch == 1 && wgs < 1024 && g.K >= 8 * BK) {
    ksplit = (int)std::min<long>((2048 + wgs - 1) / wgs, (long)g.K / (4 * BK));
    const long need = (long)ksplit * g.M * g.N * ((AC || BC) ? 2 : 1);
    while (ksplit > 1 && need > g.ws_elems) {
      --ksplit;
    }
    if ((long)ksplit * g.M * g.N * ((AC || BC) ? 2 : 1) > g.ws_elems) ksplit = 1;
  }
  int kchunk = g.K;
  if (ksplit > 1) {
    kchunk = ((g.K + ksplit - 1) / ksplit + BK - 1) / BK * BK;
    ksplit = (g.K + kchunk - 1) / kchunk;
  }
  const int remap = (g.batch == 1 && tiles >= 64) ? 1 : (g.batch > 1 && tiles > 1 && ksplit == 1) ? 2 : 0;
  const long T8 = remap == 1 ? (tiles + 7) / 8 : remap == 2 ? tiles : 0;
  const long gx = remap == 1 ? T8 * 8 : remap == 2 ? tiles * ((g.batch + 7) / 8 * 8) : tiles;
  const unsigned gy = remap == 2 ? 1u : (unsigned)g.batch;
  if (gx > 0x7fffffffL || gy > 65535u || ksplit > 65535) return fail(AFQ_E_UNSUPPORTED, "gemm_f64: grid too large");
  dim3 grid((unsigned)gx, gy, ksplit);
  double* ws = ksplit > 1 ? static_cast<double*>(g.ws) : nullptr;
  hipLaunchKernelGGL((gemm_f64_kernel<AC, BC, AT, BT, WM, WN, TM, TN, PART>), grid, dim3(C::NT), 0, s, g, tm, tn, remap,
                     T8, kchunk, ws);
  AFQ_TRY(launch_status("gemm_f64_kernel"));
  if (ksplit > 1) return gemm_splitk_reduce(g, ws, ksplit, AC || BC, s);
  return AFQ_OK;
}

// Walker-panel products (A: N x N per walker or shared, B: N x c walker
// columns, c <= 16 TN per column tile): each wave owns 16 rows of A and all
// TN column blocks, so A is never shared between waves and goes straight from
// global memory to registers in MFMA layout, one slab ahead; only the B panel
// is staged in LDS (two buffers, one barrier per slab).  Within a 16-deep slab
// lane (r, q) holds A[r][k0 + 4q + s] for MFMA step s, so each lane reads 4
// consecutive k of its row, and B rows are read at the same permuted k.  With
// A out of LDS a work-group needs a third of the LDS of the staged tiles,
// which doubles the work-groups (and loads in flight) per CU.
// grid: one dimension, tiles x batch rounded up to 8 entries, walkers mapped
// to XCDs as in gemm_f64_kernel's batched order.
template <bool AC, int NW, int TN>
__global__ __launch_bounds__(64 * NW) void panel_f64_kernel(Gemm g, int tiles_m, int tiles_n) {
  using C = Cfg<AC, true, false, false, NW, 1, 1, TN>;
  __shared__ double Bs[2][2 * BK * C::LDB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long T = (long)tiles_m * tiles_n;
  const long L = blockIdx.x;
  const long p = L >> 3;
  const long b = (L & 7) + 8 * (p / T);
  if (b >= g.batch) return;  // uniform per work-group
  const long tile = p % T;
  const int mt = (int)(tile % tiles_m), nt = (int)(tile / tiles_m);
  const int m0 = mt * 16 * NW, n0 = nt * C::BN;
  const int r = lane & 15, q = lane >> 4;
  const int row = m0 + wave * 16 + r;
  const bool rok = row < g.M;
  // a wave whose 16 rows all lie past M only helps stage B
  const bool active = m0 + wave * 16 < g.M;
  const double* A = static_cast<const double*>(g.A) + b * g.sA * (AC ? 2 : 1);
  const double* B = static_cast<const double*>(g.B) + b * g.sB * 2;

  auto load_a = [&](double (&x)[4][2], int k0) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = k0 + 4 * q + t;
      x[t][0] = x[t][1] = 0.0;
      if (rok && k < g.K) {
        const long off = g.a_sym ? sym_off(row, k) : (long)row * g.lda + k;
        if (AC) {
          const double2 v = reinterpret_cast<const double2*>(A)[off];
          x[t][0] = v.x;
          x[t][1] = v.y;
        } else {
          x[t][0] = A[off];
        }
      }
    }
  };

  v4d accR[TN], accI[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    accR[j] = v4d{0.0, 0.0, 0.0, 0.0};
    accI[j] = v4d{0.0, 0.0, 0.0, 0.0};
  }
  double ac[4][2], an[4][2], rb[C::RB][4][2];
  if (active) load_a(ac, 0);
  fetch_b<C, true, false>(rb, B, g.ldb, g.N, g.K, n0, 0, tid);
  store_b<C, true, false>(Bs[0], rb, tid);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < g.K; k0 += BK) {
    const bool more = k0 + BK < g.K;
    if (more) {
      if (active) load_a(an, k0 + BK);
      fetch_b<C, true, false>(rb, B, g.ldb, g.N, g.K, n0, k0 + BK, tid);
    }
    const double* bs = Bs[buf];
#pragma unroll
    for (int s = 0; s < 4 && active; ++s) {
      const int kr = 4 * q + s;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = j * 16 + r;
        const double br = bs[kr * C::LDB + col], bi = bs[BK * C::LDB + kr * C::LDB + col];
        accR[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[s][0], br, accR[j], 0, 0, 0);
        accI[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[s][0], bi, accI[j], 0, 0, 0);
        if (AC) {
          accR[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ac[s][1], bi, accR[j], 0, 0, 0);
          accI[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[s][1], br, accI[j], 0, 0, 0);
        }
      }
    }
    if (more) {
      store_b<C, true, false>(Bs[buf ^ 1], rb, tid);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        ac[t][0] = an[t][0];
        ac[t][1] = an[t][1];
      }
    }
    __syncthreads();
    buf ^= 1;
  }

  if (!active) return;
  const bool has_beta = (g.beta.x != 0.0 || g.beta.y != 0.0);
  double2* Cp = reinterpret_cast<double2*>(static_cast<double*>(g.C) + b * g.sC * 2);
  double2* Dp = g.D ? reinterpret_cast<double2*>(static_cast<double*>(g.D) + b * g.sD * 2) : nullptr;
#pragma unroll
  for (int j = 0; j < TN; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wave * 16 + q + 4 * v;
      const int n = n0 + j * 16 + r;
      if (m >= g.M || n >= g.N) continue;
      const long off = (long)m * g.ldc + n;
      double2 x = cmul(g.alpha, double2{accR[j][v], accI[j][v]});
      if (has_beta) {
        const double2 o = cmul(g.beta, Cp[off]);
        x.x += o.x;
        x.y += o.y;
      }
      Cp[off] = x;
      if (Dp) {
        double2* Dc = Dp + (long)m * g.ldd + n;
        double2 d = *Dc;
        d.x += x.x;
        d.y += x.y;
        *Dc = d;
      }
    }
}

template <bool AC, int NW, int TN>
int launch_panel(const Gemm& g, hipStream_t s) {
  const int tm = (g.M + 16 * NW - 1) / (16 * NW), tn = (g.N + 16 * TN - 1) / (16 * TN);
  const long gx = (long)tm * tn * ((g.batch + 7) / 8 * 8);
  if (gx > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "gemm_f64: grid too large");
  hipLaunchKernelGGL((panel_f64_kernel<AC, NW, TN>), dim3((unsigned)gx), dim3(64 * NW), 0, s, g, tm, tn);
  return launch_status("panel_f64_kernel");
}

// batched_config (below) restates these choices for gemm_f64_pair: a change
// here needs the same change there.
template <bool AC, bool BC, bool AT, bool BT>
int dispatch_shape(const Gemm& g, hipStream_t s) {
  // Large unbatched products with a wide N (the VHS product, N = norb^2): 16 waves
  // of 2x2 blocks (64 x 256 tiles) halve the operand traffic per flop of 64 x 64
  // tiles and keep 16 waves in flight per work-group (measured 1.73 ms vs 2.31 ms
  // at W=512, G=1200, N=200).
  if constexpr (!BC && !BT)  // (a complex B tile of this width would not fit in LDS)
    if (g.batch == 1 && g.N >= 4096 && g.M >= 64 && g.K >= 64) return launch<AC, BC, AT, BT, 2, 8, 2, 2>(g, s);
  // walker panels of 33..48 columns against a complex panel (the Taylor terms and
  // exp(-dt/2 h1) at O = 40): A straight to registers, 8 waves of 16 rows
  // (0.243 vs 0.266 ms per Taylor term, 0.108 vs 0.130 ms per exp_h1 product at
  // N=200, 512 walkers); wider panels stay on the LDS-staged tiles, which were
  // faster at 118 columns
  if constexpr (!AT && !BT && BC)
    if (g.batch > 1 && g.N > 32 && g.N <= 48 && g.M > 64 && !g.ws && !g.a_conj) return launch_panel<AC, 8, 3>(g, s);
  // panel tiles for N <= 64 (walker-column products)
  if (g.N <= 16) return launch<AC, BC, AT, BT, 4, 1, 1, 1>(g, s);
  if (g.N <= 32) return launch<AC, BC, AT, BT, 4, 1, 1, 2>(g, s);
  if (g.N <= 48) return launch<AC, BC, AT, BT, 4, 1, 1, 3>(g, s);
  if (g.N <= 64 && g.M > 64) return launch<AC, BC, AT, BT, 4, 1, 1, 4>(g, s);
  // walker panels of both spins (65..128 columns): 16 waves of 1 x 2 blocks
  // (0.47 ms vs 0.52 ms for 2 x 2 x 2 x 2 at N=300, 118 columns, 256 walkers)
  if (g.N <= 128 && g.M > 64 && g.batch > 1) return launch<AC, BC, AT, BT, 4, 4, 1, 2>(g, s);
  return launch<AC, BC, AT, BT, 2, 2, 2, 2>(g, s);
}

template <bool AC, bool BC>
int dispatch_t(const Gemm& g, hipStream_t s) {
  if (g.a_trans) return g.b_trans ? dispatch_shape<AC, BC, true, true>(g, s) : dispatch_shape<AC, BC, true, false>(g, s);
  return g.b_trans ? dispatch_shape<AC, BC, false, true>(g, s) : dispatch_shape<AC, BC, false, false>(g, s);
}

// The tile configuration dispatch_shape picks for a batched product without
// K split or packed A: 0 = 2x2x2x2, 1..4 = 4x1x1xTN, 5 = 4x4x1x2,
// -1 = a shape served otherwise (the wide VHS tiles, the walker-panel kernel).
int batched_config(const Gemm& g) {
  if (!g.b_cplx && !g.b_trans && g.batch == 1 && g.N >= 4096 && g.M >= 64 && g.K >= 64) return -1;
  if (!g.a_trans && !g.b_trans && g.b_cplx && g.batch > 1 && g.N > 32 && g.N <= 48 && g.M > 64 && !g.a_conj)
    return -1;
  if (g.N <= 16) return 1;
  if (g.N <= 32) return 2;
  if (g.N <= 48) return 3;
  if (g.N <= 64 && g.M > 64) return 4;
  if (g.N <= 128 && g.M > 64 && g.batch > 1) return 5;
  return 0;
}

template <bool AC, bool BC, bool AT, bool BT, int WM, int WN, int TM, int TN>
int launch_pair(const Gemm& g0, const Gemm& g1, hipStream_t s) {
  using C = Cfg<AC, BC, AT, BT, WM, WN, TM, TN>;
  const int tm0 = (g0.M + C::BM - 1) / C::BM, tn0 = (g0.N + C::BN - 1) / C::BN;
  const int tm1 = (g1.M + C::BM - 1) / C::BM, tn1 = (g1.N + C::BN - 1) / C::BN;
  const long items0 = (long)tm0 * tn0 * ((g0.batch + 7) / 8 * 8);
  const long items1 = (long)tm1 * tn1 * ((g1.batch + 7) / 8 * 8);
  if (items0 + items1 > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "gemm_f64_pair: grid too large");
  hipLaunchKernelGGL((gemm_f64_pair_kernel<AC, BC, AT, BT, WM, WN, TM, TN>), dim3((unsigned)(items0 + items1)),
                     dim3(C::NT), 0, s, g0, g1, tm0, tn0, tm1, tn1, items0);
  return launch_status("gemm_f64_pair_kernel");
}

template <bool AC, bool BC, bool AT, bool BT>
int dispatch_pair(const Gemm& g0, const Gemm& g1, int cfg, hipStream_t s) {
  switch (cfg) {
    case 1: return launch_pair<AC, BC, AT, BT, 4, 1, 1, 1>(g0, g1, s);
    case 2: return launch_pair<AC, BC, AT, BT, 4, 1, 1, 2>(g0, g1, s);
    case 3: return launch_pair<AC, BC, AT, BT, 4, 1, 1, 3>(g0, g1, s);
    case 4: return launch_pair<AC, BC, AT, BT, 4, 1, 1, 4>(g0, g1, s);
    case 5: return launch_pair<AC, BC, AT, BT, 4, 4, 1, 2>(g0, g1, s);
    default: return launch_pair<AC, BC, AT, BT, 2, 2, 2, 2>(g0, g1, s);
  }
}

}  // namespace

int gemm_splitk_reduce(const Gemm& g, const double* ws, int ksplit, bool cplx, hipStream_t s) {
  const long plane = (long)g.M * g.N;
  dim3 rg((unsigned)((plane + 255) / 256), g.batch);
  if (cplx)
    hipLaunchKernelGGL((splitk_reduce_kernel<true>), rg, dim3(256), 0, s, g, ws, ksplit);
  else
    hipLaunchKernelGGL((splitk_reduce_kernel<false>), rg, dim3(256), 0, s, g, ws, ksplit);
  return launch_status("splitk_reduce_kernel");
}

GemmJob gemm_group_job(const Gemm& g) {
  GemmJob j;
  j.g = g;
  j.tiles_m = (g.M + 63) / 64;
  j.tiles_n = (g.N + 63) / 64;
  return j;
}

long gemm_group_items(const GemmJob& j) { return (long)j.tiles_m * j.tiles_n * ((j.g.batch + 7) / 8 * 8); }

int gemm_f64_grouped(const GroupTab<GemmJob>& t, bool a_cplx, bool b_cplx, bool a_trans, bool b_trans,
                     hipStream_t s) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  const dim3 grid((unsigned)t.total), blk(256);
#define GG(AC, BC, AT, BT) \
  hipLaunchKernelGGL((gemm_f64_grouped_kernel<AC, BC, AT, BT>), grid, blk, 0, s, t.params, t.prefix, t.njobs)
  if (!a_cplx && b_cplx && a_trans && !b_trans)
    GG(false, true, true, false);
  else if (a_cplx && b_cplx && a_trans && b_trans)
    GG(true, true, true, true);
  else if (a_cplx && b_cplx && a_trans && !b_trans)
    GG(true, true, true, false);
  else if (a_cplx && b_cplx && !a_trans && !b_trans)
    GG(true, true, false, false);
  else
    return fail(AFQ_E_UNSUPPORTED, "gemm_f64_grouped: operand form");
#undef GG
  return launch_status("gemm_f64_grouped_kernel");
}

// Real C from one part of a complex product (64 x 64 tiles, no K split).
template <int PART>
int dispatch_part(const Gemm& g, hipStream_t s) {
  if (g.a_trans)
    return g.b_trans ? launch<true, true, true, true, 2, 2, 2, 2, PART>(g, s)
                     : launch<true, true, true, false, 2, 2, 2, 2, PART>(g, s);
  return g.b_trans ? launch<true, true, false, true, 2, 2, 2, 2, PART>(g, s)
                   : launch<true, true, false, false, 2, 2, 2, 2, PART>(g, s);
}

int gemm_f64(const Gemm& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return AFQ_OK;
  if (g.K <= 0) return fail(AFQ_E_INVALID, "gemm_f64: K must be positive");
  if (g.E) return fail(AFQ_E_UNSUPPORTED, "gemm_f64: no separate addend");
  if (g.c_part) {
    if (!g.a_cplx || !g.b_cplx || g.D || g.a_sym || g.a_conj || g.c_part < 1 || g.c_part > 2)
      return fail(AFQ_E_INVALID, "gemm_f64: a real-part product takes complex A and B, no D, no packed or conjugated A");
    if (g.beta.y != 0.0 || (g.c_part == 1 ? g.alpha.y : g.alpha.x) != 0.0)
      return fail(AFQ_E_INVALID, "gemm_f64: real-part product: alpha must be real (part 1) or imaginary (part 2), beta real");
    return g.c_part == 1 ? dispatch_part<1>(g, s) : dispatch_part<2>(g, s);
  }
  if (g.a_conj && !g.a_cplx) return fail(AFQ_E_INVALID, "gemm_f64: only a complex A can be conjugated");
  if (g.a_sym && (g.a_trans || g.a_sym != g.M || g.a_sym != g.K))
    return fail(AFQ_E_INVALID, "gemm_f64: a packed symmetric A is square and not transposed");
  if (g.a_cplx) return g.b_cplx ? dispatch_t<true, true>(g, s) : dispatch_t<true, false>(g, s);
  return g.b_cplx ? dispatch_t<false, true>(g, s) : dispatch_t<false, false>(g, s);
}

int gemm_f64_pair(const Gemm& g0, const Gemm& g1, hipStream_t s) {
  auto plain = [](const Gemm& g) {
    return g.batch > 1 && !g.ws && !g.a_sym && !g.D && !g.E && !g.c_part && g.K > 0 && g.M > 0 && g.N > 0 &&
           (!g.a_conj || g.a_cplx);
  };
  const bool same = g0.a_cplx == g1.a_cplx && g0.b_cplx == g1.b_cplx && g0.a_trans == g1.a_trans &&
                    g0.b_trans == g1.b_trans && g0.a_conj == g1.a_conj;
  const int cfg = batched_config(g0);
  if (plain(g0) && plain(g1) && same && cfg >= 0 && cfg == batched_config(g1)) {
    // the operand forms of the overlap, Green's-function and QR products
    const bool ac = g0.a_cplx, bc = g0.b_cplx, at = g0.a_trans, bt = g0.b_trans;
    if (!ac && bc && at && !bt) return dispatch_pair<false, true, true, false>(g0, g1, cfg, s);
    if (ac && bc && at && bt) return dispatch_pair<true, true, true, true>(g0, g1, cfg, s);
    if (ac && bc && at && !bt) return dispatch_pair<true, true, true, false>(g0, g1, cfg, s);
    if (ac && bc && !at && !bt) return dispatch_pair<true, true, false, false>(g0, g1, cfg, s);
  }
  AFQ_TRY(gemm_f64(g0, s));
  return gemm_f64(g1, s);
}

}  // namespace afq
