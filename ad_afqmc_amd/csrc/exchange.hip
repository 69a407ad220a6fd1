// Fused Coulomb/exchange measurement on matrix cores.
//
// For walker w and Cholesky index g: f_g = rot_chol_g (O x N) . G_w^T (N x O),
// c_g = tr f_g, K_g = sum_ij f_g,ij f_g,ji (ad_afqmc/wavefunctions.py:1137-1139,
// 1347-1356).  The (G, O, O) tensor of f's is never stored: one wave owns one
// O x O block f_{g,w} in registers (NB x NB MFMA 16x16 blocks), the four waves of
// a work-group take four consecutive g for the same walker, and the transposed
// pairs needed by K are matched through a 16x16 per-wave LDS scratch.  c and K
// are reduced in FP64 in every mode.
//
// Arithmetic of f:
//   F64    : v_mfma_f64_16x16x4f64
//   F32    : v_mfma_f32_16x16x4f32 (rot_chol read as FP32, G rounded to FP32)
//   BF16X3 : v_mfma_f32_16x16x32_bf16 on hi/lo bf16 splits, hi*hi + hi*lo + lo*hi
//
// Fragments go straight from global memory to registers in MFMA layout.  The
// K dimension (orbital index p) is permuted within each chunk so that every
// lane reads consecutive elements of one row: for the 4-deep MFMAs lane l of
// slab s reads p = c + KS*(l/16) + s (KS slabs per chunk); the 32-deep bf16
// MFMA already takes 8 consecutive k per lane.  The next chunk is loaded while
// the current one is multiplied.
//
// Work-group order is XCD-aware: work-groups are handed to the eight XCDs
// round-robin, so the logical index L = xcd * (T/8) + (bid/8) gives each XCD a
// contiguous range of the (gamma-group, walker, chunk) order, keeping one
// rot_chol group and the active walkers' G in that XCD's L2.
#include <type_traits>

#include "kernels.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace afq {
namespace {

enum Mode { F64 = 0, F32 = 1, BF16X3 = 2 };

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

struct Bf8 {
  v8bf hi, lo;
};
__device__ inline void split8(const float (&x)[8], Bf8& o) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const __bf16 h = (__bf16)x[t];
    o.hi[t] = h;
    o.lo[t] = (__bf16)(x[t] - (float)h);
  }
}
__device__ inline v4f mma3(const Bf8& a, const Bf8& b, v4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
  return c;
}

// Per-mode fragment of one chunk: NB row blocks of A (rot_chol rows) and of
// B (Green's function rows), KE elements per lane.
template <int MODE, int NB, int KS>
struct Frag {
  static constexpr int KE = MODE == BF16X3 ? 8 : KS;
  using R = std::conditional_t<MODE == F64, double, float>;
  R a[NB][KE];
  R br[NB][KE];
  R bi[NB][KE];
};

template <int MODE, int NB, int KS, typename TR>
__device__ inline void load_frag(Frag<MODE, NB, KS>& F, const TR* __restrict__ A, const double2* __restrict__ B, int O,
                                 int N, int c, int lane) {
  using R = typename Frag<MODE, NB, KS>::R;
  constexpr int KE = Frag<MODE, NB, KS>::KE;
  const int kb = c + KE * (lane >> 4);
  const int r16 = lane & 15;
  const bool full = (kb + KE - 1 < N);
  // Branch-free: rows >= O and columns >= N are clamped to valid addresses
  // and their values masked to zero.
#pragma unroll
  for (int bi = 0; bi < NB; ++bi) {
    const int row = bi * 16 + r16;
    const bool rok = row < O;
    const long roff = (long)(rok ? row : O - 1) * N;
#pragma unroll
    for (int s = 0; s < KE; ++s) {
      const bool ok = rok && (full || kb + s < N);
      const int col = full ? kb + s : min(kb + s, N - 1);
      const double2 v = B[roff + col];
      const TR a = A[roff + col];
      F.a[bi][s] = ok ? (R)a : (R)0;
      F.br[bi][s] = ok ? (R)v.x : (R)0;
      F.bi[bi][s] = ok ? (R)v.y : (R)0;
    }
  }
}

// NB <= 3 fits two work-groups per CU (<= 256 VGPRs per lane); NB = 4 needs the whole file.
template <int MODE, int NB, int KS, typename TR>
__global__ __launch_bounds__(256, (NB <= 3 ? 2 : 1)) void exchange_kernel(const TR* __restrict__ rc, const double2* __restrict__ Gm, int O,
                                                       int N, int Gn, int W, int nchunk, int cpg, long T8,
                                                       double2* __restrict__ coul, double2* __restrict__ kpart) {
  using Acc = std::conditional_t<MODE == F64, v4d, v4f>;
  __shared__ double2 scr[4][16 * 17];
  __shared__ double2 kred[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // XCD-aware logical index -> (gamma group, walker, chunk in group)
  const long bid = blockIdx.x;
  const long L = (bid & 7) * T8 + (bid >> 3);
  const long per_group = (long)W * cpg;
  const long grp = L / per_group;
  const long rem = L - grp * per_group;
  const int w = (int)(rem / cpg);
  const int chunk = (int)(grp * cpg + rem % cpg);
  if (chunk >= nchunk) return;  // uniform for the whole work-group
  const int g = chunk * 4 + wave;
  const bool valid = g < Gn;
  const int gg = valid ? g : Gn - 1;

  const TR* A = rc + (long)gg * O * N;
  const double2* B = Gm + (long)w * O * N;

  Acc accR[NB][NB], accI[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      accR[i][j] = Acc{0, 0, 0, 0};
      accI[i][j] = Acc{0, 0, 0, 0};
    }

  using Fr = Frag<MODE, NB, KS>;
  constexpr int CW = 4 * Fr::KE;
  Fr cur, nxt;
  load_frag<MODE, NB, KS, TR>(cur, A, B, O, N, 0, lane);
  for (int c = 0; c < N; c += CW) {
    if (c + CW < N) load_frag<MODE, NB, KS, TR>(nxt, A, B, O, N, c + CW, lane);
    if constexpr (MODE == BF16X3) {
      Bf8 fa[NB], fr[NB], fi[NB];
#pragma unroll
      for (int bi = 0; bi < NB; ++bi) {
        split8(cur.a[bi], fa[bi]);
        split8(cur.br[bi], fr[bi]);
        split8(cur.bi[bi], fi[bi]);
      }
#pragma unroll
      for (int bi = 0; bi < NB; ++bi)
#pragma unroll
        for (int bj = 0; bj < NB; ++bj) {
          accR[bi][bj] = mma3(fa[bi], fr[bj], accR[bi][bj]);
          accI[bi][bj] = mma3(fa[bi], fi[bj], accI[bi][bj]);
        }
    } else {
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int bi = 0; bi < NB; ++bi)
#pragma unroll
          for (int bj = 0; bj < NB; ++bj) {
            if constexpr (MODE == F64) {
              accR[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[bi][s], cur.br[bj][s], accR[bi][bj], 0, 0, 0);
              accI[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[bi][s], cur.bi[bj][s], accI[bi][bj], 0, 0, 0);
            } else {
              accR[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.a[bi][s], cur.br[bj][s], accR[bi][bj], 0, 0, 0);
              accI[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.a[bi][s], cur.bi[bj][s], accI[bi][bj], 0, 0, 0);
            }
          }
    }
    if (c + CW < N) cur = nxt;
  }

  // element v of lane l sits at (row, col) = (l/16 + 4v, l%16) for f64 and
  // (4(l/16) + v, l%16) for f32 accumulators
  const int cc = lane & 15;
  double kr = 0.0, ki = 0.0, cr = 0.0, ci = 0.0;
  double2* S = scr[wave];
#pragma unroll
  for (int bi = 0; bi < NB; ++bi)
#pragma unroll
    for (int bj = bi; bj < NB; ++bj) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = MODE == F64 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
        S[r * 17 + cc] = double2{(double)accR[bj][bi][v], (double)accI[bj][bi][v]};
      }
      __syncthreads();
      const double wgt = (bi == bj) ? 1.0 : 2.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = MODE == F64 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
        const int gi = bi * 16 + r, gj = bj * 16 + cc;
        if (gi < O && gj < O) {
          const double xr = accR[bi][bj][v], xi = accI[bi][bj][v];
          const double2 y = S[cc * 17 + r];
          kr += wgt * (xr * y.x - xi * y.y);
          ki += wgt * (xr * y.y + xi * y.x);
          if (bi == bj && r == cc) {
            cr += xr;
            ci += xi;
          }
        }
      }
      __syncthreads();
    }
  kr = wave_sum(kr);
  ki = wave_sum(ki);
  cr = wave_sum(cr);
  ci = wave_sum(ci);
  if (lane == 0) {
    if (valid) coul[(long)w * Gn + g] = double2{cr, ci};
    kred[wave] = valid ? double2{kr, ki} : double2{0.0, 0.0};
  }
  __syncthreads();
  if (tid == 0) {
    double2 s = {0.0, 0.0};
    for (int i = 0; i < 4; ++i) {
      s.x += kred[i].x;
      s.y += kred[i].y;
    }
    kpart[(long)w * nchunk + chunk] = s;
  }
}

template <int MODE, int NB, int KS, typename TR>
int launch_nb(const TR* rc, const double2* Gm, int O, int N, int Gn, int W, int nchunk, double2* coul, double2* kpart,
              hipStream_t s) {
  // gamma chunks per group: a group's rot_chol slice ~2.5 MB of L2
  const long chunk_bytes = 4L * O * N * (long)sizeof(TR);
  const int cpg = (int)std::max(1L, std::min((long)nchunk, (2560L * 1024) / std::max(1L, chunk_bytes)));
  const long ngroups = (nchunk + cpg - 1) / cpg;
  const long T = ngroups * cpg * (long)W;
  const long T8 = (T + 7) / 8;
  const long grid = T8 * 8;
  if (grid > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "exchange grid too large");
  hipLaunchKernelGGL((exchange_kernel<MODE, NB, KS, TR>), dim3((unsigned)grid), dim3(256), 0, s, rc, Gm, O, N, Gn, W,
                     nchunk, cpg, T8, coul, kpart);
  return launch_status("exchange_kernel");
}

template <int MODE, typename TR>
int launch_mode(const TR* rc, const double2* Gm, int O, int N, int Gn, int W, int nchunk, double2* coul,
                double2* kpart, hipStream_t s) {
  // chunk widths chosen so the FP64 fragments fit next to the accumulators:
  // NB = 3 uses 8-wide chunks (two work-groups per CU), NB = 4 (49 <= O <= 64) 4-wide ones
  switch ((O + 15) / 16) {
    case 1: return launch_nb<MODE, 1, 4, TR>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
    case 2: return launch_nb<MODE, 2, 4, TR>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
    case 3: return launch_nb<MODE, 3, 2, TR>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
    default: return launch_nb<MODE, 4, (MODE == F64 ? 1 : 4), TR>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
  }
}

}  // namespace

int exchange_nchunk(int Gn) { return (Gn + 3) / 4; }

int k_exchange(const double* rc, const double2* Gm, int O, int N, int Gn, int W, double2* coul, double2* kpart,
               int* nchunk_out, hipStream_t s) {
  if (O < 1 || O > 64) return fail(AFQ_E_UNSUPPORTED, "exchange kernel supports 1 <= nocc <= 64");
  const int nchunk = exchange_nchunk(Gn);
  *nchunk_out = nchunk;
  return launch_mode<F64, double>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
}

int k_exchange_lp(const float* rc, const double2* Gm, int O, int N, int Gn, int W, bool bf16x3, double2* coul,
                  double2* kpart, int* nchunk_out, hipStream_t s) {
  if (O < 1 || O > 64) return fail(AFQ_E_UNSUPPORTED, "exchange kernel supports 1 <= nocc <= 64");
  const int nchunk = exchange_nchunk(Gn);
  *nchunk_out = nchunk;
  if (bf16x3) return launch_mode<BF16X3, float>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
  return launch_mode<F32, float>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
}

}  // namespace afq
