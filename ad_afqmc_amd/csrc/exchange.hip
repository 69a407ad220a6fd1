// Fused Coulomb/exchange measurement (FP64 MFMA).
//
// For walker w and Cholesky index g: f_g = rot_chol_g (O x N) . G_w^T (N x O),
// c_g = tr f_g, K_g = sum_ij f_g,ij f_g,ji (ad_afqmc/wavefunctions.py:1137-1139,
// 1347-1356).  The (G, O, O) tensor of f's is never stored: one wave owns one
// O x O block f_{g,w} in registers (NB x NB MFMA 16x16 blocks), the four waves of
// a work-group take four consecutive g for the same walker, and the transposed
// pairs needed by K are matched through a 16x16 per-wave LDS scratch.
//
// Fragments go straight from global memory to registers in MFMA layout.  The
// K dimension (orbital index p) is permuted within each chunk so that lane l
// of slab s reads p = c + KS*(l/16) + s: each lane's KS slabs are consecutive
// elements of one row, for A and B alike.  The
// next chunk is loaded while the current one is multiplied (chunks of 16, or 4
// for NB = 4 to stay within the register file).
//
// Work-group order is XCD-aware: work-groups are handed to the eight XCDs
// round-robin, so the logical index L = xcd * (T/8) + (bid/8) gives each XCD a
// contiguous range of the (gamma-group, walker, chunk) order, keeping one
// rot_chol group and the active walkers' G in that XCD's L2.
#include "kernels.h"

namespace afq {
namespace {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// KS slabs of 4 per chunk (chunk width 4*KS); lane l of slab s reads p = c + KS*(l/16) + s.
template <int NB, int KS>
struct Frag {
  double a[NB][KS];
  double2 b[NB][KS];
};

template <int NB, int KS>
__device__ inline void load_frag(Frag<NB, KS>& F, const double* __restrict__ A, const double2* __restrict__ B, int O,
                                 int N, int c, int lane) {
  const int kb = c + KS * (lane >> 4);
  const int r16 = lane & 15;
  const bool full = (kb + KS - 1 < N);
#pragma unroll
  for (int bi = 0; bi < NB; ++bi) {
    const int row = bi * 16 + r16;
    if (row < O) {
      const double* ap = A + (long)row * N + kb;
      const double2* bp = B + (long)row * N + kb;
      if (full) {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          F.a[bi][s] = ap[s];
          F.b[bi][s] = bp[s];
        }
      } else {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const bool ok = kb + s < N;
          F.a[bi][s] = ok ? ap[s] : 0.0;
          F.b[bi][s] = ok ? bp[s] : double2{0.0, 0.0};
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        F.a[bi][s] = 0.0;
        F.b[bi][s] = double2{0.0, 0.0};
      }
    }
  }
}

template <int NB, int KS>
__global__ __launch_bounds__(256) void exchange_f64_kernel(const double* __restrict__ rc, const double2* __restrict__ Gm,
                                                           int O, int N, int Gn, int W, int nchunk, int cpg, long T8,
                                                           double2* __restrict__ coul, double2* __restrict__ kpart) {
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

  const double* A = rc + (long)gg * O * N;
  const double2* B = Gm + (long)w * O * N;

  v4d accR[NB][NB], accI[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      accR[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  constexpr int CW = 4 * KS;
  Frag<NB, KS> cur, nxt;
  load_frag<NB, KS>(cur, A, B, O, N, 0, lane);
  for (int c = 0; c < N; c += CW) {
    if (c + CW < N) load_frag<NB, KS>(nxt, A, B, O, N, c + CW, lane);
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int bi = 0; bi < NB; ++bi)
#pragma unroll
        for (int bj = 0; bj < NB; ++bj) {
          accR[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[bi][s], cur.b[bj][s].x, accR[bi][bj], 0, 0, 0);
          accI[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[bi][s], cur.b[bj][s].y, accI[bi][bj], 0, 0, 0);
        }
    if (c + CW < N) cur = nxt;
  }

  // K = sum over block pairs (bi <= bj) of weight * sum f_(bi,bj)[r][c] f_(bj,bi)[c][r]
  double kr = 0.0, ki = 0.0, cr = 0.0, ci = 0.0;
  double2* S = scr[wave];
  const int r0 = lane >> 4, cc = lane & 15;
#pragma unroll
  for (int bi = 0; bi < NB; ++bi)
#pragma unroll
    for (int bj = bi; bj < NB; ++bj) {
#pragma unroll
      for (int v = 0; v < 4; ++v) S[(r0 + 4 * v) * 17 + cc] = double2{accR[bj][bi][v], accI[bj][bi][v]};
      __syncthreads();
      const double wgt = (bi == bj) ? 1.0 : 2.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = r0 + 4 * v;
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

template <int NB, int KS>
int launch_nb(const double* rc, const double2* Gm, int O, int N, int Gn, int W, int nchunk, double2* coul,
              double2* kpart, hipStream_t s) {
  // gamma chunks per group: a group's rot_chol slice ~2.5 MB of L2
  const long chunk_bytes = 4L * O * N * (long)sizeof(double);
  const int cpg = (int)std::max(1L, std::min((long)nchunk, (2560L * 1024) / std::max(1L, chunk_bytes)));
  const long ngroups = (nchunk + cpg - 1) / cpg;
  const long T = ngroups * cpg * (long)W;
  const long T8 = (T + 7) / 8;
  const long grid = T8 * 8;
  if (grid > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "exchange grid too large");
  hipLaunchKernelGGL((exchange_f64_kernel<NB, KS>), dim3((unsigned)grid), dim3(256), 0, s, rc, Gm, O, N, Gn, W, nchunk,
                     cpg, T8, coul, kpart);
  return launch_status("exchange_f64_kernel");
}

}  // namespace

int exchange_nchunk(int Gn) { return (Gn + 3) / 4; }

int k_exchange(const double* rc, const double2* Gm, int O, int N, int Gn, int W, double2* coul, double2* kpart,
               int* nchunk_out, hipStream_t s) {
  if (O < 1 || O > 64) return fail(AFQ_E_UNSUPPORTED, "exchange kernel supports 1 <= nocc <= 64");
  const int nchunk = exchange_nchunk(Gn);
  *nchunk_out = nchunk;
  switch ((O + 15) / 16) {
    case 1: return launch_nb<1, 4>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
    case 2: return launch_nb<2, 4>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
    case 3: return launch_nb<3, 4>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
    default: return launch_nb<4, 1>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
  }
}

}  // namespace afq
