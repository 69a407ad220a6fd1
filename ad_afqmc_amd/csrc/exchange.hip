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
//   FP64   : v_mfma_f64_16x16x4f64 (exchange_kernel, exchange_split_kernel)
//   FP32   : v_mfma_f32_16x16x4f32 on pre-converted FP32 planes
//            (exchange_f32_kernel)
//   BF16X3 : v_mfma_f32_16x16x32_bf16 on pre-split hi/lo bf16 planes,
//            hi*hi + hi*lo + lo*hi (exchange_bf16_kernel)
//
// In the FP64 kernels fragments go straight from global memory to registers
// in MFMA layout.  The K dimension (orbital index p) is permuted within each
// chunk so that every lane reads consecutive elements of one row: lane l of
// slab s reads p = c + KS*(l/16) + s (KS slabs per chunk), and the next chunk
// is loaded while the current one is multiplied.  The FP32 and bf16 kernels
// take 8 consecutive k per lane from planes laid out for it.
//
// Work-group order is XCD-aware: work-groups are handed to the eight XCDs
// round-robin, so the logical index L = xcd * (T/8) + (bid/8) gives each XCD a
// contiguous range of the (gamma-group, walker, chunk) order, keeping one
// rot_chol group and the active walkers' G in that XCD's L2.

#include "kernels.h"
#include "group.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace afq {
namespace {


// Bytes of rot_chol a gamma group spans (its chunks run back to back on one
// XCD, so the slice stays in that XCD's L2 while all walkers pass over it;
// 1.25 and 0.625 MB measured no faster at N=300, G=2000, 256 walkers).
constexpr long kGroupSliceBytes = 2560L * 1024;

// The epilogue's transpose tile is private to each wave, and a wave's LDS
// accesses are performed in issue order: ordering them needs no work-group
// barrier, only that the compiler keep them in program order.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__host__ __device__ inline int bf16_opad_d(int O) { return 16 * ((O + 15) / 16); }
__host__ __device__ inline int bf16_npad_d(int N) { return 32 * ((N + 31) / 32); }
// Position of (row o, column k) in an Opad x Npad bf16 plane stored by 32-deep
// k step: [k / 32][o][k % 32].  Products that contract over the whole plane
// (the force bias) see the same permutation on both operands.
__host__ __device__ inline long bf16_tile(int o, int k, int Opad) {
  return ((long)(k >> 5) * Opad + o) * 32 + (k & 31);
}
// Position of (row o, column k) in an Opad x Npad FP32 plane for the 16x16x4
// f32 MFMA: 32-deep k steps of Opad x 32 elements; within a step, with
// k % 32 = 8q + 4h + e, [h][o][q][e].  In a step lane (r, q) of a 16-row
// fragment feeds k = 8q + t to MFMA t (t = 0..7), so its 8 k are the two
// 16-byte vectors h = 0, 1, and the 64 lanes of one vector load read 1 KB of
// consecutive bytes.
__host__ __device__ inline long f32_tile(int o, int k, int Opad) {
  return (long)(k >> 5) * Opad * 32 + (long)((k >> 2) & 1) * Opad * 16 + o * 16 + ((k >> 3) & 3) * 4 + (k & 3);
}

// Permutation of the four 16-byte k-groups of LDS row r: rows r and r' < 16
// of a fragment then occupy distinct 16-byte bank groups both within every 8
// consecutive rows (128 bytes of banks) and within all 16 (256 bytes).
__device__ inline int kswz(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// FP64 fragment of one chunk: NB row blocks of A (rot_chol rows) and of B
// (Green's function rows), KS elements per lane.
template <int NB, int KS>
struct Frag {
  double a[NB][KS];
  double br[NB][KS];
  double bi[NB][KS];
};

// One chunk as loaded: raw values at clamped (always valid) addresses, rows
// >= O and columns >= N included.  take() converts it to the MFMA type and
// zeroes A (rot_chol) there; that alone makes every padded product vanish:
// f_ij = 0 for i >= O, so f_ij f_ji and f_ii vanish whenever i or j >= O, and
// the clamped B values stay finite.  No loaded value is consumed before the
// MFMAs of the current chunk are issued, so the next chunk's loads overlap them
// (masking or converting at load time made the compiler wait for every load
// first).
template <int NB, int KS>
struct Raw {
  double a[NB][KS];
  double2 b[NB][KS];
};

template <int NB, int KS>
__device__ inline void load_raw(Raw<NB, KS>& X, const double* __restrict__ A, const double2* __restrict__ B, int O,
                                int N, int c, int lane) {
  const int kb = c + KS * (lane >> 4);
  const int r16 = lane & 15;
#pragma unroll
  for (int bi = 0; bi < NB; ++bi) {
    const long roff = (long)min(bi * 16 + r16, O - 1) * N;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int col = min(kb + s, N - 1);
      X.a[bi][s] = A[roff + col];
      X.b[bi][s] = B[roff + col];
    }
  }
}

template <int NB, int KS>
__device__ inline void take(Frag<NB, KS>& F, const Raw<NB, KS>& X, int O, int N, int c, int lane) {
  const int kb = c + KS * (lane >> 4);
#pragma unroll
  for (int bi = 0; bi < NB; ++bi) {
    const bool rok = bi * 16 + (lane & 15) < O;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      F.a[bi][s] = (rok && kb + s < N) ? X.a[bi][s] : 0.0;
      F.br[bi][s] = X.b[bi][s].x;
      F.bi[bi][s] = X.b[bi][s].y;
    }
  }
}

// NB <= 3 fits two work-groups per CU (<= 256 VGPRs per lane); NB = 4 would
// need the whole register file (exchange_split_kernel serves 48 < O <= 64).
template <int NB, int KS>
__global__ __launch_bounds__(256, 2) void exchange_kernel(const double* __restrict__ rc, const double2* __restrict__ Gm, int O,
                                                       int N, int Gn, int W, int nchunk, int cpg, long T8,
                                                       double2* __restrict__ coul, double2* __restrict__ kpart) {
  static_assert(NB <= 3, "48 < O <= 64 runs in exchange_split_kernel");
  using Acc = v4d;
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

  Acc accR[NB][NB], accI[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      accR[i][j] = Acc{0, 0, 0, 0};
      accI[i][j] = Acc{0, 0, 0, 0};
    }

  constexpr int CW = 4 * KS;
  Frag<NB, KS> cur;
  Raw<NB, KS> nxt;
  load_raw<NB, KS>(nxt, A, B, O, N, 0, lane);
  take<NB, KS>(cur, nxt, O, N, 0, lane);
  for (int c = 0; c < N; c += CW) {
    if (c + CW < N) load_raw<NB, KS>(nxt, A, B, O, N, c + CW, lane);
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int bi = 0; bi < NB; ++bi)
#pragma unroll
        for (int bj = 0; bj < NB; ++bj) {
          accR[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[bi][s], cur.br[bj][s], accR[bi][bj], 0, 0, 0);
          accI[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[bi][s], cur.bi[bj][s], accI[bi][bj], 0, 0, 0);
        }
    if (c + CW < N) take<NB, KS>(cur, nxt, O, N, c + CW, lane);
  }

  // element v of lane l sits at (row, col) = (l/16 + 4v, l%16)
  const int cc = lane & 15;
  double kr = 0.0, ki = 0.0, cr = 0.0, ci = 0.0;
  double2* S = scr[wave];
#pragma unroll
  for (int bi = 0; bi < NB; ++bi)
#pragma unroll
    for (int bj = bi; bj < NB; ++bj) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = (lane >> 4) + 4 * v;
        S[r * 17 + cc] = double2{(double)accR[bj][bi][v], (double)accI[bj][bi][v]};
      }
      wave_lds_sync();
      const double wgt = (bi == bj) ? 1.0 : 2.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = (lane >> 4) + 4 * v;
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
      wave_lds_sync();
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

// FP64 exchange for 48 < O <= 64: a 64 x 64 complex f-block takes 256
// accumulator registers, which leaves one wave per SIMD.  Here two waves share
// one (walker, gamma): wave half h holds the block rows 2h, 2h + 1 (2 x 4 blocks,
// 128 registers), so two waves fit per SIMD.  The eight waves of a work-group
// take four consecutive gamma of one walker, as in exchange_kernel.  For
// sum_ij f_ij f_ji, a block whose transposed partner lies in the other half is
// exchanged through LDS (each wave publishes its four cross blocks), and the
// pairs inside a half are matched through a per-wave LDS tile afterwards.
template <int KS>
__global__ __launch_bounds__(512, 1) void exchange_split_kernel(const double* __restrict__ rc,
                                                                const double2* __restrict__ Gm, int O, int N, int Gn,
                                                                int W, int nchunk, int cpg, long T8,
                                                                double2* __restrict__ coul, double2* __restrict__ kpart) {
  constexpr int TS = 16 * 17;  // one padded 16 x 16 tile
  extern __shared__ double2 xsh[];  // [8 waves][4 tiles][TS]
  __shared__ double2 kred[8], cred[8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int gi = wave >> 1, h = wave & 1;
  const long bid = blockIdx.x;
  const long L = (bid & 7) * T8 + (bid >> 3);
  const long per_group = (long)W * cpg;
  const long grp = L / per_group;
  const long rem = L - grp * per_group;
  const int w = (int)(rem / cpg);
  const int chunk = (int)(grp * cpg + rem % cpg);
  if (chunk >= nchunk) return;  // uniform for the whole work-group
  const int g = chunk * 4 + gi;
  const bool valid = g < Gn;
  const int gg = valid ? g : Gn - 1;
  const double* A = rc + (long)gg * O * N;
  const double2* B = Gm + (long)w * O * N;
  const int r16 = lane & 15, q4 = lane >> 4;

  long ra[2], rb[4];
  bool oka[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int row = 16 * (2 * h + i) + r16;
    oka[i] = row < O;
    ra[i] = (long)min(row, O - 1) * N;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    rb[j] = (long)min(16 * j + r16, O - 1) * N;
  }
  v4d accR[2][4], accI[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) accR[i][j] = accI[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  struct Fr {
    double a[2][KS], br[4][KS], bi[4][KS];
  };
  // raw values at clamped addresses; mask() zeroes A at rows >= O and
  // columns >= N once the chunk is about to be used, which makes every padded
  // product vanish (f_ij = 0 for i >= O) while the clamped B values stay finite
  auto load = [&](Fr& F, int c) {
    const int kb = c + KS * q4;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int pc = min(kb + s, N - 1);
#pragma unroll
      for (int i = 0; i < 2; ++i) F.a[i][s] = A[ra[i] + pc];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 bv = B[rb[j] + pc];
        F.br[j][s] = bv.x;
        F.bi[j][s] = bv.y;
      }
    }
  };
  auto mask = [&](Fr& F, int c) {
    const int kb = c + KS * q4;
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int i = 0; i < 2; ++i) F.a[i][s] = (oka[i] && kb + s < N) ? F.a[i][s] : 0.0;
  };
  constexpr int CW = 4 * KS;
  Fr cur, nxt;
  load(cur, 0);
  mask(cur, 0);
  for (int c = 0; c < N; c += CW) {
    if (c + CW < N) load(nxt, c + CW);
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[i][s], cur.br[j][s], accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[i][s], cur.bi[j][s], accI[i][j], 0, 0, 0);
        }
    if (c + CW < N) {
      cur = nxt;
      mask(cur, c + CW);
    }
  }

  // element v of lane l is (row, col) = (l/16 + 4v, l%16) of its 16 x 16 block
  const int cc = lane & 15;
  double kr = 0.0, ki = 0.0, cr = 0.0, ci = 0.0;
  double2* mine = xsh + (long)wave * 4 * TS;
  const double2* other = xsh + (long)(wave ^ 1) * 4 * TS;
  const int ho = 1 - h;
  // cross blocks (2h + i, 2ho + jj): publish, then pair with the partner's (2ho + jj, 2h + i)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        mine[(i * 2 + jj) * TS + (q4 + 4 * v) * 17 + cc] =
            double2{accR[i][2 * ho + jj][v], accI[i][2 * ho + jj][v]};
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = q4 + 4 * v;
        const double xr = accR[i][2 * ho + jj][v], xi = accI[i][2 * ho + jj][v];
        const double2 y = other[(jj * 2 + i) * TS + cc * 17 + r];
        kr += xr * y.x - xi * y.y;
        ki += xr * y.y + xi * y.x;
      }
  __syncthreads();
  // pairs inside the half: blocks (2h + i, 2h + j), i <= j, through the first tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = i; j < 2; ++j) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        mine[(q4 + 4 * v) * 17 + cc] = double2{accR[j][2 * h + i][v], accI[j][2 * h + i][v]};
      __syncthreads();
      const double wgt = (i == j) ? 1.0 : 2.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = q4 + 4 * v;
        const double xr = accR[i][2 * h + j][v], xi = accI[i][2 * h + j][v];
        const double2 y = mine[cc * 17 + r];
        kr += wgt * (xr * y.x - xi * y.y);
        ki += wgt * (xr * y.y + xi * y.x);
        if (i == j && r == cc) {
          cr += xr;
          ci += xi;
        }
      }
      __syncthreads();
    }
  kr = wave_sum(kr);
  ki = wave_sum(ki);
  cr = wave_sum(cr);
  ci = wave_sum(ci);
  if (lane == 0) {
    kred[wave] = valid ? double2{kr, ki} : double2{0.0, 0.0};
    cred[wave] = double2{cr, ci};
  }
  __syncthreads();
  if (tid < 4) {
    const int gq = chunk * 4 + tid;
    if (gq < Gn)
      coul[(long)w * Gn + gq] = double2{cred[2 * tid].x + cred[2 * tid + 1].x, cred[2 * tid].y + cred[2 * tid + 1].y};
  }
  if (tid == 0) {
    double2 sum = {0.0, 0.0};
    for (int q = 0; q < 8; ++q) {
      sum.x += kred[q].x;
      sum.y += kred[q].y;
    }
    kpart[(long)w * nchunk + chunk] = sum;
  }
}

// Exchange epilogue of one wave: c_g = tr f_g and K_g = sum_ab f_ab f_ba from
// the NB x NB accumulator blocks (block (i, j), i <= j, holds f[16i + r][16j + c]
// at accumulator position (r, c); block (j, i) its transposed partner), then
// coul[w, g] and the work-group's sum of K into kpart[w, chunk].  S is the
// wave's 16 x 17 transpose tile, kred 4 shared partial sums.
template <int NB>
__device__ __forceinline__ void exchange_finish(const v4f (&accR)[NB][NB], const v4f (&accI)[NB][NB], int O, int Gn,
                                                int nchunk, double2* __restrict__ coul, double2* __restrict__ kpart,
                                                int w, int chunk, int g, bool valid, double2* S, double2* kred) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q4 = lane >> 4;
  // accumulator element v of lane l sits at (4(l/16) + v, l%16)
  const int cc = lane & 15;
  double kr = 0.0, ki = 0.0, cr = 0.0, ci = 0.0;
  // off-diagonal pairs: block (bi, bj) and, in the same lane and element, its
  // transposed partner from block (bj, bi); each pair counts twice in the sum
#pragma unroll
  for (int bi = 0; bi < NB; ++bi)
#pragma unroll
    for (int bj = bi + 1; bj < NB; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gi = bi * 16 + 4 * q4 + v, gj = bj * 16 + cc;
        if (gi < O && gj < O) {
          const double xr = accR[bi][bj][v], xi = accI[bi][bj][v];
          const double yr = accR[bj][bi][v], yi = accI[bj][bi][v];
          kr += 2.0 * (xr * yr - xi * yi);
          ki += 2.0 * (xr * yi + xi * yr);
        }
      }
  // diagonal blocks through the wave's transpose tile
#pragma unroll
  for (int bi = 0; bi < NB; ++bi) {
#pragma unroll
    for (int v = 0; v < 4; ++v) S[(4 * q4 + v) * 17 + cc] = double2{(double)accR[bi][bi][v], (double)accI[bi][bi][v]};
    wave_lds_sync();
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 4 * q4 + v;
      const int gi = bi * 16 + r, gj = bi * 16 + cc;
      if (gi < O && gj < O) {
        const double xr = accR[bi][bi][v], xi = accI[bi][bi][v];
        const double2 y = S[cc * 17 + r];
        kr += xr * y.x - xi * y.y;
        ki += xr * y.y + xi * y.x;
        if (r == cc) {
          cr += xr;
          ci += xi;
        }
      }
    }
    wave_lds_sync();
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

// Split-BF16 x3 exchange on pre-split operands.  rot_chol is held as bf16
// (hi, lo) planes and each walker's Green's function as (re_hi, re_lo, im_hi,
// im_lo) planes, rows padded to 16*NB and columns to a multiple of 32 with
// zeros, so every lane loads its 8 consecutive k as one 16-byte vector and no
// bounds masks or conversions remain in the loop.  The four waves of a
// work-group share one walker: its Green's-function planes for the next 32-deep
// step are copied once per work-group from global memory straight into LDS
// (global_load_lds, 16 bytes a lane, double-buffered, one barrier per step);
// each wave reads its own rot_chol fragments into registers one step ahead.
// The three products lo*hi, hi*hi, hi*lo are issued as three passes over all
// NB x NB blocks, so consecutive MFMAs never depend on each other; the lo
// fragments of the next step are requested as soon as the first pass has
// consumed this step's, the hi ones at the start of the step.
// Block (i, j), i <= j, holds f[16i + r][16j + c] = rot_chol_i . G_j^T with the
// accumulator element at (r, c); block (j, i), j > i, is computed with the
// operands swapped (G_i as A, rot_chol_j as B), so its element at (r, c) is
// f[16j + c][16i + r]: the transposed partner of the same lane's element of
// block (i, j).  The pair products f_ab f_ba of the off-diagonal blocks then
// need no data movement; only the NB diagonal blocks go through an LDS
// transpose tile.
template <int NB>
__device__ __forceinline__ void exchange_bf16_body(const v8bf* __restrict__ rcs, const v8bf* __restrict__ gs, int O,
                                                   int Opad, int Npad, int Gn, int nchunk, double2* __restrict__ coul,
                                                   double2* __restrict__ kpart, int w, int chunk) {
  constexpr int ROWS = NB * 16;
  constexpr int PER = 4 * ROWS * 4;  // v8bf of one step's B stage: 4 planes x ROWS x 4 k-groups
  static_assert(PER % 256 == 0, "stage must split into whole 64-lane copies per wave");
  constexpr int CPW = PER / 256;  // 64-lane copies per wave per step
  __shared__ v8bf bsh[2][PER];
  __shared__ double2 scr[4][16 * 17];
  __shared__ double2 kred[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = chunk * 4 + wave;
  const bool valid = g < Gn;
  const int gg = valid ? g : Gn - 1;

  // planes in units of v8bf (8 bf16): a plane is Opad x Npad
  const long plane = (long)Opad * Npad / 8;
  const int rowv = Npad / 8;
  const int r16 = lane & 15, q4 = lane >> 4;
  // a k step of a plane is Opad rows x 4 k-groups (bf16_tile); fragment b
  // of step kv / 4 starts at kv * Opad + 64 b
  const v8bf* Ah = rcs + (long)gg * 2 * plane + r16 * 4 + q4;
  const v8bf* Al = Ah + plane;
  const v8bf* Bw = gs + (long)w * 4 * plane;
  // LDS slot s of a stage holds plane s / (4 ROWS), row (s / 4) % ROWS and
  // k-group (s & 3) ^ kswz(row % 16): the four 16-byte k-groups of a row are
  // permuted so that the 16 rows of a fragment read fall in distinct banks.
  // Copy t of this wave fills slots [64 (4 t + wave), +64), one per lane, so
  // each lane fetches the global element its slot holds.
  long src[CPW];
#pragma unroll
  for (int t = 0; t < CPW; ++t) {
    const int s = 64 * (4 * t + wave) + lane;
    const int row = (s >> 2) % ROWS;
    src[t] = (long)(s / (4 * ROWS)) * plane + (long)row * 4 + ((s & 3) ^ kswz(row & 15));
  }
  auto copy = [&](int kv, int buf) {
#pragma unroll
    for (int t = 0; t < CPW; ++t)
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(Bw + src[t] + (long)kv * ROWS),
                                       (__attribute__((address_space(3))) void*)&bsh[buf][64 * (4 * t + wave)], 16,
                                       0, 0);
  };
  // s_waitcnt vmcnt(0): the copies of this wave have landed in LDS (before
  // the barrier that publishes them to the other waves), and so have the
  // rot_chol fragments of the next step.  The builtin lets the compiler count
  // this wait, so it adds none of its own for those fragments.
  auto copies_done = [] { __builtin_amdgcn_s_waitcnt(0x0F70); };

  v4f accR[NB][NB], accI[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      accR[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
      accI[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  v8bf ah[NB], al[NB], nh[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    ah[b] = Ah[b * 64];
    al[b] = Al[b * 64];
  }
  // one product pass over all blocks: rot_chol fragments a (one bf16 part),
  // Green's-function fragments gr, gi (real and imaginary, one bf16 part)
  auto pass = [&](const v8bf(&a)[NB], const v8bf(&gr)[NB], const v8bf(&gi)[NB]) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (i <= j) {
          accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], gr[j], accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], gi[j], accI[i][j], 0, 0, 0);
        } else {
          accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gr[j], a[i], accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gi[j], a[i], accI[i][j], 0, 0, 0);
        }
      }
  };
  copy(0, 0);
  int buf = 0;
  // The issue order of a step is pinned with scheduling barriers: B reads,
  // the next step's copies and hi fragments, the lo x hi pass, the next lo
  // fragments, then the hi x hi and hi x lo passes.  All LDS reads of a step
  // precede its copies into the other buffer, so no read waits on a copy.
  for (int kv = 0; kv < rowv; kv += 4) {
    // this step's stage and fragments have arrived (the wait opens the step,
    // so that nothing requested before the loop is left to count)
    copies_done();
    __syncthreads();
    // the next step's k offset; the last step fetches its own again (into
    // the idle buffer and registers), which keeps the loop free of branches
    // and the compiler's load counting exact
    const int kn = kv + 4 < rowv ? kv + 4 : kv;
    const v8bf* B = bsh[buf] + r16 * 4 + (q4 ^ kswz(r16));  // + plane * 4 ROWS + 16 b * 4
    v8bf brh[NB], bih[NB], brl[NB], bil[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      brh[b] = B[b * 64];
      bih[b] = B[2 * 4 * ROWS + b * 64];
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      brl[b] = B[4 * ROWS + b * 64];
      bil[b] = B[3 * 4 * ROWS + b * 64];
    }
    copy(kn, buf ^ 1);
#pragma unroll
    for (int b = 0; b < NB; ++b) nh[b] = Ah[b * 64 + (long)kn * ROWS];
    __builtin_amdgcn_sched_barrier(0);
    pass(al, brh, bih);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int b = 0; b < NB; ++b) al[b] = Al[b * 64 + (long)kn * ROWS];
    __builtin_amdgcn_sched_barrier(0);
    pass(ah, brh, bih);
    pass(ah, brl, bil);
#pragma unroll
    for (int b = 0; b < NB; ++b) ah[b] = nh[b];
    __builtin_amdgcn_sched_barrier(0);
    buf ^= 1;
  }
  // no copy into LDS may be in flight when the work-group ends
  copies_done();
  exchange_finish<NB>(accR, accI, O, Gn, nchunk, coul, kpart, w, chunk, g, valid, scr[wave], kred);
}

template <int NB>
__global__ __launch_bounds__(256, 2) void exchange_bf16_kernel(const v8bf* __restrict__ rcs, const v8bf* __restrict__ gs,
                                                               int O, int Opad, int Npad, int Gn, int W, int nchunk,
                                                               int cpg, long T8, double2* __restrict__ coul,
                                                               double2* __restrict__ kpart) {
  const long bid = blockIdx.x;
  const long L = (bid & 7) * T8 + (bid >> 3);
  const long per_group = (long)W * cpg;
  const long grp = L / per_group;
  const long rem = L - grp * per_group;
  const int w = (int)(rem / cpg);
  const int chunk = (int)(grp * cpg + rem % cpg);
  if (chunk >= nchunk) return;  // uniform for the whole work-group
  exchange_bf16_body<NB>(rcs, gs, O, Opad, Npad, Gn, nchunk, coul, kpart, w, chunk);
}

// Grouped: job j owns work-groups [prefix[j], prefix[j + 1]) (8 T8 of them,
// from a multiple of 8), in exchange_bf16_kernel's XCD-aware (gamma group,
// walker, chunk) order.
template <int NB>
__global__ __launch_bounds__(256, 2) void exchange_bf16_grouped_kernel(const ExchJob* __restrict__ jobs,
                                                                       const int* __restrict__ prefix, int njobs) {
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const ExchJob e = jobs[j];
  const long local = item - prefix[j];
  const long L = (local & 7) * e.T8 + (local >> 3);
  const long per_group = (long)e.W * e.cpg;
  const long grp = L / per_group;
  const long rem = L - grp * per_group;
  const int w = (int)(rem / e.cpg);
  const int chunk = (int)(grp * e.cpg + rem % e.cpg);
  if (chunk >= e.nchunk) return;  // uniform for the whole work-group
  exchange_bf16_body<NB>(reinterpret_cast<const v8bf*>(e.rcs), reinterpret_cast<const v8bf*>(e.gs), e.O, e.Opad,
                         e.Npad, e.Gn, e.nchunk, e.coul, e.kpart, w, chunk);
}

// FP32 exchange on pre-converted operands: rot_chol as one FP32 plane per
// gamma and each walker's Green's function as (re, im) FP32 planes, both in
// the f32_tile layout, rows padded to 16 NB and columns to a multiple of 32
// with zeros.  The work-group layout is the split-BF16 kernel's: four waves,
// one gamma each, share one walker, whose Green's-function planes for the
// next 32-deep step are copied into LDS by global_load_lds (double-buffered,
// one barrier per step), in MFMA lane order, so each fragment read is 1 KB of
// consecutive LDS bytes; each wave reads its rot_chol fragments into registers
// one step ahead.  A step is 8 MFMAs (16x16x4 f32) per block pair and part;
// for each column block j the loop runs t outermost, so the 2 NB products
// between two updates of one accumulator never wait on each other.  Blocks
// below the diagonal take the operands swapped, as in the split-BF16 kernel.
__device__ __forceinline__ float f4c(const float4& v, int e) {
  return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
}

template <int NB>
__device__ __forceinline__ void exchange_f32_body(const float4* __restrict__ rct, const float4* __restrict__ gt,
                                                  int O, int Npad, int Gn, int nchunk, double2* __restrict__ coul,
                                                  double2* __restrict__ kpart, int w, int chunk) {
  constexpr int ROWS = NB * 16;
  constexpr int STEP = ROWS * 8;           // float4 of one 32-deep step of a plane
  constexpr int PER = 2 * STEP;            // float4 of one step's stage: re and im planes
  static_assert(PER % 256 == 0, "stage must split into whole 64-lane copies per wave");
  constexpr int CPW = PER / 256;           // 64-lane copies per wave per step
  __shared__ float4 bsh[2][PER];
  __shared__ double2 scr[4][16 * 17];
  __shared__ double2 kred[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = chunk * 4 + wave;
  const bool valid = g < Gn;
  const int gg = valid ? g : Gn - 1;
  const long plane = (long)ROWS * Npad / 4;  // float4 per plane
  const int nsteps = Npad / 32;
  const int r16 = lane & 15, q4 = lane >> 4;
  // fragment (block b, half h) of step kv: kv STEP + h 4 ROWS + 64 b + 4 r + q
  const float4* A = rct + (long)gg * plane + r16 * 4 + q4;
  const float4* Bw = gt + (long)w * 2 * plane;
  // LDS slot 64 u + l, u = (plane p, half h, block j) = (2 p + h) NB + j, holds
  // the float4 that lane l = 16 q + c of that fragment reads: row 16 j + c,
  // vector q.  Copy t of this wave fills slots [64 (4 t + wave), +64).
  long src[CPW];
#pragma unroll
  for (int t = 0; t < CPW; ++t) {
    const int u = 4 * t + wave;
    const int j = u % NB, h = (u / NB) & 1, p = u / (2 * NB);
    src[t] = p * plane + h * 4 * ROWS + (16 * j + r16) * 4 + q4;
  }
  auto copy = [&](int kv, int buf) {
#pragma unroll
    for (int t = 0; t < CPW; ++t)
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(Bw + src[t] + (long)kv * STEP),
                                       (__attribute__((address_space(3))) void*)&bsh[buf][64 * (4 * t + wave)], 16,
                                       0, 0);
  };
  // s_waitcnt vmcnt(0): this wave's copies and rot_chol fragments have landed
  auto copies_done = [] { __builtin_amdgcn_s_waitcnt(0x0F70); };

  v4f accR[NB][NB], accI[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      accR[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
      accI[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  float4 a[NB][2], an[NB][2];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    a[b][0] = A[b * 64];
    a[b][1] = A[4 * ROWS + b * 64];
  }
  copy(0, 0);
  int buf = 0;
  for (int kv = 0; kv < nsteps; ++kv) {
    copies_done();
    __syncthreads();
    // the last step fetches its own again (into the idle buffer and
    // registers), which keeps the loop free of branches
    const int kn = kv + 1 < nsteps ? kv + 1 : kv;
    copy(kn, buf ^ 1);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      an[b][0] = A[(long)kn * STEP + b * 64];
      an[b][1] = A[(long)kn * STEP + 4 * ROWS + b * 64];
    }
    const float4* B = bsh[buf] + lane;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const float4 br0 = B[(0 * NB + j) * 64], br1 = B[(1 * NB + j) * 64];
      const float4 bi0 = B[(2 * NB + j) * 64], bi1 = B[(3 * NB + j) * 64];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float br = f4c(t < 4 ? br0 : br1, t & 3);
        const float bi = f4c(t < 4 ? bi0 : bi1, t & 3);
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          const float av = f4c(a[i][t >> 2], t & 3);
          if (i <= j) {
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, br, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bi, accI[i][j], 0, 0, 0);
          } else {
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(br, av, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bi, av, accI[i][j], 0, 0, 0);
          }
        }
      }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      a[b][0] = an[b][0];
      a[b][1] = an[b][1];
    }
    buf ^= 1;
  }
  // no copy into LDS may be in flight when the work-group ends
  copies_done();
  exchange_finish<NB>(accR, accI, O, Gn, nchunk, coul, kpart, w, chunk, g, valid, scr[wave], kred);
}

template <int NB>
__global__ __launch_bounds__(256, 2) void exchange_f32_kernel(const float4* __restrict__ rct,
                                                              const float4* __restrict__ gt, int O, int Npad, int Gn,
                                                              int W, int nchunk, int cpg, long T8,
                                                              double2* __restrict__ coul,
                                                              double2* __restrict__ kpart) {
  const long bid = blockIdx.x;
  const long L = (bid & 7) * T8 + (bid >> 3);
  const long per_group = (long)W * cpg;
  const long grp = L / per_group;
  const long rem = L - grp * per_group;
  const int w = (int)(rem / cpg);
  const int chunk = (int)(grp * cpg + rem % cpg);
  if (chunk >= nchunk) return;  // uniform for the whole work-group
  exchange_f32_body<NB>(rct, gt, O, Npad, Gn, nchunk, coul, kpart, w, chunk);
}

// out[r][c][f32_tile(o, k)] (o < Opad, k < Npad, zero padded) = float of
// src[r][o][k] (o < O, k < N; real, or interleaved complex with c = re, im).
__global__ void split_f32_kernel(const double* __restrict__ src, int cplx, int O, int N, int Opad, int Npad,
                                 long rows, float* __restrict__ out) {
  const long per = (long)Opad * Npad;
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= rows * per) return;
  const long r = idx / per;
  const int o = (int)((idx % per) / Npad), k = (int)(idx % Npad);
  const bool in = o < O && k < N;
  const int np = cplx ? 2 : 1;
  float* base = out + r * np * per + f32_tile(o, k, Opad);
  for (int c = 0; c < np; ++c) {
    double x = 0.0;
    if (in) x = cplx ? src[2 * ((r * O + o) * (long)N + k) + c] : src[(r * O + o) * (long)N + k];
    base[c * per] = (float)x;
  }
}

// out[r][plane][tile(o, k)] (o < Opad, k < Npad, zero padded) from
// src[r][o][k] (o < O, k < N; real or interleaved complex double): planes hi,
// lo of bf16(float(x)) and bf16(float(x) - hi), for the real then the
// imaginary part.  Within a plane the elements are stored by 32-deep k step,
// then row, then k (bf16_tile), so the 16 rows x 32 k of one MFMA operand
// fragment are 1 KB of consecutive bytes (the exchange's layout).  With ilv the
// planes of a row are instead interleaved every 32 elements of that order,
// out[r][t / 32][plane][t % 32] with t = bf16_tile(o, k): the force-bias
// product reads a row's 32-deep step of all planes as whole 128-byte lines.
// Rows of out are rst elements apart (rst = 0: the row's own size), so that
// the planes of two spins can share a row (both spins' force bias as one
// product over the concatenated k steps).
__device__ __forceinline__ void split_bf16_body(const double* __restrict__ src, int cplx, int O, int N, int Opad,
                                                int Npad, long rows, int ilv, long rst, __bf16* __restrict__ out,
                                                long idx) {
  const long per = (long)Opad * Npad;
  if (idx >= rows * per) return;
  const long r = idx / per;
  const int o = (int)((idx % per) / Npad), k = (int)(idx % Npad);
  const bool in = o < O && k < N;
  const int np = cplx ? 4 : 2;
  const long t = bf16_tile(o, k, Opad);
  const long pst = ilv ? 32 : per;  // plane stride
  __bf16* base = out + r * (rst > 0 ? rst : np * per) + (ilv ? (t >> 5) * (32 * np) + (t & 31) : t);
  for (int c = 0; c < (cplx ? 2 : 1); ++c) {
    double x = 0.0;
    if (in) x = cplx ? src[2 * ((r * O + o) * (long)N + k) + c] : src[(r * O + o) * (long)N + k];
    const float xf = (float)x;
    const __bf16 h = (__bf16)xf;
    base[(2 * c) * pst] = h;
    base[(2 * c + 1) * pst] = (__bf16)(xf - (float)h);
  }
}

__global__ void split_bf16_kernel(const double* __restrict__ src, int cplx, int O, int N, int Opad, int Npad,
                                  long rows, int ilv, long rst, __bf16* __restrict__ out) {
  split_bf16_body(src, cplx, O, N, Opad, Npad, rows, ilv, rst, out, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

__global__ void split_bf16_grouped_kernel(const SplitJob* __restrict__ jobs, const int* __restrict__ prefix,
                                          int njobs) {
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const SplitJob J = jobs[j];
  split_bf16_body(J.src, J.cplx, J.O, J.N, bf16_opad_d(J.O), bf16_npad_d(J.N), J.rows, J.ilv, J.rst,
                  reinterpret_cast<__bf16*>(J.out), (long)(item - prefix[j]) * 256 + threadIdx.x);
}

// c_g = tr f_g and K_g = sum_ij f_ij f_ji from stored f blocks: f holds, for
// each walker w, gcur consecutive O x O matrices f_{g0 + g} (row-major, one
// after another).  Work-group (w, c) covers the 4-g chunk c of the pass: it
// writes coul[w, g0 + 4c + t] and kpart[w, g0 / 4 + c] = sum_t K.
__global__ __launch_bounds__(256) void exchange_wide_reduce_kernel(const double2* __restrict__ f, int O, int gcur,
                                                                   int g0, int Gn, int nchunk,
                                                                   double2* __restrict__ coul,
                                                                   double2* __restrict__ kpart) {
  __shared__ double red[4][4];
  const int w = blockIdx.x, c = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long OO = (long)O * O;
  double ksr = 0.0, ksi = 0.0;
  for (int t = 0; t < 4; ++t) {
    const int g = 4 * c + t;
    if (g >= gcur) break;  // uniform for the work-group
    const double2* fg = f + ((long)w * gcur + g) * OO;
    double kr = 0.0, ki = 0.0, cr = 0.0, ci = 0.0;
    for (long e = tid; e < OO; e += 256) {
      const int i = (int)(e / O), jj = (int)(e - (long)i * O);
      const double2 x = fg[e], y = fg[(long)jj * O + i];
      kr += x.x * y.x - x.y * y.y;
      ki += x.x * y.y + x.y * y.x;
      if (i == jj) {
        cr += x.x;
        ci += x.y;
      }
    }
    kr = wave_sum(kr);
    ki = wave_sum(ki);
    cr = wave_sum(cr);
    ci = wave_sum(ci);
    if (lane == 0) {
      red[wave][0] = kr;
      red[wave][1] = ki;
      red[wave][2] = cr;
      red[wave][3] = ci;
    }
    __syncthreads();
    if (tid == 0) {
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      for (int q = 0; q < 4; ++q)
        for (int v = 0; v < 4; ++v) s[v] += red[q][v];
      coul[(long)w * Gn + g0 + g] = double2{s[2], s[3]};
      ksr += s[0];
      ksi += s[1];
    }
    __syncthreads();
  }
  if (tid == 0) kpart[(long)w * nchunk + g0 / 4 + c] = double2{ksr, ksi};
}

template <int NB, int KS>
int launch_nb(const double* rc, const double2* Gm, int O, int N, int Gn, int W, int nchunk, double2* coul,
              double2* kpart, hipStream_t s) {
  // gamma chunks per group: a group's rot_chol slice ~2.5 MB of L2
  const long chunk_bytes = 4L * O * N * (long)sizeof(double);
  const int cpg = (int)std::max(1L, std::min((long)nchunk, (kGroupSliceBytes) / std::max(1L, chunk_bytes)));
  const long ngroups = (nchunk + cpg - 1) / cpg;
  const long T = ngroups * cpg * (long)W;
  const long T8 = (T + 7) / 8;
  const long grid = T8 * 8;
  if (grid > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "exchange grid too large");
  hipLaunchKernelGGL((exchange_kernel<NB, KS>), dim3((unsigned)grid), dim3(256), 0, s, rc, Gm, O, N, Gn, W, nchunk,
                     cpg, T8, coul, kpart);
  return launch_status("exchange_kernel");
}

template <int KS>
int launch_split(const double* rc, const double2* Gm, int O, int N, int Gn, int W, int nchunk, double2* coul,
                 double2* kpart, hipStream_t s) {
  const long chunk_bytes = 4L * O * N * (long)sizeof(double);
  const int cpg = (int)std::max(1L, std::min((long)nchunk, (kGroupSliceBytes) / std::max(1L, chunk_bytes)));
  const long ngroups = (nchunk + cpg - 1) / cpg;
  const long T = ngroups * cpg * (long)W;
  const long T8 = (T + 7) / 8;
  const long grid = T8 * 8;
  if (grid > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "exchange grid too large");
  const size_t lds = sizeof(double2) * 8 * 4 * 16 * 17;
  static bool raised = false;
  if (!raised) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(exchange_split_kernel<KS>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(AFQ_E_HIP, std::string("exchange LDS attribute: ") + hipGetErrorString(e));
    raised = true;
  }
  hipLaunchKernelGGL((exchange_split_kernel<KS>), dim3((unsigned)grid), dim3(512), lds, s, rc, Gm, O, N, Gn, W,
                     nchunk, cpg, T8, coul, kpart);
  return launch_status("exchange_split_kernel");
}

// O <= 48; chunk widths chosen so the fragments fit next to the accumulators
// (NB = 3: 8-wide chunks, two work-groups per CU)
int launch_f64(const double* rc, const double2* Gm, int O, int N, int Gn, int W, int nchunk, double2* coul,
               double2* kpart, hipStream_t s) {
  switch ((O + 15) / 16) {
    case 1: return launch_nb<1, 4>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
    case 2: return launch_nb<2, 4>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
    default: return launch_nb<3, 2>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
  }
}

}  // namespace

int exchange_nchunk(int Gn) { return (Gn + 3) / 4; }

int k_exchange(const double* rc, const double2* Gm, int O, int N, int Gn, int W, double2* coul, double2* kpart,
               int* nchunk_out, hipStream_t s) {
  if (O < 1 || O > 64) return fail(AFQ_E_UNSUPPORTED, "exchange kernel supports 1 <= nocc <= 64");
  const int nchunk = exchange_nchunk(Gn);
  *nchunk_out = nchunk;
  // 48 < O <= 64: two waves per f-block (the 64 x 64 block alone would take 256 registers)
  if (O > 48) return launch_split<2>(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
  return launch_f64(rc, Gm, O, N, Gn, W, nchunk, coul, kpart, s);
}

long exchange_wide_ws_elems(int O, int Gn, int W) {
  // a pass of gc g's (a multiple of 4, at least one chunk) within 256 MB
  const long per_g = (long)W * O * O;
  const long budget = (256L << 20) / (long)sizeof(double2);
  long gc = std::max(4L, budget / std::max(1L, per_g) / 4 * 4);
  gc = std::min(gc, 4L * exchange_nchunk(Gn));
  return gc * per_g;
}

int k_exchange_wide(const double* rc, const double2* Gm, int O, int N, int Gn, int W, double2* ws, long ws_elems,
                    double2* coul, double2* kpart, int* nchunk_out, hipStream_t s) {
  if (O < 1 || N < 1 || Gn < 1 || W < 1) return fail(AFQ_E_INVALID, "exchange: sizes must be positive");
  const int nchunk = exchange_nchunk(Gn);
  *nchunk_out = nchunk;
  const long per_g = (long)W * O * O;
  const long gc = ws_elems / per_g / 4 * 4;
  if (!ws || gc < 4) return fail(AFQ_E_INVALID, "exchange: workspace below one 4-g chunk");
  for (long g0 = 0; g0 < Gn; g0 += gc) {
    const int gcur = (int)std::min(gc, (long)Gn - g0);
    // f_{g0..g0+gcur}[w] = rot_chol rows (gcur O x N) . G_w^T (N x O)
    Gemm g;
    g.M = gcur * O;
    g.N = O;
    g.K = N;
    g.batch = W;
    g.A = rc + g0 * O * (long)N;
    g.lda = N;
    g.sA = 0;
    g.b_cplx = true;
    g.b_trans = true;
    g.B = Gm;
    g.ldb = N;
    g.sB = (long)O * N;
    g.C = ws;
    g.ldc = O;
    g.sC = (long)gcur * O * O;
    AFQ_TRY(gemm_f64(g, s));
    hipLaunchKernelGGL(exchange_wide_reduce_kernel, dim3(W, (gcur + 3) / 4), dim3(256), 0, s, ws, O, gcur, (int)g0,
                       Gn, nchunk, coul, kpart);
    AFQ_TRY(launch_status("exchange_wide_reduce_kernel"));
  }
  return AFQ_OK;
}

int bf16_opad(int O) { return bf16_opad_d(O); }
int bf16_npad(int N) { return bf16_npad_d(N); }

long split_items(const SplitJob& j) {
  return (j.rows * bf16_opad_d(j.O) * (long)bf16_npad_d(j.N) + 255) / 256;
}

int split_bf16_grouped(const GroupTab<SplitJob>& t, hipStream_t s) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  hipLaunchKernelGGL(split_bf16_grouped_kernel, dim3((unsigned)t.total), dim3(256), 0, s, t.params, t.prefix,
                     t.njobs);
  return launch_status("split_bf16_grouped_kernel");
}

int k_split_bf16(const double* src, bool cplx, int O, int N, long rows, bool ilv, uint16_t* out, hipStream_t s,
                 long row_stride) {
  const int Opad = bf16_opad(O), Npad = bf16_npad(N);
  const long n = rows * Opad * (long)Npad;
  if (n <= 0) return AFQ_OK;
  if (row_stride != 0 && row_stride < (cplx ? 4L : 2L) * Opad * Npad)
    return fail(AFQ_E_INVALID, "k_split_bf16: row stride below the row size");
  hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, cplx ? 1 : 0, O, N,
                     Opad, Npad, rows, ilv ? 1 : 0, row_stride, reinterpret_cast<__bf16*>(out));
  return launch_status("split_bf16_kernel");
}

long exchange_group_items(ExchJob* e) {
  const long chunk_bytes = 4L * 2 * e->Opad * e->Npad * 2;
  e->cpg = (int)std::max(1L, std::min((long)e->nchunk, kGroupSliceBytes / std::max(1L, chunk_bytes)));
  const long ngroups = (e->nchunk + e->cpg - 1) / e->cpg;
  e->T8 = (ngroups * e->cpg * (long)e->W + 7) / 8;
  return 8 * e->T8;
}

int exchange_bf16_grouped(const GroupTab<ExchJob>& t, int nb, hipStream_t s) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  const dim3 grid((unsigned)t.total), blk(256);
  switch (nb) {
    case 1: hipLaunchKernelGGL(exchange_bf16_grouped_kernel<1>, grid, blk, 0, s, t.params, t.prefix, t.njobs); break;
    case 2: hipLaunchKernelGGL(exchange_bf16_grouped_kernel<2>, grid, blk, 0, s, t.params, t.prefix, t.njobs); break;
    case 3: hipLaunchKernelGGL(exchange_bf16_grouped_kernel<3>, grid, blk, 0, s, t.params, t.prefix, t.njobs); break;
    case 4: hipLaunchKernelGGL(exchange_bf16_grouped_kernel<4>, grid, blk, 0, s, t.params, t.prefix, t.njobs); break;
    default: return fail(AFQ_E_UNSUPPORTED, "exchange kernel supports 1 <= nocc <= 64");
  }
  return launch_status("exchange_bf16_grouped_kernel");
}

int k_exchange_bf16(const uint16_t* rcs, const uint16_t* gs, int O, int N, int Gn, int W, double2* coul,
                    double2* kpart, int* nchunk_out, hipStream_t s) {
  if (O < 1 || O > 64) return fail(AFQ_E_UNSUPPORTED, "exchange kernel supports 1 <= nocc <= 64");
  const int nchunk = exchange_nchunk(Gn);
  *nchunk_out = nchunk;
  const int Opad = bf16_opad(O), Npad = bf16_npad(N);
  const long chunk_bytes = 4L * 2 * Opad * Npad * 2;
  const int cpg = (int)std::max(1L, std::min((long)nchunk, (kGroupSliceBytes) / std::max(1L, chunk_bytes)));
  const long ngroups = (nchunk + cpg - 1) / cpg;
  const long T = ngroups * cpg * (long)W;
  const long T8 = (T + 7) / 8;
  const long grid = T8 * 8;
  if (grid > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "exchange grid too large");
  const v8bf* a = reinterpret_cast<const v8bf*>(rcs);
  const v8bf* b = reinterpret_cast<const v8bf*>(gs);
#define EXB(NB)                                                                                                  \
  hipLaunchKernelGGL(exchange_bf16_kernel<NB>, dim3((unsigned)grid), dim3(256), 0, s, a, b, O, Opad, Npad, Gn, W, \
                     nchunk, cpg, T8, coul, kpart)
  switch (Opad / 16) {
    case 1: EXB(1); break;
    case 2: EXB(2); break;
    case 3: EXB(3); break;
    default: EXB(4); break;
  }
#undef EXB
  return launch_status("exchange_bf16_kernel");
}

int k_split_f32(const double* src, bool cplx, int O, int N, long rows, float* out, hipStream_t s) {
  const int Opad = bf16_opad(O), Npad = bf16_npad(N);
  const long n = rows * Opad * (long)Npad;
  if (n <= 0) return AFQ_OK;
  hipLaunchKernelGGL(split_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, cplx ? 1 : 0, O, N,
                     Opad, Npad, rows, out);
  return launch_status("split_f32_kernel");
}

int k_exchange_f32(const float* rct, const float* gt, int O, int N, int Gn, int W, double2* coul, double2* kpart,
                   int* nchunk_out, hipStream_t s) {
  if (O < 1 || O > 64) return fail(AFQ_E_UNSUPPORTED, "exchange kernel supports 1 <= nocc <= 64");
  const int nchunk = exchange_nchunk(Gn);
  *nchunk_out = nchunk;
  const int Opad = bf16_opad(O), Npad = bf16_npad(N);
  const long chunk_bytes = 4L * Opad * Npad * 4;
  const int cpg = (int)std::max(1L, std::min((long)nchunk, (kGroupSliceBytes) / std::max(1L, chunk_bytes)));
  const long ngroups = (nchunk + cpg - 1) / cpg;
  const long T = ngroups * cpg * (long)W;
  const long T8 = (T + 7) / 8;
  const long grid = T8 * 8;
  if (grid > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "exchange grid too large");
  const float4* a = reinterpret_cast<const float4*>(rct);
  const float4* b = reinterpret_cast<const float4*>(gt);
#define EXF(NB)                                                                                                   \
  hipLaunchKernelGGL(exchange_f32_kernel<NB>, dim3((unsigned)grid), dim3(256), 0, s, a, b, O, Npad, Gn, W, nchunk, \
                     cpg, T8, coul, kpart)
  switch (Opad / 16) {
    case 1: EXF(1); break;
    case 2: EXF(2); break;
    case 3: EXF(3); break;
    default: EXF(4); break;
  }
#undef EXF
  return launch_status("exchange_f32_kernel");
}

}  // namespace afq
