// Fused split-BF16 propagator: the whole Trotter step of a walker panel,
//     W <- E . (t0 + sum_{n=1}^{K} VHS^n t0 / n!),   t0 = E . W,
// E = exp(-dt/2 h1) of the walker's spin, VHS = i sqrt(dt) sum_g x~_g L_g
// (complex64, per walker), in one kernel (propagation.py:154-176 with the
// reference's vhs_mixed_precision split, :196-203).
//
// A work-group owns one walker's column slab (16 TN columns of one spin) for
// all N rows.  The slab of the current operand (W, t0, a Taylor term, or the
// sum) lives in LDS as four bf16 planes (re hi, re lo, im hi, im lo) with k
// (the row index) contiguous, so an MFMA B fragment is one 16-byte read; the
// 8-element groups of k are XOR-swizzled by the column so that the 16 lanes of
// a fragment read distinct banks.  Each of the 8 waves owns row blocks
// rb = wave, wave + 8, ... and keeps their outputs in registers:
//   * the one-body products (E - I) . X run on E - I pre-split into three bf16
//     parts (hi, mid, lo: 2^-27 relative) times the panel's two, four
//     products (hi.hi + hi.lo + mid.hi + lo.hi), and the identity part is
//     added in FP64 from the walker buffer: t0 = X + (E - I) X;
//   * a Taylor term is VHS . term / n on the split-BF16 x3 scheme (hi.hi +
//     lo.hi + hi.lo per real product).  The VHS product writes the VHS as four
//     bf16 planes of the full N x N matrix (from the packed lower block
//     triangle when L is symmetric, mirrored) stored by 32-deep k step, so
//     an A fragment is one 16-byte load per lane and plane, 1 KB of
//     consecutive bytes, straight from global memory (L2), and the inner
//     loop does no conversion;
//   * the running sum of the terms stays in FP32 registers (the terms are
//     complex64 values) and is added to t0 in FP64.
// t0 and t0 + sum are kept in the walker buffer itself (FP64): every lane
// writes and later re-reads only its own accumulator elements, after the
// whole slab has been staged in LDS.  HBM traffic per walker and step is the
// walker panel in and out plus one read of the VHS; the VHS re-reads of the
// other slabs and terms come from L2 (a walker's slabs run on one XCD).
#include "kernels.h"
#include "trot_fused.h"
#include "group.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

namespace afq {
namespace {

constexpr int FW = 8;  // waves per work-group

__device__ __forceinline__ v4f mfma_bf16(const v8bf& a, const v8bf& b, const v4f& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// bf16 element offset of panel entry (k, n) in a plane whose columns hold kpl
// k-entries each (kpl a multiple of 128): the 8-element groups of k are
// XOR-permuted by n mod 16, so the 16 columns of a fragment read occupy
// distinct 16-byte bank groups
__device__ __forceinline__ int pofs(int n, int k, int kpl) {
  return n * kpl + ((((k >> 3) ^ (n & 15)) << 3) | (k & 7));
}

// bf16 element offset of entry (row i, column k) in a VHS or (E - I) plane of
// rows x K32 entries stored by 32-deep k step, [k / 32][i][k % 32]: the 16
// rows x 32 k of an MFMA A fragment are 1 KB of consecutive bytes
__device__ __forceinline__ long tofs(int i, int k, int rows) {
  return ((long)(k >> 5) * rows + i) * 32 + (k & 31);
}

__device__ __forceinline__ long sym_packed_cols_d(int N) {
  const long nb = (N + 15) / 16;
  return nb * (nb + 1) / 2 * 256;
}

__device__ __forceinline__ void split1(float y, __bf16& hi, __bf16& lo) {
  hi = (__bf16)y;
  lo = (__bf16)(y - (float)hi);
}


// -x of 8 bf16 (sign bits flipped)
__device__ __forceinline__ v8bf neg8(const v8bf& x) {
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  v4u u = __builtin_bit_cast(v4u, x);
  u ^= 0x80008000u;
  return __builtin_bit_cast(v8bf, u);
}

// XCD-aware order of a job's work-groups: item i (i mod 8 = its XCD) is slab
// (i / 8) mod nsl of walker i mod 8 + 8 ((i / 8) / nsl), so all slabs of a
// walker run on one XCD and its VHS is read into one L2.
__device__ __forceinline__ void fused_item(const FusedArgs& a, long i, int* w, int* sl) {
  const int nsl = a.nsl[0] + a.nsl[1];
  const long it = i >> 3;
  *w = (int)(i & 7) + 8 * (int)(it / nsl);
  *sl = (int)(it % nsl);
}

template <int TN, int RBW>
__device__ __forceinline__ void trot_fused_body(const FusedArgs& a, int w, int sl) {
  constexpr int CS = 16 * TN;
  constexpr int KPLMAX = (RBW * FW * 16 + 63) / 64 * 64;
  // B fragments read once per k-step for all row blocks, except in the
  // shapes whose registers would then spill
  constexpr bool BONCE = !(TN == 3 && RBW >= 3) && !(TN == 2 && RBW >= 4);
  __shared__ __align__(16) __bf16 P[4 * CS * KPLMAX];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, q = lane >> 4;
  const int s = sl < a.nsl[0] ? 0 : 1;
  if (s) sl -= a.nsl[0];
  const int cs = a.c[s], c0 = sl * CS, N = a.N, kpl = a.KPL;
  const int plane = CS * kpl;
  double2* X = a.wb[s] + (long)w * N * cs;
  const __bf16* E = a.e1[s];
  const __bf16* V = a.vs + 4 * a.ep * w;
  const int nrb = max(0, (a.R - wave + FW - 1) / FW);  // this wave's row blocks: wave, wave + FW, ...

  // ---- the walker slab into the LDS planes (zero past N rows and past the buffer's columns)
  {
    const int items = (a.K32 >> 3) * CS;
    for (int e = tid; e < items; e += 64 * FW) {
      const int n = e % CS, g = e / CS, col = c0 + n;
      v8bf rh, rl, ih, il;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int m = 8 * g + t;
        double2 x = {0.0, 0.0};
        if (m < N && col < cs) x = X[(long)m * cs + col];
        __bf16 h, l;
        split1((float)x.x, h, l);
        rh[t] = h;
        rl[t] = l;
        split1((float)x.y, h, l);
        ih[t] = h;
        il[t] = l;
      }
      const int o = n * kpl + ((g ^ (n & 15)) << 3);
      *reinterpret_cast<v8bf*>(P + o) = rh;
      *reinterpret_cast<v8bf*>(P + plane + o) = rl;
      *reinterpret_cast<v8bf*>(P + 2 * plane + o) = ih;
      *reinterpret_cast<v8bf*>(P + 3 * plane + o) = il;
    }
  }
  __syncthreads();

  v4f aR[RBW][TN], aI[RBW][TN];  // product of the current phase
  v4f sR[RBW][TN], sI[RBW][TN];  // running sum of the Taylor terms

  // B fragments of column block j at k-step k0 (4 planes)
  auto bfrag = [&](int j, int k0, v8bf& brh, v8bf& brl, v8bf& bih, v8bf& bil) {
    const int n = 16 * j + r;
    const int o = n * kpl + ((((k0 >> 3) + q) ^ (n & 15)) << 3);
    brh = *reinterpret_cast<const v8bf*>(P + o);
    brl = *reinterpret_cast<const v8bf*>(P + plane + o);
    bih = *reinterpret_cast<const v8bf*>(P + 2 * plane + o);
    bil = *reinterpret_cast<const v8bf*>(P + 3 * plane + o);
  };

  // The products run k-outer, row-block-inner.  Within a 32-deep k-step a
  // wave multiplies each of its row blocks in turn, then immediately loads
  // that row block's A fragments for the next step, so every load has a whole
  // step of MFMAs (all row blocks) to arrive without a second register set;
  // the B fragments of a k-step are read from LDS once for all row blocks
  // (BONCE) or per row block.  Row blocks past this wave's count are neither
  // loaded nor multiplied (a wave-uniform test).

  // aR/aI = (E - I) . panel
  auto onebody = [&]() {
#pragma unroll
    for (int ii = 0; ii < RBW; ++ii)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        aR[ii][j] = v4f{0.f, 0.f, 0.f, 0.f};
        aI[ii][j] = v4f{0.f, 0.f, 0.f, 0.f};
      }
    v8bf c[RBW][3];
    auto load = [&](int ii, int k0) {
      const __bf16* Er = E + tofs(16 * (wave + FW * ii) + r, k0 + 8 * q, 16 * a.R);
#pragma unroll
      for (int t = 0; t < 3; ++t) c[ii][t] = *reinterpret_cast<const v8bf*>(Er + t * a.ep);
    };
#pragma unroll
    for (int ii = 0; ii < RBW; ++ii)
      if (ii < nrb) load(ii, 0);
    for (int k0 = 0; k0 < a.K32; k0 += 32) {
      const int kn = min(k0 + 32, a.K32 - 32);
      // B fragments of the k-step, read once for all of this wave's row blocks
      // (per row block where the registers are short)
      v8bf b[TN][4];
      if constexpr (BONCE) {
#pragma unroll
        for (int j = 0; j < TN; ++j) bfrag(j, k0, b[j][0], b[j][1], b[j][2], b[j][3]);
      }
#pragma unroll
      for (int ii = 0; ii < RBW; ++ii) {
        if (ii >= nrb) continue;
        if constexpr (!BONCE) {
#pragma unroll
          for (int j = 0; j < TN; ++j) bfrag(j, k0, b[j][0], b[j][1], b[j][2], b[j][3]);
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          aR[ii][j] = mfma_bf16(c[ii][0], b[j][0], aR[ii][j]);
          aR[ii][j] = mfma_bf16(c[ii][0], b[j][1], aR[ii][j]);
          aR[ii][j] = mfma_bf16(c[ii][1], b[j][0], aR[ii][j]);
          aR[ii][j] = mfma_bf16(c[ii][2], b[j][0], aR[ii][j]);
          aI[ii][j] = mfma_bf16(c[ii][0], b[j][2], aI[ii][j]);
          aI[ii][j] = mfma_bf16(c[ii][0], b[j][3], aI[ii][j]);
          aI[ii][j] = mfma_bf16(c[ii][1], b[j][2], aI[ii][j]);
          aI[ii][j] = mfma_bf16(c[ii][2], b[j][2], aI[ii][j]);
        }
        load(ii, kn);
      }
    }
  };

  // aR/aI = VHS . panel, A fragments from the VHS planes (rows and columns
  // past N hold zeros)
  auto term = [&]() {
#pragma unroll
    for (int ii = 0; ii < RBW; ++ii)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        aR[ii][j] = v4f{0.f, 0.f, 0.f, 0.f};
        aI[ii][j] = v4f{0.f, 0.f, 0.f, 0.f};
      }
    v8bf c[RBW][4];
    auto load = [&](int ii, int k0) {
      const __bf16* Vr = V + tofs(16 * (wave + FW * ii) + r, k0 + 8 * q, 16 * a.R);
#pragma unroll
      for (int t = 0; t < 4; ++t) c[ii][t] = *reinterpret_cast<const v8bf*>(Vr + t * a.ep);
    };
#pragma unroll
    for (int ii = 0; ii < RBW; ++ii)
      if (ii < nrb) load(ii, 0);
    for (int k0 = 0; k0 < a.K32; k0 += 32) {
      const int kn = min(k0 + 32, a.K32 - 32);
      // B fragments of every column block, read once per k-step for all of
      // this wave's row blocks; consecutive MFMAs go to the 2 TN independent
      // accumulators of a row block in turn
      v8bf b[TN][6];
      auto bload = [&]() {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          bfrag(j, k0, b[j][0], b[j][1], b[j][2], b[j][3]);
          b[j][4] = neg8(b[j][2]);
          b[j][5] = neg8(b[j][3]);
        }
      };
      if constexpr (BONCE) bload();
#pragma unroll
      for (int ii = 0; ii < RBW; ++ii) {
        if (ii >= nrb) continue;
        if constexpr (!BONCE) bload();
        // re += ar br - ai bi ; im += ar bi + ai br, each as hi*hi + lo*hi + hi*lo
        constexpr int PR[6][2] = {{0, 0}, {1, 0}, {0, 1}, {2, 4}, {3, 4}, {2, 5}};  // (A part, B part) of re
        constexpr int PI[6][2] = {{0, 2}, {1, 2}, {0, 3}, {2, 0}, {3, 0}, {2, 1}};  // of im
#pragma unroll
        for (int t = 0; t < 6; ++t)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            aR[ii][j] = mfma_bf16(c[ii][PR[t][0]], b[j][PR[t][1]], aR[ii][j]);
            aI[ii][j] = mfma_bf16(c[ii][PI[t][0]], b[j][PI[t][1]], aI[ii][j]);
          }
        load(ii, kn);
      }
    }
  };

  // aR/aI (accumulator layout: element v of lane (r, q) in block (rb, j) is
  // row 16 rb + 4 q + v, column 16 j + r) into the LDS planes; rows past N as zeros
  auto put = [&]() {
#pragma unroll
    for (int ii = 0; ii < RBW; ++ii) {
      const int rb = wave + FW * ii;
      if (rb >= a.R) continue;
      const int m0 = 16 * rb + 4 * q;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int n = 16 * j + r;
        v4bf rh, rl, ih, il;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const bool ok = m0 + v < N;
          __bf16 h, l;
          split1(ok ? aR[ii][j][v] : 0.f, h, l);
          rh[v] = h;
          rl[v] = l;
          split1(ok ? aI[ii][j][v] : 0.f, h, l);
          ih[v] = h;
          il[v] = l;
        }
        const int o = pofs(n, m0, kpl);
        *reinterpret_cast<v4bf*>(P + o) = rh;
        *reinterpret_cast<v4bf*>(P + plane + o) = rl;
        *reinterpret_cast<v4bf*>(P + 2 * plane + o) = ih;
        *reinterpret_cast<v4bf*>(P + 3 * plane + o) = il;
      }
    }
  };

  // dst = src (this lane's elements of the walker buffer, FP64) + add; also
  // leaves dst rounded to FP32 in aR/aI for put()
  auto combine = [&](const v4f (&addR)[RBW][TN], const v4f (&addI)[RBW][TN]) {
#pragma unroll
    for (int ii = 0; ii < RBW; ++ii) {
      const int rb = wave + FW * ii;
      if (rb >= a.R) continue;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = c0 + 16 * j + r;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = 16 * rb + 4 * q + v;
          float yr = 0.f, yi = 0.f;
          if (m < N && col < cs) {
            double2* p = X + (long)m * cs + col;
            double2 x = *p;
            x.x += (double)addR[ii][j][v];
            x.y += (double)addI[ii][j][v];
            *p = x;
            yr = (float)x.x;
            yi = (float)x.y;
          }
          aR[ii][j][v] = yr;
          aI[ii][j][v] = yi;
        }
      }
    }
  };

  // t0 = X + (E - I) X
  onebody();
  __syncthreads();  // every wave is done reading X from LDS
  combine(aR, aI);
  put();
#pragma unroll
  for (int ii = 0; ii < RBW; ++ii)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      sR[ii][j] = v4f{0.f, 0.f, 0.f, 0.f};
      sI[ii][j] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  __syncthreads();
  // term_n = VHS . term_{n-1} / n
  for (int n = 1; n < a.nterms; ++n) {
    term();
    const float inv = 1.0f / (float)n;
#pragma unroll
    for (int ii = 0; ii < RBW; ++ii)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        aR[ii][j] *= inv;
        aI[ii][j] *= inv;
        sR[ii][j] += aR[ii][j];
        sI[ii][j] += aI[ii][j];
      }
    __syncthreads();  // every wave is done reading term_{n-1}
    if (n + 1 < a.nterms) {
      put();
      __syncthreads();
    }
  }
  // Y = t0 + sum of the terms, then W <- Y + (E - I) Y
  combine(sR, sI);
  put();
  __syncthreads();
  onebody();
  combine(aR, aI);
}

template <int TN, int RBW>
__global__ __launch_bounds__(64 * FW, 1) void trot_fused_bf3_kernel(FusedArgs a) {
  int w, sl;
  fused_item(a, blockIdx.x, &w, &sl);
  if (w >= a.W) return;  // uniform for the work-group
  trot_fused_body<TN, RBW>(a, w, sl);
}

// Grouped: job j owns work-groups [prefix[j], prefix[j + 1]) (a multiple of 8
// from a multiple of 8, so the XCD order within a job is kept).
template <int TN, int RBW>
__global__ __launch_bounds__(64 * FW, 1) void trot_fused_grouped_kernel(const FusedArgs* __restrict__ jobs,
                                                                        const int* __restrict__ prefix, int njobs) {
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const FusedArgs a = jobs[j];
  int w, sl;
  fused_item(a, item - prefix[j], &w, &sl);
  if (w >= a.W) return;
  trot_fused_body<TN, RBW>(a, w, sl);
}

// One launch: the single job `a` (tab == nullptr) or the grouped table.
template <int TN, int RBW>
int launch_tn(const FusedArgs& a, const GroupTab<FusedArgs>* tab, long gx, hipStream_t st) {
  if (tab) {
    hipLaunchKernelGGL((trot_fused_grouped_kernel<TN, RBW>), dim3((unsigned)tab->total), dim3(64 * FW), 0, st,
                       tab->params, tab->prefix, tab->njobs);
    return launch_status("trot_fused_grouped_kernel");
  }
  hipLaunchKernelGGL((trot_fused_bf3_kernel<TN, RBW>), dim3((unsigned)gx), dim3(64 * FW), 0, st, a);
  return launch_status("trot_fused_bf3_kernel");
}

template <int TN>
int launch_rbw(const FusedArgs& a, const GroupTab<FusedArgs>* tab, int rbw, long gx, hipStream_t st) {
  switch (rbw) {
    case 1: return launch_tn<TN, 1>(a, tab, gx, st);
    case 2: return launch_tn<TN, 2>(a, tab, gx, st);
    case 3: return launch_tn<TN, 3>(a, tab, gx, st);
    case 4:
      // three column blocks at four row blocks per wave would exceed the LDS
      if constexpr (TN < 3) return launch_tn<TN, 4>(a, tab, gx, st);
      break;
    default: break;
  }
  return fail(AFQ_E_UNSUPPORTED, "trot_fused_bf3: no kernel for this shape");
}

// VHS -> planes, one 64-thread group per 16 x 16 block of a walker's VHS (four
// blocks per work-group): thread (r, q) reads entries (r, 4q .. 4q + 3) of the
// block (the block's 2 KB are read contiguously) and writes their bf16 parts
// at row 16 ib + r, columns 16 kb + 4q ..; for an off-diagonal packed block
// the transposed block (kb, ib) is written too, through a per-block LDS tile
// so that each thread again writes 4 consecutive entries of one row.
__device__ __forceinline__ void put4(__bf16* base, long ep, long off, const float2 (&x)[4], int nvalid) {
  __bf16 rh[4], rl[4], ih[4], il[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    split1(x[t].x, rh[t], rl[t]);
    split1(x[t].y, ih[t], il[t]);
  }
  if (nvalid == 4 && (off & 3) == 0) {
    *reinterpret_cast<v4bf*>(base + off) = v4bf{rh[0], rh[1], rh[2], rh[3]};
    *reinterpret_cast<v4bf*>(base + ep + off) = v4bf{rl[0], rl[1], rl[2], rl[3]};
    *reinterpret_cast<v4bf*>(base + 2 * ep + off) = v4bf{ih[0], ih[1], ih[2], ih[3]};
    *reinterpret_cast<v4bf*>(base + 3 * ep + off) = v4bf{il[0], il[1], il[2], il[3]};
  } else {
    for (int t = 0; t < nvalid; ++t) {
      base[off + t] = rh[t];
      base[ep + off + t] = rl[t];
      base[2 * ep + off + t] = ih[t];
      base[3 * ep + off + t] = il[t];
    }
  }
}

__device__ __forceinline__ void split_vhs_body(const VhsSplit& v, long blk) {
  __shared__ float2 tile[4][16][17];
  const int N = v.N, nb = (N + 15) / 16, K32 = 32 * ((N + 31) / 32);
  const long ep = 16L * nb * K32;
  const int sub = threadIdx.x >> 6, t = threadIdx.x & 63;
  const long nblk = v.sym ? (long)nb * (nb + 1) / 2 : (long)nb * nb;
  const long b = blk * 4 + sub;  // (walker, block) index
  const bool active = b < (long)v.W * nblk;
  const long w = active ? b / nblk : 0;
  const int bi = active ? (int)(b % nblk) : 0;
  int ib, kb;
  if (v.sym) {
    sym_block(bi, &ib, &kb);
  } else {
    ib = bi / nb;
    kb = bi % nb;
  }
  const int r = t >> 2, q = t & 3;
  float2 x[4] = {};
  if (active) {
    if (v.sym) {
      const float2* src = v.vhs + w * (sym_packed_cols_d(N)) + ((long)bi << 8) + r * 16 + 4 * q;
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = src[u];
    } else {
      const int i = 16 * ib + r;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = 16 * kb + 4 * q + u;
        if (i < N && k < N) x[u] = v.vhs[w * (long)N * N + (long)i * N + k];
      }
    }
  }
  __bf16* base = reinterpret_cast<__bf16*>(v.out) + 4 * ep * w;
  {
    const int i = 16 * ib + r, k0 = 16 * kb + 4 * q;
    if (active && i < N && k0 < N) put4(base, ep, tofs(i, k0, 16 * nb), x, min(4, N - k0));
  }
  if (!v.sym) return;
#pragma unroll
  for (int u = 0; u < 4; ++u) tile[sub][r][4 * q + u] = x[u];
  __syncthreads();
  if (!active || ib == kb) return;
  // transposed block: row 16 kb + r, columns 16 ib + 4q ..
  float2 y[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) y[u] = tile[sub][4 * q + u][r];
  const int i = 16 * kb + r, k0 = 16 * ib + 4 * q;
  if (i < N && k0 < N) put4(base, ep, tofs(i, k0, 16 * nb), y, min(4, N - k0));
}

__global__ __launch_bounds__(256) void split_vhs_kernel(VhsSplit v) { split_vhs_body(v, blockIdx.x); }

__global__ __launch_bounds__(256) void split_vhs_grouped_kernel(const VhsSplit* __restrict__ jobs,
                                                                const int* __restrict__ prefix, int njobs) {
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const VhsSplit v = jobs[j];
  split_vhs_body(v, item - prefix[j]);
}

// (E - I) split into three bf16 planes of rows x ld (zero padded), stored by
// 32-deep k step (tofs)
__global__ void split_e1_kernel(const double* E, int N, int rows, int ld, long plane, __bf16* out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)rows * ld) return;
  const int m = (int)(i / ld), k = (int)(i % ld);
  const double v = (m < N && k < N) ? E[(long)m * N + k] - (m == k ? 1.0 : 0.0) : 0.0;
  const __bf16 h = (__bf16)(float)v;
  const double r1 = v - (double)(float)h;
  const __bf16 mid = (__bf16)(float)r1;
  const double r2 = r1 - (double)(float)mid;
  const long o = tofs(m, k, rows);
  out[o] = h;
  out[plane + o] = mid;
  out[2 * plane + o] = (__bf16)(float)r2;
}

int fused_tn(const int* c, int nspin, int N) {
  // column blocks per slab: the fewest padded columns over both spins, then the
  // fewest slabs; three blocks only where their LDS planes fit (N <= 384)
  int best = 1;
  long best_cols = -1, best_sl = 0;
  for (int tn = 1; tn <= 3; ++tn) {
    if (tn == 3 && N > 384) break;
    long cols = 0, sl = 0;
    for (int s = 0; s < nspin; ++s) {
      const long k = (c[s] + 16 * tn - 1) / (16 * tn);
      sl += k;
      cols += k * 16 * tn;
    }
    if (best_cols < 0 || cols < best_cols || (cols == best_cols && sl < best_sl)) {
      best = tn;
      best_cols = cols;
      best_sl = sl;
    }
  }
  return best;
}

}  // namespace

bool trot_fused_supported(int N) { return N >= 1 && N <= 512; }

int trot_fused_tn(const int* c, int nspin, int N) { return fused_tn(c, nspin, N); }

int trot_fused_rbw(int N) { return ((N + 15) / 16 + FW - 1) / FW; }

long vhs_split_items(const VhsSplit& v) {
  const long nb = (v.N + 15) / 16;
  const long nblk = v.sym ? nb * (nb + 1) / 2 : nb * nb;
  return ((long)v.W * nblk + 3) / 4;
}

int k_split_vhs_bf16(const float2* vhs, bool sym, int N, int W, uint16_t* out, hipStream_t st) {
  const VhsSplit v{vhs, sym, N, W, out};
  const long n = vhs_split_items(v);
  if (n > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "split_vhs: grid too large");
  hipLaunchKernelGGL(split_vhs_kernel, dim3((unsigned)n), dim3(256), 0, st, v);
  return launch_status("split_vhs_kernel");
}

int split_vhs_grouped(const GroupTab<VhsSplit>& t, hipStream_t st) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  hipLaunchKernelGGL(split_vhs_grouped_kernel, dim3((unsigned)t.total), dim3(256), 0, st, t.params, t.prefix,
                     t.njobs);
  return launch_status("split_vhs_grouped_kernel");
}

long trot_fused_plane(int N) { return 16L * ((N + 15) / 16) * (32L * ((N + 31) / 32)); }

long trot_fused_e1_elems(int N) { return 3 * trot_fused_plane(N); }

long trot_fused_vhs_elems(int N, int W) { return 4 * trot_fused_plane(N) * W; }



int k_split_e1_bf16(const double* E, int N, uint16_t* out, hipStream_t st) {
  const int rows = 16 * ((N + 15) / 16), ld = 32 * ((N + 31) / 32);
  const long n = (long)rows * ld;
  hipLaunchKernelGGL(split_e1_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, E, N, rows, ld, n,
                     reinterpret_cast<__bf16*>(out));
  return launch_status("split_e1_kernel");
}

int trot_fused_args(const TrotFused& p, int tn, FusedArgs* out) {
  if (!trot_fused_supported(p.N)) return fail(AFQ_E_UNSUPPORTED, "trot_fused_bf3: N above 512");
  if (p.nspin < 1 || p.nspin > 2 || p.W < 1 || p.nterms < 1) return fail(AFQ_E_INVALID, "trot_fused_bf3: bad shape");
  if (!p.vs) return fail(AFQ_E_INVALID, "trot_fused_bf3: missing VHS planes");
  FusedArgs a{};
  const int R = (p.N + 15) / 16;
  a.N = p.N;
  a.R = R;
  a.K32 = 32 * ((p.N + 31) / 32);
  a.KPL = 128 * ((p.N + 127) / 128);
  a.W = p.W;
  a.vs = reinterpret_cast<const __bf16*>(p.vs);
  a.ep = 16L * R * a.K32;
  a.nterms = p.nterms;
  for (int s = 0; s < 2; ++s) {
    const bool on = s < p.nspin;
    a.wb[s] = on ? p.wb[s] : nullptr;
    a.c[s] = on ? p.c[s] : 0;
    a.nsl[s] = on ? (p.c[s] + 16 * tn - 1) / (16 * tn) : 0;
    a.e1[s] = reinterpret_cast<const __bf16*>(on ? p.e1[s] : nullptr);
    if (on && (!p.wb[s] || !p.e1[s] || p.c[s] < 1)) return fail(AFQ_E_INVALID, "trot_fused_bf3: missing operand");
  }
  *out = a;
  return AFQ_OK;
}

long trot_fused_items(const FusedArgs& a) { return 8L * ((a.W + 7) / 8) * (a.nsl[0] + a.nsl[1]); }

int trot_fused_grouped(const GroupTab<FusedArgs>& t, int tn, int rbw, hipStream_t st) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  const FusedArgs none{};
  switch (tn) {
    case 1: return launch_rbw<1>(none, &t, rbw, 0, st);
    case 2: return launch_rbw<2>(none, &t, rbw, 0, st);
    default: return launch_rbw<3>(none, &t, rbw, 0, st);
  }
}

int trot_fused_bf3(const TrotFused& p, hipStream_t st) {
  FusedArgs a;
  AFQ_TRY(trot_fused_args(p, fused_tn(p.c, p.nspin, p.N), &a));
  const int rbw = (a.R + FW - 1) / FW;  // the static LDS planes hold KPLMAX(rbw) >= KPL entries per column
  const long gx = trot_fused_items(a);
  if (gx > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "trot_fused_bf3: grid too large");
  switch (fused_tn(p.c, p.nspin, p.N)) {
    case 1: return launch_rbw<1>(a, nullptr, rbw, gx, st);
    case 2: return launch_rbw<2>(a, nullptr, rbw, gx, st);
    default: return launch_rbw<3>(a, nullptr, rbw, gx, st);
  }
}

}  // namespace afq
