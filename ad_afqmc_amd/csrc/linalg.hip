// Per-walker FP64 dense linear algebra kept in LDS / L2:
//   * Gauss-Jordan determinant + inverse of the O x O overlap matrix T^H W
//     (replaces jnp.linalg.det / jnp.linalg.inv, ad_afqmc/wavefunctions.py:1066,
//     1100, 1266-1268, 1306-1311)
//   * upper Cholesky factor of the O x O Gram matrix W^H W, the middle step of
//     the Cholesky-QR re-orthonormalisation of the block loop (replaces the
//     batched Householder QR of ad_afqmc/linalg_utils.py:51-72; R is
//     discarded there)
//   * CGS2 re-orthonormalisation of an N x O walker panel, kept for the
//     job-free walker entry point, where no Gram or triangle workspace exists.
#include "afq_internal.h"
#include "group.h"

namespace afq {
namespace {

__device__ inline double2 cmul(double2 a, double2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// 1 / (a.x + i a.y) without branches or divisions: conj(a) / |a|^2 on a
// scaled by the power of two 2^-e (exact) that brings max(|re|, |im|) into
// [0.5, 1), so |a|^2 can neither overflow nor underflow; 1 / |a|^2 from the
// hardware reciprocal and two Newton steps (within about an ulp).
__device__ inline double2 crecip_fast(double2 a) {
  const int e = __builtin_amdgcn_frexp_exp(fmax(fabs(a.x), fabs(a.y)));
  const double xs = __builtin_amdgcn_ldexp(a.x, -e), ys = __builtin_amdgcn_ldexp(a.y, -e);
  const double q = fma(xs, xs, ys * ys);
  double r = __builtin_amdgcn_rcp(q);
  r = fma(r, fma(-q, r, 1.0), r);
  r = fma(r, fma(-q, r, 1.0), r);
  const double rs = __builtin_amdgcn_ldexp(r, -e);
  return {xs * rs, -ys * rs};
}
__device__ inline double2 crecip(double2 a) {
  // Smith's algorithm for 1/(a.x + i a.y)
  if (fabs(a.x) >= fabs(a.y)) {
    const double r = a.y / a.x, den = a.x + a.y * r;
    return {1.0 / den, -r / den};
  }
  const double r = a.x / a.y, den = a.x * r + a.y;
  return {r / den, -1.0 / den};
}

// Gauss-Jordan determinant and inverse, one 256-thread work-group per matrix,
// partial pivoting on the largest |re|+|im| of the pivot column (LAPACK
// izamax measure).  blockIdx.y selects one of up to two matrix sets (the two
// spins' overlap matrices factor in one launch, so a CU holds two work-groups
// at W = 256).
//
// n <= 64 (gj_reg_body): the matrix stays in registers.  Lane i of every wave
// holds row i, wave w the columns 16 w .. 16 w + 15 (one panel).  Rows are
// never swapped: the pivot of column k is the first (lowest) not yet used row
// of maximal measure (a DPP maximum and a ballot).  Panel P is factored by
// wave P alone, column by column, on its own 16 columns (the pivot row through
// LDS within the wave; column k replaced by the unit vector of its pivot row,
// so one branch-free update leaves a_p / d in the pivot row and a - (f / d) a_p
// elsewhere).  The panel's steps compose to M = I + sum_t U_t e_{p_t}^T, and
// because of the unit vectors the panel's columns end up holding M e_{p_t}:
// after one barrier every other wave applies M to its columns as one rank-16
// update, a += U a[p_t] (rows p_t as they were before the panel).  Four
// barriers per matrix instead of one per column.  The result R relates to the
// inverse through the pivot sequence piv (row piv[m] was the pivot of column
// m): A^-1[pinv[r]][piv[c]] = R[r][c], pinv the inverse permutation; det A =
// sign(piv) prod of the pivots (the determinant alone runs the same sweep).
//
// n > 64 (gj_det_inv_body, n <= 256): the matrix is factored in global memory,
// in the inverse's buffer or the set's scratch (it stays in L2), with explicit
// row interchanges (the first row of maximal measure at or below k).  Thread t
// owns the columns t % 64 + 64 c for the rows t/64, t/64 + 4, ..., so the
// elimination needs no index division and reads the pivot row and factor
// column as LDS broadcasts.  With an inverse requested the sweep is
// Gauss-Jordan in place (the inverse ends with its columns permuted by the row
// interchanges, undone at the end); for the determinant alone only the
// trailing rows and columns are eliminated (the pivots, hence the determinant,
// are the same).  NMAX = kDetInvMaxN; each thread covers NMAX / 64 columns
// (and pivot-search rows); n <= 64 matrices of such a launch use dynamic LDS.
constexpr int kLdsN = 64;  // largest n factored in registers (and LDS)

__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long v = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(v & 0xffffffffLL), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// max(v, v of the lane that DPP control CTRL names), rows outside ROWS kept
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_fmax(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)(b & 0xffffffffLL), (int)(b & 0xffffffffLL), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, ROWS, 0xF, false);
  return fmax(v, __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo)));
}

// maximum of v over the wave (NaN ignored): within quads, half rows and rows,
// then rows 0-1 into row 1 and 2-3 into row 3 (row_bcast 15), rows 0-1 into
// row 3 (row_bcast 31); lane 63 holds the result
__device__ __forceinline__ double wave_fmax(double v) {
  v = dpp_fmax<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
  v = dpp_fmax<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
  v = dpp_fmax<0x141, 0xF>(v);  // row_half_mirror
  v = dpp_fmax<0x140, 0xF>(v);  // row_mirror
  v = dpp_fmax<0x142, 0xA>(v);  // row_bcast 15
  v = dpp_fmax<0x143, 0xC>(v);  // row_bcast 31
  return readlane_d(v, 63);
}

// orders a wave's LDS accesses (they are performed in issue order) without a
// work-group barrier
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void gj_reg_body(const DetInvSet& set, long b) {
  __shared__ double2 ubuf[2][16][64];   // panel columns minus the unit vectors, U[t][row], by panel parity
  __shared__ double2 prow[4][16][16];   // per wave: [column][t] of the panel's pivot rows before the panel
  __shared__ double2 rrow[16];          // the panel wave's current pivot row
  __shared__ double2 pdet[4];           // product of each panel's pivots
  __shared__ int piv[64], pinv[64];
  const int n = set.n;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int c0 = 16 * wave;  // this wave's columns c0 .. c0 + 15
  const bool want_inv = set.inv != nullptr;
  const double2* src = set.S + b * set.strideS;
  for (int e = tid; e < 4 * 16 * 16; e += 256) (&prow[0][0][0])[e] = double2{0.0, 0.0};
  double2 a[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int c = c0 + t;
    a[t] = (lane < n && c < n) ? src[lane * n + c] : double2{0.0, 0.0};
  }
  unsigned long long used = 0;  // rows already used as pivots (wave-uniform)
  const int npan = (n + 15) >> 4;
  for (int P = 0; P < npan; ++P) {
    const int np = min(16, n - 16 * P);
    double2(&U)[16][64] = ubuf[P & 1];
    if (wave == P) {
      // ---- the panel, by this wave alone: Gauss-Jordan steps on its 16 columns
      double2 dp = {1.0, 0.0};
      int pj[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        pj[t] = -1;
        if (t < np) {
          const double2 x = a[t];
          const bool cand = lane < n && !((used >> lane) & 1ull);
          const double v = cand ? fabs(x.x) + fabs(x.y) : -1.0;
          const double m = wave_fmax(v);
          // lowest row of maximal measure (a column of NaNs falls back to the
          // lowest unused row, so p is always a valid row)
          const unsigned long long best = __ballot(cand && v == m), open = __ballot(cand);
          const int p = __builtin_ctzll(best ? best : open);
          pj[t] = p;
          const double2 d = {readlane_d(x.x, p), readlane_d(x.y, p)};
          dp = cmul(dp, d);
          used |= 1ull << p;
          // column t becomes the unit vector of row p: the common update
          // below then leaves 1 / d in row p and -f / d in the others
          const bool isp = lane == p;
          a[t] = isp ? double2{1.0, 0.0} : double2{0.0, 0.0};
          if (isp) {
#pragma unroll
            for (int s = 0; s < 16; ++s) rrow[s] = a[s];
          }
          const double2 dinv = crecip_fast(d);
          const double sc = isp ? 0.0 : 1.0;
          const double2 fd = cmul(x, dinv);
          const double2 g = isp ? dinv : double2{-fd.x, -fd.y};
          wave_lds_sync();
          double2 r[16];
#pragma unroll
          for (int s = 0; s < 16; ++s) r[s] = rrow[s];
#pragma unroll
          for (int s = 0; s < 16; ++s) {
            const double2 y = a[s];
            a[s].x = fma(g.x, r[s].x, fma(-g.y, r[s].y, sc * y.x));
            a[s].y = fma(g.x, r[s].y, fma(g.y, r[s].x, sc * y.y));
          }
          wave_lds_sync();
        }
      }
      if (lane == 0) {
        pdet[P] = dp;
#pragma unroll
        for (int t = 0; t < 16; ++t)
          if (t < np) piv[16 * P + t] = pj[t];
      }
      // the panel's transform M = I + sum_t U_t e_{p_t}^T: column t of the
      // panel now holds M e_{p_t} (zero past the panel's width)
#pragma unroll
      for (int t = 0; t < 16; ++t) U[t][lane] = (lane == pj[t]) ? double2{a[t].x - 1.0, a[t].y} : a[t];
    }
    __syncthreads();
    if (wave != P) {
      int pj[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) pj[t] = t < np ? __builtin_amdgcn_readfirstlane(piv[16 * P + t]) : -1;
#pragma unroll
      for (int t = 0; t < 16; ++t)
        if (pj[t] >= 0) used |= 1ull << pj[t];
      if (c0 < n) {
        // ---- apply the panel to this wave's columns: a += U a[p_t] (rows
        // p_t before the panel, through LDS: the lane holding row p_t writes it)
        int jj = -1;
#pragma unroll
        for (int t = 0; t < 16; ++t) jj = (pj[t] == lane) ? t : jj;
        if (jj >= 0) {
#pragma unroll
          for (int s = 0; s < 16; ++s) prow[wave][s][jj] = a[s];
        }
        wave_lds_sync();
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          const double2 u = U[t][lane];
#pragma unroll
          for (int s = 0; s < 16; ++s) {
            const double2 r = prow[wave][s][t];
            a[s].x = fma(u.x, r.x, fma(-u.y, r.y, a[s].x));
            a[s].y = fma(u.x, r.y, fma(u.y, r.x, a[s].y));
          }
        }
        wave_lds_sync();
      }
    }
  }
  __syncthreads();  // piv and pdet complete
  if (tid == 0) {
    double2 dprod = {1.0, 0.0};
    for (int q = 0; q < npan; ++q) dprod = cmul(dprod, pdet[q]);
    // sign of the pivot permutation: n minus its number of cycles transpositions
    unsigned long long seen = 0;
    int cycles = 0;
    for (int s0 = 0; s0 < n; ++s0) {
      if ((seen >> s0) & 1ull) continue;
      ++cycles;
      for (int q = s0; !((seen >> q) & 1ull); q = piv[q]) seen |= 1ull << q;
    }
    const bool neg = ((n - cycles) & 1) != 0;
    set.det[b] = neg ? double2{-dprod.x, -dprod.y} : dprod;
  }
  if (!want_inv) return;
  if (tid < n) pinv[piv[tid]] = tid;
  __syncthreads();
  double2* dst = set.inv + b * set.strideInv;
  if (lane < n) {
    const long row = (long)pinv[lane] * n;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int c = c0 + t;
      if (c < n) dst[row + piv[c]] = a[t];
    }
  }
}

template <int NMAX>
__device__ __forceinline__ void gj_det_inv_body(const DetInvSet& set, long b, double2* __restrict__ lds) {
  constexpr int kColsPer = NMAX / 64;
  __shared__ double2 prow[NMAX];
  __shared__ double2 fcol[NMAX];
  __shared__ int piv[NMAX];
  __shared__ int pk;
  const int n = set.n;
  double2* __restrict__ det = set.det;
  double2* __restrict__ inv = set.inv;
  const int tid = threadIdx.x;
  const int j = tid & 63, rg = tid >> 6;
  const double2* src = set.S + b * set.strideS;
  const int nn = n * n;
  const bool want_inv = inv != nullptr;
  double2* a = lds;
  if (n > kLdsN) a = want_inv ? inv + b * set.strideInv : set.work + b * set.strideS;
  if (a != src)
    for (int i = tid; i < nn; i += 256) a[i] = src[i];
  __syncthreads();

  double2 dprod = {1.0, 0.0};
  bool neg = false;
  for (int k = 0; k < n; ++k) {
    if (tid < 64) {
      double v = -1.0;
      int idx = tid;
#pragma unroll
      for (int m = 0; m < kColsPer; ++m) {
        const int r = tid + 64 * m;
        if (r >= k && r < n) {
          const double2 x = a[r * n + k];
          const double vr = fabs(x.x) + fabs(x.y);
          if (vr > v) {  // rows ascend: a tie keeps the lower row
            v = vr;
            idx = r;
          }
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > v || (ov == v && oi < idx)) {
          v = ov;
          idx = oi;
        }
      }
      if (tid == 0) {
        pk = idx;
        piv[k] = idx;
      }
    }
    __syncthreads();
    const int p = pk;
    const double2 d = a[p * n + k];
    const double2 dinv = crecip(d);
    if (tid == 0) {
      dprod = cmul(dprod, d);
      if (p != k) neg = !neg;
    }
    // read: the pivot row (scaled) and the rows to swap (group 0), the factor
    // column after the swap (group 1)
    const int jlo = want_inv ? 0 : k;
    double2 xp[kColsPer], xk[kColsPer], fc[kColsPer];
#pragma unroll
    for (int c = 0; c < kColsPer; ++c) {
      const int jc = j + 64 * c;
      xp[c] = xk[c] = fc[c] = double2{0.0, 0.0};
      if (rg == 0 && jc >= jlo && jc < n) {
        xp[c] = a[p * n + jc];
        xk[c] = a[k * n + jc];
      } else if (rg == 1 && jc < n) {
        fc[c] = (jc == k) ? d : a[(jc == p ? k : jc) * n + k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kColsPer; ++c) {
      const int jc = j + 64 * c;
      if (rg == 0 && jc >= jlo && jc < n) {
        const double2 r = (jc == k) ? dinv : cmul(xp[c], dinv);
        prow[jc] = r;
        a[k * n + jc] = r;
        if (p != k) a[p * n + jc] = xk[c];
      } else if (rg == 1 && jc < n) {
        fcol[jc] = fc[c];
      }
    }
    __syncthreads();
    const int i0 = want_inv ? rg : k + 1 + rg;
#pragma unroll
    for (int c = 0; c < kColsPer; ++c) {
      const int jc = j + 64 * c;
      if (jc < n && jc >= jlo) {
        const double2 rk = prow[jc];
#pragma unroll 4
        for (int i = i0; i < n; i += 4) {
          if (i == k) continue;
          const double2 f = fcol[i];
          double2* x = &a[i * n + jc];
          if (jc == k) {
            *x = double2{-(f.x * rk.x - f.y * rk.y), -(f.x * rk.y + f.y * rk.x)};
          } else {
            double2 y = *x;
            y.x -= f.x * rk.x - f.y * rk.y;
            y.y -= f.x * rk.y + f.y * rk.x;
            *x = y;
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) det[b] = neg ? double2{-dprod.x, -dprod.y} : dprod;
  if (inv) {
    // undo the column permutation: each thread owns whole rows, so no barrier
    // is needed between the sequential swaps.
    for (int i = tid; i < n; i += 256) {
      for (int k = n - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k) {
          const double2 t = a[i * n + k];
          a[i * n + k] = a[i * n + p];
          a[i * n + p] = t;
        }
      }
    }
    __syncthreads();
    double2* dst = inv + b * set.strideInv;
    if (dst != a)
      for (int i = tid; i < nn; i += 256) dst[i] = a[i];
  }
}

template <int NMAX>
__global__ __launch_bounds__(256) void gj_det_inv_kernel(DetInvSet s0, DetInvSet s1) {
  extern __shared__ double2 lds_gj[];
  if constexpr (NMAX == kLdsN)
    gj_reg_body(blockIdx.y == 0 ? s0 : s1, blockIdx.x);
  else
    gj_det_inv_body<NMAX>(blockIdx.y == 0 ? s0 : s1, blockIdx.x, lds_gj);
}

// Grouped: work-group i factors matrix i - prefix[j] of job j's set.
template <int NMAX>
__global__ __launch_bounds__(256) void gj_det_inv_grouped_kernel(const DetInvSet* __restrict__ sets,
                                                                 const int* __restrict__ prefix, int njobs) {
  extern __shared__ double2 lds_gj[];
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const DetInvSet set = sets[j];
  if constexpr (NMAX == kLdsN)
    gj_reg_body(set, item - prefix[j]);
  else
    gj_det_inv_body<NMAX>(set, item - prefix[j], lds_gj);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// CGS2 on one panel per work-group, in global memory (the panel stays in L2).
// Column j: two passes of r = Q[:, :j]^H a_j, a_j -= Q r, then a_j /= |a_j|.
// The resulting R has a positive real diagonal.
__global__ __launch_bounds__(256) void cgs2_kernel(double2* __restrict__ Wp, long ld, long sW, int nr, int nc) {
  __shared__ double2 r[128];
  __shared__ double red[4];
  double2* Wb = Wp + (long)blockIdx.x * sW;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = 0; j < nc; ++j) {
    for (int pass = 0; pass < 2 && j > 0; ++pass) {
      for (int k = wave; k < j; k += 4) {
        double sr = 0.0, si = 0.0;
        for (int p = lane; p < nr; p += 64) {
          const double2 q = Wb[(long)p * ld + k], x = Wb[(long)p * ld + j];
          sr += q.x * x.x + q.y * x.y;
          si += q.x * x.y - q.y * x.x;
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) r[k] = double2{sr, si};
      }
      __syncthreads();
      for (int p = tid; p < nr; p += 256) {
        double2 x = Wb[(long)p * ld + j];
        for (int k = 0; k < j; ++k) {
          const double2 q = Wb[(long)p * ld + k], rk = r[k];
          x.x -= q.x * rk.x - q.y * rk.y;
          x.y -= q.x * rk.y + q.y * rk.x;
        }
        Wb[(long)p * ld + j] = x;
      }
      __syncthreads();
    }
    double s = 0.0;
    for (int p = tid; p < nr; p += 256) {
      const double2 x = Wb[(long)p * ld + j];
      s += x.x * x.x + x.y * x.y;
    }
    s = wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    const double scale = 1.0 / sqrt(red[0] + red[1] + red[2] + red[3]);
    for (int p = tid; p < nr; p += 256) {
      double2 x = Wb[(long)p * ld + j];
      x.x *= scale;
      x.y *= scale;
      Wb[(long)p * ld + j] = x;
    }
    __syncthreads();
  }
}

// Upper Cholesky factor R (S = R^H R, positive real diagonal) of `count`
// Hermitian positive-definite n x n matrices (n <= 256), one 256-thread
// work-group per matrix, in dynamic LDS up to n = 64 and in place in R above;
// thread t owns the columns t % 64 + 64 c for the rows t/64, t/64 + 4, ...
// The strict lower triangle of R is written as zero.
template <int NMAX>
__device__ __forceinline__ void cholesky_upper_body(const double2* __restrict__ S, long strideS, int n,
                                                    double2* __restrict__ R, long strideR, long b,
                                                    double2* __restrict__ lds) {
  constexpr int kColsPer = NMAX / 64;
  const int tid = threadIdx.x;
  const int j = tid & 63, rg = tid >> 6;
  const double2* src = S + b * strideS;
  double2* dst = R + b * strideR;
  const int nn = n * n;
  double2* a = lds;
  if constexpr (NMAX > kLdsN)
    if (n > kLdsN) a = dst;
  if (a != src)
    for (int i = tid; i < nn; i += 256) a[i] = src[i];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double dinv = 1.0 / sqrt(a[k * n + k].x);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kColsPer; ++c) {
      const int jc = j + 64 * c;
      if (rg == 0 && jc >= k && jc < n) {
        const double2 x = a[k * n + jc];
        a[k * n + jc] = (jc == k) ? double2{1.0 / dinv, 0.0} : double2{x.x * dinv, x.y * dinv};
      }
    }
    __syncthreads();
    if constexpr (NMAX == kLdsN) {
      // reads first, then the updates (as in the Gauss-Jordan sweep)
      if (j > k && j < n) {
        const double2 rkj = a[k * n + j];
        double2 xv[16], rv[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          const int i = k + 1 + rg + 4 * t;
          xv[t] = rv[t] = double2{0.0, 0.0};
          if (i <= j) {
            rv[t] = a[k * n + i];
            xv[t] = a[i * n + j];
          }
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          const int i = k + 1 + rg + 4 * t;
          if (i <= j) {  // S_ij -= conj(R_ki) R_kj
            xv[t].x -= rv[t].x * rkj.x + rv[t].y * rkj.y;
            xv[t].y -= rv[t].x * rkj.y - rv[t].y * rkj.x;
            a[i * n + j] = xv[t];
          }
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < kColsPer; ++c) {
        const int jc = j + 64 * c;
        if (jc > k && jc < n) {
          const double2 rkj = a[k * n + jc];
          for (int i = k + 1 + rg; i <= jc; i += 4) {
            const double2 rki = a[k * n + i];  // S_ij -= conj(R_ki) R_kj
            double2 x = a[i * n + jc];
            x.x -= rki.x * rkj.x + rki.y * rkj.y;
            x.y -= rki.x * rkj.y - rki.y * rkj.x;
            a[i * n + jc] = x;
          }
        }
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < nn; i += 256) {
    const int r = i / n, c = i - r * n;
    if (c < r)
      dst[i] = double2{0.0, 0.0};
    else if (dst != a)
      dst[i] = a[i];
  }
}

// grid (count, sets): matrix blockIdx.x of set blockIdx.y
template <int NMAX>
__global__ __launch_bounds__(256) void cholesky_upper_kernel(CholSet c0, CholSet c1) {
  extern __shared__ double2 lds_ch[];
  const CholSet c = blockIdx.y == 0 ? c0 : c1;
  cholesky_upper_body<NMAX>(c.S, c.strideS, c.n, c.R, c.strideR, blockIdx.x, lds_ch);
}

template <int NMAX>
__global__ __launch_bounds__(256) void cholesky_grouped_kernel(const CholSet* __restrict__ sets,
                                                              const int* __restrict__ prefix, int njobs) {
  extern __shared__ double2 lds_ch[];
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const CholSet c = sets[j];
  cholesky_upper_body<NMAX>(c.S, c.strideS, c.n, c.R, c.strideR, item - prefix[j], lds_ch);
}

// Raises a kernel's dynamic LDS limit to 96 KB (once per kernel).
int raise_lds(const void* kernel, bool* done, const char* what) {
  if (*done) return AFQ_OK;
  hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
  if (e != hipSuccess) return fail(AFQ_E_HIP, std::string(what) + " LDS attribute: " + hipGetErrorString(e));
  *done = true;
  return AFQ_OK;
}


// Grouped copies / conjugate copies: work-group i covers elements
// [(i - prefix[j]) * 1024, +1024) of job j's vector.
__global__ __launch_bounds__(256) void vec_grouped_kernel(const VecJob* __restrict__ jobs,
                                                          const int* __restrict__ prefix, int njobs) {
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const VecJob v = jobs[j];
  const long base = (long)(item - prefix[j]) * 1024;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const long i = base + t * 256 + threadIdx.x;
    if (i < v.n) {
      const double2 x = v.src[i];
      v.dst[i] = v.conj ? double2{x.x, -x.y} : x;
    }
  }
}

}  // namespace

// Dynamic LDS of the det/inverse and Cholesky kernels: the largest matrix
// factored in LDS (n <= 64; larger ones are factored in global memory).
static size_t matrix_lds_bytes(int nmax) {
  const int m = std::min(nmax, kLdsN);
  return sizeof(double2) * m * m;
}

static int check_det_inv_set(const DetInvSet& s) {
  if (s.n < 1 || s.n > kDetInvMaxN)
    return fail(AFQ_E_UNSUPPORTED, "det/inverse kernel supports 1 <= n <= " + std::to_string(kDetInvMaxN));
  if (s.n > kLdsN && !s.inv && !s.work)
    return fail(AFQ_E_INVALID, "det/inverse: n > 64 without an inverse needs a scratch matrix per entry");
  return AFQ_OK;
}

template <int NMAX>
static int launch_det_inv(const DetInvSet* sets, int nsets, int count, size_t lds, hipStream_t stream) {
  static bool raised = false;
  AFQ_TRY(raise_lds(reinterpret_cast<const void*>(gj_det_inv_kernel<NMAX>), &raised, "det/inverse"));
  hipLaunchKernelGGL(gj_det_inv_kernel<NMAX>, dim3(count, nsets), dim3(256), lds, stream, sets[0],
                     sets[nsets > 1 ? 1 : 0]);
  return launch_status("gj_det_inv_kernel");
}

int batched_det_inv_sets(const DetInvSet* sets, int nsets, int count, hipStream_t stream) {
  if (nsets < 1 || nsets > 2) return fail(AFQ_E_INVALID, "det/inverse: one or two matrix sets");
  int nmax = 0;
  for (int i = 0; i < nsets; ++i) {
    AFQ_TRY(check_det_inv_set(sets[i]));
    nmax = std::max(nmax, sets[i].n);
  }
  if (count <= 0) return AFQ_OK;
  // n <= 64: the matrices stay in registers, no dynamic LDS
  return nmax <= kLdsN ? launch_det_inv<kLdsN>(sets, nsets, count, 0, stream)
                       : launch_det_inv<kDetInvMaxN>(sets, nsets, count, matrix_lds_bytes(nmax), stream);
}

template <int NMAX>
static int launch_det_inv_grouped(const GroupTab<DetInvSet>& t, size_t lds, hipStream_t stream) {
  static bool raised = false;
  AFQ_TRY(raise_lds(reinterpret_cast<const void*>(gj_det_inv_grouped_kernel<NMAX>), &raised, "det/inverse"));
  hipLaunchKernelGGL(gj_det_inv_grouped_kernel<NMAX>, dim3((unsigned)t.total), dim3(256), lds, stream, t.params,
                     t.prefix, t.njobs);
  return launch_status("gj_det_inv_grouped_kernel");
}

int det_inv_grouped(const GroupTab<DetInvSet>& t, int nmax, hipStream_t stream) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  if (nmax < 1 || nmax > kDetInvMaxN)
    return fail(AFQ_E_UNSUPPORTED, "det/inverse kernel supports 1 <= n <= " + std::to_string(kDetInvMaxN));
  return nmax <= kLdsN ? launch_det_inv_grouped<kLdsN>(t, 0, stream)
                       : launch_det_inv_grouped<kDetInvMaxN>(t, matrix_lds_bytes(nmax), stream);
}

int batched_det_inv(const double2* S, long strideS, int n, int count, double2* det, double2* inv, long strideInv,
                    hipStream_t stream) {
  const DetInvSet set{S, strideS, n, det, inv, strideInv};
  return batched_det_inv_sets(&set, 1, count, stream);
}

template <int NMAX>
static int launch_cholesky(const CholSet* sets, int nsets, int count, size_t lds, hipStream_t stream) {
  static bool raised = false;
  AFQ_TRY(raise_lds(reinterpret_cast<const void*>(cholesky_upper_kernel<NMAX>), &raised, "Cholesky"));
  hipLaunchKernelGGL(cholesky_upper_kernel<NMAX>, dim3(count, nsets), dim3(256), lds, stream, sets[0],
                     sets[nsets > 1 ? 1 : 0]);
  return launch_status("cholesky_upper_kernel");
}

int batched_cholesky_upper_sets(const CholSet* sets, int nsets, int count, hipStream_t stream) {
  if (nsets < 1 || nsets > 2) return fail(AFQ_E_INVALID, "Cholesky: one or two matrix sets");
  int nmax = 0;
  for (int i = 0; i < nsets; ++i) {
    if (sets[i].n < 1 || sets[i].n > kDetInvMaxN)
      return fail(AFQ_E_UNSUPPORTED, "Cholesky kernel supports 1 <= n <= " + std::to_string(kDetInvMaxN));
    nmax = std::max(nmax, sets[i].n);
  }
  if (count <= 0) return AFQ_OK;
  const size_t lds = matrix_lds_bytes(nmax);
  return nmax <= kLdsN ? launch_cholesky<kLdsN>(sets, nsets, count, lds, stream)
                       : launch_cholesky<kDetInvMaxN>(sets, nsets, count, lds, stream);
}

int batched_cholesky_upper(const double2* S, long strideS, int n, int count, double2* R, long strideR,
                           hipStream_t stream) {
  const CholSet set{S, strideS, n, R, strideR};
  return batched_cholesky_upper_sets(&set, 1, count, stream);
}

template <int NMAX>
static int launch_cholesky_grouped(const GroupTab<CholSet>& t, size_t lds, hipStream_t stream) {
  static bool raised = false;
  AFQ_TRY(raise_lds(reinterpret_cast<const void*>(cholesky_grouped_kernel<NMAX>), &raised, "Cholesky"));
  hipLaunchKernelGGL(cholesky_grouped_kernel<NMAX>, dim3((unsigned)t.total), dim3(256), lds, stream, t.params,
                     t.prefix, t.njobs);
  return launch_status("cholesky_grouped_kernel");
}

int cholesky_grouped(const GroupTab<CholSet>& t, int nmax, hipStream_t stream) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  if (nmax < 1 || nmax > kDetInvMaxN)
    return fail(AFQ_E_UNSUPPORTED, "Cholesky kernel supports 1 <= n <= " + std::to_string(kDetInvMaxN));
  const size_t lds = matrix_lds_bytes(nmax);
  return nmax <= kLdsN ? launch_cholesky_grouped<kLdsN>(t, lds, stream)
                       : launch_cholesky_grouped<kDetInvMaxN>(t, lds, stream);
}

int vec_grouped(const GroupTab<VecJob>& t, hipStream_t stream) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  hipLaunchKernelGGL(vec_grouped_kernel, dim3((unsigned)t.total), dim3(256), 0, stream, t.params, t.prefix, t.njobs);
  return launch_status("vec_grouped_kernel");
}


int batched_cgs2(double2* Wp, long ld, long sW, int n_rows, int n_cols, int count, hipStream_t stream) {
  if (n_cols < 1 || n_cols > 128 || n_cols > n_rows) return fail(AFQ_E_UNSUPPORTED, "CGS2 supports 1 <= nocc <= 128 <= N");
  if (count <= 0) return AFQ_OK;
  hipLaunchKernelGGL(cgs2_kernel, dim3(count), dim3(256), 0, stream, Wp, ld, sW, n_rows, n_cols);
  return launch_status("cgs2_kernel");
}

}  // namespace afq
