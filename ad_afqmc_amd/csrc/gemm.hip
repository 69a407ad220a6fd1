// Batched FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// One 256-thread work-group computes a 64x64 tile of C; the four waves each own
// a 32x32 quadrant made of 2x2 MFMA 16x16 blocks.  K advances in slabs of 16
// staged through LDS.  Complex operands are split into real/imaginary planes
// in LDS, so a complex x real product costs two real MFMAs per block and a
// complex x complex product four, with FP64 accumulation throughout.
//
// MFMA f64 16x16x4 lane layout (gfx950): lane l feeds A[l%16][l/16] and
// B[l/16][l%16]; accumulator element v of lane l is D[l/16 + 4v][l%16].
#include "afq_internal.h"

namespace afq {
namespace {

constexpr int BM = 64, BN = 64, BK = 16, PAD = 1;

__device__ inline double2 cmul(double2 a, double2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

template <bool AC, bool BC, bool AT, bool BT>
__global__ __launch_bounds__(256) void gemm_f64_kernel(Gemm g) {
  constexpr bool CC = AC || BC;
  __shared__ double As[AC ? 2 : 1][BK][BM + PAD];
  __shared__ double Bs[BC ? 2 : 1][BK][BN + PAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const long b = blockIdx.z;
  const double* A = static_cast<const double*>(g.A) + b * g.sA * (AC ? 2 : 1);
  const double* B = static_cast<const double*>(g.B) + b * g.sB * (BC ? 2 : 1);

  v4d accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  for (int k0 = 0; k0 < g.K; k0 += BK) {
    // ---- stage A (BM x BK) into As[plane][k][m]
    if (!AT) {
      const int m = tid >> 2, kq = (tid & 3) * 4;
      const int gm = m0 + m;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = kq + t, gk = k0 + k;
        double re = 0.0, im = 0.0;
        if (gm < g.M && gk < g.K) {
          const long off = (long)gm * g.lda + gk;
          if (AC) {
            const double2 v = reinterpret_cast<const double2*>(A)[off];
            re = v.x;
            im = v.y;
          } else {
            re = A[off];
          }
        }
        As[0][k][m] = re;
        if (AC) As[AC ? 1 : 0][k][m] = im;
      }
    } else {
      const int k = tid >> 4, mq = (tid & 15) * 4;
      const int gk = k0 + k;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int m = mq + t, gm = m0 + m;
        double re = 0.0, im = 0.0;
        if (gm < g.M && gk < g.K) {
          const long off = (long)gk * g.lda + gm;
          if (AC) {
            const double2 v = reinterpret_cast<const double2*>(A)[off];
            re = v.x;
            im = v.y;
          } else {
            re = A[off];
          }
        }
        As[0][k][m] = re;
        if (AC) As[AC ? 1 : 0][k][m] = im;
      }
    }
    // ---- stage B (BK x BN) into Bs[plane][k][n]
    if (!BT) {
      const int k = tid >> 4, nq = (tid & 15) * 4;
      const int gk = k0 + k;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int n = nq + t, gn = n0 + n;
        double re = 0.0, im = 0.0;
        if (gn < g.N && gk < g.K) {
          const long off = (long)gk * g.ldb + gn;
          if (BC) {
            const double2 v = reinterpret_cast<const double2*>(B)[off];
            re = v.x;
            im = v.y;
          } else {
            re = B[off];
          }
        }
        Bs[0][k][n] = re;
        if (BC) Bs[BC ? 1 : 0][k][n] = im;
      }
    } else {
      const int n = tid >> 2, kq = (tid & 3) * 4;
      const int gn = n0 + n;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = kq + t, gk = k0 + k;
        double re = 0.0, im = 0.0;
        if (gn < g.N && gk < g.K) {
          const long off = (long)gn * g.ldb + gk;
          if (BC) {
            const double2 v = reinterpret_cast<const double2*>(B)[off];
            re = v.x;
            im = v.y;
          } else {
            re = B[off];
          }
        }
        Bs[0][k][n] = re;
        if (BC) Bs[BC ? 1 : 0][k][n] = im;
      }
    }
    __syncthreads();

#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = As[0][kr][wm * 32 + i * 16 + c];
        ai[i] = AC ? As[AC ? 1 : 0][kr][wm * 32 + i * 16 + c] : 0.0;
        br[i] = Bs[0][kr][wn * 32 + i * 16 + c];
        bi[i] = BC ? Bs[BC ? 1 : 0][kr][wn * 32 + i * 16 + c] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accR[i][j], 0, 0, 0);
          if (AC && BC) {
            accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[i], bi[j], accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], accI[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], accI[i][j], 0, 0, 0);
          } else if (AC) {
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], accI[i][j], 0, 0, 0);
          } else if (BC) {
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], accI[i][j], 0, 0, 0);
          }
        }
    }
    __syncthreads();
  }

  // ---- epilogue
  const bool has_beta = (g.beta.x != 0.0 || g.beta.y != 0.0);
  double* C = static_cast<double*>(g.C) + b * g.sC * (CC ? 2 : 1);
  double* Dp = g.D ? static_cast<double*>(g.D) + b * g.sD * (CC ? 2 : 1) : nullptr;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * v;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const long off = (long)m * g.ldc + n;
        if (CC) {
          double2 r = cmul(g.alpha, double2{accR[i][j][v], accI[i][j][v]});
          double2* Cc = reinterpret_cast<double2*>(C);
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
          if (has_beta) r += g.beta.x * C[off];
          C[off] = r;
          if (Dp) Dp[(long)m * g.ldd + n] += r;
        }
      }
}

template <bool AC, bool BC, bool AT, bool BT>
int launch(const Gemm& g, hipStream_t s) {
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, g.batch);
  hipLaunchKernelGGL((gemm_f64_kernel<AC, BC, AT, BT>), grid, dim3(256), 0, s, g);
  return launch_status("gemm_f64_kernel");
}

template <bool AC, bool BC>
int launch_t(const Gemm& g, hipStream_t s) {
  if (g.a_trans) return g.b_trans ? launch<AC, BC, true, true>(g, s) : launch<AC, BC, true, false>(g, s);
  return g.b_trans ? launch<AC, BC, false, true>(g, s) : launch<AC, BC, false, false>(g, s);
}

}  // namespace

int gemm_f64(const Gemm& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return AFQ_OK;
  if (g.K <= 0) return fail(AFQ_E_INVALID, "gemm_f64: K must be positive");
  if (g.batch > 65535 || (g.M + BM - 1) / BM > 65535)
    return fail(AFQ_E_UNSUPPORTED, "gemm_f64: grid too large");
  if (g.a_cplx) return g.b_cplx ? launch_t<true, true>(g, s) : launch_t<true, false>(g, s);
  return g.b_cplx ? launch_t<false, true>(g, s) : launch_t<false, false>(g, s);
}

}  // namespace afq
