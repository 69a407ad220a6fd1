// Split-BF16 x3 products on pre-split operands (VHS and force bias).
//
//   C[b] (M x N) = alpha * A[b] (M x K) . B[b]^T (N x K)
//
// Both operands are stored as bf16 planes with k contiguous: a real operand as
// (hi, lo) = (bf16(x), bf16(x - hi)), a complex one as (re_hi, re_lo, im_hi,
// im_lo).  Element (row r, plane q, k) of A is at A + b*sA + r*lda + q*pA + k
// (same for B).  K is a multiple of 32 and zero-padded; rows past the valid
// ones up to the next multiple of 16 exist and hold zeros, and reads of rows
// beyond that are clamped (their outputs are discarded).  Each real product is
// hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with FP32 accumulation;
// the passes run over all of a wave's TM x TN blocks so consecutive MFMAs are
// independent.  Fragments are read straight from global memory: lane l reads
// the 8 consecutive k (one 16-byte load) of row 16*block + l%16 at offset
// 8*(l/16) of each 32-deep step, which is the MFMA's input layout.
//
// Outputs: complex64 C[b*sC + m*ldc + n] = alpha * acc (the VHS product,
// feeding the complex64 Taylor terms), or complex128 C = alpha * acc + beta * C
// through K slices written to a workspace and a fixed-order reduction (the
// force bias: few output tiles, a long K).
#include "afq_internal.h"
#include "gemm_bf3.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

namespace afq {
namespace {

__device__ inline v8bf neg8(v8bf x) {
  v4u u = __builtin_bit_cast(v4u, x);
  u ^= v4u{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(v8bf, u);
}
__device__ inline v4f mma(v8bf a, v8bf b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
template <bool AC, bool BC, int WM, int WN, int TM, int TN>
__global__ __launch_bounds__(64 * WM * WN, 2) void gemm_bf3_kernel(Bf3 p, int tiles_m, int tiles_n, int xcd_remap,
                                                                   long T8, int kchunk, double* ws) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  long tile = blockIdx.x;
  if (xcd_remap) tile = (tile & 7) * T8 + (tile >> 3);
  const int mt = (int)(tile % tiles_m);
  const long ntl = tile / tiles_m;
  if (ntl >= tiles_n) return;  // uniform per work-group
  const long b = blockIdx.y;
  const int m0 = mt * 16 * TM * WM + wm * TM * 16, n0 = (int)ntl * 16 * TN * WN + wn * TN * 16;
  const int r16 = lane & 15, q4 = lane >> 4;
  const __bf16* Ab = reinterpret_cast<const __bf16*>(p.A) + b * p.sA;
  const __bf16* Bb = reinterpret_cast<const __bf16*>(p.B) + b * p.sB;
  const v8bf* pa[TM];
  const v8bf* pb[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
    pa[i] = reinterpret_cast<const v8bf*>(Ab + (long)min(m0 + i * 16 + r16, p.a_rows - 1) * p.lda) + q4;
#pragma unroll
  for (int j = 0; j < TN; ++j)
    pb[j] = reinterpret_cast<const v8bf*>(Bb + (long)min(n0 + j * 16 + r16, p.b_rows - 1) * p.ldb) + q4;
  const long pA8 = p.pA / 8, pB8 = p.pB / 8;

  v4f accR[TM][TN], accI[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      accR[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
      accI[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  const int kz = blockIdx.z;
  const int kv0 = kz * kchunk / 8, kvn = min(p.K, (kz + 1) * kchunk) / 8;
  for (int kv = kv0; kv < kvn; kv += 4) {
    v8bf ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      ar[i] = pa[i][kv];
      if (AC) ai[i] = pa[i][kv + 2 * pA8];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      br[j] = pb[j][kv];
      if (BC) bi[j] = pb[j][kv + 2 * pB8];
    }
    // hi x hi
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        accR[i][j] = mma(ar[i], br[j], accR[i][j]);
        if (AC && BC) accR[i][j] = mma(neg8(ai[i]), bi[j], accR[i][j]);
        if (BC) accI[i][j] = mma(ar[i], bi[j], accI[i][j]);
        if (AC) accI[i][j] = mma(ai[i], br[j], accI[i][j]);
      }
    // lo(A) x hi(B)
    {
      v8bf lr[TM], li[TM];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        lr[i] = pa[i][kv + pA8];
        if (AC) li[i] = pa[i][kv + 3 * pA8];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          accR[i][j] = mma(lr[i], br[j], accR[i][j]);
          if (AC && BC) accR[i][j] = mma(neg8(li[i]), bi[j], accR[i][j]);
          if (BC) accI[i][j] = mma(lr[i], bi[j], accI[i][j]);
          if (AC) accI[i][j] = mma(li[i], br[j], accI[i][j]);
        }
    }
    // hi(A) x lo(B)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      br[j] = pb[j][kv + pB8];
      if (BC) bi[j] = pb[j][kv + 3 * pB8];
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        accR[i][j] = mma(ar[i], br[j], accR[i][j]);
        if (AC && BC) accR[i][j] = mma(neg8(ai[i]), bi[j], accR[i][j]);
        if (BC) accI[i][j] = mma(ar[i], bi[j], accI[i][j]);
        if (AC) accI[i][j] = mma(ai[i], br[j], accI[i][j]);
      }
  }

  // accumulator element v of lane l is (4(l/16) + v, l%16)
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + i * 16 + 4 * q4 + v;
        const int n = n0 + j * 16 + r16;
        if (m >= p.M || n >= p.N) continue;
        const double xr = accR[i][j][v], xi = accI[i][j][v];
        if (ws) {  // raw partial of K slice kz (complex128), reduced by the split-K kernel
          reinterpret_cast<double2*>(ws)[((long)kz * p.M + m) * p.N + n] = double2{xr, xi};
          continue;
        }
        const double rr = p.alpha.x * xr - p.alpha.y * xi, ri = p.alpha.x * xi + p.alpha.y * xr;
        reinterpret_cast<float2*>(p.C)[b * p.sC + (long)m * p.ldc + n] = float2{(float)rr, (float)ri};
      }
}

template <bool AC, bool BC, int WM, int WN, int TM, int TN>
int launch(const Bf3& p, hipStream_t s) {
  const int BM = 16 * TM * WM, BN = 16 * TN * WN;
  const int tm = (p.M + BM - 1) / BM, tn = (p.N + BN - 1) / BN;
  const long tiles = (long)tm * tn;
  const int remap = (p.batch == 1 && tiles >= 64) ? 1 : 0;
  const long T8 = remap ? (tiles + 7) / 8 : 0;
  const long gx = remap ? T8 * 8 : tiles;
  if (gx > 0x7fffffffL || p.batch > 65535) return fail(AFQ_E_UNSUPPORTED, "gemm_bf3: grid too large");
  int ksplit = 1, kchunk = p.K;
  double* ws = nullptr;
  if (p.c_f64) {
    // complex128 output goes through the fixed-order split-K reduction (which
    // applies alpha and beta); split K until about 1024 work-groups are in flight
    if (p.batch != 1 || !p.ws) return fail(AFQ_E_INVALID, "gemm_bf3: complex128 output needs batch 1 and a workspace");
    ksplit = (int)std::max<long>(1, std::min<long>((1024 + tiles - 1) / tiles, p.K / 256));
    while (ksplit > 1 && (long)ksplit * p.M * p.N * 2 > p.ws_elems) --ksplit;
    if ((long)ksplit * p.M * p.N * 2 > p.ws_elems) return fail(AFQ_E_INVALID, "gemm_bf3: workspace too small");
    kchunk = ((p.K + ksplit - 1) / ksplit + 31) / 32 * 32;
    ksplit = (p.K + kchunk - 1) / kchunk;
    ws = static_cast<double*>(p.ws);
  }
  hipLaunchKernelGGL((gemm_bf3_kernel<AC, BC, WM, WN, TM, TN>), dim3((unsigned)gx, p.batch, ksplit),
                     dim3(64 * WM * WN), 0, s, p, tm, tn, remap, T8, kchunk, ws);
  AFQ_TRY(launch_status("gemm_bf3_kernel"));
  if (!ws) return AFQ_OK;
  Gemm g;
  g.M = p.M;
  g.N = p.N;
  g.batch = 1;
  g.C = p.C;
  g.ldc = p.ldc;
  g.alpha = p.alpha;
  g.beta = p.beta;
  return gemm_splitk_reduce(g, ws, ksplit, true, s);
}

// ------------------------------------------------------------------ operand preparation
// dst[r][q][k] (k < kpad, rows r < rows) from a row-major complex (or real) double
// matrix src[r][k] (k < K): bf16 planes, zero for k >= K.
__global__ void split_rows_kernel(const double* __restrict__ src, int cplx, long rows, int K, long lds, int kpad,
                                  long ldd, long pd, __bf16* __restrict__ dst) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= rows * kpad) return;
  const long r = idx / kpad;
  const int k = (int)(idx % kpad);
  double xr = 0.0, xi = 0.0;
  if (k < K) {
    if (cplx) {
      xr = src[2 * (r * lds + k)];
      xi = src[2 * (r * lds + k) + 1];
    } else {
      xr = src[r * lds + k];
    }
  }
  __bf16* d = dst + r * ldd + k;
  const float fr = (float)xr;
  const __bf16 h = (__bf16)fr;
  d[0] = h;
  d[pd] = (__bf16)(fr - (float)h);
  if (cplx) {
    const float fi = (float)xi;
    const __bf16 hi = (__bf16)fi;
    d[2 * pd] = hi;
    d[3 * pd] = (__bf16)(fi - (float)hi);
  }
}

// L (G x NN, row-major doubles) -> Lt[n][q][g] (bf16 hi/lo planes, g < gpad, zero for g >= G)
__global__ void split_chol_t_kernel(const double* __restrict__ L, int G, long NN, int gpad, long ldd, long pd,
                                    __bf16* __restrict__ dst) {
  __shared__ double t[32][33];
  const long n0 = (long)blockIdx.x * 32;
  const int g0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {
    const int g = g0 + y;
    const long n = n0 + tx;
    t[y][tx] = (g < G && n < NN) ? L[(long)g * NN + n] : 0.0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const long n = n0 + y;
    const int g = g0 + tx;
    if (n < NN && g < gpad) {
      const float x = (float)t[tx][y];
      const __bf16 h = (__bf16)x;
      __bf16* d = dst + n * ldd + g;
      d[0] = h;
      d[pd] = (__bf16)(x - (float)h);
    }
  }
}

}  // namespace

int gemm_bf3(const Bf3& p, hipStream_t s) {
  if (p.M <= 0 || p.N <= 0 || p.batch <= 0) return AFQ_OK;
  if (p.K <= 0 || p.K % 32 != 0) return fail(AFQ_E_INVALID, "gemm_bf3: K must be a positive multiple of 32");
  if (p.a_rows < 1 || p.b_rows < 1) return fail(AFQ_E_INVALID, "gemm_bf3: operand row counts must be positive");
  if (!p.a_cplx || p.b_cplx) return fail(AFQ_E_INVALID, "gemm_bf3: complex A and real B expected");
  // 2 waves of 4 x 4 blocks (64 x 128 tiles): 1.06 ms vs 1.28 ms for 2 x 2 waves of
  // 2 x 4 blocks on the VHS product at W=256, G=2000, N=300
  return launch<true, false, 1, 2, 4, 4>(p, s);
}

int k_split_rows_bf16(const double* src, bool cplx, long rows, int K, long lds, int kpad, long ldd, long pd,
                      uint16_t* dst, hipStream_t s) {
  const long n = rows * (long)kpad;
  if (n <= 0) return AFQ_OK;
  hipLaunchKernelGGL(split_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, cplx ? 1 : 0, rows, K,
                     lds, kpad, ldd, pd, reinterpret_cast<__bf16*>(dst));
  return launch_status("split_rows_kernel");
}

int k_split_chol_t_bf16(const double* L, int G, long NN, int gpad, long ldd, long pd, uint16_t* dst, hipStream_t s) {
  const long nb = (NN + 31) / 32;
  if (nb > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "split_chol_t: too large");
  hipLaunchKernelGGL(split_chol_t_kernel, dim3((unsigned)nb, (gpad + 31) / 32), dim3(256), 0, s, L, G, NN, gpad, ldd,
                     pd, reinterpret_cast<__bf16*>(dst));
  return launch_status("split_chol_t_kernel");
}

}  // namespace afq
