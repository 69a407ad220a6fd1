// Split-BF16 x3 products on pre-split operands (VHS and force bias).
//
//   C[b] (M x N) = alpha * A[b] (M x K, complex) . B[b]^T (N x K, real)
//
// Both operands are stored as bf16 planes with k contiguous in 32-deep steps: a
// real operand as (hi, lo) = (bf16(x), bf16(x - hi)), a complex one as (re_hi,
// re_lo, im_hi, im_lo).  Element (row r, plane q, k) of A is at A + b*sA +
// r*lda + (k/32)*kA + q*pA + k%32 (same for B), so the planes may be stored
// apart or interleaved step by step.  K is a multiple of 32 and zero-padded; rows past the valid
// ones up to the next multiple of 16 exist and hold zeros, and reads of rows
// beyond that are clamped (their outputs are discarded).  Each real product is
// hi*hi + lo*hi + hi*lo on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Outputs: complex64 C[b*sC + m*ldc + n] = alpha * acc (the VHS product,
// feeding the complex64 Taylor terms), or complex128 C = alpha * acc + beta * C
// through K slices written to a workspace and a fixed-order reduction (the
// force bias: few output tiles, a long K).  A K slice's partial is its FP32
// accumulator, stored as complex64 (exactly) and summed in FP64.  The K slices
// of a split product are spread over the XCDs, each slice's tiles on one XCD,
// so its A and B slabs are fetched into one L2 (tiles in the other order made
// every XCD read all of A).
#include <type_traits>

#include "afq_internal.h"
#include "gemm_bf3.h"
#include "group.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace afq {
namespace {

__device__ inline v4f mma(v8bf a, v8bf b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// Complex A against real B, staged through LDS.  The complex rows of A act as real rows 2r (re)
// and 2r + 1 (im), so a wave's 16 x 16 blocks cover re and im rows alike and
// the accumulator rows 4(l/16) + v pair up as (re, im) of one output element.
// Per 32-deep k step the work-group copies its A (BMR real rows) and B (BN
// rows) hi/lo planes to LDS once (16-byte loads), and the waves read their
// fragments from there; two LDS stages, one barrier per step.  In LDS each
// 16-row block and plane is one 1 KB fragment of 64 16-byte slots in MFMA
// lane order, slot 16 q + (r ^ (2 q + 8 plane)) holding row r, k-group q: a
// fragment read covers 16 consecutive slots per 16 lanes, and the stash
// writes (8 lanes per row, both planes) also fall in 16 distinct bank groups
// per 16 lanes (rows of 144 bytes read by row were 1.4x slower).
// Per accumulator the products run in a fixed order (hi*hi, lo*hi, hi*lo per
// step, steps ascending), so results do not depend on the tiling.
constexpr int LFRAG = 512;  // bf16 per LDS fragment: 16 rows x 32 k

// bf16 offset of (row, plane, k-group part) in a stage of 16-row fragments
__device__ __forceinline__ int lslot(int row, int plane, int part) {
  return ((row >> 4) * 2 + plane) * LFRAG + (16 * part + ((row & 15) ^ (2 * part + 8 * plane))) * 8;
}

template <int BMR, int BN, int WM, int WN>
__device__ __forceinline__ void gemm_bf3_body(const Bf3& p, int tiles_m, int tiles_n, long tile, long b, int kz,
                                              int kchunk, float* ws, __bf16* __restrict__ lds) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BMR / (16 * WM), TN = BN / (16 * WN);
  constexpr int CA = BMR * 8, CB = BN * 8;  // 16-byte chunks per stage
  constexpr int RA = (CA + NT - 1) / NT, RB = (CB + NT - 1) / NT;
  // every thread's chunks exist when the tile's chunks divide evenly: no
  // guard (and no branch) around those loads and stores
  constexpr bool EVENA = CA % NT == 0, EVENB = CB % NT == 0;
  static_assert(TM * 16 * WM == BMR && TN * 16 * WN == BN, "tile");
  __bf16* As = lds;                // [2][BMR / 16][2 planes][LFRAG]
  __bf16* Bs = lds + 2 * BMR * 64;  // [2][BN / 16][2 planes][LFRAG]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int mt = (int)(tile % tiles_m);
  const long ntl = tile / tiles_m;
  if (ntl >= tiles_n) return;  // uniform per work-group
  const int m0r = mt * BMR;  // first real row
  const int n0 = (int)ntl * BN;
  const __bf16* Ab = reinterpret_cast<const __bf16*>(p.A) + b * p.sA;
  const __bf16* Bb = reinterpret_cast<const __bf16*>(p.B) + b * p.sB;

  // global source and LDS destination of each chunk this thread copies
  const v8bf* ga[RA];
  int la[RA];
#pragma unroll
  for (int i = 0; i < RA; ++i) {
    const int c = tid + i * NT;
    const int row = c >> 3, plane = (c >> 2) & 1, part = c & 3;
    const int gr = m0r + row;
    const int w = min(gr >> 1, p.a_rows - 1);
    const int q = 2 * (gr & 1) + plane;  // re_hi, re_lo, im_hi, im_lo
    ga[i] = reinterpret_cast<const v8bf*>(Ab + (long)w * p.lda + q * p.pA + part * 8);
    la[i] = c < CA ? lslot(row, plane, part) : -1;
  }
  const v8bf* gb[RB];
  int lb[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int c = tid + i * NT;
    const int row = c >> 3, plane = (c >> 2) & 1, part = c & 3;
    const int gn = min(n0 + row, p.b_rows - 1);
    gb[i] = reinterpret_cast<const v8bf*>(Bb + (long)gn * p.ldb + plane * p.pB + part * 8);
    lb[i] = c < CB ? lslot(row, plane, part) : -1;
  }

  v4f acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  const int ks0 = kz * kchunk / 32, nk = min(p.K, (kz + 1) * kchunk) / 32;
  const long kva = p.kA >> 3, kvb = p.kB >> 3;  // k step strides in units of 8 bf16
  // Two register sets: a step's slabs are fetched two steps before they are
  // used, stashed into LDS one step before.  Past the last step the fetches
  // repeat it (no branch around a load, so the compiler's count of the loads
  // still in flight stays exact).
  v8bf ra[2][RA], rb[2][RB];
  auto fetch = [&](int ks, v8bf(&xa)[RA], v8bf(&xb)[RB]) {
    ks = min(ks, nk - 1);
#pragma unroll
    for (int i = 0; i < RA; ++i)
      if (EVENA || la[i] >= 0) xa[i] = ga[i][ks * kva];
#pragma unroll
    for (int i = 0; i < RB; ++i)
      if (EVENB || lb[i] >= 0) xb[i] = gb[i][ks * kvb];
  };
  auto stash = [&](int st, const v8bf(&xa)[RA], const v8bf(&xb)[RB]) {
    __bf16* a = As + st * BMR * 64;
    __bf16* bb = Bs + st * BN * 64;
#pragma unroll
    for (int i = 0; i < RA; ++i)
      if (EVENA || la[i] >= 0) *reinterpret_cast<v8bf*>(a + la[i]) = xa[i];
#pragma unroll
    for (int i = 0; i < RB; ++i)
      if (EVENB || lb[i] >= 0) *reinterpret_cast<v8bf*>(bb + lb[i]) = xb[i];
  };

  const int r16 = lane & 15, q4 = lane >> 4, q8 = 8 * q4;
  // this lane's slot in a hi and a lo fragment
  const int sh = (16 * q4 + (r16 ^ (2 * q4))) * 8, sl = LFRAG + (16 * q4 + (r16 ^ (2 * q4 + 8))) * 8;
  // step ks on LDS stage P (= (ks - ks0) mod 2) and register set P, whose
  // slabs were stashed the step before: fetch step ks + 2 into set P, multiply
  // stage P, stash set 1 - P (step ks + 1) into stage 1 - P
  auto step = [&](int ks, auto par) {
    constexpr int P = decltype(par)::value;
    fetch(ks + 2, ra[P], rb[P]);
    __builtin_amdgcn_sched_barrier(0);  // the fetch is issued first, not sunk towards its use
    const __bf16* a = As + P * BMR * 64 + wm * TM * 2 * LFRAG;
    const __bf16* bb = Bs + P * BN * 64 + wn * TN * 2 * LFRAG;
    v8bf ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      ah[i] = *reinterpret_cast<const v8bf*>(a + i * 2 * LFRAG + sh);
      al[i] = *reinterpret_cast<const v8bf*>(a + i * 2 * LFRAG + sl);
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      bh[j] = *reinterpret_cast<const v8bf*>(bb + j * 2 * LFRAG + sh);
      bl[j] = *reinterpret_cast<const v8bf*>(bb + j * 2 * LFRAG + sl);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = mma(ah[i], bh[j], acc[i][j]);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = mma(al[i], bh[j], acc[i][j]);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = mma(ah[i], bl[j], acc[i][j]);
    __builtin_amdgcn_sched_barrier(0);
    stash(1 - P, ra[1 - P], rb[1 - P]);
    __syncthreads();
  };

  fetch(ks0, ra[0], rb[0]);
  fetch(ks0 + 1, ra[1], rb[1]);
  stash(0, ra[0], rb[0]);
  __syncthreads();
  int ks = ks0;
  for (; ks + 1 < nk; ks += 2) {
    step(ks, std::integral_constant<int, 0>{});
    step(ks + 1, std::integral_constant<int, 1>{});
  }
  if (ks < nk) step(ks, std::integral_constant<int, 0>{});

  // accumulator element v of lane l is real row 4(l/16) + v of the block:
  // v = 0, 1 are (re, im) of one output row, v = 2, 3 of the next
  float2* C = reinterpret_cast<float2*>(p.C) + b * p.sC;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int rr = m0r + (wm * TM + i) * 16 + q8 / 2 + 2 * h;
        const int m = rr >> 1;
        const int n = n0 + (wn * TN + j) * 16 + r16;
        if (m >= p.M || n >= p.N) continue;
        const double xr = acc[i][j][2 * h], xi = acc[i][j][2 * h + 1];
        if (ws) {  // raw partial of K slice kz (the FP32 accumulator as complex64), reduced in FP64
          reinterpret_cast<float2*>(ws)[((long)kz * p.M + m) * p.N + n] = float2{acc[i][j][2 * h], acc[i][j][2 * h + 1]};
          continue;
        }
        const double orr = p.alpha.x * xr - p.alpha.y * xi, oi = p.alpha.x * xi + p.alpha.y * xr;
        C[(long)m * p.ldc + n] = float2{(float)orr, (float)oi};
      }
}

// Split-K order: item i runs on XCD i mod 8 and takes, of the K slices
// kz = XCD, XCD + 8, ..., slice (i / 8) / tiles and tile (i / 8) mod tiles
// (the M-tiles of an N-tile consecutive).  Returns false past the last slice.
__device__ __forceinline__ bool slice_item(long i, long tiles, int ksplit, long* tile, int* kz) {
  const long j = i >> 3;
  *kz = (int)(8 * (j / tiles) + (i & 7));
  *tile = j % tiles;
  return *kz < ksplit;
}

template <int BMR, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_bf3_lds_kernel(Bf3 p, int tiles_m, int tiles_n, int xcd_remap,
                                                                     long T8, int ksplit, int kchunk, float* ws) {
  extern __shared__ __align__(16) __bf16 lds_bf3[];
  long tile = blockIdx.x;
  int kz = 0;
  if (ws) {
    if (!slice_item(blockIdx.x, (long)tiles_m * tiles_n, ksplit, &tile, &kz)) return;  // uniform
  } else if (xcd_remap) {
    tile = (tile & 7) * T8 + (tile >> 3);
  }
  gemm_bf3_body<BMR, BN, WM, WN>(p, tiles_m, tiles_n, tile, blockIdx.y, kz, kchunk, ws, lds_bf3);
}

// Grouped (batch-1 products): work-group i belongs to job j with
// prefix[j] <= i < prefix[j + 1] (a multiple of 8), in gemm_bf3_lds_kernel's
// orders: split-K jobs by slice_item, the others by the XCD-aware tile remap.
__global__ __launch_bounds__(512) void gemm_bf3_grouped_kernel(const Bf3Job* __restrict__ jobs,
                                                               const int* __restrict__ prefix, int njobs) {
  extern __shared__ __align__(16) __bf16 lds_bf3[];
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const Bf3Job J = jobs[j];
  const int local = item - prefix[j];
  const long tiles = (long)J.tiles_m * J.tiles_n;
  long t = local;
  int kz = 0;
  if (J.ws) {
    if (!slice_item(local, tiles, J.ksplit, &t, &kz)) return;  // uniform
  } else {
    if (J.T8 > 0) t = (t & 7) * J.T8 + (t >> 3);
    if (t >= tiles) return;
  }
  gemm_bf3_body<128, 128, 2, 4>(J.p, J.tiles_m, J.tiles_n, t, 0, kz, J.kchunk, J.ws, lds_bf3);
}

// C = alpha sum_kz partial_kz + beta C for a complex128 output from complex64
// K-slice partials, summed in FP64 in slice order.
__device__ __forceinline__ void reduce_f32_body(const Gemm& g, const float* __restrict__ ws, int ksplit, long idx) {
  const long plane = (long)g.M * g.N;
  if (idx >= plane) return;
  const int m = (int)(idx / g.N), n = (int)(idx % g.N);
  double2 sum = {0.0, 0.0};
  for (int z = 0; z < ksplit; ++z) {
    const float2 v = reinterpret_cast<const float2*>(ws)[(long)z * plane + idx];
    sum.x += (double)v.x;
    sum.y += (double)v.y;
  }
  const double2 a = g.alpha;
  double2 r = {a.x * sum.x - a.y * sum.y, a.x * sum.y + a.y * sum.x};
  double2* Cc = static_cast<double2*>(g.C) + (long)m * g.ldc + n;
  if (g.beta.x != 0.0 || g.beta.y != 0.0) {
    const double2 o = *Cc, bt = g.beta;
    r.x += bt.x * o.x - bt.y * o.y;
    r.y += bt.x * o.y + bt.y * o.x;
  }
  *Cc = r;
}

__global__ __launch_bounds__(256) void reduce_f32_kernel(Gemm g, const float* ws, int ksplit) {
  reduce_f32_body(g, ws, ksplit, blockIdx.x * 256L + threadIdx.x);
}

__global__ __launch_bounds__(256) void reduce_f32_grouped_kernel(const ReduceJob* __restrict__ jobs,
                                                                 const int* __restrict__ prefix, int njobs) {
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  reduce_f32_body(jobs[j].g, jobs[j].ws, jobs[j].ksplit, (long)(item - prefix[j]) * 256 + threadIdx.x);
}

// K slices for a complex128 output: split until about 1024 work-groups are in
// flight; the fixed-order reduction applies alpha and beta.  A partial takes
// one complex64 (the size of one double of the workspace) per output element.
int bf3_ksplit(const Bf3& p, long tiles, int* kchunk) {
  if (p.batch != 1 || !p.ws) return fail(AFQ_E_INVALID, "gemm_bf3: complex128 output needs batch 1 and a workspace");
  int ksplit = (int)std::max<long>(1, std::min<long>((1024 + tiles - 1) / tiles, p.K / 256));
  while (ksplit > 1 && (long)ksplit * p.M * p.N > p.ws_elems) --ksplit;
  if ((long)ksplit * p.M * p.N > p.ws_elems) return fail(AFQ_E_INVALID, "gemm_bf3: workspace too small");
  *kchunk = ((p.K + ksplit - 1) / ksplit + 31) / 32 * 32;
  return (p.K + *kchunk - 1) / *kchunk;
}

// work-groups of a split-K product in slice_item's order
long slice_items(long tiles, int ksplit) { return 8L * ((ksplit + 7) / 8) * tiles; }

int bf3_reduce(const Bf3& p, float* ws, int ksplit, hipStream_t s) {
  Gemm g;
  g.M = p.M;
  g.N = p.N;
  g.batch = 1;
  g.C = p.C;
  g.ldc = p.ldc;
  g.alpha = p.alpha;
  g.beta = p.beta;
  const long n = (long)p.M * p.N;
  hipLaunchKernelGGL(reduce_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, ws, ksplit);
  return launch_status("reduce_f32_kernel");
}

template <int BMR, int BN, int WM, int WN>
int launch_lds(const Bf3& p, hipStream_t s) {
  const int tm = (2 * p.M + BMR - 1) / BMR, tn = (p.N + BN - 1) / BN;
  const long tiles = (long)tm * tn;
  const int remap = (p.batch == 1 && tiles >= 64) ? 1 : 0;
  const long T8 = remap ? (tiles + 7) / 8 : 0;
  const long gx = remap ? T8 * 8 : tiles;
  if (gx > 0x7fffffffL || p.batch > 65535) return fail(AFQ_E_UNSUPPORTED, "gemm_bf3: grid too large");
  int ksplit = 1, kchunk = p.K;
  float* ws = nullptr;
  long items = gx;
  if (p.c_f64) {
    ksplit = bf3_ksplit(p, tiles, &kchunk);
    if (ksplit < 1) return ksplit;
    ws = static_cast<float*>(p.ws);
    items = slice_items(tiles, ksplit);
    if (items > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "gemm_bf3: grid too large");
  }
  const size_t lds = sizeof(__bf16) * 2 * (BMR + BN) * 64;
  static bool raised = false;
  if (!raised && lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(gemm_bf3_lds_kernel<BMR, BN, WM, WN>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(AFQ_E_HIP, std::string("gemm_bf3 LDS attribute: ") + hipGetErrorString(e));
    raised = true;
  }
  hipLaunchKernelGGL((gemm_bf3_lds_kernel<BMR, BN, WM, WN>), dim3((unsigned)items, p.batch), dim3(64 * WM * WN), lds,
                     s, p, tm, tn, remap, T8, ksplit, kchunk, ws);
  AFQ_TRY(launch_status("gemm_bf3_lds_kernel"));
  return ws ? bf3_reduce(p, ws, ksplit, s) : AFQ_OK;
}

// ------------------------------------------------------------------ operand preparation
// dst[r][k / 32][q][k % 32] (strides ldd, kst, pd; k < kpad, rows r < rows)
// from a row-major complex (or real) double matrix src[r][k] (k < K): bf16
// planes, zero for k >= K.
__device__ __forceinline__ void split_rows_body(const double* __restrict__ src, int cplx, long rows, int K, long lds,
                                                int kpad, long ldd, long pd, long kst, __bf16* __restrict__ dst,
                                                long idx) {
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
  __bf16* d = dst + r * ldd + (k >> 5) * kst + (k & 31);
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

__global__ void split_rows_kernel(const double* __restrict__ src, int cplx, long rows, int K, long lds, int kpad,
                                  long ldd, long pd, long kst, __bf16* __restrict__ dst) {
  split_rows_body(src, cplx, rows, K, lds, kpad, ldd, pd, kst, dst, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

__global__ void split_rows_grouped_kernel(const RowSplitJob* __restrict__ jobs, const int* __restrict__ prefix,
                                          int njobs) {
  const int item = blockIdx.x;
  const int j = group_find(prefix, njobs, item);
  const RowSplitJob J = jobs[j];
  split_rows_body(J.src, J.cplx, J.rows, J.K, J.lds, J.kpad, J.ldd, J.pd, J.kst, reinterpret_cast<__bf16*>(J.dst),
                  (long)(item - prefix[j]) * 256 + threadIdx.x);
}

// L (G x NN, row-major doubles) -> Lt[n][g / 32][q][g % 32] (strides ldd,
// kst, pd; bf16 hi/lo planes, g < gpad, zero for g >= G) for output rows
// n < ncols; with sym_N > 0 output row n is packed column n of a symmetric L
// (sym_src), else column n itself.
__global__ void split_chol_t_kernel(const double* __restrict__ L, int G, long NN, long ncols, int sym_N, int gpad,
                                    long ldd, long pd, long kst, __bf16* __restrict__ dst) {
  __shared__ double t[32][33];
  const long n0 = (long)blockIdx.x * 32;
  const int g0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  {
    const long n = n0 + tx;
    const long src = n >= ncols ? -1 : sym_N > 0 ? sym_src(n, sym_N) : n;
    for (int y = ty; y < 32; y += 8) {
      const int g = g0 + y;
      t[y][tx] = (g < G && src >= 0) ? L[(long)g * NN + src] : 0.0;
    }
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const long n = n0 + y;
    const int g = g0 + tx;
    if (n < ncols && g < gpad) {
      const float x = (float)t[tx][y];
      const __bf16 h = (__bf16)x;
      __bf16* d = dst + n * ldd + (g >> 5) * kst + (g & 31);
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
  if (p.kA < 32 || p.kB < 32 || p.kA % 8 || p.kB % 8)
    return fail(AFQ_E_INVALID, "gemm_bf3: k step strides must be multiples of 8, at least 32");
  // 128 real rows (64 complex) x 128 columns per work-group, 2 x 4 waves of
  // 4 x 2 blocks: measured fastest of the LDS tilings on the VHS product
  // (W=256, N=300, G=2000: 0.63 ms; W=512, N=200, G=1200: 0.31 ms) and the force
  // bias (0.17 and 0.14 ms), against 1.12 / 0.56 / 0.29 / 0.25 ms for fragments
  // read straight from global memory
  return launch_lds<128, 128, 2, 4>(p, s);
}

int bf3_group_job(const Bf3& p, Bf3Job* out) {
  if (p.batch != 1) return fail(AFQ_E_INVALID, "gemm_bf3 grouped: batch-1 products only");
  if (p.K <= 0 || p.K % 32 != 0) return fail(AFQ_E_INVALID, "gemm_bf3: K must be a positive multiple of 32");
  if (p.kA < 32 || p.kB < 32 || p.kA % 8 || p.kB % 8)
    return fail(AFQ_E_INVALID, "gemm_bf3: k step strides must be multiples of 8, at least 32");
  Bf3Job j;
  j.p = p;
  j.tiles_m = (2 * p.M + 127) / 128;
  j.tiles_n = (p.N + 127) / 128;
  const long tiles = (long)j.tiles_m * j.tiles_n;
  j.T8 = tiles >= 64 ? (tiles + 7) / 8 : 0;
  j.tx = (int)(j.T8 > 0 ? j.T8 * 8 : tiles);
  j.ksplit = 1;
  j.kchunk = p.K;
  j.ws = nullptr;
  if (p.c_f64) {
    const int ks = bf3_ksplit(p, (long)j.tiles_m * j.tiles_n, &j.kchunk);
    if (ks < 1) return ks;
    j.ksplit = ks;
    j.ws = static_cast<float*>(p.ws);
  }
  *out = j;
  return AFQ_OK;
}

long bf3_group_items(const Bf3Job& j) {
  if (j.ws) return slice_items((long)j.tiles_m * j.tiles_n, j.ksplit);
  return ((long)j.tx + 7) / 8 * 8;
}

int reduce_f32_grouped(const GroupTab<ReduceJob>& t, hipStream_t s) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  hipLaunchKernelGGL(reduce_f32_grouped_kernel, dim3((unsigned)t.total), dim3(256), 0, s, t.params, t.prefix,
                     t.njobs);
  return launch_status("reduce_f32_grouped_kernel");
}

int gemm_bf3_grouped(const GroupTab<Bf3Job>& t, hipStream_t s) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  const size_t lds = sizeof(__bf16) * 2 * (128 + 128) * 64;
  static bool raised = false;
  if (!raised && lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(gemm_bf3_grouped_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(AFQ_E_HIP, std::string("gemm_bf3 LDS attribute: ") + hipGetErrorString(e));
    raised = true;
  }
  hipLaunchKernelGGL(gemm_bf3_grouped_kernel, dim3((unsigned)t.total), dim3(512), lds, s, t.params, t.prefix,
                     t.njobs);
  return launch_status("gemm_bf3_grouped_kernel");
}

int k_split_rows_bf16(const double* src, bool cplx, long rows, int K, long lds, int kpad, long ldd, long pd, long kst,
                      uint16_t* dst, hipStream_t s) {
  const long n = rows * (long)kpad;
  if (n <= 0) return AFQ_OK;
  hipLaunchKernelGGL(split_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, cplx ? 1 : 0, rows, K,
                     lds, kpad, ldd, pd, kst, reinterpret_cast<__bf16*>(dst));
  return launch_status("split_rows_kernel");
}

long row_split_items(const RowSplitJob& j) { return (j.rows * (long)j.kpad + 255) / 256; }

int split_rows_grouped(const GroupTab<RowSplitJob>& t, hipStream_t s) {
  if (t.njobs < 1 || t.total < 1) return AFQ_OK;
  hipLaunchKernelGGL(split_rows_grouped_kernel, dim3((unsigned)t.total), dim3(256), 0, s, t.params, t.prefix,
                     t.njobs);
  return launch_status("split_rows_grouped_kernel");
}

int k_split_chol_t_bf16(const double* L, int G, long NN, long ncols, int sym_N, int gpad, long ldd, long pd, long kst,
                        uint16_t* dst, hipStream_t s) {
  const long nb = (ncols + 31) / 32;
  if (nb > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "split_chol_t: too large");
  hipLaunchKernelGGL(split_chol_t_kernel, dim3((unsigned)nb, (gpad + 31) / 32), dim3(256), 0, s, L, G, NN, ncols,
                     sym_N, gpad, ldd, pd, kst, reinterpret_cast<__bf16*>(dst));
  return launch_status("split_chol_t_kernel");
}

}  // namespace afq
