// Elementwise / reduction kernels of the walker step (FP64), and the fused
// exchange-energy kernel.  Launch helpers are declared in kernels.h.
#include "kernels.h"
#include "group.h"

namespace afq {
namespace {

__device__ inline double2 cmul(double2 a, double2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline double2 cadd(double2 a, double2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline double2 cdiv(double2 a, double2 b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, den = b.x + b.y * r;
    return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
  }
  const double r = b.x / b.y, den = b.x * r + b.y;
  return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}
__device__ inline double2 cexp(double2 z) {
  const double e = exp(z.x);
  double s, c;
  sincos(z.y, &s, &c);
  return {e * c, e * s};
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
// Block (256-thread) sum of a complex value; every thread gets the result.
__device__ inline double2 block_sum2(double2 v, double2* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double2 r = {0.0, 0.0};
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = cadd(r, red[w]);
  return r;
}

// ------------------------------------------------------------------ setup
__global__ void sym_h1_kernel(const double* in, double* out, int n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nn = (long)n * n;
  if (i >= 2 * nn) return;
  const long s = i / nn, r = (i % nn) / n, c = i % n;
  out[i] = (in[s * nn + r * n + c] + in[s * nn + c * n + r]) / 2.0;
}

// out = alpha*(x + y) + beta*(u + v) with absent terms skipped (elementwise).
__global__ void axpby4_kernel(double* out, const double* x, const double* y, double alpha, const double* u,
                              const double* v, double beta, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = alpha * (x[i] + (y ? y[i] : 0.0));
  if (u) r += beta * (u[i] + (v ? v[i] : 0.0));
  out[i] = r;
}

// mf_shifts = i*m, mf_r = Re(i*mf) = -m, h0_prop = -h0 - sum(mf^2)/2.
__global__ void mf_kernel(const double* m, int G, double h0, double2* mf, double* mfr, double2* h0_prop) {
  __shared__ double2 red[4];
  double2 acc = {0.0, 0.0};
  for (int g = threadIdx.x; g < G; g += 256) {
    const double2 z = {0.0 * m[g], m[g]};  // 1j * m
    mf[g] = z;
    mfr[g] = -z.y;  // real part of 1j*z
    const double2 z2 = cmul(z, z);
    acc = cadd(acc, z2);
  }
  acc = block_sum2(acc, red);
  if (threadIdx.x == 0) *h0_prop = double2{-h0 - acc.x / 2.0, -acc.y / 2.0};
}

// out[i][gl][k] = L[g0+gl][i][k]  (permute a gamma chunk for the v0 GEMM)
__global__ void permute_chunk_kernel(const double* L, double* out, int n, int g0, int gch) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long total = (long)n * gch * n;
  if (idx >= total) return;
  const int k = idx % n;
  const int gl = (idx / n) % gch;
  const int i = idx / ((long)n * gch);
  out[idx] = L[((long)(g0 + gl) * n + i) * n + k];
}

__global__ void norm1_kernel(const double* A, int n, double* out) {
  __shared__ double red[4];
  double best = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(A[(long)i * n + j]);
    best = fmax(best, s);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) best = fmax(best, __shfl_xor(best, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ void scaled_identity_kernel(double* out, const double* A, double scale, int n, int mode) {
  // mode 0: out = I ; mode 1: out = scale*A
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nn = (long)n * n;
  if (i >= nn) return;
  if (mode == 0)
    out[i] = (i / n == i % n) ? 1.0 : 0.0;
  else
    out[i] = scale * A[i];
}

// ------------------------------------------------------------------ measurement
__global__ void overlap_combine_kernel(const double2* d0, const double2* d1, double2* out, int W) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  out[w] = d1 ? cmul(d0[w], d1[w]) : cmul(d0[w], d0[w]);
}

// ene1[w] = sum_s f_s sum_{ij} G_s[w,i,j] rot_h1_s[i,j]
__device__ __forceinline__ void ene1_body(const double2* G0, const double* rh0, int n0, const double2* G1,
                                          const double* rh1, int n1, int N, double f0, double2* out, long w) {
  __shared__ double2 red[4];
  double2 acc = {0.0, 0.0};
  const long l0 = (long)n0 * N;
  const double2* g0 = G0 + w * l0;
  for (long i = threadIdx.x; i < l0; i += 256) {
    acc.x += f0 * g0[i].x * rh0[i];
    acc.y += f0 * g0[i].y * rh0[i];
  }
  if (G1) {
    const long l1 = (long)n1 * N;
    const double2* g1 = G1 + w * l1;
    for (long i = threadIdx.x; i < l1; i += 256) {
      acc.x += g1[i].x * rh1[i];
      acc.y += g1[i].y * rh1[i];
    }
  }
  acc = block_sum2(acc, red);
  if (threadIdx.x == 0) out[w] = acc;
}

__global__ void ene1_kernel(const double2* G0, const double* rh0, int n0, const double2* G1, const double* rh1, int n1,
                            int N, double f0, double2* out) {
  ene1_body(G0, rh0, n0, G1, rh1, n1, N, f0, out, blockIdx.x);
}

// E = h0 + ene1 + 1/2 sum_g (c_up + c_dn)^2 - 1/2 (K_up + K_dn)
__device__ __forceinline__ void energy_finalize_body(const double2* ene1, const double2* c0, const double2* c1,
                                                     const double2* k0, const double2* k1, int Gn, int nch0, int nch1,
                                                     double h0, double2* out, long w) {
  __shared__ double2 red[4];
  double2 J = {0.0, 0.0};
  for (int g = threadIdx.x; g < Gn; g += 256) {
    const double2 a = c0[w * Gn + g];
    const double2 b = c1 ? c1[w * Gn + g] : a;
    const double2 s = cadd(a, b);
    J = cadd(J, cmul(s, s));
  }
  J = block_sum2(J, red);
  __syncthreads();
  double2 K = {0.0, 0.0};
  for (int c = threadIdx.x; c < nch0; c += 256) K = cadd(K, k0[w * nch0 + c]);
  if (k1) {
    for (int c = threadIdx.x; c < nch1; c += 256) K = cadd(K, k1[w * nch1 + c]);
  } else {
    K.x *= 2.0;
    K.y *= 2.0;
  }
  K = block_sum2(K, red);
  if (threadIdx.x == 0) {
    const double2 e1 = ene1[w];
    out[w] = double2{h0 + e1.x + 0.5 * J.x - 0.5 * K.x, e1.y + 0.5 * J.y - 0.5 * K.y};
  }
}

__global__ void energy_finalize_kernel(const double2* ene1, const double2* c0, const double2* c1, const double2* k0,
                                       const double2* k1, int Gn, int nch0, int nch1, double h0, double2* out) {
  energy_finalize_body(ene1, c0, c1, k0, k1, Gn, nch0, nch1, h0, out, blockIdx.x);
}

// ------------------------------------------------------------------ propagation
// s = -sqrt(dt) (i fb - mf); xs = x - s; shift = sum xs*mf; fbt = sum x*s - s*s/2
__device__ __forceinline__ void field_shift_body(const double2* fb, const double* x, const double2* mf, int Gn,
                                                 double sqdt, double2* xs, double2* terms, long w) {
  __shared__ double2 red[4];
  double2 sh = {0.0, 0.0}, ft = {0.0, 0.0};
  for (int g = threadIdx.x; g < Gn; g += 256) {
    const double2 f = fb[w * Gn + g];
    const double2 m = mf[g];
    const double xv = x[w * Gn + g];
    const double2 ifb = {-f.y, f.x};
    const double2 s = {-sqdt * (ifb.x - m.x), -sqdt * (ifb.y - m.y)};
    const double2 z = {xv - s.x, -s.y};
    xs[w * Gn + g] = z;
    sh = cadd(sh, cmul(z, m));
    const double2 ss = cmul(s, s);
    ft.x += xv * s.x - ss.x / 2.0;
    ft.y += xv * s.y - ss.y / 2.0;
  }
  sh = block_sum2(sh, red);
  __syncthreads();
  ft = block_sum2(ft, red);
  if (threadIdx.x == 0) {
    terms[2 * w] = sh;
    terms[2 * w + 1] = ft;
  }
}

__global__ void field_shift_kernel(const double2* fb, const double* x, const double2* mf, int Gn, double sqdt,
                                   double2* xs, double2* terms) {
  field_shift_body(fb, x, mf, Gn, sqdt, xs, terms, blockIdx.x);
}

// Phaseless importance function, clipping, weight update and the
// population-control shift (single work-group of 1024 threads).
__device__ __forceinline__ void weight_update_body(const double2* ov_new, double2* overlaps, const double2* terms,
                                                   double* weights, double* scalars, const double2* h0_prop, double dt,
                                                   int W) {
  __shared__ double red[16];
  const double sqdt = sqrt(dt);
  const double pcs = scalars[1];
  const double2 h0p = *h0_prop;
  double local = 0.0;
  for (int w = threadIdx.x; w < W; w += blockDim.x) {
    const double2 on = ov_new[w], oo = overlaps[w];
    const double2 ratio = cdiv(on, oo);
    const double2 sh = terms[2 * w], ft = terms[2 * w + 1];
    const double2 arg = {-sqdt * sh.x + ft.x + dt * (pcs + h0p.x), -sqdt * sh.y + ft.y + dt * h0p.y};
    const double2 imp = cmul(cexp(arg), ratio);
    const double2 ph_arg = cmul(cexp(double2{-sqdt * sh.x, -sqdt * sh.y}), ratio);
    const double theta = atan2(ph_arg.y, ph_arg.x);
    double ph = hypot(imp.x, imp.y) * cos(theta);
    if (isnan(ph)) ph = 0.0;
    if (ph < 1.0e-3) ph = 0.0;
    if (ph > 100.0) ph = 0.0;
    double wn = ph * weights[w];
    if (wn > 100.0) wn = 0.0;
    weights[w] = wn;
    overlaps[w] = on;
    local += wn;
  }
  local = wave_sum(local);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    scalars[1] = scalars[0] - 0.1 * log(s / W) / dt;
  }
}

__global__ __launch_bounds__(1024) void weight_update_kernel(const double2* ov_new, double2* overlaps,
                                                             const double2* terms, double* weights, double* scalars,
                                                             const double2* h0_prop, double dt, int W) {
  weight_update_body(ov_new, overlaps, terms, weights, scalars, h0_prop, dt, W);
}

// Block bookkeeping after the step loop: energy clipping around e_estimate,
// weighted block energy, and pop_control_ene_shift <- 0.9 shift + 0.1 E_block.
__device__ __forceinline__ void block_finalize_body(const double2* energies, const double* weights, double* scalars,
                                                    double clip, int W, double* out) {
  __shared__ double r1[16], r2[16];
  const double e_est = scalars[0];
  double sw = 0.0, swe = 0.0;
  for (int w = threadIdx.x; w < W; w += blockDim.x) {
    double e = energies[w].x;
    if (fabs(e - e_est) > clip) e = e_est;
    const double wt = weights[w];
    sw += wt;
    swe += e * wt;
  }
  sw = wave_sum(sw);
  swe = wave_sum(swe);
  if ((threadIdx.x & 63) == 0) {
    r1[threadIdx.x >> 6] = sw;
    r2[threadIdx.x >> 6] = swe;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      a += r1[i];
      b += r2[i];
    }
    const double be = b / a;
    out[0] = be;
    out[1] = a;
    scalars[1] = 0.9 * scalars[1] + 0.1 * be;
  }
}

__global__ __launch_bounds__(1024) void block_finalize_kernel(const double2* energies, const double* weights,
                                                              double* scalars, double clip, int W, double* out) {
  block_finalize_body(energies, weights, scalars, clip, W, out);
}

// ------------------------------------------------------------------ population control
// Comb resampling (single work-group): cum = cumsum|w| (sequential per thread
// chunk + block scan), z_i = total (i + zeta)/W, idx_i = first j with cum_j >= z_i.
__device__ __forceinline__ void sr_index_body(double* weights, const double* zeta, int W, double* cum, int32_t* idx) {
  __shared__ double part[1024];
  const int t = threadIdx.x, T = blockDim.x;
  const int chunk = (W + T - 1) / T;
  const int lo = min(W, t * chunk), hi = min(W, lo + chunk);
  double s = 0.0;
  for (int i = lo; i < hi; ++i) s += fabs(weights[i]);
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    double run = 0.0;
    for (int i = 0; i < T; ++i) {
      const double v = part[i];
      part[i] = run;
      run += v;
    }
  }
  __syncthreads();
  double run = part[t];
  for (int i = lo; i < hi; ++i) {
    run += fabs(weights[i]);
    cum[i] = run;
  }
  __syncthreads();
  const double total = cum[W - 1];
  const double z0 = *zeta;
  for (int i = t; i < W; i += T) {
    const double z = total * ((double)i + z0) / (double)W;
    int a = 0, b = W;  // first index with cum >= z
    while (a < b) {
      const int m = (a + b) >> 1;
      if (cum[m] < z)
        a = m + 1;
      else
        b = m;
    }
    idx[i] = min(a, W - 1);
  }
  __syncthreads();
  const double avg = total / W;
  for (int i = t; i < W; i += T) weights[i] = avg;
}

__global__ __launch_bounds__(1024) void sr_index_kernel(double* weights, const double* zeta, int W, double* cum,
                                                        int32_t* idx) {
  sr_index_body(weights, zeta, W, cum, idx);
}

__global__ void gather_walkers_kernel(const double2* src, double2* dst, const int32_t* idx, long per) {
  const long w = blockIdx.y;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= per) return;
  dst[w * per + i] = src[(long)idx[w] * per + i];
}

// ------------------------------------------------------------------ RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the Box-Muller transform.
__device__ inline uint4 philox4x32_10(uint4 c, uint2 k) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    c = uint4{hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0};
    k.x += W0;
    k.y += W1;
  }
  return c;
}

// counter_dev (when not null) supplies the stream counter from device memory.
__device__ __forceinline__ void normal_fill_body(double* out, long n, uint64_t seed, uint64_t counter,
                                                 const uint64_t* counter_dev, long pair) {
  if (2 * pair >= n) return;
  if (counter_dev) counter = *counter_dev;
  const uint4 r = philox4x32_10(uint4{(uint32_t)pair, (uint32_t)((uint64_t)pair >> 32), (uint32_t)counter,
                                      (uint32_t)(counter >> 32)},
                                uint2{(uint32_t)seed, (uint32_t)(seed >> 32)});
  // two 53-bit uniforms in (0, 1]
  const uint64_t a = ((uint64_t)r.x << 21) ^ ((uint64_t)r.y >> 11);
  const uint64_t b = ((uint64_t)r.z << 21) ^ ((uint64_t)r.w >> 11);
  const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(b & ((1ull << 53) - 1)) + 0.5) * (1.0 / 9007199254740992.0);
  const double rad = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  out[2 * pair] = rad * c;
  if (2 * pair + 1 < n) out[2 * pair + 1] = rad * s;
}

__global__ void normal_fill_kernel(double* out, long n, uint64_t seed, uint64_t counter, const uint64_t* counter_dev) {
  normal_fill_body(out, n, seed, counter, counter_dev, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

// SR offset zeta ~ U[0, 1) of block `*counter_dev`: Philox on a key distinct from
// the field stream's, then the counter advances to the next block.
// counter_dev (when not null) supplies the counter from device memory.
__device__ __forceinline__ void sr_zeta_body(double* zeta, uint64_t seed, uint64_t counter,
                                             const uint64_t* counter_dev) {
  const uint64_t c = counter_dev ? *counter_dev : counter;
  const uint64_t key = seed ^ 0x9E3779B97F4A7C15ull;
  const uint4 r = philox4x32_10(uint4{0xFFFFFFFFu, 0xFFFFFFFFu, (uint32_t)c, (uint32_t)(c >> 32)},
                                uint2{(uint32_t)key, (uint32_t)(key >> 32)});
  const uint64_t a = ((uint64_t)r.x << 21) ^ ((uint64_t)r.y >> 11);
  *zeta = (double)(a & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
}

__global__ void sr_zeta_kernel(double* zeta, uint64_t seed, uint64_t counter, const uint64_t* counter_dev) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sr_zeta_body(zeta, seed, counter, counter_dev);
}

__global__ void counter_incr_kernel(uint64_t* counter_dev) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter_dev += 1;
}

__global__ void to_f32_kernel(const double* src, float* dst, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// dst = E - I in FP32 for an n x n row-major E (the identity is subtracted in FP64).
__global__ void minus_identity_f32_kernel(const double* src, float* dst, int n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)n * n) return;
  const double v = src[i] - ((i / n) == (i % n) ? 1.0 : 0.0);
  dst[i] = (float)v;
}

// out[0] = max |L_g,ij - L_g,ji|, out[1] = max |L_g,ij| as IEEE bit patterns
// (non-negative doubles order like their unsigned bits); out must start at zero.
__global__ void sym_check_kernel(const double* __restrict__ L, long total, int N, unsigned long long* out) {
  const long NN = (long)N * N;
  double dmax = 0.0, amax = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long g = t / NN, q = t - g * NN;
    const int i = (int)(q / N), j = (int)(q - (long)i * N);
    const double a = L[t];
    amax = fmax(amax, fabs(a));
    if (j > i) dmax = fmax(dmax, fabs(a - L[g * NN + (long)j * N + i]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, off));
    amax = fmax(amax, __shfl_xor(amax, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&out[0], (unsigned long long)__double_as_longlong(dmax));
    atomicMax(&out[1], (unsigned long long)__double_as_longlong(amax));
  }
}

template <typename T>
__global__ void sym_pack_kernel(const double* __restrict__ L, int N, long npk, long total, T* __restrict__ out) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long g = t / npk, p = t - g * npk;
  const long src = sym_src(p, N);
  out[t] = src >= 0 ? (T)L[g * (long)N * N + src] : (T)0;
}

inline unsigned blocks_for(long n, int t) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------------------------ grouped block kernels
#define GJ_FIND                                   \
  const int item = blockIdx.x;                    \
  const int jn = group_find(prefix, njobs, item); \
  const BlockJob& J = jobs[jn];                   \
  const long local = item - prefix[jn];

__global__ void g_rng_kernel(const BlockJob* __restrict__ jobs, const int* __restrict__ prefix, int njobs) {
  GJ_FIND
  normal_fill_body(J.fields, (long)J.nsteps * J.W * J.G, J.seed, 0, J.counter, local * 256 + threadIdx.x);
}

__global__ void g_field_shift_kernel(const BlockJob* __restrict__ jobs, const int* __restrict__ prefix, int njobs,
                                     int step) {
  GJ_FIND
  field_shift_body(J.fb, J.fields + (long)step * J.W * J.G, J.mf, J.G, J.sqdt, J.xs, J.terms, local);
}

__global__ void g_overlap_kernel(const BlockJob* __restrict__ jobs, const int* __restrict__ prefix, int njobs,
                                 int to_overlaps) {
  GJ_FIND
  const long w = local * 256 + threadIdx.x;
  if (w >= J.W) return;
  (to_overlaps ? J.overlaps : J.ov_new)[w] = cmul(J.det[w], J.det[w]);
}

__global__ __launch_bounds__(1024) void g_weight_update_kernel(const BlockJob* __restrict__ jobs,
                                                               const int* __restrict__ prefix, int njobs) {
  GJ_FIND
  (void)local;
  weight_update_body(J.ov_new, J.overlaps, J.terms, J.weights, J.scalars, J.h0_prop, J.dt, J.W);
}

__global__ void g_ene1_kernel(const BlockJob* __restrict__ jobs, const int* __restrict__ prefix, int njobs) {
  GJ_FIND
  ene1_body(J.green, J.rot_h1, J.O, nullptr, nullptr, 0, J.N, 2.0, J.ene1, local);
}

__global__ void g_energy_finalize_kernel(const BlockJob* __restrict__ jobs, const int* __restrict__ prefix,
                                         int njobs) {
  GJ_FIND
  energy_finalize_body(J.ene1, J.coul, nullptr, J.kpart, nullptr, J.G, J.nch, 0, J.h0, J.energies, local);
}

__global__ __launch_bounds__(1024) void g_block_finalize_kernel(const BlockJob* __restrict__ jobs,
                                                                const int* __restrict__ prefix, int njobs) {
  GJ_FIND
  (void)local;
  block_finalize_body(J.energies, J.weights, J.scalars, J.clip, J.W, J.out);
}

__global__ __launch_bounds__(1024) void g_sr_index_kernel(const BlockJob* __restrict__ jobs,
                                                          const int* __restrict__ prefix, int njobs) {
  GJ_FIND
  (void)local;
  if (threadIdx.x == 0) sr_zeta_body(J.zeta, J.seed, 0, J.counter);
  __syncthreads();
  sr_index_body(J.weights, J.zeta, J.W, J.cum, J.idx);
}

__global__ void g_gather_kernel(const BlockJob* __restrict__ jobs, const int* __restrict__ prefix, int njobs) {
  GJ_FIND
  const long per = (long)J.N * J.O;
  const long chunks = (per + 1023) / 1024;
  const long w = local / chunks;
  const long base = (local % chunks) * 1024;
  const long src = (long)J.idx[w] * per;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const long i = base + t * 256 + threadIdx.x;
    if (i < per) J.wgather[w * per + i] = J.walkers[src + i];
  }
}

__global__ void g_block_tail_kernel(const BlockJob* __restrict__ jobs, const int* __restrict__ prefix, int njobs) {
  GJ_FIND
  (void)local;
  if (threadIdx.x != 0) return;
  *J.counter += 1;
  // the driver's energy estimate follows the block energy (driver.py:168)
  J.scalars[0] = 0.9 * J.scalars[0] + 0.1 * J.out[0];
}
#undef GJ_FIND

}  // namespace

// ------------------------------------------------------------------ launchers
long block_items(const BlockJob& j, int kind) {
  switch (kind) {
    case BI_WALKER: return j.W;
    case BI_WALKER256: return (j.W + 255) / 256;
    case BI_JOB: return 1;
    case BI_RNG: return (((long)j.nsteps * j.W * j.G + 1) / 2 + 255) / 256;
    case BI_GATHER: return (long)j.W * (((long)j.N * j.O + 1023) / 1024);
    default: return 0;
  }
}

#define GJ_LAUNCH(kern, threads, ...)                                                                       \
  do {                                                                                                   \
    if (t.njobs < 1 || t.total < 1) return AFQ_OK;                                                       \
    hipLaunchKernelGGL(kern, dim3((unsigned)t.total), dim3(threads), 0, s, t.params, t.prefix, t.njobs, \
                       ##__VA_ARGS__);                                                                   \
    return launch_status(#kern);                                                                         \
  } while (0)
int g_rng(const GroupTab<BlockJob>& t, hipStream_t s) { GJ_LAUNCH(g_rng_kernel, 256); }
int g_field_shift(const GroupTab<BlockJob>& t, int step, hipStream_t s) {
  GJ_LAUNCH(g_field_shift_kernel, 256, step);
}
int g_overlap(const GroupTab<BlockJob>& t, bool to_overlaps, hipStream_t s) {
  GJ_LAUNCH(g_overlap_kernel, 256, to_overlaps ? 1 : 0);
}
int g_weight_update(const GroupTab<BlockJob>& t, hipStream_t s) { GJ_LAUNCH(g_weight_update_kernel, 1024); }
int g_ene1(const GroupTab<BlockJob>& t, hipStream_t s) { GJ_LAUNCH(g_ene1_kernel, 256); }
int g_energy_finalize(const GroupTab<BlockJob>& t, hipStream_t s) { GJ_LAUNCH(g_energy_finalize_kernel, 256); }
int g_block_finalize(const GroupTab<BlockJob>& t, hipStream_t s) { GJ_LAUNCH(g_block_finalize_kernel, 1024); }
int g_sr_index(const GroupTab<BlockJob>& t, hipStream_t s) { GJ_LAUNCH(g_sr_index_kernel, 1024); }
int g_gather(const GroupTab<BlockJob>& t, hipStream_t s) { GJ_LAUNCH(g_gather_kernel, 256); }
int g_block_tail(const GroupTab<BlockJob>& t, hipStream_t s) { GJ_LAUNCH(g_block_tail_kernel, 64); }
#undef GJ_LAUNCH

int k_sym_h1(const double* in, double* out, int n, hipStream_t s) {
  const long tot = 2L * n * n;
  hipLaunchKernelGGL(sym_h1_kernel, dim3(blocks_for(tot, 256)), dim3(256), 0, s, in, out, n);
  return launch_status("sym_h1_kernel");
}
int k_axpby4(double* out, const double* x, const double* y, double alpha, const double* u, const double* v,
             double beta, long n, hipStream_t s) {
  hipLaunchKernelGGL(axpby4_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, x, y, alpha, u, v, beta, n);
  return launch_status("axpby4_kernel");
}
int k_mf(const double* m, int G, double h0, double2* mf, double* mfr, double2* h0_prop, hipStream_t s) {
  hipLaunchKernelGGL(mf_kernel, dim3(1), dim3(256), 0, s, m, G, h0, mf, mfr, h0_prop);
  return launch_status("mf_kernel");
}
int k_permute_chunk(const double* L, double* out, int n, int g0, int gch, hipStream_t s) {
  const long tot = (long)n * gch * n;
  hipLaunchKernelGGL(permute_chunk_kernel, dim3(blocks_for(tot, 256)), dim3(256), 0, s, L, out, n, g0, gch);
  return launch_status("permute_chunk_kernel");
}
int k_norm1(const double* A, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(norm1_kernel, dim3(1), dim3(256), 0, s, A, n, out);
  return launch_status("norm1_kernel");
}
int k_identity(double* out, int n, hipStream_t s) {
  hipLaunchKernelGGL(scaled_identity_kernel, dim3(blocks_for((long)n * n, 256)), dim3(256), 0, s, out,
                     (const double*)nullptr, 0.0, n, 0);
  return launch_status("scaled_identity_kernel");
}
int k_scale(double* out, const double* A, double scale, int n, hipStream_t s) {
  hipLaunchKernelGGL(scaled_identity_kernel, dim3(blocks_for((long)n * n, 256)), dim3(256), 0, s, out, A, scale, n,
                     1);
  return launch_status("scaled_identity_kernel");
}
int k_overlap_combine(const double2* d0, const double2* d1, double2* out, int W, hipStream_t s) {
  hipLaunchKernelGGL(overlap_combine_kernel, dim3(blocks_for(W, 256)), dim3(256), 0, s, d0, d1, out, W);
  return launch_status("overlap_combine_kernel");
}
int k_ene1(const double2* G0, const double* rh0, int n0, const double2* G1, const double* rh1, int n1, int N, double f0,
           double2* out, int W, hipStream_t s) {
  hipLaunchKernelGGL(ene1_kernel, dim3(W), dim3(256), 0, s, G0, rh0, n0, G1, rh1, n1, N, f0, out);
  return launch_status("ene1_kernel");
}
int k_energy_finalize(const double2* ene1, const double2* c0, const double2* c1, const double2* k0, const double2* k1,
                      int Gn, int nch0, int nch1, double h0, double2* out, int W, hipStream_t s) {
  hipLaunchKernelGGL(energy_finalize_kernel, dim3(W), dim3(256), 0, s, ene1, c0, c1, k0, k1, Gn, nch0, nch1, h0, out);
  return launch_status("energy_finalize_kernel");
}
int k_field_shift(const double2* fb, const double* x, const double2* mf, int Gn, double sqdt, double2* xs,
                  double2* terms, int W, hipStream_t s) {
  hipLaunchKernelGGL(field_shift_kernel, dim3(W), dim3(256), 0, s, fb, x, mf, Gn, sqdt, xs, terms);
  return launch_status("field_shift_kernel");
}
int k_weight_update(const double2* ov_new, double2* overlaps, const double2* terms, double* weights, double* scalars,
                    const double2* h0_prop, double dt, int W, hipStream_t s) {
  hipLaunchKernelGGL(weight_update_kernel, dim3(1), dim3(1024), 0, s, ov_new, overlaps, terms, weights, scalars,
                     h0_prop, dt, W);
  return launch_status("weight_update_kernel");
}
int k_block_finalize(const double2* energies, const double* weights, double* scalars, double clip, int W, double* out,
                     hipStream_t s) {
  hipLaunchKernelGGL(block_finalize_kernel, dim3(1), dim3(1024), 0, s, energies, weights, scalars, clip, W, out);
  return launch_status("block_finalize_kernel");
}
int k_sr_index(double* weights, const double* zeta, int W, double* cum, int32_t* idx, hipStream_t s) {
  hipLaunchKernelGGL(sr_index_kernel, dim3(1), dim3(1024), 0, s, weights, zeta, W, cum, idx);
  return launch_status("sr_index_kernel");
}
int k_gather_walkers(const double2* src, double2* dst, const int32_t* idx, long per, int W, hipStream_t s) {
  hipLaunchKernelGGL(gather_walkers_kernel, dim3(blocks_for(per, 256), W), dim3(256), 0, s, src, dst, idx, per);
  return launch_status("gather_walkers_kernel");
}
int k_to_f32(const double* src, float* dst, long n, hipStream_t s) {
  if (n <= 0) return AFQ_OK;
  hipLaunchKernelGGL(to_f32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, dst, n);
  return launch_status("to_f32_kernel");
}
long sym_packed_cols(int N) {
  const long nb = (N + 15) / 16;
  return nb * (nb + 1) / 2 * 256;
}
int k_sym_check(const double* L, int G, int N, unsigned long long* out2, hipStream_t s) {
  const long total = (long)G * N * N;
  if (total <= 0) return AFQ_OK;
  AFQ_HIP(hipMemsetAsync(out2, 0, 2 * sizeof(unsigned long long), s));
  const unsigned blocks = (unsigned)std::min<long>(4096, (total + 255) / 256);
  hipLaunchKernelGGL(sym_check_kernel, dim3(blocks), dim3(256), 0, s, L, total, N, out2);
  return launch_status("sym_check_kernel");
}
int k_sym_pack(const double* L, int G, int N, double* out64, float* out32, hipStream_t s) {
  const long npk = sym_packed_cols(N), total = (long)G * npk;
  if (out64)
    hipLaunchKernelGGL((sym_pack_kernel<double>), dim3(blocks_for(total, 256)), dim3(256), 0, s, L, N, npk, total,
                       out64);
  else
    hipLaunchKernelGGL((sym_pack_kernel<float>), dim3(blocks_for(total, 256)), dim3(256), 0, s, L, N, npk, total,
                       out32);
  return launch_status("sym_pack_kernel");
}
int k_minus_identity_f32(const double* src, float* dst, int n, hipStream_t s) {
  if (n <= 0) return AFQ_OK;
  hipLaunchKernelGGL(minus_identity_f32_kernel, dim3(blocks_for((long)n * n, 256)), dim3(256), 0, s, src, dst, n);
  return launch_status("minus_identity_f32_kernel");
}
int k_normal_fill(double* out, long n, uint64_t seed, uint64_t counter, hipStream_t s) {
  const long pairs = (n + 1) / 2;
  if (pairs <= 0) return AFQ_OK;
  hipLaunchKernelGGL(normal_fill_kernel, dim3(blocks_for(pairs, 256)), dim3(256), 0, s, out, n, seed, counter,
                     (const uint64_t*)nullptr);
  return launch_status("normal_fill_kernel");
}
int k_normal_fill_dev(double* out, long n, uint64_t seed, const uint64_t* counter_dev, hipStream_t s) {
  const long pairs = (n + 1) / 2;
  if (pairs <= 0) return AFQ_OK;
  hipLaunchKernelGGL(normal_fill_kernel, dim3(blocks_for(pairs, 256)), dim3(256), 0, s, out, n, seed, (uint64_t)0,
                     counter_dev);
  return launch_status("normal_fill_kernel");
}
int k_sr_zeta(double* zeta, uint64_t seed, const uint64_t* counter_dev, hipStream_t s) {
  hipLaunchKernelGGL(sr_zeta_kernel, dim3(1), dim3(64), 0, s, zeta, seed, (uint64_t)0, counter_dev);
  return launch_status("sr_zeta_kernel");
}
int k_sr_zeta_host(double* zeta, uint64_t seed, uint64_t counter, hipStream_t s) {
  hipLaunchKernelGGL(sr_zeta_kernel, dim3(1), dim3(64), 0, s, zeta, seed, counter, (const uint64_t*)nullptr);
  return launch_status("sr_zeta_kernel");
}
int k_counter_incr(uint64_t* counter_dev, hipStream_t s) {
  hipLaunchKernelGGL(counter_incr_kernel, dim3(1), dim3(64), 0, s, counter_dev);
  return launch_status("counter_incr_kernel");
}

}  // namespace afq
