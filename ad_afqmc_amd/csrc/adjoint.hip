// Reverse pass of a phaseless segment on the device (RHF trial, restricted
// walkers, FP64): the value and the gradient with respect to h1 and the
// Cholesky vectors of
//
//     F = sum_b E_b W_b / sum_b W_b
//
// over n_blocks blocks of n_steps phaseless steps with given fields and no
// population control -- the function the reference differentiates with
// jax.vjp in its "reverse" AD mode (sampling.py:300-341 _nosr_norot variant,
// driver.py:865-876).  The segment is checkpointed once per block (walkers,
// weights, population-control shift).  The steps' tapes (walkers, weights,
// overlaps, shifted fields before and after each step) are recorded for every
// block in the forward sweep when they fit the memory limit set in
// afq_segment_energy_vjp; otherwise the reverse sweep recomputes each block's
// steps from its checkpoint.  Either way the steps are then pulled back in
// reverse order.
//
// Adjoint convention: a complex intermediate z carries g_z = dF/dRe z - i dF/dIm z,
// so a holomorphic product y = A x pulls back as g_x = A^T g_y (transposes only,
// which the GEMM supports directly); a real parameter p in z gets Re(g_z dz/dp).
// Products that feed the real parameters L and rot_chol compute only the part of
// the complex product that Re(alpha A.B) needs (the GEMM's real-part mode); the
// rot_h1 accumulator stays complex and is reduced with Re at the end.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "kernels.h"
#include "pipeline.h"

namespace afq {
namespace {

__device__ inline double2 zmul(double2 a, double2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline double2 zdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
__device__ inline double2 zexp(double2 a) {
  const double e = exp(a.x);
  return {e * cos(a.y), e * sin(a.y)};
}
__device__ inline double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
// Sum over a 1024-thread work-group; every thread gets the total.
__device__ inline double wg_sum(double v, double* red) {
  v = wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
  return t;
}
inline unsigned nblk(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// Step k of the weight update, pulled back: from the adjoints of w_{k+1}
// (gw, in) and of the shift E_{k+1} (gesh, in) to the adjoints of w_k (gw,
// out) and E_k (gesh, out), plus gd = g of log(ov'/ov) and of -sqrt(dt)
// shift_term (they enter the importance function the same way), alpha = the
// adjoint of its real exponent, and dt sum(alpha) added to the h0_prop adjoint.
// Follows propagation.py:306-334 (phaseless factor, clipping, weight cap, shift).
__global__ __launch_bounds__(1024) void adj_weight_kernel(const double2* ov_k, const double2* ov_n,
                                                          const double2* terms, const double* w_k, const double* w_n,
                                                          const double* sc_k, const double2* h0p_p, double dt, int W,
                                                          double* gw, double* gesh, double2* gd, double* alpha,
                                                          double* gh0p) {
  __shared__ double red[16];
  const double sqdt = sqrt(dt);
  double loc = 0.0;
  for (int w = threadIdx.x; w < W; w += blockDim.x) loc += w_n[w];
  const double sw = wg_sum(loc, red);
  const double geshn = *gesh;
  const double pcs = sc_k[1];
  const double2 h0p = *h0p_p;
  double asum = 0.0;
  for (int w = threadIdx.x; w < W; w += blockDim.x) {
    const double gwn = gw[w] + geshn * (-0.1 / dt) / sw;
    const double2 ratio = zdiv(ov_n[w], ov_k[w]);
    const double2 sh = terms[2 * w], ft = terms[2 * w + 1];
    const double2 arg = {-sqdt * sh.x + ft.x + dt * (pcs + h0p.x), -sqdt * sh.y + ft.y + dt * h0p.y};
    const double2 imp = zmul(zexp(arg), ratio);
    const double2 ph_arg = zmul(zexp(double2{-sqdt * sh.x, -sqdt * sh.y}), ratio);
    const double theta = atan2(ph_arg.y, ph_arg.x);
    const double aimp = hypot(imp.x, imp.y);
    const double ph_raw = aimp * cos(theta);
    const bool mph = !(isnan(ph_raw) || ph_raw < 1.0e-3 || ph_raw > 100.0);
    const double ph = mph ? ph_raw : 0.0;
    const bool mw = !(ph * w_k[w] > 100.0);
    const double gph = mw ? gwn * w_k[w] : 0.0;
    const double a = mph ? gph * ph_raw : 0.0;
    const double b = mph ? gph * (-aimp * sin(theta)) : 0.0;
    gw[w] = mw ? gwn * ph : 0.0;
    gd[w] = double2{a, -b};
    alpha[w] = a;
    asum += a;
  }
  const double at = wg_sum(asum, red);
  if (threadIdx.x == 0) {
    *gesh = dt * at;
    *gh0p += dt * at;
  }
}

// g of the shifted fields from shift_term = sum_g xt_g (i m_g):  gxt = gsh * mf, gsh = -sqrt(dt) gd.
__global__ void adj_gxt_init_kernel(const double2* gd, const double2* mf, double sqdt, int G, long n, double2* gxt) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 gsh = {-sqdt * gd[i / G].x, -sqdt * gd[i / G].y};
  gxt[i] = zmul(gsh, mf[i % G]);
}

// gs = alpha xt - gxt (fb_term and xt = x - s), gfb = -i sqrt(dt) gs (s = -i sqrt(dt)(fb - m)),
// and the m adjoint from shift_term and s, summed over walkers.  A work-group
// of 1024 threads takes 64 gammas: 16 walker lanes per gamma, lane l taking
// walkers l, l + 16, ...; the lanes' partial sums are added in a fixed order.
__global__ __launch_bounds__(1024) void adj_fb_kernel(const double2* xt, const double2* gxt, const double2* gd,
                                                      const double* alpha, double sqdt, int G, int W, double2* gfb,
                                                      double* mbar) {
  __shared__ double part[16][64];
  const int gl = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int g = blockIdx.x * 64 + gl;
  double macc = 0.0;
  if (g < G) {
    for (int w = wl; w < W; w += 16) {
      const long i = (long)w * G + g;
      const double2 x = xt[i], gx = gxt[i];
      const double a = alpha[w];
      const double2 gs = {a * x.x - gx.x, a * x.y - gx.y};
      gfb[i] = double2{sqdt * gs.y, -sqdt * gs.x};
      const double2 gsh = {-sqdt * gd[w].x, -sqdt * gd[w].y};
      const double2 p = zmul(gsh, x);
      macc += -p.y - sqdt * gs.y;  // Re(i gsh xt) + Re(i sqrt(dt) gs)
    }
  }
  part[wl][gl] = macc;
  __syncthreads();
  if (wl == 0 && g < G) {
    double t = 0.0;
    for (int l = 0; l < 16; ++l) t += part[l][gl];
    mbar[g] += t;
  }
}

// X_w = c gd_w S_w^-T (O x O per walker)
__global__ void adj_scaled_transpose_kernel(const double2* Sinv, const double2* gd, double c, int O, long n,
                                            double2* X) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long oo = (long)O * O;
  const long w = i / oo, r = (i % oo) / O, q = i % O;
  const double2 v = Sinv[w * oo + q * O + r];
  X[i] = zmul(double2{c * gd[w].x, c * gd[w].y}, v);
}

// P_w += c gd_w I
__global__ void adj_diag_add_kernel(double2* P, const double2* gd, double c, int O, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long w = i / O, r = i % O;
  double2& p = P[w * (long)O * O + r * O + r];
  p.x += c * gd[w].x;
  p.y += c * gd[w].y;
}

// Block end (sampling.py:84-92 and the segment ratio F): from the adjoint of the
// shift after the block to the energy weights kappa_w = dF/dE_w (zero where the
// energy was clipped), the weight adjoints and the shift adjoint before it.
__global__ __launch_bounds__(1024) void adj_block_end_kernel(const double2* energies, const double* w_end,
                                                             const double* sc, const double* bo, const double* seg,
                                                             double clip, int W, double* gesh, double* kappa,
                                                             double* gw) {
  const double gsa = *gesh;
  const double e_est = sc[0], be = bo[0], bw = bo[1], F = seg[0], D = seg[1];
  const double gbe = 0.1 * gsa + bw / D;
  const double gbw = (be - F) / D;
  for (int w = threadIdx.x; w < W; w += blockDim.x) {
    const double e = energies[w].x;
    const bool in = !(fabs(e - e_est) > clip);
    const double ec = in ? e : e_est;
    kappa[w] = in ? gbe * w_end[w] / bw : 0.0;
    gw[w] += gbe * (ec - be) / bw + gbw;
  }
  __syncthreads();
  if (threadIdx.x == 0) *gesh = 0.9 * gsa;
}

// F = sum E_b W_b / sum W_b (and D = sum W_b)
__global__ void adj_seed_kernel(const double* bo, int nb, double* seg, double* energy_out) {
  if (threadIdx.x != 0) return;
  double num = 0.0, den = 0.0;
  for (int b = 0; b < nb; ++b) {
    num += bo[2 * b] * bo[2 * b + 1];
    den += bo[2 * b + 1];
  }
  seg[0] = num / den;
  seg[1] = den;
  *energy_out = num / den;
}

// F~ = c kappa_w f (in place, layout [g][i][w][j]) and its copy in layout [w][i][g][j].
__global__ void adj_fscale_kernel(double2* F, const double* kappa, double c, int G, int O, int Wc, double2* Fp) {
  const long n = (long)G * O * Wc * O;
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int j = (int)(idx % O);
  long t = idx / O;
  const int wl = (int)(t % Wc);
  t /= Wc;
  const int i = (int)(t % O);
  const int g = (int)(t / O);
  const double s = c * kappa[wl];
  const double2 v = {s * F[idx].x, s * F[idx].y};
  F[idx] = v;
  Fp[(((long)wl * O + i) * G + g) * O + j] = v;
}

// Cc[w][g] = c kappa_w Ct[w][g] (the Coulomb adjoint from the traces Ct)
__global__ void adj_cscale_kernel(const double2* Ct, const double* kappa, double c, int G, long n, double2* Cc) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = c * kappa[i / G];
  Cc[i] = double2{s * Ct[i].x, s * Ct[i].y};
}

// gG_w += c kappa_w rot_h1
__global__ void adj_outer_kernel(double2* gG, const double* kappa, const double* rh1, double c, long per, long n) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  gG[idx].x += c * kappa[idx / per] * rh1[idx % per];
}

__global__ void adj_conj_kernel(const double2* a, double2* b, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = double2{a[i].x, -a[i].y};
}

// out[e] = Re sum_w X[w][e]
__global__ void adj_reduce_re_kernel(const double2* X, int W, long n, double* out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int w = 0; w < W; ++w) s += X[(long)w * n + e].x;
  out[e] = s;
}

__global__ void adj_re_kernel(const double2* X, long n, double* out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) out[e] = X[e].x;
}

// M = [[A^T, E], [0, A^T]] (2n x 2n): expm(M) holds the Frechet derivative of
// expm at A^T in direction E in its upper-right block.
__global__ void adj_block_matrix_kernel(const double* A, const double* E, int n, double* M) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long m = 2L * n;
  if (idx >= m * m) return;
  const int r = (int)(idx / m), c = (int)(idx % m);
  double v = 0.0;
  if (r < n && c < n)
    v = A[(long)c * n + r];
  else if (r < n)
    v = E[(long)r * n + (c - n)];
  else if (c >= n)
    v = A[(long)(c - n) * n + (r - n)];
  M[idx] = v;
}

__global__ void adj_extract_ur_kernel(const double* M, int n, double scale, double* out) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  out[idx] = scale * M[(long)r * 2 * n + n + c];
}

// mbar_g += h0pbar m_g + t_g
__global__ void adj_mbar_final_kernel(double* mbar, const double* m, const double* gh0p, const double* t, int G) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) mbar[g] += (*gh0p) * m[g] + t[g];
}

// Lbar_g += m_g hmodbar + mbar_g rho
__global__ void adj_lfinal_kernel(double* Lbar, const double* m, const double* mbar, const double* hmodbar,
                                  const double* rho, int G, long nn) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)G * nn) return;
  const long g = idx / nn, e = idx % nn;
  Lbar[idx] += m[g] * hmodbar[e] + mbar[g] * rho[e];
}

// h1bar[s] = (X + X^T)/4 for s = 0, 1 (restricted walkers: h1s = ((h1_0 + h1_0^T)/2 + (h1_1 + h1_1^T)/2)/2)
__global__ void adj_h1_out_kernel(const double* X, int n, double* out) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nn = (long)n * n;
  if (idx >= nn) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  const double v = 0.25 * (X[idx] + X[(long)c * n + r]);
  out[idx] = v;
  out[nn + idx] = v;
}

// h1bar_s = (X + X^T)/2 (unrestricted walkers: h1s_s = (h1_s + h1_s^T)/2)
__global__ void adj_h1_out_spin_kernel(const double* X, int n, double* out) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  out[idx] = 0.5 * (X[idx] + X[(long)c * n + r]);
}

// ---------------------------------------------------------------- GEMM builder
struct GB {
  Gemm g;
  GB(int M, int N, int K, int batch = 1) {
    g.M = M;
    g.N = N;
    g.K = K;
    g.batch = batch;
  }
  GB& a(const void* p, long lda, long sA, bool cplx, bool trans = false) {
    g.A = p;
    g.lda = lda;
    g.sA = sA;
    g.a_cplx = cplx;
    g.a_trans = trans;
    return *this;
  }
  GB& b(const void* p, long ldb, long sB, bool cplx, bool trans = false) {
    g.B = p;
    g.ldb = ldb;
    g.sB = sB;
    g.b_cplx = cplx;
    g.b_trans = trans;
    return *this;
  }
  GB& c(void* p, long ldc, long sC) {
    g.C = p;
    g.ldc = ldc;
    g.sC = sC;
    return *this;
  }
  GB& d(void* p, long ldd, long sD) {
    g.D = p;
    g.ldd = ldd;
    g.sD = sD;
    return *this;
  }
  GB& alpha(double re, double im = 0.0) {
    g.alpha = double2{re, im};
    return *this;
  }
  GB& beta(double re) {
    g.beta = double2{re, 0.0};
    return *this;
  }
  GB& ws(afq_job* j) {
    g.ws = j->gemm_ws;
    g.ws_elems = j->gemm_ws_elems;
    return *this;
  }
  GB& part(int p) {  // real C: Re(alpha A.B) from the real (1) or imaginary (2) part of A.B
    g.c_part = p;
    return *this;
  }
  int run(hipStream_t st) { return gemm_f64(g, st); }
};

// Device arena carved in two passes (size, then pointers).
struct Arena {
  char* base = nullptr;
  size_t off = 0;
  template <typename T>
  T* take(size_t n) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += n * sizeof(T);
    return p;
  }
};

// Tapes of one block: walkers, weights, overlaps and scalars before each step
// and after the last (n_steps + 1 entries), shifted fields and weight terms of
// each step.
struct Tape {
  double2* W;
  double* w;
  double* sc;
  double2* ov;
  double2* xt;
  double2* terms;
};

// Spin layout.  A walker set holds spin 0's (W, N, O_0) panel and, for
// unrestricted walkers of a UHF trial, spin 1's (W, N, O_1) panel after it
// (ws elements in all); O x O per-walker matrices likewise (wq elements).
struct Spins {
  int nsp = 1;
  long O[2] = {0, 0}, pw[2] = {0, 0};  // occupied orbitals, N O_s per walker
  long wo[2] = {0, 0}, qo[2] = {0, 0};  // spin offsets in a walker set / O x O set
  long ws = 0, wq = 0, pt = 0;          // walker set, O x O set, sum_s N O_s
  void init(const afq_job* j) {
    nsp = j->restricted ? 1 : 2;
    const long N = j->N, W = j->W;
    for (int s = 0; s < nsp; ++s) {
      O[s] = j->nocc[s];
      pw[s] = N * O[s];
    }
    wo[1] = W * pw[0];
    qo[1] = W * O[0] * O[0];
    ws = W * (pw[0] + pw[1]);
    wq = W * (O[0] * O[0] + O[1] * O[1]);
    pt = pw[0] + pw[1];
  }
};

struct Buffers {
  // checkpoints and block results
  double2* ckW;
  double* ckw;
  double* cksc;
  double* bo;
  double* seg;
  // forward state
  double* fw;
  double2* fov;
  double* fsc;
  // tapes: one block's, or every block's when the forward records them
  Tape tp;
  // with recorded tapes: each block's orthonormalised end walkers and energies
  double2* qk;
  double2* ek;
  // reverse sweep
  double2 *Q, *Qc, *R, *Rinv, *detR;
  double2* Tb;
  double2 *W2, *gW, *gWk, *gW2, *gTa, *gTb;
  double2 *gvhs, *EbarW;
  double2 *gxt, *gfb, *gd, *gG, *P, *gS, *X;
  double *alpha, *gw, *kappa, *mbar, *scal;  // scal: gesh, gh0p
  double* Lbar;  // L adjoint (accumulated over the steps)
  double2* rh1barC;
  // energy chunks
  int Wc;
  double2 *F, *Fp, *Cc, *Ct;
  // setup
  double *Ebar, *v0, *v1, *hmod, *Amat, *Mblk, *Mexp, *Mwork, *hmodbar, *h1sbar, *rh1bar, *rcbar, *rho, *tg;
};

void carve(Arena& ar, Buffers& b, const afq_job* j, const Spins& sp, int nb, int ns, int K, bool keep) {
  const long N = j->N, G = j->G, W = j->W;
  const long NN = N * N, ws = sp.ws, wq = sp.wq, ns_ = sp.nsp;
  const long slots = keep ? nb : 1;
  b.ckW = ar.take<double2>(nb * ws);
  b.ckw = ar.take<double>(nb * W);
  b.cksc = ar.take<double>(2 * nb);
  b.bo = ar.take<double>(2 * nb);
  b.seg = ar.take<double>(2);
  b.fw = ar.take<double>(W);
  b.fov = ar.take<double2>(W);
  b.fsc = ar.take<double>(2);
  b.tp.W = ar.take<double2>(slots * (ns + 1) * ws);
  b.tp.w = ar.take<double>(slots * (ns + 1) * W);
  b.tp.sc = ar.take<double>(slots * 2 * (ns + 1));
  b.tp.ov = ar.take<double2>(slots * (ns + 1) * W);
  b.tp.xt = ar.take<double2>(slots * ns * W * G);
  b.tp.terms = ar.take<double2>(slots * ns * 2 * W);
  b.qk = ar.take<double2>(keep ? nb * ws : 0);
  b.ek = ar.take<double2>(keep ? nb * W : 0);
  b.Q = ar.take<double2>(ws);
  b.Qc = ar.take<double2>(ws);
  b.R = ar.take<double2>(wq);
  b.Rinv = ar.take<double2>(wq);
  b.detR = ar.take<double2>(2 * W);
  b.Tb = ar.take<double2>((K + 1) * ws);
  b.W2 = ar.take<double2>(ws);
  b.gW = ar.take<double2>(ws);
  b.gWk = ar.take<double2>(ws);
  b.gW2 = ar.take<double2>(ws);
  b.gTa = ar.take<double2>(ws);
  b.gTb = ar.take<double2>(ws);
  b.gvhs = ar.take<double2>(W * NN);
  b.EbarW = ar.take<double2>(ns_ * W * NN);
  b.gxt = ar.take<double2>(W * G);
  b.gfb = ar.take<double2>(W * G);
  b.gd = ar.take<double2>(W);
  b.gG = ar.take<double2>(ws);
  b.P = ar.take<double2>(wq);
  b.gS = ar.take<double2>(wq);
  b.X = ar.take<double2>(wq);
  b.alpha = ar.take<double>(W);
  b.gw = ar.take<double>(W);
  b.kappa = ar.take<double>(W);
  b.mbar = ar.take<double>(G);
  b.scal = ar.take<double>(4);
  b.Lbar = ar.take<double>(G * NN);
  b.rh1barC = ar.take<double2>(sp.pt);
  // exchange-adjoint chunks: F and its permuted copy at most ~2 GB each
  const long omax = std::max(sp.O[0], sp.O[1]);
  const long per_w = G * omax * omax * (long)sizeof(double2);
  b.Wc = (int)std::max(1L, std::min(W, (2048L << 20) / std::max(1L, per_w)));
  b.F = ar.take<double2>(G * omax * b.Wc * omax);
  b.Fp = ar.take<double2>(G * omax * b.Wc * omax);
  b.Cc = ar.take<double2>(W * G);
  b.Ct = ar.take<double2>(W * G);
  b.Ebar = ar.take<double>(NN);
  b.v0 = ar.take<double>(NN);
  b.v1 = ar.take<double>(NN);
  b.hmod = ar.take<double>(NN);
  b.Amat = ar.take<double>(NN);
  b.Mblk = ar.take<double>(4 * NN);
  b.Mexp = ar.take<double>(4 * NN);
  b.Mwork = ar.take<double>(12 * NN);
  b.hmodbar = ar.take<double>(ns_ * NN);
  b.h1sbar = ar.take<double>(NN);
  b.rh1bar = ar.take<double>(sp.pt);
  b.rcbar = ar.take<double>(G * sp.pt);
  b.rho = ar.take<double>(NN);
  b.tg = ar.take<double>(G);
}

#define LAUNCH(kernel, n, ...)                                                       \
  do {                                                                               \
    const long n_ = (n);                                                             \
    if (n_ > 0) hipLaunchKernelGGL(kernel, dim3(nblk(n_)), dim3(256), 0, st, __VA_ARGS__); \
    AFQ_TRY(launch_status(#kernel));                                                 \
  } while (0)

struct Seg {
  afq_job* j;
  Buffers b;
  Spins sp;
  hipStream_t st;
  int nb, ns, K;
  long N, G, W, NN;
  double dt, sqdt, e_est;
  // spin weights: restricted walkers stand for both spins (force bias, log
  // overlap, one-body and exchange energy counted twice, Coulomb 2c^2)
  double cs = 1.0, ccoul = 1.0;
  const double* fields;
  bool keep = false;  // the forward records every block's tapes
  Tape t;             // tapes of the block being pulled back

  Tape tape_at(int blk) const {
    const long s = keep ? blk : 0;
    return Tape{b.tp.W + s * (ns + 1) * sp.ws, b.tp.w + s * (ns + 1) * W, b.tp.sc + s * 2 * (ns + 1),
                b.tp.ov + s * (ns + 1) * W,       b.tp.xt + s * ns * W * G,  b.tp.terms + s * ns * 2 * W};
  }

  const double* T(int s) const { return j->mo[s]; }
  const double* E(int s) const { return j->exp_h1[s]; }
  const double* fields_at(int blk, int k) const { return fields + ((long)blk * ns + k) * W * G; }
  // spin s of a walker set / of an O x O set
  template <typename P>
  P* wsp(P* set, int s) const { return set + sp.wo[s]; }
  template <typename P>
  P* qsp(P* set, int s) const { return set + sp.qo[s]; }
  void* dn(double2* set) const { return sp.nsp == 2 ? static_cast<void*>(set + sp.wo[1]) : nullptr; }

  int cpy(void* dst, const void* src, size_t bytes) {
    AFQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
    return AFQ_OK;
  }
  int zero(void* p, size_t bytes) {
    AFQ_HIP(hipMemsetAsync(p, 0, bytes, st));
    return AFQ_OK;
  }
  int green(double2* Wp, bool want_green) { return green_pipeline(j, Wp, dn(Wp), true, want_green, st); }
  int overlap(double2* Wp, double2* ov) {
    AFQ_TRY(green_pipeline(j, Wp, dn(Wp), false, false, st));
    return overlap_from_dets(j, ov, st);
  }

  // Forward over all blocks with checkpoints (afq_block with the given fields),
  // with keep also every block's tapes, end walkers and energies.
  int forward(const void* walkers, const double* weights, double* energy_out) {
    const size_t wbytes = sizeof(double2) * sp.ws;
    double2* Wc = b.Q;  // current walkers
    double* wc = b.fw;
    double2* ov = b.fov;
    double* sc = b.fsc;
    AFQ_TRY(cpy(Wc, walkers, wbytes));
    AFQ_TRY(cpy(wc, weights, sizeof(double) * W));
    const double sc0[2] = {e_est, e_est};
    AFQ_HIP(hipMemcpyAsync(sc, sc0, sizeof(sc0), hipMemcpyHostToDevice, st));
    AFQ_TRY(overlap(Wc, ov));
    for (int blk = 0; blk < nb; ++blk) {
      AFQ_TRY(cpy(b.ckW + blk * sp.ws, Wc, wbytes));
      AFQ_TRY(cpy(b.ckw + blk * W, wc, sizeof(double) * W));
      AFQ_TRY(cpy(b.cksc + 2 * blk, sc, 2 * sizeof(double)));
      if (keep) {  // the same steps, on the block's tape
        AFQ_TRY(tape(blk));
        const Tape tb = tape_at(blk);
        AFQ_TRY(cpy(Wc, tb.W + ns * sp.ws, wbytes));
        AFQ_TRY(cpy(wc, tb.w + ns * W, sizeof(double) * W));
        AFQ_TRY(cpy(ov, tb.ov + ns * W, sizeof(double2) * W));
        AFQ_TRY(cpy(sc, tb.sc + 2 * ns, 2 * sizeof(double)));
      } else {
        for (int k = 0; k < ns; ++k) AFQ_TRY(propagate(j, Wc, dn(Wc), fields_at(blk, k), wc, ov, sc, st));
      }
      AFQ_TRY(orthonormalize(j, Wc, dn(Wc), st));
      AFQ_TRY(overlap(Wc, ov));
      AFQ_TRY(energy_pipeline(j, Wc, dn(Wc), j->energies, st));
      if (keep) {
        AFQ_TRY(cpy(b.qk + blk * sp.ws, Wc, wbytes));
        AFQ_TRY(cpy(b.ek + blk * W, j->energies, sizeof(double2) * W));
      }
      AFQ_TRY(k_block_finalize(j->energies, wc, sc, std::sqrt(2.0 / dt), (int)W, b.bo + 2 * blk, st));
    }
    hipLaunchKernelGGL(adj_seed_kernel, dim3(1), dim3(64), 0, st, b.bo, nb, b.seg, energy_out);
    return launch_status("adj_seed_kernel");
  }

  // Computes the tapes of block blk from its checkpoint.
  int tape(int blk) {
    const size_t wbytes = sizeof(double2) * sp.ws;
    const Tape tb = tape_at(blk);
    AFQ_TRY(cpy(tb.W, b.ckW + blk * sp.ws, wbytes));
    AFQ_TRY(cpy(tb.w, b.ckw + blk * W, sizeof(double) * W));
    AFQ_TRY(cpy(tb.sc, b.cksc + 2 * blk, 2 * sizeof(double)));
    AFQ_TRY(overlap(tb.W, tb.ov));
    for (int k = 0; k < ns; ++k) {
      double2* Wn = tb.W + (k + 1) * sp.ws;
      double* wn = tb.w + (k + 1) * W;
      double2* ovn = tb.ov + (k + 1) * W;
      double* scn = tb.sc + 2 * (k + 1);
      AFQ_TRY(cpy(Wn, tb.W + k * sp.ws, wbytes));
      AFQ_TRY(cpy(wn, tb.w + k * W, sizeof(double) * W));
      AFQ_TRY(cpy(ovn, tb.ov + k * W, sizeof(double2) * W));
      AFQ_TRY(cpy(scn, tb.sc + 2 * k, 2 * sizeof(double)));
      AFQ_TRY(propagate(j, Wn, dn(Wn), fields_at(blk, k), wn, ovn, scn, st));
      AFQ_TRY(cpy(tb.xt + k * W * G, j->xs, sizeof(double2) * W * G));
      AFQ_TRY(cpy(tb.terms + k * 2 * W, j->terms, sizeof(double2) * 2 * W));
    }
    return AFQ_OK;
  }

  // gWout += adjoint through G_s^T = W_s S_s^-1 and cs log det S_s (coefficient
  // gd, or none) for every spin, with j->green / j->Sinv holding G and S^-1 of
  // the walkers concerned.
  int green_backward(const double2* gG, const double2* gd, double2* gWout) {
    for (int s = 0; s < sp.nsp; ++s) {
      const long O = sp.O[s], per = sp.pw[s], OO = O * O;
      double2* P = qsp(b.P, s);
      double2* gS = qsp(b.gS, s);
      const double2* gGs = wsp(gG, s);
      double2* gWs = wsp(gWout, s);
      AFQ_TRY(GB(O, O, N, W).a(j->green[s], N, per, true).b(gGs, N, per, true, true).c(P, O, OO).run(st));
      if (gd) LAUNCH(adj_diag_add_kernel, W * O, P, gd, cs, (int)O, W * O);
      AFQ_TRY(GB(O, O, O, W).a(P, O, OO, true).b(j->Sinv[s], O, OO, true, true).c(gS, O, OO).alpha(-1.0).run(st));
      AFQ_TRY(GB(N, O, O, W)
                  .a(gGs, N, per, true, true)
                  .b(j->Sinv[s], O, OO, true, true)
                  .c(gWs, O, per)
                  .beta(1.0)
                  .run(st));
      AFQ_TRY(GB(N, O, O, W).a(T(s), O, 0, false).b(gS, O, OO, true).c(gWs, O, per).beta(1.0).run(st));
    }
    return AFQ_OK;
  }

  // Pulls kappa_w Re E_L(Q_w) back to gG (into b.gG) and the parameter accumulators;
  // j->green / j->Sinv hold G and S^-1 at Q.  wavefunctions.py:1128-1150, 1331-1363.
  // E_L = cs sum_s tr(rot_h1_s G_s^T) + (ccoul / 2) sum_g c_g^2 - (cs / 2) sum_sg K_sg,
  // c_g = sum_s tr f_sg, K_sg = sum_ij f_sg,ij f_sg,ji, f_sg = rot_chol_sg G_s^T.
  int energy_backward() {
    const int Wc = b.Wc;
    // Coulomb adjoint Cc[w][g] = ccoul kappa_w c_wg, c from one product per spin over all walkers
    for (int s = 0; s < sp.nsp; ++s)
      AFQ_TRY(GB(W, G, sp.pw[s])
                  .a(j->green[s], sp.pw[s], 0, true)
                  .b(j->rot_chol[s], sp.pw[s], 0, false, true)
                  .c(b.Ct, G, 0)
                  .beta(s > 0 ? 1.0 : 0.0)
                  .ws(j)
                  .run(st));
    LAUNCH(adj_cscale_kernel, W * G, b.Ct, b.kappa, ccoul, (int)G, W * G, b.Cc);
    for (int s = 0; s < sp.nsp; ++s) {
      const long O = sp.O[s], per = sp.pw[s];
      double* rcbar = b.rcbar + (s ? G * sp.pw[0] : 0);
      double2* rh1bar = b.rh1barC + (s ? sp.pw[0] : 0);
      for (long w0 = 0; w0 < W; w0 += Wc) {
        const int wc = (int)std::min<long>(Wc, W - w0);
        const double2* Gc = j->green[s] + w0 * per;
        double2* gG = wsp(b.gG, s) + w0 * per;
        const double2* Cc = b.Cc + w0 * G;
        const long fsz = G * O * wc * O;
        // f[g][i][(w,j)] = rc_g G_w^T, all gammas in one product: rows (g, i) of rot_chol
        AFQ_TRY(GB((int)(G * O), (int)(wc * O), N)
                    .a(j->rot_chol[s], N, 0, false)
                    .b(Gc, N, 0, true, true)
                    .c(b.F, wc * O, 0)
                    .run(st));
        LAUNCH(adj_fscale_kernel, fsz, b.F, b.kappa + w0, -cs, (int)G, (int)O, wc, b.Fp);
        // gG_w = sum_g Cc_wg rc_g - cs kappa sum_g f_g rc_g + cs kappa rot_h1
        AFQ_TRY(GB(wc, per, G).a(Cc, G, 0, true).b(j->rot_chol[s], per, 0, false).c(gG, per, 0).run(st));
        // all wc walkers in one product (row (w, i) of Fp starts at (w O + i) G O), split over K
        AFQ_TRY(GB((int)(wc * O), N, G * O)
                    .a(b.Fp, G * O, 0, true)
                    .b(j->rot_chol[s], N, 0, false)
                    .c(gG, N, 0)
                    .beta(1.0)
                    .ws(j)
                    .run(st));
        LAUNCH(adj_outer_kernel, wc * per, gG, b.kappa + w0, j->rot_h1[s], cs, per, wc * per);
        // rc_g adjoint: sum_w Cc_wg G_w - cs kappa f_wg^T G_w ; rot_h1 adjoint: cs sum_w kappa G_w
        AFQ_TRY(GB(G, per, wc).a(Cc, G, 0, true, true).b(Gc, per, 0, true).c(rcbar, per, 0).beta(1.0).part(1).run(st));
        // rows (g, j) of rot_chol: sum over (w, i) of Fp[w][i][g][j] G_w[i][n], one product
        AFQ_TRY(GB((int)(G * O), N, (int)(wc * O))
                    .a(b.Fp, G * O, 0, true, true)
                    .b(Gc, N, 0, true)
                    .c(rcbar, N, 0)
                    .beta(1.0)
                    .part(1)
                    .run(st));
        AFQ_TRY(GB(1, per, wc).a(b.kappa + w0, wc, 0, false).b(Gc, per, 0, true).c(rh1bar, per, 0).alpha(cs).beta(1.0).run(
            st));
      }
    }
    return AFQ_OK;
  }

  // One propagate_constrained step pulled back (propagation.py:274-336).
  // On entry b.gW / b.gw / scal[0] hold the adjoints of W_{k+1}, w_{k+1}, E_{k+1};
  // on exit those of W_k, w_k, E_k.
  int step_backward(int k) {
    double2* Wk = t.W + k * sp.ws;
    double2* Wn = t.W + (k + 1) * sp.ws;
    const double2* xt = t.xt + k * W * G;
    hipLaunchKernelGGL(adj_weight_kernel, dim3(1), dim3(1024), 0, st, t.ov + k * W, t.ov + (k + 1) * W,
                       t.terms + k * 2 * W, t.w + k * W, t.w + (k + 1) * W, t.sc + 2 * k, j->h0_prop, dt, (int)W,
                       b.gw, b.scal, b.gd, b.alpha, b.scal + 1);
    AFQ_TRY(launch_status("adj_weight_kernel"));
    // log ov' = cs sum_s log det(T_s^T W'_s): gW'_s += T_s (cs gd S'_s^-T)
    AFQ_TRY(green(Wn, false));
    for (int s = 0; s < sp.nsp; ++s) {
      const long O = sp.O[s], OO = O * O;
      LAUNCH(adj_scaled_transpose_kernel, W * OO, j->Sinv[s], b.gd, cs, (int)O, W * OO, qsp(b.X, s));
      AFQ_TRY(GB(N, O, O, W).a(T(s), O, 0, false).b(qsp(b.X, s), O, OO, true).c(wsp(b.gW, s), O, sp.pw[s]).beta(1.0).run(
          st));
    }
    LAUNCH(adj_gxt_init_kernel, W * G, b.gd, j->mf_shifts, sqdt, (int)G, W * G, b.gxt);
    // recompute VHS, the Taylor terms and their sum, per spin
    if (K > 0)
      AFQ_TRY(GB(W, NN, G).a(xt, G, 0, true).b(j->chol, NN, 0, false).c(j->vhs, NN, 0).alpha(0.0, sqdt).run(st));
    for (int s = 0; s < sp.nsp; ++s) {
      const long O = sp.O[s], per = sp.pw[s];
      double2* Tb0 = wsp(b.Tb, s);
      AFQ_TRY(GB(N, O, N, W).a(E(s), N, 0, false).b(wsp(Wk, s), O, per, true).c(Tb0, O, per).run(st));
      AFQ_TRY(cpy(wsp(b.W2, s), Tb0, sizeof(double2) * W * per));
      for (int n = 1; n <= K; ++n)
        AFQ_TRY(GB(N, O, N, W)
                    .a(j->vhs, N, NN, true)
                    .b(Tb0 + (n - 1) * sp.ws, O, per, true)
                    .c(Tb0 + n * sp.ws, O, per)
                    .d(wsp(b.W2, s), O, per)
                    .alpha(1.0 / n)
                    .run(st));
    }
    for (int s = 0; s < sp.nsp; ++s) {
      const long O = sp.O[s], per = sp.pw[s];
      double2* Tb0 = wsp(b.Tb, s);
      double2* EbarW = b.EbarW + s * W * NN;
      // W' = E W2
      AFQ_TRY(GB(N, O, N, W).a(E(s), N, 0, false, true).b(wsp(b.gW, s), O, per, true).c(wsp(b.gW2, s), O, per).run(st));
      AFQ_TRY(GB(N, N, O, W).a(wsp(b.gW, s), O, per, true).b(wsp(b.W2, s), O, per, true, true).c(EbarW, N, NN).beta(1.0).run(
          st));
      // Taylor: gT_K = gW2 ; gT_{n-1} = gW2 + vhs^T gT_n / n ; gvhs = sum_n gT_n T_{n-1}^T / n
      const double2* cur = wsp(b.gW2, s);
      double2* ga = wsp(b.gTa, s);
      double2* gb = wsp(b.gTb, s);
      for (int n = K; n >= 1; --n) {
        AFQ_TRY(GB(N, N, O, W)
                    .a(cur, O, per, true)
                    .b(Tb0 + (n - 1) * sp.ws, O, per, true, true)
                    .c(b.gvhs, N, NN)
                    .alpha(1.0 / n)
                    .beta(n == K && s == 0 ? 0.0 : 1.0)
                    .run(st));
        double2* nxt = (cur == ga) ? gb : ga;
        AFQ_TRY(cpy(nxt, wsp(b.gW2, s), sizeof(double2) * W * per));
        AFQ_TRY(GB(N, O, N, W).a(j->vhs, N, NN, true, true).b(cur, O, per, true).c(nxt, O, per).alpha(1.0 / n).beta(1.0).run(
            st));
        cur = nxt;
      }
      // W1 = E W_k
      AFQ_TRY(GB(N, N, O, W).a(cur, O, per, true).b(wsp(Wk, s), O, per, true, true).c(EbarW, N, NN).beta(1.0).run(st));
      AFQ_TRY(GB(N, O, N, W).a(E(s), N, 0, false, true).b(cur, O, per, true).c(wsp(b.gWk, s), O, per).run(st));
    }
    // vhs = i sqrt(dt) sum_g xt_g L_g
    if (K > 0) {
      AFQ_TRY(GB(W, G, NN).a(b.gvhs, NN, 0, true).b(j->chol, NN, 0, false, true).c(b.gxt, G, 0).alpha(0.0, sqdt).beta(1.0).ws(
          j).run(st));
      AFQ_TRY(GB(G, NN, W)
                  .a(xt, G, 0, true, true)
                  .b(b.gvhs, NN, 0, true)
                  .c(b.Lbar, NN, 0)
                  .alpha(0.0, sqdt)
                  .beta(1.0)
                  .part(2)
                  .run(st));
    }
    // shifts, force bias and the m adjoint
    hipLaunchKernelGGL(adj_fb_kernel, dim3((unsigned)((G + 63) / 64)), dim3(1024), 0, st, xt, b.gxt, b.gd, b.alpha, sqdt,
                       (int)G, (int)W, b.gfb, b.mbar);
    AFQ_TRY(launch_status("adj_fb_kernel"));
    // fb = cs sum_s rc_s . G_s at W_k
    AFQ_TRY(green(Wk, true));
    for (int s = 0; s < sp.nsp; ++s) {
      const long per = sp.pw[s];
      AFQ_TRY(GB(W, per, G).a(b.gfb, G, 0, true).b(j->rot_chol[s], per, 0, false).c(wsp(b.gG, s), per, 0).alpha(cs).run(st));
      AFQ_TRY(GB(G, per, W)
                  .a(b.gfb, G, 0, true, true)
                  .b(j->green[s], per, 0, true)
                  .c(b.rcbar + (s ? G * sp.pw[0] : 0), per, 0)
                  .alpha(cs)
                  .beta(1.0)
                  .part(1)
                  .run(st));
    }
    // Green's function and -log ov at W_k
    AFQ_TRY(green_backward(b.gG, b.gd, b.gWk));
    std::swap(b.gW, b.gWk);
    return AFQ_OK;
  }

  int reverse() {
    const size_t wbytes = sizeof(double2) * sp.ws;
    AFQ_TRY(zero(b.Lbar, sizeof(double) * G * NN));
    AFQ_TRY(zero(b.rcbar, sizeof(double) * G * sp.pt));
    AFQ_TRY(zero(b.rh1barC, sizeof(double2) * sp.pt));
    AFQ_TRY(zero(b.EbarW, sizeof(double2) * sp.nsp * W * NN));
    AFQ_TRY(zero(b.mbar, sizeof(double) * G));
    AFQ_TRY(zero(b.scal, sizeof(double) * 4));
    AFQ_TRY(zero(b.gW, wbytes));
    AFQ_TRY(zero(b.gw, sizeof(double) * W));
    for (int blk = nb - 1; blk >= 0; --blk) {
      if (!keep) AFQ_TRY(tape(blk));
      t = tape_at(blk);
      // block end: Q = qr(W_end), energies at Q (recorded, or recomputed)
      double2* Wend = t.W + ns * sp.ws;
      if (keep) {
        AFQ_TRY(cpy(b.Q, b.qk + blk * sp.ws, wbytes));
        AFQ_TRY(cpy(j->energies, b.ek + blk * W, sizeof(double2) * W));
        AFQ_TRY(green(b.Q, true));
      } else {
        AFQ_TRY(cpy(b.Q, Wend, wbytes));
        AFQ_TRY(orthonormalize(j, b.Q, dn(b.Q), st));
        AFQ_TRY(energy_pipeline(j, b.Q, dn(b.Q), j->energies, st));
      }
      hipLaunchKernelGGL(adj_block_end_kernel, dim3(1), dim3(1024), 0, st, j->energies, t.w + ns * W, t.sc,
                         b.bo + 2 * blk, b.seg, std::sqrt(2.0 / dt), (int)W, b.scal, b.kappa, b.gw);
      AFQ_TRY(launch_status("adj_block_end_kernel"));
      AFQ_TRY(energy_backward());
      AFQ_TRY(green_backward(b.gG, nullptr, b.gW));  // gQ = incoming gW + energy part
      // gauge per spin: Q = W_end R^-1 with R = Q^H W_end ; g_Wend = g_Q R^-T
      LAUNCH(adj_conj_kernel, sp.ws, b.Q, b.Qc, sp.ws);
      for (int s = 0; s < sp.nsp; ++s) {
        const long O = sp.O[s], per = sp.pw[s], OO = O * O;
        AFQ_TRY(GB(O, O, N, W).a(wsp(b.Qc, s), O, per, true, true).b(wsp(Wend, s), O, per, true).c(qsp(b.R, s), O, OO).run(
            st));
        AFQ_TRY(batched_det_inv(qsp(b.R, s), OO, (int)O, (int)W, b.detR + s * W, qsp(b.Rinv, s), OO, st));
        AFQ_TRY(GB(N, O, O, W)
                    .a(wsp(b.gW, s), O, per, true)
                    .b(qsp(b.Rinv, s), O, OO, true, true)
                    .c(wsp(b.gWk, s), O, per)
                    .run(st));
      }
      std::swap(b.gW, b.gWk);
      for (int k = ns - 1; k >= 0; --k) AFQ_TRY(step_backward(k));
    }
    return AFQ_OK;
  }

  // Setup intermediates pulled back to h1 and L (wavefunctions.py:1215-1232,
  // 1472-1496, propagation.py:378-438).
  int setup_backward(double* h1_bar, double* chol_bar) {
    double* hbar_tot = b.hmodbar + (sp.nsp - 1) * NN;  // sum over the spins' h1_mod adjoints
    AFQ_TRY(compute_v0(j, b.v0, st));
    AFQ_TRY(GB(1, NN, G).a(j->mreal + G, G, 0, false).b(j->chol, NN, 0, false).c(b.v1, NN, 0).run(st));
    for (int s = 0; s < sp.nsp; ++s) {
      double* hmodbar = b.hmodbar + s * NN;
      LAUNCH(adj_reduce_re_kernel, NN, b.EbarW + s * W * NN, (int)W, NN, b.Ebar);
      // h1_mod = (h1_0 + h1_1)/2 - v0 - v1 (restricted walkers) or h1_s - v0 - v1, A = -dt/2 h1_mod
      if (sp.nsp == 1)
        AFQ_TRY(k_axpby4(b.hmod, j->h1, j->h1 + NN, 0.5, b.v0, b.v1, -1.0, NN, st));
      else
        AFQ_TRY(k_axpby4(b.hmod, j->h1 + s * NN, nullptr, 1.0, b.v0, b.v1, -1.0, NN, st));
      AFQ_TRY(k_scale(b.Amat, b.hmod, -0.5 * dt, (int)N, st));
      LAUNCH(adj_block_matrix_kernel, 4 * NN, b.Amat, b.Ebar, (int)N, b.Mblk);
      AFQ_TRY(expm_device(b.Mblk, 2 * (int)N, b.Mexp, b.Mwork, j->dscratch, st));
      LAUNCH(adj_extract_ur_kernel, NN, b.Mexp, (int)N, -0.5 * dt, hmodbar);
      // h1s adjoint = hmodbar + T rot_h1bar
      const long per = sp.pw[s];
      double* rh1bar = b.rh1bar + (s ? sp.pw[0] : 0);
      LAUNCH(adj_re_kernel, per, b.rh1barC + (s ? sp.pw[0] : 0), per, rh1bar);
      AFQ_TRY(cpy(b.h1sbar, hmodbar, sizeof(double) * NN));
      AFQ_TRY(GB(N, N, sp.O[s]).a(T(s), sp.O[s], 0, false).b(rh1bar, N, 0, false).c(b.h1sbar, N, 0).beta(1.0).run(st));
      if (sp.nsp == 1)
        LAUNCH(adj_h1_out_kernel, NN, b.h1sbar, (int)N, h1_bar);
      else
        LAUNCH(adj_h1_out_spin_kernel, NN, b.h1sbar, (int)N, h1_bar + s * NN);
    }
    if (sp.nsp == 2) AFQ_TRY(k_axpby4(hbar_tot, b.hmodbar, b.hmodbar + NN, 1.0, nullptr, nullptr, 0.0, NN, st));
    // m adjoint: + h0_prop' m (h0_prop = -h0 + sum m^2/2) + sum L_g . hbar (v1 = -sum m_g L_g)
    AFQ_TRY(GB(G, 1, NN).a(j->chol, NN, 0, false).b(hbar_tot, 1, 0, false).c(b.tg, 1, 0).run(st));
    LAUNCH(adj_mbar_final_kernel, G, b.mbar, j->mreal, b.scal + 1, b.tg, (int)G);
    // L adjoint
    AFQ_TRY(cpy(chol_bar, b.Lbar, sizeof(double) * G * NN));
    AFQ_TRY(GB(N, N, N, G).a(hbar_tot, N, 0, false).b(j->chol, N, NN, false, true).c(chol_bar, N, NN).alpha(-0.5).beta(
        1.0).run(st));
    AFQ_TRY(GB(N, N, N, G).a(j->chol, N, NN, false, true).b(hbar_tot, N, 0, false).c(chol_bar, N, NN).alpha(-0.5).beta(
        1.0).run(st));
    for (int s = 0; s < sp.nsp; ++s)
      AFQ_TRY(GB(N, N, sp.O[s], G)
                  .a(T(s), sp.O[s], 0, false)
                  .b(b.rcbar + (s ? G * sp.pw[0] : 0), N, sp.pw[s], false)
                  .c(chol_bar, N, NN)
                  .beta(1.0)
                  .run(st));
    AFQ_TRY(k_axpby4(b.rho, j->rdm1, j->rdm1 + NN, 1.0, nullptr, nullptr, 0.0, NN, st));
    LAUNCH(adj_lfinal_kernel, G * NN, chol_bar, j->mreal, b.mbar, hbar_tot, b.rho, (int)G, NN);
    return AFQ_OK;
  }
};

#undef LAUNCH

}  // namespace
}  // namespace afq

using namespace afq;

extern "C" int afq_segment_energy_vjp(afq_job_t j, const void* walkers, const double* weights, double e_estimate,
                                      const double* fields, int32_t n_blocks, int32_t n_steps, double* energy_out,
                                      double* h1_bar, double* chol_bar, void* stream) {
  AFQ_TRY(require(j, true, true));
  if (!walkers || !weights || !fields || !energy_out) return fail(AFQ_E_INVALID, "NULL argument");
  if ((h1_bar == nullptr) != (chol_bar == nullptr))
    return fail(AFQ_E_INVALID, "h1_bar and chol_bar must both be given or both be NULL");
  if (n_blocks < 1 || n_steps < 1) return fail(AFQ_E_INVALID, "n_blocks and n_steps must be >= 1");
  if (j->rhf != j->restricted)
    return fail(AFQ_E_UNSUPPORTED,
                "segment derivative: RHF trial with restricted walkers, or UHF trial with unrestricted walkers");
  if (j->lp) return fail(AFQ_E_UNSUPPORTED, "segment derivative: FP64 jobs only");
  for (int s = 0; s < 2; ++s)
    if (j->have_rot_h1[s] || j->have_rot_chol[s] || j->have_exp[s])
      return fail(AFQ_E_STATE, "segment derivative needs intermediates built from the job's h1 and chol");
  if (j->have_mf || j->have_h0p)
    return fail(AFQ_E_STATE, "segment derivative needs intermediates built from the job's h1 and chol");
  Seg s;
  s.j = j;
  s.st = reinterpret_cast<hipStream_t>(stream);
  s.sp.init(j);
  s.nb = n_blocks;
  s.ns = n_steps;
  s.K = j->d.n_exp_terms - 1;
  s.N = j->N;
  s.G = j->G;
  s.W = j->W;
  s.NN = s.N * s.N;
  s.dt = j->d.dt;
  s.sqdt = std::sqrt(j->d.dt);
  s.e_est = e_estimate;
  s.fields = fields;
  s.cs = s.sp.nsp == 1 ? 2.0 : 1.0;
  s.ccoul = s.sp.nsp == 1 ? 4.0 : 1.0;
  // Record every block's tapes in the forward (instead of recomputing them
  // block by block in the reverse sweep) when that arena takes at most half of
  // the free device memory, or AFQ_SEGMENT_TAPE_LIMIT bytes when that is set.
  Arena sizer;
  bool keep = false;
  if (h1_bar) {
    carve(sizer, s.b, j, s.sp, n_blocks, n_steps, s.K, true);
    if (n_blocks == 1) {
      keep = true;  // a single block's tape takes the same memory either way
    } else {
      size_t free_b = 0, total_b = 0;
      AFQ_HIP(hipMemGetInfo(&free_b, &total_b));
      double limit = 0.5 * (double)free_b;
      if (const char* e = std::getenv("AFQ_SEGMENT_TAPE_LIMIT")) limit = std::min(limit, std::atof(e));
      keep = (double)sizer.off <= limit;
    }
  }
  if (!keep) {
    sizer = Arena{};
    carve(sizer, s.b, j, s.sp, n_blocks, n_steps, s.K, false);
  }
  s.keep = keep;
  void* base = nullptr;
  AFQ_TRY(dalloc(&base, sizer.off));
  Arena ar;
  ar.base = static_cast<char*>(base);
  carve(ar, s.b, j, s.sp, n_blocks, n_steps, s.K, s.keep);
  int rc = s.forward(walkers, weights, energy_out);
  if (rc == AFQ_OK && h1_bar) rc = s.reverse();
  if (rc == AFQ_OK && h1_bar) rc = s.setup_backward(h1_bar, chol_bar);
  const hipError_t e = hipStreamSynchronize(s.st);
  (void)hipFree(base);
  if (rc != AFQ_OK) return rc;
  if (e != hipSuccess) return fail(AFQ_E_HIP, std::string("afq_segment_energy_vjp: ") + hipGetErrorString(e));
  return AFQ_OK;
}
