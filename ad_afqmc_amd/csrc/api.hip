// C ABI of the engine: job lifetime, setup intermediates, measurements,
// propagation, re-orthonormalisation, population control and the device
// block loop.  Every kernel runs on the caller's stream.
#include <cmath>
#include <cstring>
#include <string>

#include "gemm_bf3.h"
#include "group.h"
#include "kernels.h"
#include "trot_fused.h"
#include "pipeline.h"

namespace afq {
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}
}  // namespace afq

using namespace afq;

namespace afq {

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

int dalloc(void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(AFQ_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  return AFQ_OK;
}

void release(afq_job* j) {
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  fr(j->h1);
  fr(j->chol);
  fr(j->mo[0]);
  if (!j->rhf) fr(j->mo[1]);
  fr(j->rdm1);
  fr(j->rot_h1[0]);
  if (!j->rhf) fr(j->rot_h1[1]);
  fr(j->rot_chol[0]);
  if (!j->rhf) fr(j->rot_chol[1]);
  fr(j->mf_shifts);
  fr(j->mreal);
  fr(j->h0_prop);
  fr(j->exp_h1[0]);
  fr(j->exp_h1[1]);
  for (int s = 0; s < 2; ++s) {
    fr(j->S[s]);
    fr(j->Sinv[s]);
    fr(j->det[s]);
    fr(j->green[s]);
  }
  fr(j->fb);
  fr(j->xs);
  fr(j->terms);
  fr(j->vhs);
  for (auto& t : j->tbuf) fr(t);
  fr(j->ov_new);
  fr(j->coul);
  fr(j->kpart);
  fr(j->ene1);
  fr(j->energies);
  fr(j->fields);
  fr(j->wgather);
  fr(j->sr_idx);
  fr(j->cum);
  fr(j->tmp_nn);
  fr(j->gemm_ws);
  fr(j->chol32);
  fr(j->cholp);
  fr(j->vhsp);
  fr(j->vhsp32);
  fr(j->exp_h1m32[0]);
  fr(j->exp_h1m32[1]);
  fr(j->rot_chol32[0]);
  if (!j->rhf) fr(j->rot_chol32[1]);
  fr(j->vhs32);
  fr(j->e1s[0]);
  fr(j->e1s[1]);
  fr(j->vsplit);
  fr(j->xwide);
  fr(j->t32[0]);
  fr(j->t32[1]);
  fr(j->rcs[0]);
  if (!j->rhf) fr(j->rcs[1]);
  fr(j->rcs_fb[0]);  // rcs_fb[1] points into it
  fr(j->gsplit);
  fr(j->rct32[0]);
  if (!j->rhf) fr(j->rct32[1]);
  fr(j->gt32);
  fr(j->lt_split);
  fr(j->xs_split);
  fr(j->dscratch);
  for (auto& p : j->prof) {
    for (auto& sp : p.spans) {
      (void)hipEventDestroy(sp.first);
      (void)hipEventDestroy(sp.second);
    }
    for (auto ev : p.pool) (void)hipEventDestroy(ev);
    p.spans.clear();
    p.pool.clear();
  }
}

// Walker panel of spin s: pointer and row stride (columns of the buffer).
inline void panel(const afq_job* j, int s, const void* up, const void* dn, const double2** p, long* ld) {
  if (j->restricted) {
    *p = static_cast<const double2*>(up);
    *ld = j->wcols;
  } else {
    *p = static_cast<const double2*>(s == 0 ? up : dn);
    *ld = j->nocc[s];
  }
}

int check_walkers(const afq_job* j, const void* up, const void* dn) {
  if (!up) return fail(AFQ_E_INVALID, "walkers_up is NULL");
  if (!j->restricted && !dn) return fail(AFQ_E_INVALID, "unrestricted walkers need walkers_dn");
  return AFQ_OK;
}

// S_s = T_s^H W_s ; det_s ; optionally S_s^-1 and G_s = (W_s S_s^-1)^T.  With
// have_inv the caller guarantees that S_s^-1 and det_s already belong to these
// walkers (computed by the previous call on them), and only G_s is formed.
int green_pipeline(afq_job* j, const void* up, const void* dn, bool want_inv, bool want_green, hipStream_t st,
                   bool have_inv) {
  const int N = j->N, W = j->W;
  if (!have_inv) {
    DetInvSet sets[2];
    Gemm gs[2];
    for (int s = 0; s < j->nspin_meas; ++s) {
      const int O = j->nocc[s];
      const double2* P;
      long ld;
      panel(j, s, up, dn, &P, &ld);
      Gemm& g = gs[s];
      g.M = O;
      g.N = O;
      g.K = N;
      g.batch = W;
      g.a_trans = true;
      g.A = j->mo[s];
      g.lda = O;
      g.sA = 0;
      g.b_cplx = true;
      g.B = P;
      g.ldb = ld;
      g.sB = (long)N * ld;
      g.C = j->S[s];
      g.ldc = O;
      g.sC = (long)O * O;
      // above n = 64 a determinant alone is factored in place in S
      sets[s] = DetInvSet{j->S[s], (long)O * O, O, j->det[s], want_inv ? j->Sinv[s] : nullptr, (long)O * O, j->S[s]};
    }
    {
      ProfScope ps(j, AFQ_K_OVLP_GEMM, st);
      AFQ_TRY(j->nspin_meas == 2 ? gemm_f64_pair(gs[0], gs[1], st) : gemm_f64(gs[0], st));
    }
    ProfScope ps(j, AFQ_K_DETINV, st);
    AFQ_TRY(batched_det_inv_sets(sets, j->nspin_meas, W, st));
  }
  if (!want_green) return AFQ_OK;
  Gemm hs[2];
  for (int s = 0; s < j->nspin_meas; ++s) {
    const int O = j->nocc[s];
    const double2* P;
    long ld;
    panel(j, s, up, dn, &P, &ld);
    Gemm& h = hs[s];
    h.M = O;
    h.N = N;
    h.K = O;
    h.batch = W;
    h.a_cplx = true;
    h.a_trans = true;
    h.A = j->Sinv[s];
    h.lda = O;
    h.sA = (long)O * O;
    h.b_cplx = true;
    h.b_trans = true;
    h.B = P;
    h.ldb = ld;
    h.sB = (long)N * ld;
    h.C = j->green[s];
    h.ldc = N;
    h.sC = (long)O * N;
  }
  ProfScope ps(j, AFQ_K_GREEN, st);
  return j->nspin_meas == 2 ? gemm_f64_pair(hs[0], hs[1], st) : gemm_f64(hs[0], st);
}

int overlap_from_dets(afq_job* j, double2* out, hipStream_t st) {
  ProfScope ps(j, AFQ_K_ENERGY_MISC, st);
  return k_overlap_combine(j->det[0], j->one_spin ? nullptr : j->det[1], out, j->W, st);
}

// Split-BF16 force bias of a UHF trial: both spins as one product over the
// concatenated zero-padded (Opad, Npad) orbital pairs, spin 1's k steps after
// spin 0's in every row of the split Green's functions and of rcs_fb (one
// allocation; rcs_fb[1] points at spin 1's part).  fb_kcat is that depth.
bool fb_cat(const afq_job* j) { return j->bf16x3 && !j->rhf; }
long fb_kspin(const afq_job* j, int s) { return (long)bf16_opad(j->nocc[s]) * bf16_npad(j->N); }
long fb_kcat(const afq_job* j) { return fb_kspin(j, 0) + (j->rhf ? 0 : fb_kspin(j, 1)); }

// fb = sum_s f_s * Gflat_s (W x O_s N) . rot_chol_s^T (O_s N x G), f = 2 for one spin channel.
int force_bias_gemm(afq_job* j, double2* out, hipStream_t st) {
  if (fb_cat(j)) {
    ProfScope ps(j, AFQ_K_FORCE_BIAS, st);
    const long K = fb_kcat(j);
    for (int s = 0; s < 2; ++s)
      AFQ_TRY(k_split_bf16(reinterpret_cast<const double*>(j->green[s]), true, j->nocc[s], j->N, j->W, true,
                           j->gsplit + (s ? 4 * fb_kspin(j, 0) : 0), st, 4 * K));
    Bf3 b;
    b.M = j->W;
    b.N = j->G;
    b.K = (int)K;
    b.a_cplx = true;
    b.A = j->gsplit;
    b.lda = 4 * K;
    b.pA = 32;
    b.kA = 4 * 32;
    b.a_rows = j->Wcap;
    b.B = j->rcs_fb[0];
    b.ldb = 2 * K;
    b.pB = 32;
    b.kB = 2 * 32;
    b.b_rows = j->G;
    b.C = out;
    b.ldc = j->G;
    b.c_f64 = true;
    b.ws = j->gemm_ws;
    b.ws_elems = j->gemm_ws_elems;
    return gemm_bf3(b, st);
  }
  for (int s = 0; s < j->nspin_meas; ++s) {
    const int O = j->nocc[s];
    Gemm g;
    g.M = j->W;
    g.N = j->G;
    g.K = O * j->N;
    g.a_cplx = true;
    g.A = j->green[s];
    g.lda = (long)O * j->N;
    g.b_trans = true;
    g.B = j->rot_chol[s];
    g.ldb = (long)O * j->N;
    g.C = out;
    g.ldc = j->G;
    g.alpha = double2{j->one_spin ? 2.0 : 1.0, 0.0};
    g.beta = double2{s > 0 ? 1.0 : 0.0, 0.0};
    g.ws = j->gemm_ws;
    g.ws_elems = j->gemm_ws_elems;
    ProfScope ps(j, AFQ_K_FORCE_BIAS, st);
    if (j->bf16x3) {
      // over the flattened, zero-padded (Opad, Npad) orbital pairs: split G_w, reuse the split rot_chol
      const long P = (long)bf16_opad(O) * bf16_npad(j->N);
      AFQ_TRY(k_split_bf16(reinterpret_cast<const double*>(j->green[s]), true, O, j->N, j->W, true, j->gsplit, st));
      Bf3 b;
      b.M = j->W;
      b.N = j->G;
      b.K = (int)P;
      b.a_cplx = true;
      b.A = j->gsplit;
      b.lda = 4 * P;
      b.pA = 32;
      b.kA = 4 * 32;
      b.a_rows = j->Wcap;
      b.B = j->rcs_fb[s];
      b.ldb = 2 * P;
      b.pB = 32;
      b.kB = 2 * 32;
      b.b_rows = j->G;
      b.alpha = g.alpha;
      b.beta = g.beta;
      b.C = out;
      b.ldc = j->G;
      b.c_f64 = true;
      b.ws = j->gemm_ws;
      b.ws_elems = j->gemm_ws_elems;
      AFQ_TRY(gemm_bf3(b, st));
    } else if (j->lp) {
      g.B = j->rot_chol32[s];
      AFQ_TRY(gemm_lp(LpKind::FORCE_BIAS, false, g, st));
    } else {
      AFQ_TRY(gemm_f64(g, st));
    }
  }
  return AFQ_OK;
}

int energy_pipeline(afq_job* j, const void* up, const void* dn, double2* out, hipStream_t st, bool have_green) {
  if (!have_green) AFQ_TRY(green_pipeline(j, up, dn, true, true, st));
  const bool two = !j->one_spin;
  {
    ProfScope ps(j, AFQ_K_ENERGY_MISC, st);
    AFQ_TRY(k_ene1(j->green[0], j->rot_h1[0], j->nocc[0], two ? j->green[1] : nullptr, two ? j->rot_h1[1] : nullptr,
                   j->nocc[1], j->N, two ? 1.0 : 2.0, j->ene1, j->W, st));
  }
  int nch[2] = {1, 1};
  for (int s = 0; s < j->nspin_meas; ++s) {
    double2* coul = j->coul + (long)s * j->W * j->G;
    double2* kpart = j->kpart + (long)s * j->W * j->kchunk_max;
    if (j->bf16x3 && j->nocc[s] <= kExchangeRegMaxO) {
      {
        ProfScope pm(j, AFQ_K_ENERGY_MISC, st);
        AFQ_TRY(k_split_bf16(reinterpret_cast<const double*>(j->green[s]), true, j->nocc[s], j->N, j->W, false,
                             j->gsplit, st));
      }
      ProfScope ps(j, AFQ_K_EXCHANGE, st);
      AFQ_TRY(k_exchange_bf16(j->rcs[s], j->gsplit, j->nocc[s], j->N, j->G, j->W, coul, kpart, &nch[s], st));
      continue;
    }
    if (j->lp && j->nocc[s] <= kExchangeRegMaxO) {  // FP32
      {
        ProfScope pm(j, AFQ_K_ENERGY_MISC, st);
        AFQ_TRY(k_split_f32(reinterpret_cast<const double*>(j->green[s]), true, j->nocc[s], j->N, j->W, j->gt32, st));
      }
      ProfScope ps(j, AFQ_K_EXCHANGE, st);
      AFQ_TRY(k_exchange_f32(j->rct32[s], j->gt32, j->nocc[s], j->N, j->G, j->W, coul, kpart, &nch[s], st));
      continue;
    }
    ProfScope ps(j, AFQ_K_EXCHANGE, st);
    if (j->nocc[s] > kExchangeRegMaxO)
      AFQ_TRY(k_exchange_wide(j->rot_chol[s], j->green[s], j->nocc[s], j->N, j->G, j->W, j->xwide, j->xwide_elems,
                              coul, kpart, &nch[s], st));
    else
      AFQ_TRY(k_exchange(j->rot_chol[s], j->green[s], j->nocc[s], j->N, j->G, j->W, coul, kpart, &nch[s], st));
  }
  ProfScope ps(j, AFQ_K_ENERGY_MISC, st);
  return k_energy_finalize(j->ene1, j->coul, two ? j->coul + (long)j->W * j->G : nullptr, j->kpart,
                           two ? j->kpart + (long)j->W * j->kchunk_max : nullptr, j->G, nch[0], nch[1], j->h0, out,
                           j->W, st);
}

// Column blocks of the propagation panel: spin s occupies columns
// [off[s], off[s] + c[s]) of a (W, N, ctot) array.
struct PanelCols {
  int c[2] = {0, 0}, off[2] = {0, 0}, ctot = 0;
  explicit PanelCols(const afq_job* j) {
    for (int s = 0; s < j->nspin_prop; ++s) {
      c[s] = j->restricted ? j->wcols : j->nocc[s];
      off[s] = ctot;
      ctot += c[s];
    }
  }
};

// dst[w] (N x c, row stride ldd, batch stride sd) = exp(-dt/2 h1)_s . src[w] (row stride lds, batch stride ss)
int apply_exp_h1(afq_job* j, const double2* src, long lds, long ss, int s, double2* dst, long ldd, long sd,
                 hipStream_t st) {
  const int N = j->N;
  Gemm g;
  g.M = N;
  g.N = (int)(j->restricted ? j->wcols : j->nocc[s]);
  g.K = N;
  g.batch = j->W;
  g.A = j->exp_h1[j->restricted ? 0 : s];
  g.lda = N;
  g.sA = 0;
  g.b_cplx = true;
  g.B = src;
  g.ldb = lds;
  g.sB = ss;
  g.C = dst;
  g.ldc = ldd;
  g.sC = sd;
  ProfScope ps(j, AFQ_K_EXPH1, st);
  return gemm_f64(g, st);
}

// dst = src + (exp(-dt/2 h1) - I) src with the product on FP32 MFMA: rounding
// E - I (entries of order dt |h1|) to FP32 perturbs the effective h1 by a
// relative 6e-8, and the identity part is carried exactly in complex128.
int apply_exp_h1_lp(afq_job* j, const double2* src, long lds, long ss, int s, double2* dst, long ldd, long sd,
                    hipStream_t st) {
  const int N = j->N;
  Gemm g;
  g.M = N;
  g.N = (int)(j->restricted ? j->wcols : j->nocc[s]);
  g.K = N;
  g.batch = j->W;
  g.A = j->exp_h1m32[j->restricted ? 0 : s];
  g.lda = N;
  g.sA = 0;
  g.b_cplx = true;
  g.B = src;
  g.ldb = lds;
  g.sB = ss;
  g.E = src;
  g.lde = lds;
  g.sE = ss;
  g.D = dst;
  g.ldd = ldd;
  g.sD = sd;
  ProfScope ps(j, AFQ_K_EXPH1, st);
  return gemm_lp(LpKind::ONE_BODY, false, g, st);
}

// Reduced precision: VHS and the Taylor terms in complex64 on FP32 / split-BF16
// MFMA, the running Taylor sum in complex128 (the reference's
// vhs_mixed_precision split, propagation.py:154-176, 196-203), and the
// exp(-dt/2 h1) applications as identity plus an FP32 product.
int trotprop_lp(afq_job* j, void* up, void* dn, const double2* xs, hipStream_t st) {
  const int N = j->N, W = j->W;
  const long NN = (long)N * N;
  {
    // symmetric L: packed columns into vhsp32, which the Taylor products read
    const bool sym = j->sym_chol;
    const long ncols = sym ? j->npk : NN;
    float2* out = sym ? j->vhsp32 : j->vhs32;
    Gemm g;
    g.M = W;
    g.N = (int)ncols;
    g.K = j->G;
    g.a_cplx = true;
    g.A = xs;
    g.lda = j->G;
    g.B = j->chol32;
    g.ldb = ncols;
    g.C = out;
    g.ldc = ncols;
    g.alpha = double2{0.0, std::sqrt(j->d.dt)};
    ProfScope ps(j, AFQ_K_VHS, st);
    if (j->bf16x3) {
      // on pre-split operands: x~ split per step, L^T split once in set_hamiltonian
      const int gpad = bf16_npad(j->G);
      AFQ_TRY(k_split_rows_bf16(reinterpret_cast<const double*>(xs), true, W, j->G, j->G, gpad, 4L * gpad, 32,
                                4 * 32, j->xs_split, st));
      Bf3 v;
      v.M = W;
      v.N = (int)ncols;
      v.K = gpad;
      v.a_cplx = true;
      v.A = j->xs_split;
      v.lda = 4L * gpad;
      v.pA = 32;
      v.kA = 4 * 32;
      v.a_rows = bf16_opad(j->Wcap);
      v.B = j->lt_split;
      v.ldb = 2L * gpad;
      v.pB = 32;
      v.kB = 2 * 32;
      v.b_rows = (int)(16 * ((ncols + 15) / 16));
      v.alpha = g.alpha;
      v.C = out;
      v.ldc = ncols;
      AFQ_TRY(gemm_bf3(v, st));
    } else {
      AFQ_TRY(gemm_lp(LpKind::VHS, false, g, st));
    }
  }
  PanelCols pc(j);
#ifndef AFQ_NO_FUSED_TROT
  if (j->e1s[0]) {
    // split-BF16 x3: one-body products, Taylor terms and running sum in one kernel
    TrotFused f;
    f.N = N;
    f.W = W;
    f.nspin = j->nspin_prop;
    f.nterms = j->d.n_exp_terms;
    for (int s = 0; s < j->nspin_prop; ++s) {
      f.wb[s] = static_cast<double2*>(s == 0 ? up : dn);
      f.c[s] = pc.c[s];
      f.e1[s] = j->e1s[s];
    }
    f.vs = j->vsplit;
    ProfScope ps(j, AFQ_K_TAYLOR, st);
    AFQ_TRY(k_split_vhs_bf16(j->sym_chol ? j->vhsp32 : j->vhs32, j->sym_chol, N, W, j->vsplit, st));
    return trot_fused_bf3(f, st);
  }
#endif
  // The walker buffers of both spins are propagated as the column blocks of
  // one (W, N, c_up + c_dn) panel, so each Taylor term is one product that
  // reads the shared VHS once.
  const long per = (long)N * pc.ctot;
  double2* t0 = j->tbuf[0];
  double2* acc = j->tbuf[2];
  for (int s = 0; s < j->nspin_prop; ++s)
    AFQ_TRY(apply_exp_h1_lp(j, static_cast<const double2*>(s == 0 ? up : dn), pc.c[s], (long)N * pc.c[s], s,
                            t0 + pc.off[s], pc.ctot, per, st));
  // the running sum starts as t0 + (first term), written by the first product
  if (j->d.n_exp_terms < 2)
    AFQ_HIP(hipMemcpyAsync(acc, t0, sizeof(double2) * per * W, hipMemcpyDeviceToDevice, st));
  int cur = -1;  // -1: previous term is t0 (complex128)
  for (int n = 1; n < j->d.n_exp_terms; ++n) {
    const int nxt = (cur == 0) ? 1 : 0;
    Gemm g;
    g.M = N;
    g.N = pc.ctot;
    g.K = N;
    g.batch = W;
    g.a_cplx = true;
    g.A = j->sym_chol ? j->vhsp32 : j->vhs32;
    g.lda = N;
    g.sA = j->sym_chol ? j->npk : NN;
    g.a_sym = j->sym_chol ? N : 0;
    g.b_cplx = true;
    g.B = cur < 0 ? static_cast<const void*>(t0) : static_cast<const void*>(j->t32[cur]);
    g.ldb = pc.ctot;
    g.sB = per;
    g.C = j->t32[nxt];
    g.ldc = pc.ctot;
    g.sC = per;
    g.D = acc;
    g.ldd = pc.ctot;
    g.sD = per;
    if (n == 1) {
      g.E = t0;
      g.lde = pc.ctot;
      g.sE = per;
    }
    g.alpha = double2{1.0 / n, 0.0};
    {
      ProfScope ps(j, AFQ_K_TAYLOR, st);
      AFQ_TRY(gemm_lp(cur < 0 ? LpKind::TAYLOR_D : LpKind::TAYLOR_F, j->bf16x3, g, st));
    }
    cur = nxt;
  }
  for (int s = 0; s < j->nspin_prop; ++s)
    AFQ_TRY(apply_exp_h1_lp(j, acc + pc.off[s], pc.ctot, per, s, static_cast<double2*>(s == 0 ? up : dn), pc.c[s],
                            (long)N * pc.c[s], st));
  return AFQ_OK;
}

// W <- E . Taylor(VHS) . E . W for each walker buffer.
int trotprop(afq_job* j, void* up, void* dn, const double2* xs, hipStream_t st) {
  if (j->lp) return trotprop_lp(j, up, dn, xs, st);
  const int N = j->N, W = j->W;
  const long NN = (long)N * N;
  {
    const bool sym = j->sym_chol;
    const long ncols = sym ? j->npk : NN;
    Gemm g;
    g.M = W;
    g.N = (int)ncols;
    g.K = j->G;
    g.a_cplx = true;
    g.A = xs;
    g.lda = j->G;
    g.B = sym ? j->cholp : j->chol;
    g.ldb = ncols;
    g.C = sym ? static_cast<void*>(j->vhsp) : static_cast<void*>(j->vhs);
    g.ldc = ncols;
    g.alpha = double2{0.0, std::sqrt(j->d.dt)};
    ProfScope ps(j, AFQ_K_VHS, st);
    AFQ_TRY(gemm_f64(g, st));
  }
  PanelCols pc(j);
  const long per = (long)N * pc.ctot;
  double2* t0 = j->tbuf[0];
  double2* t1 = j->tbuf[1];
  double2* acc = j->tbuf[2];
  for (int s = 0; s < j->nspin_prop; ++s)
    AFQ_TRY(apply_exp_h1(j, static_cast<const double2*>(s == 0 ? up : dn), pc.c[s], (long)N * pc.c[s], s,
                         t0 + pc.off[s], pc.ctot, per, st));
  AFQ_HIP(hipMemcpyAsync(acc, t0, sizeof(double2) * per * W, hipMemcpyDeviceToDevice, st));
  double2* src = t0;
  double2* dst = t1;
  for (int n = 1; n < j->d.n_exp_terms; ++n) {
    Gemm g;
    g.M = N;
    g.N = pc.ctot;
    g.K = N;
    g.batch = W;
    g.a_cplx = true;
    // symmetric L: the Taylor products read the packed VHS directly
    g.A = j->sym_chol ? static_cast<const void*>(j->vhsp) : static_cast<const void*>(j->vhs);
    g.lda = N;
    g.sA = j->sym_chol ? j->npk : NN;
    g.a_sym = j->sym_chol ? N : 0;
    g.b_cplx = true;
    g.B = src;
    g.ldb = pc.ctot;
    g.sB = per;
    g.C = dst;
    g.ldc = pc.ctot;
    g.sC = per;
    g.D = acc;
    g.ldd = pc.ctot;
    g.sD = per;
    g.alpha = double2{1.0 / n, 0.0};
    {
      ProfScope ps(j, AFQ_K_TAYLOR, st);
      AFQ_TRY(gemm_f64(g, st));
    }
    double2* t = src;
    src = dst;
    dst = t;
  }
  for (int s = 0; s < j->nspin_prop; ++s)
    AFQ_TRY(apply_exp_h1(j, acc + pc.off[s], pc.ctot, per, s, static_cast<double2*>(s == 0 ? up : dn), pc.c[s],
                         (long)N * pc.c[s], st));
  return AFQ_OK;
}

int propagate(afq_job* j, void* up, void* dn, const double* fields, double* weights, double2* overlaps,
              double* scalars, hipStream_t st, bool inv_ready, bool keep_inv) {
  AFQ_TRY(green_pipeline(j, up, dn, true, true, st, inv_ready));
  AFQ_TRY(force_bias_gemm(j, j->fb, st));
  {
    ProfScope ps(j, AFQ_K_FIELD, st);
    AFQ_TRY(k_field_shift(j->fb, fields, j->mf_shifts, j->G, std::sqrt(j->d.dt), j->xs, j->terms, j->W, st));
  }
  AFQ_TRY(trotprop(j, up, dn, j->xs, st));
  AFQ_TRY(green_pipeline(j, up, dn, keep_inv, false, st));
  AFQ_TRY(overlap_from_dets(j, j->ov_new, st));
  ProfScope ps(j, AFQ_K_WEIGHT, st);
  return k_weight_update(j->ov_new, overlaps, j->terms, weights, scalars, j->h0_prop, j->d.dt, j->W, st);
}

// Cholesky QR, twice (CholeskyQR2): S = W^H W on the matrix cores, S = R^H R
// with a positive real diagonal, W <- W R^-1.  This is the QR factorisation
// whose R has a positive real diagonal -- the one CGS2 computes -- and is
// accurate for the well-conditioned panels of a block (re-orthonormalised
// every block); the second pass restores orthogonality to rounding.  The
// first pass writes the scratch panel and the second reads it and writes W.
int cholqr2_panel(afq_job* j, double2* Wp, int c, int s, hipStream_t st) {
  const int N = j->N, W = j->W;
  const long per = (long)N * c, cc = (long)c * c;
  double2* tmp = j->wgather;
  double2* S = j->S[s];
  double2* Rinv = j->Sinv[s];
  for (int pass = 0; pass < 2; ++pass) {
    double2* in = pass == 0 ? Wp : tmp;
    double2* out = pass == 0 ? tmp : Wp;
    Gemm g;
    g.M = c;
    g.N = c;
    g.K = N;
    g.batch = W;
    g.a_cplx = true;
    g.a_trans = true;
    g.a_conj = true;
    g.A = in;
    g.lda = c;
    g.sA = per;
    g.b_cplx = true;
    g.B = in;
    g.ldb = c;
    g.sB = per;
    g.C = S;
    g.ldc = c;
    g.sC = cc;
    AFQ_TRY(gemm_f64(g, st));
    AFQ_TRY(batched_cholesky_upper(S, cc, c, W, S, cc, st));
    AFQ_TRY(batched_det_inv(S, cc, c, W, j->det[s], Rinv, cc, st));
    Gemm q;
    q.M = N;
    q.N = c;
    q.K = c;
    q.batch = W;
    q.a_cplx = true;
    q.A = in;
    q.lda = c;
    q.sA = per;
    q.b_cplx = true;
    q.B = Rinv;
    q.ldb = c;
    q.sB = cc;
    q.C = out;
    q.ldc = c;
    q.sC = per;
    AFQ_TRY(gemm_f64(q, st));
  }
  return AFQ_OK;
}

// cholqr2_panel for both spins of unrestricted walkers at once: every product
// and factorisation of a pass covers the two spins in one launch (spin 1's
// scratch panel follows spin 0's in wgather).  Same arithmetic per matrix.
int cholqr2_pair(afq_job* j, double2* W0, double2* W1, hipStream_t st) {
  const int N = j->N, W = j->W;
  double2* Wp[2] = {W0, W1};
  double2* tmp[2] = {j->wgather, j->wgather + (long)W * N * j->nocc[0]};
  for (int pass = 0; pass < 2; ++pass) {
    Gemm g[2], q[2];
    CholSet cs[2];
    DetInvSet ds[2];
    for (int s = 0; s < 2; ++s) {
      const int c = j->nocc[s];
      const long per = (long)N * c, cc = (long)c * c;
      double2* in = pass == 0 ? Wp[s] : tmp[s];
      double2* out = pass == 0 ? tmp[s] : Wp[s];
      Gemm& a = g[s];
      a.M = c;
      a.N = c;
      a.K = N;
      a.batch = W;
      a.a_cplx = true;
      a.a_trans = true;
      a.a_conj = true;
      a.A = in;
      a.lda = c;
      a.sA = per;
      a.b_cplx = true;
      a.B = in;
      a.ldb = c;
      a.sB = per;
      a.C = j->S[s];
      a.ldc = c;
      a.sC = cc;
      cs[s] = CholSet{j->S[s], cc, c, j->S[s], cc};
      ds[s] = DetInvSet{j->S[s], cc, c, j->det[s], j->Sinv[s], cc};
      Gemm& b = q[s];
      b.M = N;
      b.N = c;
      b.K = c;
      b.batch = W;
      b.a_cplx = true;
      b.A = in;
      b.lda = c;
      b.sA = per;
      b.b_cplx = true;
      b.B = j->Sinv[s];
      b.ldb = c;
      b.sB = cc;
      b.C = out;
      b.ldc = c;
      b.sC = per;
    }
    AFQ_TRY(gemm_f64_pair(g[0], g[1], st));
    AFQ_TRY(batched_cholesky_upper_sets(cs, 2, W, st));
    AFQ_TRY(batched_det_inv_sets(ds, 2, W, st));
    AFQ_TRY(gemm_f64_pair(q[0], q[1], st));
  }
  return AFQ_OK;
}

int orthonormalize(afq_job* j, void* up, void* dn, hipStream_t st) {
  ProfScope ps(j, AFQ_K_QR, st);
  if (j->restricted) return cholqr2_panel(j, static_cast<double2*>(up), j->wcols, 0, st);
  return cholqr2_pair(j, static_cast<double2*>(up), static_cast<double2*>(dn), st);
}

int require(const afq_job* j, bool meas, bool prop) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  if (!j->have_trial) return fail(AFQ_E_STATE, "trial not set");
  if (meas && !j->have_meas) return fail(AFQ_E_STATE, "measurement intermediates not built");
  if (prop && !j->have_prop) return fail(AFQ_E_STATE, "propagation intermediates not built");
  return AFQ_OK;
}

int compute_v0(afq_job* j, double* v0, hipStream_t st) {
  const int N = j->N, G = j->G;
  const long NN = (long)N * N;
  // accumulated over gamma chunks of a permuted copy: v0 = perm (N x nN) . L_chunk (nN x N)
  long gch = std::max(1L, std::min((long)G, (64L << 20) / (long)(sizeof(double) * NN)));
  double* perm = nullptr;
  AFQ_TRY(dalloc_n(&perm, (size_t)(gch * NN)));
  int rc = AFQ_OK;
  for (int g0 = 0; g0 < G && rc == AFQ_OK; g0 += (int)gch) {
    const int n = (int)std::min(gch, (long)(G - g0));
    rc = k_permute_chunk(j->chol, perm, N, g0, n, st);
    if (rc != AFQ_OK) break;
    Gemm g;
    g.M = N;
    g.N = N;
    g.K = n * N;
    g.A = perm;
    g.lda = (long)n * N;
    g.B = j->chol + (long)g0 * NN;
    g.ldb = N;
    g.C = v0;
    g.ldc = N;
    g.alpha = double2{0.5, 0.0};
    g.ws = j->gemm_ws;
    g.ws_elems = j->gemm_ws_elems;
    g.beta = double2{g0 == 0 ? 0.0 : 1.0, 0.0};
    rc = gemm_f64(g, st);
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(perm);
  return rc;
}

int expm_device(const double* A, int n, double* out, double* work, double* dscr, hipStream_t st) {
  const long nn = (long)n * n;
  double* As = work;
  double* P = work + nn;
  double* Q = work + 2 * nn;
  AFQ_TRY(k_norm1(A, n, dscr, st));
  double nrm = 0.0;
  AFQ_HIP(hipMemcpyAsync(&nrm, dscr, sizeof(double), hipMemcpyDeviceToHost, st));
  AFQ_HIP(hipStreamSynchronize(st));
  int sq = 0;
  while (nrm / std::ldexp(1.0, sq) > 0.25 && sq < 60) ++sq;
  AFQ_TRY(k_scale(As, A, std::ldexp(1.0, -sq), n, st));
  AFQ_TRY(k_identity(P, n, st));
  for (int k = 18; k >= 1; --k) {
    AFQ_TRY(k_identity(Q, n, st));
    Gemm g;
    g.M = n;
    g.N = n;
    g.K = n;
    g.A = As;
    g.lda = n;
    g.B = P;
    g.ldb = n;
    g.C = Q;
    g.ldc = n;
    g.alpha = double2{1.0 / k, 0.0};
    g.beta = double2{1.0, 0.0};
    AFQ_TRY(gemm_f64(g, st));
    std::swap(P, Q);
  }
  for (int i = 0; i < sq; ++i) {
    Gemm g;
    g.M = n;
    g.N = n;
    g.K = n;
    g.A = P;
    g.lda = n;
    g.B = P;
    g.ldb = n;
    g.C = Q;
    g.ldc = n;
    AFQ_TRY(gemm_f64(g, st));
    std::swap(P, Q);
  }
  AFQ_HIP(hipMemcpyAsync(out, P, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
  return AFQ_OK;
}

}  // namespace afq

extern "C" {

const char* afq_version(void) { return "afq 0.1.0 gfx950"; }
const char* afq_last_error(void) { return afq::g_err.c_str(); }

}  // extern "C"

namespace {

int validate_desc(const afq_job_desc* d) {
  if (d->norb < 1 || d->nchol < 1 || d->n_walkers < 1 || d->nelec_up < 1 || d->nelec_dn < 1)
    return fail(AFQ_E_INVALID, "sizes must be positive");
  if (d->nelec_up > d->norb || d->nelec_dn > d->norb) return fail(AFQ_E_INVALID, "nelec exceeds norb");
  if (d->walker_type != AFQ_WALKERS_RESTRICTED && d->walker_type != AFQ_WALKERS_UNRESTRICTED)
    return fail(AFQ_E_INVALID, "unknown walker_type");
  if (d->trial_type != AFQ_TRIAL_RHF && d->trial_type != AFQ_TRIAL_UHF) return fail(AFQ_E_INVALID, "unknown trial_type");
  if (d->trial_type == AFQ_TRIAL_RHF && d->nelec_up != d->nelec_dn)
    return fail(AFQ_E_INVALID, "RHF requires equal number of up and down electrons.");
  if (d->walker_type == AFQ_WALKERS_RESTRICTED && d->nelec_dn > d->nelec_up)
    return fail(AFQ_E_INVALID, "restricted walkers need nelec_dn <= nelec_up");
  if (d->n_exp_terms < 1) return fail(AFQ_E_INVALID, "n_exp_terms must be >= 1");
  if (!(d->dt > 0.0)) return fail(AFQ_E_INVALID, "dt must be positive");
  if (d->precision != AFQ_PREC_FP64 && d->precision != AFQ_PREC_FP32 && d->precision != AFQ_PREC_BF16X3)
    return fail(AFQ_E_INVALID, "unknown precision");
  if (d->nelec_up > kDetInvMaxN || d->nelec_dn > kDetInvMaxN)
    return fail(AFQ_E_UNSUPPORTED, "nocc > " + std::to_string(kDetInvMaxN) + " not supported");
  if (d->nchol > 65535 || d->n_walkers > 65535) return fail(AFQ_E_UNSUPPORTED, "nchol and n_walkers must be <= 65535");
  return AFQ_OK;
}

void init_job(afq_job* j, const afq_job_desc* d) {
  j->d = *d;
  j->N = d->norb;
  j->G = d->nchol;
  j->W = d->n_walkers;
  j->Wcap = d->n_walkers;
  j->nocc[0] = d->nelec_up;
  j->nocc[1] = d->nelec_dn;
  j->wcols = d->nelec_up;
  j->restricted = d->walker_type == AFQ_WALKERS_RESTRICTED;
  j->rhf = d->trial_type == AFQ_TRIAL_RHF;
  j->one_spin = j->rhf && j->restricted;
  j->nspin_meas = j->one_spin ? 1 : 2;
  j->nspin_prop = j->restricted ? 1 : 2;
  j->lp = d->precision != AFQ_PREC_FP64;
  j->bf16x3 = d->precision == AFQ_PREC_BF16X3;
  j->kchunk_max = exchange_nchunk(j->G);
}

// Allocates every device buffer of the job (dry: only sums the bytes).
int alloc_job(afq_job* j, bool dry, int64_t* total) {
  const long N = j->N, G = j->G, W = j->W;
  const long wmax = std::max(j->wcols, std::max(j->nocc[0], j->nocc[1]));
  // propagation panel: both spins' columns side by side for unrestricted walkers
  const long wprop = j->restricted ? j->wcols : (long)j->nocc[0] + j->nocc[1];
  const int kch = j->kchunk_max;
  int rc = AFQ_OK;
  *total = 0;
#define A_(p, n)                                                   \
  do {                                                             \
    *total += (int64_t)(n) * (int64_t)sizeof(*(p));                \
    if (!dry && rc == AFQ_OK) rc = dalloc_n(&(p), (size_t)(n));   \
  } while (0)
  A_(j->h1, 2 * N * N);
  A_(j->chol, G * N * N);
  A_(j->mo[0], N * j->nocc[0]);
  if (!j->rhf) A_(j->mo[1], N * j->nocc[1]);
  A_(j->rdm1, 2 * N * N);
  A_(j->rot_h1[0], j->nocc[0] * N);
  if (!j->rhf) A_(j->rot_h1[1], j->nocc[1] * N);
  A_(j->rot_chol[0], G * j->nocc[0] * N);
  if (!j->rhf) A_(j->rot_chol[1], G * j->nocc[1] * N);
  A_(j->mf_shifts, G);
  A_(j->mreal, 2 * G);
  A_(j->h0_prop, 1);
  A_(j->exp_h1[0], N * N);
  if (!j->restricted) A_(j->exp_h1[1], N * N);
  for (int s = 0; s < 2; ++s) {
    const long O = j->nocc[s];
    A_(j->S[s], W * O * O);
    A_(j->Sinv[s], W * O * O);
    A_(j->det[s], W);
    A_(j->green[s], W * O * N);
  }
  const int omax = std::max(j->nocc[0], j->nocc[1]);
  if (omax > kExchangeRegMaxO) {
    j->xwide_elems = exchange_wide_ws_elems(omax, (int)G, (int)W);
    A_(j->xwide, j->xwide_elems);
  }
  A_(j->fb, W * G);
  A_(j->xs, W * G);
  A_(j->terms, 2 * W);
  // packed symmetric VHS product (its L copy lives in chol32 / lt_split in
  // reduced precision, which have at least npk columns)
  {
    const long npk = sym_packed_cols((int)N);
    j->npk = (npk * 20 <= N * N * 17) ? npk : 0;
  }
  if (j->npk > 0) {
    if (j->lp) {
      A_(j->vhsp32, W * j->npk);
    } else {
      A_(j->cholp, G * j->npk);
      A_(j->vhsp, W * j->npk);
    }
  }
  if (j->lp) {
    // complex64 VHS and Taylor terms; the complex128 Taylor buffers hold
    // exp(-dt/2 h1) W and the running sum.  FP32: FP32 copies of L and rot_chol.
    // BF16x3: L^T and the shifted fields as split-BF16 planes for the VHS product,
    // rot_chol as split planes (exchange) and FP32 (force bias), and one walker
    // batch of split Green's functions.
    if (j->bf16x3) {
      const long np = bf16_npad(j->N), gp = bf16_npad(j->G);
      A_(j->rcs[0], G * 2 * bf16_opad(j->nocc[0]) * np);
      if (!j->rhf) A_(j->rcs[1], G * 2 * bf16_opad(j->nocc[1]) * np);
      A_(j->rcs_fb[0], G * 2 * fb_kcat(j));  // both spins' rows side by side for a UHF trial
      A_(j->gsplit, W * 4 * std::max(fb_kcat(j), (long)bf16_opad((int)std::max(j->nocc[0], j->nocc[1])) * np));
      A_(j->lt_split, 16 * ((N * N + 15) / 16) * 2 * gp);
      A_(j->xs_split, bf16_opad((int)W) * 4 * gp);
      if (trot_fused_supported((int)N)) {
        A_(j->e1s[0], trot_fused_e1_elems((int)N));
        if (!j->restricted) A_(j->e1s[1], trot_fused_e1_elems((int)N));
        A_(j->vsplit, trot_fused_vhs_elems((int)N, (int)W));
      }
    } else {
      A_(j->chol32, G * N * N);
      const long np = bf16_npad(j->N);
      int omax = 0;
      for (int s = 0; s < (j->rhf ? 1 : 2); ++s)
        if (j->nocc[s] <= kExchangeRegMaxO) {
          A_(j->rct32[s], G * bf16_opad(j->nocc[s]) * np);
          omax = std::max(omax, j->nocc[s]);
        }
      if (omax > 0) A_(j->gt32, W * 2 * bf16_opad(omax) * np);
    }
    A_(j->rot_chol32[0], G * j->nocc[0] * N);
    if (!j->rhf) A_(j->rot_chol32[1], G * j->nocc[1] * N);
    A_(j->exp_h1m32[0], N * N);
    if (!j->restricted) A_(j->exp_h1m32[1], N * N);
    A_(j->vhs32, W * N * N);
    for (auto& t : j->t32) A_(t, W * N * wprop);
    A_(j->tbuf[0], W * N * wprop);
    A_(j->tbuf[2], W * N * wprop);
  } else {
    A_(j->vhs, W * N * N);
    for (auto& t : j->tbuf) A_(t, W * N * wprop);
  }
  A_(j->ov_new, W);
  A_(j->coul, 2 * W * G);
  A_(j->kpart, 2 * W * kch);
  A_(j->ene1, W);
  A_(j->energies, W);
  // gather buffer of the SR copy; for unrestricted walkers also both spins'
  // scratch panels of the paired QR
  A_(j->wgather, W * N * (j->restricted ? wmax : (long)j->nocc[0] + j->nocc[1]));
  A_(j->sr_idx, W);
  A_(j->cum, W);
  A_(j->tmp_nn, 8 * N * N);
  // split-K partials: up to 16 slices of the (W, G) complex force bias, or of an N x N setup product
  j->gemm_ws_elems = 16 * std::max(2 * W * G, N * N);
  A_(j->gemm_ws, j->gemm_ws_elems);
  A_(j->dscratch, 8);
#undef A_
  return rc;
}

}  // namespace

extern "C" {

int64_t afq_job_footprint(const afq_job_desc* d, int32_t n_prop_steps) {
  if (!d || validate_desc(d) != AFQ_OK) return -1;
  afq_job j;
  init_job(&j, d);
  int64_t total = 0;
  alloc_job(&j, true, &total);
  if (n_prop_steps > 0) total += (int64_t)sizeof(double) * n_prop_steps * j.W * j.G;
  return total;
}

int afq_job_create(afq_job_t* out, const afq_job_desc* d) {
  if (!out || !d) return fail(AFQ_E_INVALID, "NULL argument");
  *out = nullptr;
  AFQ_TRY(validate_desc(d));
  afq_job* j = new afq_job();
  init_job(j, d);
  int64_t total = 0;
  const int rc = alloc_job(j, false, &total);
  if (rc != AFQ_OK) {
    release(j);
    delete j;
    return rc;
  }
  if (j->vsplit) {
    // plane entries past N are never written and must read as zeros
    const hipError_t e = hipMemset(j->vsplit, 0, sizeof(uint16_t) * trot_fused_vhs_elems(j->N, j->Wcap));
    if (e != hipSuccess) {
      release(j);
      delete j;
      return fail(AFQ_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    }
  }
  if (j->rhf) {
    j->mo[1] = j->mo[0];
    j->rot_h1[1] = j->rot_h1[0];
    j->rot_chol[1] = j->rot_chol[0];
    j->rot_chol32[1] = j->rot_chol32[0];
    j->rcs[1] = j->rcs[0];
    j->rcs_fb[1] = j->rcs_fb[0];
    j->rct32[1] = j->rct32[0];
  } else if (j->rcs_fb[0]) {
    j->rcs_fb[1] = j->rcs_fb[0] + 2 * fb_kspin(j, 0);
  }
  *out = j;
  return AFQ_OK;
}

int afq_job_destroy(afq_job_t j) {
  if (!j) return AFQ_OK;
  (void)hipDeviceSynchronize();
  release(j);
  delete j;
  return AFQ_OK;
}

int afq_job_set_batch(afq_job_t j, int32_t n) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  if (n < 1 || n > j->Wcap) return fail(AFQ_E_INVALID, "n_walkers outside [1, capacity]");
  j->W = n;
  return AFQ_OK;
}

int afq_job_set_hamiltonian(afq_job_t j, double h0, const double* h1, const double* chol, void* stream) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  j->h0 = h0;
  if (!h1 && !chol) return AFQ_OK;
  if (!h1 || !chol) return fail(AFQ_E_INVALID, "h1 and chol must both be given or both be NULL");
  const long N = j->N;
  AFQ_HIP(hipMemcpyAsync(j->h1, h1, sizeof(double) * 2 * N * N, hipMemcpyDefault, S(stream)));
  AFQ_HIP(hipMemcpyAsync(j->chol, chol, sizeof(double) * j->G * N * N, hipMemcpyDefault, S(stream)));
  // Symmetric L_g (to 1e-12 of its largest entry): the VHS product covers the
  // lower block triangle only and mirrors it.
  j->sym_chol = false;
  if (j->npk > 0) {
    unsigned long long* dev = reinterpret_cast<unsigned long long*>(j->dscratch + 6);
    AFQ_TRY(k_sym_check(j->chol, j->G, (int)N, dev, S(stream)));
    unsigned long long host[2] = {0, 0};
    AFQ_HIP(hipMemcpyAsync(host, dev, sizeof(host), hipMemcpyDeviceToHost, S(stream)));
    AFQ_HIP(hipStreamSynchronize(S(stream)));
    double dmax, amax;
    std::memcpy(&dmax, &host[0], sizeof(double));
    std::memcpy(&amax, &host[1], sizeof(double));
    j->sym_chol = dmax <= 1e-12 * amax;
  }
  const long ncols = j->sym_chol ? j->npk : N * N;
  if (j->bf16x3) {
    const long gp = bf16_npad(j->G);
    AFQ_TRY(k_split_chol_t_bf16(j->chol, j->G, N * N, ncols, j->sym_chol ? (int)N : 0, (int)gp, 2 * gp, 32, 2 * 32,
                                j->lt_split, S(stream)));
  } else if (j->lp) {
    if (j->sym_chol)
      AFQ_TRY(k_sym_pack(j->chol, j->G, (int)N, nullptr, j->chol32, S(stream)));
    else
      AFQ_TRY(k_to_f32(j->chol, j->chol32, (long)j->G * N * N, S(stream)));
  } else if (j->sym_chol) {
    AFQ_TRY(k_sym_pack(j->chol, j->G, (int)N, j->cholp, nullptr, S(stream)));
  }
  j->have_ham = true;
  j->have_meas = j->have_prop = false;
  return AFQ_OK;
}

int afq_job_set_trial(afq_job_t j, const double* mo_up, const double* mo_dn, const double* rdm1, void* stream) {
  if (!j || !mo_up || !rdm1) return fail(AFQ_E_INVALID, "NULL argument");
  if (!j->rhf && !mo_dn) return fail(AFQ_E_INVALID, "UHF trial needs mo_dn");
  const long N = j->N;
  AFQ_HIP(hipMemcpyAsync(j->mo[0], mo_up, sizeof(double) * N * j->nocc[0], hipMemcpyDefault, S(stream)));
  if (!j->rhf)
    AFQ_HIP(hipMemcpyAsync(j->mo[1], mo_dn, sizeof(double) * N * j->nocc[1], hipMemcpyDefault, S(stream)));
  AFQ_HIP(hipMemcpyAsync(j->rdm1, rdm1, sizeof(double) * 2 * N * N, hipMemcpyDefault, S(stream)));
  j->have_trial = true;
  j->have_meas = j->have_prop = false;
  return AFQ_OK;
}

int afq_build_measurement_intermediates(afq_job_t j, void* stream) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  if (!j->have_ham || !j->have_trial) return fail(AFQ_E_STATE, "hamiltonian and trial must be set first");
  hipStream_t st = S(stream);
  const int N = j->N;
  const long NN = (long)N * N;
  // symmetrise h1 in place (via a temporary)
  AFQ_TRY(k_sym_h1(j->h1, j->tmp_nn, N, st));
  AFQ_HIP(hipMemcpyAsync(j->h1, j->tmp_nn, sizeof(double) * 2 * NN, hipMemcpyDeviceToDevice, st));
  const int nsp = j->rhf ? 1 : 2;
  for (int s = 0; s < nsp; ++s) {
    const int O = j->nocc[s];
    const double* hsrc = j->h1 + s * NN;
    if (j->rhf) {
      AFQ_TRY(k_axpby4(j->tmp_nn, j->h1, j->h1 + NN, 0.5, nullptr, nullptr, 0.0, NN, st));
      hsrc = j->tmp_nn;
    }
    Gemm g;
    g.M = O;
    g.N = N;
    g.K = N;
    g.a_trans = true;
    g.A = j->mo[s];
    g.lda = O;
    g.B = hsrc;
    g.ldb = N;
    g.C = j->rot_h1[s];
    g.ldc = N;
    AFQ_TRY(gemm_f64(g, st));
    Gemm r;
    r.M = O;
    r.N = N;
    r.K = N;
    r.batch = j->G;
    r.a_trans = true;
    r.A = j->mo[s];
    r.lda = O;
    r.sA = 0;
    r.B = j->chol;
    r.ldb = N;
    r.sB = NN;
    r.C = j->rot_chol[s];
    r.ldc = N;
    r.sC = (long)O * N;
    AFQ_TRY(gemm_f64(r, st));
  }
  if (j->lp)
    for (int s = 0; s < nsp; ++s) {
      AFQ_TRY(k_to_f32(j->rot_chol[s], j->rot_chol32[s], (long)j->G * j->nocc[s] * N, st));
      if (j->rct32[s]) AFQ_TRY(k_split_f32(j->rot_chol[s], false, j->nocc[s], N, j->G, j->rct32[s], st));
      if (j->bf16x3) {
        AFQ_TRY(k_split_bf16(j->rot_chol[s], false, j->nocc[s], N, j->G, false, j->rcs[s], st));
        AFQ_TRY(k_split_bf16(j->rot_chol[s], false, j->nocc[s], N, j->G, true, j->rcs_fb[s], st,
                             fb_cat(j) ? 2 * fb_kcat(j) : 0));
      }
    }
  j->have_meas = true;
  return AFQ_OK;
}

int afq_build_propagation_intermediates(afq_job_t j, void* stream) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  if (!j->have_ham || !j->have_trial) return fail(AFQ_E_STATE, "hamiltonian and trial must be set first");
  hipStream_t st = S(stream);
  const int N = j->N, G = j->G;
  const long NN = (long)N * N;
  double* rho = j->tmp_nn;           // [0]
  double* v0 = j->tmp_nn + NN;       // [1]
  double* v1 = j->tmp_nn + 2 * NN;   // [2]
  double* hmod = j->tmp_nn + 3 * NN; // [3]
  double* As = j->tmp_nn + 4 * NN;   // [4]; [5..7] expm work

  // mf_shifts = i sum_pq L_g,pq rho_pq, rho = rdm1[0] + rdm1[1]
  AFQ_TRY(k_axpby4(rho, j->rdm1, j->rdm1 + NN, 1.0, nullptr, nullptr, 0.0, NN, st));
  {
    Gemm g;
    g.M = G;
    g.N = 1;
    g.K = (int)NN;
    g.A = j->chol;
    g.lda = NN;
    g.B = rho;
    g.ldb = 1;
    g.C = j->mreal;
    g.ldc = 1;
    AFQ_TRY(gemm_f64(g, st));
  }
  double* mfr = j->mreal + G;
  AFQ_TRY(k_mf(j->mreal, G, j->h0, j->mf_shifts, mfr, j->h0_prop, st));

  // v0 = 1/2 sum_g L_g L_g
  AFQ_TRY(compute_v0(j, v0, st));
  // v1 = sum_g Re(i mf_g) L_g
  {
    Gemm g;
    g.M = 1;
    g.N = (int)NN;
    g.K = G;
    g.A = mfr;
    g.lda = G;
    g.B = j->chol;
    g.ldb = NN;
    g.C = v1;
    g.ldc = NN;
    AFQ_TRY(gemm_f64(g, st));
  }
  const int nexp = j->restricted ? 1 : 2;
  for (int s = 0; s < nexp; ++s) {
    if (j->restricted)
      AFQ_TRY(k_axpby4(hmod, j->h1, j->h1 + NN, 0.5, v0, v1, -1.0, NN, st));
    else
      AFQ_TRY(k_axpby4(hmod, j->h1 + s * NN, nullptr, 1.0, v0, v1, -1.0, NN, st));
    // exp(A), A = -dt/2 h1_mod
    AFQ_TRY(k_scale(As, hmod, -0.5 * j->d.dt, N, st));
    AFQ_TRY(expm_device(As, N, j->exp_h1[s], j->tmp_nn + 5 * NN, j->dscratch, st));
    if (j->lp) AFQ_TRY(k_minus_identity_f32(j->exp_h1[s], j->exp_h1m32[s], N, st));
    if (j->e1s[s]) AFQ_TRY(k_split_e1_bf16(j->exp_h1[s], N, j->e1s[s], st));
  }
  AFQ_HIP(hipMemcpyAsync(&j->h0_prop_host, j->h0_prop, sizeof(double2), hipMemcpyDeviceToHost, st));
  AFQ_HIP(hipStreamSynchronize(st));
  j->have_prop = true;
  return AFQ_OK;
}

int afq_get_intermediate(afq_job_t j, int which, int spin, void* dst, void* stream) {
  if (!j || !dst) return fail(AFQ_E_INVALID, "NULL argument");
  if (spin < 0 || spin > 1) return fail(AFQ_E_INVALID, "spin must be 0 or 1");
  const long N = j->N, G = j->G, O = j->nocc[spin];
  const void* src = nullptr;
  size_t bytes = 0;
  switch (which) {
    case AFQ_INT_H1: src = j->h1; bytes = sizeof(double) * 2 * N * N; break;
    case AFQ_INT_ROT_H1: src = j->rot_h1[spin]; bytes = sizeof(double) * O * N; break;
    case AFQ_INT_ROT_CHOL: src = j->rot_chol[spin]; bytes = sizeof(double) * G * O * N; break;
    case AFQ_INT_MF_SHIFTS: src = j->mf_shifts; bytes = sizeof(double2) * G; break;
    case AFQ_INT_H0_PROP: src = j->h0_prop; bytes = sizeof(double2); break;
    case AFQ_INT_EXP_H1:
      src = j->exp_h1[j->restricted ? 0 : spin];
      bytes = sizeof(double) * N * N;
      break;
    default: return fail(AFQ_E_INVALID, "unknown intermediate");
  }
  AFQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, S(stream)));
  return AFQ_OK;
}

int afq_set_intermediate(afq_job_t j, int which, int spin, const void* src, void* stream) {
  if (!j || !src) return fail(AFQ_E_INVALID, "NULL argument");
  if (spin < 0 || spin > 1) return fail(AFQ_E_INVALID, "spin must be 0 or 1");
  if (j->rhf && spin == 1 && (which == AFQ_INT_ROT_H1 || which == AFQ_INT_ROT_CHOL))
    return fail(AFQ_E_INVALID, "an RHF trial has spin-0 intermediates only");
  const long N = j->N, G = j->G, O = j->nocc[spin];
  void* dst = nullptr;
  size_t bytes = 0;
  switch (which) {
    case AFQ_INT_H1: dst = j->h1; bytes = sizeof(double) * 2 * N * N; break;
    case AFQ_INT_ROT_H1: dst = j->rot_h1[spin]; bytes = sizeof(double) * O * N; break;
    case AFQ_INT_ROT_CHOL: dst = j->rot_chol[spin]; bytes = sizeof(double) * G * O * N; break;
    case AFQ_INT_MF_SHIFTS: dst = j->mf_shifts; bytes = sizeof(double2) * G; break;
    case AFQ_INT_H0_PROP: dst = j->h0_prop; bytes = sizeof(double2); break;
    case AFQ_INT_EXP_H1:
      if (j->restricted && spin == 1) return fail(AFQ_E_INVALID, "restricted walkers use one exp_h1");
      dst = j->exp_h1[spin];
      bytes = sizeof(double) * N * N;
      break;
    default: return fail(AFQ_E_INVALID, "unknown intermediate");
  }
  AFQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, S(stream)));
  switch (which) {
    case AFQ_INT_ROT_H1: j->have_rot_h1[spin] = true; break;
    case AFQ_INT_ROT_CHOL:
      j->have_rot_chol[spin] = true;
      if (j->lp) AFQ_TRY(k_to_f32(j->rot_chol[spin], j->rot_chol32[spin], G * O * N, S(stream)));
      if (j->rct32[spin])
        AFQ_TRY(k_split_f32(j->rot_chol[spin], false, (int)O, (int)N, G, j->rct32[spin], S(stream)));
      if (j->bf16x3) {
        AFQ_TRY(k_split_bf16(j->rot_chol[spin], false, (int)O, (int)N, G, false, j->rcs[spin], S(stream)));
        AFQ_TRY(k_split_bf16(j->rot_chol[spin], false, (int)O, (int)N, G, true, j->rcs_fb[spin], S(stream),
                             fb_cat(j) ? 2 * fb_kcat(j) : 0));
      }
      break;
    case AFQ_INT_MF_SHIFTS: j->have_mf = true; break;
    case AFQ_INT_H0_PROP: j->have_h0p = true; break;
    case AFQ_INT_EXP_H1:
      j->have_exp[spin] = true;
      if (j->lp) AFQ_TRY(k_minus_identity_f32(j->exp_h1[spin], j->exp_h1m32[spin], (int)N, S(stream)));
      if (j->e1s[spin]) AFQ_TRY(k_split_e1_bf16(j->exp_h1[spin], (int)N, j->e1s[spin], S(stream)));
      break;
    default: break;
  }
  const int nsp = j->rhf ? 1 : 2;
  bool meas = true;
  for (int s = 0; s < nsp; ++s) meas = meas && j->have_rot_h1[s] && j->have_rot_chol[s];
  if (meas) j->have_meas = true;
  const bool prop = j->have_mf && j->have_h0p && j->have_exp[0] && (j->restricted || j->have_exp[1]);
  if (prop) j->have_prop = true;
  return AFQ_OK;
}

int afq_calc_overlap(afq_job_t j, const void* up, const void* dn, void* ov, void* stream) {
  AFQ_TRY(require(j, false, false));
  AFQ_TRY(check_walkers(j, up, dn));
  if (!ov) return fail(AFQ_E_INVALID, "overlaps is NULL");
  AFQ_TRY(green_pipeline(j, up, dn, false, false, S(stream)));
  return overlap_from_dets(j, static_cast<double2*>(ov), S(stream));
}

int afq_calc_green(afq_job_t j, const void* up, const void* dn, void* gup, void* gdn, void* stream) {
  AFQ_TRY(require(j, false, false));
  AFQ_TRY(check_walkers(j, up, dn));
  if (!gup) return fail(AFQ_E_INVALID, "green_up is NULL");
  hipStream_t st = S(stream);
  AFQ_TRY(green_pipeline(j, up, dn, true, true, st));
  const long W = j->W, N = j->N;
  AFQ_HIP(hipMemcpyAsync(gup, j->green[0], sizeof(double2) * W * j->nocc[0] * N, hipMemcpyDeviceToDevice, st));
  if (gdn)
    AFQ_HIP(hipMemcpyAsync(gdn, j->green[j->one_spin ? 0 : 1], sizeof(double2) * W * j->nocc[1] * N,
                           hipMemcpyDeviceToDevice, st));
  return AFQ_OK;
}

int afq_calc_force_bias(afq_job_t j, const void* up, const void* dn, void* fb, void* stream) {
  AFQ_TRY(require(j, true, false));
  AFQ_TRY(check_walkers(j, up, dn));
  if (!fb) return fail(AFQ_E_INVALID, "force_bias is NULL");
  AFQ_TRY(green_pipeline(j, up, dn, true, true, S(stream)));
  return force_bias_gemm(j, static_cast<double2*>(fb), S(stream));
}

int afq_calc_energy(afq_job_t j, const void* up, const void* dn, void* e, void* stream) {
  AFQ_TRY(require(j, true, false));
  AFQ_TRY(check_walkers(j, up, dn));
  if (!e) return fail(AFQ_E_INVALID, "energies is NULL");
  return energy_pipeline(j, up, dn, static_cast<double2*>(e), S(stream));
}

int afq_apply_trotprop(afq_job_t j, void* up, void* dn, const void* xs, void* stream) {
  AFQ_TRY(require(j, false, true));
  AFQ_TRY(check_walkers(j, up, dn));
  if (!xs) return fail(AFQ_E_INVALID, "shifted_fields is NULL");
  return trotprop(j, up, dn, static_cast<const double2*>(xs), S(stream));
}

int afq_propagate_constrained(afq_job_t j, void* up, void* dn, const double* fields, double* weights, void* overlaps,
                              double* scalars, void* fb_out, void* stream) {
  AFQ_TRY(require(j, true, true));
  AFQ_TRY(check_walkers(j, up, dn));
  if (!fields || !weights || !overlaps || !scalars) return fail(AFQ_E_INVALID, "NULL argument");
  hipStream_t st = S(stream);
  AFQ_TRY(propagate(j, up, dn, fields, weights, static_cast<double2*>(overlaps), scalars, st));
  if (fb_out)
    AFQ_HIP(hipMemcpyAsync(fb_out, j->fb, sizeof(double2) * (long)j->W * j->G, hipMemcpyDeviceToDevice, st));
  return AFQ_OK;
}

int afq_orthonormalize(afq_job_t j, void* up, void* dn, void* stream) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  AFQ_TRY(check_walkers(j, up, dn));
  return orthonormalize(j, up, dn, S(stream));
}

int afq_stochastic_reconfiguration(afq_job_t j, void* up, void* dn, double* weights, const double* zeta,
                                   int32_t* idx_out, void* stream) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  AFQ_TRY(check_walkers(j, up, dn));
  if (!weights || !zeta) return fail(AFQ_E_INVALID, "NULL argument");
  hipStream_t st = S(stream);
  ProfScope ps(j, AFQ_K_SR, st);
  AFQ_TRY(k_sr_index(weights, zeta, j->W, j->cum, j->sr_idx, st));
  for (int s = 0; s < j->nspin_prop; ++s) {
    double2* Wp = static_cast<double2*>(s == 0 ? up : dn);
    const int c = j->restricted ? j->wcols : j->nocc[s];
    const long per = (long)j->N * c;
    AFQ_TRY(k_gather_walkers(Wp, j->wgather, j->sr_idx, per, j->W, st));
    AFQ_HIP(hipMemcpyAsync(Wp, j->wgather, sizeof(double2) * per * j->W, hipMemcpyDeviceToDevice, st));
  }
  if (idx_out) AFQ_HIP(hipMemcpyAsync(idx_out, j->sr_idx, sizeof(int32_t) * j->W, hipMemcpyDeviceToDevice, st));
  return AFQ_OK;
}

int afq_normal_fill(double* out, int64_t n, uint64_t seed, uint64_t counter, void* stream) {
  if (!out || n < 0) return fail(AFQ_E_INVALID, "bad argument");
  return k_normal_fill(out, n, seed, counter, S(stream));
}

int afq_sr_zeta(double* out, uint64_t seed, uint64_t counter, void* stream) {
  if (!out) return fail(AFQ_E_INVALID, "out is NULL");
  return k_sr_zeta_host(out, seed, counter, S(stream));
}

}  // extern "C"

namespace afq {

int reserve_fields(afq_job* j, long need, hipStream_t st) {
  if (need <= j->fields_cap) return AFQ_OK;
  AFQ_HIP(hipStreamSynchronize(st));
  if (j->fields) (void)hipFree(j->fields);
  j->fields = nullptr;
  j->fields_cap = 0;
  AFQ_TRY(dalloc_n(&j->fields, (size_t)need));
  j->fields_cap = need;
  return AFQ_OK;
}

// Steps over the fields already in j->fields, then QR, overlap, energy and the
// block bookkeeping (sampling.py:52-94).  The overlap matrices of the walkers a
// step produces are inverted together with their determinants, so the next
// step's force bias starts from that inverse instead of forming and factoring
// T^H W again; after the QR one factorisation serves both the overlap and the
// energy.
int block_body(afq_job* j, void* up, void* dn, double* weights, double2* ov, double* scalars, int n_prop_steps,
               double* block_out, hipStream_t st) {
  const long per_step = (long)j->W * j->G;
  for (int i = 0; i < n_prop_steps; ++i)
    AFQ_TRY(propagate(j, up, dn, j->fields + i * per_step, weights, ov, scalars, st, i > 0, i + 1 < n_prop_steps));
  AFQ_TRY(orthonormalize(j, up, dn, st));
  AFQ_TRY(green_pipeline(j, up, dn, true, true, st));
  AFQ_TRY(overlap_from_dets(j, ov, st));
  AFQ_TRY(energy_pipeline(j, up, dn, j->energies, st, true));
  ProfScope ps(j, AFQ_K_ENERGY_MISC, st);
  return k_block_finalize(j->energies, weights, scalars, std::sqrt(2.0 / j->d.dt), j->W, block_out, st);
}

// Comb resampling with a device-resident offset, then the overlaps of the new population.
int sr_and_overlap(afq_job* j, void* up, void* dn, double* weights, const double* zeta, double2* ov, hipStream_t st) {
  {
    ProfScope ps(j, AFQ_K_SR, st);
    AFQ_TRY(k_sr_index(weights, zeta, j->W, j->cum, j->sr_idx, st));
    for (int s = 0; s < j->nspin_prop; ++s) {
      double2* Wp = static_cast<double2*>(s == 0 ? up : dn);
      const int c = j->restricted ? j->wcols : j->nocc[s];
      const long per = (long)j->N * c;
      AFQ_TRY(k_gather_walkers(Wp, j->wgather, j->sr_idx, per, j->W, st));
      AFQ_HIP(hipMemcpyAsync(Wp, j->wgather, sizeof(double2) * per * j->W, hipMemcpyDeviceToDevice, st));
    }
  }
  AFQ_TRY(green_pipeline(j, up, dn, false, false, st));
  return overlap_from_dets(j, ov, st);
}

}  // namespace afq

extern "C" {

int afq_block(afq_job_t j, void* up, void* dn, double* weights, void* overlaps, double* scalars, int n_prop_steps,
              uint64_t seed, uint64_t* rng_counter, double* block_out, void* stream) {
  AFQ_TRY(require(j, true, true));
  AFQ_TRY(check_walkers(j, up, dn));
  if (!weights || !overlaps || !scalars || !rng_counter || !block_out) return fail(AFQ_E_INVALID, "NULL argument");
  if (n_prop_steps < 1) return fail(AFQ_E_INVALID, "n_prop_steps must be >= 1");
  hipStream_t st = S(stream);
  const long need = (long)j->W * j->G * n_prop_steps;
  AFQ_TRY(reserve_fields(j, need, st));
  {
    ProfScope ps(j, AFQ_K_RNG, st);
    AFQ_TRY(k_normal_fill(j->fields, need, seed, *rng_counter, st));
  }
  *rng_counter += 1;
  return block_body(j, up, dn, weights, static_cast<double2*>(overlaps), scalars, n_prop_steps, block_out, st);
}

int afq_job_reserve_fields(afq_job_t j, int32_t n_prop_steps, void* stream) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  if (n_prop_steps < 1) return fail(AFQ_E_INVALID, "n_prop_steps must be >= 1");
  return reserve_fields(j, (long)j->Wcap * j->G * n_prop_steps, S(stream));
}

int afq_sr_block(afq_job_t j, void* up, void* dn, double* weights, void* overlaps, double* scalars, int n_prop_steps,
                 uint64_t seed, uint64_t* counter_dev, double* block_out, void* stream) {
  AFQ_TRY(require(j, true, true));
  AFQ_TRY(check_walkers(j, up, dn));
  if (!weights || !overlaps || !scalars || !counter_dev || !block_out) return fail(AFQ_E_INVALID, "NULL argument");
  if (n_prop_steps < 1) return fail(AFQ_E_INVALID, "n_prop_steps must be >= 1");
  hipStream_t st = S(stream);
  const long need = (long)j->W * j->G * n_prop_steps;
  if (need > j->fields_cap)
    return fail(AFQ_E_STATE, "afq_sr_block: call afq_job_reserve_fields for this n_prop_steps first");
  {
    ProfScope ps(j, AFQ_K_RNG, st);
    AFQ_TRY(k_normal_fill_dev(j->fields, need, seed, counter_dev, st));
  }
  double2* ov = static_cast<double2*>(overlaps);
  AFQ_TRY(block_body(j, up, dn, weights, ov, scalars, n_prop_steps, block_out, st));
  double* zeta = j->dscratch + 4;
  AFQ_TRY(k_sr_zeta(zeta, seed, counter_dev, st));
  AFQ_TRY(sr_and_overlap(j, up, dn, weights, zeta, ov, st));
  return k_counter_incr(counter_dev, st);
}

int afq_job_profile(afq_job_t j, int cls, int enable) {
  if (!j) return fail(AFQ_E_INVALID, "job is NULL");
  if (cls < 0 || cls >= AFQ_K_NCLASSES) return fail(AFQ_E_INVALID, "unknown kernel class");
  j->prof[cls].on = enable != 0;
  return AFQ_OK;
}

int afq_job_profile_read(afq_job_t j, int cls, double* total_ms, int64_t* launches) {
  if (!j || !total_ms || !launches) return fail(AFQ_E_INVALID, "NULL argument");
  if (cls < 0 || cls >= AFQ_K_NCLASSES) return fail(AFQ_E_INVALID, "unknown kernel class");
  auto& p = j->prof[cls];
  double sum = 0.0;
  for (auto& sp : p.spans) {
    AFQ_HIP(hipEventSynchronize(sp.second));
    float ms = 0.f;
    AFQ_HIP(hipEventElapsedTime(&ms, sp.first, sp.second));
    sum += ms;
    p.pool.push_back(sp.first);
    p.pool.push_back(sp.second);
  }
  *total_ms = sum;
  *launches = (int64_t)p.spans.size();
  p.spans.clear();
  return AFQ_OK;
}

int afq_walkers_orthonormalize(void* walkers, int32_t n_walkers, int32_t norb, int32_t ncols, void* stream) {
  if (!walkers || n_walkers < 1 || norb < 1 || ncols < 1) return fail(AFQ_E_INVALID, "bad argument");
  return batched_cgs2(static_cast<double2*>(walkers), ncols, (long)norb * ncols, norb, ncols, n_walkers, S(stream));
}

int64_t afq_walkers_sr_workspace(int32_t n_walkers, int64_t max_walker_elems) {
  if (n_walkers < 1 || max_walker_elems < 1) return 0;
  const int64_t cum = ((int64_t)sizeof(double) * n_walkers + 255) / 256 * 256;
  const int64_t idx = ((int64_t)sizeof(int32_t) * n_walkers + 255) / 256 * 256;
  return cum + idx + (int64_t)sizeof(double2) * n_walkers * max_walker_elems;
}

int afq_walkers_sr(void* up, int64_t elems_up, void* dn, int64_t elems_dn, int32_t W, double* weights,
                   const double* zeta, int32_t* idx_out, void* workspace, void* stream) {
  if (!up || !weights || !zeta || !workspace || W < 1 || elems_up < 1 || (dn && elems_dn < 1))
    return fail(AFQ_E_INVALID, "bad argument");
  hipStream_t st = S(stream);
  char* ws = static_cast<char*>(workspace);
  double* cum = reinterpret_cast<double*>(ws);
  ws += ((int64_t)sizeof(double) * W + 255) / 256 * 256;
  int32_t* idx = reinterpret_cast<int32_t*>(ws);
  ws += ((int64_t)sizeof(int32_t) * W + 255) / 256 * 256;
  double2* gather = reinterpret_cast<double2*>(ws);
  AFQ_TRY(k_sr_index(weights, zeta, W, cum, idx, st));
  void* bufs[2] = {up, dn};
  const int64_t elems[2] = {elems_up, elems_dn};
  for (int s = 0; s < 2; ++s) {
    if (!bufs[s]) continue;
    double2* Wp = static_cast<double2*>(bufs[s]);
    AFQ_TRY(k_gather_walkers(Wp, gather, idx, elems[s], W, st));
    AFQ_HIP(hipMemcpyAsync(Wp, gather, sizeof(double2) * elems[s] * W, hipMemcpyDeviceToDevice, st));
  }
  if (idx_out) AFQ_HIP(hipMemcpyAsync(idx_out, idx, sizeof(int32_t) * W, hipMemcpyDeviceToDevice, st));
  return AFQ_OK;
}

}  // extern "C"
