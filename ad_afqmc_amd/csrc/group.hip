// Grouped blocks: one phaseless block (afq_sr_block) for several jobs at once,
// every kernel of the block launched once over all jobs.
//
// A group is built from jobs that keep their buffers for its lifetime:
// restricted walkers of an RHF trial in split-BF16 x3 with the fused
// propagator.  At creation the per-job parameters of every launch of the block
// (batched products, Gauss-Jordan and Cholesky sets, split-BF16 products,
// propagator, exchange, and the small per-walker / per-job kernels) are built
// once on the host, with a prefix array over each launch's work items, and
// uploaded; a block then issues about 150 launches whatever the number of
// jobs, each a grouped kernel that finds its job by binary search over the
// prefix.  Each job's arithmetic is the single-job block's: the same
// operands, products and reductions in the same order, so a job's results do
// not depend on the other jobs of its group.  The block makes no allocation,
// copy or memset call, so a caller may capture it in a HIP graph.
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>

#include "group.h"
#include "kernels.h"
#include "pipeline.h"
#include "trot_fused.h"

namespace afq {
namespace {

// A device copy of a host parameter vector and the prefix of its work items.
template <class P>
struct DevTab {
  P* params = nullptr;
  int* prefix = nullptr;
  int njobs = 0, total = 0;
  GroupTab<P> view() const {
    GroupTab<P> t;
    t.params = params;
    t.prefix = prefix;
    t.njobs = njobs;
    t.total = total;
    return t;
  }
};

// All tables of a group in one device allocation: table() stages them in a
// host buffer (256-byte aligned), commit() allocates once, copies once and
// points the tables at their device copies.
struct DevMem {
  void* base = nullptr;
  std::vector<unsigned char> host;
  std::vector<std::function<void(unsigned char*)>> fixups;
  ~DevMem() {
    if (base) (void)hipFree(base);
  }
  template <class T>
  void stage(const std::vector<T>& v, T** out) {
    const size_t off = (host.size() + 255) & ~size_t(255);
    host.resize(off + sizeof(T) * std::max<size_t>(1, v.size()));
    if (!v.empty()) std::memcpy(host.data() + off, v.data(), sizeof(T) * v.size());
    fixups.push_back([out, off](unsigned char* b) { *out = reinterpret_cast<T*>(b + off); });
  }
  // params and the prefix of per-job item counts
  template <class P>
  int table(const std::vector<P>& params, const std::vector<long>& items, DevTab<P>* out) {
    std::vector<int> prefix(params.size() + 1, 0);
    for (size_t i = 0; i < params.size(); ++i) {
      const long next = (long)prefix[i] + items[i];
      if (next > 0x7fffffffL) return fail(AFQ_E_UNSUPPORTED, "group: too many work-groups in one launch");
      prefix[i + 1] = (int)next;
    }
    stage(params, &out->params);
    stage(prefix, &out->prefix);
    out->njobs = (int)params.size();
    out->total = prefix.back();
    return AFQ_OK;
  }
  int commit() {
    AFQ_TRY(dalloc(&base, std::max<size_t>(host.size(), 256)));
    AFQ_HIP(hipMemcpy(base, host.data(), host.size(), hipMemcpyHostToDevice));
    for (auto& f : fixups) f(static_cast<unsigned char*>(base));
    host = {};
    fixups = {};
    return AFQ_OK;
  }
};

}  // namespace
}  // namespace afq

struct afq_group {
  int njobs = 0, nsteps = 0, nmax = 0;
  afq::DevMem mem;
  afq::DevTab<afq::BlockJob> bj[afq::BI_NKINDS];
  afq::DevTab<afq::GemmJob> ovlp, green, gram, tri;
  afq::DevTab<afq::DetInvSet> di_inv, di_det, di_qr;
  afq::DevTab<afq::CholSet> chol;
  afq::DevTab<afq::VecJob> qr_conj, qr_copy, sr_copy;
  afq::DevTab<afq::SplitJob> gsplit, gsplit_fb;  // exchange and force-bias layouts
  afq::DevTab<afq::Bf3Job> fb, vhs;
  afq::DevTab<afq::ReduceJob> fb_red;
  afq::DevTab<afq::RowSplitJob> xsplit;
  std::map<int, afq::DevTab<afq::FusedArgs>> trot;  // by 8 x column blocks + row blocks per wave
  afq::DevTab<afq::VhsSplit> vsplit;
  std::map<int, afq::DevTab<afq::ExchJob>> exch;  // by ceil(O / 16)
};

namespace afq {
namespace {

int check_job(const afq_job* j, int nsteps) {
  AFQ_TRY(require(j, true, true));
  if (!j->restricted || !j->rhf)
    return fail(AFQ_E_UNSUPPORTED, "group: jobs with restricted walkers and an RHF trial only");
  if (!j->bf16x3 || !j->e1s[0] || !j->vsplit)
    return fail(AFQ_E_UNSUPPORTED, "group: split-BF16 x3 jobs with the fused propagator only (N <= 512)");
  if (j->nocc[0] > kExchangeRegMaxO)
    return fail(AFQ_E_UNSUPPORTED, "group: nocc <= 64 only (the register-blocked exchange)");
  if ((long)j->W * j->G * nsteps > j->fields_cap)
    return fail(AFQ_E_STATE, "group: call afq_job_reserve_fields for this n_prop_steps first");
  return AFQ_OK;
}

int build(afq_group* g, afq_job* const* jobs, void* const* walkers, double* const* weights, void* const* overlaps,
          double* const* scalars, uint64_t* const* counters, double* const* outs, const uint64_t* seeds) {
  const int n = g->njobs;
  std::vector<BlockJob> bjs(n);
  std::vector<GemmJob> ov(n), gr(n), gm(n), tr(n);
  std::vector<DetInvSet> dinv(n), ddet(n), dqr(n);
  std::vector<CholSet> ch(n);
  std::vector<VecJob> vconj(n), vcopy(n), vsr(n);
  std::vector<SplitJob> sp(n), sp_fb(n);
  std::vector<Bf3Job> fbj(n), vhj(n);
  std::vector<ReduceJob> red(n);
  std::vector<RowSplitJob> rsp(n);
  std::map<int, std::vector<FusedArgs>> fa;
  std::vector<VhsSplit> vsj(n);
  std::map<int, std::vector<ExchJob>> ex;
  std::vector<long> it_gemm_ov(n), it_gemm_gr(n), it_gemm_gm(n), it_gemm_tr(n), it_w(n), it_vec(n), it_sp(n),
      it_fb(n), it_vh(n), it_red(n), it_rsp(n), it_vs(n);
  std::map<int, std::vector<long>> it_ex, it_trot;
  std::vector<long> it_bj[BI_NKINDS];
  for (auto& v : it_bj) v.resize(n);

  for (int i = 0; i < n; ++i) {
    afq_job* j = jobs[i];
    const int N = j->N, O = j->nocc[0], G = j->G, W = j->W, c = j->wcols;
    const long NN = (long)N * N, per = (long)N * c, cc = (long)c * c;
    double2* up = static_cast<double2*>(walkers[i]);
    g->nmax = std::max(g->nmax, std::max(O, c));

    BlockJob& b = bjs[i];
    std::memset(&b, 0, sizeof(b));
    b.N = N;
    b.O = c;
    b.G = G;
    b.W = W;
    b.nsteps = g->nsteps;
    b.dt = j->d.dt;
    b.sqdt = std::sqrt(j->d.dt);
    b.h0 = j->h0;
    b.clip = std::sqrt(2.0 / j->d.dt);
    b.seed = seeds[i];
    b.walkers = up;
    b.weights = weights[i];
    b.overlaps = static_cast<double2*>(overlaps[i]);
    b.scalars = scalars[i];
    b.counter = counters[i];
    b.out = outs[i];
    b.fields = j->fields;
    b.fb = j->fb;
    b.mf = j->mf_shifts;
    b.xs = j->xs;
    b.terms = j->terms;
    b.det = j->det[0];
    b.ov_new = j->ov_new;
    b.h0_prop = j->h0_prop;
    b.green = j->green[0];
    b.rot_h1 = j->rot_h1[0];
    b.ene1 = j->ene1;
    b.coul = j->coul;
    b.kpart = j->kpart;
    b.nch = exchange_nchunk(G);
    b.energies = j->energies;
    b.cum = j->cum;
    b.idx = j->sr_idx;
    b.zeta = j->dscratch + 4;
    b.wgather = j->wgather;
    for (int k = 0; k < BI_NKINDS; ++k) it_bj[k][i] = block_items(b, k);

    // S = T^T W (overlap matrices), as green_pipeline
    Gemm o;
    o.M = O;
    o.N = O;
    o.K = N;
    o.batch = W;
    o.a_trans = true;
    o.A = j->mo[0];
    o.lda = O;
    o.sA = 0;
    o.b_cplx = true;
    o.B = up;
    o.ldb = c;
    o.sB = per;
    o.C = j->S[0];
    o.ldc = O;
    o.sC = (long)O * O;
    ov[i] = gemm_group_job(o);
    it_gemm_ov[i] = gemm_group_items(ov[i]);
    // G = (W S^-1)^T
    Gemm h;
    h.M = O;
    h.N = N;
    h.K = O;
    h.batch = W;
    h.a_cplx = true;
    h.a_trans = true;
    h.A = j->Sinv[0];
    h.lda = O;
    h.sA = (long)O * O;
    h.b_cplx = true;
    h.b_trans = true;
    h.B = up;
    h.ldb = c;
    h.sB = per;
    h.C = j->green[0];
    h.ldc = N;
    h.sC = (long)O * N;
    gr[i] = gemm_group_job(h);
    it_gemm_gr[i] = gemm_group_items(gr[i]);
    dinv[i] = DetInvSet{j->S[0], (long)O * O, O, j->det[0], j->Sinv[0], (long)O * O};
    ddet[i] = DetInvSet{j->S[0], (long)O * O, O, j->det[0], nullptr, (long)O * O, j->S[0]};
    it_w[i] = W;

    // Cholesky QR (cholqr2_panel): tmp = conj(W); S = tmp^T W; S = R^H R; R^-1; tmp = W R^-1; W = tmp
    vconj[i] = VecJob{up, j->wgather, per * W, 1};
    vcopy[i] = VecJob{j->wgather, up, per * W, 0};
    vsr[i] = VecJob{j->wgather, up, per * W, 0};
    it_vec[i] = (per * W + 1023) / 1024;
    Gemm q;
    q.M = c;
    q.N = c;
    q.K = N;
    q.batch = W;
    q.a_cplx = true;
    q.a_trans = true;
    q.A = j->wgather;
    q.lda = c;
    q.sA = per;
    q.b_cplx = true;
    q.B = up;
    q.ldb = c;
    q.sB = per;
    q.C = j->S[0];
    q.ldc = c;
    q.sC = cc;
    gm[i] = gemm_group_job(q);
    it_gemm_gm[i] = gemm_group_items(gm[i]);
    ch[i] = CholSet{j->S[0], cc, c, j->S[0], cc};
    dqr[i] = DetInvSet{j->S[0], cc, c, j->det[0], j->Sinv[0], cc};
    Gemm t;
    t.M = N;
    t.N = c;
    t.K = c;
    t.batch = W;
    t.a_cplx = true;
    t.A = up;
    t.lda = c;
    t.sA = per;
    t.b_cplx = true;
    t.B = j->Sinv[0];
    t.ldb = c;
    t.sB = cc;
    t.C = j->wgather;
    t.ldc = c;
    t.sC = per;
    tr[i] = gemm_group_job(t);
    it_gemm_tr[i] = gemm_group_items(tr[i]);

    // force bias (force_bias_gemm, split-BF16 x3): split G, product with the split rot_chol
    const long P = (long)bf16_opad(O) * bf16_npad(N);
    sp[i] = SplitJob{reinterpret_cast<const double*>(j->green[0]), 1, O, N, (long)W, j->gsplit, 0};
    sp_fb[i] = sp[i];
    sp_fb[i].ilv = 1;
    it_sp[i] = split_items(sp[i]);
    Bf3 f;
    f.M = W;
    f.N = G;
    f.K = (int)P;
    f.a_cplx = true;
    f.A = j->gsplit;
    f.lda = 4 * P;
    f.pA = 32;
    f.kA = 4 * 32;
    f.a_rows = j->Wcap;
    f.B = j->rcs_fb[0];
    f.ldb = 2 * P;
    f.pB = 32;
    f.kB = 2 * 32;
    f.b_rows = G;
    f.alpha = double2{2.0, 0.0};
    f.beta = double2{0.0, 0.0};
    f.C = j->fb;
    f.ldc = G;
    f.c_f64 = true;
    f.ws = j->gemm_ws;
    f.ws_elems = j->gemm_ws_elems;
    AFQ_TRY(bf3_group_job(f, &fbj[i]));
    it_fb[i] = bf3_group_items(fbj[i]);
    red[i].g.M = W;
    red[i].g.N = G;
    red[i].g.C = j->fb;
    red[i].g.ldc = G;
    red[i].g.alpha = f.alpha;
    red[i].g.beta = f.beta;
    red[i].ws = fbj[i].ws;
    red[i].ksplit = fbj[i].ksplit;
    it_red[i] = ((long)W * G + 255) / 256;  // the complex128 output always passes through ws

    // VHS (trotprop_lp): split the shifted fields, product with the split L^T into the planes
    const int gpad = bf16_npad(G);
    rsp[i] = RowSplitJob{reinterpret_cast<const double*>(j->xs), 1, (long)W, G, (long)G, gpad, 4L * gpad, 32L,
                         4L * 32, j->xs_split};
    it_rsp[i] = row_split_items(rsp[i]);
    const long ncols = j->sym_chol ? j->npk : NN;
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
    v.alpha = double2{0.0, std::sqrt(j->d.dt)};
    v.C = j->sym_chol ? static_cast<void*>(j->vhsp32) : static_cast<void*>(j->vhs32);
    v.ldc = ncols;
    AFQ_TRY(bf3_group_job(v, &vhj[i]));
    it_vh[i] = bf3_group_items(vhj[i]);
    vsj[i] = VhsSplit{static_cast<const float2*>(v.C), j->sym_chol, N, W, j->vsplit};
    it_vs[i] = vhs_split_items(vsj[i]);

    TrotFused tf;
    tf.N = N;
    tf.W = W;
    tf.nspin = 1;
    tf.nterms = j->d.n_exp_terms;
    tf.wb[0] = up;
    tf.c[0] = c;
    tf.e1[0] = j->e1s[0];
    tf.vs = j->vsplit;
    // fused propagator, one launch per (column blocks, row blocks per wave)
    // class: each job runs the configuration its single-job launch picks
    const int tn = trot_fused_tn(tf.c, 1, N), tkey = 8 * tn + trot_fused_rbw(N);
    FusedArgs a;
    AFQ_TRY(trot_fused_args(tf, tn, &a));
    fa[tkey].push_back(a);
    it_trot[tkey].push_back(trot_fused_items(a));

    // exchange, grouped by the 16-row block count of the f-blocks
    const int nb = (O + 15) / 16;
    ExchJob e{j->rcs[0], j->gsplit, O, bf16_opad(O), bf16_npad(N), G, W, exchange_nchunk(G), j->coul, j->kpart};
    it_ex[nb].push_back(exchange_group_items(&e));
    ex[nb].push_back(e);
  }
  DevMem& m = g->mem;
  for (int k = 0; k < BI_NKINDS; ++k) AFQ_TRY(m.table(bjs, it_bj[k], &g->bj[k]));
  AFQ_TRY(m.table(ov, it_gemm_ov, &g->ovlp));
  AFQ_TRY(m.table(gr, it_gemm_gr, &g->green));
  AFQ_TRY(m.table(gm, it_gemm_gm, &g->gram));
  AFQ_TRY(m.table(tr, it_gemm_tr, &g->tri));
  AFQ_TRY(m.table(dinv, it_w, &g->di_inv));
  AFQ_TRY(m.table(ddet, it_w, &g->di_det));
  AFQ_TRY(m.table(dqr, it_w, &g->di_qr));
  AFQ_TRY(m.table(ch, it_w, &g->chol));
  AFQ_TRY(m.table(vconj, it_vec, &g->qr_conj));
  AFQ_TRY(m.table(vcopy, it_vec, &g->qr_copy));
  AFQ_TRY(m.table(vsr, it_vec, &g->sr_copy));
  AFQ_TRY(m.table(sp, it_sp, &g->gsplit));
  AFQ_TRY(m.table(sp_fb, it_sp, &g->gsplit_fb));
  AFQ_TRY(m.table(fbj, it_fb, &g->fb));
  AFQ_TRY(m.table(red, it_red, &g->fb_red));
  AFQ_TRY(m.table(rsp, it_rsp, &g->xsplit));
  AFQ_TRY(m.table(vhj, it_vh, &g->vhs));
  AFQ_TRY(m.table(vsj, it_vs, &g->vsplit));
  for (auto& kv : fa) AFQ_TRY(m.table(kv.second, it_trot[kv.first], &g->trot[kv.first]));
  for (auto& kv : ex) AFQ_TRY(m.table(kv.second, it_ex[kv.first], &g->exch[kv.first]));
  return m.commit();
}

// green_pipeline over the group: S and its determinant (inverse when want_inv),
// then G when want_green; have_inv: the inverses are already those of the walkers.
int g_green(afq_group* g, bool want_inv, bool want_green, bool have_inv, hipStream_t st) {
  if (!have_inv) {
    AFQ_TRY(gemm_f64_grouped(g->ovlp.view(), false, true, true, false, st));
    AFQ_TRY(det_inv_grouped(want_inv ? g->di_inv.view() : g->di_det.view(), g->nmax, st));
  }
  if (want_green) AFQ_TRY(gemm_f64_grouped(g->green.view(), true, true, true, true, st));
  return AFQ_OK;
}

int g_block(afq_group* g, hipStream_t st) {
  AFQ_TRY(g_rng(g->bj[BI_RNG].view(), st));
  for (int i = 0; i < g->nsteps; ++i) {
    // propagate(inv_ready = i > 0, keep_inv = i + 1 < nsteps)
    AFQ_TRY(g_green(g, true, true, i > 0, st));
    AFQ_TRY(split_bf16_grouped(g->gsplit_fb.view(), st));
    AFQ_TRY(gemm_bf3_grouped(g->fb.view(), st));
    AFQ_TRY(reduce_f32_grouped(g->fb_red.view(), st));
    AFQ_TRY(g_field_shift(g->bj[BI_WALKER].view(), i, st));
    AFQ_TRY(split_rows_grouped(g->xsplit.view(), st));
    AFQ_TRY(gemm_bf3_grouped(g->vhs.view(), st));
    AFQ_TRY(split_vhs_grouped(g->vsplit.view(), st));
    for (auto& kv : g->trot) AFQ_TRY(trot_fused_grouped(kv.second.view(), kv.first / 8, kv.first % 8, st));
    AFQ_TRY(g_green(g, i + 1 < g->nsteps, false, false, st));
    AFQ_TRY(g_overlap(g->bj[BI_WALKER256].view(), false, st));
    AFQ_TRY(g_weight_update(g->bj[BI_JOB].view(), st));
  }
  // Cholesky QR, twice
  for (int pass = 0; pass < 2; ++pass) {
    AFQ_TRY(vec_grouped(g->qr_conj.view(), st));
    AFQ_TRY(gemm_f64_grouped(g->gram.view(), true, true, true, false, st));
    AFQ_TRY(cholesky_grouped(g->chol.view(), g->nmax, st));
    AFQ_TRY(det_inv_grouped(g->di_qr.view(), g->nmax, st));
    AFQ_TRY(gemm_f64_grouped(g->tri.view(), true, true, false, false, st));
    AFQ_TRY(vec_grouped(g->qr_copy.view(), st));
  }
  AFQ_TRY(g_green(g, true, true, false, st));
  AFQ_TRY(g_overlap(g->bj[BI_WALKER256].view(), true, st));
  // local energy
  AFQ_TRY(g_ene1(g->bj[BI_WALKER].view(), st));
  AFQ_TRY(split_bf16_grouped(g->gsplit.view(), st));
  for (auto& kv : g->exch) AFQ_TRY(exchange_bf16_grouped(kv.second.view(), kv.first, st));
  AFQ_TRY(g_energy_finalize(g->bj[BI_WALKER].view(), st));
  AFQ_TRY(g_block_finalize(g->bj[BI_JOB].view(), st));
  // comb resampling and the overlaps of the new population
  AFQ_TRY(g_sr_index(g->bj[BI_JOB].view(), st));
  AFQ_TRY(g_gather(g->bj[BI_GATHER].view(), st));
  AFQ_TRY(vec_grouped(g->sr_copy.view(), st));
  AFQ_TRY(g_green(g, false, false, false, st));
  AFQ_TRY(g_overlap(g->bj[BI_WALKER256].view(), true, st));
  return g_block_tail(g->bj[BI_JOB].view(), st);
}

}  // namespace
}  // namespace afq

extern "C" {

int afq_group_create(afq_group_t* out, const afq_job_t* jobs, int32_t n_jobs, void* const* walkers,
                     double* const* weights, void* const* overlaps, double* const* scalars,
                     uint64_t* const* rng_counters, double* const* block_out, const uint64_t* seeds,
                     int32_t n_prop_steps) {
  using namespace afq;
  if (!out || !jobs || !walkers || !weights || !overlaps || !scalars || !rng_counters || !block_out || !seeds)
    return fail(AFQ_E_INVALID, "NULL argument");
  *out = nullptr;
  if (n_jobs < 1) return fail(AFQ_E_INVALID, "n_jobs must be >= 1");
  if (n_prop_steps < 1) return fail(AFQ_E_INVALID, "n_prop_steps must be >= 1");
  for (int i = 0; i < n_jobs; ++i) {
    if (!jobs[i] || !walkers[i] || !weights[i] || !overlaps[i] || !scalars[i] || !rng_counters[i] || !block_out[i])
      return fail(AFQ_E_INVALID, "NULL job, buffer or counter of job " + std::to_string(i));
    for (int k = 0; k < i; ++k)
      if (jobs[k] == jobs[i]) return fail(AFQ_E_INVALID, "a job appears twice in the group");
    AFQ_TRY(check_job(jobs[i], n_prop_steps));
  }
  std::unique_ptr<afq_group> g(new afq_group());
  g->njobs = n_jobs;
  g->nsteps = n_prop_steps;
  AFQ_TRY(build(g.get(), jobs, walkers, weights, overlaps, scalars, rng_counters, block_out, seeds));
  *out = g.release();
  return AFQ_OK;
}

int afq_group_sr_block(afq_group_t g, void* stream) {
  using namespace afq;
  if (!g) return fail(AFQ_E_INVALID, "group is NULL");
  return g_block(g, reinterpret_cast<hipStream_t>(stream));
}

int afq_group_destroy(afq_group_t g) {
  delete g;
  return AFQ_OK;
}

}  // extern "C"
