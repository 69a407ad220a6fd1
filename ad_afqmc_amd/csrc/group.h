// Grouped launches: one kernel launch covers the work of several jobs.
//
// A grouped kernel receives a device array of per-job parameter structs and a
// prefix array over the jobs' work items (work-groups, walkers, matrices):
// job j owns the flat items [prefix[j], prefix[j + 1]).  A work-group finds
// its job by binary search over prefix (the same for every thread, so it is
// wave-uniform) and runs the single-job body on that job's parameters with
// its local item index.
#pragma once
#include "afq_internal.h"
#include "gemm_bf3.h"

namespace afq {

// Device-resident parameter table of one grouped launch.
template <class P>
struct GroupTab {
  const P* params = nullptr;  // njobs entries
  const int* prefix = nullptr;  // njobs + 1 entries, prefix[0] = 0
  int njobs = 0;
  int total = 0;  // prefix[njobs]
};

__device__ inline int group_find(const int* __restrict__ prefix, int njobs, int item) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ---- grouped kernels (each declared next to its single-job form)

// Batched products on 64 x 64 tiles (2 x 2 waves of 2 x 2 MFMA blocks).
struct GemmJob {
  Gemm g;
  int tiles_m = 0, tiles_n = 0;
};
GemmJob gemm_group_job(const Gemm& g);
// work items: tiles_m * tiles_n * (g.batch rounded up to 8); a multiple of 8,
// so every job's items start on XCD 0 and the XCD order below holds per job
long gemm_group_items(const GemmJob& j);
int gemm_f64_grouped(const GroupTab<GemmJob>& t, bool a_cplx, bool b_cplx, bool a_trans, bool b_trans,
                     hipStream_t s);

// Determinant (and inverse) of each job's set of matrices (Gauss-Jordan, one
// work-group per matrix); work items: the set's matrix count.  nmax: the
// largest n.
int det_inv_grouped(const GroupTab<DetInvSet>& t, int nmax, hipStream_t s);

// Upper Cholesky factors of each job's set of Hermitian positive-definite
// matrices (n <= 64), one work-group per matrix; work items: matrices.
struct CholSet {
  const double2* S;
  long strideS;
  int n;
  double2* R;
  long strideR;
};
int cholesky_grouped(const GroupTab<CholSet>& t, int nmax, hipStream_t s);
// One or two sets of `count` matrices in one launch (both spins of a QR pass).
int batched_cholesky_upper_sets(const CholSet* sets, int nsets, int count, hipStream_t s);

// dst = src or conj(src) over n complex elements per job; work items:
// ceil(n / 1024) per job.
struct VecJob {
  const double2* src;
  double2* dst;
  long n;
  int conj;
};
int vec_grouped(const GroupTab<VecJob>& t, hipStream_t s);

// Split-BF16 x3 products of batch 1 (gemm_bf3's 128 x 128 tiles, 2 x 4
// waves); a complex128 output takes K slices into ws, reduced afterwards by
// reduce_f32_grouped.
struct Bf3Job {
  Bf3 p;
  int tiles_m = 0, tiles_n = 0, ksplit = 1, kchunk = 0;
  int tx = 0;    // work items of an unsplit product (tiles, rounded up to 8 when remapped)
  long T8 = 0;   // > 0: gemm_bf3's XCD-aware tile order (the M-tiles of an L panel on one XCD)
  float* ws = nullptr;  // split-K partials (complex64); null: unsplit
};
int bf3_group_job(const Bf3& p, Bf3Job* out);
// work items: gemm_bf3's split-K slice order, or tx rounded up to 8 (every job
// starts on XCD 0)
long bf3_group_items(const Bf3Job& j);
int gemm_bf3_grouped(const GroupTab<Bf3Job>& t, hipStream_t s);

// C = alpha sum_kz ws[kz] + beta C for batch-1 complex128 outputs from the
// complex64 K-slice partials of gemm_bf3; work items: ceil(M N / 256).
struct ReduceJob {
  Gemm g;
  const float* ws;
  int ksplit;
};
int reduce_f32_grouped(const GroupTab<ReduceJob>& t, hipStream_t s);

// Split-BF16 x3 exchange of each job (k_exchange_bf16's operands: rot_chol
// planes rcs, Green's-function planes gs); work items: W * nchunk, all jobs of
// one launch with ceil(O / 16) = nb.
struct ExchJob {
  const uint16_t* rcs;
  const uint16_t* gs;
  int O, Opad, Npad, Gn, W, nchunk;
  double2* coul;
  double2* kpart;
  int cpg = 1;  // chunks per gamma group (k_exchange_bf16's XCD-aware order)
  long T8 = 0;  // work items / 8
};
// fills cpg and T8 of e (its other fields set); returns its work items (8 T8)
long exchange_group_items(ExchJob* e);
int exchange_bf16_grouped(const GroupTab<ExchJob>& t, int nb, hipStream_t s);

// k_split_bf16 of each job's rows x (O, N) operand; work items: split_items.
struct SplitJob {
  const double* src;
  int cplx, O, N;
  long rows;
  uint16_t* out;
  int ilv;  // planes interleaved by 32-deep k step (the force-bias product's layout)
  long rst = 0;  // elements between rows of out (0: the row's own size)
};
long split_items(const SplitJob& j);
int split_bf16_grouped(const GroupTab<SplitJob>& t, hipStream_t s);

// k_split_rows_bf16 of each job's rows; work items: row_split_items.
struct RowSplitJob {
  const double* src;
  int cplx;
  long rows;
  int K;
  long lds;
  int kpad;
  long ldd, pd, kst;
  uint16_t* dst;
};
long row_split_items(const RowSplitJob& j);
int split_rows_grouped(const GroupTab<RowSplitJob>& t, hipStream_t s);

// Everything the small per-walker and per-job kernels of a phaseless block
// read or write for one job (restricted walkers of an RHF trial).
struct BlockJob {
  int N, O, G, W;  // orbitals, walker columns, Cholesky vectors, walkers
  int nsteps;      // phaseless steps per block
  double dt, sqdt, h0, clip;
  uint64_t seed;
  // caller state
  double2* walkers;
  double* weights;
  double2* overlaps;
  double* scalars;  // e_estimate, pop_control_ene_shift
  uint64_t* counter;
  double* out;  // block energy, block weight
  // job buffers
  double* fields;  // (nsteps, W, G)
  const double2* fb;
  const double2* mf;
  double2* xs;
  double2* terms;
  const double2* det;
  double2* ov_new;
  const double2* h0_prop;
  const double2* green;
  const double* rot_h1;
  double2* ene1;
  const double2* coul;
  const double2* kpart;
  int nch;
  double2* energies;
  double* cum;
  int32_t* idx;
  double* zeta;
  double2* wgather;
};
// Work items of the grouped block kernels: WALKER = one per walker,
// WALKER256 = one per 256 walkers, JOB = one per job, RNG = one per 256 field
// pairs, GATHER = one per 1024 elements of each walker.
enum BlockItems { BI_WALKER, BI_WALKER256, BI_JOB, BI_RNG, BI_GATHER, BI_NKINDS };
long block_items(const BlockJob& j, int kind);
// The block's small kernels over all jobs; `t` carries the prefix of the
// kernel's item kind.
int g_rng(const GroupTab<BlockJob>& t, hipStream_t s);                 // RNG: fields of the block
int g_field_shift(const GroupTab<BlockJob>& t, int step, hipStream_t s);  // WALKER
int g_overlap(const GroupTab<BlockJob>& t, bool to_overlaps, hipStream_t s);  // WALKER256: det^2 -> ov_new | overlaps
int g_weight_update(const GroupTab<BlockJob>& t, hipStream_t s);       // JOB
int g_ene1(const GroupTab<BlockJob>& t, hipStream_t s);                // WALKER
int g_energy_finalize(const GroupTab<BlockJob>& t, hipStream_t s);     // WALKER
int g_block_finalize(const GroupTab<BlockJob>& t, hipStream_t s);      // JOB
int g_sr_index(const GroupTab<BlockJob>& t, hipStream_t s);            // JOB: zeta of the block counter, comb indices
int g_gather(const GroupTab<BlockJob>& t, hipStream_t s);              // GATHER: walkers[idx] -> wgather
int g_block_tail(const GroupTab<BlockJob>& t, hipStream_t s);          // JOB: counter + 1, e_estimate update

}  // namespace afq
