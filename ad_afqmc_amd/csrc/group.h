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
GemmJob gemm_group_job(const Gemm& g);  // work items: tiles_m * tiles_n * g.batch
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
// splitk_reduce_grouped.  Work items: tiles_m * tiles_n * ksplit.
struct Bf3Job {
  Bf3 p;
  int tiles_m = 0, tiles_n = 0, ksplit = 1, kchunk = 0;
  double* ws = nullptr;
};
int bf3_group_job(const Bf3& p, Bf3Job* out);
int gemm_bf3_grouped(const GroupTab<Bf3Job>& t, hipStream_t s);

// C = alpha sum_kz ws[kz] + beta C for batch-1 complex outputs; work items:
// ceil(M N / 256).
struct ReduceJob {
  Gemm g;
  const double* ws;
  int ksplit;
};
int splitk_reduce_grouped(const GroupTab<ReduceJob>& t, hipStream_t s);

}  // namespace afq
