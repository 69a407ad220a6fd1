// Fused split-BF16 propagator (trot_fused.hip).
#pragma once
#include "afq_internal.h"
#include "group.h"

namespace afq {

// W <- E . Taylor(VHS) . E . W for every walker buffer, in place, in one launch.
struct TrotFused {
  int N = 0, W = 0, nspin = 1, nterms = 6;
  double2* wb[2] = {nullptr, nullptr};        // (W, N, c[s]) complex128
  int c[2] = {0, 0};
  const uint16_t* vs = nullptr;                // VHS planes re hi, re lo, im hi, im lo per walker,
                                               // each 16 ceil(N/16) x 32 ceil(N/32), zero past N,
                                               // stored by 32-deep k step ([k/32][row][k%32])
  const uint16_t* e1[2] = {nullptr, nullptr};  // k_split_e1_bf16 output of each spin's E
};
// Kernel arguments of one job (bf16 pointers as __bf16).
struct FusedArgs {
  double2* wb[2];        // walker buffers (W, N, c[s]), updated in place
  int c[2];              // columns per spin buffer
  int nsl[2];            // column slabs per spin
  int N, R, K32, KPL, W;  // rows, 16-row blocks, k extent (32-multiple), LDS column length (128-multiple)
  const __bf16* vs;      // VHS planes re hi, re lo, im hi, im lo of (16 R) x K32 per walker
  long ep;               // plane stride of vs and e1 (elements)
  const __bf16* e1[2];   // E - I of each spin: planes hi, mid, lo of (16 R) x K32, zero padded
                         // (vs and e1 planes stored by 32-deep k step)
  int nterms;            // n_exp_terms: Taylor terms 1 .. nterms - 1
};

bool trot_fused_supported(int N);
// bf16 elements of one spin's split E - I, and of the split VHS of W walkers
long trot_fused_e1_elems(int N);
long trot_fused_vhs_elems(int N, int W);

// VHS (W walkers, complex64: the packed lower block triangle (sym_packed_cols
// columns) when sym, else N x N) -> the bf16 planes trot_fused_bf3 reads (entries
// past N are not written: the planes start zeroed).
int k_split_vhs_bf16(const float2* vhs, bool sym, int N, int W, uint16_t* out, hipStream_t st);
struct VhsSplit {
  const float2* vhs;
  bool sym;
  int N, W;
  uint16_t* out;
};
long vhs_split_items(const VhsSplit& v);
int split_vhs_grouped(const GroupTab<VhsSplit>& t, hipStream_t st);

// E (N x N FP64) -> E - I as three bf16 planes (hi, mid, lo) of 16 ceil(N/16) x 32 ceil(N/32), zero padded
int k_split_e1_bf16(const double* E, int N, uint16_t* out, hipStream_t st);
int trot_fused_bf3(const TrotFused& p, hipStream_t st);
// Grouped form: per-job arguments with tn column blocks per slab (the same
// for the whole group), trot_fused_items work-groups each (a multiple of 8);
// rbw = ceil(ceil(N_max / 16) / 8).
int trot_fused_args(const TrotFused& p, int tn, FusedArgs* out);
long trot_fused_items(const FusedArgs& a);
int trot_fused_grouped(const GroupTab<FusedArgs>& t, int tn, int rbw, hipStream_t st);
// Column blocks per slab (1..3) the single-job launch picks for these
// columns, and the row blocks per wave of its LDS planes for N rows.
int trot_fused_tn(const int* c, int nspin, int N);
int trot_fused_rbw(int N);

}  // namespace afq
