// Fused split-BF16 propagator (trot_fused.hip).
#pragma once
#include "afq_internal.h"

namespace afq {

// W <- E . Taylor(VHS) . E . W for every walker buffer, in place, in one launch.
struct TrotFused {
  int N = 0, W = 0, nspin = 1, nterms = 6;
  double2* wb[2] = {nullptr, nullptr};        // (W, N, c[s]) complex128
  int c[2] = {0, 0};
  const uint16_t* vs = nullptr;                // k_split_vhs_bf16 output
  const uint16_t* e1[2] = {nullptr, nullptr};  // k_split_e1_bf16 output of each spin's E
};
bool trot_fused_supported(int N);
// bf16 elements of one spin's split E - I, and of the split VHS of W walkers
long trot_fused_e1_elems(int N);
long trot_fused_vhs_elems(int N, int W);
// VHS (W, sV) complex64 (packed lower block triangle when sym) -> four bf16
// planes (re hi, re lo, im hi, im lo) of the full 16 ceil(N/16) x 32 ceil(N/32) matrix per walker
int k_split_vhs_bf16(const float2* vhs, long sV, bool sym, int N, int W, uint16_t* out, hipStream_t st);
// E (N x N FP64) -> E - I as three bf16 planes (hi, mid, lo) of 16 ceil(N/16) x 32 ceil(N/32), zero padded
int k_split_e1_bf16(const double* E, int N, uint16_t* out, hipStream_t st);
int trot_fused_bf3(const TrotFused& p, hipStream_t st);

}  // namespace afq
