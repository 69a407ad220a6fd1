// Split-BF16 x3 products on pre-split operands (gemm_bf3.hip).
#pragma once
#include "afq_internal.h"

namespace afq {

// C[b] (M x N) = alpha * A[b] (M x K, complex) . B[b]^T (N x K, real) on bf16
// (hi, lo) planes, into complex64 C, or (c_f64, batch 1) into complex128
// C = alpha * product + beta * C through split-K partials in ws.  Offsets and
// strides count bf16 elements (operands) or output elements.  Element (row r,
// plane q, k) of A is at A + b*sA + r*lda + (k/32)*kA + q*pA + k%32: planes
// stored apart (kA = 32, pA = plane size) or interleaved by k step (pA = 32,
// kA = 32 x planes, one step of a row then fills whole 128-byte lines).
struct Bf3 {
  int M = 0, N = 0, K = 0, batch = 1;  // K: padded depth, a multiple of 32
  bool a_cplx = false, b_cplx = false;
  const void* A = nullptr;
  long lda = 0, pA = 0, sA = 0;  // row, plane and batch strides
  long kA = 32;                  // from one 32-deep k step of a row to the next
  int a_rows = 0;                // rows that may be read (>= M rounded up to 16, or clamped)
  const void* B = nullptr;
  long ldb = 0, pB = 0, sB = 0;
  long kB = 32;
  int b_rows = 0;
  double2 alpha = {1.0, 0.0};
  double2 beta = {0.0, 0.0};  // complex128 output only
  void* C = nullptr;
  long ldc = 0, sC = 0;
  bool c_f64 = false;
  void* ws = nullptr;  // doubles
  long ws_elems = 0;
};
int gemm_bf3(const Bf3& p, hipStream_t s);

// Row split: dst[r*ldd + (k/32)*kst + q*pd + k%32] for k < kpad from src rows
// of K doubles (complex when cplx: four planes, else two), zero for k >= K.
int k_split_rows_bf16(const double* src, bool cplx, long rows, int K, long lds, int kpad, long ldd, long pd, long kst,
                      uint16_t* dst, hipStream_t s);
// Cholesky vectors L (G x NN) -> dst[n*ldd + (g/32)*kst + q*pd + g%32] for
// g < gpad (zero past G) and rows n < ncols: column n of L, or with sym_N > 0
// packed column n of a symmetric L (sym_packed_cols layout).
int k_split_chol_t_bf16(const double* L, int G, long NN, long ncols, int sym_N, int gpad, long ldd, long pd, long kst,
                        uint16_t* dst, hipStream_t s);

}  // namespace afq
