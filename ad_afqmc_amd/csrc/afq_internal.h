// Internal declarations of the MI355X AFQMC engine (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "afq.h"

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace afq {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define AFQ_HIP(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::afq::fail(AFQ_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define AFQ_TRY(expr)            \
  do {                           \
    int rc_ = (expr);            \
    if (rc_ != AFQ_OK) return rc_; \
  } while (0)

inline int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(AFQ_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return AFQ_OK;
}

// ---------------------------------------------------------------- GEMM
// Batched C[b] = alpha * opA(A[b]) * opB(B[b]) + beta * C[b]   (and D[b] += result
// when D != nullptr).  Operands are real (double) or interleaved complex
// (double2); C (and D) are complex when either operand is.  Row-major:
//   A element (m,k) at A[m*lda+k]        (a_trans: A[k*lda+m])
//   B element (k,n) at B[k*ldb+n]        (b_trans: B[n*ldb+k])
//   C element (m,n) at C[m*ldc+n]
// Leading dimensions and batch strides count elements (real or complex).
// Symmetric Cholesky vectors, block-packed lower triangle.
// Column p of the packed layout is block b = p / 256, entry (r, c) = ((p % 256) / 16,
// p % 16) of the 16 x 16 orbital block (ib, jb), ib >= jb, b = ib (ib + 1) / 2 + jb.
__device__ inline void sym_block(int b, int* ib, int* jb) {
  int i = (int)((sqrtf(8.0f * b + 1.0f) - 1.0f) * 0.5f);
  while ((i + 1) * (i + 2) / 2 <= b) ++i;
  while (i * (i + 1) / 2 > b) --i;
  *ib = i;
  *jb = b - i * (i + 1) / 2;
}
// source column i*N + j of packed column p, or -1 for padding
__device__ inline long sym_src(long p, int N) {
  int ib, jb;
  sym_block((int)(p >> 8), &ib, &jb);
  const int i = 16 * ib + (int)((p >> 4) & 15), j = 16 * jb + (int)(p & 15);
  return (i < N && j < N) ? (long)i * N + j : -1;
}

// Offset of entry (i, k) of a symmetric matrix stored as its packed lower
// block triangle (entries of upper blocks are read from the mirrored block).
__device__ inline long sym_off(int i, int k) {
  const int ib = i >> 4, kb = k >> 4;
  return ib >= kb ? ((long)(ib * (ib + 1) / 2 + kb) << 8) + ((i & 15) << 4) + (k & 15)
                  : ((long)(kb * (kb + 1) / 2 + ib) << 8) + ((k & 15) << 4) + (i & 15);
}

struct Gemm {
  int M = 0, N = 0, K = 0, batch = 1;
  bool a_cplx = false, b_cplx = false, a_trans = false, b_trans = false;
  bool a_conj = false;  // complex A taken conjugated (op(A) = conj(A) or A^H)
  const void* A = nullptr;
  long lda = 0, sA = 0;
  // a_sym > 0: A is a symmetric a_sym x a_sym matrix held as its packed lower
  // block triangle (sym_packed_cols layout, lda unused, batch stride sA); !a_trans
  int a_sym = 0;
  const void* B = nullptr;
  long ldb = 0, sB = 0;
  void* C = nullptr;
  long ldc = 0, sC = 0;
  void* D = nullptr;
  long ldd = 0, sD = 0;
  // reduced-precision products only: with E set, D = E + alpha*A.B (E read,
  // complex double) instead of D += alpha*A.B; C may then be null
  const void* E = nullptr;
  long lde = 0, sE = 0;
  double2 alpha = {1.0, 0.0};
  double2 beta = {0.0, 0.0};
  // optional split-K workspace (doubles); used only for unbatched products
  // with too few output tiles to fill the GPU
  void* ws = nullptr;
  long ws_elems = 0;
  // complex A and B with real C = Re(alpha A.B) + beta C, computing only the
  // part of A.B that alpha selects: 1 = real part (alpha real), 2 = imaginary
  // part (alpha imaginary); ldc and sC then count doubles
  int c_part = 0;
};
int gemm_f64(const Gemm& g, hipStream_t stream);
// Two batched products in one launch when their operand form and tile
// configuration match (the overlap, Green's-function and QR products of both
// spins); otherwise two gemm_f64 calls.  Same results as the separate calls.
int gemm_f64_pair(const Gemm& g0, const Gemm& g1, hipStream_t stream);
// C = alpha * sum of `ksplit` split-K partial planes in ws + beta * C (+ D), fixed order.
int gemm_splitk_reduce(const Gemm& g, const double* ws, int ksplit, bool cplx, hipStream_t stream);

// Reduced-precision products of the walker step (FP32, or split-BF16 x3), FP32 accumulation:
//   VHS        A complex double (W x G) . B float (G x N^2)          -> C complex float
//   TAYLOR_D   A complex float (N x N) . B complex double (N x c)    -> C complex float, D complex double +=
//   TAYLOR_F   A complex float (N x N) . B complex float (N x c)     -> C complex float, D complex double +=
//   FORCE_BIAS A complex double (W x ON) . B float^T (G x ON)        -> C complex double
//   ONE_BODY   A float (N x N) . B complex double (N x c)            -> D = E + A.B (FP32 MFMA always)
enum class LpKind { VHS, TAYLOR_D, TAYLOR_F, FORCE_BIAS, ONE_BODY };
int gemm_lp(LpKind kind, bool bf16x3, const Gemm& g, hipStream_t stream);

// ---------------------------------------------------------------- small dense linear algebra
// In-place Gauss-Jordan with partial pivoting (|re|+|im| pivot metric, first
// maximum wins, as LAPACK's izamax) on `count` n x n complex matrices.  Writes
// det (count,) and, if inv != nullptr, the inverse.  n <= 256: up to n = 64
// the matrix is factored in LDS, above in global memory (in inv itself).
int batched_det_inv(const double2* S, long strideS, int n, int count, double2* det, double2* inv, long strideInv,
                    hipStream_t stream);
// One launch over one or two sets of `count` matrices each (e.g. both spins).
struct DetInvSet {
  const double2* S;
  long strideS;
  int n;
  double2* det;
  double2* inv;  // nullptr: determinant only
  long strideInv;
  // n > 64 without an inverse: global scratch of n x n per matrix, matrix
  // stride strideS (S itself may be given: it is then overwritten)
  double2* work = nullptr;
};
constexpr int kDetInvMaxN = 256;
int batched_det_inv_sets(const DetInvSet* sets, int nsets, int count, hipStream_t stream);

// Upper Cholesky factor R (S = R^H R, positive real diagonal, zero strict lower
// triangle) of `count` Hermitian positive-definite n x n matrices, n <= 256
// (above n = 64 factored in place in R; R may be S).
int batched_cholesky_upper(const double2* S, long strideS, int n, int count, double2* R, long strideR,
                           hipStream_t stream);

// Classical Gram-Schmidt with re-orthogonalisation (CGS2) on `count` row-major
// n_rows x n_cols complex panels (leading dimension ld, batch stride sW), in place.
int batched_cgs2(double2* Wp, long ld, long sW, int n_rows, int n_cols, int count, hipStream_t stream);

}  // namespace afq

// ---------------------------------------------------------------- job
struct afq_job {
  afq_job_desc d{};
  int N = 0, G = 0, W = 0;  // W: active walkers (<= Wcap)
  int Wcap = 0;
  int nocc[2] = {0, 0};
  int wcols = 0;             // columns of the restricted walker buffer (nelec_up)
  bool restricted = true;    // restricted walkers (one buffer)
  bool rhf = true;           // RHF trial (mo_dn == mo_up)
  bool one_spin = true;      // spin-dn quantities equal spin-up (RHF trial and restricted walkers)
  int nspin_meas = 1;        // spin channels measured (1 when one_spin)
  int nspin_prop = 1;        // walker buffers propagated

  double h0 = 0.0;
  bool have_ham = false, have_trial = false, have_meas = false, have_prop = false;
  // per-item flags for intermediates supplied through afq_set_intermediate
  bool have_rot_h1[2] = {false, false}, have_rot_chol[2] = {false, false};
  bool have_mf = false, have_h0p = false, have_exp[2] = {false, false};

  // resident data
  double* h1 = nullptr;        // (2,N,N) symmetrised after measurement build
  double* chol = nullptr;      // (G,N*N)
  double* mo[2] = {nullptr, nullptr};      // (N, nocc_s)
  double* rdm1 = nullptr;      // (2,N,N)
  double* rot_h1[2] = {nullptr, nullptr};  // (nocc_s, N)
  double* rot_chol[2] = {nullptr, nullptr}; // (G, nocc_s, N)
  double2* mf_shifts = nullptr; // (G,)
  double* mreal = nullptr;      // (2G,) m_g = sum L_g rho, and -m_g
  double2* h0_prop = nullptr;   // device scalar
  double2 h0_prop_host = {0, 0};
  double* exp_h1[2] = {nullptr, nullptr}; // (N,N)
  float* exp_h1m32[2] = {nullptr, nullptr}; // (N,N) exp_h1 - I in FP32, reduced-precision jobs

  // scratch (per job)
  double2* S[2] = {nullptr, nullptr};     // (W, nocc, nocc)
  double2* Sinv[2] = {nullptr, nullptr};  // (W, nocc, nocc)
  double2* det[2] = {nullptr, nullptr};   // (W,)
  double2* green[2] = {nullptr, nullptr}; // (W, nocc, N)
  double2* fb = nullptr;                  // (W, G)
  double2* xs = nullptr;                  // (W, G) shifted fields
  double2* terms = nullptr;               // (W, 2) shift_term, fb_term
  double2* vhs = nullptr;                 // (W, N, N)
  double2* tbuf[3] = {nullptr, nullptr, nullptr}; // (W, N, wmax) Taylor buffers
  double2* ov_new = nullptr;              // (W,)
  double2* coul = nullptr;                // (2, W, G) c_gamma per spin
  double2* kpart = nullptr;               // (2, W, nchunk) exchange partials
  double2* ene1 = nullptr;                // (W,)
  double2* energies = nullptr;            // (W,)
  double* fields = nullptr;               // (n_prop_steps_max, W, G) for afq_block
  long fields_cap = 0;
  double2* wgather = nullptr;             // (W, N, wmax) SR gather buffer
  int32_t* sr_idx = nullptr;              // (W,)
  double* cum = nullptr;                  // (W,) SR cumulative weights
  double* dscratch = nullptr;             // small device scalars
  int kchunk_max = 1;                     // exchange partials per walker
  double* tmp_nn = nullptr;               // (8, N, N) setup temporaries
  // reduced-precision mode (desc.precision != FP64)
  bool lp = false, bf16x3 = false;
  float* chol32 = nullptr;                // (G, N*N)
  float* rot_chol32[2] = {nullptr, nullptr}; // (G, nocc_s, N)
  float2* vhs32 = nullptr;                // (W, N, N)
  float2* t32[2] = {nullptr, nullptr};    // (W, N, wmax) Taylor terms
  // split-BF16 x3 exchange and force-bias operands (bf16 bit patterns):
  // rot_chol (G, 2, Opad, Npad) per spin in the exchange's tiled layout (rcs)
  // and with the planes interleaved (rcs_fb), and one walker batch of G
  // (W, 4, Opad, Npad) in either layout
  uint16_t* rcs[2] = {nullptr, nullptr};
  uint16_t* rcs_fb[2] = {nullptr, nullptr};
  uint16_t* gsplit = nullptr;
  // FP32 exchange operands (nocc <= 64): rot_chol (G, Opad, Npad) per spin and
  // one walker batch of G (W, 2, Opad, Npad), in the FP32 tiled layout
  float* rct32[2] = {nullptr, nullptr};
  float* gt32 = nullptr;
  // split-BF16 x3 VHS operands: L^T (NNr, 2, Gpad) and the shifted fields
  // (Wr, 4, Gpad); rows padded to 16, k to 32
  uint16_t* lt_split = nullptr;
  uint16_t* xs_split = nullptr;
  // split-BF16 x3 fused propagator: exp(-dt/2 h1) - I per propagated spin as
  // three bf16 planes (k_split_e1_bf16)
  uint16_t* e1s[2] = {nullptr, nullptr};
  uint16_t* vsplit = nullptr;             // (W, 4, 16 ceil(N/16), 32 ceil(N/32)) split VHS planes
  // nocc > 64: stored f blocks of the wide exchange path (k_exchange_wide)
  double2* xwide = nullptr;
  long xwide_elems = 0;
  // symmetric Cholesky vectors: the VHS product over block-packed columns
  // (sym_packed_cols), used when set_hamiltonian finds L_g = L_g^T; npk = 0 when
  // packing would not save work at this N
  long npk = 0;
  bool sym_chol = false;
  double* cholp = nullptr;                // (G, npk) FP64 packed L
  double2* vhsp = nullptr;                // (W, npk) FP64 packed VHS
  float2* vhsp32 = nullptr;               // (W, npk) reduced-precision packed VHS
  double* gemm_ws = nullptr;              // split-K workspace
  long gemm_ws_elems = 0;

  // live per-class kernel timing (afq_job_profile)
  struct Prof {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spans;
    std::vector<hipEvent_t> pool;
  } prof[AFQ_K_NCLASSES];
};

namespace afq {
// Brackets the launches issued in its scope with events when the class is enabled.
struct ProfScope {
  afq_job* j;
  int cls;
  hipStream_t st;
  hipEvent_t b = nullptr, e = nullptr;
  static hipEvent_t take(afq_job::Prof& p) {
    if (!p.pool.empty()) {
      hipEvent_t ev = p.pool.back();
      p.pool.pop_back();
      return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
  }
  ProfScope(afq_job* job, int c, hipStream_t s) : j(job), cls(c), st(s) {
    if (!j->prof[cls].on) return;
    b = take(j->prof[cls]);
    e = take(j->prof[cls]);
    if (b && e) (void)hipEventRecord(b, st);
  }
  ~ProfScope() {
    if (!b || !e) return;
    (void)hipEventRecord(e, st);
    j->prof[cls].spans.emplace_back(b, e);
  }
};
}  // namespace afq
