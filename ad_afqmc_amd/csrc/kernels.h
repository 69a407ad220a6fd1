// Launch helpers for the walker-step kernels in kernels.hip.
#pragma once
#include "afq_internal.h"

namespace afq {

int k_sym_h1(const double* in, double* out, int n, hipStream_t s);
int k_axpby4(double* out, const double* x, const double* y, double alpha, const double* u, const double* v,
             double beta, long n, hipStream_t s);
int k_mf(const double* m, int G, double h0, double2* mf, double* mfr, double2* h0_prop, hipStream_t s);
int k_permute_chunk(const double* L, double* out, int n, int g0, int gch, hipStream_t s);
int k_norm1(const double* A, int n, double* out, hipStream_t s);
int k_identity(double* out, int n, hipStream_t s);
int k_scale(double* out, const double* A, double scale, int n, hipStream_t s);

int k_overlap_combine(const double2* d0, const double2* d1, double2* out, int W, hipStream_t s);
int k_ene1(const double2* G0, const double* rh0, int n0, const double2* G1, const double* rh1, int n1, int N, double f0,
           double2* out, int W, hipStream_t s);
int exchange_nchunk(int Gn);
int k_exchange(const double* rc, const double2* Gm, int O, int N, int Gn, int W, double2* coul, double2* kpart,
               int* nchunk_out, hipStream_t s);
// nocc > 64 (any precision, FP64 arithmetic): f_g of a pass of g's for all
// walkers by the batched FP64 product into ws, then c_g and the K_g of each
// 4-g chunk; ws holds exchange_wide_ws_elems(O, Gn, W) complex elements.
long exchange_wide_ws_elems(int O, int Gn, int W);
int k_exchange_wide(const double* rc, const double2* Gm, int O, int N, int Gn, int W, double2* ws, long ws_elems,
                    double2* coul, double2* kpart, int* nchunk_out, hipStream_t s);
constexpr int kExchangeRegMaxO = 64;  // largest nocc of the register-blocked kernels
// Split-BF16 operands of the exchange and the force bias: rows x (O, N) real
// (2 planes) or complex (4 planes) doubles -> bf16 (hi, lo) planes of
// bf16_opad(O) x bf16_npad(N), zero padded, each plane tiled by 32-deep k step
// (the exchange's layout) or, with ilv, the planes of a row interleaved every
// 32 elements (the force-bias product's layout).
int bf16_opad(int O);
int bf16_npad(int N);
// row_stride: elements from one row of out to the next (0: the row's own size).
int k_split_bf16(const double* src, bool cplx, int O, int N, long rows, bool ilv, uint16_t* out, hipStream_t s,
                 long row_stride = 0);
int k_exchange_bf16(const uint16_t* rcs, const uint16_t* gs, int O, int N, int Gn, int W, double2* coul,
                    double2* kpart, int* nchunk_out, hipStream_t s);
// FP32 operands of the exchange: rows x (O, N) real (1 plane) or complex (re,
// im planes) doubles -> FP32 planes of bf16_opad(O) x bf16_npad(N), zero
// padded, in the FP32 exchange's tiled layout.
int k_split_f32(const double* src, bool cplx, int O, int N, long rows, float* out, hipStream_t s);
int k_exchange_f32(const float* rct, const float* gt, int O, int N, int Gn, int W, double2* coul, double2* kpart,
                   int* nchunk_out, hipStream_t s);
int k_to_f32(const double* src, float* dst, long n, hipStream_t s);
// Symmetric Cholesky vectors L_g = L_g^T: the VHS product runs over the 16 x 16
// orbital blocks (ib, jb) with ib >= jb only, and the Taylor products read that
// packed VHS (Gemm::a_sym).  Packed column p is entry
// ((p % 256) / 16, p % 16) of block p / 256 = ib (ib + 1) / 2 + jb (zero where
// the block overhangs N).
long sym_packed_cols(int N);
// out2[0] = max |L_ij - L_ji|, out2[1] = max |L_ij| (double bit patterns, device)
int k_sym_check(const double* L, int G, int N, unsigned long long* out2, hipStream_t s);
// (G, N*N) -> (G, sym_packed_cols(N)) as doubles (out64) or floats (out32)
int k_sym_pack(const double* L, int G, int N, double* out64, float* out32, hipStream_t s);
// dst = src - I (n x n), rounded to FP32 after the subtraction
int k_minus_identity_f32(const double* src, float* dst, int n, hipStream_t s);
int k_energy_finalize(const double2* ene1, const double2* c0, const double2* c1, const double2* k0, const double2* k1,
                      int Gn, int nch0, int nch1, double h0, double2* out, int W, hipStream_t s);

int k_field_shift(const double2* fb, const double* x, const double2* mf, int Gn, double sqdt, double2* xs,
                  double2* terms, int W, hipStream_t s);
int k_weight_update(const double2* ov_new, double2* overlaps, const double2* terms, double* weights, double* scalars,
                    const double2* h0_prop, double dt, int W, hipStream_t s);
int k_block_finalize(const double2* energies, const double* weights, double* scalars, double clip, int W, double* out,
                     hipStream_t s);
int k_sr_index(double* weights, const double* zeta, int W, double* cum, int32_t* idx, hipStream_t s);
int k_gather_walkers(const double2* src, double2* dst, const int32_t* idx, long per, int W, hipStream_t s);
int k_normal_fill(double* out, long n, uint64_t seed, uint64_t counter, hipStream_t s);
// The same stream with the counter read from device memory, the SR offset of
// that counter, and the counter increment (graph-capturable block loop).
int k_normal_fill_dev(double* out, long n, uint64_t seed, const uint64_t* counter_dev, hipStream_t s);
int k_sr_zeta(double* zeta, uint64_t seed, const uint64_t* counter_dev, hipStream_t s);
int k_counter_incr(uint64_t* counter_dev, hipStream_t s);
int k_sr_zeta_host(double* zeta, uint64_t seed, uint64_t counter, hipStream_t s);

}  // namespace afq
