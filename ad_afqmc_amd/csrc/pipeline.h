// Walker-step pipelines shared by the C ABI (api.hip) and the reverse pass
// (adjoint.hip).  All run on the given stream and use the job's scratch buffers.
#pragma once
#include "afq_internal.h"

namespace afq {

int dalloc(void** p, size_t bytes);
template <typename T>
int dalloc_n(T** p, size_t n) {
  return dalloc(reinterpret_cast<void**>(p), n * sizeof(T));
}

int check_walkers(const afq_job* j, const void* up, const void* dn);
int require(const afq_job* j, bool meas, bool prop);

// S_s = T_s^T W_s, det_s; optionally S_s^-1 (j->Sinv) and G_s = (W_s S_s^-1)^T (j->green).
// have_inv: S_s^-1 and det_s are already those of these walkers; only G_s is formed.
int green_pipeline(afq_job* j, const void* up, const void* dn, bool want_inv, bool want_green, hipStream_t st,
                   bool have_inv = false);
int overlap_from_dets(afq_job* j, double2* out, hipStream_t st);
int force_bias_gemm(afq_job* j, double2* out, hipStream_t st);
// have_green: j->green already holds the Green's functions of these walkers.
int energy_pipeline(afq_job* j, const void* up, const void* dn, double2* out, hipStream_t st,
                    bool have_green = false);
int trotprop(afq_job* j, void* up, void* dn, const double2* xs, hipStream_t st);
// inv_ready: j->Sinv/j->det belong to the incoming walkers (the previous step
// kept them); keep_inv: invert the new walkers' overlap matrices for the next step.
int propagate(afq_job* j, void* up, void* dn, const double* fields, double* weights, double2* overlaps,
              double* scalars, hipStream_t st, bool inv_ready = false, bool keep_inv = false);
int orthonormalize(afq_job* j, void* up, void* dn, hipStream_t st);

// v0 = 1/2 sum_g L_g L_g (N x N) from the job's Cholesky vectors.
int compute_v0(afq_job* j, double* v0, hipStream_t st);
// out = expm(A) for a device n x n matrix (scaling and squaring around a
// degree-18 Taylor/Horner core).  work: 3 n^2 doubles; dscr: one double.
// Synchronises the stream (the squaring count is chosen on the host).
int expm_device(const double* A, int n, double* out, double* work, double* dscr, hipStream_t st);

}  // namespace afq
