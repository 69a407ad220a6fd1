/*
 * afq.h -- C ABI of the MI355X phaseless-AFQMC walker engine (libafq.so).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (AFQ_OK = 0, negative AFQ_E_* on failure; afq_last_error() gives
 * the message of the calling thread's last failure) and is stream-ordered on
 * the caller's HIP stream (`stream` is a hipStream_t passed as void*; NULL
 * means the legacy default stream).  Calls do not synchronise the host unless
 * stated.  Array arguments are DEVICE pointers unless stated.
 *
 * Complex arrays are interleaved complex128 (re, im) in the reference's
 * row-major layouts:
 *   restricted walkers    (W, N, nelec_up)                 one buffer
 *   unrestricted walkers  (W, N, nelec_up), (W, N, nelec_dn) two buffers
 *   Green's function      (W, nocc_s, N)
 *   force bias            (W, G)
 *   overlaps / energies   (W,)
 *
 * A job owns its resident Hamiltonian (h1, Cholesky vectors), trial, derived
 * intermediates and scratch; it is not re-entrant (one stream at a time), and
 * distinct jobs may run concurrently on distinct streams.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   afq_build_measurement_intermediates -> hamiltonian.build_measurement_intermediates
 *        (ad_afqmc/hamiltonian.py:48-63; rhf/uhf._build_measurement_intermediates
 *         ad_afqmc/wavefunctions.py:1215-1232, 1472-1496)
 *   afq_build_propagation_intermediates -> hamiltonian.build_propagation_intermediates
 *        (ad_afqmc/hamiltonian.py:65-81; propagator_afqmc._build_propagation_intermediates
 *         ad_afqmc/propagation.py:378-438)
 *   afq_calc_overlap      -> wave_function.calc_overlap      (ad_afqmc/wavefunctions.py:93-163)
 *   afq_calc_green        -> rhf/uhf._calc_green             (ad_afqmc/wavefunctions.py:1089-1100, 1289-1312)
 *   afq_calc_force_bias   -> wave_function.calc_force_bias   (ad_afqmc/wavefunctions.py:202-262)
 *   afq_calc_energy       -> wave_function.calc_energy       (ad_afqmc/wavefunctions.py:264-394)
 *   afq_apply_trotprop    -> propagator_afqmc._apply_trotprop (ad_afqmc/propagation.py:154-272)
 *   afq_propagate_constrained -> propagator_afqmc.propagate_constrained (ad_afqmc/propagation.py:274-336)
 *   afq_orthonormalize    -> RHFWalkers/UHFWalkers.orthonormalize (ad_afqmc/walkers.py:155-157, 260-262;
 *                            ad_afqmc/linalg_utils.py:51-72)
 *   afq_stochastic_reconfiguration -> stochastic_reconfiguration_local
 *                            (ad_afqmc/walkers.py:139-145, 244-250; ad_afqmc/sr.py:24-48)
 *   afq_block             -> sampler._block_scan (ad_afqmc/sampling.py:52-94) with on-device fields
 *   afq_sr_block          -> sampler._sr_block_scan (ad_afqmc/sampling.py:146-168), graph-capturable
 *   afq_segment_energy_vjp -> jax.vjp of sampler.propagate_phaseless_ad_nosr_norot
 *                            (ad_afqmc/sampling.py:300-341; ad_afqmc/driver.py:865-876, 842-864)
 */
#ifndef AFQ_H
#define AFQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afq_job* afq_job_t;

enum {
  AFQ_OK = 0,
  AFQ_E_INVALID = -1,     /* bad argument / size */
  AFQ_E_NOMEM = -2,       /* device allocation failed */
  AFQ_E_HIP = -3,         /* HIP runtime error */
  AFQ_E_STATE = -4,       /* call order violated (e.g. intermediates not built) */
  AFQ_E_UNSUPPORTED = -5  /* shape or option outside what the kernels support */
};

enum { AFQ_WALKERS_RESTRICTED = 0, AFQ_WALKERS_UNRESTRICTED = 1 };
enum { AFQ_TRIAL_RHF = 0, AFQ_TRIAL_UHF = 1 };
/* Arithmetic of the VHS / Taylor / exchange contractions.  LU, overlap ratios,
 * QR and weight/energy accumulation stay FP64 in every mode. */
enum { AFQ_PREC_FP64 = 0, AFQ_PREC_FP32 = 1, AFQ_PREC_BF16X3 = 2 };

typedef struct {
  int32_t norb;        /* N */
  int32_t nelec_up;    /* occupied alpha orbitals */
  int32_t nelec_dn;    /* occupied beta orbitals (== nelec_up for an RHF trial) */
  int32_t nchol;       /* G */
  int32_t n_walkers;   /* W */
  int32_t walker_type; /* AFQ_WALKERS_* */
  int32_t trial_type;  /* AFQ_TRIAL_* */
  int32_t n_exp_terms; /* Taylor terms of exp(VHS), reference default 6 */
  int32_t precision;   /* AFQ_PREC_* */
  int32_t reserved;
  double dt;
} afq_job_desc;

/* Intermediates readable with afq_get_intermediate (for inspection / tests). */
enum {
  AFQ_INT_H1 = 0,        /* (2,N,N) f64, symmetrised */
  AFQ_INT_ROT_H1 = 1,    /* spin s: (nocc_s, N) f64 */
  AFQ_INT_ROT_CHOL = 2,  /* spin s: (G, nocc_s, N) f64 */
  AFQ_INT_MF_SHIFTS = 3, /* (G,) c128 */
  AFQ_INT_H0_PROP = 4,   /* scalar c128 */
  AFQ_INT_EXP_H1 = 5     /* spin s: (N,N) f64 (restricted walkers: spin 0 only) */
};

const char* afq_version(void);
const char* afq_last_error(void);

int afq_job_create(afq_job_t* job, const afq_job_desc* desc);
int afq_job_destroy(afq_job_t job);
/* Device bytes afq_job_create would allocate for desc (plus the afq_block
 * field buffer for n_prop_steps steps); -1 if desc is invalid.  No device call. */
int64_t afq_job_footprint(const afq_job_desc* desc, int32_t n_prop_steps);

/* Active walker count for the following calls, 1 <= n_walkers <= desc.n_walkers
 * (the capacity the job was created with).  Default: the capacity. */
int afq_job_set_batch(afq_job_t job, int32_t n_walkers);

/* Sets h0 and copies h1 (2,N,N) and chol (G,N*N) (device or host pointers)
 * into resident HBM.  h1/chol may be NULL to set h0 only (measurement-only use
 * with intermediates supplied by afq_set_intermediate).  Checks whether every
 * L_g is symmetric (to 1e-12 of max |L|; synchronises the stream once); if so,
 * the VHS product of the propagation covers the lower block triangle and
 * mirrors it. */
int afq_job_set_hamiltonian(afq_job_t job, double h0, const double* h1, const double* chol, void* stream);
/* Trial orbitals mo_up (N,nelec_up), mo_dn (N,nelec_dn; NULL for RHF) and the
 * trial 1-RDM rdm1 (2,N,N), device or host pointers. */
int afq_job_set_trial(afq_job_t job, const double* mo_up, const double* mo_dn, const double* rdm1, void* stream);

/* Builds rot_h1 / rot_chol (symmetrising h1 first, as the reference does). */
int afq_build_measurement_intermediates(afq_job_t job, void* stream);
/* Builds mf_shifts, h0_prop, exp_h1 = expm(-dt/2 h1_mod).  Synchronises the
 * stream (the scaling exponent of expm is chosen on the host). */
int afq_build_propagation_intermediates(afq_job_t job, void* stream);
/* Copies intermediate `which` (AFQ_INT_*) of spin `spin` to dst. */
int afq_get_intermediate(afq_job_t job, int which, int spin, void* dst, void* stream);
/* Overwrites intermediate `which` of spin `spin` from src (the reference's
 * measurement methods read whatever rot_h1 / rot_chol the caller's ham_data
 * holds).  For an RHF trial only spin 0 exists.  Once rot_h1 and rot_chol of
 * every spin are set the measurement intermediates count as built; once
 * mf_shifts, h0_prop and every exp_h1 are set, the propagation ones do. */
int afq_set_intermediate(afq_job_t job, int which, int spin, const void* src, void* stream);

/* Measurements.  walkers_dn is NULL for restricted walkers. */
int afq_calc_overlap(afq_job_t job, const void* walkers_up, const void* walkers_dn, void* overlaps, void* stream);
int afq_calc_green(afq_job_t job, const void* walkers_up, const void* walkers_dn, void* green_up, void* green_dn,
                   void* stream);
int afq_calc_force_bias(afq_job_t job, const void* walkers_up, const void* walkers_dn, void* force_bias,
                        void* stream);
int afq_calc_energy(afq_job_t job, const void* walkers_up, const void* walkers_dn, void* energies, void* stream);

/* W <- exp(-dt/2 h1) Taylor_n(exp(VHS)) exp(-dt/2 h1) W with
 * VHS = i sqrt(dt) sum_g shifted_fields[w,g] L_g; shifted_fields (W,G) c128. */
int afq_apply_trotprop(afq_job_t job, void* walkers_up, void* walkers_dn, const void* shifted_fields, void* stream);

/* One phaseless step.  fields (W,G) f64; weights (W,) f64 in/out; overlaps
 * (W,) c128 in/out; scalars (device f64[2]) = {e_estimate, pop_control_ene_shift},
 * the second updated in place.  force_bias_out (W,G) c128 may be NULL. */
int afq_propagate_constrained(afq_job_t job, void* walkers_up, void* walkers_dn, const double* fields,
                              double* weights, void* overlaps, double* scalars, void* force_bias_out,
                              void* stream);

/* QR re-orthonormalisation in place (R discarded): Cholesky QR twice, the
 * factorisation whose R has a positive real diagonal. */
int afq_orthonormalize(afq_job_t job, void* walkers_up, void* walkers_dn, void* stream);

/* Comb resampling by |weight| with offset zeta (device f64 scalar); permutes
 * walkers in place and sets every weight to sum|w|/W.  indices_out (W,) int32
 * may be NULL. */
int afq_stochastic_reconfiguration(afq_job_t job, void* walkers_up, void* walkers_dn, double* weights,
                                   const double* zeta, int32_t* indices_out, void* stream);

/* One sampler block: n_prop_steps phaseless steps with on-device N(0,1)
 * fields (Philox4x32-10 keyed by seed, counter advanced by *rng_counter),
 * then QR, overlap and energy measurement, energy clipping, and the
 * population-control shift update.  block_out (device f64[2]) receives
 * {block_energy, block_weight}.  *rng_counter (host) is advanced. */
int afq_block(afq_job_t job, void* walkers_up, void* walkers_dn, double* weights, void* overlaps,
              double* scalars, int n_prop_steps, uint64_t seed, uint64_t* rng_counter, double* block_out,
              void* stream);

/* One sampler block followed by comb resampling and the overlaps of the new
 * population -- the reference's _sr_block_scan with one energy block
 * (ad_afqmc/sampling.py:146-168) -- entirely on the device.  The stream
 * counter lives in device memory: *counter_dev (uint64) keys this block's
 * fields (the same Philox stream afq_block uses for that counter) and the SR
 * offset zeta ~ U[0,1), and is incremented on the device at the end.  With
 * the field buffer reserved (afq_job_reserve_fields) the call makes no host
 * synchronisation and no allocation, so it can be captured once into a HIP
 * graph and replayed.  block_out receives {block_energy, block_weight}
 * (before resampling). */
int afq_sr_block(afq_job_t job, void* walkers_up, void* walkers_dn, double* weights, void* overlaps,
                 double* scalars, int n_prop_steps, uint64_t seed, uint64_t* counter_dev, double* block_out,
                 void* stream);
/* Allocates the field buffer for n_prop_steps steps of the job's walker
 * capacity (synchronises the stream when it grows). */
int afq_job_reserve_fields(afq_job_t job, int32_t n_prop_steps, void* stream);

/* Fills out (n,) f64 with N(0,1) draws of the engine's field generator for
 * (seed, counter) -- the stream afq_block uses. */
int afq_normal_fill(double* out, int64_t n, uint64_t seed, uint64_t counter, void* stream);
/* Writes to out (device f64 scalar) the comb offset zeta ~ U[0,1) that
 * afq_sr_block draws for the block keyed by (seed, counter). */
int afq_sr_zeta(double* out, uint64_t seed, uint64_t counter, void* stream);

/* Live kernel timing.  Kernel classes of the walker step: */
enum {
  AFQ_K_VHS = 0,        /* VHS = i sqrt(dt) x~ . L GEMM */
  AFQ_K_TAYLOR = 1,     /* one Taylor term VHS . X (per spin buffer) */
  AFQ_K_EXPH1 = 2,      /* exp(-dt/2 h1) . W */
  AFQ_K_OVLP_GEMM = 3,  /* S = T^H W */
  AFQ_K_DETINV = 4,     /* Gauss-Jordan det / inverse of S */
  AFQ_K_GREEN = 5,      /* G = (W S^-1)^T */
  AFQ_K_FORCE_BIAS = 6, /* fb = G . rot_chol^T */
  AFQ_K_FIELD = 7,      /* field shifts */
  AFQ_K_WEIGHT = 8,     /* phaseless weight update */
  AFQ_K_EXCHANGE = 9,   /* fused Coulomb/exchange */
  AFQ_K_ENERGY_MISC = 10,
  AFQ_K_QR = 11,
  AFQ_K_SR = 12,
  AFQ_K_RNG = 13,
  AFQ_K_NCLASSES = 14
};
/* When enabled for a class, every launch of that class on the job's calls is
 * bracketed by hipEventRecord on the launch stream.  afq_job_profile_read
 * synchronises on the recorded events, returns the summed elapsed ms and
 * launch count of the class, and clears its record. */
int afq_job_profile(afq_job_t job, int kernel_class, int enable);
int afq_job_profile_read(afq_job_t job, int kernel_class, double* total_ms, int64_t* launches);

/* Job-free walker-batch operations (the reference's walker containers carry no
 * Hamiltonian: RHFWalkers/UHFWalkers.orthonormalize, walkers.py:155-157,
 * 260-262, and stochastic_reconfiguration_local, walkers.py:139-145, 244-250).
 * `walkers` is one (n_walkers, norb, ncols) c128 buffer. */
int afq_walkers_orthonormalize(void* walkers, int32_t n_walkers, int32_t norb, int32_t ncols, void* stream);
/* Bytes of device workspace afq_walkers_sr needs. */
int64_t afq_walkers_sr_workspace(int32_t n_walkers, int64_t max_walker_elems);
/* Comb resampling of up to two walker buffers sharing one weight vector:
 * elems_s = norb*ncols_s complex elements per walker (elems_dn = 0 and
 * walkers_dn = NULL for one buffer).  weights set to sum|w|/W. */
int afq_walkers_sr(void* walkers_up, int64_t elems_up, void* walkers_dn, int64_t elems_dn, int32_t n_walkers,
                   double* weights, const double* zeta, int32_t* indices_out, void* workspace, void* stream);

/* No counterpart in the reference.  afq_walkers_sr without the workspace
 * argument and without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results and errors as
 * afq_walkers_sr; its workspace is a 16 MiB device array of the library (one
 * per GPU, held until the process ends), so n_walkers is limited to
 * afq_walkers_sr_noalloc_max_walkers() (65535) and the walker gather runs in
 * column chunks that fit the array.  Calls that overlap on one GPU (on
 * different streams or threads) share that array and race.  The first call on
 * a device looks the array's address up (no stream work); make it before
 * capturing a graph. */
int afq_walkers_sr_noalloc(void* walkers_up, int64_t elems_up, void* walkers_dn, int64_t elems_dn, int32_t n_walkers,
                           double* weights, const double* zeta, int32_t* indices_out, void* stream);
int64_t afq_walkers_sr_noalloc_max_walkers(void);

/* ---- grouped blocks (serving) ----
 * No single counterpart in the reference, whose many-job workloads (LNO
 * fragments) run one job after another (lnoutils.py:175-177): one afq_sr_block
 * (sampling.py:52-94, 146-168 with the comb of sr.py:24-48) for every job of a
 * group, each kernel of the block launched once over all the jobs.  Jobs:
 * restricted walkers of an RHF trial, split-BF16 x3 precision, N <= 512,
 * nocc <= 64, with their fields reserved for n_prop_steps.  Per job i the group keeps the
 * caller's walkers[i] (W, N, O) c128, weights[i] (W,), overlaps[i] (W,) c128,
 * scalars[i] (e_estimate, pop_control_ene_shift), rng_counters[i] (device
 * uint64) and block_out[i] (block energy, block weight), all device buffers
 * that must outlive the group.  A grouped block updates them exactly as
 * afq_sr_block(job_i, ..., seeds[i], rng_counters[i], block_out[i]) does, then
 * advances e_estimate to 0.9 e_estimate + 0.1 E_block (driver.py:168).
 * afq_group_create copies the per-job launch parameters to the device
 * (synchronous); afq_group_sr_block is stream-ordered, allocates nothing and
 * may be captured in a HIP graph.  A job may belong to one group at a time. */
typedef struct afq_group* afq_group_t;
int afq_group_create(afq_group_t* out, const afq_job_t* jobs, int32_t n_jobs, void* const* walkers,
                     double* const* weights, void* const* overlaps, double* const* scalars,
                     uint64_t* const* rng_counters, double* const* block_out, const uint64_t* seeds,
                     int32_t n_prop_steps);
int afq_group_sr_block(afq_group_t group, void* stream);
int afq_group_destroy(afq_group_t group);

/* Value and reverse-mode gradient of the segment energy
 *     F = sum_b E_b W_b / sum_b W_b
 * over n_blocks blocks of n_steps phaseless steps (afq_block's step, QR,
 * overlap, energy and block bookkeeping) with the given fields
 * (n_blocks, n_steps, W, G) and no population control, starting from
 * `walkers` (not modified; restricted walkers (W, N, nocc), or for
 * unrestricted walkers spin up's (W, N, nocc_up) panel followed by spin
 * down's (W, N, nocc_dn) in the same buffer) and `weights` (W,), with the
 * population-control shift starting at, and the energy clipping centred on,
 * e_estimate.  The intermediates are
 * functions of the job's h1 and Cholesky vectors (they must have been built by
 * afq_build_*_intermediates, not supplied); the gradient is taken with respect
 * to those inputs:
 *   energy_out (1,) device, h1_bar (2, N, N), chol_bar (G, N*N).
 * h1_bar = chol_bar = NULL evaluates F only.  The fields are constants (the
 * reference draws them from a key that carries no tangent); clipping masks
 * are held fixed.  RHF trial with restricted walkers or UHF trial with
 * unrestricted walkers, FP64 jobs.  Allocates a
 * device workspace for the call and synchronises the stream.  With a gradient
 * requested, the steps' tapes of every block are kept from the forward sweep
 * when that workspace takes at most half of the free device memory (or at
 * most AFQ_SEGMENT_TAPE_LIMIT bytes, from the environment, when set);
 * otherwise each block is recomputed from its checkpoint in the reverse sweep
 * (O(n_blocks + n_steps) walker arrays).  Both give the same result. */
int afq_segment_energy_vjp(afq_job_t job, const void* walkers, const double* weights, double e_estimate,
                           const double* fields, int32_t n_blocks, int32_t n_steps, double* energy_out,
                           double* h1_bar, double* chol_bar, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AFQ_H */
