/*
 * rbmpp.h -- C ABI of the MI355X VMC hot path for the RBM x pair-product
 * wave function of the J1-J2 Heisenberg model.
 *
 * The entry points replace the reference's sampler / estimator / SR-step
 * seam (yusukenomura/RBM_PP, code/normal_code) one for one and are meant to
 * be bound from a Fortran driver through iso_c_binding (see
 * rbm_pp_amd/fortran/rbmpp_mod.f90).  All arrays cross the boundary in the
 * reference's own Fortran layouts (column-major, 1-based index values where
 * the reference files hold 1-based values), as plain pointers.  The handle
 * owns all device memory and one HIP stream; calls are synchronous on
 * return and return 0 on success or a negative RBMPP_E_* status (never
 * abort).  One host thread per handle.
 */
#ifndef RBMPP_H
#define RBMPP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct rbmpp_handle rbmpp_handle;

/* status codes */
#define RBMPP_OK 0
#define RBMPP_E_ARG (-1)          /* bad argument / size */
#define RBMPP_E_HIP (-2)          /* HIP runtime error */
#define RBMPP_E_NOMEM (-3)        /* device allocation failed */
#define RBMPP_E_SINGULAR (-10)    /* X singular (pivot == 0) */
#define RBMPP_E_THETA_DRIFT (-40) /* incremental theta off by >1e-5 (make_sample.f90:84-89) */
#define RBMPP_E_SIGN_DRIFT (-42)  /* incremental sign /= recomputed (make_sample.f90:90-92) */
#define RBMPP_E_CANDIDATE (-30)   /* no exchange candidate found (make_sample.f90:290) */
#define RBMPP_E_INIT (-21)        /* no initial configuration with |psi|>1e-300 (make_sample.f90:359) */
#define RBMPP_E_SDIAG (-70)       /* S diagonal checks of SR_pdposv.f90:32-33,78-100 */
#define RBMPP_E_NOT_SPD (-75)     /* Cholesky failed (pdposv info /= 0, SR_pdposv.f90:233) */
#define RBMPP_E_SLATER (-5)       /* Slater_Elm not skew-symmetric (slater.f90:51-56) */
#define RBMPP_E_STATE (-6)        /* call out of order (e.g. sr_step before measure) */

/* S-matrix precision: products of the S build on the matrix cores in FP64,
   FP32, or split BF16 x 3 (c = hi + lo, hi*hi + hi*lo + lo*hi) with FP32
   accumulation.  The sampler, X^-1 chain, centring, shift and solve are FP64
   in every mode. */
#define RBMPP_S_FP64 0
#define RBMPP_S_FP32 1
#define RBMPP_S_BF16X3 2

/* Spin-projection weights (InitSPWeight): RBMPP_PROJ_NORMAL is
   normal_code/spin_projection.f90:5-79 (beta = {0, pi}) with cos(pi/2) taken
   as exactly 0 instead of its rounding residue 6.1e-17, so that every X is
   block off-diagonal and the sampler works with the N/2 x N/2 block;
   RBMPP_PROJ_NORMAL_EXACT keeps the residue as the reference computes it;
   RBMPP_PROJ_TRANSVERSE is transverse_code/spin_projection.f90:30-38
   (beta = {-pi/2, pi/2}, cos and sin of beta/2 scaled by sqrt(2)). */
#define RBMPP_PROJ_NORMAL 0
#define RBMPP_PROJ_TRANSVERSE 1
#define RBMPP_PROJ_NORMAL_EXACT 2

typedef struct {
  int32_t L;                 /* linear size; N = L*L sites (main.f90:77,85) */
  int32_t n_sym;             /* N_MP_Trans, number of symmetry operations (main.f90:91) */
  int32_t alphac, alphar;    /* complex / real hidden-unit densities (main.f90:79) */
  int32_t n_slater;          /* N_Slater (main.f90:128) */
  int32_t spin_parity;       /* +1 / -1 (main.f90:77) */
  double J1, J2;             /* couplings (main.f90:83) */
  int32_t n_walkers;         /* independent Markov chains (one MPI rank each in the reference) */
  int32_t n_samples;         /* stored samples per walker per make_sample call */
  int32_t n_warmup;          /* sweeps before storing; <0: max(30, n_samples/100) (make_sample.f90:66) */
  int32_t n_warmup_first;    /* sweeps before storing on the first call; <0: as n_warmup */
  int32_t n_update;          /* proposals per sweep; 0: N (main.f90:139) */
  int32_t recompute_k;       /* >0: full Pf/X^-1 recompute after recompute_k accepts;
                                0: adaptive Nfastup in [1,100] (make_sample.f90:74-128) */
  int32_t s_precision;       /* RBMPP_S_* */
  int32_t fused_measure;     /* 1: measure every stored sample inside the sampler with the
                                walker's fast-updated X^-1 (refreshed every recompute_k
                                accepts); 0: recompute Pf and X^-1 per sample in
                                rbmpp_measure, as calc_phys_quantity.f90:85 does.  N <= 104 */
  uint64_t seed;             /* Philox key; walker w uses stream w */
  double p_restart;          /* reference 0.05 (make_sample.f90:61) */
  double p_global;           /* reference 3e-4 (make_sample.f90:135) */
  double cond_guard;         /* > 0: recompute Pf and X^-1 of the walker right after an
                                accepted move whose Pf ratio of any projection copy lies
                                outside (1/cond_guard, cond_guard); 0: never (the
                                reference).  A copy that passes close to a singular
                                configuration otherwise leaves the fast-updated Pf and
                                X^-1 inaccurate (up to overflow) until the next periodic
                                recompute; not in the reference. */
} rbmpp_config;

typedef struct {
  double s_factor0;          /* relative diagonal shift (SR_pdposv.f90:167) */
  double delta_tau;          /* step (main.f90:79) */
  double max_change;         /* 0.04 (main.f90:50) */
  double slater_abs_av;      /* 0.25 (module.f90:66) */
  int32_t refine_iters;      /* FP64 iterative-refinement sweeps after the Cholesky solve (0..8); with a
                                mixed-precision S the conjugate-gradient iterations captured with it */
  int32_t flags;             /* RBMPP_SR_* bits; 0 by default */
} rbmpp_sr_config;

/* rbmpp_sr_config.flags: run the SR system's launches eagerly instead of
   replaying the hipGraph they are captured into on the first call */
#define RBMPP_SR_NO_GRAPH 1
/* With S assembled from FP32 or split-BF16 products, the system is solved by
   default with conjugate gradients on the FP64 operator of the stored O rows,
   D (C^T C / norm) D + shift with C = O - kappa <O>^T (one pass over O per
   iteration), preconditioned by the Cholesky factor of the assembled matrix:
   refine_iters iterations inside the captured launch sequence, then more
   (up to 40 per round) until the relative residual is below 1e-13;
   each round ends with the true residual D g - A x, and a round that leaves
   it above the tolerance restarts the recurrence from it (up to four
   rounds), so that the step is the FP64 system's solution.  The diagonal of a mixed S is formed in
   FP64 from the O rows.  This flag refines against the assembled matrix
   instead (refine_iters sweeps). */
#define RBMPP_SR_ASSEMBLED_REFINE 2

typedef struct {
  double lambda;             /* sqrt(x^T S x) * 5/L (SR_pdposv.f90:309) */
  double delw_max;           /* max |delta_tau * x_k| (SR_pdposv.f90:329-340) */
  double scaling;            /* PP block scale (SR_pdposv.f90:40) */
  double sdiag_av;           /* mean positive diagonal (SR_pdposv.f90:133) */
  double s_factor;           /* max(1e-4, s_factor0) */
  double residual;           /* ||S x - g|| / ||g|| after refinement (S: the operator refined against) */
  double mean_abs_f;         /* mean |f| before renormalisation */
  int32_t k_nonzero;         /* SR_pdposv.f90:337-341 */
  int32_t cg_iters;          /* conjugate-gradient iterations on the O-row operator (0: FP64 S or
                                RBMPP_SR_ASSEMBLED_REFINE) */
  double precond_boost;      /* diagonal added to the assembled mixed-precision matrix to factor it
                                as the preconditioner (0 when it was positive definite) */
} rbmpp_sr_report;

/* Creates the device state.  Tables in the reference's layouts:
 *   w_map[(g-1)*N + (i-1)]        = W_map(i,1,g), 1-based site (main.f90:106)
 *   para_mp_trans[2*(g-1)+{0,1}]  = re/im of para_mp_trans(g)  (main.f90:98)
 *   orbital_idx[(j-1)*N + (i-1)]  = Orbital_Idx(i,j), 1-based  (parameter.f90:93-95)
 * Replaces the set-up of main.f90:85-259 (sizes, allocation, InitSPWeight,
 * define_sublattice) for the GPU path. */
int rbmpp_create(const rbmpp_config* cfg, const int32_t* w_map, const double* para_mp_trans,
                 const int32_t* orbital_idx, rbmpp_handle** out);
void rbmpp_destroy(rbmpp_handle* h);
const char* rbmpp_last_error(const rbmpp_handle* h);
int rbmpp_nv(const rbmpp_handle* h);

/* Parameters (the BCAST payload of main.f90:251-259) followed by
 * update_slater_elements (slater.f90:7-64) on the device.
 *   wcirr[2*((f-1)*N+(iw-1))+{0,1}] = Wcirr(iw,f);  bcirr[2*(f-1)+{0,1}]
 *   wrirr[(f-1)*N+(iw-1)] = Wrirr(iw,f);  brirr[f-1];  slater[k-1] = Slater(k) */
int rbmpp_set_params(rbmpp_handle* h, const double* wcirr, const double* bcirr,
                     const double* wrirr, const double* brirr, const double* slater);
/* gamma vector in the order of main.f90:263-294 (SR_pdposv.f90:385-412) */
int rbmpp_set_gamma(rbmpp_handle* h, const double* gamma);
/* Selects the spin-projection weights (RBMPP_PROJ_*) and rebuilds
 * Slater_Elm with the current parameters: the InitSPWeight call of
 * main.f90:202, compiled from the normal_code or the transverse_code tree.
 * rbmpp_create starts with RBMPP_PROJ_NORMAL. */
int rbmpp_set_projection(rbmpp_handle* h, int32_t projection);
int rbmpp_get_gamma(const rbmpp_handle* h, double* gamma);

/* make_sample (make_sample.f90:4-240) for all walkers; the samples stay on
 * the device.  naccept = accepted exchanges in stored sweeps, summed over
 * walkers (main.f90:348). */
int rbmpp_make_sample(rbmpp_handle* h, int32_t first_iteration, double shift, int64_t* naccept);

/* calc_phys_quantity_opt (calc_phys_quantity.f90:20-246) over the stored
 * samples followed by the reductions of main.f90:347-367.  Outputs are the
 * normalised values main.f90 holds after the reduction: etot = sum E / norm,
 * etot2, norm (un-normalised), wf_av = sum|psi|/Ns, gvec, ovec (Nv each).  O
 * and S stay on the device for rbmpp_sr_step. */
int rbmpp_measure(rbmpp_handle* h, double shift, double* etot, double* etot2, double* norm,
                  double* wf_av, double* gvec, double* ovec);

/* The SR block (SR_pdposv.f90:1-414): S build in the configured precision
 * with FP64 centring, PP block scaling, relative diagonal shift, FP64
 * Cholesky + iterative refinement (replaces pdposv), trust region, max-change
 * clip, f renormalisation.  gamma is updated in place and loaded as the new
 * parameters. */
int rbmpp_sr_step(rbmpp_handle* h, const rbmpp_sr_config* c, double* gamma,
                  rbmpp_sr_report* rep);

/* Switches the S-build precision (RBMPP_S_*) for the next SR step; no
 * counterpart in the reference (it builds S in FP64 only). */
int rbmpp_set_s_precision(rbmpp_handle* h, int32_t s_precision);

/* calc_phys_quantity_after_opt (calc_phys_quantity.f90:250-302) +
 * main.f90:528-542 reductions. */
int rbmpp_measure_energy(rbmpp_handle* h, double shift, double* etot, double* etot2,
                         double* norm, double* wf_av);

/* The measurement of the correlation run on the stored samples
 * (Calc_Correlation/calc_phys_quantity.f90:264-550, its
 * calc_phys_quantity_after_opt, with the reductions and normalisation of
 * Calc_Correlation/main.f90:543-581): energy as rbmpp_measure_energy, and
 *   wf_sgn[2]           = sum psi/|psi| / Nsample (re, im; may be NULL)
 *   chi_sp[2*(i0-1)+k-1] = Chi_sp(k, i0) / norm   (k = 1: SzSz, 2: SpSm)
 *   chi_vbs[(t-1)*Nchi+i0-1] = Chi_vbs(i0, t) / norm  (t = 1..4 bond types)
 * for the Nchi = N(N+1)/2 pairs i1 <= i2 numbered i0 = i2(i2-1)/2 + i1
 * (1-based, i2 outer), i.e. the Fortran arrays Chi_sp(2,Nchi) and
 * Chi_vbs(Nchi,4).  Needs L >= 4 and N <= 104. */
int rbmpp_measure_correlation(rbmpp_handle* h, double shift, double* etot, double* etot2, double* norm,
                              double* wf_av, double* wf_sgn, double* chi_sp, double* chi_vbs);

/* ---- tuning ----
 * Times the launch configurations the kernels accept on this handle's own
 * data and keeps the fastest for later calls; no counterpart in the
 * reference.  RBMPP_TUNE_S_BUILD (after rbmpp_measure or rbmpp_set_measured,
 * in the current S precision): output tile (128, or 256 for split BF16 x 3)
 * and split-K count of the S build.  RBMPP_TUNE_MEASURE (stored samples):
 * workgroups of the per-sample estimator, timed on up to 32768 stored samples
 * (it overwrites the measurement buffers: measure again afterwards).  The
 * sampler has no launch choice to tune: one walker per workgroup of one
 * wave per projection copy, one workgroup per CU (its LDS). */
#define RBMPP_TUNE_S_BUILD 1
#define RBMPP_TUNE_MEASURE 2
typedef struct {
  int32_t s_tile, s_nsplit;                  /* chosen */
  int32_t s_default_tile, s_default_nsplit;  /* built-in defaults, for comparison */
  double s_build_s;                          /* S build with the choice, seconds */
  int32_t measure_grid, measure_configurations;
  double measure_s;                          /* estimator with the choice, seconds */
} rbmpp_tune_report;
int rbmpp_tune(rbmpp_handle* h, int32_t what, rbmpp_tune_report* rep);

/* ---- several models in one sampler launch ----
 * A batch of handles (e.g. lattices of different sizes and couplings, each
 * with its own parameters, walkers and SR state) whose make_sample runs as
 * one persistent launch over all their walkers: each workgroup takes the
 * model its walker belongs to.  Every handle then measures and takes its SR
 * step on its own (rbmpp_measure, rbmpp_sr_step), and may still sample alone.
 * The batch covers the normal projection's block form (16 copies, N <= 256)
 * and the skew form above N = 104 (up to 256); no fused measurement.  Walker w of handle
 * i draws from the same Philox stream as in rbmpp_make_sample, so a batch
 * call reproduces the handles' separate calls exactly.  No counterpart in
 * the reference (one lattice per run, main.f90:75-141). */
typedef struct rbmpp_batch rbmpp_batch;
int rbmpp_batch_create(rbmpp_handle* const* handles, int32_t n, rbmpp_batch** out);
/* first_iteration[i], shift[i], naccept[i] per handle, as rbmpp_make_sample */
int rbmpp_batch_make_sample(rbmpp_batch* b, const int32_t* first_iteration, const double* shift,
                            int64_t* naccept);
/* device time of the last batch launch, seconds */
double rbmpp_batch_sample_seconds(const rbmpp_batch* b);
const char* rbmpp_batch_last_error(const rbmpp_batch* b);
void rbmpp_batch_destroy(rbmpp_batch* b);

/* ---- parity / test entry points on caller-given configurations ---- */
/* Sample store: x[(s)*N + i] = +-1, s < n (walker-major: s = w*n_samples + k). */
int rbmpp_set_samples(rbmpp_handle* h, int32_t n, const int32_t* x);
int rbmpp_get_samples(const rbmpp_handle* h, int32_t* x, double* psi);
int64_t rbmpp_n_samples(const rbmpp_handle* h);
/* Per-sample values of the last measurement (calc_phys_quantity_opt,
 * main.f90:351-367, before the sums): psi[2s+{0,1}], the reweighted local
 * energy eloc[2s+{0,1}] = rw[s] * sum_x' H psi(x'), rw[s] = 1/(|psi| + shift)
 * (calc_phys_quantity.f90:85-92).
 * Any pointer may be NULL.  Valid after rbmpp_measure / rbmpp_make_sample
 * with fused measurement. */
int rbmpp_get_measured(const rbmpp_handle* h, double* psi, double* eloc, double* rw);
/* Loads a measurement in place of rbmpp_measure, for rbmpp_sr_step: the
 * per-configuration values of rbmpp_get_measured plus the O rows
 * (orows[s*Nv + k] = O_loc(k) of configuration s, already reweighted as
 * calc_phys_quantity.f90:187-219 stores them; complex models: rows 2s (re) and
 * 2s+1 (im)), then the sums of main.f90:351-367.  With orows == NULL the
 * values are synthetic instead (psi = rw = 1, eloc and O uniform in (-1, 1)
 * from `seed`, the f_ii columns zero), so that the SR step can be timed at
 * shapes the sampler is not run at.  n may exceed the sample store.  No
 * counterpart in the reference. */
int rbmpp_set_measured(rbmpp_handle* h, int64_t n, const double* psi, const double* eloc, const double* rw,
                       const double* orows, uint64_t seed);
/* Walker state after make_sample: x[w*N+i], pf[w*C+q], psi[2w+{0,1}] and
 * (optional) InvM in the reference's layout invm[w][q][(j)*N+(i)] = InvM(i+1,j+1,q+1). */
int rbmpp_get_walkers(const rbmpp_handle* h, int32_t* x, double* pf, double* psi, double* invm,
                      int64_t* naccept_per_walker);
/* Ele_Idx of every walker after make_sample: ele_idx[w*N + i] = 1-based site
 * of electron i+1 (up electrons first, then down; interface_for_PP.f90:5-40).
 * The walker's InvM of rbmpp_get_walkers is indexed by these electrons. */
int rbmpp_get_walker_electrons(const rbmpp_handle* h, int32_t* ele_idx);
/* Full evaluation (calc_theta, calculate_MAll, sign_PP, calc_amplitude_RBM_PP)
 * of n configurations: psi[s][0..C] complex, pf[s][C], invm optional. */
int rbmpp_eval(rbmpp_handle* h, int32_t n, const int32_t* x, double* psi, double* pf,
               double* invm);
/* Pf ratios (calculate_new_PfM_two) for one exchange per configuration:
 * i1[s], i2[s] 0-based sites with opposite spins; outputs pf_new[s][C] and
 * psi_new[s] (the amplitude of the exchanged configuration). */
int rbmpp_ratio_exchange(rbmpp_handle* h, int32_t n, const int32_t* x, const int32_t* i1,
                         const int32_t* i2, double* pf_new, double* psi_new);
/* calc_local_energy_correlation + O_k of calc_phys_quantity.f90:114-222 for n
 * configurations, not reweighted: eloc[s] complex (sum_x' H psi(x')), psi[s]
 * complex, oloc[s][Nv] complex (d psi / d gamma_k); oloc may be NULL. */
int rbmpp_local(rbmpp_handle* h, int32_t n, const int32_t* x, double* eloc, double* psi,
                double* oloc);
/* rbmpp_local with oloc = NULL (local energies and amplitudes of n
 * configurations, not reweighted), allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Its
 * device staging (the configurations, amplitudes, local energies, statuses and
 * the estimator's per-workgroup scratch) lives in one fixed array of 64 MiB per
 * GPU, held from the first rbmpp_create on that GPU until the process ends.
 * It takes at least (64 MiB - 32 MiB - 2 KiB) / (5 N + 44) configurations per
 * call (61,677 at N = 100, 25,341 at N = 256); a larger n returns RBMPP_E_ARG
 * with the limit in the message.  Calls that overlap on one GPU (from any
 * handle or thread) race on the array and must be serialised by the caller;
 * the array keeps the last call's contents. */
int rbmpp_local_noalloc(rbmpp_handle* h, int32_t n, const int32_t* x, double* eloc, double* psi);
/* The centred S = <O O^T>/norm - <O><O>^T of the last rbmpp_measure (before
 * the SR step's scaling and shift), Nv x Nv column-major, in the configured
 * precision; no counterpart in the reference (its Smat is the raw sum). */
int rbmpp_get_smat(rbmpp_handle* h, double* s);
/* phase timings of the last calls in seconds:
 * {slater, sample, measure, s_build, solve, sr_total, recompute_count, reserved} */
int rbmpp_timings(const rbmpp_handle* h, double* t);

#ifdef __cplusplus
}
#endif
#endif
