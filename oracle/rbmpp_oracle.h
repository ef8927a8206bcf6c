/*
 * CPU oracle for the RBM x pair-product VMC hot path (J1-J2 Heisenberg model).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran routines (yusukenomura/RBM_PP, code/normal_code), used as the
 * checker for the HIP implementation and as the timed CPU baseline.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (rbm_pp_amd/librbmpp.so) never links or calls it.
 *
 * Conventions (all 0-based):
 *   site index  s = ix*L + iy                      (bond_map, calc_phys_quantity.f90:394-422)
 *   copy index  q = g*2 + sidx, sidx=0 <-> beta=0   (sub.f90:30-33)
 *   theta index j = q*alpha + f                     (RBM.f90:46-64)
 *   W_map[g*N+i]   = W_map(i+1,1,g+1)-1             (main.f90:106)
 *   orb[i*N+j]     = Orbital_Idx(i+1,j+1)-1         (parameter.f90:93-95)
 *   Wr[f*N+iw]     = Wrirr(iw+1,f+1)                (module.f90:30)
 *   Wc[2*(f*N+iw)] = re/im of Wcirr(iw+1,f+1)
 *   se[q*4N^2 + j*2N + i]   = Slater_Elm(i+1,j+1,q+1)  (column-major, as slater.f90)
 *   invm[q*N^2 + j*N + i]   = InvM(i+1,j+1,q+1)        (column-major, as matrix.f90)
 *   complex values are interleaved (re, im) doubles.
 */
#ifndef RBMPP_ORACLE_H
#define RBMPP_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct om_model om_model;
typedef struct om_state om_state;
typedef struct om_rng om_rng;

/* ---- random number streams ---- */
om_rng* om_rng_mt(int32_t seed);                                  /* mt.f sgrnd/grnd */
om_rng* om_rng_philox(uint64_t seed, uint32_t stream, uint64_t counter);
void    om_rng_free(om_rng* r);
double  om_rng_u01(om_rng* r);
void    om_rng_fill(om_rng* r, int n, double* out);
uint64_t om_rng_counter(const om_rng* r);
void    om_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- model ---- */
om_model* om_create(int L, int P, int alphac, int alphar, int n_slater,
                    double J1, double J2, int spin_parity,
                    const int* w_map, const double* chi, const int* orb_idx);
void om_free(om_model* m);
int  om_nv(const om_model* m);
int  om_nv_rbm(const om_model* m);
int  om_set_params(om_model* m, const double* Wc, const double* bc,
                   const double* Wr, const double* br, const double* f);
int  om_set_gamma(om_model* m, const double* gamma);
/* 0: normal_code weights; 1: transverse_code/spin_projection.f90:30-38. */
int  om_set_projection(om_model* m, int proj);
int  om_get_gamma(const om_model* m, double* gamma);
const double* om_slater_elm(const om_model* m);
void om_sp_weights(const om_model* m, double* out /* cos,sin,cc,cs,ss,w  x2 */);

/* ---- linear algebra used by calculate_MAll ---- */
double om_pfaffian(int n, const double* A);            /* column-major, Parlett-Reid */
int    om_inverse(int n, const double* A, double* Ainv); /* LU with partial pivoting */

/* ---- wave-function evaluation on one configuration ---- */
om_state* om_state_create(const om_model* m);
void om_state_free(om_state* s);
int  om_eval(const om_model* m, const int* x, om_state* s, om_rng* r);
int  om_state_get(const om_model* m, const om_state* s, int* x, double* thetac,
                  double* thetar, int* ele_idx, int* ele_cfg, int* sgn, double* pf,
                  double* invm, double* psi_wo, double* psi);
void om_calc_theta(const om_model* m, const int* x, double* thetac, double* thetar);
void om_sign_pp(const om_model* m, const int* ele_idx, int* sgn);
int  om_ratio_exchange(const om_model* m, const om_state* s, int i1, int i2,
                       double* ratio, double* pf_new, double* psi_new);
int  om_accept_exchange(const om_model* m, om_state* s, int i1, int i2, om_rng* r);
/* pfupdate_two.f90 with the reference's arguments (0-based electrons and
   sites, InvM column-major per copy as above); timing comparisons only */
void om_pfm_two(const om_model* m, int msa, int msb, const int* tmp_idx, const double* invm,
                const double* pf_old, double* pf_new);
void om_invm_two(const om_model* m, int msa, int msb, const int* ele_idx, double* invm);
double om_eloc(const om_model* m, const om_state* s, double* eloc /*2*/, om_rng* r);
int  om_oloc(const om_model* m, const om_state* s, double* oloc /*2*Nv*/);

/* ---- measurement over a set of samples (calc_phys_quantity_opt) ---- */
int om_measure(const om_model* m, int nsample, const int* xs, double shift, int want_s,
               double* sums /* Etot,Etot2,norm,wf_av */, double* gvec, double* ovec,
               double* smat, double* eloc_out, double* psi_out, double* ostore);
int om_measure_energy(const om_model* m, int nsample, const int* xs, double shift,
                      double* sums);
/* Correlation run (Calc_Correlation/calc_phys_quantity.f90:264-550) */
double om_bond_sdots(const om_model* m, const om_state* s, double* eloc /*2*/, double* sdots /*8N*/);
int om_local_correlation(const om_model* m, const om_state* s, double* chi /*4*N(N+1)/2*/);
int om_measure_correlation(const om_model* m, int nsample, const int* xs, double shift,
                           double* sums /*6*/, double* chi_sp /*N(N+1)*/, double* chi_vbs /*2N(N+1)*/);

/* periodic recompute with the maintained electron order (make_sample.f90:74-128);
   drift = {max |dPf|/<|Pf|>, max |dInvM|/max|InvM|} */
int om_recompute(const om_model* m, om_state* s, double* drift);

/* ---- sampler (make_sample) ---- */
typedef struct {
  int nsample;      /* stored samples for this chain                      */
  int nwarmup;      /* thermalisation sweeps (sweeps i0=-nwarmup..0)       */
  int nupdate;      /* proposals per sweep (reference: N)                  */
  int fixed_k;      /* >0: recompute when accepts since recompute > fixed_k;
                       0: reference adaptive Nfastup                       */
  int ref_draws;    /* 1: consume grnd where the reference does (mt.f mode) */
  double p_restart; /* reference 0.05 */
  double p_global;  /* reference 3e-4 */
  double shift;
  double cond_guard; /* > 0: recompute right after an accepted move whose Pf
                        ratio of any copy lies outside (1/cond_guard,
                        cond_guard) (not in the reference: 0 restates it) */
} om_sample_cfg;
int om_make_sample(const om_model* m, om_state* s, om_rng* r, int first,
                   const om_sample_cfg* c, int* nfastup, int* x_out, double* psi_out,
                   double* pf_out, long* naccept, long* nrecompute);

/* ---- SR step (SR_pdposv.f90 + main.f90 reductions) ---- */
typedef struct {
  double s_factor0;
  double delta_tau;
  double max_change;
  double slater_abs_av;
} om_sr_cfg;
int om_sr_step(const om_model* m, const om_sr_cfg* c, double etot, double norm,
               const double* gvec_in, const double* ovec, const double* smat_raw,
               double* gamma, double* report /* 8 doubles */, double* x_out);

/* ---- helpers ---- */
int om_bond(int L, int type, int i1);
#ifdef __cplusplus
}
#endif
#endif
