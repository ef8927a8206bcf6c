/*
 * CPU oracle: plain-C restatement of the reference RBM x pair-product VMC
 * routines (yusukenomura/RBM_PP, code/normal_code).  TEST INFRASTRUCTURE
 * ONLY -- the checker for the HIP path and the timed CPU baseline; never
 * part of the product library.  Each function cites the reference routine
 * it restates.  Index conventions are listed in rbmpp_oracle.h.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   mt19937       -- first 10^4 grnd values from the compiled reference mt.f
 *   philox        -- Random123 known-answer vectors
 *   om_pfaffian   -- brute-force matchings (n<=8), Pf^2 = det, 4x4 closed form
 *   om_inverse    -- A*Ainv = I
 *   theta / sign / slater / amplitude / ratio / update / eloc / oloc
 *                 -- compiled reference routines on identical inputs
 *                    (oracle/_ref), fast update vs full recompute,
 *                    exact 4x4 enumeration against an independent sparse
 *                    Hamiltonian, central finite differences of psi
 *   om_sr_step    -- residual of the solve, reference invariants
 *                    (SR_pdposv.f90:140,313-323,414)
 */
#include "rbmpp_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

/* ------------------------------------------------------------------ */
/* random streams                                                      */
/* ------------------------------------------------------------------ */
struct om_rng {
  int kind; /* 0 = mt.f, 1 = philox4x32-10 */
  uint32_t mt[624];
  int mti;
  uint32_t key[2];
  uint32_t stream;
  uint64_t ctr;
};

/* mt.f:6-29 (sgrnd): Knuth 69069 LCG seeding of the 624-word state. */
static void mt_sgrnd(om_rng* r, int32_t seed) {
  r->mt[0] = (uint32_t)seed;
  for (int i = 1; i < 624; i++) r->mt[i] = 69069u * r->mt[i - 1];
  r->mti = 624;
}

/* mt.f:32-102 (grnd): MT19937 word, tempered, divided by 2^32-1. */
static double mt_grnd(om_rng* r) {
  const uint32_t UM = 0x80000000u, LM = 0x7fffffffu, MATA = 0x9908b0dfu;
  uint32_t y;
  if (r->mti >= 624) {
    if (r->mti == 625) mt_sgrnd(r, 4357);
    int kk;
    for (kk = 0; kk < 624 - 397; kk++) {
      y = (r->mt[kk] & UM) | (r->mt[kk + 1] & LM);
      r->mt[kk] = r->mt[kk + 397] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
    }
    for (; kk < 623; kk++) {
      y = (r->mt[kk] & UM) | (r->mt[kk + 1] & LM);
      r->mt[kk] = r->mt[kk + (397 - 624)] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
    }
    y = (r->mt[623] & UM) | (r->mt[0] & LM);
    r->mt[623] = r->mt[396] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
    r->mti = 0;
  }
  y = r->mt[r->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return (double)y / 4294967295.0;
}

/* Philox4x32-10 (Salmon et al., SC'11).  This is the stream the GPU sampler
 * uses; the oracle carries it so that sampler trajectories can be compared
 * draw for draw. */
void om_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int rnd = 0; rnd < 10; rnd++) {
    if (rnd) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double philox_u01(om_rng* r) {
  uint32_t ctr[4] = {(uint32_t)r->ctr, (uint32_t)(r->ctr >> 32), r->stream, 0u};
  uint32_t o[4];
  om_philox4x32_10(ctr, r->key, o);
  r->ctr++;
  uint64_t z = ((uint64_t)o[1] << 32) | o[0];
  return (double)(z >> 11) * 0x1.0p-53;
}

om_rng* om_rng_mt(int32_t seed) {
  om_rng* r = (om_rng*)calloc(1, sizeof(om_rng));
  r->kind = 0;
  mt_sgrnd(r, seed);
  return r;
}
om_rng* om_rng_philox(uint64_t seed, uint32_t stream, uint64_t counter) {
  om_rng* r = (om_rng*)calloc(1, sizeof(om_rng));
  r->kind = 1;
  r->key[0] = (uint32_t)seed;
  r->key[1] = (uint32_t)(seed >> 32);
  r->stream = stream;
  r->ctr = counter;
  return r;
}
void om_rng_free(om_rng* r) { free(r); }
double om_rng_u01(om_rng* r) { return r->kind == 0 ? mt_grnd(r) : philox_u01(r); }
void om_rng_fill(om_rng* r, int n, double* out) {
  for (int i = 0; i < n; i++) out[i] = om_rng_u01(r);
}
uint64_t om_rng_counter(const om_rng* r) { return r->ctr; }

/* make_sample.f90:254 -- nran(i) = mod(int(dble(i)*grnd()),i)+1, here 0-based */
static int nran0(om_rng* r, int i) { return ((int)((double)i * om_rng_u01(r))) % i; }

/* ------------------------------------------------------------------ */
/* model                                                                */
/* ------------------------------------------------------------------ */
struct om_model {
  int L, N, P, C, ac, ar, npair, nsl, nv, nv_rbm, mc, mr;
  double J1, J2;
  int sp;
  int* wmap;           /* [P][N] */
  cplx* chi;           /* [P] */
  int* orb;            /* [N][N] */
  unsigned char* subB; /* [N] */
  double sp_cos[2], sp_sin[2], sp_cc[2], sp_cs[2], sp_ss[2], sp_w[2];
  cplx *Wc, *bc;
  double *Wr, *br, *f;
  double* se; /* [C][2N][2N] column-major per copy */
};

struct om_state {
  int *x, *ele_idx, *ele_cfg, *sgn;
  cplx* thetac;
  double* thetar;
  double *pf, *invm;
  cplx *psi_wo, *psi;
  /* scratch */
  cplx* thc2;
  double* thr2;
  double *pf2, *invm2, *bufA, *bufB, *va, *vb, *vp, *vq, *vs, *vt;
  int *tmp_idx, *tmp_cfg, *sgn2, *x2;
  cplx *psi_wo2, *psi2;
};

#define SE(m, q, i, j) ((m)->se[(size_t)(q) * 4 * (m)->N * (m)->N + (size_t)(j) * 2 * (m)->N + (i)])
#define IM(s, m, q, i, j) ((s)[(size_t)(q) * (m)->N * (m)->N + (size_t)(j) * (m)->N + (i)])

int om_bond(int L, int type, int i1) {
  /* calc_phys_quantity.f90:382-427 (bond_map), 0-based type and site */
  int ix = i1 / L, iy = i1 - ix * L;
  if (type == 0) {
    ix++; if (ix >= L) ix -= L;
  } else if (type == 1) {
    iy++; if (iy >= L) iy -= L;
  } else if (type == 2) {
    ix++; if (ix >= L) ix -= L;
    iy++; if (iy >= L) iy -= L;
  } else {
    ix++; if (ix >= L) ix -= L;
    iy--; if (iy < 0) iy += L;
  }
  return ix * L + iy;
}

/* spin_projection.f90:5-79 (InitSPWeight) with N_SP_Gauss_Leg = 2:
 * beta = {0, pi}; cos/sin of beta/2 and products; weights (1, spin_parity).
 * proj = 1: the transverse_code variant (transverse_code/spin_projection.f90:
 * 30-38), beta = {-pi/2, pi/2} and sqrt(2) cos/sin of beta/2. */
static void init_sp_weight(om_model* m, int proj) {
  const double pi = acos(-1.0);
  const double beta[2] = {proj ? -0.5 * pi : 0.0, proj ? 0.5 * pi : pi};
  const double amp = proj ? sqrt(2.0) : 1.0;
  for (int s = 0; s < 2; s++) {
    m->sp_cos[s] = proj ? amp * cos(0.5 * beta[s]) : cos(0.5 * beta[s]);
    m->sp_sin[s] = proj ? amp * sin(0.5 * beta[s]) : sin(0.5 * beta[s]);
    m->sp_cc[s] = m->sp_cos[s] * m->sp_cos[s];
    m->sp_cs[s] = m->sp_cos[s] * m->sp_sin[s];
    m->sp_ss[s] = m->sp_sin[s] * m->sp_sin[s];
  }
  m->sp_w[0] = 1.0;
  m->sp_w[1] = (m->sp == 1) ? 1.0 : -1.0;
}

void om_sp_weights(const om_model* m, double* out) {
  for (int s = 0; s < 2; s++) {
    out[6 * s + 0] = m->sp_cos[s];
    out[6 * s + 1] = m->sp_sin[s];
    out[6 * s + 2] = m->sp_cc[s];
    out[6 * s + 3] = m->sp_cs[s];
    out[6 * s + 4] = m->sp_ss[s];
    out[6 * s + 5] = m->sp_w[s];
  }
}

om_model* om_create(int L, int P, int alphac, int alphar, int n_slater, double J1,
                    double J2, int spin_parity, const int* w_map, const double* chi,
                    const int* orb_idx) {
  om_model* m = (om_model*)calloc(1, sizeof(om_model));
  int N = L * L;
  m->L = L; m->N = N; m->P = P; m->C = 2 * P;
  m->ac = alphac; m->ar = alphar; m->npair = N / 2; m->nsl = n_slater;
  /* main.f90:134-137 */
  m->mc = alphac * P * 2;
  m->mr = alphar * P * 2;
  m->nv = alphac * 2 * (N + 1) + alphar * (N + 1) + n_slater;
  m->nv_rbm = m->nv - n_slater;
  m->J1 = J1; m->J2 = J2; m->sp = spin_parity;
  m->wmap = (int*)malloc(sizeof(int) * P * N);
  memcpy(m->wmap, w_map, sizeof(int) * P * N);
  m->chi = (cplx*)malloc(sizeof(cplx) * P);
  for (int g = 0; g < P; g++) m->chi[g] = CMPLX(chi[2 * g], chi[2 * g + 1]);
  m->orb = (int*)malloc(sizeof(int) * N * N);
  memcpy(m->orb, orb_idx, sizeof(int) * N * N);
  /* interface_for_PP.f90:110-136 (define_sublattice) */
  m->subB = (unsigned char*)malloc(N);
  for (int s = 0; s < N; s++) m->subB[s] = (unsigned char)(((s / L) + (s % L)) & 1);
  init_sp_weight(m, 0);
  m->Wc = (cplx*)calloc((size_t)alphac * N + 1, sizeof(cplx));
  m->bc = (cplx*)calloc((size_t)alphac + 1, sizeof(cplx));
  m->Wr = (double*)calloc((size_t)alphar * N + 1, sizeof(double));
  m->br = (double*)calloc((size_t)alphar + 1, sizeof(double));
  m->f = (double*)calloc((size_t)n_slater, sizeof(double));
  m->se = (double*)calloc((size_t)m->C * 4 * N * N, sizeof(double));
  return m;
}

void om_free(om_model* m) {
  if (!m) return;
  free(m->wmap); free(m->chi); free(m->orb); free(m->subB);
  free(m->Wc); free(m->bc); free(m->Wr); free(m->br); free(m->f); free(m->se);
  free(m);
}
int om_nv(const om_model* m) { return m->nv; }
int om_nv_rbm(const om_model* m) { return m->nv_rbm; }
const double* om_slater_elm(const om_model* m) { return m->se; }

/* slater.f90:7-64 (update_slater_elements).  Returns -1 when the result is
 * not skew-symmetric to 1e-10 (slater.f90:51-56). */
static int update_slater_elements(om_model* m) {
  int N = m->N;
  for (int g = 0; g < m->P; g++) {
    for (int sidx = 0; sidx < 2; sidx++) {
      int q = 2 * g + sidx;
      double cs = m->sp_cs[sidx], cc = m->sp_cc[sidx], ss = m->sp_ss[sidx];
      for (int ri = 0; ri < N; ri++) {
        int tri = m->wmap[g * N + ri];
        for (int rj = 0; rj < N; rj++) {
          int trj = m->wmap[g * N + rj];
          int i1 = m->orb[tri * N + trj];
          int i2 = m->orb[trj * N + tri];
          double slt_ij = m->f[i1] * 1.0;
          double slt_ji = m->f[i2] * 1.0;
          SE(m, q, rj, ri) = -(slt_ij - slt_ji) * cs;
          SE(m, q, rj + N, ri) = slt_ij * cc + slt_ji * ss;
          SE(m, q, rj, ri + N) = -slt_ij * ss - slt_ji * cc;
          SE(m, q, rj + N, ri + N) = (slt_ij - slt_ji) * cs;
        }
      }
      for (int i2 = 0; i2 < 2 * N; i2++)
        for (int i1 = 0; i1 < 2 * N; i1++)
          if (fabs(SE(m, q, i1, i2) + SE(m, q, i2, i1)) > 1.0e-10) return -1;
    }
  }
  return 0;
}

int om_set_params(om_model* m, const double* Wc, const double* bc, const double* Wr,
                  const double* br, const double* f) {
  int N = m->N;
  for (int k = 0; k < m->ac * N; k++) m->Wc[k] = CMPLX(Wc[2 * k], Wc[2 * k + 1]);
  for (int k = 0; k < m->ac; k++) m->bc[k] = CMPLX(bc[2 * k], bc[2 * k + 1]);
  if (m->ar) {
    memcpy(m->Wr, Wr, sizeof(double) * m->ar * N);
    memcpy(m->br, br, sizeof(double) * m->ar);
  }
  memcpy(m->f, f, sizeof(double) * m->nsl);
  return update_slater_elements(m);
}

/* Parameter vector gamma in the order of main.f90:263-294. */
int om_get_gamma(const om_model* m, double* gam) {
  int k = 0, N = m->N;
  for (int f = 0; f < m->ac; f++) { gam[k++] = creal(m->bc[f]); gam[k++] = cimag(m->bc[f]); }
  for (int f = 0; f < m->ac; f++)
    for (int iw = 0; iw < N; iw++) {
      gam[k++] = creal(m->Wc[f * N + iw]);
      gam[k++] = cimag(m->Wc[f * N + iw]);
    }
  for (int f = 0; f < m->ar; f++) gam[k++] = m->br[f];
  for (int f = 0; f < m->ar; f++)
    for (int iw = 0; iw < N; iw++) gam[k++] = m->Wr[f * N + iw];
  for (int o = 0; o < m->nsl; o++) gam[k++] = m->f[o];
  return k;
}

/* Spin-projection variant: 0 normal_code, 1 transverse_code (the weights of
 * InitSPWeight and the Slater_Elm built from them). */
int om_set_projection(om_model* m, int proj) {
  init_sp_weight(m, proj);
  return update_slater_elements(m);
}

/* SR_pdposv.f90:385-412 (unpacking gamma) */
int om_set_gamma(om_model* m, const double* gam) {
  int k = 0, N = m->N;
  for (int f = 0; f < m->ac; f++) { m->bc[f] = CMPLX(gam[k], gam[k + 1]); k += 2; }
  for (int f = 0; f < m->ac; f++)
    for (int iw = 0; iw < N; iw++) { m->Wc[f * N + iw] = CMPLX(gam[k], gam[k + 1]); k += 2; }
  for (int f = 0; f < m->ar; f++) m->br[f] = gam[k++];
  for (int f = 0; f < m->ar; f++)
    for (int iw = 0; iw < N; iw++) m->Wr[f * N + iw] = gam[k++];
  for (int o = 0; o < m->nsl; o++) m->f[o] = gam[k++];
  return update_slater_elements(m);
}

/* ------------------------------------------------------------------ */
/* RBM angles                                                           */
/* ------------------------------------------------------------------ */
/* RBM.f90:31-91 (calc_theta) */
void om_calc_theta(const om_model* m, const int* x, double* thc, double* thr) {
  int N = m->N;
  cplx* tc = (cplx*)thc;
  for (int g = 0; g < m->P; g++) {
    int j = g * 2 * m->ac;
    for (int s = 1; s >= -1; s -= 2)
      for (int f = 0; f < m->ac; f++, j++) {
        cplx t = m->bc[f];
        for (int i = 0; i < N; i++) {
          int iw = m->wmap[g * N + i];
          t = t + m->Wc[f * N + iw] * (double)x[i] * (double)s;
        }
        tc[j] = t;
      }
  }
  for (int g = 0; g < m->P; g++) {
    int j = g * 2 * m->ar;
    for (int s = 1; s >= -1; s -= 2)
      for (int f = 0; f < m->ar; f++, j++) {
        double t = m->br[f];
        for (int i = 0; i < N; i++) {
          int iw = m->wmap[g * N + i];
          t = t + m->Wr[f * N + iw] * (double)x[i] * (double)s;
        }
        thr[j] = t;
      }
  }
}

/* RBM.f90:95-148 (update_theta): sigma is the configuration before the move */
static void update_theta(const om_model* m, const int* x, int i1, int i2, cplx* tc,
                         double* tr) {
  int N = m->N;
  for (int g = 0; g < m->P; g++) {
    int j = g * 2 * m->ac;
    int iw1 = m->wmap[g * N + i1], iw2 = m->wmap[g * N + i2];
    for (int s = 1; s >= -1; s -= 2)
      for (int f = 0; f < m->ac; f++, j++) {
        tc[j] = tc[j] - 2.0 * m->Wc[f * N + iw1] * (double)x[i1] * (double)s;
        tc[j] = tc[j] - 2.0 * m->Wc[f * N + iw2] * (double)x[i2] * (double)s;
      }
  }
  for (int g = 0; g < m->P; g++) {
    int j = g * 2 * m->ar;
    int iw1 = m->wmap[g * N + i1], iw2 = m->wmap[g * N + i2];
    for (int s = 1; s >= -1; s -= 2)
      for (int f = 0; f < m->ar; f++, j++) {
        tr[j] = tr[j] - 2.0 * m->Wr[f * N + iw1] * (double)x[i1] * (double)s;
        tr[j] = tr[j] - 2.0 * m->Wr[f * N + iw2] * (double)x[i2] * (double)s;
      }
  }
}

/* ------------------------------------------------------------------ */
/* electron bookkeeping and fermion sign                                */
/* ------------------------------------------------------------------ */
/* interface_for_PP.f90:5-33 (x_to_Ele_Idx_Ele_Cfg) */
static int x_to_ele(const om_model* m, const int* x, int* idx, int* cfg) {
  int i1 = 0, i2 = m->npair;
  for (int i = 0; i < m->N; i++) {
    if (x[i] == 1) { idx[i1] = i; cfg[i] = i1; i1++; }
    else if (x[i] == -1) { idx[i2] = i; cfg[i] = i2; i2++; }
    else return -1;
  }
  if (i1 != m->npair || i2 != m->N) return -2;
  return 0;
}

/* interface_for_PP.f90:71-106 (sign_PP) */
void om_sign_pp(const om_model* m, const int* ele_idx, int* sgn) {
  int N = m->N;
  int* tmp = (int*)malloc(sizeof(int) * N);
  for (int g = 0; g < m->P; g++) {
    for (int i = 0; i < N; i++) tmp[i] = m->wmap[g * N + ele_idx[i]];
    int s = 1;
    for (int i1 = 1; i1 < N; i1++) {
      if (i1 >= m->npair && m->subB[tmp[i1]]) s = -s;
      for (int i2 = 0; i2 < i1; i2++)
        if (tmp[i2] > tmp[i1]) s = -s;
    }
    sgn[g] = s;
  }
  free(tmp);
}

/* ------------------------------------------------------------------ */
/* Pfaffian and inverse                                                 */
/* ------------------------------------------------------------------ */
/* Pfaffian of a skew-symmetric matrix by the Parlett-Reid LTL^T reduction
 * with partial pivoting (M. Wimmer, ACM TOMS 38, 30 (2012), Alg. 923; the
 * reference calls PFAPACK dskpfa('U','P'), matrix.f90:133-151). */
double om_pfaffian(int n, const double* A0) {
  if (n % 2) return 0.0;
  double* A = (double*)malloc(sizeof(double) * n * n);
  double* tau = (double*)malloc(sizeof(double) * (n + 1));
  memcpy(A, A0, sizeof(double) * n * n);
#define A_(i, j) A[(size_t)(j) * n + (i)]
  double pf = 1.0;
  for (int k = 0; k < n - 1; k += 2) {
    int kp = k + 1;
    double best = fabs(A_(k + 1, k));
    for (int i = k + 2; i < n; i++)
      if (fabs(A_(i, k)) > best) { best = fabs(A_(i, k)); kp = i; }
    if (kp != k + 1) {
      for (int j = k; j < n; j++) { double t = A_(k + 1, j); A_(k + 1, j) = A_(kp, j); A_(kp, j) = t; }
      for (int i = k; i < n; i++) { double t = A_(i, k + 1); A_(i, k + 1) = A_(i, kp); A_(i, kp) = t; }
      pf = -pf;
    }
    if (A_(k + 1, k) != 0.0) {
      pf *= A_(k, k + 1);
      if (k + 2 < n) {
        for (int i = k + 2; i < n; i++) tau[i] = A_(k, i) / A_(k, k + 1);
        for (int j = k + 2; j < n; j++)
          for (int i = k + 2; i < n; i++) {
            A_(i, j) = A_(i, j) + tau[i] * A_(j, k + 1);
            A_(i, j) = A_(i, j) - A_(i, k + 1) * tau[j];
          }
      }
    } else {
      pf = 0.0;
      break;
    }
  }
#undef A_
  free(A); free(tau);
  return pf;
}

/* Inverse by LU with partial pivoting (the reference uses LAPACK
 * dgetrf+dgetri, matrix.f90:155-177).  Returns -1 if singular. */
int om_inverse(int n, const double* A0, double* X) {
  double* A = (double*)malloc(sizeof(double) * n * n);
  int* piv = (int*)malloc(sizeof(int) * n);
  memcpy(A, A0, sizeof(double) * n * n);
#define A_(i, j) A[(size_t)(j) * n + (i)]
  int rc = 0;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = fabs(A_(k, k));
    for (int i = k + 1; i < n; i++)
      if (fabs(A_(i, k)) > best) { best = fabs(A_(i, k)); p = i; }
    piv[k] = p;
    if (best == 0.0) { rc = -1; break; }
    if (p != k)
      for (int j = 0; j < n; j++) { double t = A_(k, j); A_(k, j) = A_(p, j); A_(p, j) = t; }
    double d = A_(k, k);
    for (int i = k + 1; i < n; i++) A_(i, k) /= d;
    for (int j = k + 1; j < n; j++) {
      double akj = A_(k, j);
      if (akj != 0.0)
        for (int i = k + 1; i < n; i++) A_(i, j) -= A_(i, k) * akj;
    }
  }
  if (rc == 0) {
    /* solve A X = I column by column with the recorded row swaps */
    for (int c = 0; c < n; c++) {
      double* x = X + (size_t)c * n;
      for (int i = 0; i < n; i++) x[i] = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < n; k++)
        if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
      for (int i = 0; i < n; i++) {
        double s = x[i];
        for (int k = 0; k < i; k++) s -= A_(i, k) * x[k];
        x[i] = s;
      }
      for (int i = n - 1; i >= 0; i--) {
        double s = x[i];
        for (int k = i + 1; k < n; k++) s -= A_(i, k) * x[k];
        x[i] = s / A_(i, i);
      }
    }
  }
#undef A_
  free(A); free(piv);
  return rc;
}

/* matrix.f90:6-110 (calculate_MAll + _child): X(i,j) = -Slater_Elm(r_i,r_j,q),
 * Pf and the antisymmetrised (transposed) inverse.  Returns <0 on a singular
 * X or an inverse that is not skew-symmetric (matrix.f90:101-104). */
static int calculate_mall(const om_model* m, const int* ele_idx, double* pf, double* invm,
                          double* bufA, double* bufB) {
  int N = m->N, np = m->npair;
  for (int q = 0; q < m->C; q++) {
    for (int j = 0; j < N; j++) {
      int rsj = ele_idx[j] + (j / np) * N;
      for (int i = 0; i < N; i++) {
        int rsi = ele_idx[i] + (i / np) * N;
        bufA[(size_t)j * N + i] = -SE(m, q, rsi, rsj);
      }
    }
    pf[q] = om_pfaffian(N, bufA);
    if (om_inverse(N, bufA, bufB) != 0) return -10;
    double r1 = 0.0;
    for (int k = 0; k < N * N; k++)
      if (fabs(bufB[k]) > r1) r1 = fabs(bufB[k]);
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        double r2 = 0.5 * (bufB[(size_t)i * N + j] - bufB[(size_t)j * N + i]);
        double r3 = 0.5 * (bufB[(size_t)i * N + j] + bufB[(size_t)j * N + i]);
        if (fabs(r3 / r1) > 2.0e-3) return -11;
        IM(invm, m, q, i, j) = r2;
      }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* amplitude                                                            */
/* ------------------------------------------------------------------ */
/* sub.f90:4-89 (calc_amplitude_RBM_PP).  When r is an mt.f stream, one draw
 * is consumed for the reference's probabilistic self-check (sub.f90:72). */
static void calc_amplitude(const om_model* m, const cplx* thc, const double* thr,
                           const double* pf, const int* sgn, cplx* psi_wo, cplx* psi,
                           om_rng* r) {
  for (int q = 0; q < m->C; q++) {
    int g = q / 2, sidx = q % 2;
    cplx c1 = 1.0;
    for (int k = 0; k < m->ac; k++) c1 = c1 * 0.5;
    for (int f = 0; f < m->ac; f++) {
      cplx t = thc[q * m->ac + f];
      double re = creal(t), im = cimag(t);
      c1 = c1 * CMPLX(cosh(re) * cos(im), sinh(re) * sin(im));
    }
    double r1 = 1.0;
    for (int f = 0; f < m->ar; f++) r1 = r1 * cosh(thr[q * m->ar + f]);
    c1 = c1 * r1 * (double)sgn[g] * m->chi[g] * m->sp_w[sidx];
    psi_wo[q] = c1;
    psi[q + 1] = c1 * pf[q];
  }
  psi[0] = 0.0;
  for (int q = 0; q < m->C; q++) psi[0] = psi[0] + psi[q + 1];
  if (r && r->kind == 0) (void)mt_grnd(r);
}

/* ------------------------------------------------------------------ */
/* state                                                                */
/* ------------------------------------------------------------------ */
om_state* om_state_create(const om_model* m) {
  int N = m->N, C = m->C;
  om_state* s = (om_state*)calloc(1, sizeof(om_state));
  s->x = (int*)calloc(N, sizeof(int));
  s->x2 = (int*)calloc(N, sizeof(int));
  s->ele_idx = (int*)calloc(N, sizeof(int));
  s->ele_cfg = (int*)calloc(N, sizeof(int));
  s->tmp_idx = (int*)calloc(N, sizeof(int));
  s->tmp_cfg = (int*)calloc(N, sizeof(int));
  s->sgn = (int*)calloc(m->P, sizeof(int));
  s->sgn2 = (int*)calloc(m->P, sizeof(int));
  s->thetac = (cplx*)calloc(m->mc + 1, sizeof(cplx));
  s->thc2 = (cplx*)calloc(m->mc + 1, sizeof(cplx));
  s->thetar = (double*)calloc(m->mr + 1, sizeof(double));
  s->thr2 = (double*)calloc(m->mr + 1, sizeof(double));
  s->pf = (double*)calloc(C, sizeof(double));
  s->pf2 = (double*)calloc(C, sizeof(double));
  s->invm = (double*)calloc((size_t)C * N * N, sizeof(double));
  s->invm2 = (double*)calloc((size_t)C * N * N, sizeof(double));
  s->bufA = (double*)calloc((size_t)N * N, sizeof(double));
  s->bufB = (double*)calloc((size_t)N * N, sizeof(double));
  s->va = (double*)calloc(N, sizeof(double));
  s->vb = (double*)calloc(N, sizeof(double));
  s->vp = (double*)calloc(N, sizeof(double));
  s->vq = (double*)calloc(N, sizeof(double));
  s->vs = (double*)calloc(N, sizeof(double));
  s->vt = (double*)calloc(N, sizeof(double));
  s->psi_wo = (cplx*)calloc(C, sizeof(cplx));
  s->psi_wo2 = (cplx*)calloc(C, sizeof(cplx));
  s->psi = (cplx*)calloc(C + 1, sizeof(cplx));
  s->psi2 = (cplx*)calloc(C + 1, sizeof(cplx));
  return s;
}

void om_state_free(om_state* s) {
  if (!s) return;
  free(s->x); free(s->x2); free(s->ele_idx); free(s->ele_cfg); free(s->tmp_idx);
  free(s->tmp_cfg); free(s->sgn); free(s->sgn2); free(s->thetac); free(s->thc2);
  free(s->thetar); free(s->thr2); free(s->pf); free(s->pf2); free(s->invm);
  free(s->invm2); free(s->bufA); free(s->bufB); free(s->va); free(s->vb); free(s->vp);
  free(s->vq); free(s->vs); free(s->vt); free(s->psi_wo); free(s->psi_wo2);
  free(s->psi); free(s->psi2);
  free(s);
}

/* Full evaluation of a configuration: calc_theta, x_to_Ele_Idx_Ele_Cfg,
 * calculate_MAll, sign_PP, calc_amplitude_RBM_PP (make_sample.f90:55-60). */
int om_eval(const om_model* m, const int* x, om_state* s, om_rng* r) {
  memcpy(s->x, x, sizeof(int) * m->N);
  om_calc_theta(m, s->x, (double*)s->thetac, s->thetar);
  int rc = x_to_ele(m, s->x, s->ele_idx, s->ele_cfg);
  if (rc) return rc;
  rc = calculate_mall(m, s->ele_idx, s->pf, s->invm, s->bufA, s->bufB);
  if (rc) return rc;
  om_sign_pp(m, s->ele_idx, s->sgn);
  calc_amplitude(m, s->thetac, s->thetar, s->pf, s->sgn, s->psi_wo, s->psi, r);
  return 0;
}

int om_state_get(const om_model* m, const om_state* s, int* x, double* thc, double* thr,
                 int* ele_idx, int* ele_cfg, int* sgn, double* pf, double* invm,
                 double* psi_wo, double* psi) {
  int N = m->N, C = m->C;
  if (x) memcpy(x, s->x, sizeof(int) * N);
  if (thc) memcpy(thc, s->thetac, sizeof(cplx) * m->mc);
  if (thr) memcpy(thr, s->thetar, sizeof(double) * m->mr);
  if (ele_idx) memcpy(ele_idx, s->ele_idx, sizeof(int) * N);
  if (ele_cfg) memcpy(ele_cfg, s->ele_cfg, sizeof(int) * N);
  if (sgn) memcpy(sgn, s->sgn, sizeof(int) * m->P);
  if (pf) memcpy(pf, s->pf, sizeof(double) * C);
  if (invm) memcpy(invm, s->invm, sizeof(double) * (size_t)C * N * N);
  if (psi_wo) memcpy(psi_wo, s->psi_wo, sizeof(cplx) * C);
  if (psi) memcpy(psi, s->psi, sizeof(cplx) * (C + 1));
  return 0;
}

/* ------------------------------------------------------------------ */
/* fast two-electron updates                                            */
/* ------------------------------------------------------------------ */
/* pfupdate_two.f90:27-88 (calculate_new_PfM_two_child) */
static double pfm_two_ratio(const om_model* m, int q, int msa, int msb, const int* tmp,
                            const double* invm, double* va, double* vb) {
  int N = m->N, np = m->npair;
  int rsa = tmp[msa] + (msa / np) * N;
  int rsb = tmp[msb] + (msb / np) * N;
  for (int i = 0; i < N; i++) {
    int rsi = tmp[i] + (i / np) * N;
    va[i] = SE(m, q, rsi, rsa);
    vb[i] = SE(m, q, rsi, rsb);
  }
  double vec_ba = vb[msa];
  double invm_ab = IM(invm, m, q, msb, msa);
  double p_a = 0, p_b = 0, q_a = 0, q_b = 0, bMa = 0;
  for (int i = 0; i < N; i++) {
    p_a = p_a + IM(invm, m, q, i, msa) * va[i];
    p_b = p_b + IM(invm, m, q, i, msb) * va[i];
    q_a = q_a + IM(invm, m, q, i, msa) * vb[i];
    q_b = q_b + IM(invm, m, q, i, msb) * vb[i];
  }
  for (int i = 0; i < N; i++) {
    double rtmp = 0.0;
    for (int j = 0; j < N; j++) rtmp = rtmp + IM(invm, m, q, j, i) * va[j];
    bMa = bMa + vb[i] * rtmp;
  }
  return invm_ab * vec_ba + invm_ab * bMa + p_a * q_b - p_b * q_a;
}

/* pfupdate_two.f90:113-224 (calculate_new_InvM_two_child); ele_idx is the
 * configuration after the move. */
static void invm_two_update(const om_model* m, int q, int msa, int msb, const int* ele_idx,
                            double* invm, double* vS, double* vT, double* vP, double* vQ) {
  int N = m->N, np = m->npair;
  int rsa = ele_idx[msa] + (msa / np) * N;
  int rsb = ele_idx[msb] + (msb / np) * N;
  for (int i = 0; i < N; i++) {
    int rsi = ele_idx[i] + (i / np) * N;
    vS[i] = SE(m, q, rsi, rsa);
    vT[i] = SE(m, q, rsi, rsb);
  }
  for (int i = 0; i < N; i++) {
    double p = 0.0, qq = 0.0;
    for (int j = 0; j < N; j++) {
      p = p + IM(invm, m, q, j, i) * vS[j];
      qq = qq + IM(invm, m, q, j, i) * vT[j];
    }
    vP[i] = p;
    vQ[i] = qq;
  }
  double bMa = 0.0;
  for (int i = 0; i < N; i++) bMa = bMa + vT[i] * vP[i];
  double a = -vP[msa], b = vP[msb], c = vQ[msa], d = -vQ[msb];
  double e = -bMa - vT[msa];
  double f = IM(invm, m, q, msb, msa);
  double Det = a * d - b * c - e * f;
  double InvDet = 1.0 / Det;
  for (int i = 0; i < N; i++) {
    vS[i] = InvDet * IM(invm, m, q, i, msa);
    vT[i] = InvDet * IM(invm, m, q, i, msb);
  }
  for (int i = 0; i < N; i++) {
    double p_i = vP[i], q_i = vQ[i], s_i = vS[i], t_i = vT[i];
    for (int j = 0; j < N; j++) {
      double p_j = vP[j], q_j = vQ[j], s_j = vS[j], t_j = vT[j];
      IM(invm, m, q, j, i) = IM(invm, m, q, j, i) + a * (q_i * t_j - q_j * t_i) +
                             b * (q_i * s_j - q_j * s_i) + c * (p_i * t_j - p_j * t_i) +
                             d * (p_i * s_j - p_j * s_i) + e * Det * (s_i * t_j - s_j * t_i) +
                             f * InvDet * (p_i * q_j - q_i * p_j);
    }
    IM(invm, m, q, msa, i) = IM(invm, m, q, msa, i) - c * t_i - d * s_i - f * InvDet * q_i;
    IM(invm, m, q, msb, i) = IM(invm, m, q, msb, i) - a * t_i - b * s_i + f * InvDet * p_i;
  }
  for (int i = 0; i < N; i++) {
    double p_i = vP[i], q_i = vQ[i], s_i = vS[i], t_i = vT[i];
    IM(invm, m, q, i, msa) = IM(invm, m, q, i, msa) + c * t_i + d * s_i + f * InvDet * q_i;
    IM(invm, m, q, i, msb) = IM(invm, m, q, i, msb) + a * t_i + b * s_i - f * InvDet * p_i;
  }
  IM(invm, m, q, msa, msb) = IM(invm, m, q, msa, msb) - f * InvDet;
  IM(invm, m, q, msb, msa) = IM(invm, m, q, msb, msa) + f * InvDet;
}

/* calculate_new_PfM_two and calculate_new_InvM_two on caller arrays, with the
 * reference's arguments (pfupdate_two.f90:5-25, 93-111): 0-based electrons
 * msa, msb, the electron-to-site list after the move, InvM of every copy in
 * the reference's column-major layout.  For timing the restatement against
 * the compiled reference on identical inputs. */
void om_pfm_two(const om_model* m, int msa, int msb, const int* tmp_idx, const double* invm,
                const double* pf_old, double* pf_new) {
  double va[m->N], vb[m->N];
  for (int q = 0; q < m->C; q++) pf_new[q] = pfm_two_ratio(m, q, msa, msb, tmp_idx, invm, va, vb) * pf_old[q];
}

void om_invm_two(const om_model* m, int msa, int msb, const int* ele_idx, double* invm) {
  double vs[m->N], vt[m->N], vp[m->N], vq[m->N];
  for (int q = 0; q < m->C; q++) invm_two_update(m, q, msa, msb, ele_idx, invm, vs, vt, vp, vq);
}

/* Proposed exchange of the spins at sites i1, i2 (must differ): ratio per copy,
 * new Pfaffians and new amplitude (make_sample.f90:167-198).  Leaves the
 * proposal in the state's scratch (thc2, thr2, tmp_idx, pf2, sgn2, psi2). */
static void propose_exchange(const om_model* m, om_state* s, int i1, int i2, double* ratio,
                             om_rng* r) {
  memcpy(s->thc2, s->thetac, sizeof(cplx) * m->mc);
  memcpy(s->thr2, s->thetar, sizeof(double) * m->mr);
  update_theta(m, s->x, i1, i2, s->thc2, s->thr2);
  int i3 = s->ele_cfg[i1], i4 = s->ele_cfg[i2];
  memcpy(s->tmp_idx, s->ele_idx, sizeof(int) * m->N);
  s->tmp_idx[i3] = i2;
  s->tmp_idx[i4] = i1;
  for (int q = 0; q < m->C; q++) {
    double rt = pfm_two_ratio(m, q, i3, i4, s->tmp_idx, s->invm, s->va, s->vb);
    if (ratio) ratio[q] = rt;
    s->pf2[q] = rt * s->pf[q];
  }
  /* make_sample.f90:195-197: sign rule for an exchange */
  for (int g = 0; g < m->P; g++) {
    int t = -s->sgn[g];
    if (m->subB[i1]) t = -t;
    if (m->subB[i2]) t = -t;
    s->sgn2[g] = t;
  }
  calc_amplitude(m, s->thc2, s->thr2, s->pf2, s->sgn2, s->psi_wo2, s->psi2, r);
}

int om_ratio_exchange(const om_model* m, const om_state* s0, int i1, int i2, double* ratio,
                      double* pf_new, double* psi_new) {
  om_state* s = (om_state*)s0; /* only scratch fields are written */
  if (s->x[i1] == s->x[i2]) return -1;
  propose_exchange(m, s, i1, i2, ratio, NULL);
  if (pf_new) memcpy(pf_new, s->pf2, sizeof(double) * m->C);
  if (psi_new) memcpy(psi_new, s->psi2, sizeof(cplx) * (m->C + 1));
  return 0;
}

/* Accept branch of make_sample.f90:204-219 */
static void accept_exchange(const om_model* m, om_state* s, int i1, int i2) {
  int i3 = s->ele_cfg[i1], i4 = s->ele_cfg[i2];
  s->x[i1] = -s->x[i1];
  s->x[i2] = -s->x[i2];
  memcpy(s->thetac, s->thc2, sizeof(cplx) * m->mc);
  memcpy(s->thetar, s->thr2, sizeof(double) * m->mr);
  s->ele_cfg[i1] = i4;
  s->ele_cfg[i2] = i3;
  s->ele_idx[i3] = i2;
  s->ele_idx[i4] = i1;
  memcpy(s->pf, s->pf2, sizeof(double) * m->C);
  for (int q = 0; q < m->C; q++)
    invm_two_update(m, q, i3, i4, s->ele_idx, s->invm, s->vs, s->vt, s->vp, s->vq);
  memcpy(s->sgn, s->sgn2, sizeof(int) * m->P);
  memcpy(s->psi_wo, s->psi_wo2, sizeof(cplx) * m->C);
  memcpy(s->psi, s->psi2, sizeof(cplx) * (m->C + 1));
}

int om_accept_exchange(const om_model* m, om_state* s, int i1, int i2, om_rng* r) {
  if (s->x[i1] == s->x[i2]) return -1;
  propose_exchange(m, s, i1, i2, NULL, r);
  accept_exchange(m, s, i1, i2);
  return 0;
}

/* ------------------------------------------------------------------ */
/* estimators                                                           */
/* ------------------------------------------------------------------ */
/* calc_phys_quantity.f90:306-378 (calc_local_energy_correlation).  E_loc in
 * the reference's units (Pauli sigma, i.e. 4 S.S), not yet reweighted. */
double om_eloc(const om_model* m, const om_state* s0, double* eloc_out, om_rng* r) {
  om_state* s = (om_state*)s0;
  cplx psi0 = s->psi[0];
  cplx e = 0.0;
  for (int t = 0; t < 4; t++) {
    double J = (t <= 1) ? m->J1 : m->J2;
    double sg = (t <= 1) ? -1.0 : 1.0;
    for (int i1 = 0; i1 < m->N; i1++) {
      int i2 = om_bond(m->L, t, i1);
      e = e + J * (double)s->x[i1] * (double)s->x[i2] * psi0;
      if (s->x[i1] != s->x[i2]) {
        propose_exchange(m, s, i1, i2, NULL, r);
        e = e + 2.0 * J * sg * s->psi2[0];
      }
    }
  }
  if (eloc_out) { eloc_out[0] = creal(e); eloc_out[1] = cimag(e); }
  return creal(e);
}

/* slater.f90:94-170 (calculate_derivative_Pfaffian_child) for one copy,
 * accumulated with weight psi_wo_pf(q) as calc_phys_quantity.f90:207-219. */
static void pp_derivative(const om_model* m, const om_state* s, int q, double* deriv) {
  int N = m->N, np = m->npair;
  int g = q / 2, sidx = q % 2;
  double cs = s->pf[q] * m->sp_cs[sidx];
  double cc = s->pf[q] * m->sp_cc[sidx];
  double ss = s->pf[q] * m->sp_ss[sidx];
  for (int o = 0; o < m->nsl; o++) deriv[o] = 0.0;
#define TOI(i, j) (m->orb[m->wmap[g * N + s->ele_idx[i]] * N + m->wmap[g * N + s->ele_idx[j]]])
  for (int msj = 0; msj < np; msj++)
    for (int msi = 0; msi < np; msi++) {
      int o;
      o = TOI(msi, msj);
      deriv[o] = deriv[o] - IM(s->invm, m, q, msi, msj) * cs * 1.0;
      o = TOI(msi, msj + np);
      deriv[o] = deriv[o] + IM(s->invm, m, q, msi, msj + np) * cc * 1.0;
      o = TOI(msi + np, msj);
      deriv[o] = deriv[o] - IM(s->invm, m, q, msi + np, msj) * ss * 1.0;
      o = TOI(msi + np, msj + np);
      deriv[o] = deriv[o] + IM(s->invm, m, q, msi + np, msj + np) * cs * 1.0;
    }
#undef TOI
}

/* calc_phys_quantity.f90:114-222: O_loc(k) = d psi / d gamma_k (not divided by
 * psi, not yet reweighted), complex, Nv entries. */
int om_oloc(const om_model* m, const om_state* s, double* oloc_out) {
  int N = m->N, nv = m->nv, C = m->C;
  cplx* o = (cplx*)oloc_out;
  const cplx ci = CMPLX(0.0, 1.0);
  for (int k = 0; k < nv; k++) o[k] = 0.0;
  /* complex biases */
  for (int f = 0; f < m->ac; f++) {
    int k1 = 2 * f;
    for (int q = 0; q < C; q++) {
      cplx t = s->thetac[q * m->ac + f];
      cplx c2 = CMPLX(tanh(creal(t)), tan(cimag(t)));
      cplx c3 = CMPLX(1.0, tanh(creal(t)) * tan(cimag(t)));
      o[k1] = o[k1] + (c2 / c3) * s->psi[q + 1];
    }
    o[k1 + 1] = o[k1] * ci;
  }
  /* complex W */
  for (int f = 0; f < m->ac; f++) {
    int off = 2 * m->ac + f * N * 2;
    for (int q = 0; q < C; q++) {
      int g = q / 2;
      double sp = (q % 2 == 0) ? 1.0 : -1.0;
      cplx t = s->thetac[q * m->ac + f];
      cplx c2 = CMPLX(tanh(creal(t)), tan(cimag(t)));
      cplx c3 = CMPLX(1.0, tanh(creal(t)) * tan(cimag(t)));
      cplx c1 = (c2 / c3) * s->psi[q + 1];
      for (int i = 0; i < N; i++) {
        int iw = m->wmap[g * N + i];
        int k1 = off + 2 * iw;
        o[k1] = o[k1] + c1 * (double)s->x[i] * sp;
        if (g == m->P - 1) o[k1 + 1] = o[k1] * ci;
      }
    }
  }
  /* real biases */
  for (int f = 0; f < m->ar; f++) {
    int k1 = 2 * m->ac * (N + 1) + f;
    for (int q = 0; q < C; q++) o[k1] = o[k1] + tanh(s->thetar[q * m->ar + f]) * s->psi[q + 1];
  }
  /* real W */
  for (int f = 0; f < m->ar; f++) {
    int off = 2 * m->ac * (N + 1) + m->ar + f * N;
    for (int q = 0; q < C; q++) {
      int g = q / 2;
      double sp = (q % 2 == 0) ? 1.0 : -1.0;
      cplx c1 = tanh(s->thetar[q * m->ar + f]) * s->psi[q + 1];
      for (int i = 0; i < N; i++) {
        int iw = m->wmap[g * N + i];
        o[off + iw] = o[off + iw] + c1 * (double)s->x[i] * sp;
      }
    }
  }
  /* pair-product part */
  int off = 2 * m->ac * (N + 1) + m->ar * (N + 1);
  double* deriv = (double*)malloc(sizeof(double) * m->nsl * C);
  for (int q = 0; q < C; q++) pp_derivative(m, s, q, deriv + (size_t)q * m->nsl);
  for (int orb = 0; orb < m->nsl; orb++)
    for (int q = 0; q < C; q++)
      o[off + orb] = o[off + orb] + deriv[(size_t)q * m->nsl + orb] * s->psi_wo[q];
  free(deriv);
  return 0;
}

/* calc_phys_quantity.f90:20-246 (calc_phys_quantity_opt): un-normalised sums
 * over the given samples.  sums = {Etot, Etot2, norm, wf_av}.  smat (Nv*Nv,
 * optional) is the raw <O O^T> of calc_phys_quantity.f90:240; ostore
 * (optional, [2*nsample][Nv]) holds the reweighted O rows (re, im). */
int om_measure(const om_model* m, int ns, const int* xs, double shift, int want_s,
               double* sums, double* gvec, double* ovec, double* smat, double* eloc_out,
               double* psi_out, double* ostore) {
  int N = m->N, nv = m->nv;
  om_state* s = om_state_create(m);
  cplx* oloc = (cplx*)malloc(sizeof(cplx) * nv);
  double* ost = ostore;
  int own = 0;
  if (want_s && !ost) { ost = (double*)malloc(sizeof(double) * 2 * (size_t)ns * nv); own = 1; }
  double Et = 0, Et2 = 0, nrm = 0, wf = 0;
  for (int k = 0; k < nv; k++) { gvec[k] = 0; ovec[k] = 0; }
  int rc = 0;
  for (int is = 0; is < ns; is++) {
    rc = om_eval(m, xs + (size_t)is * N, s, NULL);
    if (rc) break;
    cplx psi0 = s->psi[0];
    double rw = 1.0 / (cabs(psi0) + shift);
    double el[2];
    om_eloc(m, s, el, NULL);
    cplx e = CMPLX(el[0], el[1]);
    e = e * rw;
    Et = Et + rw * creal(conj(psi0) * e);
    Et2 = Et2 + creal(conj(e) * e);
    nrm = nrm + (cabs(psi0) * rw) * (cabs(psi0) * rw);
    wf = wf + cabs(psi0);
    om_oloc(m, s, (double*)oloc);
    for (int k = 0; k < nv; k++) oloc[k] = oloc[k] * rw;
    for (int k = 0; k < nv; k++) {
      gvec[k] = creal(gvec[k] + 2.0 * conj(e) * oloc[k]);
      ovec[k] = ovec[k] + rw * creal(conj(psi0) * oloc[k]);
      if (ost) {
        ost[(size_t)(2 * is) * nv + k] = creal(oloc[k]);
        ost[(size_t)(2 * is + 1) * nv + k] = cimag(oloc[k]);
      }
    }
    if (eloc_out) { eloc_out[2 * is] = creal(e); eloc_out[2 * is + 1] = cimag(e); }
    if (psi_out) { psi_out[2 * is] = creal(psi0); psi_out[2 * is + 1] = cimag(psi0); }
  }
  if (rc == 0 && want_s && smat) {
    for (int a = 0; a < nv; a++)
      for (int b = 0; b < nv; b++) smat[(size_t)b * nv + a] = 0.0;
    for (int row = 0; row < 2 * ns; row++) {
      const double* o = ost + (size_t)row * nv;
      for (int b = 0; b < nv; b++) {
        double ob = o[b];
        if (ob == 0.0) continue;
        double* col = smat + (size_t)b * nv;
        for (int a = 0; a < nv; a++) col[a] += o[a] * ob;
      }
    }
  }
  sums[0] = Et; sums[1] = Et2; sums[2] = nrm; sums[3] = wf;
  if (own) free(ost);
  free(oloc);
  om_state_free(s);
  return rc;
}

/* calc_phys_quantity.f90:250-302 (calc_phys_quantity_after_opt) */
int om_measure_energy(const om_model* m, int ns, const int* xs, double shift, double* sums) {
  int N = m->N;
  om_state* s = om_state_create(m);
  double Et = 0, Et2 = 0, nrm = 0, wf = 0;
  int rc = 0;
  for (int is = 0; is < ns; is++) {
    rc = om_eval(m, xs + (size_t)is * N, s, NULL);
    if (rc) break;
    cplx psi0 = s->psi[0];
    double rw = 1.0 / (cabs(psi0) + shift);
    double el[2];
    om_eloc(m, s, el, NULL);
    cplx e = CMPLX(el[0], el[1]) * rw;
    Et = Et + rw * creal(conj(psi0) * e);
    Et2 = Et2 + creal(conj(e) * e);
    nrm = nrm + (cabs(psi0) * rw) * (cabs(psi0) * rw);
    wf = wf + cabs(psi0);
  }
  sums[0] = Et; sums[1] = Et2; sums[2] = nrm; sums[3] = wf;
  om_state_free(s);
  return rc;
}

/* Calc_Correlation/calc_phys_quantity.f90:348-413 (its
 * calc_local_energy_correlation): the local energy together with the per-bond
 * SdotS(i1, t) = x_i1 x_i2 psi + [x_i1 /= x_i2] 2 sgn_t psi(x'), sgn_t = -1
 * for the J1 bonds and +1 for the J2 bonds.  sdots[2*(t*N + i1) + {0,1}]. */
double om_bond_sdots(const om_model* m, const om_state* s0, double* eloc_out, double* sdots) {
  om_state* s = (om_state*)s0;
  const int N = m->N;
  cplx psi0 = s->psi[0];
  cplx e = 0.0;
  for (int t = 0; t < 4; t++) {
    double J = (t <= 1) ? m->J1 : m->J2;
    double sg = (t <= 1) ? -1.0 : 1.0;
    for (int i1 = 0; i1 < N; i1++) {
      int i2 = om_bond(m->L, t, i1);
      cplx sd = (double)s->x[i1] * (double)s->x[i2] * psi0;
      e = e + J * (double)s->x[i1] * (double)s->x[i2] * psi0;
      if (s->x[i1] != s->x[i2]) {
        propose_exchange(m, s, i1, i2, NULL, NULL);
        e = e + 2.0 * J * sg * s->psi2[0];
        sd = sd + 2.0 * sg * s->psi2[0];
      }
      sdots[2 * (t * N + i1)] = creal(sd);
      sdots[2 * (t * N + i1) + 1] = cimag(sd);
    }
  }
  if (eloc_out) { eloc_out[0] = creal(e); eloc_out[1] = cimag(e); }
  return creal(e);
}

/* Calc_Correlation/calc_phys_quantity.f90:469-550 (calc_local_correlation):
 * for the pairs i1 <= i2, numbered i0 = i2 (i2 + 1) / 2 + i1 (i2 outer, as
 * the reference's loops), chi[4*i0 + {0,1}] = Chi_loc_x(1, i0) = x_i1 x_i2 psi
 * and chi[4*i0 + {2,3}] = Chi_loc_x(2, i0): psi on the diagonal, and
 * s_i1 s_i2 psi(x') (s = -1 on sublattice B) for an exchange of
 * antiparallel spins. */
int om_local_correlation(const om_model* m, const om_state* s0, double* chi) {
  om_state* s = (om_state*)s0;
  const int N = m->N;
  cplx psi0 = s->psi[0];
  size_t i0 = 0;
  for (int i2 = 0; i2 < N; i2++)
    for (int i1 = 0; i1 <= i2; i1++, i0++) {
      cplx zz = (double)s->x[i1] * (double)s->x[i2] * psi0;
      cplx pm = (i1 == i2) ? psi0 : 0.0;
      if (s->x[i1] != s->x[i2]) {
        propose_exchange(m, s, i1, i2, NULL, NULL);
        double sg = 1.0;
        if (m->subB[i1]) sg = -sg;
        if (m->subB[i2]) sg = -sg;
        pm = sg * s->psi2[0];
      }
      chi[4 * i0] = creal(zz); chi[4 * i0 + 1] = cimag(zz);
      chi[4 * i0 + 2] = creal(pm); chi[4 * i0 + 3] = cimag(pm);
    }
  return 0;
}

/* Calc_Correlation/calc_phys_quantity.f90:264-341 (calc_phys_quantity_after_opt
 * of the correlation run): un-normalised sums over the samples.
 * sums = {Etot, Etot2, norm, wf_av, Re wf_sgn, Im wf_sgn};
 * chi_sp[2*i0 + k] = Chi_sp(k+1, i0); chi_vbs[t*Nchi + i0] = Chi_vbs(i0, t+1). */
int om_measure_correlation(const om_model* m, int ns, const int* xs, double shift, double* sums,
                           double* chi_sp, double* chi_vbs) {
  const int N = m->N;
  const size_t nchi = (size_t)N * (N + 1) / 2;
  om_state* s = om_state_create(m);
  double* sd = (double*)malloc(sizeof(double) * 8 * N);
  double* cl = (double*)malloc(sizeof(double) * 4 * nchi);
  double Et = 0, Et2 = 0, nrm = 0, wf = 0;
  cplx wsg = 0.0;
  for (size_t k = 0; k < 2 * nchi; k++) chi_sp[k] = 0.0;
  for (size_t k = 0; k < 4 * nchi; k++) chi_vbs[k] = 0.0;
  int rc = 0;
  for (int is = 0; is < ns; is++) {
    rc = om_eval(m, xs + (size_t)is * N, s, NULL);
    if (rc) break;
    cplx psi0 = s->psi[0];
    double rw = 1.0 / (cabs(psi0) + shift);
    double el[2];
    om_bond_sdots(m, s, el, sd);
    cplx e = CMPLX(el[0], el[1]) * rw;
    Et = Et + rw * creal(conj(psi0) * e);
    Et2 = Et2 + creal(conj(e) * e);
    nrm = nrm + (cabs(psi0) * rw) * (cabs(psi0) * rw);
    wf = wf + cabs(psi0);
    wsg = wsg + psi0 / cabs(psi0);
    for (int k = 0; k < 8 * N; k++) sd[k] *= rw;
    for (int t = 0; t < 4; t++) {
      size_t i0 = 0;
      for (int i2 = 0; i2 < N; i2++)
        for (int i1 = 0; i1 <= i2; i1++, i0++) {
          cplx a = CMPLX(sd[2 * (t * N + i1)], sd[2 * (t * N + i1) + 1]);
          cplx b = CMPLX(sd[2 * (t * N + i2)], sd[2 * (t * N + i2) + 1]);
          chi_vbs[t * nchi + i0] += creal(conj(a) * b);
        }
    }
    om_local_correlation(m, s, cl);
    for (size_t i0 = 0; i0 < nchi; i0++)
      for (int k = 0; k < 2; k++) {
        cplx c = CMPLX(cl[4 * i0 + 2 * k], cl[4 * i0 + 2 * k + 1]);
        chi_sp[2 * i0 + k] += rw * rw * creal(conj(psi0) * c);
      }
  }
  sums[0] = Et; sums[1] = Et2; sums[2] = nrm; sums[3] = wf;
  sums[4] = creal(wsg); sums[5] = cimag(wsg);
  free(cl);
  free(sd);
  om_state_free(s);
  return rc;
}

/* ------------------------------------------------------------------ */
/* sampler                                                              */
/* ------------------------------------------------------------------ */
/* make_sample.f90:298-362 (make_initial_x) */
static int make_initial_x(const om_model* m, om_state* s, om_rng* r, om_rng* rd) {
  int N = m->N;
  for (int it = 0; it < 1000000; it++) {
    int cnt = 0, ok = 0;
    for (int i = 0; i < N; i++) s->x2[i] = -1;
    for (int k = 0; k < 1000000; k++) {
      int i = nran0(r, N);
      if (s->x2[i] == -1) { cnt++; s->x2[i] = 1; }
      if (cnt == N / 2) { ok = 1; break; }
    }
    if (!ok) return -20;
    int rc = om_eval(m, s->x2, s, rd);
    if (rc) return rc;
    if (cabs(s->psi[0]) > 1e-300) return 0;
  }
  return -21;
}

/* make_sample.f90:244-294 (make_spin_flip_candidate) */
static int spin_flip_candidate(const om_model* m, const int* x, om_rng* r, int* o1, int* o2) {
  int L = m->L, N = m->N;
  int dist_max = nran0(r, L - 1) + 2;
  for (int it = 0; it < 1000000; it++) {
    int i1 = nran0(r, N);
    int ix = i1 / L, iy = i1 - ix * L;
    ix = ix + nran0(r, dist_max);
    iy = iy + nran0(r, dist_max);
    if (ix >= L) ix -= L;
    if (iy >= L) iy -= L;
    int i2 = ix * L + iy;
    if (x[i1] != x[i2]) { *o1 = i1; *o2 = i2; return 0; }
  }
  return -30;
}

/* Full recompute of the state from s->x, as make_sample.f90:78-97 */
static int recompute(const om_model* m, om_state* s, om_rng* r, double* dpf, double* dinv) {
  int N = m->N, C = m->C;
  memcpy(s->thc2, s->thetac, sizeof(cplx) * m->mc);
  memcpy(s->thr2, s->thetar, sizeof(double) * m->mr);
  memcpy(s->sgn2, s->sgn, sizeof(int) * m->P);
  om_calc_theta(m, s->x, (double*)s->thetac, s->thetar);
  om_sign_pp(m, s->ele_idx, s->sgn);
  for (int j = 0; j < m->mc; j++)
    if (cabs(s->thetac[j] - s->thc2[j]) > 1.0e-5) return -40;
  for (int j = 0; j < m->mr; j++)
    if (fabs(s->thetar[j] - s->thr2[j]) > 1.0e-5) return -41;
  for (int g = 0; g < m->P; g++)
    if (s->sgn[g] != s->sgn2[g]) return -42;
  memcpy(s->invm2, s->invm, sizeof(double) * (size_t)C * N * N);
  memcpy(s->pf2, s->pf, sizeof(double) * C);
  int rc = calculate_mall(m, s->ele_idx, s->pf, s->invm, s->bufA, s->bufB);
  if (rc) return rc;
  calc_amplitude(m, s->thetac, s->thetar, s->pf, s->sgn, s->psi_wo, s->psi, r);
  double r2 = 0.0, dp = -1.0, di = -1.0;
  for (int q = 0; q < C; q++) r2 = r2 + fabs(s->pf[q]);
  r2 = r2 / (double)C;
  for (int q = 0; q < C; q++) {
    double d = fabs(s->pf2[q] - s->pf[q]) / r2;
    if (d > dp) dp = d;
  }
  for (int q = 0; q < C; q++) {
    double r3 = 0.0;
    const double* a = s->invm + (size_t)q * N * N;
    const double* b = s->invm2 + (size_t)q * N * N;
    for (int k = 0; k < N * N; k++) if (fabs(a[k]) > r3) r3 = fabs(a[k]);
    for (int k = 0; k < N * N; k++) {
      double d = fabs(b[k] - a[k]) / r3;
      if (d > di) di = d;
    }
  }
  *dpf = dp;
  *dinv = di;
  return 0;
}

int om_recompute(const om_model* m, om_state* s, double* drift /* pf, invm */) {
  double dpf = 0, dinv = 0;
  int rc = recompute(m, s, NULL, &dpf, &dinv);
  if (drift) { drift[0] = dpf; drift[1] = dinv; }
  return rc;
}

/* make_sample.f90:4-240 (make_sample) for one chain.  With ref_draws=1 and an
 * mt.f stream it consumes random numbers exactly where the reference does;
 * with ref_draws=0 it consumes them exactly where the GPU sweep kernel does
 * (no self-check draws), so a Philox stream reproduces a GPU walker. */
int om_make_sample(const om_model* m, om_state* s, om_rng* r, int first,
                   const om_sample_cfg* c, int* nfastup, int* x_out, double* psi_out,
                   double* pf_out, long* naccept, long* nrecompute) {
  int N = m->N, C = m->C, rc;
  om_rng* rd = c->ref_draws ? r : NULL; /* stream for the amplitude self-check */
  long nacc = 0, nrec = 0;
  double rt[C]; /* Pf ratio of every copy for the current proposal */
  if (first) {
    rc = make_initial_x(m, s, r, rd);
    if (rc) return rc;
  } else {
    memcpy(s->x2, s->x, sizeof(int) * N);
    rc = om_eval(m, s->x2, s, rd);
    if (rc) return rc;
    if (om_rng_u01(r) < c->p_restart) {
      rc = make_initial_x(m, s, r, rd);
      if (rc) return rc;
    }
  }
  int nacc_tmp = 0;
  for (int i0 = -c->nwarmup; i0 <= c->nsample; i0++) {
    int limit = c->fixed_k > 0 ? c->fixed_k : *nfastup;
    if (nacc_tmp > limit) {
      double dpf, dinv;
      rc = recompute(m, s, rd, &dpf, &dinv);
      if (rc) return rc;
      nrec++;
      nacc_tmp = 0;
      if (c->fixed_k <= 0) {
        if (c->ref_draws) (void)om_rng_u01(r); /* make_sample.f90:116 */
        if (dinv > 1.0e-5 || dpf > 1.0e-5) {
          long v = lround((double)*nfastup / 1.2);
          *nfastup = (int)(v < 1 ? 1 : v);
          if (c->ref_draws && dinv > 10.0 && dpf > 10.0) (void)om_rng_u01(r);
        } else {
          long v = lround((double)*nfastup * 1.1);
          *nfastup = (int)(v > 100 ? 100 : v);
        }
      }
    }
    for (int iu = 0; iu < c->nupdate; iu++) {
      if (om_rng_u01(r) < c->p_global) {
        /* make_sample.f90:135-158 (global spin flip) */
        for (int i = 0; i < N; i++) s->x2[i] = -s->x[i];
        om_state* t = om_state_create(m);
        rc = om_eval(m, s->x2, t, rd);
        if (rc) { om_state_free(t); return rc; }
        double r1 = (cabs(t->psi[0]) + c->shift) / (cabs(s->psi[0]) + c->shift);
        if (om_rng_u01(r) < r1 * r1) {
          if (c->ref_draws) (void)om_rng_u01(r);
          memcpy(s->x, t->x, sizeof(int) * N);
          memcpy(s->thetac, t->thetac, sizeof(cplx) * m->mc);
          memcpy(s->thetar, t->thetar, sizeof(double) * m->mr);
          memcpy(s->ele_idx, t->ele_idx, sizeof(int) * N);
          memcpy(s->ele_cfg, t->ele_cfg, sizeof(int) * N);
          memcpy(s->pf, t->pf, sizeof(double) * C);
          memcpy(s->invm, t->invm, sizeof(double) * (size_t)C * N * N);
          memcpy(s->sgn, t->sgn, sizeof(int) * m->P);
          memcpy(s->psi_wo, t->psi_wo, sizeof(cplx) * C);
          memcpy(s->psi, t->psi, sizeof(cplx) * (C + 1));
        }
        om_state_free(t);
      }
      int i1, i2;
      rc = spin_flip_candidate(m, s->x, r, &i1, &i2);
      if (rc) return rc;
      propose_exchange(m, s, i1, i2, rt, rd);
      double r1 = (cabs(s->psi2[0]) + c->shift) / (cabs(s->psi[0]) + c->shift);
      if (om_rng_u01(r) < r1 * r1) {
        if (i0 > 0) nacc++;
        nacc_tmp++;
        accept_exchange(m, s, i1, i2);
        if (c->cond_guard > 0.0) {
          int hit = 0;
          for (int q = 0; q < C; q++) {
            const double a = fabs(rt[q]);
            if (!(a * c->cond_guard > 1.0 && a < c->cond_guard)) hit = 1;
          }
          if (hit) {
            double dpf, dinv;
            rc = recompute(m, s, NULL, &dpf, &dinv);
            if (rc) return rc;
            nrec++;
            nacc_tmp = 0;
          }
        }
      }
    }
    if (i0 > 0) {
      int is = i0 - 1;
      if (x_out) memcpy(x_out + (size_t)is * N, s->x, sizeof(int) * N);
      if (psi_out) { psi_out[2 * is] = creal(s->psi[0]); psi_out[2 * is + 1] = cimag(s->psi[0]); }
      if (pf_out) memcpy(pf_out + (size_t)is * C, s->pf, sizeof(double) * C);
    }
  }
  if (naccept) *naccept = nacc;
  if (nrecompute) *nrecompute = nrec;
  return 0;
}

/* ------------------------------------------------------------------ */
/* SR step                                                              */
/* ------------------------------------------------------------------ */
/* Cholesky solve of an SPD system (the reference uses ScaLAPACK pdposv('U'),
 * SR_pdposv.f90:232, checked against LAPACK dposv, SR.f90:1-28). */
static int chol_solve(int n, double* A /* col-major, overwritten */, double* b) {
#define A_(i, j) A[(size_t)(j) * n + (i)]
  for (int j = 0; j < n; j++) {
    double d = A_(j, j);
    for (int k = 0; k < j; k++) d -= A_(j, k) * A_(j, k);
    if (!(d > 0.0)) return -60;
    d = sqrt(d);
    A_(j, j) = d;
    for (int i = j + 1; i < n; i++) {
      double v = A_(i, j);
      for (int k = 0; k < j; k++) v -= A_(i, k) * A_(j, k);
      A_(i, j) = v / d;
    }
  }
  for (int i = 0; i < n; i++) {
    double v = b[i];
    for (int k = 0; k < i; k++) v -= A_(i, k) * b[k];
    b[i] = v / A_(i, i);
  }
  for (int i = n - 1; i >= 0; i--) {
    double v = b[i];
    for (int k = i + 1; k < n; k++) v -= A_(k, i) * b[k];
    b[i] = v / A_(i, i);
  }
#undef A_
  return 0;
}

/* SR_pdposv.f90:1-414 preceded by the normalisation of main.f90:354-367.
 * Inputs are the un-normalised sums of calc_phys_quantity_opt except etot,
 * which is already Etot/norm.  report = {lambda*5/L, delW_max, scaling,
 * Sdiag_av, s_factor, k_nonzero, lambda_raw, mean|f| before renorm}. */
int om_sr_step(const om_model* m, const om_sr_cfg* c, double etot, double norm,
               const double* gvec_in, const double* ovec_in, const double* smat_raw,
               double* gam, double* rep, double* x_out) {
  int nv = m->nv, nvr = m->nv_rbm, N = m->N, nsl = m->nsl;
  double* S = (double*)malloc(sizeof(double) * (size_t)nv * nv);
  double* Ssave = (double*)malloc(sizeof(double) * (size_t)nv * nv);
  double* g = (double*)malloc(sizeof(double) * nv);
  double* ov = (double*)malloc(sizeof(double) * nv);
  int rc = 0;
#define S_(i, j) S[(size_t)(j) * nv + (i)]
  for (int k = 0; k < nv; k++) { g[k] = gvec_in[k] / norm; ov[k] = ovec_in[k] / norm; }
  for (int k = 0; k < nv; k++) g[k] = g[k] - 2.0 * etot * ov[k];
  for (int k2 = 0; k2 < nv; k2++)
    for (int k1 = 0; k1 < nv; k1++) S_(k1, k2) = smat_raw[(size_t)k2 * nv + k1] / norm - ov[k1] * ov[k2];
  double avR = 0, avP = 0;
  int i1 = 0, i2 = 0;
  for (int k = 0; k < nv; k++)
    if (S_(k, k) > 0.0) {
      if (k < nvr) { i1++; avR += S_(k, k); } else { i2++; avP += S_(k, k); }
    }
  if (i1 != nvr) { rc = -70; goto done; }
  if (i2 != nsl - nsl / N) { rc = -71; goto done; }
  avR = (nvr > 0) ? avR / i1 : 1.0;
  avP = avP / i2;
  double scal = sqrt(avR / avP);
  for (int k = nvr; k < nv; k++) g[k] *= scal;
  for (int k2 = nvr; k2 < nv; k2++)
    for (int k1 = 0; k1 < nv; k1++) S_(k1, k2) *= scal;
  for (int k1 = nvr; k1 < nv; k1++)
    for (int k2 = 0; k2 < nv; k2++) S_(k1, k2) *= scal;
  /* SR_pdposv.f90:78-100 structural checks */
  for (int k = 0; k < nv; k++) {
    if (S_(k, k) < 0.0) { rc = -72; goto done; }
    int is_fii = 0;
    for (int i = 0; i < N; i++)
      if (k == m->orb[i * N + i] + nv - nsl) is_fii = 1;
    if (S_(k, k) == 0.0 && !is_fii) { rc = -73; goto done; }
    if (S_(k, k) > 0.0 && is_fii) { rc = -74; goto done; }
  }
  double Sav = 0;
  int cnt = 0;
  for (int k = 0; k < nv; k++)
    if (S_(k, k) > 0.0) { Sav += S_(k, k); cnt++; }
  Sav /= cnt;
  double sf = c->s_factor0 > 1e-4 ? c->s_factor0 : 1e-4;
  for (int k = 0; k < nv; k++) S_(k, k) = S_(k, k) + Sav * sf;
  memcpy(Ssave, S, sizeof(double) * (size_t)nv * nv);
  if (chol_solve(nv, S, g) != 0) { rc = -75; goto done; }
  double lam = 0;
  for (int k2 = 0; k2 < nv; k2++)
    for (int k1 = 0; k1 < nv; k1++) lam += g[k1] * Ssave[(size_t)k2 * nv + k1] * g[k2];
  lam = sqrt(lam);
  double lam_raw = lam;
  lam = lam * 5.0 / (double)m->L;
  if (lam > 1.0)
    for (int k = 0; k < nv; k++) g[k] /= lam;
  for (int k = nvr; k < nv; k++) g[k] *= scal;
  double dmax = 0;
  int knz = 0;
  for (int k = 0; k < nv; k++) {
    if (fabs(c->delta_tau * g[k]) > dmax) dmax = fabs(c->delta_tau * g[k]);
    if (fabs(g[k]) > 1e-12) knz++;
  }
  for (int k = 0; k < nv; k++) {
    if (dmax > c->max_change) gam[k] = gam[k] - c->delta_tau * g[k] * (c->max_change / dmax);
    else gam[k] = gam[k] - c->delta_tau * g[k];
  }
  int off = nv - nsl;
  double r1 = 0;
  for (int k = off; k < nv; k++) r1 += fabs(gam[k]);
  r1 /= nsl;
  for (int k = off; k < nv; k++) gam[k] = gam[k] * c->slater_abs_av / r1;
  if (x_out) memcpy(x_out, g, sizeof(double) * nv);
  rep[0] = lam; rep[1] = dmax; rep[2] = scal; rep[3] = Sav; rep[4] = sf;
  rep[5] = knz; rep[6] = lam_raw; rep[7] = r1;
#undef S_
done:
  free(S); free(Ssave); free(g); free(ov);
  return rc;
}
