// Per-configuration estimators on gfx950: full recompute of Pf and X^-1
// (calculate_MAll, matrix.f90:6-110), local energy of the J1-J2 model
// (calc_local_energy_correlation, calc_phys_quantity.f90:306-378) and the
// log-derivative vector O_k (calc_phys_quantity.f90:114-222, slater.f90:94-170)
// as a hand-written adjoint: RBM entries from tanh(theta) * psi_q, pair
// entries from the projection-weighted X^-1 entries scattered into the
// orbital bins.  One workgroup per configuration (grid-stride).
//
// The Gauss-Jordan result is kept in its pivoted order A = (P X P^T)^-1;
// X^-1(i,j) = A(ip[i], ip[j]).  Every exchange ratio needs P = X^-1 a:
// lane r owns P_{perm[r]} = -sum_c A(c, r) a_{perm[c]} (skew symmetry turns
// the product into a contiguous row read of LDS).
#include "rbmpp_kernels.h"
#include "rbmpp_measure.h"

template <int NMAX, bool LDSMAT>
struct LocalSmem {
  static constexpr int NW = 8;
  static constexpr int LD = NMAX < RBMPP_NLDS ? NMAX : RBMPP_NLDS;
  static constexpr int MAT = LDSMAT ? LD * LD : 1;
  double A[MAT];
  double vab[NW][2][NMAX];
  double opp[2][4 * NMAX];          // pair-product O, re / im
  double2 psib[4 * NMAX];           // sum over copies of psi(x') per antiparallel bond
  double thr[RBMPP_MRMAX];
  double2 thc[RBMPP_MCMAX];
  double pf[RBMPP_CMAX];
  double2 psiq[RBMPP_CMAX], psiwo[RBMPP_CMAX];
  double gjbuf[4 * NMAX];
  double red[32];
  double2 psi;
  int x[NMAX], eidx[NMAX], ecfg[NMAX], perm[NMAX], ip[NMAX];
  short bi1[4 * NMAX], bi2[4 * NMAX];
  int sgn[RBMPP_CMAX / 2];
  unsigned sgnbits[RBMPP_CMAX / 2];
  int redi[32];
  int piv, nb, status;
};

template <int NMAX, bool LDSMAT>
__global__ __launch_bounds__(512) void k_local(LocalArgs a) {
  constexpr int KMAX = NMAX / RBMPP_WAVE;
  __shared__ LocalSmem<NMAX, LDSMAT> s;
  const DevModel& m = *a.m;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & (RBMPP_WAVE - 1), wid = tid >> 6, nwv = nth >> 6;
  const int N = m.N, C = m.C, P = m.P, np = m.npair;
  const bool want_eloc = a.mode & 1, want_o = a.mode & 2, want_inv = a.mode & 4, custom = a.mode & 8;
  double* A = LDSMAT ? s.A : a.scratch + (size_t)blockIdx.x * N * N;
  const int ld = N;

  for (int cfg = blockIdx.x; cfg < a.nconf; cfg += gridDim.x) {
    __syncthreads();
    for (int i = tid; i < N; i += nth) s.x[i] = a.xs[(size_t)cfg * N + i];
    if (tid == 0) s.status = WS_OK;
    for (int k = tid; k < 4 * NMAX; k += nth) {
      s.opp[0][k] = 0.0; s.opp[1][k] = 0.0;
      s.psib[k] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (tid == 0) {
      // x_to_Ele_Idx_Ele_Cfg (interface_for_PP.f90:5-33)
      int u = 0, d = np;
      for (int i = 0; i < N; i++) {
        if (s.x[i] == 1) { s.eidx[u] = i; s.ecfg[i] = u; u++; }
        else { s.eidx[d] = i; s.ecfg[i] = d; d++; }
      }
      if (u != np) s.status = WS_INIT;
      // antiparallel bonds in the order of calc_phys_quantity.f90:332-339
      int nb = 0;
      if (custom) {
        s.bi1[0] = (short)a.ex_i1[cfg]; s.bi2[0] = (short)a.ex_i2[cfg]; nb = 1;
      } else if (want_eloc) {
        nb = bond_list(m, s);
      }
      s.nb = nb;
    }
    __syncthreads();
    if (s.status != WS_OK) {
      if (tid == 0 && a.status) a.status[cfg] = s.status;
      continue;
    }
    wg_theta_sign(m, s, 0);
    for (int q = tid; q < C; q += nth) s.psiwo[q] = psi_wo_pf(m, q, s.thc, s.thr, s.sgn[q >> 1]);
    const int nb = s.nb;

    for (int q = 0; q < C; q++) {
      const int g = q >> 1, sidx = q & 1;
      const double* S = m.se + (size_t)q * 4 * N * N;
      __syncthreads();
      for (int e = tid; e < N * N; e += nth) {
        const int i = e / N, j = e - i * N;
        const int ri = s.eidx[i] + (i >= np ? N : 0);
        const int rj = s.eidx[j] + (j >= np ? N : 0);
        A[(size_t)i * ld + j] = S[(size_t)ri * 2 * N + rj];
      }
      __syncthreads();
      const double pfq = wg_skew_inverse(A, N, ld, s.perm, s.gjbuf, s.red, s.redi, &s.piv);
      if (pfq == 0.0) {
        if (tid == 0) s.status = WS_SINGULAR;
        __syncthreads();
        break;
      }
      for (int r = tid; r < N; r += nth) s.ip[s.perm[r]] = r;
      if (tid == 0) {
        s.pf[q] = pfq;
        s.psiq[q] = cscale(s.psiwo[q], pfq);
      }
      __syncthreads();
      // X^-1(i,j) antisymmetrised: InvM(j,i) (matrix.f90:97-106)
#define XINV(i, j) (0.5 * (A[(size_t)s.ip[i] * ld + s.ip[j]] - A[(size_t)s.ip[j] * ld + s.ip[i]]))
      if (want_inv) {
        double* out = a.invm + ((size_t)cfg * C + q) * N * N;
        for (int e = tid; e < N * N; e += nth) {
          const int j = e / N, i = e - j * N;  // out[j*N+i] = InvM(i,j) = Xinv(j,i)
          out[e] = XINV(j, i);
        }
      }
      if (want_o) wg_pp_oloc(m, s, q, pfq, s.psiwo[q], tid, nth, [&](int u, int v) -> double { return XINV(u, v); });
      // exchange ratios (calculate_new_PfM_two_child, pfupdate_two.f90:27-88)
      for (int b0 = 0; b0 < nb; b0 += nwv) {
        const int b = b0 + wid;
        double* va = &s.vab[wid][0][0];
        double* vb = &s.vab[wid][1][0];
        int i1 = 0, i2 = 0, msa = 0, msb = 0;
        if (b < nb) {
          i1 = s.bi1[b]; i2 = s.bi2[b];
          msa = s.ecfg[i1]; msb = s.ecfg[i2];
          const int rsa = i2 + (msa >= np ? N : 0), rsb = i1 + (msb >= np ? N : 0);
          const double* colA = S + (size_t)rsa * 2 * N;
          const double* colB = S + (size_t)rsb * 2 * N;
          // a, b in pivoted order: va[c] = a_{perm[c]}
          for (int c = lane; c < N; c += RBMPP_WAVE) {
            const int i = s.perm[c];
            int t = s.eidx[i];
            if (i == msa) t = i2;
            else if (i == msb) t = i1;
            const int r = t + (i >= np ? N : 0);
            va[c] = colA[r];
            vb[c] = colB[r];
          }
        }
        __syncthreads();
        if (b < nb) {
          double Pk[KMAX];
#pragma unroll
          for (int k = 0; k < KMAX; k++) Pk[k] = 0.0;
          double qa = 0.0, qb = 0.0;
          const int ra = s.ip[msa], rb = s.ip[msb];
          for (int c = 0; c < N; c++) {
            const double ac = va[c];
            const double* rowc = A + (size_t)c * ld;
#pragma unroll
            for (int k = 0; k < KMAX; k++) {
              const int r = lane + k * RBMPP_WAVE;
              if (r < N) Pk[k] = fma(-rowc[r], ac, Pk[k]);
            }
          }
          // Q(msa), Q(msb) = (X^-1 b)(msa/msb) = -sum_c A(c, ip) b_{perm[c]}
          double bdotp = 0.0;
          for (int c = lane; c < N; c += RBMPP_WAVE) {
            qa = fma(-A[(size_t)c * ld + ra], vb[c], qa);
            qb = fma(-A[(size_t)c * ld + rb], vb[c], qb);
          }
#pragma unroll
          for (int k = 0; k < KMAX; k++) {
            const int r = lane + k * RBMPP_WAVE;
            if (r < N) bdotp = fma(vb[r], Pk[k], bdotp);
          }
          qa = wave_sum(qa);
          qb = wave_sum(qb);
          const double bMa = wave_sum(bdotp);
          // p_a = P(msa): held by lane ra % 64, slot ra / 64
          double pa = 0.0, pb = 0.0;
#pragma unroll
          for (int k = 0; k < KMAX; k++) {
            const double va_k = __shfl(Pk[k], ra & 63, RBMPP_WAVE);
            const double vb_k = __shfl(Pk[k], rb & 63, RBMPP_WAVE);
            if ((ra >> 6) == k) pa = va_k;
            if ((rb >> 6) == k) pb = vb_k;
          }
          if (lane == 0) {
            const double invm_ab = XINV(msa, msb);       // InvM(msb, msa)
            const double vec_ba = vb[ra];                // b_{msa}
            const double ratio = invm_ab * vec_ba + invm_ab * bMa + pa * qb - pb * qa;
            const double pfn = ratio * pfq;
            int sfac = -1;
            if (m.subB[i1]) sfac = -sfac;
            if (m.subB[i2]) sfac = -sfac;
            const double2 c = psi_wo_pf_exchanged(m, q, s.thc, s.thr, sfac * s.sgn[g], i1, i2,
                                                  s.x[i1], s.x[i2]);
            s.psib[b] = cadd(s.psib[b], cscale(c, pfn));
            if (custom && a.ex_pf) a.ex_pf[(size_t)cfg * C + q] = pfn;
          }
        }
        __syncthreads();
      }
#undef XINV
    }
    if (s.status != WS_OK) {
      if (tid == 0 && a.status) a.status[cfg] = s.status;
      continue;
    }
    __syncthreads();
    // amplitude, local energy, reweighting
    if (tid == 0) {
      double2 p = make_double2(0.0, 0.0);
      for (int q = 0; q < C; q++) p = cadd(p, s.psiq[q]);
      s.psi = p;
      const double rw = a.shift >= 0.0 ? 1.0 / (cmod(p) + a.shift) : 1.0;
      a.psi[cfg] = p;
      if (a.rw) a.rw[cfg] = rw;
      if (want_eloc && !custom) a.eloc[cfg] = cscale(eloc_from_bonds(m, s, p), rw);
      if (custom && a.ex_psi) a.ex_psi[cfg] = s.psib[0];
      s.red[0] = rw;
    }
    for (int q = tid; q < C; q += nth) {
      if (a.pf) a.pf[(size_t)cfg * C + q] = s.pf[q];
      if (a.psi_q) a.psi_q[(size_t)cfg * C + q] = s.psiq[q];
    }
    __syncthreads();
    if (want_o) {
      const bool cplx_rows = !m.real_only;
      double* ore = a.orows + (size_t)(cplx_rows ? 2 * cfg : cfg) * m.nv;
      wg_orow(m, s, s.red[0], ore, ore + m.nv, cplx_rows);
    }
    if (tid == 0 && a.status) a.status[cfg] = WS_OK;
  }
}

template <int NMAX, bool LDSMAT>
static void launch_local_t(const LocalArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL((k_local<NMAX, LDSMAT>), dim3(grid), dim3(512), 0, st, a);
}

int launch_local(const DevModel& hm, const LocalArgs& a, int grid, hipStream_t st) {
  const int N = hm.N;
  if (hm.C > RBMPP_CMAX || hm.mr > RBMPP_MRMAX || hm.mc > RBMPP_MCMAX) return -1;
  if (hm.nsl > 4 * N || (N & 1)) return -1;
  if (a.corr_part) {
    // the correlation run: the pair pass of the block-form or the register
    // estimator
    if (N > RBMPP_NLDS || hm.L < 4) return -1;
    if (hm.block && hm.C == RBMPP_CMAX) {
      if (!a.scratch) return -1;
      launch_meas_blk_corr(a, grid, st);
    } else {
      if (N & 3) return -1;
      launch_meas_corr(a, grid, st);
    }
  } else if (hm.block && hm.C == RBMPP_CMAX && N <= RBMPP_NLDS) {
    if (!a.scratch) return -1;
    launch_meas_blk(a, grid, st);
  } else if (N <= RBMPP_NLDS) launch_meas(a, grid, st);
  else if (N <= 256) {
    if (!a.scratch) return -1;
    // the callers' scratch holds N x N doubles per workgroup
    if (hm.block && hm.C == RBMPP_CMAX && meas_blk_big_scratch(N) <= (size_t)N * N) launch_meas_blk_big(a, grid, st);
    else launch_local_t<256, false>(a, grid, st);
  } else return -1;
  return 0;
}
