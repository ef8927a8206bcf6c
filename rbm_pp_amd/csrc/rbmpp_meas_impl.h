#pragma once
// Included by rbmpp_meas.hip (k_meas<false>) and rbmpp_corr.hip (k_meas<true>):
// one instantiation per translation unit, so that the correlation pass does
// not change the register allocation of the plain estimator.
//
// Per-configuration estimators for N <= 104 on gfx950: Pf and X^-1 of every
// projection copy (calculate_MAll, matrix.f90:6-110), the J1-J2 local energy
// (calc_local_energy_correlation, calc_phys_quantity.f90:306-378) and the
// log-derivatives O_k (calc_phys_quantity.f90:114-222, slater.f90:94-170).
// One workgroup of 1024 threads per configuration (grid-stride).
//
// Per copy:
//  * X is gathered into registers, 8 x 2 elements per thread (the matrix
//    zero-padded to 128 x 128), and inverted there by
//    Gauss-Jordan exchange steps on 2 x 2 pivot blocks.  Pivots are chosen
//    logically (the first remaining index c and the remaining r maximising
//    |X(c, r)|, the Parlett-Reid rule of dskpfa, matrix.f90:133-151) instead
//    of by swapping rows, so every step costs two workgroup barriers and the
//    result is X^-1 in the original order.  Each exchange keeps the matrix
//    exactly skew-symmetric (its products are rounded symmetrically), so
//    InvM = 0.5 (X^-1(j,i) - X^-1(i,j)) of matrix.f90:97-106 is X^-1 itself.
//  * The pair-product log-derivatives are scattered from the registers.
//  * X^-1 goes to LDS once; each wave then takes two antiparallel bonds per
//    pass, forms P = X^-1 a for both with one read of X^-1, and evaluates the
//    exchange ratio of pfupdate_two.f90:27-88 with wave reductions.
//
// The correlation run (k_meas<true>, Calc_Correlation/calc_phys_quantity.f90:
// 264-550) needs the exchange ratio of every antiparallel pair, (N/2)^2 of
// them.  Instead of P = X^-1 a per pair (N^2 each), the rows G(s, .) of
// Slater_Elm that an electron moving to site s brings (the row of s for the
// spin opposite to its occupant, restricted to the occupied columns) are
// multiplied once per copy: Y = G InvM^T (N^3), and for the pairs
// Z(u, d) = G(s_u, .) . Y(s_d, .) (N^3 / 4).  Every ratio is then a few
// entries of Y, Z, InvM and Slater_Elm (the two replaced columns enter as
// corrections).  The sums of Chi_sp, Chi_vbs and wf_sgn are kept per
// workgroup.
#include "rbmpp_kernels.h"
#include "rbmpp_measure.h"
#include "rbmpp_skew.h"

namespace {

using namespace rbmpp_skew;

constexpr int MNT = 1024;               // threads per workgroup
constexpr int MNW = MNT / RBMPP_WAVE;   // waves
constexpr int MNMAX = RBMPP_NLDS;       // largest N
constexpr int MLD = MNMAX + 1;          // LDS leading dimension bound (n + 1)

struct MeasSmem {
  double inv[MNMAX * MLD];       // inv[j * (n+1) + i] = InvM(j, i) = X^-1(i, j)
  double2 avec[MNW][MNMAX];      // a vectors of the bonds in flight (4 per bond wave)
  double rowbuf[2][2][MPAD];     // pivot rows c and r, double buffered by step
  double opp[2][4 * MNMAX];      // pair-product O, re / im
  double2 psib[4 * MNMAX];       // per antiparallel bond: sum over copies of psi(x')
  double rat[4 * MNMAX];         // per antiparallel bond: Pf ratio of the current copy
  double thr[RBMPP_MRMAX];
  double2 thc[RBMPP_MCMAX];
  double pf[RBMPP_CMAX];
  double2 psiq[RBMPP_CMAX], psiwo[RBMPP_CMAX];
  double selv[2];
  int sel[2];
  int x[MNMAX], eidx[MNMAX], ecfg[MNMAX], site[MNMAX];
  short bi1[4 * MNMAX], bi2[4 * MNMAX];
  int sgn[RBMPP_CMAX / 2];
  unsigned sgnbits[RBMPP_CMAX / 2];
  int nb, status;
  double2 psi_cur;  // psi(x) of the configuration, for the correlation sums
};

// cold helpers kept out of line, so their transcendental temporaries do not
// weigh on the register allocation of the sweep and bond loops
__device__ __attribute__((noinline)) double2 exchanged_amp(const DevModel& m, int q, const MeasSmem& s,
                                                           int sgn, int i1, int i2) {
  return psi_wo_pf_exchanged(m, q, s.thc, s.thr, sgn, i1, i2, s.x[i1], s.x[i2]);
}
__device__ __attribute__((noinline)) void theta_sign(const DevModel& m, MeasSmem& s) { wg_theta_sign(m, s, 0); }
__device__ __attribute__((noinline)) void o_row(const DevModel& m, const MeasSmem& s, double rw, double* ore,
                                                double* oim, bool cplx_rows) {
  wg_orow(m, s, rw, ore, oim, cplx_rows);
}
__device__ __attribute__((noinline)) int bonds_of(const DevModel& m, MeasSmem& s) { return bond_list(m, s); }
__device__ __attribute__((noinline)) double2 eloc_of(const DevModel& m, const MeasSmem& s, double2 p) {
  return eloc_from_bonds(m, s, p);
}

// Exchange ratios of the antiparallel pairs of copy q from Y (in s.inv:
// Y(s, k) at s * ld + k) and InvM(u, np + d) (iud[u * np + d]); lane d takes
// the down electron np + d, wave w the up electrons u = 4w .. 4w + 3.  The
// ratio is that of calculate_new_PfM_two_child (pfupdate_two.f90:27-88) for
// the exchange of the spins at i1 < i2 (electrons a at i1, b at i2), with
// a = g_rsa + corrections at a, b (rsa = the new row of a), likewise b, and
// Y = InvM g, Z = g_rsb . Y_rsa.  psi(x') of the copy is added to psip.
__device__ __attribute__((noinline)) void corr_pairs(const DevModel& m, const MeasSmem& s, const double* S,
                                                     int q, double pfq, double2* psip, int lane, int wid) {
  const int N = m.N, np = m.npair, ld = N + 1, two = 2 * N;
  if (lane >= np || 4 * wid >= np) return;
  const double* ys = s.inv;
  const double* iud = (const double*)&s.avec[0][0];
  const int d = lane, sd = s.eidx[np + d];
  const double* yd = ys + sd * ld;
  int uu[4];
  const double* gu[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    uu[j] = 4 * wid + j;
    const int su = uu[j] < np ? s.eidx[uu[j]] : 0;
    gu[j] = S + (size_t)(su + N) * two;  // an up site takes a down electron: row su + N
  }
  double z[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < N; k++) {
    const double yk = yd[k];
    const int ck = s.eidx[k] + (k >= np ? N : 0);
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (uu[j] < np) z[j] = fma(gu[j][ck], yk, z[j]);
  }
  const int g = q >> 1;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int u = uu[j];
    if (u >= np) break;  // wave-uniform
    const int su = s.eidx[u];
    const bool upfirst = su < sd;
    const int i1 = upfirst ? su : sd, i2 = upfirst ? sd : su;
    const int ea = upfirst ? u : np + d, eb = upfirst ? np + d : u;
    const int rsa = i2 + (s.x[i2] == 1 ? N : 0), rsb = i1 + (s.x[i1] == 1 ? N : 0);
    const int ca = i1 + (ea >= np ? N : 0), cb = i2 + (eb >= np ? N : 0);
    const double D = upfirst ? z[j] : -z[j];  // Z is skew in its two sites
    const double iab = upfirst ? iud[u * np + d] : -iud[u * np + d];
    const double iba = -iab;
    const double* Sa = S + (size_t)rsa * two;
    const double* Sb = S + (size_t)rsb * two;
    const double srr = Sa[rsb];
    const double daa = -Sa[ca], dab = srr - Sa[cb];
    const double dbb = -Sb[cb], dba = -srr - Sb[ca];
    const double y2a = ys[i2 * ld + ea], y2b = ys[i2 * ld + eb];
    const double y1a = ys[i1 * ld + ea], y1b = ys[i1 * ld + eb];
    const double pa = -y2a - iab * dab, pb = -y2b - iba * daa;
    const double qa = -y1a - iab * dbb, qb = -y1b - iba * dba;
    const double gp = -D + daa * y1a + dab * y1b;
    const double bma = gp + dba * pa + dbb * pb;
    const double rat = iba * (-srr) + iba * bma + pa * qb - pb * qa;
    int sfac = -1;  // make_sample.f90:195-197
    if (m.subB[i1]) sfac = -sfac;
    if (m.subB[i2]) sfac = -sfac;
    const double2 c = exchanged_amp(m, q, s, sfac * s.sgn[g], i1, i2);
    psip[u * np + d] = cadd(psip[u * np + d], cscale(c, rat * pfq));
  }
}

template <bool CORR>
__global__ __launch_bounds__(MNT) void k_meas(LocalArgs a) {
  __shared__ MeasSmem s;
  const DevModel& m = *a.m;
  const int tid = threadIdx.x, nth = MNT;
  const int lane = tid & (RBMPP_WAVE - 1), wid = tid >> 6;
  const int tr = tid >> 6, tc = tid & 63;  // register tile coordinates
  const int N = m.N, C = m.C, np = m.npair;
  const int ld = N + 1;
  const bool want_eloc = a.mode & 1, want_o = a.mode & 2, want_inv = a.mode & 4, custom = a.mode & 8;
  // pivot-row entries past the register tile are read but never written
  for (int k = tid; k < 2 * 2 * MPAD; k += nth) (&s.rowbuf[0][0][0])[k] = 0.0;
  const bool h0 = lane < N;               // lane owns index lane
  const bool h1 = lane + RBMPP_WAVE < N;  // lane also owns index lane + 64
  // correlation sums of this workgroup, and the pair amplitudes
  double* cpart = nullptr;
  double2* psip = nullptr;
  int nchi = 0;
  if constexpr (CORR) {
    nchi = corr_nchi(N);
    cpart = a.corr_part + (size_t)blockIdx.x * corr_len(N);
    psip = (double2*)(cpart + corr_sums(N));
    for (int k = tid; k < corr_sums(N); k += nth) cpart[k] = 0.0;
  }

  for (int cfg = blockIdx.x; cfg < a.nconf; cfg += gridDim.x) {
    __syncthreads();
    for (int i = tid; i < N; i += nth) s.x[i] = a.xs[(size_t)cfg * N + i];
    if constexpr (CORR)
      for (int e = tid; e < np * np; e += nth) psip[e] = make_double2(0.0, 0.0);
    if (tid == 0) s.status = WS_OK;
    for (int k = tid; k < 4 * MNMAX; k += nth) {
      s.opp[0][k] = 0.0;
      s.opp[1][k] = 0.0;
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
      int nb = 0;
      if (custom) {
        s.bi1[0] = (short)a.ex_i1[cfg];
        s.bi2[0] = (short)a.ex_i2[cfg];
        nb = 1;
      } else if (want_eloc) {
        nb = bonds_of(m, s);
      }
      s.nb = nb;
    }
    __syncthreads();
    if (s.status != WS_OK) {
      if (tid == 0 && a.status) a.status[cfg] = s.status;
      continue;
    }
    theta_sign(m, s);
    for (int q = tid; q < C; q += nth) s.psiwo[q] = psi_wo_pf(m, q, s.thc, s.thr, s.sgn[q >> 1]);
    const int nb = s.nb;

    for (int q = 0; q < C; q++) {
      const int g = q >> 1, sidx = q & 1;
      const double* S = m.se + (size_t)q * 4 * N * N;
      __syncthreads();  // the previous copy's readers of inv and site are done
      for (int e = tid; e < N; e += nth) s.site[e] = m.wmap[g * N + s.eidx[e]];
      // X(i,j) = -Slater_Elm(r_i, r_j) = se[q][r_i][r_j] (matrix.f90:72-78)
      tile_t x;
      {
        int rj[RB];
#pragma unroll
        for (int b = 0; b < RB; b++) {
          const int j = tc + 64 * b;
          rj[b] = j < N ? s.eidx[j] + (j >= np ? N : 0) : -1;
        }
#pragma unroll
        for (int aa = 0; aa < RA; aa++) {
          const int i = tr + 16 * aa;
          const int ri = i < N ? s.eidx[i] + (i >= np ? N : 0) : -1;
#pragma unroll
          for (int b = 0; b < RB; b++)
            TILE(x, aa, b) = (ri >= 0 && rj[b] >= 0) ? S[(size_t)ri * 2 * N + rj[b]] : 0.0;
        }
      }
      const double pfq = reg_skew_pf_inverse<true>(x, N, &s.rowbuf[0][0][0], s.sel, s.selv);
      if (pfq == 0.0) {
        if (tid == 0) s.status = WS_SINGULAR;
        __syncthreads();
        break;
      }
      // inv[i * ld + j] = InvM(i, j) = X^-1(j, i) = -X^-1(i, j): lanes write
      // consecutive addresses
#pragma unroll
      for (int aa = 0; aa < RA; aa++) {
        const int i = tr + 16 * aa;
#pragma unroll
        for (int b = 0; b < RB; b++) {
          const int j = tc + 64 * b;
          if (i < N && j < N) s.inv[i * ld + j] = -TILE(x, aa, b);
        }
      }
      if (want_inv) {
        // reference layout: out[j*N + i] = InvM(i,j) = X^-1(j,i)
        double* out = a.invm + ((size_t)cfg * C + q) * N * N;
#pragma unroll
        for (int aa = 0; aa < RA; aa++)
#pragma unroll
          for (int b = 0; b < RB; b++) {
            const int i = tr + 16 * aa, j = tc + 64 * b;
            if (i < N && j < N) out[(size_t)i * N + j] = TILE(x, aa, b);
          }
      }
      if (tid == 0) {
        s.pf[q] = pfq;
        s.psiq[q] = cscale(s.psiwo[q], pfq);
      }
      __syncthreads();  // inv complete
      if (want_o) {
        // X^-1(u, v) = -InvM(u, v) = -inv[u * ld + v]
        wg_pp_oloc(m, s, q, pfq, s.psiwo[q], tid, nth, [&](int u, int v) -> double { return -s.inv[u * ld + v]; });
      }
      // exchange ratios (calculate_new_PfM_two_child, pfupdate_two.f90:27-88):
      // waves 0..BW-1 take BT bonds per pass (bonds BT*p .. BT*p+BT-1 for
      // p = w, w + BW, ...); every row of X^-1 read from LDS feeds BT dot
      // products.  The pass is latency-bound: two bonds on all 16 waves beat
      // four bonds on 8.
      constexpr int BT = 2, BW = MNW * 2 / BT;
      double* av = (double*)&s.avec[0][0] + (size_t)wid * BT * MNMAX;  // a_t(j) at av[BT*j + t]
      for (int p = wid; wid < BW && BT * p < nb; p += BW) {
        int bi[BT], i1[BT], i2[BT], msa[BT], msb[BT], rsa[BT], rsb[BT];
#pragma unroll
        for (int t = 0; t < BT; t++) {
          bi[t] = min(BT * p + t, nb - 1);  // a missing bond repeats the last one (not written)
          i1[t] = s.bi1[bi[t]];
          i2[t] = s.bi2[bi[t]];
          msa[t] = s.ecfg[i1[t]];
          msb[t] = s.ecfg[i2[t]];
          rsa[t] = i2[t] + (msa[t] >= np ? N : 0);  // new row of electron msa
          rsb[t] = i1[t] + (msb[t] >= np ? N : 0);
        }
        // a (to LDS) and b (registers) at electrons j = lane, lane + 64
        double bv[BT][2];
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
          const int j = lane + RBMPP_WAVE * hh;
          double av_[BT];
#pragma unroll
          for (int t = 0; t < BT; t++) { av_[t] = 0.0; bv[t][hh] = 0.0; }
          if (j < N) {
            const int off = j >= np ? N : 0;
#pragma unroll
            for (int t = 0; t < BT; t++) {
              int tt = s.eidx[j];
              if (j == msa[t]) tt = i2[t];
              else if (j == msb[t]) tt = i1[t];
              const int rr = tt + off;
              av_[t] = S[(size_t)rsa[t] * 2 * N + rr];
              bv[t][hh] = S[(size_t)rsb[t] * 2 * N + rr];
            }
#pragma unroll
            for (int t = 0; t < BT; t += 2) *(double2*)&av[BT * j + t] = make_double2(av_[t], av_[t + 1]);
          }
        }
        wave_sync();
        // P_t(i) = sum_j InvM(j, i) a_t(j) for i = lane, lane + 64
        double Pt[BT][2];
#pragma unroll
        for (int t = 0; t < BT; t++) { Pt[t][0] = 0.0; Pt[t][1] = 0.0; }
#pragma unroll 4
        for (int j = 0; j < N; j++) {
          const double* row = s.inv + j * ld;
          const double m0 = h0 ? row[lane] : 0.0;
          const double m1 = h1 ? row[lane + RBMPP_WAVE] : 0.0;
#pragma unroll
          for (int t = 0; t < BT; t += 2) {
            const double2 aj = *(const double2*)&av[BT * j + t];
            Pt[t][0] = fma(m0, aj.x, Pt[t][0]);
            Pt[t + 1][0] = fma(m0, aj.y, Pt[t + 1][0]);
            Pt[t][1] = fma(m1, aj.x, Pt[t][1]);
            Pt[t + 1][1] = fma(m1, aj.y, Pt[t + 1][1]);
          }
        }
#pragma unroll
        for (int t = 0; t < BT; t++) {
          if (BT * p + t >= nb) break;  // wave-uniform
          // b^T P, and Q(msa), Q(msb) = -sum_j InvM(msa/msb, j) b(j) (skew symmetry)
          const double* ra = s.inv + msa[t] * ld;
          const double* rb = s.inv + msb[t] * ld;
          double bp = bv[t][0] * Pt[t][0];
          double qa_ = h0 ? ra[lane] * bv[t][0] : 0.0;
          double qb_ = h0 ? rb[lane] * bv[t][0] : 0.0;
          if (h1) {
            bp = fma(bv[t][1], Pt[t][1], bp);
            qa_ = fma(ra[lane + RBMPP_WAVE], bv[t][1], qa_);
            qb_ = fma(rb[lane + RBMPP_WAVE], bv[t][1], qb_);
          }
          // wave sums (DPP, valid in lane 63) and the entries of P and b at
          // the electrons msa, msb read from their owning lanes
          const double bMa = wave_sum_to63(bp);
          const double qa = -wave_sum_to63(qa_);
          const double qb = -wave_sum_to63(qb_);
          const int la = uniform_i(msa[t]), lb = uniform_i(msb[t]);
          const double pa = readlane_d(la >= RBMPP_WAVE ? Pt[t][1] : Pt[t][0], la & 63);
          const double pb = readlane_d(lb >= RBMPP_WAVE ? Pt[t][1] : Pt[t][0], lb & 63);
          const double vec_ba = readlane_d(la >= RBMPP_WAVE ? bv[t][1] : bv[t][0], la & 63);
          if (lane == 63) {
            const double invm_ab = s.inv[lb * ld + la];  // InvM(msb, msa)
            s.rat[bi[t]] = invm_ab * vec_ba + invm_ab * bMa + pa * qb - pb * qa;
          }
        }
        wave_sync();  // av is rewritten by the next pass
      }
      __syncthreads();
      // psi(x') of this copy for every bond, one thread per bond
      for (int b = tid; b < nb; b += nth) {
        const int j1 = s.bi1[b], j2 = s.bi2[b];
        int sfac = -1;
        if (m.subB[j1]) sfac = -sfac;
        if (m.subB[j2]) sfac = -sfac;
        const double pfn = s.rat[b] * pfq;
        const double2 c = exchanged_amp(m, q, s, sfac * s.sgn[g], j1, j2);
        s.psib[b] = cadd(s.psib[b], cscale(c, pfn));
        if (custom && a.ex_pf) a.ex_pf[(size_t)cfg * C + q] = pfn;
      }
      if constexpr (CORR) {
        __syncthreads();  // the bond pass is done with avec and inv
        double* iud = (double*)&s.avec[0][0];  // InvM(u, np + d), up x down block
        for (int e = tid; e < np * np; e += nth) {
          const int u = e / np;
          iud[e] = s.inv[u * ld + np + (e - u * np)];
        }
        // Y(s, k) = sum_i InvM(i, k) Slater[col_i][R(s)] = sum_i InvM(k, i) G(s, i)
        // for sites s = lane, lane + 64 and electrons k = wid + 16 kk; R(s)
        // is the row of site s for the spin opposite to its occupant and
        // col_i the column of electron i (G(s, i) = Slater[R(s)][col_i])
        constexpr int KK = (MNMAX + 15) / 16;
        double y0[KK], y1[KK];
#pragma unroll
        for (int kk = 0; kk < KK; kk++) { y0[kk] = 0.0; y1[kk] = 0.0; }
        const int R0 = h0 ? lane + (s.x[lane] == 1 ? N : 0) : 0;
        const int R1 = h1 ? lane + RBMPP_WAVE + (s.x[lane + RBMPP_WAVE] == 1 ? N : 0) : 0;
        for (int i = 0; i < N; i++) {
          const double* srow = S + (size_t)(s.eidx[i] + (i >= np ? N : 0)) * 2 * N;
          const double g0 = h0 ? srow[R0] : 0.0;
          const double g1 = h1 ? srow[R1] : 0.0;
          const double* irow = s.inv + i * ld;
#pragma unroll
          for (int kk = 0; kk < KK; kk++) {
            const int k = wid + 16 * kk;
            const double mv = k < N ? irow[k] : 0.0;
            y0[kk] = fma(mv, g0, y0[kk]);
            y1[kk] = fma(mv, g1, y1[kk]);
          }
        }
        __syncthreads();  // every read of InvM is done
#pragma unroll
        for (int kk = 0; kk < KK; kk++) {
          const int k = wid + 16 * kk;
          if (k < N) {
            if (h0) s.inv[lane * ld + k] = y0[kk];
            if (h1) s.inv[(lane + RBMPP_WAVE) * ld + k] = y1[kk];
          }
        }
        __syncthreads();  // Y complete
        corr_pairs(m, s, S, q, pfq, psip, lane, wid);
      }
    }
    __syncthreads();
    if (s.status != WS_OK) {
      if (tid == 0 && a.status) a.status[cfg] = s.status;
      continue;
    }
    // amplitude, local energy, reweighting
    if (tid == 0) {
      double2 p = make_double2(0.0, 0.0);
      for (int q = 0; q < C; q++) p = cadd(p, s.psiq[q]);
      const double rw = a.shift >= 0.0 ? 1.0 / (cmod(p) + a.shift) : 1.0;
      a.psi[cfg] = p;
      if constexpr (CORR) s.psi_cur = p;
      if (a.rw) a.rw[cfg] = rw;
      if (want_eloc && !custom) a.eloc[cfg] = cscale(eloc_of(m, s, p), rw);
      if (custom && a.ex_psi) a.ex_psi[cfg] = s.psib[0];
      s.selv[0] = rw;
    }
    for (int q = tid; q < C; q += nth) {
      if (a.pf) a.pf[(size_t)cfg * C + q] = s.pf[q];
      if (a.psi_q) a.psi_q[(size_t)cfg * C + q] = s.psiq[q];
    }
    __syncthreads();
    if (want_o) {
      const bool cplx_rows = !m.real_only;
      double* ore = a.orows + (size_t)(cplx_rows ? 2 * cfg : cfg) * m.nv;
      o_row(m, s, s.selv[0], ore, ore + m.nv, cplx_rows);
    }
    if constexpr (CORR) {
      const double2 p = s.psi_cur;
      const double rw = s.selv[0], w2 = rw * rw;
      corr_accumulate(m, s, p, rw, nb, cpart, nchi);
      double* sp2 = cpart + nchi;
      // exchange part of Chi_loc(2, .) for the antiparallel pairs, by the
      // thread that accumulated the pair: s_i1 s_i2 psi(x')
      if (lane < np) {
        const int sd = s.eidx[np + lane];
        for (int j = 0; j < 4; j++) {
          const int u = 4 * wid + j;
          if (u >= np) break;
          const int su = s.eidx[u];
          const int i1 = min(su, sd), i2 = max(su, sd);
          const double sg = (m.subB[i1] != m.subB[i2]) ? -1.0 : 1.0;
          const double2 c = psip[u * np + lane];
          sp2[i2 * (i2 + 1) / 2 + i1] += w2 * (sg * (p.x * c.x + p.y * c.y));
        }
      }
    }
    if (tid == 0 && a.status) a.status[cfg] = WS_OK;
  }
}

}  // namespace
