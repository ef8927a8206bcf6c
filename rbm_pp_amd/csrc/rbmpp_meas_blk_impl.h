#pragma once
// Included by rbmpp_meas_blk.hip (k_meas_blk<false>) and rbmpp_corr_blk.hip
// (k_meas_blk<true>): one instantiation per translation unit, so that the
// correlation pass does not change the register allocation of the plain
// estimator.
//
// Per-configuration estimators in block form (DevModel::block, normal
// projection): every X is [[0, F], [-F^T, 0]] with F the np x np block of up
// rows and down columns, so Pf X = (-1)^(np(np-1)/2) det F and
// X^-1 = [[0, -F^-T], [F^-1, 0]] (calculate_MAll, matrix.f90:6-110, for this
// structure).  One workgroup of 1024 threads per configuration (grid-stride):
//  * each of the 16 waves inverts one copy's F by Gauss-Jordan in registers
//    (rbmpp_block.h: lane i holds row i; partial pivoting over the rows not
//    yet used, chosen logically), with no workgroup barrier, and writes F^-1
//    to the workgroup's global scratch;
//  * then copy by copy F^-1 is staged in LDS for the exchange ratio of every
//    antiparallel bond (calculate_new_PfM_two_child, pfupdate_two.f90:27-88,
//    as det F'/det F for F' = F + e_u r^T + c e_d^T) and for the
//    pair-product log-derivatives (slater.f90:136-167).
// The energy and O rows follow calc_phys_quantity.f90:20-246 as in the skew
// estimator (rbmpp_meas_impl.h).
//
// The exchange ratios come from Yh = Gdd F^-1, one (N/2)^3 product per copy,
// and two (N/2)-long dot products per exchange (blk_ratio).  The correlation
// run (k_meas_blk<true>, Calc_Correlation/calc_phys_quantity.f90:264-550)
// needs the ratio of every antiparallel pair, (N/2)^2 of them, from the same
// Yh (corr_pairs_blk).  The sums of Chi_sp, Chi_vbs and wf_sgn are kept per
// workgroup as in the skew kernel.
#include "rbmpp_block.h"
#include "rbmpp_kernels.h"
#include "rbmpp_measure.h"

namespace {

using rbmpp_block::copy_block_inverse;

constexpr int BNT = 1024;                // threads per workgroup
constexpr int BNW = BNT / RBMPP_WAVE;    // waves: one per projection copy
constexpr int NMX = RBMPP_NLDS;          // largest N
constexpr int HMX = rbmpp_block::HMX;    // largest np = N/2

struct BlkSmem {
  double binv[HMX * HMX];        // F^-1 of the current copy: B[d][u] at d * np + u
  int piv[BNW][HMX];             // pivot row (lane) of each elimination step
  union {
    struct {
      double stage[BNW][8 * HMX];  // write-back stage of each wave's elimination
    } e;
    // per copy, after the elimination: Gdd and Yh = Gdd F^-1 (copy_yh), and
    // Guu for the correlation run's pairs (corr_pairs_blk)
    struct {
      double gdd[HMX * HMX], guu[HMX * HMX], yh[HMX * HMX];
    } c;
  } u;
  double opp[2][4 * NMX];        // pair-product O, re / im
  double2 psib[4 * NMX];         // per antiparallel bond: sum over copies of psi(x')
  double rat[4 * NMX];           // per antiparallel bond: Pf ratio of the current copy
  double thr[RBMPP_MRMAX];
  double2 thc[RBMPP_MCMAX];
  double pf[RBMPP_CMAX];
  double2 psiq[RBMPP_CMAX], psiwo[RBMPP_CMAX];
  double selv[2];
  int x[NMX], eidx[NMX], ecfg[NMX], site[NMX];
  short bi1[4 * NMX], bi2[4 * NMX];
  int sgn[RBMPP_CMAX / 2];
  unsigned sgnbits[RBMPP_CMAX / 2];
  int nb, status;
  double2 psi_cur;  // psi(x) of the configuration, for the correlation sums
  int wc[3 * BNW];        // per-wave counts of the electron and bond lists
  short bpos[4 * NMX];    // list position of bond (t, i1), -1 when parallel
  double red[2 * BNW];    // per-wave partial sums of E_loc
};
static_assert(sizeof(BlkSmem) <= 160 * 1024, "the block estimator's LDS fits a gfx950 workgroup");

// cold helpers kept out of line (their transcendental temporaries stay out of
// the register allocation of the elimination and bond loops)
__device__ __attribute__((noinline)) double2 exchanged_amp(const DevModel& m, int q, const BlkSmem& s, int sgn,
                                                           int i1, int i2) {
  return psi_wo_pf_exchanged(m, q, s.thc, s.thr, sgn, i1, i2, s.x[i1], s.x[i2]);
}
__device__ __attribute__((noinline)) void theta_sign(const DevModel& m, BlkSmem& s) { wg_theta_sign(m, s, 0); }
__device__ __attribute__((noinline)) void o_row(const DevModel& m, const BlkSmem& s, double rw, double* ore,
                                                double* oim, bool cplx_rows) {
  wg_orow(m, s, rw, ore, oim, cplx_rows);
}

// Exchange ratios in block form.  An exchange of the up electron u (site su)
// and the down electron d (site sd) moves u to sd and d to su:
// F' = F + e_u r^T + c e_d^T with
//   r = R_sd - R_su + delta e_d,  R_s(d') = Slater[s][s_d' + N] (the row of F
//       an up electron at s has), delta = Slater[sd][su+N] - Slater[sd][sd+N],
//   c = C_su - C_sd - eps e_u,    C_s(u') = Slater[s_u'][s + N] (the column
//       a down electron at s has), eps = Slater[su][su+N] - Slater[su][sd+N].
// With B = F^-1, Gdd(d, d') = R_sd(d'), Guu(u', u) = C_su(u'), Yh = Gdd B,
// Xh(d, u) = B[d] . Guu[., u], W(d, u) = Yh[d] . Guu[., u], and R_su B = e_u,
// B C_sd = e_d, Yh C_sd = Gdd(d, d):
//   1 + y_u    = Yh(d, u) + delta B(d, u)              (y = r^T B)
//   1 + B[d].c = Xh(d, u) - eps B(d, u)
//   y.c        = W(d, u) - Gdd(d, d) - delta + delta Xh(d, u) - eps (1 + y_u)
// and det F'/det F = (1 + y_u)(1 + B[d].c) - (y.c) B(d, u), the ratio of
// pfupdate_two.f90:27-88 for the block structure: one (N/2)^3 product per
// copy and two (N/2)-long dot products per exchange instead of a pass over
// F^-1 per exchange.
__device__ __forceinline__ double blk_ratio(double yh, double xh, double w, double b, double delta, double eps,
                                            double gdd) {
  const double a1 = yh + delta * b;
  const double a2 = xh - eps * b;
  const double yc = w - gdd - delta + delta * xh - eps * a1;
  return a1 * a2 - yc * b;
}

typedef double f64x4_t __attribute__((ext_vector_type(4)));

// Gdd and Yh = Gdd F^-1 of copy S (F^-1 in s.binv), and with guu also Guu,
// by the whole workgroup; ends in step.  The product runs on the FP64 matrix
// cores: np is padded to 64 with zeros, wave w computes the 16 x 16 tile
// (w / 4, w % 4) of Yh by 16 v_mfma_f64_16x16x4 (lane l supplies row l % 16
// of Gdd and column l % 16 of F^-1 at k = 4 s + l / 16; output register r
// is element (l / 16 + 4 r, l % 16) of the tile), reading each operand once
// from LDS instead of once per output entry.
__device__ __attribute__((noinline)) void copy_yh(const DevModel& m, BlkSmem& s, const double* S, bool guu) {
  const int N = m.N, np = m.npair, two = 2 * N, tid = threadIdx.x;
  const int npp = np * np;
  for (int e = tid; e < npp; e += BNT) {
    const int r = e / np, c = e - r * np;
    s.u.c.gdd[e] = S[(size_t)s.eidx[np + r] * two + s.eidx[np + c] + N];
    if (guu) s.u.c.guu[e] = S[(size_t)s.eidx[r] * two + s.eidx[c] + N];
  }
  __syncthreads();
  static_assert(HMX <= 64 && BNW == 16, "Yh tiles: np <= 64, one 16 x 16 tile per wave");
  const int lane = tid & (RBMPP_WAVE - 1), wid = tid >> 6;
  const int r0 = 16 * (wid >> 2), c0 = 16 * (wid & 3);
  if (r0 < np && c0 < np) {  // wave-uniform
    const int mr = lane & 15, ks = lane >> 4;
    const int ra = r0 + mr, cb = c0 + mr;
    f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < np; k0 += 4) {
      const int k = k0 + ks;
      const double a = (ra < np && k < np) ? s.u.c.gdd[ra * np + k] : 0.0;
      const double b = (cb < np && k < np) ? s.binv[k * np + cb] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int d = r0 + ks + 4 * r;
      if (d < np && cb < np) s.u.c.yh[d * np + cb] = acc[r];
    }
  }
  __syncthreads();
}

// Exchange ratios of the antiparallel bonds b < nb of the copy into s.rat,
// four bonds per wave and pass (lanes over the up electrons u'): Xh and W of
// the four by two four-way wave sums, so that the four bonds' gathers and
// reductions overlap; Guu[., u] is read from row su + N of Slater_Elm (with
// the block weights Slater[su+N][su'] = -Slater[su'][su+N] exactly).  Needs Yh.
__device__ __attribute__((noinline)) void bond_ratios(const DevModel& m, BlkSmem& s, const double* S, int nb) {
  constexpr int BT = 4;
  const int N = m.N, np = m.npair, two = 2 * N;
  const int lane = threadIdx.x & (RBMPP_WAVE - 1), wid = threadIdx.x >> 6;
  const bool live = lane < np;
  const int su_l = live ? s.eidx[lane] : 0;  // site of up electron `lane`
  for (int p = wid; BT * p < nb; p += BNW) {
    int bb[BT], u[BT], d[BT];
    double g[BT], bv[BT], yv[BT], dl[BT], ep[BT], gd[BT];
#pragma unroll
    for (int t = 0; t < BT; t++) {
      bb[t] = min(BT * p + t, nb - 1);  // a missing bond repeats the last one (not written)
      const int i1 = s.bi1[bb[t]], i2 = s.bi2[bb[t]];
      const int msa = s.ecfg[i1], msb = s.ecfg[i2];
      const bool aup = msa < np;
      u[t] = __builtin_amdgcn_readfirstlane(aup ? msa : msb);
      d[t] = __builtin_amdgcn_readfirstlane((aup ? msb : msa) - np);
      const int su = __builtin_amdgcn_readfirstlane(aup ? i1 : i2);
      const int sd = __builtin_amdgcn_readfirstlane(aup ? i2 : i1);
      g[t] = live ? -S[(size_t)(su + N) * two + su_l] : 0.0;
      bv[t] = live ? s.binv[d[t] * np + lane] : 0.0;
      yv[t] = live ? s.u.c.yh[d[t] * np + lane] : 0.0;
      // the scalars of blk_ratio, loaded here so that their latency overlaps
      // the gathers and the wave sums
      const double* Su = S + (size_t)su * two;
      const double* Sd = S + (size_t)sd * two;
      gd[t] = Sd[sd + N];
      dl[t] = Sd[su + N] - gd[t];
      ep[t] = Su[su + N] - Su[sd + N];
    }
    // lanes 15, 47, 31, 63 hold bonds 0-3
    const double xv = wave_sum4(bv[0] * g[0], bv[1] * g[1], bv[2] * g[2], bv[3] * g[3]);
    const double wv = wave_sum4(yv[0] * g[0], yv[1] * g[1], yv[2] * g[2], yv[3] * g[3]);
    constexpr int SUMLANE[BT] = {15, 47, 31, 63};
#pragma unroll
    for (int t = 0; t < BT; t++) {
      if (BT * p + t >= nb) break;  // wave-uniform
      const double xh = readlane_d(xv, SUMLANE[t]), w = readlane_d(wv, SUMLANE[t]);
      const double bdu = readlane_d(bv[t], u[t]), ydu = readlane_d(yv[t], u[t]);
      if (lane == t) s.rat[bb[t]] = blk_ratio(ydu, xh, w, bdu, dl[t], ep[t], gd[t]);
    }
  }
}

// Exchange ratio of every antiparallel pair of the copy and psi(x') of the
// copy, added to psip[d * np + u]; called by the whole workgroup with Gdd,
// Guu and Yh of the copy in place (copy_yh).  Thread e owns pairs e,
// e + 1024, ... in every copy, so psip needs no synchronisation.
__device__ __attribute__((noinline)) void corr_pairs_blk(const DevModel& m, BlkSmem& s, const double* S, int q,
                                                         double pfq, double2* psip) {
  const int N = m.N, np = m.npair, two = 2 * N, tid = threadIdx.x;
  const int npp = np * np;
  const int g = q >> 1;
  for (int e = tid; e < npp; e += BNT) {
    const int d = e / np, u = e - d * np;
    const double* bd = s.binv + d * np;
    const double* yd = s.u.c.yh + d * np;
    double xh = 0.0, w = 0.0;
    for (int k = 0; k < np; k++) {
      const double gk = s.u.c.guu[k * np + u];
      xh = fma(bd[k], gk, xh);
      w = fma(yd[k], gk, w);
    }
    const int su = s.eidx[u], sd = s.eidx[np + d];
    const double* Su = S + (size_t)su * two;
    const double* Sd = S + (size_t)sd * two;
    const double rat = blk_ratio(yd[u], xh, w, bd[u], Sd[su + N] - Sd[sd + N], Su[su + N] - Su[sd + N], Sd[sd + N]);
    const int i1 = min(su, sd), i2 = max(su, sd);
    int sfac = -1;  // make_sample.f90:195-197
    if (m.subB[i1]) sfac = -sfac;
    if (m.subB[i2]) sfac = -sfac;
    const double2 c = exchanged_amp(m, q, s, sfac * s.sgn[g], i1, i2);
    psip[e] = cadd(psip[e], cscale(c, rat * pfq));
  }
}

template <bool CORR>
__global__ __launch_bounds__(BNT) void k_meas_blk(LocalArgs a) {
  __shared__ BlkSmem s;
  const DevModel& m = *a.m;
  const int tid = threadIdx.x, nth = BNT;
  const int lane = tid & (RBMPP_WAVE - 1), wid = tid >> 6;
  const int N = m.N, C = m.C, np = m.npair;
  const bool want_eloc = a.mode & 1, want_o = a.mode & 2, want_inv = a.mode & 4, custom = a.mode & 8;
  const double pfsign = ((np * (np - 1) / 2) & 1) ? -1.0 : 1.0;  // Pf X = pfsign det F
  double* Bw = a.scratch + (size_t)blockIdx.x * C * np * np;  // F^-1 of every copy
  // correlation sums of this workgroup, and the pair amplitudes (pair e =
  // d * np + u at psip[e])
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
    for (int k = tid; k < 4 * NMX; k += nth) {
      s.opp[0][k] = 0.0;
      s.opp[1][k] = 0.0;
      s.psib[k] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    // electron lists and the antiparallel bonds (or the one custom exchange)
    wg_ele_bonds(m, s, want_eloc && !custom, s.wc, s.bpos);
    if (tid == 0) {
      if (custom) {
        s.bi1[0] = (short)a.ex_i1[cfg];
        s.bi2[0] = (short)a.ex_i2[cfg];
        s.nb = 1;
      } else if (!want_eloc) {
        s.nb = 0;
      }
    }
    __syncthreads();
    if (s.status != WS_OK) {
      if (tid == 0 && a.status) a.status[cfg] = s.status;
      continue;
    }
    theta_sign(m, s);
    for (int q = tid; q < C; q += nth) s.psiwo[q] = psi_wo_pf(m, q, s.thc, s.thr, s.sgn[q >> 1]);
    // ---- det and F^-1 of every copy, one wave each
    if (wid < C) {
      const int q = wid;
      const double det = copy_block_inverse(m.se + (size_t)q * 4 * N * N, N, np, s.eidx, s.piv[wid],
                                            s.u.e.stage[wid], Bw + (size_t)q * np * np, lane);
      if (lane == 0) {
        s.pf[q] = pfsign * det;
        if (det == 0.0) s.status = WS_SINGULAR;
      }
    }
    __syncthreads();
    if (s.status != WS_OK) {
      if (tid == 0 && a.status) a.status[cfg] = s.status;
      continue;
    }
    const int nb = s.nb;
    for (int q = 0; q < C; q++) {
      const int g = q >> 1, sidx = q & 1;
      const double* S = m.se + (size_t)q * 4 * N * N;
      const double pfq = s.pf[q];
      __syncthreads();  // the previous copy's readers of binv and site are done
      for (int e = tid; e < np * np; e += nth) s.binv[e] = Bw[(size_t)q * np * np + e];
      for (int e = tid; e < N; e += nth) s.site[e] = m.wmap[g * N + s.eidx[e]];
      if (tid == 0) s.psiq[q] = cscale(s.psiwo[q], pfq);
      __syncthreads();
      if (want_inv) {
        // reference layout out[i*N + j] = X^-1(i, j): X^-1(u, np+d) = -F^-1(d, u),
        // X^-1(np+d, u) = F^-1(d, u), zero same-spin blocks
        double* out = a.invm + ((size_t)cfg * C + q) * N * N;
        for (int e = tid; e < N * N; e += nth) {
          const int i = e / N, j = e - i * N;
          double v = 0.0;
          if (i < np && j >= np) v = -s.binv[(j - np) * np + i];
          else if (i >= np && j < np) v = s.binv[(i - np) * np + j];
          out[e] = v;
        }
      }
      if (want_o) {
        // X^-1(u, np+d) = -F^-1(d, u), X^-1(np+d, u) = F^-1(d, u); the
        // same-spin entries have weight cs = 0
        const double* Bq = s.binv;
        wg_pp_oloc(m, s, q, pfq, s.psiwo[q], tid, nth, [&](int u, int v) -> double {
          return u < np ? -Bq[(v - np) * np + u] : Bq[(u - np) * np + v];
        });
      }
      // exchange ratios of the antiparallel bonds (blk_ratio)
      if (nb > 0 || CORR) {
        copy_yh(m, s, S, CORR);
        bond_ratios(m, s, S, nb);
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
      if constexpr (CORR) corr_pairs_blk(m, s, S, q, pfq, psip);
    }
    __syncthreads();
    // amplitude, local energy, reweighting
    {
      // every thread forms p the same way (the 16 terms in copy order)
      double2 p = make_double2(0.0, 0.0);
      for (int q = 0; q < C; q++) p = cadd(p, s.psiq[q]);
      const double rw = a.shift >= 0.0 ? 1.0 / (cmod(p) + a.shift) : 1.0;
      double2 el = make_double2(0.0, 0.0);
      if (want_eloc && !custom) el = wg_eloc(m, s, p, s.bpos, s.red);
      if (tid == 0) {
        a.psi[cfg] = p;
        if constexpr (CORR) s.psi_cur = p;
        if (a.rw) a.rw[cfg] = rw;
        if (want_eloc && !custom) a.eloc[cfg] = cscale(el, rw);
        if (custom && a.ex_psi) a.ex_psi[cfg] = s.psib[0];
        s.selv[0] = rw;
      }
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
      const double rw = s.selv[0];
      corr_accumulate(m, s, p, rw, nb, cpart, nchi);
      // exchange part of Chi_loc(2, .) for the antiparallel pairs, by the
      // thread that accumulated the pair: s_i1 s_i2 psi(x')
      double* sp2 = cpart + nchi;
      const double w2 = rw * rw;
      for (int e = tid; e < np * np; e += nth) {
        const int d = e / np, u = e - d * np;
        const int su = s.eidx[u], sd = s.eidx[np + d];
        const int i1 = min(su, sd), i2 = max(su, sd);
        const double sg = (m.subB[i1] != m.subB[i2]) ? -1.0 : 1.0;
        const double2 c = psip[e];
        sp2[i2 * (i2 + 1) / 2 + i1] += w2 * (sg * (p.x * c.x + p.y * c.y));
      }
    }
    if (tid == 0 && a.status) a.status[cfg] = WS_OK;
  }
}

}  // namespace
