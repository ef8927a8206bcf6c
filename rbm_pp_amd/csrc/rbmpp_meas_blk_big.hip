// Per-configuration estimator in block form for 104 < N <= 256 (np = N/2 up
// to 128): Pf and X^-1 of every copy (calculate_MAll, matrix.f90:6-110, for
// X = [[0, F], [-F^T, 0]]: Pf X = (-1)^(np(np-1)/2) det F,
// X^-1 = [[0, -F^-T], [F^-1, 0]]), the J1-J2 local energy
// (calc_phys_quantity.f90:306-378) and the log-derivatives O_k
// (calc_phys_quantity.f90:114-222, slater.f90:94-170).  One workgroup of 1024
// threads per configuration (grid-stride).  A row of F per lane does not fit
// the registers here, so the copies are taken one after the other by the
// whole workgroup: F in LDS, Gauss-Jordan with partial pivoting
// (rbmpp_block::wg_block_inverse_lds), F^-1 restaged into LDS in natural
// order; then the pair-product O_k from it, Yh = Gdd F^-1 (one np^3 product
// per copy, Gdd gathered from the Slater table) into the workgroup's global
// scratch, and the exchange ratio of every antiparallel bond from two
// np-long dot products (blk_ratio, the algebra of pfupdate_two.f90:27-88 for
// the block structure), four bonds per wave and pass.  The per-bond
// amplitude sums, the ratios and the pair-product O_k bins live in the
// scratch as well: F alone fills most of the LDS.
#include "rbmpp_block.h"
#include "rbmpp_kernels.h"
#include "rbmpp_measure.h"

namespace {

constexpr int GNT = 1024;              // threads per workgroup
constexpr int GNW = GNT / RBMPP_WAVE;  // waves
constexpr int GNMX = 256;              // largest N
constexpr int GHMX = GNMX / 2;         // largest np

// doubles of global scratch per workgroup: F^-1 (natural order) and Yh,
// np^2 each; the ratios of the current copy (4N); the per-bond sums of psi(x')
// (4N double2); the pair-product O_k bins, re / im (two rows of 4 GNMX)
__host__ __device__ inline size_t big_scratch(int N) {
  const size_t np = N / 2;
  return 2 * np * np + 12 * (size_t)N + 8 * (size_t)GNMX;
}

struct BigSmem {
  double T[GHMX * GHMX];  // F, then F^-1 (natural order: B[d][u] at d * np + u)
  double rowb[GHMX], colb[GHMX];
  int used[GHMX], perm[GHMX], pinv[GHMX];
  double red[2 * GNW];
  int s_p, s_cnt;
  double s_pv;
  double thr[RBMPP_MRMAX];
  double2 thc[RBMPP_MCMAX];
  double pf[RBMPP_CMAX];
  double2 psiq[RBMPP_CMAX], psiwo[RBMPP_CMAX];
  double selv[2];
  int x[GNMX], eidx[GNMX], ecfg[GNMX];
  short bi1[4 * GNMX], bi2[4 * GNMX];
  short bpos[4 * GNMX];
  int sgn[RBMPP_CMAX / 2];
  unsigned sgnbits[RBMPP_CMAX / 2];
  int wc[3 * GNW];
  int nb, status;
};
static_assert(sizeof(BigSmem) <= 160 * 1024, "the estimator's LDS fits a gfx950 workgroup");

// the members the shared helpers (wg_pp_oloc, wg_eloc, wg_orow) reach through
// S: the shared-memory block plus the scratch-resident bins and sums
struct BigView {
  double (*opp)[4 * GNMX];
  double2* psib;
  const int* x;
  const int* ecfg;
  const double* thr;
  const double2* thc;
  const double2* psiq;
};

__device__ __attribute__((noinline)) double2 big_exchanged_amp(const DevModel& m, int q, const BigSmem& s, int sgn,
                                                               int i1, int i2) {
  return psi_wo_pf_exchanged(m, q, s.thc, s.thr, sgn, i1, i2, s.x[i1], s.x[i2]);
}
__device__ __attribute__((noinline)) void big_theta_sign(const DevModel& m, BigSmem& s) { wg_theta_sign(m, s, 0); }

// ratio of pfupdate_two.f90:27-88 for the block structure (see blk_ratio of
// the N <= 104 estimator): det F'/det F from Yh(d, u), Xh(d, u), W(d, u),
// B(d, u) and the three Slater scalars of the exchange
__device__ __forceinline__ double big_ratio(double yh, double xh, double w, double b, double delta, double eps,
                                            double gdd) {
  const double a1 = yh + delta * b;
  const double a2 = xh - eps * b;
  const double yc = w - gdd - delta + delta * xh - eps * a1;
  return a1 * a2 - yc * b;
}

typedef double f64x4_t __attribute__((ext_vector_type(4)));

// Yh = Gdd F^-1 of copy S into yh (global scratch), Gdd(d, d') =
// Slater[site_(np+d)][site_(np+d') + N], on the FP64 matrix cores: np is
// covered by 16 x 16 tiles (zero beyond np), wave w takes row strip w / 2
// and the four column tiles 4 (w % 2) + j, so that one operand fragment of
// Gdd (gathered from the Slater table) feeds four v_mfma_f64_16x16x4 (lane l
// supplies row l % 16 of Gdd and column l % 16 of F^-1 at k = 4 s + l / 16;
// output register r is element (l / 16 + 4 r, l % 16) of a tile).  Ends in
// step.
__device__ __attribute__((noinline)) void big_yh(const DevModel& m, const BigSmem& s, const double* S, double* yh) {
  const int N = m.N, np = m.npair, two = 2 * N, tid = threadIdx.x;
  const int lane = tid & (RBMPP_WAVE - 1), wid = tid >> 6;
  const int r0 = 16 * (wid >> 1);
  if (r0 < np) {  // wave-uniform
    const int mr = lane & 15, ks = lane >> 4;
    const int ra = r0 + mr;
    const double* srow = S + (size_t)s.eidx[np + min(ra, np - 1)] * two + N;
    f64x4_t acc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) acc[j] = f64x4_t{0.0, 0.0, 0.0, 0.0};
    const int cbase = 16 * 4 * (wid & 1) + mr;
    for (int k0 = 0; k0 < np; k0 += 4) {
      const int k = k0 + ks;
      const double a = (ra < np && k < np) ? srow[s.eidx[np + k]] : 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int cb = cbase + 16 * j;
        const double b = (cb < np && k < np) ? s.T[k * np + cb] : 0.0;
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int cb = cbase + 16 * j;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int d = r0 + ks + 4 * r;
        if (d < np && cb < np) yh[(size_t)d * np + cb] = acc[j][r];
      }
    }
  }
  __syncthreads();
}

// Exchange ratios of the antiparallel bonds b < nb of the copy into rat
// (global scratch), four bonds per wave and pass, lanes over the up
// electrons u' and u' + 64: Xh = B[d] . Guu[., u] and W = Yh[d] . Guu[., u]
// by two four-way wave sums; Guu[., u] is read from row su + N of the Slater
// table (Slater[su+N][su'] = -Slater[su'][su+N] exactly with the block
// weights).
__device__ __attribute__((noinline)) void big_bond_ratios(const DevModel& m, const BigSmem& s, const double* S,
                                                          const double* yh, int nb, double* rat) {
  constexpr int BT = 4;
  const int N = m.N, np = m.npair, two = 2 * N;
  const int lane = threadIdx.x & (RBMPP_WAVE - 1), wid = threadIdx.x >> 6;
  const bool lv0 = lane < np, lv1 = lane + RBMPP_WAVE < np;
  const int su0 = lv0 ? s.eidx[lane] : 0, su1 = lv1 ? s.eidx[lane + RBMPP_WAVE] : 0;
  for (int p = wid; BT * p < nb; p += GNW) {
    int bb[BT], u[BT];
    double b0[BT], b1[BT], y0[BT], y1[BT], xs[BT], ws[BT], dl[BT], ep[BT], gd[BT];
#pragma unroll
    for (int t = 0; t < BT; t++) {
      bb[t] = min(BT * p + t, nb - 1);  // a missing bond repeats the last one (not written)
      const int i1 = s.bi1[bb[t]], i2 = s.bi2[bb[t]];
      const int msa = s.ecfg[i1], msb = s.ecfg[i2];
      const bool aup = msa < np;
      u[t] = __builtin_amdgcn_readfirstlane(aup ? msa : msb);
      const int d = __builtin_amdgcn_readfirstlane((aup ? msb : msa) - np);
      const int su = __builtin_amdgcn_readfirstlane(aup ? i1 : i2);
      const int sd = __builtin_amdgcn_readfirstlane(aup ? i2 : i1);
      const double* G = S + (size_t)(su + N) * two;
      const double g0 = lv0 ? -G[su0] : 0.0, g1 = lv1 ? -G[su1] : 0.0;
      b0[t] = lv0 ? s.T[d * np + lane] : 0.0;
      b1[t] = lv1 ? s.T[d * np + lane + RBMPP_WAVE] : 0.0;
      y0[t] = lv0 ? yh[(size_t)d * np + lane] : 0.0;
      y1[t] = lv1 ? yh[(size_t)d * np + lane + RBMPP_WAVE] : 0.0;
      xs[t] = fma(b1[t], g1, b0[t] * g0);
      ws[t] = fma(y1[t], g1, y0[t] * g0);
      const double* Su = S + (size_t)su * two;
      const double* Sd = S + (size_t)sd * two;
      gd[t] = Sd[sd + N];
      dl[t] = Sd[su + N] - gd[t];
      ep[t] = Su[su + N] - Su[sd + N];
    }
    // lanes 15, 47, 31, 63 hold bonds 0-3
    const double xv = wave_sum4(xs[0], xs[1], xs[2], xs[3]);
    const double wv = wave_sum4(ws[0], ws[1], ws[2], ws[3]);
    constexpr int SUMLANE[BT] = {15, 47, 31, 63};
#pragma unroll
    for (int t = 0; t < BT; t++) {
      if (BT * p + t >= nb) break;  // wave-uniform
      const double xh = readlane_d(xv, SUMLANE[t]), w = readlane_d(wv, SUMLANE[t]);
      const int ul = u[t] & (RBMPP_WAVE - 1);
      const bool hi = u[t] >= RBMPP_WAVE;
      const double bdu = readlane_d(hi ? b1[t] : b0[t], ul), ydu = readlane_d(hi ? y1[t] : y0[t], ul);
      if (lane == t) rat[bb[t]] = big_ratio(ydu, xh, w, bdu, dl[t], ep[t], gd[t]);
    }
  }
}

__global__ __launch_bounds__(GNT) void k_meas_blk_big(LocalArgs a) {
  __shared__ BigSmem s;
  const DevModel& m = *a.m;
  const int tid = threadIdx.x, nth = GNT;
  const int N = m.N, C = m.C, np = m.npair, npp = np * np;
  const bool want_eloc = a.mode & 1, want_o = a.mode & 2, want_inv = a.mode & 4, custom = a.mode & 8;
  const double pfsign = ((np * (np - 1) / 2) & 1) ? -1.0 : 1.0;  // Pf X = pfsign det F
  double* scr = a.scratch + (size_t)blockIdx.x * big_scratch(N);
  double* Bg = scr;                        // F^-1 of the current copy, natural order
  double* yh = Bg + (size_t)npp;           // Yh of the current copy
  double* rat = yh + (size_t)npp;          // [4N] ratios of the current copy
  double2* psib = (double2*)(rat + 4 * N);  // [4N] sums over copies of psi(x')
  double (*opp)[4 * GNMX] = (double (*)[4 * GNMX])(rat + 12 * N);  // [2][4N used] bins
  BigView v{opp, psib, s.x, s.ecfg, s.thr, s.thc, s.psiq};

  for (int cfg = blockIdx.x; cfg < a.nconf; cfg += gridDim.x) {
    __syncthreads();
    for (int i = tid; i < N; i += nth) s.x[i] = a.xs[(size_t)cfg * N + i];
    if (tid == 0) s.status = WS_OK;
    for (int k = tid; k < 4 * N; k += nth) {
      opp[0][k] = 0.0;
      opp[1][k] = 0.0;
      psib[k] = make_double2(0.0, 0.0);
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
    big_theta_sign(m, s);
    for (int q = tid; q < C; q += nth) s.psiwo[q] = psi_wo_pf(m, q, s.thc, s.thr, s.sgn[q >> 1]);
    const int nb = s.nb;
    bool singular = false;
    for (int q = 0; q < C; q++) {
      const int g = q >> 1;
      const double* S = m.se + (size_t)q * 4 * N * N;
      __syncthreads();  // the previous copy's readers of T are done
      const double det = rbmpp_block::wg_block_inverse_lds(S, N, np, s.eidx, s.T, s.rowb, s.colb, s.used, s.perm,
                                                           s.pinv, s.red, &s.s_p, &s.s_pv, &s.s_cnt, Bg, nullptr);
      if (det == 0.0) {  // the same in every thread
        singular = true;
        break;
      }
      const double pfq = pfsign * det;
      __syncthreads();  // Bg written by the workgroup; T is read back in natural order
      for (int e = tid; e < npp; e += nth) s.T[e] = Bg[e];
      if (tid == 0) {
        s.pf[q] = pfq;
        s.psiq[q] = cscale(s.psiwo[q], pfq);
      }
      __syncthreads();
      if (want_inv) {
        // reference layout out[i*N + j] = X^-1(i, j): X^-1(u, np+d) = -F^-1(d, u),
        // X^-1(np+d, u) = F^-1(d, u), zero same-spin blocks
        double* out = a.invm + ((size_t)cfg * C + q) * N * N;
        for (int e = tid; e < N * N; e += nth) {
          const int i = e / N, j = e - i * N;
          double val = 0.0;
          if (i < np && j >= np) val = -s.T[(j - np) * np + i];
          else if (i >= np && j < np) val = s.T[(i - np) * np + j];
          out[e] = val;
        }
      }
      if (want_o) {
        const double* Bq = s.T;
        wg_pp_oloc(m, v, q, pfq, s.psiwo[q], tid, nth, [&](int uu, int vv) -> double {
          return uu < np ? -Bq[(vv - np) * np + uu] : Bq[(uu - np) * np + vv];
        });
      }
      if (nb > 0) {
        big_yh(m, s, S, yh);
        big_bond_ratios(m, s, S, yh, nb, rat);
      }
      __syncthreads();
      // psi(x') of this copy for every bond, one thread per bond
      for (int b = tid; b < nb; b += nth) {
        const int j1 = s.bi1[b], j2 = s.bi2[b];
        int sfac = -1;
        if (m.subB[j1]) sfac = -sfac;
        if (m.subB[j2]) sfac = -sfac;
        const double pfn = rat[b] * pfq;
        const double2 c = big_exchanged_amp(m, q, s, sfac * s.sgn[g], j1, j2);
        psib[b] = cadd(psib[b], cscale(c, pfn));
        if (custom && a.ex_pf) a.ex_pf[(size_t)cfg * C + q] = pfn;
      }
    }
    if (singular) {
      if (tid == 0 && a.status) a.status[cfg] = WS_SINGULAR;
      continue;
    }
    __syncthreads();
    // amplitude, local energy, reweighting
    {
      double2 p = make_double2(0.0, 0.0);
      for (int q = 0; q < C; q++) p = cadd(p, s.psiq[q]);
      const double rw = a.shift >= 0.0 ? 1.0 / (cmod(p) + a.shift) : 1.0;
      double2 el = make_double2(0.0, 0.0);
      if (want_eloc && !custom) el = wg_eloc(m, v, p, s.bpos, s.red);
      if (tid == 0) {
        a.psi[cfg] = p;
        if (a.rw) a.rw[cfg] = rw;
        if (want_eloc && !custom) a.eloc[cfg] = cscale(el, rw);
        if (custom && a.ex_psi) a.ex_psi[cfg] = psib[0];
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
      wg_orow(m, v, s.selv[0], ore, ore + m.nv, cplx_rows);
    }
    if (tid == 0 && a.status) a.status[cfg] = WS_OK;
  }
}

}  // namespace

size_t meas_blk_big_scratch(int N) { return big_scratch(N); }

void launch_meas_blk_big(const LocalArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_meas_blk_big, dim3(grid), dim3(GNT), 0, st, a);
}
