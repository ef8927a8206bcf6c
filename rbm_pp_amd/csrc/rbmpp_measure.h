// Estimator pieces shared by the per-configuration kernel (rbmpp_local.hip)
// and the measurement fused into the sampler (rbmpp_sampler.hip).  S is a
// shared-memory block with members x, thr, thc, psiq (psi_q including Pf),
// psib (per antiparallel bond: sum over copies of psi(x')), bi1, bi2, opp.
#pragma once
#include "rbmpp_device.h"

// partner site of bond type t at site i1 (bond_map, calc_phys_quantity.f90:382-427)
__device__ __forceinline__ int bond_partner(int L, int t, int i1) {
  int ix = i1 / L, iy = i1 - ix * L;
  if (t == 0) { ix = (ix + 1) % L; }
  else if (t == 1) { iy = (iy + 1) % L; }
  else if (t == 2) { ix = (ix + 1) % L; iy = (iy + 1) % L; }
  else { ix = (ix + 1) % L; iy = (iy + L - 1) % L; }
  return ix * L + iy;
}

// antiparallel bonds in the order of calc_phys_quantity.f90:332-339
// (call from one thread); returns their number
template <class S>
__device__ int bond_list(const DevModel& m, S& s) {
  int nb = 0;
  for (int t = 0; t < 4; t++)
    for (int i1 = 0; i1 < m.N; i1++) {
      const int i2 = bond_partner(m.L, t, i1);
      if (s.x[i1] != s.x[i2]) { s.bi1[nb] = (short)i1; s.bi2[nb] = (short)i2; nb++; }
    }
  return nb;
}

// Electron lists (x_to_Ele_Idx_Ele_Cfg, interface_for_PP.f90:5-33) and, with
// bonds, the antiparallel bonds in the order of calc_phys_quantity.f90:332-339,
// by a workgroup of at least 4N threads: thread i takes site i and bond
// (t, i1) = (i / N, i % N); ranks come from wave ballots and the per-wave
// counts in wc (3 x waves ints of LDS).  bpos[e] receives the list position
// of bond e (-1 when parallel); s.nb the count.  Sets s.status on a
// configuration outside the Sz = 0 sector.  Ends in step.
template <class S>
__device__ void wg_ele_bonds(const DevModel& m, S& s, bool bonds, int* wc, short* bpos) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
  const int N = m.N, np = m.npair;
  const unsigned long long below = (1ull << lane) - 1ull;
  const bool site = tid < N;
  const bool up = site && s.x[tid] == 1;
  const unsigned long long bu = __ballot(up), bd = __ballot(site && !up);
  bool anti = false;
  int i1 = 0, i2 = 0;
  const bool inb = bonds && tid < 4 * N;
  if (inb) {
    const int t = tid / N;
    i1 = tid - t * N;
    i2 = bond_partner(m.L, t, i1);
    anti = s.x[i1] != s.x[i2];
  }
  const unsigned long long ba = __ballot(anti);
  if (lane == 0) {
    wc[wid] = __popcll(bu);
    wc[nw + wid] = __popcll(bd);
    wc[2 * nw + wid] = __popcll(ba);
  }
  __syncthreads();
  if (site) {
    int ou = 0, od = 0;
    for (int w = 0; w < wid; w++) { ou += wc[w]; od += wc[nw + w]; }
    if (up) {
      const int u = ou + __popcll(bu & below);
      if (u < np) { s.eidx[u] = tid; s.ecfg[tid] = u; }
    } else {
      const int d = np + od + __popcll(bd & below);
      if (d < N) { s.eidx[d] = tid; s.ecfg[tid] = d; }
    }
  }
  if (inb) {
    int b = -1;
    if (anti) {
      int o = 0;
      for (int w = 0; w < wid; w++) o += wc[2 * nw + w];
      b = o + __popcll(ba & below);
      s.bi1[b] = (short)i1;
      s.bi2[b] = (short)i2;
    }
    bpos[tid] = (short)b;
  }
  if (tid == 0) {
    int nu = 0, nb = 0;
    for (int w = 0; w < nw; w++) { nu += wc[w]; nb += wc[2 * nw + w]; }
    if (nu != np) s.status = WS_INIT;
    if (bonds) s.nb = nb;
  }
  __syncthreads();
}

// E_loc (calc_phys_quantity.f90:332-375) by the whole workgroup: thread e <
// 4N takes bond (t, i1) = (e / N, e % N), J s_i1 s_i2 psi plus, for the
// antiparallel bond at bpos[e] >= 0, 2 J sgn psi(x'); the terms are summed in
// a fixed order (wave sums, then the waves in order; red: 2 x waves doubles
// of LDS).  Every thread returns the sum.
template <class S>
__device__ double2 wg_eloc(const DevModel& m, const S& s, double2 p, const short* bpos, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
  const int N = m.N;
  double2 e = make_double2(0.0, 0.0);
  if (tid < 4 * N) {
    const int t = tid / N, i1 = tid - t * N, i2 = bond_partner(m.L, t, i1);
    const double J = t <= 1 ? m.J1 : m.J2;
    const double sg = t <= 1 ? -1.0 : 1.0;
    e = cscale(p, J * (double)s.x[i1] * (double)s.x[i2]);
    const int b = bpos[tid];
    if (b >= 0) e = cadd(e, cscale(s.psib[b], 2.0 * J * sg));
  }
  const double ex = wave_sum_to63(e.x), ey = wave_sum_to63(e.y);
  if (lane == 63) {
    red[2 * wid] = ex;
    red[2 * wid + 1] = ey;
  }
  __syncthreads();
  double2 r = make_double2(0.0, 0.0);
  for (int w = 0; w < nw; w++) r = cadd(r, make_double2(red[2 * w], red[2 * w + 1]));
  return r;
}

// Pair-product log-derivatives of copy q (calculate_derivative_Pfaffian_child,
// slater.f90:136-167, weighted by psi_wo_pf(q) = wq as calc_phys_quantity.f90:
// 209-216), summed in a fixed order: entry (u, v) of X^-1 lands in the
// orbital of (site(v), site(u)) with weight -cs, -ss, cc or cs by the spins
// of electrons u and v, so thread t0 + k tn owns orbital o and adds, over the
// site pairs (A, B) of o under the copy's operation (DevModel::pbin_*, in
// their fixed order), w(u, v) xinv(u, v) for v = Ele_Cfg(A), u = Ele_Cfg(B);
// opp[.][o] += sum wq.  xinv(u, v) returns X^-1(u, v) and is called for
// nonzero weights only.  Every orbital keeps its owner in every copy, so the
// sums need no atomics and are the same from run to run.
template <class S, class XI>
__device__ void wg_pp_oloc(const DevModel& m, S& s, int q, double pfq, double2 wq, int t0, int tn, XI xinv) {
  const int g = q >> 1, sidx = q & 1, N = m.N, np = m.npair;
  const double cs = pfq * m.sp_cs[sidx], cc = pfq * m.sp_cc[sidx], ss = pfq * m.sp_ss[sidx];
  const int* off = m.pbin_off + (size_t)g * (m.nsl + 1);
  const unsigned* pr = m.pbin_pairs + (size_t)g * N * (N - 1);
  for (int o = t0; o < m.nsl; o += tn) {
    double acc = 0.0;
    const int k1 = off[o + 1];
#pragma unroll 4
    for (int k = off[o]; k < k1; k++) {
      const unsigned ab = pr[k];
      const int v = s.ecfg[ab & 0xffffu], u = s.ecfg[ab >> 16];
      const double w = u < np ? (v < np ? -cs : -ss) : (v < np ? cc : cs);
      if (w != 0.0) acc = fma(w, xinv(u, v), acc);
    }
    s.opp[0][o] += acc * wq.x;
    if (!m.real_only) s.opp[1][o] += acc * wq.y;
  }
}

// E_loc = sum_bonds J s_i s_j psi + sum_antiparallel 2 J sgn psi(x')
// (calc_phys_quantity.f90:332-375), accumulated in the reference's order
template <class S>
__device__ double2 eloc_from_bonds(const DevModel& m, const S& s, double2 p) {
  double2 e = make_double2(0.0, 0.0);
  int b = 0;
  for (int t = 0; t < 4; t++) {
    const double J = t <= 1 ? m.J1 : m.J2;
    const double sg = t <= 1 ? -1.0 : 1.0;
    for (int i1 = 0; i1 < m.N; i1++) {
      const int i2 = bond_partner(m.L, t, i1);
      const double d = J * (double)s.x[i1] * (double)s.x[i2];
      e = cadd(e, cscale(p, d));
      if (s.x[i1] != s.x[i2]) { e = cadd(e, cscale(s.psib[b], 2.0 * J * sg)); b++; }
    }
  }
  return e;
}

// One O row: O_loc(k) = d psi / d gamma_k (calc_phys_quantity.f90:114-222)
// times rw.  The pair-product entries come from s.opp (re, im).
template <class S>
__device__ void wg_orow(const DevModel& m, const S& s, double rw, double* ore, double* oim, bool cplx_rows) {
  const int N = m.N, C = m.C, nv = m.nv;
  const int ac = m.ac, ar = m.ar;
  const int offr = 2 * ac * (N + 1);
  for (int k = threadIdx.x; k < nv; k += blockDim.x) {
    double2 o = make_double2(0.0, 0.0);
    if (k < 2 * ac) {
      // complex bias f: k = 2f (re), 2f+1 (im) = i * re
      const int f = k >> 1;
      for (int q = 0; q < C; q++) {
        const double2 t = s.thc[q * ac + f];
        const double th = tanh(t.x), tn = tan(t.y);
        // (th + i tn) / (1 + i th tn)
        const double den = 1.0 + (th * tn) * (th * tn);
        const double2 c2 = make_double2((th + tn * (th * tn)) / den, (tn - th * (th * tn)) / den);
        o = cadd(o, cmul(c2, s.psiq[q]));
      }
      if (k & 1) o = make_double2(-o.y, o.x);
    } else if (k < offr) {
      const int kk = k - 2 * ac, f = kk / (2 * N), iw = (kk - f * 2 * N) >> 1;
      for (int q = 0; q < C; q++) {
        const int g = q >> 1;
        const double sp = (q & 1) ? -1.0 : 1.0;
        const double2 t = s.thc[q * ac + f];
        const double th = tanh(t.x), tn = tan(t.y);
        const double den = 1.0 + (th * tn) * (th * tn);
        const double2 c2 = make_double2((th + tn * (th * tn)) / den, (tn - th * (th * tn)) / den);
        const double2 c1 = cmul(c2, s.psiq[q]);
        const int i = m.invmap[g * N + iw];
        o = cadd(o, cscale(c1, (double)s.x[i] * sp));
      }
      if (kk & 1) o = make_double2(-o.y, o.x);
    } else if (k < offr + ar) {
      const int f = k - offr;
      for (int q = 0; q < C; q++) o = cadd(o, cscale(s.psiq[q], tanh(s.thr[q * ar + f])));
    } else if (k < m.nv_rbm) {
      const int kk = k - offr - ar, f = kk / N, iw = kk - f * N;
      for (int q = 0; q < C; q++) {
        const int g = q >> 1;
        const double sp = (q & 1) ? -1.0 : 1.0;
        const int i = m.invmap[g * N + iw];
        o = cadd(o, cscale(cscale(s.psiq[q], tanh(s.thr[q * ar + f])), (double)s.x[i] * sp));
      }
    } else {
      const int orb = k - m.nv_rbm;
      o = make_double2(s.opp[0][orb], s.opp[1][orb]);
    }
    ore[k] = o.x * rw;
    if (cplx_rows) oim[k] = o.y * rw;
  }
}

// One configuration's share of the correlation run's sums
// (Calc_Correlation/calc_phys_quantity.f90:318-338), by the whole workgroup:
// SdotS per bond (calc_local_energy_correlation, :348-413) reweighted by rw,
// Chi_vbs += Re(conj SdotS(i1,t) SdotS(i2,t)), Chi_sp(1, .) and the diagonal
// of Chi_sp(2, .) (calc_local_correlation, :469-550) times rw^2, and wf_sgn.
// cpart is the workgroup's slice (corr_len layout); the SdotS table takes
// s.opp.  The exchange part of Chi_sp(2, .) (off-diagonal) is left to the
// caller, which holds the pair amplitudes; every entry this adds to is owned by
// one thread of the workgroup in every configuration.
template <class S>
__device__ void corr_accumulate(const DevModel& m, S& s, double2 p, double rw, int nb, double* cpart, int nchi) {
  const int N = m.N, tid = threadIdx.x, nth = blockDim.x;
  const double w2 = rw * rw, ap2 = p.x * p.x + p.y * p.y;
  double2* sdb = (double2*)&s.opp[0][0];  // SdotS(i1, t) at sdb[t * N + i1]
  for (int e = tid; e < 4 * N; e += nth) {
    const int t = e / N, i1 = e - t * N;
    sdb[e] = cscale(p, (double)(s.x[i1] * s.x[bond_partner(m.L, t, i1)]));
  }
  __syncthreads();
  for (int b = tid; b < nb; b += nth) {
    const int i1 = s.bi1[b], i2 = s.bi2[b];
    int t = 0;
    while (t < 3 && bond_partner(m.L, t, i1) != i2) t++;  // partners are distinct for L >= 3
    sdb[t * N + i1] = cadd(sdb[t * N + i1], cscale(s.psib[b], t <= 1 ? -2.0 : 2.0));
  }
  __syncthreads();
  for (int e = tid; e < 4 * N; e += nth) sdb[e] = cscale(sdb[e], rw);
  __syncthreads();
  double* sp1 = cpart;
  double* sp2 = cpart + nchi;
  double* vbs = cpart + 2 * nchi;
  for (int e = tid; e < N * N; e += nth) {
    const int i2 = e / N, i1 = e - i2 * N;
    if (i1 > i2) continue;
    const int i0 = i2 * (i2 + 1) / 2 + i1;
    sp1[i0] += w2 * ((double)(s.x[i1] * s.x[i2]) * ap2);
    if (i1 == i2) sp2[i0] += w2 * ap2;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const double2 u = sdb[t * N + i1], v = sdb[t * N + i2];
      vbs[t * nchi + i0] += u.x * v.x + u.y * v.y;
    }
  }
  if (tid == 0) {
    const double ap = cmod(p);
    cpart[6 * nchi] += p.x / ap;
    cpart[6 * nchi + 1] += p.y / ap;
  }
}
