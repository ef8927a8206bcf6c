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
