// Metropolis sampler for the RBM x pair-product wave function on gfx950.
//
// One workgroup owns one walker (Markov chain) for a whole make_sample call
// (make_sample.f90:4-240): thermalisation and stored sweeps run inside one
// launch.  Wave w of the workgroup owns projection copy q = w: per proposal
// it streams that copy's X^-1 (N x N FP64) once to form P = X^-T a and
// Q = X^-T b, which give the Pfaffian ratio of the spin exchange
// (pfupdate_two.f90:27-88); on acceptance the same wave applies the rank-4
// antisymmetric update (pfupdate_two.f90:113-224) in one read-modify-write
// pass, reusing P and Q.  All threads draw the same Philox numbers, so the
// control flow is uniform across the workgroup.
#include "rbmpp_kernels.h"

// ------------------------------------------------------------------
// update_slater_elements (slater.f90:7-64).  se[q][c][r] = Slater_Elm(r,c,q).
__global__ void k_slater_elements(DevModel m, const double* __restrict__ f, double* se, int* bad) {
  const int N = m.N;
  const int q = blockIdx.y;
  const int g = q >> 1, sidx = q & 1;
  const double cs = m.sp_cs[sidx], cc = m.sp_cc[sidx], ss = m.sp_ss[sidx];
  const size_t two = 2 * (size_t)N;
  double* S = se + (size_t)q * two * two;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N * N; e += gridDim.x * blockDim.x) {
    const int ri = e / N, rj = e - ri * N;
    const int tri = m.wmap[g * N + ri], trj = m.wmap[g * N + rj];
    const double slt_ij = f[m.orb[tri * N + trj]];
    const double slt_ji = f[m.orb[trj * N + tri]];
    const double d = slt_ij - slt_ji;
    // Slater_Elm(rj, ri), (rj+N, ri), (rj, ri+N), (rj+N, ri+N)
    S[(size_t)ri * two + rj] = -d * cs;
    S[(size_t)ri * two + rj + N] = slt_ij * cc + slt_ji * ss;
    S[(size_t)(ri + N) * two + rj] = -slt_ij * ss - slt_ji * cc;
    S[(size_t)(ri + N) * two + rj + N] = d * cs;
    // the mirrored entries are written by the (rj, ri) thread; skew symmetry
    // is exact by construction (negation commutes with rounding)
    (void)bad;
  }
}

void launch_slater(const DevModel& m, const double* f, double* se, int* bad, hipStream_t s) {
  dim3 grid((m.N * m.N + 255) / 256, m.C);
  hipLaunchKernelGGL(k_slater_elements, grid, dim3(256), 0, s, m, f, se, bad);
}

// ------------------------------------------------------------------
// shared-memory layout of the sampler workgroup
template <int NMAX, bool LDSMAT>
struct SweepSmem {
  static constexpr int NW = 16;
  static constexpr int PQ = 2 * RBMPP_CMAX * NMAX;  // per-copy P, Q
  static constexpr int AB = 2 * NW * NMAX;          // per-wave gathered a, b (then s, t)
  static constexpr int MAT = LDSMAT ? (NMAX < RBMPP_NLDS ? NMAX * NMAX : RBMPP_NLDS * RBMPP_NLDS) : 0;
  static constexpr int UNION = (PQ + AB) > MAT ? (PQ + AB) : MAT;
  double u[UNION];
  double thr[RBMPP_MRMAX];
  double2 thc[RBMPP_MCMAX];
  double pf[RBMPP_CMAX], pfn[RBMPP_CMAX], bma[RBMPP_CMAX], vba[RBMPP_CMAX], invab[RBMPP_CMAX];
  double2 psiq[RBMPP_CMAX];
  double gjbuf[4 * NMAX];
  double red[32];
  double2 psi;
  int x[NMAX], eidx[NMAX], ecfg[NMAX], perm[NMAX];
  int sgn[RBMPP_CMAX / 2];
  unsigned sgnbits[RBMPP_CMAX / 2];
  int redi[32];
  int piv, flag, status;
  unsigned long long ctr;
};

// ------------------------------------------------------------------
// Full recompute of a walker from s.x and the current electron order
// s.eidx (make_sample.f90:78-97): theta (calc_theta), sign_PP, per-copy
// Pf and X^-1 (calculate_MAll), amplitude.  With check=1 the recomputed
// theta and sign are compared with the incremental values
// (make_sample.f90:84-92) and the Pf / X^-1 drift is returned.
template <int NMAX, bool LDSMAT>
__device__ void wg_recompute(const DevModel& m, SweepSmem<NMAX, LDSMAT>& s, double* Mw,
                             double* gscratch, int check, double* drift_pf, double* drift_inv) {
  const int tid = threadIdx.x, nth = blockDim.x;
  const int N = m.N, C = m.C, P = m.P;
  wg_theta_sign(m, s, check);
  // Pf and X^-1 of each copy
  double* A = LDSMAT ? s.u : gscratch;
  const int ld = N;
  double dpf = 0.0, dinv = 0.0;
  for (int q = 0; q < C; q++) {
    const double* S = m.se + (size_t)q * 4 * N * N;
    __syncthreads();
    // X(i,j) = -Slater_Elm(r_i, r_j) = Slater_Elm(r_j, r_i) (matrix.f90:72-78)
    for (int e = tid; e < N * N; e += nth) {
      const int i = e / N, j = e - i * N;
      const int ri = s.eidx[i] + (i >= m.npair ? N : 0);
      const int rj = s.eidx[j] + (j >= m.npair ? N : 0);
      A[(size_t)i * ld + j] = S[(size_t)ri * 2 * N + rj];
    }
    __syncthreads();
    const double pfq = wg_skew_inverse(A, N, ld, s.perm, s.gjbuf, s.red, s.redi, &s.piv);
    if (pfq == 0.0) {
      if (tid == 0) s.status = WS_SINGULAR;
      __syncthreads();
      return;
    }
    // inverse permutation into gjbuf (as ints)
    int* ip = (int*)s.gjbuf;
    for (int r = tid; r < N; r += nth) ip[s.perm[r]] = r;
    __syncthreads();
    // M[j][i] = InvM(j,i) = 0.5 (Xinv(i,j) - Xinv(j,i))   (matrix.f90:97-106)
    double* Mq = Mw + (size_t)q * N * N;
    double lmax = 0.0, ldiff = 0.0;
    for (int e = tid; e < N * N; e += nth) {
      const int j = e / N, i = e - j * N;
      const double xij = A[(size_t)ip[i] * ld + ip[j]];
      const double xji = A[(size_t)ip[j] * ld + ip[i]];
      const double v = 0.5 * (xij - xji);
      if (check) ldiff = fmax(ldiff, fabs(Mq[e] - v));
      lmax = fmax(lmax, fabs(v));
      Mq[e] = v;
    }
    if (check) {
      // block max of (ldiff, lmax)
      double a = ldiff, b = lmax;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        a = fmax(a, __shfl_xor(a, o, RBMPP_WAVE));
        b = fmax(b, __shfl_xor(b, o, RBMPP_WAVE));
      }
      const int lane = tid & 63, wid = tid >> 6;
      __syncthreads();
      if (lane == 0) { s.red[2 * wid] = a; s.red[2 * wid + 1] = b; }
      __syncthreads();
      if (tid == 0) {
        double A1 = 0.0, B1 = 0.0;
        for (int w = 0; w < (nth >> 6); w++) { A1 = fmax(A1, s.red[2 * w]); B1 = fmax(B1, s.red[2 * w + 1]); }
        s.red[30] = A1 / B1;
      }
      __syncthreads();
      dinv = fmax(dinv, s.red[30]);
    }
    if (tid == 0) {
      s.pfn[q] = s.pf[q];  // old value, for the drift measure
      s.pf[q] = pfq;
    }
  }
  __syncthreads();
  if (check) {
    double av = 0.0;
    for (int q = 0; q < C; q++) av += fabs(s.pf[q]);
    av /= (double)C;
    for (int q = 0; q < C; q++) dpf = fmax(dpf, fabs(s.pfn[q] - s.pf[q]) / av);
  }
  // amplitude (sub.f90:28-59)
  for (int q = tid; q < C; q += nth) s.psiq[q] = cscale(psi_wo_pf(m, q, s.thc, s.thr, s.sgn[q >> 1]), s.pf[q]);
  __syncthreads();
  if (tid == 0) {
    double2 p = make_double2(0.0, 0.0);
    for (int q = 0; q < C; q++) p = cadd(p, s.psiq[q]);
    s.psi = p;
  }
  __syncthreads();
  if (drift_pf) *drift_pf = dpf;
  if (drift_inv) *drift_inv = dinv;
}

// x_to_Ele_Idx_Ele_Cfg (interface_for_PP.f90:5-33), serial on thread 0
template <int NMAX, bool LDSMAT>
__device__ void wg_x_to_ele(const DevModel& m, SweepSmem<NMAX, LDSMAT>& s) {
  if (threadIdx.x == 0) {
    int i1 = 0, i2 = m.npair;
    for (int i = 0; i < m.N; i++) {
      if (s.x[i] == 1) { s.eidx[i1] = i; s.ecfg[i] = i1; i1++; }
      else { s.eidx[i2] = i; s.ecfg[i] = i2; i2++; }
    }
  }
  __syncthreads();
}

// make_initial_x (make_sample.f90:298-362): random Sz=0 configuration with
// |psi| > 1e-300.  The serial draws run on thread 0; the counter is then
// broadcast so every thread's stream copy stays in step.
template <int NMAX, bool LDSMAT>
__device__ int wg_initial_x(const DevModel& m, SweepSmem<NMAX, LDSMAT>& s, Rng& rng, double* Mw,
                            double* gscratch) {
  for (int attempt = 0; attempt < 1000; attempt++) {
    if (threadIdx.x == 0) {
      for (int i = 0; i < m.N; i++) s.x[i] = -1;
      int cnt = 0;
      for (int k = 0; k < 1000000 && cnt < m.N / 2; k++) {
        const int i = rng.nran0(m.N);
        if (s.x[i] == -1) { s.x[i] = 1; cnt++; }
      }
      s.ctr = rng.ctr;
      s.flag = (cnt == m.N / 2);
    }
    __syncthreads();
    rng.ctr = s.ctr;
    if (!s.flag) return WS_INIT;
    wg_x_to_ele(m, s);
    wg_recompute(m, s, Mw, gscratch, 0, nullptr, nullptr);
    if (s.status) return s.status;
    if (cabs2(s.psi) > 1e-300) return WS_OK;
  }
  return WS_INIT;
}

// ------------------------------------------------------------------
template <int NMAX, bool LDSMAT>
__global__ __launch_bounds__(1024) void k_sweep(SweepArgs a) {
  constexpr int KMAX = NMAX / RBMPP_WAVE;
  __shared__ SweepSmem<NMAX, LDSMAT> s;
  const DevModel& m = a.m;
  const int w = blockIdx.x;
  if (w >= a.nwalk) return;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & (RBMPP_WAVE - 1), wid = tid >> 6, nwv = nth >> 6;
  const int N = m.N, C = m.C, P = m.P;
  const size_t NN = (size_t)N * N;
  double* Mw = a.ws.minv + (size_t)w * C * NN;
  double* gscr = a.ws.scratch ? a.ws.scratch + (size_t)w * NN : nullptr;

  Rng rng;
  rng.k0 = a.key0; rng.k1 = a.key1; rng.stream = (uint32_t)w;
  rng.ctr = a.ws.ctr[w];
  if (tid == 0) s.status = WS_OK;
  for (int i = tid; i < C * m.ar; i += nth) s.thr[i] = a.ws.thr[(size_t)w * m.mr + i];
  for (int i = tid; i < C * m.ac; i += nth) s.thc[i] = a.ws.thc[(size_t)w * m.mc + i];
  for (int q = tid; q < C; q += nth) s.pf[q] = a.ws.pf[(size_t)w * C + q];
  for (int g = tid; g < P; g += nth) s.sgn[g] = a.ws.sgn[(size_t)w * P + g];
  __syncthreads();

  int status = WS_OK;
  // ---- set-up (make_sample.f90:50-62)
  if (a.first) {
    status = wg_initial_x(m, s, rng, Mw, gscr);
  } else {
    for (int i = tid; i < N; i += nth) s.x[i] = a.ws.x[(size_t)w * N + i];
    __syncthreads();
    wg_x_to_ele(m, s);
    wg_recompute(m, s, Mw, gscr, 0, nullptr, nullptr);
    status = s.status;
    if (status == WS_OK && rng.u01() < a.p_restart) status = wg_initial_x(m, s, rng, Mw, gscr);
  }
  int nacc_tmp = 0;
  int nfast = a.ws.nfastup[w];
  long long naccept = 0;
  int nrec = 0;
  double* PQ = s.u;                                   // [C][2][NMAX]
  double* AB = s.u + 2 * RBMPP_CMAX * NMAX;           // [NW][2][NMAX]
  double* va = AB + (size_t)wid * 2 * NMAX;
  double* vb = va + NMAX;

  for (int i0 = -a.nwarm; i0 <= a.nsamp && status == WS_OK; i0++) {
    // ---- periodic recompute (make_sample.f90:74-128)
    const int limit = a.fixed_k > 0 ? a.fixed_k : nfast;
    if (nacc_tmp > limit) {
      double dpf = 0.0, dinv = 0.0;
      wg_recompute(m, s, Mw, gscr, 1, &dpf, &dinv);
      status = s.status;
      nrec++;
      nacc_tmp = 0;
      if (a.fixed_k <= 0) {
        if (dinv > 1.0e-5 || dpf > 1.0e-5) nfast = max(1, (int)llround((double)nfast / 1.2));
        else nfast = min(100, (int)llround((double)nfast * 1.1));
      }
      if (status) break;
    }
    for (int iu = 0; iu < a.nupdate; iu++) {
      // ---- global spin flip (make_sample.f90:135-158)
      if (rng.u01() < a.p_global) {
        const double2 psi_old = s.psi;
        for (int i = tid; i < N; i += nth) s.x[i] = -s.x[i];
        __syncthreads();
        wg_x_to_ele(m, s);
        wg_recompute(m, s, Mw, gscr, 0, nullptr, nullptr);
        if (s.status) { status = s.status; break; }
        const double r1 = (cabs2(s.psi) + a.shift) / (cabs2(psi_old) + a.shift);
        if (!(rng.u01() < r1 * r1)) {
          for (int i = tid; i < N; i += nth) s.x[i] = -s.x[i];
          __syncthreads();
          wg_x_to_ele(m, s);
          wg_recompute(m, s, Mw, gscr, 0, nullptr, nullptr);
          if (s.status) { status = s.status; break; }
        }
      }
      // ---- exchange candidate (make_sample.f90:244-294)
      int i1 = -1, i2 = -1;
      {
        const int dist = rng.nran0(m.L - 1) + 2;
        for (int it = 0; it < 1000000; it++) {
          const int c1 = rng.nran0(N);
          int ix = c1 / m.L, iy = c1 - (c1 / m.L) * m.L;
          ix += rng.nran0(dist);
          iy += rng.nran0(dist);
          if (ix >= m.L) ix -= m.L;
          if (iy >= m.L) iy -= m.L;
          const int c2 = ix * m.L + iy;
          if (s.x[c1] != s.x[c2]) { i1 = c1; i2 = c2; break; }
        }
      }
      if (i1 < 0) { status = WS_CANDIDATE; break; }
      const int msa = s.ecfg[i1], msb = s.ecfg[i2];
      const int x1 = s.x[i1], x2 = s.x[i2];
      int sfac = -1;
      if (m.subB[i1]) sfac = -sfac;
      if (m.subB[i2]) sfac = -sfac;
      const int rsa = i2 + (msa >= m.npair ? N : 0);  // new row of electron msa
      const int rsb = i1 + (msb >= m.npair ? N : 0);
      // ---- ratio per copy (wave per copy)
      for (int q0 = 0; q0 < C; q0 += nwv) {
        const int q = q0 + wid;
        if (q < C) {
          const double* S = m.se + (size_t)q * 4 * N * N;
          const double* colA = S + (size_t)rsa * 2 * N;
          const double* colB = S + (size_t)rsb * 2 * N;
          for (int i = lane; i < N; i += RBMPP_WAVE) {
            int t = s.eidx[i];
            if (i == msa) t = i2;
            else if (i == msb) t = i1;
            const int r = t + (i >= m.npair ? N : 0);
            va[i] = colA[r];
            vb[i] = colB[r];
          }
        }
        __syncthreads();
        if (q < C) {
          const double* Mq = Mw + (size_t)q * NN;
          double Pk[KMAX], Qk[KMAX];
#pragma unroll
          for (int k = 0; k < KMAX; k++) { Pk[k] = 0.0; Qk[k] = 0.0; }
#pragma unroll 4
          for (int j = 0; j < N; j++) {
            const double aj = va[j], bj = vb[j];
            const double* row = Mq + (size_t)j * N;
#pragma unroll
            for (int k = 0; k < KMAX; k++) {
              const int i = lane + k * RBMPP_WAVE;
              const double mv = (i < N) ? row[i] : 0.0;
              Pk[k] = fma(mv, aj, Pk[k]);
              Qk[k] = fma(mv, bj, Qk[k]);
            }
          }
          double* Pq = PQ + (size_t)q * 2 * NMAX;
          double* Qq = Pq + NMAX;
          double part = 0.0;
#pragma unroll
          for (int k = 0; k < KMAX; k++) {
            const int i = lane + k * RBMPP_WAVE;
            if (i < N) { Pq[i] = Pk[k]; Qq[i] = Qk[k]; part = fma(vb[i], Pk[k], part); }
          }
          const double bMa = wave_sum(part);
          if (lane == 0) s.bma[q] = bMa;
        }
        __syncthreads();
        if (q < C && lane == 0) {
          const double* Mq = Mw + (size_t)q * NN;
          const double* Pq = PQ + (size_t)q * 2 * NMAX;
          const double* Qq = Pq + NMAX;
          const double invm_ab = Mq[(size_t)msb * N + msa];
          const double vec_ba = vb[msa];
          const double bMa = s.bma[q];
          const double ratio = invm_ab * vec_ba + invm_ab * bMa + Pq[msa] * Qq[msb] - Pq[msb] * Qq[msa];
          s.vba[q] = vec_ba;
          s.invab[q] = invm_ab;
          s.pfn[q] = ratio * s.pf[q];
          const double2 c = psi_wo_pf_exchanged(m, q, s.thc, s.thr, sfac * s.sgn[q >> 1], i1, i2, x1, x2);
          s.psiq[q] = cscale(c, s.pfn[q]);
        }
        __syncthreads();
      }
      // ---- Metropolis test (make_sample.f90:203-204)
      double2 psin = make_double2(0.0, 0.0);
      for (int q = 0; q < C; q++) psin = cadd(psin, s.psiq[q]);
      const double r1 = (cabs2(psin) + a.shift) / (cabs2(s.psi) + a.shift);
      if (rng.u01() < r1 * r1) {
        // ---- accept: rank-4 update of every copy's X^-1
        for (int q0 = 0; q0 < C; q0 += nwv) {
          const int q = q0 + wid;
          double Ai[KMAX], Bi[KMAX], Ci[KMAX], Di[KMAX], R1i[KMAX], R2i[KMAX];
          double* Mq = Mw + (size_t)q * NN;
          const double* Pq = PQ + (size_t)q * 2 * NMAX;
          const double* Qq = Pq + NMAX;
          double a_ = 0, b_ = 0, c_ = 0, d_ = 0, fID = 0;
          if (q < C) {
            const double Pa = Pq[msa], Pb = Pq[msb], Qa = Qq[msa], Qb = Qq[msb];
            a_ = -Pa; b_ = Pb; c_ = Qa; d_ = -Qb;
            const double e_ = -s.bma[q] - s.vba[q];
            const double f_ = s.invab[q];
            const double Det = a_ * d_ - b_ * c_ - e_ * f_;
            const double InvDet = 1.0 / Det;
            fID = f_ * InvDet;
            const double eDet = e_ * Det;
            // s_i = InvDet * InvM(i,msa) = -InvDet * InvM(msa,i); t_i likewise
            for (int i = lane; i < N; i += RBMPP_WAVE) {
              va[i] = -InvDet * Mq[(size_t)msa * N + i];
              vb[i] = -InvDet * Mq[(size_t)msb * N + i];
            }
#pragma unroll
            for (int k = 0; k < KMAX; k++) {
              const int i = lane + k * RBMPP_WAVE;
              Ai[k] = Bi[k] = Ci[k] = Di[k] = R1i[k] = R2i[k] = 0.0;
              if (i < N) {
                const double p_i = Pq[i], q_i = Qq[i];
                const double s_i = -InvDet * Mq[(size_t)msa * N + i];
                const double t_i = -InvDet * Mq[(size_t)msb * N + i];
                Ai[k] = a_ * q_i + c_ * p_i + eDet * s_i;
                Bi[k] = b_ * q_i + d_ * p_i - eDet * t_i;
                Ci[k] = -a_ * t_i - b_ * s_i + fID * p_i;
                Di[k] = -c_ * t_i - d_ * s_i - fID * q_i;
                if (i == msa) { Ai[k] += c_; Bi[k] += d_; Ci[k] += fID; }
                if (i == msb) { Ai[k] += a_; Bi[k] += b_; Di[k] -= fID; }
                R1i[k] = c_ * t_i + d_ * s_i + fID * q_i;
                R2i[k] = a_ * t_i + b_ * s_i - fID * p_i;
              }
            }
          }
          __syncthreads();
          if (q < C) {
            for (int j = 0; j < N; j++) {
              const double tj = vb[j], sj = va[j], qj = Qq[j], pj = Pq[j];
              double* row = Mq + (size_t)j * N;
#pragma unroll
              for (int k = 0; k < KMAX; k++) {
                const int i = lane + k * RBMPP_WAVE;
                if (i < N) {
                  double v = row[i];
                  v = fma(Ai[k], tj, v);
                  v = fma(Bi[k], sj, v);
                  v = fma(Ci[k], qj, v);
                  v = fma(Di[k], pj, v);
                  if (j == msa) v -= R1i[k];
                  if (j == msb) v -= R2i[k];
                  if (j == msa && i == msb) v -= fID;
                  if (j == msb && i == msa) v += fID;
                  row[i] = v;
                }
              }
            }
            // angles of this copy (update_theta)
            const int g = q >> 1;
            const double sp = (q & 1) ? -1.0 : 1.0;
            const int iw1 = m.wmap[g * N + i1], iw2 = m.wmap[g * N + i2];
            for (int f = lane; f < m.ar; f += RBMPP_WAVE) {
              double t = s.thr[q * m.ar + f];
              t = t - 2.0 * m.Wr[f * N + iw1] * (double)x1 * sp;
              t = t - 2.0 * m.Wr[f * N + iw2] * (double)x2 * sp;
              s.thr[q * m.ar + f] = t;
            }
            for (int f = lane; f < m.ac; f += RBMPP_WAVE) {
              double2 t = s.thc[q * m.ac + f];
              const double2 w1 = m.Wc[f * N + iw1], w2 = m.Wc[f * N + iw2];
              const double a1 = 2.0 * (double)x1 * sp, a2 = 2.0 * (double)x2 * sp;
              t.x = t.x - w1.x * a1; t.y = t.y - w1.y * a1;
              t.x = t.x - w2.x * a2; t.y = t.y - w2.y * a2;
              s.thc[q * m.ac + f] = t;
            }
            if (lane == 0) s.pf[q] = s.pfn[q];
          }
          __syncthreads();
        }
        if (tid == 0) {
          s.x[i1] = -x1;
          s.x[i2] = -x2;
          s.ecfg[i1] = msb;
          s.ecfg[i2] = msa;
          s.eidx[msa] = i2;
          s.eidx[msb] = i1;
          for (int g = 0; g < P; g++) s.sgn[g] = sfac * s.sgn[g];
          s.psi = psin;
        }
        __syncthreads();
        if (i0 > 0) naccept++;
        nacc_tmp++;
      }
    }
    if (status) break;
    // ---- store the sample (make_sample.f90:228-232)
    if (i0 > 0) {
      const size_t row = (size_t)w * a.nsamp + (i0 - 1);
      for (int i = tid; i < N; i += nth) a.xs[row * N + i] = (signed char)s.x[i];
      if (tid == 0) a.psis[row] = s.psi;
    }
  }
  __syncthreads();
  // ---- write the walker state back
  for (int i = tid; i < N; i += nth) {
    a.ws.x[(size_t)w * N + i] = (signed char)s.x[i];
    a.ws.eidx[(size_t)w * N + i] = s.eidx[i];
    a.ws.ecfg[(size_t)w * N + i] = s.ecfg[i];
  }
  for (int i = tid; i < C * m.ar; i += nth) a.ws.thr[(size_t)w * m.mr + i] = s.thr[i];
  for (int i = tid; i < C * m.ac; i += nth) a.ws.thc[(size_t)w * m.mc + i] = s.thc[i];
  for (int q = tid; q < C; q += nth) a.ws.pf[(size_t)w * C + q] = s.pf[q];
  for (int g = tid; g < P; g += nth) a.ws.sgn[(size_t)w * P + g] = s.sgn[g];
  if (tid == 0) {
    a.ws.psi[w] = s.psi;
    a.ws.ctr[w] = rng.ctr;
    a.ws.nacc_tmp[w] = nacc_tmp;
    a.ws.nfastup[w] = nfast;
    a.ws.status[w] = status;
    a.ws.naccept[w] = naccept;
    a.ws.nrecompute[w] = nrec;
  }
}

template <int NMAX, bool LDSMAT>
static void launch_sweep_t(const SweepArgs& a, hipStream_t st) {
  const int nw = a.m.C < 16 ? a.m.C : 16;
  hipLaunchKernelGGL((k_sweep<NMAX, LDSMAT>), dim3(a.nwalk), dim3(64 * nw), 0, st, a);
}

int launch_sweep(const SweepArgs& a, hipStream_t st) {
  const int N = a.m.N;
  if (a.m.C > RBMPP_CMAX || a.m.mr > RBMPP_MRMAX || a.m.mc > RBMPP_MCMAX) return -1;
  if (N <= 64) launch_sweep_t<64, true>(a, st);
  else if (N <= RBMPP_NLDS) launch_sweep_t<128, true>(a, st);
  else if (N <= 256) launch_sweep_t<256, false>(a, st);
  else return -1;
  return 0;
}
