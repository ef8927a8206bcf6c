// Metropolis sampler for the RBM x pair-product wave function on gfx950.
//
// One workgroup owns one walker (Markov chain) for a whole make_sample call
// (make_sample.f90:4-240): thermalisation and stored sweeps run inside one
// launch.  Wave w of the workgroup owns projection copy q = w: per proposal
// it streams that copy's X^-1 once to form P = X^-T a and Q = X^-T b, which
// give the Pfaffian ratio of the spin exchange (pfupdate_two.f90:27-88); on
// acceptance the same wave applies the rank-4 antisymmetric update
// (pfupdate_two.f90:113-224) in one read-modify-write pass, reusing P and Q.
// X^-1 is held in folded skew storage (N^2/2 FP64, rbmpp_device.h), so each
// pass moves half the bytes of the full matrix and the walkers resident on
// the chip keep their X^-1 within the last-level cache.  In block form
// (normal projection, BLK) each wave keeps F^-1 of its copy ((N/2)^2 FP64),
// the ratio is det F'/det F, and an accepted move's rank-2 update is applied
// by the next proposal's ratio pass as it reads each row (one pass over F^-1
// per proposal).  All threads draw the same Philox numbers, so the control
// flow is uniform across the workgroup and runs on the scalar unit.
#include <algorithm>
#include <cstdlib>

#include "rbmpp_kernels.h"
#include "rbmpp_measure.h"
#include "rbmpp_block.h"
#include "rbmpp_skew.h"

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
    // the mirrored entries are written by the (rj, ri) thread.  The same-spin
    // blocks are exact negations of their mirrors (d changes sign exactly);
    // the off-diagonal ones are when cs = 0 (one of cc, ss is zero, so each
    // entry is one rounded product), which the block-form ratios rely on
    (void)bad;
  }
}

void launch_slater(const DevModel& m, const double* f, double* se, int* bad, hipStream_t s) {
  dim3 grid((m.N * m.N + 255) / 256, m.C);
  hipLaunchKernelGGL(k_slater_elements, grid, dim3(256), 0, s, m, f, se, bad);
}

// ------------------------------------------------------------------
// shared-memory layout of the sampler workgroup.  Each wave (one copy)
// keeps a per-row quad {s_j|a_j, t_j|b_j, p_j, q_j} for broadcasts; the quads
// share storage with the LDS recompute matrix, which is only live between
// proposals.  RL >= 0 selects the slim layout of the block form without the
// fused measurement: no measurement buffers, the recompute's stages in the
// union, amplitude tables for C * ar <= 256, C * ac <= 64 and
// ar N + 2 ac N <= 512 (the RBM tables stay in global memory past that), and
// the orbital map and f for Slater_Elm lookups in LDS.  RL > 0 keeps rows
// [RR, RR + RL) of every copy's F^-1 in the union between proposals; the
// recompute's stages overlap them (they are reloaded after every recompute).
template <int NMAX, bool LDSMAT, int RL = -1>
struct SweepSmem {
  static constexpr bool SLIM = RL >= 0;
  static constexpr int NW = RBMPP_CMAX;
  static constexpr int HS = NMAX / 2;  // row stride of the resident F^-1 rows
  static constexpr int VEC = (SLIM && NW * RL * HS > NW * 4 * NMAX) ? NW * RL * HS : NW * 4 * NMAX;
  static constexpr int MRL = SLIM ? 256 : RBMPP_MRMAX;  // real angles held
  static constexpr int MCL = SLIM ? 64 : RBMPP_MCMAX;   // complex angles held
  static constexpr int MAT = LDSMAT ? (NMAX < RBMPP_NLDS ? NMAX * NMAX : RBMPP_NLDS * RBMPP_NLDS) : 0;
  // fused measurement: one copy's X^-1 (at offset 0) + per-wave a, b vectors
  static constexpr int MWAVES = 8;
  // (slim: the block recompute's write-back stages)
  static constexpr int MEAS = LDSMAT ? (SLIM ? NW * 8 * HS : MAT + MWAVES * 2 * NMAX) : 0;
  static constexpr int UNION = VEC > MEAS ? VEC : MEAS;
  static constexpr int MB = (LDSMAT && !SLIM) ? 4 * NMAX : 1;  // bonds / orbital bins of the fused measurement
  // Slater_Elm lookups (slim): orbital index of every site pair, f
  static constexpr int SO = SLIM ? NMAX * NMAX : 1;
  static constexpr int SF = SLIM ? 4 * NMAX : 1;
  double u[UNION];
  double2 psib[MB];
  double opp[2][MB];
  short bi1[MB], bi2[MB];
  int nb;
  double thr[MRL];
  double2 thc[MCL];
  double pf[RBMPP_CMAX], pfn[RBMPP_CMAX];
  double2 psiq[RBMPP_CMAX];
  double gjbuf[4 * NMAX];
  double red[32];
  double2 psi;
  int x[NMAX], eidx[NMAX], ecfg[NMAX], perm[NMAX];
  int sgn[RBMPP_CMAX / 2];
  unsigned sgnbits[RBMPP_CMAX / 2];
  int redi[32];
  int piv, flag, status;
  int guard;  // set by a copy whose updated X^-1 exceeds the conditioning bound
  int gsel[2];     // pivot index / value of the register Gauss-Jordan
  double gselv[2];
  // RBM tables staged for the proposals' amplitudes (Wr, Wc as doubles;
  // W_map; sublattice flags), used when they fit
  static constexpr int TW = LDSMAT ? (SLIM ? 512 : 1536) : 2;
  static constexpr int TM = LDSMAT ? (RBMPP_CMAX / 2) * NMAX : 1;
  alignas(16) double tw[TW];
  double sf[SF];
  unsigned short sorb[SO];
  int tmap[TM];
  unsigned char tsub[LDSMAT ? NMAX : 1];
  unsigned long long ctr;
};


template <int NMAX, bool LDSMAT, int RL>
__device__ __attribute__((noinline)) void wg_recompute(const DevModel& m, SweepSmem<NMAX, LDSMAT, RL>& s, double* Mw,
                             double* gscratch, int check, double* drift_pf, double* drift_inv) {
  const int tid = threadIdx.x, nth = blockDim.x;
  const int N = m.N, C = m.C;
  wg_theta_sign(m, s, check);
  // Pf and X^-1 of each copy: with 16 waves and N <= 112 the register-tile
  // Gauss-Jordan (rbmpp_skew.h), otherwise row-swapping Gauss-Jordan on a
  // matrix in LDS (or global scratch)
  constexpr bool SLIM = SweepSmem<NMAX, LDSMAT, RL>::SLIM;  // launched for block-form models only
  const bool regpath = SLIM || (LDSMAT && nth == 1024 && N <= 112);
  double* A = LDSMAT ? s.u : gscratch;
  const int ld = N;
  double dpf = 0.0, dinv = 0.0;
  if (regpath) {
    for (int k = tid; k < 4 * rbmpp_skew::MPAD; k += nth) s.u[k] = 0.0;  // pivot-row buffers
  }
  const int H = N >> 1;
  if (SLIM || (regpath && m.block)) {
    // block form: every wave inverts its copy's F (rbmpp_block.h) into
    // Mq[d * np + u] = F^-1(d, u), all copies at once
    const int lane = tid & 63, wid = tid >> 6, np = m.npair;
    // pivot rows in gjbuf (4 NMAX doubles = 16 x NMAX/2 ints), the write-back
    // stages in u (16 x 8 x NMAX/2 doubles)
    static_assert(!LDSMAT || (SweepSmem<NMAX, LDSMAT, RL>::UNION >= RBMPP_CMAX * 8 * (NMAX / 2) &&
                              4 * NMAX * 2 >= RBMPP_CMAX * (NMAX / 2)),
                  "stages and pivot rows of the block recompute fit u and gjbuf");
    int* piv = (int*)s.gjbuf + (size_t)wid * (NMAX / 2);
    double* stage = s.u + (size_t)wid * 8 * (NMAX / 2);
    __syncthreads();  // s.u is free
    if (wid < C) {
      const int q = wid;
      double dr[2] = {0.0, 0.0};
      const double det = rbmpp_block::copy_block_inverse(m.se + (size_t)q * 4 * N * N, N, np, s.eidx, piv, stage,
                                                         Mw + (size_t)q * fold_size(N), lane, check ? dr : nullptr);
      if (lane == 0) {
        s.pfn[q] = s.pf[q];  // old value, for the drift measure
        s.pf[q] = (((np * (np - 1) / 2) & 1) ? -1.0 : 1.0) * det;
        s.red[q] = dr[1] > 0.0 ? dr[0] / dr[1] : 0.0;
        if (det == 0.0) s.status = WS_SINGULAR;
      }
    }
    __syncthreads();
    if (s.status) return;
    if (check)
      for (int q = 0; q < C; q++) dinv = fmax(dinv, s.red[q]);
  } else if (NMAX > 128 && m.block) {
    // block form at N > 104: one copy at a time with F in LDS (s.u holds
    // (NMAX / 2)^2 doubles); pivot row and column in gjbuf, then the index
    // arrays
    constexpr int HM = NMAX / 2;
    static_assert(SweepSmem<NMAX, LDSMAT, RL>::UNION >= HM * HM && 4 * NMAX >= 2 * HM + (3 * HM + 1) / 2,
                  "F of the largest copy fits u, its pivot buffers gjbuf");
    int* ib = (int*)(s.gjbuf + 2 * HM);
    const int np = m.npair;
    const double pfsign = ((np * (np - 1) / 2) & 1) ? -1.0 : 1.0;
    for (int q = 0; q < C; q++) {
      double dr[2] = {0.0, 0.0};
      const double det = rbmpp_block::wg_block_inverse_lds(m.se + (size_t)q * 4 * N * N, N, np, s.eidx, s.u, s.gjbuf,
                                              s.gjbuf + HM, ib, ib + HM, ib + 2 * HM, s.red, &s.gsel[0],
                                              &s.gselv[0], &s.gsel[1], Mw + (size_t)q * fold_size(N),
                                              check ? dr : nullptr);
      if (tid == 0) {
        s.pfn[q] = s.pf[q];  // old value, for the drift measure
        s.pf[q] = pfsign * det;
        if (det == 0.0) s.status = WS_SINGULAR;
      }
      if (check) dinv = fmax(dinv, dr[1] > 0.0 ? dr[0] / dr[1] : 0.0);
      __syncthreads();
      if (s.status) return;
    }
  } else if constexpr (!SLIM)
  for (int q = 0; q < C; q++) {
    const double* S = m.se + (size_t)q * 4 * N * N;
    double* Mq = Mw + (size_t)q * fold_size(N);
    double lmax = 0.0, ldiff = 0.0, pfq = 0.0;
    __syncthreads();
    if (regpath) {
      using namespace rbmpp_skew;
      const int tr = tid >> 6, tc = tid & 63;
      tile_t x;
      int rj[RB];
#pragma unroll
      for (int b = 0; b < RB; b++) {
        const int j = tc + 64 * b;
        rj[b] = j < N ? s.eidx[j] + (j >= m.npair ? N : 0) : -1;
      }
#pragma unroll
      for (int aa = 0; aa < RA; aa++) {
        const int i = tr + 16 * aa;
        const int ri = i < N ? s.eidx[i] + (i >= m.npair ? N : 0) : -1;
#pragma unroll
        for (int b = 0; b < RB; b++)
          TILE(x, aa, b) = (ri >= 0 && rj[b] >= 0) ? S[(size_t)ri * 2 * N + rj[b]] : 0.0;
      }
      pfq = reg_skew_pf_inverse<false>(x, N, s.u, s.gsel, s.gselv);
      if (pfq == 0.0) {
        if (tid == 0) s.status = WS_SINGULAR;
        __syncthreads();
        return;
      }
      // M(r,c) = InvM(r,c) = X^-1(c,r) = -X^-1(r,c) (exactly skew); the
      // element of each pair's folded slot is written by its holder
#pragma unroll
      for (int aa = 0; aa < RA; aa++) {
        const int r = tr + 16 * aa;
#pragma unroll
        for (int b = 0; b < RB; b++) {
          const int c = tc + 64 * b;
          if (r < N && c < N) {
            const double v = -TILE(x, aa, b);
            lmax = fmax(lmax, fabs(v));
            int d = c - r;
            if (d < 0) d += N;
            if (d > 0 && (d < H || (d == H && r < H))) {
              const size_t slot = (size_t)(d - 1) * N + r;
              if (check) ldiff = fmax(ldiff, fabs(Mq[slot] - v));
              Mq[slot] = v;
            }
          }
        }
      }
    } else {
      // X(i,j) = -Slater_Elm(r_i, r_j) = Slater_Elm(r_j, r_i) (matrix.f90:72-78)
      for (int e = tid; e < N * N; e += nth) {
        const int i = e / N, j = e - i * N;
        const int ri = s.eidx[i] + (i >= m.npair ? N : 0);
        const int rj = s.eidx[j] + (j >= m.npair ? N : 0);
        A[(size_t)i * ld + j] = S[(size_t)ri * 2 * N + rj];
      }
      __syncthreads();
      pfq = wg_skew_inverse(A, N, ld, s.perm, s.gjbuf, s.red, s.redi, &s.piv);
      if (pfq == 0.0) {
        if (tid == 0) s.status = WS_SINGULAR;
        __syncthreads();
        return;
      }
      // inverse permutation into gjbuf (as ints)
      int* ip = (int*)s.gjbuf;
      for (int r = tid; r < N; r += nth) ip[s.perm[r]] = r;
      __syncthreads();
      // M(r,c) = InvM(r,c) = 0.5 (Xinv(c,r) - Xinv(r,c)) (matrix.f90:97-106),
      // written to the folded slots (or, in block form, InvM(u, np + d) =
      // F^-1(d, u) at d * np + u)
      if (m.block) {
        const int np = m.npair;
        for (int e = tid; e < np * np; e += nth) {
          const int d = e / np, u = e - d * np;
          const double xdu = A[(size_t)ip[np + d] * ld + ip[u]];
          const double xud = A[(size_t)ip[u] * ld + ip[np + d]];
          const double v = 0.5 * (xdu - xud);
          if (check) ldiff = fmax(ldiff, fabs(Mq[e] - v));
          lmax = fmax(lmax, fabs(v));
          Mq[e] = v;
        }
      } else for (int e = tid; e < N * H; e += nth) {
        const int d = e / N + 1, r = e - (d - 1) * N;
        if (d == H && r >= H) continue;
        int c = r + d;
        if (c >= N) c -= N;
        const double xcr = A[(size_t)ip[c] * ld + ip[r]];
        const double xrc = A[(size_t)ip[r] * ld + ip[c]];
        const double v = 0.5 * (xcr - xrc);
        if (check) ldiff = fmax(ldiff, fabs(Mq[e] - v));
        lmax = fmax(lmax, fabs(v));
        Mq[e] = v;
      }
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
template <int NMAX, bool LDSMAT, int RL>
__device__ void wg_x_to_ele(const DevModel& m, SweepSmem<NMAX, LDSMAT, RL>& s) {
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
template <int NMAX, bool LDSMAT, int RL>
__device__ __attribute__((noinline)) int wg_initial_x(const DevModel& m, SweepSmem<NMAX, LDSMAT, RL>& s, Rng& rng, double* Mw,
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
    if (cmod(s.psi) > 1e-300) return WS_OK;
  }
  return WS_INIT;
}

// ------------------------------------------------------------------
// ------------------------------------------------------------------
// Measurement of the stored sample with the walker's own X^-1 (fused
// alternative to calc_phys_quantity_opt, calc_phys_quantity.f90:69-234, which
// recomputes Pf and X^-1 from scratch; here they come from the sampler's
// fast-update chain, refreshed by the full recompute every K accepts).  Per
// copy the N x N X^-1 is staged in LDS; half of the waves evaluate the Pf
// ratio of every antiparallel bond (calculate_new_PfM_two_child), the other
// half scatter the Pfaffian derivative into the orbital bins
// (calculate_derivative_Pfaffian_child).  Writes the reweighted E_loc,
// amplitude, reweight factor and O row(s) of sample `row`.
template <int NMAX, bool LDSMAT, int RL>
__device__ __attribute__((noinline)) void wg_measure_sample(const DevModel& m, SweepSmem<NMAX, LDSMAT, RL>& s,
                                                            const double* Mw, size_t row, const SweepArgs& a) {
  constexpr int KMAX = NMAX / RBMPP_WAVE;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & (RBMPP_WAVE - 1);
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6), nwv = nth >> 6;
  const int N = m.N, C = m.C, np = m.npair;
  const size_t NN = (size_t)N * N;
  const int nbw = nwv >= 2 ? min(nwv / 2, SweepSmem<NMAX, LDSMAT, RL>::MWAVES) : 1;  // bond waves
  if (tid == 0) s.nb = bond_list(m, s);
  for (int k = tid; k < 4 * NMAX; k += nth) {
    s.psib[k] = make_double2(0.0, 0.0);
    s.opp[0][k] = 0.0;
    s.opp[1][k] = 0.0;
  }
  __syncthreads();
  const int nb = s.nb;
  double* Ms = s.u;
  for (int q = 0; q < C; q++) {
    const int g = q >> 1, sidx = q & 1;
    const double* Mq = Mw + (size_t)q * fold_size(N);
    const double* Sq = m.se + (size_t)q * 4 * NN;
    const double pfq = s.pf[q];
    __syncthreads();
    // unfold this copy's X^-1 into a full N x N block
    if (m.block) {
      // InvM(u, np + d) = F^-1(d, u), InvM(np + d, u) = -F^-1(d, u), same-spin blocks 0
      for (int e = tid; e < N * N; e += nth) {
        const int r = e / N, c = e - r * N;
        double v = 0.0;
        if (r < np && c >= np) v = Mq[(size_t)(c - np) * np + r];
        else if (r >= np && c < np) v = -Mq[(size_t)(r - np) * np + c];
        Ms[e] = v;
      }
    } else {
      for (int e = tid; e < N * (N / 2); e += nth) {
        const int d = e / N + 1, r = e - (d - 1) * N;
        if (d == N / 2 && r >= N / 2) continue;
        int c = r + d;
        if (c >= N) c -= N;
        const double v = Mq[e];
        Ms[(size_t)r * N + c] = v;
        Ms[(size_t)c * N + r] = -v;
      }
      for (int r = tid; r < N; r += nth) Ms[(size_t)r * N + r] = 0.0;
    }
    __syncthreads();
    const bool pp_wave = (nwv >= 2) ? (wid >= nbw) : true;
    const bool bond_wave = (nwv >= 2) ? (wid < nbw) : true;
    if (pp_wave) {
      // pair-product O_k of the copy (slater.f90:136-167) from InvM(a, b) =
      // Ms[a * N + b] = X^-1(b, a)
      const double2 wq = psi_wo_pf(m, q, s.thc, s.thr, s.sgn[g]);
      const int t0 = (nwv >= 2) ? (tid - nbw * RBMPP_WAVE) : tid;
      const int tn = (nwv >= 2) ? (nth - nbw * RBMPP_WAVE) : nth;
      wg_pp_oloc(m, s, q, pfq, wq, t0, tn, [&](int u, int v) -> double { return Ms[(size_t)v * N + u]; });
    }
    if (bond_wave && wid < nbw) {
      double* va = s.u + SweepSmem<NMAX, LDSMAT, RL>::MAT + (size_t)wid * 2 * NMAX;
      double* vb = va + NMAX;
      for (int b = wid; b < nb; b += nbw) {
        const int i1 = s.bi1[b], i2 = s.bi2[b];
        const int msa = s.ecfg[i1], msb = s.ecfg[i2];
        const int rsa = i2 + (msa >= np ? N : 0), rsb = i1 + (msb >= np ? N : 0);
        const double* colA = Sq + (size_t)rsa * 2 * N;
        const double* colB = Sq + (size_t)rsb * 2 * N;
        for (int i = lane; i < N; i += RBMPP_WAVE) {
          int t = s.eidx[i];
          if (i == msa) t = i2;
          else if (i == msb) t = i1;
          const int r = t + (i >= np ? N : 0);
          va[i] = colA[r];
          vb[i] = colB[r];
        }
        wave_sync();
        // P(i) = sum_j InvM(j,i) a_j (lane-owned i); Q(msa), Q(msb) by reduction
        double Pk[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; k++) Pk[k] = 0.0;
        for (int j = 0; j < N; j++) {
          const double aj = va[j];
          const double* rowj = Ms + (size_t)j * N;
#pragma unroll
          for (int k = 0; k < KMAX; k++) {
            const int i = lane + k * RBMPP_WAVE;
            if (i < N) Pk[k] = fma(rowj[i], aj, Pk[k]);
          }
        }
        double qa = 0.0, qb = 0.0, bdotp = 0.0;
        for (int j = lane; j < N; j += RBMPP_WAVE) {
          qa = fma(Ms[(size_t)j * N + msa], vb[j], qa);
          qb = fma(Ms[(size_t)j * N + msb], vb[j], qb);
        }
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
          const int i = lane + k * RBMPP_WAVE;
          if (i < N) bdotp = fma(vb[i], Pk[k], bdotp);
        }
        qa = wave_sum(qa);
        qb = wave_sum(qb);
        const double bMa = wave_sum(bdotp);
        double pa = 0.0, pb = 0.0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
          const double ta = __shfl(Pk[k], msa & 63, RBMPP_WAVE);
          const double tb = __shfl(Pk[k], msb & 63, RBMPP_WAVE);
          if ((msa >> 6) == k) pa = ta;
          if ((msb >> 6) == k) pb = tb;
        }
        const double invm_ab = Ms[(size_t)msb * N + msa];
        const double vec_ba = vb[msa];
        const double ratio = invm_ab * vec_ba + invm_ab * bMa + pa * qb - pb * qa;
        int sfac = -1;
        if (m.subB[i1]) sfac = -sfac;
        if (m.subB[i2]) sfac = -sfac;
        const double2 c = psi_wo_pf_exchanged(m, q, s.thc, s.thr, sfac * s.sgn[g], i1, i2, s.x[i1], s.x[i2]);
        if (lane == 0) s.psib[b] = cadd(s.psib[b], cscale(c, ratio * pfq));
        wave_sync();
      }
    }
  }
  __syncthreads();
  for (int q = tid; q < C; q += nth) s.psiq[q] = cscale(psi_wo_pf(m, q, s.thc, s.thr, s.sgn[q >> 1]), s.pf[q]);
  __syncthreads();
  if (tid == 0) {
    double2 p = make_double2(0.0, 0.0);
    for (int q = 0; q < C; q++) p = cadd(p, s.psiq[q]);
    const double rw = 1.0 / (cmod(p) + a.shift);
    a.m_psi[row] = p;
    a.m_rw[row] = rw;
    a.m_eloc[row] = cscale(eloc_from_bonds(m, s, p), rw);
    s.red[0] = rw;
  }
  __syncthreads();
  const bool cplx_rows = !m.real_only;
  double* ore = a.m_orows + (cplx_rows ? 2 * row : row) * (size_t)m.nv;
  wg_orow(m, s, s.red[0], ore, ore + m.nv, cplx_rows);
  __syncthreads();
}

// ------------------------------------------------------------------
// Lane layout of the streaming loops: lane l owns the column pair
// (2l, 2l+1) of every 128-column block, so one 16-byte load per lane and
// row covers a whole 100-wide row of X^-1 in one instruction.  MEAS
// instantiates the measurement fused into the store of each sample (kept
// out of the plain sampler so it does not weigh on its register allocation).
// BLK: the copies are in block form (DevModel::block) and each wave keeps its
// copy's F^-1 instead of the folded X^-1.
// Block form, one wave: the rank-2 update of an accepted exchange applied to
// F^-1 in place, B'[d][u] = B[d][u] + B[d][ua] al_u + x_d be_u with
// x_d = B[d].c (old B).  Lane u holds al_u, be_u, c_u; ua is wave-uniform.
// The sweep applies the same update inside the next ratio pass; this pass
// serves the points where F^-1 must be current without a proposal (before a
// checked recompute, the fused measurement and the end of the launch).
__device__ __attribute__((noinline)) void blk_update_pass(double* Mq, int np, int lane, double cv, int ua,
                                                          double al, double be) {
  constexpr int UN = 8;
  ua = __builtin_amdgcn_readfirstlane(ua);
  const __amdgpu_buffer_rsrc_t rs = block_rsrc(Mq, np * np * 8);
  const int vo = lane < np ? 8 * lane : RBMPP_OOB;
  const int rowb = 8 * np;
  int d = 0;
  for (; d + UN <= np; d += UN) {
    double g[UN];
#pragma unroll
    for (int u = 0; u < UN; u++) g[u] = bload1(rs, vo, (d + u) * rowb);
#pragma unroll
    for (int u = 0; u < UN; u++) {
      const double xd = readlane_d(wave_sum_to63(g[u] * cv), 63);
      const double cd = readlane_d(g[u], ua);
      bstore1(fma(cd, al, fma(xd, be, g[u])), rs, vo, (d + u) * rowb);
    }
  }
  for (; d < np; d++) {
    const double g = bload1(rs, vo, d * rowb);
    const double xd = readlane_d(wave_sum_to63(g * cv), 63);
    const double cd = readlane_d(g, ua);
    bstore1(fma(cd, al, fma(xd, be, g)), rs, vo, d * rowb);
  }
}

// The same for np up to 128 (N > 104): lane l holds columns 2l and 2l + 1
// of every row (one 16-byte access); cv, al, be per column likewise.
__device__ __attribute__((noinline)) void blk_update_pass2(double* Mq, int np, int lane, double2 cv, int ua,
                                                           double2 al, double2 be) {
  ua = __builtin_amdgcn_readfirstlane(ua);
  const __amdgpu_buffer_rsrc_t rs = block_rsrc(Mq, np * np * 8);
  const int vo = 2 * lane < np ? 16 * lane : RBMPP_OOB;
  const int rowb = 8 * np, ul = ua >> 1;
  const bool uy = ua & 1;
  for (int d = 0; d < np; d++) {
    double2 g = bload2(rs, vo, d * rowb);
    const double xd = readlane_d(wave_sum_to63(fma(g.y, cv.y, g.x * cv.x)), 63);
    const double cd = readlane_d(uy ? g.y : g.x, ul);
    g.x = fma(cd, al.x, fma(xd, be.x, g.x));
    g.y = fma(cd, al.y, fma(xd, be.y, g.y));
    bstore2(g, rs, vo, d * rowb);
  }
}

// The pending update of one row g = B[d][lane] (x_d by a wave sum), and of
// four rows with their four wave sums taken together.
__device__ __forceinline__ double blk_upd1(double g, double cv, int ua, double al, double be) {
  const double xd = readlane_d(wave_sum_to63(g * cv), 63);
  return fma(readlane_d(g, ua), al, fma(xd, be, g));
}
__device__ __forceinline__ void blk_upd4(double& g0, double& g1, double& g2, double& g3, double cv, int ua,
                                         double al, double be) {
  const double v = wave_sum4(g0 * cv, g1 * cv, g2 * cv, g3 * cv);
  const double x0 = readlane_d(v, 15), x1 = readlane_d(v, 47), x2 = readlane_d(v, 31), x3 = readlane_d(v, 63);
  g0 = fma(readlane_d(g0, ua), al, fma(x0, be, g0));
  g1 = fma(readlane_d(g1, ua), al, fma(x1, be, g1));
  g2 = fma(readlane_d(g2, ua), al, fma(x2, be, g2));
  g3 = fma(readlane_d(g3, ua), al, fma(x3, be, g3));
}

// Every thread of the sweep draws the same Philox numbers and reads the same
// shared configuration, so these values are equal in all lanes.  Saying so
// keeps the counter, the draws and the branches on them on the scalar unit
// (a branch on a shared-memory load is otherwise taken as divergent, and
// everything after it, the Philox rounds included, moves to vector registers).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  return ((uint64_t)(uint32_t)uni((int)(uint32_t)(v >> 32)) << 32) | (uint32_t)uni((int)(uint32_t)v);
}

// make_sample for walker w of a.ws (one workgroup of one wave per copy).
// Called by k_sweep (one model) and k_sweep_batch (models of several
// lattices in one launch); s is the workgroup's shared memory.
//
// RR, RL > 0 (block form, no fused measurement, RBMPP_CMAX copies): rows
// [0, RR) of each copy's F^-1 stay in the wave's registers (lane u holds
// column u) and rows [RR, RR + RL) in LDS for the whole launch; only the rest
// is streamed from global memory by each ratio pass.  Global memory is made
// current before a recompute with drift check and at the end of the launch,
// and the resident rows are reloaded after every recompute.
template <int NMAX, bool LDSMAT, bool MEAS, bool BLK, int RR = 0, int RL = -1>
__device__ __forceinline__ void sweep_walker(const SweepArgs& a, const int w, SweepSmem<NMAX, LDSMAT, RL>& s) {
  constexpr bool RES = RR + RL > 0;
  static_assert(!RES || (BLK && !MEAS && LDSMAT), "resident F^-1 rows: block form without the fused measurement");
  static_assert(RR % 4 == 0, "register rows are updated four at a time");
  constexpr int HS = SweepSmem<NMAX, LDSMAT, RL>::HS;
  constexpr int KD = NMAX >= 128 ? NMAX / 128 : 1;
  constexpr int UN = 8;  // rows in flight per wave
  const DevModel& m = *a.m;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & (RBMPP_WAVE - 1), wid = tid >> 6;
  const int N = m.N, C = m.C, P = m.P;
  const size_t NN = (size_t)N * N;
  double* Mw = a.ws.minv + (size_t)w * C * fold_size(N);
  double* gscr = a.ws.scratch ? a.ws.scratch + (size_t)w * NN : nullptr;

  Rng rng;
  rng.k0 = a.key0; rng.k1 = a.key1; rng.stream = (uint32_t)w;
  rng.ctr = a.ws.ctr[w];
  if (tid == 0) {
    s.status = WS_OK;
    s.guard = 0;
  }
  for (int i = tid; i < C * m.ar; i += nth) s.thr[i] = a.ws.thr[(size_t)w * m.mr + i];
  for (int i = tid; i < C * m.ac; i += nth) s.thc[i] = a.ws.thc[(size_t)w * m.mc + i];
  for (int q = tid; q < C; q += nth) s.pf[q] = a.ws.pf[(size_t)w * C + q];
  for (int g = tid; g < P; g += nth) s.sgn[g] = a.ws.sgn[(size_t)w * P + g];
  RbmTables tb = model_tables(m);
  // slim layout: Slater_Elm(a, b + N) of copy q from f and the orbital map in
  // LDS instead of the global table (block form: one of cc, ss is zero)
  constexpr bool SLIM = SweepSmem<NMAX, LDSMAT, RL>::SLIM;
  const bool slt = SLIM && m.nsl <= SweepSmem<NMAX, LDSMAT, RL>::SF && N <= NMAX;
  if constexpr (SLIM) {
    if (slt) {
      for (int i = tid; i < N * N; i += nth) s.sorb[i] = (unsigned short)m.orb[i];
      for (int k = tid; k < m.nsl; k += nth) s.sf[k] = m.f[k];
    }
  }
  if constexpr (LDSMAT) {
    using SS = SweepSmem<NMAX, LDSMAT, RL>;
    const int nwr = m.ar * N, nwc = 2 * m.ac * N;
    if (nwr + nwc <= SS::TW && P * N <= SS::TM) {
      for (int i = tid; i < nwr; i += nth) s.tw[i] = m.Wr[i];
      for (int i = tid; i < nwc; i += nth) s.tw[nwr + i] = ((const double*)m.Wc)[i];
      for (int i = tid; i < P * N; i += nth) s.tmap[i] = m.wmap[i];
      for (int i = tid; i < N; i += nth) s.tsub[i] = m.subB[i];
      tb = {s.tw, (const double2*)(s.tw + nwr), s.tmap, s.tsub};
    }
  }
  __syncthreads();

  int status = WS_OK;
  // ---- set-up (make_sample.f90:50-62)
  if (a.first) {
    // the out-of-line set-up draws from a copy, so that the sweep's own
    // generator never has its address taken and stays in scalar registers
    Rng r0 = rng;
    status = uni(wg_initial_x(m, s, r0, Mw, gscr));
    rng.ctr = uni64(r0.ctr);
  } else {
    for (int i = tid; i < N; i += nth) s.x[i] = a.ws.x[(size_t)w * N + i];
    __syncthreads();
    wg_x_to_ele(m, s);
    wg_recompute(m, s, Mw, gscr, 0, nullptr, nullptr);
    status = uni(s.status);
    if (status == WS_OK && uni(rng.u01() < a.p_restart)) {
      Rng r0 = rng;
      status = uni(wg_initial_x(m, s, r0, Mw, gscr));
      rng.ctr = uni64(r0.ctr);
    }
  }
  rng.ctr = uni64(rng.ctr);
  // The sweep's draws come from a per-wave table of 64 consecutive counters
  // of the walker's stream, one per lane, formed on the vector unit when the
  // counter leaves it and read by readlane: the same numbers in the same
  // order as Rng::u01, without every wave of the workgroup running the same
  // Philox rounds on the scalar unit it shares with the others.
  double rtab = 0.0;
  uint64_t rbase = rng.ctr + RBMPP_WAVE;  // empty
  auto draw = [&]() -> double {
    if (rng.ctr - rbase >= (uint64_t)RBMPP_WAVE) {
      rbase = rng.ctr;
      rtab = philox_u01(rng.k0, rng.k1, rng.stream, rbase + (uint64_t)lane);
    }
    const double u = readlane_d(rtab, (int)(rng.ctr - rbase));
    rng.ctr++;
    return u;
  };
  // Rng::nran0: (int)(n u) lies in [0, n], so the modulo is one subtraction
  auto nran = [&](int n) -> int {
    const int r = (int)((double)n * draw());
    return r >= n ? r - n : r;
  };
  // c / L for the candidate's site arithmetic: the high word of c * ceil(2^32 / L)
  // is exact for L <= 256 and c < L^2
  const uint32_t mL = (uint32_t)(((1ull << 32) + (uint64_t)m.L - 1) / (uint64_t)m.L);
  int nacc_tmp = 0;
  int nfast = a.ws.nfastup[w];
  long long naccept = 0;
  int nrec = 0;
  // this wave's projection copy and its broadcast quads
  const int q = __builtin_amdgcn_readfirstlane(wid);  // wave-uniform
  double* wv = s.u + (size_t)wid * 4 * NMAX;  // (unused by the resident form)
  double* Mq = Mw + (size_t)q * fold_size(N);
  const double* Sq = m.se + (size_t)q * 4 * NN;
  const int H = N >> 1;
  // block form: an accepted exchange leaves its update of F^-1 pending
  // (coefficients of lane u: pal, pbe, pcv; wave-uniform pua) and the next
  // ratio pass applies it to each row as it reads it, so that an accept
  // costs one store pass instead of a read-modify-write pass more.  A
  // recompute without drift check drops it (it rebuilds F^-1 from x).
  bool pend = false;
  int pua = 0;
  double pal = 0.0, pbe = 0.0, pcv = 0.0;
  // N > 104 (NMAX = 256): two columns per lane
  double2 pal2 = make_double2(0.0, 0.0), pbe2 = make_double2(0.0, 0.0), pcv2 = make_double2(0.0, 0.0);
  // the pending update applied by a pass of its own
  auto flush_pending = [&]() {
    if constexpr (NMAX > 128) blk_update_pass2(Mq, m.npair, lane, pcv2, pua, pal2, pbe2);
    else blk_update_pass(Mq, m.npair, lane, pcv, pua, pal, pbe);
  };
  // resident rows: breg (registers), lres (LDS, row d - RR at (d - RR) * HS);
  // the host launches this form only with npair >= RR + RL
  double breg[RR > 0 ? RR : 1];
  double* lres = s.u + (size_t)q * (RL > 0 ? RL : 0) * HS;
  const int bnp = m.npair;
  const bool blive = lane < bnp;
  const int blo = blive ? lane : 0;  // an in-range column for the lanes past npair
  auto res_load = [&]() {
    if constexpr (RES) {
      const __amdgpu_buffer_rsrc_t rs = block_rsrc(Mq, bnp * bnp * 8);
      const int vo = blive ? 8 * lane : RBMPP_OOB;
#pragma unroll
      for (int d = 0; d < RR; d++) breg[d] = bload1(rs, vo, d * 8 * bnp);
#pragma unroll
      for (int d = 0; d < RL; d++) {
        const double v = bload1(rs, vo, (RR + d) * 8 * bnp);
        if (lane < HS) lres[d * HS + lane] = v;
      }
    }
  };
  // global memory made current: the pending update applied everywhere and
  // the resident rows written back
  auto res_flush = [&]() {
    if constexpr (RES) {
      const __amdgpu_buffer_rsrc_t rs = block_rsrc(Mq, bnp * bnp * 8);
      const int vo = blive ? 8 * lane : RBMPP_OOB;
      const int rowb = 8 * bnp;
      if (pend) {
#pragma unroll
        for (int d = 0; d < RR; d += 4) blk_upd4(breg[d], breg[d + 1], breg[d + 2], breg[d + 3], pcv, pua, pal, pbe);
      }
#pragma unroll
      for (int d = 0; d < RR; d++) bstore1(breg[d], rs, vo, d * rowb);
      for (int d = 0; d < RL; d++) {
        const double v = lres[d * HS + blo];
        double g = blive ? v : 0.0;
        if (pend) g = blk_upd1(g, pcv, pua, pal, pbe);
        bstore1(g, rs, vo, (RR + d) * rowb);
      }
      if (pend)
        for (int d = RR + RL; d < bnp; d++)
          bstore1(blk_upd1(bload1(rs, vo, d * rowb), pcv, pua, pal, pbe), rs, vo, d * rowb);
      pend = false;
    }
  };
  res_load();

  for (int i0 = -a.nwarm; i0 <= a.nsamp && status == WS_OK; i0++) {
    // ---- periodic recompute (make_sample.f90:74-128)
    const int limit = a.fixed_k > 0 ? a.fixed_k : nfast;
    if (nacc_tmp > limit) {
      if constexpr (BLK) {
        // the drift check compares with the fast-updated F^-1
        if constexpr (RES) res_flush();
        else if (pend && q < C) flush_pending();
        pend = false;
      }
      double dpf = 0.0, dinv = 0.0;
      wg_recompute(m, s, Mw, gscr, 1, &dpf, &dinv);
      status = uni(s.status);
      nrec++;
      nacc_tmp = 0;
      if (a.fixed_k <= 0) {
        if (uni(dinv > 1.0e-5 || dpf > 1.0e-5)) nfast = max(1, (int)llround((double)nfast / 1.2));
        else nfast = min(100, (int)llround((double)nfast * 1.1));
      }
      if (status) break;
      res_load();
    }
    for (int iu = 0; iu < a.nupdate; iu++) {
      rng.ctr = uni64(rng.ctr);
      // ---- global spin flip (make_sample.f90:135-158)
      if (uni(draw() < a.p_global)) {
        const double2 psi_old = s.psi;
        for (int i = tid; i < N; i += nth) s.x[i] = -s.x[i];
        __syncthreads();
        wg_x_to_ele(m, s);
        wg_recompute(m, s, Mw, gscr, 0, nullptr, nullptr);
        pend = false;
        if (uni(s.status)) { status = uni(s.status); break; }
        const double r1 = (cmod(s.psi) + a.shift) / (cmod(psi_old) + a.shift);
        if (!uni(draw() < r1 * r1)) {
          for (int i = tid; i < N; i += nth) s.x[i] = -s.x[i];
          __syncthreads();
          wg_x_to_ele(m, s);
          wg_recompute(m, s, Mw, gscr, 0, nullptr, nullptr);
          if (uni(s.status)) { status = uni(s.status); break; }
        }
        res_load();
      }
      // ---- exchange candidate (make_sample.f90:244-294)
      int i1 = -1, i2 = -1;
      {
        const int dist = nran(m.L - 1) + 2;
        for (int it = 0; it < 1000000; it++) {
          const int c1 = nran(N);
          int ix = (int)__umulhi((uint32_t)c1, mL), iy = c1 - ix * m.L;
          ix += nran(dist);
          iy += nran(dist);
          if (ix >= m.L) ix -= m.L;
          if (iy >= m.L) iy -= m.L;
          const int c2 = ix * m.L + iy;
          if (uni(s.x[c1] != s.x[c2])) { i1 = c1; i2 = c2; break; }
        }
      }
      if (i1 < 0) { status = WS_CANDIDATE; break; }
      const int msa = s.ecfg[i1], msb = s.ecfg[i2];
      const int x1 = s.x[i1], x2 = s.x[i2];
      int sfac = -1;
      if (tb.subB[i1]) sfac = -sfac;
      if (tb.subB[i2]) sfac = -sfac;
      const int rsa = i2 + (msa >= m.npair ? N : 0);  // new row of electron msa
      const int rsb = i1 + (msb >= m.npair ? N : 0);

      // ---- Pfaffian ratio of copy q (calculate_new_PfM_two_child)
      double2 Pk[KD], Qk[KD], bk[KD];
      double bMa = 0.0, vec_ba = 0.0, invm_ab = 0.0, pa = 0.0, pb = 0.0, qa = 0.0, qb = 0.0;
      double rq = 1.0;  // this copy's Pf ratio
      // block form: y = r^T F^-1 (lane u), row db of F^-1, c(u), and the scalars
      double by = 0.0, brow = 0.0, bcv = 0.0, bgam = 0.0, bdel = 0.0, beps = 0.0, bbet = 0.0;
      double2 by2 = make_double2(0.0, 0.0), brow2 = make_double2(0.0, 0.0), bcv2 = make_double2(0.0, 0.0);
      int bua = 0;
      if constexpr (BLK) {
        if (q < C) {
          // X = [[0, F], [-F^T, 0]] with F(u, d) = Slater[site_u][site_(np+d) + N],
          // Pf X = (-1)^(np(np-1)/2) det F.  The exchange moves the up electron
          // ua from site s1 to s2 and the down electron db from s2 to s1:
          // F' = F + e_ua r^T + c e_db^T, r(d) = F'(ua, d) - F(ua, d),
          // c(u) = F'(u, db) - F(u, db) (c(ua) = 0), and with B = F^-1 and
          // y = r^T B the ratio of pfupdate_two.f90:27-88 becomes
          // det F'/det F = (1 + y_ua)(1 + B[db].c) - (y.c) B[db][ua].
          if constexpr (NMAX <= 128) {
          const int np = m.npair;
          const bool aup = msa < np;
          bua = __builtin_amdgcn_readfirstlane(aup ? msa : msb);
          const int db = __builtin_amdgcn_readfirstlane((aup ? msb : msa) - np);
          const int s1 = aup ? i1 : i2, s2 = aup ? i2 : i1;
          const double* S1 = Sq + (size_t)s1 * 2 * N;
          const double* S2 = Sq + (size_t)s2 * 2 * N;
          // c(u) = Slater[s_u][s1+N] - Slater[s_u][s2+N] read from rows s1+N
          // and s2+N: with cos sin = 0 one of the two weights of every
          // off-diagonal entry is zero, so Slater[r+N][c] is the exact
          // negation of Slater[c][r+N] and the difference is bit-identical,
          // from two rows instead of one row per lane
          const double* T1 = Sq + (size_t)(s1 + N) * 2 * N;
          const double* T2 = Sq + (size_t)(s2 + N) * 2 * N;
          double rv = 0.0;
          if (slt) {
            // Sl(a, b) = Slater_Elm(a, b + N) = cc f[orb(w a, w b)] + ss f[orb(w b, w a)]
            // (slater.f90:7-64) with the zero weight's term left out (it
            // adds an exact zero); Slater_Elm(a + N, b) = -Sl(b, a)
            const int g = q >> 1;
            const double wc = m.sp_cc[q & 1], ws = m.sp_ss[q & 1];
            const int* wm = tb.wmap + g * N;
            auto sl = [&](int wa, int wb) -> double {
              return wc != 0.0 ? s.sf[s.sorb[wa * N + wb]] * wc : s.sf[s.sorb[wb * N + wa]] * ws;
            };
            const int w1 = wm[s1], w2 = wm[s2];
            if (lane < np) {
              const int wd = wm[s.eidx[np + lane]], wu = wm[s.eidx[lane]];
              rv = lane == db ? sl(w2, w1) - sl(w1, w2) : sl(w2, wd) - sl(w1, wd);
              bcv = lane == bua ? 0.0 : sl(wu, w1) - sl(wu, w2);
            }
          } else if (lane < np) {
            const int sd = s.eidx[np + lane], su = s.eidx[lane];
            rv = lane == db ? S2[s1 + N] - S1[s2 + N] : S2[sd + N] - S1[sd + N];
            bcv = lane == bua ? 0.0 : T2[su] - T1[su];
          }
          double* rt = wv;  // r(d) for every d, zero past np (the resident form reads lane d)
          if constexpr (!RES) {
            rt[lane] = rv;
            wave_sync();
          }
          const __amdgpu_buffer_rsrc_t rs = block_rsrc(Mq, np * np * 8);
          const int vo = lane < np ? 8 * lane : RBMPP_OOB;
          const int rowb = 8 * np;
          double y = 0.0;
          if constexpr (RES) {
            // the same pass over the three tiers: registers, LDS, global
            if (pend) {
#pragma unroll
              for (int d = 0; d < RR; d += 4)
                blk_upd4(breg[d], breg[d + 1], breg[d + 2], breg[d + 3], pcv, pua, pal, pbe);
            }
            double bsel = 0.0;
#pragma unroll
            for (int d = 0; d < RR; d++) {
              y = fma(readlane_d(rv, d), breg[d], y);
              bsel = d == db ? breg[d] : bsel;
            }
#pragma unroll
            for (int d0 = 0; d0 < RL; d0 += 4) {
              double g[4];
#pragma unroll
              for (int j = 0; j < 4; j++) {
                if (d0 + j < RL) {
                  const double v = lres[(d0 + j) * HS + blo];
                  g[j] = blive ? v : 0.0;
                }
              }
              if (pend) {
                if (d0 + 4 <= RL) {
                  blk_upd4(g[0], g[1], g[2], g[3], pcv, pua, pal, pbe);
                } else {
#pragma unroll
                  for (int j = 0; j < 4; j++)
                    if (d0 + j < RL) g[j] = blk_upd1(g[j], pcv, pua, pal, pbe);
                }
#pragma unroll
                for (int j = 0; j < 4; j++)
                  if (d0 + j < RL && blive) lres[(d0 + j) * HS + lane] = g[j];
              }
#pragma unroll
              for (int j = 0; j < 4; j++)
                if (d0 + j < RL) y = fma(readlane_d(rv, RR + d0 + j), g[j], y);
            }
            constexpr int UG = 16;  // streamed rows in flight
            int d = RR + RL;
            for (; d + UG <= np; d += UG) {
              double g[UG];
#pragma unroll
              for (int u = 0; u < UG; u++) g[u] = bload1(rs, vo, (d + u) * rowb);
              if (pend) {
#pragma unroll
                for (int u = 0; u < UG; u += 4) blk_upd4(g[u], g[u + 1], g[u + 2], g[u + 3], pcv, pua, pal, pbe);
#pragma unroll
                for (int u = 0; u < UG; u++) bstore1(g[u], rs, vo, (d + u) * rowb);
              }
#pragma unroll
              for (int u = 0; u < UG; u++) y = fma(readlane_d(rv, d + u), g[u], y);
            }
            for (; d < np; d++) {
              double g = bload1(rs, vo, d * rowb);
              if (pend) {
                g = blk_upd1(g, pcv, pua, pal, pbe);
                bstore1(g, rs, vo, d * rowb);
              }
              y = fma(readlane_d(rv, d), g, y);
            }
            // row db as updated above (LDS and global rows: each lane reads
            // back its own store)
            if (db < RR) {
              brow = bsel;
            } else if (db < RR + RL) {
              const double v = lres[(db - RR) * HS + blo];
              brow = blive ? v : 0.0;
            } else {
              brow = bload1(rs, vo, db * rowb);
            }
          } else {
          // one pass over the rows of F^-1: the pending update of the last
          // accept first (stored back), then y += r(d) B[d]; row db is read
          // again after the pass (a reload is cheaper than selecting it from
          // every row)
          constexpr int UB = 25;  // rows in flight (N = 100: two groups)
          int d = 0;
          for (; d + UB <= np; d += UB) {
            double g[UB];
#pragma unroll
            for (int u = 0; u < UB; u++) g[u] = bload1(rs, vo, (d + u) * rowb);
            if (pend) {
              // x_d = B[d].c four rows at a time (wave_sum4), the rest singly
              constexpr int U4 = UB / 4 * 4;
#pragma unroll
              for (int u = 0; u < U4; u += 4) {
                const double v = wave_sum4(g[u] * pcv, g[u + 1] * pcv, g[u + 2] * pcv, g[u + 3] * pcv);
                const double xd[4] = {readlane_d(v, 15), readlane_d(v, 47), readlane_d(v, 31), readlane_d(v, 63)};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                  const double cd = readlane_d(g[u + j], pua);
                  g[u + j] = fma(cd, pal, fma(xd[j], pbe, g[u + j]));
                  bstore1(g[u + j], rs, vo, (d + u + j) * rowb);
                }
              }
#pragma unroll
              for (int u = U4; u < UB; u++) {
                const double xd = readlane_d(wave_sum_to63(g[u] * pcv), 63);
                const double cd = readlane_d(g[u], pua);
                g[u] = fma(cd, pal, fma(xd, pbe, g[u]));
                bstore1(g[u], rs, vo, (d + u) * rowb);
              }
            }
#pragma unroll
            for (int u = 0; u < UB; u++) y = fma(rt[d + u], g[u], y);
          }
          for (; d < np; d++) {
            double g = bload1(rs, vo, d * rowb);
            if (pend) {
              const double xd = readlane_d(wave_sum_to63(g * pcv), 63);
              const double cd = readlane_d(g, pua);
              g = fma(cd, pal, fma(xd, pbe, g));
              bstore1(g, rs, vo, d * rowb);
            }
            y = fma(rt[d], g, y);
          }
          // row db as updated above (each lane reads back its own store)
          brow = bload1(rs, vo, db * rowb);
          }
          pend = false;
          by = y;
          bgam = readlane_d(y, bua);
          bbet = readlane_d(brow, bua);
          {
            const double v = wave_sum2(brow * bcv, y * bcv);
            bdel = readlane_d(v, 31);
            beps = readlane_d(v, 63);
          }
          } else {
          // N > 104 (np up to 128): lane l holds columns 2l and 2l + 1 of F^-1
          // (one 16-byte load per row), r(d) and c(u) likewise; the same
          // algebra, with rows read from the global Slater table
          const int np = m.npair;
          const bool aup = msa < np;
          bua = __builtin_amdgcn_readfirstlane(aup ? msa : msb);
          const int db = __builtin_amdgcn_readfirstlane((aup ? msb : msa) - np);
          const int s1 = aup ? i1 : i2, s2 = aup ? i2 : i1;
          const double* S1 = Sq + (size_t)s1 * 2 * N;
          const double* S2 = Sq + (size_t)s2 * 2 * N;
          const double* T1 = Sq + (size_t)(s1 + N) * 2 * N;
          const double* T2 = Sq + (size_t)(s2 + N) * 2 * N;
          const bool live = 2 * lane < np;  // np = N / 2 is even
          double2 rv2 = make_double2(0.0, 0.0);
          if (live) {
            const int e0 = 2 * lane, e1 = e0 + 1;
            const int sd0 = s.eidx[np + e0], sd1 = s.eidx[np + e1];
            const int su0 = s.eidx[e0], su1 = s.eidx[e1];
            const double rdb = S2[s1 + N] - S1[s2 + N];
            rv2.x = e0 == db ? rdb : S2[sd0 + N] - S1[sd0 + N];
            rv2.y = e1 == db ? rdb : S2[sd1 + N] - S1[sd1 + N];
            bcv2.x = e0 == bua ? 0.0 : T2[su0] - T1[su0];
            bcv2.y = e1 == bua ? 0.0 : T2[su1] - T1[su1];
          }
          double* rt = wv;  // r(d) for every d (zero past np)
          *(double2*)&rt[2 * lane] = rv2;
          wave_sync();
          const __amdgpu_buffer_rsrc_t rs = block_rsrc(Mq, np * np * 8);
          const int vo = live ? 16 * lane : RBMPP_OOB;
          const int rowb = 8 * np, pl = pua >> 1;
          const bool py = pua & 1;
          double2 y = make_double2(0.0, 0.0);
          constexpr int UB = 8;  // rows in flight
          int d = 0;
          for (; d + UB <= np; d += UB) {
            double2 g[UB];
#pragma unroll
            for (int u = 0; u < UB; u++) g[u] = bload2(rs, vo, (d + u) * rowb);
            if (pend) {
#pragma unroll
              for (int u = 0; u < UB; u += 4) {
                const double v = wave_sum4(fma(g[u].y, pcv2.y, g[u].x * pcv2.x),
                                           fma(g[u + 1].y, pcv2.y, g[u + 1].x * pcv2.x),
                                           fma(g[u + 2].y, pcv2.y, g[u + 2].x * pcv2.x),
                                           fma(g[u + 3].y, pcv2.y, g[u + 3].x * pcv2.x));
                const double xd[4] = {readlane_d(v, 15), readlane_d(v, 47), readlane_d(v, 31), readlane_d(v, 63)};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                  const double cd = readlane_d(py ? g[u + j].y : g[u + j].x, pl);
                  g[u + j].x = fma(cd, pal2.x, fma(xd[j], pbe2.x, g[u + j].x));
                  g[u + j].y = fma(cd, pal2.y, fma(xd[j], pbe2.y, g[u + j].y));
                  bstore2(g[u + j], rs, vo, (d + u + j) * rowb);
                }
              }
            }
#pragma unroll
            for (int u = 0; u < UB; u++) {
              const double rd = rt[d + u];
              y.x = fma(rd, g[u].x, y.x);
              y.y = fma(rd, g[u].y, y.y);
            }
          }
          for (; d < np; d++) {
            double2 g = bload2(rs, vo, d * rowb);
            if (pend) {
              const double xd = readlane_d(wave_sum_to63(fma(g.y, pcv2.y, g.x * pcv2.x)), 63);
              const double cd = readlane_d(py ? g.y : g.x, pl);
              g.x = fma(cd, pal2.x, fma(xd, pbe2.x, g.x));
              g.y = fma(cd, pal2.y, fma(xd, pbe2.y, g.y));
              bstore2(g, rs, vo, d * rowb);
            }
            const double rd = rt[d];
            y.x = fma(rd, g.x, y.x);
            y.y = fma(rd, g.y, y.y);
          }
          // row db as updated above (each lane reads back its own store)
          brow2 = bload2(rs, vo, db * rowb);
          pend = false;
          by2 = y;
          const int bl = bua >> 1;
          bgam = readlane_d((bua & 1) ? y.y : y.x, bl);
          bbet = readlane_d((bua & 1) ? brow2.y : brow2.x, bl);
          {
            const double v = wave_sum2(fma(brow2.y, bcv2.y, brow2.x * bcv2.x), fma(y.y, bcv2.y, y.x * bcv2.x));
            bdel = readlane_d(v, 31);
            beps = readlane_d(v, 63);
          }
          }
          const double ratio = (1.0 + bgam) * (1.0 + bdel) - beps * bbet;
          rq = ratio;
          const double pfn = ratio * s.pf[q];
          const double2 c = psi_wo_pf_exchanged_wave(m, tb, q, s.thc, s.thr, sfac * s.sgn[q >> 1], i1, i2, x1, x2, lane);
          if (lane == 0) {
            s.pfn[q] = pfn;
            s.psiq[q] = cscale(c, pfn);
          }
        }
      } else if (q < C) {
        const double* colA = Sq + (size_t)rsa * 2 * N;
        const double* colB = Sq + (size_t)rsb * 2 * N;
        // (a_j, b_j) for j in [-N/2, 3N/2) at ab[j + N/2] (periodic in N), so
        // that the shifted reads of the folded rows below need no wrap
        double2* ab = (double2*)wv;
#pragma unroll
        for (int k = 0; k < KD; k++) {
          const int c = 2 * lane + 128 * k;
          double2 av = make_double2(0.0, 0.0), bv = make_double2(0.0, 0.0);
          if (c < N) {
#pragma unroll
            for (int e = 0; e < 2; e++) {
              const int i = c + e;
              int t = s.eidx[i];
              if (i == msa) t = i2;
              else if (i == msb) t = i1;
              const int r = t + (i >= m.npair ? N : 0);
              const double aa = colA[r], bb = colB[r];
              if (e == 0) { av.x = aa; bv.x = bb; } else { av.y = aa; bv.y = bb; }
            }
            const double2 e0 = make_double2(av.x, bv.x), e1 = make_double2(av.y, bv.y);
            ab[c + H] = e0;
            ab[c + 1 + H] = e1;
            const int cw = c >= H ? c - H : c + N + H;  // the periodic image (N/2 is even)
            ab[cw] = e0;
            ab[cw + 1] = e1;
          }
          bk[k] = bv;
          Pk[k] = make_double2(0.0, 0.0);
          Qk[k] = make_double2(0.0, 0.0);
        }
        invm_ab = fold_get(Mq, N, msb, msa);
        wave_sync();
        // P_c = sum_j M(j,c) a_j and Q_c (with b) over the folded rows: row d
        // holds G(d,c) = M(c,c+d), which contributes -G(d,c) a_{c+d}, and at
        // the rotated column c-d the element M(c-d,c), which contributes
        // +G(d,c-d) a_{c-d} (the same row again: an L1 hit).  The last row
        // holds each pair (c, c+N/2) once, at c < N/2.
        {
          const __amdgpu_buffer_rsrc_t rs = block_rsrc(Mq, (int)(fold_size(N) * 8));
          int vo[KD], cb[KD];  // the lane's column pair (dead lanes: a valid one, results unused)
#pragma unroll
          for (int k = 0; k < KD; k++) {
            const int c = 2 * lane + 128 * k;
            vo[k] = c < N ? 8 * c : RBMPP_OOB;
            cb[k] = c < N ? c : N - 2;
          }
          const int rowb = 8 * N;
          // byte offset of column (t mod N) for t in (-N, N): the unsigned
          // minimum of t and t + N picks the non-negative one
          auto wrap8 = [N](int t8) -> int { return (int)min((unsigned)t8, (unsigned)(t8 + 8 * N)); };
          static_assert(UN % 2 == 0, "batches start at odd rows");
          int d = 1;
          for (; d + UN <= H; d += UN) {
            // rows d..d+UN-1, d odd: row d+u is even for odd u, and then the
            // rotated pair (c-d-u, c+1-d-u) is an aligned pair of one row
            double2 gl[UN][KD], gr[UN][KD];
#pragma unroll
            for (int u = 0; u < UN; u++)
#pragma unroll
              for (int k = 0; k < KD; k++) {
                const int soff = (d + u - 1) * rowb;
                const int t0 = 8 * (cb[k] - (d + u));
                gl[u][k] = bload2(rs, vo[k], soff);
                if (u & 1) {
                  gr[u][k] = bload2(rs, wrap8(t0), soff);
                } else {
                  gr[u][k] = make_double2(bload1(rs, wrap8(t0), soff), bload1(rs, wrap8(t0 + 8), soff));
                }
              }
#pragma unroll
            for (int k = 0; k < KD; k++) {
              const double2* ap = ab + (cb[k] + H + d);             // a_{c+d+u} at ap[u]
              const double2* am = ab + (cb[k] + H - d - (UN - 1));  // a_{c-d-u} at am[UN-1-u]
#pragma unroll
              for (int u = 0; u < UN; u++) {
                const double2 ap0 = ap[u], ap1 = ap[u + 1], am0 = am[UN - 1 - u], am1 = am[UN - u];
                Pk[k].x = fma(-gl[u][k].x, ap0.x, Pk[k].x); Pk[k].x = fma(gr[u][k].x, am0.x, Pk[k].x);
                Qk[k].x = fma(-gl[u][k].x, ap0.y, Qk[k].x); Qk[k].x = fma(gr[u][k].x, am0.y, Qk[k].x);
                Pk[k].y = fma(-gl[u][k].y, ap1.x, Pk[k].y); Pk[k].y = fma(gr[u][k].y, am1.x, Pk[k].y);
                Qk[k].y = fma(-gl[u][k].y, ap1.y, Qk[k].y); Qk[k].y = fma(gr[u][k].y, am1.y, Qk[k].y);
              }
            }
          }
          for (; d < H; d++) {
#pragma unroll
            for (int k = 0; k < KD; k++) {
              const int soff = (d - 1) * rowb;
              const int t0 = 8 * (cb[k] - d);
              const double2 gl = bload2(rs, vo[k], soff);
              const double g0 = bload1(rs, wrap8(t0), soff), g1 = bload1(rs, wrap8(t0 + 8), soff);
              const double2 ap0 = ab[cb[k] + H + d], ap1 = ab[cb[k] + H + d + 1];
              const double2 am0 = ab[cb[k] + H - d], am1 = ab[cb[k] + H - d + 1];
              Pk[k].x = fma(-gl.x, ap0.x, Pk[k].x); Pk[k].x = fma(g0, am0.x, Pk[k].x);
              Qk[k].x = fma(-gl.x, ap0.y, Qk[k].x); Qk[k].x = fma(g0, am0.y, Qk[k].x);
              Pk[k].y = fma(-gl.y, ap1.x, Pk[k].y); Pk[k].y = fma(g1, am1.x, Pk[k].y);
              Qk[k].y = fma(-gl.y, ap1.y, Qk[k].y); Qk[k].y = fma(g1, am1.y, Qk[k].y);
            }
          }
          // last row: columns c < N/2 read their own slot (pair (c, c+N/2)),
          // columns c >= N/2 the slot c-N/2 (pair (c-N/2, c)); N/2 is even, so
          // a lane's column pair falls on one side
#pragma unroll
          for (int k = 0; k < KD; k++) {
            const bool live = vo[k] != RBMPP_OOB, lo = cb[k] < H;
            const int soff = (H - 1) * rowb;
            const double2 gl = bload2(rs, live && lo ? vo[k] : RBMPP_OOB, soff);
            const double2 gr = bload2(rs, live && !lo ? 8 * (cb[k] - H) : RBMPP_OOB, soff);
            const int o = (lo ? cb[k] + H : cb[k] - H) + H;
            const double2 a0 = ab[o], a1 = ab[o + 1];
            const double vx = gr.x - gl.x, vy = gr.y - gl.y;  // one of the two is zero
            Pk[k].x = fma(vx, a0.x, Pk[k].x); Qk[k].x = fma(vx, a0.y, Qk[k].x);
            Pk[k].y = fma(vy, a1.x, Pk[k].y); Qk[k].y = fma(vy, a1.y, Qk[k].y);
          }
        }
        vec_ba = ab[msa + H].y;
        wave_sync();  // the (a, b) table is read before (P, Q) replaces it
        double2* pq = (double2*)wv;  // (P_c, Q_c)
        double part = 0.0;
#pragma unroll
        for (int k = 0; k < KD; k++) {
          const int c = 2 * lane + 128 * k;
          if (c < N) {
            pq[c] = make_double2(Pk[k].x, Qk[k].x);
            pq[c + 1] = make_double2(Pk[k].y, Qk[k].y);
            part = fma(bk[k].x, Pk[k].x, part);
            part = fma(bk[k].y, Pk[k].y, part);
          }
        }
        bMa = wave_sum(part);
        wave_sync();
        pa = pq[msa].x; qa = pq[msa].y;
        pb = pq[msb].x; qb = pq[msb].y;
        const double ratio = invm_ab * vec_ba + invm_ab * bMa + pa * qb - pb * qa;
        rq = ratio;
        const double pfn = ratio * s.pf[q];
        const double2 c = psi_wo_pf_exchanged_wave(m, tb, q, s.thc, s.thr, sfac * s.sgn[q >> 1], i1, i2, x1, x2, lane);
        if (lane == 0) {
          s.pfn[q] = pfn;
          s.psiq[q] = cscale(c, pfn);
        }
      }
      __syncthreads();
      // ---- Metropolis test (make_sample.f90:203-204)
      double2 psin = make_double2(0.0, 0.0);
      for (int qq = 0; qq < C; qq++) psin = cadd(psin, s.psiq[qq]);
      const double r1 = (cmod(psin) + a.shift) / (cmod(s.psi) + a.shift);
      if (uni(draw() < r1 * r1)) {
        // ---- accept: rank-4 update of this copy's X^-1 (calculate_new_InvM_two_child)
        if (q < C) {
          if constexpr (BLK) {
            // Sherman-Morrison-Woodbury for F' = F + e_ua r^T + c e_db^T:
            // B'[d][u] = B[d][u] + B[d][ua] al_u + x_d be_u with x = B c (old B),
            // al = -((1 + B[db].c) y - (y.c) B[db]) / R,
            // be = -((1 + y_ua) B[db] - B[db][ua] y) / R,
            // left pending for the next pass over F^-1
            const double ia = -1.0 / rq;
            if constexpr (NMAX > 128) {
              pal2 = make_double2(ia * ((1.0 + bdel) * by2.x - beps * brow2.x),
                                  ia * ((1.0 + bdel) * by2.y - beps * brow2.y));
              pbe2 = make_double2(ia * ((1.0 + bgam) * brow2.x - bbet * by2.x),
                                  ia * ((1.0 + bgam) * brow2.y - bbet * by2.y));
              pcv2 = bcv2;
            } else {
              pal = ia * ((1.0 + bdel) * by - beps * brow);
              pbe = ia * ((1.0 + bgam) * brow - bbet * by);
              pcv = bcv;
            }
            pua = bua;
            pend = true;
          } else {
          const double a_ = -pa, b_ = pb, c_ = qa, d_ = -qb;
          const double e_ = -bMa - vec_ba;
          const double f_ = invm_ab;
          const double Det = a_ * d_ - b_ * c_ - e_ * f_;
          const double InvDet = 1.0 / Det;
          const double fID = f_ * InvDet;
          const double eDet = e_ * Det;
          // Delta M(j,c) = A_c t_j + B_c s_j + C_c q_j + D_c p_j with the
          // column coefficients (A, B, C, D) published per column and the row
          // scalars of the lane's own rows.  The reference's separate fix-ups
          // of rows msa and msb (pfupdate_two.f90:382-383,397-398) are +D_c
          // and +C_c: they enter as p_j + [j == msa] and q_j + [j == msb].
          // The pass runs one column per lane (c = lane + 64 k) so that every
          // store moves 64 bits: a wider store needs a wait state before its
          // data registers are rewritten, which the compiler was seen to drop
          // in this loop.
          constexpr int KC = 2 * KD;
          double2* cf = (double2*)wv;  // (A_c, B_c) at cf[2c], (C_c, D_c) at cf[2c+1]
          const double2* pq = (const double2*)wv;  // (P_c, Q_c) of the ratio pass
          double ps[KC], pt[KC], pp[KC], pqv[KC];
#pragma unroll
          for (int k = 0; k < KC; k++) {
            const int c = lane + RBMPP_WAVE * k;
            ps[k] = pt[k] = pp[k] = pqv[k] = 0.0;
            if (c < N) {
              // s_i = InvDet InvM(i,msa) = -InvDet InvM(msa,i); t_i likewise
              ps[k] = -InvDet * fold_get(Mq, N, msa, c);
              pt[k] = -InvDet * fold_get(Mq, N, msb, c);
              pp[k] = pq[c].x;
              pqv[k] = pq[c].y;
            }
          }
          wave_sync();  // the (P, Q) table is read before the coefficients replace it
#pragma unroll
          for (int k = 0; k < KC; k++) {
            const int c = lane + RBMPP_WAVE * k;
            if (c < N) {
              const double p_i = pp[k], q_i = pqv[k], s_i = ps[k], t_i = pt[k];
              double A = a_ * q_i + c_ * p_i + eDet * s_i;
              double B = b_ * q_i + d_ * p_i - eDet * t_i;
              double Cc = -a_ * t_i - b_ * s_i + fID * p_i;
              double D = -c_ * t_i - d_ * s_i - fID * q_i;
              if (c == msa) { A += c_; B += d_; Cc += fID; }
              if (c == msb) { A += a_; B += b_; D -= fID; }
              cf[2 * c] = make_double2(A, B);
              cf[2 * c + 1] = make_double2(Cc, D);
              pp[k] = p_i + (c == msa ? 1.0 : 0.0);
              pqv[k] = q_i + (c == msb ? 1.0 : 0.0);
            }
          }
          wave_sync();
          // one read-modify-write pass over the folded rows: slot (d, j) is
          // M(j, j+d); the last row only for j < N/2
          {
            const __amdgpu_buffer_rsrc_t rs = block_rsrc(Mq, (int)(fold_size(N) * 8));
            int vo[KC], vl[KC], cl[KC];
#pragma unroll
            for (int k = 0; k < KC; k++) {
              const int c = lane + RBMPP_WAVE * k;
              vo[k] = c < N ? 8 * c : RBMPP_OOB;
              vl[k] = c < H ? 8 * c : RBMPP_OOB;
              cl[k] = c < N ? c : 0;
            }
            const int rowb = 8 * N;
            int d = 1;
            for (; d + UN - 1 <= H; d += UN) {
              double g[UN][KC];
#pragma unroll
              for (int u = 0; u < UN; u++)
#pragma unroll
                for (int k = 0; k < KC; k++) g[u][k] = bload1(rs, d + u == H ? vl[k] : vo[k], (d + u - 1) * rowb);
#pragma unroll
              for (int u = 0; u < UN; u++)
#pragma unroll
                for (int k = 0; k < KC; k++) {
                  int p0 = cl[k] + (d + u);
                  if (p0 >= N) p0 -= N;
                  const double2 ab0 = cf[2 * p0], cd0 = cf[2 * p0 + 1];
                  double v = g[u][k];
                  v = fma(ab0.x, pt[k], v);
                  v = fma(ab0.y, ps[k], v);
                  v = fma(cd0.x, pqv[k], v);
                  v = fma(cd0.y, pp[k], v);
                  bstore1(v, rs, d + u == H ? vl[k] : vo[k], (d + u - 1) * rowb);
                }
            }
            for (; d <= H; d++) {
#pragma unroll
              for (int k = 0; k < KC; k++) {
                const int vv = d == H ? vl[k] : vo[k];
                int p0 = cl[k] + d;
                if (p0 >= N) p0 -= N;
                const double2 ab0 = cf[2 * p0], cd0 = cf[2 * p0 + 1];
                double v = bload1(rs, vv, (d - 1) * rowb);
                v = fma(ab0.x, pt[k], v);
                v = fma(ab0.y, ps[k], v);
                v = fma(cd0.x, pqv[k], v);
                v = fma(cd0.y, pp[k], v);
                bstore1(v, rs, vv, (d - 1) * rowb);
              }
            }
          }
          }  // folded form
          // conditioning guard: a Pf ratio outside (1/g, g) means this copy
          // passes close to a singular configuration
          if (a.cond_guard > 0.0 && lane == 0) {
            const double ar = fabs(rq);
            if (!(ar * a.cond_guard > 1.0 && ar < a.cond_guard)) s.guard = 1;
          }
          // angles of this copy (update_theta)
          const int g = q >> 1;
          const double sp = (q & 1) ? -1.0 : 1.0;
          const int iw1 = tb.wmap[g * N + i1], iw2 = tb.wmap[g * N + i2];
          for (int f = lane; f < m.ar; f += RBMPP_WAVE) {
            double t = s.thr[q * m.ar + f];
            t = t - 2.0 * tb.Wr[f * N + iw1] * (double)x1 * sp;
            t = t - 2.0 * tb.Wr[f * N + iw2] * (double)x2 * sp;
            s.thr[q * m.ar + f] = t;
          }
          for (int f = lane; f < m.ac; f += RBMPP_WAVE) {
            double2 t = s.thc[q * m.ac + f];
            const double2 w1 = tb.Wc[f * N + iw1], w2 = tb.Wc[f * N + iw2];
            const double a1 = 2.0 * (double)x1 * sp, a2 = 2.0 * (double)x2 * sp;
            t.x = t.x - w1.x * a1; t.y = t.y - w1.y * a1;
            t.x = t.x - w2.x * a2; t.y = t.y - w2.y * a2;
            s.thc[q * m.ac + f] = t;
          }
          if (lane == 0) s.pf[q] = s.pfn[q];
        }
        __syncthreads();
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
        if (uni(s.guard)) {
          // a copy passed close to a singular configuration: its
          // fast-updated X^-1 would lose accuracy
          wg_recompute(m, s, Mw, gscr, 0, nullptr, nullptr);
          pend = false;
          status = uni(s.status);
          nrec++;
          nacc_tmp = 0;
          if (tid == 0) s.guard = 0;
          __syncthreads();
          if (status) break;
          res_load();
        }
      }
    }
    if (status) break;
    // ---- store the sample (make_sample.f90:228-232)
    if (i0 > 0) {
      const size_t row = (size_t)w * a.nsamp + (i0 - 1);
      for (int i = tid; i < N; i += nth) a.xs[row * N + i] = (signed char)s.x[i];
      if (tid == 0) a.psis[row] = s.psi;
      if constexpr (LDSMAT && MEAS) {
        if constexpr (BLK) {
          if (pend && q < C) flush_pending();
          pend = false;
          __syncthreads();  // every copy's F^-1 is current
        }
        wg_measure_sample(m, s, Mw, row, a);
      }
    }
  }
  if constexpr (RES) res_flush();
  else if constexpr (BLK) {
    if (pend && q < C) flush_pending();
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
  for (int qq = tid; qq < C; qq += nth) a.ws.pf[(size_t)w * C + qq] = s.pf[qq];
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

template <int NMAX, bool LDSMAT, bool MEAS, bool BLK, int RR = 0, int RL = -1>
__global__ __launch_bounds__(1024) void k_sweep(SweepArgs a) {
  __shared__ SweepSmem<NMAX, LDSMAT, RL> s;
  const int w = blockIdx.x;
  if (w >= a.nwalk) return;
  sweep_walker<NMAX, LDSMAT, MEAS, BLK, RR, RL>(a, w, s);
}

// Block-form models with 16 copies and N <= 104 use the slim layout (Slater_Elm
// lookups in LDS); the resident form also keeps rows [0, 16) of every copy's
// F^-1 in LDS, the rest streamed.  (Rows kept in registers as well spill: the
// sweep already fills the 128 registers that a 1024-thread workgroup allows.)
constexpr int RES_NMAX = 104, RES_RR = 0, RES_RL = 16;
static_assert(sizeof(SweepSmem<RES_NMAX, true, RES_RL>) <= 160 * 1024, "resident rows fit the LDS");
static_assert(sizeof(SweepSmem<RES_NMAX, true, 0>) <= 160 * 1024, "the slim layout fits the LDS");

// Walkers of several models (lattices of different sizes, couplings and
// parameters) in one launch: workgroup b belongs to the entry whose block
// range holds it and runs that model's specialisation of the sweep over the
// same shared-memory block.
static constexpr size_t BATCH_SMEM =
    std::max(std::max(sizeof(SweepSmem<64, true>), sizeof(SweepSmem<128, true>)), sizeof(SweepSmem<256, false>));

__global__ __launch_bounds__(1024) void k_sweep_batch(const SweepBatchEntry* __restrict__ ent, int nent) {
  __shared__ __align__(16) unsigned char smem[BATCH_SMEM];
  const int b = blockIdx.x;
  int k = 0;
  while (k + 1 < nent && b >= ent[k + 1].first_block) k++;
  k = __builtin_amdgcn_readfirstlane(k);
  const SweepArgs a = ent[k].a;
  const int w = b - ent[k].first_block;
  if (w >= a.nwalk) return;
  switch (ent[k].kind) {
    case SWEEP_BLK64: sweep_walker<64, true, false, true>(a, w, *reinterpret_cast<SweepSmem<64, true>*>(smem)); break;
    case SWEEP_BLK128: sweep_walker<128, true, false, true>(a, w, *reinterpret_cast<SweepSmem<128, true>*>(smem)); break;
    case SWEEP_BLK256: sweep_walker<256, false, false, true>(a, w, *reinterpret_cast<SweepSmem<256, false>*>(smem)); break;
    default: sweep_walker<256, false, false, false>(a, w, *reinterpret_cast<SweepSmem<256, false>*>(smem)); break;
  }
}

int sweep_kind(const DevModel& hm) {
  const int N = hm.N;
  if (hm.C > RBMPP_CMAX || hm.mr > RBMPP_MRMAX || hm.mc > RBMPP_MCMAX || (N & 1)) return -1;
  const bool blk = hm.block && hm.C == RBMPP_CMAX;
  if (N <= RBMPP_NLDS && !blk) return -1;  // skew form at N <= 104: not in the batch kernel
  if (N <= 64) return SWEEP_BLK64;
  if (N <= RBMPP_NLDS) return SWEEP_BLK128;
  if (N <= 256) return blk ? SWEEP_BLK256 : SWEEP_SKEW256;
  return -1;
}

void launch_sweep_batch(const SweepBatchEntry* d_ent, int nent, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_sweep_batch, dim3(nblocks), dim3(1024), 0, st, d_ent, nent);
}

template <int NMAX, bool LDSMAT>
static void launch_sweep_t(int C, bool blk, const SweepArgs& a, hipStream_t st) {
  // one wave per projection copy
  if constexpr (LDSMAT) {
    if (blk) {
      if (a.measure)
        hipLaunchKernelGGL((k_sweep<NMAX, LDSMAT, true, true>), dim3(a.nwalk), dim3(64 * C), 0, st, a);
      else
        hipLaunchKernelGGL((k_sweep<NMAX, LDSMAT, false, true>), dim3(a.nwalk), dim3(64 * C), 0, st, a);
      return;
    }
    if (a.measure) {
      hipLaunchKernelGGL((k_sweep<NMAX, LDSMAT, true, false>), dim3(a.nwalk), dim3(64 * C), 0, st, a);
      return;
    }
  }
  hipLaunchKernelGGL((k_sweep<NMAX, LDSMAT, false, false>), dim3(a.nwalk), dim3(64 * C), 0, st, a);
}

// RBMPP_SWEEP_RESIDENT=1 in the environment selects the resident form where
// it applies.  It is not the default: at 10x10 it takes 3.31 s per 80-sweep
// launch against 2.68 s for the streaming form (the pass over the LDS rows
// adds to the latency of the streamed rows instead of hiding it).
static bool resident_enabled() {
  const char* e = getenv("RBMPP_SWEEP_RESIDENT");
  return e && e[0] == '1';
}
// RBMPP_SWEEP_SLIM=0 selects the full layout (Slater_Elm from the global
// table) for comparisons
static bool slim_enabled() {
  const char* e = getenv("RBMPP_SWEEP_SLIM");
  return !(e && e[0] == '0');
}

int launch_sweep(const DevModel& hm, const SweepArgs& a, hipStream_t st) {
  const int N = hm.N;
  if (hm.C > RBMPP_CMAX || hm.mr > RBMPP_MRMAX || hm.mc > RBMPP_MCMAX || (N & 1)) return -1;
  if (N > RBMPP_NLDS && !a.ws.scratch) return -1;
  // block form: per-wave register recompute at N <= RBMPP_NLDS, F in LDS above
  const bool blk = hm.block && hm.C == RBMPP_CMAX;
  using RS = SweepSmem<RES_NMAX, true, RES_RL>;
  const bool slim = blk && !a.measure && N <= RES_NMAX && hm.C * hm.ar <= RS::MRL && hm.C * hm.ac <= RS::MCL;
  if (slim && hm.npair >= RES_RR + RES_RL && resident_enabled()) {
    hipLaunchKernelGGL((k_sweep<RES_NMAX, true, false, true, RES_RR, RES_RL>), dim3(a.nwalk), dim3(1024), 0, st, a);
    return 0;
  }
  if (slim && slim_enabled()) {
    hipLaunchKernelGGL((k_sweep<RES_NMAX, true, false, true, 0, 0>), dim3(a.nwalk), dim3(1024), 0, st, a);
    return 0;
  }
  if (N <= 64) launch_sweep_t<64, true>(hm.C, blk, a, st);
  else if (N <= RBMPP_NLDS) launch_sweep_t<128, true>(hm.C, blk, a, st);
  else if (N <= 256) {
    if (blk) {
      if (a.measure) return -1;  // the host keeps the skew form with the fused measurement
      hipLaunchKernelGGL((k_sweep<256, false, false, true>), dim3(a.nwalk), dim3(64 * hm.C), 0, st, a);
    } else {
      launch_sweep_t<256, false>(hm.C, false, a, st);
    }
  } else return -1;
  return 0;
}
