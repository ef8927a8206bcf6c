// det F and F^-1 of the block F of a block-form copy (X = [[0, F], [-F^T, 0]],
// DevModel::block) by one wave (np <= 52) or by a workgroup with F in LDS
// (np <= 128), shared by the estimators and the sampler's recompute.
#pragma once
#include "rbmpp_device.h"
#include "rbmpp_kernels.h"
#include "rbmpp_skew.h"

namespace rbmpp_block {

constexpr int HMX = RBMPP_NLDS / 2;  // largest np = N/2

// det F and F^-1 of copy S by one wave (NB >= np registers per lane).
// F(i, j) = Slater[site_i][site_(np+j) + N]; with the block weights (0 and 1)
// Slater_Elm is exactly skew there, so the lanes read it as
// -Slater[site_(np+j) + N][site_i], one row per j.  In-place Gauss-Jordan
// exchange steps on the pivots (p_k, k), p_k the remaining row with the
// largest |F(., k)| (smallest index on ties); the pivot row reaches the other
// lanes by readlane.  Each step also rotates the lane's row by one register,
// so that column k is always register 0 and every register index is a
// constant (a runtime index into a register array would go through scratch
// memory).  Afterwards row p_k, column m of the registers is F^-1(k, p_m),
// written to Bout[k * np + p_m] through the wave's LDS stage (8 * np
// doubles); piv holds np ints of LDS.  det F is the product of the pivots times
// the sign of the row order.  Returns 0 for an exactly singular pivot
// (uniform).  With drift != nullptr, drift[0] and drift[1] receive (in every
// lane) the largest |new - old| over the entries of Bout and the largest |new|.
template <int NB>
__device__ __forceinline__ double wave_block_inverse(const double* __restrict__ S, int N, int np,
                                                     const int* eidx, int* piv, double* stage, double* Bout,
                                                     int lane, double* drift) {
  double a[NB];
  const int si = lane < np ? eidx[lane] : 0;
#pragma unroll
  for (int j = 0; j < NB; j++) {
    double v = 0.0;
    if (j < np) {
      const double* row = S + (size_t)(eidx[np + j] + N) * 2 * N;
      if (lane < np) v = -row[si];
    }
    a[j] = v;
  }
  unsigned long long rem = np >= 64 ? ~0ull : ((1ull << np) - 1ull);
  double det = 1.0;
  int neg = 0, myk = 0;
  for (int k = 0; k < np; k++) {
    // register j holds column (k + j) mod NB
    const double ak = a[0];
    double bv = ((rem >> lane) & 1ull) ? fabs(ak) : -1.0;
    int bi = lane;
    rbmpp_skew::wave_amax_to63(bv, bi);
    const int p = __builtin_amdgcn_readlane(bi, 63);
    const double pv = readlane_d(ak, p);
    if (pv == 0.0) return 0.0;
    det *= pv;
    neg ^= __builtin_popcountll(rem & ((1ull << p) - 1ull)) & 1;
    rem &= ~(1ull << p);
    const bool isp = lane == p;
    if (isp) myk = k;
    if (lane == 0) piv[k] = p;
    const double inv = 1.0 / pv;
    // other rows: a - (a_k / pv) * pivot row; the pivot row: a / pv; the new
    // column k (-a_k / pv, or 1 / pv in the pivot row) moves to the last register
    const double t = isp ? 0.0 : ak * inv;
    const double sc = isp ? inv : 1.0;
    const double colk = isp ? inv : -(ak * inv);
#pragma unroll
    for (int j = 0; j + 1 < NB; j++) {
      const double pj = readlane_d(a[j + 1], p);
      a[j] = fma(-t, pj, a[j + 1] * sc);
    }
    a[NB - 1] = colk;
  }
  // back to register m = column m
  for (int r = np; r < NB; r++) {
    const double a0 = a[0];
#pragma unroll
    for (int j = 0; j + 1 < NB; j++) a[j] = a[j + 1];
    a[NB - 1] = a0;
  }
  wave_sync();  // piv is complete
  // write-back through the stage, 8 rows of F^-1 at a time: the lanes of
  // those rows put them in column order, then the wave stores the rows
  // (contiguous in Bout) with consecutive lanes on consecutive addresses
  double ldiff = 0.0, lmax = 0.0;
  for (int r0 = 0; r0 < np; r0 += 8) {
    if (lane < np && myk >= r0 && myk < r0 + 8) {
#pragma unroll
      for (int m = 0; m < NB; m++)
        if (m < np) stage[(myk - r0) * np + piv[m]] = a[m];
    }
    wave_sync();
    const int cnt = min(8, np - r0) * np;
    for (int e = lane; e < cnt; e += RBMPP_WAVE) {
      double* dst = Bout + (size_t)r0 * np + e;
      const double v = stage[e];
      if (drift) {
        ldiff = fmax(ldiff, fabs(*dst - v));
        lmax = fmax(lmax, fabs(v));
      }
      *dst = v;
    }
    wave_sync();  // the stage is rewritten by the next rows
  }
  if (drift) {
    drift[0] = wave_max(ldiff);
    drift[1] = wave_max(lmax);
  }
  return neg ? -det : det;
}

template <int NB>
__device__ __attribute__((noinline)) double block_inverse_nb(const double* S, int N, int np, const int* eidx,
                                                             int* piv, double* stage, double* Bout, int lane,
                                                             double* drift) {
  return wave_block_inverse<NB>(S, N, np, eidx, piv, stage, Bout, lane, drift);
}
__device__ __forceinline__ double copy_block_inverse(const double* S, int N, int np, const int* eidx, int* piv,
                                                     double* stage, double* Bout, int lane,
                                                     double* drift = nullptr) {
  if (np <= 8) return block_inverse_nb<8>(S, N, np, eidx, piv, stage, Bout, lane, drift);
  if (np <= 18) return block_inverse_nb<18>(S, N, np, eidx, piv, stage, Bout, lane, drift);
  if (np <= 32) return block_inverse_nb<32>(S, N, np, eidx, piv, stage, Bout, lane, drift);
  if (np <= 50) return block_inverse_nb<50>(S, N, np, eidx, piv, stage, Bout, lane, drift);
  return block_inverse_nb<HMX>(S, N, np, eidx, piv, stage, Bout, lane, drift);
}

// det F and F^-1 of one copy in block form for N > 104 (np up to 128, where
// a row per lane does not fit the registers), by the whole workgroup with F
// in LDS: T(u, d) = F(u, d) = Slater[site_u][site_(np+d) + N], then
// Gauss-Jordan exchange steps on the pivots (p_k, k), p_k the unused row with
// the largest |T(., k)| (smallest index on ties), as the per-wave elimination
// of rbmpp_block.h: pivot search by one wave, then the pivot row and column
// staged and the whole matrix updated, three barriers a step.  Afterwards
// F^-1(k, u) = T(perm[k], pinv[u]), written to Bout row-major.  det F is the
// product of the pivots times the sign of the row order.  Returns 0 for an
// exactly singular pivot (every thread).  With drift, drift[0] and drift[1]
// get the largest |new - old| and |new| over Bout (every thread).
// LDS: T np^2 doubles, rowb and colb np doubles each, used, perm, pinv np ints
// each, red 32 doubles, s_p / s_pv / s_cnt one value each.
static __device__ __attribute__((noinline)) double wg_block_inverse_lds(const double* __restrict__ S, int N, int np,
                                                                 const int* eidx, double* T, double* rowb,
                                                                 double* colb, int* used, int* perm, int* pinv,
                                                                 double* red, int* s_p, double* s_pv, int* s_cnt,
                                                                 double* Bout, double* drift) {
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & (RBMPP_WAVE - 1), wid = tid >> 6;
  const int npp = np * np;
  for (int e = tid; e < npp; e += nth) {
    const int u = e / np, d = e - u * np;
    T[e] = S[(size_t)eidx[u] * 2 * N + eidx[np + d] + N];
  }
  for (int r = tid; r < np; r += nth) used[r] = 0;
  __syncthreads();
  double det = 1.0;
  int neg = 0;
  for (int k = 0; k < np; k++) {
    if (wid == 0) {
      // rows lane and lane + 64: the unused one with the largest |T(., k)|
      double bv = -1.0;
      int bi = np;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int r = lane + RBMPP_WAVE * h;
        if (r < np && !used[r]) {
          const double v = fabs(T[r * np + k]);
          if (v > bv) { bv = v; bi = r; }
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, RBMPP_WAVE);
        const int oi = __shfl_xor(bi, o, RBMPP_WAVE);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      // unused rows below the pivot: the sign of the row order
      const int r0 = lane, r1 = lane + RBMPP_WAVE;
      const int below = __popcll(__ballot(r0 < np && r0 < bi && !used[r0])) +
                        __popcll(__ballot(r1 < np && r1 < bi && !used[r1]));
      if (lane == 0) {
        if (bi < np) {
          *s_p = bi;
          *s_pv = T[bi * np + k];
          *s_cnt = below;
          used[bi] = 1;
          perm[k] = bi;
        } else {  // no finite candidate (NaN entries): reported as singular
          *s_p = 0;
          *s_pv = 0.0;
          *s_cnt = 0;
        }
      }
    }
    __syncthreads();
    const int p = *s_p;
    const double pv = *s_pv;
    if (pv == 0.0) return 0.0;  // the same value in every thread
    det *= pv;
    neg ^= *s_cnt & 1;
    for (int c = tid; c < np; c += nth) rowb[c] = T[p * np + c];
    for (int r = tid; r < np; r += nth) colb[r] = T[r * np + k];
    __syncthreads();
    const double inv = 1.0 / pv;
    // other rows: T - (T(., k) / pv) * pivot row, column k -> -T(., k) / pv;
    // the pivot row: / pv, column k -> 1 / pv
    for (int e = tid; e < npp; e += nth) {
      const int r = e / np, c = e - r * np;
      double v;
      if (r == p) v = c == k ? inv : rowb[c] * inv;
      else v = c == k ? -(colb[r] * inv) : fma(-(colb[r] * inv), rowb[c], T[e]);
      T[e] = v;
    }
    __syncthreads();
  }
  for (int k = tid; k < np; k += nth) pinv[perm[k]] = k;
  __syncthreads();
  double ldiff = 0.0, lmax = 0.0;
  for (int e = tid; e < npp; e += nth) {
    const int k = e / np, u = e - k * np;
    const double v = T[perm[k] * np + pinv[u]];
    if (drift) {
      ldiff = fmax(ldiff, fabs(Bout[e] - v));
      lmax = fmax(lmax, fabs(v));
    }
    Bout[e] = v;
  }
  if (drift) {
    ldiff = wave_max(ldiff);
    lmax = wave_max(lmax);
    if (lane == 0) { red[2 * wid] = ldiff; red[2 * wid + 1] = lmax; }
    __syncthreads();
    double a = 0.0, b = 0.0;
    for (int w = 0; w < (nth >> 6); w++) { a = fmax(a, red[2 * w]); b = fmax(b, red[2 * w + 1]); }
    drift[0] = a;
    drift[1] = b;
    __syncthreads();  // red is reused by the caller
  }
  return neg ? -det : det;
}

}  // namespace rbmpp_block
