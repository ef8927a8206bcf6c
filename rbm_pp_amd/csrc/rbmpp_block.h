// det F and F^-1 of the block F of a block-form copy (X = [[0, F], [-F^T, 0]],
// DevModel::block) by one wave, shared by the estimator and the sampler's
// recompute.
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

}  // namespace rbmpp_block
