// Pfaffian and inverse of a skew-symmetric matrix of order n <= 112 held in
// registers across a workgroup of 1024 threads (16 waves), shared by the
// per-sample estimator and the sampler's recompute.
#pragma once
#include "rbmpp_device.h"

namespace rbmpp_skew {

constexpr int MPAD = 128;  // padded order of the pivot-row buffers

// LDS scratch of reg_skew_pf_inverse: two double-buffered pivot rows of MPAD
// doubles each (rowbuf[4 * MPAD], zero-filled once past the tile's 112 rows),
// the pivot index and value (sel[2], selv[2]).
// Register tile of the Gauss-Jordan sweep: element (i, j) of the zero-padded
// 112 x 128 matrix is x[i >> 4][j >> 6] of lane (j & 63) in wave (i & 15);
// each thread holds RA x RB = 7 x 2 elements.
constexpr int RA = 7, RB = 2;
static_assert(16 * RA >= 112 && 64 * RB >= 112, "register tile covers orders up to 112");

// The tile is one vector value: a runtime (wave-uniform) row or column index
// becomes an element extract instead of a dynamically indexed private array,
// which would live in scratch memory.
typedef double tile_t __attribute__((ext_vector_type(RA * RB)));
#define TILE(x, a, b) (x)[(a) * RB + (b)]

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform_d(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Eliminated indices of the padded order 128, as four wave-uniform words.
struct Done128 {
  uint32_t w0, w1, w2, w3;
  __device__ __forceinline__ uint32_t word(int k) const {
    return k == 0 ? w0 : (k == 1 ? w1 : (k == 2 ? w2 : w3));
  }
  __device__ __forceinline__ int first_free() const {
    int c = MPAD;
    if (~w3) c = 96 + __builtin_ctz(~w3);
    if (~w2) c = 64 + __builtin_ctz(~w2);
    if (~w1) c = 32 + __builtin_ctz(~w1);
    if (~w0) c = __builtin_ctz(~w0);
    return c;
  }
  // remaining indices strictly between lo and hi (lo < hi)
  __device__ __forceinline__ int free_between(int lo, int hi) const {
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int b0 = max(lo + 1 - 32 * k, 0), b1 = min(hi - 32 * k, 32);
      if (b1 > b0) {
        const uint32_t w = b1 - b0 == 32 ? 0xffffffffu : (((1u << (b1 - b0)) - 1u) << b0);
        cnt += __builtin_popcount(~word(k) & w);
      }
    }
    return cnt;
  }
  __device__ __forceinline__ void mark(int i) {
    const uint32_t bit = 1u << (i & 31);
    const int k = i >> 5;
    w0 |= k == 0 ? bit : 0u;
    w1 |= k == 1 ? bit : 0u;
    w2 |= k == 2 ? bit : 0u;
    w3 |= k == 3 ? bit : 0u;
  }
};

// One DPP step of the (|v|, index) arg-max: take the candidate moved in by
// CTRL when it is larger, or equal with a smaller index.  Lanes whose source
// lies outside the pattern keep their own candidate.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void amax_step(double& bv, int& bi) {
  const long long b = __builtin_bit_cast(long long, bv);
  const int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
  const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xf, false);
  const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xf, false);
  const int oi = __builtin_amdgcn_update_dpp(bi, bi, CTRL, ROWMASK, 0xf, false);
  const double ov = __builtin_bit_cast(double, ((long long)ohi << 32) | (unsigned)olo);
  if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
}
// Arg-max over the wave, valid in lane 63: row_shr 1, 2, 4, 8 reduce each row
// of 16 lanes into its last lane, row_bcast 15 and 31 carry rows into the next.
__device__ __forceinline__ void wave_amax_to63(double& bv, int& bi) {
  amax_step<0x111, 0xf>(bv, bi);
  amax_step<0x112, 0xf>(bv, bi);
  amax_step<0x114, 0xf>(bv, bi);
  amax_step<0x118, 0xf>(bv, bi);
  amax_step<0x142, 0xa>(bv, bi);
  amax_step<0x143, 0xc>(bv, bi);
}

// Pfaffian and inverse of the skew-symmetric n x n matrix held in the
// register tile; on return the tile holds X^-1.  Each step eliminates the
// pivot block {c, r}, r the first index of the largest |X(c, j)| among the
// remaining j (the Parlett-Reid rule).  With a = X(c, r) the exchange gives
//   X'(c,r) = -1/a, X'(r,c) = 1/a, X'(c,j) = -X(r,j)/a, X'(r,j) = X(c,j)/a,
//   X'(i,c) = -X(i,r)/a, X'(i,r) = X(i,c)/a,
//   X'(i,j) = X(i,j) + (X(i,c) X(r,j) - X(i,r) X(c,j))/a,
// and Pf(X) = (-1)^{#remaining between c and r} a Pf(Schur complement).
// EXACT keeps the tile exactly skew (the two products of each element are
// rounded apart, so X'(j,i) = -X'(i,j) and the diagonal stays 0); otherwise
// the update is two FMAs on prescaled pivot rows.
// Returns 0 for an exactly singular pivot.  Called by all 1024 threads.
template <bool EXACT>
__device__ __forceinline__ double reg_skew_pf_inverse(tile_t& x, int n, double* rowbuf, int* sel, double* selv) {
  const int tid = threadIdx.x, tr = uniform_i(tid >> 6), tc = tid & 63;
  Done128 done;
  {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int lo = 32 * k;
      w[k] = n >= lo + 32 ? 0u : (n <= lo ? 0xffffffffu : ~((1u << (n - lo)) - 1u));
    }
    done = Done128{w[0], w[1], w[2], w[3]};
  }
  double pf = 1.0;
  for (int step = 0; 2 * step < n; step++) {
    const int sb = step & 1;
    const int c = done.first_free();
    // pivot search along row c by the wave holding it, which also publishes
    // row c (with 1 in place of X(c,c) = 0: see the update below)
    if (tr == (c & 15)) {
      const int ac = c >> 4;
      double bv = -1.0;
      int bi = MPAD;
#pragma unroll
      for (int b = 0; b < RB; b++) {
        const int j = tc + 64 * b;
        double v = 0.0;
#pragma unroll
        for (int aa = 0; aa < RA; aa++)
          if (aa == ac) v = TILE(x, aa, b);
        rowbuf[(2 * sb + 0) * MPAD + j] = j == c ? 1.0 : v;
        const uint32_t w = tc < 32 ? done.word(2 * b) : done.word(2 * b + 1);
        const bool ok = !((w >> (tc & 31)) & 1u) && j != c;
        if (ok && fabs(v) > bv) { bv = fabs(v); bi = j; }
      }
      wave_amax_to63(bv, bi);
      const int r = __builtin_amdgcn_readlane(bi, 63);
      wave_sync();
      if (tc == 0) {
        sel[sb] = r;
        selv[sb] = r < MPAD ? rowbuf[(2 * sb + 0) * MPAD + r] : 0.0;
      }
    }
    __syncthreads();
    const int r = uniform_i(sel[sb]);
    const double a = uniform_d(selv[sb]);
    if (r >= n || a == 0.0) return 0.0;  // uniform
    // row r = -(column r), published by the lanes holding column r
#pragma unroll
    for (int b = 0; b < RB; b++)
      if (tc + 64 * b == r) {
#pragma unroll
        for (int aa = 0; aa < RA; aa++) {
          const int i = tr + 16 * aa;
          rowbuf[(2 * sb + 1) * MPAD + i] = i == r ? 1.0 : -TILE(x, aa, b);
        }
      }
    __syncthreads();
    pf *= (done.free_between(c, r) & 1) ? -a : a;
    const double inva = 1.0 / a;
    double rcj[RB], rrj[RB], rci[RA], rri[RA];
#pragma unroll
    for (int b = 0; b < RB; b++) {
      rcj[b] = rowbuf[(2 * sb + 0) * MPAD + tc + 64 * b];
      rrj[b] = rowbuf[(2 * sb + 1) * MPAD + tc + 64 * b];
    }
#pragma unroll
    for (int aa = 0; aa < RA; aa++) {
      rci[aa] = rowbuf[(2 * sb + 0) * MPAD + tr + 16 * aa];
      rri[aa] = rowbuf[(2 * sb + 1) * MPAD + tr + 16 * aa];
    }
    double rcs[RB], rrs[RB];
#pragma unroll
    for (int b = 0; b < RB; b++) { rcs[b] = rcj[b] * inva; rrs[b] = rrj[b] * inva; }
    // Schur-complement update of every element.  With the published rows
    // holding 1 at (c,c) and (r,r), the same expression also gives the new
    // pivot columns: X(i,c) + (X(r,i) * 1 - X(c,i) * X(r,c)) / a = X(r,i)/a =
    // -X(i,r)/a, and X(i,r) + (X(r,i) * X(c,r) - X(c,i) * 1) / a = X(i,c)/a.
#pragma unroll
    for (int aa = 0; aa < RA; aa++)
#pragma unroll
      for (int b = 0; b < RB; b++) {
        if constexpr (EXACT) {
          double t;
          {
#pragma clang fp contract(off)
            t = rri[aa] * rcj[b] - rci[aa] * rrj[b];
          }
          TILE(x, aa, b) = fma(t, inva, TILE(x, aa, b));
        } else {
          TILE(x, aa, b) = fma(rri[aa], rcs[b], fma(-rci[aa], rrs[b], TILE(x, aa, b)));
        }
      }
    // the pivot rows, set apart by the (at most two) waves holding them
    if (tr == (c & 15)) {
      const int ac = c >> 4;
#pragma unroll
      for (int aa = 0; aa < RA; aa++)
        if (aa == ac) {
#pragma unroll
          for (int b = 0; b < RB; b++) {
            const int j = tc + 64 * b;
            TILE(x, aa, b) = j == c ? 0.0 : (j == r ? -inva : -(EXACT ? rrj[b] * inva : rrs[b]));
          }
        }
    }
    if (tr == (r & 15)) {
      const int ar = r >> 4;
#pragma unroll
      for (int aa = 0; aa < RA; aa++)
        if (aa == ar) {
#pragma unroll
          for (int b = 0; b < RB; b++) {
            const int j = tc + 64 * b;
            TILE(x, aa, b) = j == r ? 0.0 : (j == c ? inva : (EXACT ? rcj[b] * inva : rcs[b]));
          }
        }
    }
    done.mark(c);
    done.mark(r);
  }
  return pf;
}

}  // namespace rbmpp_skew
