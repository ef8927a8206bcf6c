// Device-side building blocks shared by the sampler, estimator and SR
// kernels: model descriptor, Philox stream, amplitude of one projection
// copy, and the workgroup-cooperative Pfaffian + inverse of a skew-symmetric
// matrix.  gfx950 (CDNA4), wave64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RBMPP_WAVE 64

struct DevModel {
  int L, N, P, C, ac, ar, npair, nsl, nv, nv_rbm, mc, mr;
  int real_only;  // alphac == 0 and all para_mp_trans real: psi and O are real
  // spin weights with cos*sin = 0 for both copies of a pair: X is exactly
  // [[0, F], [-F^T, 0]] (F = up x down block) and the sampler keeps
  // F^-1 (N/2 x N/2) per copy instead of the folded X^-1
  int block;
  double J1, J2;
  double sp_cs[2], sp_cc[2], sp_ss[2], sp_w[2];
  const int* wmap;            // [P][N]   image of site i under operation g
  const int* invmap;          // [P][N]   inverse permutation
  const int* orb;             // [N][N]   orbital index of f_ij
  const unsigned char* subB;  // [N]
  const double2* chi;         // [P]     para_mp_trans
  const double* se;           // [C][2N][2N]  se[q][c][r] = Slater_Elm(r,c,q)
  const double* f;            // [nsl] pair-orbital parameters (Slater_Elm is built from them)
  const double* Wr;           // [ar][N]
  const double* br;           // [ar]
  const double2* Wc;          // [ac][N]
  const double2* bc;          // [ac]
  // ordered site pairs (A, B), A != B, of every orbital bin o under operation
  // g, orb[wmap_g(A)][wmap_g(B)] = o, in a fixed order: pairs
  // pbin_pairs[g][pbin_off[g][o] .. pbin_off[g][o + 1]), packed A | B << 16
  const int* pbin_off;        // [P][nsl + 1]
  const unsigned* pbin_pairs; // [P][N (N - 1)]
};

// ---------------------------------------------------------------- complex
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double cmod(double2 a) { return hypot(a.x, a.y); }

// ------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon et al., SC'11): counter (ctr_lo, ctr_hi, stream, 0),
// key (seed_lo, seed_hi).  One double per counter value, 53 random bits.
__device__ __forceinline__ double philox_u01(uint32_t k0, uint32_t k1, uint32_t stream, uint64_t ctr) {
  uint32_t c0 = (uint32_t)ctr, c1 = (uint32_t)(ctr >> 32), c2 = stream, c3 = 0u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    // high and low words of the 32 x 32 -> 64-bit products as two 32-bit
    // multiplies, which the scalar unit has (a 64-bit product is vector only)
    const uint32_t h0 = __umulhi(0xD2511F53u, c0), l0 = 0xD2511F53u * c0;
    const uint32_t h1 = __umulhi(0xCD9E8D57u, c2), l1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = h1 ^ c1 ^ k0;
    const uint32_t n2 = h0 ^ c3 ^ k1;
    c1 = l1; c3 = l0; c0 = n0; c2 = n2;
  }
  uint64_t z = ((uint64_t)c1 << 32) | c0;
  return (double)(z >> 11) * 0x1.0p-53;
}

struct Rng {
  uint32_t k0, k1, stream;
  uint64_t ctr;
  __device__ __forceinline__ double u01() { return philox_u01(k0, k1, stream, ctr++); }
  // make_sample.f90:254  nran(i) = mod(int(dble(i)*grnd()),i)+1, here 0-based
  __device__ __forceinline__ int nran0(int n) { return ((int)((double)n * u01())) % n; }
};

// ------------------------------------------------ raw buffer streaming
// One copy's N x N block of X^-1 through a wave-uniform buffer resource: the
// lane's column pair is a VGPR byte offset, the row a scalar offset, so a
// streamed row costs no 64-bit address arithmetic.  Lanes past the block's
// width get an offset beyond num_records: the hardware returns zeros for
// their loads and drops their stores (robust buffer access).
typedef unsigned int rbmpp_u32x4 __attribute__((ext_vector_type(4)));
constexpr int RBMPP_OOB = 0x40000000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_rsrc(const double* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ double2 bload2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ __forceinline__ void bstore2(double2 v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(rbmpp_u32x4, v), r, voff, soff, 0);
}
__device__ __forceinline__ double bload1(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void bstore1(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

// ------------------------------------------ folded skew-symmetric storage
// A skew-symmetric N x N matrix M (N divisible by 4) stored as its strictly
// upper part in cyclic-diagonal order: N/2 rows of N doubles, slot
// (d-1)*N + i holding M(i, (i+d) mod N) for d = 1..N/2; in the last row
// (d = N/2) only i < N/2 is used, the pair (i, i+N/2) being stored once.
// M(i,i) = 0 and M(j,i) = -M(i,j) are implied, so the two triangles cannot
// drift apart and a streamed pass moves half the bytes of the full matrix.
__device__ __forceinline__ size_t fold_size(int N) { return (size_t)N * (N / 2); }
// slot of the pair (r, c), r != c; *neg set when the slot holds M(c, r)
__device__ __forceinline__ int fold_slot(int N, int r, int c, bool* neg) {
  const int H = N >> 1;
  int d = c - r;
  if (d < 0) d += N;
  if (d < H || (d == H && r < H)) {
    *neg = false;
    return (d - 1) * N + r;
  }
  *neg = true;
  return (N - d - 1) * N + c;
}
__device__ __forceinline__ double fold_get(const double* G, int N, int r, int c) {
  if (r == c) return 0.0;
  bool neg;
  const double v = G[fold_slot(N, r, c, &neg)];
  return neg ? -v : v;
}

// ----------------------------------------------------------- reductions
// orders LDS traffic between the lanes of one wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, RBMPP_WAVE);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, RBMPP_WAVE));
  return v;
}
// v + (v moved by the DPP pattern CTRL); lanes without a source add 0
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_add(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return v + __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
// v * (v moved by the DPP pattern CTRL); lanes without a source multiply by 1
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_mul(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0x3ff00000, (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return v * __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
// Product over the wave, valid in lane 63 (the stages of wave_sum_to63)
__device__ __forceinline__ double wave_prod_to63(double v) {
  v = dpp_mul<0x111, 0xf>(v);
  v = dpp_mul<0x112, 0xf>(v);
  v = dpp_mul<0x114, 0xf>(v);
  v = dpp_mul<0x118, 0xf>(v);
  v = dpp_mul<0x142, 0xa>(v);
  v = dpp_mul<0x143, 0xc>(v);
  return v;
}
// Sum over the wave, valid in lane 63 (row_shr 1, 2, 4, 8 within rows of 16
// lanes, then row_bcast 15 and 31 across rows).
__device__ __forceinline__ double wave_sum_to63(double v) {
  v = dpp_add<0x111, 0xf>(v);
  v = dpp_add<0x112, 0xf>(v);
  v = dpp_add<0x114, 0xf>(v);
  v = dpp_add<0x118, 0xf>(v);
  v = dpp_add<0x142, 0xa>(v);
  v = dpp_add<0x143, 0xc>(v);
  return v;
}
// Four wave sums at once (gfx950 lane swaps, every lane active).
// v_permlane32_swap(x, y) gives [x lanes 0-31 | y lanes 0-31] and
// [x lanes 32-63 | y lanes 32-63]; v_permlane16_swap(x, y) gives
// [x row 0 | y row 0 | x row 2 | y row 2] and the odd rows likewise (rows of
// 16 lanes).  Adding each pair folds a and b into one register (a on lanes
// 0-31, b on 32-63), c and d into another, then the four into rows 0-3 as
// a, c, b, d; the row stages of wave_sum_to63 finish each row in its last
// lane: lane 15 holds the sum of a, lane 47 that of b, lane 31 that of c,
// lane 63 that of d.
template <bool SIXTEEN>
__device__ __forceinline__ double swap_add(double x, double y) {
  const unsigned long long bx = __builtin_bit_cast(unsigned long long, x);
  const unsigned long long by = __builtin_bit_cast(unsigned long long, y);
  const unsigned xl = (unsigned)bx, xh = (unsigned)(bx >> 32), yl = (unsigned)by, yh = (unsigned)(by >> 32);
  unsigned f0, f1, h0, h1;
  if constexpr (SIXTEEN) {
    const auto lo = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
    f0 = lo[0]; f1 = lo[1]; h0 = hi[0]; h1 = hi[1];
  } else {
    const auto lo = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
    f0 = lo[0]; f1 = lo[1]; h0 = hi[0]; h1 = hi[1];
  }
  const double first = __builtin_bit_cast(double, ((unsigned long long)h0 << 32) | f0);
  const double second = __builtin_bit_cast(double, ((unsigned long long)h1 << 32) | f1);
  return first + second;
}
__device__ __forceinline__ double wave_sum4(double a, double b, double c, double d) {
  double v = swap_add<true>(swap_add<false>(a, b), swap_add<false>(c, d));
  v = dpp_add<0x111, 0xf>(v);
  v = dpp_add<0x112, 0xf>(v);
  v = dpp_add<0x114, 0xf>(v);
  v = dpp_add<0x118, 0xf>(v);
  return v;
}

// Two wave sums at once: the 32-lane swap puts a on lanes 0-31 and b on
// 32-63, the row stages and the row_bcast:15 stage of wave_sum_to63 finish
// them: lane 31 holds the sum of a, lane 63 that of b.
__device__ __forceinline__ double wave_sum2(double a, double b) {
  double v = swap_add<false>(a, b);
  v = dpp_add<0x111, 0xf>(v);
  v = dpp_add<0x112, 0xf>(v);
  v = dpp_add<0x114, 0xf>(v);
  v = dpp_add<0x118, 0xf>(v);
  v = dpp_add<0x142, 0xa>(v);
  return v;
}

// the value of lane `l` (wave-uniform) in every lane
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// The RBM tables an amplitude update reads: the model's own (global memory)
// or copies staged in LDS by the kernel
struct RbmTables {
  const double* Wr;
  const double2* Wc;
  const int* wmap;
  const unsigned char* subB;
};
__device__ __forceinline__ RbmTables model_tables(const DevModel& m) { return {m.Wr, m.Wc, m.wmap, m.subB}; }

// amplitude of projection copy q without the Pfaffian (sub.f90:30-48), from
// angle arrays thc[q*ac+f], thr[q*ar+f] and the sign of operation g.
__device__ __forceinline__ double2 psi_wo_pf(const DevModel& m, int q, const double2* thc,
                                             const double* thr, int sgn) {
  int g = q >> 1, sidx = q & 1;
  double2 c = make_double2(1.0, 0.0);
  for (int k = 0; k < m.ac; k++) c = cscale(c, 0.5);
  for (int f = 0; f < m.ac; f++) {
    double2 t = thc[q * m.ac + f];
    double sn, cs, sh = sinh(t.x), ch = cosh(t.x);
    sincos(t.y, &sn, &cs);
    c = cmul(c, make_double2(ch * cs, sh * sn));
  }
  double r1 = 1.0;
  for (int f = 0; f < m.ar; f++) r1 = r1 * cosh(thr[q * m.ar + f]);
  c = cscale(c, r1);
  c = cscale(c, (double)sgn);
  c = cmul(c, m.chi[g]);
  c = cscale(c, m.sp_w[sidx]);
  return c;
}

// Same, for the configuration after exchanging the spins at sites i1, i2
// (update_theta, RBM.f90:95-148): angles are updated on the fly.
__device__ __forceinline__ double2 psi_wo_pf_exchanged(const DevModel& m, const RbmTables& tb, int q,
                                                       const double2* thc, const double* thr, int sgn_new,
                                                       int i1, int i2, int x1, int x2) {
  int g = q >> 1, sidx = q & 1;
  double s = sidx ? -1.0 : 1.0;
  int iw1 = tb.wmap[g * m.N + i1], iw2 = tb.wmap[g * m.N + i2];
  double2 c = make_double2(1.0, 0.0);
  for (int k = 0; k < m.ac; k++) c = cscale(c, 0.5);
  for (int f = 0; f < m.ac; f++) {
    double2 t = thc[q * m.ac + f];
    double2 w1 = tb.Wc[f * m.N + iw1], w2 = tb.Wc[f * m.N + iw2];
    double a1 = 2.0 * (double)x1 * s, a2 = 2.0 * (double)x2 * s;
    t.x = t.x - w1.x * a1; t.y = t.y - w1.y * a1;
    t.x = t.x - w2.x * a2; t.y = t.y - w2.y * a2;
    double sn, cs, sh = sinh(t.x), ch = cosh(t.x);
    sincos(t.y, &sn, &cs);
    c = cmul(c, make_double2(ch * cs, sh * sn));
  }
  double r1 = 1.0;
  for (int f = 0; f < m.ar; f++) {
    double t = thr[q * m.ar + f];
    t = t - 2.0 * tb.Wr[f * m.N + iw1] * (double)x1 * s;
    t = t - 2.0 * tb.Wr[f * m.N + iw2] * (double)x2 * s;
    r1 = r1 * cosh(t);
  }
  c = cscale(c, r1);
  c = cscale(c, (double)sgn_new);
  c = cmul(c, m.chi[g]);
  c = cscale(c, m.sp_w[sidx]);
  return c;
}

__device__ __forceinline__ double2 psi_wo_pf_exchanged(const DevModel& m, int q, const double2* thc,
                                                       const double* thr, int sgn_new, int i1, int i2,
                                                       int x1, int x2) {
  return psi_wo_pf_exchanged(m, model_tables(m), q, thc, thr, sgn_new, i1, i2, x1, x2);
}

// Same, evaluated by one wave: lane f takes hidden units f, f+64, ... and the
// factors are multiplied across the wave (butterfly), so the cosh and the
// table reads of all hidden units proceed in parallel.  Every lane returns
// the amplitude.  The products are those of the serial form in another
// order (the 1/2 of each complex factor is exact).
__device__ __forceinline__ double2 psi_wo_pf_exchanged_wave(const DevModel& m, const RbmTables& tb, int q,
                                                            const double2* thc, const double* thr, int sgn_new,
                                                            int i1, int i2, int x1, int x2, int lane) {
  const int g = q >> 1, sidx = q & 1;
  const double s = sidx ? -1.0 : 1.0;
  const int iw1 = tb.wmap[g * m.N + i1], iw2 = tb.wmap[g * m.N + i2];
  const double a1 = 2.0 * (double)x1 * s, a2 = 2.0 * (double)x2 * s;
  double2 c = make_double2(1.0, 0.0);
  for (int f = lane; f < m.ac; f += RBMPP_WAVE) {
    double2 t = thc[q * m.ac + f];
    const double2 w1 = tb.Wc[f * m.N + iw1], w2 = tb.Wc[f * m.N + iw2];
    t.x = t.x - w1.x * a1; t.y = t.y - w1.y * a1;
    t.x = t.x - w2.x * a2; t.y = t.y - w2.y * a2;
    double sn, cs;
    const double sh = sinh(t.x), ch = cosh(t.x);
    sincos(t.y, &sn, &cs);
    c = cmul(c, make_double2(0.5 * (ch * cs), 0.5 * (sh * sn)));
  }
  double r1 = 1.0;
  for (int f = lane; f < m.ar; f += RBMPP_WAVE) {
    double t = thr[q * m.ar + f];
    t = t - 2.0 * tb.Wr[f * m.N + iw1] * (double)x1 * s;
    t = t - 2.0 * tb.Wr[f * m.N + iw2] * (double)x2 * s;
    r1 = r1 * cosh(t);
  }
  // the real factors by DPP stages into lane 63; the complex ones (models
  // with alphac > 0) by a butterfly
  r1 = readlane_d(wave_prod_to63(r1), 63);
  if (m.ac > 0) {
#pragma unroll
    for (int o = 1; o < RBMPP_WAVE; o <<= 1) {
      const double2 oc = make_double2(__shfl_xor(c.x, o, RBMPP_WAVE), __shfl_xor(c.y, o, RBMPP_WAVE));
      c = cmul(c, oc);
    }
  }
  c = cscale(c, r1);
  c = cscale(c, (double)sgn_new);
  c = cmul(c, m.chi[g]);
  c = cscale(c, m.sp_w[sidx]);
  return c;
}

// --------------------------------------------------------------------
// Workgroup-cooperative Pfaffian and inverse of a skew-symmetric n x n
// matrix A (row-major, leading dimension ld; LDS or global).  Gauss-Jordan
// sweeps over 2x2 pivot blocks with symmetric pivoting on the largest
// |A(k,j)|, j>k (the Parlett-Reid pivot rule of dskpfa('U','P'),
// matrix.f90:133-151).  The Schur-complement update of each sweep is the
// Parlett-Reid step, so Pf = (-1)^{#swaps} prod_k A(k,k+1); the completed
// sweeps leave the inverse of the permuted matrix in A.
//   perm[n] : returns the symmetric permutation (row r of A <-> original perm[r])
//   buf[4*ld] : workspace; red[2*nwaves]: reduction scratch
// Returns 0 when a pivot block is exactly singular (Pf = 0).
// Must be called by all threads of the workgroup.
__device__ __forceinline__ double wg_skew_inverse(double* A, int n, int ld, int* perm, double* buf,
                                                  double* red_v, int* red_i, int* s_piv) {
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & (RBMPP_WAVE - 1), wid = tid / RBMPP_WAVE;
  double* R0 = buf;
  double* R1 = buf + ld;
  double* C0 = buf + 2 * ld;
  double* C1 = buf + 3 * ld;
  for (int i = tid; i < n; i += nth) perm[i] = i;
  double pf = 1.0;
  __syncthreads();
  for (int k = 0; k < n - 1; k += 2) {
    // pivot search over row k, columns k+1..n-1 (first index of the maximum)
    if (wid == 0) {
      double bv = -1.0;
      int bi = n;
      for (int j = k + 1 + lane; j < n; j += RBMPP_WAVE) {
        double v = fabs(A[(size_t)k * ld + j]);
        if (v > bv) { bv = v; bi = j; }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        double ov = __shfl_xor(bv, o, RBMPP_WAVE);
        int oi = __shfl_xor(bi, o, RBMPP_WAVE);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      if (lane == 0) *s_piv = bi;
    }
    __syncthreads();
    const int p = *s_piv;
    if (p != k + 1) {
      for (int j = tid; j < n; j += nth) {
        double t = A[(size_t)(k + 1) * ld + j];
        A[(size_t)(k + 1) * ld + j] = A[(size_t)p * ld + j];
        A[(size_t)p * ld + j] = t;
      }
      __syncthreads();
      for (int i = tid; i < n; i += nth) {
        double t = A[(size_t)i * ld + k + 1];
        A[(size_t)i * ld + k + 1] = A[(size_t)i * ld + p];
        A[(size_t)i * ld + p] = t;
      }
      if (tid == 0) { int t = perm[k + 1]; perm[k + 1] = perm[p]; perm[p] = t; }
      pf = -pf;
    }
    __syncthreads();
    const double e01 = A[(size_t)k * ld + k + 1];
    const double e10 = A[(size_t)(k + 1) * ld + k];
    if (e01 == 0.0 || e10 == 0.0) return 0.0;
    pf *= e01;
    const double i01 = 1.0 / e01, i10 = 1.0 / e10;
    for (int j = tid; j < n; j += nth) {
      R0[j] = A[(size_t)k * ld + j];
      R1[j] = A[(size_t)(k + 1) * ld + j];
      C0[j] = A[(size_t)j * ld + k];
      C1[j] = A[(size_t)j * ld + k + 1];
    }
    __syncthreads();
    // new pivot rows: row k <- R1/e10, row k+1 <- R0/e01; new pivot columns:
    // col k <- -C1/e01, col k+1 <- -C0/e10; pivot block <- E^-1; the rest
    // A_ij -= C0_i * R1_j / e10 + C1_i * R0_j / e01.
    const int nn = n * n;
    for (int e = tid; e < nn; e += nth) {
      int i = e / n, j = e - i * n;
      bool ik = (i == k) || (i == k + 1), jk = (j == k) || (j == k + 1);
      double v;
      if (ik && jk) {
        v = (i == j) ? 0.0 : ((i == k) ? i10 : i01);
      } else if (ik) {
        v = (i == k) ? R1[j] * i10 : R0[j] * i01;
      } else if (jk) {
        v = (j == k) ? -C1[i] * i01 : -C0[i] * i10;
      } else {
        v = A[(size_t)i * ld + j] - C0[i] * (R1[j] * i10) - C1[i] * (R0[j] * i01);
      }
      A[(size_t)i * ld + j] = v;
    }
    __syncthreads();
  }
  (void)red_v; (void)red_i;
  return pf;
}

// calc_theta (RBM.f90:31-91) and sign_PP (interface_for_PP.f90:71-106) of the
// configuration s.x with electron order s.eidx.  With check=1 the results are
// compared with the incrementally maintained values (make_sample.f90:84-92)
// and s.status is set on a mismatch.  S is any shared-memory block with
// members x, eidx, thr, thc, sgn, sgnbits, status.
template <class S>
__device__ void wg_theta_sign(const DevModel& m, S& s, int check) {
  const int tid = threadIdx.x, nth = blockDim.x;
  const int N = m.N, C = m.C, P = m.P;
  for (int idx = tid; idx < C * m.ar; idx += nth) {
    const int q = idx / m.ar, f = idx - q * m.ar, g = q >> 1;
    const double sp = (q & 1) ? -1.0 : 1.0;
    double t = m.br[f];
    for (int i = 0; i < N; i++) t = t + m.Wr[f * N + m.wmap[g * N + i]] * (double)s.x[i] * sp;
    if (check && fabs(t - s.thr[idx]) > 1.0e-5) s.status = -40;
    s.thr[idx] = t;
  }
  for (int idx = tid; idx < C * m.ac; idx += nth) {
    const int q = idx / m.ac, f = idx - q * m.ac, g = q >> 1;
    const double sp = (q & 1) ? -1.0 : 1.0;
    double2 t = m.bc[f];
    for (int i = 0; i < N; i++) {
      const double2 w = m.Wc[f * N + m.wmap[g * N + i]];
      const double a = (double)s.x[i] * sp;
      t.x = t.x + w.x * a;
      t.y = t.y + w.y * a;
    }
    if (check && cmod(make_double2(t.x - s.thc[idx].x, t.y - s.thc[idx].y)) > 1.0e-5) s.status = -40;
    s.thc[idx] = t;
  }
  for (int g = tid; g < P; g += nth) s.sgnbits[g] = 0u;
  __syncthreads();
  for (int e = tid; e < P * N; e += nth) {
    const int g = e / N, i1 = e - g * N;
    if (i1 == 0) continue;
    const int t1 = m.wmap[g * N + s.eidx[i1]];
    unsigned par = (i1 >= m.npair && m.subB[t1]) ? 1u : 0u;
    for (int i2 = 0; i2 < i1; i2++) par ^= (m.wmap[g * N + s.eidx[i2]] > t1) ? 1u : 0u;
    if (par) atomicXor(&s.sgnbits[g], 1u);
  }
  __syncthreads();
  for (int g = tid; g < P; g += nth) {
    const int sg = (s.sgnbits[g] & 1u) ? -1 : 1;
    if (check && sg != s.sgn[g]) s.status = -42;
    s.sgn[g] = sg;
  }
  __syncthreads();
}
