// Reductions over the measured samples and the linear algebra of the SR step
// on gfx950: PP-block scaling and diagonal shift, blocked FP64 Cholesky
// (replaces ScaLAPACK pdposv, SR_pdposv.f90:232), triangular solves,
// residual / lambda products.  The S build itself is in rbmpp_syrk.hip.
#include "rbmpp_sr.h"

#include <algorithm>

// ---------------------------------------------------------------- sums
// main.f90:351-361 operands: per-block partials of
// {Etot, Etot2, norm, wf_av} (calc_phys_quantity.f90:105-109).
__global__ void k_scalar_partials(int ns, const double2* __restrict__ psi, const double* __restrict__ rw,
                                  const double2* __restrict__ eloc, double* part) {
  __shared__ double red[4][8];
  double e = 0, e2 = 0, nr = 0, wf = 0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
    const double2 p = psi[s], el = eloc[s];
    const double r = rw[s], ap = hypot(p.x, p.y);
    e += r * (p.x * el.x + p.y * el.y);
    e2 += el.x * el.x + el.y * el.y;
    nr += (ap * r) * (ap * r);
    wf += ap;
  }
  e = wave_sum(e); e2 = wave_sum(e2); nr = wave_sum(nr); wf = wave_sum(wf);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { red[0][wid] = e; red[1][wid] = e2; red[2][wid] = nr; red[3][wid] = wf; }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[threadIdx.x][w];
    part[blockIdx.x * 4 + threadIdx.x] = t;
  }
}

// g(k) = sum_s 2 Re(conj(E_s) O_s(k)),  ovec(k) = sum_s rw_s Re(conj(psi_s) O_s(k))
// (calc_phys_quantity.f90:228-230).  grid.y splits the samples; partials
// [grid.y][2][nv] are summed in a fixed order afterwards.
__global__ void k_gvec_partials(int ns, int nv, int cplx_rows, const double* __restrict__ orows,
                                const double2* __restrict__ psi, const double* __restrict__ rw,
                                const double2* __restrict__ eloc, double* part) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nv) return;
  const int chunk = (ns + gridDim.y - 1) / gridDim.y;
  const int s0 = blockIdx.y * chunk, s1 = min(ns, s0 + chunk);
  double g = 0, o = 0;
  for (int s = s0; s < s1; s++) {
    const double2 el = eloc[s], p = psi[s];
    const double r = rw[s];
    const double ore = orows[(size_t)(cplx_rows ? 2 * s : s) * nv + k];
    const double oim = cplx_rows ? orows[(size_t)(2 * s + 1) * nv + k] : 0.0;
    g += 2.0 * (el.x * ore + el.y * oim);
    o += r * (p.x * ore + p.y * oim);
  }
  part[((size_t)blockIdx.y * 2) * nv + k] = g;
  part[((size_t)blockIdx.y * 2 + 1) * nv + k] = o;
}

__global__ void k_sum_partials(int nv, int nchunk, const double* __restrict__ part, double* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 2 * nv) return;
  const int which = k / nv, kk = k - which * nv;
  double t = 0;
  for (int c = 0; c < nchunk; c++) t += part[((size_t)c * 2 + which) * nv + kk];
  out[k] = t;
}

// centring coefficient of every O row: kappa = rw Re(psi) (re row),
// rw Im(psi) (im row), so that c = O_row - kappa * obar
__global__ void k_row_kappa(int ns, int cplx_rows, const double2* __restrict__ psi,
                            const double* __restrict__ rw, double* kappa) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  if (cplx_rows) {
    kappa[2 * s] = rw[s] * psi[s].x;
    kappa[2 * s + 1] = rw[s] * psi[s].y;
  } else {
    kappa[s] = rw[s] * psi[s].x;
  }
}

// y = A x (row-major, one wave per row)
__global__ void k_gemv(int n, const double* __restrict__ A, const double* __restrict__ x, double* y) {
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double t = 0;
  for (int j = lane; j < n; j += 64) t = fma(A[(size_t)row * n + j], x[j], t);
  t = wave_sum(t);
  if (lane == 0) y[row] = t;
}

// ------------------------------------------------------- blocked Cholesky
// Right-looking, lower triangle of a row-major SPD matrix, block NB=64.
#define CH_NB 64

__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Factor the diagonal block [k0, k0+nb)^2 and invert its factor with one
// wave.  Factor: lane i keeps row i in registers (rows and columns past nb
// hold the identity); step j takes the pivot from lane j, scales column j and
// applies the rank-1 update, reading L(k, j) from lane k (the textbook
// right-looking loop, the same operations in the same order).  Column j of L
// also goes to LDS.  Inverse X = L^-1: lane c forms column c by column-wise
// forward substitution on e_c, reading column k of L from LDS as a broadcast.
// L goes back into A's lower triangle, X (row-major, zero above the diagonal
// and past nb) into dinv.  A pivot <= 0 or NaN sets info (first failing
// pivot, 1-based) and leaves the block as it is.
__global__ __launch_bounds__(64) void k_potrf_inv(int n, int k0, double* A, double* dinv, int* info) {
  __shared__ double Lc[CH_NB][CH_NB + 1];  // Lc[j][i] = L(i, j)
  const int nb = min(CH_NB, n - k0), i = threadIdx.x;
  double r[CH_NB];
  if (i < nb) {
    const double* Ar = A + (size_t)(k0 + i) * n + k0;
#pragma unroll
    for (int c = 0; c < CH_NB; c++) r[c] = c < nb ? Ar[c] : 0.0;
  } else {
#pragma unroll
    for (int c = 0; c < CH_NB; c++) r[c] = c == i ? 1.0 : 0.0;
  }
#pragma unroll
  for (int j = 0; j < CH_NB; j++) {
    const double p = lane_value(r[j], j);
    if (!(p > 0.0)) {
      if (i == 0 && *info == 0) *info = k0 + j + 1;
      return;
    }
    const double d = sqrt(p);
    // lane j carries the pivot; its row's entries right of the diagonal
    // are never read
    const double lij = i == j ? d : (i > j ? r[j] / d : 0.0);
    Lc[j][i] = lij;
#pragma unroll
    for (int k = j + 1; k < CH_NB; k++) r[k] = fma(-lij, lane_value(lij, k), r[k]);
  }
  __syncthreads();
  double x[CH_NB];
#pragma unroll
  for (int k = 0; k < CH_NB; k++) x[k] = k == i ? 1.0 : 0.0;
#pragma unroll
  for (int k = 0; k < CH_NB; k++) {
    x[k] = x[k] / Lc[k][k];
#pragma unroll
    for (int rr = k + 1; rr < CH_NB; rr++) x[rr] = fma(-Lc[k][rr], x[k], x[rr]);
  }
  double* D = dinv + (size_t)(k0 / CH_NB) * CH_NB * CH_NB;
#pragma unroll
  for (int rr = 0; rr < CH_NB; rr++) D[rr * CH_NB + i] = (i <= rr && rr < nb) ? x[rr] : 0.0;
  for (int e = i; e < CH_NB * CH_NB; e += 64) {
    const int rr = e / CH_NB, c = e - rr * CH_NB;
    if (rr < nb && c <= rr) A[(size_t)(k0 + rr) * n + k0 + c] = Lc[c][rr];
  }
}

typedef double f64x4_t __attribute__((ext_vector_type(4)));

// Panel below the diagonal block, L21 = A21 L11^-T = A21 X^T: 64 rows per
// workgroup on the FP64 matrix cores (16 x 16 x 4 blocks, 2 x 2 per wave;
// lane l supplies row l % 16 at k = 4 s + l / 16 of both operands, output
// register r is element (l / 16 + 4 r, l % 16) of a block).
__global__ __launch_bounds__(256) void k_trsm_panel(int n, int k0, double* A, const double* __restrict__ dinv) {
  constexpr int TS = 64;
  __shared__ double Ai[TS][CH_NB + 1];
  __shared__ double Xs[TS][CH_NB + 1];
  const int nb = min(CH_NB, n - k0);
  const int r0 = k0 + nb + blockIdx.x * TS;
  const double* D = dinv + (size_t)(k0 / CH_NB) * CH_NB * CH_NB;
  for (int e = threadIdx.x; e < TS * CH_NB; e += 256) {
    const int r = e / CH_NB, k = e - r * CH_NB;
    const int gi = r0 + r;
    Ai[r][k] = (gi < n && k < nb) ? A[(size_t)gi * n + k0 + k] : 0.0;
    Xs[r][k] = D[e];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  const int m = lane & 15, ks = lane >> 4;
  f64x4_t acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = f64x4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int s4 = 0; s4 < CH_NB / 4; s4++) {
    const int k = 4 * s4 + ks;
    double a[2], b[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      a[x] = Ai[wi * 32 + 16 * x + m][k];
      b[x] = Xs[wj * 32 + 16 * x + m][k];  // B(k, j) = X(j, k)
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = r0 + wi * 32 + 16 * x + ks + 4 * r;
        const int j = wj * 32 + 16 * y + m;
        if (gi < n && j < nb) A[(size_t)gi * n + k0 + j] = acc[x][y][r];
      }
}

// trailing update of the lower triangle, A_ij -= sum_k L_ik L_jk for
// i >= j >= t0 and k over the panel [k0, k0+nb): 64 x 64 tiles on the FP64
// matrix cores (v_mfma_f64_16x16x4), both row panels staged in LDS; each wave
// a 32 x 32 quadrant of 2 x 2 blocks.  MFMA lane l supplies row l % 16 at
// k = 4 s + l / 16 of both operands; its output register r is element
// (l / 16 + 4 r, l % 16) of the block.
// With col_only, workgroup p takes the tile (p, 0) of the first tile column
// only (the look-ahead of launch_cholesky).
__global__ __launch_bounds__(256) void k_syrk_trailing(int n, int k0, int t0, int ntile, int col_only, double* A) {
  constexpr int TS = 64;
  __shared__ double Li[TS][CH_NB + 1];
  __shared__ double Lj[TS][CH_NB + 1];
  int p = blockIdx.x, tj = 0;  // lower-triangle tile pairs ti >= tj
  if (!col_only)
    while (p >= ntile - tj) { p -= ntile - tj; tj++; }
  const int ti = tj + p;
  const int nb = min(CH_NB, n - k0);
  for (int e = threadIdx.x; e < TS * CH_NB; e += 256) {
    const int r = e / CH_NB, k = e - r * CH_NB;
    const int gi = t0 + ti * TS + r, gj = t0 + tj * TS + r;
    Li[r][k] = (gi < n && k < nb) ? A[(size_t)gi * n + k0 + k] : 0.0;
    Lj[r][k] = (gj < n && k < nb) ? A[(size_t)gj * n + k0 + k] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  const int m = lane & 15, ks = lane >> 4;
  f64x4_t acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = f64x4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int s4 = 0; s4 < CH_NB / 4; s4++) {
    const int k = 4 * s4 + ks;
    double a[2], b[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      a[x] = Li[wi * 32 + 16 * x + m][k];
      b[x] = Lj[wj * 32 + 16 * x + m][k];
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = t0 + ti * TS + wi * 32 + 16 * x + ks + 4 * r;
        const int gj = t0 + tj * TS + wj * 32 + 16 * y + m;
        if (gi < n && gj <= gi) A[(size_t)gi * n + gj] -= acc[x][y][r];
      }
}

// ------------------------------------------------- device SR pipeline
// Scalars of the SR system, from the diagonal of S (SR_pdposv.f90:18-40,
// 78-140): block scale scal = sqrt(mean positive RBM diagonal / mean positive
// PP diagonal), Sdiag_av = mean positive diagonal of D S D, shift =
// Sdiag_av * s_factor; and the structural checks of :32-33, 78-100 as flags.
// One workgroup, partial sums per thread then a fixed-order tree (the result
// does not depend on timing).
//   sc: [0] scal [1] shift [2] Sdiag_av [3] s_factor (input) [4] lambda^2
//       [5] |r|^2 [6] |g|^2
//   fl: [0] i1 [1] i2 [2] first bad index [3] kind (1 negative, 2 zero not
//       f_ii, 3 positive f_ii) [4] Cholesky info
__global__ __launch_bounds__(1024) void k_sr_prepare(int nv, int nvr, const double* __restrict__ S,
                                                     const unsigned char* __restrict__ fii, double* sc, int* fl) {
  __shared__ double rs[4][1024];
  __shared__ int ri[2][1024];
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = nv;
  double sR = 0, sP = 0, cR = 0, cP = 0;
  int i1 = 0, i2 = 0;
  __syncthreads();
  for (int k = tid; k < nv; k += 1024) {
    const double d = S[(size_t)k * nv + k];
    const bool f = fii[k] != 0;
    if (d > 0.0) {
      if (k < nvr) { sR += d; cR += 1; i1++; } else { sP += d; cP += 1; i2++; }
    }
    if (d < 0.0 || !(d == d) || (d == 0.0 && !f) || (d > 0.0 && f)) atomicMin(&bad, k);
  }
  rs[0][tid] = sR; rs[1][tid] = sP; rs[2][tid] = cR; rs[3][tid] = cP;
  ri[0][tid] = i1; ri[1][tid] = i2;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (tid < o) {
      for (int q = 0; q < 4; q++) rs[q][tid] += rs[q][tid + o];
      ri[0][tid] += ri[0][tid + o];
      ri[1][tid] += ri[1][tid + o];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double avR = nvr > 0 ? rs[0][0] / rs[2][0] : 1.0;
    const double avP = rs[1][0] / rs[3][0];
    const double scal = sqrt(avR / avP);
    const double sav = (rs[0][0] + scal * scal * rs[1][0]) / (rs[2][0] + rs[3][0]);
    sc[0] = scal;
    sc[2] = sav;
    sc[1] = sav * sc[3];
    fl[0] = ri[0][0];
    fl[1] = ri[1][0];
    fl[2] = bad < nv ? bad : -1;
    int kind = 0;
    if (bad < nv) {
      const double d = S[(size_t)bad * nv + bad];
      kind = (d < 0.0 || !(d == d)) ? 1 : (d == 0.0 ? 2 : 3);
    }
    fl[3] = kind;
    fl[4] = 0;  // Cholesky info, set by k_potrf_diag
  }
}

// A = L = D S D + shift I, D = diag(1 (RBM), scal (PP)) (SR_pdposv.f90:56-73,
// 167-170); x = D g
__global__ void k_sr_system(int nv, int nvr, const double* __restrict__ sc, const double* __restrict__ S,
                            double* __restrict__ A, double* __restrict__ L, const double* __restrict__ g,
                            double* __restrict__ x) {
  const double scal = sc[0], shift = sc[1], boost = sc[12];
  const size_t n2 = (size_t)nv * nv;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / nv), j = (int)(e - (size_t)i * nv);
    double v = S[e];
    if (i >= nvr) v *= scal;
    if (j >= nvr) v *= scal;
    if (i == j) v += shift;
    A[e] = v;
    L[e] = i == j ? v + boost : v;
    if (e < (size_t)nv) x[e] = e >= (size_t)nvr ? g[e] * scal : g[e];
  }
}

// Substitutions with the Cholesky factor in steps of TS_NB = 4 CH_NB rows:
// every workgroup of a step solves the step's diagonal block itself (four
// CH_NB blocks through their inverses X_k and the L blocks between them, all
// from L2), so each step is one launch and the launches over the rest of the
// factor use many workgroups.  The diagonal entries read are the step's own,
// which no workgroup of the launch writes.  Every inner product runs over a
// fixed number of terms, fully unrolled and without per-term branches, so
// that its loads are independent of each other: X_k is stored with zeros
// above its diagonal and beyond the factor's end, the step's solution is
// zero beyond its end, and a term beyond the end reads the last row or
// column inside it and adds that times zero.
#define TS_NB (4 * CH_NB)

// Forward, step [k0, k0 + nb): y = L_kk^-1 b_k (written to out by workgroup
// 0), then every row below, b_i -= sum_j L(i, k0 + j) y_j, one row per wave
// and four rows in flight (lanes over j, 512-byte row segments).
__global__ __launch_bounds__(256) void k_trsv_fwd(int n, int k0, const double* __restrict__ L,
                                                  const double* __restrict__ dinv, double* b, double* out) {
  __shared__ double bb[TS_NB], yy[TS_NB];
  const int nb = min(TS_NB, n - k0), t = threadIdx.x, lane = t & 63;
  bb[t] = t < nb ? b[k0 + t] : 0.0;
  yy[t] = 0.0;
  __syncthreads();
  for (int s = 0; s < 4 && CH_NB * s < nb; s++) {
    const int o = CH_NB * s;
    if (t < CH_NB) {
      const double* D = dinv + (size_t)(k0 / CH_NB + s) * CH_NB * CH_NB + t * CH_NB;
      double y = 0.0;
#pragma unroll
      for (int j = 0; j < CH_NB; j++) y = fma(D[j], bb[o + j], y);
      yy[o + t] = y;
    }
    __syncthreads();
    if (t >= o + CH_NB && t < nb) {
      const double* row = L + (size_t)(k0 + t) * n + k0 + o;
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < CH_NB; j++) acc = fma(row[j], yy[o + j], acc);
      bb[t] -= acc;
    }
    __syncthreads();
  }
  if (blockIdx.x == 0 && t < nb) out[k0 + t] = yy[t];
  int cj[4];
#pragma unroll
  for (int c = 0; c < 4; c++) cj[c] = min(lane + CH_NB * c, nb - 1);
  const int nw = gridDim.x * 4;
  for (int i0 = k0 + nb + blockIdx.x * 4 + (t >> 6); i0 < n; i0 += 4 * nw) {
    double acc[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = min(i0 + r * nw, n - 1);  // rows past the end repeat the last row; not stored
      const double* row = L + (size_t)i * n + k0;
      acc[r] = 0.0;
#pragma unroll
      for (int c = 0; c < 4; c++) acc[r] = fma(row[cj[c]], yy[lane + CH_NB * c], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const double v = wave_sum(acc[r]);
      const int i = i0 + r * nw;
      if (lane == 0 && i < n) b[i] -= v;
    }
  }
}

// Back substitution L^T x = y, step [k0, k0 + nb): x = L_kk^-T y_k (written
// to out by workgroup 0), then the rows above, y_i -= sum_j L(k0 + j, i) x_j:
// workgroup g takes i = 64 g + lane, wave w the terms j in [64 w, 64 w + 64),
// added in a fixed order.
__global__ __launch_bounds__(256) void k_trsv_bwd(int n, int k0, const double* __restrict__ L,
                                                  const double* __restrict__ dinv, double* y, double* out) {
  __shared__ double yb[TS_NB], xk[TS_NB], red[4][64];
  const int nb = min(TS_NB, n - k0), t = threadIdx.x, lane = t & 63, w = t >> 6;
  yb[t] = t < nb ? y[k0 + t] : 0.0;
  xk[t] = 0.0;
  __syncthreads();
  for (int s = (nb + CH_NB - 1) / CH_NB - 1; s >= 0; s--) {
    const int o = CH_NB * s, ns = min(CH_NB, nb - o);
    if (t < CH_NB) {
      // X_k^T row t: the entries X_k(j, t), zero for j < t
      const double* D = dinv + (size_t)(k0 / CH_NB + s) * CH_NB * CH_NB + t;
      double x = 0.0;
#pragma unroll
      for (int j = 0; j < CH_NB; j++) x = fma(D[j * CH_NB], yb[o + j], x);
      xk[o + t] = x;
    }
    __syncthreads();
    if (t < o) {
      const double* col = L + (size_t)(k0 + o) * n + k0 + t;
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < CH_NB; j++) acc = fma(col[(size_t)min(j, ns - 1) * n], xk[o + j], acc);
      yb[t] -= acc;
    }
    __syncthreads();
  }
  if (blockIdx.x == 0 && t < nb) out[k0 + t] = xk[t];
  const int i = blockIdx.x * 64 + lane;
  double acc = 0.0;
  if (i < k0 && CH_NB * w < nb) {
    const double* col = L + (size_t)(k0 + CH_NB * w) * n + i;
    const int last = min(CH_NB, nb - CH_NB * w) - 1;
#pragma unroll
    for (int j = 0; j < CH_NB; j++) acc = fma(col[(size_t)min(j, last) * n], xk[CH_NB * w + j], acc);
  }
  red[w][lane] = acc;
  __syncthreads();
  if (w == 0 && i < k0) y[i] -= ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// r = D g - A x (one wave per row); with acc: x += 0 (plain residual)
__global__ void k_residual(int n, int nvr, const double* __restrict__ sc, const double* __restrict__ A,
                           const double* __restrict__ x, const double* __restrict__ g, double* __restrict__ r) {
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double t = 0;
  for (int j = lane; j < n; j += 64) t = fma(A[(size_t)row * n + j], x[j], t);
  t = wave_sum(t);
  if (lane == 0) r[row] = (row >= nvr ? g[row] * sc[0] : g[row]) - t;
}

__global__ void k_axpy1(int n, const double* __restrict__ r, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += r[i];
}

// lambda^2 = x.(A x) and |r|^2, |Dg|^2 with r = D g - A x; ax holds A x.
// One workgroup, fixed-order tree.
__global__ __launch_bounds__(1024) void k_sr_final(int n, int nvr, double* __restrict__ sc,
                                                   const double* __restrict__ x, const double* __restrict__ ax,
                                                   const double* __restrict__ g) {
  __shared__ double rs[3][1024];
  const int tid = threadIdx.x;
  const double scal = sc[0];
  double l2 = 0, r2 = 0, g2 = 0;
  for (int k = tid; k < n; k += 1024) {
    const double gk = k >= nvr ? g[k] * scal : g[k];
    l2 = fma(x[k], ax[k], l2);
    r2 = fma(gk - ax[k], gk - ax[k], r2);
    g2 = fma(gk, gk, g2);
  }
  rs[0][tid] = l2; rs[1][tid] = r2; rs[2][tid] = g2;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (tid < o)
      for (int q = 0; q < 3; q++) rs[q][tid] += rs[q][tid + o];
    __syncthreads();
  }
  if (tid == 0) { sc[4] = rs[0][0]; sc[5] = rs[1][0]; sc[6] = rs[2][0]; }
}

// ------------------------------------------------------------ launchers
void launch_scalar_partials(int ns, const double2* psi, const double* rw, const double2* eloc, double* part,
                            int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_scalar_partials, dim3(nblk), dim3(512), 0, st, ns, psi, rw, eloc, part);
}
void launch_gvec_partials(int ns, int nv, int cplx_rows, const double* orows, const double2* psi,
                          const double* rw, const double2* eloc, double* part, int nchunk, hipStream_t st) {
  hipLaunchKernelGGL(k_gvec_partials, dim3((nv + 255) / 256, nchunk), dim3(256), 0, st, ns, nv, cplx_rows,
                     orows, psi, rw, eloc, part);
  hipLaunchKernelGGL(k_sum_partials, dim3((2 * nv + 255) / 256), dim3(256), 0, st, nv, nchunk, part,
                     part + (size_t)nchunk * 2 * nv);
}
void launch_row_kappa(int ns, int cplx_rows, const double2* psi, const double* rw, double* kappa,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_row_kappa, dim3((ns + 255) / 256), dim3(256), 0, st, ns, cplx_rows, psi, rw, kappa);
}
void launch_gemv(int n, const double* A, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_gemv, dim3((n + 3) / 4), dim3(256), 0, st, n, A, x, y);
}
void launch_cholesky(int n, double* A, double* dinv, int* info, hipStream_t st, const CholAhead* ahead) {
  if (!ahead) {
    for (int k0 = 0; k0 < n; k0 += CH_NB) {
      hipLaunchKernelGGL(k_potrf_inv, dim3(1), dim3(64), 0, st, n, k0, A, dinv, info);
      const int t0 = k0 + CH_NB;
      if (t0 >= n) break;
      const int rows = n - t0;
      hipLaunchKernelGGL(k_trsm_panel, dim3((rows + 63) / 64), dim3(256), 0, st, n, k0, A, (const double*)dinv);
      const int nt = (rows + 63) / 64;
      hipLaunchKernelGGL(k_syrk_trailing, dim3(nt * (nt + 1) / 2), dim3(256), 0, st, n, k0, t0, nt, 0, A);
    }
    return;
  }
  // Look-ahead: after the panel of block k, the trailing update of block
  // column k + 1 alone; block k + 1 is then factored on the second stream
  // while the rest of the trailing update runs, and the next panel waits for
  // both.  Every launch touches what it did in the one-stream order.
  hipLaunchKernelGGL(k_potrf_inv, dim3(1), dim3(64), 0, st, n, 0, A, dinv, info);
  for (int k0 = 0; k0 + CH_NB < n; k0 += CH_NB) {
    const int t0 = k0 + CH_NB, rows = n - t0, nt = (rows + 63) / 64;
    hipLaunchKernelGGL(k_trsm_panel, dim3((rows + 63) / 64), dim3(256), 0, st, n, k0, A, (const double*)dinv);
    hipLaunchKernelGGL(k_syrk_trailing, dim3(nt), dim3(256), 0, st, n, k0, t0, nt, 1, A);
    hipEventRecord(ahead->col_done, st);
    hipStreamWaitEvent(ahead->st2, ahead->col_done, 0);
    hipLaunchKernelGGL(k_potrf_inv, dim3(1), dim3(64), 0, ahead->st2, n, t0, A, dinv, info);
    hipEventRecord(ahead->diag_done, ahead->st2);
    if (nt > 1)
      hipLaunchKernelGGL(k_syrk_trailing, dim3((nt - 1) * nt / 2), dim3(256), 0, st, n, k0, t0 + CH_NB, nt - 1, 0, A);
    hipStreamWaitEvent(st, ahead->diag_done, 0);
  }
}

void launch_chol_solve(int n, const double* L, const double* dinv, double* x, double* tmp, hipStream_t st) {
  // forward: b = x (updated below each step) -> y in tmp
  for (int k0 = 0; k0 < n; k0 += TS_NB) {
    const int rows = std::max(n - k0 - TS_NB, 1);
    const int grid = std::min(256, (rows + 31) / 32);
    hipLaunchKernelGGL(k_trsv_fwd, dim3(grid), dim3(256), 0, st, n, k0, L, dinv, x, tmp);
  }
  // backward: y = tmp (updated above each step) -> x
  const int nstep = (n + TS_NB - 1) / TS_NB;
  for (int blk = nstep - 1; blk >= 0; blk--) {
    const int k0 = blk * TS_NB;
    hipLaunchKernelGGL(k_trsv_bwd, dim3(std::max(1, (k0 + 63) / 64)), dim3(256), 0, st, n, k0, L, dinv, tmp, x);
  }
}

// ------------------------------------------- the FP64 operator of the O rows
// y = (D S D + shift I) x with S = inv_norm C^T C and C = O - kappa obar^T,
// the centred rows taken from the FP64 O rows themselves, so that a
// refinement against it converges to the FP64 system's solution whatever the
// precision of the assembled S.  sc[7] = inv_norm, sc[8] holds obar . D x.

__global__ __launch_bounds__(1024) void k_op_obar(int nv, int nvr, double* __restrict__ sc,
                                                  const double* __restrict__ obar, const double* __restrict__ x) {
  __shared__ double rs[1024];
  const int tid = threadIdx.x;
  const double scal = sc[0];
  double t = 0.0;
  for (int j = tid; j < nv; j += 1024) t = fma(obar[j], j >= nvr ? x[j] * scal : x[j], t);
  rs[tid] = t;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (tid < o) rs[tid] += rs[tid + o];
    __syncthreads();
  }
  if (tid == 0) sc[8] = rs[0];
}

// One pass over the O rows: workgroup c takes rows [c per, (c + 1) per) in
// batches of RB and forms t_k = O_k . D x - kappa_k (obar . D x) from the rows
// held in registers (thread tid holds columns tid + 1024 i, i < NPT), then
// accumulates part[c][j] = sum_k O_kj t_k and part[c][nv] = sum_k kappa_k t_k
// from the same registers, each in a fixed order.  With PF the next batch's
// rows are loaded while the current batch's dot products are reduced (where
// the registers allow it).
template <int NPT, int RB, bool PF>
__global__ __launch_bounds__(1024) void k_op_pass(int rows, int nv, int nvr, const double* __restrict__ sc,
                                                  const double* __restrict__ O, const double* __restrict__ kappa,
                                                  const double* __restrict__ x, double* __restrict__ part) {
  static_assert(RB == 1 || RB == 2 || RB == 4, "rows per batch: one wave sum, or a two- or four-way one");
  __shared__ double red[2][RB][16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double scal = sc[0], od = sc[8];
  double dv[NPT], acc[NPT];
#pragma unroll
  for (int i = 0; i < NPT; i++) {
    const int j = tid + 1024 * i;
    dv[i] = j < nv ? (j >= nvr ? x[j] * scal : x[j]) : 0.0;
    acc[i] = 0.0;
  }
  const int per = (rows + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * per, r1 = min(rows, r0 + per);
  double o[RB][NPT];
  double on[PF ? RB : 1][PF ? NPT : 1];
  auto load = [&](int k0, double (&dst)[RB][NPT]) {
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const bool live = k0 + r < r1;
      const double* Or = O + (size_t)(live ? k0 + r : 0) * nv;
#pragma unroll
      for (int i = 0; i < NPT; i++) {
        const int j = tid + 1024 * i;
        dst[r][i] = (live && j < nv) ? Or[j] : 0.0;
      }
    }
  };
  double kt = 0.0;
  int buf = 0;
  if constexpr (PF) {
    if (r0 < r1) load(r0, on);
  }
  for (int k0 = r0; k0 < r1; k0 += RB) {
    if constexpr (PF) {
#pragma unroll
      for (int r = 0; r < RB; r++)
#pragma unroll
        for (int i = 0; i < NPT; i++) o[r][i] = on[r][i];
      if (k0 + RB < r1) load(k0 + RB, on);
    } else {
      load(k0, o);
    }
    double d[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      d[r] = 0.0;
#pragma unroll
      for (int i = 0; i < NPT; i++) d[r] = fma(o[r][i], dv[i], d[r]);
    }
    if constexpr (RB == 4) {
      const double v = wave_sum4(d[0], d[1], d[2], d[3]);  // lanes 15, 47, 31, 63
      if (lane == 15) red[buf][0][wid] = v;
      if (lane == 47) red[buf][1][wid] = v;
      if (lane == 31) red[buf][2][wid] = v;
      if (lane == 63) red[buf][3][wid] = v;
    } else if constexpr (RB == 2) {
      const double v = wave_sum2(d[0], d[1]);  // lanes 31, 63
      if (lane == 31) red[buf][0][wid] = v;
      if (lane == 63) red[buf][1][wid] = v;
    } else {
      const double v = wave_sum_to63(d[0]);
      if (lane == 63) red[buf][0][wid] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RB; r++) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < 16; w++) s += red[buf][r][w];
      const bool live = k0 + r < r1;
      const double tk = live ? s - kappa[k0 + r] * od : 0.0;
#pragma unroll
      for (int i = 0; i < NPT; i++) acc[i] = fma(o[r][i], tk, acc[i]);
      if (tid == 0 && live) kt = fma(kappa[k0 + r], tk, kt);
    }
    // red[buf] is written again two batches on, after every thread has
    // passed the next batch's barrier (and so has read it)
    buf ^= 1;
  }
  double* pc = part + (size_t)blockIdx.x * (nv + 1);
#pragma unroll
  for (int i = 0; i < NPT; i++) {
    const int j = tid + 1024 * i;
    if (j < nv) pc[j] = acc[i];
  }
  if (tid == 0) pc[nv] = kt;
}

// y = D inv_norm (sum_c part_c - obar sum_c part_c[nv]) + shift x, and with
// r: r = D g - y
__global__ void k_op_finish(int nv, int nvr, const double* __restrict__ sc, const double* __restrict__ obar,
                            const double* __restrict__ part, const double* __restrict__ x,
                            const double* __restrict__ g, double* __restrict__ y, double* __restrict__ r) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  double w = 0.0, kt = 0.0;
  for (int c = 0; c < SR_ROW_CHUNKS; c++) {
    w += part[(size_t)c * (nv + 1) + j];
    kt += part[(size_t)c * (nv + 1) + nv];
  }
  const double scal = sc[0], shift = sc[1], inv_norm = sc[7];
  const double d = j >= nvr ? scal : 1.0;
  const double yj = d * (inv_norm * fma(-obar[j], kt, w)) + shift * x[j];
  y[j] = yj;
  if (r) r[j] = (j >= nvr ? g[j] * scal : g[j]) - yj;
}

template <int NPT, int RB, bool PF>
static void launch_op_pass(const SrPipeline& p, const double* v, hipStream_t st) {
  hipLaunchKernelGGL((k_op_pass<NPT, RB, PF>), dim3(SR_ROW_CHUNKS), dim3(1024), 0, st, p.rows, p.nv, p.nvr,
                     (const double*)p.sc, p.orows, p.kappa, v, p.part);
}

// y = A v for the operator of the O rows; with r: r = D g - y
static void launch_orow_op(const SrPipeline& p, const double* v, double* y, double* r, hipStream_t st) {
  const int n = p.nv;
  hipLaunchKernelGGL(k_op_obar, dim3(1), dim3(1024), 0, st, n, p.nvr, p.sc, p.obar, v);
  const int npt = (n + 1023) / 1024;
  if (npt <= 1) launch_op_pass<1, 4, true>(p, v, st);
  else if (npt <= 2) launch_op_pass<2, 4, true>(p, v, st);
  else if (npt <= 4) launch_op_pass<4, 2, true>(p, v, st);
  else if (npt <= 6) launch_op_pass<6, 2, false>(p, v, st);
  else if (npt <= 8) launch_op_pass<8, 2, false>(p, v, st);
  else if (npt <= 12) launch_op_pass<12, 1, false>(p, v, st);
  else if (npt <= 16) launch_op_pass<16, 1, false>(p, v, st);
  else launch_op_pass<20, 1, false>(p, v, st);
  hipLaunchKernelGGL(k_op_finish, dim3((n + 255) / 256), dim3(256), 0, st, n, p.nvr, (const double*)p.sc, p.obar,
                     (const double*)p.part, v, p.g, y, r);
}

// ---------------------------------- conjugate gradients on the O-row operator
// preconditioned by the Cholesky factor of the assembled matrix.  Scalar
// slots: sc[9], sc[11] alternate as r.z of successive iterations, sc[10]
// holds p.q.  A step whose p.q or r.z is not positive (converged, or lost
// to rounding) is skipped rather than divided by.
__global__ void k_copy(int n, const double* __restrict__ a, double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(1024) void k_dot(int n, const double* __restrict__ a, const double* __restrict__ b,
                                              double* __restrict__ out) {
  __shared__ double rs[1024];
  const int tid = threadIdx.x;
  double t = 0.0;
  for (int k = tid; k < n; k += 1024) t = fma(a[k], b[k], t);
  rs[tid] = t;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (tid < o) rs[tid] += rs[tid + o];
    __syncthreads();
  }
  if (tid == 0) *out = rs[0];
}

// x += alpha p, r -= alpha q, alpha = rz / pq
__global__ void k_cg_step(int n, const double* __restrict__ rz, const double* __restrict__ pq,
                          const double* __restrict__ pd, const double* __restrict__ q, double* __restrict__ x,
                          double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = (*pq > 0.0 && *rz > 0.0) ? *rz / *pq : 0.0;
  x[i] = fma(a, pd[i], x[i]);
  r[i] = fma(-a, q[i], r[i]);
}

// p = z + beta p, beta = rz_new / rz_old
__global__ void k_cg_dir(int n, const double* __restrict__ rz_new, const double* __restrict__ rz_old,
                         const double* __restrict__ z, double* __restrict__ pd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double b = (*rz_old > 0.0 && *rz_new > 0.0) ? *rz_new / *rz_old : 0.0;
  pd[i] = fma(b, pd[i], z[i]);
}

// ax = D g - r (A x carried by the recurrence)
__global__ void k_b_minus(int n, int nvr, const double* __restrict__ sc, const double* __restrict__ g,
                          const double* __restrict__ r, double* __restrict__ ax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ax[i] = (i >= nvr ? g[i] * sc[0] : g[i]) - r[i];
}

void launch_true_residual(const SrPipeline& p, hipStream_t st) {
  launch_orow_op(p, p.x, p.ax, p.r, st);  // r = b - A x
}

void launch_pcg_direction(const SrPipeline& p, hipStream_t st) {
  const int n = p.nv, nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, st, n, (const double*)p.r, p.z);
  launch_chol_solve(n, p.L, p.dinv, p.z, p.tmp, st);
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, st, n, (const double*)p.z, p.pd);
  hipLaunchKernelGGL(k_dot, dim3(1), dim3(1024), 0, st, n, (const double*)p.r, (const double*)p.z, p.sc + 9);
}

void launch_pcg_start(const SrPipeline& p, hipStream_t st) {
  launch_true_residual(p, st);
  launch_pcg_direction(p, st);
}

void launch_pcg_iters(const SrPipeline& p, int k0, int k1, hipStream_t st) {
  const int n = p.nv, nb = (n + 255) / 256;
  for (int k = k0; k < k1; k++) {
    double* rz_old = p.sc + ((k & 1) ? 11 : 9);
    double* rz_new = p.sc + ((k & 1) ? 9 : 11);
    launch_orow_op(p, p.pd, p.ax, nullptr, st);  // q = A p
    hipLaunchKernelGGL(k_dot, dim3(1), dim3(1024), 0, st, n, (const double*)p.pd, (const double*)p.ax, p.sc + 10);
    hipLaunchKernelGGL(k_cg_step, dim3(nb), dim3(256), 0, st, n, (const double*)rz_old, (const double*)(p.sc + 10),
                       (const double*)p.pd, (const double*)p.ax, p.x, p.r);
    hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, st, n, (const double*)p.r, p.z);
    launch_chol_solve(n, p.L, p.dinv, p.z, p.tmp, st);
    hipLaunchKernelGGL(k_dot, dim3(1), dim3(1024), 0, st, n, (const double*)p.r, (const double*)p.z, rz_new);
    hipLaunchKernelGGL(k_cg_dir, dim3(nb), dim3(256), 0, st, n, (const double*)rz_new, (const double*)rz_old,
                       (const double*)p.z, p.pd);
  }
}

void launch_pcg_final(const SrPipeline& p, hipStream_t st) {
  const int n = p.nv;
  hipLaunchKernelGGL(k_b_minus, dim3((n + 255) / 256), dim3(256), 0, st, n, p.nvr, (const double*)p.sc, p.g,
                     (const double*)p.r, p.ax);
  hipLaunchKernelGGL(k_sr_final, dim3(1), dim3(1024), 0, st, n, p.nvr, p.sc, p.x, p.ax, p.g);
}

size_t orow_op_scratch(int rows, int nv) { return (size_t)SR_ROW_CHUNKS * (nv + 1); }

// The diagonal of S in FP64, S_jj = inv_norm sum_k C_kj^2, from the per-chunk
// sums of launch_center_pack added in chunk order, written over the diagonal
// of an S assembled from FP32 or split-BF16 products: the PP scaling, the
// shift and the diagonal checks then see FP64 values, and the operator of the
// O rows and the assembled matrix share D and the shift.
__global__ void k_diag_finish(int nv, double inv_norm, const double* __restrict__ part, double* __restrict__ S) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  double a = 0.0;
  for (int c = 0; c < SR_ROW_CHUNKS; c++) a += part[(size_t)c * (nv + 1) + j];
  S[(size_t)j * nv + j] = a * inv_norm;
}

void launch_s_diag_finish(int nv, double inv_norm, const double* part, double* S, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_finish, dim3((nv + 255) / 256), dim3(256), 0, st, nv, inv_norm, part, S);
}

void launch_sr_pipeline(const SrPipeline& p, hipStream_t st) {
  const int n = p.nv;
  hipLaunchKernelGGL(k_sr_prepare, dim3(1), dim3(1024), 0, st, n, p.nvr, p.S, p.fii, p.sc, p.fl);
  hipLaunchKernelGGL(k_sr_system, dim3(1024), dim3(256), 0, st, n, p.nvr, (const double*)p.sc, p.S, p.A, p.L, p.g,
                     p.x);
  launch_cholesky(n, p.L, p.dinv, p.fl + 4, st, p.ahead.st2 ? &p.ahead : nullptr);
  launch_chol_solve(n, p.L, p.dinv, p.x, p.ax, st);
  if (p.orows) {
    // conjugate gradients on the FP64 operator of the O rows, preconditioned
    // by the factor of the assembled matrix, from its solution; lambda from
    // the recurrence's A x
    launch_pcg_start(p, st);
    launch_pcg_iters(p, 0, p.refine_iters, st);
    launch_pcg_final(p, st);
    return;
  }
  for (int it = 0; it < p.refine_iters; it++) {
    hipLaunchKernelGGL(k_residual, dim3((n + 3) / 4), dim3(256), 0, st, n, p.nvr, (const double*)p.sc, p.A, p.x,
                       p.g, p.r);
    launch_chol_solve(n, p.L, p.dinv, p.r, p.ax, st);
    hipLaunchKernelGGL(k_axpy1, dim3((n + 255) / 256), dim3(256), 0, st, n, p.r, p.x);
  }
  hipLaunchKernelGGL(k_gemv, dim3((n + 3) / 4), dim3(256), 0, st, n, p.A, p.x, p.ax);
  hipLaunchKernelGGL(k_sr_final, dim3(1), dim3(1024), 0, st, n, p.nvr, p.sc, p.x, p.ax, p.g);
}
