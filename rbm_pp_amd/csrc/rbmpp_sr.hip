// Reductions over the measured samples and the linear algebra of the SR step
// on gfx950: centred S build (SYRK over the O rows on the matrix cores, FP64
// centring, FP64 / FP32 / split-BF16x3 products), PP-block scaling and diagonal shift, blocked FP64 Cholesky
// (replaces ScaLAPACK pdposv, SR_pdposv.f90:232), triangular solves,
// residual / lambda products.
#include "rbmpp_sr.h"
#include <type_traits>

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

// ---------------------------------------------------------------- S build
// S = <O O^T>/norm - <O><O>^T (SR_pdposv.f90:11-15) as a SYRK on the matrix
// cores over the centred rows c = O_row - kappa_row * obar.  The centring is
// done in FP64 while a 32-row stage is staged into LDS; the products then run
// in the configured precision:
//   MODE 0: FP64 MFMA (v_mfma_f64_16x16x4_f64), FP64 accumulation;
//   MODE 1: FP32 MFMA (v_mfma_f32_16x16x4_f32), FP32 accumulation;
//   MODE 2: split BF16 x 3 (v_mfma_f32_16x16x32_bf16): c = hi + lo with
//           hi = bf16(c), lo = bf16(c - hi); c c' ~ hi hi' + hi lo' + lo hi',
//           FP32 accumulation.
// FP32 accumulators are added into FP64 ones every SYRK_FLUSH stages; row
// splits (grid.y) are summed in FP64 by k_syrk_finish.  grid.x enumerates the
// 128 x 128 upper-triangle tile pairs; the 4 waves own 64 x 64 quadrants of
// 16 x 16 MFMA blocks.
constexpr int ST = 128, SK = 32, SYRK_FLUSH = 32;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int MODE>
struct SyrkLds {
  // MODE 0/1: panels [k][column] (+pad); MODE 2: hi/lo panels [column][k]
  static constexpr int KP = SK + 8;
  typedef typename std::conditional<MODE == 0, double, float>::type T;
  union {
    struct { T a[SK][ST + 4]; T b[SK][ST + 4]; } kc;
    struct { __bf16 ahi[ST][KP], alo[ST][KP], bhi[ST][KP], blo[ST][KP]; } bf;
  };
};

template <int MODE>
__global__ __launch_bounds__(256) void k_syrk_mfma(int nrows, int nv, const double* __restrict__ orows,
                                                   const double* __restrict__ kappa,
                                                   const double* __restrict__ obar, int ntile, double* part) {
  __shared__ SyrkLds<MODE> L;
  int p = blockIdx.x, ti = 0;
  while (p >= ntile - ti) { p -= ntile - ti; ti++; }
  const int tj = ti + p;
  const bool diag = ti == tj;
  const int chunk = ((nrows + gridDim.y - 1) / gridDim.y + SK - 1) / SK * SK;
  const int r0 = blockIdx.y * chunk, r1 = min(nrows, r0 + chunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  // staging: column c of both panels, rows hk*16 .. hk*16+15 of the stage
  const int c = tid & (ST - 1), hk = tid >> 7;
  const int ci = ti * ST + c, cj = tj * ST + c;
  const double oa = ci < nv ? obar[ci] : 0.0, ob = cj < nv ? obar[cj] : 0.0;

  f64x4 acc64[4][4];
  f32x4 acc32[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      acc64[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
      acc32[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  int nstage = 0;
  // the stage's rows are loaded one stage ahead, so the global loads of
  // stage s+1 are in flight while the matrix cores work on stage s
  double va[16], vb[16];
  auto load_stage = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int row = k0 + hk * 16 + u;
      double x = 0.0, y = 0.0;
      if (row < r1) {
        const double kap = kappa[row];
        const double* orow = orows + (size_t)row * nv;
        if (ci < nv) x = orow[ci] - kap * oa;
        if (!diag && cj < nv) y = orow[cj] - kap * ob;
      }
      va[u] = x;
      vb[u] = y;
    }
  };
  if (r0 < r1) load_stage(r0);
  for (int k0 = r0; k0 < r1; k0 += SK) {
    // ---- stage rows k0..k0+SK-1 (centred in FP64, converted once)
    if constexpr (MODE == 2) {
#pragma unroll
      for (int g = 0; g < 2; g++) {
        bf16x8 h, l, hb, lb;
#pragma unroll
        for (int t = 0; t < 8; t++) {
          const double x = va[8 * g + t], y = vb[8 * g + t];
          const __bf16 xh = (__bf16)(float)x, yh = (__bf16)(float)y;
          h[t] = xh;
          l[t] = (__bf16)(float)(x - (double)(float)xh);
          hb[t] = yh;
          lb[t] = (__bf16)(float)(y - (double)(float)yh);
        }
        *(bf16x8*)&L.bf.ahi[c][hk * 16 + 8 * g] = h;
        *(bf16x8*)&L.bf.alo[c][hk * 16 + 8 * g] = l;
        if (!diag) {
          *(bf16x8*)&L.bf.bhi[c][hk * 16 + 8 * g] = hb;
          *(bf16x8*)&L.bf.blo[c][hk * 16 + 8 * g] = lb;
        }
      }
    } else {
      typedef typename SyrkLds<MODE>::T T;
#pragma unroll
      for (int u = 0; u < 16; u++) {
        L.kc.a[hk * 16 + u][c] = (T)va[u];
        if (!diag) L.kc.b[hk * 16 + u][c] = (T)vb[u];
      }
    }
    __syncthreads();
    if (k0 + SK < r1) load_stage(k0 + SK);
    // ---- products of this wave's 4 x 4 blocks
    const int ib = wi * 64 + (lane & 15), jb = wj * 64 + (lane & 15);
    if constexpr (MODE == 2) {
      const int kk = 8 * (lane >> 4);
      bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int x = 0; x < 4; x++) {
        ah[x] = *(const bf16x8*)&L.bf.ahi[ib + 16 * x][kk];
        al[x] = *(const bf16x8*)&L.bf.alo[ib + 16 * x][kk];
        bh[x] = diag ? *(const bf16x8*)&L.bf.ahi[jb + 16 * x][kk] : *(const bf16x8*)&L.bf.bhi[jb + 16 * x][kk];
        bl[x] = diag ? *(const bf16x8*)&L.bf.alo[jb + 16 * x][kk] : *(const bf16x8*)&L.bf.blo[jb + 16 * x][kk];
      }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
          acc32[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc32[x][y], 0, 0, 0);
          acc32[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc32[x][y], 0, 0, 0);
          acc32[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc32[x][y], 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int s = 0; s < SK / 4; s++) {
        const int kr = 4 * s + (lane >> 4);
        typename SyrkLds<MODE>::T av[4], bv[4];
#pragma unroll
        for (int x = 0; x < 4; x++) {
          av[x] = L.kc.a[kr][ib + 16 * x];
          bv[x] = diag ? L.kc.a[kr][jb + 16 * x] : L.kc.b[kr][jb + 16 * x];
        }
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
          for (int y = 0; y < 4; y++) {
            if constexpr (MODE == 0)
              acc64[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc64[x][y], 0, 0, 0);
            else
              acc32[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x], bv[y], acc32[x][y], 0, 0, 0);
          }
      }
    }
    if constexpr (MODE != 0) {
      if (++nstage == SYRK_FLUSH) {
        nstage = 0;
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
          for (int y = 0; y < 4; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              acc64[x][y][r] += (double)acc32[x][y][r];
              acc32[x][y][r] = 0.f;
            }
      }
    }
    __syncthreads();
  }
  if constexpr (MODE != 0) {
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc64[x][y][r] += (double)acc32[x][y][r];
  }
  // ---- FP64 partial tile of this split.  Output element (row, col) of a
  // 16 x 16 block: FP32 MFMA lane l, register r -> (4 (l/16) + r, l % 16);
  // FP64 MFMA -> (l/16 + 4 r, l % 16).
  const size_t tile = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  double* out = part + tile * ST * ST;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = wi * 64 + 16 * x + (MODE == 0 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r);
        const int j = wj * 64 + 16 * y + (lane & 15);
        out[i * ST + j] = acc64[x][y][r];
      }
}

// Sum the split partials of each tile pair, scale by 1/norm and write both
// triangles of S (row-major nv x nv).
__global__ void k_syrk_finish(int nv, int ntile, int npairs, int nsplit, const double* __restrict__ part,
                              double inv_norm, double* S) {
  int p = blockIdx.x, ti = 0;
  while (p >= ntile - ti) { p -= ntile - ti; ti++; }
  const int tj = ti + p;
  for (int e = threadIdx.x; e < ST * ST; e += blockDim.x) {
    const int a = e / ST, b = e - a * ST;
    const int i = ti * ST + a, j = tj * ST + b;
    if (i >= nv || j >= nv) continue;
    double t = 0;
    for (int z = 0; z < nsplit; z++) t += part[((size_t)z * npairs + blockIdx.x) * ST * ST + e];
    t *= inv_norm;
    S[(size_t)i * nv + j] = t;
    S[(size_t)j * nv + i] = t;
  }
}

// ------------------------------------------------------- SR system set-up
// Ssh = D S D + shift I with D = diag(1 (RBM), scal (PP)) (SR_pdposv.f90:56-73,
// 167-170); returns into A (row-major).
__global__ void k_scale_shift(int nv, int nv_rbm, double scal, double shift, const double* __restrict__ S,
                              double* A) {
  const size_t n2 = (size_t)nv * nv;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / nv), j = (int)(e - (size_t)i * nv);
    double v = S[e];
    if (i >= nv_rbm) v *= scal;
    if (j >= nv_rbm) v *= scal;
    if (i == j) v += shift;
    A[e] = v;
  }
}

__global__ void k_diag(int nv, const double* __restrict__ S, double* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) d[i] = S[(size_t)i * nv + i];
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
// Right-looking, lower triangle of a row-major SPD matrix, block NB=32.
#define CH_NB 32

// factor the diagonal block [k0, k0+nb)
__global__ __launch_bounds__(256) void k_potrf_diag(int n, int k0, double* A, int* info) {
  __shared__ double B[CH_NB][CH_NB + 1];
  const int nb = min(CH_NB, n - k0);
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e - i * nb;
    B[i][j] = A[(size_t)(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    if (B[j][j] <= 0.0 || !(B[j][j] == B[j][j])) {
      if (threadIdx.x == 0) *info = k0 + j + 1;
      return;
    }
    __syncthreads();
    const double d = sqrt(B[j][j]);
    __syncthreads();
    if (threadIdx.x == 0) B[j][j] = d;
    for (int i = j + 1 + threadIdx.x; i < nb; i += blockDim.x) B[i][j] /= d;
    __syncthreads();
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
      const int i = e / nb, l = e - i * nb;
      if (i > j && l > j && l <= i) B[i][l] -= B[i][j] * B[l][j];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e - i * nb;
    if (j <= i) A[(size_t)(k0 + i) * n + k0 + j] = B[i][j];
  }
}

// panel: rows i >= k0+nb, columns [k0, k0+nb): L_ik = A_ik L_kk^-T
__global__ __launch_bounds__(256) void k_trsm_panel(int n, int k0, double* A) {
  __shared__ double Lk[CH_NB][CH_NB + 1];
  const int nb = min(CH_NB, n - k0);
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e - i * nb;
    Lk[i][j] = A[(size_t)(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  const int i = k0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[CH_NB];
#pragma unroll
  for (int j = 0; j < CH_NB; j++) r[j] = (j < nb) ? A[(size_t)i * n + k0 + j] : 0.0;
#pragma unroll
  for (int j = 0; j < CH_NB; j++) {
    if (j < nb) {
      double v = r[j];
#pragma unroll
      for (int k = 0; k < CH_NB; k++)
        if (k < j) v -= r[k] * Lk[j][k];
      r[j] = v / Lk[j][j];
    }
  }
#pragma unroll
  for (int j = 0; j < CH_NB; j++)
    if (j < nb) A[(size_t)i * n + k0 + j] = r[j];
}

// trailing update of the lower triangle: A_ij -= sum_k L_ik L_jk, i,j >= t0
__global__ __launch_bounds__(256) void k_syrk_trailing(int n, int k0, int t0, int ntile, double* A) {
  constexpr int TS = 64;
  __shared__ double Li[TS][CH_NB + 1];
  __shared__ double Lj[TS][CH_NB + 1];
  int p = blockIdx.x, tj = 0;  // lower-triangle tile pairs ti >= tj
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
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int r = ty * 4 + a, c = tx * 4 + b;
      const int gi = t0 + ti * TS + r, gj = t0 + tj * TS + c;
      if (gi < n && gj < n && gj <= gi) {
        double t = 0;
#pragma unroll 8
        for (int k = 0; k < CH_NB; k++) t = fma(Li[r][k], Lj[c][k], t);
        A[(size_t)gi * n + gj] -= t;
      }
    }
}

// L y = b then L^T x = y, single workgroup, in place in x
__global__ __launch_bounds__(1024) void k_chol_solve(int n, const double* __restrict__ L, double* x) {
  __shared__ double red[16];
  __shared__ double xi;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int i = 0; i < n; i++) {
    double t = 0;
    for (int k = threadIdx.x; k < i; k += blockDim.x) t = fma(L[(size_t)i * n + k], x[k], t);
    t = wave_sum(t);
    if (lane == 0) red[wid] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0;
      for (int w = 0; w < nw; w++) s += red[w];
      x[i] = (x[i] - s) / L[(size_t)i * n + i];
    }
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; i--) {
    double t = 0;
    for (int k = i + 1 + threadIdx.x; k < n; k += blockDim.x) t = fma(L[(size_t)k * n + i], x[k], t);
    t = wave_sum(t);
    if (lane == 0) red[wid] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0;
      for (int w = 0; w < nw; w++) s += red[w];
      x[i] = (x[i] - s) / L[(size_t)i * n + i];
    }
    __syncthreads();
  }
  (void)xi;
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
int syrk_tiles(int nv) { return (nv + ST - 1) / ST; }
int syrk_tile_size() { return ST; }
void launch_syrk(int precision, int nrows, int nv, const double* orows, const double* kappa, const double* obar,
                 int nsplit, double* part, double inv_norm, double* S, hipStream_t st) {
  const int nt = syrk_tiles(nv), np = nt * (nt + 1) / 2;
  const dim3 grid(np, nsplit);
  if (precision == 2)
    hipLaunchKernelGGL(k_syrk_mfma<2>, grid, dim3(256), 0, st, nrows, nv, orows, kappa, obar, nt, part);
  else if (precision == 1)
    hipLaunchKernelGGL(k_syrk_mfma<1>, grid, dim3(256), 0, st, nrows, nv, orows, kappa, obar, nt, part);
  else
    hipLaunchKernelGGL(k_syrk_mfma<0>, grid, dim3(256), 0, st, nrows, nv, orows, kappa, obar, nt, part);
  hipLaunchKernelGGL(k_syrk_finish, dim3(np), dim3(256), 0, st, nv, nt, np, nsplit, part, inv_norm, S);
}
void launch_scale_shift(int nv, int nv_rbm, double scal, double shift, const double* S, double* A,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_scale_shift, dim3(1024), dim3(256), 0, st, nv, nv_rbm, scal, shift, S, A);
}
void launch_diag(int nv, const double* S, double* d, hipStream_t st) {
  hipLaunchKernelGGL(k_diag, dim3((nv + 255) / 256), dim3(256), 0, st, nv, S, d);
}
void launch_gemv(int n, const double* A, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_gemv, dim3((n + 3) / 4), dim3(256), 0, st, n, A, x, y);
}
void launch_cholesky(int n, double* A, int* info, hipStream_t st) {
  for (int k0 = 0; k0 < n; k0 += CH_NB) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(256), 0, st, n, k0, A, info);
    const int t0 = k0 + CH_NB;
    if (t0 >= n) break;
    const int rows = n - t0;
    hipLaunchKernelGGL(k_trsm_panel, dim3((rows + 255) / 256), dim3(256), 0, st, n, k0, A);
    const int nt = (rows + 63) / 64;
    hipLaunchKernelGGL(k_syrk_trailing, dim3(nt * (nt + 1) / 2), dim3(256), 0, st, n, k0, t0, nt, A);
  }
}
void launch_chol_solve(int n, const double* L, double* x, hipStream_t st) {
  hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(1024), 0, st, n, L, x);
}
