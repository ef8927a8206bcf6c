// Reductions over the measured samples and the linear algebra of the SR step
// on gfx950: centred S build (SYRK over the O rows, FP64 centring, FP64 or
// FP32 products), PP-block scaling and diagonal shift, blocked FP64 Cholesky
// (replaces ScaLAPACK pdposv, SR_pdposv.f90:232), triangular solves,
// residual / lambda products.
#include "rbmpp_sr.h"

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
// Upper-triangle 64x64 tiles of sum_rows c c^T, c = O_row - kappa_row * obar
// (centring in FP64 before any rounding to the product precision:
// S = <O O^T>/norm - <O><O>^T, SR_pdposv.f90:11-15).  grid.x enumerates the
// tile pairs ti <= tj, grid.y splits the rows; each split writes an FP64
// partial tile.  T = double: FP64 FMA; T = float: FP32 products, FP32
// accumulation inside a split, FP64 across splits.
template <typename T>
__global__ __launch_bounds__(256) void k_syrk_centred(int nrows, int nv, const double* __restrict__ orows,
                                                      const double* __restrict__ kappa,
                                                      const double* __restrict__ obar, int ntile,
                                                      double* part) {
  constexpr int TS = 64, KC = 32;
  __shared__ T As[KC][TS + 1];
  __shared__ T Bs[KC][TS + 1];
  // decode upper-triangle tile pair
  int p = blockIdx.x, ti = 0;
  while (p >= ntile - ti) { p -= ntile - ti; ti++; }
  const int tj = ti + p;
  const int chunk = (nrows + gridDim.y - 1) / gridDim.y;
  const int r0 = blockIdx.y * chunk, r1 = min(nrows, r0 + chunk);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = (T)0;
  for (int k0 = r0; k0 < r1; k0 += KC) {
    for (int e = threadIdx.x; e < KC * TS; e += 256) {
      const int kk = e / TS, c = e - kk * TS;
      const int row = k0 + kk;
      const int ci = ti * TS + c, cj = tj * TS + c;
      double va = 0.0, vb = 0.0;
      if (row < r1) {
        const double kap = kappa[row];
        if (ci < nv) va = orows[(size_t)row * nv + ci] - kap * obar[ci];
        if (cj < nv) vb = orows[(size_t)row * nv + cj] - kap * obar[cj];
      }
      As[kk][c] = (T)va;
      Bs[kk][c] = (T)vb;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < KC; kk++) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; a++) av[a] = As[kk][ty * 4 + a];
#pragma unroll
      for (int b = 0; b < 4; b++) bv[b] = Bs[kk][tx * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
  const size_t tile = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  double* out = part + tile * TS * TS;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) out[(ty * 4 + a) * TS + tx * 4 + b] = (double)acc[a][b];
}

// Sum the split partials of each tile pair, scale by 1/norm and write both
// triangles of S (row-major nv x nv).
__global__ void k_syrk_finish(int nv, int ntile, int npairs, int nsplit, const double* __restrict__ part,
                              double inv_norm, double* S) {
  constexpr int TS = 64;
  int p = blockIdx.x, ti = 0;
  while (p >= ntile - ti) { p -= ntile - ti; ti++; }
  const int tj = ti + p;
  for (int e = threadIdx.x; e < TS * TS; e += blockDim.x) {
    const int a = e / TS, b = e - a * TS;
    const int i = ti * TS + a, j = tj * TS + b;
    if (i >= nv || j >= nv) continue;
    double t = 0;
    for (int z = 0; z < nsplit; z++) t += part[((size_t)z * npairs + blockIdx.x) * TS * TS + e];
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
int syrk_tiles(int nv) { return (nv + 63) / 64; }
void launch_syrk(int precision, int nrows, int nv, const double* orows, const double* kappa, const double* obar,
                 int nsplit, double* part, double inv_norm, double* S, hipStream_t st) {
  const int nt = syrk_tiles(nv), np = nt * (nt + 1) / 2;
  if (precision == 1)
    hipLaunchKernelGGL(k_syrk_centred<float>, dim3(np, nsplit), dim3(256), 0, st, nrows, nv, orows, kappa,
                       obar, nt, part);
  else
    hipLaunchKernelGGL(k_syrk_centred<double>, dim3(np, nsplit), dim3(256), 0, st, nrows, nv, orows, kappa,
                       obar, nt, part);
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
