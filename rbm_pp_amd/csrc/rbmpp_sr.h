// Launchers of the reduction, S-build and Cholesky kernels (rbmpp_sr.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rbmpp_device.h"

void launch_scalar_partials(int ns, const double2* psi, const double* rw, const double2* eloc, double* part,
                            int nblk, hipStream_t st);
// partials [nchunk][2][nv] followed by the sums {g[nv], ovec[nv]} at part + nchunk*2*nv
void launch_gvec_partials(int ns, int nv, int cplx_rows, const double* orows, const double2* psi,
                          const double* rw, const double2* eloc, double* part, int nchunk, hipStream_t st);
void launch_row_kappa(int ns, int cplx_rows, const double2* psi, const double* rw, double* kappa,
                      hipStream_t st);
// S build from the centred, packed O rows (rbmpp_syrk.hip)
struct SyrkPlan {
  int mode;          // RBMPP_S_FP64 / _FP32 / _BF16X3
  int kb;            // rows per K-block (16 FP64, 32 FP32 / BF16)
  int tile;          // output tile edge: 128, or 256 (split BF16 x 3 at large nv)
  int nvp;           // nv rounded up to the tile
  int ntile, npair;  // column tiles, upper-triangle tile pairs
  int nkb;           // K-blocks
  int nsplit, kb_per_split;
  size_t packed_bytes;   // nkb * nvp * 128
  size_t part_doubles;   // nsplit * npair * tile * tile
};
// tile <= 0: chosen from precision and nv; nsplit <= 0: about two resident
// workgroups per CU
SyrkPlan syrk_plan(int precision, int nrows, int nv, int cus, int tile = 0, int nsplit = 0);
// row chunks of the passes over the O rows whose column sums are kept per
// chunk and added in a fixed order: [SR_ROW_CHUNKS][nv + 1] partials
constexpr int SR_ROW_CHUNKS = 256;
// centre and pack the O rows for the products; for FP32 and split BF16 x 3
// also the FP64 column sums of squares per chunk into dpart
// ([SR_ROW_CHUNKS][nv + 1], finished by launch_s_diag_finish)
void launch_center_pack(const SyrkPlan& pl, int nrows, int nv, const double* orows, const double* kappa,
                        const double* obar, unsigned char* packed, double* dpart, hipStream_t st);
void launch_syrk_packed(const SyrkPlan& pl, int nv, const unsigned char* packed, double* part, double inv_norm,
                        double* S, hipStream_t st);
void launch_gemv(int n, const double* A, const double* x, double* y, hipStream_t st);
// A second stream (default priority) and two events for the look-ahead of
// launch_cholesky; the sequence joins it back into the calling stream, so it
// can be captured into a graph from that stream.
struct CholAhead {
  hipStream_t st2 = nullptr;
  hipEvent_t col_done = nullptr, diag_done = nullptr;
};
// blocked Cholesky of the lower triangle in place (64-wide blocks); dinv gets
// the inverses of the 64 x 64 diagonal blocks of the factor.  With ahead,
// each diagonal block is factored on ahead->st2 under the previous block's
// trailing update.
void launch_cholesky(int n, double* A, double* dinv, int* info, hipStream_t st, const CholAhead* ahead = nullptr);
// L L^T x = b in place in x (blocked substitutions through the diagonal-block
// inverses; tmp is an n-vector of scratch)
void launch_chol_solve(int n, const double* L, const double* dinv, double* x, double* tmp, hipStream_t st);
// doubles of dinv for an n x n factor
inline size_t chol_dinv_size(int n) { return (size_t)((n + 63) / 64) * 64 * 64; }

// The SR system on the device (scalars, D S D + shift I, Cholesky, solve,
// refinement, lambda and residual), every launch on one stream with no host
// synchronisation, so that the sequence can be captured in a graph.
struct SrPipeline {
  int nv, nvr, refine_iters;
  const double* S;           // centred S, row-major nv x nv
  const unsigned char* fii;  // [nv] 1 for the f_ii parameters
  const double* g;           // [nv] g - 2 E <O> (unscaled)
  double *A, *L;             // nv x nv: D S D + shift I and its factor
  double *x, *r, *ax;        // [nv]: solution, residual / correction, A x (also solve scratch)
  double* dinv;              // inverses of the factor's diagonal blocks (chol_dinv_size)
  double* sc;                // [16] scalars: k_sr_prepare's, sc[7] = 1/norm, sc[8..11] scratch of the
                             // O-row operator and the conjugate gradients, sc[12] extra diagonal of L
  int* fl;                   // [8] flags and Cholesky info
  CholAhead ahead;           // st2 set: the Cholesky's look-ahead stream and events
  // orows != nullptr: the refinement residuals and lambda use the FP64
  // operator of the O rows, D (C^T C / norm) D + shift I with
  // C = O - kappa obar^T, instead of the assembled A (for S assembled from
  // FP32 or split-BF16 products)
  const double* orows = nullptr;  // [rows][nv]
  const double* kappa = nullptr;  // [rows]
  const double* obar = nullptr;   // [nv]
  int rows = 0;
  double* part = nullptr;         // orow_op_scratch(rows, nv) doubles: per-chunk column sums
  double *z = nullptr, *pd = nullptr, *tmp = nullptr;  // [nv] conjugate-gradient vectors, solve scratch
};
// With orows set, refine_iters counts conjugate-gradient iterations on the
// operator of the O rows, preconditioned by the assembled matrix's factor
// (whose diagonal gets sc[12] added: a boost when the assembled matrix is
// not positive definite); launch_pcg_iters and launch_pcg_final continue
// such a solve from iteration k0.
void launch_sr_pipeline(const SrPipeline& p, hipStream_t st);
void launch_pcg_iters(const SrPipeline& p, int k0, int k1, hipStream_t st);
// (re)start: r = D g - A x with the operator of the O rows, z = M^-1 r, p = z;
// launch_pcg_iters continues from k0 = 0.  launch_pcg_start is the two
// halves in order: launch_true_residual (r and A x = D g - r, one pass over
// the O rows) and launch_pcg_direction (z, p and r.z from that r).
void launch_pcg_start(const SrPipeline& p, hipStream_t st);
void launch_true_residual(const SrPipeline& p, hipStream_t st);
void launch_pcg_direction(const SrPipeline& p, hipStream_t st);
void launch_pcg_final(const SrPipeline& p, hipStream_t st);
// doubles of scratch of the O-row operator
size_t orow_op_scratch(int rows, int nv);
// S_jj = inv_norm sum_c part[c][j] over the diagonal of S (the FP64 column
// sums of squares of launch_center_pack, added in chunk order)
void launch_s_diag_finish(int nv, double inv_norm, const double* part, double* S, hipStream_t st);
