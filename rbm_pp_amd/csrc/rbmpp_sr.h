// Launchers of the reduction, S-build and Cholesky kernels (rbmpp_sr.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "rbmpp_device.h"

void launch_scalar_partials(int ns, const double2* psi, const double* rw, const double2* eloc, double* part,
                            int nblk, hipStream_t st);
// partials [nchunk][2][nv] followed by the sums {g[nv], ovec[nv]} at part + nchunk*2*nv
void launch_gvec_partials(int ns, int nv, int cplx_rows, const double* orows, const double2* psi,
                          const double* rw, const double2* eloc, double* part, int nchunk, hipStream_t st);
void launch_row_kappa(int ns, int cplx_rows, const double2* psi, const double* rw, double* kappa,
                      hipStream_t st);
int syrk_tiles(int nv);
int syrk_tile_size();
void launch_syrk(int precision, int nrows, int nv, const double* orows, const double* kappa, const double* obar,
                 int nsplit, double* part, double inv_norm, double* S, hipStream_t st);
void launch_scale_shift(int nv, int nv_rbm, double scal, double shift, const double* S, double* A,
                        hipStream_t st);
void launch_diag(int nv, const double* S, double* d, hipStream_t st);
void launch_gemv(int n, const double* A, const double* x, double* y, hipStream_t st);
void launch_cholesky(int n, double* A, int* info, hipStream_t st);
void launch_chol_solve(int n, const double* L, double* x, hipStream_t st);
