// Kernel argument blocks and launch helpers shared by the host C-ABI and the
// kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbmpp_device.h"

#define RBMPP_CMAX 16   // projection copies per walker held in LDS (2 x 8 operations)
#define RBMPP_MRMAX 512 // real RBM angles per walker
#define RBMPP_MCMAX 256 // complex RBM angles per walker
#define RBMPP_NLDS 104  // largest N whose N x N matrices are staged in LDS

// walker status codes (mirrors RBMPP_E_* of rbmpp.h)
#define WS_OK 0
#define WS_SINGULAR (-10)
#define WS_CANDIDATE (-30)
#define WS_INIT (-21)
#define WS_THETA (-40)
#define WS_SIGN (-42)

struct WalkerState {
  signed char* x;          // [W][N]
  int* eidx;               // [W][N]  Ele_Idx (0-based sites)
  int* ecfg;               // [W][N]  Ele_Cfg
  double* thr;             // [W][Mr]
  double2* thc;            // [W][Mc]
  double* pf;              // [W][C]
  int* sgn;                // [W][P]
  double2* psi;            // [W]
  double* minv;            // [W][C][N*N/2] InvM of copy q in folded skew storage (fold_slot)
  double* scratch;         // [W][N][N]     recompute matrix when N > RBMPP_NLDS
  int* nacc_tmp;           // [W] accepts since the last recompute
  int* nfastup;            // [W] adaptive recompute period
  unsigned long long* ctr; // [W] Philox counter
  int* status;             // [W]
  long long* naccept;      // [W]
  int* nrecompute;         // [W]
};

struct SweepArgs {
  const DevModel* m;  // device copy of the model descriptor
  WalkerState ws;
  int nwalk, nsamp, nwarm, nupdate, fixed_k, first;
  double shift, p_restart, p_global;
  // > 0: recompute Pf and X^-1 right after an accepted move whose Pf ratio
  // of any copy lies outside (1/cond_guard, cond_guard) (0: never, as the
  // reference)
  double cond_guard;
  uint32_t key0, key1;
  signed char* xs;  // [W*nsamp][N] stored samples
  double2* psis;    // [W*nsamp]
  // measurement fused into the sampler (at every stored sweep): reweighted
  // local energy, amplitude, reweight factor and O row(s) per sample
  int measure;
  double2 *m_psi, *m_eloc;
  double *m_rw, *m_orows;
};

// per-configuration evaluation (eval / local energy / O_k / measurement)
struct LocalArgs {
  const DevModel* m;  // device copy of the model descriptor
  int nconf;
  const signed char* xs;  // [nconf][N]
  int mode;               // bit0: eloc, bit1: O_k, bit2: write invm, bit3: custom exchanges
  double shift;           // reweighting offset; <0: no reweighting (raw d psi, sum H psi)
  const int* ex_i1;       // [nconf] custom exchange (mode bit3)
  const int* ex_i2;
  double* scratch;        // [grid][N][N] recompute matrix when N > RBMPP_NLDS
  // outputs
  double2* psi;           // [nconf]
  double* pf;             // [nconf][C]
  double2* psi_q;         // [nconf][C] (optional)
  double* invm;           // [nconf][C][N][N] reference layout InvM(i,j) at j*N+i (optional)
  double2* eloc;          // [nconf]
  double* orows;          // O rows: real models [nconf][nv]; complex [2*nconf][nv] (re, im rows)
  double* rw;             // [nconf] reweight factor 1/(|psi|+shift)
  double2* ex_psi;        // [nconf] amplitude after the custom exchange
  double* ex_pf;          // [nconf][C]
  int* status;            // [nconf]
  // correlation run (launch_meas_corr): per-workgroup partial sums
  // [grid][corr_len(N)], laid out by corr_len below
  double* corr_part;
};

// Per-workgroup slice of the correlation sums, nchi = N(N+1)/2 pairs i1 <= i2
// numbered i0 = i2(i2+1)/2 + i1: [0,nchi) Chi_sp(1,i0), [nchi,2nchi)
// Chi_sp(2,i0), [(2+t)nchi, (3+t)nchi) Chi_vbs(i0,t+1) for t < 4, then
// wf_sgn (re, im), then 2 (N/2)^2 doubles of scratch for the pair amplitudes.
__host__ __device__ inline int corr_nchi(int N) { return N * (N + 1) / 2; }
__host__ __device__ inline int corr_sums(int N) { return 6 * corr_nchi(N) + 2; }
__host__ __device__ inline int corr_len(int N) { return corr_sums(N) + 2 * (N / 2) * (N / 2); }

void launch_slater(const DevModel& m, const double* f, double* se, int* bad, hipStream_t s);
int launch_sweep(const DevModel& hm, const SweepArgs& a, hipStream_t s);

// one entry of a multi-model sampler launch: the model's sweep arguments, the
// first workgroup of its walkers and the kernel specialisation it needs
enum { SWEEP_BLK64 = 0, SWEEP_BLK128 = 1, SWEEP_SKEW256 = 2, SWEEP_BLK256 = 3 };
struct SweepBatchEntry {
  SweepArgs a;
  int first_block;
  int kind;
};
// the specialisation a model needs in the multi-model launch, or -1 (sizes
// or projections the batch kernel does not cover)
int sweep_kind(const DevModel& hm);
void launch_sweep_batch(const SweepBatchEntry* d_ent, int nent, int nblocks, hipStream_t s);
int launch_local(const DevModel& hm, const LocalArgs& a, int grid, hipStream_t s);
// the N <= RBMPP_NLDS estimator (register Gauss-Jordan, 1024 threads)
void launch_meas(const LocalArgs& a, int grid, hipStream_t s);
// block form (DevModel::block, C == RBMPP_CMAX, N <= RBMPP_NLDS): a.scratch
// holds grid x C x (N/2)^2 doubles for the F^-1 of every copy
void launch_meas_blk(const LocalArgs& a, int grid, hipStream_t s);
// block form for RBMPP_NLDS < N <= 256 (C == RBMPP_CMAX): a.scratch holds
// grid x meas_blk_big_scratch(N) doubles (at most grid x N x N)
void launch_meas_blk_big(const LocalArgs& a, int grid, hipStream_t s);
size_t meas_blk_big_scratch(int N);
// the same estimator with the correlation measurement (N <= RBMPP_NLDS)
void launch_meas_corr(const LocalArgs& a, int grid, hipStream_t s);
// the block-form estimator with the correlation measurement; a.scratch as for
// launch_meas_blk, a.corr_part [grid][corr_len(N)]
void launch_meas_blk_corr(const LocalArgs& a, int grid, hipStream_t s);
// out[k] = sum_c part[c * stride + k] for k < len, c < nchunk (fixed order)
void launch_sum_strided(int len, int nchunk, int stride, const double* part, double* out, hipStream_t s);
