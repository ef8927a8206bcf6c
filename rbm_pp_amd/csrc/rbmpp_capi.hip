// Host side of the C ABI (include/rbmpp.h): device state of one model, the
// make_sample / measure / SR-step phases, and the parity entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <climits>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rbmpp.h"
#include "rbmpp_kernels.h"
#include "rbmpp_sr.h"

namespace {

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct rbmpp_handle {
  rbmpp_config cfg;
  int L, N, P, C, ac, ar, nsl, nv, nv_rbm, mc, mr, npair, W, nspw;
  int real_only;
  int device = 0;  // HIP device current at rbmpp_create
  hipStream_t st = nullptr;
  CholAhead ahead;  // the SR Cholesky's look-ahead stream and events (joined back into st)
  std::string err;
  std::vector<int> wmap, invmap, orb;
  std::vector<unsigned char> subB;
  std::vector<double> chi, gamma;
  double sp_cs[2], sp_cc[2], sp_ss[2], sp_w[2];
  int projection = 0;  // RBMPP_PROJ_*
  std::vector<void*> allocs;
  // tables and parameters
  int *d_wmap = nullptr, *d_invmap = nullptr, *d_orb = nullptr;
  int* d_pboff = nullptr;        // DevModel::pbin_off
  unsigned* d_pbpairs = nullptr; // DevModel::pbin_pairs
  unsigned char* d_subB = nullptr;
  double2* d_chi = nullptr;
  double* d_se = nullptr;
  double *d_Wr = nullptr, *d_br = nullptr, *d_f = nullptr;
  double2 *d_Wc = nullptr, *d_bc = nullptr;
  // walkers and samples
  WalkerState ws;
  signed char* d_xs = nullptr;
  double2* d_psis = nullptr;
  int64_t n_stored = 0;
  int64_t n_meas = 0;  // configurations in the measurement buffers (rows of O / 1 or 2)
  int walkers_ready = 0;
  int minv_block = 0;  // ws.minv holds F^-1 per copy (block form) instead of folded X^-1
  int fused_valid = 0;     // measurement buffers hold the fused results of the stored samples
  double fused_shift = 0;  // reweighting offset they were made with
  // measurement
  size_t meas_cap = 0;  // configurations the measurement buffers hold
  double2 *d_mpsi = nullptr, *d_eloc = nullptr;
  double *d_rw = nullptr, *d_orows = nullptr, *d_kappa = nullptr, *d_part = nullptr;
  double* d_scr = nullptr;  // recompute scratch of the local kernel (N > RBMPP_NLDS)
  // correlation run: per-workgroup partial sums and their total
  double *d_cpart = nullptr, *d_csum = nullptr;
  size_t cpart_cap = 0;
  double* d_bscr = nullptr;  // F^-1 of every copy per workgroup of the block-form estimator
  int* d_status = nullptr;
  int local_grid = 0;
  size_t part_cap = 0;
  // SR
  double *d_S = nullptr, *d_A = nullptr, *d_L = nullptr, *d_vec = nullptr;
  int* d_info = nullptr;
  double* d_sc = nullptr;  // SR scalars (k_sr_prepare)
  double* d_dinv = nullptr;  // inverses of the Cholesky factor's diagonal blocks
  int* d_fl = nullptr;     // SR flags
  unsigned char* d_fii = nullptr;  // [nv] 1 for the f_ii parameters
  hipGraph_t sr_graph = nullptr;   // captured SR pipeline
  hipGraphExec_t sr_exec = nullptr;
  int sr_graph_iters = -1;         // refine_iters it was captured with
  // and the O-row operator it refines against (null: the assembled matrix)
  const double *sr_graph_orows = nullptr, *sr_graph_kappa = nullptr;
  int sr_graph_rows = 0;
  double* sr_graph_part = nullptr;
  double* d_op = nullptr;  // scratch of the O-row operator
  int sr_pcg_iters = 0;    // conjugate-gradient iterations of the last SR step (0: FP64 S)
  double sr_boost = 0;     // extra diagonal its preconditioner needed
  size_t op_cap = 0;
  unsigned char* d_packed = nullptr;  // centred O rows packed for the S build
  size_t packed_cap = 0;
  int n_cu = 256;                     // compute units of the device
  // choices of rbmpp_tune (0: built-in defaults)
  int tune_tile[3] = {0, 0, 0}, tune_nsplit[3] = {0, 0, 0};  // per S precision
  int tune_meas_grid = 0;
  DevModel* d_model = nullptr;  // device copy of devmodel(h)
  int measured = 0;
  int s_valid = 0;
  double m_norm = 0, m_etot = 0, m_etot2 = 0, m_wfav = 0;
  std::vector<double> m_g, m_o;
  double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

// Device staging of rbmpp_local_noalloc: one fixed array per GPU, held until
// the process ends.  In host code its name is only a stand-in; the device
// address of each GPU's copy comes from hipGetSymbolAddress (noalloc_setup,
// called by rbmpp_create with the handle's device current) and is the only
// form in which host code uses the array.
static constexpr size_t NOALLOC_BYTES = (size_t)64 << 20;
static constexpr size_t NOALLOC_WG_BYTES = (size_t)32 << 20;  // per-workgroup estimator scratch
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static unsigned char* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_noalloc_mu;

static int fail(rbmpp_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, x)                                                          \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess)                                                     \
      return fail(h, RBMPP_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int dalloc(rbmpp_handle* h, T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(h, RBMPP_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  h->allocs.push_back((void*)*p);
  return 0;
}

// Records the device address of this GPU's staging array (once per device).
static int noalloc_setup(rbmpp_handle* h) {
  if (h->device < 0 || h->device >= NOALLOC_MAX_DEVICES) return fail(h, RBMPP_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[h->device]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(h, RBMPP_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_noalloc_ptr[h->device] = (unsigned char*)ptr;
  return 0;
}

static DevModel devmodel(const rbmpp_handle* h) {
  DevModel m;
  m.L = h->L; m.N = h->N; m.P = h->P; m.C = h->C; m.ac = h->ac; m.ar = h->ar;
  m.npair = h->npair; m.nsl = h->nsl; m.nv = h->nv; m.nv_rbm = h->nv_rbm; m.mc = h->mc; m.mr = h->mr;
  m.real_only = h->real_only;
  // block form: the sampler keeps F^-1 per copy (N <= 104 with every
  // sampler variant; up to N = 256 without the fused measurement, which
  // exists in skew form only there)
  m.block = (h->sp_cs[0] == 0.0 && h->sp_cs[1] == 0.0 && h->C == RBMPP_CMAX &&
             (h->N <= RBMPP_NLDS || (h->N <= 256 && !h->cfg.fused_measure))) ? 1 : 0;
  m.J1 = h->cfg.J1; m.J2 = h->cfg.J2;
  for (int s = 0; s < 2; s++) {
    m.sp_cs[s] = h->sp_cs[s]; m.sp_cc[s] = h->sp_cc[s]; m.sp_ss[s] = h->sp_ss[s]; m.sp_w[s] = h->sp_w[s];
  }
  m.wmap = h->d_wmap; m.invmap = h->d_invmap; m.orb = h->d_orb; m.subB = h->d_subB; m.chi = h->d_chi;
  m.se = h->d_se; m.f = h->d_f; m.Wr = h->d_Wr; m.br = h->d_br; m.Wc = h->d_Wc; m.bc = h->d_bc;
  m.pbin_off = h->d_pboff; m.pbin_pairs = h->d_pbpairs;
  return m;
}

// InitSPWeight (spin_projection.f90:5-79): the two-point rule at beta = {0, pi};
// transverse_code/spin_projection.f90:30-38 takes beta = {-pi/2, pi/2} and
// scales cos and sin of beta/2 by sqrt(2).  para_spin_proj is the same in both.
// RBMPP_PROJ_NORMAL takes cos(pi/2) as exactly 0: the reference's value is
// its rounding residue 6.1e-17, which puts entries of relative size 6e-17 into
// the same-spin blocks of the beta = pi copies; without them every X is
// exactly block off-diagonal and the sampler runs in determinant form.
// RBMPP_PROJ_NORMAL_EXACT keeps the residue (and the general skew path).
static void set_sp_weights(rbmpp_handle* h, int projection) {
  const double pi = std::acos(-1.0);
  const bool tr = projection == RBMPP_PROJ_TRANSVERSE;
  const double beta[2] = {tr ? -0.5 * pi : 0.0, tr ? 0.5 * pi : pi};
  for (int s = 0; s < 2; s++) {
    double c = std::cos(0.5 * beta[s]), sn = std::sin(0.5 * beta[s]);
    if (tr) { c *= std::sqrt(2.0); sn *= std::sqrt(2.0); }
    if (projection == RBMPP_PROJ_NORMAL && s == 1) c = 0.0;
    h->sp_cc[s] = c * c; h->sp_cs[s] = c * sn; h->sp_ss[s] = sn * sn;
  }
  h->projection = projection;
}

static void tic(rbmpp_handle* h) { hipEventRecord(h->e0, h->st); }
static double toc(rbmpp_handle* h) {
  hipEventRecord(h->e1, h->st);
  hipEventSynchronize(h->e1);
  float ms = 0;
  hipEventElapsedTime(&ms, h->e0, h->e1);
  return ms * 1e-3;
}

static int ensure_measure(rbmpp_handle* h, size_t n);

// int32 configurations -> the estimators' int8 layout
__global__ void k_x32_to_x8(size_t n, const int32_t* __restrict__ x, signed char* __restrict__ y) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    y[k] = (signed char)x[k];
}

// out = {first configuration with a nonzero status or -1, its status}
__global__ void k_first_status(int n, const int* __restrict__ st, int* __restrict__ out) {
  __shared__ int first;
  if (threadIdx.x == 0) first = n;
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x)
    if (st[k] != 0) atomicMin(&first, k);
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = first < n ? first : -1;
    out[1] = first < n ? st[first] : 0;
  }
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C" {

const char* rbmpp_last_error(const rbmpp_handle* h) { return h ? h->err.c_str() : "null handle"; }
int rbmpp_nv(const rbmpp_handle* h) { return h ? h->nv : 0; }

void rbmpp_destroy(rbmpp_handle* h) {
  if (!h) return;
  if (h->st) hipStreamSynchronize(h->st);
  for (void* p : h->allocs) hipFree(p);
  if (h->e0) hipEventDestroy(h->e0);
  if (h->e1) hipEventDestroy(h->e1);
  if (h->sr_exec) hipGraphExecDestroy(h->sr_exec);
  if (h->sr_graph) hipGraphDestroy(h->sr_graph);
  if (h->ahead.st2) hipStreamSynchronize(h->ahead.st2);
  if (h->ahead.col_done) hipEventDestroy(h->ahead.col_done);
  if (h->ahead.diag_done) hipEventDestroy(h->ahead.diag_done);
  if (h->ahead.st2) hipStreamDestroy(h->ahead.st2);
  if (h->st) hipStreamDestroy(h->st);
  delete h;
}

int rbmpp_create(const rbmpp_config* cfg, const int32_t* w_map, const double* para_mp_trans,
                 const int32_t* orbital_idx, rbmpp_handle** out) {
  if (!cfg || !w_map || !para_mp_trans || !orbital_idx || !out) return RBMPP_E_ARG;
  *out = nullptr;
  rbmpp_handle* h = new rbmpp_handle();
  h->cfg = *cfg;
  const int L = cfg->L, N = L * L, P = cfg->n_sym;
  h->L = L; h->N = N; h->P = P; h->C = 2 * P; h->ac = cfg->alphac; h->ar = cfg->alphar;
  h->nsl = cfg->n_slater; h->npair = N / 2;
  h->mc = h->ac * P * 2; h->mr = h->ar * P * 2;
  h->nv = h->ac * 2 * (N + 1) + h->ar * (N + 1) + h->nsl;
  h->nv_rbm = h->nv - h->nsl;
  h->W = cfg->n_walkers; h->nspw = cfg->n_samples;
  auto bad = [&](const char* m) { int r = fail(h, RBMPP_E_ARG, m); *out = h; return r; };
  if (L < 2 || N % 2) return bad("L must give an even number of sites");
  if (P < 1 || h->C > RBMPP_CMAX) return bad("n_sym must be in [1, 8]");
  if (h->ac < 0 || h->ar < 0 || h->mr > RBMPP_MRMAX || h->mc > RBMPP_MCMAX) return bad("alpha out of range");
  // the estimators bin the pair-product derivative into at most 4N orbitals
  // (launch_local rejects more); a larger table only serves loaded
  // measurements (rbmpp_set_measured)
  if (h->nsl < 1 || h->nsl > N * N) return bad("n_slater must be in [1, N^2]");
  if (N > 256) return bad("N > 256 not supported");
  if (h->W < 1 || h->nspw < 0) return bad("n_walkers / n_samples");
  if (cfg->spin_parity != 1 && cfg->spin_parity != -1) return bad("spin_parity must be +-1");
  if (cfg->s_precision < 0 || cfg->s_precision > 2) return bad("s_precision");
  // tables: 1-based Fortran layouts -> 0-based
  h->wmap.resize((size_t)P * N);
  h->invmap.resize((size_t)P * N);
  for (int g = 0; g < P; g++) {
    std::vector<int> seen(N, 0);
    for (int i = 0; i < N; i++) {
      const int v = w_map[(size_t)g * N + i] - 1;
      if (v < 0 || v >= N || seen[v]) return bad("W_map is not a permutation (main.f90:110-119)");
      seen[v] = 1;
      h->wmap[(size_t)g * N + i] = v;
      h->invmap[(size_t)g * N + v] = i;
    }
  }
  h->orb.resize((size_t)N * N);
  for (int j = 0; j < N; j++)
    for (int i = 0; i < N; i++) {
      const int v = orbital_idx[(size_t)j * N + i] - 1;
      if (v < 0 || v >= h->nsl) return bad("Orbital_Idx out of range (parameter.f90:98)");
      h->orb[(size_t)i * N + j] = v;
    }
  h->chi.assign(para_mp_trans, para_mp_trans + 2 * P);
  h->real_only = (h->ac == 0);
  for (int g = 0; g < P; g++)
    if (h->chi[2 * g + 1] != 0.0) h->real_only = 0;
  h->subB.resize(N);
  for (int s = 0; s < N; s++) h->subB[s] = (unsigned char)(((s / L) + (s % L)) & 1);
  set_sp_weights(h, RBMPP_PROJ_NORMAL);
  h->sp_w[0] = 1.0;
  h->sp_w[1] = cfg->spin_parity == 1 ? 1.0 : -1.0;
  h->gamma.assign(h->nv, 0.0);
  *out = h;

  HIPCHK(h, hipGetDevice(&h->device));
  HIPCHK(h, hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, h->device));
  HIPCHK(h, hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
  HIPCHK(h, hipStreamCreateWithFlags(&h->ahead.st2, hipStreamNonBlocking));
  HIPCHK(h, hipEventCreateWithFlags(&h->ahead.col_done, hipEventDisableTiming));
  HIPCHK(h, hipEventCreateWithFlags(&h->ahead.diag_done, hipEventDisableTiming));
  HIPCHK(h, hipEventCreate(&h->e0));
  HIPCHK(h, hipEventCreate(&h->e1));
  int rc = 0;
  const size_t W = h->W, C = h->C, NN = (size_t)N * N;
  if ((rc = dalloc(h, &h->d_wmap, (size_t)P * N)) || (rc = dalloc(h, &h->d_invmap, (size_t)P * N)) ||
      (rc = dalloc(h, &h->d_orb, NN)) || (rc = dalloc(h, &h->d_subB, N)) || (rc = dalloc(h, &h->d_chi, P)) ||
      (rc = dalloc(h, &h->d_se, C * 4 * NN)) || (rc = dalloc(h, &h->d_Wr, (size_t)h->ar * N)) ||
      (rc = dalloc(h, &h->d_br, h->ar)) || (rc = dalloc(h, &h->d_Wc, (size_t)h->ac * N)) ||
      (rc = dalloc(h, &h->d_bc, h->ac)) || (rc = dalloc(h, &h->d_f, h->nsl)) ||
      (rc = dalloc(h, &h->d_pboff, (size_t)P * (h->nsl + 1))) ||
      (rc = dalloc(h, &h->d_pbpairs, (size_t)P * N * (N - 1))))
    return rc;
  WalkerState& ws = h->ws;
  if ((rc = dalloc(h, &ws.x, W * N)) || (rc = dalloc(h, &ws.eidx, W * N)) || (rc = dalloc(h, &ws.ecfg, W * N)) ||
      (rc = dalloc(h, &ws.thr, W * std::max(h->mr, 1))) || (rc = dalloc(h, &ws.thc, W * std::max(h->mc, 1))) ||
      (rc = dalloc(h, &ws.pf, W * C)) || (rc = dalloc(h, &ws.sgn, W * P)) || (rc = dalloc(h, &ws.psi, W)) ||
      (rc = dalloc(h, &ws.minv, W * C * (NN / 2))) || (rc = dalloc(h, &ws.nacc_tmp, W)) ||
      (rc = dalloc(h, &ws.nfastup, W)) || (rc = dalloc(h, &ws.ctr, W)) || (rc = dalloc(h, &ws.status, W)) ||
      (rc = dalloc(h, &ws.naccept, W)) || (rc = dalloc(h, &ws.nrecompute, W)))
    return rc;
  ws.scratch = nullptr;
  if (N > RBMPP_NLDS && (rc = dalloc(h, &ws.scratch, W * NN))) return rc;
  const size_t ns = (size_t)W * std::max(h->nspw, 1);
  if ((rc = dalloc(h, &h->d_xs, ns * N)) || (rc = dalloc(h, &h->d_psis, ns))) return rc;
  if ((rc = dalloc(h, &h->d_model, 1))) return rc;
  {
    const DevModel hm = devmodel(h);
    HIPCHK(h, hipMemcpyAsync(h->d_model, &hm, sizeof(DevModel), hipMemcpyHostToDevice, h->st));
  }
  HIPCHK(h, hipMemcpyAsync(h->d_wmap, h->wmap.data(), sizeof(int) * P * N, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(h->d_invmap, h->invmap.data(), sizeof(int) * P * N, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(h->d_orb, h->orb.data(), sizeof(int) * NN, hipMemcpyHostToDevice, h->st));
  // the site pairs of every orbital bin under each operation, for the
  // fixed-order sums of the pair-product log-derivatives
  std::vector<int> pboff((size_t)P * (h->nsl + 1), 0);
  std::vector<unsigned> pbpairs((size_t)P * N * (N - 1));
  for (int g = 0; g < P; g++) {
    int* off = pboff.data() + (size_t)g * (h->nsl + 1);
    const int* wm = h->wmap.data() + (size_t)g * N;
    for (int A = 0; A < N; A++)
      for (int B = 0; B < N; B++)
        if (A != B) off[h->orb[(size_t)wm[A] * N + wm[B]] + 1]++;
    for (int o = 0; o < h->nsl; o++) off[o + 1] += off[o];
    std::vector<int> fill(off, off + h->nsl);
    unsigned* pr = pbpairs.data() + (size_t)g * N * (N - 1);
    for (int A = 0; A < N; A++)
      for (int B = 0; B < N; B++)
        if (A != B) pr[fill[h->orb[(size_t)wm[A] * N + wm[B]]]++] = (unsigned)A | ((unsigned)B << 16);
  }
  HIPCHK(h, hipMemcpyAsync(h->d_pboff, pboff.data(), sizeof(int) * pboff.size(), hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(h->d_pbpairs, pbpairs.data(), sizeof(unsigned) * pbpairs.size(), hipMemcpyHostToDevice,
                           h->st));
  HIPCHK(h, hipMemcpyAsync(h->d_subB, h->subB.data(), N, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(h->d_chi, h->chi.data(), sizeof(double) * 2 * P, hipMemcpyHostToDevice, h->st));
  // walker bookkeeping: Nfastup = min(100, 2N) (main.f90:309), counters 0
  std::vector<int> nf(W, std::min(100, 2 * N)), zi(W, 0);
  std::vector<unsigned long long> zc(W, 0ull);
  HIPCHK(h, hipMemcpyAsync(ws.nfastup, nf.data(), sizeof(int) * W, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(ws.nacc_tmp, zi.data(), sizeof(int) * W, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(ws.ctr, zc.data(), sizeof(unsigned long long) * W, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  if ((rc = noalloc_setup(h))) return rc;
  return RBMPP_OK;
}

int rbmpp_set_projection(rbmpp_handle* h, int32_t projection) {
  if (!h || (projection != RBMPP_PROJ_NORMAL && projection != RBMPP_PROJ_TRANSVERSE &&
             projection != RBMPP_PROJ_NORMAL_EXACT))
    return RBMPP_E_ARG;
  set_sp_weights(h, projection);
  const DevModel hm = devmodel(h);
  HIPCHK(h, hipMemcpyAsync(h->d_model, &hm, sizeof(DevModel), hipMemcpyHostToDevice, h->st));
  const std::vector<double> g = h->gamma;  // Slater_Elm depends on the weights
  return rbmpp_set_gamma(h, g.data());
}

int rbmpp_set_params(rbmpp_handle* h, const double* wcirr, const double* bcirr, const double* wrirr,
                     const double* brirr, const double* slater) {
  if (!h || !slater) return RBMPP_E_ARG;
  const int N = h->N;
  std::vector<double> g(h->nv);
  int k = 0;
  for (int f = 0; f < h->ac; f++) { g[k++] = bcirr[2 * f]; g[k++] = bcirr[2 * f + 1]; }
  for (int f = 0; f < h->ac; f++)
    for (int iw = 0; iw < N; iw++) { g[k++] = wcirr[2 * (f * N + iw)]; g[k++] = wcirr[2 * (f * N + iw) + 1]; }
  for (int f = 0; f < h->ar; f++) g[k++] = brirr[f];
  for (int f = 0; f < h->ar; f++)
    for (int iw = 0; iw < N; iw++) g[k++] = wrirr[f * N + iw];
  for (int o = 0; o < h->nsl; o++) g[k++] = slater[o];
  return rbmpp_set_gamma(h, g.data());
}

int rbmpp_set_gamma(rbmpp_handle* h, const double* gamma) {
  if (!h || !gamma) return RBMPP_E_ARG;
  const int N = h->N;
  h->gamma.assign(gamma, gamma + h->nv);
  std::vector<double> Wr((size_t)h->ar * N), br(h->ar), Wc((size_t)2 * h->ac * N), bc(2 * h->ac), f(h->nsl);
  int k = 0;
  for (int q = 0; q < h->ac; q++) { bc[2 * q] = gamma[k++]; bc[2 * q + 1] = gamma[k++]; }
  for (int q = 0; q < h->ac; q++)
    for (int iw = 0; iw < N; iw++) { Wc[2 * (q * N + iw)] = gamma[k++]; Wc[2 * (q * N + iw) + 1] = gamma[k++]; }
  for (int q = 0; q < h->ar; q++) br[q] = gamma[k++];
  for (int q = 0; q < h->ar; q++)
    for (int iw = 0; iw < N; iw++) Wr[q * N + iw] = gamma[k++];
  for (int o = 0; o < h->nsl; o++) f[o] = gamma[k++];
  if (h->ar) {
    HIPCHK(h, hipMemcpyAsync(h->d_Wr, Wr.data(), sizeof(double) * Wr.size(), hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(h->d_br, br.data(), sizeof(double) * br.size(), hipMemcpyHostToDevice, h->st));
  }
  if (h->ac) {
    HIPCHK(h, hipMemcpyAsync(h->d_Wc, Wc.data(), sizeof(double) * Wc.size(), hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(h->d_bc, bc.data(), sizeof(double) * bc.size(), hipMemcpyHostToDevice, h->st));
  }
  HIPCHK(h, hipMemcpyAsync(h->d_f, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice, h->st));
  tic(h);
  launch_slater(devmodel(h), h->d_f, h->d_se, nullptr, h->st);
  HIPCHK(h, hipGetLastError());
  h->t[0] = toc(h);
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->s_valid = 0;
  h->fused_valid = 0;  // the fused values belong to the previous parameters
  return RBMPP_OK;
}

int rbmpp_get_gamma(const rbmpp_handle* h, double* gamma) {
  if (!h || !gamma) return RBMPP_E_ARG;
  std::copy(h->gamma.begin(), h->gamma.end(), gamma);
  return RBMPP_OK;
}

// sweep arguments of one make_sample call (make_sample.f90:4-240 for all walkers)
static int sweep_args(rbmpp_handle* h, int32_t first_iteration, double shift, bool allow_fused, SweepArgs& a,
                      bool& fused) {
  if (!first_iteration && !h->walkers_ready) return fail(h, RBMPP_E_STATE, "make_sample: walkers not initialised");
  a.m = h->d_model;
  a.ws = h->ws;
  a.nwalk = h->W;
  a.nsamp = h->nspw;
  int nwarm = h->cfg.n_warmup;
  if (first_iteration && h->cfg.n_warmup_first >= 0) nwarm = h->cfg.n_warmup_first;
  if (nwarm < 0) nwarm = std::max(30, h->nspw / 100);  // make_sample.f90:66
  a.nwarm = nwarm;
  a.nupdate = h->cfg.n_update > 0 ? h->cfg.n_update : h->N;
  a.fixed_k = h->cfg.recompute_k;
  a.first = first_iteration ? 1 : 0;
  a.shift = shift;
  a.p_restart = h->cfg.p_restart;
  a.p_global = h->cfg.p_global;
  a.cond_guard = h->cfg.cond_guard;
  a.key0 = (uint32_t)h->cfg.seed;
  a.key1 = (uint32_t)(h->cfg.seed >> 32);
  a.xs = h->d_xs;
  a.psis = h->d_psis;
  a.measure = 0;
  a.m_psi = nullptr; a.m_eloc = nullptr; a.m_rw = nullptr; a.m_orows = nullptr;
  h->fused_valid = 0;
  fused = allow_fused && h->cfg.fused_measure && h->N <= RBMPP_NLDS && h->nspw > 0;
  if (fused) {
    int rc = ensure_measure(h, (size_t)h->W * h->nspw);
    if (rc) return rc;
    a.measure = 1;
    a.m_psi = h->d_mpsi; a.m_eloc = h->d_eloc; a.m_rw = h->d_rw; a.m_orows = h->d_orows;
  }
  return 0;
}

// after a sweep launch: walker statuses, acceptance and recompute counts
static int sweep_finish(rbmpp_handle* h, bool fused, double shift, int64_t* naccept) {
  std::vector<int> stat(h->W);
  std::vector<long long> nacc(h->W);
  std::vector<int> nrec(h->W);
  HIPCHK(h, hipMemcpy(stat.data(), h->ws.status, sizeof(int) * h->W, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(nacc.data(), h->ws.naccept, sizeof(long long) * h->W, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(nrec.data(), h->ws.nrecompute, sizeof(int) * h->W, hipMemcpyDeviceToHost));
  long long tot = 0, trec = 0;
  for (int w = 0; w < h->W; w++) {
    if (stat[w] != 0) {
      char buf[128];
      snprintf(buf, sizeof buf, "make_sample: walker %d status %d", w, stat[w]);
      return fail(h, stat[w], buf);
    }
    tot += nacc[w];
    trec += nrec[w];
  }
  h->t[6] = (double)trec;
  if (naccept) *naccept = tot;
  h->walkers_ready = 1;
  h->fused_valid = fused ? 1 : 0;
  h->fused_shift = shift;
  h->n_stored = (int64_t)h->W * h->nspw;
  if (fused) h->n_meas = h->n_stored;
  h->measured = 0;
  return RBMPP_OK;
}

int rbmpp_make_sample(rbmpp_handle* h, int32_t first_iteration, double shift, int64_t* naccept) {
  if (!h) return RBMPP_E_ARG;
  SweepArgs a;
  bool fused = false;
  int rc = sweep_args(h, first_iteration, shift, true, a, fused);
  if (rc) return rc;
  tic(h);
  {
    const DevModel hm = devmodel(h);
    if (launch_sweep(hm, a, h->st) != 0) return fail(h, RBMPP_E_ARG, "sweep: unsupported size");
    h->minv_block = hm.block;  // layout of ws.minv written by this launch
  }
  HIPCHK(h, hipGetLastError());
  h->t[1] = toc(h);
  return sweep_finish(h, fused, shift, naccept);
}

// frees a buffer made by dalloc
static void dfree(rbmpp_handle* h, void* p) {
  if (!p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
  if (it != h->allocs.end()) h->allocs.erase(it);
  hipFree(p);
}

// (re)allocate the measurement buffers for n configurations: all or nothing,
// so a failed request leaves the previous buffers in place
static int ensure_measure(rbmpp_handle* h, size_t n) {
  if (n <= h->meas_cap && h->d_mpsi) return 0;
  int rc;
  const size_t rows = h->real_only ? n : 2 * n;
  double2 *mpsi = nullptr, *eloc = nullptr;
  double *rw = nullptr, *orows = nullptr, *kappa = nullptr;
  int* status = nullptr;
  if ((rc = dalloc(h, &mpsi, n)) || (rc = dalloc(h, &eloc, n)) || (rc = dalloc(h, &rw, n)) ||
      (rc = dalloc(h, &orows, rows * h->nv)) || (rc = dalloc(h, &kappa, rows)) || (rc = dalloc(h, &status, n))) {
    for (void* q : {(void*)mpsi, (void*)eloc, (void*)rw, (void*)orows, (void*)kappa, (void*)status}) dfree(h, q);
    (void)hipGetLastError();
    return rc;
  }
  for (void* q : {(void*)h->d_mpsi, (void*)h->d_eloc, (void*)h->d_rw, (void*)h->d_orows, (void*)h->d_kappa,
                  (void*)h->d_status})
    dfree(h, q);
  h->d_mpsi = mpsi; h->d_eloc = eloc; h->d_rw = rw; h->d_orows = orows; h->d_kappa = kappa; h->d_status = status;
  h->meas_cap = n;
  // the local kernel is grid-strided; N > RBMPP_NLDS needs one N x N
  // recompute scratch per block, sized once for the largest grid
  h->local_grid = h->N > RBMPP_NLDS ? 1024 : 4096;
  if (h->N > RBMPP_NLDS && !h->d_scr &&
      (rc = dalloc(h, &h->d_scr, (size_t)h->local_grid * h->N * h->N)))
    return rc;
  return 0;
}

static int ensure_part(rbmpp_handle* h, size_t n) {
  if (n <= h->part_cap) return 0;
  double* p = nullptr;
  int rc = dalloc(h, &p, n);
  if (rc) return rc;
  dfree(h, h->d_part);
  h->d_part = p;
  h->part_cap = n;
  return 0;
}

// workgroups of the block-form estimator (each holds C (N/2)^2 doubles of scratch)
static const int BLK_GRID = 1024;
// workgroups of the correlation run (each keeps its own partial sums); the
// block form launches at most BLK_GRID, and every slice summed must be written
static const int CORR_GRID = 1024;
static_assert(CORR_GRID <= BLK_GRID, "the correlation run's workgroups all run in block form too");

static int run_local(rbmpp_handle* h, int n, const signed char* d_x, int mode, double shift, const int* ex1,
                     const int* ex2, double* d_invm, double* d_pf, double2* d_psiq, double2* d_expsi,
                     double* d_expf, double* d_corr = nullptr, int grid_cap = 0) {
  int rc = ensure_measure(h, (size_t)n);
  if (rc) return rc;
  h->fused_valid = 0;  // the buffers are about to be overwritten
  LocalArgs a;
  memset(&a, 0, sizeof a);
  a.m = h->d_model;
  a.nconf = n;
  a.xs = d_x;
  a.mode = mode;
  a.shift = shift;
  a.ex_i1 = ex1;
  a.ex_i2 = ex2;
  a.scratch = h->d_scr;
  a.psi = h->d_mpsi;
  a.pf = d_pf;
  a.psi_q = d_psiq;
  a.invm = d_invm;
  a.eloc = h->d_eloc;
  a.orows = h->d_orows;
  a.rw = h->d_rw;
  a.ex_psi = d_expsi;
  a.ex_pf = d_expf;
  a.status = h->d_status;
  a.corr_part = d_corr;
  int cap = grid_cap > 0 ? grid_cap : h->local_grid;
  if (grid_cap <= 0 && h->tune_meas_grid > 0) cap = std::min(cap, h->tune_meas_grid);
  const DevModel hm0 = devmodel(h);
  if (hm0.block && hm0.N <= RBMPP_NLDS) {
    // the block-form estimator keeps F^-1 of every copy per workgroup
    cap = std::min(cap, BLK_GRID);
    if (!h->d_bscr && (rc = dalloc(h, &h->d_bscr, (size_t)BLK_GRID * h->C * (h->N / 2) * (h->N / 2)))) return rc;
    a.scratch = h->d_bscr;
  }
  const int grid = std::min(n, cap);
  // the kernel reads the descriptor from device memory; real_only may have
  // been switched for this call
  {
    const DevModel hm = devmodel(h);
    HIPCHK(h, hipMemcpyAsync(h->d_model, &hm, sizeof(DevModel), hipMemcpyHostToDevice, h->st));
  }
  if (launch_local(devmodel(h), a, grid, h->st) != 0) return fail(h, RBMPP_E_ARG, "local: unsupported size");
  HIPCHK(h, hipGetLastError());
  std::vector<int> stat(n);
  HIPCHK(h, hipMemcpyAsync(stat.data(), h->d_status, sizeof(int) * n, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  for (int s = 0; s < n; s++)
    if (stat[s] != 0) {
      char buf[128];
      snprintf(buf, sizeof buf, "configuration %d: status %d", s, stat[s]);
      return fail(h, stat[s], buf);
    }
  return 0;
}

// sums over the measured configurations (main.f90:351-367)
static int reduce_measure(rbmpp_handle* h, int n, bool with_o) {
  const int nblk = 256;
  const int nchunk = std::max(1, std::min(256, n / 64));
  int rc = ensure_part(h, std::max<size_t>((size_t)nblk * 4, (size_t)(nchunk + 1) * 2 * h->nv));
  if (rc) return rc;
  launch_scalar_partials(n, h->d_mpsi, h->d_rw, h->d_eloc, h->d_part, nblk, h->st);
  std::vector<double> part((size_t)nblk * 4);
  HIPCHK(h, hipMemcpyAsync(part.data(), h->d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  double s[4] = {0, 0, 0, 0};
  for (int b = 0; b < nblk; b++)
    for (int k = 0; k < 4; k++) s[k] += part[(size_t)b * 4 + k];
  h->m_norm = s[2];
  h->m_etot = s[0] / s[2];
  h->m_etot2 = s[1] / s[2];
  h->m_wfav = s[3] / (double)n;
  if (with_o) {
    launch_gvec_partials(n, h->nv, h->real_only ? 0 : 1, h->d_orows, h->d_mpsi, h->d_rw, h->d_eloc, h->d_part,
                         nchunk, h->st);
    HIPCHK(h, hipGetLastError());
    std::vector<double> go(2 * h->nv);
    HIPCHK(h, hipMemcpyAsync(go.data(), h->d_part + (size_t)nchunk * 2 * h->nv, sizeof(double) * 2 * h->nv,
                             hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->m_g.assign(go.begin(), go.begin() + h->nv);
    h->m_o.assign(go.begin() + h->nv, go.end());
    for (int k = 0; k < h->nv; k++) { h->m_g[k] /= h->m_norm; h->m_o[k] /= h->m_norm; }
  }
  return 0;
}

int rbmpp_measure(rbmpp_handle* h, double shift, double* etot, double* etot2, double* norm, double* wf_av,
                  double* gvec, double* ovec) {
  if (!h) return RBMPP_E_ARG;
  if (h->n_stored <= 0) return fail(h, RBMPP_E_STATE, "measure: no samples");
  const int n = (int)h->n_stored;
  tic(h);
  int rc = 0;
  if (!(h->fused_valid && h->fused_shift == shift))
    rc = run_local(h, n, h->d_xs, 1 | 2, shift, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  h->n_meas = n;
  rc = reduce_measure(h, n, true);
  if (rc) return rc;
  h->t[2] = toc(h);
  if (etot) *etot = h->m_etot;
  if (etot2) *etot2 = h->m_etot2;
  if (norm) *norm = h->m_norm;
  if (wf_av) *wf_av = h->m_wfav;
  if (gvec) std::copy(h->m_g.begin(), h->m_g.end(), gvec);
  if (ovec) std::copy(h->m_o.begin(), h->m_o.end(), ovec);
  h->measured = 1;
  h->s_valid = 0;
  return RBMPP_OK;
}

int rbmpp_measure_energy(rbmpp_handle* h, double shift, double* etot, double* etot2, double* norm,
                         double* wf_av) {
  if (!h) return RBMPP_E_ARG;
  if (h->n_stored <= 0) return fail(h, RBMPP_E_STATE, "measure_energy: no samples");
  const int n = (int)h->n_stored;
  tic(h);
  int rc = 0;
  if (!(h->fused_valid && h->fused_shift == shift))
    rc = run_local(h, n, h->d_xs, 1, shift, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  h->n_meas = n;
  rc = reduce_measure(h, n, false);
  if (rc) return rc;
  h->t[2] = toc(h);
  if (etot) *etot = h->m_etot;
  if (etot2) *etot2 = h->m_etot2;
  if (norm) *norm = h->m_norm;
  if (wf_av) *wf_av = h->m_wfav;
  return RBMPP_OK;
}

int rbmpp_measure_correlation(rbmpp_handle* h, double shift, double* etot, double* etot2, double* norm,
                              double* wf_av, double* wf_sgn, double* chi_sp, double* chi_vbs) {
  if (!h || !chi_sp || !chi_vbs) return RBMPP_E_ARG;
  if (h->n_stored <= 0) return fail(h, RBMPP_E_STATE, "measure_correlation: no samples");
  if (h->N > RBMPP_NLDS || h->L < 4)
    return fail(h, RBMPP_E_ARG, "measure_correlation: needs 4 <= L and N <= 104");
  const int n = (int)h->n_stored, N = h->N;
  const int grid = std::min(n, CORR_GRID);
  const size_t len = (size_t)corr_len(N), nsum = (size_t)corr_sums(N);
  int rc;
  if (h->cpart_cap < (size_t)grid * len) {
    if ((rc = dalloc(h, &h->d_cpart, (size_t)grid * len))) return rc;
    h->cpart_cap = (size_t)grid * len;
  }
  if (!h->d_csum && (rc = dalloc(h, &h->d_csum, nsum))) return rc;
  tic(h);
  rc = run_local(h, n, h->d_xs, 1, shift, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                 h->d_cpart, grid);
  if (rc) return rc;
  h->n_meas = n;
  rc = reduce_measure(h, n, false);
  if (rc) return rc;
  launch_sum_strided((int)nsum, grid, (int)len, h->d_cpart, h->d_csum, h->st);
  HIPCHK(h, hipGetLastError());
  std::vector<double> sums(nsum);
  HIPCHK(h, hipMemcpyAsync(sums.data(), h->d_csum, sizeof(double) * nsum, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->t[2] = toc(h);
  // normalisation of Calc_Correlation/main.f90:548-581
  const size_t nchi = (size_t)corr_nchi(N);
  const double nrm = h->m_norm;
  for (size_t i0 = 0; i0 < nchi; i0++) {
    chi_sp[2 * i0] = sums[i0] / nrm;
    chi_sp[2 * i0 + 1] = sums[nchi + i0] / nrm;
  }
  for (size_t k = 0; k < 4 * nchi; k++) chi_vbs[k] = sums[2 * nchi + k] / nrm;
  if (wf_sgn) {
    wf_sgn[0] = sums[6 * nchi] / (double)n;
    wf_sgn[1] = sums[6 * nchi + 1] / (double)n;
  }
  if (etot) *etot = h->m_etot;
  if (etot2) *etot2 = h->m_etot2;
  if (norm) *norm = h->m_norm;
  if (wf_av) *wf_av = h->m_wfav;
  return RBMPP_OK;
}

// the f_ii mask of the SR checks (SR_pdposv.f90:78-100) on the device
static int ensure_fii(rbmpp_handle* h) {
  if (h->d_fii) return 0;
  const int nv = h->nv, N = h->N, nsl = h->nsl;
  int rc = dalloc(h, &h->d_fii, (size_t)nv);
  if (rc) return rc;
  std::vector<unsigned char> fii(nv, 0);
  for (int i = 0; i < N; i++) fii[h->orb[(size_t)i * N + i] + nv - nsl] = 1;
  HIPCHK(h, hipMemcpyAsync(h->d_fii, fii.data(), nv, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return 0;
}

// Seeded synthetic measurement (rbmpp_set_measured with orows == NULL):
// psi = 1, rw = 1, eloc and the O entries uniform in (-1, 1) from a hash of
// (seed, row, column); the f_ii columns are zero, as in a measured O.
__device__ __forceinline__ double synth_u(uint64_t seed, uint64_t a) {
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (a + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__global__ void k_synth_measurement(int64_t n, int nv, int cplx_rows, uint64_t seed,
                                    const unsigned char* __restrict__ fii, double2* psi, double2* eloc, double* rw,
                                    double* orows) {
  const size_t rows = (size_t)n * (cplx_rows ? 2 : 1);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < rows * nv; e += stride) {
    const int k = (int)(e % nv);
    orows[e] = fii[k] ? 0.0 : synth_u(seed, e);
  }
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < (size_t)n; s += stride) {
    psi[s] = make_double2(1.0, 0.0);
    rw[s] = 1.0;
    eloc[s] = make_double2(synth_u(seed ^ 0x5555555555555555ull, s), 0.0);
  }
}

// scratch of the FP64 O-row passes (diagonal of a mixed S, SR refinement)
static int ensure_op(rbmpp_handle* h, int rows) {
  const size_t need = orow_op_scratch(rows, h->nv);
  if (need <= h->op_cap) return 0;
  double* op = nullptr;
  int rc = dalloc(h, &op, need);
  if (rc) return rc;
  dfree(h, h->d_op);
  h->d_op = op;
  h->op_cap = need;
  return 0;
}

// centred S = <O O^T>/norm - <O><O>^T in the configured precision: the O
// rows are centred once in FP64 and packed in the precision of the products
// (rbmpp_syrk.hip), then multiplied on the matrix cores; a mixed-precision
// S gets its diagonal from an FP64 pass over the O rows
static int build_s(rbmpp_handle* h) {
  const int nv = h->nv, n = (int)h->n_meas;
  int rc;
  if (!h->d_S && ((rc = dalloc(h, &h->d_S, (size_t)nv * nv)) || (rc = dalloc(h, &h->d_A, (size_t)nv * nv)) ||
                  (rc = dalloc(h, &h->d_L, (size_t)nv * nv)) || (rc = dalloc(h, &h->d_vec, (size_t)8 * nv)) ||
                  (rc = dalloc(h, &h->d_info, 1)) || (rc = dalloc(h, &h->d_sc, 16)) || (rc = dalloc(h, &h->d_fl, 8)) ||
                  (rc = dalloc(h, &h->d_dinv, chol_dinv_size(nv)))))
    return rc;
  const int rows = h->real_only ? n : 2 * n;
  const int prec = h->cfg.s_precision;
  const SyrkPlan pl = syrk_plan(prec, rows, nv, h->n_cu, h->tune_tile[prec], h->tune_nsplit[prec]);
  if ((rc = ensure_part(h, pl.part_doubles))) return rc;
  if (prec != RBMPP_S_FP64 && (rc = ensure_op(h, rows))) return rc;
  if (pl.packed_bytes > h->packed_cap) {
    unsigned char* pk = nullptr;
    if ((rc = dalloc(h, &pk, pl.packed_bytes))) return rc;
    dfree(h, h->d_packed);
    h->d_packed = pk;
    h->packed_cap = pl.packed_bytes;
  }
  // obar to device (kept after the S build for the SR refinement)
  double* d_obar = h->d_vec + 4 * nv;
  HIPCHK(h, hipMemcpyAsync(d_obar, h->m_o.data(), sizeof(double) * nv, hipMemcpyHostToDevice, h->st));
  launch_row_kappa(n, h->real_only ? 0 : 1, h->d_mpsi, h->d_rw, h->d_kappa, h->st);
  tic(h);
  // a mixed-precision S gets its FP64 diagonal from the packing pass's column sums
  double* dpart = prec != RBMPP_S_FP64 ? h->d_op : nullptr;
  launch_center_pack(pl, rows, nv, h->d_orows, h->d_kappa, d_obar, h->d_packed, dpart, h->st);
  launch_syrk_packed(pl, nv, h->d_packed, h->d_part, 1.0 / h->m_norm, h->d_S, h->st);
  if (dpart) launch_s_diag_finish(nv, 1.0 / h->m_norm, dpart, h->d_S, h->st);
  HIPCHK(h, hipGetLastError());
  h->t[3] = toc(h);
  h->s_valid = 1;
  return 0;
}

int rbmpp_get_smat(rbmpp_handle* h, double* s) {
  if (!h || !s) return RBMPP_E_ARG;
  if (!h->measured) return fail(h, RBMPP_E_STATE, "get_smat: no measurement");
  if (!h->s_valid) {
    int rc = build_s(h);
    if (rc) return rc;
  }
  HIPCHK(h, hipMemcpyAsync(s, h->d_S, sizeof(double) * h->nv * h->nv, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RBMPP_OK;
}

int rbmpp_sr_step(rbmpp_handle* h, const rbmpp_sr_config* c, double* gamma, rbmpp_sr_report* rep) {
  if (!h || !c || !gamma) return RBMPP_E_ARG;
  if (!h->measured) return fail(h, RBMPP_E_STATE, "sr_step before measure");
  if (c->refine_iters < 0 || c->refine_iters > 8) return fail(h, RBMPP_E_ARG, "refine_iters must be in [0, 8]");
  const int nv = h->nv, nvr = h->nv_rbm, N = h->N, nsl = h->nsl;
  int rc = build_s(h);
  if (rc) return rc;
  if ((rc = ensure_fii(h))) return rc;
  hipEvent_t ea, eb;
  HIPCHK(h, hipEventCreate(&ea));
  HIPCHK(h, hipEventCreate(&eb));
  HIPCHK(h, hipEventRecord(ea, h->st));
  // SR_pdposv.f90:4-6 : g <- g - 2 E <O>; s_factor = max(1e-4, s_factor0)
  std::vector<double> g(nv);
  for (int k = 0; k < nv; k++) g[k] = h->m_g[k] - 2.0 * h->m_etot * h->m_o[k];
  const double sf = std::max(1e-4, c->s_factor0);
  SrPipeline p;
  p.nv = nv; p.nvr = nvr; p.refine_iters = c->refine_iters;
  p.S = h->d_S; p.fii = h->d_fii; p.A = h->d_A; p.L = h->d_L;
  p.x = h->d_vec; p.r = h->d_vec + nv; p.ax = h->d_vec + 2 * nv;
  p.z = h->d_vec + 5 * nv; p.pd = h->d_vec + 6 * nv; p.tmp = h->d_vec + 7 * nv;
  double* d_g = h->d_vec + 3 * nv;
  p.g = d_g;
  p.sc = h->d_sc; p.fl = h->d_fl; p.dinv = h->d_dinv;
  p.ahead = h->ahead;
  // a mixed-precision S is refined against the FP64 operator of the O rows
  if (h->cfg.s_precision != RBMPP_S_FP64 && !(c->flags & RBMPP_SR_ASSEMBLED_REFINE)) {
    const int rows = h->real_only ? (int)h->n_meas : 2 * (int)h->n_meas;
    if ((rc = ensure_op(h, rows))) return rc;
    p.orows = h->d_orows; p.kappa = h->d_kappa; p.obar = h->d_vec + 4 * nv;
    p.rows = rows; p.part = h->d_op;
  }
  const double inv_norm = 1.0 / h->m_norm;
  HIPCHK(h, hipMemcpyAsync(d_g, g.data(), sizeof(double) * nv, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(h->d_sc + 3, &sf, sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(h->d_sc + 7, &inv_norm, sizeof(double), hipMemcpyHostToDevice, h->st));
  // the launch sequence (scalars, system, Cholesky, solves, refinement,
  // lambda) has no host synchronisation; it is captured once into a graph
  // and replayed (RBMPP_SR_NO_GRAPH runs it eagerly)
  auto run_pipeline = [&]() -> int {
    if (c->flags & RBMPP_SR_NO_GRAPH) {
      launch_sr_pipeline(p, h->st);
      HIPCHK(h, hipGetLastError());
      return 0;
    }
    if (!h->sr_exec || h->sr_graph_iters != c->refine_iters || h->sr_graph_orows != p.orows ||
        h->sr_graph_kappa != p.kappa || h->sr_graph_rows != p.rows || h->sr_graph_part != p.part) {
      if (h->sr_exec) { hipGraphExecDestroy(h->sr_exec); h->sr_exec = nullptr; }
      if (h->sr_graph) { hipGraphDestroy(h->sr_graph); h->sr_graph = nullptr; }
      HIPCHK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
      launch_sr_pipeline(p, h->st);
      const hipError_t le = hipGetLastError();
      HIPCHK(h, hipStreamEndCapture(h->st, &h->sr_graph));
      HIPCHK(h, le);
      HIPCHK(h, hipGraphInstantiate(&h->sr_exec, h->sr_graph, nullptr, nullptr, 0));
      h->sr_graph_iters = c->refine_iters;
      h->sr_graph_orows = p.orows;
      h->sr_graph_kappa = p.kappa;
      h->sr_graph_rows = p.rows;
      h->sr_graph_part = p.part;
    }
    HIPCHK(h, hipGraphLaunch(h->sr_exec, h->st));
    return 0;
  };
  double sc[8];
  int fl[8];
  auto read_scalars = [&]() -> int {
    HIPCHK(h, hipMemcpyAsync(sc, h->d_sc, sizeof sc, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipMemcpyAsync(fl, h->d_fl, sizeof fl, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
  };
  // With the operator of the O rows, an assembled mixed-precision matrix
  // that is not positive definite only preconditions: its factor is retried
  // with a larger diagonal (10, 100, 1000, 10^4 times the shift), and the
  // conjugate gradients still solve the FP64 system.
  double boost = 0.0;
  for (int attempt = 0;; attempt++) {
    HIPCHK(h, hipMemcpyAsync(h->d_sc + 12, &boost, sizeof(double), hipMemcpyHostToDevice, h->st));
    if ((rc = run_pipeline())) return rc;
    if ((rc = read_scalars())) return rc;
    if (fl[4] == 0 || !p.orows || fl[3] != 0 || attempt == 4) break;
    boost = (boost == 0.0 ? 10.0 : 10.0 * boost / sc[1]) * sc[1];
  }
  // more conjugate-gradient iterations (up to 40 per round) until
  // the residual against the operator of the O rows is below 1e-13.  The
  // recurrence's residual can drift from D g - A x over many iterations (a
  // poor, boosted preconditioner of an ill-conditioned system), so each
  // round ends with the true residual, which also gives lambda and the
  // reported residual; a round that leaves it above the tolerance restarts
  // the recurrence from it (up to four rounds).
  int pcg_iters = c->refine_iters;
  const double tol = 1e-13;
  // iterations per residual check: one where a pass over the O rows (at
  // least 256 MB) outweighs the host round trip of the check, else four
  const int chunk = (double)p.rows * nv * sizeof(double) >= 256.0 * (1 << 20) ? 1 : 4;
  auto rel_res = [&]() { return std::sqrt(sc[5] / std::max(sc[6], 1e-300)); };
  if (p.orows && fl[4] == 0 && fl[3] == 0)
    for (int round = 0; round < 4; round++) {
      if (round > 0) launch_pcg_direction(p, h->st);  // z = M^-1 r, p = z from the true residual
      int k = round == 0 ? c->refine_iters : 0, k_end = k + 40;
      while (rel_res() > tol && k < k_end) {
        launch_pcg_iters(p, k, k + chunk, h->st);
        launch_pcg_final(p, h->st);
        HIPCHK(h, hipGetLastError());
        k += chunk;
        pcg_iters += chunk;
        if ((rc = read_scalars())) return rc;
      }
      launch_true_residual(p, h->st);  // r = D g - A x
      launch_pcg_final(p, h->st);      // A x = D g - r: lambda and the true residual
      HIPCHK(h, hipGetLastError());
      if ((rc = read_scalars())) return rc;
      if (rel_res() <= tol) break;
    }
  std::vector<double> x(nv);
  HIPCHK(h, hipMemcpyAsync(x.data(), p.x, sizeof(double) * nv, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipEventRecord(eb, h->st));
  HIPCHK(h, hipEventSynchronize(eb));
  float ms = 0;
  hipEventElapsedTime(&ms, ea, eb);
  h->t[4] = ms * 1e-3;
  hipEventDestroy(ea);
  hipEventDestroy(eb);
  h->sr_pcg_iters = p.orows ? pcg_iters : 0;
  h->sr_boost = boost;
  // SR_pdposv.f90:32-33, 78-100 : the structural checks of the diagonal
  if (fl[0] != nvr) return fail(h, RBMPP_E_SDIAG, "i1/=Nv_RBM (SR_pdposv.f90:32): " + std::to_string(fl[0]) + " positive");
  if (fl[1] != nsl - nsl / N)
    return fail(h, RBMPP_E_SDIAG, "i2 /= N_Slater - N_Slater/N (SR_pdposv.f90:33): " + std::to_string(fl[1]) + " positive");
  if (fl[3] == 1) return fail(h, RBMPP_E_SDIAG, "diagonal element of S matrix is not positive (S[" + std::to_string(fl[2]) + "])");
  if (fl[3] == 2) return fail(h, RBMPP_E_SDIAG, "zero S diagonal for other than f_ii (S[" + std::to_string(fl[2]) + "])");
  if (fl[3] == 3) return fail(h, RBMPP_E_SDIAG, "positive S diagonal for f_ii (S[" + std::to_string(fl[2]) + "])");
  if (fl[4] != 0) return fail(h, RBMPP_E_NOT_SPD, "Cholesky: matrix not positive definite at pivot " + std::to_string(fl[4]));
  const double scal = sc[0], sav = sc[2];
  const double resid = std::sqrt(sc[5] / std::max(sc[6], 1e-300));
  // SR_pdposv.f90:251-311 : lambda = sqrt(x^T A x), trust region 5/L
  const double lam = std::sqrt(sc[4]) * 5.0 / (double)h->L;
  if (lam > 1.0)
    for (int k = 0; k < nv; k++) x[k] /= lam;
  for (int k = nvr; k < nv; k++) x[k] *= scal;
  // SR_pdposv.f90:329-366 : max-change clip, update, f renormalisation
  double dmax = 0;
  int knz = 0;
  for (int k = 0; k < nv; k++) {
    if (std::fabs(c->delta_tau * x[k]) > dmax) dmax = std::fabs(c->delta_tau * x[k]);
    if (std::fabs(x[k]) > 1e-12) knz++;
  }
  for (int k = 0; k < nv; k++) {
    if (dmax > c->max_change) gamma[k] = gamma[k] - c->delta_tau * x[k] * (c->max_change / dmax);
    else gamma[k] = gamma[k] - c->delta_tau * x[k];
  }
  double r1 = 0;
  for (int k = nv - nsl; k < nv; k++) r1 += std::fabs(gamma[k]);
  r1 /= nsl;
  for (int k = nv - nsl; k < nv; k++) gamma[k] = gamma[k] * c->slater_abs_av / r1;
  if (rep) {
    rep->lambda = lam; rep->delw_max = dmax; rep->scaling = scal; rep->sdiag_av = sav; rep->s_factor = sf;
    rep->residual = resid; rep->mean_abs_f = r1; rep->k_nonzero = knz; rep->cg_iters = h->sr_pcg_iters;
    rep->precond_boost = h->sr_boost;
  }
  rc = rbmpp_set_gamma(h, gamma);
  h->t[5] = h->t[3] + h->t[4];
  return rc;
}

// ---------------------------------------------------------- test entries
// device temporaries of one call, freed by free_tmp; a failed allocation
// returns RBMPP_E_NOMEM before anything is launched on the buffer
extern "C++" template <typename T>
static int tmp_alloc(rbmpp_handle* h, T** p, size_t n, std::vector<void*>& tmp) {
  *p = nullptr;
  hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();  // clear the sticky error of the failed request
    return fail(h, RBMPP_E_NOMEM, std::string("hipMalloc (") + std::to_string(n * sizeof(T)) +
                                      " bytes): " + hipGetErrorString(e));
  }
  tmp.push_back((void*)*p);
  return 0;
}

static int upload_x(rbmpp_handle* h, int n, const int32_t* x, signed char** d_out, std::vector<void*>& tmp) {
  std::vector<signed char> xb((size_t)n * h->N);
  for (size_t k = 0; k < xb.size(); k++) {
    if (x[k] != 1 && x[k] != -1) return fail(h, RBMPP_E_ARG, "x must be +-1");
    xb[k] = (signed char)x[k];
  }
  for (int s = 0; s < n; s++) {
    int up = 0;
    for (int i = 0; i < h->N; i++) up += x[(size_t)s * h->N + i] == 1;
    if (up != h->npair) return fail(h, RBMPP_E_ARG, "configuration not in the Sz=0 sector");
  }
  signed char* d = nullptr;
  int rc = tmp_alloc(h, &d, xb.size(), tmp);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(d, xb.data(), xb.size(), hipMemcpyHostToDevice, h->st));
  *d_out = d;
  return 0;
}

static void free_tmp(std::vector<void*>& tmp) {
  for (void* p : tmp) hipFree(p);
  tmp.clear();
}

// synchronous device-to-host copy with the status mapped to RBMPP_E_HIP
static int copy_out(rbmpp_handle* h, void* dst, const void* src, size_t bytes) {
  HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return 0;
}

int rbmpp_set_samples(rbmpp_handle* h, int32_t n, const int32_t* x) {
  if (!h || !x || n <= 0) return RBMPP_E_ARG;
  if ((int64_t)n > (int64_t)h->W * std::max(h->nspw, 1)) return fail(h, RBMPP_E_ARG, "too many samples");
  std::vector<signed char> xb((size_t)n * h->N);
  for (size_t k = 0; k < xb.size(); k++) {
    if (x[k] != 1 && x[k] != -1) return fail(h, RBMPP_E_ARG, "x must be +-1");
    xb[k] = (signed char)x[k];
  }
  HIPCHK(h, hipMemcpyAsync(h->d_xs, xb.data(), xb.size(), hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->n_stored = n;
  h->fused_valid = 0;
  h->measured = 0;
  return RBMPP_OK;
}

int64_t rbmpp_n_samples(const rbmpp_handle* h) { return h ? h->n_stored : 0; }

int rbmpp_get_samples(const rbmpp_handle* h0, int32_t* x, double* psi) {
  rbmpp_handle* h = const_cast<rbmpp_handle*>(h0);
  if (!h) return RBMPP_E_ARG;
  const size_t n = (size_t)h->n_stored;
  if (x) {
    std::vector<signed char> xb(n * h->N);
    HIPCHK(h, hipMemcpy(xb.data(), h->d_xs, xb.size(), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < xb.size(); k++) x[k] = xb[k];
  }
  if (psi) HIPCHK(h, hipMemcpy(psi, h->d_psis, sizeof(double2) * n, hipMemcpyDeviceToHost));
  return RBMPP_OK;
}

int rbmpp_get_measured(const rbmpp_handle* h0, double* psi, double* eloc, double* rw) {
  rbmpp_handle* h = const_cast<rbmpp_handle*>(h0);
  if (!h) return RBMPP_E_ARG;
  const size_t n = (size_t)h->n_meas;
  if (n == 0 || n > h->meas_cap || !h->d_mpsi) return fail(h, RBMPP_E_STATE, "get_measured: no measurement");
  if (psi) HIPCHK(h, hipMemcpy(psi, h->d_mpsi, sizeof(double2) * n, hipMemcpyDeviceToHost));
  if (eloc) HIPCHK(h, hipMemcpy(eloc, h->d_eloc, sizeof(double2) * n, hipMemcpyDeviceToHost));
  if (rw) HIPCHK(h, hipMemcpy(rw, h->d_rw, sizeof(double) * n, hipMemcpyDeviceToHost));
  return RBMPP_OK;
}

int rbmpp_set_measured(rbmpp_handle* h, int64_t n, const double* psi, const double* eloc, const double* rw,
                       const double* orows, uint64_t seed) {
  if (!h || n <= 0 || n > (int64_t)(INT32_MAX / 2)) return RBMPP_E_ARG;
  if (orows && (!psi || !eloc || !rw)) return fail(h, RBMPP_E_ARG, "set_measured: psi, eloc and rw are needed with orows");
  int rc = ensure_measure(h, (size_t)n);
  if (rc) return rc;
  h->fused_valid = 0;
  const size_t rows = (size_t)n * (h->real_only ? 1 : 2);
  if (orows) {
    HIPCHK(h, hipMemcpyAsync(h->d_mpsi, psi, sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(h->d_eloc, eloc, sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(h->d_rw, rw, sizeof(double) * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(h->d_orows, orows, sizeof(double) * rows * h->nv, hipMemcpyHostToDevice, h->st));
  } else {
    if ((rc = ensure_fii(h))) return rc;
    hipLaunchKernelGGL(k_synth_measurement, dim3(4096), dim3(256), 0, h->st, n, h->nv, h->real_only ? 0 : 1, seed,
                       (const unsigned char*)h->d_fii, h->d_mpsi, h->d_eloc, h->d_rw, h->d_orows);
    HIPCHK(h, hipGetLastError());
  }
  h->n_meas = n;
  rc = reduce_measure(h, (int)n, true);
  if (rc) return rc;
  h->measured = 1;
  h->s_valid = 0;
  return RBMPP_OK;
}

int rbmpp_get_walkers(const rbmpp_handle* h0, int32_t* x, double* pf, double* psi, double* invm,
                      int64_t* naccept_per_walker) {
  rbmpp_handle* h = const_cast<rbmpp_handle*>(h0);
  if (!h) return RBMPP_E_ARG;
  const size_t W = h->W, N = h->N, C = h->C;
  if (x) {
    std::vector<signed char> xb(W * N);
    HIPCHK(h, hipMemcpy(xb.data(), h->ws.x, xb.size(), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < xb.size(); k++) x[k] = xb[k];
  }
  if (pf) HIPCHK(h, hipMemcpy(pf, h->ws.pf, sizeof(double) * W * C, hipMemcpyDeviceToHost));
  if (psi) HIPCHK(h, hipMemcpy(psi, h->ws.psi, sizeof(double2) * W, hipMemcpyDeviceToHost));
  if (invm) {
    // folded device slots ((d-1)*N + r holds InvM(r, (r+d) mod N), d = 1..N/2)
    // -> reference layout InvM(i,j) at j*N+i
    const size_t H = N / 2, FS = N * H;
    std::vector<double> m(W * C * FS);
    HIPCHK(h, hipMemcpy(m.data(), h->ws.minv, sizeof(double) * m.size(), hipMemcpyDeviceToHost));
    for (size_t b = 0; b < W * C; b++) {
      double* out = invm + b * N * N;
      if (h->minv_block) {
        // block form: F^-1(d, u) at d*H + u is InvM(u, H+d); InvM(H+d, u) is its
        // negative and the same-spin blocks are zero
        for (size_t k = 0; k < N * N; k++) out[k] = 0.0;
        for (size_t d = 0; d < H; d++)
          for (size_t u = 0; u < H; u++) {
            const double v = m[b * FS + d * H + u];
            out[(H + d) * N + u] = v;   // InvM(u, H+d)
            out[u * N + H + d] = -v;    // InvM(H+d, u)
          }
        continue;
      }
      for (size_t r = 0; r < N; r++) out[r * N + r] = 0.0;
      for (size_t d = 1; d <= H; d++)
        for (size_t r = 0; r < (d == H ? H : N); r++) {
          const size_t c2 = (r + d) % N;
          const double v = m[b * FS + (d - 1) * N + r];
          out[c2 * N + r] = v;   // InvM(r, c2)
          out[r * N + c2] = -v;  // InvM(c2, r)
        }
    }
  }
  if (naccept_per_walker) {
    std::vector<long long> v(W);
    HIPCHK(h, hipMemcpy(v.data(), h->ws.naccept, sizeof(long long) * W, hipMemcpyDeviceToHost));
    for (size_t w = 0; w < W; w++) naccept_per_walker[w] = v[w];
  }
  return RBMPP_OK;
}

int rbmpp_get_walker_electrons(const rbmpp_handle* h0, int32_t* ele_idx) {
  rbmpp_handle* h = const_cast<rbmpp_handle*>(h0);
  if (!h || !ele_idx) return RBMPP_E_ARG;
  if (!h->walkers_ready) return fail(h, RBMPP_E_STATE, "get_walker_electrons: walkers not initialised");
  const size_t n = (size_t)h->W * h->N;
  std::vector<int> e(n);
  int rc = copy_out(h, e.data(), h->ws.eidx, sizeof(int) * n);
  if (rc) return rc;
  for (size_t k = 0; k < n; k++) ele_idx[k] = e[k] + 1;  // 1-based sites, as Ele_Idx
  return RBMPP_OK;
}

int rbmpp_eval(rbmpp_handle* h, int32_t n, const int32_t* x, double* psi, double* pf, double* invm) {
  if (!h || !x || n <= 0) return RBMPP_E_ARG;
  std::vector<void*> tmp;
  signed char* d_x = nullptr;
  int rc = upload_x(h, n, x, &d_x, tmp);
  double *d_pf = nullptr, *d_inv = nullptr;
  double2* d_pq = nullptr;
  const size_t C = h->C, NN = (size_t)h->N * h->N;
  if (!rc) rc = tmp_alloc(h, &d_pf, (size_t)n * C, tmp);
  if (!rc) rc = tmp_alloc(h, &d_pq, (size_t)n * C, tmp);
  if (!rc && invm) rc = tmp_alloc(h, &d_inv, (size_t)n * C * NN, tmp);
  if (!rc) rc = run_local(h, n, d_x, invm ? 4 : 0, -1.0, nullptr, nullptr, d_inv, d_pf, d_pq, nullptr, nullptr);
  std::vector<double2> p0, pq;
  if (!rc) {
    p0.resize(n);
    pq.resize((size_t)n * C);
    rc = copy_out(h, p0.data(), h->d_mpsi, sizeof(double2) * n);
    if (!rc) rc = copy_out(h, pq.data(), d_pq, sizeof(double2) * n * C);
    if (!rc && pf) rc = copy_out(h, pf, d_pf, sizeof(double) * n * C);
    if (!rc && invm) rc = copy_out(h, invm, d_inv, sizeof(double) * n * C * NN);
  }
  if (!rc) {
    if (psi)
      for (int s = 0; s < n; s++) {
        psi[(size_t)s * 2 * (C + 1)] = p0[s].x;
        psi[(size_t)s * 2 * (C + 1) + 1] = p0[s].y;
        for (size_t q = 0; q < C; q++) {
          psi[(size_t)s * 2 * (C + 1) + 2 * (q + 1)] = pq[s * C + q].x;
          psi[(size_t)s * 2 * (C + 1) + 2 * (q + 1) + 1] = pq[s * C + q].y;
        }
      }
  }
  free_tmp(tmp);
  return rc;
}

int rbmpp_ratio_exchange(rbmpp_handle* h, int32_t n, const int32_t* x, const int32_t* i1, const int32_t* i2,
                         double* pf_new, double* psi_new) {
  if (!h || !x || !i1 || !i2 || n <= 0) return RBMPP_E_ARG;
  for (int s = 0; s < n; s++) {
    if (i1[s] < 0 || i1[s] >= h->N || i2[s] < 0 || i2[s] >= h->N)
      return fail(h, RBMPP_E_ARG, "exchange site out of range");
    if (x[(size_t)s * h->N + i1[s]] == x[(size_t)s * h->N + i2[s]])
      return fail(h, RBMPP_E_ARG, "exchange sites must carry opposite spins");
  }
  std::vector<void*> tmp;
  signed char* d_x = nullptr;
  int rc = upload_x(h, n, x, &d_x, tmp);
  int *d1 = nullptr, *d2 = nullptr;
  double* d_pf = nullptr;
  double2* d_ps = nullptr;
  const size_t C = h->C;
  if (!rc) rc = tmp_alloc(h, &d1, (size_t)n, tmp);
  if (!rc) rc = tmp_alloc(h, &d2, (size_t)n, tmp);
  if (!rc) rc = tmp_alloc(h, &d_pf, (size_t)n * C, tmp);
  if (!rc) rc = tmp_alloc(h, &d_ps, (size_t)n, tmp);
  if (!rc && hipMemcpyAsync(d1, i1, sizeof(int) * n, hipMemcpyHostToDevice, h->st) != hipSuccess)
    rc = fail(h, RBMPP_E_HIP, "ratio_exchange: copy of i1");
  if (!rc && hipMemcpyAsync(d2, i2, sizeof(int) * n, hipMemcpyHostToDevice, h->st) != hipSuccess)
    rc = fail(h, RBMPP_E_HIP, "ratio_exchange: copy of i2");
  if (!rc) rc = run_local(h, n, d_x, 8, -1.0, d1, d2, nullptr, nullptr, nullptr, d_ps, d_pf);
  if (!rc && pf_new) rc = copy_out(h, pf_new, d_pf, sizeof(double) * n * C);
  if (!rc && psi_new) rc = copy_out(h, psi_new, d_ps, sizeof(double2) * n);
  free_tmp(tmp);
  return rc;
}

int rbmpp_local(rbmpp_handle* h, int32_t n, const int32_t* x, double* eloc, double* psi, double* oloc) {
  if (!h || !x || n <= 0) return RBMPP_E_ARG;
  // the complex O layout is needed for the API even for real models
  const int saved = h->real_only;
  std::vector<void*> tmp;
  signed char* d_x = nullptr;
  int rc = upload_x(h, n, x, &d_x, tmp);
  if (!rc && oloc && saved && h->meas_cap < (size_t)2 * n) {
    // make room for 2 rows per configuration
    rc = ensure_measure(h, (size_t)2 * n);
  }
  if (!rc) {
    h->real_only = oloc ? 0 : saved;
    rc = run_local(h, n, d_x, 1 | (oloc ? 2 : 0), -1.0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                   nullptr);
    h->real_only = saved;
  }
  if (!rc && eloc) rc = copy_out(h, eloc, h->d_eloc, sizeof(double2) * n);
  if (!rc && psi) rc = copy_out(h, psi, h->d_mpsi, sizeof(double2) * n);
  if (!rc && oloc) {
    const size_t nv = h->nv;
    std::vector<double> rows(2 * nv * n);
    rc = copy_out(h, rows.data(), h->d_orows, sizeof(double) * rows.size());
    if (!rc)
      for (int s = 0; s < n; s++)
        for (size_t k = 0; k < nv; k++) {
          oloc[((size_t)s * nv + k) * 2] = rows[(2 * (size_t)s) * nv + k];
          oloc[((size_t)s * nv + k) * 2 + 1] = rows[(2 * (size_t)s + 1) * nv + k];
        }
  }
  h->measured = 0;
  free_tmp(tmp);
  return rc;
}

int rbmpp_local_noalloc(rbmpp_handle* h, int32_t n, const int32_t* x, double* eloc, double* psi) {
  if (!h || !x || n <= 0) return RBMPP_E_ARG;
  const int N = h->N;
  for (size_t k = 0; k < (size_t)n * N; k++)
    if (x[k] != 1 && x[k] != -1) return fail(h, RBMPP_E_ARG, "x must be +-1");
  for (int s = 0; s < n; s++) {
    int up = 0;
    for (int i = 0; i < N; i++) up += x[(size_t)s * N + i] == 1;
    if (up != h->npair) return fail(h, RBMPP_E_ARG, "configuration not in the Sz=0 sector");
  }
  unsigned char* base = (h->device >= 0 && h->device < NOALLOC_MAX_DEVICES) ? g_noalloc_ptr[h->device] : nullptr;
  if (!base) return fail(h, RBMPP_E_STATE, "local_noalloc: staging array of this device not set up");
  const DevModel hm = devmodel(h);
  // per-workgroup scratch of the estimator that will run: F^-1 of every copy
  // (block form) or the N x N recompute matrix (N > RBMPP_NLDS)
  const size_t np = (size_t)h->npair;
  const size_t wg = (hm.block && N <= RBMPP_NLDS) ? (size_t)h->C * np * np * sizeof(double)
                                                  : (N > RBMPP_NLDS ? (size_t)N * N * sizeof(double) : 0);
  const int gcap = wg ? (int)std::min<size_t>(1024, NOALLOC_WG_BYTES / wg) : 1024;
  // per configuration: x as int32 and int8, psi and eloc (complex), rw, status
  const size_t per_cfg = (size_t)N * 5 + 2 * sizeof(double2) + sizeof(double) + sizeof(int);
  const size_t fixed = 256 + round256((size_t)gcap * wg) + 7 * 256;
  const int64_t nmax = (int64_t)((NOALLOC_BYTES - fixed) / per_cfg);
  if ((int64_t)n > nmax)
    return fail(h, RBMPP_E_ARG, "local_noalloc: n = " + std::to_string(n) + " exceeds the staging limit of " +
                                    std::to_string(nmax) + " configurations at N = " + std::to_string(N));
  const int grid = std::min(n, gcap);
  size_t off = 0;
  int* d_hdr = (int*)(base + off); off += 256;
  double* d_wg = (double*)(base + off); off += round256((size_t)grid * wg);
  int32_t* d_x32 = (int32_t*)(base + off); off += round256((size_t)n * N * sizeof(int32_t));
  signed char* d_x8 = (signed char*)(base + off); off += round256((size_t)n * N);
  double2* d_psi = (double2*)(base + off); off += round256((size_t)n * sizeof(double2));
  double2* d_eloc = (double2*)(base + off); off += round256((size_t)n * sizeof(double2));
  double* d_rw = (double*)(base + off); off += round256((size_t)n * sizeof(double));
  int* d_st = (int*)(base + off); off += round256((size_t)n * sizeof(int));
  if (off > NOALLOC_BYTES) return fail(h, RBMPP_E_ARG, "local_noalloc: staging layout exceeds the array");
  // nothing in the array needs clearing: every workgroup writes its scratch
  // before reading it, every configuration's status and outputs are written
  // by the estimator, the header by k_first_status
  HIPCHK(h, hipMemcpyAsync(d_x32, x, (size_t)n * N * sizeof(int32_t), hipMemcpyHostToDevice, h->st));
  hipLaunchKernelGGL(k_x32_to_x8, dim3(256), dim3(256), 0, h->st, (size_t)n * N, (const int32_t*)d_x32, d_x8);
  HIPCHK(h, hipGetLastError());
  LocalArgs a;
  memset(&a, 0, sizeof a);
  a.m = h->d_model;
  a.nconf = n;
  a.xs = d_x8;
  a.mode = 1;       // local energy, amplitude (as rbmpp_local without O)
  a.shift = -1.0;   // not reweighted
  a.scratch = wg ? d_wg : nullptr;
  a.psi = d_psi;
  a.eloc = d_eloc;
  a.rw = d_rw;
  a.status = d_st;
  HIPCHK(h, hipMemcpyAsync(h->d_model, &hm, sizeof(DevModel), hipMemcpyHostToDevice, h->st));
  if (launch_local(hm, a, grid, h->st) != 0) return fail(h, RBMPP_E_ARG, "local_noalloc: unsupported size");
  HIPCHK(h, hipGetLastError());
  hipLaunchKernelGGL(k_first_status, dim3(1), dim3(256), 0, h->st, n, (const int*)d_st, d_hdr);
  HIPCHK(h, hipGetLastError());
  int hdr[2] = {0, 0};
  HIPCHK(h, hipMemcpyAsync(hdr, d_hdr, sizeof hdr, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (hdr[0] >= 0) {
    char buf[128];
    snprintf(buf, sizeof buf, "configuration %d: status %d", hdr[0], hdr[1]);
    return fail(h, hdr[1], buf);
  }
  if (eloc) HIPCHK(h, hipMemcpyAsync(eloc, d_eloc, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, h->st));
  if (psi) HIPCHK(h, hipMemcpyAsync(psi, d_psi, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RBMPP_OK;
}

int rbmpp_set_s_precision(rbmpp_handle* h, int32_t s_precision) {
  if (!h || s_precision < 0 || s_precision > 2) return RBMPP_E_ARG;
  h->cfg.s_precision = s_precision;
  h->s_valid = 0;
  return RBMPP_OK;
}

// ------------------------------------------------------------ tuning
int rbmpp_tune(rbmpp_handle* h, int32_t what, rbmpp_tune_report* rep) {
  if (!h) return RBMPP_E_ARG;
  if (rep) memset(rep, 0, sizeof *rep);
  int rc = 0;
  if (what & RBMPP_TUNE_S_BUILD) {
    if (!h->measured) return fail(h, RBMPP_E_STATE, "tune: the S build is tuned on a measurement");
    const int prec = h->cfg.s_precision, rows = h->real_only ? (int)h->n_meas : 2 * (int)h->n_meas;
    const SyrkPlan base = syrk_plan(prec, rows, h->nv, h->n_cu);
    std::vector<int> tiles = {128};
    if (prec == RBMPP_S_BF16X3) tiles.push_back(256);
    double best = 1e300;
    int bt = 0, bs = 0;
    for (int tile : tiles) {
      const SyrkPlan p0 = syrk_plan(prec, rows, h->nv, h->n_cu, tile);
      // split counts around the default (multiples of 8 for the XCD order)
      std::vector<int> splits = {1, p0.nsplit};
      for (int f : {8, 16, 32}) splits.push_back(f);
      for (int ns : splits) {
        const SyrkPlan p = syrk_plan(prec, rows, h->nv, h->n_cu, tile, ns);
        if (p.nsplit != ns && ns != p0.nsplit) continue;  // clamped to the rows: same as another candidate
        h->tune_tile[prec] = tile;
        h->tune_nsplit[prec] = p.nsplit;
        double t = 1e300;
        for (int r = 0; r < 2; r++) {
          if ((rc = build_s(h))) return rc;
          t = std::min(t, h->t[3]);
        }
        if (t < best) { best = t; bt = tile; bs = p.nsplit; }
      }
    }
    h->tune_tile[prec] = bt;
    h->tune_nsplit[prec] = bs;
    h->s_valid = 0;
    if (rep) {
      rep->s_tile = bt; rep->s_nsplit = bs; rep->s_build_s = best;
      rep->s_default_tile = base.tile; rep->s_default_nsplit = base.nsplit;
    }
  }
  if (what & RBMPP_TUNE_MEASURE) {
    if (h->n_stored <= 0) return fail(h, RBMPP_E_STATE, "tune: the estimator is tuned on stored samples");
    const int n = (int)std::min<int64_t>(h->n_stored, 32768);
    const bool was_measured = h->measured;
    double best = 1e300;
    int bg = 0;
    for (int grid : {256, 512, 1024, 2048, 4096}) {
      h->tune_meas_grid = grid;
      double t = 1e300;
      for (int r = 0; r < 2; r++) {
        tic(h);
        if ((rc = run_local(h, n, h->d_xs, 1 | 2, 0.0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr)))
          return rc;
        t = std::min(t, toc(h));
      }
      if (t < best) { best = t; bg = grid; }
    }
    h->tune_meas_grid = bg;
    // the measurement buffers now hold the tuning runs
    if (was_measured) h->measured = 0;
    if (rep) { rep->measure_grid = bg; rep->measure_s = best; rep->measure_configurations = n; }
  }
  return RBMPP_OK;
}

// ------------------------------------------------------------ batches
struct rbmpp_batch {
  std::vector<rbmpp_handle*> hs;
  std::vector<int> kind;
  SweepBatchEntry* d_ent = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int device = 0;
  double t_sample = 0;
  std::string err;
};

static int bfail(rbmpp_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

#define BHIPCHK(b, x)                                                          \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess)                                                     \
      return bfail(b, RBMPP_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

void rbmpp_batch_destroy(rbmpp_batch* b) {
  if (!b) return;
  if (b->st) hipStreamSynchronize(b->st);
  if (b->d_ent) hipFree(b->d_ent);
  if (b->e0) hipEventDestroy(b->e0);
  if (b->e1) hipEventDestroy(b->e1);
  if (b->st) hipStreamDestroy(b->st);
  delete b;
}

const char* rbmpp_batch_last_error(const rbmpp_batch* b) { return b ? b->err.c_str() : "null batch"; }

int rbmpp_batch_create(rbmpp_handle* const* handles, int32_t n, rbmpp_batch** out) {
  if (!handles || n < 1 || !out) return RBMPP_E_ARG;
  *out = nullptr;
  rbmpp_batch* b = new rbmpp_batch();
  *out = b;
  for (int i = 0; i < n; i++) {
    rbmpp_handle* h = handles[i];
    if (!h) return bfail(b, RBMPP_E_ARG, "null handle in batch");
    for (int j = 0; j < i; j++)
      if (handles[j] == h) return bfail(b, RBMPP_E_ARG, "a handle appears twice in the batch");
    if (i > 0 && h->device != handles[0]->device) return bfail(b, RBMPP_E_ARG, "batch handles on different devices");
    const int kind = sweep_kind(devmodel(h));
    if (kind < 0)
      return bfail(b, RBMPP_E_ARG, "handle " + std::to_string(i) +
                                       ": the batch covers the block form (normal projection, 16 copies) up to "
                                       "N = 256 and the skew form from N = 106 to 256");
    b->hs.push_back(h);
    b->kind.push_back(kind);
  }
  b->device = handles[0]->device;
  BHIPCHK(b, hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
  BHIPCHK(b, hipEventCreate(&b->e0));
  BHIPCHK(b, hipEventCreate(&b->e1));
  if (hipMalloc((void**)&b->d_ent, sizeof(SweepBatchEntry) * n) != hipSuccess) {
    b->d_ent = nullptr;
    (void)hipGetLastError();
    return bfail(b, RBMPP_E_NOMEM, "batch: hipMalloc of the entry table");
  }
  return RBMPP_OK;
}

int rbmpp_batch_make_sample(rbmpp_batch* b, const int32_t* first_iteration, const double* shift, int64_t* naccept) {
  if (!b || !first_iteration || !shift) return RBMPP_E_ARG;
  const int n = (int)b->hs.size();
  // the lattices with the most work per walker (N^3) first, so that their
  // walkers start in the first wave of workgroups; results do not depend on
  // the order (every walker draws from its own stream)
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->hs[x]->N > b->hs[y]->N; });
  std::vector<SweepBatchEntry> ent(n);
  int blocks = 0;
  for (int e = 0; e < n; e++) {
    const int i = order[e];
    rbmpp_handle* h = b->hs[i];
    bool fused = false;
    const int rc = sweep_args(h, first_iteration[i], shift[i], false, ent[e].a, fused);
    if (rc) return bfail(b, rc, "handle " + std::to_string(i) + ": " + h->err);
    ent[e].first_block = blocks;
    ent[e].kind = b->kind[i];
    blocks += h->W;
    // the model descriptor and walker state of each handle are current on
    // its own stream
    BHIPCHK(b, hipStreamSynchronize(h->st));
  }
  BHIPCHK(b, hipMemcpyAsync(b->d_ent, ent.data(), sizeof(SweepBatchEntry) * n, hipMemcpyHostToDevice, b->st));
  BHIPCHK(b, hipEventRecord(b->e0, b->st));
  launch_sweep_batch(b->d_ent, n, blocks, b->st);
  BHIPCHK(b, hipGetLastError());
  BHIPCHK(b, hipEventRecord(b->e1, b->st));
  BHIPCHK(b, hipEventSynchronize(b->e1));
  float ms = 0;
  hipEventElapsedTime(&ms, b->e0, b->e1);
  b->t_sample = ms * 1e-3;
  for (int i = 0; i < n; i++) {
    rbmpp_handle* h = b->hs[i];
    h->minv_block = devmodel(h).block;
    h->t[1] = b->t_sample;
    int64_t na = 0;
    const int rc = sweep_finish(h, false, shift[i], &na);
    if (rc) return bfail(b, rc, "handle " + std::to_string(i) + ": " + h->err);
    if (naccept) naccept[i] = na;
  }
  return RBMPP_OK;
}

double rbmpp_batch_sample_seconds(const rbmpp_batch* b) { return b ? b->t_sample : 0.0; }

int rbmpp_timings(const rbmpp_handle* h, double* t) {
  if (!h || !t) return RBMPP_E_ARG;
  for (int k = 0; k < 8; k++) t[k] = h->t[k];
  return RBMPP_OK;
}

}  // extern "C"
