// S build of the SR step on the matrix cores (SR_pdposv.f90:11-15 with the
// dgemm of calc_phys_quantity.f90:240): S = sum_r c_r c_r^T / norm over the
// centred O rows c_r = O_r - kappa_r obar, one symmetric triangle.
//
// k_center_pack centres every row once in FP64 and stores the result in the
// precision of the products, transposed into K-blocks so that each column
// contributes 128 contiguous bytes per block:
//   MODE 0 (FP64):      16 rows per block, 16 doubles per column;
//   MODE 1 (FP32):      32 rows per block, 32 floats per column;
//   MODE 2 (BF16 x 3):  32 rows per block, 32 bf16 hi then 32 bf16 lo, with
//                       c = hi + lo, hi = bf16(c), lo = bf16(c - hi).
// The packed matrix is P[kb][col][128 B] with col < nvp (nv rounded up to the
// 128-column tile, padding zero) and the last block's missing rows zero.
//
// k_syrk_packed computes the 128 x 128 upper-triangle tiles of S over a range
// of K-blocks (split-K).  One workgroup of 4 waves per tile and split, each
// wave a 64 x 64 quadrant of 16 x 16 MFMA blocks; a K-block of both column
// panels (2 x 16 KB) is staged in LDS, double buffered, while the next one is
// loaded.  Per lane the MFMA operands are contiguous along k (a double4, two
// float4 or a bf16x8 pair), so FP64 and FP32 products use the lane's four k
// values over four MFMAs (k slot s of lane l holds k = 4 (l/16) + s for A and
// B alike).  FP32 accumulators are added into FP64 ones every FLUSH (128) K-blocks;
// split partials are summed in FP64, in a fixed order, by k_syrk_reduce.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbmpp_sr.h"

namespace {

constexpr int TILE = 128;      // output tile edge
constexpr int COLB = 128;      // bytes per column per K-block
constexpr int LDP = COLB + 16; // LDS pitch of a column (bank spread)
#ifndef RBMPP_SYRK_FLUSH
#define RBMPP_SYRK_FLUSH 128
#endif
constexpr int FLUSH = RBMPP_SYRK_FLUSH;  // K-blocks between FP32 -> FP64 flushes

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int kblock_rows(int mode) { return mode == 0 ? 16 : 32; }

// ------------------------------------------------------------ centre + pack
// Workgroup (chunk, column block) takes 128 columns of the K-blocks
// [chunk * per, (chunk + 1) * per): rows are read coalesced along the
// columns (the next block's into registers while the current one is
// packed), centred in FP64 into LDS, then written out column by column (128
// contiguous bytes each).  With DIAG (the mixed precisions) it also sums the
// squares of the FP64 centred values per column over its K-blocks, in a
// fixed order, into dpart[chunk * (nv + 1) + col]: the FP64 diagonal of S
// without another pass over the O rows.
template <int MODE, bool DIAG>
__global__ __launch_bounds__(256) void k_center_pack(int nrows, int nv, int nvp, int nkb, int per,
                                                     const double* __restrict__ orows,
                                                     const double* __restrict__ kappa,
                                                     const double* __restrict__ obar, unsigned char* __restrict__ P,
                                                     double* __restrict__ dpart) {
  constexpr int KB = kblock_rows(MODE);
  constexpr int NL = KB * TILE / 256;  // tile elements loaded per thread
  __shared__ double t[KB][TILE + 1];
  __shared__ double dsum[2][TILE];
  const int c0 = blockIdx.x * TILE, chunk = blockIdx.y;
  const int tid = threadIdx.x;
  const int kb0 = chunk * per, kb1 = min(nkb, kb0 + per);
  // element e = tid + 256 i of a tile: row e / TILE, column e % TILE
  const int lc = tid & (TILE - 1), lr = tid >> 7;
  const int col = c0 + lc;
  const double ob = col < nv ? obar[col] : 0.0;
  double nxt[NL];
  auto load = [&](int kb) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int row = kb * KB + lr + 2 * i;
      nxt[i] = (row < nrows && col < nv) ? orows[(size_t)row * nv + col] - kappa[row] * ob : 0.0;
    }
  };
  double dacc = 0.0;  // column lc, rows lr, lr + 2, ... of every block
  if (kb0 < kb1) load(kb0);
  for (int kb = kb0; kb < kb1; kb++) {
#pragma unroll
    for (int i = 0; i < NL; i++) t[lr + 2 * i][lc] = nxt[i];
    if constexpr (DIAG) {
#pragma unroll
      for (int i = 0; i < NL; i++) dacc = fma(nxt[i], nxt[i], dacc);
    }
    __syncthreads();
    if (kb + 1 < kb1) load(kb + 1);
    unsigned char* base = P + ((size_t)kb * nvp + c0) * COLB;
    // 128 columns x 8 chunks of 16 bytes
    for (int e = tid; e < TILE * 8; e += 256) {
      const int c = e >> 3, q = e & 7;
      u32x4 w;
      if constexpr (MODE == 0) {
        const double d0 = t[2 * q][c], d1 = t[2 * q + 1][c];
        w = u32x4{(unsigned)__double_as_longlong(d0), (unsigned)(__double_as_longlong(d0) >> 32),
                  (unsigned)__double_as_longlong(d1), (unsigned)(__double_as_longlong(d1) >> 32)};
      } else if constexpr (MODE == 1) {
        w = u32x4{__float_as_uint((float)t[4 * q][c]), __float_as_uint((float)t[4 * q + 1][c]),
                  __float_as_uint((float)t[4 * q + 2][c]), __float_as_uint((float)t[4 * q + 3][c])};
      } else {
        // chunks 0..3: hi of k = 8q .. 8q+7; chunks 4..7: lo of k = 8(q-4) ..
        const int k0 = 8 * (q & 3);
        unsigned short h[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const double v = t[k0 + j][c];
          const __bf16 hi = (__bf16)(float)v;
          const __bf16 sel = (q < 4) ? hi : (__bf16)(float)(v - (double)(float)hi);
          h[j] = __builtin_bit_cast(unsigned short, sel);
        }
        w = u32x4{(unsigned)h[0] | ((unsigned)h[1] << 16), (unsigned)h[2] | ((unsigned)h[3] << 16),
                  (unsigned)h[4] | ((unsigned)h[5] << 16), (unsigned)h[6] | ((unsigned)h[7] << 16)};
      }
      *(u32x4*)(base + (size_t)c * COLB + 16 * q) = w;
    }
    __syncthreads();  // t is rewritten by the next block
  }
  if constexpr (DIAG) {
    dsum[lr][lc] = dacc;
    __syncthreads();
    if (tid < TILE && col < nv) dpart[(size_t)chunk * (nv + 1) + col] = dsum[0][tid] + dsum[1][tid];
  }
}

// ------------------------------------------------------------------ SYRK
struct SyrkSmem {
  unsigned char p[2][2][TILE * LDP];  // [stage][panel][column * LDP + byte]
};

template <int MODE>
__global__ __launch_bounds__(256, 1) void k_syrk_packed(const unsigned char* __restrict__ P, int nkb, int nvp,
                                                        int ntile, int npair, int nsplit, int kb_per_split,
                                                        int nlogical, double* __restrict__ part) {
  __shared__ SyrkSmem L;
  // XCD-aware order: hardware workgroup g runs on XCD g % 8; logical
  // workgroups [x T/8, (x+1) T/8) go to XCD x, so that one XCD's L2 serves
  // whole splits (all tile pairs of the same rows)
  const int g = blockIdx.x, per = gridDim.x >> 3;
  const int lg = (g & 7) * per + (g >> 3);
  if (lg >= nlogical) return;
  const int pair = lg % npair, split = lg / npair;
  int p = pair, ti = 0;
  while (p >= ntile - ti) { p -= ntile - ti; ti++; }
  const int tj = ti + p;
  const bool diag = ti == tj;
  const int kb0 = split * kb_per_split, kb1 = min(nkb, kb0 + kb_per_split);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  const int npan = diag ? 1 : 2;

  f64x4 acc[4][4];
  f32x4 a32[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
      a32[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
    }

  // global -> register staging of one K-block: 4 chunks of 16 B per panel
  u32x4 st[2][4];
  auto gload = [&](int kb) {
#pragma unroll
    for (int pn = 0; pn < 2; pn++) {
      if (pn < npan) {
        const int col0 = (pn == 0 ? ti : tj) * TILE;
        const unsigned char* src = P + ((size_t)kb * nvp + col0) * COLB;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int q = tid + 256 * u;  // column q >> 3, chunk q & 7
          st[pn][u] = *(const u32x4*)(src + (size_t)q * 16);
        }
      }
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int pn = 0; pn < 2; pn++) {
      if (pn < npan) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int q = tid + 256 * u;
          *(u32x4*)&L.p[buf][pn][(q >> 3) * LDP + 16 * (q & 7)] = st[pn][u];
        }
      }
    }
  };

  const int ia = wi * 64 + (lane & 15), jb = wj * 64 + (lane & 15);
  const int ks = lane >> 4;  // k slot of the lane
  int nflush = 0;
  if (kb0 < kb1) {
    gload(kb0);
    lstore(0);
  }
  __syncthreads();
  int buf = 0;
  for (int kb = kb0; kb < kb1; kb++) {
    const bool more = kb + 1 < kb1;
    if (more) gload(kb + 1);  // in flight during the products
    const unsigned char* pa = L.p[buf][0];
    const unsigned char* pb = L.p[buf][diag ? 0 : 1];
    if constexpr (MODE == 0) {
      // column c: 16 doubles; lane holds k = 4 ks .. 4 ks + 3
      f64x4 av[4], bv[4];
#pragma unroll
      for (int x = 0; x < 4; x++) {
        av[x] = *(const f64x4*)(pa + (ia + 16 * x) * LDP + 32 * ks);
        bv[x] = *(const f64x4*)(pb + (jb + 16 * x) * LDP + 32 * ks);
      }
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
          for (int y = 0; y < 4; y++)
            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x][s], bv[y][s], acc[x][y], 0, 0, 0);
    } else if constexpr (MODE == 1) {
      // column c: 32 floats; lane holds k = 4 ks .. +3 and 16 + 4 ks .. +3
#pragma unroll
      for (int h = 0; h < 2; h++) {
        f32x4 av[4], bv[4];
#pragma unroll
        for (int x = 0; x < 4; x++) {
          av[x] = *(const f32x4*)(pa + (ia + 16 * x) * LDP + 64 * h + 16 * ks);
          bv[x] = *(const f32x4*)(pb + (jb + 16 * x) * LDP + 64 * h + 16 * ks);
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++)
              a32[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x][s], bv[y][s], a32[x][y], 0, 0, 0);
      }
    } else {
      // column c: 32 bf16 hi then 32 lo; lane holds k = 8 ks .. 8 ks + 7
      bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int x = 0; x < 4; x++) {
        ah[x] = *(const bf16x8*)(pa + (ia + 16 * x) * LDP + 16 * ks);
        al[x] = *(const bf16x8*)(pa + (ia + 16 * x) * LDP + 64 + 16 * ks);
        bh[x] = *(const bf16x8*)(pb + (jb + 16 * x) * LDP + 16 * ks);
        bl[x] = *(const bf16x8*)(pb + (jb + 16 * x) * LDP + 64 + 16 * ks);
      }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
          a32[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], a32[x][y], 0, 0, 0);
          a32[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], a32[x][y], 0, 0, 0);
          a32[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], a32[x][y], 0, 0, 0);
        }
    }
    if constexpr (MODE != 0) {
      if (++nflush == FLUSH) {
        nflush = 0;
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
          for (int y = 0; y < 4; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              acc[x][y][r] += (double)a32[x][y][r];
              a32[x][y][r] = 0.f;
            }
      }
    }
    if (more) lstore(buf ^ 1);  // the other buffer was last read before the previous barrier
    __syncthreads();
    buf ^= 1;
  }
  if constexpr (MODE != 0) {
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[x][y][r] += (double)a32[x][y][r];
  }
  // partial tile [split][pair][128 x 128]: FP64 MFMA lane l, register r ->
  // (l/16 + 4 r, l % 16); FP32 / BF16 -> (4 (l/16) + r, l % 16)
  double* out = part + ((size_t)split * npair + pair) * TILE * TILE;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = wi * 64 + 16 * x + (MODE == 0 ? ks + 4 * r : 4 * ks + r);
        const int j = wj * 64 + 16 * y + (lane & 15);
        out[i * TILE + j] = acc[x][y][r];
      }
}

// Split BF16 x 3 on 256 x 256 tiles: 8 waves of 64 x 128 (4 x 8 blocks),
// twice the products per staged byte of the 128 x 128 kernel.  The FP32
// accumulators (128 per lane) leave no room for FP64 ones, so every FLUSH
// K-blocks they are added into the workgroup's own FP64 partial tile in
// global memory (read-modify-write by the lane that owns each entry; the
// first flush writes).
constexpr int T2 = 256;
struct SyrkSmem256 {
  unsigned char p[2][2][T2 * LDP];
};

__global__ __launch_bounds__(512, 1) void k_syrk_bf16_256(const unsigned char* __restrict__ P, int nkb, int nvp,
                                                          int ntile, int npair, int nsplit, int kb_per_split,
                                                          int nlogical, double* __restrict__ part) {
  __shared__ SyrkSmem256 L;
  const int g = blockIdx.x, per = gridDim.x >> 3;
  const int lg = (g & 7) * per + (g >> 3);
  if (lg >= nlogical) return;
  const int pair = lg % npair, split = lg / npair;
  int p = pair, ti = 0;
  while (p >= ntile - ti) { p -= ntile - ti; ti++; }
  const int tj = ti + p;
  const bool diag = ti == tj;
  const int kb0 = split * kb_per_split, kb1 = min(nkb, kb0 + kb_per_split);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave >> 1, wj = wave & 1;  // rows wi * 64, columns wj * 128
  const int npan = diag ? 1 : 2;
  f32x4 a32[4][8];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 8; y++) a32[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
  u32x4 st[2][4];
  auto gload = [&](int kb) {
#pragma unroll
    for (int pn = 0; pn < 2; pn++) {
      if (pn < npan) {
        const unsigned char* src = P + ((size_t)kb * nvp + (pn == 0 ? ti : tj) * T2) * COLB;
#pragma unroll
        for (int u = 0; u < 4; u++) st[pn][u] = *(const u32x4*)(src + (size_t)(tid + 512 * u) * 16);
      }
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int pn = 0; pn < 2; pn++) {
      if (pn < npan) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int q = tid + 512 * u;
          *(u32x4*)&L.p[buf][pn][(q >> 3) * LDP + 16 * (q & 7)] = st[pn][u];
        }
      }
    }
  };
  // the partial tile through a buffer resource: one per-lane byte offset,
  // the (x, y, r) part of each address a scalar constant
  const __amdgpu_buffer_rsrc_t prs = block_rsrc(part + ((size_t)split * npair + pair) * T2 * T2, T2 * T2 * 8);
  const int m = lane & 15, ks = lane >> 4;
  const int vbase = ((wi * 64 + 4 * ks) * T2 + wj * 128 + m) * 8;
  bool written = false;
  auto flush = [&]() {
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 8; y++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int soff = ((16 * x + r) * T2 + 16 * y) * 8;
          double v = (double)a32[x][y][r];
          if (written) v += bload1(prs, vbase, soff);
          bstore1(v, prs, vbase, soff);
          a32[x][y][r] = 0.f;
        }
    written = true;
  };
  const int ia = wi * 64 + m, jb = wj * 128 + m;
  int nflush = 0;
  if (kb0 < kb1) {
    gload(kb0);
    lstore(0);
  }
  __syncthreads();
  int buf = 0;
  for (int kb = kb0; kb < kb1; kb++) {
    const bool more = kb + 1 < kb1;
    if (more) gload(kb + 1);
    const unsigned char* pa = L.p[buf][0];
    const unsigned char* pb = L.p[buf][diag ? 0 : 1];
    bf16x8 ah[4], al[4];
#pragma unroll
    for (int x = 0; x < 4; x++) {
      ah[x] = *(const bf16x8*)(pa + (ia + 16 * x) * LDP + 16 * ks);
      al[x] = *(const bf16x8*)(pa + (ia + 16 * x) * LDP + 64 + 16 * ks);
    }
#pragma unroll
    for (int yh = 0; yh < 4; yh++) {
      bf16x8 bh[2], bl[2];
#pragma unroll
      for (int y = 0; y < 2; y++) {
        bh[y] = *(const bf16x8*)(pb + (jb + 16 * (2 * yh + y)) * LDP + 16 * ks);
        bl[y] = *(const bf16x8*)(pb + (jb + 16 * (2 * yh + y)) * LDP + 64 + 16 * ks);
      }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
          f32x4& c = a32[x][2 * yh + y];
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], c, 0, 0, 0);
        }
    }
    if (++nflush == FLUSH) {
      nflush = 0;
      flush();
    }
    if (more) lstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  if (nflush > 0 || !written) flush();
}

// Sum of the split partials of each tile pair (fixed order), scaled by
// 1/norm, written to both triangles of S (row-major nv x nv).  In diagonal
// tiles only the entries with i <= j are used, so S is exactly symmetric.
__global__ void k_syrk_reduce(int nv, int ts, int ntile, int npair, int nsplit, const double* __restrict__ part,
                              double inv_norm, double* __restrict__ S) {
  int p = blockIdx.x, ti = 0;
  while (p >= ntile - ti) { p -= ntile - ti; ti++; }
  const int tj = ti + p;
  const size_t tt = (size_t)ts * ts;
  for (int e = threadIdx.x; e < ts * ts; e += blockDim.x) {
    const int a = e / ts, b = e - a * ts;
    const int i = ti * ts + a, j = tj * ts + b;
    if (i >= nv || j >= nv || i > j) continue;
    double t = 0;
    for (int z = 0; z < nsplit; z++) t += part[((size_t)z * npair + blockIdx.x) * tt + e];
    t *= inv_norm;
    S[(size_t)i * nv + j] = t;
    S[(size_t)j * nv + i] = t;
  }
}

}  // namespace

// ---------------------------------------------------------------- host side
SyrkPlan syrk_plan(int precision, int nrows, int nv, int cus, int tile, int nsplit_want) {
  SyrkPlan pl;
  pl.mode = precision;
  pl.kb = kblock_rows(precision);
  // 256 x 256 tiles (split BF16 x 3 only) once the 128-tile grid has several
  // tile pairs per CU
  if (tile <= 0) tile = (precision == 2 && nv >= 1536) ? T2 : TILE;
  if (precision != 2) tile = TILE;
  pl.tile = tile;
  pl.nvp = (nv + tile - 1) / tile * tile;
  pl.ntile = pl.nvp / tile;
  pl.npair = pl.ntile * (pl.ntile + 1) / 2;
  pl.nkb = (nrows + pl.kb - 1) / pl.kb;
  // about two workgroups per CU, a multiple of 8 splits (one set per XCD)
  // unless the tile pairs alone fill the chip; at least 64 K-blocks a split
  const int per_cu = tile == T2 ? 1 : 2;  // resident workgroups per CU
  int nsplit = (per_cu * cus + pl.npair - 1) / pl.npair;
  if (nsplit > 1) nsplit = (nsplit + 7) / 8 * 8;
  if (nsplit_want > 0) nsplit = nsplit_want;
  nsplit = std::max(1, std::min(nsplit, pl.nkb / 64));
  pl.kb_per_split = (pl.nkb + nsplit - 1) / nsplit;
  pl.nsplit = (pl.nkb + pl.kb_per_split - 1) / pl.kb_per_split;
  if (pl.nsplit < 1) pl.nsplit = 1;
  pl.packed_bytes = (size_t)pl.nkb * pl.nvp * COLB;
  pl.part_doubles = (size_t)pl.nsplit * pl.npair * tile * tile;
  return pl;
}

void launch_center_pack(const SyrkPlan& pl, int nrows, int nv, const double* orows, const double* kappa,
                        const double* obar, unsigned char* packed, double* dpart, hipStream_t st) {
  // SR_ROW_CHUNKS chunks of K-blocks (every chunk writes its diagonal
  // partials, zero for an empty one)
  const int per = (pl.nkb + SR_ROW_CHUNKS - 1) / SR_ROW_CHUNKS;
  const dim3 grid(pl.nvp / TILE, SR_ROW_CHUNKS);
  if (pl.mode == 2)
    hipLaunchKernelGGL((k_center_pack<2, true>), grid, dim3(256), 0, st, nrows, nv, pl.nvp, pl.nkb, per, orows,
                       kappa, obar, packed, dpart);
  else if (pl.mode == 1)
    hipLaunchKernelGGL((k_center_pack<1, true>), grid, dim3(256), 0, st, nrows, nv, pl.nvp, pl.nkb, per, orows,
                       kappa, obar, packed, dpart);
  else
    hipLaunchKernelGGL((k_center_pack<0, false>), grid, dim3(256), 0, st, nrows, nv, pl.nvp, pl.nkb, per, orows,
                       kappa, obar, packed, dpart);
}

void launch_syrk_packed(const SyrkPlan& pl, int nv, const unsigned char* packed, double* part, double inv_norm,
                        double* S, hipStream_t st) {
  const int nlogical = pl.npair * pl.nsplit;
  const int ngrid = (nlogical + 7) / 8 * 8;
  if (pl.tile == T2)
    hipLaunchKernelGGL(k_syrk_bf16_256, dim3(ngrid), dim3(512), 0, st, packed, pl.nkb, pl.nvp, pl.ntile, pl.npair,
                       pl.nsplit, pl.kb_per_split, nlogical, part);
  else if (pl.mode == 2)
    hipLaunchKernelGGL(k_syrk_packed<2>, dim3(ngrid), dim3(256), 0, st, packed, pl.nkb, pl.nvp, pl.ntile, pl.npair,
                       pl.nsplit, pl.kb_per_split, nlogical, part);
  else if (pl.mode == 1)
    hipLaunchKernelGGL(k_syrk_packed<1>, dim3(ngrid), dim3(256), 0, st, packed, pl.nkb, pl.nvp, pl.ntile, pl.npair,
                       pl.nsplit, pl.kb_per_split, nlogical, part);
  else
    hipLaunchKernelGGL(k_syrk_packed<0>, dim3(ngrid), dim3(256), 0, st, packed, pl.nkb, pl.nvp, pl.ntile, pl.npair,
                       pl.nsplit, pl.kb_per_split, nlogical, part);
  hipLaunchKernelGGL(k_syrk_reduce, dim3(pl.npair), dim3(256), 0, st, nv, pl.tile, pl.ntile, pl.npair, pl.nsplit,
                     part, inv_norm, S);
}
