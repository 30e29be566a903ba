// Split-bf16 products ("BF16X3") on gfx950 matrix cores.
//
// Each FP64 operand element is split into two bf16 numbers, x ~ hi + lo with
// hi = bf16(x) and lo = bf16(x - hi) (|x - hi - lo| <= ~2^-17 |x|), and each
// product term is formed as hi*hi + hi*lo + lo*hi (the lo*lo term, ~2^-16 of
// the product, is dropped) with v_mfma_f32_16x16x32_bf16, accumulated in FP32.
//
// A pack pass writes both splits of op(A) and op(B)^T in the order the matrix
// cores consume them, rows padded to the 256-row tile and k to the 32-deep
// MFMA step (zero fill), so the product kernel has no bounds checks and one
// operand orientation:  C(m, n) = alpha * sum_k Ap(m, k) Bp(n, k) + beta * C(m, n).
//
// MFMA 16x16x32 bf16 operand layout: lane l holds 8 consecutive k of row
// (A) or column (B) l % 16, starting at k = 8 (l / 16); the FP32 result
// element r of lane l is C(4 (l / 16) + r, l % 16).
#include <algorithm>
#include <string>

#include "gemm.hpp"
#include "ladder_order.hpp"

namespace etg {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int kTile = 256;      // output tile (rows and columns)
constexpr int kDepth = 32;      // k per step (one MFMA step)
constexpr int kXcds = 8;
constexpr int kGroupM = 8;

// ---------------------------------------------------------------------------
// pack: both splits of op(X) (rows_p x Kp per batch entry) in the order the
// matrix cores consume them: for row block rb (16 rows) and k step ks (32 k)
// one 2 KB chunk [hi: 64 lanes x 8 bf16][lo: 64 lanes x 8 bf16], lane =
// r % 16 + 16 ((k % 32) / 8), element k % 8; chunks ordered (rb, ks) with ks
// fastest.  A fragment is then one contiguous 1 KB block (16 B per lane).
//   kcontig:  element (r, k) = X[k + r ld]     (A with op = T, B with op = N)
//   else:     element (r, k) = X[r + k ld]
// ---------------------------------------------------------------------------
// One workgroup packs a tile of 64 rows (4 row blocks) x 32 k (one step): each
// thread reads 8 consecutive k of one row (kcontig: 4 threads span a row's 256
// B; otherwise 64 threads span 64 rows of one k), splits them and places hi and
// lo in the tile's four chunks in LDS; the chunks then leave as contiguous
// 2 KB blocks.
__global__ __launch_bounds__(256) void pack_split_bf16_kernel(const double* __restrict__ X, int64_t ld,
                                                              int64_t strideX, int kcontig, int64_t rows,
                                                              int64_t K, int64_t rows_p, int64_t Kp,
                                                              __bf16* __restrict__ P) {
  __shared__ u4 tile[4 * 128];       // 4 chunks of 2 KB
  const int64_t KS = Kp / kDepth;
  const int64_t ks = blockIdx.x;
  const int64_t bz = blockIdx.z;
  const double* Xb = X + bz * strideX;
  u4* Pb = reinterpret_cast<u4*>(P + bz * rows_p * Kp * 2);
  const int t = threadIdx.x;
  const int row = kcontig ? (t >> 2) : (t & 63);
  const int kq = kcontig ? (t & 3) : (t >> 6);
  const int lane = (row & 15) + 16 * kq;
  // tiles of 64 rows: blockIdx.y, then every gridDim.y-th one (more than 65535 tiles)
  for (int64_t tr = blockIdx.y; tr < rows_p / 64; tr += gridDim.y) {
    const int64_t rb0 = tr * 4;
    const int64_t r = rb0 * 16 + row;
    bf8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t k = ks * kDepth + 8 * kq + e;
      double x = 0.0;
      if (r < rows && k < K) x = kcontig ? Xb[k + r * ld] : Xb[r + k * ld];
      const __bf16 xh = __bf16(float(x));
      h[e] = xh;
      l[e] = __bf16(float(x - double(float(xh))));
    }
    __syncthreads();                             // the previous tile has left LDS
    tile[(row >> 4) * 128 + lane] = __builtin_bit_cast(u4, h);
    tile[(row >> 4) * 128 + 64 + lane] = __builtin_bit_cast(u4, l);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int q = t + 256 * i;                 // u4 unit of the tile: chunk q / 128
      Pb[((rb0 + (q >> 7)) * KS + ks) * 128 + (q & 127)] = tile[q];
    }
  }
}

// ---------------------------------------------------------------------------
// The ladder's g+- written straight as split-bf16 packs.  For the pairs k =
// (a, b) of a G block, M_k(c, d) = G(c + v a, d + v b) gives
// g+_cd = M(c, d) + M(d, c) and g-_cd = M(c, d) - M(d, c); the packs are the B
// operands (rows = pairs, k = the tile-blocked (c, d) index of
// ladder_order.hpp) of the products w+- = t+- g+-.  One workgroup per (16 x 16
// tile pair ti <= tj of the (c, d) plane, 16 pairs): it stages the tile and
// its mirror for its 16 pairs in LDS (both read along the contiguous index),
// then writes whole 2 KB chunks of both packs: a tile is a whole number of
// 32-deep steps, and the 16 pairs are one 16-row block.  G comes from a
// split-bf16 product in FP32 (its accumulation type, so nothing is lost); the
// splits are those of the pack kernel applied to the FP64 sums.
// ---------------------------------------------------------------------------
constexpr int kPairTile = 16 * 17 + 1;   // floats per staged pair tile (odd strides)

__global__ __launch_bounds__(256) void ladder_g_pm_split_kernel(
    const float* __restrict__ G, int64_t n_v, int64_t ldG, const int32_t* __restrict__ pa,
    const int32_t* __restrict__ pb, int64_t npairs, int64_t KSP, int64_t KSM, int64_t row_blocks,
    u4* __restrict__ Pp, u4* __restrict__ Pm) {
  // T[p][x][y] = M_k(d0 + y, c0 + x): the mirrored tile (for a diagonal tile, the tile itself);
  // D[p][x][y] = M_k(c0 + y, d0 + x): the tile (off-diagonal tiles only)
  __shared__ float T[16 * kPairTile];
  __shared__ float D[16 * kPairTile];
  const int t = threadIdx.x;
  const int tj = small_tri_root(int(blockIdx.x)), ti = int(blockIdx.x) - tj * (tj + 1) / 2;
  const int64_t c0 = int64_t(kLadderTile) * ti, d0 = int64_t(kLadderTile) * tj;
  const bool diag = ti == tj;
  const int w = int(ladder_col_width(tj, n_v));
  const int64_t ksp0 = ladder_tile_offset(ti, tj, n_v, true) / kDepth;
  const int64_t ksm0 = ladder_tile_offset(ti, tj, n_v, false) / kDepth;
  // 32-deep steps of the tile in g+ and g-
  const int nsp = int(diag ? ladder_diag_size(w, true) / kDepth : w / 2);
  const int nsm = int(diag ? ladder_diag_size(w, false) / kDepth : w / 2);
  for (int64_t rb = blockIdx.y; rb < row_blocks; rb += gridDim.y) {
    __syncthreads();                     // the previous row block has left LDS
    {
      const int x = t >> 4, y = t & 15;
#pragma unroll 4
      for (int p = 0; p < 16; ++p) {
        const int64_t k = rb * 16 + p;
        float val = 0.0f, vd = 0.0f;
        if (k < npairs) {
          const float* M = G + n_v * pa[k] + ldG * n_v * pb[k];   // M(c, d) = M[c + ldG d]
          if (d0 + y < n_v && c0 + x < n_v) val = M[(d0 + y) + ldG * (c0 + x)];
          if (!diag && c0 + y < n_v && d0 + x < n_v) vd = M[(c0 + y) + ldG * (d0 + x)];
        }
        T[p * kPairTile + x * 17 + y] = val;
        if (!diag) D[p * kPairTile + x * 17 + y] = vd;
      }
    }
    __syncthreads();
    // item: (sign, step ks of the tile, lane slot ls = p + 16 q): 8 consecutive
    // blocked indices 32 ks + 8 q + e of pair p, one 16 B lane of the chunk
    const float* Dx = diag ? T : D;      // x1 = M(c, d) = Dx[p][d'][c']
    for (int i = t; i < (nsp + nsm) * 64; i += 256) {
      const bool plus = i < nsp * 64;
      const int j = plus ? i : i - nsp * 64;
      const int ks = j >> 6, ls = j & 63, p = ls & 15, q = ls >> 4;
      const bool row_ok = rb * 16 + p < npairs;
      bf8 h, l;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        int cl, dl;
        const bool ok = ladder_tile_pair(32 * ks + 8 * q + e, diag, w, plus, cl, dl);
        double g = 0.0;
        if (ok && row_ok && c0 + cl < n_v && d0 + dl < n_v) {
          const double x1 = Dx[p * kPairTile + dl * 17 + cl];  // M(c, d)
          const double x2 = T[p * kPairTile + cl * 17 + dl];   // M(d, c)
          g = plus ? x1 + x2 : x1 - x2;
        }
        const __bf16 gh = __bf16(float(g));
        h[e] = gh;
        l[e] = __bf16(float(g - double(float(gh))));
      }
      u4* P = plus ? Pp : Pm;
      const int64_t chunk = rb * (plus ? KSP : KSM) + (plus ? ksp0 : ksm0) + ks;
      P[chunk * 128 + ls] = __builtin_bit_cast(u4, h);
      P[chunk * 128 + 64 + ls] = __builtin_bit_cast(u4, l);
    }
  }
}

struct Params {
  const __bf16 *A, *B;               // fragment-ordered packs
  int64_t KS;                        // k steps (of 32) in a packed row block
  int64_t strideAp, strideBp;        // per-batch element strides of the packs
  int64_t m, n;
  double alpha, beta;
  double* C;
  float* Cf;                         // C in FP32 instead (beta = 0)
  int64_t ldc, strideC;
  double* W;                         // split-K partials (m x n per split and batch entry)
  int64_t ks_per_split;
  int tiles_m, tiles_n;
  int vec_rows;                      // W / C / Cf rows written four at a time (alignment checked on the host)
};

__device__ inline void map_tile(int bid, int tiles_m, int tiles_n, int& tm, int& tn) {
  const int nb = tiles_m * tiles_n;
  const int per = nb / kXcds, rem = nb % kXcds;
  const int xcd = bid % kXcds, idx = bid / kXcds;
  const int logical = (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
  const int group = kGroupM * tiles_n;
  const int first_m = (logical / group) * kGroupM;
  const int gsize = min(tiles_m - first_m, kGroupM);
  tm = first_m + (logical % group) % gsize;
  tn = (logical % group) / gsize;
}

__device__ inline f4 mfma(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Workgroup tile 256 x 256, 8 waves in 2 x 4 (two per SIMD), each 128 x 64 of
// the output (8 x 4 blocks; FP32 accumulators in VGPRs).  Both operand slabs
// of a 32-k step (16 row blocks x 2 splits each: 64 chunks of 1 KB) go from
// global memory straight into LDS (buffer_load_dwordx4 with the lds bit: lane
// l's 16 B land at the chunk's LDS base + 16 l), double buffered (128 KB).
// The copies use buffer (MUBUF) rather than global instructions: the compiler
// treats a global-to-LDS copy in flight like a flat access and then waits for
// every outstanding LDS read (lgkmcnt(0)) wherever one result is needed, while
// after a buffer copy it counts the reads in order.  A wave's step is 96 MFMAs
// (1536 cycles) against 24 KB of LDS fragment reads: all four B fragments
// first, then the A fragments of one row block ahead of the 12 products that
// use them, so the products start after the first reads land and each later
// pair of reads has a row block of products to arrive.  Per output element and
// k the workgroup moves 1/64 operand bytes.
// ---------------------------------------------------------------------------
constexpr int kProductThreads = 512;

__global__ __launch_bounds__(kProductThreads, 1) void gemm_bf16x3_kernel(Params p) {
  __shared__ u4 lds[2][4096];        // per buffer: A chunks 0..31, B chunks 32..63 (1 KB each)

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;     // rows 128 wm.., columns 64 wn..
  int tile_m, tile_n;
  map_tile(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  const int64_t bz = blockIdx.z, split = blockIdx.y;
  const int64_t ks0 = split * p.ks_per_split;
  const int64_t ks1 = min(p.KS, ks0 + p.ks_per_split);
  const int64_t nsteps = ks1 - ks0;
  const int64_t astride = p.KS * 128;          // u4 units from one row block to the next

  // waves 0-3 copy A, waves 4-7 copy B: 8 chunks each per step, c = 8 (wave % 4)
  // + q: row block c / 2, split c % 2 of the operand's 32 chunks.  Chunk
  // addresses: a wave-uniform buffer base (this wave's first chunk of the tile)
  // + a 32-bit offset (the host keeps 4 row blocks of a pack below 2 GB).
  const bool copies_a = wave < 4;
  const int cbase = 8 * (wave & 3);
  const u4* src = reinterpret_cast<const u4*>(copies_a ? p.A + bz * p.strideAp : p.B + bz * p.strideBp) +
                  int64_t(copies_a ? tile_m : tile_n) * 16 * astride + (cbase >> 1) * astride;
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<u4*>(src), (short)0, 0x7fffffff, 0x00020000);
  auto load = [&](int buf, int64_t ks) {
    u4* dst = &lds[buf][(copies_a ? 0 : 2048) + cbase * 64];
    const uint32_t koff = uint32_t(ks) * 2048u + uint32_t(lane) * 16u;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t off = uint32_t(((q >> 1) * astride + (q & 1) * 64) * 16) + koff;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void*)(dst + q * 64), 16, off, 0, 0, 0);
    }
  };

  f4 acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  // Per step: the wave's four B fragments (hi, lo) and row block 0's A
  // fragments, then row block by row block the next block's two A reads ahead
  // of the current block's 12 products (hi hi, hi lo, lo hi over the four
  // columns; four independent accumulators between dependent products).  The
  // scheduling groups pin that order (the compiler otherwise issues the reads
  // just before their use); measured 2-10 % faster than reading all 24
  // fragments up front.
  auto compute = [&](int buf) {
    const u4* As = &lds[buf][0];
    const u4* Bs = &lds[buf][2048];
    bf8 ah[2], al[2], bh[4], bl[4];
    auto rd_b = [&](int j, int sp) {
      return __builtin_bit_cast(bf8, Bs[((wn * 4 + j) * 2 + sp) * 64 + lane]);
    };
    auto rd_a = [&](int i, int sp) {
      return __builtin_bit_cast(bf8, As[((wm * 8 + i) * 2 + sp) * 64 + lane]);
    };
    ah[0] = rd_a(0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) bh[j] = rd_b(j, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) bl[j] = rd_b(j, 1);
    al[0] = rd_a(0, 1);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int c = i & 1;
      if (i + 1 < 8) {
        ah[c ^ 1] = rd_a(i + 1, 0);
        al[c ^ 1] = rd_a(i + 1, 1);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(ah[c], bh[j], acc[i][j]);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(ah[c], bl[j], acc[i][j]);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(al[c], bh[j], acc[i][j]);
      // groups: LDS reads (0x100), then matrix-core products (0x008)
      if (i == 0) __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);
      else if (i + 1 < 8) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 12, 0);
      __builtin_amdgcn_sched_barrier(0);
    }
  };

  // Step s computes from buffer s % 2 while the slabs of step s + 1 land in the
  // other buffer (issued at the start of step s, after the barrier that ended
  // every wave's reads of that buffer).
  if (nsteps > 0) load(0, ks0);
  for (int64_t s = 0; s < nsteps; ++s) {
    __builtin_amdgcn_s_waitcnt(0);             // this wave's copies of step s have landed
    __syncthreads();                           // everyone's have; buffer (s + 1) % 2 is free
    if (s + 1 < nsteps) load(int((s + 1) & 1), ks0 + s + 1);
    compute(int(s & 1));
  }

  const int lrow = lane >> 4, lcol = lane & 15;
  const int64_t r0 = int64_t(tile_m) * kTile + wm * 128, c0 = int64_t(tile_n) * kTile + wn * 64;
  if (p.W) {
    double* W = p.W + (split + bz * gridDim.y) * p.m * p.n;
    if (p.vec_rows) {
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t row = r0 + i * 16 + 4 * lrow;
          const int64_t col = c0 + j * 16 + lcol;
          if (row < p.m && col < p.n) {
            double2* w = reinterpret_cast<double2*>(W + row + col * p.m);
            w[0] = double2{double(acc[i][j][0]), double(acc[i][j][1])};
            w[1] = double2{double(acc[i][j][2]), double(acc[i][j][3])};
          }
        }
      return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = r0 + i * 16 + 4 * lrow + r;
          const int64_t col = c0 + j * 16 + lcol;
          if (row < p.m && col < p.n) W[row + col * p.m] = double(acc[i][j][r]);
        }
    return;
  }
  if (p.Cf) {
    float* C = p.Cf + bz * p.strideC;
    if (p.vec_rows) {
      // a lane's four accumulators of a block are four consecutive rows of one column
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t row = r0 + i * 16 + 4 * lrow;
          const int64_t col = c0 + j * 16 + lcol;
          if (row < p.m && col < p.n) {
            f4 v;
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = float(p.alpha * double(acc[i][j][r]));
            *reinterpret_cast<f4*>(C + row + col * p.ldc) = v;
          }
        }
      return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = r0 + i * 16 + 4 * lrow + r;
          const int64_t col = c0 + j * 16 + lcol;
          if (row < p.m && col < p.n) C[row + col * p.ldc] = float(p.alpha * double(acc[i][j][r]));
        }
    return;
  }
  double* C = p.C + bz * p.strideC;
  if (p.vec_rows) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t row = r0 + i * 16 + 4 * lrow;
        const int64_t col = c0 + j * 16 + lcol;
        if (row < p.m && col < p.n) {
          double2* c = reinterpret_cast<double2*>(C + row + col * p.ldc);
          double2 v0 = {p.alpha * double(acc[i][j][0]), p.alpha * double(acc[i][j][1])};
          double2 v1 = {p.alpha * double(acc[i][j][2]), p.alpha * double(acc[i][j][3])};
          if (p.beta != 0.0) {
            const double2 o0 = c[0], o1 = c[1];
            v0.x += p.beta * o0.x; v0.y += p.beta * o0.y;
            v1.x += p.beta * o1.x; v1.y += p.beta * o1.y;
          }
          c[0] = v0;
          c[1] = v1;
        }
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = r0 + i * 16 + 4 * lrow + r;
        const int64_t col = c0 + j * 16 + lcol;
        if (row < p.m && col < p.n) {
          double* c = C + row + col * p.ldc;
          const double v = p.alpha * double(acc[i][j][r]);
          *c = (p.beta == 0.0) ? v : v + p.beta * *c;
        }
      }
}

// C = alpha * sum_s W_s + beta * C
__global__ void splitk_reduce_bf(const double* __restrict__ W, int splits, int64_t m, int64_t n,
                                 double alpha, double beta, double* C, float* Cf, int64_t ldc,
                                 int64_t strideC) {
  const int64_t bz = blockIdx.z;
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  const int64_t row = idx % m, col = idx / m;
  const double* w = W + bz * splits * m * n + idx;
  double s = 0.0;
  for (int q = 0; q < splits; ++q) s += w[int64_t(q) * m * n];
  const double v = alpha * s;
  if (Cf) {
    Cf[bz * strideC + row + col * ldc] = float(v);
    return;
  }
  double* c = C + bz * strideC + row + col * ldc;
  *c = (beta == 0.0) ? v : v + beta * *c;
}

void pack(hipStream_t s, const double* X, int64_t ld, int64_t strideX, bool kcontig, int64_t rows,
          int64_t K, int64_t rows_p, int64_t Kp, int64_t batch, __bf16* P) {
  if (rows_p % 64 != 0 || batch > 65535 || Kp / kDepth > 0x7fffffff)
    fail("split-bf16 pack: operand shape outside the pack grid");
  dim3 grid(unsigned(Kp / kDepth), unsigned(std::min<int64_t>(rows_p / 64, 65535)), unsigned(batch));
  pack_split_bf16_kernel<<<grid, 256, 0, s>>>(X, ld, strideX, kcontig ? 1 : 0, rows, K, rows_p, Kp, P);
  ETG_LAUNCH("pack_split_bf16");
}

// the packed operand A of g: a cached pack under g's tag, or a new pack
__bf16* packed_a(hipStream_t stream, const Gemm& g, GemmWorkspace* ws, int64_t rows_p, int64_t Kp) {
  const int64_t a_batch = g.strideA == 0 ? 1 : g.batch;
  const size_t a_elems = size_t(a_batch) * size_t(rows_p) * size_t(Kp);
  if (g.a_tag != 0)
    for (PackedOperand& e : ws->cached)
      if (e.tag == g.a_tag && e.A == g.A && e.ta == g.ta && e.m == g.m && e.k == g.k &&
          e.lda == g.lda && e.strideA == g.strideA && e.batch == g.batch) {
        e.used = ++ws->clock;
        return e.buf.as<__bf16>();
      }
  DeviceBuffer* dst = &ws->packA;
  if (g.a_tag != 0) {
    PackedOperand* lru = &ws->cached[0];
    for (PackedOperand& e : ws->cached)
      if (e.used < lru->used) lru = &e;
    PackedOperand& e = *lru;
    e.used = ++ws->clock;
    e.tag = g.a_tag; e.A = g.A; e.ta = g.ta; e.m = g.m; e.k = g.k;
    e.lda = g.lda; e.strideA = g.strideA; e.batch = g.batch;
    dst = &e.buf;
  }
  if (dst->resize(2 * a_elems * sizeof(__bf16))) ++ws->gen;
  __bf16* P = dst->as<__bf16>();
  pack(stream, g.A, g.lda, g.strideA, g.ta, g.m, g.k, rows_p, Kp, a_batch, P);
  return P;
}

}  // namespace

int64_t split_bf16_pack_elems(int64_t rows, int64_t k) {
  return 2 * ceil_div(rows, kTile) * kTile * ceil_div(k, kDepth) * kDepth;
}

void ladder_g_pm_split(hipStream_t s, const float* G, int64_t n_v, int64_t lenA, const int32_t* pair_a,
                       const int32_t* pair_b, int64_t npairs, __bf16* Pp, __bf16* Pm) {
  if (npairs <= 0) return;
  const int64_t nt = ladder_tiles(n_v);
  const int64_t row_blocks = ceil_div(npairs, kTile) * kTile / 16;
  dim3 grid(unsigned(nt * (nt + 1) / 2), unsigned(std::min<int64_t>(row_blocks, 65535)));
  ladder_g_pm_split_kernel<<<grid, 256, 0, s>>>(G, n_v, n_v * lenA, pair_a, pair_b, npairs,
                                                ladder_blocked_len(n_v, true) / kDepth,
                                                ladder_blocked_len(n_v, false) / kDepth, row_blocks,
                                                reinterpret_cast<u4*>(Pp), reinterpret_cast<u4*>(Pm));
  ETG_LAUNCH("ladder_g_pm_split");
}

GemmLaunch gemm_bf16x3(hipStream_t stream, const Gemm& g, GemmWorkspace* ws) {
  if (ws == nullptr) fail("gemm: split-bf16 products need a workspace");
  const int64_t m_p = ceil_div(g.m, kTile) * kTile, n_p = ceil_div(g.n, kTile) * kTile;
  const int64_t Kp = ceil_div(g.k, kDepth) * kDepth;
  const int64_t KS = Kp / kDepth;
  const int64_t b_batch = g.strideB == 0 ? 1 : g.batch;
  __bf16* Ap = packed_a(stream, g, ws, m_p, Kp);
  const __bf16* Bp = static_cast<const __bf16*>(g.b_packed);
  if (Bp == nullptr) {
    const size_t b_elems = size_t(b_batch) * size_t(n_p) * size_t(Kp);
    if (ws->packB.resize(2 * b_elems * sizeof(__bf16))) ++ws->gen;
    __bf16* P = ws->packB.as<__bf16>();
    pack(stream, g.B, g.ldb, g.strideB, !g.tb, g.n, g.k, n_p, Kp, b_batch, P);
    Bp = P;
  } else if (g.strideB != 0) {
    fail("gemm: a prepacked B operand is shared by the whole batch");
  }

  // the kernel's copy offsets are 32-bit: four row blocks of a pack below 2 GB
  if (KS > (int64_t(1) << 18) - 64) fail("gemm: split-bf16 product deeper than the copy offsets reach");
  Params p{};
  p.A = Ap; p.B = Bp;
  p.KS = KS;
  p.strideAp = g.strideA == 0 ? 0 : 2 * m_p * Kp;
  p.strideBp = g.strideB == 0 ? 0 : 2 * n_p * Kp;
  p.m = g.m; p.n = g.n; p.alpha = g.alpha; p.beta = g.beta;
  p.C = g.C; p.Cf = g.Cf; p.ldc = g.ldc; p.strideC = g.strideC;
  p.tiles_m = int(m_p / kTile);
  p.tiles_n = int(n_p / kTile);
  p.ks_per_split = KS;
  // One workgroup per CU.  Split K when the output has few tiles: the smallest
  // split count whose rounds of 256 workgroups are at least 90 % full (at most
  // 256 splits of at least 16 k steps each; partials of at most 256 tiles).
  const int64_t tiles = int64_t(p.tiles_m) * p.tiles_n * g.batch;
  int splits = 1;
  if (tiles < 2048 && KS >= 32) {
    const int64_t cus = 256, max_s = std::min<int64_t>(256, KS / 16);
    int best = 1;
    double best_eff = 0.0;
    for (int64_t s = 1; s <= max_s; ++s) {
      const double wgs = double(tiles * s);
      const double eff = (wgs / cus) / double(ceil_div(tiles * s, cus));
      if (eff > best_eff + 1e-9) { best_eff = eff; best = int(s); }
      if (eff >= 0.9) { best = int(s); break; }
    }
    splits = best;
  }
  if (splits > 1) {
    const int64_t kps = ceil_div(KS, splits);
    splits = int(ceil_div(KS, kps));
    p.ks_per_split = kps;
    if (ws->buf.resize(size_t(splits) * size_t(g.m) * size_t(g.n) * size_t(g.batch) * sizeof(double)))
      ++ws->gen;
    p.W = ws->buf.as<double>();
  }
  // rows four at a time: m, ldc and the batch stride multiples of 4 elements and
  // the output 16 B (FP32) or 32 B (FP64) aligned; split-K partials (ld m, in the
  // 256 B aligned workspace) need m a multiple of 4
  if (splits > 1) {
    p.vec_rows = g.m % 4 == 0 && reinterpret_cast<uintptr_t>(p.W) % 32 == 0;
  } else {
    const uintptr_t base = g.Cf ? reinterpret_cast<uintptr_t>(g.Cf) : reinterpret_cast<uintptr_t>(g.C);
    p.vec_rows = g.m % 4 == 0 && g.ldc % 4 == 0 && (g.batch == 1 || g.strideC % 4 == 0) &&
                 base % (g.Cf ? 16 : 32) == 0;
  }
  dim3 grid(unsigned(tiles / g.batch), unsigned(splits), unsigned(g.batch));
  gemm_bf16x3_kernel<<<grid, kProductThreads, 0, stream>>>(p);
  ETG_LAUNCH("gemm_bf16x3");
  if (splits > 1) {
    const int64_t mn = g.m * g.n;
    dim3 rgrid(unsigned(ceil_div(mn, 256)), 1, unsigned(g.batch));
    splitk_reduce_bf<<<rgrid, 256, 0, stream>>>(p.W, splits, g.m, g.n, g.alpha, g.beta, g.C, g.Cf, g.ldc,
                                                g.strideC);
    ETG_LAUNCH("splitk_reduce");
  }
  GemmLaunch info;
  info.kernel = "gemm_bf16x3<256x256>";
  info.splits = splits;
  return info;
}

}  // namespace etg
