// Split-bf16 products ("BF16X3") on gfx950 matrix cores.
//
// Each FP64 operand element is split into two bf16 numbers, x ~ hi + lo with
// hi = bf16(x) and lo = bf16(x - hi) (|x - hi - lo| <= ~2^-17 |x|), and each
// product term is formed as hi*hi + hi*lo + lo*hi (the lo*lo term, ~2^-16 of
// the product, is dropped) with v_mfma_f32_16x16x32_bf16, accumulated in FP32.
//
// A pack pass writes both splits of op(A) and op(B)^T contiguous along k, with
// rows padded to the 128-row tile and k padded to the 32-deep MFMA step (zero
// fill), so the product kernel has no bounds checks and a single operand
// orientation:  C(m, n) = alpha * sum_k Ap(m, k) Bp(n, k) + beta * C(m, n).
//
// MFMA 16x16x32 bf16 operand layout: lane l holds 8 consecutive k of row
// (A) or column (B) l % 16, starting at k = 8 (l / 16); the FP32 result
// element r of lane l is C(4 (l / 16) + r, l % 16).
#include <algorithm>
#include <string>

#include "gemm.hpp"

namespace etg {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kTile = 128;      // output tile (rows and columns)
constexpr int kDepth = 32;      // k per stage (one MFMA step)
constexpr int kLd = 40;         // LDS row stride in bf16 (80 B: 16-B aligned rows, skewed banks)
constexpr int kXcds = 8;
constexpr int kGroupM = 8;

// ---------------------------------------------------------------------------
// pack: P_hi / P_lo (rows_p x Kp, k fastest, per batch entry) from op(X)
//   kcontig:  element (r, k) = X[k + r ld]     (A with op = T, B with op = N)
//   else:     element (r, k) = X[r + k ld]
// ---------------------------------------------------------------------------
__global__ void pack_split_bf16_kernel(const double* __restrict__ X, int64_t ld, int64_t strideX,
                                       int kcontig, int64_t rows, int64_t K, int64_t rows_p,
                                       int64_t Kp, __bf16* __restrict__ hi, __bf16* __restrict__ lo) {
  const int64_t nchunk = Kp / 8;
  const int64_t total = rows_p * nchunk;
  const int64_t bz = blockIdx.z;
  const double* Xb = X + bz * strideX;
  __bf16* hb = hi + bz * rows_p * Kp;
  __bf16* lb = lo + bz * rows_p * Kp;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    int64_t r, kc;
    if (kcontig) { kc = idx % nchunk; r = idx / nchunk; } else { r = idx % rows_p; kc = idx / rows_p; }
    bf8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t k = 8 * kc + e;
      double x = 0.0;
      if (r < rows && k < K) x = kcontig ? Xb[k + r * ld] : Xb[r + k * ld];
      const __bf16 xh = __bf16(float(x));
      h[e] = xh;
      l[e] = __bf16(float(x - double(float(xh))));
    }
    *reinterpret_cast<bf8*>(hb + r * Kp + 8 * kc) = h;
    *reinterpret_cast<bf8*>(lb + r * Kp + 8 * kc) = l;
  }
}

struct Params {
  const __bf16 *Ahi, *Alo, *Bhi, *Blo;
  int64_t Kp, strideAp, strideBp;   // packed row length, per-batch element strides
  int64_t m, n;
  double alpha, beta;
  double* C;
  int64_t ldc, strideC;
  double* W;                         // split-K partials (m x n per split and batch entry)
  int64_t k_per_split;               // multiple of kDepth
  int tiles_m, tiles_n;
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

// 4 waves in 2 x 2, each a 64 x 64 sub-tile of 4 x 4 MFMA blocks.  One stage
// holds 32 k of both splits of a 128-row slab of each operand: 4 x 128 x 32
// bf16; stages are double buffered in LDS with one barrier per stage.  A
// stage's matrix-core work (48 MFMAs per wave) is short next to a load that
// misses L2, so global loads run two stages ahead through two register sets:
// the loads of step s + 2 are issued at step s.
__global__ __launch_bounds__(kThreads, 2) void gemm_bf16x3_kernel(Params p) {
  __shared__ __bf16 As[2][2][kTile * kLd];
  __shared__ __bf16 Bs[2][2][kTile * kLd];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int tile_m, tile_n;
  map_tile(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  const int64_t m0 = int64_t(tile_m) * kTile, n0 = int64_t(tile_n) * kTile;
  const int64_t bz = blockIdx.z, split = blockIdx.y;
  const int64_t kbeg = split * p.k_per_split;
  const int64_t kend = min(p.Kp, kbeg + p.k_per_split);
  const int64_t nsteps = (kend - kbeg) / kDepth;

  // global -> LDS: 1024 chunks of 16 B per operand and stage (2 splits x 128
  // rows x 4 chunks); thread tid moves chunks tid + 256 i, i = 0..3.
  // chunk i = 0..3 of this thread: split i / 2 (hi, lo), row (tid / 4) + 64 (i % 2),
  // 8 k from 8 (tid % 4): one per-thread offset, the rest uniform
  const int trow = tid >> 2, tk = 8 * (tid & 3);
  const int64_t toff = int64_t(trow) * p.Kp + tk;
  const __bf16* baseA[2] = {p.Ahi + bz * p.strideAp + m0 * p.Kp, p.Alo + bz * p.strideAp + m0 * p.Kp};
  const __bf16* baseB[2] = {p.Bhi + bz * p.strideBp + n0 * p.Kp, p.Blo + bz * p.strideBp + n0 * p.Kp};
  const int64_t half = 64 * p.Kp;
  const int tdst = trow * kLd + tk;
  u4 ra[2][4], rb[2][4];
  auto load = [&](int set, int64_t k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t o = toff + (i & 1) * half + k0;
      ra[set][i] = *reinterpret_cast<const u4*>(baseA[i >> 1] + o);
      rb[set][i] = *reinterpret_cast<const u4*>(baseB[i >> 1] + o);
    }
  };
  auto store = [&](int set, int buf) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int d = (i >> 1) * (kTile * kLd) + (i & 1) * (64 * kLd) + tdst;
      *reinterpret_cast<u4*>(&As[buf][0][0] + d) = ra[set][i];
      *reinterpret_cast<u4*>(&Bs[buf][0][0] + d) = rb[set][i];
    }
  };

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  const int frow = lane & 15, fk = 8 * (lane >> 4);
  auto compute = [&](int buf) {
    bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = (wm * 64 + i * 16 + frow) * kLd + fk;
      ah[i] = *reinterpret_cast<const bf8*>(&As[buf][0][r]);
      al[i] = *reinterpret_cast<const bf8*>(&As[buf][1][r]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = (wn * 64 + j * 16 + frow) * kLd + fk;
      bh[j] = *reinterpret_cast<const bf8*>(&Bs[buf][0][c]);
      bl[j] = *reinterpret_cast<const bf8*>(&Bs[buf][1][c]);
    }
    // three passes over the 16 blocks keep dependent products 16 MFMAs apart
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(ah[i], bh[j], acc[i][j]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(ah[i], bl[j], acc[i][j]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(al[i], bh[j], acc[i][j]);
  };
  // Step s computes on LDS buffer s % 2, issues the loads of step s + 2 into
  // register set s % 2 (whose slab is already in LDS) and stores set (s + 1) % 2,
  // loaded one step earlier, into the other buffer.  Loads past the last step
  // reload its slab: no branch on the loads.
  auto slab = [&](int64_t s) { return kbeg + min(s, nsteps - 1) * kDepth; };
  if (nsteps > 0) {
    load(0, slab(0));
    load(1, slab(1));
    store(0, 0);
  }
  __syncthreads();
  for (int64_t s = 0; s < nsteps; s += 2) {
    load(0, slab(s + 2));
    compute(0);
    store(1, 1);
    __syncthreads();
    if (s + 1 >= nsteps) break;
    load(1, slab(s + 3));
    compute(1);
    store(0, 0);
    __syncthreads();
  }

  const int lrow = lane >> 4, lcol = lane & 15;
  if (p.W) {
    double* W = p.W + (split + bz * gridDim.y) * p.m * p.n;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wm * 64 + i * 16 + 4 * lrow + r;
          const int64_t col = n0 + wn * 64 + j * 16 + lcol;
          if (row < p.m && col < p.n) W[row + col * p.m] = double(acc[i][j][r]);
        }
    return;
  }
  double* C = p.C + bz * p.strideC;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 64 + i * 16 + 4 * lrow + r;
        const int64_t col = n0 + wn * 64 + j * 16 + lcol;
        if (row < p.m && col < p.n) {
          double* c = C + row + col * p.ldc;
          const double v = p.alpha * double(acc[i][j][r]);
          *c = (p.beta == 0.0) ? v : v + p.beta * *c;
        }
      }
}

// C = alpha * sum_s W_s + beta * C
__global__ void splitk_reduce_bf(const double* __restrict__ W, int splits, int64_t m, int64_t n,
                                 double alpha, double beta, double* C, int64_t ldc, int64_t strideC) {
  const int64_t bz = blockIdx.z;
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  const int64_t row = idx % m, col = idx / m;
  const double* w = W + bz * splits * m * n + idx;
  double s = 0.0;
  for (int q = 0; q < splits; ++q) s += w[int64_t(q) * m * n];
  double* c = C + bz * strideC + row + col * ldc;
  const double v = alpha * s;
  *c = (beta == 0.0) ? v : v + beta * *c;
}

void pack(hipStream_t s, const double* X, int64_t ld, int64_t strideX, bool kcontig, int64_t rows,
          int64_t K, int64_t rows_p, int64_t Kp, int64_t batch, __bf16* hi, __bf16* lo) {
  const int64_t total = rows_p * (Kp / 8);
  dim3 grid(unsigned(std::min<int64_t>((total + 255) / 256, 1 << 16)), 1, unsigned(batch));
  pack_split_bf16_kernel<<<grid, 256, 0, s>>>(X, ld, strideX, kcontig ? 1 : 0, rows, K, rows_p, Kp,
                                              hi, lo);
  ETG_LAUNCH("pack_split_bf16");
}

}  // namespace

GemmLaunch gemm_bf16x3(hipStream_t stream, const Gemm& g, GemmWorkspace* ws) {
  if (ws == nullptr) fail("gemm: split-bf16 products need a workspace");
  const int64_t m_p = ceil_div(g.m, kTile) * kTile, n_p = ceil_div(g.n, kTile) * kTile;
  const int64_t Kp = ceil_div(g.k, kDepth) * kDepth;
  // an operand shared by every batch entry (stride 0) is packed once
  const int64_t a_batch = g.strideA == 0 ? 1 : g.batch, b_batch = g.strideB == 0 ? 1 : g.batch;
  const size_t a_elems = size_t(a_batch) * size_t(m_p) * size_t(Kp);
  const size_t b_elems = size_t(b_batch) * size_t(n_p) * size_t(Kp);
  // A: reuse a pack made under the same caller tag for the same operand
  PackedOperand* hit = nullptr;
  if (g.a_tag != 0)
    for (PackedOperand& e : ws->cached)
      if (e.tag == g.a_tag && e.A == g.A && e.ta == g.ta && e.m == g.m && e.k == g.k &&
          e.lda == g.lda && e.strideA == g.strideA && e.batch == g.batch)
        hit = &e;
  __bf16* Ahi;
  if (hit) {
    hit->used = ++ws->clock;
    Ahi = hit->buf.as<__bf16>();
  } else {
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
    Ahi = dst->as<__bf16>();
    pack(stream, g.A, g.lda, g.strideA, g.ta, g.m, g.k, m_p, Kp, a_batch, Ahi, Ahi + a_elems);
  }
  __bf16* Alo = Ahi + a_elems;
  if (ws->packB.resize(2 * b_elems * sizeof(__bf16))) ++ws->gen;
  __bf16* Bhi = ws->packB.as<__bf16>();
  __bf16* Blo = Bhi + b_elems;
  pack(stream, g.B, g.ldb, g.strideB, !g.tb, g.n, g.k, n_p, Kp, b_batch, Bhi, Blo);

  Params p{};
  p.Ahi = Ahi; p.Alo = Alo; p.Bhi = Bhi; p.Blo = Blo;
  p.Kp = Kp;
  p.strideAp = g.strideA == 0 ? 0 : m_p * Kp;
  p.strideBp = g.strideB == 0 ? 0 : n_p * Kp;
  p.m = g.m; p.n = g.n; p.alpha = g.alpha; p.beta = g.beta;
  p.C = g.C; p.ldc = g.ldc; p.strideC = g.strideC;
  p.tiles_m = int(m_p / kTile);
  p.tiles_n = int(n_p / kTile);
  p.k_per_split = Kp;
  // split K when the output has few tiles (two workgroups per CU on 256 CUs)
  const int64_t tiles = int64_t(p.tiles_m) * p.tiles_n * g.batch;
  int splits = 1;
  if (tiles < 1024 && Kp >= 4 * 512) {
    splits = int(std::min<int64_t>(ceil_div(1024, tiles), Kp / 512));
    splits = std::max(1, std::min(splits, 1024));
    const int64_t kps = ceil_div(ceil_div(Kp, splits), kDepth) * kDepth;
    splits = int(ceil_div(Kp, kps));
    p.k_per_split = kps;
    if (ws->buf.resize(size_t(splits) * size_t(g.m) * size_t(g.n) * size_t(g.batch) * sizeof(double)))
      ++ws->gen;
    p.W = ws->buf.as<double>();
  }
  dim3 grid(unsigned(tiles / g.batch), unsigned(splits), unsigned(g.batch));
  gemm_bf16x3_kernel<<<grid, kThreads, 0, stream>>>(p);
  ETG_LAUNCH("gemm_bf16x3");
  if (splits > 1) {
    const int64_t mn = g.m * g.n;
    dim3 rgrid(unsigned(ceil_div(mn, 256)), 1, unsigned(g.batch));
    splitk_reduce_bf<<<rgrid, 256, 0, stream>>>(p.W, splits, g.m, g.n, g.alpha, g.beta, g.C, g.ldc,
                                                g.strideC);
    ETG_LAUNCH("splitk_reduce");
  }
  GemmLaunch info;
  info.kernel = "gemm_bf16x3<128x128>";
  info.splits = splits;
  return info;
}

}  // namespace etg
