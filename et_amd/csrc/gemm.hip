// Matrix-core products for gfx950 (CDNA4).
//
// One workgroup of 256 threads (4 wave64s in a 2 x 2 arrangement) computes a
// BM x BN tile of C.  Operand tiles of BK columns are staged through LDS with
// double buffering: the global loads for step k+1 are issued before the
// matrix-core work of step k and written to the other LDS buffer after it, so
// a single barrier per step suffices.  Each wave owns a (BM/2) x (BN/2)
// sub-tile made of 16 x 16 MFMA blocks.
//
// MFMA 16x16x4 operand layout (one element per lane, f64 and f32 alike):
//   A[i][k] in lane i + 16 k,  B[k][j] in lane j + 16 k.
// Result D (element r of the lane's 4-vector, column lane % 16) differs:
//   f32: row 4 (lane/16) + r        f64: row (lane/16) + 4 r
//
// Tile order: workgroups are dispatched round-robin over the 8 XCDs, so the
// linear block id is first remapped to give each XCD a contiguous range of
// tiles, and that range is walked in column groups of GROUP_M tile-rows so
// that tiles sharing A rows / B columns run close together in time on the
// same L2.
#include "gemm.hpp"

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace etg {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kXcds = 8;
constexpr int kGroupM = 8;

struct Params {
  int64_t m, n, k;
  double alpha, beta;
  const double* A;
  int64_t lda, strideA;
  const double* B;
  int64_t ldb, strideB;
  double* C;
  int64_t ldc, strideC;
  // split-K: partial sums go to W (m x n per split, ld = m) when W != nullptr
  double* W;
  int64_t k_per_split;
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

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef d4 acc_t;
  static __device__ inline acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline double cvt(double x) { return x; }
  static __device__ inline int row(int lrow, int r) { return lrow + 4 * r; }
};

template <>
struct Mfma<float> {
  typedef f4 acc_t;
  static __device__ inline acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline float cvt(double x) { return static_cast<float>(x); }
  static __device__ inline int row(int lrow, int r) { return 4 * lrow + r; }
};

// Raw buffer access (gfx9 descriptor word 3).  Offsets at or past kOutOfRange
// are beyond the record count and read as 0.
constexpr uint32_t kOutOfRange = 0x80000000u;
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const double* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, int(kOutOfRange - 1),
                                           0x00020000);
}
__device__ inline double load_f64(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
}

// LDS tile orientation follows the operand's contiguous dimension in HBM so
// that both the global loads and the LDS stores are unit-stride:
//   contiguous along k  -> tile[mn][k], leading dim BK + 1 (odd stride: the 16
//                          lanes of a fragment read hit distinct banks);
//   contiguous along mn -> tile[k][mn], leading dim BM + 128 B (adjacent k rows of
//                          a fragment read land on opposite bank halves).
template <typename T, bool KCONTIG, int BMN, int BK>
struct TileLayout {
  static constexpr int LD = KCONTIG ? BK + 1 : BMN + int(128 / sizeof(T));
  static constexpr int SIZE = KCONTIG ? BMN * LD : BK * LD;
  static __device__ inline int at(int mn, int k) { return KCONTIG ? mn * LD + k : k * LD + mn; }
};

// T: arithmetic type of the products (double or float); operands are FP64 in memory.
// Two workgroups per CU: the FP64 128 x 128 tile fits its 128 accumulator
// registers and operand staging in 256 VGPRs and 72 KB of LDS, so one
// workgroup's barrier and LDS refill overlap the other's matrix-core work.
template <typename T, bool TA, bool TB, int BM, int BN, int BK, bool WIDE>
__global__ __launch_bounds__(kThreads, 2) void gemm_mfma16(Params p) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int WM = BM / 2, WN = BN / 2;  // wave sub-tile
  constexpr int TMs = WM / 16, TNs = WN / 16;
  // op(A) = A^T means A is stored k x m: contiguous along k.  op(B) = B means k x n: along k.
  using LA = TileLayout<T, TA, BM, BK>;
  using LB = TileLayout<T, !TB, BN, BK>;
  constexpr int A_PER = BM * BK / kThreads, B_PER = BN * BK / kThreads;
  static_assert(A_PER * kThreads == BM * BK && B_PER * kThreads == BN * BK, "tile/threads");

  __shared__ T As[2][LA::SIZE];
  __shared__ T Bs[2][LB::SIZE];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int tile_m, tile_n;
  map_tile(blockIdx.x, p.tiles_m, p.tiles_n, tile_m, tile_n);
  const int64_t m0 = int64_t(tile_m) * BM, n0 = int64_t(tile_n) * BN;
  const int64_t bz = blockIdx.z;
  const int64_t split = blockIdx.y;
  const double* A = p.A + bz * p.strideA;
  const double* B = p.B + bz * p.strideB;
  const int64_t kbeg = split * p.k_per_split;
  const int64_t kend = min(p.k, kbeg + p.k_per_split);

  T ra[A_PER], rb[B_PER];

  // Element r of this thread's share of a tile sits at (mn, k) inside the tile.
  // Tiles are read with buffer loads: an element outside the matrix or beyond
  // the split's k range gets an offset past the buffer's record count and the
  // hardware returns 0 for it.  Every load is unconditional, so the k loop has
  // no branches and the accumulators stay in place across iterations.
  // Byte offsets are 31-bit.  With WIDE = false each operand has one buffer
  // base per k step and a per-element offset (leading dimensions < 2^21);
  // with WIDE = true element r gets its own base, r * SA (SB) operand columns
  // after element 0, and the per-thread offset spans at most 16 columns
  // (leading dimensions < 2^24) at the cost of more scalar work per step.
  auto a_coord = [&](int r, int& mm, int& kk) {
    const int idx = tid + r * kThreads;
    if (TA) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
  };
  auto b_coord = [&](int r, int& nn, int& kk) {
    const int idx = tid + r * kThreads;
    if (TB) { nn = idx % BN; kk = idx / BN; } else { kk = idx % BK; nn = idx / BK; }
  };
  constexpr int SA = TA ? kThreads / BK : kThreads / BM;    // operand columns between elements
  constexpr int SB = TB ? kThreads / BN : kThreads / BK;
  static_assert(kThreads % BK == 0 && kThreads % BM == 0 && kThreads % BN == 0, "tile/threads");
  const uint32_t lda = uint32_t(p.lda), ldb = uint32_t(p.ldb);
  uint32_t a_off[A_PER], b_off[B_PER];   // in-range byte offsets (kOutOfRange outside the matrix)
#pragma unroll
  for (int r = 0; r < A_PER; ++r) {
    int mm, kk;
    a_coord(r, mm, kk);
    const int mo = WIDE ? mm - r * (TA ? SA : 0) : mm, ko = WIDE ? kk - r * (TA ? 0 : SA) : kk;
    a_off[r] = (m0 + mm < p.m) ? 8u * (TA ? uint32_t(mo) * lda + ko : uint32_t(ko) * lda + mo)
                               : kOutOfRange;
  }
#pragma unroll
  for (int r = 0; r < B_PER; ++r) {
    int nn, kk;
    b_coord(r, nn, kk);
    const int no = WIDE ? nn - r * (TB ? 0 : SB) : nn, ko = WIDE ? kk - r * (TB ? SB : 0) : kk;
    b_off[r] = (n0 + nn < p.n) ? 8u * (TB ? uint32_t(ko) * ldb + no : uint32_t(no) * ldb + ko)
                               : kOutOfRange;
  }
  const double* A0 = A + (TA ? m0 * p.lda : m0);
  const double* B0 = B + (TB ? n0 : n0 * p.ldb);
  auto load = [&](int64_t k0) {
    const double* Ak = A0 + (TA ? k0 : k0 * p.lda);
    const double* Bk = B0 + (TB ? k0 * p.ldb : k0);
    const __amdgpu_buffer_rsrc_t a_rsrc = make_rsrc(Ak), b_rsrc = make_rsrc(Bk);
    const int64_t left = kend - k0;
    const int kleft = left < BK ? int(left) : BK;
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      int mm, kk;
      a_coord(r, mm, kk);
      const uint32_t off = kk < kleft ? a_off[r] : kOutOfRange;
      ra[r] = M::cvt(load_f64(WIDE ? make_rsrc(Ak + int64_t(r) * SA * p.lda) : a_rsrc, off));
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      int nn, kk;
      b_coord(r, nn, kk);
      const uint32_t off = kk < kleft ? b_off[r] : kOutOfRange;
      rb[r] = M::cvt(load_f64(WIDE ? make_rsrc(Bk + int64_t(r) * SB * p.ldb) : b_rsrc, off));
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      int mm, kk;
      a_coord(r, mm, kk);
      As[buf][LA::at(mm, kk)] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      int nn, kk;
      b_coord(r, nn, kk);
      Bs[buf][LB::at(nn, kk)] = rb[r];
    }
  };

  acc_t acc[TMs][TNs];
#pragma unroll
  for (int i = 0; i < TMs; ++i)
#pragma unroll
    for (int j = 0; j < TNs; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int64_t nsteps = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
  const int64_t klast = kbeg + (nsteps - 1) * BK;
  if (nsteps > 0) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  const int lrow = lane >> 4, lcol = lane & 15;
  for (int64_t s = 0; s < nsteps; ++s) {
    const int buf = int(s & 1);
    // The last step reloads its own tile into the idle buffer: no branch.
    load(min(kbeg + (s + 1) * BK, klast));
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[TMs], b[TNs];
#pragma unroll
      for (int i = 0; i < TMs; ++i) a[i] = As[buf][LA::at(wm * WM + i * 16 + lcol, kk + lrow)];
#pragma unroll
      for (int j = 0; j < TNs; ++j) b[j] = Bs[buf][LB::at(wn * WN + j * 16 + lcol, kk + lrow)];
#pragma unroll
      for (int i = 0; i < TMs; ++i)
#pragma unroll
        for (int j = 0; j < TNs; ++j) acc[i][j] = M::run(a[i], b[j], acc[i][j]);
    }
    store(buf ^ 1);
    __syncthreads();
  }

  // epilogue
  if (p.W) {
    double* W = p.W + (split + bz * gridDim.y) * p.m * p.n;
#pragma unroll
    for (int i = 0; i < TMs; ++i)
#pragma unroll
      for (int j = 0; j < TNs; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wm * WM + i * 16 + M::row(lrow, r);
          const int64_t col = n0 + wn * WN + j * 16 + lcol;
          if (row < p.m && col < p.n) W[row + col * p.m] = double(acc[i][j][r]);
        }
    return;
  }
  double* C = p.C + bz * p.strideC;
#pragma unroll
  for (int i = 0; i < TMs; ++i)
#pragma unroll
    for (int j = 0; j < TNs; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * WM + i * 16 + M::row(lrow, r);
        const int64_t col = n0 + wn * WN + j * 16 + lcol;
        if (row < p.m && col < p.n) {
          double* c = C + row + col * p.ldc;
          const double v = p.alpha * double(acc[i][j][r]);
          *c = (p.beta == 0.0) ? v : v + p.beta * *c;
        }
      }
}

// C = beta * C (a product with k = 0).
__global__ void scale_matrix(double* C, int64_t m, int64_t n, int64_t ldc, int64_t strideC,
                             double beta) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  double* c = C + int64_t(blockIdx.z) * strideC + (idx % m) + (idx / m) * ldc;
  *c = (beta == 0.0) ? 0.0 : beta * *c;
}

// C = alpha * sum_s W_s + beta * C for split-K partials.
__global__ void splitk_reduce(const double* __restrict__ W, int splits, int64_t m, int64_t n,
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

template <typename T, bool TA, bool TB, int BM, int BN, bool WIDE>
const char* kernel_name() {
  static const std::string name = std::string("gemm_mfma16<") + (sizeof(T) == 8 ? "f64," : "f32,") +
                                  (TA ? "T" : "N") + (TB ? "T" : "N") + "," + std::to_string(BM) +
                                  "x" + std::to_string(BN) + (WIDE ? ",wide" : "") + ">";
  return name.c_str();
}

template <typename T, int BM, int BN, int BK, bool WIDE>
const char* launch_cfg(hipStream_t stream, const Gemm& g, Params p, int splits) {
  p.tiles_m = int(ceil_div(g.m, BM));
  p.tiles_n = int(ceil_div(g.n, BN));
  dim3 grid(p.tiles_m * p.tiles_n, splits, unsigned(g.batch));
  const char* name;
  if (!g.ta && !g.tb) {
    gemm_mfma16<T, false, false, BM, BN, BK, WIDE><<<grid, kThreads, 0, stream>>>(p);
    name = kernel_name<T, false, false, BM, BN, WIDE>();
  } else if (!g.ta && g.tb) {
    gemm_mfma16<T, false, true, BM, BN, BK, WIDE><<<grid, kThreads, 0, stream>>>(p);
    name = kernel_name<T, false, true, BM, BN, WIDE>();
  } else if (g.ta && !g.tb) {
    gemm_mfma16<T, true, false, BM, BN, BK, WIDE><<<grid, kThreads, 0, stream>>>(p);
    name = kernel_name<T, true, false, BM, BN, WIDE>();
  } else {
    gemm_mfma16<T, true, true, BM, BN, BK, WIDE><<<grid, kThreads, 0, stream>>>(p);
    name = kernel_name<T, true, true, BM, BN, WIDE>();
  }
  ETG_LAUNCH("gemm_mfma16");
  return name;
}

template <typename T, int BM, int BN, int BK>
const char* launch_tile(hipStream_t stream, const Gemm& g, const Params& p, int splits) {
  // one buffer base per operand covers 128 columns when the leading dimensions are below 2^21
  const bool wide = std::max(g.lda, g.ldb) >= (int64_t(1) << 21);
  return wide ? launch_cfg<T, BM, BN, BK, true>(stream, g, p, splits)
              : launch_cfg<T, BM, BN, BK, false>(stream, g, p, splits);
}

struct Choice {
  int bm = 128, splits = 1;
};

// only_bm != 0 restricts the search to that tile size
Choice choose_tiling(const Gemm& g, double peak, bool can_split, int only_bm = 0) {
  struct Cfg { int bm; int64_t slots; double eff; };
  const Cfg cfgs[2] = {{128, 512, 0.83}, {64, 1024, 0.62}};
  const double hbm = 5.0e12;                       // sustained bytes/s for the split-K round trip
  const double max_ws = 2.0e9;                     // bytes of split-K partials
  Choice best;
  double best_t = 1e300;
  for (const Cfg& c : cfgs) {
    if (only_bm != 0 && c.bm != only_bm) continue;
    const int64_t tiles = ceil_div(g.m, c.bm) * ceil_div(g.n, c.bm) * g.batch;
    static const int kSplits[] = {1,  2,  3,  4,  5,  6,  7,  8,  9,   10,  11,  12,  13,  14,  16,
                                  18, 20, 24, 28, 32, 40, 48, 64, 96, 128, 192, 256, 384, 512, 1024};
    for (int s : kSplits) {
      if (s > 1 && (!can_split || g.k < int64_t(s) * 256)) break;
      const double mn = double(g.m) * double(g.n) * double(g.batch);
      if (s > 1 && s * mn * 8.0 > max_ws) break;
      const int64_t kps = ceil_div(ceil_div(g.k, s), 16) * 16;
      const int64_t rounds = ceil_div(tiles * s, c.slots);
      const double t_round = double(c.slots) * 2.0 * c.bm * c.bm * double(kps) / (c.eff * peak);
      double t = double(rounds) * t_round;
      if (s > 1) t += (2.0 * s + 2.0) * mn * 8.0 / hbm + 4e-6;
      if (t < best_t * 0.98) {
        best_t = t;
        best.bm = c.bm;
        best.splits = s;
      }
    }
  }
  return best;
}

// Runs the product with a given tile size and split-K count.
template <typename T>
GemmLaunch run_choice(hipStream_t stream, const Gemm& g, GemmWorkspace* ws, Choice ch) {
  Params p{};
  p.m = g.m; p.n = g.n; p.k = g.k;
  p.alpha = g.alpha; p.beta = g.beta;
  p.A = g.A; p.lda = g.lda; p.strideA = g.strideA;
  p.B = g.B; p.ldb = g.ldb; p.strideB = g.strideB;
  p.C = g.C; p.ldc = g.ldc; p.strideC = g.strideC;
  p.W = nullptr;
  p.k_per_split = g.k;
  int splits = ch.splits;
  if (splits > 1) {
    const int64_t kps = ceil_div(ceil_div(g.k, splits), 16) * 16;
    splits = int(ceil_div(g.k, kps));
    p.k_per_split = kps;
    const size_t need = size_t(splits) * size_t(g.m) * size_t(g.n) * size_t(g.batch) * sizeof(double);
    if (ws->buf.resize(need)) ++ws->gen;
    p.W = ws->buf.as<double>();
  }
  GemmLaunch info;
  info.splits = splits;
  if (ch.bm == 128) info.kernel = launch_tile<T, 128, 128, 16>(stream, g, p, splits);
  else info.kernel = launch_tile<T, 64, 64, 16>(stream, g, p, splits);
  if (splits > 1) {
    const int64_t mn = g.m * g.n;
    dim3 grid(unsigned(ceil_div(mn, 256)), 1, unsigned(g.batch));
    splitk_reduce<<<grid, 256, 0, stream>>>(p.W, splits, g.m, g.n, g.alpha, g.beta, g.C, g.ldc,
                                            g.strideC);
    ETG_LAUNCH("splitk_reduce");
  }
  return info;
}

// ---- measured tile choice ---------------------------------------------------
// The cost model below ranks tile size and split-K count; the first time a
// product shape is seen outside graph capture, the model's pick and its
// neighbours (the other tile size's best, half and twice the split count) are
// timed on the real operands into a scratch output, and the fastest is kept
// for the rest of the process.  Products under 2 GFLOP (tens of microseconds)
// and shapes whose output exceeds 1 GiB keep the model's pick; batch counts
// share a choice within a power-of-two bucket (the trial-vector groups of a
// Davidson solve shrink one by one).
struct TuneKey {
  int prec;
  bool ta, tb, wide, shared_a, shared_b;
  int64_t m, n, k, batch;
  bool operator<(const TuneKey& o) const {
    return std::tie(prec, ta, tb, wide, shared_a, shared_b, m, n, k, batch) <
           std::tie(o.prec, o.ta, o.tb, o.wide, o.shared_a, o.shared_b, o.m, o.n, o.k, o.batch);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, Choice> g_tuned;
std::atomic<bool> g_autotune{true};

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  ETG_HIP(hipStreamIsCapturing(s, &st));
  return st != hipStreamCaptureStatusNone;
}

template <typename T>
Choice tuned_choice(hipStream_t stream, const Gemm& g, GemmWorkspace* ws, Choice model, double peak) {
  if (ws == nullptr || !g_autotune.load()) return model;
  if (2.0 * double(g.m) * double(g.n) * double(g.k) * double(g.batch) < 2.0e9) return model;
  int64_t bucket = 1;
  while (bucket < g.batch) bucket <<= 1;
  const TuneKey key{int(sizeof(T)), g.ta, g.tb, std::max(g.lda, g.ldb) >= (int64_t(1) << 21),
                    g.batch > 1 && g.strideA == 0, g.batch > 1 && g.strideB == 0, g.m, g.n, g.k, bucket};
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) return it->second;
  }
  const double out_bytes = 8.0 * double(g.m) * double(g.n) * double(g.batch);
  if (capturing(stream) || out_bytes > double(int64_t(1) << 30)) return model;
  std::vector<Choice> cand{model};
  auto add = [&](Choice c) {
    for (const Choice& e : cand)
      if (e.bm == c.bm && e.splits == c.splits) return;
    cand.push_back(c);
  };
  add(choose_tiling(g, peak, true, model.bm == 128 ? 64 : 128));
  if (model.splits > 1) add(Choice{model.bm, std::max(1, model.splits / 2)});
  if (g.k >= int64_t(2 * model.splits) * 256 && 2.0 * model.splits * out_bytes <= 2.0e9)
    add(Choice{model.bm, 2 * model.splits});
  Choice best = model;
  if (cand.size() > 1) {
    if (ws->tune.resize(size_t(out_bytes))) ++ws->gen;
    Gemm t = g;
    t.C = ws->tune.as<double>();
    t.ldc = g.m;
    t.strideC = g.m * g.n;
    t.beta = 0.0;
    std::vector<hipEvent_t> ev(2 * cand.size());
    for (auto& e : ev) ETG_HIP(hipEventCreate(&e));
    run_choice<T>(stream, t, ws, cand[0]);            // warm caches and page tables
    for (size_t c = 0; c < cand.size(); ++c) {
      ETG_HIP(hipEventRecord(ev[2 * c], stream));
      run_choice<T>(stream, t, ws, cand[c]);
      ETG_HIP(hipEventRecord(ev[2 * c + 1], stream));
    }
    ETG_HIP(hipEventSynchronize(ev.back()));
    float best_ms = 0.0f;
    for (size_t c = 0; c < cand.size(); ++c) {
      float ms = 0.0f;
      ETG_HIP(hipEventElapsedTime(&ms, ev[2 * c], ev[2 * c + 1]));
      if (c == 0 || ms < best_ms) {
        best_ms = ms;
        best = cand[c];
      }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tuned[key] = best;
  return best;
}

template <typename T>
GemmLaunch launch_rows(hipStream_t stream, const Gemm& g, GemmWorkspace* ws) {
  // Tile size and split-K count from a cost model: a launch runs in rounds of
  // `slots` concurrent tiles (two 128 x 128 or four 64 x 64 workgroups per CU
  // on 256 CUs), each round at the tile's sustained matrix-core efficiency;
  // a split adds its partial sums' round trip through HBM.  The model's pick
  // is then checked against its neighbours by measurement (tuned_choice).
  const double peak = sizeof(T) == 8 ? 78.6e12 : 157.3e12;
  Choice ch = choose_tiling(g, peak, ws != nullptr);
  ch = tuned_choice<T>(stream, g, ws, ch, peak);
  return run_choice<T>(stream, g, ws, ch);
}

// A product whose last 128-row tile would hold at most 64 rows (the ladder's
// o(o+1)/2 = 1176 = 9 x 128 + 24 pairs, or n_J = 1600 = 12 x 128 + 64) would
// spend a whole tile row on them: beyond 1e10 FLOP those rows run as a product
// of their own, on the tiling chosen for that shape.
template <typename T>
GemmLaunch launch_typed(hipStream_t stream, const Gemm& g, GemmWorkspace* ws) {
  const int64_t rest_rows = g.m % 128;
  if (g.m > 128 && rest_rows != 0 && rest_rows <= 64 && gemm_flops(g) >= 1e10) {
    Gemm top = g, rest = g;
    top.m = g.m - rest_rows;
    rest.m = rest_rows;
    rest.A = g.ta ? g.A + top.m * g.lda : g.A + top.m;
    rest.C = g.C + top.m;
    const GemmLaunch info = launch_rows<T>(stream, top, ws);
    launch_rows<T>(stream, rest, ws);
    return info;
  }
  return launch_rows<T>(stream, g, ws);
}

// ---- tall products with leading dimensions beyond the matrix-core kernel ----
// Davidson's trial space at o = 40, v = 152 has n_t1 + n_t2 > 2^24 rows, more
// than the 31-bit per-thread buffer offsets of the tile kernel span over 16
// columns.  Its products are tall and skinny and bound by HBM, so two plain
// FP64 kernels with 64-bit addressing serve them:
//   gemm_tall_tn: C (m x n) = alpha A^T B + beta C, A (K x m), B (K x n), K long,
//     m <= 128, n <= 16: each block takes a range of rows in chunks of 32,
//     staged through LDS row-major (a chunk row holds the m values of A and
//     the n of B); wave w accumulates columns 4w .. 4w+3 of C and lane l rows
//     2l, 2l+1, so per row an FMA-pair reads 16 B of A (lanes contiguous, no
//     bank conflict) and a broadcast of B; the next chunk's loads are in flight
//     while the current one is multiplied.  Per-block partial sums, then a
//     deterministic sum over the blocks (for m <= 16 a kernel with one
//     output per thread instead);
//   gemm_tall_nn: C (M x n) = alpha A B + beta C, A (M x k), B (k x n) small:
//     B in LDS, one row of C per thread, columns of A read coalesced.
constexpr int kTallRows = 32;
constexpr int kTallThreads = 256;
constexpr int kTallM = 128;                        // rows of C (A columns)
constexpr int kTallN = 16;                         // columns of C (B columns)
constexpr int kTallLdA = kTallM + 2;               // LDS row stride of A (doubles; keeps 16-B alignment)

__global__ __launch_bounds__(kTallThreads) void gemm_tall_tn_kernel(
    const double* __restrict__ A, int64_t lda, const double* __restrict__ B, int64_t ldb, int64_t K, int m,
    int n, int64_t rows_per_block, double* __restrict__ part) {
  __shared__ double As[kTallRows * kTallLdA];      // [row][i]
  __shared__ double Bs[kTallRows * kTallN];        // [row][j]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = int64_t(blockIdx.x) * rows_per_block;
  const int64_t r1 = min(K, r0 + rows_per_block);
  // loads of one chunk: thread t takes row t % 32 of columns t / 32 + 8 q
  const int lr = tid & (kTallRows - 1), lc = tid >> 5;
  double ra[kTallM / 8], rb[kTallN / 8];
  auto load = [&](int64_t rb0) {
    const int64_t row = rb0 + lr;
    const bool ok = row < r1;
#pragma unroll
    for (int q = 0; q < kTallM / 8; ++q) {
      const int i = lc + 8 * q;
      ra[q] = (ok && i < m) ? A[row + int64_t(i) * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kTallN / 8; ++q) {
      const int j = lc + 8 * q;
      rb[q] = (ok && j < n) ? B[row + int64_t(j) * ldb] : 0.0;
    }
  };
  auto store = [&] {
#pragma unroll
    for (int q = 0; q < kTallM / 8; ++q) As[lr * kTallLdA + lc + 8 * q] = ra[q];
#pragma unroll
    for (int q = 0; q < kTallN / 8; ++q) Bs[lr * kTallN + lc + 8 * q] = rb[q];
  };
  double acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  if (r0 < r1) load(r0);
  for (int64_t rb0 = r0; rb0 < r1; rb0 += kTallRows) {
    __syncthreads();                               // the previous chunk's reads are done
    store();
    __syncthreads();
    if (rb0 + kTallRows < r1) load(rb0 + kTallRows);
#pragma unroll 8
    for (int rr = 0; rr < kTallRows; ++rr) {
      const double2 a = *reinterpret_cast<const double2*>(&As[rr * kTallLdA + 2 * lane]);
      const double* b = &Bs[rr * kTallN + 4 * wave];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        acc[0][c] += a.x * b[c];
        acc[1][c] += a.y * b[c];
      }
    }
  }
  const int nout = m * n;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 2 * lane + a, j = 4 * wave + c;
      if (i < m && j < n) part[int64_t(blockIdx.x) * nout + i + j * m] = acc[a][c];
    }
}

// The same partial sums for m, n <= 16: a chunk of 32 rows of A and B staged
// in LDS, thread t accumulating C(t % 16, t / 16) over the chunk (one FMA per
// row and thread; the 128 x 16 kernel would spend 2048 per row).
__global__ __launch_bounds__(kTallThreads) void gemm_tall_tn_small_kernel(
    const double* __restrict__ A, int64_t lda, const double* __restrict__ B, int64_t ldb, int64_t K, int m,
    int n, int64_t rows_per_block, double* __restrict__ part) {
  __shared__ double As[kTallRows * 17];            // [row][i]
  __shared__ double Bs[kTallRows * 17];            // [row][j]
  const int tid = threadIdx.x;
  const int ci = tid & 15, cj = tid >> 4;
  const int64_t r0 = int64_t(blockIdx.x) * rows_per_block;
  const int64_t r1 = min(K, r0 + rows_per_block);
  double ra[2], rb[2];
  auto load = [&](int64_t rb0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + kTallThreads * q;       // row e % 32, column e / 32
      const int64_t row = rb0 + (e & (kTallRows - 1));
      const int col = e / kTallRows;
      const bool ok = row < r1;
      ra[q] = (ok && col < m) ? A[row + int64_t(col) * lda] : 0.0;
      rb[q] = (ok && col < n) ? B[row + int64_t(col) * ldb] : 0.0;
    }
  };
  double acc = 0.0;
  if (r0 < r1) load(r0);
  for (int64_t rb0 = r0; rb0 < r1; rb0 += kTallRows) {
    __syncthreads();                               // the previous chunk's reads are done
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + kTallThreads * q;
      As[(e & (kTallRows - 1)) * 17 + e / kTallRows] = ra[q];
      Bs[(e & (kTallRows - 1)) * 17 + e / kTallRows] = rb[q];
    }
    __syncthreads();
    if (rb0 + kTallRows < r1) load(rb0 + kTallRows);
#pragma unroll 8
    for (int rr = 0; rr < kTallRows; ++rr) acc += As[rr * 17 + ci] * Bs[rr * 17 + cj];
  }
  if (ci < m && cj < n) part[int64_t(blockIdx.x) * (m * n) + ci + cj * m] = acc;
}

// C(i, j) (or C(j, i) with transpose) = alpha sum_b part_b(i, j) + beta C.
// 16 outputs per workgroup, each summed by 16 threads over every 16th block
// and the 16 partial sums then added in a fixed order (deterministic).
__global__ __launch_bounds__(256) void gemm_tall_reduce_kernel(const double* __restrict__ part, int64_t nblocks,
                                                               int m, int n, double alpha, double beta,
                                                               double* __restrict__ C, int64_t ldc, int transpose) {
  __shared__ double red[16][17];
  const int lo = threadIdx.x & 15, q = threadIdx.x >> 4;
  const int o = blockIdx.x * 16 + lo;
  const int mn = m * n;
  double s = 0.0;
  if (o < mn)
    for (int64_t b = q; b < nblocks; b += 16) s += part[b * mn + o];
  red[q][lo] = s;
  __syncthreads();
  if (q != 0 || o >= mn) return;
  s = 0.0;
  for (int r = 0; r < 16; ++r) s += red[r][lo];
  const int i = o % m, j = o / m;
  double* c = transpose ? C + j + int64_t(i) * ldc : C + i + int64_t(j) * ldc;
  *c = (beta == 0.0) ? alpha * s : alpha * s + beta * *c;
}

__global__ __launch_bounds__(kTallThreads) void gemm_tall_nn_kernel(
    const double* __restrict__ A, int64_t lda, const double* __restrict__ B, int64_t ldb, double* __restrict__ C,
    int64_t ldc, int64_t M, int k, int n, double alpha, double beta) {
  extern __shared__ double Bs[];                    // k x n
  for (int e = threadIdx.x; e < k * n; e += kTallThreads) Bs[e] = B[(e % k) + int64_t(e / k) * ldb];
  __syncthreads();
  for (int64_t row = int64_t(blockIdx.x) * kTallThreads + threadIdx.x; row < M;
       row += int64_t(gridDim.x) * kTallThreads) {
    for (int j0 = 0; j0 < n; j0 += 16) {
      const int nj = min(16, n - j0);
      double acc[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) acc[jj] = 0.0;
      // eight columns of A per round: their loads are in flight together
      int i = 0;
      for (; i + 8 <= k; i += 8) {
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = A[row + int64_t(i + u) * lda];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int jj = 0; jj < 16; ++jj)
            if (jj < nj) acc[jj] += a[u] * Bs[i + u + k * (j0 + jj)];
      }
      for (; i < k; ++i) {
        const double a = A[row + int64_t(i) * lda];
#pragma unroll
        for (int jj = 0; jj < 16; ++jj)
          if (jj < nj) acc[jj] += a * Bs[i + k * (j0 + jj)];
      }
      for (int jj = 0; jj < nj; ++jj) {
        double* c = C + row + int64_t(j0 + jj) * ldc;
        *c = (beta == 0.0) ? alpha * acc[jj] : alpha * acc[jj] + beta * *c;
      }
    }
  }
}

GemmLaunch gemm_tall(hipStream_t stream, const Gemm& g, GemmWorkspace* ws) {
  if (g.batch != 1) fail("gemm: leading dimension too large for a batched product");
  if (g.ta && !g.tb) {
    if (ws == nullptr) fail("gemm: tall products need a workspace");
    // the kernel takes at most 128 x 16 outputs: the longer side of C goes to its
    // 128-row side (C^T = B^T A when n > m), larger C in 128 x 16 blocks
    const bool swap = g.n > g.m;
    const double* P = swap ? g.B : g.A;
    const double* Q = swap ? g.A : g.B;
    const int64_t ldp = swap ? g.ldb : g.lda, ldq = swap ? g.lda : g.ldb;
    const int64_t rows = swap ? g.n : g.m, cols = swap ? g.m : g.n;
    const int64_t nblocks = std::max<int64_t>(1, std::min<int64_t>(2048, ceil_div(g.k, 4096)));
    const int64_t per = ceil_div(ceil_div(g.k, nblocks), kTallRows) * kTallRows;
    const int64_t nb = ceil_div(g.k, per);
    if (ws->buf.resize(sizeof(double) * size_t(nb * kTallM * kTallN))) ++ws->gen;
    for (int64_t j0 = 0; j0 < cols; j0 += kTallN)
      for (int64_t i0 = 0; i0 < rows; i0 += kTallM) {
        const int mi = int(std::min<int64_t>(kTallM, rows - i0)), nj = int(std::min<int64_t>(kTallN, cols - j0));
        if (mi <= 16)
          gemm_tall_tn_small_kernel<<<unsigned(nb), kTallThreads, 0, stream>>>(
              P + i0 * ldp, ldp, Q + j0 * ldq, ldq, g.k, mi, nj, per, ws->buf.as<double>());
        else
          gemm_tall_tn_kernel<<<unsigned(nb), kTallThreads, 0, stream>>>(P + i0 * ldp, ldp, Q + j0 * ldq, ldq,
                                                                          g.k, mi, nj, per, ws->buf.as<double>());
        ETG_LAUNCH("gemm_tall_tn");
        double* C0 = swap ? g.C + j0 + i0 * g.ldc : g.C + i0 + j0 * g.ldc;
        gemm_tall_reduce_kernel<<<unsigned(ceil_div(int64_t(mi) * nj, 16)), 256, 0, stream>>>(
            ws->buf.as<double>(), nb, mi, nj, g.alpha, g.beta, C0, g.ldc, swap ? 1 : 0);
        ETG_LAUNCH("gemm_tall_reduce");
      }
    return GemmLaunch{"gemm_tall_tn", int(nb)};
  }
  if (!g.ta && !g.tb && g.k * g.n <= 8192) {
    const unsigned grid = unsigned(std::min<int64_t>(ceil_div(g.m, kTallThreads), 4096));
    gemm_tall_nn_kernel<<<grid, kTallThreads, sizeof(double) * size_t(g.k * g.n), stream>>>(
        g.A, g.lda, g.B, g.ldb, g.C, g.ldc, g.m, int(g.k), int(g.n), g.alpha, g.beta);
    ETG_LAUNCH("gemm_tall_nn");
    return GemmLaunch{"gemm_tall_nn", 1};
  }
  fail("gemm: leading dimension too large");
}

}  // namespace

void set_gemm_autotune(bool on) { g_autotune.store(on); }

int64_t gemm_tuned_shapes() {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  return int64_t(g_tuned.size());
}

GemmLaunch gemm(hipStream_t stream, const Gemm& g, Prec prec, GemmWorkspace* ws) {
  if (g.m < 0 || g.n < 0 || g.k < 0 || g.batch < 1) fail("gemm: negative dimension");
  if (g.m == 0 || g.n == 0) return GemmLaunch{};
  if (g.b_packed != nullptr && (prec != Prec::BF16X3 || g.tb || g.k == 0 ||
                                std::max(g.lda, g.ldb) >= (int64_t(1) << 24)))
    fail("gemm: a prepacked B operand needs a split-bf16 product with op(B) = B");
  if (g.Cf != nullptr && (prec != Prec::BF16X3 || g.beta != 0.0 || g.k == 0 ||
                          std::max(g.lda, g.ldb) >= (int64_t(1) << 24)))
    fail("gemm: an FP32 output needs a split-bf16 product with beta = 0");
  const int64_t rowsA = g.ta ? g.k : g.m;
  const int64_t rowsB = g.tb ? g.n : g.k;
  if (g.lda < std::max<int64_t>(1, rowsA)) fail("gemm: lda too small");
  if (g.ldb < std::max<int64_t>(1, rowsB)) fail("gemm: ldb too small");
  if (g.ldc < std::max<int64_t>(1, g.m)) fail("gemm: ldc too small");
  if (g.batch > 65535) fail("gemm: batch too large");
  if (g.k == 0) {
    const int64_t mn = g.m * g.n;
    dim3 grid(unsigned(ceil_div(mn, 256)), 1, unsigned(g.batch));
    scale_matrix<<<grid, 256, 0, stream>>>(g.C, g.m, g.n, g.ldc, g.strideC, g.beta);
    ETG_LAUNCH("scale_matrix");
    return GemmLaunch{"scale_matrix", 1};
  }
  // per-thread byte offsets inside a tile span 16 operand columns and are 31-bit:
  // longer leading dimensions go to the tall FP64 kernels (any precision asked)
  if (std::max(g.lda, g.ldb) >= (int64_t(1) << 24)) return gemm_tall(stream, g, ws);
  if (prec == Prec::BF16X3 && split_bf16_pays(g)) return gemm_bf16x3(stream, g, ws);
  if (prec == Prec::F32) return launch_typed<float>(stream, g, ws);
  return launch_typed<double>(stream, g, ws);
}

}  // namespace etg
