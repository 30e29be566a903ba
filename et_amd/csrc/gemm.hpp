// Dense matrix products on CDNA4 matrix cores.
//
// Column-major (Fortran) convention throughout, as the reference's BLAS
// calls use it: op(A) is m x k, op(B) is k x n, C is m x n,
//   C = alpha * op(A) * op(B) + beta * C,
// op(X) = X ("N") or X^T ("T").  Batched products take element strides.
#pragma once

#include "common.hpp"

namespace etg {

// Arithmetic used for the products.
//  F64  : FP64 operands, v_mfma_f64_16x16x4 (exact FP64 arithmetic).
//  F32  : FP64 operands rounded to FP32 on load, v_mfma_f32_16x16x4, FP32 accumulation.
//  BF16X3: FP64 operands split into bf16 hi + lo (x ~ hi + lo) by a pack pass;
//         three products hi*hi + hi*lo + lo*hi per term with
//         v_mfma_f32_16x16x32_bf16, FP32 accumulation.
enum class Prec : int { F64 = 0, F32 = 1, BF16X3 = 2 };

struct Gemm {
  bool ta = false, tb = false;
  int64_t m = 0, n = 0, k = 0;
  double alpha = 1.0, beta = 0.0;
  const double* A = nullptr;
  int64_t lda = 0, strideA = 0;
  const double* B = nullptr;
  int64_t ldb = 0, strideB = 0;
  double* C = nullptr;
  int64_t ldc = 0, strideC = 0;
  int64_t batch = 1;
  // Nonzero: the caller guarantees that A holds the same data whenever it passes
  // this tag, so the split-bf16 path may reuse the packed A of an earlier call.
  uint64_t a_tag = 0;
  // Split-bf16 products only: op(B) given as the pack of op(B)^T (n rows, k
  // columns, split_bf16_pack_elems(n, k) bf16 values, fragment order), shared by
  // the batch; B and ldb are then not read.
  const void* b_packed = nullptr;
  // Split-bf16 products only: C written in FP32 (the products' accumulation
  // type) to Cf instead of C; beta must be 0.
  float* Cf = nullptr;
  // Split-bf16 products only: run the split-bf16 kernel whatever the shape
  // (diagnostics; see split_bf16_pays).
  bool split_any_shape = false;
};

// Whether a product asked for in split-bf16 runs on the split-bf16 kernel.  The
// kernel first packs both operands (one pass over each).  A product with few
// outputs per operand element, m n < 64 (m + n), is bound by HBM even on the
// FP64 matrix cores (at most ~32 FLOP per operand byte); when its operands also
// hold 2^24 elements or more (128 MB in FP64), where the pack passes cost more
// than the whole FP64 product, it runs in FP64 instead: faster, and exact.
// Smaller products keep the precision asked for.  Products with a prepacked B
// or an FP32 output always take the split-bf16 kernel.
inline bool split_bf16_pays(const Gemm& g) {
  if (g.b_packed != nullptr || g.Cf != nullptr || g.split_any_shape) return true;
  const double m = double(g.m), n = double(g.n);
  const double operand_elems = (m + n) * double(g.k) * double(g.batch);
  return m * n >= 64.0 * (m + n) || operand_elems < double(int64_t(1) << 24);
}

// A packed split-bf16 operand kept for reuse under a caller tag.
struct PackedOperand {
  uint64_t tag = 0;
  bool ta = false;
  int64_t m = 0, k = 0, lda = 0, strideA = 0, batch = 0;
  const double* A = nullptr;
  uint64_t used = 0;                 // last use (least recently used entry is replaced)
  DeviceBuffer buf;
};

// Scratch for split-K partial sums and the split-bf16 operand packs; grown on
// demand.  `gen` counts the times one of its buffers moved (launch sequences
// captured into a graph hold these addresses).
struct GemmWorkspace {
  DeviceBuffer buf, packA, packB;
  DeviceBuffer partials;             // per-entry products summed by the caller
  DeviceBuffer tune;                 // scratch output of the tile measurements
  static constexpr int kCached = 4;
  PackedOperand cached[kCached];
  uint64_t clock = 0;
  uint64_t gen = 0;
};

// Which kernel instantiation a product ran on (for the per-kernel ledger).
struct GemmLaunch {
  const char* kernel = "";   // e.g. "gemm_mfma16<f64,NN,128x128>"
  int splits = 1;            // > 1: split-K partial sums + reduction kernel
};

// Measured tile choice for FP64/FP32 products (on by default, process-wide):
// the first product of a shape outside graph capture times the cost model's
// pick against its neighbours and keeps the fastest.
void set_gemm_autotune(bool on);
int64_t gemm_tuned_shapes();

// Launches the product on `stream`.  Validates shapes and leading dimensions
// on the host (throws etg::Error on inconsistent arguments).
GemmLaunch gemm(hipStream_t stream, const Gemm& g, Prec prec, GemmWorkspace* ws);

// The split-bf16 path (packs both operands into ws, then the bf16 matrix-core kernel).
GemmLaunch gemm_bf16x3(hipStream_t stream, const Gemm& g, GemmWorkspace* ws);

// bf16 values of the split-bf16 pack of a rows x k operand (both splits, padded
// to the product kernel's 256-row tile and 32-deep step).
int64_t split_bf16_pack_elems(int64_t rows, int64_t k);

// The ladder's g+- of one G block as split-bf16 packs for the products
// w+- = t+- g+- (B operands: rows = the block's npairs pairs, k = the tile-blocked
// (c, d) index of ladder_order.hpp, ladder_blocked_len(n_v, +-) long).  G is the
// (n_v lenA) x (n_v lenB) block in FP32, pair_a / pair_b the pairs' local a and b.
void ladder_g_pm_split(hipStream_t s, const float* G, int64_t n_v, int64_t lenA, const int32_t* pair_a,
                       const int32_t* pair_b, int64_t npairs, __bf16* Pp, __bf16* Pm);

// Convenience: C = alpha op(A) op(B) + beta C, no batching.
inline GemmLaunch gemm(hipStream_t s, char ta, char tb, int64_t m, int64_t n, int64_t k,
                       double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                       double beta, double* C, int64_t ldc, Prec prec, GemmWorkspace* ws) {
  Gemm g;
  g.ta = (ta == 'T' || ta == 't');
  g.tb = (tb == 'T' || tb == 't');
  g.m = m; g.n = n; g.k = k;
  g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  return gemm(s, g, prec, ws);
}

// Gemm descriptor for C = alpha op(A) op(B) + beta C without batching.
inline Gemm make_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                      int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  Gemm g;
  g.ta = (ta == 'T' || ta == 't');
  g.tb = (tb == 'T' || tb == 't');
  g.m = m; g.n = n; g.k = k;
  g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  return g;
}

// Floating-point operations the product performs (2 m n k per batch entry).
inline double gemm_flops(const Gemm& g) {
  return 2.0 * double(g.m) * double(g.n) * double(g.k) * double(g.batch);
}

}  // namespace etg
