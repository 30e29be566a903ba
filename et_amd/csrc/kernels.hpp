// Bandwidth-bound kernels of the CCSD residual: index permutations,
// (un)packing of the symmetric doubles vector, the +/- combinations of the
// ladder term, and deterministic FP64 reductions.  All arrays are
// column-major; "packed" means the eT doubles layout
//   idx(X, Y) = max(X,Y) (max(X,Y) + 1) / 2 + min(X,Y),   X = a + n_v i.
#pragma once

#include "common.hpp"

namespace etg {

// y[o0,o1,o2,o3] = alpha * x[...] + beta * y, where output index k runs over
// dims out_dim[k] and reads x at sum_k o_k * src_stride[k].
void permute4(hipStream_t s, const double* x, double* y, const int64_t out_dim[4],
              const int64_t src_stride[4], double alpha, double beta);

// Helper: y = alpha * permute(x) + beta * y for x of dims d (column-major) and
// y laid out in the order `perm` (1-based digits, e.g. "1324": y(p,r,q,s) = x(p,q,r,s)).
void sort4(hipStream_t s, const double* x, const int64_t d[4], const char* perm, double* y,
           double alpha = 1.0, double beta = 0.0);

// full N x N symmetric matrix from packed storage
void squareup(hipStream_t s, const double* packed, double* full, int64_t N);

// packed[idx(i,j)] = alpha * full(i,j) for i >= j (+ beta * packed), diagonal times diag_scale
void packin(hipStream_t s, const double* full, double* packed, int64_t N, double alpha, double beta,
            double diag_scale);

// y = x + x^T (N x N), out of place
void symmetric_sum(hipStream_t s, const double* x, double* y, int64_t N);

// y (packed) += alpha (1 - delta/2) [f1 x(p,q,r,s) + f2 x(p,s,r,q) + f1 x(r,s,p,q) + f2 x(r,q,p,s)]
// for x a full (v,o,v,o) array; f1 = 1, f2 = 0 gives the plain symmetrisation.
// Diagonal elements receive alpha * x(p,q,p,q).
void symmetrize_and_pack(hipStream_t s, const double* x, double* packed, int64_t n_v, int64_t n_o,
                         double alpha, double f1, double f2);

// Ladder (+/-) helpers ------------------------------------------------------
// t+ (o+ x v+): (1 - d_cd/2)(t_cidj + t_cjdi), rows ij (i <= j), cols cd (c <= d)
// t- (o- x v-): t_cidj - t_cjdi, rows i < j, cols c < d
void ladder_t_pm(hipStream_t s, const double* t2, double* tp, double* tm, int64_t n_o, int64_t n_v);
// the same in the tile-blocked (c, d) order of the split-bf16 ladder (ladder_order.hpp):
// tp is op x ladder_blocked_len(n_v, true), tm is om x ladder_blocked_len(n_v, false)
void ladder_t_pm_blocked(hipStream_t s, const double* t2, double* tp, double* tm, int64_t n_o,
                         int64_t n_v);

// From G(c, a', d, b') = sum_J L_ac L_bd of one (A, B) block build
// g+ (v+ x npairs) = G(c,a,d,b) + G(d,a,c,b) and g- (v- x npairs) = G(c,a,d,b) - G(d,a,c,b)
// for the block's (a, b) pairs (pair_a, pair_b are local indices into A and B).
void ladder_g_pm(hipStream_t s, const double* G, int64_t n_v, int64_t lenA, int64_t lenB,
                 const int32_t* pair_a, const int32_t* pair_b, int64_t npairs, double* gp, double* gm);

// packed Omega += (1 - delta/2) W_abij with W_abij = w+ +- w- looked up through
// col[a + n_v b] (pair column for a >= b).
void ladder_gather(hipStream_t s, const double* wp, const double* wm, const int32_t* col,
                   double* omega2, int64_t n_o, int64_t n_v);

// Reductions (deterministic two-pass, FP64) ---------------------------------
// sum_aibj (t_aibj + t_ai t_bj)(2 g_iajb - g_ibja); t2 packed, g_ovov (o,v,o,v)
void ccsd_energy(hipStream_t s, const double* t1, const double* t2, const double* g_ovov,
                 int64_t n_o, int64_t n_v, double* partial, double* result);
// ||x||_2^2 into result
void sum_squares(hipStream_t s, const double* x, int64_t n, double* partial, double* result);
// result = x . y
void dot(hipStream_t s, const double* x, const double* y, int64_t n, double* partial, double* result);
constexpr int kReduceBlocks = 1024;

// Amplitude-space helpers ----------------------------------------------------
// MP2 guess: t2[aibj] = g_vovo(a,i,b,j) / (e_i - e_a + e_j - e_b)
void mp2_guess(hipStream_t s, const double* g_vovo, const double* eps, double* t2, int64_t n_o,
               int64_t n_v);
// orbital differences for singles (v,o) and packed doubles
void orbital_differences(hipStream_t s, const double* eps, double* x, int64_t n_o, int64_t n_v,
                         int64_t n);
// excited-state projections of n_vec vectors of length n (eT layout); core: n_o flags
// on the device.  mode 0 (core-valence separation): elements with no core occupied
// index are zeroed; mode 1 (remove core): elements with a core occupied index are
// zeroed; mode 2 (ionisation): elements with no index on the last virtual (the bath
// orbital) are zeroed, core unused
void cvs_project(hipStream_t s, double* x, int64_t n_vec, int64_t n_t1, int64_t n, int64_t n_v,
                 const unsigned char* core, int mode);
// x (+)= alpha * y
void axpy(hipStream_t s, int64_t n, double alpha, const double* x, double* y);
void fill(hipStream_t s, double* x, int64_t n, double value);

}  // namespace etg

namespace etg {

// ---- adjoints of the reorder / packing kernels (derivatives of the residual) ----
// x (n elements) *= beta
void scale(hipStream_t s, double* x, int64_t n, double beta);
// y = x + x^T + beta * y
void symmetric_sum_acc(hipStream_t s, const double* x, double* y, int64_t N, double beta);
// adjoint of squareup: packed[k] += full(X,Y) + full(Y,X)   (full(X,X) on the diagonal)
void packsum(hipStream_t s, const double* full, double* packed, int64_t N);
// adjoint of packin: full(Y, X) += alpha * s_k * packed[k], X >= Y
void unpack_upper_add(hipStream_t s, const double* packed, double* full, int64_t N, double alpha,
                      double diag_scale);
// adjoint of symmetrize_and_pack: x(R,C) += alpha * (...)
void symmetrize_and_pack_adjoint(hipStream_t s, const double* packed, double* x, int64_t n_v,
                                 int64_t n_o, double alpha, double f1, double f2);
// adjoints of the ladder helpers
void ladder_t_pm_adjoint(hipStream_t s, const double* tp, const double* tm, double* t2, int64_t n_o,
                         int64_t n_v);
void ladder_g_pm_adjoint(hipStream_t s, const double* gp, const double* gm, double* G, int64_t n_v,
                         int64_t a0, int64_t lenA, int64_t b0, int64_t lenB, const int32_t* col,
                         int64_t col0);
void ladder_gather_adjoint(hipStream_t s, const double* omega2, const int32_t* pcol_a,
                           const int32_t* pcol_b, double* wp, double* wm, int64_t n_o, int64_t n_v);
// T1-related small kernels (n = n_o + n_v)
void t1_xy(hipStream_t s, const double* t1, double* X, double* Y, int64_t n_o, int64_t n_v,
           double diag);
void t1_xy_adjoint(hipStream_t s, const double* Xb, const double* Yb, double* t1b, int64_t n_o,
                   int64_t n_v);
void trace_oo(hipStream_t s, const double* Loo, double* XJ, int64_t nJ, int64_t n_o);
void trace_oo_adjoint(hipStream_t s, const double* XJb, double* Loob, int64_t nJ, int64_t n_o);
void fock_ov_vector(hipStream_t s, const double* F, double* f, int64_t n_o, int64_t n_v);
void fock_ov_vector_adjoint(hipStream_t s, const double* fb, double* Fb, int64_t n_o, int64_t n_v);
// e = -residual / eps (x diag_factor on the packed doubles diagonal); x += e
void quasi_newton_step(hipStream_t s, const double* eps, const double* residual, double* x, double* e,
                       int64_t n_o, int64_t n_v, int64_t n, double diag_factor);
// dst (m x n, ld ldd) += sum over b < nbatch of the contiguous m x n blocks part_b
void batch_sum_acc(hipStream_t s, const double* part, int64_t nbatch, int64_t m, int64_t n, double* dst,
                   int64_t ldd);
// dst (m x n) = alpha * src + beta * dst
void copy_sub(hipStream_t s, const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
              int64_t n, double alpha, double beta);

}  // namespace etg
