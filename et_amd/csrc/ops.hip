// DirectOps (runs the residual pipeline) and Tape (records it and replays its
// derivative forward or backward).  See ops.hpp.
#include "ops.hpp"

#include <algorithm>
#include <cstring>

#include "ccsd.hpp"
#include "kernels.hpp"

namespace etg {

// ---------------------------------------------------------------------------
// DirectOps
// ---------------------------------------------------------------------------

double* DirectOps::buf(int slot, const char* name, int64_t elems) {
  return wf_.scratch(slot, elems, name);
}

void DirectOps::gemm(const Gemm& g, Prec prec) { wf_.run_gemm(g, prec); }

void DirectOps::sort4(const double* x, const int64_t d[4], const char* perm, double* y, double alpha,
                      double beta) {
  etg::sort4(wf_.stream(), x, d, perm, y, alpha, beta);
}

void DirectOps::squareup(const double* packed, double* full, int64_t N) {
  etg::squareup(wf_.stream(), packed, full, N);
}

void DirectOps::packin(const double* full, double* packed, int64_t N, double alpha, double beta,
                       double diag) {
  etg::packin(wf_.stream(), full, packed, N, alpha, beta, diag);
}

void DirectOps::symsum(const double* x, double* y, int64_t N) {
  etg::symmetric_sum(wf_.stream(), x, y, N);
}

void DirectOps::sympack(const double* x, double* packed, int64_t n_v, int64_t n_o, double alpha,
                        double f1, double f2) {
  etg::symmetrize_and_pack(wf_.stream(), x, packed, n_v, n_o, alpha, f1, f2);
}

void DirectOps::copy_sub(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
                         int64_t n, double alpha, double beta) {
  etg::copy_sub(wf_.stream(), src, lds, dst, ldd, m, n, alpha, beta);
}

void DirectOps::fill0(double* x, int64_t n) { etg::fill(wf_.stream(), x, n, 0.0); }

void DirectOps::t1_xy(const double* t1, double* X, double* Y) {
  etg::t1_xy(wf_.stream(), t1, X, Y, wf_.n_o, wf_.n_v, 1.0);
}

void DirectOps::trace_oo(const double* Loo, double* XJ) {
  etg::trace_oo(wf_.stream(), Loo, XJ, wf_.n_J, wf_.n_o);
}

void DirectOps::fock_ov_vec(const double* F, double* f) {
  etg::fock_ov_vector(wf_.stream(), F, f, wf_.n_o, wf_.n_v);
}

void DirectOps::ladder(const double* t2, const double* LvvT, double* om2, Prec prec) {
  wf_.ladder_primal(t2, LvvT, om2, prec);
}

// ---------------------------------------------------------------------------
// Tape: recording
// ---------------------------------------------------------------------------

double* Tape::buf(int slot, const char* name, int64_t elems) {
  (void)slot;
  auto it = named_.find(name);
  if (it != named_.end()) {
    if (it->second.bytes() < size_t(elems) * sizeof(double))
      fail(std::string("tape buffer grew: ") + name);
    return it->second.as<double>();
  }
  DeviceBuffer b(size_t(std::max<int64_t>(elems, 1)) * sizeof(double));
  double* p = b.as<double>();
  named_.emplace(name, std::move(b));
  register_buffer(p, elems, false, name);
  return p;
}

void Tape::register_buffer(const double* base, int64_t elems, bool constant, const char* name) {
  auto b = std::make_unique<Buf>();
  b->base = base;
  b->elems = elems;
  b->constant = constant;
  b->name = name;
  bufs_.push_back(std::move(b));
}

void Tape::clear() {
  ops_.clear();
  bufs_.clear();
  named_.clear();
}

int Tape::find(const double* p) const {
  for (size_t i = 0; i < bufs_.size(); ++i) {
    const Buf& b = *bufs_[i];
    if (p >= b.base && p < b.base + std::max<int64_t>(b.elems, 1)) return int(i);
  }
  fail("tape: operand not in a registered buffer");
}

double* Tape::deriv(const double* p) {
  Buf& b = *bufs_[size_t(find(p))];
  if (b.constant) return nullptr;
  if (b.d.bytes() < size_t(b.elems) * sizeof(double)) fail("tape: derivatives not initialised");
  return b.d.as<double>() + (p - b.base);
}

void Tape::zero_derivatives() {
  for (auto& bp : bufs_) {
    Buf& b = *bp;
    if (b.constant) continue;
    b.d.resize(size_t(std::max<int64_t>(b.elems, 1)) * sizeof(double));
    fill(wf_.stream(), b.d.as<double>(), b.elems, 0.0);
  }
}

void Tape::gemm(const Gemm& g, Prec prec) {
  Op op;
  op.kind = GEMM;
  op.g = g;
  op.prec = prec;
  ops_.push_back(op);
  DirectOps::gemm(g, prec);
}

void Tape::sort4(const double* x, const int64_t d[4], const char* perm, double* y, double alpha,
                 double beta) {
  Op op;
  op.kind = SORT4;
  op.x = x;
  op.y = y;
  std::memcpy(op.d, d, sizeof(op.d));
  std::strncpy(op.perm, perm, 4);
  op.alpha = alpha;
  op.beta = beta;
  ops_.push_back(op);
  DirectOps::sort4(x, d, perm, y, alpha, beta);
}

void Tape::squareup(const double* packed, double* full, int64_t N) {
  Op op;
  op.kind = SQUAREUP;
  op.x = packed;
  op.y = full;
  op.N = N;
  ops_.push_back(op);
  DirectOps::squareup(packed, full, N);
}

void Tape::packin(const double* full, double* packed, int64_t N, double alpha, double beta,
                  double diag) {
  Op op;
  op.kind = PACKIN;
  op.x = full;
  op.y = packed;
  op.N = N;
  op.alpha = alpha;
  op.beta = beta;
  op.diag = diag;
  ops_.push_back(op);
  DirectOps::packin(full, packed, N, alpha, beta, diag);
}

void Tape::symsum(const double* x, double* y, int64_t N) {
  Op op;
  op.kind = SYMSUM;
  op.x = x;
  op.y = y;
  op.N = N;
  ops_.push_back(op);
  DirectOps::symsum(x, y, N);
}

void Tape::sympack(const double* x, double* packed, int64_t n_v, int64_t n_o, double alpha,
                   double f1, double f2) {
  Op op;
  op.kind = SYMPACK;
  op.x = x;
  op.y = packed;
  op.nv = n_v;
  op.no = n_o;
  op.alpha = alpha;
  op.f1 = f1;
  op.f2 = f2;
  ops_.push_back(op);
  DirectOps::sympack(x, packed, n_v, n_o, alpha, f1, f2);
}

void Tape::copy_sub(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t n,
                    double alpha, double beta) {
  Op op;
  op.kind = COPY_SUB;
  op.x = src;
  op.y = dst;
  op.lds = lds;
  op.ldd = ldd;
  op.m = m;
  op.n = n;
  op.alpha = alpha;
  op.beta = beta;
  ops_.push_back(op);
  DirectOps::copy_sub(src, lds, dst, ldd, m, n, alpha, beta);
}

void Tape::fill0(double* x, int64_t n) {
  Op op;
  op.kind = FILL0;
  op.y = x;
  op.n = n;
  ops_.push_back(op);
  DirectOps::fill0(x, n);
}

void Tape::t1_xy(const double* t1, double* X, double* Y) {
  Op op;
  op.kind = T1_XY;
  op.x = t1;
  op.y = X;
  op.y2 = Y;
  ops_.push_back(op);
  DirectOps::t1_xy(t1, X, Y);
}

void Tape::trace_oo(const double* Loo, double* XJ) {
  Op op;
  op.kind = TRACE_OO;
  op.x = Loo;
  op.y = XJ;
  ops_.push_back(op);
  DirectOps::trace_oo(Loo, XJ);
}

void Tape::fock_ov_vec(const double* F, double* f) {
  Op op;
  op.kind = FOCK_OV_VEC;
  op.x = F;
  op.y = f;
  ops_.push_back(op);
  DirectOps::fock_ov_vec(F, f);
}

void Tape::ladder(const double* t2, const double* LvvT, double* om2, Prec prec) {
  Op op;
  op.kind = LADDER;
  op.x = t2;
  op.x2 = LvvT;
  op.y = om2;
  op.prec = prec;
  ops_.push_back(op);
  DirectOps::ladder(t2, LvvT, om2, prec);
}

// ---------------------------------------------------------------------------
// Tape: replays
// ---------------------------------------------------------------------------

namespace {
int64_t dense_size(const int64_t d[4]) { return d[0] * d[1] * d[2] * d[3]; }

// C = beta * C over the product's output (and batch)
void scale_output(CCSD& wf, const Gemm& g, double* C, double beta, Prec prec) {
  if (beta == 1.0) return;
  Gemm z = g;
  z.k = 0;
  z.C = C;
  z.beta = beta;
  wf.run_gemm(z, prec);
}
}  // namespace

void Tape::tangent(Prec prec) {
  hipStream_t s = wf_.stream();
  for (const Op& op : ops_) {
    switch (op.kind) {
      case GEMM: {
        const Gemm& g = op.g;
        double* dC = deriv(g.C);
        const double* dA = deriv(g.A);
        const double* dB = deriv(g.B);
        scale_output(wf_, g, dC, g.beta, prec);
        if (dA) {
          Gemm h = g;
          h.A = dA;
          h.C = dC;
          h.beta = 1.0;
          wf_.run_gemm(h, prec);
        }
        if (dB) {
          Gemm h = g;
          h.B = dB;
          h.C = dC;
          h.beta = 1.0;
          wf_.run_gemm(h, prec);
        }
        break;
      }
      case SORT4: {
        double* dy = deriv(op.y);
        const double* dx = deriv(op.x);
        if (dx) etg::sort4(s, dx, op.d, op.perm, dy, op.alpha, op.beta);
        else scale(s, dy, dense_size(op.d), op.beta);
        break;
      }
      case SQUAREUP: {
        const double* dp = deriv(op.x);
        double* df = deriv(op.y);
        if (dp) etg::squareup(s, dp, df, op.N);
        else fill(s, df, op.N * op.N, 0.0);
        break;
      }
      case PACKIN: {
        const double* df = deriv(op.x);
        double* dp = deriv(op.y);
        if (df) etg::packin(s, df, dp, op.N, op.alpha, op.beta, op.diag);
        else scale(s, dp, op.N * (op.N + 1) / 2, op.beta);
        break;
      }
      case SYMSUM: {
        const double* dx = deriv(op.x);
        double* dy = deriv(op.y);
        if (dx) etg::symmetric_sum(s, dx, dy, op.N);
        else fill(s, dy, op.N * op.N, 0.0);
        break;
      }
      case SYMPACK: {
        const double* dx = deriv(op.x);
        if (dx) etg::symmetrize_and_pack(s, dx, deriv(op.y), op.nv, op.no, op.alpha, op.f1, op.f2);
        break;
      }
      case COPY_SUB: {
        const double* dsrc = deriv(op.x);
        double* ddst = deriv(op.y);
        if (dsrc) etg::copy_sub(s, dsrc, op.lds, ddst, op.ldd, op.m, op.n, op.alpha, op.beta);
        else etg::copy_sub(s, ddst, op.ldd, ddst, op.ldd, op.m, op.n, 0.0, op.beta);
        break;
      }
      case FILL0:
        fill(s, deriv(op.y), op.n, 0.0);
        break;
      case T1_XY:
        etg::t1_xy(s, deriv(op.x), deriv(op.y), deriv(op.y2), wf_.n_o, wf_.n_v, 0.0);
        break;
      case TRACE_OO: {
        const double* dL = deriv(op.x);
        if (dL) etg::trace_oo(s, dL, deriv(op.y), wf_.n_J, wf_.n_o);
        else fill(s, deriv(op.y), wf_.n_J, 0.0);
        break;
      }
      case FOCK_OV_VEC:
        etg::fock_ov_vector(s, deriv(op.x), deriv(op.y), wf_.n_o, wf_.n_v);
        break;
      case LADDER:
        wf_.ladder_tangent(op.x, deriv(op.x), op.x2, deriv(op.x2), deriv(op.y), prec);
        break;
    }
  }
}

void Tape::adjoint(Prec prec) {
  hipStream_t s = wf_.stream();
  for (auto it = ops_.rbegin(); it != ops_.rend(); ++it) {
    const Op& op = *it;
    switch (op.kind) {
      case GEMM: {
        const Gemm& g = op.g;
        const double* Cb = deriv(g.C);
        double* Ab = deriv(g.A);
        double* Bb = deriv(g.B);
        const char ta = g.ta ? 'T' : 'N', tb = g.tb ? 'T' : 'N';
        (void)ta;
        (void)tb;
        auto run_pair = [&](bool for_A, int64_t b0, int64_t nb) {
          Gemm h;
          h.batch = nb;
          const double* A = g.A + b0 * g.strideA;
          const double* B = g.B + b0 * g.strideB;
          const double* C = Cb + b0 * g.strideC;
          if (for_A) {
            double* Ao = Ab + b0 * g.strideA;
            if (!g.ta) {   // A_bar (m x k) += alpha Cbar op(B)^T
              h.ta = false; h.tb = !g.tb;
              h.m = g.m; h.n = g.k; h.k = g.n;
              h.A = C; h.lda = g.ldc; h.strideA = g.strideC;
              h.B = B; h.ldb = g.ldb; h.strideB = g.strideB;
            } else {       // A_bar (k x m) += alpha op(B) Cbar^T
              h.ta = g.tb; h.tb = true;
              h.m = g.k; h.n = g.m; h.k = g.n;
              h.A = B; h.lda = g.ldb; h.strideA = g.strideB;
              h.B = C; h.ldb = g.ldc; h.strideB = g.strideC;
            }
            h.C = Ao; h.ldc = g.lda; h.strideC = g.strideA;
          } else {
            double* Bo = Bb + b0 * g.strideB;
            if (!g.tb) {   // B_bar (k x n) += alpha op(A)^T Cbar
              h.ta = !g.ta; h.tb = false;
              h.m = g.k; h.n = g.n; h.k = g.m;
              h.A = A; h.lda = g.lda; h.strideA = g.strideA;
              h.B = C; h.ldb = g.ldc; h.strideB = g.strideC;
            } else {       // B_bar (n x k) += alpha Cbar^T op(A)
              h.ta = true; h.tb = g.ta;
              h.m = g.n; h.n = g.k; h.k = g.m;
              h.A = C; h.lda = g.ldc; h.strideA = g.strideC;
              h.B = A; h.ldb = g.lda; h.strideB = g.strideA;
            }
            h.C = Bo; h.ldc = g.ldb; h.strideC = g.strideB;
          }
          h.alpha = g.alpha;
          h.beta = 1.0;
          wf_.run_gemm(h, prec);
        };
        if (Ab) {
          if (g.batch > 1 && g.strideA == 0) {
            for (int64_t b = 0; b < g.batch; ++b) run_pair(true, b, 1);
          } else {
            run_pair(true, 0, g.batch);
          }
        }
        if (Bb) {
          if (g.batch > 1 && g.strideB == 0) {
            for (int64_t b = 0; b < g.batch; ++b) run_pair(false, b, 1);
          } else {
            run_pair(false, 0, g.batch);
          }
        }
        scale_output(wf_, g, deriv(g.C), g.beta, prec);
        break;
      }
      case SORT4: {
        double* yb = deriv(op.y);
        double* xb = deriv(op.x);
        if (xb) {
          int64_t e[4];
          char inv[5] = {0, 0, 0, 0, 0};
          for (int k = 0; k < 4; ++k) {
            const int q = op.perm[k] - '1';
            e[k] = op.d[q];
            inv[q] = char('1' + k);
          }
          etg::sort4(s, yb, e, inv, xb, op.alpha, 1.0);
        }
        scale(s, yb, dense_size(op.d), op.beta);
        break;
      }
      case SQUAREUP: {
        double* pb = deriv(op.x);
        double* fb = deriv(op.y);
        if (pb) packsum(s, fb, pb, op.N);
        fill(s, fb, op.N * op.N, 0.0);
        break;
      }
      case PACKIN: {
        double* fb = deriv(op.x);
        double* pb = deriv(op.y);
        if (fb) unpack_upper_add(s, pb, fb, op.N, op.alpha, op.diag);
        scale(s, pb, op.N * (op.N + 1) / 2, op.beta);
        break;
      }
      case SYMSUM: {
        double* xb = deriv(op.x);
        double* yb = deriv(op.y);
        if (xb) symmetric_sum_acc(s, yb, xb, op.N, 1.0);
        fill(s, yb, op.N * op.N, 0.0);
        break;
      }
      case SYMPACK: {
        double* xb = deriv(op.x);
        if (xb) symmetrize_and_pack_adjoint(s, deriv(op.y), xb, op.nv, op.no, op.alpha, op.f1, op.f2);
        break;
      }
      case COPY_SUB: {
        double* sb = deriv(op.x);
        double* db = deriv(op.y);
        if (sb) etg::copy_sub(s, db, op.ldd, sb, op.lds, op.m, op.n, op.alpha, 1.0);
        etg::copy_sub(s, db, op.ldd, db, op.ldd, op.m, op.n, 0.0, op.beta);
        break;
      }
      case FILL0:
        fill(s, deriv(op.y), op.n, 0.0);
        break;
      case T1_XY: {
        double* tb = deriv(op.x);
        double* Xb = deriv(op.y);
        double* Yb = deriv(op.y2);
        if (tb) t1_xy_adjoint(s, Xb, Yb, tb, wf_.n_o, wf_.n_v);
        fill(s, Xb, wf_.n_mo * wf_.n_mo, 0.0);
        fill(s, Yb, wf_.n_mo * wf_.n_mo, 0.0);
        break;
      }
      case TRACE_OO: {
        double* Lb = deriv(op.x);
        double* Xb = deriv(op.y);
        if (Lb) trace_oo_adjoint(s, Xb, Lb, wf_.n_J, wf_.n_o);
        fill(s, Xb, wf_.n_J, 0.0);
        break;
      }
      case FOCK_OV_VEC: {
        double* Fb = deriv(op.x);
        double* fb = deriv(op.y);
        if (Fb) fock_ov_vector_adjoint(s, fb, Fb, wf_.n_o, wf_.n_v);
        fill(s, fb, wf_.n_o * wf_.n_v, 0.0);
        break;
      }
      case LADDER:
        wf_.ladder_adjoint(op.x, deriv(op.x), op.x2, deriv(op.x2), deriv(op.y), prec);
        break;
    }
  }
}

}  // namespace etg
