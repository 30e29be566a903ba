// The residual pipeline (T1 dressing -> T1 Fock -> Omega) is written once
// against this interface.  DirectOps runs it (reusing scratch slots);
// Tape runs it once to record every operation with its operands kept in
// distinct HBM buffers, then replays the record forward (tangent: the
// right Jacobian transformation) or backward (adjoint: the left one).
// Every operation is linear in each operand except the products, which are
// bilinear, so each kernel has a tangent and an adjoint rule of its own.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"

namespace etg {

class CCSD;

// Scratch slots of the direct program ((ov)^2-sized, then small ones).
enum ScratchSlot {
  SL_T = 0, SL_U, SL_A, SL_B, SL_C, SL_D, SL_E, SL_F, kBigSlots,
  SL_SMALL0 = kBigSlots, SL_SMALL1, SL_SMALL2, SL_SMALL3, kAllSlots
};

class Ops {
 public:
  virtual ~Ops() = default;
  // scratch buffer: `slot` is the reuse slot of the direct program, `name` its identity on a tape
  virtual double* buf(int slot, const char* name, int64_t elems) = 0;
  virtual void gemm(const Gemm& g, Prec prec) = 0;
  // y = alpha * permute(x) + beta * y   (kernels.hpp sort4 semantics)
  virtual void sort4(const double* x, const int64_t d[4], const char* perm, double* y, double alpha,
                     double beta) = 0;
  virtual void squareup(const double* packed, double* full, int64_t N) = 0;
  virtual void packin(const double* full, double* packed, int64_t N, double alpha, double beta,
                      double diag) = 0;
  virtual void symsum(const double* x, double* y, int64_t N) = 0;
  virtual void sympack(const double* x, double* packed, int64_t n_v, int64_t n_o, double alpha,
                       double f1, double f2) = 0;
  // dst (m x n, ld ldd) = alpha * src (ld lds) + beta * dst
  virtual void copy_sub(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
                        int64_t n, double alpha, double beta) = 0;
  virtual void fill0(double* x, int64_t n) = 0;
  // X = I - t1 (vo block), Y = I + t1^T (ov block), n_mo x n_mo
  virtual void t1_xy(const double* t1, double* X, double* Y) = 0;
  // XJ(J) = sum_k Loo(J, k, k)
  virtual void trace_oo(const double* Loo, double* XJ) = 0;
  // f(b + v j) = F(j, o + b)
  virtual void fock_ov_vec(const double* F, double* f) = 0;
  // packed Omega doubles += ladder term sum_cd g_acbd t_cidj (g from the (J, c, a) vv block)
  virtual void ladder(const double* t2, const double* LvvT, double* om2, Prec prec) = 0;

  void gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
            int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
            Prec prec) {
    Gemm g;
    g.ta = (ta == 'T');
    g.tb = (tb == 'T');
    g.m = m; g.n = n; g.k = k;
    g.alpha = alpha; g.beta = beta;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
    gemm(g, prec);
  }
  void gemm_batched(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                    const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb,
                    int64_t sB, double beta, double* C, int64_t ldc, int64_t sC, int64_t batch,
                    Prec prec) {
    Gemm g;
    g.ta = (ta == 'T');
    g.tb = (tb == 'T');
    g.m = m; g.n = n; g.k = k;
    g.alpha = alpha; g.beta = beta;
    g.A = A; g.lda = lda; g.strideA = sA;
    g.B = B; g.ldb = ldb; g.strideB = sB;
    g.C = C; g.ldc = ldc; g.strideC = sC;
    g.batch = batch;
    gemm(g, prec);
  }
};

// Runs the operations directly on the handle's stream.
class DirectOps : public Ops {
 public:
  explicit DirectOps(CCSD& wf) : wf_(wf) {}
  double* buf(int slot, const char* name, int64_t elems) override;
  void gemm(const Gemm& g, Prec prec) override;
  using Ops::gemm;
  void sort4(const double* x, const int64_t d[4], const char* perm, double* y, double alpha,
             double beta) override;
  void squareup(const double* packed, double* full, int64_t N) override;
  void packin(const double* full, double* packed, int64_t N, double alpha, double beta,
              double diag) override;
  void symsum(const double* x, double* y, int64_t N) override;
  void sympack(const double* x, double* packed, int64_t n_v, int64_t n_o, double alpha, double f1,
               double f2) override;
  void copy_sub(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t n,
                double alpha, double beta) override;
  void fill0(double* x, int64_t n) override;
  void t1_xy(const double* t1, double* X, double* Y) override;
  void trace_oo(const double* Loo, double* XJ) override;
  void fock_ov_vec(const double* F, double* f) override;
  void ladder(const double* t2, const double* LvvT, double* om2, Prec prec) override;

 protected:
  CCSD& wf_;
};

// Records while running; replays forward (tangent) or backward (adjoint).
class Tape : public DirectOps {
 public:
  explicit Tape(CCSD& wf) : DirectOps(wf) {}
  double* buf(int slot, const char* name, int64_t elems) override;
  void gemm(const Gemm& g, Prec prec) override;
  using Ops::gemm;
  void sort4(const double* x, const int64_t d[4], const char* perm, double* y, double alpha,
             double beta) override;
  void squareup(const double* packed, double* full, int64_t N) override;
  void packin(const double* full, double* packed, int64_t N, double alpha, double beta,
              double diag) override;
  void symsum(const double* x, double* y, int64_t N) override;
  void sympack(const double* x, double* packed, int64_t n_v, int64_t n_o, double alpha, double f1,
               double f2) override;
  void copy_sub(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t n,
                double alpha, double beta) override;
  void fill0(double* x, int64_t n) override;
  void t1_xy(const double* t1, double* X, double* Y) override;
  void trace_oo(const double* Loo, double* XJ) override;
  void fock_ov_vec(const double* F, double* f) override;
  void ladder(const double* t2, const double* LvvT, double* om2, Prec prec) override;

  // Buffers that exist outside the tape.  Constant ones carry no derivative;
  // `internal` ones are written by the recorded program itself, so their
  // derivatives live only inside a pass, as those of the tape's own buffers.
  void register_buffer(const double* base, int64_t elems, bool constant, const char* name,
                       bool internal = false);
  void clear();
  // Every derivative buffer holds up to `nb` vectors: vector k of the
  // derivative of primal element p sits at deriv(p) + k * dstride(p).
  // Forward replay: derivatives of every recorded output along the nb input
  // derivatives already placed in the derivative buffers.
  void tangent(Prec prec, int64_t nb = 1);
  // Backward replay: adjoints of every recorded input from the nb adjoints
  // placed in the output's derivative buffer.
  void adjoint(Prec prec, int64_t nb = 1);
  // Room for nb derivative vectors per buffer; true when a buffer moved.
  bool reserve_derivatives(int64_t nb);
  // After recording: which buffers carry a derivative (an intermediate written
  // only from constant data carries none, and the passes skip its terms), which
  // primal values the passes read (operands of products whose other operand
  // carries a derivative), and which of those are rebuilt on demand by their
  // bandwidth operations from data that stays in HBM instead of being stored.
  // Every other intermediate's memory is released.  `rebuild_slots` buffers of
  // the largest rebuilt size hold the rebuilt values during a pass.
  void finalize(int rebuild_slots = 4);
  // Address of the primal values at p during a pass (p itself unless p's buffer is rebuilt).
  const double* primal(const double* p);
  // Zeroes the first nb derivative vectors of the buffers that exist outside
  // the tape (allocating on first use); the passes zero each intermediate's
  // derivative where its lifetime starts.
  void zero_derivatives(int64_t nb = 1);
  // Derivative buffer matching a primal pointer (nullptr for constants), vector 0.
  double* deriv(const double* p);
  // Element distance between consecutive derivative vectors of p's buffer.
  int64_t dstride(const double* p) const;
  // Elements of one derivative vector over all buffers (memory per batched vector).
  int64_t derivative_elems();
  // The same without sharing (one derivative buffer per intermediate).
  int64_t unshared_derivative_elems() const;
  // Elements of the recorded intermediates (the primal record).
  int64_t record_elems() const;
  bool empty() const { return ops_.empty(); }

  enum Kind { GEMM, SORT4, SQUAREUP, PACKIN, SYMSUM, SYMPACK, COPY_SUB, FILL0, T1_XY, TRACE_OO,
              FOCK_OV_VEC, LADDER };
  struct Op {
    Kind kind;
    Gemm g;
    Prec prec = Prec::F64;
    const double* x = nullptr;
    double* y = nullptr;
    const double* x2 = nullptr;
    double* y2 = nullptr;
    int64_t d[4] = {0, 0, 0, 0};
    char perm[5] = {0, 0, 0, 0, 0};
    double alpha = 1.0, beta = 0.0, f1 = 0.0, f2 = 0.0, diag = 1.0;
    int64_t N = 0, m = 0, n = 0, lds = 0, ldd = 0, nv = 0, no = 0;
  };

 private:
  // The derivative of an intermediate (a buffer the tape allocated) is live
  // from the first to the last operation that touches the buffer; that
  // interval is the same for the forward and the backward replay, so
  // intermediates whose intervals do not overlap share one derivative slot.
  // The amplitudes, the residual and h (whose derivatives are the passes'
  // inputs and outputs) keep derivative buffers of their own.
  struct Buf {
    const double* base;
    int64_t elems;
    bool constant;
    bool intermediate = false;   // derivative lives inside a pass only
    bool active = true;          // carries a derivative (set by finalize)
    bool rebuild = false;        // primal values rebuilt on demand by `recipe`
    std::vector<int> recipe;     // the operations that write the buffer, in order
    int last_write = -1;
    int first = -1, last = -1;   // operation interval (intermediates)
    int slot = -1;               // shared derivative slot (intermediates)
    std::string name;
    DeviceBuffer d;  // derivative (the passes' inputs and outputs)
  };
  int find(const double* p) const;
  // operands an operation reads (up to 3) and writes (up to 2), null-padded
  static void op_io(const Op& op, const double* in[3], const double* out[2]);
  // rebuilt primal values: slot of buffer k (rebuilt if not current), pinned for the current operation
  int ensure(int k);
  void run_recipe_op(const Op& op, int k, double* base);
  void reset_rebuilt();
  void unpin_all();
  std::vector<char> op_active_;          // some output of the operation carries a derivative
  struct RebuildSlot {
    DeviceBuffer mem;
    int owner = -1;
    uint64_t stamp = 0;
    int pins = 0;
  };
  std::vector<RebuildSlot> rslots_;
  uint64_t rclock_ = 0;
  bool finalized_ = false;
  void tangent_gemm(const Gemm& g, Prec prec, int64_t nb);
  void adjoint_gemm(const Gemm& g, Prec prec, int64_t nb);
  void plan_derivatives();
  void zero_starting(const std::vector<int>& bufs, int64_t nb);
  std::vector<Op> ops_;
  std::vector<std::unique_ptr<Buf>> bufs_;
  std::map<std::string, DeviceBuffer> named_;
  size_t planned_ops_ = size_t(-1);     // ops_.size() when the slots were planned
  std::vector<DeviceBuffer> slots_;
  std::vector<int64_t> slot_elems_;
  std::vector<std::vector<int>> starts_fwd_, starts_bwd_;   // per operation: lifetimes starting there
};

}  // namespace etg
