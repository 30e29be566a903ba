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
  register_buffer(p, elems, false, name, true);
  return p;
}

void Tape::register_buffer(const double* base, int64_t elems, bool constant, const char* name,
                           bool internal) {
  auto b = std::make_unique<Buf>();
  b->base = base;
  b->elems = elems;
  b->constant = constant;
  b->intermediate = internal && !constant;
  b->name = name;
  bufs_.push_back(std::move(b));
}

void Tape::clear() {
  ops_.clear();
  bufs_.clear();
  named_.clear();
  slots_.clear();
  slot_elems_.clear();
  rslots_.clear();
  op_active_.clear();
  finalized_ = false;
  planned_ops_ = size_t(-1);
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
  if (b.constant || !b.active) return nullptr;
  const size_t need = size_t(std::max<int64_t>(b.elems, 1)) * sizeof(double);
  if (b.intermediate) {
    if (b.slot < 0 || slots_[size_t(b.slot)].bytes() < need) fail("tape: derivatives not initialised");
    return slots_[size_t(b.slot)].as<double>() + (p - b.base);
  }
  if (b.d.bytes() < need) fail("tape: derivatives not initialised");
  return b.d.as<double>() + (p - b.base);
}

int64_t Tape::dstride(const double* p) const {
  return std::max<int64_t>(bufs_[size_t(find(p))]->elems, 1);
}

// Lifetimes of the intermediates' derivatives and their slots: intervals in
// order of their first operation, each placed in the smallest free slot that
// holds it, else the largest free slot (grown), else a new slot.
void Tape::plan_derivatives() {
  if (planned_ops_ == ops_.size()) return;
  for (auto& bp : bufs_) bp->first = bp->last = -1;
  auto touch = [&](const double* p, int i) {
    if (p == nullptr) return;
    Buf& b = *bufs_[size_t(find(p))];
    if (b.constant || !b.intermediate || !b.active) return;
    if (b.first < 0) b.first = i;
    b.last = i;
  };
  for (size_t i = 0; i < ops_.size(); ++i) {
    const Op& op = ops_[i];
    if (op.kind == GEMM) {
      touch(op.g.A, int(i));
      touch(op.g.B, int(i));
      touch(op.g.C, int(i));
    } else {
      touch(op.x, int(i));
      touch(op.x2, int(i));
      touch(op.y, int(i));
      touch(op.y2, int(i));
    }
  }
  std::vector<int> order;
  for (size_t k = 0; k < bufs_.size(); ++k)
    if (bufs_[k]->intermediate && bufs_[k]->active && bufs_[k]->first >= 0) order.push_back(int(k));
  std::sort(order.begin(), order.end(), [&](int a, int b) { return bufs_[size_t(a)]->first < bufs_[size_t(b)]->first; });
  slots_.clear();
  slot_elems_.clear();
  std::vector<int> busy_until;          // last operation of each slot's current occupant
  starts_fwd_.assign(ops_.size(), {});
  starts_bwd_.assign(ops_.size(), {});
  for (auto& bp : bufs_) bp->slot = -1;
  for (int k : order) {
    Buf& b = *bufs_[size_t(k)];
    const int64_t e = std::max<int64_t>(b.elems, 1);
    int fit = -1, grow = -1;
    for (size_t s = 0; s < slot_elems_.size(); ++s) {
      if (busy_until[s] >= b.first) continue;
      if (slot_elems_[s] >= e) {
        if (fit < 0 || slot_elems_[s] < slot_elems_[size_t(fit)]) fit = int(s);
      } else if (grow < 0 || slot_elems_[s] > slot_elems_[size_t(grow)]) {
        grow = int(s);
      }
    }
    int s = fit >= 0 ? fit : grow;
    if (s < 0) {
      s = int(slot_elems_.size());
      slot_elems_.push_back(0);
      busy_until.push_back(-1);
    }
    slot_elems_[size_t(s)] = std::max(slot_elems_[size_t(s)], e);
    busy_until[size_t(s)] = b.last;
    b.slot = s;
    starts_fwd_[size_t(b.first)].push_back(k);
    starts_bwd_[size_t(b.last)].push_back(k);
  }
  slots_.resize(slot_elems_.size());
  planned_ops_ = ops_.size();
}

int64_t Tape::derivative_elems() {
  plan_derivatives();
  int64_t e = 0;
  for (const auto& bp : bufs_)
    if (!bp->constant && !bp->intermediate && bp->active) e += std::max<int64_t>(bp->elems, 1);
  for (int64_t s : slot_elems_) e += s;
  return e;
}

int64_t Tape::unshared_derivative_elems() const {
  int64_t e = 0;
  for (const auto& bp : bufs_)
    if (!bp->constant && bp->active) e += std::max<int64_t>(bp->elems, 1);
  return e;
}

int64_t Tape::record_elems() const {
  int64_t e = 0;
  for (const auto& kv : named_) e += int64_t(kv.second.bytes() / sizeof(double));
  for (const auto& r : rslots_) e += int64_t(r.mem.bytes() / sizeof(double));
  return e;
}

bool Tape::reserve_derivatives(int64_t nb) {
  plan_derivatives();
  bool moved = false;
  for (auto& bp : bufs_) {
    Buf& b = *bp;
    if (b.constant || b.intermediate || !b.active) continue;
    moved |= b.d.resize(size_t(nb) * size_t(std::max<int64_t>(b.elems, 1)) * sizeof(double));
  }
  for (size_t s = 0; s < slots_.size(); ++s)
    moved |= slots_[s].resize(size_t(nb) * size_t(slot_elems_[s]) * sizeof(double));
  return moved;
}

void Tape::zero_derivatives(int64_t nb) {
  reserve_derivatives(nb);
  for (auto& bp : bufs_) {
    Buf& b = *bp;
    if (b.constant || b.intermediate || !b.active) continue;
    fill(wf_.stream(), b.d.as<double>(), nb * std::max<int64_t>(b.elems, 1), 0.0);
  }
}

// nb derivative vectors of each listed intermediate (contiguous from its slot's start).
void Tape::zero_starting(const std::vector<int>& bufs, int64_t nb) {
  for (int k : bufs) {
    const Buf& b = *bufs_[size_t(k)];
    fill(wf_.stream(), slots_[size_t(b.slot)].as<double>(), nb * std::max<int64_t>(b.elems, 1), 0.0);
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
// Tape: what the passes need of the record
// ---------------------------------------------------------------------------

void Tape::op_io(const Op& op, const double* in[3], const double* out[2]) {
  in[0] = in[1] = in[2] = nullptr;
  out[0] = out[1] = nullptr;
  switch (op.kind) {
    case GEMM:
      in[0] = op.g.A;
      in[1] = op.g.B;
      if (op.g.beta != 0.0) in[2] = op.g.C;
      out[0] = op.g.C;
      break;
    case SORT4:
    case PACKIN:
    case COPY_SUB:
      in[0] = op.x;
      if (op.beta != 0.0) in[1] = op.y;
      out[0] = op.y;
      break;
    case SQUAREUP:
    case SYMSUM:
    case TRACE_OO:
    case FOCK_OV_VEC:
      in[0] = op.x;
      out[0] = op.y;
      break;
    case SYMPACK:                       // accumulates into the packed output
      in[0] = op.x;
      in[1] = op.y;
      out[0] = op.y;
      break;
    case FILL0:
      out[0] = op.y;
      break;
    case T1_XY:
      in[0] = op.x;
      out[0] = op.y;
      out[1] = op.y2;
      break;
    case LADDER:
      in[0] = op.x;
      in[1] = op.x2;
      in[2] = op.y;
      out[0] = op.y;
      break;
  }
}

void Tape::finalize(int rebuild_slots) {
  const int nops = int(ops_.size());
  const int nbuf = int(bufs_.size());
  // 1. derivatives: the passes' inputs (amplitudes, h) carry one; an output
  //    carries one if any operation that writes it reads one that does
  std::vector<char> active(size_t(nbuf), 0);
  for (int k = 0; k < nbuf; ++k) {
    const Buf& b = *bufs_[size_t(k)];
    active[size_t(k)] = !b.constant && !b.intermediate;
  }
  for (auto& bp : bufs_) {
    bp->recipe.clear();
    bp->last_write = -1;
    bp->rebuild = false;
  }
  op_active_.assign(size_t(nops), 0);
  std::vector<std::vector<int>> reads(static_cast<size_t>(nops)), writes(static_cast<size_t>(nops));
  for (int i = 0; i < nops; ++i) {
    const double* in[3];
    const double* out[2];
    op_io(ops_[size_t(i)], in, out);
    bool live = false;
    for (const double* p : in)
      if (p) {
        const int k = find(p);
        reads[size_t(i)].push_back(k);
        live |= bool(active[size_t(k)]);
      }
    for (const double* p : out)
      if (p) {
        const int k = find(p);
        writes[size_t(i)].push_back(k);
        if (live) active[size_t(k)] = 1;
        bufs_[size_t(k)]->recipe.push_back(i);
        bufs_[size_t(k)]->last_write = i;
      }
  }
  for (int k = 0; k < nbuf; ++k) bufs_[size_t(k)]->active = bufs_[size_t(k)]->constant ? false : bool(active[size_t(k)]);
  for (int i = 0; i < nops; ++i)
    for (int k : writes[size_t(i)]) op_active_[size_t(i)] |= char(active[size_t(k)]);
  // 2. primal values the passes read: a product's operand pairs with the other
  //    operand's derivative (tangent and adjoint alike); the ladder reads its
  //    amplitudes and integrals
  std::vector<char> needed(size_t(nbuf), 0);
  std::vector<int> first_use(size_t(nbuf), nops);
  for (int i = 0; i < nops; ++i) {
    const Op& op = ops_[size_t(i)];
    if (!op_active_[size_t(i)]) continue;
    if (op.kind == GEMM) {
      const int a = find(op.g.A), b = find(op.g.B);
      if (bufs_[size_t(b)]->active) { needed[size_t(a)] = 1; first_use[size_t(a)] = std::min(first_use[size_t(a)], i); }
      if (bufs_[size_t(a)]->active) { needed[size_t(b)] = 1; first_use[size_t(b)] = std::min(first_use[size_t(b)], i); }
    } else if (op.kind == LADDER) {
      for (const double* p : {op.x, op.x2}) {
        const int k = find(p);
        needed[size_t(k)] = 1;
        first_use[size_t(k)] = std::min(first_use[size_t(k)], i);
      }
    }
  }
  // the passes read each primal operand as its last write left it
  for (int k = 0; k < nbuf; ++k)
    if (needed[size_t(k)] && bufs_[size_t(k)]->last_write > first_use[size_t(k)])
      fail("tape: a product reads a buffer that is written again later: " + bufs_[size_t(k)]->name);
  // 3. rebuilt instead of stored: an intermediate written only by bandwidth
  //    operations whose inputs stay in HBM or are rebuilt themselves, each input
  //    final when it is read (taken in the order of the buffers' first writes)
  auto bandwidth_op = [](Kind kd) {
    return kd == SORT4 || kd == SQUAREUP || kd == COPY_SUB || kd == FILL0 || kd == SYMSUM || kd == PACKIN ||
           kd == SYMPACK;
  };
  std::vector<int> order;
  for (int k = 0; k < nbuf; ++k)
    if (bufs_[size_t(k)]->intermediate && !bufs_[size_t(k)]->recipe.empty() && named_.count(bufs_[size_t(k)]->name))
      order.push_back(k);
  std::sort(order.begin(), order.end(),
            [&](int a, int b) { return bufs_[size_t(a)]->recipe[0] < bufs_[size_t(b)]->recipe[0]; });
  int64_t rebuild_elems = 0;
  for (int k : order) {
    Buf& b = *bufs_[size_t(k)];
    bool ok = true;
    for (int i : b.recipe) {
      if (!bandwidth_op(ops_[size_t(i)].kind)) { ok = false; break; }
      for (int q : reads[size_t(i)]) {
        if (q == k) continue;
        const Buf& in = *bufs_[size_t(q)];
        const bool stays = !named_.count(in.name) || in.rebuild;   // registered outside the tape, or rebuilt
        if (!stays || in.last_write >= i) { ok = false; break; }
      }
      if (!ok) break;
    }
    b.rebuild = ok;
    if (ok) rebuild_elems = std::max(rebuild_elems, std::max<int64_t>(b.elems, 1));
  }
  // 4. keep what is read and cannot be rebuilt; release the rest
  for (int k : order) {
    Buf& b = *bufs_[size_t(k)];
    if (needed[size_t(k)] && !b.rebuild) continue;
    named_[b.name].release();
  }
  for (int k = 0; k < nbuf; ++k) {           // intermediates with no write (never read as primal)
    Buf& b = *bufs_[size_t(k)];
    if (b.intermediate && b.recipe.empty() && named_.count(b.name) && !needed[size_t(k)]) named_[b.name].release();
  }
  rslots_.clear();
  rslots_.resize(rebuild_elems > 0 ? size_t(std::max(2, rebuild_slots)) : 0);
  for (auto& r : rslots_) r.mem.resize(sizeof(double) * size_t(rebuild_elems));
  planned_ops_ = size_t(-1);
  finalized_ = true;
}

void Tape::reset_rebuilt() {
  for (auto& r : rslots_) {
    r.owner = -1;
    r.pins = 0;
    r.stamp = 0;
  }
  rclock_ = 0;
}

void Tape::unpin_all() {
  for (auto& r : rslots_) r.pins = 0;
}

const double* Tape::primal(const double* p) {
  if (!finalized_) return p;
  const int k = find(p);
  const Buf& b = *bufs_[size_t(k)];
  if (!b.rebuild) {
    if (b.intermediate && named_.count(b.name) && named_[b.name].bytes() == 0)
      fail("tape: primal values of " + b.name + " were released");
    return p;
  }
  const int s = ensure(k);
  return rslots_[size_t(s)].mem.as<double>() + (p - b.base);
}

// Slot holding buffer k's primal values: the slot it is in, else the least
// recently used unpinned slot, filled by replaying k's operations (their kernel
// launches follow every earlier read of the slot on the stream).
int Tape::ensure(int k) {
  for (size_t s = 0; s < rslots_.size(); ++s)
    if (rslots_[s].owner == k) {
      rslots_[s].stamp = ++rclock_;
      ++rslots_[s].pins;
      return int(s);
    }
  int pick = -1;
  for (size_t s = 0; s < rslots_.size(); ++s) {
    if (rslots_[s].pins > 0) continue;
    if (pick < 0 || rslots_[s].stamp < rslots_[size_t(pick)].stamp) pick = int(s);
  }
  if (pick < 0) fail("tape: every rebuild slot is in use");
  RebuildSlot& r = rslots_[size_t(pick)];
  r.owner = -1;
  ++r.pins;
  const Buf& b = *bufs_[size_t(k)];
  if (size_t(std::max<int64_t>(b.elems, 1)) * sizeof(double) > r.mem.bytes()) fail("tape: rebuild slot too small for " + b.name);
  for (int i : b.recipe) run_recipe_op(ops_[size_t(i)], k, r.mem.as<double>());
  r.owner = k;
  r.stamp = ++rclock_;
  return pick;
}

void Tape::run_recipe_op(const Op& op, int k, double* base) {
  hipStream_t s = wf_.stream();
  const Buf& b = *bufs_[size_t(k)];
  auto out = [&](const double* p) { return base + (p - b.base); };
  std::vector<int> pinned;
  auto in = [&](const double* p) -> const double* {
    if (p >= b.base && p < b.base + std::max<int64_t>(b.elems, 1)) return out(p);
    const int q = find(p);
    if (!bufs_[size_t(q)]->rebuild) return p;
    const int sl = ensure(q);
    pinned.push_back(sl);
    return rslots_[size_t(sl)].mem.as<double>() + (p - bufs_[size_t(q)]->base);
  };
  switch (op.kind) {
    case SORT4: etg::sort4(s, in(op.x), op.d, op.perm, out(op.y), op.alpha, op.beta); break;
    case SQUAREUP: etg::squareup(s, in(op.x), out(op.y), op.N); break;
    case COPY_SUB: etg::copy_sub(s, in(op.x), op.lds, out(op.y), op.ldd, op.m, op.n, op.alpha, op.beta); break;
    case FILL0: fill(s, out(op.y), op.n, 0.0); break;
    case SYMSUM: etg::symmetric_sum(s, in(op.x), out(op.y), op.N); break;
    case PACKIN: etg::packin(s, in(op.x), out(op.y), op.N, op.alpha, op.beta, op.diag); break;
    case SYMPACK: etg::symmetrize_and_pack(s, in(op.x), out(op.y), op.nv, op.no, op.alpha, op.f1, op.f2); break;
    default: fail("tape: not a bandwidth operation in a rebuild recipe");
  }
  for (int sl : pinned) --rslots_[size_t(sl)].pins;    // queued reads precede later writes on the stream
}

// ---------------------------------------------------------------------------
// Tape: replays
// ---------------------------------------------------------------------------

namespace {
int64_t dense_size(const int64_t d[4]) { return d[0] * d[1] * d[2] * d[3]; }

}  // namespace

// dC_k = beta dC_k + alpha (op(dA_k) op(B) + op(A) op(dB_k)), k < nb.  A product
// of one matrix pair runs as one strided-batched product over the vectors (the
// primal operand at stride 0); an already batched product runs once per vector.
void Tape::tangent_gemm(const Gemm& g, Prec prec, int64_t nb) {
  double* dC = deriv(g.C);
  const double* dA = deriv(g.A);
  const double* dB = deriv(g.B);
  const int64_t sC = dstride(g.C), sA = dA ? dstride(g.A) : 0, sB = dB ? dstride(g.B) : 0;
  const bool vec_batch = g.batch == 1 && nb > 1;
  const int64_t reps = vec_batch ? 1 : nb;
  // primal operands: B pairs with dA, A with dB
  const double* pA = dB ? primal(g.A) : g.A;
  const double* pB = dA ? primal(g.B) : g.B;
  for (int64_t k = 0; k < reps; ++k) {
    Gemm base = g;
    base.a_tag = 0;
    base.A = pA;
    base.B = pB;
    base.C = dC + k * sC;
    if (vec_batch) {
      base.batch = nb;
      base.strideC = sC;
      base.strideA = 0;
      base.strideB = 0;
    }
    if (g.beta != 1.0) {
      Gemm z = base;
      z.k = 0;
      z.beta = g.beta;
      wf_.run_gemm(z, prec);
    }
    if (dA) {
      Gemm h = base;
      h.A = dA + k * sA;
      if (vec_batch) h.strideA = sA;
      h.beta = 1.0;
      wf_.run_gemm(h, prec);
    }
    if (dB) {
      Gemm h = base;
      h.B = dB + k * sB;
      if (vec_batch) h.strideB = sB;
      h.beta = 1.0;
      wf_.run_gemm(h, prec);
    }
  }
}

// A_bar_k += alpha C_bar_k op(B)^T (or its transpose), B_bar_k likewise, then
// C_bar_k *= beta; batched over the vectors as in tangent_gemm.
void Tape::adjoint_gemm(const Gemm& g0, Prec prec, int64_t nb) {
  double* Cb0 = deriv(g0.C);
  double* Ab0 = deriv(g0.A);
  double* Bb0 = deriv(g0.B);
  // primal operands: A_bar reads B, B_bar reads A
  Gemm g = g0;
  g.a_tag = 0;
  if (Bb0) g.A = primal(g0.A);
  if (Ab0) g.B = primal(g0.B);
  const int64_t sC = dstride(g0.C), sA = Ab0 ? dstride(g0.A) : 0, sB = Bb0 ? dstride(g0.B) : 0;
  const bool vec_batch = g.batch == 1 && nb > 1;
  const int64_t reps = vec_batch ? 1 : nb;
  for (int64_t v = 0; v < reps; ++v) {
    const double* Cb = Cb0 + v * sC;
    double* Ab = Ab0 ? Ab0 + v * sA : nullptr;
    double* Bb = Bb0 ? Bb0 + v * sB : nullptr;
    // adjoint product for entries b0 .. b0 + nbat of the recorded batch
    auto pair = [&](bool for_A, int64_t b0, int64_t nbat) {
      Gemm h;
      h.batch = nbat;
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
      if (vec_batch) {             // one product per vector: adjoints at their strides, primal shared
        h.batch = nb;
        h.strideA = (h.A == C) ? sC : 0;
        h.strideB = (h.B == C) ? sC : 0;
        h.strideC = for_A ? sA : sB;
      }
      return h;
    };
    // An operand shared by every entry of a recorded batch (stride 0) collects
    // the sum of the entries' adjoints: one batched product into per-entry
    // partials, then their sum (or entry by entry when the partials would be large).
    auto shared = [&](bool for_A) {
      Gemm h = pair(for_A, 0, g.batch);
      const int64_t mn = h.m * h.n;
      if (double(g.batch) * double(mn) > double(int64_t(1) << 27)) {
        for (int64_t b = 0; b < g.batch; ++b) wf_.run_gemm(pair(for_A, b, 1), prec);
        return;
      }
      double* part = wf_.gemm_partials(g.batch * mn);
      double* out = h.C;
      const int64_t ldo = h.ldc;
      h.C = part;
      h.ldc = h.m;
      h.strideC = mn;
      h.beta = 0.0;
      wf_.run_gemm(h, prec);
      batch_sum_acc(wf_.stream(), part, g.batch, h.m, h.n, out, ldo);
    };
    if (Ab) {
      if (g.batch > 1 && g.strideA == 0) shared(true);
      else wf_.run_gemm(pair(true, 0, g.batch), prec);
    }
    if (Bb) {
      if (g.batch > 1 && g.strideB == 0) shared(false);
      else wf_.run_gemm(pair(false, 0, g.batch), prec);
    }
    if (g.beta != 1.0) {
      Gemm z = g;
      z.a_tag = 0;
      z.k = 0;
      z.C = Cb0 + v * sC;
      z.beta = g.beta;
      if (vec_batch) {
        z.batch = nb;
        z.strideC = sC;
      }
      wf_.run_gemm(z, prec);
    }
  }
}

void Tape::tangent(Prec prec, int64_t nb) {
  hipStream_t s = wf_.stream();
  auto D = [&](const double* p, int64_t k) -> double* {
    double* d = deriv(p);
    return d ? d + k * dstride(p) : nullptr;
  };
  plan_derivatives();
  reset_rebuilt();
  for (size_t i = 0; i < ops_.size(); ++i) {
    const Op& op = ops_[i];
    zero_starting(starts_fwd_[i], nb);
    if (!op_active_.empty() && !op_active_[i]) continue;   // no output carries a derivative
    unpin_all();
    if (op.kind == GEMM) {
      tangent_gemm(op.g, prec, nb);
      continue;
    }
    if (op.kind == LADDER) {
      wf_.ladder_tangent(op.x, deriv(op.x), dstride(op.x), op.x2, deriv(op.x2), dstride(op.x2),
                         deriv(op.y), dstride(op.y), nb, prec);
      continue;
    }
    for (int64_t k = 0; k < nb; ++k) {
      switch (op.kind) {
        case SORT4: {
          double* dy = D(op.y, k);
          const double* dx = D(op.x, k);
          if (dx) etg::sort4(s, dx, op.d, op.perm, dy, op.alpha, op.beta);
          else scale(s, dy, dense_size(op.d), op.beta);
          break;
        }
        case SQUAREUP: {
          const double* dp = D(op.x, k);
          double* df = D(op.y, k);
          if (dp) etg::squareup(s, dp, df, op.N);
          else fill(s, df, op.N * op.N, 0.0);
          break;
        }
        case PACKIN: {
          const double* df = D(op.x, k);
          double* dp = D(op.y, k);
          if (df) etg::packin(s, df, dp, op.N, op.alpha, op.beta, op.diag);
          else scale(s, dp, op.N * (op.N + 1) / 2, op.beta);
          break;
        }
        case SYMSUM: {
          const double* dx = D(op.x, k);
          double* dy = D(op.y, k);
          if (dx) etg::symmetric_sum(s, dx, dy, op.N);
          else fill(s, dy, op.N * op.N, 0.0);
          break;
        }
        case SYMPACK: {
          const double* dx = D(op.x, k);
          if (dx) etg::symmetrize_and_pack(s, dx, D(op.y, k), op.nv, op.no, op.alpha, op.f1, op.f2);
          break;
        }
        case COPY_SUB: {
          const double* dsrc = D(op.x, k);
          double* ddst = D(op.y, k);
          if (dsrc) etg::copy_sub(s, dsrc, op.lds, ddst, op.ldd, op.m, op.n, op.alpha, op.beta);
          else etg::copy_sub(s, ddst, op.ldd, ddst, op.ldd, op.m, op.n, 0.0, op.beta);
          break;
        }
        case FILL0:
          fill(s, D(op.y, k), op.n, 0.0);
          break;
        case T1_XY:
          etg::t1_xy(s, D(op.x, k), D(op.y, k), D(op.y2, k), wf_.n_o, wf_.n_v, 0.0);
          break;
        case TRACE_OO: {
          const double* dL = D(op.x, k);
          if (dL) etg::trace_oo(s, dL, D(op.y, k), wf_.n_J, wf_.n_o);
          else fill(s, D(op.y, k), wf_.n_J, 0.0);
          break;
        }
        case FOCK_OV_VEC:
          etg::fock_ov_vector(s, D(op.x, k), D(op.y, k), wf_.n_o, wf_.n_v);
          break;
        case GEMM:
        case LADDER:
          break;
      }
    }
  }
}

void Tape::adjoint(Prec prec, int64_t nb) {
  hipStream_t s = wf_.stream();
  auto D = [&](const double* p, int64_t k) -> double* {
    double* d = deriv(p);
    return d ? d + k * dstride(p) : nullptr;
  };
  plan_derivatives();
  reset_rebuilt();
  for (size_t i = ops_.size(); i-- > 0;) {
    const Op& op = ops_[i];
    zero_starting(starts_bwd_[i], nb);
    if (!op_active_.empty() && !op_active_[i]) continue;   // no output carries a derivative
    unpin_all();
    if (op.kind == GEMM) {
      adjoint_gemm(op.g, prec, nb);
      continue;
    }
    if (op.kind == LADDER) {
      wf_.ladder_adjoint(op.x, deriv(op.x), dstride(op.x), op.x2, deriv(op.x2), dstride(op.x2),
                         deriv(op.y), dstride(op.y), nb, prec);
      continue;
    }
    for (int64_t k = 0; k < nb; ++k) {
      switch (op.kind) {
        case SORT4: {
          double* yb = D(op.y, k);
          double* xb = D(op.x, k);
          if (xb) {
            int64_t e[4];
            char inv[5] = {0, 0, 0, 0, 0};
            for (int q = 0; q < 4; ++q) {
              const int r = op.perm[q] - '1';
              e[q] = op.d[r];
              inv[r] = char('1' + q);
            }
            etg::sort4(s, yb, e, inv, xb, op.alpha, 1.0);
          }
          scale(s, yb, dense_size(op.d), op.beta);
          break;
        }
        case SQUAREUP: {
          double* pb = D(op.x, k);
          double* fb = D(op.y, k);
          if (pb) packsum(s, fb, pb, op.N);
          fill(s, fb, op.N * op.N, 0.0);
          break;
        }
        case PACKIN: {
          double* fb = D(op.x, k);
          double* pb = D(op.y, k);
          if (fb) unpack_upper_add(s, pb, fb, op.N, op.alpha, op.diag);
          scale(s, pb, op.N * (op.N + 1) / 2, op.beta);
          break;
        }
        case SYMSUM: {
          double* xb = D(op.x, k);
          double* yb = D(op.y, k);
          if (xb) symmetric_sum_acc(s, yb, xb, op.N, 1.0);
          fill(s, yb, op.N * op.N, 0.0);
          break;
        }
        case SYMPACK: {
          double* xb = D(op.x, k);
          if (xb) symmetrize_and_pack_adjoint(s, D(op.y, k), xb, op.nv, op.no, op.alpha, op.f1, op.f2);
          break;
        }
        case COPY_SUB: {
          double* sb = D(op.x, k);
          double* db = D(op.y, k);
          if (sb) etg::copy_sub(s, db, op.ldd, sb, op.lds, op.m, op.n, op.alpha, 1.0);
          etg::copy_sub(s, db, op.ldd, db, op.ldd, op.m, op.n, 0.0, op.beta);
          break;
        }
        case FILL0:
          fill(s, D(op.y, k), op.n, 0.0);
          break;
        case T1_XY: {
          double* tb = D(op.x, k);
          double* Xb = D(op.y, k);
          double* Yb = D(op.y2, k);
          if (tb) t1_xy_adjoint(s, Xb, Yb, tb, wf_.n_o, wf_.n_v);
          fill(s, Xb, wf_.n_mo * wf_.n_mo, 0.0);
          fill(s, Yb, wf_.n_mo * wf_.n_mo, 0.0);
          break;
        }
        case TRACE_OO: {
          double* Lb = D(op.x, k);
          double* Xb = D(op.y, k);
          if (Lb) trace_oo_adjoint(s, Xb, Lb, wf_.n_J, wf_.n_o);
          fill(s, Xb, wf_.n_J, 0.0);
          break;
        }
        case FOCK_OV_VEC: {
          double* Fb = D(op.x, k);
          double* fb = D(op.y, k);
          if (Fb) fock_ov_vector_adjoint(s, fb, Fb, wf_.n_o, wf_.n_v);
          fill(s, fb, wf_.n_o * wf_.n_v, 0.0);
          break;
        }
        case GEMM:
        case LADDER:
          break;
      }
    }
  }
}

}  // namespace etg
