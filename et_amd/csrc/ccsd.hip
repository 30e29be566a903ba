// CCSD ground-state residual, energy, T1 transformation, Fock matrix and the
// Jacobian transformations on one MI355X.  The sequence of contractions
// follows eT's factorisation (src/wavefunctions/ccsd/omega_ccsd.F90 and the
// doubles/ccs submodules it calls); every product runs on the matrix-core
// kernels in gemm.hip and every reorder on the bandwidth kernels in
// kernels.hip.  The pipeline (T1 dressing -> Fock -> Omega) is written once
// against the Ops interface: run directly it is the residual; recorded on a
// Tape it is replayed forward for rho = A c and backward for sigma = A^T b.
// All tensors stay in HBM between calls.
#include "ccsd.hpp"

#include <algorithm>
#include <cmath>

#include "kernels.hpp"
#include "ladder_order.hpp"

namespace etg {

CCSD::CCSD(int64_t no, int64_t nv, int64_t nJ) : n_o(no), n_v(nv), n_J(nJ) {
  if (no < 1 || nv < 1 || nJ < 1) fail("etg_create: n_o, n_v and n_J must be positive");
  n_mo = n_o + n_v;
  n_t1 = n_o * n_v;
  n_t2 = n_t1 * (n_t1 + 1) / 2;
  ETG_HIP(hipStreamCreate(&stream_));
  const int64_t J = n_J, o = n_o, v = n_v, N = n_t1, n = n_mo;
  Lmo_oo_.resize(sizeof(double) * J * o * o);
  Lmo_ov_.resize(sizeof(double) * J * o * v);
  Lmo_vo_.resize(sizeof(double) * J * v * o);
  Lmo_vv_.resize(sizeof(double) * J * v * v);
  Lt_oo_.resize(sizeof(double) * J * o * o);
  Lt_vo_.resize(sizeof(double) * J * v * o);
  Lt_vv_.resize(sizeof(double) * J * v * v);
  Lt_vvT_.resize(sizeof(double) * J * v * v);
  h_.resize(sizeof(double) * n * n);
  eps_.resize(sizeof(double) * n);
  F_.resize(sizeof(double) * n * n);
  t_.resize(sizeof(double) * (n_t1 + n_t2));
  omega_.resize(sizeof(double) * (n_t1 + n_t2));
  g_ovov_.resize(sizeof(double) * N * N);
  red_partial_.resize(sizeof(double) * kReduceBlocks);
  red_result_.resize(sizeof(double) * 8);
  work_.resize(kBigSlots);
  // small scratch, sized once for every use in the programs below
  small_[0].resize(sizeof(double) * n * n);
  small_[1].resize(sizeof(double) * n * n);
  small_[2].resize(sizeof(double) * std::max(std::max(n * n, J), o * o * o * o));
  small_[3].resize(sizeof(double) * std::max(std::max(J * N, J * o * o), o * o * o * o));
  // default ladder batch: G blocks of at most ~12 GB
  const double budget = 12e9 / 8.0;
  int64_t nb = int64_t(std::sqrt(budget) / double(v));
  set_ladder_batch(std::max<int64_t>(1, std::min<int64_t>(nb, v)));
}

CCSD::~CCSD() {
  drop_transform_graphs();
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
  for (auto& r : ledger_recs_) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto e : event_pool_) (void)hipEventDestroy(e);
}

void CCSD::require(bool ok, const char* what) const {
  if (!ok) fail(what);
}

void CCSD::synchronize() { ETG_HIP(hipStreamSynchronize(stream_)); }

double* CCSD::scratch(int slot, int64_t elems, const char* name) {
  if (slot < kBigSlots) {
    // (ov)^2 elements, or more when n_o > n_v (o^3 v and o^4 blocks share the slots)
    const int64_t cap = std::max(n_t1 * n_t1, n_o * n_o * n_o * std::max(n_o, n_v));
    if (elems > cap) fail(std::string("scratch slot too small for ") + name);
    work_[size_t(slot)].resize(sizeof(double) * cap);
    return work_[size_t(slot)].as<double>();
  }
  DeviceBuffer& b = small_[slot - SL_SMALL0];
  if (size_t(elems) * sizeof(double) > b.bytes()) fail(std::string("small scratch too small for ") + name);
  return b.as<double>();
}

// ---------------------------------------------------------------------------
// products and the per-kernel ledger
// ---------------------------------------------------------------------------

hipEvent_t CCSD::take_event() {
  if (!event_pool_.empty()) {
    hipEvent_t e = event_pool_.back();
    event_pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  ETG_HIP(hipEventCreate(&e));
  return e;
}

void CCSD::set_ledger(bool on) {
  synchronize();
  for (auto& r : ledger_recs_) {
    event_pool_.push_back(r.start);
    event_pool_.push_back(r.stop);
  }
  ledger_recs_.clear();
  ledger_on_ = on;
}

std::vector<CCSD::LedgerEntry> CCSD::ledger() {
  synchronize();
  std::vector<LedgerEntry> out;
  for (const auto& r : ledger_recs_) {
    float ms = 0.f;
    ETG_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    auto it = std::find_if(out.begin(), out.end(),
                           [&](const LedgerEntry& e) { return e.kernel == r.kernel; });
    if (it == out.end()) {
      out.push_back(LedgerEntry{r.kernel, 0, 0.0, 0.0, 0.0});
      it = out.end() - 1;
    }
    it->launches += 1;
    it->seconds += double(ms) * 1e-3;
    it->flops += r.flops;
    it->bytes += r.bytes;
  }
  return out;
}

namespace {
Gemm tagged(Gemm g, uint64_t tag, bool on) {
  g.a_tag = on ? tag : 0;
  return g;
}
}  // namespace

double* CCSD::gemm_partials(int64_t elems) {
  if (gws_cur_->partials.resize(sizeof(double) * size_t(std::max<int64_t>(1, elems)))) ++gws_cur_->gen;
  return gws_cur_->partials.as<double>();
}

void CCSD::run_gemm(const Gemm& g, Prec prec) {
  hipEvent_t a = nullptr, b = nullptr;
  if (ledger_on_) {
    a = take_event();
    b = take_event();
    ETG_HIP(hipEventRecord(a, stream_));
  }
  GemmLaunch info;
  if (!pack_cache_ && g.a_tag != 0) {
    Gemm h = g;
    h.a_tag = 0;
    info = etg::gemm(stream_, h, prec, gws_cur_);
  } else {
    info = etg::gemm(stream_, g, prec, gws_cur_);
  }
  if (ledger_on_) {
    ETG_HIP(hipEventRecord(b, stream_));
    // algorithmic HBM bytes: read A and B once (once in all for an operand shared
    // by the batch), write C once (and read it if beta != 0)
    const double mn = double(g.m) * double(g.n), nbat = double(g.batch);
    const double bytes = 8.0 * (double(g.m) * double(g.k) * (g.strideA == 0 ? 1.0 : nbat) +
                                double(g.k) * double(g.n) * (g.strideB == 0 ? 1.0 : nbat) +
                                (g.beta != 0.0 ? 2.0 : 1.0) * mn * nbat);
    ledger_recs_.push_back(LedgerRec{info.kernel, gemm_flops(g), bytes, a, b});
  }
  last_omega_flops += gemm_flops(g);
}

// ---------------------------------------------------------------------------
// inputs
// ---------------------------------------------------------------------------

namespace {
// L(J, p, q) full -> block (J, p in [p0, p0+np), q in [q0, q0+nq)) with 2D host->device copies
void copy_block(hipStream_t s, const double* L, int64_t nJ, int64_t n_mo, int64_t p0, int64_t np,
                int64_t q0, int64_t nq, double* dst) {
  // each (J, p-range) column strip for fixed q is contiguous in both
  ETG_HIP(hipMemcpy2DAsync(dst, size_t(nJ * np) * sizeof(double),
                           L + nJ * (p0 + n_mo * q0), size_t(nJ * n_mo) * sizeof(double),
                           size_t(nJ * np) * sizeof(double), size_t(nq), hipMemcpyHostToDevice, s));
}
}  // namespace

void CCSD::set_cholesky(const double* L) {
  require(L != nullptr, "etg_set_cholesky: null pointer");
  reset_record();
  const int64_t o = n_o, v = n_v, J = n_J;
  copy_block(stream_, L, J, n_mo, 0, o, 0, o, Lmo_oo_.as<double>());
  copy_block(stream_, L, J, n_mo, 0, o, o, v, Lmo_ov_.as<double>());
  copy_block(stream_, L, J, n_mo, o, v, 0, o, Lmo_vo_.as<double>());
  copy_block(stream_, L, J, n_mo, o, v, o, v, Lmo_vv_.as<double>());
  // g_ovov = L_ov^T L_ov depends only on the undressed ov block
  DirectOps X(*this);
  X.gemm('T', 'N', o * v, o * v, J, 1.0, Lmo_ov_.as<double>(), J, Lmo_ov_.as<double>(), J, 0.0,
         g_ovov_.as<double>(), o * v, Prec::F64);
  have_L_ = true;
  if (!have_t_) {
    fill(stream_, t_.as<double>(), n_t1 + n_t2, 0.0);
    have_t_ = true;
  }
  construct_t1_cholesky();
  synchronize();
}

void CCSD::set_oei(const double* h) {
  require(h != nullptr, "etg_set_oei: null pointer");
  reset_record();
  ETG_HIP(hipMemcpyAsync(h_.as<double>(), h, sizeof(double) * n_mo * n_mo, hipMemcpyHostToDevice,
                         stream_));
  synchronize();
  have_h_ = true;
  have_F_ = false;
}

void CCSD::set_orbital_energies(const double* eps) {
  require(eps != nullptr, "etg_set_orbital_energies: null pointer");
  ETG_HIP(hipMemcpyAsync(eps_.as<double>(), eps, sizeof(double) * n_mo, hipMemcpyHostToDevice,
                         stream_));
  synchronize();
  have_eps_ = true;
}

void CCSD::set_amplitudes(const double* t) {
  require(t != nullptr, "etg_set_amplitudes: null pointer");
  reset_record();
  ETG_HIP(hipMemcpyAsync(t_.as<double>(), t, sizeof(double) * (n_t1 + n_t2),
                         hipMemcpyHostToDevice, stream_));
  synchronize();
  have_t_ = true;
  have_omega_ = have_update_ = false;
}

void CCSD::get_amplitudes(double* t) {
  require(t != nullptr, "etg_get_amplitudes: null pointer");
  ETG_HIP(hipMemcpyAsync(t, t_.as<double>(), sizeof(double) * (n_t1 + n_t2),
                         hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void CCSD::get_omega(double* out) {
  require(out != nullptr, "etg_get_omega: null pointer");
  ETG_HIP(hipMemcpyAsync(out, omega_.as<double>(), sizeof(double) * (n_t1 + n_t2),
                         hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void CCSD::get_fock(double* out) {
  require(have_F_, "etg_get_fock: Fock matrix not constructed");
  ETG_HIP(hipMemcpyAsync(out, F_.as<double>(), sizeof(double) * n_mo * n_mo,
                         hipMemcpyDeviceToHost, stream_));
  synchronize();
}

// ---------------------------------------------------------------------------
// T1 transformation of the Cholesky vectors (t1_ccs.F90:208-499)
// ---------------------------------------------------------------------------

void CCSD::program_t1_cholesky(Ops& X) {
  const int64_t o = n_o, v = n_v, J = n_J;
  const double* t1 = t_.as<double>();
  double* Loo = Lt_oo_.as<double>();
  double* Lvo = Lt_vo_.as<double>();
  double* Lvv = Lt_vv_.as<double>();
  // oo: L_ij = L_ij + sum_a L_ia t_aj          ((J i) x a) (a x j)
  X.copy_sub(Lmo_oo_.as<double>(), J * o * o, Loo, J * o * o, J * o * o, 1, 1.0, 0.0);
  X.gemm('N', 'N', J * o, o, v, 1.0, Lmo_ov_.as<double>(), J * o, t1, v, 1.0, Loo, J * o, Prec::F64);
  // vo: L_ai = L_ai + sum_b L_ab t_bi - sum_j t_aj L'_ji   (L' = dressed oo block)
  X.copy_sub(Lmo_vo_.as<double>(), J * v * o, Lvo, J * v * o, J * v * o, 1, 1.0, 0.0);
  X.gemm('N', 'N', J * v, o, v, 1.0, Lmo_vv_.as<double>(), J * v, t1, v, 1.0, Lvo, J * v, Prec::F64);
  X.gemm_batched('N', 'T', J, v, o, -1.0, Loo, J, J * o, t1, v, 0, 1.0, Lvo, J, J * v, o, Prec::F64);
  // vv: L_ab = L_ab - sum_j t_aj L_jb         batched over b: (J x j)(j x a)
  X.copy_sub(Lmo_vv_.as<double>(), J * v * v, Lvv, J * v * v, J * v * v, 1, 1.0, 0.0);
  X.gemm_batched('N', 'T', J, v, o, -1.0, Lmo_ov_.as<double>(), J, J * o, t1, v, 0, 1.0, Lvv, J,
                 J * v, v, Prec::F64);
  // (J, c, a) copy of the vv block: a-batches of the ladder become contiguous columns
  const int64_t d[4] = {J, v, v, 1};
  X.sort4(Lvv, d, "1324", Lt_vvT_.as<double>(), 1.0, 0.0);
}

void CCSD::construct_t1_cholesky() {
  require(have_L_, "construct_t1_cholesky: Cholesky vectors not set");
  DirectOps X(*this);
  program_t1_cholesky(X);
  have_F_ = false;
}

// ---------------------------------------------------------------------------
// T1 Fock matrix (fock_ccs.F90:242-1109): F = h^T1 + sum_k (2 g_pqkk - g_pkkq)
// ---------------------------------------------------------------------------

void CCSD::program_fock(Ops& X) {
  const int64_t o = n_o, v = n_v, J = n_J, n = n_mo;
  double* F = F_.as<double>();
  const double* Lov = Lmo_ov_.as<double>();
  const double* Loo = Lt_oo_.as<double>();
  // h^T1 = X h Y^T,  X = I - t1, Y = I + t1^T     (t1_ccs.F90:34-95, oei_ccs.F90:35-97)
  double* Xm = X.buf(SL_SMALL0, "fock_X", n * n);
  double* Ym = X.buf(SL_SMALL1, "fock_Y", n * n);
  double* Z = X.buf(SL_SMALL2, "fock_hY", n * n);
  X.t1_xy(t_.as<double>(), Xm, Ym);
  X.gemm('N', 'T', n, n, n, 1.0, h_.as<double>(), n, Ym, n, 0.0, Z, n, Prec::F64);
  X.gemm('N', 'N', n, n, n, 1.0, Xm, n, Z, n, 0.0, F, n, Prec::F64);
  // Coulomb: F_pq += 2 sum_J L_pq X_J,  X_J = sum_k L_kk
  double* XJ = X.buf(SL_SMALL2, "fock_XJ", J);
  X.trace_oo(Loo, XJ);
  double* c = X.buf(SL_SMALL0, "fock_coul_oo", o * o);
  X.gemm('T', 'N', o * o, 1, J, 2.0, Loo, J, XJ, J, 0.0, c, o * o, Prec::F64);
  X.copy_sub(c, o, F, n, o, o, 1.0, 1.0);
  c = X.buf(SL_SMALL0, "fock_coul_ov", o * v);
  X.gemm('T', 'N', o * v, 1, J, 2.0, Lov, J, XJ, J, 0.0, c, o * v, Prec::F64);
  X.copy_sub(c, o, F + n * o, n, o, v, 1.0, 1.0);
  c = X.buf(SL_SMALL0, "fock_coul_vo", v * o);
  X.gemm('T', 'N', v * o, 1, J, 2.0, Lt_vo_.as<double>(), J, XJ, J, 0.0, c, v * o, Prec::F64);
  X.copy_sub(c, v, F + o, n, v, o, 1.0, 1.0);
  c = X.buf(SL_SMALL0, "fock_coul_vv", v * v);
  X.gemm('T', 'N', v * v, 1, J, 2.0, Lt_vv_.as<double>(), J, XJ, J, 0.0, c, v * v, Prec::F64);
  X.copy_sub(c, v, F + o + n * o, n, v, v, 1.0, 1.0);
  // Exchange: F_pq -= sum_{J,k} L_pk L_kq with left factors reordered to (J, k, p)
  double* Lleft = X.buf(SL_SMALL3, "fock_Lleft_oo", J * o * o);
  const int64_t d_oo[4] = {J, o, o, 1};
  X.sort4(Loo, d_oo, "1324", Lleft, 1.0, 0.0);                        // (J, k, i)
  X.gemm('T', 'N', o, o, J * o, -1.0, Lleft, J * o, Loo, J * o, 1.0, F, n, Prec::F64);
  X.gemm('T', 'N', o, v, J * o, -1.0, Lleft, J * o, Lov, J * o, 1.0, F + n * o, n, Prec::F64);
  Lleft = X.buf(SL_SMALL3, "fock_Lleft_vo", J * o * v);
  const int64_t d_vo[4] = {J, v, o, 1};
  X.sort4(Lt_vo_.as<double>(), d_vo, "1324", Lleft, 1.0, 0.0);       // (J, k, a)
  X.gemm('T', 'N', v, o, J * o, -1.0, Lleft, J * o, Loo, J * o, 1.0, F + o, n, Prec::F64);
  X.gemm('T', 'N', v, v, J * o, -1.0, Lleft, J * o, Lov, J * o, 1.0, F + o + n * o, n, Prec::F64);
}

void CCSD::construct_fock() {
  require(have_L_ && have_h_, "construct_fock: Cholesky vectors and one-electron integrals required");
  DirectOps X(*this);
  program_fock(X);
  have_F_ = true;
}

// ---------------------------------------------------------------------------
// energy (mean_value_ccsd.F90:36-99), MP2 guess, orbital differences
// ---------------------------------------------------------------------------

double CCSD::calculate_energy() {
  require(have_L_ && have_t_, "calculate_energy: Cholesky vectors and amplitudes required");
  double* res = red_result_.as<double>();
  ccsd_energy(stream_, t_.as<double>(), t_.as<double>() + n_t1, g_ovov_.as<double>(), n_o, n_v,
              red_partial_.as<double>(), res);
  double e = 0.0;
  ETG_HIP(hipMemcpyAsync(&e, res, sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  return e;
}

double CCSD::omega_norm() {
  double* res = red_result_.as<double>() + 1;
  sum_squares(stream_, omega_.as<double>(), n_t1 + n_t2, red_partial_.as<double>(), res);
  double e = 0.0;
  ETG_HIP(hipMemcpyAsync(&e, res, sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  return std::sqrt(e);
}

void CCSD::set_t2_to_mp2_guess() {
  require(have_L_ && have_eps_, "set_t2_to_mp2_guess: Cholesky vectors and orbital energies required");
  reset_record();
  const int64_t N = n_t1;
  double* g = scratch(SL_A, N * N, "mp2_g_vovo");
  DirectOps X(*this);
  X.gemm('T', 'N', N, N, n_J, 1.0, Lt_vo_.as<double>(), n_J, Lt_vo_.as<double>(), n_J, 0.0, g, N,
         Prec::F64);
  mp2_guess(stream_, g, eps_.as<double>(), t_.as<double>() + n_t1, n_o, n_v);
  synchronize();
  have_t_ = true;
  have_omega_ = have_update_ = false;
}

void CCSD::get_orbital_differences(double* host, int64_t n) {
  require(have_eps_, "get_orbital_differences: orbital energies not set");
  require(n == n_t1 || n == n_t1 + n_t2, "get_orbital_differences: n must be n_t1 or n_t1 + n_t2");
  diff_.resize(sizeof(double) * (n_t1 + n_t2));
  double* d = diff_.as<double>();
  orbital_differences(stream_, eps_.as<double>(), d, n_o, n_v, n);
  ETG_HIP(hipMemcpyAsync(host, d, sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
  synchronize();
}

// ---------------------------------------------------------------------------
// Omega (omega_ccsd.F90:39-110)
// ---------------------------------------------------------------------------

void CCSD::program_omega(Ops& X, Prec prec) {
  const int64_t N = n_t1;
  // t_aibj (full) and u_aibj = 2 t_aibj - t_ajbi      (omega_ccsd.F90:881-914)
  double* T = X.buf(SL_T, "t_aibj", N * N);
  double* U = X.buf(SL_U, "u_aibj", N * N);
  X.squareup(t_.as<double>() + n_t1, T, N);
  const int64_t d[4] = {n_v, n_o, n_v, n_o};
  X.sort4(T, d, "1234", U, 2.0, 0.0);
  X.sort4(T, d, "1432", U, -1.0, 1.0);
  // singles: F_ai + doubles A1, B1, C1
  omega_singles(X, prec, U);
  // doubles: E2, symmetric sum, B2, 1/2 diagonal, pack, then C2/D2 and A2
  omega_e2_b2(X, prec, T, U);
  omega_c2_d2(X, prec, T, U);
  X.ladder(t_.as<double>() + n_t1, Lt_vvT_.as<double>(), omega_.as<double>() + n_t1, prec);
}

void CCSD::construct_omega(Prec prec) {
  require(have_L_ && have_t_, "construct_omega: Cholesky vectors and amplitudes required");
  DirectOps X(*this);
  if (!have_F_) construct_fock();
  last_omega_flops = 0.0;
  program_omega(X, prec);
  have_omega_ = true;
}

void CCSD::omega_singles(Ops& X, Prec prec, const double* U) {
  const int64_t o = n_o, v = n_v, J = n_J, N = n_t1, n = n_mo;
  double* om1 = omega_.as<double>();
  // Omega_ai = F_ai                                    (omega_ccs.F90:62-86)
  X.copy_sub(F_.as<double>() + o, n, om1, v, v, o, 1.0, 0.0);
  // A1: X_Jbi = sum_cj L^mo_Jcj u_cjbi ; Omega_ai += sum_Jb L^T1_Jab X_Jbi   (omega_doubles.F90:36-139)
  double* Xj = X.buf(SL_SMALL3, "a1_X", J * N);
  X.gemm('N', 'N', J, N, N, 1.0, Lmo_vo_.as<double>(), J, U, N, 0.0, Xj, J, prec);
  X.gemm('T', 'N', v, o, J * v, 1.0, Lt_vvT_.as<double>(), J * v, Xj, J * v, 1.0, om1, v, prec);
  // B1: Omega_ai -= sum_jbk u_ajbk g_kbji             (omega_doubles.F90:141-205)
  double* g_kbji = X.buf(SL_A, "b1_g_kbji", o * v * o * o);
  double* g_jbki = X.buf(SL_B, "b1_g_jbki", o * v * o * o);
  X.gemm('T', 'N', o * v, o * o, J, 1.0, Lmo_ov_.as<double>(), J, Lt_oo_.as<double>(), J, 0.0,
         g_kbji, o * v, prec);
  const int64_t d[4] = {o, v, o, o};
  X.sort4(g_kbji, d, "3214", g_jbki, 1.0, 0.0);
  X.gemm('N', 'N', v, o, o * o * v, -1.0, U, v, g_jbki, o * o * v, 1.0, om1, v, prec);
  // C1: Omega_ai += sum_bj u_aibj F_jb                (omega_doubles.F90:207-250)
  double* f = X.buf(SL_C, "c1_f", N);
  X.fock_ov_vec(F_.as<double>(), f);
  X.gemm('N', 'N', N, 1, N, 1.0, U, N, f, N, 1.0, om1, N, prec);
}

void CCSD::omega_e2_b2(Ops& X, Prec prec, const double* T, const double* U) {
  const int64_t o = n_o, v = n_v, J = n_J, N = n_t1, n = n_mo;
  const double* F = F_.as<double>();
  const double* g_ovov = g_ovov_.as<double>();
  const int64_t d[4] = {v, o, v, o};
  // E2 (omega_ccsd.F90:712-878)
  double* u_e2 = X.buf(SL_A, "e2_u", N * N);   // u_kl^bd = 2 t_bkdl - t_bldk = u(b,k,d,l)
  X.sort4(U, d, "1432", u_e2, 1.0, 0.0);
  double* om = X.buf(SL_B, "e2_omega", N * N);
  X.fill0(om, N * N);
  double* Xbc = X.buf(SL_SMALL0, "e2_X", v * v);
  double* Ykj = X.buf(SL_SMALL1, "e2_Y", o * o);
  X.copy_sub(F + o + n * o, n, Xbc, v, v, v, 1.0, 0.0);
  X.gemm('N', 'N', v, v, o * v * o, -1.0, u_e2, v, g_ovov, o * v * o, 1.0, Xbc, v, prec);
  X.gemm('N', 'N', v, o * v * o, v, 1.0, Xbc, v, T, v, 1.0, om, v, prec);
  X.copy_sub(F, n, Ykj, o, o, o, 1.0, 0.0);
  X.gemm('N', 'N', o, o, v * o * v, 1.0, g_ovov, o, u_e2, v * o * v, 1.0, Ykj, o, prec);
  X.gemm('N', 'N', v * o * v, o, o, -1.0, T, v * o * v, Ykj, o, 1.0, om, v * o * v, prec);
  // symmetric sum
  double* om2 = X.buf(SL_C, "e2_omega_sym", N * N);
  X.symsum(om, om2, N);
  // B2 in (a,b,i,j) order (omega_ccsd.F90:483-583)
  double* om_abij = X.buf(SL_A, "b2_omega_abij", N * N);
  X.sort4(om2, d, "1324", om_abij, 1.0, 0.0);
  double* t_abij = X.buf(SL_B, "b2_t_abij", N * N);
  X.sort4(T, d, "1324", t_abij, 1.0, 0.0);
  double* g_vovo = X.buf(SL_C, "b2_g_vovo", N * N);
  X.gemm('T', 'N', N, N, J, 1.0, Lt_vo_.as<double>(), J, Lt_vo_.as<double>(), J, 0.0, g_vovo, N, prec);
  X.sort4(g_vovo, d, "1324", om_abij, 1.0, 1.0);
  double* g_kilj = X.buf(SL_SMALL3, "b2_g_kilj", o * o * o * o);
  double* X_klij = X.buf(SL_SMALL2, "b2_X", o * o * o * o);
  X.gemm('T', 'N', o * o, o * o, J, 1.0, Lt_oo_.as<double>(), J, Lt_oo_.as<double>(), J, 0.0, g_kilj,
         o * o, prec);
  const int64_t d_oooo[4] = {o, o, o, o};
  X.sort4(g_kilj, d_oooo, "1324", X_klij, 1.0, 0.0);
  double* g_klcd = X.buf(SL_C, "b2_g_klcd", N * N);
  const int64_t d_ovov[4] = {o, v, o, v};
  X.sort4(g_ovov, d_ovov, "1324", g_klcd, 1.0, 0.0);
  X.gemm('N', 'N', o * o, o * o, v * v, 1.0, g_klcd, o * o, t_abij, v * v, 1.0, X_klij, o * o, prec);
  X.gemm('N', 'N', v * v, o * o, o * o, 1.0, t_abij, v * v, X_klij, o * o, 1.0, om_abij, v * v, prec);
  // back to (a,i,b,j), 1/2 on the diagonal and pack into omega2   (omega_ccsd.F90:97-102)
  double* om_aibj = X.buf(SL_C, "b2_omega_aibj", N * N);
  const int64_t d_vvoo[4] = {v, v, o, o};
  X.sort4(om_abij, d_vvoo, "1324", om_aibj, 1.0, 0.0);
  X.packin(om_aibj, omega_.as<double>() + n_t1, N, 1.0, 0.0, 0.5);
}

void CCSD::omega_c2_d2(Ops& X, Prec prec, const double* Tfull, const double* U) {
  // omega_ccsd.F90:586-709: four (ov) x (ov) products
  const int64_t o = n_o, v = n_v, J = n_J, N = n_t1;
  const int64_t d[4] = {v, o, v, o};
  double* Tm = X.buf(SL_A, "c2_T", N * N);                // T[(p,q),(r,s)] = t_psrq
  X.sort4(Tfull, d, "1432", Tm, 1.0, 0.0);
  double* g_vvoo = X.buf(SL_B, "c2_g_vvoo", N * N);
  double* C = X.buf(SL_C, "c2_C", N * N);
  double* D = X.buf(SL_D, "c2_D", N * N);
  X.gemm('T', 'N', v * v, o * o, J, 1.0, Lt_vv_.as<double>(), J, Lt_oo_.as<double>(), J, 0.0, g_vvoo,
         v * v, prec);                                       // g_abij
  const int64_t d_vvoo[4] = {v, v, o, o};
  X.sort4(g_vvoo, d_vvoo, "2314", C, 1.0, 0.0);             // C(b,i,a,j) = g_abij
  X.sort4(C, d, "1234", D, -1.0, 0.0);                      // D = -C
  double* g_voov = X.buf(SL_B, "c2_g_voov", N * N);
  X.gemm('T', 'N', N, N, J, 1.0, Lt_vo_.as<double>(), J, Lmo_ov_.as<double>(), J, 0.0, g_voov, N,
         prec);                                              // g_aijb
  const int64_t d_voov[4] = {v, o, o, v};
  X.sort4(g_voov, d_voov, "4312", D, 2.0, 1.0);             // D(c,k,a,i) += 2 g_aikc
  double* G = X.buf(SL_B, "c2_G", N * N);
  const int64_t d_ovov[4] = {o, v, o, v};
  X.sort4(g_ovov_.as<double>(), d_ovov, "2341", G, 1.0, 0.0);   // G(c,l,d,k) = g_kcld
  X.gemm('N', 'N', N, N, N, -0.5, G, N, Tm, N, 1.0, C, N, prec);
  double* Gc = X.buf(SL_E, "c2_Gc", N * N);                 // 2 G(p,s,r,q) - G(p,q,r,s)
  X.sort4(G, d, "1432", Gc, 2.0, 0.0);
  X.sort4(G, d, "1234", Gc, -1.0, 1.0);
  X.gemm('N', 'N', N, N, N, 0.5, Gc, N, U, N, 1.0, D, N, prec);
  double* W = X.buf(SL_B, "c2_W", N * N);
  X.gemm('N', 'N', N, N, N, -1.5, Tm, N, C, N, 0.0, W, N, prec);
  X.sympack(W, omega_.as<double>() + n_t1, v, o, 1.0, 1.0 / 3.0, 2.0 / 3.0);
  double* V = X.buf(SL_B, "c2_V", N * N);
  X.gemm('N', 'N', N, N, N, 0.5, U, N, D, N, 0.0, V, N, prec);
  X.sympack(V, omega_.as<double>() + n_t1, v, o, 1.0, 1.0, 0.0);
}

// ---------------------------------------------------------------------------
// A2 ladder (omega_ccsd.F90:113-480) with +/- symmetrisation
//   W_abij = sum_cd g_acbd t_cidj = w+_{ij,ab} + w-_{ij,ab}
//   w+ = 1/2 sum_{c<=d} t+_{ij,cd} g+_{cd,ab},  w- = 1/2 sum_{c<d} t-_{ij,cd} g-_{cd,ab}
// g is generated per (a-batch, b-batch) block as G = L_vvT(:, A)^T L_vvT(:, B)
// and never stored whole; w+- for all pairs are gathered into packed Omega at the end.
// The tangent and adjoint recompute G block by block.
// ---------------------------------------------------------------------------

void CCSD::set_ladder_batch(int64_t nb) {
  if (nb < 1) fail("ladder batch must be positive");
  synchronize();
  drop_transform_graphs();              // captured passes hold the old plan's buffers
  ladder_nb_ = std::min(nb, n_v);
  build_ladder_plan();
}

void CCSD::build_ladder_plan() {
  const int64_t v = n_v, nb = ladder_nb_;
  blocks_.clear();
  std::vector<int32_t> pa, pb, col(size_t(v * v), -1), ca, cb;
  int64_t col0 = 0, maxpairs = 0;
  const int64_t nbatch = ceil_div(v, nb);
  for (int64_t A = 0; A < nbatch; ++A) {
    const int64_t a0 = A * nb, la = std::min(nb, v - a0);
    for (int64_t B = 0; B <= A; ++B) {
      const int64_t b0 = B * nb, lb = std::min(nb, v - b0);
      Block blk{a0, la, b0, lb, col0, 0};
      for (int64_t bb = 0; bb < lb; ++bb)
        for (int64_t aa = 0; aa < la; ++aa) {
          const int64_t a = a0 + aa, b = b0 + bb;
          if (a < b) continue;
          pa.push_back(int32_t(aa));
          pb.push_back(int32_t(bb));
          ca.push_back(int32_t(a));
          cb.push_back(int32_t(b));
          col[size_t(a + v * b)] = int32_t(col0 + blk.npairs);
          col[size_t(b + v * a)] = int32_t(col0 + blk.npairs);
          ++blk.npairs;
        }
      col0 += blk.npairs;
      maxpairs = std::max(maxpairs, blk.npairs);
      blocks_.push_back(blk);
    }
  }
  if (col0 != v * (v + 1) / 2) fail("ladder plan: pair count mismatch");
  auto upload = [](DeviceBuffer& d, const std::vector<int32_t>& h) {
    d.resize(sizeof(int32_t) * std::max<size_t>(1, h.size()));
    ETG_HIP(hipMemcpy(d.as<int32_t>(), h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice));
  };
  upload(pair_a_, pa);
  upload(pair_b_, pb);
  upload(pair_col_, col);
  upload(pcol_a_, ca);
  upload(pcol_b_, cb);
  const int64_t op = n_o * (n_o + 1) / 2, om = n_o * (n_o - 1) / 2;
  const int64_t vp = v * (v + 1) / 2, vm = v * (v - 1) / 2;
  // t+- hold either order of the contraction index (triangle or tile-blocked)
  lsz_tp_ = op * std::max(vp, ladder_blocked_len(v, true));
  lsz_tm_ = std::max<int64_t>(1, om * std::max(vm, ladder_blocked_len(v, false)));
  lsz_wp_ = op * vp;
  lsz_wm_ = std::max<int64_t>(1, om * vp);
  tp_.resize(sizeof(double) * lsz_tp_);
  tm_.resize(sizeof(double) * lsz_tm_);
  wp_.resize(sizeof(double) * lsz_wp_);
  wm_.resize(sizeof(double) * lsz_wm_);
  gblk_.resize(sizeof(double) * v * v * std::min(nb, v) * std::min(nb, v));
  gp_.resize(sizeof(double) * vp * maxpairs);
  gm_.resize(sizeof(double) * std::max<int64_t>(1, vm * maxpairs));
  ladder_maxpairs_ = maxpairs;
  gpk_.release();                       // sized by the first split-bf16 ladder
  gmk_.release();
  reserved_nb_ = 0;                     // per-vector sizes changed: size the batches again
  nb_cap_ = 0;
}

// The split-bf16 packs of g+- for one block (grown to the plan's largest block
// on first use; true when they moved).
bool CCSD::reserve_ladder_packs() {
  const int64_t v = n_v;
  bool moved = gpk_.resize(2 * size_t(split_bf16_pack_elems(ladder_maxpairs_, ladder_blocked_len(v, true))));
  moved |= gmk_.resize(2 * size_t(split_bf16_pack_elems(ladder_maxpairs_, ladder_blocked_len(v, false))));
  return moved;
}

// G of ladder block b: G(c + v aa, d + v bb) = sum_J L_t1(J, c, a0 + aa)
// L_t1(J, d, b0 + bb), in FP32 when `fp32` (the split-bf16 path).  The pairs
// of a diagonal block (a0 == b0) have aa >= bb, so it is formed in column strips
// of 8 b's, each from the row block of its first b on (a little over half of the
// full block's work).  `tag` names the full block's A operand for the pack cache
// (0: no tag).
void CCSD::ladder_G(const Block& b, const double* LT, Prec prec, bool fp32, uint64_t tag) {
  const int64_t v = n_v, J = n_J, ldg = v * b.la;
  auto product = [&](int64_t r0, int64_t c0, int64_t nc, uint64_t t) {
    Gemm gg = make_gemm('T', 'N', v * (b.la - r0), v * nc, J, 1.0, LT + J * v * (b.a0 + r0), J,
                        LT + J * v * (b.b0 + c0), J, 0.0, gblk_.as<double>() + v * r0 + ldg * v * c0, ldg);
    if (fp32) {
      gg.C = nullptr;
      gg.Cf = gblk_.as<float>() + v * r0 + ldg * v * c0;
    }
    run_gemm(tagged(gg, t, t != 0 && pack_cache_), prec);
  };
  if (b.a0 != b.b0) {
    product(0, 0, b.lb, tag);
    return;
  }
  constexpr int64_t kStrip = 8;
  for (int64_t c0 = 0; c0 < b.lb; c0 += kStrip) product(c0, c0, std::min(kStrip, b.lb - c0), 0);
}

// Per block: G, then g+- and the products w+- = 1/2 t+- g+-.  The split-bf16
// form writes g+- straight as packs in the tile-blocked (c, d) order, with t+-
// in the same order; the FP64 and FP32 forms use FP64 g+- in the triangle order.
void CCSD::ladder_primal(const double* t2, const double* LT, double* om2, Prec prec) {
  const int64_t o = n_o, v = n_v;
  const int64_t op = o * (o + 1) / 2, om = o * (o - 1) / 2;
  const bool split = prec == Prec::BF16X3;
  const int64_t kp = split ? ladder_blocked_len(v, true) : v * (v + 1) / 2;
  const int64_t km = split ? ladder_blocked_len(v, false) : v * (v - 1) / 2;
  const bool minus = om > 0 && v > 1;
  hipStream_t s = stream_;
  if (split && reserve_ladder_packs()) drop_transform_graphs();
  const uint64_t gen = ++ladder_gen_ << 3;     // t+/- are rebuilt below: new pack tags
  if (split)
    ladder_t_pm_blocked(s, t2, tp_.as<double>(), tm_.as<double>(), o, v);
  else
    ladder_t_pm(s, t2, tp_.as<double>(), tm_.as<double>(), o, v);
  for (const Block& b : blocks_) {
    ladder_G(b, LT, prec, split, gen | 7);
    Gemm gpl = make_gemm('N', 'N', op, b.npairs, kp, 0.5, tp_.as<double>(), op, gp_.as<double>(), kp, 0.0,
                         wp_.as<double>() + op * b.col0, op);
    Gemm gmi = make_gemm('N', 'N', om, b.npairs, km, 0.5, tm_.as<double>(), om, gm_.as<double>(), km, 0.0,
                         wm_.as<double>() + om * b.col0, om);
    if (split) {
      ladder_g_pm_split(s, gblk_.as<float>(), v, b.la, pair_a_.as<int32_t>() + b.col0,
                        pair_b_.as<int32_t>() + b.col0, b.npairs, gpk_.as<__bf16>(), gmk_.as<__bf16>());
      gpl.B = gmi.B = nullptr;
      gpl.b_packed = gpk_.as<__bf16>();
      gmi.b_packed = gmk_.as<__bf16>();
    } else {
      ladder_g_pm(s, gblk_.as<double>(), v, b.la, b.lb, pair_a_.as<int32_t>() + b.col0,
                  pair_b_.as<int32_t>() + b.col0, b.npairs, gp_.as<double>(), gm_.as<double>());
    }
    run_gemm(tagged(gpl, gen | 1, pack_cache_), prec);
    if (minus) run_gemm(tagged(gmi, gen | 2, pack_cache_), prec);
  }
  ladder_gather(s, wp_.as<double>(), wm_.as<double>(), pair_col_.as<int32_t>(), om2, o, v);
}

// Per-vector work of the ladder derivatives: t+/- and w+/- of a derivative
// vector and the two (ov)^2 arrays of its T1 part (elements).
int64_t CCSD::ladder_vector_elems() const {
  return lsz_tp_ + lsz_tm_ + lsz_wp_ + lsz_wm_ + 2 * n_t1 * n_t1;
}

bool CCSD::reserve_ladder_vectors(int64_t nb) {
  bool moved = false;
  auto grow = [&](DeviceBuffer& b, int64_t e) { moved |= b.resize(sizeof(double) * size_t(nb * e)); };
  grow(dtp_, lsz_tp_);
  grow(dtm_, lsz_tm_);
  grow(dwp_, lsz_wp_);
  grow(dwm_, lsz_wm_);
  grow(lz1_, n_t1 * n_t1);
  grow(lz2_, n_t1 * n_t1);
  return moved;
}

// T1 part of the ladder derivative.  The dressed block depends on t1 through
// L_t1(J, a, c) = L(J, a, c) - sum_k t_ak L(J, k, c) (ov block undressed), so
//   sum_cd dg_acbd t_cidj = Z_aibj + Z_bjai,   Z_aibj = -sum_k dt_ak W_kbij,
//   W_kbij = sum_cd G_kcbd t_cidj,   G_kcbd = sum_J L(J, k, c) L_t1(J, b, d),
// with W (o^3 v) built once per record (save_jacobian_d2_intermediates,
// jacobian_ccsd.F90:1984, keeps the same X_kbij).  G is formed in slices of d.
void CCSD::build_ladder_t1_intermediate() {
  const int64_t o = n_o, v = n_v, J = n_J, N = n_t1;
  hipStream_t s = stream_;
  DirectOps X(*this);
  lw_.resize(sizeof(double) * size_t(std::max<int64_t>(1, o * v * o * o)));
  double* W = lw_.as<double>();
  double* Tfull = scratch(SL_A, N * N, "ladder_t1_t_full");
  double* Tcd = scratch(SL_B, N * N, "ladder_t1_t_cdij");
  squareup(s, t_.as<double>() + n_t1, Tfull, N);
  const int64_t dT[4] = {v, o, v, o};
  sort4(s, Tfull, dT, "1324", Tcd, 1.0, 0.0);                  // (c, d, i, j)
  // slices of d sized to the (ov)^2 scratch slots (o v^2 <= (ov)^2, so one d always fits)
  const int64_t cap = std::max(N * N, o * o * o * std::max(o, v));
  const int64_t slice = std::max<int64_t>(1, std::min<int64_t>(v, cap / (o * v * v)));
  double* G = scratch(SL_C, o * v * v * slice, "ladder_t1_G");
  double* Gs = scratch(SL_D, o * v * v * slice, "ladder_t1_G_kbcd");
  for (int64_t d0 = 0; d0 < v; d0 += slice) {
    const int64_t nd = std::min(slice, v - d0);
    // G(k, c, b, d) = sum_J L(J, k, c) L_t1(J, b, d),  d in [d0, d0 + nd)
    X.gemm('T', 'N', o * v, v * nd, J, 1.0, Lmo_ov_.as<double>(), J, Lt_vv_.as<double>() + J * v * d0, J,
           0.0, G, o * v, Prec::F64);
    const int64_t dG[4] = {o, v, v, nd};
    sort4(s, G, dG, "1324", Gs, 1.0, 0.0);                      // (k, b, c, d)
    // W(kb, ij) += G(kb, cd) T(cd, ij) over this slice of d
    X.gemm('N', 'N', o * v, o * o, v * nd, 1.0, Gs, o * v, Tcd + v * d0, v * v, d0 == 0 ? 0.0 : 1.0, W,
           o * v, Prec::F64);
  }
  have_lw_ = true;
}

// Tangent of the ladder term for nb vectors (derivative k of t at dt2 - n_t1 +
// k s_dt2, into dom2 + k s_dom2):
//   d(om2) = ladder(dt2, g) + [T1 part through W].
// g+/- of each (a, b) block is built once and applied to every vector (the
// vector dimension is the batch of the products).  The derivative of the
// dressed vv block is not read: the T1 part above is the whole of it.
void CCSD::ladder_tangent(const double* t2, const double* dt2, int64_t s_dt2, const double* LT,
                          const double* dLT, int64_t s_dL, double* dom2, int64_t s_dom2, int64_t nb,
                          Prec prec) {
  (void)dLT;
  (void)s_dL;
  if (dt2 == nullptr || dom2 == nullptr) fail("ladder tangent: missing derivative");
  if (t2 != t_.as<double>() + n_t1 || LT != Lt_vvT_.as<double>() || !have_lw_)
    fail("ladder tangent: the record's ladder operands are required");
  const int64_t o = n_o, v = n_v, N = n_t1;
  const int64_t op = o * (o + 1) / 2, om = o * (o - 1) / 2;
  const bool split = prec == Prec::BF16X3;
  const int64_t kp = split ? ladder_blocked_len(v, true) : v * (v + 1) / 2;
  const int64_t km = split ? ladder_blocked_len(v, false) : v * (v - 1) / 2;
  const bool minus = om > 0 && v > 1;
  hipStream_t s = stream_;
  reserve_ladder_vectors(nb);
  if (split) reserve_ladder_packs();    // sized before any capture by transform_batch
  DirectOps X(*this);
  double* dtp = dtp_.as<double>();
  double* dtm = dtm_.as<double>();
  double* dwp = dwp_.as<double>();
  double* dwm = dwm_.as<double>();
  auto batched = [&](Gemm g, int64_t sA, int64_t sB, int64_t sC) {
    g.batch = nb;
    g.strideA = sA;
    g.strideB = sB;
    g.strideC = sC;
    run_gemm(g, prec);
  };
  for (int64_t k = 0; k < nb; ++k) {
    if (split)
      ladder_t_pm_blocked(s, dt2 + k * s_dt2, dtp + k * lsz_tp_, dtm + k * lsz_tm_, o, v);
    else
      ladder_t_pm(s, dt2 + k * s_dt2, dtp + k * lsz_tp_, dtm + k * lsz_tm_, o, v);
  }
  for (const Block& b : blocks_) {
    ladder_G(b, LT, prec, split, 0);
    Gemm gpl = make_gemm('N', 'N', op, b.npairs, kp, 0.5, dtp, op, gp_.as<double>(), kp, 0.0,
                         dwp + op * b.col0, op);
    Gemm gmi = make_gemm('N', 'N', om, b.npairs, km, 0.5, dtm, om, gm_.as<double>(), km, 0.0,
                         dwm + om * b.col0, om);
    if (split) {
      ladder_g_pm_split(s, gblk_.as<float>(), v, b.la, pair_a_.as<int32_t>() + b.col0,
                        pair_b_.as<int32_t>() + b.col0, b.npairs, gpk_.as<__bf16>(), gmk_.as<__bf16>());
      gpl.B = gmi.B = nullptr;
      gpl.b_packed = gpk_.as<__bf16>();
      gmi.b_packed = gmk_.as<__bf16>();
    } else {
      ladder_g_pm(s, gblk_.as<double>(), v, b.la, b.lb, pair_a_.as<int32_t>() + b.col0,
                  pair_b_.as<int32_t>() + b.col0, b.npairs, gp_.as<double>(), gm_.as<double>());
    }
    batched(gpl, lsz_tp_, 0, lsz_wp_);
    if (minus) batched(gmi, lsz_tm_, 0, lsz_wm_);
  }
  for (int64_t k = 0; k < nb; ++k)
    ladder_gather(s, dwp + k * lsz_wp_, dwm + k * lsz_wm_, pair_col_.as<int32_t>(), dom2 + k * s_dom2,
                  o, v);
  // T1 part: Z(a, b, i, j) = -dt1(a, k) W(k, bij), then (a, i, b, j), then symmetrised into packed
  double* Z = lz1_.as<double>();
  double* Zf = lz2_.as<double>();
  batched(make_gemm('N', 'N', v, v * o * o, o, -1.0, dt2 - n_t1, v, lw_.as<double>(), o, 0.0, Z, v),
          s_dt2, 0, N * N);
  const int64_t dZ[4] = {v, v, o, o};
  for (int64_t k = 0; k < nb; ++k) {
    sort4(s, Z + k * N * N, dZ, "1324", Zf + k * N * N, 1.0, 0.0);
    symmetrize_and_pack(s, Zf + k * N * N, dom2 + k * s_dom2, v, o, 1.0, 1.0, 0.0);
  }
}

// Adjoint of the ladder term for nb vectors: from om2b + k s_om2b, accumulates
// into t2b + k s_t2b (doubles) and the t1 adjoint ahead of it (T1 part through W);
// the dressed vv block receives nothing.  g+/- recomputed once per block.
void CCSD::ladder_adjoint(const double* t2, double* t2b, int64_t s_t2b, const double* LT, double* LTb,
                          int64_t s_Lb, const double* om2b, int64_t s_om2b, int64_t nb, Prec prec) {
  (void)LTb;
  (void)s_Lb;
  if (t2b == nullptr || om2b == nullptr) fail("ladder adjoint: missing derivative");
  if (t2 != t_.as<double>() + n_t1 || LT != Lt_vvT_.as<double>() || !have_lw_)
    fail("ladder adjoint: the record's ladder operands are required");
  const int64_t o = n_o, v = n_v, N = n_t1;
  const int64_t op = o * (o + 1) / 2, om = o * (o - 1) / 2;
  const int64_t vp = v * (v + 1) / 2, vm = v * (v - 1) / 2;
  hipStream_t s = stream_;
  reserve_ladder_vectors(nb);
  DirectOps X(*this);
  double* tpb = dtp_.as<double>();
  double* tmb = dtm_.as<double>();
  double* wpb = dwp_.as<double>();
  double* wmb = dwm_.as<double>();
  auto batched = [&](Gemm g, int64_t sA, int64_t sB, int64_t sC) {
    g.batch = nb;
    g.strideA = sA;
    g.strideB = sB;
    g.strideC = sC;
    run_gemm(g, prec);
  };
  ladder_t_pm(s, t2, tp_.as<double>(), tm_.as<double>(), o, v);
  for (int64_t k = 0; k < nb; ++k)
    ladder_gather_adjoint(s, om2b + k * s_om2b, pcol_a_.as<int32_t>(), pcol_b_.as<int32_t>(),
                          wpb + k * lsz_wp_, wmb + k * lsz_wm_, o, v);
  fill(s, tpb, nb * lsz_tp_, 0.0);
  if (om > 0 && vm > 0) fill(s, tmb, nb * lsz_tm_, 0.0);
  for (const Block& b : blocks_) {
    ladder_G(b, LT, prec, false, 0);
    ladder_g_pm(s, gblk_.as<double>(), v, b.la, b.lb, pair_a_.as<int32_t>() + b.col0,
                pair_b_.as<int32_t>() + b.col0, b.npairs, gp_.as<double>(), gm_.as<double>());
    // w+ = 1/2 t+ g+  ->  t+_bar += 1/2 w+_bar g+^T
    batched(make_gemm('N', 'T', op, vp, b.npairs, 0.5, wpb + op * b.col0, op, gp_.as<double>(), vp,
                      1.0, tpb, op), lsz_wp_, 0, lsz_tp_);
    if (om > 0 && vm > 0)
      batched(make_gemm('N', 'T', om, vm, b.npairs, 0.5, wmb + om * b.col0, om, gm_.as<double>(), vm,
                        1.0, tmb, om), lsz_wm_, 0, lsz_tm_);
  }
  for (int64_t k = 0; k < nb; ++k)
    ladder_t_pm_adjoint(s, tpb + k * lsz_tp_, tmb + k * lsz_tm_, t2b + k * s_t2b, o, v);
  // T1 part: Zf_bar from the packed adjoint, to (a, b, i, j), then t1_bar -= Z_bar W^T
  double* Zb = lz1_.as<double>();
  double* Zfb = lz2_.as<double>();
  fill(s, Zfb, nb * N * N, 0.0);
  const int64_t dZf[4] = {v, o, v, o};
  for (int64_t k = 0; k < nb; ++k) {
    symmetrize_and_pack_adjoint(s, om2b + k * s_om2b, Zfb + k * N * N, v, o, 1.0, 1.0, 0.0);
    sort4(s, Zfb + k * N * N, dZf, "1324", Zb + k * N * N, 1.0, 0.0);
  }
  batched(make_gemm('N', 'T', v, o, v * o * o, -1.0, Zb, v, lw_.as<double>(), o, 1.0, t2b - n_t1, v),
          N * N, 0, s_t2b);
}

// ---------------------------------------------------------------------------
// Jacobian transformations (jacobian_ccsd.F90:74-178, jacobian_transpose_ccsd.F90:129-236)
//   A_mu,nu = <mu~| exp(-T) [H, tau_nu] exp(T) |R>,  tau_aibj = E_ai E_bj,
// so A = (dOmega/dt) D2 where D2 doubles the doubles diagonal of the trial vector
// (eT scales c_aiai by two, jacobian_ccsd.F90:123).  rho = A c is the forward
// (tangent) replay of the recorded residual program along D2 c; sigma = A^T b
// is its backward (adjoint) replay from b, followed by D2.
// ---------------------------------------------------------------------------

namespace {
// vector blockIdx.y of a batch at stride n_t1 + n_t2
__global__ void scale_doubles_diagonal_kernel(double* x, int64_t n_t1, int64_t n, double factor) {
  double* xv = x + int64_t(blockIdx.y) * n;
  for (int64_t d = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; d < n_t1;
       d += int64_t(gridDim.x) * blockDim.x)
    xv[n_t1 + d * (d + 1) / 2 + d] *= factor;
}
}  // namespace

void CCSD::prepare_for_jacobian() {
  require(have_L_ && have_t_ && have_h_, "prepare_for_jacobian: Cholesky vectors, h and amplitudes required");
  reset_record();
  synchronize();
  auto tape = std::make_unique<Tape>(*this);
  const int64_t J = n_J, o = n_o, v = n_v, n = n_mo;
  tape->register_buffer(Lmo_oo_.as<double>(), J * o * o, true, "L_mo oo");
  tape->register_buffer(Lmo_ov_.as<double>(), J * o * v, true, "L_mo ov");
  tape->register_buffer(Lmo_vo_.as<double>(), J * v * o, true, "L_mo vo");
  tape->register_buffer(Lmo_vv_.as<double>(), J * v * v, true, "L_mo vv");
  tape->register_buffer(h_.as<double>(), n * n, false, "h");   // one-electron perturbations
  tape->register_buffer(g_ovov_.as<double>(), n_t1 * n_t1, true, "g_ovov");
  tape->register_buffer(t_.as<double>(), n_t1 + n_t2, false, "t");
  tape->register_buffer(F_.as<double>(), n * n, false, "F", true);
  tape->register_buffer(omega_.as<double>(), n_t1 + n_t2, false, "omega");
  tape->register_buffer(Lt_oo_.as<double>(), J * o * o, false, "L_t1 oo", true);
  tape->register_buffer(Lt_vo_.as<double>(), J * v * o, false, "L_t1 vo", true);
  tape->register_buffer(Lt_vv_.as<double>(), J * v * v, false, "L_t1 vv", true);
  tape->register_buffer(Lt_vvT_.as<double>(), J * v * v, false, "L_t1 vvT", true);
  program_t1_cholesky(*tape);
  program_fock(*tape);
  program_omega(*tape, Prec::F64);
  build_ladder_t1_intermediate();       // t and L_t1 of this record
  synchronize();
  // keep only the primal values the passes read and cannot rebuild cheaply
  tape->finalize();
  // the record holds its own intermediates: the direct program's (ov)^2 slots
  // are given back (and taken again by the next direct residual)
  for (DeviceBuffer& w : work_) w.release();
  drop_transform_graphs();
  tape_ = std::move(tape);
  have_F_ = true;
  have_omega_ = true;
}

void CCSD::jacobian_transform(const double* c, double* rho) {
  require(c != nullptr && rho != nullptr, "jacobian_transformation: null pointer");
  transform_batch(0, 1, c, rho, Prec::F64);
}

void CCSD::jacobian_transpose_transform(const double* b, double* sigma) {
  require(b != nullptr && sigma != nullptr, "jacobian_transpose_transformation: null pointer");
  transform_batch(1, 1, b, sigma, Prec::F64);
}

// Vectors per transformation pass for the current record: as many as fit in
// 35 % of the free HBM (the record's derivative buffers and the ladder's
// per-vector work, per vector), at most kMaxTransformBatch or the caller's limit.
int64_t CCSD::transform_batch_capacity() {
  size_t free_b = 0, total_b = 0;
  ETG_HIP(hipMemGetInfo(&free_b, &total_b));
  const double per_vec = 8.0 * double(tape_->derivative_elems() + ladder_vector_elems() + 2 * (n_t1 + n_t2));
  int64_t nb = int64_t(0.35 * double(free_b) / per_vec);
  nb = std::max<int64_t>(1, std::min<int64_t>(nb, kMaxTransformBatch));
  if (transform_batch_limit_ > 0) nb = std::min(nb, transform_batch_limit_);
  return nb;
}

void CCSD::jacobian_memory(int64_t* record, int64_t* per_vector, int64_t* per_vector_unshared) {
  require(tape_ != nullptr, "jacobian_memory: prepare_for_jacobian has not been called");
  const int64_t extra = ladder_vector_elems() + 2 * (n_t1 + n_t2);
  if (record) *record = 8 * tape_->record_elems();
  if (per_vector) *per_vector = 8 * (tape_->derivative_elems() + extra);
  if (per_vector_unshared) *per_vector_unshared = 8 * (tape_->unshared_derivative_elems() + extra);
}

void CCSD::set_transform_batch(int64_t nb) {
  if (nb < 0) fail("set_transform_batch: negative vector count");
  transform_batch_limit_ = nb;
  nb_cap_ = 0;                          // sized again at the next transformation
}

// One transformation pass over the nb vectors staged in jin_ (stride n) into
// jout_ (device buffers of fixed address, so the sequence can be captured once
// and replayed).
void CCSD::transform_body(int side, Prec prec, int64_t nb) {
  // the passes use their own product workspace: its blocks must not move
  // between a capture and the replays, whatever other products need
  struct Use {
    CCSD& w;
    explicit Use(CCSD& x) : w(x) { w.gws_cur_ = &w.gws_pass_; }
    ~Use() { w.gws_cur_ = &w.gws_; }
  } use(*this);
  const int64_t n = n_t1 + n_t2;
  const dim3 g(unsigned(ceil_div(n_t1, 256)), unsigned(nb));
  tape_->zero_derivatives(nb);
  // derivative vectors of t and omega are contiguous: their buffers hold n elements
  if (side == 0) {
    double* dt = tape_->deriv(t_.as<double>());
    ETG_HIP(hipMemcpyAsync(dt, jin_.as<double>(), sizeof(double) * n * nb, hipMemcpyDeviceToDevice,
                           stream_));
    scale_doubles_diagonal_kernel<<<g, 256, 0, stream_>>>(dt, n_t1, n, 2.0);
    ETG_LAUNCH("scale_doubles_diagonal");
    tape_->tangent(prec, nb);
    ETG_HIP(hipMemcpyAsync(jout_.as<double>(), tape_->deriv(omega_.as<double>()), sizeof(double) * n * nb,
                           hipMemcpyDeviceToDevice, stream_));
  } else {
    ETG_HIP(hipMemcpyAsync(tape_->deriv(omega_.as<double>()), jin_.as<double>(), sizeof(double) * n * nb,
                           hipMemcpyDeviceToDevice, stream_));
    tape_->adjoint(prec, nb);
    double* tb = tape_->deriv(t_.as<double>());
    scale_doubles_diagonal_kernel<<<g, 256, 0, stream_>>>(tb, n_t1, n, 2.0);
    ETG_LAUNCH("scale_doubles_diagonal");
    ETG_HIP(hipMemcpyAsync(jout_.as<double>(), tb, sizeof(double) * n * nb, hipMemcpyDeviceToDevice, stream_));
  }
}

void CCSD::set_cvs(int64_t n_core, const int64_t* core_mos, bool remove_core) {
  if (n_core == 0) {
    cvs_ = false;
    return;
  }
  ionization_ = false;
  require(n_core > 0 && core_mos != nullptr, "set_cvs: core MOs required");
  std::vector<unsigned char> flags(size_t(n_o), 0);
  for (int64_t c = 0; c < n_core; ++c) {
    require(core_mos[c] >= 0 && core_mos[c] < n_o, "set_cvs: a core MO must be an occupied orbital");
    flags[size_t(core_mos[c])] = 1;
  }
  cvs_core_.resize(size_t(n_o));
  ETG_HIP(hipMemcpyAsync(cvs_core_.as<unsigned char>(), flags.data(), size_t(n_o), hipMemcpyHostToDevice, stream_));
  synchronize();
  cvs_ = true;
  remove_core_ = remove_core;
}

void CCSD::set_ionization(bool on) {
  ionization_ = on;
  if (on) cvs_ = false;
}

void CCSD::project_cvs(double* x, int64_t n_vec) {
  if (ionization_)
    cvs_project(stream_, x, n_vec, n_t1, n_t1 + n_t2, n_v, nullptr, 2);
  else if (cvs_)
    cvs_project(stream_, x, n_vec, n_t1, n_t1 + n_t2, n_v, cvs_core_.as<unsigned char>(), remove_core_ ? 1 : 0);
}

int64_t CCSD::newton_raphson_update(double rel_threshold, int64_t max_iterations, Prec prec,
                                    double* micro_norms) {
  require(have_omega_, "newton_raphson_update: the residual has not been constructed");
  require(tape_ != nullptr, "newton_raphson_update: prepare_for_jacobian has not been called at these amplitudes");
  require(rel_threshold > 0.0, "newton_raphson_update: the relative threshold must be positive");
  const int64_t n = n_t1 + n_t2;
  const double norm = omega_norm();
  nr_rhs_.resize(sizeof(double) * size_t(n));
  update_.resize(sizeof(double) * size_t(n));
  ETG_HIP(hipMemcpyAsync(nr_rhs_.as<double>(), omega_.as<double>(), sizeof(double) * n, hipMemcpyDeviceToDevice,
                         stream_));
  scale(stream_, nr_rhs_.as<double>(), n, -1.0);
  const double threshold = rel_threshold * norm;
  const int64_t its = linear_davidson(0, nr_rhs_.as<double>(), nullptr, threshold, max_iterations, 50, prec,
                                      update_.as<double>(), micro_norms, std::min(1e-11, threshold));
  const dim3 g(unsigned(ceil_div(n_t1, 256)), 1);
  scale_doubles_diagonal_kernel<<<g, 256, 0, stream_>>>(update_.as<double>(), n_t1, n, 2.0);
  ETG_LAUNCH("scale_doubles_diagonal");
  axpy(stream_, n, 1.0, update_.as<double>(), t_.as<double>());
  // the DIIS step that follows records (Omega, t)
  ETG_HIP(hipMemcpyAsync(update_.as<double>(), omega_.as<double>(), sizeof(double) * n, hipMemcpyDeviceToDevice,
                         stream_));
  synchronize();
  have_update_ = true;
  have_omega_ = false;
  reset_record();
  return its;
}

// The record and everything captured from it are dropped together.
void CCSD::reset_record() {
  drop_transform_graphs();
  tape_.reset();
  have_lw_ = false;
  nb_cap_ = 0;
  reserved_nb_ = 0;
}

void CCSD::drop_transform_graphs() {
  for (auto& e : tgraphs_)
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
  tgraphs_.clear();
}

// R_k = A X_k (side 0) or A^T X_k (side 1), k < n_vec, columns of length
// n_t1 + n_t2; X and R are host or device pointers (the copies use unified
// addressing).  Columns of X are read before the same columns of R are
// written.  The vectors go through the passes in groups of up to nb_cap_, each
// group as the batch dimension of every product.  The first pass of a
// (side, precision, group size) runs eagerly and sizes every buffer; the same
// sequence is then captured into a HIP graph and replayed for later groups of
// that size (and later calls on the same record).  A graph is captured again
// when the pass workspace has moved since its capture.
void CCSD::transform_batch(int side, int64_t n_vec, const double* X, double* R, Prec prec) {
  require(tape_ != nullptr, side == 0 ? "jacobian_transformation: prepare_for_jacobian has not been called"
                                      : "jacobian_transpose_transformation: prepare_for_jacobian has not been called");
  require(side == 0 || side == 1, "jacobian_transformations: side must be 0 (right) or 1 (left)");
  require(n_vec >= 0, "jacobian_transformations: negative vector count");
  require(n_vec == 0 || (X != nullptr && R != nullptr), "jacobian_transformations: null pointer");
  if (n_vec == 0) return;
  const int64_t n = n_t1 + n_t2;
  if (tape_->dstride(t_.as<double>()) != n || tape_->dstride(omega_.as<double>()) != n)
    fail("jacobian_transformations: amplitude derivative layout");
  if (graph_tape_ != tape_.get()) {
    drop_transform_graphs();
    graph_tape_ = tape_.get();
  }
  if (nb_cap_ == 0) nb_cap_ = transform_batch_capacity();
  const int64_t group = std::min(nb_cap_, n_vec);
  if (group > reserved_nb_) {
    // room for `group` vectors everywhere; graphs captured before a move are stale
    bool moved = tape_->reserve_derivatives(group);
    moved |= reserve_ladder_vectors(group);
    moved |= jin_.resize(sizeof(double) * n * group);
    moved |= jout_.resize(sizeof(double) * n * group);
    if (moved) drop_transform_graphs();
    reserved_nb_ = group;
  }
  if (prec == Prec::BF16X3 && reserve_ladder_packs()) drop_transform_graphs();
  const bool graphs = use_graphs_ && !ledger_on_;
  pack_cache_ = false;                  // packs go to fixed workspaces inside the passes
  for (int64_t k0 = 0; k0 < n_vec; k0 += group) {
    const int64_t nb = std::min(group, n_vec - k0);
    ETG_HIP(hipMemcpyAsync(jin_.as<double>(), X + k0 * n, sizeof(double) * n * nb, hipMemcpyDefault,
                           stream_));
    TransformGraph* tg = nullptr;
    for (auto& e : tgraphs_)
      if (e.side == side && e.prec == prec && e.nb == nb) tg = &e;
    if (tg != nullptr && tg->ws_gen != gws_pass_.gen) {   // the workspace moved: capture again
      drop_transform_graphs();
      tg = nullptr;
    }
    if (graphs && tg != nullptr) {
      ETG_HIP(hipGraphLaunch(tg->exec, stream_));
    } else {
      transform_body(side, prec, nb);
      if (graphs) {
        synchronize();                  // buffers are sized; capture the same sequence
        const uint64_t gen = gws_pass_.gen;
        hipGraph_t graph = nullptr;
        ETG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
          transform_body(side, prec, nb);
        } catch (...) {
          (void)hipStreamEndCapture(stream_, &graph);
          if (graph) (void)hipGraphDestroy(graph);
          pack_cache_ = true;
          throw;
        }
        ETG_HIP(hipStreamEndCapture(stream_, &graph));
        if (gws_pass_.gen != gen) {     // cannot happen after an identical eager pass
          (void)hipGraphDestroy(graph);
          pack_cache_ = true;
          fail("jacobian_transformations: workspace moved during capture");
        }
        TransformGraph e;
        e.side = side;
        e.prec = prec;
        e.nb = nb;
        e.ws_gen = gen;
        ETG_HIP(hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        tgraphs_.push_back(e);
      }
    }
    ETG_HIP(hipMemcpyAsync(R + k0 * n, jout_.as<double>(), sizeof(double) * n * nb, hipMemcpyDefault,
                           stream_));
  }
  pack_cache_ = true;
  synchronize();
}

}  // namespace etg
