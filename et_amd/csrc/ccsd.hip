// CCSD ground-state residual, energy, T1 transformation and Fock matrix on
// one MI355X.  The sequence of contractions follows eT's factorisation
// (src/wavefunctions/ccsd/omega_ccsd.F90 and the doubles/ccs submodules it
// calls); every product runs on the matrix-core kernels in gemm.hip and every
// reorder on the bandwidth kernels in kernels.hip.  All tensors stay in HBM
// between calls; host pointers are only read/written at the boundary.
#include "ccsd.hpp"

#include <algorithm>
#include <cmath>

#include "kernels.hpp"

namespace etg {

namespace {
enum Slot { S_T = 0, S_U, S_A, S_B, S_C, S_D, S_E, S_F, kSlots };
}

CCSD::CCSD(int64_t no, int64_t nv, int64_t nJ) : n_o(no), n_v(nv), n_J(nJ) {
  if (no < 1 || nv < 1 || nJ < 1) fail("etg_create: n_o, n_v and n_J must be positive");
  n_mo = n_o + n_v;
  n_t1 = n_o * n_v;
  n_t2 = n_t1 * (n_t1 + 1) / 2;
  ETG_HIP(hipStreamCreate(&stream_));
  const int64_t J = n_J, o = n_o, v = n_v;
  Lmo_oo_.resize(sizeof(double) * J * o * o);
  Lmo_ov_.resize(sizeof(double) * J * o * v);
  Lmo_vo_.resize(sizeof(double) * J * v * o);
  Lmo_vv_.resize(sizeof(double) * J * v * v);
  Lt_oo_.resize(sizeof(double) * J * o * o);
  Lt_vo_.resize(sizeof(double) * J * v * o);
  Lt_vv_.resize(sizeof(double) * J * v * v);
  Lt_vvT_.resize(sizeof(double) * J * v * v);
  h_.resize(sizeof(double) * n_mo * n_mo);
  eps_.resize(sizeof(double) * n_mo);
  F_.resize(sizeof(double) * n_mo * n_mo);
  t_.resize(sizeof(double) * (n_t1 + n_t2));
  omega_.resize(sizeof(double) * (n_t1 + n_t2));
  g_ovov_.resize(sizeof(double) * n_t1 * n_t1);
  red_partial_.resize(sizeof(double) * kReduceBlocks);
  red_result_.resize(sizeof(double) * 8);
  work_.resize(kSlots);
  // small scratch, sized once for every use below
  const int64_t N = n_t1, n = n_mo;
  tmp_small_[0].resize(sizeof(double) * n * n);
  tmp_small_[1].resize(sizeof(double) * n * n);
  tmp_small_[2].resize(sizeof(double) * std::max(std::max(n * n, J), o * o * o * o));
  tmp_small_[3].resize(sizeof(double) * std::max(std::max(J * N, J * o * o), o * o * o * o));
  // default ladder batch: G blocks of at most ~12 GB
  const double budget = 12e9 / 8.0;
  int64_t nb = int64_t(std::sqrt(budget) / double(v));
  set_ladder_batch(std::max<int64_t>(1, std::min<int64_t>(nb, v)));
}

CCSD::~CCSD() {
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

double* CCSD::work(int slot) {
  // (ov)^2 elements, or more when n_o > n_v (the o^3 v and o^4 integral blocks share the slots)
  const int64_t elems = std::max(n_t1 * n_t1, n_o * n_o * n_o * std::max(n_o, n_v));
  work_[slot].resize(sizeof(double) * elems);
  return work_[slot].as<double>();
}

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

void CCSD::record_gemm(const GemmLaunch& info, double flops, hipEvent_t a, hipEvent_t b) {
  ETG_HIP(hipEventRecord(b, stream_));
  ledger_recs_.push_back(LedgerRec{info.kernel, flops, a, b});
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
      out.push_back(LedgerEntry{r.kernel, 0, 0.0, 0.0});
      it = out.end() - 1;
    }
    it->launches += 1;
    it->seconds += double(ms) * 1e-3;
    it->flops += r.flops;
  }
  return out;
}

void CCSD::gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                Prec prec) {
  const double flops = 2.0 * double(m) * double(n) * double(k);
  hipEvent_t a = nullptr, b = nullptr;
  if (ledger_on_) {
    a = take_event();
    b = take_event();
    ETG_HIP(hipEventRecord(a, stream_));
  }
  const GemmLaunch info =
      etg::gemm(stream_, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, &gws_);
  if (ledger_on_) record_gemm(info, flops, a, b);
  last_omega_flops += flops;
}

void CCSD::gemm_batched(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
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
  hipEvent_t a = nullptr, b = nullptr;
  if (ledger_on_) {
    a = take_event();
    b = take_event();
    ETG_HIP(hipEventRecord(a, stream_));
  }
  const GemmLaunch info = etg::gemm(stream_, g, prec, &gws_);
  if (ledger_on_) record_gemm(info, gemm_flops(g), a, b);
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
  const int64_t o = n_o, v = n_v, J = n_J;
  copy_block(stream_, L, J, n_mo, 0, o, 0, o, Lmo_oo_.as<double>());
  copy_block(stream_, L, J, n_mo, 0, o, o, v, Lmo_ov_.as<double>());
  copy_block(stream_, L, J, n_mo, o, v, 0, o, Lmo_vo_.as<double>());
  copy_block(stream_, L, J, n_mo, o, v, o, v, Lmo_vv_.as<double>());
  // g_ovov = L_ov^T L_ov depends only on the undressed ov block
  gemm('T', 'N', o * v, o * v, J, 1.0, Lmo_ov_.as<double>(), J, Lmo_ov_.as<double>(), J, 0.0,
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
  ETG_HIP(hipMemcpyAsync(h_.as<double>(), h, sizeof(double) * n_mo * n_mo, hipMemcpyHostToDevice,
                         stream_));
  synchronize();
  have_h_ = true;
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
  ETG_HIP(hipMemcpyAsync(t_.as<double>(), t, sizeof(double) * (n_t1 + n_t2),
                         hipMemcpyHostToDevice, stream_));
  synchronize();
  have_t_ = true;
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

void CCSD::construct_t1_cholesky() {
  require(have_L_, "construct_t1_cholesky: Cholesky vectors not set");
  const int64_t o = n_o, v = n_v, J = n_J;
  const double* t1 = t_.as<double>();
  hipStream_t s = stream_;
  const size_t b_oo = sizeof(double) * J * o * o, b_vo = sizeof(double) * J * v * o,
               b_vv = sizeof(double) * J * v * v;
  // oo: L_ij = L_ij + sum_a L_ia t_aj          ((J i) x a) (a x j)
  ETG_HIP(hipMemcpyAsync(Lt_oo_.as<double>(), Lmo_oo_.as<double>(), b_oo, hipMemcpyDeviceToDevice, s));
  gemm('N', 'N', J * o, o, v, 1.0, Lmo_ov_.as<double>(), J * o, t1, v, 1.0, Lt_oo_.as<double>(),
       J * o, Prec::F64);
  // vo: L_ai = L_ai + sum_b L_ab t_bi - sum_j t_aj L'_ji   (L' = dressed oo block)
  ETG_HIP(hipMemcpyAsync(Lt_vo_.as<double>(), Lmo_vo_.as<double>(), b_vo, hipMemcpyDeviceToDevice, s));
  gemm('N', 'N', J * v, o, v, 1.0, Lmo_vv_.as<double>(), J * v, t1, v, 1.0, Lt_vo_.as<double>(),
       J * v, Prec::F64);
  gemm_batched('N', 'T', J, v, o, -1.0, Lt_oo_.as<double>(), J, J * o, t1, v, 0, 1.0,
               Lt_vo_.as<double>(), J, J * v, o, Prec::F64);
  // vv: L_ab = L_ab - sum_j t_aj L_jb         batched over b: (J x j)(j x a)
  ETG_HIP(hipMemcpyAsync(Lt_vv_.as<double>(), Lmo_vv_.as<double>(), b_vv, hipMemcpyDeviceToDevice, s));
  gemm_batched('N', 'T', J, v, o, -1.0, Lmo_ov_.as<double>(), J, J * o, t1, v, 0, 1.0,
               Lt_vv_.as<double>(), J, J * v, v, Prec::F64);
  // (J, c, a) copy of the vv block: a-batches of the ladder become contiguous columns
  const int64_t d[4] = {J, v, v, 1};
  sort4(s, Lt_vv_.as<double>(), d, "1324", Lt_vvT_.as<double>());
  have_F_ = false;
}

// ---------------------------------------------------------------------------
// T1 Fock matrix (fock_ccs.F90:242-1109): F = h^T1 + sum_k (2 g_pqkk - g_pkkq)
// ---------------------------------------------------------------------------

namespace {
__global__ void t1_xy_kernel(const double* t1, double* X, double* Y, int64_t n_o, int64_t n_v) {
  const int64_t n = n_o + n_v;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n * n;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t p = k % n, q = k / n;
    double x = (p == q) ? 1.0 : 0.0, y = x;
    if (p >= n_o && q < n_o) x = -t1[(p - n_o) + n_v * q];  // X(o+a, i) = -t_ai
    if (p < n_o && q >= n_o) y = t1[(q - n_o) + n_v * p];   // Y(i, o+a) = t_ai
    X[k] = x;
    Y[k] = y;
  }
}

__global__ void trace_oo_kernel(const double* Loo, int64_t nJ, int64_t n_o, double* XJ) {
  for (int64_t J = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; J < nJ;
       J += int64_t(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int64_t k = 0; k < n_o; ++k) s += Loo[J + nJ * (k + n_o * k)];
    XJ[J] = s;
  }
}

// vector f(b + v j) = F(j, o + b)
__global__ void fock_ov_vector_kernel(const double* F, int64_t n_mo, int64_t n_o, int64_t n_v,
                                      double* f) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n_o * n_v;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = k % n_v, j = k / n_v;
    f[k] = F[j + n_mo * (n_o + b)];
  }
}

// dst(r, c) = alpha * src(r, c) + beta * dst   for an m x n sub-matrix
__global__ void copy_sub_kernel(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
                                int64_t n, double alpha, double beta) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < m * n;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = k % m, c = k / m;
    const double v = alpha * src[r + lds * c];
    dst[r + ldd * c] = (beta == 0.0) ? v : v + beta * dst[r + ldd * c];
  }
}

void copy_sub(hipStream_t s, const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
              int64_t n, double alpha = 1.0, double beta = 0.0) {
  const int64_t tot = m * n;
  if (tot == 0) return;
  copy_sub_kernel<<<unsigned(std::min<int64_t>(ceil_div(tot, 256), 1 << 20)), 256, 0, s>>>(
      src, lds, dst, ldd, m, n, alpha, beta);
  ETG_LAUNCH("copy_sub");
}
}  // namespace

void CCSD::construct_fock() {
  require(have_L_ && have_h_, "construct_fock: Cholesky vectors and one-electron integrals required");
  const int64_t o = n_o, v = n_v, J = n_J, n = n_mo;
  hipStream_t s = stream_;
  double* X = tmp_small_[0].as<double>();
  double* Y = tmp_small_[1].as<double>();
  double* Z = tmp_small_[2].as<double>();
  double* F = F_.as<double>();
  // h^T1 = X h Y^T
  t1_xy_kernel<<<unsigned(ceil_div(n * n, 256)), 256, 0, s>>>(t_.as<double>(), X, Y, o, v);
  ETG_LAUNCH("t1_xy");
  gemm('N', 'N', n, n, n, 1.0, X, n, h_.as<double>(), n, 0.0, Z, n, Prec::F64);
  gemm('N', 'T', n, n, n, 1.0, Z, n, Y, n, 0.0, F, n, Prec::F64);
  // Coulomb: F_pq += 2 sum_J L_pq X_J,  X_J = sum_k L_kk
  double* XJ = tmp_small_[2].as<double>();
  trace_oo_kernel<<<unsigned(ceil_div(J, 256)), 256, 0, s>>>(Lt_oo_.as<double>(), J, o, XJ);
  ETG_LAUNCH("trace_oo");
  const double* Lov = Lmo_ov_.as<double>();
  gemm('T', 'N', o * o, 1, J, 2.0, Lt_oo_.as<double>(), J, XJ, J, 0.0, X, o * o, Prec::F64);
  copy_sub(s, X, o, F, n, o, o, 1.0, 1.0);
  gemm('T', 'N', o * v, 1, J, 2.0, Lov, J, XJ, J, 0.0, X, o * v, Prec::F64);
  copy_sub(s, X, o, F + n * o, n, o, v, 1.0, 1.0);
  gemm('T', 'N', v * o, 1, J, 2.0, Lt_vo_.as<double>(), J, XJ, J, 0.0, X, v * o, Prec::F64);
  copy_sub(s, X, v, F + o, n, v, o, 1.0, 1.0);
  gemm('T', 'N', v * v, 1, J, 2.0, Lt_vv_.as<double>(), J, XJ, J, 0.0, X, v * v, Prec::F64);
  copy_sub(s, X, v, F + o + n * o, n, v, v, 1.0, 1.0);
  // Exchange: F_pq -= sum_{J,k} L_pk L_kq with left factors reordered to (J, k, p)
  double* Lleft = tmp_small_[3].as<double>();
  const int64_t d_oo[4] = {J, o, o, 1};
  sort4(s, Lt_oo_.as<double>(), d_oo, "1324", Lleft);               // (J, k, i)
  gemm('T', 'N', o, o, J * o, -1.0, Lleft, J * o, Lt_oo_.as<double>(), J * o, 1.0, F, n, Prec::F64);
  gemm('T', 'N', o, v, J * o, -1.0, Lleft, J * o, Lov, J * o, 1.0, F + n * o, n, Prec::F64);
  const int64_t d_vo[4] = {J, v, o, 1};
  sort4(s, Lt_vo_.as<double>(), d_vo, "1324", Lleft);               // (J, k, a)
  gemm('T', 'N', v, o, J * o, -1.0, Lleft, J * o, Lt_oo_.as<double>(), J * o, 1.0, F + o, n, Prec::F64);
  gemm('T', 'N', v, v, J * o, -1.0, Lleft, J * o, Lov, J * o, 1.0, F + o + n * o, n, Prec::F64);
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
  const int64_t N = n_t1;
  double* g = work(S_A);
  gemm('T', 'N', N, N, n_J, 1.0, Lt_vo_.as<double>(), n_J, Lt_vo_.as<double>(), n_J, 0.0, g, N,
       Prec::F64);
  mp2_guess(stream_, g, eps_.as<double>(), t_.as<double>() + n_t1, n_o, n_v);
  synchronize();
  have_t_ = true;
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

void CCSD::construct_omega(Prec prec) {
  require(have_L_ && have_t_, "construct_omega: Cholesky vectors and amplitudes required");
  if (!have_F_) construct_fock();
  last_omega_flops = 0.0;
  const int64_t N = n_t1;
  hipStream_t s = stream_;
  // t_aibj (full) and u_aibj = 2 t_aibj - t_ajbi      (omega_ccsd.F90:881-914)
  double* T = work(S_T);
  double* U = work(S_U);
  squareup(s, t_.as<double>() + n_t1, T, N);
  const int64_t d[4] = {n_v, n_o, n_v, n_o};
  sort4(s, T, d, "1234", U, 2.0, 0.0);
  sort4(s, T, d, "1432", U, -1.0, 1.0);
  // singles: F_ai + doubles A1, B1, C1
  omega_singles(prec);
  // doubles: E2, symmetric sum, B2, 1/2 diagonal, pack, then C2/D2 and A2
  omega_doubles_e2_b2(prec);
  omega_c2_d2(prec);
  omega_a2(prec);
}

void CCSD::omega_singles(Prec prec) {
  const int64_t o = n_o, v = n_v, J = n_J, N = n_t1, n = n_mo;
  hipStream_t s = stream_;
  double* om1 = omega_.as<double>();
  const double* U = work(S_U);
  // Omega_ai = F_ai                                    (omega_ccs.F90:62-86)
  copy_sub(s, F_.as<double>() + o, n, om1, v, v, o);
  // A1: X_Jbi = sum_cj L^mo_Jcj u_cjbi ; Omega_ai += sum_Jb L^T1_Jab X_Jbi   (omega_doubles.F90:36-139)
  double* X = tmp_small_[3].as<double>();
  gemm('N', 'N', J, N, N, 1.0, Lmo_vo_.as<double>(), J, U, N, 0.0, X, J, prec);
  gemm('T', 'N', v, o, J * v, 1.0, Lt_vvT_.as<double>(), J * v, X, J * v, 1.0, om1, v, prec);
  // B1: Omega_ai -= sum_jbk u_ajbk g_kbji             (omega_doubles.F90:141-205)
  double* g_kbji = work(S_A);
  double* g_jbki = work(S_B);
  gemm('T', 'N', o * v, o * o, J, 1.0, Lmo_ov_.as<double>(), J, Lt_oo_.as<double>(), J, 0.0,
       g_kbji, o * v, prec);
  const int64_t d[4] = {o, v, o, o};
  sort4(s, g_kbji, d, "3214", g_jbki);
  gemm('N', 'N', v, o, o * o * v, -1.0, U, v, g_jbki, o * o * v, 1.0, om1, v, prec);
  // C1: Omega_ai += sum_bj u_aibj F_jb                (omega_doubles.F90:207-250)
  double* f = work(S_C);
  fock_ov_vector_kernel<<<unsigned(ceil_div(N, 256)), 256, 0, s>>>(F_.as<double>(), n, o, v, f);
  ETG_LAUNCH("fock_ov_vector");
  gemm('N', 'N', N, 1, N, 1.0, U, N, f, N, 1.0, om1, N, prec);
}

void CCSD::omega_doubles_e2_b2(Prec prec) {
  const int64_t o = n_o, v = n_v, J = n_J, N = n_t1, n = n_mo;
  hipStream_t s = stream_;
  const double* T = work(S_T);
  const double* U = work(S_U);
  const double* F = F_.as<double>();
  const double* g_ovov = g_ovov_.as<double>();
  const int64_t d[4] = {v, o, v, o};
  // E2 (omega_ccsd.F90:712-878)
  double* u_e2 = work(S_A);                 // u_kl^bd = 2 t_bkdl - t_bldk = u(b,k,d,l)
  sort4(s, U, d, "1432", u_e2);
  double* om = work(S_B);
  fill(s, om, N * N, 0.0);
  double* Xbc = tmp_small_[0].as<double>();
  double* Ykj = tmp_small_[1].as<double>();
  copy_sub(s, F + o + n * o, n, Xbc, v, v, v);
  gemm('N', 'N', v, v, o * v * o, -1.0, u_e2, v, g_ovov, o * v * o, 1.0, Xbc, v, prec);
  gemm('N', 'N', v, o * v * o, v, 1.0, Xbc, v, T, v, 1.0, om, v, prec);
  copy_sub(s, F, n, Ykj, o, o, o);
  gemm('N', 'N', o, o, v * o * v, 1.0, g_ovov, o, u_e2, v * o * v, 1.0, Ykj, o, prec);
  gemm('N', 'N', v * o * v, o, o, -1.0, T, v * o * v, Ykj, o, 1.0, om, v * o * v, prec);
  // symmetric sum
  double* om2 = work(S_C);
  symmetric_sum(s, om, om2, N);
  // B2 in (a,b,i,j) order (omega_ccsd.F90:483-583)
  double* om_abij = work(S_A);
  sort4(s, om2, d, "1324", om_abij);
  double* t_abij = work(S_B);
  sort4(s, T, d, "1324", t_abij);
  double* g_vovo = work(S_C);
  gemm('T', 'N', N, N, J, 1.0, Lt_vo_.as<double>(), J, Lt_vo_.as<double>(), J, 0.0, g_vovo, N, prec);
  sort4(s, g_vovo, d, "1324", om_abij, 1.0, 1.0);
  double* g_kilj = tmp_small_[3].as<double>();
  double* X_klij = tmp_small_[2].as<double>();
  gemm('T', 'N', o * o, o * o, J, 1.0, Lt_oo_.as<double>(), J, Lt_oo_.as<double>(), J, 0.0, g_kilj,
       o * o, prec);
  const int64_t d_oooo[4] = {o, o, o, o};
  sort4(s, g_kilj, d_oooo, "1324", X_klij);
  double* g_klcd = work(S_C);
  const int64_t d_ovov[4] = {o, v, o, v};
  sort4(s, g_ovov, d_ovov, "1324", g_klcd);
  gemm('N', 'N', o * o, o * o, v * v, 1.0, g_klcd, o * o, t_abij, v * v, 1.0, X_klij, o * o, prec);
  gemm('N', 'N', v * v, o * o, o * o, 1.0, t_abij, v * v, X_klij, o * o, 1.0, om_abij, v * v, prec);
  // back to (a,i,b,j), 1/2 on the diagonal and pack into omega2   (omega_ccsd.F90:97-102)
  double* om_aibj = work(S_C);
  const int64_t d_vvoo[4] = {v, v, o, o};
  sort4(s, om_abij, d_vvoo, "1324", om_aibj);
  packin(s, om_aibj, omega_.as<double>() + n_t1, N, 1.0, 0.0, 0.5);
}

void CCSD::omega_c2_d2(Prec prec) {
  // omega_ccsd.F90:586-709: four (ov) x (ov) products
  const int64_t o = n_o, v = n_v, J = n_J, N = n_t1;
  hipStream_t s = stream_;
  const double* Tfull = work(S_T);
  const double* U = work(S_U);
  const int64_t d[4] = {v, o, v, o};
  double* Tm = work(S_A);                    // T[(p,q),(r,s)] = t_psrq
  sort4(s, Tfull, d, "1432", Tm);
  double* gtmp = work(S_B);
  double* C = work(S_C);
  double* D = work(S_D);
  gemm('T', 'N', v * v, o * o, J, 1.0, Lt_vv_.as<double>(), J, Lt_oo_.as<double>(), J, 0.0, gtmp,
       v * v, prec);                         // g_abij
  const int64_t d_vvoo[4] = {v, v, o, o};
  sort4(s, gtmp, d_vvoo, "2314", C);         // C(b,i,a,j) = g_abij
  sort4(s, C, d, "1234", D, -1.0, 0.0);      // D = -C
  gemm('T', 'N', N, N, J, 1.0, Lt_vo_.as<double>(), J, Lmo_ov_.as<double>(), J, 0.0, gtmp, N,
       prec);                                // g_aijb
  const int64_t d_voov[4] = {v, o, o, v};
  sort4(s, gtmp, d_voov, "4312", D, 2.0, 1.0);   // D(c,k,a,i) += 2 g_aikc
  double* G = work(S_B);
  const int64_t d_ovov[4] = {o, v, o, v};
  sort4(s, g_ovov_.as<double>(), d_ovov, "2341", G);   // G(c,l,d,k) = g_kcld
  gemm('N', 'N', N, N, N, -0.5, G, N, Tm, N, 1.0, C, N, prec);
  double* Gc = work(S_E);                    // 2 G(p,s,r,q) - G(p,q,r,s)
  sort4(s, G, d, "1432", Gc, 2.0, 0.0);
  sort4(s, G, d, "1234", Gc, -1.0, 1.0);
  gemm('N', 'N', N, N, N, 0.5, Gc, N, U, N, 1.0, D, N, prec);
  double* W = work(S_B);
  gemm('N', 'N', N, N, N, -1.5, Tm, N, C, N, 0.0, W, N, prec);
  symmetrize_and_pack(s, W, omega_.as<double>() + n_t1, v, o, 1.0, 1.0 / 3.0, 2.0 / 3.0);
  gemm('N', 'N', N, N, N, 0.5, U, N, D, N, 0.0, W, N, prec);
  symmetrize_and_pack(s, W, omega_.as<double>() + n_t1, v, o, 1.0, 1.0, 0.0);
}

// ---------------------------------------------------------------------------
// A2 ladder (omega_ccsd.F90:113-480) with +/- symmetrisation
//   W_abij = sum_cd g_acbd t_cidj = w+_{ij,ab} + w-_{ij,ab}
//   w+ = 1/2 sum_{c<=d} t+_{ij,cd} g+_{cd,ab},  w- = 1/2 sum_{c<d} t-_{ij,cd} g-_{cd,ab}
// g is generated per (a-batch, b-batch) block as G = L_vvT(:, A)^T L_vvT(:, B)
// and never stored whole; w+- for all pairs are gathered into packed Omega at the end.
// ---------------------------------------------------------------------------

void CCSD::set_ladder_batch(int64_t nb) {
  if (nb < 1) fail("ladder batch must be positive");
  synchronize();
  ladder_nb_ = std::min(nb, n_v);
  build_ladder_plan();
}

void CCSD::build_ladder_plan() {
  const int64_t v = n_v, nb = ladder_nb_;
  blocks_.clear();
  std::vector<int32_t> pa, pb, col(size_t(v * v), -1);
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
  pair_a_.resize(sizeof(int32_t) * pa.size());
  pair_b_.resize(sizeof(int32_t) * pb.size());
  pair_col_.resize(sizeof(int32_t) * col.size());
  ETG_HIP(hipMemcpy(pair_a_.as<int32_t>(), pa.data(), sizeof(int32_t) * pa.size(), hipMemcpyHostToDevice));
  ETG_HIP(hipMemcpy(pair_b_.as<int32_t>(), pb.data(), sizeof(int32_t) * pb.size(), hipMemcpyHostToDevice));
  ETG_HIP(hipMemcpy(pair_col_.as<int32_t>(), col.data(), sizeof(int32_t) * col.size(), hipMemcpyHostToDevice));
  const int64_t op = n_o * (n_o + 1) / 2, om = n_o * (n_o - 1) / 2;
  const int64_t vp = v * (v + 1) / 2, vm = v * (v - 1) / 2;
  tp_.resize(sizeof(double) * op * vp);
  tm_.resize(sizeof(double) * std::max<int64_t>(1, om * vm));
  wp_.resize(sizeof(double) * op * vp);
  wm_.resize(sizeof(double) * std::max<int64_t>(1, om * vp));
  gblk_.resize(sizeof(double) * v * v * std::min(nb, v) * std::min(nb, v));
  gp_.resize(sizeof(double) * vp * maxpairs);
  gm_.resize(sizeof(double) * std::max<int64_t>(1, vm * maxpairs));
}

void CCSD::omega_a2(Prec prec) {
  const int64_t o = n_o, v = n_v, J = n_J;
  const int64_t op = o * (o + 1) / 2, om = o * (o - 1) / 2;
  const int64_t vp = v * (v + 1) / 2, vm = v * (v - 1) / 2;
  hipStream_t s = stream_;
  ladder_t_pm(s, t_.as<double>() + n_t1, tp_.as<double>(), tm_.as<double>(), o, v);
  const double* LT = Lt_vvT_.as<double>();
  for (const Block& b : blocks_) {
    gemm('T', 'N', v * b.la, v * b.lb, J, 1.0, LT + J * v * b.a0, J, LT + J * v * b.b0, J, 0.0,
         gblk_.as<double>(), v * b.la, prec);
    ladder_g_pm(s, gblk_.as<double>(), v, b.la, b.lb, pair_a_.as<int32_t>() + b.col0,
                pair_b_.as<int32_t>() + b.col0, b.npairs, gp_.as<double>(), gm_.as<double>());
    gemm('N', 'N', op, b.npairs, vp, 0.5, tp_.as<double>(), op, gp_.as<double>(), vp, 0.0,
         wp_.as<double>() + op * b.col0, op, prec);
    if (om > 0 && vm > 0)
      gemm('N', 'N', om, b.npairs, vm, 0.5, tm_.as<double>(), om, gm_.as<double>(), vm, 0.0,
           wm_.as<double>() + om * b.col0, om, prec);
  }
  ladder_gather(s, wp_.as<double>(), wm_.as<double>(), pair_col_.as<int32_t>(),
                omega_.as<double>() + n_t1, o, v);
}

}  // namespace etg
