// CCSD multiplier (Lambda) equations on the device:  t-bar^T A + eta = 0.
//
// eT's DIIS multiplier solver (run_diis_cc_multipliers,
// src/solvers/cc/diis_cc_multipliers_class.F90:195-299) iterates
//   R = A^T t-bar + eta          construct_multiplier_equation_ccs (multiplier_equation_ccs.F90:58-94)
//   t-bar += -R / eps; e = -R/eps    quasi-Newton updater without amplitude scaling
//   t-bar = DIIS(e, t-bar)
// starting from t-bar1 = t1, t-bar2 = 2 t_aibj - t_ajbi
// (set_initial_multipliers_guess_ccsd, ccsd_class.F90:423-475).  A^T b is the
// recorded residual program run backwards (the left Jacobian transformation).
#include <cmath>

#include "ccsd.hpp"
#include "kernels.hpp"

namespace etg {
namespace {

constexpr int kBlock = 256;

__device__ inline int64_t tri_root_m(int64_t k) {
  int64_t x = int64_t((sqrt(8.0 * double(k) + 1.0) - 1.0) * 0.5);
  while (x * (x + 1) / 2 > k) --x;
  while ((x + 1) * (x + 2) / 2 <= k) ++x;
  return x;
}

unsigned grid_for_m(int64_t n) {
  return unsigned(std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16));
}

// eta_ai = 2 F_ia; eta_aibj = 4 g_iajb - 2 g_ibja, packed (construct_eta_ccsd,
// multiplier_equation_ccsd.F90:56-110).  F: n_mo x n_mo T1 Fock matrix;
// g: (o, v, o, v) = g_iajb of the ov Cholesky block.
__global__ void eta_kernel(const double* __restrict__ F, const double* __restrict__ g,
                           double* __restrict__ eta, int64_t n_o, int64_t n_v) {
  const int64_t N = n_o * n_v, n_mo = n_o + n_v;
  const int64_t n = N + N * (N + 1) / 2;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    if (k < N) {
      const int64_t a = k % n_v, i = k / n_v;
      eta[k] = 2.0 * F[i + (n_o + a) * n_mo];
    } else {
      const int64_t q = k - N;
      const int64_t X = tri_root_m(q), Y = q - X * (X + 1) / 2;
      const int64_t a = X % n_v, i = X / n_v, b = Y % n_v, j = Y / n_v;
      const double g_iajb = g[(i + n_o * a) + N * (j + n_o * b)];
      const double g_ibja = g[(i + n_o * b) + N * (j + n_o * a)];
      eta[k] = 4.0 * g_iajb - 2.0 * g_ibja;
    }
  }
}

// t-bar = (t1, 2 t_aibj - t_ajbi) (construct_packed_contravariant, reordering.F90:4542-4591)
__global__ void contravariant_guess_kernel(const double* __restrict__ t, double* __restrict__ tbar,
                                           int64_t n_o, int64_t n_v) {
  const int64_t N = n_o * n_v;
  const int64_t n = N + N * (N + 1) / 2;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    if (k < N) {
      tbar[k] = t[k];
    } else {
      const int64_t q = k - N;
      const int64_t X = tri_root_m(q), Y = q - X * (X + 1) / 2;
      const int64_t a = X % n_v, i = X / n_v, b = Y % n_v, j = Y / n_v;
      const int64_t P = a + n_v * j, Q = b + n_v * i;          // (aj), (bi)
      const int64_t hi = P > Q ? P : Q, lo = P > Q ? Q : P;
      tbar[k] = 2.0 * t[k] - t[N + hi * (hi + 1) / 2 + lo];
    }
  }
}

}  // namespace

void CCSD::set_initial_multipliers_guess() {
  require(have_t_, "set_initial_multipliers_guess: amplitudes required");
  const int64_t n = n_t1 + n_t2;
  tbar_.resize(sizeof(double) * n);
  contravariant_guess_kernel<<<grid_for_m(n), kBlock, 0, stream_>>>(t_.as<double>(), tbar_.as<double>(),
                                                                     n_o, n_v);
  ETG_LAUNCH("contravariant_guess");
  synchronize();
  have_tbar_ = true;
  have_mres_ = have_mupdate_ = false;
}

void CCSD::set_multipliers(const double* host) {
  require(host != nullptr, "set_multipliers: null pointer");
  const int64_t n = n_t1 + n_t2;
  tbar_.resize(sizeof(double) * n);
  ETG_HIP(hipMemcpyAsync(tbar_.as<double>(), host, sizeof(double) * n, hipMemcpyHostToDevice, stream_));
  synchronize();
  have_tbar_ = true;
  have_mres_ = have_mupdate_ = false;
}

void CCSD::get_multipliers(double* host) {
  require(have_tbar_, "get_multipliers: no multipliers set");
  require(host != nullptr, "get_multipliers: null pointer");
  ETG_HIP(hipMemcpyAsync(host, tbar_.as<double>(), sizeof(double) * (n_t1 + n_t2),
                         hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void CCSD::construct_eta(double* host) {
  require(have_L_ && have_F_, "construct_eta: Cholesky vectors and the T1 Fock matrix required");
  const int64_t n = n_t1 + n_t2;
  eta_.resize(sizeof(double) * n);
  eta_kernel<<<grid_for_m(n), kBlock, 0, stream_>>>(F_.as<double>(), g_ovov_.as<double>(),
                                                    eta_.as<double>(), n_o, n_v);
  ETG_LAUNCH("eta");
  if (host)
    ETG_HIP(hipMemcpyAsync(host, eta_.as<double>(), sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void CCSD::construct_multiplier_equation(double* host, Prec prec) {
  require(have_tbar_, "construct_multiplier_equation: no multipliers (set an initial guess first)");
  require(tape_ != nullptr,
          "construct_multiplier_equation: prepare_for_jacobian has not been called at these amplitudes");
  const int64_t n = n_t1 + n_t2;
  construct_eta(nullptr);
  mres_.resize(sizeof(double) * n);
  transform_batch(1, 1, tbar_.as<double>(), mres_.as<double>(), prec);
  axpy(stream_, n, 1.0, eta_.as<double>(), mres_.as<double>());
  if (host)
    ETG_HIP(hipMemcpyAsync(host, mres_.as<double>(), sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
  synchronize();
  have_mres_ = true;
}

double CCSD::multiplier_residual_norm() {
  require(have_mres_, "multiplier_residual_norm: the multiplier equation has not been constructed");
  double* res = red_result_.as<double>() + 2;
  sum_squares(stream_, mres_.as<double>(), n_t1 + n_t2, red_partial_.as<double>(), res);
  double e = 0.0;
  ETG_HIP(hipMemcpyAsync(&e, res, sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  return std::sqrt(e);
}

void CCSD::multiplier_quasi_newton_update() {
  require(have_mres_, "multiplier_quasi_newton_update: the multiplier equation has not been constructed");
  require(have_eps_, "multiplier_quasi_newton_update: orbital energies required");
  const int64_t n = n_t1 + n_t2;
  mupdate_.resize(sizeof(double) * n);
  quasi_newton_step(stream_, eps_.as<double>(), mres_.as<double>(), tbar_.as<double>(),
                    mupdate_.as<double>(), n_o, n_v, n, 1.0);
  have_mres_ = false;          // the residual belongs to the multipliers before the step
  have_mupdate_ = true;
}

void CCSD::multiplier_diis_update() {
  require(have_mupdate_, "multiplier_diis_update: no update vector (call the quasi-Newton step first)");
  diis_step(diis_, mupdate_.as<double>(), tbar_.as<double>());
  have_mupdate_ = false;
}

// CC mean value of a one-electron operator X (MO basis, n_mo x n_mo) with the
// Lagrangian density: <X> = <HF|(1 + Lambda) e^-T X e^T|HF> (construct_gs_density,
// mean_value_ccs.F90:94-122, mean_value_doubles.F90), which is dL/de at e = 0 for
// h -> h + e X, L = E + t-bar . Omega at fixed t and t-bar:
//   dE/de = 2 sum_i X_ii + 2 sum_ai X_ia t_ai   (F_ia(e) = e X_ia on the HF reference),
//   t-bar . dOmega/de is one forward replay of the recorded residual program
//   along dh = X (the record keeps h differentiable).
double CCSD::one_electron_expectation(const double* X) {
  require(X != nullptr, "one_electron_expectation: null operator");
  require(tape_ != nullptr, "one_electron_expectation: prepare_for_jacobian has not been called at these amplitudes");
  require(have_tbar_, "one_electron_expectation: multipliers required");
  const int64_t n = n_t1 + n_t2, nm = n_mo;
  std::vector<double> t1(size_t(std::max<int64_t>(1, n_t1)));
  ETG_HIP(hipMemcpyAsync(t1.data(), t_.as<double>(), sizeof(double) * n_t1, hipMemcpyDeviceToHost, stream_));
  tape_->zero_derivatives(1);
  double* dh = tape_->deriv(h_.as<double>());
  require(dh != nullptr, "one_electron_expectation: the record holds h as a constant");
  ETG_HIP(hipMemcpyAsync(dh, X, sizeof(double) * nm * nm, hipMemcpyDefault, stream_));
  tape_->tangent(Prec::F64, 1);
  double* res = red_result_.as<double>() + 3;
  dot(stream_, tbar_.as<double>(), tape_->deriv(omega_.as<double>()), n, red_partial_.as<double>(), res);
  double lag = 0.0;
  ETG_HIP(hipMemcpyAsync(&lag, res, sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  double e = 0.0;
  for (int64_t i = 0; i < n_o; ++i) e += 2.0 * X[i + nm * i];
  for (int64_t i = 0; i < n_o; ++i)
    for (int64_t a = 0; a < n_v; ++a) e += 2.0 * X[i + nm * (n_o + a)] * t1[size_t(a + n_v * i)];
  return e + lag;
}

// D_pq = dL/dh_pq, so that <X> = sum_pq D_pq X_pq for every one-electron X: the
// backward replay from t-bar on the residual gives d(t-bar . Omega)/dh in h's
// adjoint; the energy adds 2 on the occupied diagonal and 2 t_ai at (i, a).
void CCSD::one_electron_density(double* D) {
  require(D != nullptr, "one_electron_density: null output");
  require(tape_ != nullptr, "one_electron_density: prepare_for_jacobian has not been called at these amplitudes");
  require(have_tbar_, "one_electron_density: multipliers required");
  const int64_t n = n_t1 + n_t2, nm = n_mo;
  std::vector<double> t1(size_t(std::max<int64_t>(1, n_t1)));
  ETG_HIP(hipMemcpyAsync(t1.data(), t_.as<double>(), sizeof(double) * n_t1, hipMemcpyDeviceToHost, stream_));
  tape_->zero_derivatives(1);
  double* hb = tape_->deriv(h_.as<double>());
  require(hb != nullptr, "one_electron_density: the record holds h as a constant");
  ETG_HIP(hipMemcpyAsync(tape_->deriv(omega_.as<double>()), tbar_.as<double>(), sizeof(double) * n,
                         hipMemcpyDeviceToDevice, stream_));
  tape_->adjoint(Prec::F64, 1);
  ETG_HIP(hipMemcpyAsync(D, hb, sizeof(double) * nm * nm, hipMemcpyDeviceToHost, stream_));
  synchronize();
  for (int64_t i = 0; i < n_o; ++i) D[i + nm * i] += 2.0;
  for (int64_t i = 0; i < n_o; ++i)
    for (int64_t a = 0; a < n_v; ++a) D[i + nm * (n_o + a)] += 2.0 * t1[size_t(a + n_v * i)];
}

}  // namespace etg
