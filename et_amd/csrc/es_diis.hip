// eT's DIIS solver for excited states on the device (run_diis_cc_es,
// src/solvers/cc/diis_cc_es_class.F90:260-492).
//
// Each state j keeps its own trial vector X_j and DIIS space in HBM.  Per
// iteration, the unconverged states go through one transformation pass
// together (R = A X or A^T X, the vectors as the batch dimension), then each
//   E_j = X_j . R_j,  R_j -= E_j X_j,  converged if ||R_j|| and |E_j - E_j(prev)|
//   are below the thresholds (the residual alone in the first iteration),
// and otherwise
//   R_j = -R_j / (eps - E_j)        shifted preconditioner (precondition_tool_class.F90:100-164)
//   X_j += R_j, X_j = DIIS(R_j, X_j) over the state's records, X_j /= ||X_j||.
// Start vectors: unit vectors on the lowest orbital differences (es_start_vector_tool,
// es_start_vector_tool_class.F90:60-111) or the caller's.  On convergence the roots
// are sorted by energy.  Only scalars reach the host.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "ccsd.hpp"
#include "kernels.hpp"

namespace etg {
namespace {

constexpr int kBlock = 256;

unsigned grid_es(int64_t n) { return unsigned(std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16)); }

// R = -R / (d - shift), or -R where |d - shift| < 1e-8.
__global__ void shifted_precondition_kernel(double* __restrict__ R, const double* __restrict__ d, int64_t n,
                                            double shift) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x) {
    const double den = d[k] - shift;
    R[k] = fabs(den) < 1e-8 ? -R[k] : -R[k] / den;
  }
}

}  // namespace

int64_t CCSD::diis_excited_states(int side, int64_t n_states, double residual_threshold,
                                  double energy_threshold, int64_t max_iterations, int64_t diis_dimension,
                                  Prec prec, const double* start, double* energies, double* vectors,
                                  double* residual_norms, double* energy_history) {
  require(tape_ != nullptr, "diis_excited_states: prepare_for_jacobian has not been called");
  require(have_eps_, "diis_excited_states: orbital energies required");
  require(side == 0 || side == 1, "diis_excited_states: side must be 0 (right) or 1 (left)");
  require(energies != nullptr, "diis_excited_states: null energies");
  const int64_t n = n_t1 + n_t2, s = n_states;
  require(s >= 1 && s < n, "diis_excited_states: n_states must be in [1, n_t1 + n_t2)");
  require(max_iterations >= 1, "diis_excited_states: max_iterations must be positive");
  es_X_.resize(sizeof(double) * size_t(n) * size_t(s));
  es_R_.resize(sizeof(double) * size_t(n) * size_t(s));
  diff_.resize(sizeof(double) * size_t(n));
  double* X = es_X_.as<double>();
  double* R = es_R_.as<double>();
  double* d = diff_.as<double>();
  orbital_differences(stream_, eps_.as<double>(), d, n_o, n_v, n);
  auto dot_host = [&](const double* a, const double* b) {
    double* res = red_result_.as<double>() + 3;
    dot(stream_, a, b, n, red_partial_.as<double>(), res);
    double v = 0.0;
    ETG_HIP(hipMemcpyAsync(&v, res, sizeof(double), hipMemcpyDeviceToHost, stream_));
    synchronize();
    return v;
  };
  if (start) {
    ETG_HIP(hipMemcpyAsync(X, start, sizeof(double) * n * s, hipMemcpyDefault, stream_));
    for (int64_t j = 0; j < s; ++j) {
      const double nx = std::sqrt(dot_host(X + j * n, X + j * n));
      if (!(nx > 0.0)) fail("diis_excited_states: zero start vector");
      scale(stream_, X + j * n, n, 1.0 / nx);
    }
  } else {
    std::vector<double> hd(static_cast<size_t>(n));
    ETG_HIP(hipMemcpyAsync(hd.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
    synchronize();
    std::vector<int64_t> idx(static_cast<size_t>(n));
    std::iota(idx.begin(), idx.end(), int64_t(0));
    std::partial_sort(idx.begin(), idx.begin() + s, idx.end(), [&](int64_t a, int64_t b) {
      return hd[size_t(a)] < hd[size_t(b)] || (hd[size_t(a)] == hd[size_t(b)] && a < b);
    });
    fill(stream_, X, n * s, 0.0);
    const double one = 1.0;
    for (int64_t j = 0; j < s; ++j)
      ETG_HIP(hipMemcpyAsync(X + j * n + idx[size_t(j)], &one, sizeof(double), hipMemcpyHostToDevice, stream_));
    synchronize();
  }
  es_diis_.resize(static_cast<size_t>(s));
  for (DiisSpace& sp : es_diis_) diis_space_initialize(stream_, sp, n, diis_dimension);

  std::vector<double> E(static_cast<size_t>(s), 0.0), E_prev(static_cast<size_t>(s), 0.0),
      norms(static_cast<size_t>(s), 0.0);
  std::vector<char> converged(static_cast<size_t>(s), 0);
  int64_t it = 0;
  while (std::find(converged.begin(), converged.end(), 0) != converged.end()) {
    if (it >= max_iterations) fail("diis_excited_states: Did not converge in the max number of iterations.");
    ++it;
    std::vector<int64_t> active;
    for (int64_t j = 0; j < s; ++j)
      if (!converged[size_t(j)]) active.push_back(j);
    const int64_t na = int64_t(active.size());
    for (int64_t q = 0; q < na; ++q)
      ETG_HIP(hipMemcpyAsync(R + q * n, X + active[size_t(q)] * n, sizeof(double) * n, hipMemcpyDeviceToDevice,
                             stream_));
    transform_batch(side, na, R, R, prec);       // columns are read before they are written
    project_cvs(R, na);                          // diis_cc_es_class.F90 projector
    for (int64_t q = 0; q < na; ++q) {
      const int64_t j = active[size_t(q)];
      double* Rj = R + q * n;
      double* Xj = X + j * n;
      E[size_t(j)] = dot_host(Xj, Rj);
      axpy(stream_, n, -E[size_t(j)], Xj, Rj);
      norms[size_t(j)] = std::sqrt(dot_host(Rj, Rj));
      const bool res_ok = norms[size_t(j)] < residual_threshold;
      const bool e_ok = energy_threshold <= 0.0 || std::fabs(E[size_t(j)] - E_prev[size_t(j)]) < energy_threshold;
      converged[size_t(j)] = (res_ok && (e_ok || it == 1)) ? 1 : 0;
      if (!converged[size_t(j)]) {
        shifted_precondition_kernel<<<grid_es(n), kBlock, 0, stream_>>>(Rj, d, n, E[size_t(j)]);
        ETG_LAUNCH("shifted_precondition");
        axpy(stream_, n, 1.0, Rj, Xj);
        diis_step(es_diis_[size_t(j)], Rj, Xj);
        const double nx = std::sqrt(dot_host(Xj, Xj));
        scale(stream_, Xj, n, 1.0 / nx);
      }
    }
    for (int64_t j = 0; j < s; ++j) {
      if (residual_norms) residual_norms[(it - 1) * s + j] = norms[size_t(j)];
      if (energy_history) energy_history[(it - 1) * s + j] = E[size_t(j)];
    }
    E_prev = E;
  }
  std::vector<int64_t> order(static_cast<size_t>(s));
  std::iota(order.begin(), order.end(), int64_t(0));
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return E[size_t(a)] < E[size_t(b)]; });
  for (int64_t q = 0; q < s; ++q) {
    energies[q] = E[size_t(order[size_t(q)])];
    if (vectors)
      ETG_HIP(hipMemcpyAsync(vectors + q * n, X + order[size_t(q)] * n, sizeof(double) * n, hipMemcpyDefault,
                             stream_));
  }
  synchronize();
  es_diis_.clear();
  return it;
}

}  // namespace etg
