// eT's linear Davidson solver for one linear equation A X = F on the device
// (run_linear_davidson_solver, src/solvers/linear_davidson_solver_class.F90:122-270,
// with the linear Davidson tool, src/solver_tools/davidson/linear_davidson_tool_class.F90,
// and the trial handling of davidson_tool_class.F90:312-351, 863-953).
//
// Start: the caller's vector, or the preconditioner guess X0 = F / (eps - 0).
// Iteration: when the space would exceed max_dim, it restarts on the current
// solution; the new trial is orthonormalised against the space (twice; dropped
// below the linear-dependence threshold); its transform A c (right) or A^T c
// (left) is added; the reduced equation (V^T A V) y = V^T F is solved; the
// residual R = A V y - F is tested against the threshold, and otherwise
// -R / (eps - 0), normalised, is the next trial.  lindep: the linear-dependence
// threshold on a normalised trial's projected norm (<= 0: min(1e-11, 0.1 threshold)).  V, A V and every vector stay
// in HBM; the reduced matrices (max_dim <= 128) go to the host for the solve.
#include <algorithm>
#include <cmath>

#include "ccsd.hpp"
#include "kernels.hpp"

namespace etg {
namespace {

constexpr int kBlock = 256;
constexpr int64_t kMaxLinearDim = 128;    // reduced dimension bound (host solve)

unsigned grid_ld(int64_t n) { return unsigned(std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16)); }

// x = prefactor * x / d  (d = eps differences, shift 0; entries with |d| < 1e-8 keep x)
__global__ void precondition_kernel(double* __restrict__ x, const double* __restrict__ d, int64_t n,
                                    double prefactor) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x)
    x[k] = fabs(d[k]) < 1e-8 ? prefactor * x[k] : prefactor * x[k] / d[k];
}

// Gaussian elimination with partial pivoting (dgesv's method) on the host, k <= 128.
bool solve_dense(std::vector<double> A, std::vector<double>& b, int64_t k) {
  for (int64_t c = 0; c < k; ++c) {
    int64_t p = c;
    for (int64_t r = c + 1; r < k; ++r)
      if (std::fabs(A[size_t(r + c * k)]) > std::fabs(A[size_t(p + c * k)])) p = r;
    if (A[size_t(p + c * k)] == 0.0) return false;
    if (p != c) {
      for (int64_t j = 0; j < k; ++j) std::swap(A[size_t(c + j * k)], A[size_t(p + j * k)]);
      std::swap(b[size_t(c)], b[size_t(p)]);
    }
    for (int64_t r = c + 1; r < k; ++r) {
      const double f = A[size_t(r + c * k)] / A[size_t(c + c * k)];
      for (int64_t j = c; j < k; ++j) A[size_t(r + j * k)] -= f * A[size_t(c + j * k)];
      b[size_t(r)] -= f * b[size_t(c)];
    }
  }
  for (int64_t r = k - 1; r >= 0; --r) {
    double s = b[size_t(r)];
    for (int64_t j = r + 1; j < k; ++j) s -= A[size_t(r + j * k)] * b[size_t(j)];
    b[size_t(r)] = s / A[size_t(r + r * k)];
  }
  return true;
}

}  // namespace

int64_t CCSD::linear_davidson(int side, const double* F, const double* start, double residual_threshold,
                              int64_t max_iterations, int64_t max_dim, Prec prec, double* X_out,
                              double* residual_norms, double lindep) {
  require(tape_ != nullptr, "linear_davidson: prepare_for_jacobian has not been called");
  require(have_eps_, "linear_davidson: orbital energies required");
  require(side == 0 || side == 1, "linear_davidson: side must be 0 (right) or 1 (left)");
  require(F != nullptr && X_out != nullptr, "linear_davidson: null vector");
  require(max_iterations >= 1, "linear_davidson: max_iterations must be positive");
  const int64_t n = n_t1 + n_t2;
  if (max_dim <= 0) max_dim = 50;
  max_dim = std::min<int64_t>(std::min<int64_t>(max_dim, kMaxLinearDim), n);
  require(max_dim >= 2, "linear_davidson: maximum reduced dimension below 2");
  if (!(lindep > 0.0)) lindep = std::min(1e-11, 0.1 * residual_threshold);

  // HBM: V, AV (n x max_dim), the right-hand side, the trial and residual vectors
  dav_V_.resize(sizeof(double) * size_t(n) * size_t(max_dim));
  dav_AV_.resize(sizeof(double) * size_t(n) * size_t(max_dim));
  dav_Y_.resize(sizeof(double) * size_t(n) * 3);
  dav_small_.resize(sizeof(double) * size_t(2 * max_dim + 2));
  diff_.resize(sizeof(double) * size_t(n));
  double* V = dav_V_.as<double>();
  double* AV = dav_AV_.as<double>();
  double* Fd = dav_Y_.as<double>();
  double* c = Fd + n;                        // the new trial
  double* R = Fd + 2 * n;                    // solution, then residual
  double* col = dav_small_.as<double>();     // V^T x (max_dim), x^T AV (max_dim)
  double* d = diff_.as<double>();
  orbital_differences(stream_, eps_.as<double>(), d, n_o, n_v, n);
  ETG_HIP(hipMemcpyAsync(Fd, F, sizeof(double) * n, hipMemcpyDefault, stream_));

  auto dot_host = [&](const double* a, const double* b) {
    double* res = red_result_.as<double>() + 3;
    dot(stream_, a, b, n, red_partial_.as<double>(), res);
    double v = 0.0;
    ETG_HIP(hipMemcpyAsync(&v, res, sizeof(double), hipMemcpyDeviceToHost, stream_));
    synchronize();
    return v;
  };
  // y (host, k) -> out = M y for M = V or AV
  auto combine = [&](const double* M, const std::vector<double>& y, int64_t k, double* out) {
    ETG_HIP(hipMemcpyAsync(col, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, stream_));
    run_gemm(make_gemm('N', 'N', n, 1, k, 1.0, M, n, col, k, 0.0, out, n), Prec::F64);
  };

  if (start) {
    ETG_HIP(hipMemcpyAsync(c, start, sizeof(double) * n, hipMemcpyDefault, stream_));
  } else {
    ETG_HIP(hipMemcpyAsync(c, Fd, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
    precondition_kernel<<<grid_ld(n), kBlock, 0, stream_>>>(c, d, n, 1.0);
    ETG_LAUNCH("precondition");
  }

  std::vector<double> G(size_t(max_dim * max_dim), 0.0), Fr(size_t(max_dim), 0.0), y;
  int64_t k = 0;                             // space dimension
  for (int64_t it = 1; it <= max_iterations; ++it) {
    if (k + 1 > max_dim) {                   // restart on the current solution
      combine(V, y, k, c);
      k = 0;
    }
    // orthonormalise the new trial against the space (two passes)
    for (int pass = 0; pass < 2 && k > 0; ++pass) {
      run_gemm(make_gemm('T', 'N', k, 1, n, 1.0, V, n, c, n, 0.0, col, k), Prec::F64);
      run_gemm(make_gemm('N', 'N', n, 1, k, -1.0, V, n, col, k, 1.0, c, n), Prec::F64);
    }
    const double nc = std::sqrt(dot_host(c, c));
    if (!(nc > lindep)) fail("linear_davidson: the new trial vector is linearly dependent on the space");
    scale(stream_, c, n, 1.0 / nc);
    ETG_HIP(hipMemcpyAsync(V + k * n, c, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
    transform_batch(side, 1, V + k * n, AV + k * n, prec);
    // new row and column of G = V^T A V, new element of V^T F
    run_gemm(make_gemm('T', 'N', k + 1, 1, n, 1.0, V, n, AV + k * n, n, 0.0, col, k + 1), Prec::F64);
    run_gemm(make_gemm('T', 'N', k + 1, 1, n, 1.0, AV, n, V + k * n, n, 0.0, col + max_dim, k + 1), Prec::F64);
    std::vector<double> gc(size_t(k + 1)), gr(size_t(k + 1));
    ETG_HIP(hipMemcpyAsync(gc.data(), col, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, stream_));
    ETG_HIP(hipMemcpyAsync(gr.data(), col + max_dim, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, stream_));
    synchronize();
    const double fk = dot_host(V + k * n, Fd);
    // the reduced matrix of the new dimension (k + 1), column-major
    std::vector<double> Gn(size_t((k + 1) * (k + 1)));
    for (int64_t j = 0; j < k; ++j)
      for (int64_t i = 0; i < k; ++i) Gn[size_t(i + j * (k + 1))] = G[size_t(i + j * max_dim)];
    for (int64_t i = 0; i <= k; ++i) {
      Gn[size_t(i + k * (k + 1))] = gc[size_t(i)];     // V_i^T A V_k
      Gn[size_t(k + i * (k + 1))] = gr[size_t(i)];     // V_k^T A V_i
    }
    Fr[size_t(k)] = fk;
    ++k;
    for (int64_t j = 0; j < k; ++j)
      for (int64_t i = 0; i < k; ++i) G[size_t(i + j * max_dim)] = Gn[size_t(i + j * k)];
    y.assign(Fr.begin(), Fr.begin() + k);
    if (!solve_dense(Gn, y, k)) fail("linear_davidson: could not solve the reduced linear equation");
    // residual R = AV y - F
    combine(AV, y, k, R);
    axpy(stream_, n, -1.0, Fd, R);
    const double rn = std::sqrt(dot_host(R, R));
    if (residual_norms) residual_norms[it - 1] = rn;
    if (rn <= residual_threshold) {
      combine(V, y, k, R);
      ETG_HIP(hipMemcpyAsync(X_out, R, sizeof(double) * n, hipMemcpyDefault, stream_));
      synchronize();
      return it;
    }
    // next trial: -R / (eps - 0), normalised before the orthonormalisation, so that
    // the linear-dependence test compares the fraction left after projection
    ETG_HIP(hipMemcpyAsync(c, R, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
    precondition_kernel<<<grid_ld(n), kBlock, 0, stream_>>>(c, d, n, -1.0);
    ETG_LAUNCH("precondition");
    scale(stream_, c, n, 1.0 / std::sqrt(dot_host(c, c)));
  }
  fail("linear_davidson: was not able to converge the equations in the given number of maximum iterations.");
}

}  // namespace etg
