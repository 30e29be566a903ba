// Ground-state amplitude update and DIIS extrapolation on the device, in FP64.
//
// eT's solver loop (run_diis_cc_gs, diis_cc_gs_class.F90:253-293) calls
//   t_updater%update(wf, amplitudes, omega)   -> quasi-Newton step
//   diis%update(omega, amplitudes)            -> DIIS extrapolation
// Both work here on the handle's HBM-resident vectors: the update vector, the
// DIIS records (errors and parameters), the DIIS matrix G and its weights
// never leave the device; only the solver's status flag is read back.
#include <algorithm>
#include <cmath>

#include "ccsd.hpp"
#include "kernels.hpp"

namespace etg {
namespace {

constexpr int kBlock = 256;

__device__ inline int64_t tri_root_u(int64_t k) {
  int64_t x = int64_t((sqrt(8.0 * double(k) + 1.0) - 1.0) * 0.5);
  while (x * (x + 1) / 2 > k) --x;
  while ((x + 1) * (x + 2) / 2 <= k) ++x;
  return x;
}

// e = -omega / (eps_a - eps_i [+ eps_b - eps_j]), times diag_factor on the packed
// doubles diagonal; t += e.  quasi_newton_updater_class.F90:81-118 (denominators
// from get_orbital_differences, doubles_class.F90:421-470); diag_factor = 2 is
// scale_amplitudes_ccsd (ccsd_class.F90:648-671) applied to the update (ground
// state), 1 the multiplier updater without scaling
// (cc_multipliers_solver_factory_class.F90:249-251).
__global__ void quasi_newton_kernel(const double* __restrict__ eps, const double* __restrict__ omega,
                                    double* __restrict__ t, double* __restrict__ e, int64_t n_o,
                                    int64_t n_v, int64_t n, double diag_factor) {
  const int64_t N = n_o * n_v;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    double d, f = 1.0;
    if (k < N) {
      d = eps[n_o + k % n_v] - eps[k / n_v];
    } else {
      const int64_t q = k - N;
      const int64_t X = tri_root_u(q), Y = q - X * (X + 1) / 2;
      d = eps[n_o + X % n_v] - eps[X / n_v] + eps[n_o + Y % n_v] - eps[Y / n_v];
      if (X == Y) f = diag_factor;
    }
    const double u = -f * omega[k] / d;
    e[k] = u;
    t[k] += u;
  }
}

// G(s, i) = G(i, s) for the row s just computed.
__global__ void mirror_row_kernel(double* G, int ld, int s, int dim) {
  const int i = threadIdx.x;
  if (i < dim) G[i + s * ld] = G[s + i * ld];
}

// Solves the padded DIIS equation (diis_tool_class.F90:592-640)
//   [ G  -1 ] [w]   [ 0 ]
//   [ -1  0 ] [l] = [-1 ]
// by Gaussian elimination with partial pivoting (the elimination dgesv does),
// one thread, FP64.  info = 0 on success, k > 0 if pivot k is exactly zero.
__global__ void diis_solve_kernel(const double* __restrict__ G, int ld, int dim,
                                  double* __restrict__ w, int* __restrict__ info) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  constexpr int kMax = CCSD::kMaxDiis + 1;
  double A[kMax][kMax + 1];
  const int n = dim + 1;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j)
      A[i][j] = (i < dim && j < dim) ? G[i + j * ld] : ((i == dim && j == dim) ? 0.0 : -1.0);
    A[i][n] = (i == dim) ? -1.0 : 0.0;
  }
  int status = 0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
    if (A[p][c] == 0.0) {
      status = c + 1;
      break;
    }
    if (p != c)
      for (int j = c; j <= n; ++j) {
        const double tmp = A[c][j];
        A[c][j] = A[p][j];
        A[p][j] = tmp;
      }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int j = c; j <= n; ++j) A[r][j] -= f * A[c][j];
    }
  }
  if (status == 0) {
    for (int i = n - 1; i >= 0; --i) {
      double s = A[i][n];
      for (int j = i + 1; j < n; ++j) s -= A[i][j] * A[j][n];
      A[i][n] = s / A[i][i];
    }
    for (int i = 0; i < dim; ++i) w[i] = A[i][n];
  }
  *info = status;
}

// x = sum_s w_s X_s over the dim records (diis_tool construct_extrapolated_vector).
// Leaves x unchanged when the solve failed.
__global__ void extrapolate_kernel(const double* __restrict__ X, int64_t n, int dim,
                                   const double* __restrict__ w, const int* __restrict__ info,
                                   double* __restrict__ x) {
  if (*info != 0) return;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < dim; ++r) s += w[r] * X[k + int64_t(r) * n];
    x[k] = s;
  }
}

// CROP (diis_tool_class.F90:763-805): the newest error record E_slot becomes the
// extrapolated error <e> = sum_r w_r E_r.  Each element is read and written by one
// thread, so the record is updated in place.
__global__ void crop_error_kernel(double* E, int64_t n, int dim, int slot, const double* __restrict__ w,
                                  const int* __restrict__ info) {
  if (*info != 0) return;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < dim; ++r) s += w[r] * E[k + int64_t(r) * n];
    E[k + int64_t(slot) * n] = s;
  }
}

unsigned grid_for_n(int64_t n) {
  return unsigned(std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16));
}

}  // namespace

void quasi_newton_step(hipStream_t s, const double* eps, const double* residual, double* x, double* e,
                       int64_t n_o, int64_t n_v, int64_t n, double diag_factor) {
  quasi_newton_kernel<<<grid_for_n(n), kBlock, 0, s>>>(eps, residual, x, e, n_o, n_v, n, diag_factor);
  ETG_LAUNCH("quasi_newton");
}

void diis_space_initialize(hipStream_t s, DiisSpace& d, int64_t n, int64_t dimension, bool crop) {
  if (dimension < 1 || dimension > CCSD::kMaxDiis) fail("diis_initialize: dimension must be in [1, 32]");
  d.dim = int(dimension);
  d.crop = crop;
  d.count = 0;
  d.n = n;
  d.e.resize(sizeof(double) * size_t(n) * size_t(dimension));
  d.x.resize(sizeof(double) * size_t(n) * size_t(dimension));
  d.G.resize(sizeof(double) * CCSD::kMaxDiis * CCSD::kMaxDiis);
  d.w.resize(sizeof(double) * CCSD::kMaxDiis);
  d.info.resize(sizeof(int));
  fill(s, d.G.as<double>(), CCSD::kMaxDiis * CCSD::kMaxDiis, 0.0);
}

void CCSD::diis_initialize(int64_t dimension, bool crop) {
  diis_space_initialize(stream_, diis_, n_t1 + n_t2, dimension, crop);
}

// finalize_diis_tool (diis_tool_class.F90:280-299): the records are released
void CCSD::diis_finalize() {
  synchronize();
  diis_.e.release();
  diis_.x.release();
  diis_.dim = 0;
  diis_.count = 0;
}

void CCSD::quasi_newton_update() {
  require(have_eps_ && have_t_, "quasi_newton_update: orbital energies and amplitudes required");
  require(have_omega_, "quasi_newton_update: the residual has not been constructed");
  const int64_t n = n_t1 + n_t2;
  update_.resize(sizeof(double) * n);
  quasi_newton_kernel<<<grid_for_n(n), kBlock, 0, stream_>>>(eps_.as<double>(), omega_.as<double>(),
                                                             t_.as<double>(), update_.as<double>(),
                                                             n_o, n_v, n, 2.0);
  ETG_LAUNCH("quasi_newton");
  have_update_ = true;
  have_omega_ = false;        // the residual belongs to the amplitudes before the step
  reset_record();
}

void CCSD::diis_update() {
  require(have_update_, "diis_update: no update vector (call quasi_newton_update first)");
  diis_step(diis_, update_.as<double>(), t_.as<double>());
  have_update_ = false;
  reset_record();
}

void CCSD::diis_step(DiisSpace& d, const double* e, double* x) {
  require(d.dim > 0, "diis_update: DIIS has not been initialised");
  const int64_t n = d.n;
  // record slot: the oldest record is replaced once the space is full
  // (diis_tool_class.F90 save_errors_and_parameters / modify_mappings_to_remove_oldest_record)
  const int slot = int(d.count % d.dim);
  const int dim = int(std::min<int64_t>(d.count + 1, d.dim));
  double* E = d.e.as<double>();
  double* X = d.x.as<double>();
  double* G = d.G.as<double>();
  ETG_HIP(hipMemcpyAsync(E + int64_t(slot) * n, e, sizeof(double) * n, hipMemcpyDeviceToDevice,
                         stream_));
  ETG_HIP(hipMemcpyAsync(X + int64_t(slot) * n, x, sizeof(double) * n, hipMemcpyDeviceToDevice,
                         stream_));
  // new row of G (construct_new_G_elements, diis_tool_class.F90:503-546)
  auto new_row = [&] {
    for (int i = 0; i < dim; ++i)
      dot(stream_, E + int64_t(slot) * n, E + int64_t(i) * n, n, red_partial_.as<double>(),
          G + slot + int64_t(i) * kMaxDiis);
    mirror_row_kernel<<<1, kMaxDiis, 0, stream_>>>(G, kMaxDiis, slot, dim);
    ETG_LAUNCH("diis_mirror_row");
  };
  new_row();
  diis_solve_kernel<<<1, 1, 0, stream_>>>(G, kMaxDiis, dim, d.w.as<double>(), d.info.as<int>());
  ETG_LAUNCH("diis_solve");
  extrapolate_kernel<<<grid_for_n(n), kBlock, 0, stream_>>>(X, n, dim, d.w.as<double>(),
                                                            d.info.as<int>(), x);
  ETG_LAUNCH("diis_extrapolate");
  if (d.crop) {
    // the newest records become <x> and <e>, and G's row follows <e>
    ETG_HIP(hipMemcpyAsync(X + int64_t(slot) * n, x, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
    crop_error_kernel<<<grid_for_n(n), kBlock, 0, stream_>>>(E, n, dim, slot, d.w.as<double>(), d.info.as<int>());
    ETG_LAUNCH("diis_crop_error");
    new_row();
  }
  int info = 0;
  ETG_HIP(hipMemcpyAsync(&info, d.info.as<int>(), sizeof(int), hipMemcpyDeviceToHost, stream_));
  synchronize();
  if (info != 0) fail("diis_update: could not solve the DIIS matrix equation (zero pivot)");
  ++d.count;
}

}  // namespace etg
