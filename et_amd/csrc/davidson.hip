// Davidson eigensolver for the CC Jacobian with every vector and all reduced-
// space linear algebra on the device, in FP64.
//
// Follows eT's eigen Davidson solver (run_eigen_davidson_solver,
// src/solvers/eigen_davidson_solver_class.F90:122-318, and the tools in
// src/solver_tools/davidson/): trial vectors V and their transforms AV in HBM,
// reduced matrix G = V^T A V, its lowest eigenpairs, residuals
// R = (A X - w X) / ||X|| (construct_residual_eigen_davidson_tool,
// eigen_davidson_tool_class.F90:382-461), diagonal preconditioning by the
// orbital differences, Gram-Schmidt orthonormalisation against the space
// (davidson_tool_class.F90:354), restart on the current Ritz vectors when the
// reduced dimension would exceed its maximum (default max(100, 10 n_states),
// cc_es_amplitudes_solver_factory_class.F90:307), linear-dependence threshold
// min(1e-11, 0.1 threshold) (:315).
//
// The reduced eigenproblem (non-symmetric, dimension <= 128) is solved by one
// wavefront: Householder reduction to upper Hessenberg form, the implicitly
// double-shifted QR iteration for the eigenvalues (francis.hpp), then inverse
// iteration on G - w I with a partially pivoted LU for the wanted eigenvectors
// (for a complex pair a +- ib, the real 2k system of u + i v).  Complex pairs are carried as (u, v) column pairs with
// real and imaginary residuals, as eT's Davidson tool does.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "ccsd.hpp"
#include "francis.hpp"
#include "kernels.hpp"

namespace etg {
namespace {

constexpr int kMaxRed = 128;
constexpr int kLanes = 64;

__device__ inline double sgn(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// The reduced eigensolver's team: one wavefront.
struct WaveTeam {
  __device__ int rank() const { return int(threadIdx.x); }
  __device__ int size() const { return kLanes; }
  __device__ void sync() const { __syncthreads(); }
};

// H(i, j) of the k x k working matrix in LDS, column-major with the odd leading
// dimension kLdH: lanes that walk a row touch distinct banks.
constexpr int kLdH = kMaxRed + 1;
#define HH(i, j) H[(i) + (j) * kLdH]

// Inverse iteration on the nm x nm matrix M (ld ldm; LDS or global): LU with
// partial pivoting in place, then three solves M y = y starting from a fixed
// pattern, each normalised to unit 2-norm (largest component positive).  For
// real roots M = G - (w + d) I; for a complex pair a +- ib the real 2k system
// [[G - a I, b I], [-b I, G - a I]] acting on [u; v] (eigenvector u + i v).
__device__ void inverse_iteration(double* M, int ldm, int nm, double anorm, int* piv, double* y,
                                  int seed, int lane) {
#define MM(i, j) M[(i) + (j) * ldm]
  for (int c = 0; c < nm; ++c) {
    // pivot: the first row of largest magnitude, found across the lanes (each
    // lane scans every kLanes-th row, then the wave keeps the larger value and,
    // on a tie, the smaller row)
    int p = c;
    double best = fabs(MM(c, c));
    for (int r = c + 1 + lane; r < nm; r += kLanes) {
      const double v = fabs(MM(r, c));
      if (v > best) {
        best = v;
        p = r;
      }
    }
    for (int off = kLanes / 2; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, kLanes);
      const int op = __shfl_xor(p, off, kLanes);
      if (ob > best || (ob == best && op < p)) {
        best = ob;
        p = op;
      }
    }
    __syncthreads();
    if (lane == 0) {
      piv[c] = p;
      if (MM(p, c) == 0.0) MM(p, c) = 1e-300 + 1e-16 * anorm;
    }
    __syncthreads();
    if (p != c)
      for (int j = lane; j < nm; j += kLanes) {
        const double tmp = MM(c, j);
        MM(c, j) = MM(p, j);
        MM(p, j) = tmp;
      }
    __syncthreads();
    const double d = MM(c, c);
    for (int r = c + 1 + lane; r < nm; r += kLanes) {
      const double f = MM(r, c) / d;
      MM(r, c) = f;
      // four columns per round: their loads go out before the stores
      int j = c + 1;
      for (; j + 4 <= nm; j += 4) {
        const double p0 = MM(c, j), p1 = MM(c, j + 1), p2 = MM(c, j + 2), p3 = MM(c, j + 3);
        const double a0 = MM(r, j), a1 = MM(r, j + 1), a2 = MM(r, j + 2), a3 = MM(r, j + 3);
        MM(r, j) = a0 - f * p0;
        MM(r, j + 1) = a1 - f * p1;
        MM(r, j + 2) = a2 - f * p2;
        MM(r, j + 3) = a3 - f * p3;
      }
      for (; j < nm; ++j) MM(r, j) -= f * MM(c, j);
    }
    __syncthreads();
  }
  for (int i = lane; i < nm; i += kLanes) y[i] = 1.0 + 0.01 * ((i * 7 + seed * 3) % 11);
  __syncthreads();
  for (int iter = 0; iter < 3; ++iter) {
    if (lane == 0)
      for (int c = 0; c < nm; ++c) {
        const int p = piv[c];
        const double tmp = y[c];
        y[c] = y[p];
        y[p] = tmp;
      }
    __syncthreads();
    // triangular solves column by column, the updates of each column across the lanes
    for (int c = 0; c < nm; ++c) {                     // unit lower L
      const double yc = y[c];
      for (int r = c + 1 + lane; r < nm; r += kLanes) y[r] -= MM(r, c) * yc;
      __syncthreads();
    }
    for (int c = nm - 1; c >= 0; --c) {                // upper U
      const double yc = y[c] / MM(c, c);
      for (int r = lane; r < c; r += kLanes) y[r] -= MM(r, c) * yc;
      __syncthreads();
      if (lane == 0) y[c] = yc;
      __syncthreads();
    }
    if (lane == 0) {
      double nrm = 0.0, big = 0.0;
      int ib = 0;
      for (int i = 0; i < nm; ++i) {
        nrm += y[i] * y[i];
        if (fabs(y[i]) > big) { big = fabs(y[i]); ib = i; }
      }
      nrm = sqrt(nrm);
      const double sg = y[ib] < 0.0 ? -1.0 : 1.0;
      for (int i = 0; i < nm; ++i) y[i] = sg * y[i] / nrm;
    }
    __syncthreads();
  }
#undef MM
}

__global__ __launch_bounds__(kLanes) void reduced_eigen_kernel(const double* __restrict__ G, int k,
                                                               int ldg, int n_want,
                                                               double* __restrict__ wr_out,
                                                               double* __restrict__ wi_out,
                                                               double* __restrict__ X, int ldx,
                                                               double* __restrict__ work,
                                                               int* __restrict__ info) {
  __shared__ double H[kLdH * kMaxRed];
  __shared__ double wr[kMaxRed], wi[kMaxRed], ort[kMaxRed], y[2 * kMaxRed];
  __shared__ int order[kMaxRed], piv[2 * kMaxRed];
  __shared__ int status;
  const int lane = threadIdx.x;
  if (lane == 0) status = 0;
  for (int q = lane; q < k * k; q += kLanes) HH(q % k, q / k) = G[q % k + (q / k) * int64_t(ldg)];
  __syncthreads();

  // ---- Householder reduction to upper Hessenberg form (orthes) ----
  for (int m = 1; m < k - 1; ++m) {
    double scale = 0.0;
    for (int i = m; i < k; ++i) scale += fabs(HH(i, m - 1));
    if (scale == 0.0) continue;
    double h = 0.0;
    for (int i = m; i < k; ++i) {
      const double u = HH(i, m - 1) / scale;
      h += u * u;
    }
    const double g = -sgn(sqrt(h), HH(m, m - 1));
    h -= HH(m, m - 1) / scale * g;
    __syncthreads();
    for (int i = m + lane; i < k; i += kLanes) ort[i] = HH(i, m - 1) / scale - (i == m ? g : 0.0);
    __syncthreads();
    for (int j = m + lane; j < k; j += kLanes) {           // (I - u u^T / h) H
      double f = 0.0;
      for (int i = m; i < k; ++i) f += ort[i] * HH(i, j);
      f /= h;
      for (int i = m; i < k; ++i) HH(i, j) -= f * ort[i];
    }
    __syncthreads();
    for (int i = lane; i < k; i += kLanes) {               // H (I - u u^T / h)
      double f = 0.0;
      for (int j = m; j < k; ++j) f += ort[j] * HH(i, j);
      f /= h;
      for (int j = m; j < k; ++j) HH(i, j) -= f * ort[j];
    }
    __syncthreads();
    for (int i = m + lane; i < k; i += kLanes) HH(i, m - 1) = (i == m) ? scale * g : 0.0;
    __syncthreads();
  }

  // ---- eigenvalues of the Hessenberg matrix (implicit double-shift QR) ----
  double anorm = 0.0;
  {
    const int st = francis::hessenberg_eigenvalues(WaveTeam(), H, kLdH, k, wr, wi, &anorm);
    if (lane == 0 && st != 0) status = st;
  }
  __syncthreads();

  // ---- order by real part ----
  if (lane == 0) {
    for (int i = 0; i < k; ++i) order[i] = i;
    for (int i = 1; i < k; ++i) {
      const int o = order[i];
      int j = i - 1;
      while (j >= 0 && (wr[order[j]] > wr[o] || (wr[order[j]] == wr[o] && wi[order[j]] > wi[o]))) {
        order[j + 1] = order[j];
        --j;
      }
      order[j + 1] = o;
    }
    for (int i = 0; i < k; ++i) {
      wr_out[i] = wr[order[i]];
      wi_out[i] = wi[order[i]];
    }
  }
  __syncthreads();

  // ---- eigenvectors by inverse iteration ----
  // Real root: column w.  Complex pair a -+ ib (ordered a - ib first): columns
  // w and w + 1 hold u and v of the eigenvector u + i v of a + ib.
  for (int w = 0; w < n_want && status == 0; ++w) {
    const double lam = wr[order[w]], mu = wi[order[w]];
    if (mu == 0.0) {
      const double shift = lam + 1e-13 * fmax(anorm / k, fabs(lam));
      // eight elements per lane and round: their global loads are in flight together
      for (int q0 = 0; q0 < k * k; q0 += 8 * kLanes) {
        double g[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int q = q0 + u * kLanes + lane;
          g[u] = q < k * k ? G[q % k + (q / k) * int64_t(ldg)] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int q = q0 + u * kLanes + lane;
          if (q < k * k) {
            const int i = q % k, j = q / k;
            HH(i, j) = g[u] - (i == j ? shift : 0.0);
          }
        }
      }
      __syncthreads();
      inverse_iteration(H, kLdH, k, anorm, piv, y, w, lane);
      if (lane == 0) {
        // orthogonal to earlier vectors of (numerically) the same eigenvalue
        for (int u = 0; u < w; ++u) {
          if (wi[order[u]] != 0.0 || fabs(wr[order[u]] - lam) > 1e-8 * fmax(1.0, fabs(lam))) continue;
          double dp = 0.0;
          for (int i = 0; i < k; ++i) dp += X[i + u * int64_t(ldx)] * y[i];
          for (int i = 0; i < k; ++i) y[i] -= dp * X[i + u * int64_t(ldx)];
        }
        double nrm = 0.0;
        for (int i = 0; i < k; ++i) nrm += y[i] * y[i];
        nrm = sqrt(nrm);
        for (int i = 0; i < k; ++i) y[i] /= nrm;
      }
      __syncthreads();
      for (int i = lane; i < k; i += kLanes) X[i + w * int64_t(ldx)] = y[i];
      __syncthreads();
    } else if (mu < 0.0) {
      const double a = lam + 1e-13 * fmax(anorm / k, fabs(lam)), b = -mu;
      const int n2 = 2 * k;
      for (int q = lane; q < n2 * n2; q += kLanes) {
        const int i = q % n2, j = q / n2;
        const int bi = i / k, bj = j / k, ii = i % k, jj = j % k;
        double v;
        if (bi == bj) v = G[ii + jj * int64_t(ldg)] - (ii == jj ? a : 0.0);
        else v = (ii == jj) ? (bi == 0 ? b : -b) : 0.0;
        work[i + int64_t(j) * n2] = v;
      }
      __syncthreads();
      inverse_iteration(work, n2, n2, anorm, piv, y, w, lane);
      for (int i = lane; i < k; i += kLanes) {
        X[i + w * int64_t(ldx)] = y[i];
        if (w + 1 < n_want) X[i + (w + 1) * int64_t(ldx)] = y[k + i];
      }
      __syncthreads();
      ++w;                                  // the partner a + ib is done
    }
  }
  if (lane == 0) *info = status;
}
#undef HH

constexpr int kBlock = 256;
unsigned grid_for_d(int64_t n) { return unsigned(std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16)); }

// R_j = AY_j - w_j Y_j - c_j Y_{p_j}  (n x s, column-major, ld n).  Real root:
// c_j = 0.  Complex pair a -+ ib in columns (j, j+1) = (u, v): R_j = Au - a u + b v,
// R_{j+1} = Av - a v - b u (eigen_davidson_tool construct_re/im_residual).
// coef = [w (s) | c (s) | p (s)] on the device.
__global__ void ritz_residual_kernel(const double* __restrict__ AY, const double* __restrict__ Y,
                                     const double* __restrict__ coef, int64_t n, int s,
                                     double* __restrict__ R) {
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n * s;
       q += int64_t(gridDim.x) * blockDim.x) {
    const int j = int(q / n);
    const int64_t i = q - int64_t(j) * n;
    const int p = int(coef[2 * s + j]);
    R[q] = AY[q] - coef[j] * Y[q] - coef[s + j] * Y[i + int64_t(p) * n];
  }
}

// x_i = r_i / (w - d_i), guarded where the denominator vanishes (diagonal preconditioner)
// x = r / (w - d) elementwise (x may be r itself)
__global__ void precondition_kernel(const double* r, const double* __restrict__ d, double w, int64_t n,
                                    double* x) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double den = w - d[i];
    if (fabs(den) < 1e-8) den = den < 0.0 ? -1e-8 : 1e-8;
    x[i] = r[i] / den;
  }
}

}  // namespace

namespace {

// Doubles of the reduced eigensolver's device workspace: eigenvalues (2 x 128),
// the real 2k x 2k systems of complex pairs (4 x 128^2), the status flag.
constexpr size_t kEigenWork = 2 * kMaxRed + 4 * kMaxRed * kMaxRed + 1;

// The same workspace as a file-scope array, for the entry point that allocates
// nothing (one copy per GPU, 526 KB, held until the process ends).  Host code
// uses only the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[kEigenWork];

// Launches the solve with its workspace at `base` (kEigenWork doubles) and
// returns wr, wi to the host; throws on QR non-convergence.
void reduced_eigen_launch(hipStream_t stream, double* base, const double* G_dev, int64_t k, int64_t ldg,
                          int64_t n_want, double* wr, double* wi, double* X_dev, int64_t ldx) {
  if (k < 1 || k > kMaxRed) fail("reduced eigenproblem: dimension must be in [1, 128]");
  if (n_want < 0 || n_want > k) fail("reduced eigenproblem: invalid number of wanted roots");
  if (ldg < k || ldx < k) fail("reduced eigenproblem: leading dimension below the dimension");
  double* dwr = base;
  double* dwi = dwr + kMaxRed;
  double* work = dwi + kMaxRed;             // (2k)^2: complex-pair systems
  int* dinfo = reinterpret_cast<int*>(work + 4 * kMaxRed * kMaxRed);
  reduced_eigen_kernel<<<1, kLanes, 0, stream>>>(G_dev, int(k), int(ldg), int(n_want), dwr, dwi, X_dev,
                                                 int(ldx), work, dinfo);
  ETG_LAUNCH("reduced_eigen");
  int info = 0;
  ETG_HIP(hipMemcpyAsync(wr, dwr, sizeof(double) * k, hipMemcpyDeviceToHost, stream));
  ETG_HIP(hipMemcpyAsync(wi, dwi, sizeof(double) * k, hipMemcpyDeviceToHost, stream));
  ETG_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, stream));
  ETG_HIP(hipStreamSynchronize(stream));
  if (info == 1) fail("reduced eigenproblem: QR iteration did not converge");
}

// Device address of `scratch` on the current device, looked up once per device.
double* scratch_address() {
  static std::mutex mu;
  static std::vector<double*> addr;
  int dev = 0;
  ETG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (size_t(dev) >= addr.size()) addr.resize(size_t(dev) + 1, nullptr);
  if (addr[size_t(dev)] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      fail(std::string("reduced eigenproblem: no device address for the static workspace: ") +
           hipGetErrorString(e));
    addr[size_t(dev)] = static_cast<double*>(p);
  }
  return addr[size_t(dev)];
}

}  // namespace

void reduced_eigen_solve(hipStream_t stream, DeviceBuffer& workspace, const double* G_dev, int64_t k,
                         int64_t ldg, int64_t n_want, double* wr, double* wi, double* X_dev,
                         int64_t ldx) {
  if (k < 1 || k > kMaxRed) fail("reduced eigenproblem: dimension must be in [1, 128]");
  workspace.resize(sizeof(double) * kEigenWork);
  reduced_eigen_launch(stream, workspace.as<double>(), G_dev, k, ldg, n_want, wr, wi, X_dev, ldx);
}

void reduced_eigen_solve_noalloc(hipStream_t stream, const double* G_dev, int64_t k, int64_t ldg,
                                 int64_t n_want, double* wr, double* wi, double* X_dev, int64_t ldx) {
  if (k < 1 || k > kMaxRed)
    fail("reduced eigenproblem (no-allocation entry): dimension " + std::to_string(k) +
         " outside [1, 128], the size of its static workspace");
  reduced_eigen_launch(stream, scratch_address(), G_dev, k, ldg, n_want, wr, wi, X_dev, ldx);
}

void CCSD::reduced_eigen(const double* G_dev, int64_t k, int64_t ldg, int64_t n_want, double* wr,
                         double* wi, double* X_dev, int64_t ldx) {
  reduced_eigen_solve(stream_, red_eig_, G_dev, k, ldg, n_want, wr, wi, X_dev, ldx);
}

int64_t CCSD::davidson(int side, int64_t n_states, double threshold, int64_t max_iterations,
                       int64_t max_dim, Prec prec, const double* start, double* energies,
                       double* vectors, double* residual_norms) {
  require(tape_ != nullptr, "davidson: prepare_for_jacobian has not been called");
  require(have_eps_, "davidson: orbital energies required");
  require(side == 0 || side == 1, "davidson: side must be 0 (right) or 1 (left)");
  struct Clock {                          // wall time of the solve (davidson_timing)
    double& total;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    ~Clock() { total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  } clock{dav_timing_[1]};
  const int64_t n = n_t1 + n_t2;
  require(n_states >= 1 && n_states <= std::min<int64_t>(n - 1, 62), "davidson: n_states must be in [1, 62]");
  if (max_dim <= 0) max_dim = std::max<int64_t>(100, 10 * n_states);
  max_dim = std::min<int64_t>(std::min<int64_t>(max_dim, kMaxRed), n);
  require(max_dim >= 2 * (n_states + 1),
          "davidson: maximum reduced dimension below twice the number of states plus two");
  const double lindep = std::min(1e-11, 0.1 * threshold);
  const int64_t s = n_states, s1 = n_states + 1;    // a complex pair may straddle the last root

  // HBM: Y, AY, R (n x s1), G (max_dim^2), Xr and the projection coefficients
  // (max_dim x s1 each), [w | c | partner] (3 s1), Gram-Schmidt coefficients.
  // V and AV (n x max_dim): in HBM when both fit in half of the free memory
  // (the transformation passes size their vector groups from the rest),
  // otherwise AV, and then V as well, in pinned host memory that the kernels
  // read and write over PCIe (etg_set_davidson_space overrides the choice).
  dav_Y_.resize(sizeof(double) * size_t(n) * size_t(3 * s1));
  dav_small_.resize(sizeof(double) * size_t(max_dim * max_dim + 2 * max_dim * s1 + 3 * s1 + max_dim + 2 * s1 * s1));
  {
    const size_t space = sizeof(double) * size_t(n) * size_t(max_dim);
    size_t free_b = 0, total_b = 0;
    ETG_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t held = (dav_V_.on_host() ? 0 : dav_V_.bytes()) + (dav_AV_.on_host() ? 0 : dav_AV_.bytes());
    const double avail = 0.5 * double(free_b + held);
    bool v_host = false, av_host = false;
    if (dav_space_ == 2) {
      v_host = av_host = true;
    } else if (dav_space_ == 0) {
      av_host = 2.0 * double(space) > avail;
      v_host = double(space) > avail;
    }
    // release first so that the other buffer can take the memory
    if (dav_V_.on_host() != v_host) dav_V_.release();
    if (dav_AV_.on_host() != av_host) dav_AV_.release();
    // pinned host memory can be refused (the system's limit on locked pages): the
    // buffer then goes to HBM after all
    auto place = [&](DeviceBuffer& b, bool host) {
      if (host) {
        try {
          b.resize_host(space);
          return;
        } catch (const Error&) {
          (void)hipGetLastError();
          b.release();
        }
      }
      b.resize(space);
    };
    place(dav_AV_, av_host);
    place(dav_V_, v_host);
  }
  double* V = dav_V_.as<double>();
  double* AV = dav_AV_.as<double>();
  double* Y = dav_Y_.as<double>();
  double* AY = Y + n * s1;
  double* Rb = Y + 2 * n * s1;
  double* G = dav_small_.as<double>();
  double* Xr = G + max_dim * max_dim;
  double* cmat = Xr + max_dim * s1;          // projection coefficients (max_dim x s1)
  double* coef = cmat + max_dim * s1;        // [w | c | partner] (3 s1)
  double* cvec = coef + 3 * s1;              // Gram-Schmidt coefficients (max_dim)
  double* gram = cvec + max_dim;             // Gram matrix of a block (s1 x s1)
  double* tmat = gram + s1 * s1;             // inverse Cholesky factor of a block (s1 x s1)
  diff_.resize(sizeof(double) * n);
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
  // Cholesky QR of q columns at X: W = X T with T = R^-1, R^T R = X^T X (and
  // WA = XA T alongside when XA is given).  It gives the Q of Gram-Schmidt when
  // every column keeps more than 1e-4 of its norm against the ones before it
  // (Cholesky pivots above 1e-8 of the diagonal); false otherwise, with nothing
  // written.
  auto cholqr = [&](const double* X, const double* XA, int64_t q, double* W, double* WA) {
    run_gemm(make_gemm('T', 'N', q, q, n, 1.0, X, n, X, n, 0.0, gram, q), Prec::F64);
    std::vector<double> hg(size_t(q * q)), R(size_t(q * q), 0.0), T(size_t(q * q), 0.0);
    ETG_HIP(hipMemcpyAsync(hg.data(), gram, sizeof(double) * q * q, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (int64_t c = 0; c < q; ++c) {
      for (int64_t i = 0; i < c; ++i) {
        double v = hg[i + c * q];
        for (int64_t l = 0; l < i; ++l) v -= R[l + i * q] * R[l + c * q];
        R[i + c * q] = v / R[i + i * q];
      }
      double dcc = hg[c + c * q];
      for (int64_t l = 0; l < c; ++l) dcc -= R[l + c * q] * R[l + c * q];
      if (!(dcc > 1e-8 * hg[c + c * q])) return false;
      R[c + c * q] = std::sqrt(dcc);
    }
    for (int64_t c = 0; c < q; ++c) {
      T[c + c * q] = 1.0 / R[c + c * q];
      for (int64_t i = c - 1; i >= 0; --i) {
        double v = 0.0;
        for (int64_t l = i + 1; l <= c; ++l) v += R[i + l * q] * T[l + c * q];
        T[i + c * q] = -v / R[i + i * q];
      }
    }
    ETG_HIP(hipMemcpyAsync(tmat, T.data(), sizeof(double) * q * q, hipMemcpyHostToDevice, stream_));
    run_gemm(make_gemm('N', 'N', n, q, q, 1.0, X, n, tmat, q, 0.0, W, n), Prec::F64);
    if (XA) run_gemm(make_gemm('N', 'N', n, q, q, 1.0, XA, n, tmat, q, 0.0, WA, n), Prec::F64);
    synchronize();                        // T (host) is read by the copy above
    return true;
  };
  if (start) {
    // caller's start vectors (e.g. a reduced-precision solution), orthonormalised
    // in HBM (the residual block) and then placed in the space
    ETG_HIP(hipMemcpyAsync(Rb, start, sizeof(double) * n * s, hipMemcpyDefault, stream_));
    for (int64_t j = 0; j < s; ++j) {
      for (int pass = 0; pass < 2; ++pass)
        for (int64_t i = 0; i < j; ++i) axpy(stream_, n, -dot_host(Rb + i * n, Rb + j * n), Rb + i * n, Rb + j * n);
      const double nv = std::sqrt(dot_host(Rb + j * n, Rb + j * n));
      if (!(nv > 1e-10)) fail("davidson: start vectors are linearly dependent");
      scale(stream_, Rb + j * n, n, 1.0 / nv);
    }
    ETG_HIP(hipMemcpyAsync(V, Rb, sizeof(double) * n * s, hipMemcpyDefault, stream_));
  } else {
    // unit vectors on the lowest orbital differences
    std::vector<double> hd(n);
    ETG_HIP(hipMemcpyAsync(hd.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
    synchronize();
    std::vector<int64_t> idx(n);
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    std::partial_sort(idx.begin(), idx.begin() + s, idx.end(),
                      [&](int64_t a, int64_t b) { return hd[a] < hd[b] || (hd[a] == hd[b] && a < b); });
    fill(stream_, Rb, n * s, 0.0);
    const double one = 1.0;
    for (int64_t j = 0; j < s; ++j)
      ETG_HIP(hipMemcpyAsync(Rb + j * n + idx[j], &one, sizeof(double), hipMemcpyHostToDevice, stream_));
    ETG_HIP(hipMemcpyAsync(V, Rb, sizeof(double) * n * s, hipMemcpyDefault, stream_));
  }
  synchronize();
  int64_t k = 0, m = s;                   // space dimension, new vectors at V[:, k..k+m)
  std::vector<double> wr(max_dim), wi(max_dim), norms(s1), hcoef(3 * s1);
  double best = HUGE_VAL;                 // reduced precision: smallest largest-residual so far
  int64_t since_best = 0;
  const Prec start_prec = dav_start_prec_ < 0 ? prec : static_cast<Prec>(dav_start_prec_);
  // ETG_DAVIDSON_LOG=1: one line per iteration on stderr (progress of long solves)
  static const bool log_iterations = [] {
    const char* e = std::getenv("ETG_DAVIDSON_LOG");
    return e != nullptr && e[0] == '1';
  }();
  for (int64_t it = 1; it <= max_iterations; ++it) {
    // transforms of the new trial vectors (the start block in its own precision)
    const auto tt0 = std::chrono::steady_clock::now();
    transform_batch(side, m, V + k * n, AV + k * n, it == 1 ? start_prec : prec);
    dav_timing_[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - tt0).count();
    dav_timing_[2] += double(m);
    project_cvs(AV + k * n, m);              // eigen_davidson_solver_class.F90:187
    // G = V^T AV: new columns for all rows, new rows for the old columns
    const int64_t kn = k + m;
    run_gemm(make_gemm('T', 'N', kn, m, n, 1.0, V, n, AV + k * n, n, 0.0, G + k * max_dim, max_dim),
             Prec::F64);
    if (k > 0)
      run_gemm(make_gemm('T', 'N', m, k, n, 1.0, V + k * n, n, AV, n, 0.0, G + k, max_dim), Prec::F64);
    k = kn;
    const int64_t nw = std::min(k, s1);
    reduced_eigen(G, k, max_dim, nw, wr.data(), wi.data(), Xr, max_dim);
    // tracked columns: real roots, or complex pairs (u, v) in two columns
    int64_t se = 0;
    while (se < s && se < nw) {
      if (wi[se] < 0.0 && se + 1 < nw) {
        const double b = -wi[se];
        hcoef[se] = hcoef[se + 1] = wr[se];
        hcoef[s1 + se] = -b;
        hcoef[s1 + se + 1] = b;
        hcoef[2 * s1 + se] = double(se + 1);
        hcoef[2 * s1 + se + 1] = double(se);
        se += 2;
      } else {                            // real root (or an unpaired half: its real part)
        hcoef[se] = wr[se];
        hcoef[s1 + se] = 0.0;
        hcoef[2 * s1 + se] = double(se);
        se += 1;
      }
    }
    std::vector<double> packed(3 * se);
    for (int64_t j = 0; j < se; ++j) {
      packed[j] = hcoef[j];
      packed[se + j] = hcoef[s1 + j];
      packed[2 * se + j] = hcoef[2 * s1 + j];
    }
    // Ritz vectors and residuals
    run_gemm(make_gemm('N', 'N', n, se, k, 1.0, V, n, Xr, max_dim, 0.0, Y, n), Prec::F64);
    run_gemm(make_gemm('N', 'N', n, se, k, 1.0, AV, n, Xr, max_dim, 0.0, AY, n), Prec::F64);
    ETG_HIP(hipMemcpyAsync(coef, packed.data(), sizeof(double) * 3 * se, hipMemcpyHostToDevice, stream_));
    ritz_residual_kernel<<<grid_for_d(n * se), kBlock, 0, stream_>>>(AY, Y, coef, n, int(se), Rb);
    ETG_LAUNCH("ritz_residual");
    synchronize();                        // packed (host) is read by the copy above
    bool converged = true;
    std::vector<double> ynorm(se);
    for (int64_t j = 0; j < se; ++j) ynorm[j] = dot_host(Y + j * n, Y + j * n);
    for (int64_t j = 0; j < se; ++j) {
      const int64_t p = int64_t(hcoef[2 * s1 + j]);
      const double yy = std::sqrt(ynorm[j] + (p != j ? ynorm[p] : 0.0));
      norms[j] = std::sqrt(dot_host(Rb + j * n, Rb + j * n)) / yy;
      if (norms[j] >= threshold) converged = false;
    }
    for (int64_t j = 0; j < s; ++j) {
      energies[j] = hcoef[j];
      if (residual_norms) residual_norms[(it - 1) * s + j] = norms[j];
    }
    if (log_iterations) {
      double worst = 0.0;
      for (int64_t j = 0; j < s; ++j) worst = std::max(worst, norms[j]);
      std::fprintf(stderr, "davidson: iteration %lld, space %lld, %lld transformed, max residual %.3e\n",
                   (long long)it, (long long)k, (long long)m, worst);
      std::fflush(stderr);
    }
    // With reduced-precision transformations the residuals bottom out at the
    // transforms' rounding level: the solve also ends once the largest residual
    // has not improved by 10 % in 6 iterations (returned as -iterations).
    bool stalled = false;
    if (prec != Prec::F64) {
      double worst = 0.0;
      for (int64_t j = 0; j < s; ++j) worst = std::max(worst, norms[j]);
      if (worst < 0.9 * best) {
        best = worst;
        since_best = 0;
      } else {
        stalled = ++since_best >= 6;
      }
    }
    if (converged || stalled) {
      if (vectors) {
        for (int64_t j = 0; j < s; ++j) scale(stream_, Y + j * n, n, 1.0 / std::sqrt(ynorm[j]));
        ETG_HIP(hipMemcpyAsync(vectors, Y, sizeof(double) * n * s, hipMemcpyDefault, stream_));
        synchronize();
      }
      return converged ? it : -it;
    }
    // restart on the Ritz vectors when the space would overflow
    int64_t n_new = 0;
    for (int64_t j = 0; j < se; ++j) n_new += norms[j] >= threshold;
    if (k + n_new > max_dim) {
      // orthonormalise the Ritz vectors (and their transforms alike), then they
      // become the space: Cholesky QR twice (through V and AV, which are
      // replaced anyway), or one vector at a time in HBM
      if (se > 1 && cholqr(Y, AY, se, V, AV) && cholqr(V, AV, se, Y, AY)) {
        run_gemm(make_gemm('T', 'N', se, se, n, 1.0, Y, n, AY, n, 0.0, G, max_dim), Prec::F64);
        ETG_HIP(hipMemcpyAsync(V, Y, sizeof(double) * n * se, hipMemcpyDefault, stream_));
        ETG_HIP(hipMemcpyAsync(AV, AY, sizeof(double) * n * se, hipMemcpyDefault, stream_));
        k = se;
      }
    }
    if (k + n_new > max_dim) {
      for (int64_t j = 0; j < se; ++j) {
        for (int pass = 0; pass < 2; ++pass)
          for (int64_t i = 0; i < j; ++i) {
            const double c = dot_host(Y + i * n, Y + j * n);
            axpy(stream_, n, -c, Y + i * n, Y + j * n);
            axpy(stream_, n, -c, AY + i * n, AY + j * n);
          }
        const double nv = std::sqrt(dot_host(Y + j * n, Y + j * n));
        scale(stream_, Y + j * n, n, 1.0 / nv);
        scale(stream_, AY + j * n, n, 1.0 / nv);
      }
      run_gemm(make_gemm('T', 'N', se, se, n, 1.0, Y, n, AY, n, 0.0, G, max_dim), Prec::F64);
      ETG_HIP(hipMemcpyAsync(V, Y, sizeof(double) * n * se, hipMemcpyDefault, stream_));
      ETG_HIP(hipMemcpyAsync(AV, AY, sizeof(double) * n * se, hipMemcpyDefault, stream_));
      k = se;
    }
    // preconditioned residuals of the unconverged roots, normalised, in the
    // first columns of the residual block (column j moves to nc <= j, already read)
    int64_t nc = 0;
    for (int64_t j = 0; j < se; ++j) {
      if (norms[j] < threshold) continue;
      double* x = Rb + nc * n;
      precondition_kernel<<<grid_for_d(n), kBlock, 0, stream_>>>(Rb + j * n, d, hcoef[j], n, x);
      ETG_LAUNCH("precondition");
      const double nx = std::sqrt(dot_host(x, x));
      if (nx == 0.0) continue;
      scale(stream_, x, n, 1.0 / nx);
      ++nc;
    }
    // orthogonalised against the space as one block (two passes: one read of V
    // per product however many vectors are new) ...
    if (k > 0 && nc > 0)
      for (int pass = 0; pass < 2; ++pass) {
        run_gemm(make_gemm('T', 'N', k, nc, n, 1.0, V, n, Rb, n, 0.0, cmat, max_dim), Prec::F64);
        run_gemm(make_gemm('N', 'N', n, nc, k, -1.0, V, n, cmat, max_dim, 1.0, Rb, n), Prec::F64);
      }
    // ... then against each other: Cholesky QR twice when no vector comes near
    // the span of the ones before it (through Y, free until the next Ritz
    // vectors), otherwise one vector at a time (two passes), dropping linearly
    // dependent ones
    m = 0;
    if (nc > 1 && cholqr(Rb, nullptr, nc, Y, nullptr) && cholqr(Y, nullptr, nc, Rb, nullptr)) m = nc;
    for (int64_t c = m; c < nc; ++c) {
      double* x = Rb + c * n;
      if (m > 0)
        for (int pass = 0; pass < 2; ++pass) {
          run_gemm(make_gemm('T', 'N', m, 1, n, 1.0, Rb, n, x, n, 0.0, cvec, max_dim), Prec::F64);
          run_gemm(make_gemm('N', 'N', n, 1, m, -1.0, Rb, n, cvec, max_dim, 1.0, x, n), Prec::F64);
        }
      const double nx = std::sqrt(dot_host(x, x));
      if (nx < lindep) continue;
      scale(stream_, x, n, 1.0 / nx);
      if (c != m)
        ETG_HIP(hipMemcpyAsync(Rb + m * n, x, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
      ++m;
    }
    if (m == 0) fail("davidson: no new trial vectors (linear dependence) before convergence");
    ETG_HIP(hipMemcpyAsync(V + k * n, Rb, sizeof(double) * n * m, hipMemcpyDefault, stream_));
  }
  fail("davidson: did not converge in the maximum number of iterations");
}

}  // namespace etg
