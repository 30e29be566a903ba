// Bandwidth-bound kernels of the CCSD residual (see kernels.hpp).
#include "kernels.hpp"

#include <cmath>
#include <cstring>

namespace etg {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int64_t per_block = kBlock, int64_t cap = 1 << 20) {
  int64_t g = ceil_div(n, per_block);
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return unsigned(g);
}

// Largest X with X (X + 1) / 2 <= k.
__device__ inline int64_t tri_root(int64_t k) {
  int64_t x = int64_t((sqrt(8.0 * double(k) + 1.0) - 1.0) * 0.5);
  while (x * (x + 1) / 2 > k) --x;
  while ((x + 1) * (x + 2) / 2 <= k) ++x;
  return x;
}

__device__ inline int64_t pidx(int64_t x, int64_t y) {
  const int64_t hi = x > y ? x : y, lo = x > y ? y : x;
  return hi * (hi + 1) / 2 + lo;
}

__global__ void permute4_kernel(const double* __restrict__ x, double* __restrict__ y, int64_t e0,
                                int64_t e1, int64_t e2, int64_t e3, int64_t s0, int64_t s1,
                                int64_t s2, int64_t s3, double alpha, double beta) {
  const int64_t n = e0 * e1 * e2 * e3;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = idx;
    const int64_t o0 = r % e0; r /= e0;
    const int64_t o1 = r % e1; r /= e1;
    const int64_t o2 = r % e2;
    const int64_t o3 = r / e2;
    const double v = alpha * x[o0 * s0 + o1 * s1 + o2 * s2 + o3 * s3];
    y[idx] = (beta == 0.0) ? v : v + beta * y[idx];
  }
}

__global__ void squareup_kernel(const double* __restrict__ packed, double* __restrict__ full,
                                int64_t N) {
  const int64_t n = N * N;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    full[idx] = packed[pidx(idx % N, idx / N)];
  }
}

// packed element k = (X, Y), X >= Y, reads full(Y, X) (column X is contiguous in Y)
__global__ void packin_kernel(const double* __restrict__ full, double* __restrict__ packed,
                              int64_t N, double alpha, double beta, double diag_scale) {
  const int64_t n = N * (N + 1) / 2;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t X = tri_root(k), Y = k - X * (X + 1) / 2;
    double v = alpha * full[Y + N * X];
    if (X == Y) v *= diag_scale;
    packed[k] = (beta == 0.0) ? v : v + beta * packed[k];
  }
}

// y = x + x^T through 32 x 32 LDS tiles
__global__ void symmetric_sum_kernel(const double* __restrict__ x, double* __restrict__ y,
                                     int64_t N) {
  __shared__ double tile[32][33];
  const int64_t bx = blockIdx.x, by = blockIdx.y;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  // load transposed block (by, bx) of x: rows by*32.., cols bx*32..
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = by * 32 + tx, c = bx * 32 + k;
    tile[k][tx] = (r < N && c < N) ? x[r + N * c] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = bx * 32 + tx, c = by * 32 + k;
    if (r < N && c < N) y[r + N * c] = x[r + N * c] + tile[tx][k];
  }
}

__global__ void symmetrize_and_pack_kernel(const double* __restrict__ x, double* __restrict__ y,
                                           int64_t n_v, int64_t n_o, double alpha, double f1,
                                           double f2) {
  const int64_t N = n_v * n_o;
  const int64_t n = N * (N + 1) / 2;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t X = tri_root(k), Y = k - X * (X + 1) / 2;  // (r,s) = X, (p,q) = Y
    if (X == Y) {
      y[k] += alpha * x[Y + N * Y];
      continue;
    }
    const int64_t p = Y % n_v, q = Y / n_v, r = X % n_v, s = X / n_v;
    double v = f1 * (x[Y + N * X] + x[X + N * Y]);
    if (f2 != 0.0) v += f2 * (x[(p + n_v * s) + N * (r + n_v * q)] + x[(r + n_v * q) + N * (p + n_v * s)]);
    y[k] += alpha * v;
  }
}

__global__ void ladder_t_pm_kernel(const double* __restrict__ t2, double* __restrict__ tp,
                                   double* __restrict__ tm, int64_t n_o, int64_t n_v) {
  const int64_t op = n_o * (n_o + 1) / 2, vp = n_v * (n_v + 1) / 2;
  const int64_t om = n_o * (n_o - 1) / 2, vm = n_v * (n_v - 1) / 2;
  const int64_t n = op * vp;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t rij = idx % op, rcd = idx / op;
    const int64_t j = tri_root(rij), i = rij - j * (j + 1) / 2;
    const int64_t d = tri_root(rcd), c = rcd - d * (d + 1) / 2;
    const double x1 = t2[pidx(c + n_v * i, d + n_v * j)];
    const double x2 = t2[pidx(c + n_v * j, d + n_v * i)];
    tp[idx] = (c == d ? 0.5 : 1.0) * (x1 + x2);
    if (i < j && c < d) {
      const int64_t mij = j * (j - 1) / 2 + i, mcd = d * (d - 1) / 2 + c;
      tm[mij + om * mcd] = x1 - x2;
    }
  }
  (void)vm;
}

__global__ void ladder_g_pm_kernel(const double* __restrict__ G, int64_t n_v, int64_t lenA,
                                   int64_t lenB, const int32_t* __restrict__ pa,
                                   const int32_t* __restrict__ pb, int64_t npairs,
                                   double* __restrict__ gp, double* __restrict__ gm) {
  const int64_t vp = n_v * (n_v + 1) / 2, vm = n_v * (n_v - 1) / 2;
  const int64_t ldG = n_v * lenA;
  const int64_t n = vp * npairs;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t rcd = idx % vp, k = idx / vp;
    const int64_t d = tri_root(rcd), c = rcd - d * (d + 1) / 2;
    const int64_t a = pa[k], b = pb[k];
    // G(c, a, d, b) and G(d, a, c, b)
    const double x1 = G[(c + n_v * a) + ldG * (d + n_v * b)];
    const double x2 = G[(d + n_v * a) + ldG * (c + n_v * b)];
    gp[idx] = x1 + x2;
    if (c < d) gm[(d * (d - 1) / 2 + c) + vm * k] = x1 - x2;
  }
}

__global__ void ladder_gather_kernel(const double* __restrict__ wp, const double* __restrict__ wm,
                                     const int32_t* __restrict__ col, double* __restrict__ om,
                                     int64_t n_o, int64_t n_v) {
  const int64_t N = n_o * n_v;
  const int64_t n = N * (N + 1) / 2;
  const int64_t op = n_o * (n_o + 1) / 2, omn = n_o * (n_o - 1) / 2;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t X = tri_root(k), Y = k - X * (X + 1) / 2;  // X = ai, Y = bj
    const int64_t a = X % n_v, i = X / n_v, b = Y % n_v, j = Y / n_v;
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    const int64_t cl = col[a + n_v * b];
    double w = wp[(hi * (hi + 1) / 2 + lo) + op * cl];
    if (i != j && a != b) {
      const double sgn = ((i < j) == (a > b)) ? 1.0 : -1.0;
      w += sgn * wm[(hi * (hi - 1) / 2 + lo) + omn * cl];
    }
    om[k] += (X == Y) ? 0.5 * w : w;
  }
}

__device__ inline double block_sum(double v) {
  __shared__ double red[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < kBlock / 64; ++q) s += red[q];
  return s;  // valid in thread 0
}

__global__ void energy_kernel(const double* __restrict__ t1, const double* __restrict__ t2,
                              const double* __restrict__ g, int64_t n_o, int64_t n_v,
                              double* __restrict__ partial) {
  const int64_t N = n_o * n_v;
  const int64_t n = N * N;
  double acc = 0.0;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t X = idx % N, Y = idx / N;
    const int64_t a = X % n_v, i = X / n_v, b = Y % n_v, j = Y / n_v;
    const double tau = t2[pidx(X, Y)] + t1[X] * t1[Y];
    const double g_iajb = g[i + n_o * a + N * (j + n_o * b)];
    const double g_ibja = g[i + n_o * b + N * (j + n_o * a)];
    acc += tau * (2.0 * g_iajb - g_ibja);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void sumsq_kernel(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                             double* __restrict__ partial) {
  double acc = 0.0;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x)
    acc += x[idx] * (y ? y[idx] : x[idx]);
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void finish_sum_kernel(const double* __restrict__ partial, int n, double* result) {
  double acc = 0.0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) acc += partial[q];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) *result = s;
}

__global__ void mp2_guess_kernel(const double* __restrict__ g, const double* __restrict__ eps,
                                 double* __restrict__ t2, int64_t n_o, int64_t n_v) {
  const int64_t N = n_o * n_v;
  const int64_t n = N * (N + 1) / 2;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t X = tri_root(k), Y = k - X * (X + 1) / 2;
    const int64_t a = X % n_v, i = X / n_v, b = Y % n_v, j = Y / n_v;
    const double d = eps[i] - eps[n_o + a] + eps[j] - eps[n_o + b];
    t2[k] = g[X + N * Y] / d;
  }
}

__global__ void orbital_differences_kernel(const double* __restrict__ eps, double* __restrict__ x,
                                           int64_t n_o, int64_t n_v, int64_t n) {
  const int64_t N = n_o * n_v;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    if (k < N) {
      x[k] = eps[n_o + k % n_v] - eps[k / n_v];
    } else {
      const int64_t q = k - N;
      const int64_t X = tri_root(q), Y = q - X * (X + 1) / 2;
      x[k] = eps[n_o + X % n_v] - eps[X / n_v] + eps[n_o + Y % n_v] - eps[Y / n_v];
    }
  }
}

__global__ void axpy_kernel(int64_t n, double alpha, const double* __restrict__ x,
                            double* __restrict__ y) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x)
    y[k] += alpha * x[k];
}

__global__ void fill_kernel(double* __restrict__ x, int64_t n, double value) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x)
    x[k] = value;
}

}  // namespace

void permute4(hipStream_t s, const double* x, double* y, const int64_t e[4], const int64_t st[4],
              double alpha, double beta) {
  const int64_t n = e[0] * e[1] * e[2] * e[3];
  if (n == 0) return;
  permute4_kernel<<<grid_for(n), kBlock, 0, s>>>(x, y, e[0], e[1], e[2], e[3], st[0], st[1], st[2],
                                                 st[3], alpha, beta);
  ETG_LAUNCH("permute4");
}

void sort4(hipStream_t s, const double* x, const int64_t d[4], const char* perm, double* y,
           double alpha, double beta) {
  if (std::strlen(perm) != 4) fail("sort4: permutation must have 4 digits");
  const int64_t in_stride[4] = {1, d[0], d[0] * d[1], d[0] * d[1] * d[2]};
  int64_t e[4], st[4];
  int seen = 0;
  for (int k = 0; k < 4; ++k) {
    const int q = perm[k] - '1';
    if (q < 0 || q > 3 || (seen & (1 << q))) fail(std::string("sort4: bad permutation ") + perm);
    seen |= 1 << q;
    e[k] = d[q];
    st[k] = in_stride[q];
  }
  permute4(s, x, y, e, st, alpha, beta);
}

void squareup(hipStream_t s, const double* packed, double* full, int64_t N) {
  squareup_kernel<<<grid_for(N * N), kBlock, 0, s>>>(packed, full, N);
  ETG_LAUNCH("squareup");
}

void packin(hipStream_t s, const double* full, double* packed, int64_t N, double alpha, double beta,
            double diag_scale) {
  packin_kernel<<<grid_for(N * (N + 1) / 2), kBlock, 0, s>>>(full, packed, N, alpha, beta, diag_scale);
  ETG_LAUNCH("packin");
}

void symmetric_sum(hipStream_t s, const double* x, double* y, int64_t N) {
  if (x == y) fail("symmetric_sum: out of place only");
  dim3 grid(unsigned(ceil_div(N, 32)), unsigned(ceil_div(N, 32)));
  symmetric_sum_kernel<<<grid, dim3(32, 8), 0, s>>>(x, y, N);
  ETG_LAUNCH("symmetric_sum");
}

void symmetrize_and_pack(hipStream_t s, const double* x, double* packed, int64_t n_v, int64_t n_o,
                         double alpha, double f1, double f2) {
  const int64_t N = n_v * n_o;
  symmetrize_and_pack_kernel<<<grid_for(N * (N + 1) / 2), kBlock, 0, s>>>(x, packed, n_v, n_o,
                                                                          alpha, f1, f2);
  ETG_LAUNCH("symmetrize_and_pack");
}

void ladder_t_pm(hipStream_t s, const double* t2, double* tp, double* tm, int64_t n_o, int64_t n_v) {
  const int64_t n = (n_o * (n_o + 1) / 2) * (n_v * (n_v + 1) / 2);
  ladder_t_pm_kernel<<<grid_for(n), kBlock, 0, s>>>(t2, tp, tm, n_o, n_v);
  ETG_LAUNCH("ladder_t_pm");
}

void ladder_g_pm(hipStream_t s, const double* G, int64_t n_v, int64_t lenA, int64_t lenB,
                 const int32_t* pair_a, const int32_t* pair_b, int64_t npairs, double* gp, double* gm) {
  const int64_t n = (n_v * (n_v + 1) / 2) * npairs;
  if (n == 0) return;
  ladder_g_pm_kernel<<<grid_for(n), kBlock, 0, s>>>(G, n_v, lenA, lenB, pair_a, pair_b, npairs, gp, gm);
  ETG_LAUNCH("ladder_g_pm");
}

void ladder_gather(hipStream_t s, const double* wp, const double* wm, const int32_t* col,
                   double* omega2, int64_t n_o, int64_t n_v) {
  const int64_t N = n_o * n_v;
  ladder_gather_kernel<<<grid_for(N * (N + 1) / 2), kBlock, 0, s>>>(wp, wm, col, omega2, n_o, n_v);
  ETG_LAUNCH("ladder_gather");
}

void ccsd_energy(hipStream_t s, const double* t1, const double* t2, const double* g_ovov,
                 int64_t n_o, int64_t n_v, double* partial, double* result) {
  energy_kernel<<<kReduceBlocks, kBlock, 0, s>>>(t1, t2, g_ovov, n_o, n_v, partial);
  ETG_LAUNCH("energy");
  finish_sum_kernel<<<1, kBlock, 0, s>>>(partial, kReduceBlocks, result);
  ETG_LAUNCH("finish_sum");
}

void sum_squares(hipStream_t s, const double* x, int64_t n, double* partial, double* result) {
  sumsq_kernel<<<kReduceBlocks, kBlock, 0, s>>>(x, nullptr, n, partial);
  ETG_LAUNCH("sum_squares");
  finish_sum_kernel<<<1, kBlock, 0, s>>>(partial, kReduceBlocks, result);
  ETG_LAUNCH("finish_sum");
}

void dot(hipStream_t s, const double* x, const double* y, int64_t n, double* partial, double* result) {
  sumsq_kernel<<<kReduceBlocks, kBlock, 0, s>>>(x, y, n, partial);
  ETG_LAUNCH("dot");
  finish_sum_kernel<<<1, kBlock, 0, s>>>(partial, kReduceBlocks, result);
  ETG_LAUNCH("finish_sum");
}

void mp2_guess(hipStream_t s, const double* g_vovo, const double* eps, double* t2, int64_t n_o,
               int64_t n_v) {
  const int64_t N = n_o * n_v;
  mp2_guess_kernel<<<grid_for(N * (N + 1) / 2), kBlock, 0, s>>>(g_vovo, eps, t2, n_o, n_v);
  ETG_LAUNCH("mp2_guess");
}

void orbital_differences(hipStream_t s, const double* eps, double* x, int64_t n_o, int64_t n_v,
                         int64_t n) {
  orbital_differences_kernel<<<grid_for(n), kBlock, 0, s>>>(eps, x, n_o, n_v, n);
  ETG_LAUNCH("orbital_differences");
}

void axpy(hipStream_t s, int64_t n, double alpha, const double* x, double* y) {
  if (n == 0) return;
  axpy_kernel<<<grid_for(n), kBlock, 0, s>>>(n, alpha, x, y);
  ETG_LAUNCH("axpy");
}

void fill(hipStream_t s, double* x, int64_t n, double value) {
  if (n == 0) return;
  fill_kernel<<<grid_for(n), kBlock, 0, s>>>(x, n, value);
  ETG_LAUNCH("fill");
}

}  // namespace etg
