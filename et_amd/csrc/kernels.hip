// Bandwidth-bound kernels of the CCSD residual (see kernels.hpp).
#include "kernels.hpp"
#include "ladder_order.hpp"

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

// I: index type (32-bit when the array has fewer than 2^32 elements: the
// element-wise divisions then cost a fraction of the 64-bit ones)
template <typename I>
__global__ void permute4_kernel(const double* __restrict__ x, double* __restrict__ y, I e0, I e1, I e2,
                                I e3, I s0, I s1, I s2, I s3, double alpha, double beta) {
  const I n = e0 * e1 * e2 * e3;
  for (I idx = I(blockIdx.x) * blockDim.x + threadIdx.x; idx < n; idx += I(gridDim.x) * blockDim.x) {
    I r = idx;
    const I o0 = r % e0; r /= e0;
    const I o1 = r % e1; r /= e1;
    const I o2 = r % e2;
    const I o3 = r / e2;
    const double v = alpha * x[o0 * s0 + o1 * s1 + o2 * s2 + o3 * s3];
    y[idx] = (beta == 0.0) ? v : v + beta * y[idx];
  }
}

// Permutation whose output-fastest index is not the input-fastest one: 32 x 32
// tiles of the (output dim 0, output dim p) plane go through LDS, where p is the
// output position of the input-fastest index, so reads (along dim p) and writes
// (along dim 0) are both unit-stride.  q, r: the other two output dims.
__global__ void permute4_tiled_kernel(const double* __restrict__ x, double* __restrict__ y,
                                      int64_t e0, int64_t ep, int64_t eq, int64_t er, int64_t s0,
                                      int64_t sq, int64_t sr, int64_t yp, int64_t yq, int64_t yr,
                                      double alpha, double beta) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t nt0 = (e0 + 31) / 32;
  const int64_t t0 = blockIdx.x % nt0, tp = blockIdx.x / nt0;
  for (int64_t o = blockIdx.y; o < eq * er; o += gridDim.y) {
    const int64_t iq = o % eq, ir = o / eq;
    const double* xb = x + iq * sq + ir * sr;
    double* yb = y + iq * yq + ir * yr;
    for (int k = ty; k < 32; k += 8) {
      const int64_t i0 = t0 * 32 + k, ip = tp * 32 + tx;
      tile[k][tx] = (i0 < e0 && ip < ep) ? xb[i0 * s0 + ip] : 0.0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
      const int64_t ip = tp * 32 + k, i0 = t0 * 32 + tx;
      if (i0 < e0 && ip < ep) {
        double* dst = yb + i0 + ip * yp;
        const double v = alpha * tile[tx][k];
        *dst = (beta == 0.0) ? v : v + beta * *dst;
      }
    }
    __syncthreads();
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

// The same through 32 x 32 LDS tiles: one workgroup per tile pair (Y block <=
// X block) reads packed elements (X, Y), Y <= X, along Y and writes full(Y, X)
// along Y and its mirror full(X, Y) along X.
__global__ __launch_bounds__(256) void squareup_tiled_kernel(const double* __restrict__ packed,
                                                             double* __restrict__ full, int64_t N) {
  __shared__ double P[32][33];                          // P[xl][yl] = packed(X0 + xl, Y0 + yl)
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t xb = tri_root(blockIdx.x), yb = int64_t(blockIdx.x) - xb * (xb + 1) / 2;
  const int64_t X0 = xb * 32, Y0 = yb * 32;
  for (int k = ty; k < 32; k += 8) {
    const int64_t X = X0 + k, Y = Y0 + tx;
    P[k][tx] = (X < N && Y <= X) ? packed[X * (X + 1) / 2 + Y] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t X = X0 + k, Y = Y0 + tx;              // full(Y, X), along Y
    if (X < N && Y <= X) full[Y + N * X] = P[k][tx];
    const int64_t Xr = X0 + tx, Yc = Y0 + k;            // full(Xr, Yc), along Xr
    if (Xr < N && Yc <= Xr) full[Xr + N * Yc] = P[tx][k];
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

// The same through 32 x 32 LDS tiles: one workgroup per tile pair (Y block <=
// X block) stages x(Y, X) (read along Y), x(X, Y) (along X) and, when f2 != 0,
// the two exchange terms x(p + v s, r + v q) (along p) and x(r + v q, p + v s)
// (along r) for Y = p + v q, X = r + v s, then adds to packed element
// X (X + 1) / 2 + Y, contiguous in Y for each X.
__global__ __launch_bounds__(256) void symmetrize_and_pack_tiled_kernel(const double* __restrict__ x,
                                                                        double* __restrict__ y, int64_t n_v,
                                                                        int64_t N, double alpha, double f1,
                                                                        double f2) {
  __shared__ double A[32][33], B[32][33];              // A[xl][yl] = x(Y, X), B[yl][xl] = x(X, Y)
  __shared__ double C[32][33], D[32][33];              // C[xl][yl] = x(p + v s, r + v q), D[yl][xl] = mirror
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t xb = tri_root(blockIdx.x), yb = int64_t(blockIdx.x) - xb * (xb + 1) / 2;
  const int64_t X0 = xb * 32, Y0 = yb * 32;
  const bool exch = f2 != 0.0;
  for (int k = ty; k < 32; k += 8) {
    const int64_t Yr = Y0 + tx, Xc = X0 + k;            // A[k][tx] = x(Y0 + tx, X0 + k)
    const bool okA = Yr < N && Xc < N;
    A[k][tx] = okA ? x[Yr + N * Xc] : 0.0;
    const int64_t Xr = X0 + tx, Yc = Y0 + k;            // B[k][tx] = x(X0 + tx, Y0 + k)
    const bool okB = Xr < N && Yc < N;
    B[k][tx] = okB ? x[Xr + N * Yc] : 0.0;
    if (exch) {
      // C[k][tx]: Y = Yr = p + v q, X = Xc = r + v s  ->  x(p + v s, r + v q)
      const int64_t p = Yr % n_v, q = Yr / n_v, r = Xc % n_v, sx = Xc / n_v;
      C[k][tx] = okA ? x[(p + n_v * sx) + N * (r + n_v * q)] : 0.0;
      // D[k][tx]: X = Xr = r' + v s', Y = Yc = p' + v q'  ->  x(r' + v q', p' + v s')
      const int64_t p2 = Yc % n_v, q2 = Yc / n_v, r2 = Xr % n_v, s2 = Xr / n_v;
      D[k][tx] = okB ? x[(r2 + n_v * q2) + N * (p2 + n_v * s2)] : 0.0;
    }
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t X = X0 + k, Y = Y0 + tx;
    if (X >= N || Y > X) continue;
    double v;
    if (X == Y) {
      v = alpha * A[k][tx];
    } else {
      v = f1 * (A[k][tx] + B[tx][k]);
      if (exch) v += f2 * (C[k][tx] + D[tx][k]);
      v *= alpha;
    }
    y[X * (X + 1) / 2 + Y] += v;
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

// t+ (op x KP) and t- (om x KM) with the contraction index in the tile-blocked
// order of ladder_order.hpp (zero at padding positions).
__global__ void ladder_t_pm_blocked_kernel(const double* __restrict__ t2, double* __restrict__ tp,
                                           double* __restrict__ tm, int64_t n_o, int64_t n_v, int64_t KP,
                                           int64_t KM) {
  const int64_t op = n_o * (n_o + 1) / 2, om = n_o * (n_o - 1) / 2;
  const int64_t np = op * KP, n = np + om * KM;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const bool plus = idx < np;
    const int64_t q = plus ? idx : idx - np;
    const int64_t nij = plus ? op : om;
    const int64_t rij = q % nij, kk = q / nij;
    // plus: i <= j at j (j + 1) / 2 + i; minus: i < j at j (j - 1) / 2 + i
    const int64_t j = plus ? tri_root(rij) : tri_root(rij) + 1;
    const int64_t i = rij - (plus ? j * (j + 1) / 2 : j * (j - 1) / 2);
    int64_t c = 0, d = 0;
    double val = 0.0;
    if (ladder_pair_of(kk, n_v, plus, c, d)) {
      const double x1 = t2[pidx(c + n_v * i, d + n_v * j)];
      const double x2 = t2[pidx(c + n_v * j, d + n_v * i)];
      val = plus ? (c == d ? 0.5 : 1.0) * (x1 + x2) : x1 - x2;
    }
    (plus ? tp : tm)[q] = val;
  }
}

// g+ (v(v+1)/2 x npairs) and g- (v(v-1)/2 x npairs) of a G block.  For pair k = (a, b) the v x v block
// M(c, d) = G(c + v a, d + v b) gives g+_cd = M(c,d) + M(d,c) (c <= d) and
// g-_cd = M(c,d) - M(d,c) (c < d); one workgroup per tile pair (ti <= tj) reads
// the tiles M(ti, tj) and M(tj, ti) unit-stride into LDS.
__global__ void ladder_g_pm_tiled_kernel(const double* __restrict__ G, int64_t n_v, int64_t lenA,
                                         const int32_t* __restrict__ pa, const int32_t* __restrict__ pb,
                                         int64_t npairs, double* __restrict__ gp, double* __restrict__ gm) {
  __shared__ double A[32][33], B[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t vp = n_v * (n_v + 1) / 2, vm = n_v * (n_v - 1) / 2;
  const int64_t ldG = n_v * lenA;
  const int64_t tj = tri_root(blockIdx.x), ti = int64_t(blockIdx.x) - tj * (tj + 1) / 2;
  const int64_t c0 = ti * 32, d0 = tj * 32;
  for (int64_t k = blockIdx.y; k < npairs; k += gridDim.y) {
    const int64_t a = pa[k], b = pb[k];
    const double* M = G + n_v * a + ldG * n_v * b;           // M(c, d) = M[c + ldG d]
    for (int q = ty; q < 32; q += 8) {
      const int64_t c = c0 + tx, d = d0 + q;                 // A[tx][q] = M(c0+tx, d0+q)
      A[tx][q] = (c < n_v && d < n_v) ? M[c + ldG * d] : 0.0;
      const int64_t c2 = d0 + tx, d2 = c0 + q;               // B[tx][q] = M(d0+tx, c0+q)
      B[tx][q] = (c2 < n_v && d2 < n_v) ? M[c2 + ldG * d2] : 0.0;
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
      const int64_t c = c0 + tx, d = d0 + q;
      if (c <= d && d < n_v) {
        const double x1 = A[tx][q], x2 = B[q][tx];           // M(c, d), M(d, c)
        gp[(d * (d + 1) / 2 + c) + vp * k] = x1 + x2;
        if (c < d) gm[(d * (d - 1) / 2 + c) + vm * k] = x1 - x2;
      }
    }
    __syncthreads();
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

// Core projections; core[i] != 0 marks core occupied MOs; blockIdx.y is the vector.
// remove_core = 0, core-valence separation (cvs_projection_ccs, ccs_class.F90:1219-1250,
// cvs_projection_doubles, doubles_class.F90:287-328): zero every singles element whose
// occupied index is not a core MO and every packed doubles element (ai, bj) with neither
// i nor j a core MO.  remove_core = 1 (rm_core_projection_ccs, ccs_class.F90:1251-1280,
// rm_core_projection_doubles, doubles_class.F90:330-370): zero exactly the others.
// mode 2, ionisation (ip_projection_ccs, ccs_class.F90:1611-1640, ip_projection_doubles,
// doubles_class.F90:376-418): keep only singles into the last virtual (the bath orbital)
// and doubles with a or b the bath orbital; core is not read.
__global__ void cvs_project_kernel(double* __restrict__ x, int64_t n_t1, int64_t n, int64_t n_v,
                                   const unsigned char* __restrict__ core, int mode) {
  double* xv = x + int64_t(blockIdx.y) * n;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    bool hit;                                // an occupied core index (modes 0, 1) or the bath (2)
    if (k < n_t1) {
      hit = mode == 2 ? k % n_v == n_v - 1 : core[k / n_v] != 0;
    } else {
      const int64_t q = k - n_t1;
      const int64_t X = tri_root(q), Y = q - X * (X + 1) / 2;
      hit = mode == 2 ? (X % n_v == n_v - 1 || Y % n_v == n_v - 1) : (core[X / n_v] != 0 || core[Y / n_v] != 0);
    }
    if (hit == (mode == 1)) xv[k] = 0.0;
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
  int p = -1;                                   // output position of the input-fastest index
  for (int k = 0; k < 4; ++k)
    if (st[k] == 1 && e[k] > 1) p = k;
  if (p > 0 && e[0] >= 8 && e[p] >= 8) {
    int oth[2], no = 0;
    for (int k = 1; k < 4; ++k)
      if (k != p) oth[no++] = k;
    const int64_t ostride[4] = {1, e[0], e[0] * e[1], e[0] * e[1] * e[2]};
    const int64_t tiles = ceil_div(e[0], 32) * ceil_div(e[p], 32);
    const int64_t others = e[oth[0]] * e[oth[1]];
    dim3 grid(unsigned(tiles), unsigned(std::min<int64_t>(others, 65535)));
    permute4_tiled_kernel<<<grid, kBlock, 0, s>>>(x, y, e[0], e[p], e[oth[0]], e[oth[1]], st[0],
                                                  st[oth[0]], st[oth[1]], ostride[p], ostride[oth[0]],
                                                  ostride[oth[1]], alpha, beta);
    ETG_LAUNCH("permute4_tiled");
    return;
  }
  // (32-bit indices: the grid-stride step stays below 2^32 - n as well)
  if (n + int64_t(kBlock) * grid_for(n) <= int64_t(UINT32_MAX))
    permute4_kernel<uint32_t><<<grid_for(n), kBlock, 0, s>>>(
        x, y, uint32_t(e[0]), uint32_t(e[1]), uint32_t(e[2]), uint32_t(e[3]), uint32_t(st[0]), uint32_t(st[1]),
        uint32_t(st[2]), uint32_t(st[3]), alpha, beta);
  else
    permute4_kernel<int64_t><<<grid_for(n), kBlock, 0, s>>>(x, y, e[0], e[1], e[2], e[3], st[0], st[1], st[2],
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
  if (N >= 32) {
    const int64_t nb = ceil_div(N, 32);
    squareup_tiled_kernel<<<unsigned(nb * (nb + 1) / 2), kBlock, 0, s>>>(packed, full, N);
    ETG_LAUNCH("squareup_tiled");
    return;
  }
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
  if (N >= 32) {
    const int64_t nb = ceil_div(N, 32);
    symmetrize_and_pack_tiled_kernel<<<unsigned(nb * (nb + 1) / 2), kBlock, 0, s>>>(x, packed, n_v, N, alpha,
                                                                                   f1, f2);
    ETG_LAUNCH("symmetrize_and_pack_tiled");
    return;
  }
  symmetrize_and_pack_kernel<<<grid_for(N * (N + 1) / 2), kBlock, 0, s>>>(x, packed, n_v, n_o,
                                                                          alpha, f1, f2);
  ETG_LAUNCH("symmetrize_and_pack");
}

void ladder_t_pm(hipStream_t s, const double* t2, double* tp, double* tm, int64_t n_o, int64_t n_v) {
  const int64_t n = (n_o * (n_o + 1) / 2) * (n_v * (n_v + 1) / 2);
  ladder_t_pm_kernel<<<grid_for(n), kBlock, 0, s>>>(t2, tp, tm, n_o, n_v);
  ETG_LAUNCH("ladder_t_pm");
}

void ladder_t_pm_blocked(hipStream_t s, const double* t2, double* tp, double* tm, int64_t n_o,
                         int64_t n_v) {
  const int64_t KP = ladder_blocked_len(n_v, true), KM = ladder_blocked_len(n_v, false);
  const int64_t n = (n_o * (n_o + 1) / 2) * KP + (n_o * (n_o - 1) / 2) * KM;
  ladder_t_pm_blocked_kernel<<<grid_for(n), kBlock, 0, s>>>(t2, tp, tm, n_o, n_v, KP, KM);
  ETG_LAUNCH("ladder_t_pm_blocked");
}

void ladder_g_pm(hipStream_t s, const double* G, int64_t n_v, int64_t lenA, int64_t lenB,
                 const int32_t* pair_a, const int32_t* pair_b, int64_t npairs, double* gp, double* gm) {
  const int64_t n = (n_v * (n_v + 1) / 2) * npairs;
  if (n == 0) return;
  (void)lenB;
  const int64_t nt = (n_v + 31) / 32;
  dim3 grid(unsigned(nt * (nt + 1) / 2), unsigned(std::min<int64_t>(npairs, 65535)));
  ladder_g_pm_tiled_kernel<<<grid, kBlock, 0, s>>>(G, n_v, lenA, pair_a, pair_b, npairs, gp, gm);
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

void cvs_project(hipStream_t s, double* x, int64_t n_vec, int64_t n_t1, int64_t n, int64_t n_v,
                 const unsigned char* core, int mode) {
  for (int64_t v0 = 0; v0 < n_vec; v0 += 65535) {
    const dim3 g(grid_for(n, kBlock, 1 << 14), unsigned(std::min<int64_t>(n_vec - v0, 65535)));
    cvs_project_kernel<<<g, kBlock, 0, s>>>(x + v0 * n, n_t1, n, n_v, core, mode);
    ETG_LAUNCH("cvs_project");
  }
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

// ===========================================================================
// Adjoint kernels and T1 helpers
// ===========================================================================
namespace etg {
namespace {

__device__ inline int64_t tri_root2(int64_t k) {
  int64_t x = int64_t((sqrt(8.0 * double(k) + 1.0) - 1.0) * 0.5);
  while (x * (x + 1) / 2 > k) --x;
  while ((x + 1) * (x + 2) / 2 <= k) ++x;
  return x;
}

__device__ inline int64_t pidx2(int64_t x, int64_t y) {
  const int64_t hi = x > y ? x : y, lo = x > y ? y : x;
  return hi * (hi + 1) / 2 + lo;
}

#define GRID_LOOP(var, n)                                                      \
  for (int64_t var = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; var < (n); \
       var += int64_t(gridDim.x) * blockDim.x)

__global__ void scale_kernel(double* x, int64_t n, double beta) {
  GRID_LOOP(k, n) x[k] = (beta == 0.0) ? 0.0 : beta * x[k];
}

__global__ void symmetric_sum_acc_kernel(const double* __restrict__ x, double* __restrict__ y,
                                         int64_t N, double beta) {
  __shared__ double tile[32][33];
  const int64_t bx = blockIdx.x, by = blockIdx.y;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = by * 32 + tx, c = bx * 32 + k;
    tile[k][tx] = (r < N && c < N) ? x[r + N * c] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = bx * 32 + tx, c = by * 32 + k;
    if (r < N && c < N) {
      const double v = x[r + N * c] + tile[tx][k];
      y[r + N * c] = (beta == 0.0) ? v : v + beta * y[r + N * c];
    }
  }
}

__global__ void packsum_kernel(const double* __restrict__ full, double* __restrict__ packed,
                               int64_t N) {
  GRID_LOOP(k, N * (N + 1) / 2) {
    const int64_t X = tri_root2(k), Y = k - X * (X + 1) / 2;
    double v = full[X + N * Y];
    if (X != Y) v += full[Y + N * X];
    packed[k] += v;
  }
}

__global__ void unpack_upper_add_kernel(const double* __restrict__ packed, double* __restrict__ full,
                                        int64_t N, double alpha, double diag_scale) {
  GRID_LOOP(k, N * (N + 1) / 2) {
    const int64_t X = tri_root2(k), Y = k - X * (X + 1) / 2;
    double v = alpha * packed[k];
    if (X == Y) v *= diag_scale;
    full[Y + N * X] += v;
  }
}

// x(R, C) += alpha * [R == C ? p{R,R} : f1 p{R,C} + f2 p{U,W}],
// U = (R_a, C_i), W = (C_a, R_i)  (the element the psrq terms of the forward pass read)
__global__ void sympack_adjoint_kernel(const double* __restrict__ p, double* __restrict__ x,
                                       int64_t n_v, int64_t n_o, double alpha, double f1, double f2) {
  const int64_t N = n_v * n_o;
  GRID_LOOP(idx, N * N) {
    const int64_t R = idx % N, C = idx / N;
    double v;
    if (R == C) {
      v = p[pidx2(R, R)];
    } else {
      v = f1 * p[pidx2(R, C)];
      if (f2 != 0.0) {
        const int64_t Ra = R % n_v, Ri = R / n_v, Ca = C % n_v, Ci = C / n_v;
        const int64_t U = Ra + n_v * Ci, W = Ca + n_v * Ri;
        if (U != W) v += f2 * p[pidx2(U, W)];
      }
    }
    x[idx] += alpha * v;
  }
}

// The f2 = 0 form through 32 x 32 LDS tiles (as squareup: one workgroup per
// tile pair reads packed (X, Y), Y <= X, along Y and updates x(Y, X) along Y
// and x(X, Y) along X).
__global__ __launch_bounds__(256) void sympack_adjoint_tiled_kernel(const double* __restrict__ p,
                                                                    double* __restrict__ x, int64_t N,
                                                                    double alpha, double f1) {
  __shared__ double P[32][33];                          // P[xl][yl] = p(X0 + xl, Y0 + yl)
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t xb = tri_root2(blockIdx.x), yb = int64_t(blockIdx.x) - xb * (xb + 1) / 2;
  const int64_t X0 = xb * 32, Y0 = yb * 32;
  for (int k = ty; k < 32; k += 8) {
    const int64_t X = X0 + k, Y = Y0 + tx;
    P[k][tx] = (X < N && Y <= X) ? p[X * (X + 1) / 2 + Y] * (X == Y ? alpha : alpha * f1) : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t X = X0 + k, Y = Y0 + tx;              // x(Y, X), along Y
    if (X < N && Y <= X) x[Y + N * X] += P[k][tx];
    const int64_t Xr = X0 + tx, Yc = Y0 + k;            // x(Xr, Yc), along Xr (the diagonal once)
    if (Xr < N && Yc < Xr) x[Xr + N * Yc] += P[tx][k];
  }
}

// t2[{(c,i),(d,j)}] += coef * tp[ij, cd] + sgn * tm[ij, cd]
__global__ void ladder_t_pm_adjoint_kernel(const double* __restrict__ tp, const double* __restrict__ tm,
                                           double* __restrict__ t2, int64_t n_o, int64_t n_v) {
  const int64_t Nov = n_o * n_v;
  const int64_t op = n_o * (n_o + 1) / 2, om = n_o * (n_o - 1) / 2;
  GRID_LOOP(k, Nov * (Nov + 1) / 2) {
    const int64_t X = tri_root2(k), Y = k - X * (X + 1) / 2;
    const int64_t c = X % n_v, i = X / n_v, d = Y % n_v, j = Y / n_v;
    const int64_t il = i < j ? i : j, ih = i < j ? j : i;
    const int64_t cl = c < d ? c : d, ch = c < d ? d : c;
    const double coef = (c == d) ? 1.0 : ((i == j) ? 2.0 : 1.0);
    double v = coef * tp[(ih * (ih + 1) / 2 + il) + op * (ch * (ch + 1) / 2 + cl)];
    if (i != j && c != d) {
      const double sgn = ((c < d) == (i < j)) ? 1.0 : -1.0;
      v += sgn * tm[(ih * (ih - 1) / 2 + il) + om * (ch * (ch - 1) / 2 + cl)];
    }
    t2[k] += v;
  }
}

// G(c, a', d, b') = [a >= b] (c == d ? 2 gp[cc,k] : gp[cd,k] +- gm[cd,k]), k = col[a + v b] - col0
__global__ void ladder_g_pm_adjoint_kernel(const double* __restrict__ gp, const double* __restrict__ gm,
                                           double* __restrict__ G, int64_t n_v, int64_t a0,
                                           int64_t lenA, int64_t b0, int64_t lenB,
                                           const int32_t* __restrict__ col, int64_t col0) {
  const int64_t vp = n_v * (n_v + 1) / 2, vm = n_v * (n_v - 1) / 2;
  const int64_t rows = n_v * lenA;
  GRID_LOOP(idx, rows * n_v * lenB) {
    const int64_t r = idx % rows, q = idx / rows;
    const int64_t c = r % n_v, aa = r / n_v, d = q % n_v, bb = q / n_v;
    const int64_t a = a0 + aa, b = b0 + bb;
    double v = 0.0;
    if (a >= b) {
      const int64_t k = int64_t(col[a + n_v * b]) - col0;
      if (c == d) {
        v = 2.0 * gp[(c * (c + 1) / 2 + c) + vp * k];
      } else {
        const int64_t lo = c < d ? c : d, hi = c < d ? d : c;
        v = gp[(hi * (hi + 1) / 2 + lo) + vp * k];
        const double m = gm[(hi * (hi - 1) / 2 + lo) + vm * k];
        v += (c < d) ? m : -m;
      }
    }
    G[idx] = v;
  }
}

// adjoint of ladder_gather: for stored (i' <= j', a' >= b'):
// wp += s1 om[P1] + s2 om[P2],  wm += s1 om[P1] - s2 om[P2]
// P1 = {(a',i'),(b',j')}, P2 = {(a',j'),(b',i')}, s = 1/2 on the packed diagonal
__global__ void ladder_gather_adjoint_kernel(const double* __restrict__ om,
                                             const int32_t* __restrict__ pa,
                                             const int32_t* __restrict__ pb, double* __restrict__ wp,
                                             double* __restrict__ wm, int64_t n_o, int64_t n_v) {
  const int64_t op = n_o * (n_o + 1) / 2, omn = n_o * (n_o - 1) / 2;
  const int64_t vp = n_v * (n_v + 1) / 2;
  GRID_LOOP(idx, op * vp) {
    const int64_t r = idx % op, cl = idx / op;
    const int64_t j = tri_root2(r), i = r - j * (j + 1) / 2;       // i <= j
    const int64_t a = pa[cl], b = pb[cl];                          // a >= b
    const int64_t X1 = a + n_v * i, Y1 = b + n_v * j;
    const int64_t X2 = a + n_v * j, Y2 = b + n_v * i;
    const int64_t k1 = pidx2(X1, Y1), k2 = pidx2(X2, Y2);
    const double v1 = (X1 == Y1 ? 0.5 : 1.0) * om[k1];
    const double v2 = (k2 != k1) ? (X2 == Y2 ? 0.5 : 1.0) * om[k2] : 0.0;
    wp[idx] = v1 + v2;
    if (i != j) wm[(j * (j - 1) / 2 + i) + omn * cl] = (a != b) ? v1 - v2 : 0.0;
  }
}

__global__ void t1_xy_kernel2(const double* t1, double* X, double* Y, int64_t n_o, int64_t n_v,
                              double diag) {
  const int64_t n = n_o + n_v;
  GRID_LOOP(k, n * n) {
    const int64_t p = k % n, q = k / n;
    double x = (p == q) ? diag : 0.0, y = x;
    if (p >= n_o && q < n_o) x = -t1[(p - n_o) + n_v * q];  // X(o+a, i) = -t_ai
    if (p < n_o && q >= n_o) y = t1[(q - n_o) + n_v * p];   // Y(i, o+a) = t_ai
    X[k] = x;
    Y[k] = y;
  }
}

__global__ void t1_xy_adjoint_kernel(const double* Xb, const double* Yb, double* t1b, int64_t n_o,
                                     int64_t n_v) {
  const int64_t n = n_o + n_v;
  GRID_LOOP(k, n_o * n_v) {
    const int64_t a = k % n_v, i = k / n_v;
    t1b[k] += -Xb[(n_o + a) + n * i] + Yb[i + n * (n_o + a)];
  }
}

__global__ void trace_oo_kernel2(const double* Loo, double* XJ, int64_t nJ, int64_t n_o) {
  GRID_LOOP(J, nJ) {
    double s = 0.0;
    for (int64_t k = 0; k < n_o; ++k) s += Loo[J + nJ * (k + n_o * k)];
    XJ[J] = s;
  }
}

__global__ void trace_oo_adjoint_kernel(const double* XJb, double* Loob, int64_t nJ, int64_t n_o) {
  GRID_LOOP(idx, nJ * n_o) {
    const int64_t J = idx % nJ, k = idx / nJ;
    Loob[J + nJ * (k + n_o * k)] += XJb[J];
  }
}

__global__ void fock_ov_vector_kernel2(const double* F, double* f, int64_t n_o, int64_t n_v) {
  const int64_t n = n_o + n_v;
  GRID_LOOP(k, n_o * n_v) {
    const int64_t b = k % n_v, j = k / n_v;
    f[k] = F[j + n * (n_o + b)];
  }
}

__global__ void fock_ov_vector_adjoint_kernel(const double* fb, double* Fb, int64_t n_o, int64_t n_v) {
  const int64_t n = n_o + n_v;
  GRID_LOOP(k, n_o * n_v) {
    const int64_t b = k % n_v, j = k / n_v;
    Fb[j + n * (n_o + b)] += fb[k];
  }
}

__global__ void copy_sub_kernel2(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
                                 int64_t n, double alpha, double beta) {
  GRID_LOOP(k, m * n) {
    const int64_t r = k % m, c = k / m;
    const double v = alpha * src[r + lds * c];
    dst[r + ldd * c] = (beta == 0.0) ? v : v + beta * dst[r + ldd * c];
  }
}

// dst(m x n, ld ldd) += sum_b part[b m n + r + m c], b < nbatch
__global__ void batch_sum_acc_kernel(const double* __restrict__ part, int64_t nbatch, int64_t m,
                                     int64_t n, double* __restrict__ dst, int64_t ldd) {
  const int64_t mn = m * n;
  GRID_LOOP(k, mn) {
    double v = 0.0;
    for (int64_t b = 0; b < nbatch; ++b) v += part[k + b * mn];
    dst[(k % m) + ldd * (k / m)] += v;
  }
}

inline unsigned g2(int64_t n) { return grid_for(n); }

}  // namespace

void batch_sum_acc(hipStream_t s, const double* part, int64_t nbatch, int64_t m, int64_t n, double* dst,
                   int64_t ldd) {
  if (m * n == 0) return;
  batch_sum_acc_kernel<<<g2(m * n), kBlock, 0, s>>>(part, nbatch, m, n, dst, ldd);
  ETG_LAUNCH("batch_sum_acc");
}

void scale(hipStream_t s, double* x, int64_t n, double beta) {
  if (n == 0 || beta == 1.0) return;
  scale_kernel<<<g2(n), kBlock, 0, s>>>(x, n, beta);
  ETG_LAUNCH("scale");
}

void symmetric_sum_acc(hipStream_t s, const double* x, double* y, int64_t N, double beta) {
  if (x == y) fail("symmetric_sum_acc: out of place only");
  dim3 grid(unsigned(ceil_div(N, 32)), unsigned(ceil_div(N, 32)));
  symmetric_sum_acc_kernel<<<grid, dim3(32, 8), 0, s>>>(x, y, N, beta);
  ETG_LAUNCH("symmetric_sum_acc");
}

void packsum(hipStream_t s, const double* full, double* packed, int64_t N) {
  packsum_kernel<<<g2(N * (N + 1) / 2), kBlock, 0, s>>>(full, packed, N);
  ETG_LAUNCH("packsum");
}

void unpack_upper_add(hipStream_t s, const double* packed, double* full, int64_t N, double alpha,
                      double diag_scale) {
  unpack_upper_add_kernel<<<g2(N * (N + 1) / 2), kBlock, 0, s>>>(packed, full, N, alpha, diag_scale);
  ETG_LAUNCH("unpack_upper_add");
}

void symmetrize_and_pack_adjoint(hipStream_t s, const double* packed, double* x, int64_t n_v,
                                 int64_t n_o, double alpha, double f1, double f2) {
  const int64_t N = n_v * n_o;
  if (f2 == 0.0 && N >= 32) {
    const int64_t nb = ceil_div(N, 32);
    sympack_adjoint_tiled_kernel<<<unsigned(nb * (nb + 1) / 2), kBlock, 0, s>>>(packed, x, N, alpha, f1);
    ETG_LAUNCH("sympack_adjoint_tiled");
    return;
  }
  sympack_adjoint_kernel<<<g2(N * N), kBlock, 0, s>>>(packed, x, n_v, n_o, alpha, f1, f2);
  ETG_LAUNCH("sympack_adjoint");
}

void ladder_t_pm_adjoint(hipStream_t s, const double* tp, const double* tm, double* t2, int64_t n_o,
                         int64_t n_v) {
  const int64_t N = n_o * n_v;
  ladder_t_pm_adjoint_kernel<<<g2(N * (N + 1) / 2), kBlock, 0, s>>>(tp, tm, t2, n_o, n_v);
  ETG_LAUNCH("ladder_t_pm_adjoint");
}

void ladder_g_pm_adjoint(hipStream_t s, const double* gp, const double* gm, double* G, int64_t n_v,
                         int64_t a0, int64_t lenA, int64_t b0, int64_t lenB, const int32_t* col,
                         int64_t col0) {
  const int64_t n = n_v * lenA * n_v * lenB;
  if (n == 0) return;
  ladder_g_pm_adjoint_kernel<<<g2(n), kBlock, 0, s>>>(gp, gm, G, n_v, a0, lenA, b0, lenB, col, col0);
  ETG_LAUNCH("ladder_g_pm_adjoint");
}

void ladder_gather_adjoint(hipStream_t s, const double* omega2, const int32_t* pcol_a,
                           const int32_t* pcol_b, double* wp, double* wm, int64_t n_o, int64_t n_v) {
  const int64_t n = (n_o * (n_o + 1) / 2) * (n_v * (n_v + 1) / 2);
  ladder_gather_adjoint_kernel<<<g2(n), kBlock, 0, s>>>(omega2, pcol_a, pcol_b, wp, wm, n_o, n_v);
  ETG_LAUNCH("ladder_gather_adjoint");
}

void t1_xy(hipStream_t s, const double* t1, double* X, double* Y, int64_t n_o, int64_t n_v,
           double diag) {
  const int64_t n = n_o + n_v;
  t1_xy_kernel2<<<g2(n * n), kBlock, 0, s>>>(t1, X, Y, n_o, n_v, diag);
  ETG_LAUNCH("t1_xy");
}

void t1_xy_adjoint(hipStream_t s, const double* Xb, const double* Yb, double* t1b, int64_t n_o,
                   int64_t n_v) {
  t1_xy_adjoint_kernel<<<g2(n_o * n_v), kBlock, 0, s>>>(Xb, Yb, t1b, n_o, n_v);
  ETG_LAUNCH("t1_xy_adjoint");
}

void trace_oo(hipStream_t s, const double* Loo, double* XJ, int64_t nJ, int64_t n_o) {
  trace_oo_kernel2<<<g2(nJ), kBlock, 0, s>>>(Loo, XJ, nJ, n_o);
  ETG_LAUNCH("trace_oo");
}

void trace_oo_adjoint(hipStream_t s, const double* XJb, double* Loob, int64_t nJ, int64_t n_o) {
  trace_oo_adjoint_kernel<<<g2(nJ * n_o), kBlock, 0, s>>>(XJb, Loob, nJ, n_o);
  ETG_LAUNCH("trace_oo_adjoint");
}

void fock_ov_vector(hipStream_t s, const double* F, double* f, int64_t n_o, int64_t n_v) {
  fock_ov_vector_kernel2<<<g2(n_o * n_v), kBlock, 0, s>>>(F, f, n_o, n_v);
  ETG_LAUNCH("fock_ov_vector");
}

void fock_ov_vector_adjoint(hipStream_t s, const double* fb, double* Fb, int64_t n_o, int64_t n_v) {
  fock_ov_vector_adjoint_kernel<<<g2(n_o * n_v), kBlock, 0, s>>>(fb, Fb, n_o, n_v);
  ETG_LAUNCH("fock_ov_vector_adjoint");
}

void copy_sub(hipStream_t s, const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m,
              int64_t n, double alpha, double beta) {
  if (m * n == 0) return;
  copy_sub_kernel2<<<g2(m * n), kBlock, 0, s>>>(src, lds, dst, ldd, m, n, alpha, beta);
  ETG_LAUNCH("copy_sub");
}

}  // namespace etg
