// Molecular producer of the CCSD inputs on the device: AO integrals over
// contracted Gaussians -> RHF -> pivoted Cholesky decomposition of the AO
// electron-repulsion matrix -> MO Cholesky vectors, h_mo and orbital energies.
//
// eT obtains these from Libint (src/libint/eri.cpp:33-89), its RHF solver
// (src/solvers/hf/scf_solver_class.F90) and its Cholesky decomposition of the
// AO integrals (src/solvers/cholesky/eri_cd_class.F90:241, AO->MO :3132).
//
// Integrals: McMurchie-Davidson Hermite expansion.  A primitive pair
// x^i e^{-a r_A^2} x^j e^{-b r_B^2} = sum_t E^{ij}_t Lambda_t(p, P); overlap,
// kinetic and nuclear attraction follow from E and the Hermite Coulomb
// integrals R^n_{tuv}(alpha, X, Y, Z), the two-electron integral is
//   (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) sum_{tuv} E^{ab}_{tuv}
//             sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'} R_{t+t',u+u',v+v'}(pq/(p+q), P-Q).
// The Boys function F_n(T) is summed from its series for T < 40 with downward
// recursion, and from its asymptotic form with upward recursion above.
// Shells are spherical for l = 2 (real solid harmonics m = -2..2 over
// individually normalised Cartesians), Cartesian = spherical for l <= 1.
//
// The dense AO matrix (pq|rs) (n^2 x n^2) is built in HBM by one wavefront per
// unique shell quartet; the pivoted Cholesky decomposition runs on it with the
// largest remaining diagonal as pivot (ties to the lowest index), each new
// vector one matrix-vector product against the earlier ones; RHF builds J and
// K from the dense integrals, as eT's Hartree-Fock does from its AO integrals.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "frontend.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace etg {
namespace fe {

constexpr int kMaxL = 2;                            // s, p, d
constexpr int kMaxCart = 6;
constexpr int kMaxHerm = 2 * kMaxL + 1;             // t = 0..4 per pair
constexpr int kMaxLtot = 4 * kMaxL;                 // 8
constexpr int kCube = kMaxLtot + 1;                 // R cube edge

__host__ __device__ inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ inline int nsph(int l) { return l < 2 ? ncart(l) : 2 * l + 1; }

// Cartesian components in the order x^lx y^ly z^lz, lx from l down, then ly from l - lx down.
__host__ __device__ inline void cart(int l, int c, int& lx, int& ly, int& lz) {
  int k = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      if (k == c) {
        lx = x;
        ly = y;
        lz = l - x - y;
        return;
      }
      ++k;
    }
  lx = ly = lz = 0;
}

// Spherical d functions over Cartesians (xx, xy, xz, yy, yz, zz) normalised
// individually: xy, yz, (2zz - xx - yy)/2, xz, (xx - yy) sqrt(3)/2.
__host__ __device__ inline double sph_coef(int l, int i, int c) {
  if (l < 2) return i == c ? 1.0 : 0.0;
  const double s3 = 0.86602540378443864676;        // sqrt(3)/2
  switch (i) {
    case 0: return c == 1 ? 1.0 : 0.0;
    case 1: return c == 4 ? 1.0 : 0.0;
    case 2: return c == 5 ? 1.0 : ((c == 0 || c == 3) ? -0.5 : 0.0);
    case 3: return c == 2 ? 1.0 : 0.0;
    default: return c == 0 ? s3 : (c == 3 ? -s3 : 0.0);
  }
}

// A shell's functions and its Cartesian -> basis coefficients: spherical d (above) or,
// for a Cartesian shell, the individually normalised Cartesian components themselves.
__host__ __device__ inline int nfun(int l, bool cartesian) { return cartesian ? ncart(l) : nsph(l); }
__host__ __device__ inline double tcoef(int l, bool cartesian, int i, int c) {
  return cartesian ? (i == c ? 1.0 : 0.0) : sph_coef(l, i, c);
}

// E^{ij}_t for i <= la, j <= lb, t <= i + j:  E[(i (lb+1) + j) (la+lb+1) + t]
__host__ __device__ inline void hermite_e(int la, int lb, double a, double b, double Ax, double Bx,
                                          double* E) {
  const double p = a + b, mu = a * b / p, X = Ax - Bx;
  const double XPA = -b / p * X, XPB = a / p * X, h = 0.5 / p;
  const int L = la + lb + 1;
  for (int k = 0; k < (la + 1) * (lb + 1) * L; ++k) E[k] = 0.0;
  auto at = [&](int i, int j, int t) -> double& { return E[(i * (lb + 1) + j) * L + t]; };
  at(0, 0, 0) = exp(-mu * X * X);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {
          v = XPA * at(i - 1, j, t);
          if (t > 0) v += h * at(i - 1, j, t - 1);
          if (t + 1 <= i - 1 + j) v += (t + 1) * at(i - 1, j, t + 1);
        } else {
          v = XPB * at(i, j - 1, t);
          if (t > 0) v += h * at(i, j - 1, t - 1);
          if (t + 1 <= i + j - 1) v += (t + 1) * at(i, j - 1, t + 1);
        }
        at(i, j, t) = v;
      }
    }
}

// F_n(T), n = 0..nmax (nmax <= 16)
__host__ __device__ inline void boys(int nmax, double T, double* F) {
  const double e = exp(-T);
  if (T < 40.0) {
    // F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * T / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[nmax] = e * sum;
    for (int m = nmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
  } else {
    F[0] = 0.5 * sqrt(3.14159265358979323846 / T);   // erf(sqrt T) = 1 to 1e-18
    for (int m = 0; m < nmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
  }
}

inline int cube(int t, int u, int v) { return (t * kCube + u) * kCube + v; }
__device__ inline int dcube(int t, int u, int v) { return (t * kCube + u) * kCube + v; }

// R^0_{tuv} (t+u+v <= L) into R (cube layout) on the host; work: 2 cubes
void hermite_r_host(int L, double alpha, double X, double Y, double Z, double* R, double* work) {
  double F[kMaxLtot + 1];
  boys(L, alpha * (X * X + Y * Y + Z * Z), F);
  double* cur = work;
  double* nxt = work + kCube * kCube * kCube;
  // level n = L: only R^L_000
  cur[cube(0, 0, 0)] = std::pow(-2.0 * alpha, L) * F[L];
  for (int n = L - 1; n >= 0; --n) {
    const int tot = L - n;
    for (int t = 0; t <= tot; ++t)
      for (int u = 0; u + t <= tot; ++u)
        for (int v = 0; v + u + t <= tot; ++v) {
          double val;
          if (t > 0) {
            val = X * cur[cube(t - 1, u, v)];
            if (t > 1) val += (t - 1) * cur[cube(t - 2, u, v)];
          } else if (u > 0) {
            val = Y * cur[cube(t, u - 1, v)];
            if (u > 1) val += (u - 1) * cur[cube(t, u - 2, v)];
          } else if (v > 0) {
            val = Z * cur[cube(t, u, v - 1)];
            if (v > 1) val += (v - 1) * cur[cube(t, u, v - 2)];
          } else {
            val = std::pow(-2.0 * alpha, n) * F[n];
          }
          nxt[cube(t, u, v)] = val;
        }
    std::swap(cur, nxt);
  }
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u + t <= L; ++u)
      for (int v = 0; v + u + t <= L; ++v) R[cube(t, u, v)] = cur[cube(t, u, v)];
}

// ---------------------------------------------------------------------------
// shells
// ---------------------------------------------------------------------------

struct DevShell {
  double x, y, z;
  int l, nprim, prim0, f0;
  bool cart;
};

double dfact(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// per-component scale turning the x^l normalisation into a unit norm for the component
__host__ __device__ inline double comp_scale(int l, int c) {
  int lx, ly, lz;
  cart(l, c, lx, ly, lz);
  auto df = [](int n) {
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
  };
  return sqrt(df(2 * l - 1) / (df(2 * lx - 1) * df(2 * ly - 1) * df(2 * lz - 1)));
}

void normalise(HostShell& s) {
  const int l = s.l;
  const size_t np = s.exps.size();
  std::vector<double> c(np);
  for (size_t i = 0; i < np; ++i) {
    const double a = s.exps[i];
    c[i] = s.coefs[i] * std::pow(2.0 * a / M_PI, 0.75) * std::pow(4.0 * a, 0.5 * l) / std::sqrt(dfact(2 * l - 1));
  }
  double sum = 0.0;
  for (size_t i = 0; i < np; ++i)
    for (size_t j = 0; j < np; ++j) {
      const double p = s.exps[i] + s.exps[j];
      sum += c[i] * c[j] * std::pow(M_PI / p, 1.5) * dfact(2 * l - 1) / std::pow(2.0 * p, l);
    }
  for (size_t i = 0; i < np; ++i) s.coefs[i] = c[i] / std::sqrt(sum);
}

// ---------------------------------------------------------------------------
// one-electron integrals (host)
// ---------------------------------------------------------------------------

void one_electron(const std::vector<HostShell>& sh, const std::vector<double>& Z,
                  const std::vector<double>& xyz, int64_t n, double* S, double* T, double* V) {
  std::vector<double> work(2 * kCube * kCube * kCube), R(kCube * kCube * kCube);
  for (const HostShell& A : sh)
    for (const HostShell& B : sh) {
      const int la = A.l, lb = B.l, na = ncart(la), nb = ncart(lb);
      double sb[kMaxCart][kMaxCart] = {}, tb[kMaxCart][kMaxCart] = {}, vb[kMaxCart][kMaxCart] = {};
      double Ex[3][(kMaxL + 1) * (kMaxL + 3) * (2 * kMaxL + 3)];
      double Eh[3][(kMaxL + 1) * (kMaxL + 1) * kMaxHerm];
      for (size_t i = 0; i < A.exps.size(); ++i)
        for (size_t j = 0; j < B.exps.size(); ++j) {
          const double a = A.exps[i], b = B.exps[j], p = a + b;
          const double cc = A.coefs[i] * B.coefs[j];
          for (int d = 0; d < 3; ++d) {
            hermite_e(la, lb + 2, a, b, A.c[d], B.c[d], Ex[d]);
            hermite_e(la, lb, a, b, A.c[d], B.c[d], Eh[d]);
          }
          const int Lb2 = la + lb + 3;         // t extent of the lb + 2 tables
          auto s1 = [&](int d, int ii, int jj) { return Ex[d][(ii * (lb + 3) + jj) * Lb2]; };
          const double pref = cc * std::pow(M_PI / p, 1.5);
          double P[3];
          for (int d = 0; d < 3; ++d) P[d] = (a * A.c[d] + b * B.c[d]) / p;
          for (int ca = 0; ca < na; ++ca)
            for (int cb = 0; cb < nb; ++cb) {
              int ax[3], bx[3];
              cart(la, ca, ax[0], ax[1], ax[2]);
              cart(lb, cb, bx[0], bx[1], bx[2]);
              double s[3], k[3];
              for (int d = 0; d < 3; ++d) {
                const int jj = bx[d];
                s[d] = s1(d, ax[d], jj);
                double v = -2.0 * b * b * s1(d, ax[d], jj + 2) + b * (2 * jj + 1) * s1(d, ax[d], jj);
                if (jj >= 2) v -= 0.5 * jj * (jj - 1) * s1(d, ax[d], jj - 2);
                k[d] = v;
              }
              sb[ca][cb] += pref * s[0] * s[1] * s[2];
              tb[ca][cb] += pref * (k[0] * s[1] * s[2] + s[0] * k[1] * s[2] + s[0] * s[1] * k[2]);
            }
          // nuclear attraction
          const int L = la + lb;
          const double vpref = cc * 2.0 * M_PI / p;
          for (size_t C = 0; C < Z.size(); ++C) {
            hermite_r_host(L, p, P[0] - xyz[3 * C], P[1] - xyz[3 * C + 1], P[2] - xyz[3 * C + 2], R.data(),
                           work.data());
            for (int ca = 0; ca < na; ++ca)
              for (int cb = 0; cb < nb; ++cb) {
                int ax[3], bx[3];
                cart(la, ca, ax[0], ax[1], ax[2]);
                cart(lb, cb, bx[0], bx[1], bx[2]);
                double acc = 0.0;
                const int Lh = la + lb + 1;
                for (int t = 0; t <= ax[0] + bx[0]; ++t)
                  for (int u = 0; u <= ax[1] + bx[1]; ++u)
                    for (int v = 0; v <= ax[2] + bx[2]; ++v)
                      acc += Eh[0][(ax[0] * (lb + 1) + bx[0]) * Lh + t] * Eh[1][(ax[1] * (lb + 1) + bx[1]) * Lh + u] *
                             Eh[2][(ax[2] * (lb + 1) + bx[2]) * Lh + v] * R[cube(t, u, v)];
                vb[ca][cb] -= Z[C] * vpref * acc;
              }
          }
        }
      // Cartesian -> spherical: out(i, j) = sum_ab T_a(i, a) x(a, b) T_b(j, b), T = sph * comp_scale
      for (int i = 0; i < nfun(la, A.cart); ++i)
        for (int j = 0; j < nfun(lb, B.cart); ++j) {
          double s = 0.0, t = 0.0, v = 0.0;
          for (int ca = 0; ca < na; ++ca) {
            const double ta = tcoef(la, A.cart, i, ca) * comp_scale(la, ca);
            if (ta == 0.0) continue;
            for (int cb = 0; cb < nb; ++cb) {
              const double tbv = tcoef(lb, B.cart, j, cb) * comp_scale(lb, cb);
              if (tbv == 0.0) continue;
              s += ta * tbv * sb[ca][cb];
              t += ta * tbv * tb[ca][cb];
              v += ta * tbv * vb[ca][cb];
            }
          }
          const int64_t k = (A.f0 + i) + n * (B.f0 + j);
          S[k] = s;
          T[k] = t;
          V[k] = v;
        }
    }
}

// <mu| r_c - O_c |nu>: along c the primitive pair gives sqrt(pi/p)(E^{ij}_1 + X_PC E^{ij}_0),
// along the other two directions the overlaps.  D: 3 matrices n x n (x, y, z).
void dipole(const std::vector<HostShell>& sh, const double* O, int64_t n, double* D) {
  for (const HostShell& A : sh)
    for (const HostShell& B : sh) {
      const int la = A.l, lb = B.l, na = ncart(la), nb = ncart(lb);
      double blk[3][kMaxCart][kMaxCart] = {};
      double E[3][(kMaxL + 1) * (kMaxL + 1) * kMaxHerm];
      const int Lh = la + lb + 1;
      for (size_t i = 0; i < A.exps.size(); ++i)
        for (size_t j = 0; j < B.exps.size(); ++j) {
          const double a = A.exps[i], b = B.exps[j], p = a + b;
          const double pref = A.coefs[i] * B.coefs[j] * std::pow(M_PI / p, 1.5);
          for (int d = 0; d < 3; ++d) hermite_e(la, lb, a, b, A.c[d], B.c[d], E[d]);
          for (int ca = 0; ca < na; ++ca)
            for (int cb = 0; cb < nb; ++cb) {
              int ax[3], bx[3];
              cart(la, ca, ax[0], ax[1], ax[2]);
              cart(lb, cb, bx[0], bx[1], bx[2]);
              double s0[3], m1[3];
              for (int d = 0; d < 3; ++d) {
                const double* e = &E[d][(ax[d] * (lb + 1) + bx[d]) * Lh];
                const double XPC = (a * A.c[d] + b * B.c[d]) / p - O[d];
                s0[d] = e[0];
                m1[d] = (ax[d] + bx[d] >= 1 ? e[1] : 0.0) + XPC * e[0];
              }
              blk[0][ca][cb] += pref * m1[0] * s0[1] * s0[2];
              blk[1][ca][cb] += pref * s0[0] * m1[1] * s0[2];
              blk[2][ca][cb] += pref * s0[0] * s0[1] * m1[2];
            }
        }
      for (int c = 0; c < 3; ++c)
        for (int i = 0; i < nfun(la, A.cart); ++i)
          for (int j = 0; j < nfun(lb, B.cart); ++j) {
            double v = 0.0;
            for (int ca = 0; ca < na; ++ca) {
              const double ta = tcoef(la, A.cart, i, ca) * comp_scale(la, ca);
              if (ta == 0.0) continue;
              for (int cb = 0; cb < nb; ++cb) v += ta * tcoef(lb, B.cart, j, cb) * comp_scale(lb, cb) * blk[c][ca][cb];
            }
            D[int64_t(c) * n * n + (A.f0 + i) + n * (B.f0 + j)] = v;
          }
    }
}

// ---------------------------------------------------------------------------
// two-electron integrals (device): one wavefront per unique shell quartet
// ---------------------------------------------------------------------------

constexpr int kEriThreads = 64;
constexpr int kEsz = (kMaxL + 1) * (kMaxL + 1) * kMaxHerm;   // 45 per dimension

__device__ inline int64_t tri_root_f(int64_t k) {
  int64_t x = int64_t((sqrt(8.0 * double(k) + 1.0) - 1.0) * 0.5);
  while (x * (x + 1) / 2 > k) --x;
  while ((x + 1) * (x + 2) / 2 <= k) ++x;
  return x;
}

// G (n^2 x n^2, column-major): G[(p + n q) + n^2 (r + n s)] = (pq|rs)
__global__ __launch_bounds__(kEriThreads) void eri_kernel(const DevShell* __restrict__ shells,
                                                          const int32_t* __restrict__ pair_a,
                                                          const int32_t* __restrict__ pair_b,
                                                          const double* __restrict__ exps,
                                                          const double* __restrict__ coefs,
                                                          int64_t npairs, int64_t n, int64_t q0,
                                                          double* __restrict__ G) {
  __shared__ double Eab[3][kEsz], Ecd[3][kEsz];
  __shared__ double Rc[2][kCube * kCube * kCube];
  __shared__ double acc[kMaxCart * kMaxCart * kMaxCart * kMaxCart];
  __shared__ double buf[kMaxCart * kMaxCart * kMaxCart * kMaxCart];
  __shared__ double Fb[kMaxLtot + 1];
  __shared__ double sc[4];
  const int tid = threadIdx.x;
  const int64_t q = q0 + int64_t(blockIdx.x);
  const int64_t P = tri_root_f(q), Q = q - P * (P + 1) / 2;     // P >= Q
  if (P >= npairs) return;
  const DevShell A = shells[pair_a[P]], B = shells[pair_b[P]];
  const DevShell C = shells[pair_a[Q]], D = shells[pair_b[Q]];
  const int la = A.l, lb = B.l, lc = C.l, ld = D.l;
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  const int Lab = la + lb, Lcd = lc + ld, L = Lab + Lcd;
  const int nq = na * nb * nc * nd;
  for (int k = tid; k < nq; k += kEriThreads) acc[k] = 0.0;
  const double AB[3] = {A.x, A.y, A.z}, BB[3] = {B.x, B.y, B.z};
  const double CC[3] = {C.x, C.y, C.z}, DD[3] = {D.x, D.y, D.z};
  for (int i = 0; i < A.nprim; ++i)
    for (int j = 0; j < B.nprim; ++j) {
      const double a = exps[A.prim0 + i], b = exps[B.prim0 + j], p = a + b;
      const double cab = coefs[A.prim0 + i] * coefs[B.prim0 + j];
      __syncthreads();
      if (tid < 3) hermite_e(la, lb, a, b, AB[tid], BB[tid], Eab[tid]);
      for (int k = 0; k < C.nprim; ++k)
        for (int l = 0; l < D.nprim; ++l) {
          const double c = exps[C.prim0 + k], d = exps[D.prim0 + l], qq = c + d;
          const double ccd = coefs[C.prim0 + k] * coefs[D.prim0 + l];
          const double alpha = p * qq / (p + qq);
          double PQ[3];
          for (int x = 0; x < 3; ++x)
            PQ[x] = (a * AB[x] + b * BB[x]) / p - (c * CC[x] + d * DD[x]) / qq;
          __syncthreads();                  // previous quartet's readers are done
          if (tid < 3) hermite_e(lc, ld, c, d, CC[tid], DD[tid], Ecd[tid]);
          if (tid == 3) boys(L, alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]), Fb);
          if (tid == 4) sc[0] = 2.0 * 17.493418327624862846 / (p * qq * sqrt(p + qq)) * cab * ccd;  // 2 pi^{5/2}
          __syncthreads();
          // R^n_{tuv} from n = L down to 0, one level per step (levels alternate buffers)
          int cur = 0;
          if (tid == 0) {
            double f = 1.0;
            for (int m = 0; m < L; ++m) f *= -2.0 * alpha;
            Rc[0][0] = f * Fb[L];
          }
          for (int nlev = L - 1; nlev >= 0; --nlev) {
            __syncthreads();
            const int tot = L - nlev;
            const int cnt = (tot + 1) * (tot + 2) * (tot + 3) / 6;
            for (int e = tid; e < cnt; e += kEriThreads) {
              // e -> (t, u, v) with t + u + v <= tot
              int t = 0, rem = e;
              while (rem >= (tot - t + 1) * (tot - t + 2) / 2) {
                rem -= (tot - t + 1) * (tot - t + 2) / 2;
                ++t;
              }
              int u = 0;
              while (rem >= tot - t - u + 1) {
                rem -= tot - t - u + 1;
                ++u;
              }
              const int v = rem;
              const double* R1 = Rc[cur];
              double val;
              if (t > 0) {
                val = PQ[0] * R1[dcube(t - 1, u, v)];
                if (t > 1) val += (t - 1) * R1[dcube(t - 2, u, v)];
              } else if (u > 0) {
                val = PQ[1] * R1[dcube(t, u - 1, v)];
                if (u > 1) val += (u - 1) * R1[dcube(t, u - 2, v)];
              } else if (v > 0) {
                val = PQ[2] * R1[dcube(t, u, v - 1)];
                if (v > 1) val += (v - 1) * R1[dcube(t, u, v - 2)];
              } else {
                double f = 1.0;
                for (int m = 0; m < nlev; ++m) f *= -2.0 * alpha;
                val = f * Fb[nlev];
              }
              Rc[cur ^ 1][dcube(t, u, v)] = val;
            }
            cur ^= 1;
          }
          __syncthreads();
          const double* R = Rc[cur];
          const double pref = sc[0];
          const int Hab = Lab + 1, Hcd = Lcd + 1;
          for (int k2 = tid; k2 < nq; k2 += kEriThreads) {
            int rem = k2;
            const int cd_ = rem % nd; rem /= nd;
            const int cc_ = rem % nc; rem /= nc;
            const int cb_ = rem % nb; rem /= nb;
            const int ca_ = rem;
            int ax[3], bx[3], cx[3], dx[3];
            cart(la, ca_, ax[0], ax[1], ax[2]);
            cart(lb, cb_, bx[0], bx[1], bx[2]);
            cart(lc, cc_, cx[0], cx[1], cx[2]);
            cart(ld, cd_, dx[0], dx[1], dx[2]);
            const double* e1x = &Eab[0][(ax[0] * (lb + 1) + bx[0]) * Hab];
            const double* e1y = &Eab[1][(ax[1] * (lb + 1) + bx[1]) * Hab];
            const double* e1z = &Eab[2][(ax[2] * (lb + 1) + bx[2]) * Hab];
            const double* e2x = &Ecd[0][(cx[0] * (ld + 1) + dx[0]) * Hcd];
            const double* e2y = &Ecd[1][(cx[1] * (ld + 1) + dx[1]) * Hcd];
            const double* e2z = &Ecd[2][(cx[2] * (ld + 1) + dx[2]) * Hcd];
            double s = 0.0;
            for (int t = 0; t <= ax[0] + bx[0]; ++t)
              for (int u = 0; u <= ax[1] + bx[1]; ++u)
                for (int v = 0; v <= ax[2] + bx[2]; ++v) {
                  const double eb = e1x[t] * e1y[u] * e1z[v];
                  double inner = 0.0;
                  for (int t2 = 0; t2 <= cx[0] + dx[0]; ++t2)
                    for (int u2 = 0; u2 <= cx[1] + dx[1]; ++u2)
                      for (int v2 = 0; v2 <= cx[2] + dx[2]; ++v2) {
                        const double ek = e2x[t2] * e2y[u2] * e2z[v2];
                        const double r = R[dcube(t + t2, u + u2, v + v2)];
                        inner += ((t2 + u2 + v2) & 1) ? -ek * r : ek * r;
                      }
                  s += eb * inner;
                }
            acc[k2] += pref * s;
          }
        }
    }
  __syncthreads();
  // Cartesian -> spherical, one index at a time (acc -> buf -> acc -> buf -> acc)
  const int sa = nfun(la, A.cart), sb = nfun(lb, B.cart), scn = nfun(lc, C.cart), sd = nfun(ld, D.cart);
  // index d
  for (int k = tid; k < na * nb * nc * sd; k += kEriThreads) {
    const int l = k % sd, rest = k / sd;             // rest = (a, b, c)
    double v = 0.0;
    for (int c = 0; c < nd; ++c) v += tcoef(ld, D.cart, l, c) * comp_scale(ld, c) * acc[rest * nd + c];
    buf[k] = v;
  }
  __syncthreads();
  // index c: buf (a, b, c, l) -> acc (a, b, k, l)
  for (int k = tid; k < na * nb * scn * sd; k += kEriThreads) {
    const int l = k % sd, kk = (k / sd) % scn, ab = k / (sd * scn);
    double v = 0.0;
    for (int c = 0; c < nc; ++c) v += tcoef(lc, C.cart, kk, c) * comp_scale(lc, c) * buf[(ab * nc + c) * sd + l];
    acc[k] = v;
  }
  __syncthreads();
  // index b: acc (a, b, k, l) -> buf (a, j, k, l)
  for (int k = tid; k < na * sb * scn * sd; k += kEriThreads) {
    const int kl = k % (scn * sd), j = (k / (scn * sd)) % sb, a = k / (scn * sd * sb);
    double v = 0.0;
    for (int b = 0; b < nb; ++b) v += tcoef(lb, B.cart, j, b) * comp_scale(lb, b) * acc[(a * nb + b) * scn * sd + kl];
    buf[k] = v;
  }
  __syncthreads();
  // index a: buf (a, j, k, l) -> acc (i, j, k, l)
  for (int k = tid; k < sa * sb * scn * sd; k += kEriThreads) {
    const int jkl = k % (sb * scn * sd), i = k / (sb * scn * sd);
    double v = 0.0;
    for (int a = 0; a < na; ++a) v += tcoef(la, A.cart, i, a) * comp_scale(la, a) * buf[a * sb * scn * sd + jkl];
    acc[k] = v;
  }
  __syncthreads();
  // scatter with the 8-fold permutational symmetry; where shells (or shell
  // pairs) coincide only the canonical component writes, so every entry and its
  // images come from one value
  const int64_t n2 = n * n;
  const bool same_ab = pair_a[P] == pair_b[P], same_cd = pair_a[Q] == pair_b[Q], same_pq = P == Q;
  for (int k = tid; k < sa * sb * scn * sd; k += kEriThreads) {
    const int l = k % sd, kk = (k / sd) % scn, j = (k / (sd * scn)) % sb, i = k / (sd * scn * sb);
    if ((same_ab && i < j) || (same_cd && kk < l) || (same_pq && i * sb + j < kk * sd + l)) continue;
    const int64_t p = A.f0 + i, qf = B.f0 + j, r = C.f0 + kk, s = D.f0 + l;
    const double v = acc[k];
    const int64_t pq = p + n * qf, qp = qf + n * p, rs = r + n * s, sr = s + n * r;
    G[pq + n2 * rs] = v;
    G[qp + n2 * rs] = v;
    G[pq + n2 * sr] = v;
    G[qp + n2 * sr] = v;
    G[rs + n2 * pq] = v;
    G[sr + n2 * pq] = v;
    G[rs + n2 * qp] = v;
    G[sr + n2 * qp] = v;
  }
}

// ---------------------------------------------------------------------------
// pivoted Cholesky helpers
// ---------------------------------------------------------------------------

constexpr int kBlk = 256;

// d[k] = G[k + n2 k]
__global__ void diag_kernel(const double* __restrict__ G, int64_t n2, double* __restrict__ d) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n2; k += int64_t(gridDim.x) * blockDim.x)
    d[k] = G[k + n2 * k];
}

// largest d (lowest index on ties): one block
__global__ void argmax_kernel(const double* __restrict__ d, int64_t n, double* __restrict__ out_v,
                              int64_t* __restrict__ out_i) {
  __shared__ double sv[kBlk];
  __shared__ int64_t si[kBlk];
  double bv = -1.0;
  int64_t bi = 0;
  for (int64_t k = threadIdx.x; k < n; k += kBlk) {
    const double v = d[k];
    if (v > bv) {
      bv = v;
      bi = k;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = kBlk / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) {
      const double v = sv[threadIdx.x + s];
      const int64_t i = si[threadIdx.x + s];
      if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && i < si[threadIdx.x])) {
        sv[threadIdx.x] = v;
        si[threadIdx.x] = i;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out_v = sv[0];
    *out_i = si[0];
  }
}

// v = col / sqrt(dmax); d -= v^2; d[piv] = 0
__global__ void cd_update_kernel(double* __restrict__ col, double* __restrict__ d, int64_t n2, int64_t piv,
                                 double dmax) {
  const double inv = 1.0 / sqrt(dmax);
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n2; k += int64_t(gridDim.x) * blockDim.x) {
    const double v = col[k] * inv;
    col[k] = v;
    d[k] = (k == piv) ? 0.0 : d[k] - v * v;
  }
}

// L^J <- (L^J + L^J^T) / 2 for the n x n block of each vector
__global__ void symmetrise_kernel(double* __restrict__ L, int64_t n, int64_t nvec) {
  const int64_t n2 = n * n;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n2 * nvec;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t J = k / n2, r = k % n2, p = r % n, q = r / n;
    if (p >= q) continue;
    double* x = L + J * n2;
    const double v = 0.5 * (x[p + n * q] + x[q + n * p]);
    x[p + n * q] = v;
    x[q + n * p] = v;
  }
}

unsigned grid_of(int64_t n) { return unsigned(std::min<int64_t>((n + kBlk - 1) / kBlk, 1 << 16)); }

// ---------------------------------------------------------------------------
// Coulomb and exchange matrices from the dense AO integrals (eT's Hartree-Fock
// builds its Fock matrix from the AO integrals, not from Cholesky vectors).
// One workgroup per output element, a fixed summation order per thread and a
// fixed tree across the workgroup: reproducible bit for bit.  Both read the
// n^4 integrals once per build (HBM-bound).
// ---------------------------------------------------------------------------

__device__ double workgroup_sum(double v) {
  __shared__ double part[kBlk];
  part[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlk / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  return part[0];
}

// J_pq = sum_rs (pq|rs) D_rs: column pq of the symmetric n^2 x n^2 matrix against D
__global__ __launch_bounds__(kBlk) void coulomb_kernel(const double* __restrict__ G, const double* __restrict__ D,
                                                       int64_t nn, double alpha, double* __restrict__ out) {
  const int64_t pq = blockIdx.x;
  const double* col = G + pq * nn;
  double acc = 0.0;
  for (int64_t rs = threadIdx.x; rs < nn; rs += kBlk) acc += col[rs] * D[rs];
  const double sum = workgroup_sum(acc);
  if (threadIdx.x == 0) out[pq] += alpha * sum;
}

// K_pq = sum_rs (pr|qs) D_rs = sum_{r,s} (rp|sq) D_rs: for each s the n integrals
// (rp|sq), r = 0..n-1, are contiguous in the matrix
__global__ __launch_bounds__(kBlk) void exchange_kernel(const double* __restrict__ G, const double* __restrict__ D,
                                                        int64_t n, double alpha, double* __restrict__ out) {
  const int64_t pq = blockIdx.x, p = pq % n, q = pq / n, nn = n * n;
  double acc = 0.0;
  for (int64_t rs = threadIdx.x; rs < nn; rs += kBlk) {
    const int64_t r = rs % n, s = rs / n;
    acc += G[r + n * p + nn * (s + n * q)] * D[rs];
  }
  const double sum = workgroup_sum(acc);
  if (threadIdx.x == 0) out[pq] += alpha * sum;
}

}  // namespace fe

// ---------------------------------------------------------------------------
// Molecule
// ---------------------------------------------------------------------------

namespace {
void check_rocblas(rocblas_status st, const char* what) {
  if (st != rocblas_status_success) fail(std::string(what) + ": rocBLAS/rocSOLVER status " + std::to_string(int(st)));
}
}  // namespace

Molecule::Molecule(int64_t n_atoms, const double* charges, const double* xyz, int64_t n_shells,
                   const int64_t* shell_atom, const int64_t* shell_l, const int64_t* shell_nprim,
                   const double* exponents, const double* coefficients, int64_t n_electrons) {
  using namespace fe;
  if (n_atoms < 1 || n_shells < 1) fail("molecule: no atoms or shells");
  if (n_electrons < 2 || n_electrons % 2) fail("molecule: closed-shell RHF needs an even electron count");
  Z_.assign(charges, charges + n_atoms);
  xyz_.assign(xyz, xyz + 3 * n_atoms);
  int64_t f = 0, pr = 0;
  for (int64_t s = 0; s < n_shells; ++s) {
    if (shell_atom[s] < 0 || shell_atom[s] >= n_atoms) fail("molecule: shell atom index out of range");
    if (shell_l[s] < 0 || shell_l[s] > kMaxL) fail("molecule: angular momentum above d");
    if (shell_nprim[s] < 1 || shell_nprim[s] > 64) fail("molecule: primitive count out of range");
    fe::HostShell h;
    for (int d = 0; d < 3; ++d) h.c[d] = xyz[3 * shell_atom[s] + d];
    h.l = int(shell_l[s]);
    h.exps.assign(exponents + pr, exponents + pr + shell_nprim[s]);
    h.coefs.assign(coefficients + pr, coefficients + pr + shell_nprim[s]);
    for (double e : h.exps)
      if (!(e > 0.0)) fail("molecule: exponents must be positive");
    fe::normalise(h);
    h.f0 = int(f);
    f += nsph(h.l);
    pr += shell_nprim[s];
    shells_.push_back(h);
  }
  n_ = f;
  n_occ_ = n_electrons / 2;
  if (n_occ_ >= n_) fail("molecule: no virtual orbitals");
  e_nuc_ = 0.0;
  for (int64_t a = 0; a < n_atoms; ++a)
    for (int64_t b = 0; b < a; ++b) {
      const double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1], dz = xyz[3 * a + 2] - xyz[3 * b + 2];
      const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (!(r > 1e-8)) fail("molecule: coincident nuclei");
      e_nuc_ += charges[a] * charges[b] / r;
    }
  ETG_HIP(hipStreamCreate(&stream_));
}

Molecule::~Molecule() {
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Molecule::one_electron_integrals(double* S, double* T, double* V) const {
  fe::one_electron(shells_, Z_, xyz_, n_, S, T, V);
}

void Molecule::dipole_integrals(const double* origin, double* D) const {
  fe::dipole(shells_, origin, n_, D);
}

void Molecule::build_eri() {
  using namespace fe;
  const int64_t ns = int64_t(shells_.size()), n2 = n_ * n_;
  std::vector<DevShell> ds(static_cast<size_t>(ns));
  std::vector<double> ex, co;
  for (int64_t s = 0; s < ns; ++s) {
    const HostShell& h = shells_[size_t(s)];
    ds[size_t(s)] = DevShell{h.c[0], h.c[1], h.c[2], h.l, int(h.exps.size()), int(ex.size()), h.f0, h.cart};
    ex.insert(ex.end(), h.exps.begin(), h.exps.end());
    co.insert(co.end(), h.coefs.begin(), h.coefs.end());
  }
  std::vector<int32_t> pa, pb;
  for (int64_t a = 0; a < ns; ++a)
    for (int64_t b = 0; b <= a; ++b) {
      pa.push_back(int32_t(a));
      pb.push_back(int32_t(b));
    }
  const int64_t npairs = int64_t(pa.size());
  const int64_t nquart = npairs * (npairs + 1) / 2;
  DeviceBuffer dshell(sizeof(DevShell) * ds.size()), dpa(sizeof(int32_t) * pa.size()), dpb(sizeof(int32_t) * pb.size());
  DeviceBuffer dex(sizeof(double) * ex.size()), dco(sizeof(double) * co.size());
  ETG_HIP(hipMemcpy(dshell.as<void>(), ds.data(), sizeof(DevShell) * ds.size(), hipMemcpyHostToDevice));
  ETG_HIP(hipMemcpy(dpa.as<void>(), pa.data(), sizeof(int32_t) * pa.size(), hipMemcpyHostToDevice));
  ETG_HIP(hipMemcpy(dpb.as<void>(), pb.data(), sizeof(int32_t) * pb.size(), hipMemcpyHostToDevice));
  ETG_HIP(hipMemcpy(dex.as<void>(), ex.data(), sizeof(double) * ex.size(), hipMemcpyHostToDevice));
  ETG_HIP(hipMemcpy(dco.as<void>(), co.data(), sizeof(double) * co.size(), hipMemcpyHostToDevice));
  G_.resize(sizeof(double) * size_t(n2) * size_t(n2));
  // one workgroup per quartet, in launches of at most 2^25 workgroups (a
  // dispatch counts its work-items in 32 bits: 2^25 x 64 stays below 2^32);
  // ETG_ERI_LAUNCH_GROUPS lowers the limit (checks of the chunked launches)
  int64_t max_groups = int64_t(1) << 25;
  if (const char* env = std::getenv("ETG_ERI_LAUNCH_GROUPS")) {
    const long long v = std::atoll(env);
    if (v > 0) max_groups = std::min<int64_t>(max_groups, v);
  }
  for (int64_t q0 = 0; q0 < nquart; q0 += max_groups) {
    const int64_t groups = std::min(max_groups, nquart - q0);
    eri_kernel<<<unsigned(groups), kEriThreads, 0, stream_>>>(dshell.as<DevShell>(), dpa.as<int32_t>(),
                                                             dpb.as<int32_t>(), dex.as<double>(), dco.as<double>(),
                                                             npairs, n_, q0, G_.as<double>());
    ETG_LAUNCH("eri");
  }
  ETG_HIP(hipStreamSynchronize(stream_));
}

void Molecule::copy_eri(double* host) const {
  if (G_.bytes() == 0) fail("molecule: the AO integrals are not held (built and released by run)");
  ETG_HIP(hipMemcpy(host, G_.as<double>(), sizeof(double) * size_t(n_ * n_) * size_t(n_ * n_), hipMemcpyDeviceToHost));
}

void Molecule::decompose(double threshold) {
  using namespace fe;
  if (!(threshold > 0.0)) fail("molecule: the Cholesky threshold must be positive");
  const int64_t n2 = n_ * n_;
  const double* G = G_.as<double>();
  DeviceBuffer d(sizeof(double) * n2), rv(sizeof(double)), ri(sizeof(int64_t));
  diag_kernel<<<grid_of(n2), kBlk, 0, stream_>>>(G, n2, d.as<double>());
  ETG_LAUNCH("cd_diag");
  int64_t cap = std::min<int64_t>(n2, std::max<int64_t>(64, 16 * n_));
  DeviceBuffer Lv(sizeof(double) * n2 * cap);
  GemmWorkspace ws;
  n_J_ = 0;
  for (;;) {
    argmax_kernel<<<1, kBlk, 0, stream_>>>(d.as<double>(), n2, rv.as<double>(), ri.as<int64_t>());
    ETG_LAUNCH("cd_argmax");
    double dmax = 0.0;
    int64_t piv = 0;
    ETG_HIP(hipMemcpyAsync(&dmax, rv.as<double>(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    ETG_HIP(hipMemcpyAsync(&piv, ri.as<int64_t>(), sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    ETG_HIP(hipStreamSynchronize(stream_));
    if (!(dmax > threshold) || n_J_ == n2) break;
    if (n_J_ == cap) {                      // grow the vector store
      const int64_t ncap = std::min<int64_t>(n2, 2 * cap);
      DeviceBuffer nL(sizeof(double) * n2 * ncap);
      ETG_HIP(hipMemcpyAsync(nL.as<double>(), Lv.as<double>(), sizeof(double) * n2 * cap, hipMemcpyDeviceToDevice, stream_));
      ETG_HIP(hipStreamSynchronize(stream_));
      Lv = std::move(nL);
      cap = ncap;
    }
    double* col = Lv.as<double>() + n2 * n_J_;
    ETG_HIP(hipMemcpyAsync(col, G + n2 * piv, sizeof(double) * n2, hipMemcpyDeviceToDevice, stream_));
    if (n_J_ > 0)   // col -= L(:, 0:J) L(piv, 0:J)^T
      etg::gemm(stream_, 'N', 'T', n2, 1, n_J_, -1.0, Lv.as<double>(), n2, Lv.as<double>() + piv, n2, 1.0, col, n2,
                Prec::F64, &ws);
    cd_update_kernel<<<grid_of(n2), kBlk, 0, stream_>>>(col, d.as<double>(), n2, piv, dmax);
    ETG_LAUNCH("cd_update");
    ++n_J_;
  }
  symmetrise_kernel<<<grid_of(n2 * n_J_), kBlk, 0, stream_>>>(Lv.as<double>(), n_, n_J_);
  ETG_LAUNCH("cd_symmetrise");
  Lao_.resize(sizeof(double) * n2 * std::max<int64_t>(1, n_J_));
  ETG_HIP(hipMemcpyAsync(Lao_.as<double>(), Lv.as<double>(), sizeof(double) * n2 * n_J_, hipMemcpyDeviceToDevice, stream_));
  ETG_HIP(hipStreamSynchronize(stream_));
}

// Symmetric eigen-decomposition on the device: A (n x n) -> eigenvectors, w ascending.
namespace {
void syevd(rocblas_handle h, double* A, int64_t n, double* w, DeviceBuffer& E, DeviceBuffer& info) {
  E.resize(sizeof(double) * n);
  info.resize(sizeof(rocblas_int));
  check_rocblas(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, rocblas_int(n), A, rocblas_int(n), w,
                                 E.as<double>(), info.as<rocblas_int>()),
                "dsyevd");
}

__global__ void scale_columns_kernel(double* U, const double* w, int64_t n) {   // U(:, j) /= sqrt(w_j)
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n * n; k += int64_t(gridDim.x) * blockDim.x)
    U[k] /= sqrt(w[k / n]);
}

__global__ void fock_kernel(const double* H, const double* J, const double* K, double* F, int64_t nn) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nn; k += int64_t(gridDim.x) * blockDim.x)
    F[k] = H[k] + 2.0 * J[k] - K[k];
}
}  // namespace

void Molecule::scf(const double* S_host, const double* H_host, double conv, int64_t max_iterations) {
  using namespace fe;
  const int64_t n = n_, nn = n * n, o = n_occ_, J = n_J_;
  const bool exact_jk = G_.bytes() != 0;
  rocblas_handle rh;
  check_rocblas(rocblas_create_handle(&rh), "rocblas_create_handle");
  struct Guard {
    rocblas_handle h;
    ~Guard() { (void)rocblas_destroy_handle(h); }
  } guard{rh};
  check_rocblas(rocblas_set_stream(rh, stream_), "rocblas_set_stream");
  DeviceBuffer S(sizeof(double) * nn), H(sizeof(double) * nn), X(sizeof(double) * nn), w(sizeof(double) * n);
  DeviceBuffer F(sizeof(double) * nn), Fp(sizeof(double) * nn), Tm(sizeof(double) * nn), Cd(sizeof(double) * nn);
  DeviceBuffer D(sizeof(double) * nn), Jm(sizeof(double) * nn), Km(sizeof(double) * nn), u(sizeof(double) * std::max<int64_t>(1, J));
  DeviceBuffer Y(sizeof(double) * n * o * std::max<int64_t>(1, J)), E, info, err(sizeof(double) * nn);
  GemmWorkspace ws;
  ETG_HIP(hipMemcpyAsync(S.as<double>(), S_host, sizeof(double) * nn, hipMemcpyHostToDevice, stream_));
  ETG_HIP(hipMemcpyAsync(H.as<double>(), H_host, sizeof(double) * nn, hipMemcpyHostToDevice, stream_));
  // X = U s^{-1/2}
  ETG_HIP(hipMemcpyAsync(X.as<double>(), S.as<double>(), sizeof(double) * nn, hipMemcpyDeviceToDevice, stream_));
  syevd(rh, X.as<double>(), n, w.as<double>(), E, info);
  scale_columns_kernel<<<grid_of(nn), kBlk, 0, stream_>>>(X.as<double>(), w.as<double>(), n);
  ETG_LAUNCH("scf_x");
  auto mm = [&](char ta, char tb, const double* A, const double* B, double* C, double alpha, double beta) {
    etg::gemm(stream_, ta, tb, n, n, n, alpha, A, n, B, n, beta, C, n, Prec::F64, &ws);
  };
  auto diag_fock = [&]() {   // Fp = X^T F X, eigenvectors -> C = X Cp, eps
    mm('N', 'N', F.as<double>(), X.as<double>(), Tm.as<double>(), 1.0, 0.0);
    mm('T', 'N', X.as<double>(), Tm.as<double>(), Fp.as<double>(), 1.0, 0.0);
    syevd(rh, Fp.as<double>(), n, w.as<double>(), E, info);
    mm('N', 'N', X.as<double>(), Fp.as<double>(), Cd.as<double>(), 1.0, 0.0);
  };
  ETG_HIP(hipMemcpyAsync(F.as<double>(), H.as<double>(), sizeof(double) * nn, hipMemcpyDeviceToDevice, stream_));
  std::vector<std::vector<double>> errs, focks;
  std::vector<double> hF(nn), hD(nn), hH(H_host, H_host + nn), hErr(nn);
  double e_old = 0.0;
  const double* Lao = Lao_.as<double>();
  if (given_orbitals_)
    ETG_HIP(hipMemcpyAsync(Cd.as<double>(), C_.data(), sizeof(double) * nn, hipMemcpyHostToDevice, stream_));
  for (int64_t it = 0; it < max_iterations; ++it) {
    if (!given_orbitals_) diag_fock();
    // D = C_occ C_occ^T
    etg::gemm(stream_, 'N', 'T', n, n, o, 1.0, Cd.as<double>(), n, Cd.as<double>(), n, 0.0, D.as<double>(), n,
              Prec::F64, &ws);
    if (exact_jk) {
      // J_pq = sum_rs (pq|rs) D_rs,  K_pq = sum_rs (pr|qs) D_rs  from the AO integrals
      fill(stream_, Jm.as<double>(), nn, 0.0);
      fill(stream_, Km.as<double>(), nn, 0.0);
      coulomb_kernel<<<unsigned(nn), kBlk, 0, stream_>>>(G_.as<double>(), D.as<double>(), nn, 1.0, Jm.as<double>());
      ETG_LAUNCH("scf_coulomb");
      exchange_kernel<<<unsigned(nn), kBlk, 0, stream_>>>(G_.as<double>(), D.as<double>(), n, 1.0, Km.as<double>());
      ETG_LAUNCH("scf_exchange");
    } else {
      // J = sum_J L^J (L^J . D)
      etg::gemm(stream_, 'T', 'N', J, 1, nn, 1.0, Lao, nn, D.as<double>(), nn, 0.0, u.as<double>(), J, Prec::F64, &ws);
      etg::gemm(stream_, 'N', 'N', nn, 1, J, 1.0, Lao, nn, u.as<double>(), J, 0.0, Jm.as<double>(), nn, Prec::F64, &ws);
      // K = sum_J (L^J C_o)(L^J C_o)^T
      Gemm g = make_gemm('N', 'N', n, o, n, 1.0, Lao, n, Cd.as<double>(), n, 0.0, Y.as<double>(), n);
      g.batch = J;
      g.strideA = nn;
      g.strideB = 0;
      g.strideC = n * o;
      etg::gemm(stream_, g, Prec::F64, &ws);
      etg::gemm(stream_, 'N', 'T', n, n, o * J, 1.0, Y.as<double>(), n, Y.as<double>(), n, 0.0, Km.as<double>(), n,
                Prec::F64, &ws);
    }
    fock_kernel<<<grid_of(nn), kBlk, 0, stream_>>>(H.as<double>(), Jm.as<double>(), Km.as<double>(), F.as<double>(), nn);
    ETG_LAUNCH("scf_fock");
    // err = F D S - S D F
    mm('N', 'N', D.as<double>(), S.as<double>(), Tm.as<double>(), 1.0, 0.0);
    mm('N', 'N', F.as<double>(), Tm.as<double>(), err.as<double>(), 1.0, 0.0);
    mm('N', 'N', D.as<double>(), F.as<double>(), Tm.as<double>(), 1.0, 0.0);
    mm('N', 'N', S.as<double>(), Tm.as<double>(), err.as<double>(), -1.0, 1.0);
    ETG_HIP(hipMemcpyAsync(hF.data(), F.as<double>(), sizeof(double) * nn, hipMemcpyDeviceToHost, stream_));
    ETG_HIP(hipMemcpyAsync(hD.data(), D.as<double>(), sizeof(double) * nn, hipMemcpyDeviceToHost, stream_));
    ETG_HIP(hipMemcpyAsync(hErr.data(), err.as<double>(), sizeof(double) * nn, hipMemcpyDeviceToHost, stream_));
    ETG_HIP(hipStreamSynchronize(stream_));
    double e = e_nuc_, emax = 0.0;
    for (int64_t k = 0; k < nn; ++k) {
      e += hD[k] * (hH[k] + hF[k]);
      emax = std::max(emax, std::fabs(hErr[k]));
    }
    ++scf_iterations_;
    if (given_orbitals_) {   // skipped SCF: the energy of the given orbitals, which must be converged
      if (!(emax < conv)) fail("scf: cannot skip scf when gradient has not converged");
      e_hf_ = e;
      h_ao_.assign(hH.begin(), hH.end());
      return;
    }
    if (emax < conv && std::fabs(e - e_old) < 1e-13) {
      e_hf_ = e;
      diag_fock();
      ETG_HIP(hipStreamSynchronize(stream_));
      C_.assign(size_t(nn), 0.0);
      eps_.assign(size_t(n), 0.0);
      ETG_HIP(hipMemcpy(C_.data(), Cd.as<double>(), sizeof(double) * nn, hipMemcpyDeviceToHost));
      ETG_HIP(hipMemcpy(eps_.data(), w.as<double>(), sizeof(double) * n, hipMemcpyDeviceToHost));
      h_ao_.assign(hH.begin(), hH.end());
      return;
    }
    e_old = e;
    // DIIS over the last 8 Fock matrices (host: n x n)
    errs.push_back(hErr);
    focks.push_back(hF);
    if (errs.size() > 8) {
      errs.erase(errs.begin());
      focks.erase(focks.begin());
    }
    const int m = int(errs.size());
    std::vector<double> B(static_cast<size_t>((m + 1) * (m + 1)), -1.0), rhs(static_cast<size_t>(m + 1), 0.0);
    B[size_t(m * (m + 1) + m)] = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int64_t k = 0; k < nn; ++k) s += errs[size_t(i)][size_t(k)] * errs[size_t(j)][size_t(k)];
        B[size_t(i * (m + 1) + j)] = s;
      }
    rhs[size_t(m)] = -1.0;
    // Gaussian elimination with partial pivoting on the (m+1) system (row-major B)
    const int N = m + 1;
    for (int c = 0; c < N; ++c) {
      int p = c;
      for (int r = c + 1; r < N; ++r)
        if (std::fabs(B[size_t(r * N + c)]) > std::fabs(B[size_t(p * N + c)])) p = r;
      if (B[size_t(p * N + c)] == 0.0) fail("scf: singular DIIS matrix");
      if (p != c) {
        for (int k = 0; k < N; ++k) std::swap(B[size_t(c * N + k)], B[size_t(p * N + k)]);
        std::swap(rhs[size_t(c)], rhs[size_t(p)]);
      }
      for (int r = c + 1; r < N; ++r) {
        const double fct = B[size_t(r * N + c)] / B[size_t(c * N + c)];
        for (int k = c; k < N; ++k) B[size_t(r * N + k)] -= fct * B[size_t(c * N + k)];
        rhs[size_t(r)] -= fct * rhs[size_t(c)];
      }
    }
    std::vector<double> x(static_cast<size_t>(N));
    for (int r = N - 1; r >= 0; --r) {
      double s = rhs[size_t(r)];
      for (int k = r + 1; k < N; ++k) s -= B[size_t(r * N + k)] * x[size_t(k)];
      x[size_t(r)] = s / B[size_t(r * N + r)];
    }
    std::fill(hF.begin(), hF.end(), 0.0);
    for (int i = 0; i < m; ++i)
      for (int64_t k = 0; k < nn; ++k) hF[size_t(k)] += x[size_t(i)] * focks[size_t(i)][size_t(k)];
    ETG_HIP(hipMemcpyAsync(F.as<double>(), hF.data(), sizeof(double) * nn, hipMemcpyHostToDevice, stream_));
    ETG_HIP(hipStreamSynchronize(stream_));
  }
  fail("scf: RHF did not converge in the maximum number of iterations");
}

void Molecule::run(double cholesky_threshold, double scf_threshold, int64_t max_scf_iterations) {
  const int64_t nn = n_ * n_;
  std::vector<double> S(static_cast<size_t>(nn)), T(static_cast<size_t>(nn)), V(static_cast<size_t>(nn)), H(static_cast<size_t>(nn));
  one_electron_integrals(S.data(), T.data(), V.data());
  for (int64_t k = 0; k < nn; ++k) H[size_t(k)] = T[size_t(k)] + V[size_t(k)];
  build_eri();
  decompose(cholesky_threshold);
  scf(S.data(), H.data(), scf_threshold, max_scf_iterations);   // J and K from the dense integrals
  mo_transform();
  G_.release();                         // the MO Cholesky vectors replace the dense matrix
  done_ = true;
}

void Molecule::set_orbitals(const double* C, const double* eps) {
  if (!C || !eps) fail("set_orbitals: null pointer");
  const size_t nn = size_t(n_) * size_t(n_);
  C_.assign(C, C + nn);
  eps_.assign(eps, eps + size_t(n_));
  given_orbitals_ = true;
}

void Molecule::set_cartesian(bool cartesian) {
  if (done_ || G_.bytes() != 0 || n_J_ != 0) fail("set_cartesian: the integrals have already been built");
  int64_t f = 0;
  for (fe::HostShell& h : shells_) {
    h.cart = cartesian;
    h.f0 = int(f);
    f += fe::nfun(h.l, cartesian);
  }
  n_ = f;
  if (n_occ_ >= n_) fail("molecule: no virtual orbitals");
}

void Molecule::set_frozen_core(int64_t n_frozen) {
  if (n_frozen < 0 || n_frozen >= n_occ_) fail("set_frozen_core: the number of frozen orbitals must be in [0, n_occ)");
  n_frozen_core_ = n_frozen;
}

void Molecule::mo_transform() {
  using namespace fe;
  const int64_t n = n_, nn = n * n, J = n_J_;
  GemmWorkspace ws;
  DeviceBuffer C(sizeof(double) * nn), T(sizeof(double) * nn * std::max<int64_t>(1, J)), M(sizeof(double) * nn * std::max<int64_t>(1, J));
  ETG_HIP(hipMemcpyAsync(C.as<double>(), C_.data(), sizeof(double) * nn, hipMemcpyHostToDevice, stream_));
  // T^J = L^J C, M^J = C^T T^J  (batched over J)
  Gemm g = make_gemm('N', 'N', n, n, n, 1.0, Lao_.as<double>(), n, C.as<double>(), n, 0.0, T.as<double>(), n);
  g.batch = J; g.strideA = nn; g.strideB = 0; g.strideC = nn;
  etg::gemm(stream_, g, Prec::F64, &ws);
  g = make_gemm('T', 'N', n, n, n, 1.0, C.as<double>(), n, T.as<double>(), n, 0.0, M.as<double>(), n);
  g.batch = J; g.strideA = 0; g.strideB = nn; g.strideC = nn;
  etg::gemm(stream_, g, Prec::F64, &ws);
  // (p, q, J) -> (J, p, q)
  Lmo_.resize(sizeof(double) * nn * std::max<int64_t>(1, J));
  const int64_t dims[4] = {n, n, J, 1};
  sort4(stream_, M.as<double>(), dims, "3124", Lmo_.as<double>(), 1.0, 0.0);
  // h_mo = C^T h C
  DeviceBuffer h(sizeof(double) * nn), t(sizeof(double) * nn), hm(sizeof(double) * nn);
  ETG_HIP(hipMemcpyAsync(h.as<double>(), h_ao_.data(), sizeof(double) * nn, hipMemcpyHostToDevice, stream_));
  if (n_frozen_core_ > 0) {
    // h += 2 J[D_c] - K[D_c],  D_c = C_c C_c^T over the frozen MOs (construct_mo_fock_fc_term_hf,
    // frozen_orbital_hf.F90:547-600), from the AO integrals as in the RHF Fock build
    const int64_t c = n_frozen_core_;
    DeviceBuffer D(sizeof(double) * nn);
    etg::gemm(stream_, 'N', 'T', n, n, c, 1.0, C.as<double>(), n, C.as<double>(), n, 0.0, D.as<double>(), n,
              Prec::F64, &ws);
    if (G_.bytes() == 0) fail("frozen core: the AO integrals are no longer held");
    coulomb_kernel<<<unsigned(nn), kBlk, 0, stream_>>>(G_.as<double>(), D.as<double>(), nn, 2.0, h.as<double>());
    ETG_LAUNCH("frozen_core_coulomb");
    exchange_kernel<<<unsigned(nn), kBlk, 0, stream_>>>(G_.as<double>(), D.as<double>(), n, -1.0, h.as<double>());
    ETG_LAUNCH("frozen_core_exchange");
  }
  etg::gemm(stream_, 'N', 'N', n, n, n, 1.0, h.as<double>(), n, C.as<double>(), n, 0.0, t.as<double>(), n, Prec::F64, &ws);
  etg::gemm(stream_, 'T', 'N', n, n, n, 1.0, C.as<double>(), n, t.as<double>(), n, 0.0, hm.as<double>(), n, Prec::F64, &ws);
  h_mo_.assign(size_t(nn), 0.0);
  ETG_HIP(hipMemcpyAsync(h_mo_.data(), hm.as<double>(), sizeof(double) * nn, hipMemcpyDeviceToHost, stream_));
  ETG_HIP(hipStreamSynchronize(stream_));
  Lao_.release();
}

void Molecule::results(double* e_nuc, double* e_hf, double* eps, double* h_mo, double* L_mo, double* C) const {
  if (!done_) fail("molecule: not run");
  const int64_t nn = n_ * n_;
  if (e_nuc) *e_nuc = e_nuc_;
  if (e_hf) *e_hf = e_hf_;
  if (eps) std::memcpy(eps, eps_.data(), sizeof(double) * size_t(n_));
  if (h_mo) std::memcpy(h_mo, h_mo_.data(), sizeof(double) * size_t(nn));
  if (C) std::memcpy(C, C_.data(), sizeof(double) * size_t(nn));
  if (L_mo)
    ETG_HIP(hipMemcpy(L_mo, Lmo_.as<double>(), sizeof(double) * size_t(nn) * size_t(n_J_), hipMemcpyDefault));
}

}  // namespace etg
