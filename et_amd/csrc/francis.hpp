// Eigenvalues of a real upper Hessenberg matrix by the implicitly
// double-shifted QR iteration (J. G. F. Francis, Comput. J. 4 (1961) 265-271;
// the step and the deflation loop as stated in Golub & Van Loan, Matrix
// Computations, 4th ed., Algorithms 7.5.1 and 7.5.2).  Only the eigenvalues
// are wanted, so each similarity transformation is applied to the active
// diagonal block alone.
//
// The routine is written once for a team of lanes: on the device the team is
// one wavefront (64 lanes, scalar control computed redundantly by every lane
// from the same LDS values, row and column sweeps split across lanes, a
// barrier between dependent phases); on the host it is one lane with no-op
// barriers, so the same code is checked against LAPACK from the CPU.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define ETG_HD __host__ __device__
#else
#define ETG_HD
#endif

namespace etg {
namespace francis {

// Reflector I - tau v v^T with v = (1, v1, v2) that maps a (length nr = 2 or 3)
// onto alpha e1.  tau = 0 when a is already a multiple of e1.
struct Reflector {
  double tau, v1, v2, alpha;
};

ETG_HD inline Reflector make_reflector(double a0, double a1, double a2, int nr) {
  Reflector r;
  const double tail = nr == 3 ? std::sqrt(a1 * a1 + a2 * a2) : std::fabs(a1);
  if (tail == 0.0) {
    r.tau = 0.0;
    r.v1 = r.v2 = 0.0;
    r.alpha = a0;
    return r;
  }
  const double len = std::sqrt(a0 * a0 + tail * tail);
  const double beta = a0 >= 0.0 ? -len : len;            // opposite sign of a0: no cancellation
  r.tau = (beta - a0) / beta;
  const double inv = 1.0 / (a0 - beta);
  r.v1 = a1 * inv;
  r.v2 = nr == 3 ? a2 * inv : 0.0;
  r.alpha = beta;
  return r;
}

// Eigenvalues of the 2 x 2 block [[a, b], [c, d]]: real pair (w1, w2, im = 0)
// or complex pair re +- i im (im > 0).
ETG_HD inline void block2_eigenvalues(double a, double b, double c, double d, double* re1, double* re2,
                                      double* im) {
  const double p = 0.5 * (a - d);
  const double disc = p * p + b * c;
  if (disc >= 0.0) {
    const double root = std::sqrt(disc);
    const double q = p >= 0.0 ? p + root : p - root;      // |q| = |p| + root
    if (q == 0.0) {
      *re1 = *re2 = d;
    } else {
      *re1 = d + q;
      *re2 = d - b * c / q;                               // (p + r)(p - r) = -bc
    }
    *im = 0.0;
  } else {
    *re1 = *re2 = d + p;
    *im = std::sqrt(-disc);
  }
}

// Team of lanes: rank, size, barrier.
struct HostTeam {
  int rank() const { return 0; }
  int size() const { return 1; }
  void sync() const {}
};

// Eigenvalues of the k x k upper Hessenberg matrix H (column-major, leading
// dimension ldh; overwritten) into wr, wi (complex pairs as re +- i im with the
// negative imaginary part first).  Returns 0, or 1 when an eigenvalue needs
// more than kMaxSweeps double-shift steps.  *hnorm receives the sum of the
// absolute values of H's entries on entry (the scale of the small-subdiagonal
// test and of the later inverse iteration).
template <class Team>
ETG_HD int hessenberg_eigenvalues(const Team& team, double* H, int ldh, int k, double* wr, double* wi,
                                  double* hnorm) {
  constexpr int kMaxSweeps = 80;
  constexpr double kEps = 2.220446049250313e-16;
#define ETG_H(i, j) H[(i) + (j) * ldh]
  const int me = team.rank(), lanes = team.size();
  double scale = 0.0;
  for (int j = 0; j < k; ++j)
    for (int i = 0; i <= (j + 1 < k ? j + 1 : k - 1); ++i) scale += std::fabs(ETG_H(i, j));
  *hnorm = scale;
  if (scale == 0.0) scale = 1.0;
  int hi = k - 1;             // the active block ends at row/column hi
  int sweeps = 0;             // double-shift steps since the last deflation
  while (hi >= 0) {
    // smallest lo such that H(lo:hi, lo:hi) has no negligible subdiagonal entry
    int lo = hi;
    while (lo > 0) {
      double ref = std::fabs(ETG_H(lo, lo)) + std::fabs(ETG_H(lo - 1, lo - 1));
      if (ref == 0.0) ref = scale;
      if (std::fabs(ETG_H(lo, lo - 1)) <= kEps * ref) break;
      --lo;
    }
    team.sync();
    if (lo > 0 && me == 0) ETG_H(lo, lo - 1) = 0.0;
    team.sync();
    if (lo == hi) {                                       // one eigenvalue deflates
      if (me == 0) {
        wr[hi] = ETG_H(hi, hi);
        wi[hi] = 0.0;
      }
      hi -= 1;
      sweeps = 0;
      team.sync();
      continue;
    }
    if (lo == hi - 1) {                                   // a 2 x 2 block deflates
      double r1, r2, im;
      block2_eigenvalues(ETG_H(hi - 1, hi - 1), ETG_H(hi - 1, hi), ETG_H(hi, hi - 1), ETG_H(hi, hi), &r1, &r2,
                         &im);
      if (me == 0) {
        if (im == 0.0) {
          wr[hi - 1] = r1;
          wr[hi] = r2;
          wi[hi - 1] = wi[hi] = 0.0;
        } else {
          wr[hi - 1] = wr[hi] = r1;
          wi[hi - 1] = -im;
          wi[hi] = im;
        }
      }
      hi -= 2;
      sweeps = 0;
      team.sync();
      continue;
    }
    if (sweeps == kMaxSweeps) return 1;
    ++sweeps;
    // shifts: the eigenvalues of the trailing 2 x 2 block (sum s, product p);
    // every 10th step two unequal real shifts h + xi and h - xi / 2 off the
    // trailing entries instead.  They break the cycles of the standard shift
    // (a cyclic permutation matrix never deflates with it), and being unequal
    // they do not tie on eigenvalues placed symmetrically about h.
    // First column of (H - s1 I)(H - s2 I) on the block, written with differences
    // of diagonal entries so that it does not cancel when the block is close to
    // a multiple of I (a converging multiple eigenvalue): with the trailing
    // block [[a, b], [c, d]], (h00 - s1)(h00 - s2) = (h00 - a)(h00 - d) - b c.
    const double h00 = ETG_H(lo, lo), h10 = ETG_H(lo + 1, lo), h01 = ETG_H(lo, lo + 1);
    const double h11 = ETG_H(lo + 1, lo + 1);
    double x, y;
    if (sweeps % 10 == 0) {
      const double h = ETG_H(hi, hi);
      const double xi = std::fabs(ETG_H(hi, hi - 1)) + std::fabs(ETG_H(hi - 1, hi - 2));
      const double e = h00 - h;                           // shifts h + xi, h - xi / 2
      x = (e - xi) * (e + 0.5 * xi) + h01 * h10;
      y = h10 * (e + (h11 - h) - 0.5 * xi);
    } else {
      const double a = ETG_H(hi - 1, hi - 1), b = ETG_H(hi - 1, hi), c = ETG_H(hi, hi - 1), d = ETG_H(hi, hi);
      x = (h00 - a) * (h00 - d) - b * c + h01 * h10;
      y = h10 * ((h00 - a) + (h11 - d));
    }
    double z = h10 * ETG_H(lo + 2, lo + 1);
    const double sz = std::fabs(x) + std::fabs(y) + std::fabs(z);
    if (sz != 0.0) {
      x /= sz;
      y /= sz;
      z /= sz;
    }
    // chase the bulge down the block
    for (int c = lo; c < hi; ++c) {
      const int nr = hi - c + 1 < 3 ? hi - c + 1 : 3;
      double a0 = x, a1 = y, a2 = z;
      if (c > lo) {
        a0 = ETG_H(c, c - 1);
        a1 = ETG_H(c + 1, c - 1);
        a2 = nr == 3 ? ETG_H(c + 2, c - 1) : 0.0;
      }
      const Reflector r = make_reflector(a0, a1, a2, nr);
      team.sync();                                        // every lane has read the column
      if (r.tau == 0.0) continue;
      if (c > lo && me == 0) {
        ETG_H(c, c - 1) = r.alpha;
        ETG_H(c + 1, c - 1) = 0.0;
        if (nr == 3) ETG_H(c + 2, c - 1) = 0.0;
      }
      // rows c .. c + nr - 1, columns c .. hi, from the left
      for (int j = c + me; j <= hi; j += lanes) {
        double w = ETG_H(c, j) + r.v1 * ETG_H(c + 1, j);
        if (nr == 3) w += r.v2 * ETG_H(c + 2, j);
        w *= r.tau;
        ETG_H(c, j) -= w;
        ETG_H(c + 1, j) -= w * r.v1;
        if (nr == 3) ETG_H(c + 2, j) -= w * r.v2;
      }
      team.sync();
      // columns c .. c + nr - 1, rows lo .. min(c + 3, hi), from the right
      const int last = c + 3 < hi ? c + 3 : hi;
      for (int i = lo + me; i <= last; i += lanes) {
        double w = ETG_H(i, c) + r.v1 * ETG_H(i, c + 1);
        if (nr == 3) w += r.v2 * ETG_H(i, c + 2);
        w *= r.tau;
        ETG_H(i, c) -= w;
        ETG_H(i, c + 1) -= w * r.v1;
        if (nr == 3) ETG_H(i, c + 2) -= w * r.v2;
      }
      team.sync();
    }
  }
  return 0;
#undef ETG_H
}

}  // namespace francis
}  // namespace etg
