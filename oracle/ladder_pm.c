/*
 * Gathers and scatter of the CCSD ladder term's +/- combinations, for the
 * oracle's batched ladder (oracle/ccsd.py omega_a2_batched).  TEST
 * INFRASTRUCTURE ONLY: part of the CPU restatement of eT's algorithm that
 * checks the HIP implementation and times eT's factorisation on the host.
 *
 * eT forms, per (a-batch, b-batch) block of g_acbd (omega_ccsd.F90:292-458,
 * construct_plus_minus_1324_from_full / _RFP in packed_array_utilities_r.F90),
 *   g+_{cd,ab} = g_acbd + g_adbc (c <= d),   g-_{cd,ab} = g_acbd - g_adbc (c < d),
 * multiplies them with t+/- and adds the products into the packed residual
 * (add_double_packed_plus_minus, packed_array_utilities_r.F90:335-399):
 *   omega(ai, bj) += w+ + w-,  omega(aj, bi) += w+ - w-   (a <= b, i <= j).
 * Conventions here: pairs (a <= b) and (c <= d) in the packed order
 * b(b+1)/2 + a; minus pairs (c < d) in the order d(d-1)/2 + c.
 *
 * Built by oracle/build.py (gcc -O3 -fopenmp) into oracle/_build/libladder_pm.so.
 */
#include <stdint.h>

/* G: block of g_acbd for a in [a0, a0 + la), b in [b0, b0 + lb), row-major
 * (la v) x (lb v) with G[(a - a0) v + c][(b - b0) v + d].  For each pair p
 * (a = pa[p], b = pb[p], global, a <= b, a in the a-range, b in the b-range)
 * writes column p of gp (v(v+1)/2 rows) and gm (v(v-1)/2 rows). */
void ladder_gather_pm(const double* G, int64_t v, int64_t a0, int64_t la, int64_t b0, int64_t lb,
                      int64_t npairs, const int64_t* pa, const int64_t* pb, double* gp, double* gm) {
  const int64_t vp = v * (v + 1) / 2, vm = v * (v - 1) / 2, ld = lb * v;
  (void)la;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < npairs; ++p) {
    const int64_t ra = (pa[p] - a0) * v, rb = (pb[p] - b0) * v;
    double* colp = gp + p * vp;
    double* colm = gm + p * vm;
    for (int64_t d = 0; d < v; ++d) {
      for (int64_t c = 0; c <= d; ++c) {
        const double acbd = G[(ra + c) * ld + rb + d];
        const double adbc = G[(ra + d) * ld + rb + c];
        colp[d * (d + 1) / 2 + c] = acbd + adbc;
        if (c < d) colm[d * (d - 1) / 2 + c] = acbd - adbc;
      }
    }
  }
}

/* omega2 (packed, length n(n+1)/2 with n = v o) += the ladder term from
 * wp (op x vp, column-major: row ij = j(j+1)/2 + i, i <= j; column ab) and
 * wm (om x vm: row j(j-1)/2 + i, i < j; column b(b-1)/2 + a, a < b). */
void ladder_scatter_pm(const double* wp, const double* wm, int64_t o, int64_t v, double* omega2) {
  const int64_t op = o * (o + 1) / 2, om = o * (o - 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < v; ++b) {
    for (int64_t a = 0; a <= b; ++a) {
      const int64_t cp = b * (b + 1) / 2 + a, cm = b * (b - 1) / 2 + a;
      for (int64_t j = 0; j < o; ++j) {
        for (int64_t i = 0; i <= j; ++i) {
          const double p = wp[cp * op + j * (j + 1) / 2 + i];
          const double m = (a < b && i < j) ? wm[cm * om + j * (j - 1) / 2 + i] : 0.0;
          const int64_t ai = a + v * i, bj = b + v * j, aj = a + v * j, bi = b + v * i;
          const int64_t x1 = ai > bj ? ai : bj, y1 = ai > bj ? bj : ai;
          const int64_t x2 = aj > bi ? aj : bi, y2 = aj > bi ? bi : aj;
          omega2[x1 * (x1 + 1) / 2 + y1] += p + m;
          omega2[x2 * (x2 + 1) / 2 + y2] += p - m;
        }
      }
    }
  }
}

/* ladder_scatter_pm for one block of pairs: wp (op x npairs) and wm (om x npairs),
 * row-major, column p for the pair (pa[p], pb[p]), a <= b; the a = b columns of wp
 * carry the factor 1/2 here; wm is read for a < b only. */
void ladder_scatter_block(const double* wp, const double* wm, int64_t npairs, const int64_t* pa,
                          const int64_t* pb, int64_t o, int64_t v, double* omega2) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < npairs; ++q) {
    const int64_t a = pa[q], b = pb[q];
    const double f = (a == b) ? 0.5 : 1.0;
    for (int64_t j = 0; j < o; ++j) {
      for (int64_t i = 0; i <= j; ++i) {
        const double p = f * wp[(j * (j + 1) / 2 + i) * npairs + q];
        const double m = (a < b && i < j) ? wm[(j * (j - 1) / 2 + i) * npairs + q] : 0.0;
        const int64_t ai = a + v * i, bj = b + v * j, aj = a + v * j, bi = b + v * i;
        const int64_t x1 = ai > bj ? ai : bj, y1 = ai > bj ? bj : ai;
        const int64_t x2 = aj > bi ? aj : bi, y2 = aj > bi ? bi : aj;
        omega2[x1 * (x1 + 1) / 2 + y1] += p + m;
        omega2[x2 * (x2 + 1) / 2 + y2] += p - m;
      }
    }
  }
}

/* dst (C order, shape n0 x n1 x n2 x n3) = src at element strides s0..s3. */
void permute4_copy(const double* src, const int64_t* shape, const int64_t* strides, double* dst) {
  const int64_t n0 = shape[0], n1 = shape[1], n2 = shape[2], n3 = shape[3];
  const int64_t s0 = strides[0], s1 = strides[1], s2 = strides[2], s3 = strides[3];
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t i0 = 0; i0 < n0; ++i0)
    for (int64_t i1 = 0; i1 < n1; ++i1) {
      double* d = dst + (i0 * n1 + i1) * n2 * n3;
      const double* s = src + i0 * s0 + i1 * s1;
      for (int64_t i2 = 0; i2 < n2; ++i2)
        for (int64_t i3 = 0; i3 < n3; ++i3) d[i2 * n3 + i3] = s[i2 * s2 + i3 * s3];
    }
}

/* full (n x n, C order) from the packed symmetric storage p[max(i,j)(max+1)/2 + min]. */
void sym_unpack(const double* p, int64_t n, double* full) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) {
      const int64_t x = i > j ? i : j, y = i > j ? j : i;
      full[i * n + j] = p[x * (x + 1) / 2 + y];
    }
}

/* packed[i(i+1)/2 + j] = full[i][j] for i >= j (full n x n, C order). */
void sym_pack_lower(const double* full, int64_t n, double* packed) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j <= i; ++j) packed[i * (i + 1) / 2 + j] = full[i * n + j];
}
