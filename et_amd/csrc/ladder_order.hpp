// Order of the ladder's contraction index (c, d) on the split-bf16 path.
//
// The ladder contracts t+-(ij, cd) with g+-(cd, ab) over the pairs c <= d (g+)
// and c < d (g-).  Any order of the pairs gives the same sum as long as t+-
// and g+- use the same one.  The FP64 and FP32 forms use the packed triangle
// d (d + 1) / 2 + c.  The split-bf16 form uses a tile-blocked order in which
// every 32-deep matrix-core step falls inside one tile of the (c, d) plane, so
// a kernel that reads two mirrored tiles of a g block can write the
// fragment-ordered bf16 splits of g+- directly:
//   tile columns tj = 0, 1, ... (d = 16 tj + d', d' < w(tj): 16, or for the
//   last column the rest of n_v rounded up to even) in turn; inside column tj
//   first the off-diagonal tiles ti < tj (16 w pairs each, index c' + 16 d'
//   for c = 16 ti + c'), then the diagonal tile (the triangle c' <= d' for g+
//   at d' (d' + 1) / 2 + c', c' < d' for g- at d' (d' - 1) / 2 + c',
//   zero-padded to a multiple of 32: 160 and 128 entries for w = 16).
// Pairs with c or d at or beyond n_v are padding as well; t+- is zero there,
// so the padding adds nothing to the products.
#pragma once

#include <cmath>
#include <cstdint>

#ifndef ETG_HD
#if defined(__HIPCC__)
#define ETG_HD __host__ __device__
#else
#define ETG_HD
#endif
#endif

namespace etg {

constexpr int kLadderTile = 16;

ETG_HD inline int64_t ladder_pad32(int64_t x) { return (x + 31) / 32 * 32; }

ETG_HD inline int64_t ladder_tiles(int64_t n_v) { return (n_v + kLadderTile - 1) / kLadderTile; }

// d-width of tile column tj (even)
ETG_HD inline int64_t ladder_col_width(int64_t tj, int64_t n_v) {
  const int64_t w = n_v - kLadderTile * tj;
  return w >= kLadderTile ? kLadderTile : w + (w & 1);
}

// entries of a diagonal tile of width w (padded)
ETG_HD inline int64_t ladder_diag_size(int64_t w, bool plus) {
  return ladder_pad32(plus ? w * (w + 1) / 2 : w * (w - 1) / 2);
}

// first index of tile column tj (all of its tiles full width below the last column)
ETG_HD inline int64_t ladder_col_offset(int64_t tj, bool plus) {
  return 128 * tj * (tj - 1) + ladder_diag_size(kLadderTile, plus) * tj;
}

// first index of tile (ti, tj), ti <= tj
ETG_HD inline int64_t ladder_tile_offset(int64_t ti, int64_t tj, int64_t n_v, bool plus) {
  return ladder_col_offset(tj, plus) + kLadderTile * ladder_col_width(tj, n_v) * ti;
}

// length of the blocked index (a multiple of 32)
ETG_HD inline int64_t ladder_blocked_len(int64_t n_v, bool plus) {
  const int64_t L = ladder_tiles(n_v) - 1;
  if (L < 0) return 0;
  const int64_t w = ladder_col_width(L, n_v);
  return ladder_tile_offset(L, L, n_v, plus) + ladder_diag_size(w, plus);
}

// largest r with r (r + 1) / 2 <= x, for small x >= 0
ETG_HD inline int small_tri_root(int x) {
  int r = int((sqrt(8.0 * double(x) + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > x) --r;
  while ((r + 1) * (r + 2) / 2 <= x) ++r;
  return r;
}

// position loc inside a tile -> (c', d'); false for a padding position of a
// diagonal tile of width w
ETG_HD inline bool ladder_tile_pair(int loc, bool diag, int w, bool plus, int& cl, int& dl) {
  if (!diag) {
    cl = loc % kLadderTile;
    dl = loc / kLadderTile;
    return true;
  }
  if (plus) {
    dl = small_tri_root(loc);
    cl = loc - dl * (dl + 1) / 2;
    return loc < w * (w + 1) / 2;
  }
  dl = small_tri_root(loc) + 1;
  cl = loc - dl * (dl - 1) / 2;
  return loc < w * (w - 1) / 2;
}

// blocked index kk -> (c, d); false for a padding position
ETG_HD inline bool ladder_pair_of(int64_t kk, int64_t n_v, bool plus, int64_t& c, int64_t& d) {
  const int64_t L = ladder_tiles(n_v) - 1;
  int64_t tj;
  if (kk >= ladder_col_offset(L, plus)) {
    tj = L;
  } else {
    // the largest t with ladder_col_offset(t) = 128 t^2 + (S - 128) t <= kk
    const double S = double(ladder_diag_size(kLadderTile, plus));
    tj = int64_t((-(S - 128.0) + sqrt((S - 128.0) * (S - 128.0) + 512.0 * double(kk))) / 256.0);
    while (tj > 0 && ladder_col_offset(tj, plus) > kk) --tj;
    while (tj + 1 < L && ladder_col_offset(tj + 1, plus) <= kk) ++tj;
  }
  const int64_t w = ladder_col_width(tj, n_v);
  const int64_t r = kk - ladder_col_offset(tj, plus);
  const int64_t ti = r < kLadderTile * w * tj ? r / (kLadderTile * w) : tj;
  const int loc = int(r - kLadderTile * w * ti);
  int cl, dl;
  if (!ladder_tile_pair(loc, ti == tj, int(w), plus, cl, dl)) return false;
  c = kLadderTile * ti + cl;
  d = kLadderTile * tj + dl;
  return c < n_v && d < n_v;
}

}  // namespace etg
