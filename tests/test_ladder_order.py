"""The split-bf16 ladder's tile-blocked (c, d) order (et_amd/csrc/ladder_order.hpp),
built for the host and checked for what the products rely on: the blocked index
visits every pair c <= d (g+) or c < d (g-) exactly once, the rest is padding,
every tile starts on a 32-deep matrix-core step, and the padding stays small.
The device kernels (t+- and the g+- packs) use the same header."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, 'et_amd', 'csrc', 'ladder_order.hpp')

SRC = r'''
#include "ladder_order.hpp"
extern "C" int64_t blocked_len(int64_t n_v, int plus) { return etg::ladder_blocked_len(n_v, plus != 0); }
extern "C" int64_t tile_offset(int64_t ti, int64_t tj, int64_t n_v, int plus) {
  return etg::ladder_tile_offset(ti, tj, n_v, plus != 0);
}
// pairs of every blocked index: c, d (or -1 for padding)
extern "C" void pairs(int64_t n_v, int plus, int64_t* c, int64_t* d) {
  const int64_t K = etg::ladder_blocked_len(n_v, plus != 0);
  for (int64_t kk = 0; kk < K; ++kk) {
    int64_t cc = -1, dd = -1;
    if (!etg::ladder_pair_of(kk, n_v, plus != 0, cc, dd)) cc = dd = -1;
    c[kk] = cc;
    d[kk] = dd;
  }
}
'''


@pytest.fixture(scope='module')
def lib():
    d = tempfile.mkdtemp(prefix='ladder_order_')
    src = os.path.join(d, 'l.cpp')
    so = os.path.join(d, 'libl.so')
    with open(src, 'w') as f:
        f.write(SRC)
    subprocess.run(['g++', '-O2', '-std=c++17', '-shared', '-fPIC', '-I', os.path.dirname(HDR), src, '-o', so],
                   check=True)
    L = ctypes.CDLL(so)
    L.blocked_len.restype = ctypes.c_int64
    L.blocked_len.argtypes = [ctypes.c_int64, ctypes.c_int]
    L.tile_offset.restype = ctypes.c_int64
    L.tile_offset.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
    L.pairs.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    return L


@pytest.mark.parametrize('n_v', [1, 2, 3, 7, 15, 16, 17, 33, 37, 64, 152, 304, 401])
@pytest.mark.parametrize('plus', [1, 0])
def test_blocked_order_is_a_bijection_onto_the_pairs(lib, n_v, plus):
    K = lib.blocked_len(n_v, plus)
    assert K % 32 == 0
    c = np.empty(K, dtype=np.int64)
    d = np.empty(K, dtype=np.int64)
    lib.pairs(n_v, plus, c.ctypes.data, d.ctypes.data)
    real = c >= 0
    cc, dd = c[real], d[real]
    assert np.all(cc <= dd) if plus else np.all(cc < dd)
    assert np.all(dd < n_v)
    code = cc + n_v * dd
    assert len(np.unique(code)) == len(code)                     # each pair once
    want = n_v * (n_v + 1) // 2 if plus else n_v * (n_v - 1) // 2
    assert len(code) == want                                     # and every pair
    # every tile of the (c, d) plane starts on a 32-deep step
    nt = (n_v + 15) // 16
    for tj in range(nt):
        for ti in range(tj + 1):
            assert lib.tile_offset(ti, tj, n_v, plus) % 32 == 0
    # padding: diagonal tiles padded to 32, the last tile column to even width
    if n_v >= 64:
        assert K <= 1.025 * want + 160


def test_blocked_order_pads_little_at_the_benchmark_sizes(lib):
    for n_v, plus, K in ((400, 1, 80800), (400, 0, 80000), (152, 1, 11872), (304, 1, 46816)):
        assert lib.blocked_len(n_v, plus) == K
