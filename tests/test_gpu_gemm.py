"""Matrix-core product kernel against numpy (FP64 exact to round-off, FP32 to its tolerance)."""
import numpy as np
import pytest

from et_amd import ccsd

pytestmark = pytest.mark.gpu

SHAPES = [
    (1, 1, 1), (16, 16, 4), (17, 33, 5), (64, 64, 16), (100, 70, 37), (128, 128, 16),
    (129, 257, 31), (300, 200, 500), (48, 48, 4000),     # split-K
    (1000, 1, 1000),                                       # GEMV-shaped
    (640, 520, 130),
]


@pytest.mark.parametrize('ta', ['N', 'T'])
@pytest.mark.parametrize('tb', ['N', 'T'])
@pytest.mark.parametrize('m,n,k', SHAPES)
def test_fp64_gemm(ta, tb, m, n, k):
    rng = np.random.default_rng(m * 131 + n * 7 + k)
    A = rng.normal(size=(k, m) if ta == 'T' else (m, k))
    B = rng.normal(size=(n, k) if tb == 'T' else (k, n))
    C = rng.normal(size=(m, n))
    out = ccsd.device_gemm(ta, tb, 0.7, A, B, -0.3, C)
    ref = 0.7 * (A.T if ta == 'T' else A) @ (B.T if tb == 'T' else B) - 0.3 * C
    scale = np.sqrt(k) + 1.0
    assert np.abs(out - ref).max() < 1e-14 * scale * 4


@pytest.mark.parametrize('m,n,k', [(100, 70, 37), (300, 200, 500), (48, 48, 4000)])
def test_fp32_gemm(m, n, k):
    rng = np.random.default_rng(5)
    A = rng.normal(size=(m, k))
    B = rng.normal(size=(k, n))
    C = np.zeros((m, n))
    out = ccsd.device_gemm('N', 'N', 1.0, A, B, 0.0, C, precision=ccsd.FP32)
    ref = A @ B
    # FP32 rounding of the operands and FP32 accumulation: ~ sqrt(k) 2^-24 relative per term
    tol = 4.0 * np.sqrt(k) * 2.0 ** -24 * np.abs(A).max() * np.abs(B).max() * np.sqrt(k)
    assert np.abs(out - ref).max() < tol


@pytest.mark.parametrize('ta', ['N', 'T'])
@pytest.mark.parametrize('m', [152, 192, 130])
def test_products_with_a_short_last_row_tile(ta, m):
    """m % 128 <= 64 with 1e10 FLOP or more: the last rows run as a product of their own
    (offset A rows or columns and C rows); FP64 and FP32 against numpy, beta != 0."""
    n, k = 2560, 16384
    assert 2 * m * n * k >= 1e10
    rng = np.random.default_rng(m)
    A = rng.normal(size=(k, m) if ta == 'T' else (m, k))
    B = rng.normal(size=(k, n))
    C = rng.normal(size=(m, n))
    opA = A.T if ta == 'T' else A
    ref = 0.6 * opA @ B + 0.25 * C
    out = ccsd.device_gemm(ta, 'N', 0.6, A, B, 0.25, C)
    assert np.abs(out - ref).max() < 1e-14 * np.sqrt(k) * 8
    out32 = ccsd.device_gemm(ta, 'N', 0.6, A, B, 0.25, C, precision=ccsd.FP32)
    assert np.abs(out32 - ref).max() < 4.0 * k * 2.0 ** -24 * np.abs(A).max() * np.abs(B).max()


def test_zero_k_scales_c():
    C = np.arange(12.0).reshape(3, 4)
    out = ccsd.device_gemm('N', 'N', 1.0, np.zeros((3, 0)), np.zeros((0, 4)), 2.0, C)
    np.testing.assert_array_equal(out, 2.0 * C)


BIG = (1 << 21) + 3      # leading dimension past one buffer base per tile (per-element bases)


@pytest.mark.parametrize('ta,tb,m,n,k', [
    ('N', 'N', BIG, 3, 8),       # A: lda = BIG, contiguous along m
    ('T', 'N', 5, 6, BIG),       # A and B contiguous along k, lda = ldb = BIG (split-K)
    ('N', 'T', 7, BIG, 5),       # B stored n x k, ldb = BIG
    ('T', 'T', 4, BIG, 6),
])
def test_fp64_gemm_large_leading_dimension(ta, tb, m, n, k):
    rng = np.random.default_rng(11)
    A = rng.normal(size=(k, m) if ta == 'T' else (m, k))
    B = rng.normal(size=(n, k) if tb == 'T' else (k, n))
    C = rng.normal(size=(m, n))
    out = ccsd.device_gemm(ta, tb, 1.1, A, B, 0.5, C)
    ref = 1.1 * (A.T if ta == 'T' else A) @ (B.T if tb == 'T' else B) + 0.5 * C
    assert np.abs(out - ref).max() < 1e-14 * (np.sqrt(k) + 1.0) * 8


@pytest.mark.parametrize('ta', ['N', 'T'])
@pytest.mark.parametrize('tb', ['N', 'T'])
@pytest.mark.parametrize('m,n,k', [(1, 1, 1), (17, 33, 5), (100, 70, 37), (129, 257, 31),
                                   (300, 200, 500), (48, 48, 4000), (640, 520, 130)])
def test_bf16x3_gemm(ta, tb, m, n, k):
    """Split-bf16 products against an elementwise bound.

    x = hi + lo + d with |d| <= 2^-16 |x| and |lo| <= 2^-8 |x|, so a product term
    loses at most 3 * 2^-16 |a b| (lo*lo dropped, two splitting residuals); FP32
    accumulation adds at most k 2^-24 sum |a b|.  A missing hi*lo or lo*hi term,
    or a wrong lane layout, gives errors of order 2^-8 |a b| and fails.
    """
    rng = np.random.default_rng(m * 17 + n * 3 + k)
    A = rng.normal(size=(k, m) if ta == 'T' else (m, k))
    B = rng.normal(size=(n, k) if tb == 'T' else (k, n))
    C = rng.normal(size=(m, n))
    opA = A.T if ta == 'T' else A
    opB = B.T if tb == 'T' else B
    out = ccsd.device_gemm(ta, tb, 0.7, A, B, -0.3, C, precision=ccsd.BF16X3_ANY_SHAPE)
    ref = 0.7 * opA @ opB - 0.3 * C
    bound = 0.7 * (4 * 2.0 ** -16 + k * 2.0 ** -24) * (np.abs(opA) @ np.abs(opB)) + 1e-15 * np.abs(C)
    assert np.all(np.abs(out - ref) <= bound)


@pytest.mark.parametrize('ta,tb,m,n,k', [('N', 'N', (1 << 22) + 1, 3, 5), ('T', 'N', 3, (1 << 22) + 1, 5)])
def test_bf16x3_gemm_with_more_than_65535_row_tiles(ta, tb, m, n, k):
    """An operand of more than 2^22 rows: the pack walks its 64-row tiles beyond the
    65535 of one grid dimension (the T1 part of the Jacobian products at o=48, v=400)."""
    rng = np.random.default_rng(7)
    A = rng.normal(size=(k, m) if ta == 'T' else (m, k))
    B = rng.normal(size=(k, n))
    C = np.zeros((m, n))
    opA = A.T if ta == 'T' else A
    out = ccsd.device_gemm(ta, tb, 1.0, A, B, 0.0, C, precision=ccsd.BF16X3_ANY_SHAPE)
    ref = opA @ B
    bound = (4 * 2.0 ** -16 + k * 2.0 ** -24) * (np.abs(opA) @ np.abs(B)) + 1e-300
    assert np.all(np.abs(out - ref) <= bound)


@pytest.mark.parametrize('m,n,k,split', [(8, 8, 1 << 20, False), (40, 3, 400000, False),
                                         (48, 48, 4000, True), (8, 8, (1 << 20) - 1, True),
                                         (129, 257, 31, True), (300, 200, 500, True), (128, 128, 8000, True)])
def test_split_bf16_products_bound_by_hbm_run_in_fp64(m, n, k, split):
    """Products asked for in split-bf16 with m n < 64 (m + n) and operands of 2^24
    elements or more run on the FP64 kernel (bit-identical to the FP64 product); the
    others on the split-bf16 kernel (different from FP64 in the last bits, within the
    split-bf16 bound)."""
    assert (m * n >= 64 * (m + n) or (m + n) * k < (1 << 24)) == split
    rng = np.random.default_rng(m + 7 * n + k)
    A = rng.normal(size=(m, k))
    B = rng.normal(size=(k, n))
    C = rng.normal(size=(m, n))
    ccsd.device_gemm('N', 'N', 0.9, A, B, 0.4, C, precision=ccsd.FP64)   # settles the measured tile choice
    f64 = ccsd.device_gemm('N', 'N', 0.9, A, B, 0.4, C, precision=ccsd.FP64)
    bf = ccsd.device_gemm('N', 'N', 0.9, A, B, 0.4, C, precision=ccsd.BF16X3)
    if split:
        assert not np.array_equal(bf, f64)
        bound = 0.9 * (4 * 2.0 ** -16 + k * 2.0 ** -24) * (np.abs(A) @ np.abs(B)) + 1e-15 * np.abs(C)
        assert np.all(np.abs(bf - (0.9 * A @ B + 0.4 * C)) <= bound)
    else:
        assert np.array_equal(bf, f64)


@pytest.mark.parametrize('ta,tb,m,n,k', [('N', 'N', 48, 48, 4000), ('T', 'N', 300, 200, 500),
                                         ('N', 'T', 640, 520, 130), ('T', 'T', 129, 257, 3000)])
def test_measured_tile_choice_keeps_results(ta, tb, m, n, k):
    """The measured tile choice times candidates into a scratch output: the product itself
    (with beta != 0) is the same as with the cost model's pick, to FP64 rounding."""
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = rng.normal(size=(k, m) if ta == 'T' else (m, k))
    B = rng.normal(size=(n, k) if tb == 'T' else (k, n))
    C = rng.normal(size=(m, n))
    ref = 0.7 * (A.T if ta == 'T' else A) @ (B.T if tb == 'T' else B) - 0.3 * C
    scale = 4e-14 * (np.sqrt(k) + 1.0)
    try:
        ccsd.set_gemm_autotune(False)
        model = ccsd.device_gemm(ta, tb, 0.7, A, B, -0.3, C)
        before = ccsd.set_gemm_autotune(True)
        tuned = ccsd.device_gemm(ta, tb, 0.7, A, B, -0.3, C)
        again = ccsd.device_gemm(ta, tb, 0.7, A, B, -0.3, C)
        after = ccsd.set_gemm_autotune(True)
    finally:
        ccsd.set_gemm_autotune(True)
    assert np.abs(model - ref).max() < scale
    assert np.abs(tuned - ref).max() < scale
    assert np.array_equal(tuned, again)          # the kept choice is reused
    assert after >= before


@pytest.mark.parametrize('ta,tb,m,n,k', [('T', 'N', 3, 2, (1 << 24) + 3),       # inner products
                                         ('T', 'N', 16, 16, (1 << 24) + 3),     # one output per thread
                                         ('T', 'N', 20, 17, (1 << 24) + 3),     # two 16-column blocks
                                         ('T', 'N', 5, 40, (1 << 24) + 3),      # C^T = B^T A
                                         ('N', 'N', (1 << 24) + 3, 2, 3)])      # tall output
def test_tall_products_beyond_the_tile_kernel_offsets(ta, tb, m, n, k):
    """Leading dimensions >= 2^24 (the Davidson space of (H2O)_8) go to the tall
    FP64 kernels: same result as numpy, with beta != 0."""
    rng = np.random.default_rng(11)
    A = rng.normal(size=(k, m) if ta == 'T' else (m, k))
    B = rng.normal(size=(k, n))
    C = rng.normal(size=(m, n))
    out = ccsd.device_gemm(ta, tb, 0.5, A, B, 2.0, C)
    ref = 0.5 * (A.T if ta == 'T' else A) @ B + 2.0 * C
    assert np.abs(out - ref).max() < 1e-13 * (np.sqrt(max(k, 1)) + 1.0) * np.abs(ref).max()
