"""etg_reduced_eigen_noalloc: the reduced-space eigensolver on caller-owned
device memory with its workspace in a static device array, against the
allocating entry point etg_test_reduced_eigen on the same matrices.

Both run the same kernel on the same input, so eigenvalues and vectors must be
identical bit for bit.  Two calls in a row, the second smaller, show that the
shared array keeps nothing a later call depends on; leading dimensions above
the dimension exercise the strided reads and writes.  A dimension above the
array's size is rejected before any device work.
"""
import ctypes

import numpy as np
import pytest

from et_amd import ccsd

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    """The HIP runtime the library itself loaded (same process-wide instance)."""
    ccsd.load_library()
    with open('/proc/self/maps') as f:
        paths = {ln.split()[-1] for ln in f if 'libamdhip64' in ln and ln.split()[-1].startswith('/')}
    assert paths, 'libetgpu.so did not load a HIP runtime'
    hip = ctypes.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def davidson_like(k, seed):
    """A reduced matrix with a real, well separated low spectrum (V^T A V of a
    nearly symmetric A)."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(k, k)))
    d = np.sort(rng.uniform(0.2, 2.0, size=k))
    return Q @ np.diag(d) @ Q.T + 1e-3 * rng.normal(size=(k, k))


def noalloc_solve(hip, G, n_want, ldg, ldx):
    lib = ccsd.load_library()
    k = G.shape[0]
    Gp = np.zeros((ldg, k), order='F')
    Gp[:k, :] = G
    dG, dX = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dG), Gp.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dX), 8 * ldx * max(1, n_want)) == 0
    try:
        assert hip.hipMemcpy(dG, Gp.ctypes.data, Gp.nbytes, H2D) == 0
        wr, wi = np.empty(k), np.empty(k)
        rc = lib.etg_reduced_eigen_noalloc(
            ctypes.c_int64(k), dG, ctypes.c_int64(ldg), ctypes.c_int64(n_want),
            wr.ctypes.data_as(ctypes.c_void_p), wi.ctypes.data_as(ctypes.c_void_p), dX, ctypes.c_int64(ldx))
        assert rc == 0, lib.etg_last_error().decode()
        Xp = np.empty((ldx, max(1, n_want)), order='F')
        assert hip.hipMemcpy(Xp.ctypes.data, dX, Xp.nbytes, D2H) == 0
        return wr, wi, Xp[:k, :n_want]
    finally:
        hip.hipFree(dG)
        hip.hipFree(dX)


def test_noalloc_reduced_eigen_matches_allocating_entry():
    hip = hip_runtime()
    for k, n_want, seed in ((96, 12, 3), (23, 4, 4)):          # the second call smaller
        print('noalloc reduced eigen: k=%d n_want=%d seed=%d' % (k, n_want, seed), flush=True)
        G = davidson_like(k, seed)
        wr_ref, wi_ref, X_ref = ccsd.reduced_eigen(G, n_want)
        wr, wi, X = noalloc_solve(hip, G, n_want, ldg=k + 3, ldx=k + 2)
        assert np.array_equal(wr, wr_ref) and np.array_equal(wi, wi_ref)
        assert np.array_equal(X, X_ref)
        # and they are eigenpairs of G
        assert np.abs(G @ X - X * wr[:n_want]).max() < 1e-10


def test_noalloc_reduced_eigen_rejects_oversize():
    lib = ccsd.load_library()
    wr, wi = np.empty(200), np.empty(200)
    rc = lib.etg_reduced_eigen_noalloc(ctypes.c_int64(129), ctypes.c_void_p(8), ctypes.c_int64(129),
                                       ctypes.c_int64(1), wr.ctypes.data_as(ctypes.c_void_p),
                                       wi.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(8),
                                       ctypes.c_int64(129))
    assert rc != 0
    assert '128' in lib.etg_last_error().decode()
