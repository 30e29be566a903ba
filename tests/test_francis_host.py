"""The reduced eigensolver's Hessenberg QR step (et_amd/csrc/francis.hpp) built
for the host as one lane and checked against LAPACK (numpy.linalg.eigvals):
the device kernel runs the same source with 64 lanes."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest
import scipy.linalg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, 'et_amd', 'csrc', 'francis.hpp')

SRC = r'''
#include "francis.hpp"
extern "C" int host_eigenvalues(double* H, int k, double* wr, double* wi, double* hnorm) {
  etg::francis::HostTeam team;
  return etg::francis::hessenberg_eigenvalues(team, H, k, k, wr, wi, hnorm);
}
'''


@pytest.fixture(scope='module')
def lib():
    d = tempfile.mkdtemp(prefix='francis_')
    src = os.path.join(d, 'f.cpp')
    so = os.path.join(d, 'libf.so')
    with open(src, 'w') as f:
        f.write(SRC)
    subprocess.run(['g++', '-O2', '-std=c++17', '-shared', '-fPIC', '-I', os.path.dirname(HDR), src, '-o', so],
                   check=True)
    L = ctypes.CDLL(so)
    L.host_eigenvalues.restype = ctypes.c_int
    return L


def eig_hessenberg(lib, A):
    k = A.shape[0]
    H = np.asfortranarray(scipy.linalg.hessenberg(A) if k > 2 else A.copy())
    wr = np.zeros(k)
    wi = np.zeros(k)
    hn = ctypes.c_double()
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    info = lib.host_eigenvalues(p(H), k, p(wr), p(wi), ctypes.byref(hn))
    return info, wr + 1j * wi


def match(got, ref, tol):
    got = sorted(got, key=lambda z: (round(z.real, 8), z.imag))
    ref = sorted(ref, key=lambda z: (round(z.real, 8), z.imag))
    return max(abs(a - b) for a, b in zip(got, ref)) <= tol


@pytest.mark.parametrize('k', [1, 2, 3, 4, 7, 16, 33, 64, 128])
def test_random_nonsymmetric(lib, k):
    rng = np.random.default_rng(k)
    for trial in range(6):
        A = rng.standard_normal((k, k))
        info, w = eig_hessenberg(lib, A)
        assert info == 0
        ref = np.linalg.eigvals(A)
        scale = np.abs(A).sum()
        assert match(w, ref, 1e-11 * scale), (k, trial)
        # complex pairs: negative imaginary part first, equal real parts
        for i in range(k):
            if w[i].imag < 0:
                assert abs(w[i + 1] - np.conj(w[i])) < 1e-14 * scale


def test_davidson_like(lib):
    # nearly upper triangular reduced matrices of a converging Davidson space,
    # a close pair, an exact multiple root and a complex pair
    rng = np.random.default_rng(7)
    for k in (12, 36, 96):
        d = np.sort(rng.uniform(0.2, 0.9, k))
        d[3] = d[2] + 1e-9
        d[5] = d[4]
        A = np.diag(d) + 1e-6 * rng.standard_normal((k, k))
        A[k - 2:, k - 2:] = [[0.95, 0.3], [-0.2, 0.95]]
        info, w = eig_hessenberg(lib, A)
        assert info == 0
        assert match(w, np.linalg.eigvals(A), 1e-12)


def test_special_matrices(lib):
    cases = [np.zeros((5, 5)), np.eye(6) * 3.0, np.diag(np.arange(8.0)),
             np.triu(np.ones((6, 6))),
             np.roll(np.eye(10), 1, axis=0),               # cyclic shift: roots of unity
             np.array([[0.0, 1.0], [-1.0, 0.0]]),
             np.array([[2.0, 1e-300], [1e-300, 2.0]])]
    for A in cases:
        info, w = eig_hessenberg(lib, A)
        assert info == 0
        assert match(w, np.linalg.eigvals(A), 1e-12 * max(1.0, np.abs(A).sum()))


def test_block2_formula(lib):
    # 2 x 2 blocks through the deflation path: real with cancellation, complex
    rng = np.random.default_rng(3)
    for _ in range(200):
        A = rng.standard_normal((2, 2)) * 10.0 ** rng.integers(-8, 8)
        info, w = eig_hessenberg(lib, A)
        assert info == 0
        assert match(w, np.linalg.eigvals(A), 1e-13 * np.abs(A).sum())


def test_unit_circle_spectra(lib):
    # cyclic permutations and random orthogonal matrices: every eigenvalue on
    # the unit circle, where the standard shift alone stagnates
    rng = np.random.default_rng(11)
    for k in range(3, 41):
        for A in (np.roll(np.eye(k), 1, axis=0), np.linalg.qr(rng.standard_normal((k, k)))[0]):
            info, w = eig_hessenberg(lib, A)
            assert info == 0, k
            assert match(w, np.linalg.eigvals(A), 1e-11 * k)


def test_multiplets_close_to_scalar_blocks(lib):
    """Exactly degenerate multiplets of Davidson-like reduced matrices: the block
    of a converging multiple root is close to a multiple of I, where the first
    column of the double-shift polynomial must be formed without cancellation
    (a degenerate triple of EOM roots stalled the iteration before)."""
    rng = np.random.default_rng(21)
    for trial in range(300):
        k = int(rng.integers(4, 100))
        base = np.sort(rng.uniform(0.2, 2.0, k))
        A = np.diag(base) + 1e-8 * rng.standard_normal((k, k))
        for _ in range(2):
            i = int(rng.integers(0, k - 3))
            A[i:i + 3, i:i + 3] = base[i] * np.eye(3) + 1e-12 * rng.standard_normal((3, 3))
        if trial % 2:
            S = np.eye(k) + 0.1 * rng.standard_normal((k, k))
            A = S @ A @ np.linalg.inv(S)
        info, w = eig_hessenberg(lib, A)
        assert info == 0, trial
        ref = np.linalg.eigvals(A)
        # multiple roots are determined to ~ sqrt(eps) of the coupling inside the block
        assert match(w, ref, 1e-9 * np.abs(A).max() * k), trial
