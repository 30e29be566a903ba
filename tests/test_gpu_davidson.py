"""Device Davidson: the reduced-space eigensolver against numpy, and the lowest
Jacobian roots against the oracle's dense A (right and left).  The H2O-He EOM
energies against eT are in tests/test_golden_h2o_he.py."""
import numpy as np
import pytest

from et_amd import ccsd, solvers
from et_amd import synthetic as syn
from oracle import ccsd as O

pytestmark = pytest.mark.gpu


def real_spectrum(k, seed):
    """Non-symmetric matrix S D S^-1 with a real spread spectrum (Davidson-like)."""
    rng = np.random.default_rng(seed)
    D = np.diag(np.sort(rng.uniform(0.2, 3.0, k)))
    S = np.eye(k) + 0.2 * rng.normal(size=(k, k)) / np.sqrt(k)
    return S @ D @ np.linalg.inv(S)


def check_pairs(G, wr, wi, X):
    """Columns of X: real eigenvectors, or (u, v) of u + i v for a pair a -+ ib."""
    j = 0
    while j < X.shape[1]:
        if wi[j] == 0.0:
            x = X[:, j]
            assert abs(np.linalg.norm(x) - 1.0) < 1e-12
            assert np.linalg.norm(G @ x - wr[j] * x) < 1e-10 * np.abs(G).max()
            j += 1
        else:
            assert wi[j] < 0 and wi[j + 1] == -wi[j] and wr[j + 1] == wr[j]
            if j + 1 == X.shape[1]:
                break
            z = X[:, j] + 1j * X[:, j + 1]
            lam = wr[j] - 1j * wi[j]
            assert np.linalg.norm(G @ z - lam * z) < 1e-10 * np.abs(G).max() * np.linalg.norm(z)
            j += 2


@pytest.mark.parametrize('k,seed', [(1, 1), (2, 2), (3, 3), (7, 4), (40, 5), (97, 6), (128, 7)])
def test_reduced_eigen_matches_numpy(k, seed):
    G = real_spectrum(k, seed)
    n_want = min(k, 6)
    wr, wi, X = ccsd.reduced_eigen(G, n_want)
    ref = np.linalg.eigvals(G)
    ref = ref[np.lexsort((ref.imag, ref.real))]
    assert np.abs(wr + 1j * wi - ref).max() < 1e-11 * np.abs(ref).max()
    check_pairs(G, wr, wi, X)


@pytest.mark.parametrize('k,seed', [(2, 12), (30, 9), (40, 5), (128, 7)])
def test_reduced_eigen_with_complex_pairs(k, seed):
    """Matrices whose lowest roots include complex pairs (as transient Ritz values do)."""
    rng = np.random.default_rng(seed)
    G = np.diag(np.sort(rng.uniform(0.2, 3.0, k))) + 0.05 * rng.normal(size=(k, k))
    if k == 2:
        G = np.array([[1.0, 1.0], [-1.0, 1.0]])          # 1 -+ i
    n_want = min(k, 7)
    wr, wi, X = ccsd.reduced_eigen(G, n_want)
    ref = np.linalg.eigvals(G)
    ref = ref[np.lexsort((ref.imag, ref.real))]
    assert np.abs(wr + 1j * wi - ref).max() < 1e-10 * np.abs(ref).max()
    assert np.any(wi[:n_want] != 0.0)
    check_pairs(G, wr, wi, X)


@pytest.mark.parametrize('side', ['right', 'left'])
def test_davidson_lowest_roots_match_dense_jacobian(side):
    s = syn.system(3, 3, 4, 9, t2_sigma=0.03, t1_sigma=0.03, J_decay=6.0, L_scale=0.3)
    A = O.jacobian_matrix(s['L'], s['h'], s['t1'], s['t2'], 3)
    ref = np.sort(np.linalg.eigvals(A).real)[:3]
    with ccsd.CCSD(3, 4, 9) as wf:
        wf.set_cholesky(s['L'])
        wf.set_oei(s['h'])
        wf.set_orbital_energies(s['eps'])
        wf.set_amplitudes(np.concatenate([s['t1'].ravel(order='F'), s['t2']]))
        wf.construct_t1_cholesky()
        dav = solvers.EigenDavidson(wf, 3, side, residual_threshold=1e-9, max_dim=30)
        w = dav.run()
        X = dav.eigenvectors
    np.testing.assert_allclose(w, ref, rtol=0, atol=1e-9)
    M = A if side == 'right' else A.T
    for j in range(3):
        assert np.linalg.norm(M @ X[:, j] - w[j] * X[:, j]) < 1e-8


@pytest.mark.parametrize('side', ['right', 'left'])
def test_davidson_space_in_host_memory(side):
    """V and A V in pinned host memory (read and written by the kernels over PCIe)
    give the roots of the HBM placement, with restarts of the space (max_dim 14,
    four new vectors per iteration) and the block Gram-Schmidt of the new trial
    vectors."""
    n_o, n_v, n_J = 3, 6, 12
    s = syn.system(8, n_o, n_v, n_J, t2_sigma=0.03, t1_sigma=0.03, J_decay=6.0, L_scale=0.3)
    A = O.jacobian_matrix(s['L'], s['h'], s['t1'], s['t2'], n_o)
    ref = np.sort(np.linalg.eigvals(A).real)[:3]
    out = {}
    with ccsd.CCSD(n_o, n_v, n_J) as wf:
        wf.set_cholesky(s['L'])
        wf.set_oei(s['h'])
        wf.set_orbital_energies(s['eps'])
        wf.set_amplitudes(np.concatenate([s['t1'].ravel(order='F'), s['t2']]))
        wf.construct_t1_cholesky()
        for where in ('hbm', 'host'):
            wf.set_davidson_space(where)
            print('space', where, flush=True)
            dav = solvers.EigenDavidson(wf, 3, side, residual_threshold=1e-8, max_dim=14)
            out[where] = (np.asarray(dav.run()), len(dav.history))
        wf.set_davidson_space('auto')
    np.testing.assert_allclose(out['hbm'][0], ref, rtol=0, atol=1e-8)
    assert np.abs(out['host'][0] - out['hbm'][0]).max() < 1e-11
    assert out['host'][1] == out['hbm'][1]
