"""Device Davidson: the reduced-space eigensolver against numpy, and the lowest
Jacobian roots against the oracle's dense A (right and left).  The H2O-He EOM
energies against eT are in tests/test_golden_h2o_he.py."""
import numpy as np
import pytest

from et_amd import ccsd, solvers
from et_amd import synthetic as syn
from oracle import ccsd as O

pytestmark = pytest.mark.gpu


def davidson_like(k, seed):
    """Diagonally dominant non-symmetric matrix with a spread, real spectrum."""
    rng = np.random.default_rng(seed)
    return np.diag(np.sort(rng.uniform(0.2, 3.0, k))) + 0.05 * rng.normal(size=(k, k))


@pytest.mark.parametrize('k,seed', [(1, 1), (2, 2), (3, 3), (7, 4), (40, 5), (97, 6), (128, 7)])
def test_reduced_eigen_matches_numpy(k, seed):
    G = davidson_like(k, seed)
    n_want = min(k, 6)
    wr, wi, X = ccsd.reduced_eigen(G, n_want)
    ref = np.linalg.eigvals(G)
    ref = ref[np.lexsort((ref.imag, ref.real))]
    assert np.abs(wr - ref.real).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(wi - ref.imag).max() < 1e-11 * np.abs(ref).max()
    for j in range(n_want):
        x = X[:, j]
        assert abs(np.linalg.norm(x) - 1.0) < 1e-12
        assert np.linalg.norm(G @ x - wr[j] * x) < 1e-11 * np.abs(G).max()


def test_reduced_eigen_with_complex_pairs_above_the_wanted_roots():
    rng = np.random.default_rng(9)
    k = 30
    G = davidson_like(k, 9)
    G[20:22, 20:22] = [[5.0, 1.0], [-1.0, 5.0]]        # a complex pair 5 +- i, far above
    G[25:27, 25:27] = [[7.0, 2.0], [-3.0, 7.0]]
    wr, wi, X = ccsd.reduced_eigen(G, 4)
    ref = np.linalg.eigvals(G)
    ref = ref[np.lexsort((ref.imag, ref.real))]
    assert np.abs(wr + 1j * wi - ref).max() < 1e-10 * np.abs(ref).max()
    with pytest.raises(ccsd.ETGError):
        ccsd.reduced_eigen(np.array([[0.0, 1.0], [-1.0, 0.0]]), 1)   # wanted root complex


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
        dav = solvers.EigenDavidson(wf, 3, side, residual_threshold=1e-9, max_dim=12)
        w = dav.run()
        X = dav.eigenvectors
    np.testing.assert_allclose(w, ref, rtol=0, atol=1e-9)
    M = A if side == 'right' else A.T
    for j in range(3):
        assert np.linalg.norm(M @ X[:, j] - w[j] * X[:, j]) < 1e-8
