"""Pins the CPU oracle (oracle/ccsd.py) independently of its own formulas.

* brute force: exp(-T) H exp(T)|HF> in the full determinant space
  (tests/brute_force_cc.py) with explicit T1 and T2 gives the singles and the
  alpha-beta doubles projections; eT's residual must equal them, with the
  factor 1/2 on the packed doubles diagonal (omega_ccsd.F90:100), and the
  energy functional must equal <HF|Hbar|HF> - E_ref;
* closed forms at t = 0 and for the MP2 guess;
* the +/- ladder against the naive sum_cd g_acbd t_cidj;
* the blockwise T1 dressing against X L Y^T;
* DIIS against the closed-form two-vector extrapolation.
"""
import numpy as np
import pytest

import brute_force_cc as bf
from et_amd import synthetic as syn
from oracle import ccsd as O


def tiny(seed, n_o, n_v, n_J, t_sigma=0.05):
    s = syn.system(seed, n_o, n_v, n_J, t2_sigma=t_sigma, t1_sigma=t_sigma, J_decay=4.0,
                   L_scale=0.3)
    return s['L'], s['h'], s['eps'], s['t1'], s['t2']


@pytest.mark.parametrize('n_o,n_v,n_J,seed', [(1, 2, 3, 1), (2, 2, 4, 2), (2, 3, 6, 7),
                                              (3, 2, 5, 3), (2, 4, 5, 11), (3, 3, 7, 5)])
def test_residual_matches_determinant_space(n_o, n_v, n_J, seed):
    L, h, eps, t1, t2 = tiny(seed, n_o, n_v, n_J)
    Lt = O.t1_cholesky(L, t1, n_o)
    F = O.fock_t1(h, Lt, t1, n_o)
    om = O.omega(L, Lt, F, t1, t2, n_o)
    g = np.einsum('Jpq,Jrs->pqrs', L, L)
    t = O.unpack_t2(t2, n_v, n_o)
    space, w = bf.similarity_transformed_hf(h, g, t1, t, n_o)
    E, R1, R2 = bf.projections(space, w)
    n_t1 = n_o * n_v
    np.testing.assert_allclose(om[:n_t1].reshape(n_v, n_o, order='F'), R1, rtol=0, atol=1e-13)
    R2m = O.aibj_to_mat(R2)
    d = np.arange(n_t1)
    R2m[d, d] *= 0.5
    np.testing.assert_allclose(om[n_t1:], O.packin(R2m), rtol=0, atol=1e-13)
    e_ref = bf.nuclear_free_reference_energy(h, g, n_o)
    assert abs((E - e_ref) - O.correlation_energy(Lt, t1, t2, n_o)) < 1e-13
    assert abs(e_ref - O.hf_energy(h, L, n_o)) < 1e-12


def test_residual_at_zero_amplitudes_is_fock_and_integrals():
    """Omega(t=0): singles F_ai, doubles g_aibj with 1/2 on the diagonal (omega_ccsd.F90:97-100)."""
    n_o, n_v, n_J = 3, 5, 9
    L, h, eps, _, _ = tiny(4, n_o, n_v, n_J)
    t1 = np.zeros((n_v, n_o))
    t2 = np.zeros(O.n_t2(n_o, n_v))
    Lt = O.t1_cholesky(L, t1, n_o)
    F = O.fock_t1(h, Lt, t1, n_o)
    np.testing.assert_allclose(F, np.diag(eps), atol=1e-13)          # canonical reference
    om = O.omega(L, Lt, F, t1, t2, n_o)
    n_t1 = n_o * n_v
    np.testing.assert_allclose(om[:n_t1], 0.0, atol=1e-14)
    g = np.einsum('Jpq,Jrs->pqrs', L, L)[n_o:, :n_o, n_o:, :n_o]
    gm = O.aibj_to_mat(g)
    d = np.arange(n_t1)
    gm[d, d] *= 0.5
    np.testing.assert_allclose(om[n_t1:], O.packin(gm), atol=1e-14)


def test_mp2_guess_energy_is_mp2_energy():
    n_o, n_v, n_J = 3, 6, 10
    L, h, eps, _, _ = tiny(9, n_o, n_v, n_J)
    Lt = O.t1_cholesky(L, np.zeros((n_v, n_o)), n_o)
    t2 = O.mp2_guess(Lt, eps, n_o)
    e = O.correlation_energy(Lt, np.zeros((n_v, n_o)), t2, n_o)
    g = np.einsum('Jpq,Jrs->pqrs', L, L)
    e_mp2 = 0.0
    for i in range(n_o):
        for j in range(n_o):
            for a in range(n_v):
                for b in range(n_v):
                    iajb = g[i, n_o + a, j, n_o + b]
                    ibja = g[i, n_o + b, j, n_o + a]
                    e_mp2 += iajb * (2 * iajb - ibja) / (eps[i] + eps[j] - eps[n_o + a] - eps[n_o + b])
    assert abs(e - e_mp2) < 1e-13 * max(1.0, abs(e_mp2))


@pytest.mark.parametrize('n_o,n_v', [(1, 1), (1, 3), (2, 2), (3, 5), (4, 3)])
def test_plus_minus_ladder_equals_naive(n_o, n_v):
    L, h, eps, t1, t2 = tiny(13, n_o, n_v, 6, t_sigma=0.1)
    Lt = O.t1_cholesky(L, t1, n_o)      # t1 != 0: L_vv is not symmetric
    a = O.omega_a2_plus_minus(Lt, t2, n_o, n_v)
    b = O.omega_a2_naive(Lt, t2, n_o, n_v)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


def test_t1_cholesky_is_similarity_sandwich():
    n_o, n_v = 3, 4
    L, h, eps, t1, t2 = tiny(21, n_o, n_v, 5, t_sigma=0.2)
    n = n_o + n_v
    X = np.eye(n)
    Y = np.eye(n)
    X[n_o:, :n_o] = -t1
    Y[:n_o, n_o:] = t1.T
    ref = np.einsum('pr,Jrs,qs->Jpq', X, L, Y)
    np.testing.assert_allclose(O.t1_cholesky(L, t1, n_o), ref, atol=1e-14)


def test_diis_two_vectors_closed_form():
    rng = np.random.default_rng(0)
    e1, e2 = rng.normal(size=7), rng.normal(size=7)
    x1, x2 = rng.normal(size=5), rng.normal(size=5)
    d = O.DIIS(8)
    d.update(e1, x1)
    x = d.update(e2, x2)
    # minimise |w e1 + (1-w) e2|^2
    w = float((e2 - e1) @ e2 / ((e2 - e1) @ (e2 - e1)))
    np.testing.assert_allclose(x, w * x1 + (1 - w) * x2, atol=1e-13)


def test_ground_state_converges_to_determinant_space_stationary_point():
    n_o, n_v, n_J = 2, 3, 6
    L, h, eps, _, _ = tiny(31, n_o, n_v, n_J)
    t1, t2, e, hist = O.solve_ground_state(L, h, eps, n_o, residual_threshold=1e-11)
    g = np.einsum('Jpq,Jrs->pqrs', L, L)
    space, w = bf.similarity_transformed_hf(h, g, t1, O.unpack_t2(t2, n_v, n_o), n_o)
    E, R1, R2 = bf.projections(space, w)
    assert np.abs(R1).max() < 1e-10 and np.abs(R2).max() < 1e-10
    assert abs(E - bf.nuclear_free_reference_energy(h, g, n_o) - e) < 1e-12
    assert hist[0][1] > hist[-1][1]


def test_complex_step_jacobian_equals_central_differences():
    n_o, n_v = 2, 3
    L, h, eps, t1, t2 = tiny(41, n_o, n_v, 6, t_sigma=0.05)
    n = n_o * n_v + O.n_t2(n_o, n_v)
    c = np.random.default_rng(3).normal(size=n)
    rho = O.jacobian_transform(L, h, t1, t2, n_o, c)
    cs = c.copy()
    cs[O.doubles_diagonal(n_o * n_v)] *= 2.0
    hstep = 1e-4
    n_t1 = n_o * n_v

    def om(x):
        return O.residual(L, h, t1 + x * cs[:n_t1].reshape(n_v, n_o, order='F'), t2 + x * cs[n_t1:], n_o)
    # fourth-order central difference (Omega is a low-degree polynomial in x)
    fd = (8 * (om(hstep) - om(-hstep)) - (om(2 * hstep) - om(-2 * hstep))) / (12 * hstep)
    np.testing.assert_allclose(rho, fd, rtol=0, atol=1e-9 * np.abs(rho).max())


def test_jacobian_diagonal_convention_on_determinant_space():
    """A_mu,nu = <mu~|[Hbar, E_ck E_dl]|HF>: for a doubles-diagonal nu (ck = dl) the operator
    E_ck E_ck is twice the amplitude's T2 weight, hence the factor 2 on c's diagonal."""
    n_o, n_v = 2, 2
    L, h, eps, t1, t2 = tiny(43, n_o, n_v, 5)
    n_t1 = n_o * n_v
    n = n_t1 + O.n_t2(n_o, n_v)
    e = np.zeros(n)
    k = O.doubles_diagonal(n_t1)[1]
    e[k] = 1.0
    rho = O.jacobian_transform(L, h, t1, t2, n_o, e)
    step = 1e-30
    t2c = t2 + 0j
    t2c[k - n_t1] += 2j * step       # amplitude change 2 eps <-> operator E_ai E_ai with weight eps
    direct = O.residual(L, h, t1 + 0j, t2c, n_o).imag / step
    np.testing.assert_allclose(rho, direct, rtol=0, atol=1e-13)


@pytest.mark.parametrize('n_o,n_v,n_J,seed', [(2, 3, 6, 1), (4, 7, 12, 3), (5, 4, 10, 4), (6, 11, 20, 6)])
def test_batched_ladder_equals_single_batch(n_o, n_v, n_J, seed):
    """eT's (a-batch, b-batch) ladder with the C gathers and scatter (oracle/ladder_pm.c)
    equals the single-batch +/- form for every batch length, including 1 and n_v."""
    s = syn.system(seed, n_o, n_v, n_J, t2_sigma=0.03, t1_sigma=0.03, J_decay=6.0, L_scale=0.3)
    Lt = O.t1_cholesky(s['L'], s['t1'], n_o)
    ref = O.omega_a2_plus_minus(Lt, s['t2'], n_o, n_v)
    for batch in sorted({1, 2, max(1, n_v // 2), n_v}):
        got = O.omega_a2_batched(Lt, s['t2'], n_o, n_v, batch)
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max(), batch
    F = O.fock_t1(s['h'], Lt, s['t1'], n_o)
    full = O.omega(s['L'], Lt, F, s['t1'], s['t2'], n_o)
    assert np.abs(O.omega(s['L'], Lt, F, s['t1'], s['t2'], n_o, ladder_batch=2) - full).max() \
        <= 1e-14 * np.abs(full).max()
