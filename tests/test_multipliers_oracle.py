"""Pins of the oracle's multiplier (Lambda) pieces and of its ground state against eT's vectors.

eta: eT's eta_mu = <R| e^-T [H, tau_mu] e^T |R> is the derivative of the CC energy
along tau_mu; in eT's packed amplitude basis (doubles diagonal counted once,
A = (dOmega/dt) D2) that is eta = D2 dE/dt, checked by complex-step derivatives
of the oracle's energy.
Multipliers: t-bar^T A + eta = 0 is stationarity of the Lagrangian E + t-bar . Omega
in t, checked along random directions by complex steps (independent of A and eta).
Ground state: the oracle's converged H2O-He amplitudes equal eT's converged ones
(tests/golden/h2o_he_v1_0_vectors.npz, eT's tests/restart_ccsd_v1.0) elementwise,
up to the sign convention of each orbital.
"""
import os

import numpy as np

from et_amd import synthetic as syn
from et_amd import workloads
from oracle import ccsd as O

import signs

V10 = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_v1_0_vectors.npz')


def small(seed, n_o=3, n_v=5, n_J=9):
    return syn.system(seed, n_o, n_v, n_J, t2_sigma=0.05, t1_sigma=0.05, J_decay=6.0, L_scale=0.3)


def energy(s, t1, t2):
    Lt = O.t1_cholesky(s['L'], t1, s['n_o'])
    return O.correlation_energy(Lt, t1, t2, s['n_o'])


def test_eta_is_the_energy_gradient_in_the_packed_basis():
    s = small(1)
    n_o, n_v = s['n_o'], s['n_v']
    n1 = n_o * n_v
    n = n1 + O.n_t2(n_o, n_v)
    h = 1e-30
    grad = np.empty(n)
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        t1 = s['t1'] + 1j * h * e[:n1].reshape(n_v, n_o, order='F')
        t2 = s['t2'] + 1j * h * e[n1:]
        grad[k] = np.imag(energy(s, t1, t2)) / h
    grad[O.doubles_diagonal(n1)] *= 2.0                 # D2 dE/dt
    eta = O.construct_eta(s['L'], s['h'], s['t1'], n_o)
    assert np.abs(eta - grad).max() < 1e-12 * np.abs(grad).max()


def test_dense_multipliers_make_the_lagrangian_stationary():
    s = small(2, n_o=2, n_v=3, n_J=6)
    n_o, n_v = s['n_o'], s['n_v']
    n1 = n_o * n_v
    tbar = O.multipliers_dense(s['L'], s['h'], s['t1'], s['t2'], n_o)
    rng = np.random.default_rng(3)
    h = 1e-30
    for _ in range(3):
        c = rng.normal(size=tbar.size)
        t1 = s['t1'] + 1j * h * c[:n1].reshape(n_v, n_o, order='F')
        t2 = s['t2'] + 1j * h * c[n1:]
        dE = np.imag(energy(s, t1, t2)) / h
        dL = dE + tbar @ (np.imag(O.residual(s['L'], s['h'], t1, t2, n_o)) / h)
        assert abs(dL) < 1e-11 * max(abs(dE), 1.0)


def test_oracle_ground_state_equals_et_converged_amplitudes():
    v = np.load(V10)
    m = workloads.h2o_he_inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    assert np.abs(m['eps'] - v['orbital_energies']).max() < 1e-10
    t1, t2, e, _ = O.solve_ground_state(m['L_mo'], m['h_mo'], m['eps'], n_o, residual_threshold=1e-11)
    ours = np.concatenate([t1.ravel(order='F'), t2])
    s = signs.orbital_signs(v['t'], ours, n_o, n_v)
    assert np.abs(v['t'] - signs.apply_signs(ours, s, n_o, n_v)).max() < 1e-10
