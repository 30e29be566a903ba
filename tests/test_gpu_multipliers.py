"""GPU multiplier (Lambda) equations through the C ABI.

eta, the initial guess and R = A^T t-bar + eta against the oracle; the device
DIIS multiplier solver against the oracle's dense solution on tiny systems and,
for H2O-He / cc-pVDZ, against eT's own converged multipliers
(tests/golden/h2o_he_v1_0_vectors.npz, eT's tests/restart_ccsd_v1.0), up to the
sign convention of each orbital (fixed from the amplitudes).
"""
import os

import numpy as np
import pytest

from et_amd import ccsd, solvers, workloads
from et_amd import synthetic as syn
from oracle import ccsd as O

import signs

pytestmark = pytest.mark.gpu

V10 = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_v1_0_vectors.npz')

CASES = [(2, 3, 6, 2, None), (3, 4, 9, 3, 2), (4, 3, 7, 4, None)]


def make(n_o, n_v, n_J, seed):
    return syn.system(seed, n_o, n_v, n_J, t2_sigma=0.03, t1_sigma=0.03, J_decay=6.0, L_scale=0.3)


def load(s, batch=None):
    wf = ccsd.CCSD(s['n_o'], s['n_v'], s['n_J'])
    wf.set_cholesky(s['L'])
    wf.set_oei(s['h'])
    wf.set_orbital_energies(s['eps'])
    wf.set_amplitudes(np.concatenate([s['t1'].ravel(order='F'), s['t2']]))
    wf.construct_t1_cholesky()
    if batch:
        wf.set_ladder_batch(batch)
    return wf


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize('n_o,n_v,n_J,seed,batch', CASES)
def test_eta_guess_and_multiplier_equation_match_oracle(n_o, n_v, n_J, seed, batch):
    print('case', n_o, n_v, n_J, seed, batch, flush=True)
    s = make(n_o, n_v, n_J, seed)
    A = O.jacobian_matrix(s['L'], s['h'], s['t1'], s['t2'], n_o)
    eta_ref = O.construct_eta(s['L'], s['h'], s['t1'], n_o)
    rng = np.random.default_rng(seed)
    with load(s, batch) as wf:
        wf.prepare_for_multiplier_equation()
        assert rel(wf.construct_eta(), eta_ref) < 1e-13
        wf.set_initial_multipliers_guess()
        assert rel(wf.get_multipliers(), O.initial_multipliers(s['t1'], s['t2'], n_o)) < 1e-15
        tbar = rng.normal(size=A.shape[0])
        wf.set_multipliers(tbar)
        R = wf.construct_multiplier_equation()
        assert rel(R, A.T @ tbar + eta_ref) < 1e-11
        assert abs(wf.multiplier_residual_norm() - np.linalg.norm(R)) < 1e-12 * np.linalg.norm(R)


@pytest.mark.parametrize('n_o,n_v,n_J,seed,batch', CASES[:2])
def test_multiplier_solver_matches_dense_oracle(n_o, n_v, n_J, seed, batch):
    print('case', n_o, n_v, n_J, seed, batch, flush=True)
    s = make(n_o, n_v, n_J, seed)
    ref = O.multipliers_dense(s['L'], s['h'], s['t1'], s['t2'], n_o)
    with load(s, batch) as wf:
        tbar = solvers.DiisCCMultipliers(wf, residual_threshold=1e-12, max_iterations=200).run()
    assert rel(tbar, ref) < 1e-9


def test_multipliers_reproduce_et_converged_multipliers():
    """H2O-He / cc-pVDZ: GPU ground state and multipliers against eT's converged vectors."""
    v = np.load(V10)
    m = workloads.h2o_he_inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['L_mo'].shape[0]) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        t = wf.get_amplitudes()
        tbar = solvers.DiisCCMultipliers(wf, residual_threshold=1e-11).run()
    s = signs.orbital_signs(v['t'], t, n_o, n_v)
    assert np.abs(v['t'] - signs.apply_signs(t, s, n_o, n_v)).max() < 1e-10
    assert np.abs(v['tbar'] - signs.apply_signs(tbar, s, n_o, n_v)).max() < 1e-9


def test_mixed_precision_multipliers_reproduce_et():
    """Split-bf16, FP32, then FP64 left transformations: the FP64 phase decides
    convergence, and the multipliers equal eT's converged ones (1e-9 after orbital
    signs) and the FP64 solver's (1e-10)."""
    v = np.load(V10)
    m = workloads.h2o_he_inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['L_mo'].shape[0]) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        t = wf.get_amplitudes()
        ref = solvers.DiisCCMultipliers(wf, residual_threshold=1e-11).run()
        mixed = solvers.MixedPrecisionMultipliers(wf, residual_threshold=1e-11)
        tbar = mixed.run()
    precisions = [p for p, _ in mixed.history]
    print('phases', [(p, precisions.count(p)) for p in (ccsd.BF16X3, ccsd.FP32, ccsd.FP64)], mixed.noise)
    assert precisions[0] == ccsd.BF16X3 and precisions[-1] == ccsd.FP64
    assert mixed.noise[ccsd.BF16X3] > mixed.noise[ccsd.FP32] > 0
    assert mixed.history[-1][1] < 1e-11
    assert np.abs(tbar - ref).max() < 1e-10
    s = signs.orbital_signs(v['t'], t, n_o, n_v)
    assert np.abs(v['tbar'] - signs.apply_signs(tbar, s, n_o, n_v)).max() < 1e-9
