"""Device quasi-Newton step and DIIS extrapolation against the oracle's restatement.

The oracle's DIIS and quasi-Newton step reproduce eT's 19-iteration H2O-He
trajectory row by row (tests/test_golden_h2o_he.py), which pins them; here the
device versions, driven through the C ABI, follow the oracle's trajectory on
synthetic systems, including DIIS spaces that overflow (oldest record dropped).
"""
import numpy as np
import pytest

from et_amd import ccsd
from et_amd import solvers
from et_amd import synthetic as syn
from oracle import ccsd as O

pytestmark = pytest.mark.gpu


def make(n_o, n_v, n_J, seed):
    return syn.system(seed, n_o, n_v, n_J, t2_sigma=0.02, t1_sigma=0.02, J_decay=6.0, L_scale=0.3)


def load(s):
    wf = ccsd.CCSD(s['n_o'], s['n_v'], s['n_J'])
    wf.set_cholesky(s['L'])
    wf.set_oei(s['h'])
    wf.set_orbital_energies(s['eps'])
    return wf


def test_quasi_newton_step_matches_oracle():
    s = make(3, 7, 12, 3)
    n_o, n_v = 3, 7
    amps = np.concatenate([s['t1'].ravel(order='F'), s['t2']])
    with load(s) as wf:
        wf.set_amplitudes(amps)
        wf.construct_t1_cholesky()
        wf.construct_fock()
        om = wf.construct_omega()
        wf.quasi_newton_update()
        got = wf.get_amplitudes()
        with pytest.raises(ccsd.ETGError):          # the residual is stale after the step
            wf.quasi_newton_update()
    ref, _ = O.quasi_newton_step(amps, om, s['eps'], n_o, n_o * n_v)
    assert np.abs(got - ref).max() < 1e-14 * np.abs(ref).max()


def oracle_trajectory(s, iterations, dimension):
    n_o = s['n_o']
    n_v = s['n_v']
    L, h, eps = s['L'], s['h'], s['eps']
    t1 = np.zeros((n_v, n_o))
    L_t1 = O.t1_cholesky(L, t1, n_o)
    t2 = O.mp2_guess(L_t1, eps, n_o)
    diis = O.DIIS(dimension)
    history = []
    for _ in range(iterations):
        F = O.fock_t1(h, L_t1, t1, n_o)
        e = O.correlation_energy(L_t1, t1, t2, n_o)
        om = O.omega(L, L_t1, F, t1, t2, n_o)
        history.append((e, float(np.linalg.norm(om))))
        amps = np.concatenate([t1.ravel(order='F'), t2])
        amps, r = O.quasi_newton_step(amps, om, eps, n_o, n_o * n_v)
        amps = diis.update(r, amps)
        t1 = amps[:n_o * n_v].reshape(n_v, n_o, order='F')
        t2 = amps[n_o * n_v:]
        L_t1 = O.t1_cholesky(L, t1, n_o)
    return history


@pytest.mark.parametrize('dimension', [3, 8])
def test_device_diis_trajectory_matches_oracle(dimension):
    print('dimension', dimension, flush=True)
    s = make(3, 8, 14, 5)
    iterations = 10
    ref = oracle_trajectory(s, iterations, dimension)
    with load(s) as wf:
        wf.hf_energy = 0.0
        solver = solvers.DiisCCGroundState(wf, residual_threshold=0.0, max_iterations=iterations,
                                           diis_dimension=dimension)
        with pytest.raises(ccsd.ETGError):
            solver.run()
        got = solver.history
    assert len(got) == iterations
    for (e, n, _), (re, rn) in zip(got, ref):
        assert abs(e - re) < 1e-12
        assert abs(n - rn) < 1e-9 * rn
    assert ref[-1][1] < 0.1 * ref[0][1]           # the trajectory does converge


def test_diis_requires_initialisation_and_update():
    s = make(2, 3, 5, 4)
    with load(s) as wf:
        wf.set_initial_amplitudes_guess()
        with pytest.raises(ccsd.ETGError):
            wf.diis_update()
        wf.diis_initialize(4)
        with pytest.raises(ccsd.ETGError):
            wf.diis_update()                          # no quasi-Newton step yet
        with pytest.raises(ccsd.ETGError):
            wf.diis_initialize(0)
