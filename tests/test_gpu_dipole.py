"""CC mean values of one-electron operators on the device.

* <X> with the Lagrangian density through the C ABI equals the oracle's
  complex-step derivative of L = E + t-bar . Omega along h -> h + e X, at the
  device's own amplitudes and multipliers (synthetic system, random X), and the
  CC density from one backward replay gives sum_pq D_pq X_pq equal to it (1e-12);
* the whole chain -- producer (STO-3G), ground state, multipliers, dipole
  integrals, mean value -- reproduces eT's tests/ccsd_dipole: E_HF, n_J, E_CCSD,
  the multiplier solver's run (eT's default linear Davidson: 14 iterations, every
  residual norm to the printed digits; its multipliers equal the DIIS solver's), and
  the electronic, nuclear and total dipole in Debye.
"""
import functools
import json
import os

import numpy as np
import pytest

from et_amd import synthetic as syn
from oracle import ccsd as O

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_sto3g_dipole.json')


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FIXTURE))


def test_expectation_value_matches_oracle_complex_step():
    from et_amd import ccsd, solvers
    s = syn.system(7, 3, 5, 10, t2_sigma=0.03, t1_sigma=0.03, J_decay=6.0, L_scale=0.3)
    n_o, n_v = 3, 5
    n = n_o + n_v
    with ccsd.CCSD(n_o, n_v, s['n_J']) as wf:
        wf.set_cholesky(s['L'])
        wf.set_oei(s['h'])
        wf.set_orbital_energies(s['eps'])
        wf.set_amplitudes(np.concatenate([s['t1'].ravel(order='F'), s['t2']]))
        wf.construct_t1_cholesky()
        wf.prepare_for_jacobian()
        rng = np.random.default_rng(3)
        tbar = rng.normal(size=wf.n_gs_amplitudes) * 0.05
        wf.set_multipliers(tbar)
        t = wf.get_amplitudes()
        D = wf.one_electron_density()
        for k in range(3):
            X = rng.normal(size=(n, n))
            X = X + X.T
            got = wf.one_electron_expectation(X)
            ref = O.one_electron_expectation(s['L'], s['h'], t[:n_o * n_v].reshape(n_v, n_o, order='F'),
                                             t[n_o * n_v:], tbar, n_o, X)
            assert abs(got - ref) < 1e-11 * max(1.0, abs(ref)), k
            # the density from one backward replay gives the same value for every X
            assert abs(np.sum(D * X) - got) < 1e-12 * max(1.0, abs(got)), k


def test_dipole_chain_reproduces_et():
    from et_amd import ccsd, molecule as M, properties, solvers
    f = fixture()
    g = f['golden']
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, f['basis'], f['bohr_to_angstrom']) as mol:
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-12)
        assert abs(r['e_nuc'] - g['nuclear_repulsion']) < 1e-11
        assert abs(r['e_hf'] - g['hf_energy']) < 1e-10
        assert r['n_J'] == g['n_cholesky']
        n_o = r['n_o']
        n_v = r['L_mo'].shape[1] - n_o
        with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
            wf.set_cholesky(r['L_mo'])
            wf.set_oei(r['h_mo'])
            wf.set_orbital_energies(r['eps'])
            solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
            assert abs(r['e_hf'] + wf.correlation_energy - g['ccsd_energy']) < 1e-10
            # eT's default CCSD multiplier solver, as in the eT run: linear Davidson
            dav = solvers.DavidsonCCMultipliers(wf, residual_threshold=1e-11)
            tbar = dav.run()
            mu = properties.dipole_moment(wf, mol, r['C'])
            ref_tbar = solvers.DiisCCMultipliers(wf, residual_threshold=1e-12, max_iterations=200).run()
    ref_norms = g['multiplier_davidson_residuals']
    print('multiplier Davidson', len(dav.history), 'iterations (eT %d)' % len(ref_norms))
    assert len(dav.history) == len(ref_norms)
    for it, (r, r_ref) in enumerate(zip(dav.history, ref_norms), 1):
        assert abs(r - r_ref) <= 6e-4 * r_ref + 1e-13, (it, r, r_ref)
    assert np.abs(tbar - ref_tbar).max() < 1e-10
    for c, comp in enumerate('xyz'):
        ref = g['dipole_debye'][comp]
        assert abs(mu['nuclear_debye'][c] - ref['nuclear']) < 1e-9
        assert abs(mu['electronic_debye'][c] - ref['electronic']) < 1e-8, comp
        assert abs(mu['total_debye'][c] - ref['total']) < 1e-8, comp
