"""Ionisation energies with eT's bath orbital: tests/ccsd_right_ip_energies, ccsd_left_ip_energies,
ccsd_right_ip_diis and ccsd_left_ip_diis.

H2O-He / cc-pVDZ through the device producer at Cholesky threshold 1e-11 (eT: 345 vectors),
the bath orbital appended as the last virtual (zero integrals, orbital energy 0), the FP64
ground state (eT's E_CCSD, 1e-10: the bath orbital does not change it), then two ionised
states on each side by the eigen Davidson and the DIIS excited-state solvers with eT's
ionisation projection and start vectors: eT's iteration counts and ionisation energies
(1e-10 Eh); the eigenvectors have no element without the bath orbital.
"""
import functools
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_ip.json')


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FIXTURE))


def test_ionisation_energies_reproduce_et():
    from et_amd import ccsd, molecule as M, solvers
    f = fixture()
    runs = f['runs']
    g = runs['right_davidson']
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, None, f['bohr_to_angstrom']) as mol:
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-11)
    print('E_nuc %.12f E_HF %.12f n_J %d (eT %d)' % (r['e_nuc'], r['e_hf'], r['n_J'], g['n_cholesky']), flush=True)
    assert r['n_J'] == g['n_cholesky']
    assert abs(r['e_nuc'] - g['nuclear_repulsion']) < 1e-9
    assert abs(r['e_hf'] - g['hf_energy']) < 1e-9
    b = M.with_bath_orbital(r)
    n_o = b['n_o']
    n_v = b['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, b['n_J']) as wf:
        wf.set_cholesky(b['L_mo'])
        wf.set_oei(b['h_mo'])
        wf.set_orbital_energies(b['eps'])
        wf.hf_energy = b['e_hf']
        e = solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        print('E_CCSD %.12f (eT %.12f)' % (e, g['ccsd_energy']), flush=True)
        assert abs(e - g['ccsd_energy']) < 1e-10
        results = {}
        for key in ('right_davidson', 'left_davidson', 'right_diis', 'left_diis'):
            side, kind = key.split('_')
            if kind == 'davidson':
                sv = solvers.EigenDavidson(wf, f['n_states'], side, residual_threshold=f['residual_threshold'],
                                           ionization=True)
            else:
                sv = solvers.DiisCCExcitedStates(wf, f['n_states'], side, residual_threshold=f['residual_threshold'],
                                                 ionization=True)
            w = sv.run()
            results[key] = (w, len(sv.history), sv.eigenvectors)
            print(key, w, len(sv.history), 'iterations; eT', runs[key]['ionisation_energies'],
                  runs[key]['iterations'], flush=True)
    n_t1 = n_o * n_v
    q = np.arange(n_t1 * (n_t1 + 1) // 2)
    Xi = ((np.sqrt(8.0 * q + 1) - 1) // 2).astype(np.int64)
    Yi = q - Xi * (Xi + 1) // 2
    no_bath = (Xi % n_v != n_v - 1) & (Yi % n_v != n_v - 1)
    for key, (w, its, X) in results.items():
        np.testing.assert_allclose(w, runs[key]['ionisation_energies'], rtol=0, atol=1e-10)
        assert its == runs[key]['iterations'], key
        assert np.abs(X[:n_t1][np.arange(n_t1) % n_v != n_v - 1]).max() == 0.0
        assert np.abs(X[n_t1:][no_bath]).max() == 0.0
