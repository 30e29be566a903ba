"""Core-valence separation with two nearly degenerate core MOs: eT's tests/ccsd_cvs_eom_degeneracy.

N2 / cc-pVDZ through the device producer at Cholesky threshold 1e-12 (eT: 354 vectors), the
FP64 ground state (eT's E_CCSD), then four right and four left CVS roots out of both N 1s MOs
(core MOs 0 and 1, eT's start vectors a = 0, 1 over both cores) by the eigen Davidson solver at
residual 1e-10: two exactly degenerate pairs, eT's roots to 1e-9 Eh on both sides.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'n2_cvs_degeneracy.json')


def test_degenerate_cvs_roots_reproduce_et():
    from et_amd import ccsd, molecule as M, solvers
    f = json.load(open(FIXTURE))
    g = f['golden']
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, None, f['bohr_to_angstrom']) as mol:
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-11)
    print('E_nuc %.12f E_HF %.12f n_J %d (eT %d)' % (r['e_nuc'], r['e_hf'], r['n_J'], g['n_cholesky']), flush=True)
    assert r['n_J'] == g['n_cholesky']
    assert abs(r['e_nuc'] - g['nuclear_repulsion']) < 1e-9
    assert abs(r['e_hf'] - g['hf_energy']) < 1e-9
    n_o = r['n_o']
    with ccsd.CCSD(n_o, r['L_mo'].shape[1] - n_o, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        wf.hf_energy = r['e_hf']
        e = solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        print('E_CCSD %.12f (eT %.12f)' % (e, g['ccsd_energy']), flush=True)
        assert abs(e - g['ccsd_energy']) < 1e-9
        for side in ('right', 'left'):
            sv = solvers.EigenDavidson(wf, f['n_states'], side, residual_threshold=1e-10, core_mos=f['core_mos'])
            w = sv.run()
            print(side, w, len(sv.history), 'iterations; eT', g[side]['excitation_energies'], g[side]['iterations'],
                  flush=True)
            np.testing.assert_allclose(w, g[side]['excitation_energies'], rtol=0, atol=1e-9)


def test_cvs_with_core_valence_basis_reproduces_et():
    """eT's tests/ccsd_right_cvs_ccX-dz: H2O with ccX-DZ on O and cc-pVDZ on H (one basis per
    element, 48 AOs), Cholesky threshold 1e-10 (eT: 546 vectors), the ground state, and the
    right CVS root out of the O 1s: E_HF (1e-9), E_CCSD (1e-8; eT stops at a residual of 1e-7)
    and the root (1e-6; eT stops at a residual of 1e-5)."""
    from et_amd import ccsd, molecule as M, solvers
    f = json.load(open(os.path.join(os.path.dirname(FIXTURE), 'h2o_ccx_cvs.json')))
    g = f['golden']
    basis = {}
    for sym, _, _, _, name in f['geometry_angstrom_basis']:
        basis[sym] = M.load_basis(name)[sym]
    atoms = [(s, (x, y, z)) for s, x, y, z, _ in f['geometry_angstrom_basis']]
    with M.Molecule(atoms, basis, f['bohr_to_angstrom']) as mol:
        assert mol.dims()['n_ao'] == g['n_ao']
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-10)
    print('E_HF %.12f (eT %.12f) n_J %d (eT %d)' % (r['e_hf'], g['hf_energy'], r['n_J'], g['n_cholesky']), flush=True)
    assert r['n_J'] == g['n_cholesky']
    assert abs(r['e_hf'] - g['hf_energy']) < 1e-9
    n_o = r['n_o']
    with ccsd.CCSD(n_o, r['L_mo'].shape[1] - n_o, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        wf.hf_energy = r['e_hf']
        e = solvers.DiisCCGroundState(wf, residual_threshold=1e-9).run()
        w = solvers.EigenDavidson(wf, 1, 'right', residual_threshold=1e-7, core_mos=f['core_mos']).run()
    print('E_CCSD %.12f (eT %.12f); root %s (eT %s)' % (e, g['ccsd_energy'], w, g['excitation_energies']))
    assert abs(e - g['ccsd_energy']) < 1e-8
    np.testing.assert_allclose(w, g['excitation_energies'], rtol=0, atol=1e-6)
