"""More of eT's CCSD test molecules through the device producer and ground state.

tests/golden/cc_pvdz_molecules.json holds eT's own results for H2O (tests/ccsd_lr_polarizability,
Cholesky threshold 1e-11) and ozone (tests/ccsd_gs_energy_nr_crop, 1e-10), both cc-pVDZ:
nuclear repulsion, RHF energy, number of Cholesky vectors, CCSD energy, and for H2O the
16-row DIIS table.  The producer (integrals, pivoted Cholesky, RHF) and the FP64 DIIS
ground state on its vectors reproduce them; the water's DIIS energies follow eT's row
by row.
"""
import functools
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'cc_pvdz_molecules.json')


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FIXTURE))


@pytest.mark.parametrize('name', ['h2o', 'o3'])
def test_producer_and_ground_state_reproduce_et(name):
    from et_amd import ccsd, molecule as M, solvers
    f = fixture()['molecules'][name]
    g = f['golden']
    print('case', name, flush=True)
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, None, fixture()['bohr_to_angstrom']) as mol:
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-10)
    print('E_nuc %.12f E_HF %.12f n_J %d (eT %d)' % (r['e_nuc'], r['e_hf'], r['n_J'], g['n_cholesky']))
    assert r['n_o'] == g['n_o']
    assert r['n_J'] == g['n_cholesky']
    assert abs(r['e_nuc'] - g['nuclear_repulsion']) < 1e-9
    assert abs(r['e_hf'] - g['hf_energy']) < 1e-8
    n_v = r['L_mo'].shape[1] - r['n_o']
    with ccsd.CCSD(r['n_o'], n_v, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        wf.hf_energy = r['e_hf']
        gs = solvers.DiisCCGroundState(wf, residual_threshold=1e-10, max_iterations=60)
        e = gs.run()
    print('E_CCSD %.12f (eT %.12f), %d iterations' % (e, g['ccsd_energy'], len(gs.history)))
    assert abs(e - g['ccsd_energy']) < 1e-8
    if 'ccsd_iterations' in g:
        ref = np.array([row[1] for row in g['ccsd_iterations']])
        ours = np.array([h[0] for h in gs.history])
        assert len(ours) == len(ref)
        assert np.abs(ours - ref).max() < 1e-9
