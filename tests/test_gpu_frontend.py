"""The molecular producer on the device against the oracle's and eT's numbers.

H2O-He / cc-pVDZ (eT's tests/ccsd_gs_energy system, fixture
tests/golden/h2o_he_ccpvdz.json): the one-electron integrals and the dense
(pq|rs) matrix equal the oracle's McMurchie-Davidson integrals elementwise;
E_nuc, E_HF and the number of Cholesky vectors at eT's threshold equal eT's
printed values; and the CCSD ground state run on the producer's Cholesky
vectors reaches eT's E_CCSD.
"""
import functools
import json
import os

import numpy as np
import pytest

from oracle import frontend as fe

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_ccpvdz.json')


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FIXTURE))


def molecule():
    from et_amd import molecule as M
    f = fixture()
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    return M.Molecule(atoms, f['basis'], f['bohr_to_angstrom'])


@functools.lru_cache(maxsize=1)
def oracle_integrals():
    f = fixture()
    shells, atoms = fe.build_molecule(f['geometry_angstrom'], f['basis'], f['bohr_to_angstrom'])
    S, T, V = fe.one_electron(shells, atoms)
    return S, T, V, fe.eri(shells)


def test_one_electron_integrals_match_oracle():
    S0, T0, V0, _ = oracle_integrals()
    with molecule() as m:
        S, T, V = m.one_electron()
    assert np.abs(S - S0).max() < 1e-13
    assert np.abs(T - T0).max() < 1e-12
    assert np.abs(V - V0).max() < 1e-11


def test_electron_repulsion_matrix_matches_oracle():
    *_, G0 = oracle_integrals()
    n = G0.shape[0]
    with molecule() as m:
        G = m.ao_eri()
    ref = G0.reshape(n * n, n * n, order='F')
    assert np.abs(G - ref).max() < 1e-12
    assert np.array_equal(G, G.T)                  # (pq|rs) = (rs|pq) written exactly


def test_chunked_integral_launches_write_the_same_matrix(monkeypatch):
    """Large molecules launch the quartets in chunks (a dispatch counts its work-items
    in 32 bits); chunks of 997 quartets give the one-launch matrix bit for bit."""
    with molecule() as m:
        G = m.ao_eri()
    monkeypatch.setenv('ETG_ERI_LAUNCH_GROUPS', '997')
    with molecule() as m:
        Gc = m.ao_eri()
    assert np.array_equal(G, Gc)


def test_producer_reproduces_et_hf_energy_and_cholesky_count():
    f = fixture()
    g = f['golden']
    with molecule() as m:
        r = m.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-12)
    assert abs(r['e_nuc'] - g['nuclear_repulsion']) < 1e-11
    assert abs(r['e_hf'] - g['hf_energy']) < 1e-10
    assert r['n_J'] == g['n_cholesky']
    # MO orthonormality of the producer's orbitals in the AO metric, and h_mo symmetric
    S0, *_ = oracle_integrals()
    C = r['C']
    assert np.abs(C.T @ S0 @ C - np.eye(C.shape[0])).max() < 1e-10
    assert np.abs(r['h_mo'] - r['h_mo'].T).max() < 1e-12


def test_producer_cholesky_vectors_reproduce_et_ccsd_energy():
    from et_amd import ccsd, solvers
    f = fixture()
    g = f['golden']
    with molecule() as m:
        r = m.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-12)
    n_o = r['n_o']
    n_v = r['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        e = r['e_hf'] + wf.correlation_energy
    assert abs(e - g['ccsd_energy']) < 1e-10


@pytest.mark.gpu
def test_cartesian_gaussians_reproduce_et():
    """eT's tests/ccsd_right_es_energies_cartgauss: H2O-He / cc-pVDZ with Cartesian d functions
    (30 AOs), Cholesky threshold 1e-12: the AO count, E_nuc, E_HF (1e-9), eT's 362 vectors,
    E_CCSD (1e-10) and the two right EOM-CCSD roots (1e-9), all different from the spherical
    basis's; the overlap matrix is positive definite with unit diagonal."""
    import json
    from et_amd import ccsd, molecule as M, solvers
    f = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_cartesian.json')))
    g = f['golden']
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, None, f['bohr_to_angstrom'], cartesian=True) as mol:
        assert mol.dims()['n_ao'] == g['n_ao']
        S = mol.one_electron()[0]
        assert np.abs(np.diag(S) - 1.0).max() < 1e-12
        assert np.linalg.eigvalsh(S).min() > 0.0
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-10)
    print('E_nuc %.12f E_HF %.12f n_J %d (eT %d)' % (r['e_nuc'], r['e_hf'], r['n_J'], g['n_cholesky']), flush=True)
    assert abs(r['e_nuc'] - g['nuclear_repulsion']) < 1e-9
    assert abs(r['e_hf'] - g['hf_energy']) < 1e-9
    assert r['n_J'] == g['n_cholesky']
    n_o = r['n_o']
    with ccsd.CCSD(n_o, r['L_mo'].shape[1] - n_o, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        wf.hf_energy = r['e_hf']
        e = solvers.DiisCCGroundState(wf, residual_threshold=1e-10).run()
        w = solvers.EigenDavidson(wf, 2, 'right', residual_threshold=1e-9).run()
    print('E_CCSD %.12f (eT %.12f); roots %s (eT %s)' % (e, g['ccsd_energy'], w, g['excitation_energies']))
    assert abs(e - g['ccsd_energy']) < 1e-10
    np.testing.assert_allclose(w, g['excitation_energies'], rtol=0, atol=1e-9)
