"""A complex-conjugate pair of EOM-CCSD roots: eT's tests/ccsd_right_es_energies_imaginary.

HOF / aug-cc-pVDZ at Cholesky threshold 1e-6 through the device producer, the FP64
ground state and the right Davidson solver for four roots.  eT finds two real roots and
the pair 0.327739 +- 0.000164 i (its last Davidson iteration); its own check compares
to 1e-4 Eh.  Here: E_nuc, E_HF and E_CCSD; the two real roots (1e-5); the pair's
real part reported for both states (1e-5); and the pair's imaginary part recovered from
the returned (u, v) columns, whose span is invariant under A: A [u v] = [u v] M with
M = [[a, b], [-b, a]], |b| within 2e-5 of eT's.
"""
import functools
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'hof_complex_pair.json')


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FIXTURE))


def test_complex_pair_of_right_roots_matches_et():
    from et_amd import ccsd, molecule as M, solvers
    f = fixture()
    g = f['golden']
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, M.load_basis('aug-cc-pvdz'), f['bohr_to_angstrom']) as mol:
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-9)
    print('E_nuc %.12f E_HF %.12f n_J %d (eT %d)' % (r['e_nuc'], r['e_hf'], r['n_J'], g['n_cholesky']),
          flush=True)
    assert r['n_o'] == g['n_o']
    assert abs(r['e_nuc'] - g['nuclear_repulsion']) < 1e-9
    assert abs(r['e_hf'] - g['hf_energy']) < 1e-5
    n_v = r['L_mo'].shape[1] - r['n_o']
    with ccsd.CCSD(r['n_o'], n_v, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        wf.hf_energy = r['e_hf']
        e = solvers.DiisCCGroundState(wf, residual_threshold=1e-9).run()
        assert abs(e - g['ccsd_energy']) < 1e-5
        sv = solvers.EigenDavidson(wf, f['n_states'], 'right', residual_threshold=1e-6)
        w = sv.run()
        X = sv.eigenvectors
        AX = wf.jacobian_transformations(X[:, 2:4], 'right')
    print('roots', w, 'eT', g['excitation_energies'], len(sv.history), 'iterations')
    np.testing.assert_allclose(w[:2], g['excitation_energies'][:2], rtol=0, atol=1e-5)
    assert w[2] == w[3]                                 # a pair reports its real part twice
    assert abs(w[2] - g['excitation_energies'][2]) < 1e-5
    UV = X[:, 2:4]
    Mx = np.linalg.lstsq(UV, AX, rcond=None)[0]
    assert np.linalg.norm(AX - UV @ Mx) < 1e-5 * np.linalg.norm(AX)   # the span is invariant
    lam = np.linalg.eigvals(Mx)
    print('pair from the 2 x 2 block:', lam, 'eT', g['last_iteration_re'][2], g['last_iteration_im'][2])
    assert abs(lam.real.mean() - g['last_iteration_re'][2]) < 1e-5
    assert abs(abs(lam.imag).max() - abs(g['last_iteration_im'][2])) < 2e-5
    assert abs(lam.imag).min() > 1e-4                   # a genuine complex pair


def test_root_next_to_a_complex_pair_follows_et():
    """eT's tests/ccsd_es_false_positive_defect: HOF / aug-cc-pVDZ at Cholesky threshold 1e-6,
    three right roots at residual 1e-4, the third next to a complex pair (eT warns about the
    pair (3, 4)).  At this Cholesky threshold the producer's E_HF is 1.9e-6 from eT's (the
    same 420 vectors), so the comparison is at that level: E_CCSD (2e-6), the roots (2e-6),
    eT's 25 Davidson iterations and the last iteration's residual norms (8 %)."""
    from et_amd import ccsd, molecule as M, solvers
    f = json.load(open(os.path.join(os.path.dirname(FIXTURE), 'hof_false_positive.json')))
    g = f['golden']
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, M.load_basis('aug-cc-pvdz'), f['bohr_to_angstrom']) as mol:
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-9)
    print('E_HF %.12f (eT %.12f) n_J %d (eT %d)' % (r['e_hf'], g['hf_energy'], r['n_J'], g['n_cholesky']),
          flush=True)
    assert r['n_J'] == g['n_cholesky']
    assert abs(r['e_hf'] - g['hf_energy']) < 5e-6
    n_o = r['n_o']
    with ccsd.CCSD(n_o, r['L_mo'].shape[1] - n_o, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        wf.hf_energy = r['e_hf']
        e = solvers.DiisCCGroundState(wf, residual_threshold=1e-9).run()
        sv = solvers.EigenDavidson(wf, f['n_states'], 'right', residual_threshold=f['residual_threshold'])
        w = sv.run()
    print('E_CCSD %.12f (eT %.12f); roots %s (eT %s), %d iterations, last residuals %s' % (
        e, g['ccsd_energy'], w, g['excitation_energies'], len(sv.history), sv.history[-1]))
    assert abs(e - g['ccsd_energy']) < 2e-6
    np.testing.assert_allclose(w, g['excitation_energies'], rtol=0, atol=2e-6)
    assert len(sv.history) == g['iterations']
    np.testing.assert_allclose(sv.history[-1], g['last_iteration_residual'], rtol=0.08)
