"""eT's Newton-Raphson CCSD ground state: tests/ccsd_gs_energy_nr and tests/ccsd_gs_energy_nr_crop.

The molecule goes through the device producer at Cholesky threshold 1e-10 (eT's number of
vectors), then NewtonRaphsonCCGroundState with eT's settings:
  H2O-He / cc-pVDZ: relative micro threshold 1e-2, residual and energy thresholds 1e-10,
      DIIS dimension 8; eT needs six macro iterations, the fifth meets the residual
      threshold but not the energy threshold;
  ozone / cc-pVDZ: residual threshold 1e-8, the CROP variant of DIIS with dimension 3;
      eT needs five macro iterations.
The macro energies follow eT's (1e-10; 1e-8 for ozone, whose RHF eT converges only to a
gradient of 1e-9), |Omega| and every linear iteration's residual norm the 4 digits eT
prints (while above 1e-9: smaller ones are set by round-off and the Cholesky vectors), and
each update takes eT's number of linear iterations.
"""
import functools
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
CASES = {'h2o_he': ('h2o_he_newton_raphson.json', 1e-10), 'o3_crop': ('o3_newton_raphson_crop.json', 1e-8)}


@functools.lru_cache(maxsize=None)
def fixture(name):
    return json.load(open(os.path.join(GOLDEN, CASES[name][0])))


@pytest.mark.parametrize('name', sorted(CASES))
def test_newton_raphson_ground_state_follows_et(name):
    from et_amd import ccsd, molecule as M, solvers
    print('case', name, flush=True)
    f = fixture(name)
    g = f['golden']
    e_tol = CASES[name][1]
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, None, f['bohr_to_angstrom']) as mol:
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-10)
    print('E_nuc %.12f E_HF %.12f n_J %d (eT %d)' % (r['e_nuc'], r['e_hf'], r['n_J'], g['n_cholesky']), flush=True)
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
        nr = solvers.NewtonRaphsonCCGroundState(wf, residual_threshold=f['residual_threshold'],
                                                energy_threshold=f['energy_threshold'],
                                                diis_dimension=f['diis_dimension'], crop=f['crop'],
                                                rel_threshold=f['relative_micro_threshold'], log=print)
        e = nr.run()
    for k, m in enumerate(nr.micro):
        print('update %d: %s' % (k + 1, ' '.join('%.4e' % x for x in m)))
    print('E_CCSD %.12f (eT %.12f)' % (e, g['ccsd_energy']))
    assert abs(e - g['ccsd_energy']) < e_tol
    macro = np.array(g['macro_iterations'])
    assert len(nr.history) == len(macro)
    ours = np.array(nr.history)
    assert np.abs(ours[:, 0] - macro[:, 1]).max() < e_tol
    big = macro[:, 2] > 1e-9
    np.testing.assert_allclose(ours[big, 1], macro[big, 2], rtol=1e-3)
    assert [len(m) for m in nr.micro] == [len(m) for m in g['micro_residual_norms']]
    for m, ref in zip(nr.micro, g['micro_residual_norms']):
        m, ref = np.array(m), np.array(ref)
        np.testing.assert_allclose(m[ref > 1e-9], ref[ref > 1e-9], rtol=1e-3)


def test_newton_raphson_with_split_bf16_linear_solves():
    """The linear solves of a Newton-Raphson step need only ||A u + Omega|| < 1e-2 ||Omega||, far
    above the split-bf16 transformations' relative error (~2e-5), so they can run in split-bf16
    while the residual and the convergence test stay FP64: H2O-He (eT's ccsd_gs_energy_nr
    settings) reaches the FP64 solver's energy (1e-10) in as many macro iterations, +-1."""
    from et_amd import ccsd, molecule as M, solvers
    f = fixture('h2o_he')
    g = f['golden']
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, None, f['bohr_to_angstrom']) as mol:
        r = mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-10)
    n_v = r['L_mo'].shape[1] - r['n_o']
    runs = {}
    for name, prec in (('fp64', ccsd.FP64), ('bf16x3', ccsd.BF16X3)):
        with ccsd.CCSD(r['n_o'], n_v, r['n_J']) as wf:
            wf.set_cholesky(r['L_mo'])
            wf.set_oei(r['h_mo'])
            wf.set_orbital_energies(r['eps'])
            wf.hf_energy = r['e_hf']
            nr = solvers.NewtonRaphsonCCGroundState(wf, residual_threshold=f['residual_threshold'],
                                                    energy_threshold=f['energy_threshold'],
                                                    rel_threshold=f['relative_micro_threshold'], precision=prec)
            runs[name] = (nr.run(), len(nr.history), [len(m) for m in nr.micro])
        print(name, runs[name], flush=True)
    assert abs(runs['bf16x3'][0] - runs['fp64'][0]) < 1e-10
    assert abs(runs['bf16x3'][0] - g['ccsd_energy']) < 1e-10
    assert abs(runs['bf16x3'][1] - runs['fp64'][1]) <= 1
