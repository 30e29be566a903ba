"""Restarting GPU calculations from eT's files (H2O-He / cc-pVDZ, eT's tests/restart_ccsd_v1.0).

eT's v1.0 restart directory, rebuilt from the fixtures and converted to the current
format, supplies everything a restart reads:
* orbital_coefficients: the producer takes eT's orbitals as they are (SCF skipped);
  E_HF equals eT's, so the MO basis is eT's own (no orbital sign ambiguity);
* t: the ground-state solver restarted from it starts on eT's converged amplitudes
  (read bit for bit) and converges at once to eT's CCSD energy; the amplitudes
  saved back read bit for bit;
* tbar: likewise for the multipliers, which stay eT's to 1e-9;
* r_00k / l_00k: the Davidson solvers restarted from them converge in fewer
  iterations than from orbital differences, to the excitation energies the files
  carry; the left solver restarted from the right vectors alone (eT's right -> left
  map, pinned on the CPU against its 4-index definition) converges to the same
  energies;
* a singles-only t file (eT's CC2 v1.0 file) gives the no-restart guess, and a
  singles-only tbar file keeps its singles with the doubles of the guess, which
  equal the host's packed contravariant map of t2.
"""
import functools
import json
import os
import shutil

import numpy as np
import pytest

import v1_0_files as V

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_ccpvdz.json')


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FIXTURE))


@pytest.fixture(scope='module')
def et_dir(tmp_path_factory):
    from et_amd import restart as R
    base = tmp_path_factory.mktemp('v1_0')
    new = str(base / 'new')
    R.convert_v1_0_directory(V.build(str(base / 'old')), new)
    return new


@pytest.fixture(scope='module')
def inputs(et_dir):
    from et_amd import molecule as M
    f = fixture()
    atoms = [(s, (x, y, z)) for s, x, y, z in f['geometry_angstrom']]
    with M.Molecule(atoms, f['basis'], f['bohr_to_angstrom']) as mol:
        mol.read_orbitals(os.path.join(et_dir, 'orbital_coefficients'))
        return mol.run(cholesky_threshold=f['cholesky_threshold'], scf_threshold=1e-8)


def wavefunction(m):
    from et_amd import ccsd
    n_o = m['n_o']
    wf = ccsd.CCSD(n_o, m['L_mo'].shape[1] - n_o, m['n_J'])
    wf.set_cholesky(m['L_mo'])
    wf.set_oei(m['h_mo'])
    wf.set_orbital_energies(m['eps'])
    wf.hf_energy = m['e_hf']
    return wf


def test_producer_takes_ets_orbitals(et_dir, inputs):
    from et_amd import restart as R
    g = fixture()['golden']
    C, eps = R.OrbitalFile(os.path.join(et_dir, 'orbital_coefficients')).read()
    assert inputs['scf_iterations'] == 1
    assert np.array_equal(inputs['C'], C) and np.array_equal(inputs['eps'], eps)
    assert abs(inputs['e_hf'] - g['hf_energy']) < 1e-9
    assert inputs['n_J'] == g['n_cholesky']


def test_ground_state_and_multipliers_restart_from_ets_files(et_dir, inputs, tmp_path):
    from et_amd import solvers
    g = fixture()['golden']
    v = V.vectors()
    with wavefunction(inputs) as wf:
        wf.initialize_files(et_dir)
        wf.set_initial_amplitudes_guess(restart=True)
        assert np.array_equal(wf.get_amplitudes(), v['t'])
        gs = solvers.DiisCCGroundState(wf, residual_threshold=1e-9)
        gs.run(restart=True)
        print('restarted ground state: %d iterations, |Omega| %.2e' % (len(gs.history), gs.history[0][1]))
        assert len(gs.history) == 1
        assert abs(wf.energy - g['ccsd_energy']) < 1e-9

        mult = solvers.DiisCCMultipliers(wf, residual_threshold=1e-9)
        tbar = mult.run(restart=True)
        print('restarted multipliers: %d iterations, |R| %.2e' % (len(mult.history), mult.history[0]))
        assert len(mult.history) == 1
        assert np.abs(tbar - v['tbar']).max() < 1e-9
        # eT's default multiplier solver (linear Davidson), as eT's v1.0 test ran it
        dav = solvers.DavidsonCCMultipliers(wf, residual_threshold=1e-9)
        tbar = dav.run(restart=True)
        assert len(dav.history) == 1
        assert np.abs(tbar - v['tbar']).max() < 1e-9

        out = wf.initialize_files(str(tmp_path / 'out'))
        wf.save_amplitudes()
        wf.save_multipliers()
        x, e, n = out.t.read()
        assert n == wf.n_gs_amplitudes and e == wf.energy and np.array_equal(x, wf.get_amplitudes())
        x, e, n = out.tbar.read()
        assert np.array_equal(x, wf.get_multipliers())


def test_excited_states_restart_from_ets_files(et_dir, inputs, tmp_path):
    from et_amd import solvers
    w_ref = V.vectors()['excitation_energies']
    right_only = str(tmp_path / 'right_only')
    os.makedirs(right_only)
    for k in (1, 2, 3):
        shutil.copy(os.path.join(et_dir, 'r_%03d' % k), right_only)
    with wavefunction(inputs) as wf:
        wf.initialize_files(et_dir)
        solvers.DiisCCGroundState(wf, residual_threshold=1e-10).run(restart=True)
        its = {}
        for side in ('right', 'left'):
            cold = solvers.EigenDavidson(wf, 3, side, residual_threshold=1e-9)
            w_cold = cold.run()
            warm = solvers.EigenDavidson(wf, 3, side, residual_threshold=1e-9, restart=True)
            w_warm = warm.run()
            its[side] = (len(cold.history), len(warm.history))
            np.testing.assert_allclose(w_warm, w_ref, rtol=0, atol=1e-10)
            np.testing.assert_allclose(w_cold, w_ref, rtol=0, atol=1e-10)
        wf.initialize_files(right_only)
        cross = solvers.EigenDavidson(wf, 3, 'left', residual_threshold=1e-9, restart=True)
        w_cross = cross.run()
        its['left from right'] = (its['left'][0], len(cross.history))
        print('Davidson iterations (orbital differences, restart):', its)
        np.testing.assert_allclose(w_cross, w_ref, rtol=0, atol=1e-10)
        for side in ('right', 'left'):
            assert its[side][1] < its[side][0], side
        # the saved states read back with their energies
        out = wf.initialize_files(str(tmp_path / 'out'))
        wf.save_excited_state(cross.eigenvectors, 1, 3, 'left', w_cross)
        X, e = wf.read_excited_state(1, 3, 'left')
        assert np.array_equal(X, cross.eigenvectors) and np.array_equal(e, w_cross)
        assert out.left(3).exists() and not out.right(1).exists()


def test_singles_only_files(inputs, tmp_path):
    from et_amd import restart as R
    with wavefunction(inputs) as wf:
        n1 = wf.n_t1
        old = V.build(str(tmp_path / 'cc2_old'), 'cc2')
        new = str(tmp_path / 'cc2_new')
        R.convert_v1_0_directory(old, new)
        wf.set_initial_amplitudes_guess()
        cold = wf.get_amplitudes()
        wf.set_amplitudes(np.full(wf.n_gs_amplitudes, 0.01))
        files = wf.initialize_files(new)
        assert files.t.read()[2] == n1
        wf.set_initial_amplitudes_guess(restart=True)
        assert np.array_equal(wf.get_amplitudes(), cold)

        rng = np.random.default_rng(4)
        t1bar = rng.normal(size=n1)
        R.AmplitudeFile(files.tbar.path, [n1]).save(t1bar)
        wf.set_initial_multipliers_guess(restart=True)
        tbar = wf.get_multipliers()
        assert np.array_equal(tbar[:n1], t1bar)
        np.testing.assert_allclose(tbar[n1:], R.packed_contravariant(cold[n1:], wf.n_v, wf.n_o),
                                   rtol=0, atol=1e-15)
