"""eT's own golden values for H2O-He / cc-pVDZ (tests/ccsd_gs_energy in the reference).

The fixture tests/golden/h2o_he_ccpvdz.json holds the basis, geometry and the
reference outputs (energies, the 19-iteration DIIS table, excitation energies)
with their citations; tests/golden/make_h2o_he_fixture.py wrote it from eT's
test files.  The molecular inputs come from oracle/frontend.py (integrals ->
RHF -> pivoted Cholesky at eT's threshold 1e-12).

CPU: the oracle reproduces E_nuc, E_HF, the number of Cholesky vectors, every
row of eT's DIIS iteration table and E_CCSD.
GPU: the same solve through the C ABI reproduces E_CCSD and the table.
"""
import functools
import json
import os

import numpy as np
import pytest

from oracle import ccsd as O
from oracle import frontend as fe

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_ccpvdz.json')


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FIXTURE))


@functools.lru_cache(maxsize=1)
def inputs():
    return fe.molecular_inputs(fixture())


def test_reference_energies_and_cholesky_count():
    g = fixture()['golden']
    m = inputs()
    assert abs(m['e_nuc'] - g['nuclear_repulsion']) < 1e-11
    assert abs(m['e_hf'] - g['hf_energy']) < 1e-10
    assert m['n_J'] == g['n_cholesky']
    F0 = O.fock_t1(m['h_mo'], m['L_mo'], np.zeros((fixture()['n_v'], m['n_o'])), m['n_o'])
    assert np.abs(F0 - np.diag(m['eps'])).max() < 1e-10      # canonical RHF through the Cholesky vectors


def test_committed_inputs_match_the_front_end():
    """tests/golden/h2o_he_ccpvdz_inputs.npz (read by et_amd.workloads) is the front end's output."""
    from et_amd import workloads
    m = inputs()
    w = workloads.h2o_he_inputs()
    assert w['n_o'] == m['n_o'] and w['L_mo'].shape == m['L_mo'].shape
    assert np.abs(w['L_mo'] - m['L_mo']).max() < 1e-14
    assert np.abs(w['h_mo'] - m['h_mo']).max() < 1e-14
    assert np.abs(w['eps'] - m['eps']).max() < 1e-14
    assert abs(w['e_hf'] - m['e_hf']) < 1e-12


def test_replicated_system_is_block_diagonal():
    """Two copies: the oracle's energy of the MP2 guess is twice the single one."""
    from et_amd import workloads
    w = workloads.h2o_he_inputs()
    r = workloads.replicate(w, 2)
    e1 = O.correlation_energy(w['L_mo'], np.zeros((w['L_mo'].shape[1] - w['n_o'], w['n_o'])),
                              O.mp2_guess(w['L_mo'], w['eps'], w['n_o']), w['n_o'])
    e2 = O.correlation_energy(r['L_mo'], np.zeros((r['L_mo'].shape[1] - r['n_o'], r['n_o'])),
                              O.mp2_guess(r['L_mo'], r['eps'], r['n_o']), r['n_o'])
    assert abs(e2 - 2 * e1) < 1e-12


def test_oracle_reproduces_et_diis_trajectory():
    g = fixture()['golden']
    m = inputs()
    t1, t2, e, hist = O.solve_ground_state(m['L_mo'], m['h_mo'], m['eps'], m['n_o'],
                                           residual_threshold=1e-10)
    table = g['ccsd_diis_iterations']
    assert len(hist) == len(table)
    for (it, e_ref, norm_ref, _), (e_corr, norm) in zip(table, hist):
        assert abs(m['e_hf'] + e_corr - e_ref) < 1e-10, it
        # eT prints |omega| with 4 significant digits
        assert abs(norm - norm_ref) <= 0.51e-4 * 10 ** np.floor(np.log10(norm_ref)) * 10, it
    assert abs(m['e_hf'] + e - g['ccsd_energy']) < 1e-10
    assert abs(e - g['ccsd_correlation_energy']) < 1e-10


@pytest.mark.gpu
def test_gpu_ground_state_reproduces_et():
    from et_amd import ccsd, solvers
    g = fixture()['golden']
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        wf.hf_energy = m['e_hf']
        solver = solvers.DiisCCGroundState(wf, residual_threshold=1e-10)
        e = solver.run()
    table = g['ccsd_diis_iterations']
    assert len(solver.history) == len(table)
    for (it, e_ref, norm_ref, _), (e_it, norm, _) in zip(table, solver.history):
        assert abs(e_it - e_ref) < 1e-10, it
    assert abs(e - g['ccsd_energy']) < 1e-10


@pytest.mark.gpu
def test_gpu_replicated_ground_state_is_size_extensive():
    """Three non-interacting copies: E_corr = 3 x eT's, along the single system's 19 iterations."""
    from et_amd import ccsd, solvers, workloads
    g = fixture()['golden']
    r = workloads.replicate(workloads.h2o_he_inputs(), 3)
    n_v = r['L_mo'].shape[1] - r['n_o']
    with ccsd.CCSD(r['n_o'], n_v, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        solver = solvers.DiisCCGroundState(wf, residual_threshold=1e-10 * np.sqrt(3))
        solver.run()
        e_corr = wf.correlation_energy
    assert len(solver.history) == len(g['ccsd_diis_iterations'])
    assert abs(e_corr - 3 * g['ccsd_correlation_energy']) < 3e-10


@pytest.mark.gpu
def test_gpu_mixed_precision_ground_state_reproduces_et():
    """Split-bf16 and FP32 residual phases, finished on FP64 residuals: eT's E_CCSD to 1e-10."""
    from et_amd import ccsd, solvers
    g = fixture()['golden']
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        wf.hf_energy = m['e_hf']
        solver = solvers.MixedPrecisionGroundState(wf, residual_threshold=1e-10)
        e = solver.run()
    phases = [p for p, _, _ in solver.history]
    assert phases[0] == ccsd.BF16X3 and phases[-1] == ccsd.FP64 and ccsd.FP32 in phases
    assert solver.noise[ccsd.BF16X3] > solver.noise[ccsd.FP32] > 0
    assert abs(e - g['ccsd_energy']) < 1e-10


@functools.lru_cache(maxsize=1)
def ground_state():
    m = inputs()
    t1, t2, e, hist = O.solve_ground_state(m['L_mo'], m['h_mo'], m['eps'], m['n_o'],
                                           residual_threshold=1e-11)
    return t1, t2


def test_oracle_jacobian_reproduces_et_excitation_energies():
    """Lowest eigenvalues of the oracle's A (complex-step derivative of Omega along D2 c)
    equal eT's EOM-CCSD excitation energies (right and left transformations share A)."""
    g = fixture()['golden']
    m = inputs()
    t1, t2 = ground_state()
    diag = O.orbital_differences(m['eps'], m['n_o'])
    w = O.davidson_lowest(lambda x: O.jacobian_transform(m['L_mo'], m['h_mo'], t1, t2, m['n_o'], x),
                          diag, 2, tol=1e-9)
    np.testing.assert_allclose(w, g['right_excitation_energies'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(w, g['left_excitation_energies'], rtol=0, atol=1e-10)


@pytest.mark.gpu
def test_gpu_eom_reproduces_et_excitation_energies():
    """GPU ground state, then right and left Davidson with the GPU Jacobian transformations."""
    from et_amd import ccsd, solvers
    g = fixture()['golden']
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        right = solvers.EigenDavidson(wf, 2, 'right', residual_threshold=1e-8).run()
        left = solvers.EigenDavidson(wf, 2, 'left', residual_threshold=1e-8).run()
    np.testing.assert_allclose(right, g['right_excitation_energies'], rtol=0, atol=1e-10)
    np.testing.assert_allclose(left, g['left_excitation_energies'], rtol=0, atol=1e-10)


def v1_0_excitation_energies():
    path = os.path.join(os.path.dirname(FIXTURE), 'h2o_he_v1_0_vectors.npz')
    return np.load(path)['excitation_energies']


@pytest.mark.gpu
def test_gpu_mixed_precision_eom_reproduces_et():
    """Split-bf16 and FP32 Davidson phases, finished by FP64 Davidson from their vectors
    (residuals <= 1e-9): eT's three lowest excitation energies (tests/restart_ccsd_v1.0) to
    1e-10 Eh, right and left; a second FP64 solve started from the converged vectors
    converges at once.  The split-bf16 phase followed by the refined phase (FP64 start
    block, split-bf16 expansions) gives the same roots, and an FP64 transformation of its
    vectors confirms residuals below 1e-9."""
    from et_amd import ccsd, solvers
    w_ref = v1_0_excitation_energies()
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        for side in ('right', 'left'):
            sv = solvers.MixedPrecisionDavidson(wf, 3, side, residual_threshold=1e-9,
                                                schedule=(ccsd.BF16X3, ccsd.FP32), refine=False)
            w = sv.run()
            np.testing.assert_allclose(w, w_ref, rtol=0, atol=1e-10)
            assert [p for p, *_ in sv.phases] == [ccsd.BF16X3, ccsd.FP32, ccsd.FP64]
            assert sv.noise[ccsd.BF16X3] > sv.noise[ccsd.FP32] > 0
            again, hist = wf.davidson(3, side, 1e-8, 10, 0, ccsd.FP64, start=sv.eigenvectors)
            assert len(hist) == 1 and not wf.davidson_stalled
            np.testing.assert_allclose(again, w_ref, rtol=0, atol=1e-10)
            # split-bf16, then the refined phase (FP64 start block, split-bf16 expansions):
            # eT's roots, and an FP64 transformation of the vectors confirms the residual
            sv = solvers.MixedPrecisionDavidson(wf, 3, side, residual_threshold=1e-9)
            w = sv.run()
            print(side, 'refined phases', sv.phases, [c.max() for c in sv.checks], flush=True)
            np.testing.assert_allclose(w, w_ref, rtol=0, atol=1e-10)
            assert sv.checks and sv.checks[-1].max() < 1e-9
            assert str(sv.phases[-1][0]).startswith('refined')


@pytest.mark.gpu
def test_gpu_eom_with_reduced_space_replacement_reproduces_et():
    """eT's tests/ccsd_right_es_replace_red_space: two right roots to residual 1e-10 with the
    reduced space capped at 20 (restarts on the Ritz vectors) give eT's energies
    (result/ccsd_right_es_replace_red_space.out:858-859) to 1e-10 Eh."""
    from et_amd import ccsd, solvers
    g = fixture()['golden']
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        dav = solvers.EigenDavidson(wf, 2, 'right', residual_threshold=1e-10, max_dim=20)
        w = dav.run()
    print('iterations', len(dav.history))
    assert len(dav.history) > 9                    # more trial vectors than the capped space holds
    np.testing.assert_allclose(w, [0.247194255570, 0.312716698776], rtol=0, atol=1e-10)
    np.testing.assert_allclose(w, g['right_excitation_energies'], rtol=0, atol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize('side', ['right', 'left'])
def test_gpu_diis_excited_states_reproduce_et(side):
    """eT's DIIS excited-state solver (tests/ccsd_right_es_diis, tests/ccsd_left_es_diis: two
    states, residual and energy thresholds 1e-10): the same number of iterations, every
    iteration's energies (1e-8 Eh; the first ones are diagonal elements of A, 1e-10) and
    residual norms (as printed, 4 digits, where above 1e-6), and the final roots -- the
    first and the third root, where the start vectors lead -- to 1e-10 Eh."""
    from et_amd import ccsd, solvers
    ref = json.load(open(os.path.join(os.path.dirname(FIXTURE), 'h2o_he_es_diis.json')))[side]
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        es = solvers.DiisCCExcitedStates(wf, 2, side, residual_threshold=1e-10, energy_threshold=1e-10)
        w = es.run()
    print(side, 'iterations', len(es.history), 'eT', ref['iterations'], w)
    np.testing.assert_allclose(w, ref['excitation_energies'], rtol=0, atol=1e-10)
    assert len(es.history) == ref['iterations']
    for it, state, e_ref, r_ref in ref['table']:
        e = es.energy_history[it - 1][state - 1]
        r = es.history[it - 1][state - 1]
        assert abs(e - e_ref) < (1e-10 if it == 1 else 1e-8), (it, state, e, e_ref)
        if r_ref > 1e-6:
            assert abs(r - r_ref) <= 6e-4 * r_ref, (it, state, r, r_ref)


@pytest.mark.gpu
def test_gpu_cvs_excited_states_reproduce_et():
    """Core-valence separation (eT's `core excitation: {1}`, the oxygen 1s): the DIIS
    excited-state solver as in tests/ccsd_right_cvs_es_energies_diis -- eT's 27 iterations,
    every iteration's energies (the first are diagonal elements of A, 1e-10; then 1e-8) and
    residual norms (4 digits where above 1e-6), the roots -- and the eigen Davidson solver as
    in tests/ccsd_right_cvs_es_energies (eT's 22 iterations), both to eT's roots (1e-9 Eh at
    19.7 Eh); the Davidson eigenvectors hold no element outside the CVS space."""
    from et_amd import ccsd, solvers
    fx = json.load(open(os.path.join(os.path.dirname(FIXTURE), 'h2o_he_es_diis.json')))
    ref, ref_dav = fx['right_cvs'], fx['right_cvs_davidson']
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        es = solvers.DiisCCExcitedStates(wf, 2, 'right', residual_threshold=1e-10, energy_threshold=1e-10,
                                         core_mos=[0])
        w = es.run()
        dav = solvers.EigenDavidson(wf, 2, 'right', residual_threshold=1e-10, core_mos=[0])
        w_dav = dav.run()
        X = dav.eigenvectors
    print('CVS DIIS iterations', len(es.history), 'eT', ref['iterations'], w)
    print('CVS Davidson iterations', len(dav.history), 'eT', ref_dav['iterations'], w_dav)
    np.testing.assert_allclose(w, ref['excitation_energies'], rtol=0, atol=1e-9)
    np.testing.assert_allclose(w_dav, ref_dav['excitation_energies'], rtol=0, atol=1e-9)
    assert len(es.history) == ref['iterations']
    assert len(dav.history) == ref_dav['iterations']
    for it, state, e_ref, r_ref in ref['table']:
        e = es.energy_history[it - 1][state - 1]
        r = es.history[it - 1][state - 1]
        assert abs(e - e_ref) < (1e-10 if it == 1 else 1e-8), (it, state, e, e_ref)
        if r_ref > 1e-6:
            assert abs(r - r_ref) <= 6e-4 * r_ref, (it, state, r, r_ref)
    # the eigenvectors hold only excitations out of the oxygen 1s (occupied MO 0)
    n_t1 = n_o * n_v
    occ_of = lambda X_idx: X_idx // n_v
    i1 = np.arange(n_t1) // n_v
    assert np.abs(X[:n_t1][i1 != 0]).max() == 0.0
    q = np.arange(X.shape[0] - n_t1)
    Xi = ((np.sqrt(8.0 * q + 1) - 1) // 2).astype(np.int64)
    Yi = q - Xi * (Xi + 1) // 2
    outside = (occ_of(Xi) != 0) & (occ_of(Yi) != 0)
    assert np.abs(X[n_t1:][outside]).max() == 0.0
    assert np.abs(X[n_t1:][~outside]).max() > 0.0


@pytest.mark.gpu
def test_gpu_remove_core_reproduces_et():
    """eT's `remove core: {1}` (tests/ccsd_es_remove_core): the eigen Davidson solver with every
    transformed vector projected away from excitations out of the oxygen 1s, from the usual
    start vectors: eT's iteration count and roots (1e-10), no eigenvector element with the
    core MO among its occupied indices."""
    from et_amd import ccsd, solvers
    ref = json.load(open(os.path.join(os.path.dirname(FIXTURE), 'h2o_he_es_diis.json')))['right_remove_core_davidson']
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
        wf.set_cholesky(m['L_mo'])
        wf.set_oei(m['h_mo'])
        wf.set_orbital_energies(m['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-11).run()
        dav = solvers.EigenDavidson(wf, 2, 'right', residual_threshold=1e-10, remove_core=[0])
        w = dav.run()
        X = dav.eigenvectors
    print('remove-core Davidson iterations', len(dav.history), 'eT', ref['iterations'], w)
    np.testing.assert_allclose(w, ref['excitation_energies'], rtol=0, atol=1e-10)
    assert len(dav.history) == ref['iterations']
    n_t1 = n_o * n_v
    assert np.abs(X[:n_t1][np.arange(n_t1) // n_v == 0]).max() == 0.0
    q = np.arange(X.shape[0] - n_t1)
    Xi = ((np.sqrt(8.0 * q + 1) - 1) // 2).astype(np.int64)
    Yi = q - Xi * (Xi + 1) // 2
    assert np.abs(X[n_t1:][(Xi // n_v == 0) | (Yi // n_v == 0)]).max() == 0.0


@pytest.mark.gpu
def test_gpu_crop_multipliers_and_ground_state_agree_with_diis():
    """eT's CROP variant of DIIS (`crop` in `solver cc gs` / `solver cc multipliers`) changes the
    path to the solution, not the solution: on H2O-He the CROP ground state (DIIS dimension 3,
    as eT's ccsd_gs_energy_nr_crop) and CROP multipliers reach the plain DIIS solvers' energy
    (1e-10) and multipliers (1e-9)."""
    from et_amd import ccsd, solvers
    m = inputs()
    n_o = m['n_o']
    n_v = m['L_mo'].shape[1] - n_o
    out = {}
    for crop, dim in ((False, 8), (True, 3)):
        with ccsd.CCSD(n_o, n_v, m['n_J']) as wf:
            wf.set_cholesky(m['L_mo'])
            wf.set_oei(m['h_mo'])
            wf.set_orbital_energies(m['eps'])
            gs = solvers.DiisCCGroundState(wf, residual_threshold=1e-11, diis_dimension=dim, crop=crop)
            gs.run()
            mu = solvers.DiisCCMultipliers(wf, residual_threshold=1e-11, diis_dimension=dim, crop=crop)
            tbar = mu.run()
            out[crop] = (wf.correlation_energy, tbar, len(gs.history), len(mu.history))
        print('crop' if crop else 'diis', 'E_corr %.12f, %d + %d iterations' % (out[crop][0], out[crop][2], out[crop][3]))
    assert abs(out[True][0] - out[False][0]) < 1e-10
    assert np.abs(out[True][1] - out[False][1]).max() < 1e-9
