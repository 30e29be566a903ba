"""GPU Jacobian transformations through the C ABI against the oracle.

rho = A c is checked against the oracle's complex-step derivative of the
residual along D2 c (exact to rounding, pinned by eT's golden excitation
energies); sigma = A^T b against the transpose of the oracle's full A on
tiny systems, and on larger ones through <b, A c> = <A^T b, c>.  Tolerance:
1e-11 relative to the largest element (FP64 throughout).
"""
import numpy as np
import pytest

from et_amd import ccsd
from et_amd import synthetic as syn
from oracle import ccsd as O

pytestmark = pytest.mark.gpu

CASES = [
    # n_o, n_v, n_J, seed, ladder batch
    (1, 2, 4, 1, None),
    (2, 3, 6, 2, None),
    (3, 4, 9, 3, 2),
    (4, 3, 7, 4, None),       # n_o > n_v
    (3, 7, 12, 5, 3),
]


def make(n_o, n_v, n_J, seed):
    return syn.system(seed, n_o, n_v, n_J, t2_sigma=0.03, t1_sigma=0.03, J_decay=6.0, L_scale=0.3)


def load(s, batch=None):
    wf = ccsd.CCSD(s['n_o'], s['n_v'], s['n_J'])
    wf.set_cholesky(s['L'])
    wf.set_oei(s['h'])
    wf.set_orbital_energies(s['eps'])
    wf.set_amplitudes(np.concatenate([s['t1'].ravel(order='F'), s['t2']]))
    wf.construct_t1_cholesky()
    if batch:
        wf.set_ladder_batch(batch)
    return wf


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize('n_o,n_v,n_J,seed,batch', CASES)
def test_right_jacobian_matches_complex_step(n_o, n_v, n_J, seed, batch):
    print('case', n_o, n_v, n_J, seed, batch, flush=True)
    s = make(n_o, n_v, n_J, seed)
    n = n_o * n_v + O.n_t2(n_o, n_v)
    rng = np.random.default_rng(seed)
    with load(s, batch) as wf:
        om_direct = wf.construct_omega()
        wf.prepare_for_jacobian()
        om_tape = wf.get_omega()
        assert rel(om_tape, om_direct) < 1e-13          # the recorded program is the residual
        for _ in range(2):
            c = rng.normal(size=n)
            rho = wf.jacobian_transformation(c)
            ref = O.jacobian_transform(s['L'], s['h'], s['t1'], s['t2'], n_o, c)
            assert rel(rho, ref) < 1e-11


@pytest.mark.parametrize('n_o,n_v,n_J,seed,batch', CASES[:4])
def test_left_jacobian_is_transpose_of_oracle_matrix(n_o, n_v, n_J, seed, batch):
    print('case', n_o, n_v, n_J, seed, batch, flush=True)
    s = make(n_o, n_v, n_J, seed)
    A = O.jacobian_matrix(s['L'], s['h'], s['t1'], s['t2'], n_o)
    rng = np.random.default_rng(100 + seed)
    with load(s, batch) as wf:
        wf.prepare_for_jacobian()
        for _ in range(2):
            b = rng.normal(size=A.shape[0])
            sigma = wf.jacobian_transpose_transformation(b)
            assert rel(sigma, A.T @ b) < 1e-11


def test_adjoint_identity_larger_system():
    n_o, n_v, n_J = 6, 17, 30
    s = make(n_o, n_v, n_J, 9)
    n = n_o * n_v + O.n_t2(n_o, n_v)
    rng = np.random.default_rng(7)
    with load(s, 5) as wf:
        wf.prepare_for_jacobian()
        for _ in range(3):
            b, c = rng.normal(size=n), rng.normal(size=n)
            Ac = wf.jacobian_transformation(c)
            Atb = wf.jacobian_transpose_transformation(b)
            lhs, rhs = float(b @ Ac), float(Atb @ c)
            assert abs(lhs - rhs) < 1e-11 * (np.abs(b).sum() * np.abs(Ac).max())
        ref = O.jacobian_transform(s['L'], s['h'], s['t1'], s['t2'], n_o, c)
        assert rel(Ac, ref) < 1e-11


def test_split_bf16_transformations_over_ladder_tiles():
    """Split-bf16 right and left transformations of a system whose ladder spans several
    16 x 16 tiles of the (c, d) plane (n_v = 37: a narrow last tile column) and several
    ladder blocks stay within the products' rounding of the FP64 ones."""
    n_o, n_v, n_J = 4, 37, 40
    s = make(n_o, n_v, n_J, 11)
    n = n_o * n_v + O.n_t2(n_o, n_v)
    X = np.random.default_rng(12).normal(size=(n, 3))
    with load(s, 13) as wf:
        wf.prepare_for_jacobian()
        for side in ('right', 'left'):
            R = wf.jacobian_transformations(X, side)
            Rp = wf.jacobian_transformations(X, side, precision=ccsd.BF16X3)
            assert 0 < rel(Rp, R) < 2e-4, side


def test_jacobian_requires_preparation():
    s = make(2, 3, 5, 4)
    with load(s) as wf:
        with pytest.raises(ccsd.ETGError):
            wf.jacobian_transformation(np.zeros(wf.n_gs_amplitudes))
        wf.prepare_for_jacobian()
        wf.jacobian_transformation(np.zeros(wf.n_gs_amplitudes))
        wf.set_amplitudes(wf.get_amplitudes())          # invalidates the record
        with pytest.raises(ccsd.ETGError):
            wf.jacobian_transpose_transformation(np.zeros(wf.n_gs_amplitudes))


def test_batched_transformations_match_single_calls():
    """etg_jacobian_transformations: FP64 columns equal the single-vector calls; the
    reduced-precision forms stay within their rounding of the FP64 ones."""
    s = make(3, 7, 12, 5)
    n = 3 * 7 + O.n_t2(3, 7)
    rng = np.random.default_rng(21)
    X = rng.normal(size=(n, 3))
    with load(s, 3) as wf:
        wf.prepare_for_jacobian()
        for side, single in (('right', wf.jacobian_transformation),
                             ('left', wf.jacobian_transpose_transformation)):
            R = wf.jacobian_transformations(X, side)
            for k in range(3):
                assert rel(R[:, k], single(X[:, k])) < 1e-14
            for prec, tol in ((ccsd.FP32, 1e-5), (ccsd.BF16X3, 2e-4)):
                Rp = wf.jacobian_transformations(X, side, precision=prec)
                assert 0 < rel(Rp, R) < tol
        assert wf.jacobian_transformations(X[:, :0], 'right').shape == (n, 0)


def test_graph_replays_equal_eager_transformations():
    """Captured transformation passes: every replay equals the eager pass exactly, with eager
    work and synchronisation between replays, after a new record and a new ladder batching."""
    s = make(3, 7, 12, 5)
    n = 3 * 7 + O.n_t2(3, 7)
    X = np.random.default_rng(33).normal(size=(n, 3))
    with load(s, 3) as wf:
        wf.prepare_for_jacobian()
        ref = {}
        wf.set_transform_graphs(False)
        for side in ('right', 'left'):
            for prec in (ccsd.FP64, ccsd.BF16X3):
                ref[side, prec] = wf.jacobian_transformations(X, side, precision=prec)
        wf.set_transform_graphs(True)
        for rep in range(3):
            for side in ('right', 'left'):
                for prec in (ccsd.FP64, ccsd.BF16X3):
                    R = wf.jacobian_transformations(X, side, precision=prec)
                    assert np.array_equal(R, ref[side, prec]), (rep, side, prec)
            wf.construct_omega()                      # eager work between the replays
            wf.synchronize()
        # a new record (amplitudes changed) and a new ladder plan drop the captured passes
        t = wf.get_amplitudes()
        t[5] += 1e-3
        wf.set_amplitudes(t)
        wf.construct_t1_cholesky()
        wf.set_ladder_batch(2)
        wf.prepare_for_jacobian()
        for side, single in (('right', wf.jacobian_transformation),
                             ('left', wf.jacobian_transpose_transformation)):
            R = wf.jacobian_transformations(X, side)
            for k in range(3):
                assert rel(R[:, k], single(X[:, k])) < 1e-14
            wf.set_transform_graphs(False)
            E = wf.jacobian_transformations(X, side)
            wf.set_transform_graphs(True)
            assert np.array_equal(R, E)


def test_vector_groups_match_one_vector_passes():
    """Trial vectors as the batch dimension of the passes: 7 vectors in groups of
    1 (one vector per pass), 3 (3 + 3 + 1, two group sizes) and automatic (one group)
    agree to FP64 rounding (to the products' rounding for split-bf16); replays of
    each group size's graph equal the first pass exactly."""
    s = make(4, 9, 15, 6)
    n = 4 * 9 + O.n_t2(4, 9)
    X = np.random.default_rng(44).normal(size=(n, 7))
    with load(s, 4) as wf:
        wf.prepare_for_jacobian()
        ref = {}
        wf.set_transform_batch(1)
        for side in ('right', 'left'):
            for prec in (ccsd.FP64, ccsd.BF16X3):
                ref[side, prec] = wf.jacobian_transformations(X, side, precision=prec)
        for limit in (3, 0):
            wf.set_transform_batch(limit)
            for side in ('right', 'left'):
                for prec, tol in ((ccsd.FP64, 1e-13), (ccsd.BF16X3, 1e-5)):
                    R = wf.jacobian_transformations(X, side, precision=prec)
                    assert rel(R, ref[side, prec]) < tol, (limit, side, prec)
                    again = wf.jacobian_transformations(X, side, precision=prec)
                    wf.construct_omega()          # eager work between the replays
                    wf.synchronize()
                    assert np.array_equal(again, R), (limit, side, prec)
                    # a group smaller than the largest one reuses the reserved buffers
                    part = wf.jacobian_transformations(X[:, 2:4], side, precision=prec)
                    assert rel(part, R[:, 2:4]) < tol, (limit, side, prec)


def test_left_transformation_against_central_differences_on_water_dimer():
    """(H2O)_2 / cc-pVDZ from the device producer at its converged ground state: along 32
    seeded random directions c (with random b), <A^T b, c> from the left transformation
    equals b . (Omega(t + h D2 c) - Omega(t - h D2 c)) / 2h with h = 1e-4, to the O(h^2)
    error of the central difference."""
    from et_amd import molecule as M, solvers
    with M.Molecule(M.water_lattice((2, 1, 1), 2.9, 0)) as mol:
        r = mol.run(cholesky_threshold=1e-8, scf_threshold=1e-10)
    n_o = r['n_o']
    n_v = r['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        solvers.DiisCCGroundState(wf, residual_threshold=1e-10).run()
        t0 = wf.get_amplitudes()
        n_t1 = wf.n_t1
        diag = n_t1 + np.arange(n_t1) * (np.arange(n_t1) + 3) // 2     # packed (ai, ai)
        wf.prepare_for_jacobian('left')
        rng = np.random.default_rng(2024)
        C = rng.normal(size=(wf.n_gs_amplitudes, 32))
        B = rng.normal(size=(wf.n_gs_amplitudes, 32))
        AtB = wf.jacobian_transformations(B, 'left')
        h = 1e-4
        worst = 0.0
        for k in range(32):
            d = C[:, k].copy()
            d[diag] *= 2.0                                              # D2 c
            om = []
            for sgn in (1.0, -1.0):
                wf.set_amplitudes(t0 + sgn * h * d)
                wf.construct_t1_cholesky()
                wf.construct_fock()
                om.append(wf.construct_omega())
            fd = B[:, k] @ (om[0] - om[1]) / (2 * h)
            an = AtB[:, k] @ C[:, k]
            worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
        assert worst < 1e-6, worst


def test_derivative_memory_is_shared_between_lifetimes():
    """Derivatives of intermediates whose lifetimes do not overlap share memory: a
    trial vector of a pass needs less than one derivative buffer per intermediate,
    and the passes still equal the complex-step / dense-transpose results above."""
    s = make(6, 17, 40, 11)
    with load(s, 5) as wf:
        wf.prepare_for_jacobian()
        m = wf.jacobian_memory()
        print('record %d B, per vector %d B (unshared %d B)' % (m['record'], m['per_vector'],
                                                              m['per_vector_unshared']))
        assert 0 < m['per_vector'] < 0.6 * m['per_vector_unshared']
        assert m['record'] > 8 * wf.n_gs_amplitudes
        rng = np.random.default_rng(0)
        c, b = rng.normal(size=(2, wf.n_gs_amplitudes))
        ref = O.jacobian_transform(s['L'], s['h'], s['t1'], s['t2'], 6, c)
        # passes in a row reuse the slots: each starts from zeroed derivatives
        for _ in range(2):
            rho = wf.jacobian_transformation(c)
            sigma = wf.jacobian_transpose_transformation(b)
            assert rel(rho, ref) < 1e-11
            assert abs(b @ rho - sigma @ c) < 1e-11 * max(abs(b @ rho), 1.0)
