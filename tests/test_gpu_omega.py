"""GPU residual path (through the C ABI) against the oracle on the same seeded inputs.

Tolerances: FP64 results agree with the oracle to round-off of the products
(relative 1e-12 of the largest element); the FP32 path to 1e-5 relative.
"""
import numpy as np
import pytest

from et_amd import ccsd
from et_amd import synthetic as syn
from oracle import ccsd as O

pytestmark = pytest.mark.gpu

CASES = [
    # n_o, n_v, n_J, seed, ladder batch
    (2, 3, 6, 1, None),
    (3, 5, 9, 2, 2),
    (4, 7, 13, 3, 3),
    (5, 12, 30, 4, 5),
    (7, 4, 11, 5, None),       # n_o > n_v
    (6, 20, 40, 6, 7),
    (8, 33, 60, 7, 16),
]


def make(n_o, n_v, n_J, seed, t1_sigma=0.03, t2_sigma=0.02):
    s = syn.system(seed, n_o, n_v, n_J, t2_sigma=t2_sigma, t1_sigma=t1_sigma, J_decay=8.0,
                   L_scale=0.3)
    return s


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


def oracle_state(s):
    n_o = s['n_o']
    Lt = O.t1_cholesky(s['L'], s['t1'], n_o)
    F = O.fock_t1(s['h'], Lt, s['t1'], n_o)
    return Lt, F


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize('n_o,n_v,n_J,seed,batch', CASES)
def test_omega_fp64_matches_oracle(n_o, n_v, n_J, seed, batch):
    s = make(n_o, n_v, n_J, seed)
    print('case', n_o, n_v, n_J, seed, batch, flush=True)
    Lt, F = oracle_state(s)
    ref = O.omega(s['L'], Lt, F, s['t1'], s['t2'], n_o)
    with load(s, batch) as wf:
        Fg = wf.construct_fock()
        np.testing.assert_allclose(Fg, F, rtol=0, atol=1e-12 * np.abs(F).max())
        om = wf.construct_omega(ccsd.FP64)
        assert rel_err(om, ref) < 1e-12
        assert abs(wf.omega_norm() - np.linalg.norm(ref)) < 1e-12 * np.linalg.norm(ref)
        e = wf.calculate_energy()
        e_ref = O.correlation_energy(Lt, s['t1'], s['t2'], n_o)
        assert abs(e - e_ref) < 1e-12 * max(1.0, abs(e_ref))


@pytest.mark.parametrize('n_o,n_v,n_J,seed,batch', CASES[2:5])
def test_omega_fp32_within_tolerance(n_o, n_v, n_J, seed, batch):
    s = make(n_o, n_v, n_J, seed)
    Lt, F = oracle_state(s)
    ref = O.omega(s['L'], Lt, F, s['t1'], s['t2'], n_o)
    with load(s, batch) as wf:
        om = wf.construct_omega(ccsd.FP32)
    assert rel_err(om, ref) < 1e-5


@pytest.mark.parametrize('n_o,n_v,n_J,seed,batch', CASES[2:])
def test_omega_bf16x3_within_tolerance(n_o, n_v, n_J, seed, batch):
    """Split-bf16 products: ~2^-16 relative per product term, FP32 accumulation.  The cases
    with n_v = 20 and 33 take the ladder's tile-blocked (c, d) order through off-diagonal
    16 x 16 tiles and a narrower last tile column, over several ladder blocks."""
    s = make(n_o, n_v, n_J, seed)
    Lt, F = oracle_state(s)
    ref = O.omega(s['L'], Lt, F, s['t1'], s['t2'], n_o)
    with load(s, batch) as wf:
        om = wf.construct_omega(ccsd.BF16X3)
        om64 = wf.construct_omega(ccsd.FP64)
    err = rel_err(om, ref)
    assert err < 2e-4
    assert rel_err(om64, ref) < 1e-12 < err      # the split-bf16 residual is not the FP64 one


def test_ladder_batching_does_not_change_omega():
    s = make(5, 17, 23, 8)
    with load(s) as wf:
        a = wf.construct_omega()
        for nb in (1, 4, 6, 17):
            wf.set_ladder_batch(nb)
            b = wf.construct_omega()
            assert rel_err(b, a) < 1e-13


def test_mp2_guess_and_orbital_differences():
    s = make(4, 9, 16, 9, t1_sigma=0.0)
    n_o = s['n_o']
    with load(s) as wf:
        wf.set_initial_amplitudes_guess()
        t = wf.get_amplitudes()
        Lt = O.t1_cholesky(s['L'], np.zeros_like(s['t1']), n_o)
        ref = O.mp2_guess(Lt, s['eps'], n_o)
        assert np.abs(t[:wf.n_t1]).max() == 0.0
        assert rel_err(t[wf.n_t1:], ref) < 1e-12
        d = wf.get_orbital_differences()
        np.testing.assert_allclose(d, O.orbital_differences(s['eps'], n_o), rtol=1e-15, atol=1e-15)


def test_t1_transformation_matches_oracle_through_fock():
    s = make(3, 8, 12, 10, t1_sigma=0.1)
    Lt, F = oracle_state(s)
    with load(s) as wf:
        Fg = wf.construct_fock()
    np.testing.assert_allclose(Fg, F, rtol=0, atol=1e-12 * np.abs(F).max())


def test_errors_are_reported_not_fatal():
    with pytest.raises(ccsd.ETGError):
        ccsd.CCSD(0, 3, 4)
    s = make(2, 3, 5, 11)
    wf = ccsd.CCSD(2, 3, 5)
    with pytest.raises(ccsd.ETGError):
        wf.calculate_energy()          # no Cholesky vectors yet
    wf.set_cholesky(s['L'])
    with pytest.raises(ccsd.ETGError):
        wf.get_orbital_differences(7)  # neither n_t1 nor n_t1 + n_t2
    wf.close()
