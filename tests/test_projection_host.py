"""Host-side pieces of the excited-state projections and the bath orbital, on the CPU.

Start vectors: eT's CVS start indices (set_cvs_start_indices_ccs, ccs_class.F90:1393-1438) walk
a = 1, 2, ... and, for each a, the core MOs in order, at singles index n_v (i - 1) + a; the
ionisation start indices (set_ip_start_indices_ccs, ccs_class.F90:1585-1608) are
(n_o - I) n_v + n_v for I = 1..n_states.  Here they are written out by hand for a small case
(n_o = 4, n_v = 3) in eT's 1-based numbering and compared with the unit vectors the wavefunction
builds.  The bath orbital (make_bath_orbital_ccs) appends one MO with zero coefficients, zero
integrals and orbital energy 0.
"""
import types

import numpy as np

from et_amd import ccsd, molecule


def _wf(n_o, n_v):
    n_t1 = n_o * n_v
    return types.SimpleNamespace(n_o=n_o, n_v=n_v, n_t1=n_t1, n_t2=n_t1 * (n_t1 + 1) // 2)


def _ones(X):
    return [int(np.flatnonzero(X[:, k])[0]) for k in range(X.shape[1])]


def test_cvs_start_vectors_follow_et_order():
    wf = _wf(4, 3)
    # eT, core MOs {2, 4} (1-based), 5 states: (a, i) = (1,2), (1,4), (2,2), (2,4), (3,2)
    et_1based = [3 * (2 - 1) + 1, 3 * (4 - 1) + 1, 3 * (2 - 1) + 2, 3 * (4 - 1) + 2, 3 * (2 - 1) + 3]
    X = ccsd.CCSD.cvs_start_vectors(wf, 5, [1, 3])
    assert X.shape == (wf.n_t1 + wf.n_t2, 5)
    assert _ones(X) == [k - 1 for k in et_1based]
    assert np.all(X.sum(axis=0) == 1.0)


def test_ip_start_vectors_follow_et_order():
    wf = _wf(4, 3)
    et_1based = [(4 - I) * 3 + 3 for I in (1, 2, 3)]        # bath = last virtual, HOMO first
    X = ccsd.CCSD.ip_start_vectors(wf, 3)
    assert _ones(X) == [k - 1 for k in et_1based]


def test_bath_orbital_is_appended_with_zero_integrals():
    rng = np.random.default_rng(0)
    n, nJ = 5, 7
    r = {'L_mo': np.asfortranarray(rng.standard_normal((nJ, n, n))), 'h_mo': np.asfortranarray(rng.standard_normal((n, n))),
         'eps': np.sort(rng.standard_normal(n)), 'C': np.asfortranarray(rng.standard_normal((n, n))), 'n_o': 2,
         'n_J': nJ, 'e_hf': -1.0}
    b = molecule.with_bath_orbital(r)
    assert b['L_mo'].shape == (nJ, n + 1, n + 1) and b['h_mo'].shape == (n + 1, n + 1)
    assert np.array_equal(b['L_mo'][:, :n, :n], r['L_mo']) and not b['L_mo'][:, n, :].any() and not b['L_mo'][:, :, n].any()
    assert np.array_equal(b['h_mo'][:n, :n], r['h_mo']) and not b['h_mo'][n].any() and not b['h_mo'][:, n].any()
    assert b['eps'][n] == 0.0 and np.array_equal(b['eps'][:n], r['eps'])
    assert not b['C'][:, n].any() and b['n_o'] == 2 and b['e_hf'] == -1.0
    assert b['L_mo'].flags['F_CONTIGUOUS'] and b['h_mo'].flags['F_CONTIGUOUS']


def test_cartesian_default_follows_et_list():
    assert molecule.cartesian_by_default('3-21G') and molecule.cartesian_by_default('6-31g**')
    assert not molecule.cartesian_by_default('cc-pVDZ') and not molecule.cartesian_by_default('aug-cc-pvdz')
