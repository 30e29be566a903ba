"""The oracle's CC mean-value dipole against eT's printed one.

eT's tests/ccsd_dipole: H2O + He / STO-3G, Cholesky threshold 1e-11
(fixture tests/golden/h2o_he_sto3g_dipole.json).  The oracle's producer
reproduces E_nuc, E_HF and the number of Cholesky vectors, its ground state
E_CCSD, and its Lagrangian expectation value of the dipole operator (dense
multipliers, complex step of the residual along the dipole integrals) the
electronic dipole table in Debye.
"""
import functools
import json
import os

import numpy as np

from oracle import ccsd as O
from oracle import frontend as fe

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_sto3g_dipole.json')


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FIXTURE))


@functools.lru_cache(maxsize=1)
def solved():
    f = fixture()
    m = fe.molecular_inputs(f)
    t1, t2, e, _ = O.solve_ground_state(m['L_mo'], m['h_mo'], m['eps'], m['n_o'], residual_threshold=1e-12)
    tbar = O.multipliers_dense(m['L_mo'], m['h_mo'], t1, t2, m['n_o'])
    shells, atoms = fe.build_molecule(f['geometry_angstrom'], f['basis'], f['bohr_to_angstrom'])
    D_ao = fe.dipole_integrals(shells)
    return m, t1, t2, tbar, D_ao, atoms


def test_producer_and_ground_state_reproduce_et():
    g = fixture()['golden']
    m, t1, t2, *_ = solved()
    assert abs(m['e_nuc'] - g['nuclear_repulsion']) < 1e-11
    assert abs(m['e_hf'] - g['hf_energy']) < 1e-10
    assert m['n_J'] == g['n_cholesky']
    Lt = O.t1_cholesky(m['L_mo'], t1, m['n_o'])
    e_corr = O.correlation_energy(Lt, t1, t2, m['n_o'])
    assert abs(m['e_hf'] + e_corr - g['ccsd_energy']) < 1e-10


def test_lagrangian_dipole_reproduces_et():
    g = fixture()['golden']
    m, t1, t2, tbar, D_ao, atoms = solved()
    C = m['C']
    debye = g['au_to_debye']                      # eT's factor (its printed inverse has 9 digits)
    for c, comp in enumerate('xyz'):
        X = C.T @ D_ao[c] @ C
        electronic = -O.one_electron_expectation(m['L_mo'], m['h_mo'], t1, t2, tbar, m['n_o'], X)
        nuclear = sum(z * r[c] for z, r in atoms)
        assert abs(nuclear * debye - g['dipole_debye'][comp]['nuclear']) < 1e-9
        assert abs(electronic * debye - g['dipole_debye'][comp]['electronic']) < 1e-8, comp
