"""Molecular workloads for the ground-state solvers.

`h2o_he_inputs` reads the H2O-He / cc-pVDZ inputs (MO Cholesky vectors, MO
one-electron Hamiltonian, RHF orbital energies; eT's tests/ccsd_gs_energy
system) from the committed data file tests/golden/h2o_he_ccpvdz_inputs.npz.

`replicate` builds K copies of a system at infinite separation: block-diagonal
Cholesky vectors (each copy keeps its own n_J vectors), occupied orbitals of
all copies first, then the virtual ones.  Closed-shell CCSD is size-extensive,
so the correlation energy of the replica is exactly K times that of one copy,
and the DIIS iteration of eT's ground-state solver follows the single system's
trajectory (every inner product is K times the single one).
"""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H2O_HE = os.path.join(ROOT, 'tests', 'golden', 'h2o_he_ccpvdz_inputs.npz')


def unpack_upper(L_tri, n_mo):
    """(n_J, n_mo (n_mo+1)/2) upper-triangle pairs (column-major order) -> (n_J, n_mo, n_mo)."""
    n_J = L_tri.shape[0]
    iu = np.triu_indices(n_mo)
    order = np.lexsort((iu[0], iu[1]))
    p, q = iu[0][order], iu[1][order]
    L = np.zeros((n_J, n_mo, n_mo), order='F')
    L[:, p, q] = L_tri
    L[:, q, p] = L_tri
    return L


def h2o_he_inputs(path=H2O_HE):
    d = np.load(path)
    n_mo = d['h_mo'].shape[0]
    return dict(L_mo=unpack_upper(d['L_tri'], n_mo), h_mo=np.asfortranarray(d['h_mo']),
                eps=d['eps'].copy(), e_hf=float(d['e_hf']), e_nuc=float(d['e_nuc']),
                n_o=int(d['n_o']))


def replicate(m, copies):
    """K non-interacting copies of the system m (dict as returned by h2o_he_inputs)."""
    L, h, eps, n_o = m['L_mo'], m['h_mo'], m['eps'], m['n_o']
    n_J, n_mo = L.shape[0], L.shape[1]
    n_v = n_mo - n_o
    K = int(copies)
    No, Nv = K * n_o, K * n_v
    N = No + Nv
    Lr = np.zeros((K * n_J, N, N), order='F')
    hr = np.zeros((N, N), order='F')
    er = np.zeros(N)
    for c in range(K):
        idx = np.concatenate([c * n_o + np.arange(n_o), No + c * n_v + np.arange(n_v)])
        Lr[c * n_J:(c + 1) * n_J][:, idx[:, None], idx[None, :]] = L
        hr[np.ix_(idx, idx)] = h
        er[idx] = eps
    return dict(L_mo=Lr, h_mo=hr, eps=er, e_hf=K * m['e_hf'], e_nuc=K * m['e_nuc'], n_o=No,
                n_J=K * n_J, copies=K)
