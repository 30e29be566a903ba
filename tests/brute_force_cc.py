"""Brute-force coupled-cluster quantities in the full determinant space.

Independent of eT's factorisation: builds second-quantised operators as dense
matrices over all S_z = 0 determinants of a tiny system, forms
H = sum h_pq E_pq + 1/2 sum g_pqrs (E_pq E_rs - delta_qr E_ps) and
T = sum t_ai E_ai + 1/2 sum t_aibj E_ai E_bj, and evaluates
exp(-T) H exp(T) |HF> by power series (T is nilpotent).  Projections onto
the alpha-spin single a+_{a alpha} a_{i alpha}|HF> and the alpha-beta double
a+_{a alpha} a_{i alpha} a+_{b beta} a_{j beta}|HF> give the singles residual
and the (ai,bj)-symmetric doubles residual in the amplitude representation
in which T2 = 1/2 sum t_aibj E_ai E_bj.
"""
import itertools
import math

import numpy as np


class DeterminantSpace:
    def __init__(self, n_mo, n_o):
        self.n_mo = n_mo
        self.n_o = n_o
        strings = [sum(1 << k for k in c) for c in itertools.combinations(range(n_mo), n_o)]
        # spin orbital k (alpha) -> bit k, k (beta) -> bit n_mo + k
        self.dets = [a | (b << n_mo) for a in strings for b in strings]
        self.index = {d: k for k, d in enumerate(self.dets)}
        self.dim = len(self.dets)
        self.hf = self.index[(1 << n_o) - 1 | (((1 << n_o) - 1) << n_mo)]
        self._E = {}

    def excite(self, p, q):
        """Dense matrix of a+_p a_q over spin orbitals p, q."""
        M = np.zeros((self.dim, self.dim))
        for col, d in enumerate(self.dets):
            if not (d >> q) & 1:
                continue
            sign = (-1) ** bin(d & ((1 << q) - 1)).count('1')
            d1 = d ^ (1 << q)
            if (d1 >> p) & 1:
                continue
            sign *= (-1) ** bin(d1 & ((1 << p) - 1)).count('1')
            d2 = d1 | (1 << p)
            M[self.index[d2], col] += sign
        return M

    def E(self, p, q):
        key = (p, q)
        if key not in self._E:
            n = self.n_mo
            self._E[key] = self.excite(p, q) + self.excite(n + p, n + q)
        return self._E[key]


def hamiltonian(space, h, g):
    n = space.n_mo
    H = np.zeros((space.dim, space.dim))
    for p in range(n):
        for q in range(n):
            if h[p, q] != 0.0:
                H += h[p, q] * space.E(p, q)
    for p, q, r, s in itertools.product(range(n), repeat=4):
        c = 0.5 * g[p, q, r, s]
        if c == 0.0:
            continue
        H += c * (space.E(p, q) @ space.E(r, s))
        if q == r:
            H -= c * space.E(p, s)
    return H


def cluster_operator(space, t1, t_aibj):
    n_o = space.n_o
    n_v = space.n_mo - n_o
    T = np.zeros((space.dim, space.dim))
    for a in range(n_v):
        for i in range(n_o):
            T += t1[a, i] * space.E(n_o + a, i)
    for a, i, b, j in itertools.product(range(n_v), range(n_o), range(n_v), range(n_o)):
        c = 0.5 * t_aibj[a, i, b, j]
        if c != 0.0:
            T += c * (space.E(n_o + a, i) @ space.E(n_o + b, j))
    return T


def apply_exp(T, v, sign=1.0, max_order=64):
    out = v.copy()
    term = v.copy()
    for k in range(1, max_order):
        term = sign * (T @ term) / k
        if not np.any(term):
            break
        out += term
    return out


def similarity_transformed_hf(h, g, t1, t_aibj, n_o):
    """Returns (space, exp(-T) H exp(T) |HF>) as a vector."""
    space = DeterminantSpace(h.shape[0], n_o)
    H = hamiltonian(space, h, g)
    T = cluster_operator(space, t1, t_aibj)
    hf = np.zeros(space.dim)
    hf[space.hf] = 1.0
    w = apply_exp(T, hf, 1.0)
    w = H @ w
    w = apply_exp(T, w, -1.0)
    return space, w


def projections(space, w):
    """(E, R1[a,i], R2[a,i,b,j]) from the transformed vector."""
    n = space.n_mo
    n_o = space.n_o
    n_v = n - n_o
    hf = np.zeros(space.dim)
    hf[space.hf] = 1.0
    E = float(w[space.hf])
    R1 = np.zeros((n_v, n_o))
    R2 = np.zeros((n_v, n_o, n_v, n_o))
    for a in range(n_v):
        for i in range(n_o):
            ket = space.excite(n_o + a, i) @ hf
            R1[a, i] = ket @ w
    for a, i, b, j in itertools.product(range(n_v), range(n_o), range(n_v), range(n_o)):
        ket = space.excite(n_o + a, i) @ (space.excite(n + n_o + b, n + j) @ hf)
        R2[a, i, b, j] = ket @ w
    return E, R1, R2


def nuclear_free_reference_energy(h, g, n_o):
    o = slice(0, n_o)
    return float(2.0 * np.trace(h[o, o])
                 + 2.0 * np.einsum('iijj->', g[o, o, o, o])
                 - np.einsum('ijji->', g[o, o, o, o]))


__all__ = ['similarity_transformed_hf', 'projections', 'nuclear_free_reference_energy', 'math']
