"""Orbital sign alignment between two sets of CC vectors in the same MO basis.

Canonical RHF orbitals are fixed up to a sign each; a singles element t_ai
changes by s_a s_i and a doubles element t_aibj by s_a s_i s_b s_j.  The sign
bits solve a linear system over GF(2) built from the largest elements of a
reference vector pair (singles and doubles), eliminated in order of magnitude.
"""
import numpy as np


def _pairs(n_t1):
    X = np.repeat(np.arange(n_t1), np.arange(1, n_t1 + 1))
    Y = np.concatenate([np.arange(x + 1) for x in range(n_t1)])
    return X, Y


def orbital_signs(ref, ours, n_o, n_v, threshold=1e-6):
    """Signs s (length n_o + n_v, occupied first) with ref ~ S(s) ours."""
    n_t1 = n_o * n_v
    n_mo = n_o + n_v
    rows, rhs, mags = [], [], []

    def add(orbs, r, o):
        if abs(r) > threshold and abs(o) > threshold:
            bits = 0
            for p in orbs:
                bits ^= 1 << int(p)
            rows.append(bits)
            rhs.append(int(np.sign(r) != np.sign(o)))
            mags.append(abs(r))

    for k in range(n_t1):
        a, i = k % n_v, k // n_v
        add((n_o + a, i), ref[k], ours[k])
    X, Y = _pairs(n_t1)
    big = np.nonzero(np.abs(ref[n_t1:]) > threshold)[0]
    for q in big:
        x, y = X[q], Y[q]
        add((n_o + x % n_v, x // n_v, n_o + y % n_v, y // n_v), ref[n_t1 + q], ours[n_t1 + q])
    basis = {}                                      # pivot bit -> (row, rhs)
    for k in np.argsort(mags)[::-1]:
        row, b = rows[k], rhs[k]
        while row:
            p = row.bit_length() - 1
            if p not in basis:
                basis[p] = (row, b)
                break
            prow, pb = basis[p]
            row ^= prow
            b ^= pb
    x = np.zeros(n_mo, dtype=int)
    for p in sorted(basis):                         # back substitution, low pivots first
        row, b = basis[p]
        rest = row & ~(1 << p)
        v = b
        q = 0
        while rest:
            if rest & 1:
                v ^= x[q]
            rest >>= 1
            q += 1
        x[p] = v
    return 1 - 2 * x


def apply_signs(vec, s, n_o, n_v):
    """S(s) vec for a vector in the amplitude layout (singles (v, o), packed doubles)."""
    n_t1 = n_o * n_v
    s_ai = s[n_o + np.arange(n_t1) % n_v] * s[np.arange(n_t1) // n_v]
    X, Y = _pairs(n_t1)
    return np.concatenate([s_ai * vec[:n_t1], s_ai[X] * s_ai[Y] * vec[n_t1:]])
