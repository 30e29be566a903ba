"""Deterministic synthetic CCSD inputs from a counter-based generator.

Every random number is a pure function of (seed, stream, counter): the
counter is hashed with the splitmix64 finaliser, the top 53 bits give a
uniform in (0, 1), and the normals at counters 2k and 2k+1 are the Box-Muller
pair r cos(theta), r sin(theta) with r = sqrt(-2 ln u(2k)), theta = 2 pi u(2k+1).
Any element can therefore be regenerated independently (on the host or in a
device kernel) without replaying a sequential stream.

Streams: 1 = Cholesky vectors, 2 = doubles amplitudes, 3 = occupied orbital
energies, 4 = virtual orbital energies, 5 = singles amplitudes.
"""
import concurrent.futures
import os

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _key(seed, stream):
    with np.errstate(over='ignore'):
        return _mix(np.uint64(seed) * _GOLD + np.uint64(stream) + np.uint64(1))


def uniform(seed, stream, counters):
    """Uniform (0,1) doubles for an array of uint64 counters."""
    c = np.asarray(counters, dtype=np.uint64)
    with np.errstate(over='ignore'):
        z = _mix(c * _GOLD + _key(seed, stream))
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed, stream, counters):
    """Standard normals, one per counter (any order)."""
    c = np.asarray(counters, dtype=np.uint64)
    pair = c >> np.uint64(1)
    u1 = uniform(seed, stream, pair * np.uint64(2))
    u2 = uniform(seed, stream, pair * np.uint64(2) + np.uint64(1))
    r = np.sqrt(-2.0 * np.log(u1))
    th = 2.0 * np.pi * u2
    return np.where((c & np.uint64(1)) == 0, r * np.cos(th), r * np.sin(th))


def normal_range(seed, stream, start, n, threads=None):
    """normal(seed, stream, start + arange(n)) computed pairwise and in parallel chunks."""
    out = np.empty(n)
    if n == 0:
        return out
    chunk = 1 << 22

    def work(lo):
        hi = min(n, lo + chunk)
        c0, c1 = start + lo, start + hi                  # counters [c0, c1)
        p0, p1 = c0 // 2, (c1 + 1) // 2                  # Box-Muller pairs covering them
        pairs = np.arange(p0, p1, dtype=np.uint64)
        u1 = uniform(seed, stream, pairs * np.uint64(2))
        u2 = uniform(seed, stream, pairs * np.uint64(2) + np.uint64(1))
        r = np.sqrt(-2.0 * np.log(u1))
        th = 2.0 * np.pi * u2
        z = np.empty(2 * (p1 - p0))
        z[0::2] = r * np.cos(th)
        z[1::2] = r * np.sin(th)
        out[lo:hi] = z[c0 - 2 * p0: c1 - 2 * p0]

    starts = range(0, n, chunk)
    if len(starts) == 1:
        work(0)
    else:
        with concurrent.futures.ThreadPoolExecutor(max_workers=threads or min(16, os.cpu_count() or 1)) as ex:
            list(ex.map(work, starts))
    return out


def cholesky_vectors(seed, n_mo, n_J, J_decay=400.0, scale=0.5, dtype=np.float64):
    """L[J, p, q] = scale * exp(-J/J_decay) * N(J * n_pairs + pair(p, q)).

    pair(p, q) = max(p,q)(max(p,q)+1)/2 + min(p,q), n_pairs = n_mo(n_mo+1)/2:
    symmetric in p <-> q, so g_pqrs = g_qprs = g_rspq.  Returned in Fortran
    order (J fastest), the layout the C ABI takes.
    """
    npair = n_mo * (n_mo + 1) // 2
    z = normal_range(seed, 1, 0, n_J * npair).reshape(n_J, npair)
    z *= (scale * np.exp(-np.arange(n_J) / J_decay))[:, None]
    p = np.arange(n_mo)
    hi = np.maximum.outer(p, p)
    idx = hi * (hi + 1) // 2 + np.minimum.outer(p, p)
    L = np.empty((n_J, n_mo, n_mo), dtype=dtype, order='F')
    step = max(1, (1 << 24) // (n_mo * n_mo))

    def fill(J0):
        L[J0:J0 + step] = z[J0:J0 + step][:, idx]

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(fill, range(0, n_J, step)))
    return L


def orbital_energies(seed, n_o, n_v, occ=(-1.5, -0.3), vir=(0.1, 3.0)):
    e_o = occ[0] + (occ[1] - occ[0]) * uniform(seed, 3, np.arange(n_o))
    e_v = vir[0] + (vir[1] - vir[0]) * uniform(seed, 4, np.arange(n_v))
    return np.concatenate([np.sort(e_o), np.sort(e_v)])


def doubles_amplitudes(seed, n_o, n_v, sigma=1e-3):
    """Packed t2 (eT layout), t2[k] = sigma * N(k): symmetric under (ai) <-> (bj)."""
    ov = n_o * n_v
    n = ov * (ov + 1) // 2
    t = normal_range(seed, 2, 0, n)
    t *= sigma
    return t


def singles_amplitudes(seed, n_o, n_v, sigma=1e-2):
    return (sigma * normal(seed, 5, np.arange(n_o * n_v, dtype=np.uint64))).reshape(n_v, n_o, order='F')


def canonical_core_hamiltonian(L, eps, n_o):
    """h = diag(eps) - sum_k (2 g_pqkk - g_pkkq), so that the Fock matrix at t1 = 0 is diag(eps)."""
    n_J, n = L.shape[0], L.shape[1]
    X = np.einsum('Jkk->J', L[:, :n_o, :n_o])
    coul = np.tensordot(X, L, axes=(0, 0))
    left = np.ascontiguousarray(L[:, :, :n_o].transpose(1, 0, 2)).reshape(n, n_J * n_o)
    right = np.ascontiguousarray(L[:, :n_o, :]).reshape(n_J * n_o, n)
    G = 2.0 * coul - left @ right
    return np.asfortranarray(np.diag(eps) - G)


def system(seed, n_o, n_v, n_J, t2_sigma=1e-3, t1_sigma=0.0, J_decay=400.0, L_scale=0.5):
    """A complete synthetic system: dict with L (J,n,n), h, eps, t1 (v,o), t2 (packed)."""
    n_mo = n_o + n_v
    L = cholesky_vectors(seed, n_mo, n_J, J_decay=J_decay, scale=L_scale)
    eps = orbital_energies(seed, n_o, n_v)
    h = canonical_core_hamiltonian(L, eps, n_o)
    t2 = doubles_amplitudes(seed, n_o, n_v, sigma=t2_sigma)
    t1 = (singles_amplitudes(seed, n_o, n_v, sigma=t1_sigma) if t1_sigma
          else np.zeros((n_v, n_o), order='F'))
    return dict(n_o=n_o, n_v=n_v, n_J=n_J, L=L, h=h, eps=eps, t1=t1, t2=t2)
