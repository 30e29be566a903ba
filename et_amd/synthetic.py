"""Deterministic synthetic CCSD inputs from a counter-based generator.

Every random number is a pure function of (seed, stream, counter): the
counter is hashed with the splitmix64 finaliser, the top 53 bits give a
uniform in (0, 1), and a standard normal is cos(2 pi u2) sqrt(-2 ln u1) of
the uniforms at counters 2k and 2k+1.  Any element can therefore be
regenerated independently (on the host or in a device kernel) without
replaying a sequential stream.

Streams: 1 = Cholesky vectors, 2 = doubles amplitudes, 3 = occupied orbital
energies, 4 = virtual orbital energies, 5 = singles amplitudes.
"""
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
    """Standard normals, one per counter k (uses uniforms at 2k and 2k+1)."""
    c = np.asarray(counters, dtype=np.uint64)
    u1 = uniform(seed, stream, c * np.uint64(2))
    u2 = uniform(seed, stream, c * np.uint64(2) + np.uint64(1))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def cholesky_vectors(seed, n_mo, n_J, J_decay=400.0, scale=0.5, dtype=np.float64):
    """L[J, p, q] = scale * exp(-J/J_decay) * N(counter(J, min(p,q), max(p,q))).

    Symmetric in p <-> q, so g_pqrs = g_qprs = g_rspq.  Returned in Fortran
    order (J fastest), the layout the C ABI takes.
    """
    L = np.empty((n_J, n_mo, n_mo), dtype=dtype, order='F')
    p = np.arange(n_mo, dtype=np.uint64)
    lo = np.minimum.outer(p, p)
    hi = np.maximum.outer(p, p)
    pair = hi * np.uint64(n_mo) + lo                     # [p, q]
    n2 = np.uint64(n_mo * n_mo)
    for J in range(n_J):
        z = normal(seed, 1, np.uint64(J) * n2 + pair)
        L[J] = (scale * np.exp(-J / J_decay)) * z
    return L


def orbital_energies(seed, n_o, n_v, occ=(-1.5, -0.3), vir=(0.1, 3.0)):
    e_o = occ[0] + (occ[1] - occ[0]) * uniform(seed, 3, np.arange(n_o))
    e_v = vir[0] + (vir[1] - vir[0]) * uniform(seed, 4, np.arange(n_v))
    return np.concatenate([np.sort(e_o), np.sort(e_v)])


def doubles_amplitudes(seed, n_o, n_v, sigma=1e-3):
    """Packed t2 (eT layout), t2[k] = sigma * N(k): symmetric under (ai) <-> (bj)."""
    ov = n_o * n_v
    n = ov * (ov + 1) // 2
    return sigma * normal(seed, 2, np.arange(n, dtype=np.uint64))


def singles_amplitudes(seed, n_o, n_v, sigma=1e-2):
    return (sigma * normal(seed, 5, np.arange(n_o * n_v, dtype=np.uint64))).reshape(n_v, n_o, order='F')


def canonical_core_hamiltonian(L, eps, n_o):
    """h = diag(eps) - sum_k (2 g_pqkk - g_pkkq), so that the Fock matrix at t1 = 0 is diag(eps)."""
    X = np.einsum('Jkk->J', L[:, :n_o, :n_o])
    G = 2.0 * np.einsum('Jpq,J->pq', L, X) - np.einsum('Jpk,Jkq->pq', L[:, :, :n_o], L[:, :n_o, :])
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
