"""Molecular producer of the residual path's inputs: AO integrals -> RHF -> Cholesky -> MO vectors.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  eT obtains these inputs
from Libint (src/libint/eri.cpp:33-89), its RHF solver
(src/solvers/hf/scf_solver_class.F90) and the pivoted Cholesky decomposition
of the AO integrals (src/solvers/cholesky/eri_cd_class.F90:241, AO->MO
transformation :3132).  This module restates those producers compactly for
small molecules so that eT's own golden energies can pin the residual path
end to end:

* integrals: McMurchie-Davidson Hermite expansion (overlap, kinetic, nuclear
  attraction, electron repulsion) over contracted Cartesian Gaussians, with
  spherical (5d) d shells as eT uses them;
* Boys function from the regularised lower incomplete gamma function with
  downward recursion;
* RHF with DIIS on the commutator FDS - SDF;
* pivoted Cholesky decomposition of the (pq|rs) matrix to a diagonal threshold.
"""
import itertools
import math

import numpy as np
from scipy.special import gammainc, gamma

Z_OF = {'H': 1, 'He': 2, 'Li': 3, 'Be': 4, 'B': 5, 'C': 6, 'N': 7, 'O': 8, 'F': 9, 'Ne': 10}


def cart_components(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2)))


def sph_transform(l):
    """Rows: real spherical functions in terms of individually normalised Cartesians."""
    if l == 0:
        return np.ones((1, 1))
    if l == 1:
        return np.eye(3)
    if l == 2:
        # order of cart_components(2): xx, xy, xz, yy, yz, zz
        s3 = math.sqrt(3.0) / 2.0
        return np.array([
            [0, 1, 0, 0, 0, 0],                 # xy
            [0, 0, 0, 0, 1, 0],                 # yz
            [-0.5, 0, 0, -0.5, 0, 1.0],         # (2zz - xx - yy)/2
            [0, 0, 1, 0, 0, 0],                 # xz
            [s3, 0, 0, -s3, 0, 0],              # (xx - yy) sqrt(3)/2
        ])
    raise NotImplementedError('l > 2')


class Shell:
    def __init__(self, center, l, exps, coefs):
        self.center = np.asarray(center, dtype=float)
        self.l = int(l)
        self.exps = np.asarray(exps, dtype=float)
        self.carts = cart_components(self.l)
        # primitive normalisation of x^l (each component renormalised below)
        a = self.exps
        norm = (2 * a / np.pi) ** 0.75 * (4 * a) ** (self.l / 2) / math.sqrt(dfact(2 * self.l - 1))
        c = np.asarray(coefs, dtype=float) * norm
        # contracted normalisation of the x^l component
        s = 0.0
        for i in range(len(a)):
            for j in range(len(a)):
                p = a[i] + a[j]
                s += c[i] * c[j] * (np.pi / p) ** 1.5 * dfact(2 * self.l - 1) / (2 * p) ** self.l
        self.coefs = c / math.sqrt(s)
        # per-component factors turning the x^l normalisation into unit norm for each component
        self.comp_scale = np.array([math.sqrt(dfact(2 * self.l - 1) /
                                              (dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1)))
                                    for lx, ly, lz in self.carts])
        self.T = sph_transform(self.l) * self.comp_scale[None, :]
        self.nfunc = self.T.shape[0]


def boys(nmax, T):
    """F_n(T) for n = 0..nmax; T any shape.  Returns array (nmax+1, *T.shape)."""
    T = np.asarray(T, dtype=float)
    F = np.empty((nmax + 1,) + T.shape)
    small = T < 1e-12
    Ts = np.where(small, 1.0, T)
    a = nmax + 0.5
    F[nmax] = np.where(small, 1.0 / (2 * nmax + 1) - T / (2 * nmax + 3),
                       gamma(a) * gammainc(a, Ts) / (2.0 * Ts ** a))
    e = np.exp(-T)
    for n in range(nmax - 1, -1, -1):
        F[n] = (2.0 * T * F[n + 1] + e) / (2 * n + 1)
    return F


def hermite_E(la, lb, a, b, Ax, Bx):
    """E[i, j, t] (i<=la, j<=lb, t<=i+j) for primitive pairs (arrays a, b broadcast)."""
    p = a + b
    mu = a * b / p
    X = Ax - Bx
    XPA = -b / p * X
    XPB = a / p * X
    E = np.zeros((la + 1, lb + 1, la + lb + 2) + np.shape(p))
    E[0, 0, 0] = np.exp(-mu * X * X)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t]
                    if t > 0:
                        v = v + E[i - 1, j, t - 1] / (2 * p)
                    v = v + (t + 1) * E[i - 1, j, t + 1]
                else:
                    v = XPB * E[i, j - 1, t]
                    if t > 0:
                        v = v + E[i, j - 1, t - 1] / (2 * p)
                    v = v + (t + 1) * E[i, j - 1, t + 1]
                E[i, j, t] = v
    return E[:, :, :la + lb + 1]


def hermite_R(L, alpha, X, Y, Z):
    """R[t, u, v] (t+u+v <= L) at n = 0, arrays broadcast over primitive combinations."""
    shp = np.broadcast(alpha, X, Y, Z).shape
    F = boys(L, alpha * (X * X + Y * Y + Z * Z))
    Rn = {}
    for n in range(L + 1):
        Rn[(n, 0, 0, 0)] = (-2.0 * alpha) ** n * F[n]
    # build increasing total order
    for tot in range(1, L + 1):
        for n in range(L - tot + 1):
            for t in range(tot + 1):
                for u in range(tot - t + 1):
                    v = tot - t - u
                    if t > 0:
                        val = X * Rn[(n + 1, t - 1, u, v)]
                        if t > 1:
                            val = val + (t - 1) * Rn[(n + 1, t - 2, u, v)]
                    elif u > 0:
                        val = Y * Rn[(n + 1, t, u - 1, v)]
                        if u > 1:
                            val = val + (u - 1) * Rn[(n + 1, t, u - 2, v)]
                    else:
                        val = Z * Rn[(n + 1, t, u, v - 1)]
                        if v > 1:
                            val = val + (v - 1) * Rn[(n + 1, t, u, v - 2)]
                    Rn[(n, t, u, v)] = val
    R = np.zeros((L + 1, L + 1, L + 1) + shp)
    for (n, t, u, v), val in Rn.items():
        if n == 0:
            R[t, u, v] = val
    return R


class PairData:
    """Hermite expansion of all Cartesian component pairs of a shell pair, contracted-ready."""

    def __init__(self, A, B):
        a = A.exps[:, None]
        b = B.exps[None, :]
        self.p = (a + b).ravel()
        self.P = ((a[..., None] * A.center + b[..., None] * B.center) / (a + b)[..., None]).reshape(-1, 3)
        self.cc = (A.coefs[:, None] * B.coefs[None, :]).ravel()
        L = A.l + B.l
        self.L = L
        Ex = [hermite_E(A.l, B.l, a, b, A.center[d], B.center[d]).reshape(A.l + 1, B.l + 1, L + 1, -1)
              for d in range(3)]
        ncart = len(A.carts) * len(B.carts)
        E = np.zeros((self.p.size, ncart, L + 1, L + 1, L + 1))
        k = 0
        for (ax, ay, az) in A.carts:
            for (bx, by, bz) in B.carts:
                E[:, k] = (Ex[0][ax, bx][:, None, None, :].transpose(3, 0, 1, 2)
                           * Ex[1][ay, by][None, :, None, :].transpose(3, 0, 1, 2)
                           * Ex[2][az, bz][None, None, :, :].transpose(3, 0, 1, 2))
                k += 1
        self.E = E
        self.Ex = Ex
        self.A, self.B = A, B


def one_electron(shells, atoms):
    """S, T, V in the spherical AO basis; atoms = [(Z, xyz)]."""
    nf = sum(s.nfunc for s in shells)
    S = np.zeros((nf, nf))
    T = np.zeros((nf, nf))
    V = np.zeros((nf, nf))
    offs = np.cumsum([0] + [s.nfunc for s in shells])
    for ia, A in enumerate(shells):
        for ib, B in enumerate(shells):
            pd = PairData(A, B)
            p = pd.p
            pref = pd.cc * (np.pi / p) ** 1.5
            # overlap and kinetic from 1D Hermite E at t = 0 (kinetic via the b-shifted overlaps)
            a_ = np.repeat(A.exps, len(B.exps))
            b_ = np.tile(B.exps, len(A.exps))
            Eb2 = [hermite_E(A.l, B.l + 2, A.exps[:, None], B.exps[None, :], A.center[d], B.center[d])
                   .reshape(A.l + 1, B.l + 3, -1, a_.size)[:, :, 0, :] for d in range(3)]
            s_blk = np.zeros((len(A.carts), len(B.carts)))
            t_blk = np.zeros_like(s_blk)
            for i, ca in enumerate(A.carts):
                for j, cb in enumerate(B.carts):
                    s1 = [Eb2[d][ca[d], cb[d]] for d in range(3)]
                    k1 = []
                    for d in range(3):
                        jj = cb[d]
                        v = -2.0 * b_ ** 2 * Eb2[d][ca[d], jj + 2] + b_ * (2 * jj + 1) * Eb2[d][ca[d], jj]
                        if jj >= 2:
                            v = v - 0.5 * jj * (jj - 1) * Eb2[d][ca[d], jj - 2]
                        k1.append(v)
                    s_blk[i, j] = np.sum(pref * s1[0] * s1[1] * s1[2])
                    t_blk[i, j] = np.sum(pref * (k1[0] * s1[1] * s1[2] + s1[0] * k1[1] * s1[2]
                                                 + s1[0] * s1[1] * k1[2]))
            v_blk = np.zeros_like(s_blk)
            for Zc, C in atoms:
                R = hermite_R(pd.L, p, pd.P[:, 0] - C[0], pd.P[:, 1] - C[1], pd.P[:, 2] - C[2])
                # sum_tuv E_tuv R_tuv
                val = np.einsum('kctuv,tuvk->c', pd.E * (pd.cc * 2 * np.pi / p)[:, None, None, None, None], R)
                v_blk -= Zc * val.reshape(len(A.carts), len(B.carts))
            sa, sb = slice(offs[ia], offs[ia + 1]), slice(offs[ib], offs[ib + 1])
            S[sa, sb] = A.T @ s_blk @ B.T.T
            T[sa, sb] = A.T @ t_blk @ B.T.T
            V[sa, sb] = A.T @ v_blk @ B.T.T
    return S, T, V


def dipole_integrals(shells, origin=(0.0, 0.0, 0.0)):
    """<mu| r_c - O_c |nu> for c = x, y, z in the spherical AO basis, (3, n, n).

    Along c the primitive pair gives sqrt(pi/p) (E^{ij}_1 + X_PC E^{ij}_0) (the first
    Hermite moment), the other two directions their overlaps (eT's dipole
    integrals come from Libint, src/libint/eri.cpp; this is the same quantity).
    """
    nf = sum(s.nfunc for s in shells)
    offs = np.cumsum([0] + [s.nfunc for s in shells])
    D = np.zeros((3, nf, nf))
    O = np.asarray(origin, dtype=float)
    for ia, A in enumerate(shells):
        for ib, B in enumerate(shells):
            blk = np.zeros((3, len(A.carts), len(B.carts)))
            for i_a, a in enumerate(A.exps):
                for i_b, b in enumerate(B.exps):
                    p = a + b
                    P = (a * A.center + b * B.center) / p
                    cc = A.coefs[i_a] * B.coefs[i_b] * (np.pi / p) ** 1.5
                    E = [hermite_E(A.l, B.l, a, b, A.center[d], B.center[d]) for d in range(3)]
                    for ka, ca in enumerate(A.carts):
                        for kb, cb in enumerate(B.carts):
                            s0 = [E[d][ca[d], cb[d], 0] for d in range(3)]
                            m1 = [(E[d][ca[d], cb[d], 1] if ca[d] + cb[d] >= 1 else 0.0)
                                  + (P[d] - O[d]) * s0[d] for d in range(3)]
                            for c in range(3):
                                others = [s0[d] for d in range(3) if d != c]
                                blk[c, ka, kb] += cc * m1[c] * others[0] * others[1]
            sa, sb = slice(offs[ia], offs[ia + 1]), slice(offs[ib], offs[ib + 1])
            for c in range(3):
                D[c, sa, sb] = A.T @ blk[c] @ B.T.T
    return D


def eri(shells):
    """(pq|rs) in the spherical AO basis, full 4-index array."""
    nf = sum(s.nfunc for s in shells)
    offs = np.cumsum([0] + [s.nfunc for s in shells])
    G = np.zeros((nf, nf, nf, nf))
    pairs = {}
    for ia in range(len(shells)):
        for ib in range(ia + 1):
            pairs[(ia, ib)] = PairData(shells[ia], shells[ib])
    keys = sorted(pairs)
    for n1, (ia, ib) in enumerate(keys):
        P = pairs[(ia, ib)]
        for (ic, id_) in keys[:n1 + 1]:
            Q = pairs[(ic, id_)]
            L = P.L + Q.L
            p = P.p[:, None]
            q = Q.p[None, :]
            alpha = p * q / (p + q)
            XPQ = P.P[:, None, :] - Q.P[None, :, :]
            R = hermite_R(L, alpha, XPQ[..., 0], XPQ[..., 1], XPQ[..., 2])     # (L+1)^3 x nP x nQ
            pref = 2.0 * np.pi ** 2.5 / (p * q * np.sqrt(p + q)) * P.cc[:, None] * Q.cc[None, :]
            La, Lc = P.L, Q.L
            sign = np.array([(-1) ** (t + u + v) for t in range(Lc + 1) for u in range(Lc + 1)
                             for v in range(Lc + 1)]).reshape(Lc + 1, Lc + 1, Lc + 1)
            Ecd = Q.E * sign[None, None]
            idx = np.arange(La + 1)
            jdx = np.arange(Lc + 1)
            T_ = idx[:, None, None, None, None, None] + jdx[None, None, None, :, None, None]
            U_ = idx[None, :, None, None, None, None] + jdx[None, None, None, None, :, None]
            V_ = idx[None, None, :, None, None, None] + jdx[None, None, None, None, None, :]
            Rg = R[T_, U_, V_]                                            # (La+1)^3 (Lc+1)^3 nP nQ
            Rg = Rg.reshape((La + 1) ** 3, (Lc + 1) ** 3, p.shape[0], q.shape[1])
            Ea = P.E.reshape(P.E.shape[0], P.E.shape[1], -1)              # nP, ncart_ab, H
            Ec = Ecd.reshape(Ecd.shape[0], Ecd.shape[1], -1)              # nQ, ncart_cd, H'
            val = np.einsum('PaH,HKPQ,PQ,QcK->ac', Ea, Rg, pref, Ec, optimize=True)
            A, B, C, D = shells[ia], shells[ib], shells[ic], shells[id_]
            val = val.reshape(len(A.carts), len(B.carts), len(C.carts), len(D.carts))
            val = np.einsum('ia,jb,kc,ld,abcd->ijkl', A.T, B.T, C.T, D.T, val, optimize=True)
            sa = slice(offs[ia], offs[ia + 1])
            sb = slice(offs[ib], offs[ib + 1])
            sc = slice(offs[ic], offs[ic + 1])
            sd = slice(offs[id_], offs[id_ + 1])
            for (x, y, X, Y) in ((sa, sb, 0, 1), (sb, sa, 1, 0)):
                vv = val if (X, Y) == (0, 1) else val.transpose(1, 0, 2, 3)
                G[x, y, sc, sd] = vv
                G[x, y, sd, sc] = vv.transpose(0, 1, 3, 2)
                G[sc, sd, x, y] = vv.transpose(2, 3, 0, 1)
                G[sd, sc, x, y] = vv.transpose(3, 2, 0, 1)
    return G


def build_molecule(geometry_angstrom, basis, bohr_to_angstrom):
    shells, atoms = [], []
    for sym, x, y, z in geometry_angstrom:
        xyz = np.array([x, y, z]) / bohr_to_angstrom
        atoms.append((Z_OF[sym], xyz))
        for sh in basis[sym]:
            shells.append(Shell(xyz, sh['l'], sh['exps'], sh['coefs']))
    return shells, atoms


def nuclear_repulsion(atoms):
    e = 0.0
    for (Za, A), (Zb, B) in itertools.combinations(atoms, 2):
        e += Za * Zb / np.linalg.norm(A - B)
    return e


def rhf(S, H, G, n_occ, e_nuc, conv=1e-12, max_it=200):
    """Restricted HF with DIIS; returns (E, C, eps)."""
    s, U = np.linalg.eigh(S)
    X = U / np.sqrt(s)
    F = H.copy()
    e_old = 0.0
    errs, focks = [], []
    for it in range(max_it):
        Fp = X.T @ F @ X
        eps, Cp = np.linalg.eigh(Fp)
        C = X @ Cp
        D = C[:, :n_occ] @ C[:, :n_occ].T
        J = np.einsum('pqrs,rs->pq', G, D)
        K = np.einsum('prqs,rs->pq', G, D)
        F = H + 2.0 * J - K
        E = float(np.sum(D * (H + F))) + e_nuc
        err = F @ D @ S - S @ D @ F
        errs.append(err)
        focks.append(F)
        if len(errs) > 8:
            errs.pop(0)
            focks.pop(0)
        if np.abs(err).max() < conv and abs(E - e_old) < 1e-13:
            Fp = X.T @ F @ X
            eps, Cp = np.linalg.eigh(Fp)
            return E, X @ Cp, eps
        e_old = E
        n = len(errs)
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        for i in range(n):
            for j in range(n):
                B[i, j] = np.sum(errs[i] * errs[j])
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        w = np.linalg.solve(B, rhs)[:n]
        F = sum(wi * Fi for wi, Fi in zip(w, focks))
    raise RuntimeError('RHF did not converge')


def pivoted_cholesky(G, threshold):
    """L (n_J, n, n) with (pq|rs) ~ sum_J L_pq L_rs, pivoting on the largest remaining diagonal."""
    n = G.shape[0]
    M = G.reshape(n * n, n * n)
    d = np.diag(M).copy()
    vecs = []
    while True:
        piv = int(np.argmax(d))
        if d[piv] <= threshold:
            break
        col = M[:, piv].copy()
        for v in vecs:
            col -= v * v[piv]
        v = col / math.sqrt(d[piv])
        vecs.append(v)
        d -= v * v
        d[piv] = 0.0
    L = np.array(vecs).reshape(len(vecs), n, n)
    return 0.5 * (L + L.transpose(0, 2, 1))


def molecular_inputs(fixture):
    """(n_o, E_nuc, E_HF, eps, h_mo, L_mo) for a fixture dict."""
    shells, atoms = build_molecule(fixture['geometry_angstrom'], fixture['basis'],
                                   fixture['bohr_to_angstrom'])
    S, T, V = one_electron(shells, atoms)
    G = eri(shells)
    e_nuc = nuclear_repulsion(atoms)
    n_el = sum(z for z, _ in atoms)
    n_o = n_el // 2
    E, C, eps = rhf(S, T + V, G, n_o, e_nuc)
    L_ao = pivoted_cholesky(G, fixture['cholesky_threshold'])
    L_mo = np.einsum('mp,Jmn,nq->Jpq', C, L_ao, C, optimize=True)
    h_mo = C.T @ (T + V) @ C
    return dict(n_o=n_o, e_nuc=e_nuc, e_hf=E, eps=eps, h_mo=h_mo, L_mo=L_mo, C=C, n_J=L_ao.shape[0])
