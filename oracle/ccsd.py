"""FP64 numpy restatement of eT's closed-shell CCSD ground-state path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the checker for the HIP
implementation, never the thing measured or shipped.

Array conventions follow eT (all indices 0-based here):

* Cholesky vectors ``L[J, p, q]`` over the full MO range, occupied first
  (eT keeps them as L(J, p, q), J fastest: src/eri_tool/abstract_eri_cholesky_class.F90:76-119).
* Singles ``t1[a, i]`` (eT ``wf%t1(n_v, n_o)``).
* Doubles packed as in eT: compound ``ai = a + n_v*i`` and the packed index
  ``max(ai,bj)*(max(ai,bj)+1)/2 + min(ai,bj)``
  (src/tools/reordering.F90:4174-4198, src/wavefunctions/doubles/doubles_class.F90:454).
* 4-index tensors are held in eT's index order, e.g. ``t_aibj[a, i, b, j]``.
* Residual vector ``omega = [omega1 (v,o) column-major, packed omega2]``
  (src/wavefunctions/ccsd/omega_ccsd.F90:53-55).
"""
import numpy as np


# ----------------------------------------------------------------------------
# index helpers (src/tools/reordering.F90:4174-4293, packed_array_utilities_r.F90)
# ----------------------------------------------------------------------------

def packed_index(n):
    """idx[p, q] of the symmetric packed storage used by eT (reordering.F90:4193)."""
    p = np.arange(n)
    hi = np.maximum.outer(p, p)
    lo = np.minimum.outer(p, p)
    return hi * (hi + 1) // 2 + lo


def _fast(*arrays):
    """Real FP64 arrays take the OpenMP helpers of ladder_pm.c (same values as the
    numpy forms below, which complex-step derivatives use)."""
    return all(a.dtype == np.float64 for a in arrays)


def squareup(packed, n):
    """squareup_real (reordering.F90:4174): packed -> full symmetric (n, n)."""
    if _fast(packed) and n > 64:
        import ctypes
        full = np.empty((n, n))
        p = np.ascontiguousarray(packed)
        dp = ctypes.POINTER(ctypes.c_double)
        _ladder_c().sym_unpack(p.ctypes.data_as(dp), n, full.ctypes.data_as(dp))
        return full
    return packed[packed_index(n)]


def packin(full):
    """packin_real (reordering.F90:4267): lower triangle full[i, j], i >= j -> packed."""
    n = full.shape[0]
    if _fast(full) and n > 64:
        import ctypes
        f = np.ascontiguousarray(full)
        out = np.empty(n * (n + 1) // 2)
        dp = ctypes.POINTER(ctypes.c_double)
        _ladder_c().sym_pack_lower(f.ctypes.data_as(dp), n, out.ctypes.data_as(dp))
        return out
    i, j = np.tril_indices(n)
    out = np.empty(n * (n + 1) // 2, dtype=full.dtype)
    out[i * (i + 1) // 2 + j] = full[i, j]
    return out


def contiguous(x):
    """C-ordered copy of a (strided) 4-index array."""
    if (x.ndim != 4 or not _fast(x) or x.size < (1 << 16) or x.flags.c_contiguous
            or any(st <= 0 or st % 8 for st in x.strides)):
        return np.ascontiguousarray(x)
    import ctypes
    out = np.empty(x.shape)
    shape = np.array(x.shape, dtype=np.int64)
    strides = np.array([st // 8 for st in x.strides], dtype=np.int64)
    ip, dp = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)
    base = x.ctypes.data
    _ladder_c().permute4_copy(ctypes.cast(base, dp), shape.ctypes.data_as(ip), strides.ctypes.data_as(ip),
                              out.ctypes.data_as(dp))
    return out


def mat_to_aibj(m, n_v, n_o):
    """(ov, ov) matrix with row ai = a + n_v*i -> tensor [a, i, b, j]."""
    return m.reshape(n_o, n_v, n_o, n_v).transpose(1, 0, 3, 2)


def aibj_to_mat(x):
    """tensor [a, i, b, j] -> (ov, ov) matrix with row ai = a + n_v*i."""
    n_v, n_o = x.shape[0], x.shape[1]
    return contiguous(x.transpose(1, 0, 3, 2)).reshape(n_o * n_v, n_o * n_v)


def unpack_t2(t2, n_v, n_o):
    return mat_to_aibj(squareup(t2, n_v * n_o), n_v, n_o)


def pack_aibj(x):
    return packin(aibj_to_mat(x))


def n_t2(n_o, n_v):
    return n_o * n_v * (n_o * n_v + 1) // 2


# ----------------------------------------------------------------------------
# integrals from Cholesky vectors (src/eri_tool/eri_tool_class.F90:92-163)
# ----------------------------------------------------------------------------

def eri(L, s1, s2, s3, s4):
    """g_pqrs = sum_J L_pq^J L_rs^J over the MO ranges s1..s4 (DGEMM 'T','N', K = n_J)."""
    a = L[:, s1, s2]
    b = L[:, s3, s4]
    n_J = L.shape[0]
    g = a.reshape(n_J, -1).T @ b.reshape(n_J, -1)
    return g.reshape(a.shape[1], a.shape[2], b.shape[1], b.shape[2])


# ----------------------------------------------------------------------------
# T1 transformation (src/wavefunctions/ccs/t1_ccs.F90)
# ----------------------------------------------------------------------------

def t1_transform(Z, t1, n_o):
    """t1_transform_ccs (t1_ccs.F90:34-95): Z <- X Z Y^T, X = I - t1, Y = I + t1^T."""
    n = Z.shape[0]
    dt = np.result_type(Z, t1)
    X = np.eye(n, dtype=dt)
    Y = np.eye(n, dtype=dt)
    X[n_o:, :n_o] = -t1
    Y[:n_o, n_o:] = t1.T
    return X @ Z @ Y.T


def t1_cholesky(L_mo, t1, n_o):
    """construct_t1_cholesky_ccs (t1_ccs.F90:208-499), block by block as eT does it.

    oo: L_ij += sum_a L_ia t_aj                                   (:249-308)
    vo: L_ai += sum_b L_ab t_bi - sum_j t_aj (L_ji + sum_b L_jb t_bi)   (:311-432)
    vv: L_ab -= sum_j t_aj L_jb                                   (:435-499)
    ov: unchanged.
    """
    o = slice(0, n_o)
    v = slice(n_o, L_mo.shape[1])
    L = L_mo.astype(np.result_type(L_mo, t1))
    L[:, o, o] = L_mo[:, o, o] + np.einsum('Jia,aj->Jij', L_mo[:, o, v], t1)
    Lp_ji = L_mo[:, o, o] + np.einsum('Jjb,bi->Jji', L_mo[:, o, v], t1)
    L[:, v, o] = (L_mo[:, v, o] + np.einsum('Jab,bi->Jai', L_mo[:, v, v], t1)
                  - np.einsum('aj,Jji->Jai', t1, Lp_ji))
    L[:, v, v] = L_mo[:, v, v] - np.einsum('aj,Jjb->Jab', t1, L_mo[:, o, v])
    return L


def fock_t1(h_mo, L_t1, t1, n_o):
    """construct_fock_ccsd (fock_ccsd.F90:34-79 -> fock_ccs.F90:242-1109).

    F_pq = h^T1_pq + sum_k (2 g_pqkk - g_pkkq), integrals from the T1 vectors,
    h^T1 = X h Y^T (oei_ccs.F90:35-97).
    """
    o = slice(0, n_o)
    h = t1_transform(h_mo, t1, n_o)
    X_J = np.einsum('Jkk->J', L_t1[:, o, o])
    coul = 2.0 * np.einsum('Jpq,J->pq', L_t1, X_J)
    exch = np.einsum('Jpk,Jkq->pq', L_t1[:, :, o], L_t1[:, o, :])
    return h + coul - exch


def hf_energy(h_mo, L_mo, n_o):
    """E_HF = sum_i (h_ii + F_ii) for the canonical reference."""
    o = slice(0, n_o)
    F = fock_t1(h_mo, L_mo, np.zeros((L_mo.shape[1] - n_o, n_o)), n_o)
    return float(np.trace(h_mo[o, o]) + np.trace(F[o, o]))


# ----------------------------------------------------------------------------
# energy, guess, denominators (mean_value_ccsd.F90, ccsd_class.F90, doubles_class.F90)
# ----------------------------------------------------------------------------

def correlation_energy(L_t1, t1, t2, n_o):
    """calculate_energy_ccsd (mean_value_ccsd.F90:36-99).

    E_corr = sum_aibj (t_aibj + t_ai t_bj)(2 g_iajb - g_ibja), g from L_t1.
    """
    n_v = L_t1.shape[1] - n_o
    o = slice(0, n_o)
    v = slice(n_o, n_o + n_v)
    g = eri(L_t1, o, v, o, v)                                 # g_iajb
    tau = unpack_t2(t2, n_v, n_o) + np.einsum('ai,bj->aibj', t1, t1)
    Lg = 2.0 * g - g.transpose(0, 3, 2, 1)                    # 2 g_iajb - g_ibja
    e = np.einsum('aibj,iajb->', tau, Lg)
    return e if np.iscomplexobj(e) else float(e)          # complex-safe for complex steps


def mp2_guess(L_t1, eps, n_o):
    """set_t2_to_cc2_guess (ccsd_class.F90:478-533): t_aibj = g_aibj / (e_i - e_a + e_j - e_b)."""
    n_v = L_t1.shape[1] - n_o
    o = slice(0, n_o)
    v = slice(n_o, n_o + n_v)
    g = eri(L_t1, v, o, v, o)
    e_o, e_v = eps[o], eps[v]
    d = (e_o[None, :, None, None] - e_v[:, None, None, None]
         + e_o[None, None, None, :] - e_v[None, None, :, None])
    return pack_aibj(g / d)


def orbital_differences(eps, n_o):
    """get_orbital_differences_doubles (doubles_class.F90:421-470) incl. singles (ccs)."""
    n_v = eps.shape[0] - n_o
    e_o, e_v = eps[:n_o], eps[n_o:]
    d1 = (e_v[:, None] - e_o[None, :])
    d2 = d1[:, :, None, None] + d1[None, None, :, :]
    return np.concatenate([d1.ravel(order='F'), pack_aibj(d2)])


def doubles_diagonal(n_t1):
    """Positions of the doubles diagonal (ai = bj) in [t1 | packed t2]."""
    d = np.arange(n_t1)
    return n_t1 + d * (d + 1) // 2 + d


def scale_doubles_diagonal(x, n_t1, factor=2.0):
    """scale_amplitudes_ccsd (ccsd_class.F90:648-671): packed doubles diagonal *= factor."""
    y = x.copy()
    n = int(n_t1)
    d = np.arange(n)
    y[n_t1 + d * (d + 1) // 2 + d] *= factor
    return y


# ----------------------------------------------------------------------------
# Omega (src/wavefunctions/ccsd/omega_ccsd.F90, doubles/omega_doubles.F90, ccs/omega_ccs.F90)
# ----------------------------------------------------------------------------

def construct_u(t_aibj):
    """construct_u_aibj_ccsd (omega_ccsd.F90:881-914): u_aibj = 2 t_aibj - t_ajbi."""
    return 2.0 * t_aibj - t_aibj.transpose(0, 3, 2, 1)


def omega_singles(L_mo, L_t1, F, u, n_o):
    """Omega_ai: CCS A1 (F_ai, omega_ccs.F90:62-86) + doubles A1, B1, C1 (omega_doubles.F90)."""
    n_v = u.shape[0]
    o = slice(0, n_o)
    v = slice(n_o, n_o + n_v)
    om = F[v, o].copy()
    # A1 (omega_doubles.F90:36-139): X_Jbi = sum_cj L_Jcj u_cjbi (L_mo vo block),
    # Omega_ai += sum_Jb L^T1_Jab X_Jbi
    X = np.einsum('Jcj,cjbi->Jbi', L_mo[:, v, o], u, optimize=True)
    om += np.einsum('Jab,Jbi->ai', L_t1[:, v, v], X, optimize=True)
    # B1 (omega_doubles.F90:141-205): Omega_ai -= sum_jbk u_ajbk g_kbji
    g_kbji = eri(L_t1, o, v, o, o)
    om -= np.einsum('ajbk,kbji->ai', u, g_kbji, optimize=True)
    # C1 (omega_doubles.F90:207-250): Omega_ai += sum_bj u_aibj F_jb
    om += np.einsum('aibj,jb->ai', u, F[o, v], optimize=True)
    return om


def omega_e2(F, L_t1, t, n_o):
    """omega_ccsd_e2 (omega_ccsd.F90:712-878); result ordered as eT leaves it (pre-symmetrisation)."""
    n_v = t.shape[0]
    o = slice(0, n_o)
    v = slice(n_o, n_o + n_v)
    u_bldk = 2.0 * t.transpose(0, 3, 2, 1) - t               # u_kl^bd = 2 t_bkdl - t_bldk
    g_ldkc = eri(L_t1, o, v, o, v)
    X_bc = F[v, v] - np.einsum('bldk,ldkc->bc', u_bldk, g_ldkc, optimize=True)
    om = np.einsum('bc,cjai->bjai', X_bc, t, optimize=True)     # E2.1, ordered (b,j,a,i)
    Y_kj = F[o, o] + np.einsum('kcld,cldj->kj', g_ldkc, u_bldk, optimize=True)
    om -= np.einsum('aibk,kj->aibj', t, Y_kj, optimize=True)    # E2.2
    return om


def omega_b2(L_t1, t, n_o):
    """omega_ccsd_b2 (omega_ccsd.F90:483-583) in aibj order.

    Omega_aibj += g_aibj + sum_kl t_akbl (g_kilj + sum_cd t_cidj g_kcld)
    """
    n_v = t.shape[0]
    o = slice(0, n_o)
    v = slice(n_o, n_o + n_v)
    X_klij = (eri(L_t1, o, o, o, o).transpose(0, 2, 1, 3)
              + np.einsum('kcld,cidj->klij', eri(L_t1, o, v, o, v), t, optimize=True))
    return eri(L_t1, v, o, v, o) + np.einsum('akbl,klij->aibj', t, X_klij, optimize=True)


def symmetrize_and_pack(x_mat, n_v, n_o, contra=False, psrq=False):
    """symmetrize_and_pack (packed_array_utilities_r.F90:586-716) of a full (ov,ov) matrix."""
    x = mat_to_aibj(x_mat, n_v, n_o)                           # x[p,q,r,s]
    x_psrq = x.transpose(0, 3, 2, 1)                           # x(p,s,r,q)
    if not contra:
        y = x_psrq if psrq else x
    else:
        f1, f2 = (1.0 / 3.0, 2.0 / 3.0) if psrq else (2.0 / 3.0, 1.0 / 3.0)
        y = f1 * x + f2 * x_psrq
    ym = aibj_to_mat(y)
    s = ym + ym.T
    d = np.arange(s.shape[0])
    s[d, d] = x_mat[d, d]
    return packin(s)


def symmetric_product(alpha, A, B):
    """alpha * A @ B, the (ov) x (ov) DSYMM products of omega_ccsd_c2_d2 (omega_ccsd.F90:647-688)."""
    return alpha * (A @ B)


def omega_c2_d2(L_t1, t2, n_o, n_v):
    """omega_ccsd_c2_d2 (omega_ccsd.F90:586-709): the four (ov)x(ov) DSYMMs.

    T[(p,q),(r,s)] = t_psrq and U[(p,q),(r,s)] = u_pqrs are the two triangles of
    eT's t2_u2 (construct_1432_and_contra, packed_array_utilities_r.F90:467).
    Returns the packed contribution.
    """
    o = slice(0, n_o)
    v = slice(n_o, n_o + n_v)
    t = unpack_t2(t2, n_v, n_o)
    T = aibj_to_mat(t.transpose(0, 3, 2, 1))
    U = aibj_to_mat(construct_u(t))
    del t
    # C_term[(c,k),(a,i)] = g_acki ; D_term = 2 g_aikc - g_acki   (:632-640)
    g_vvoo = eri(L_t1, v, v, o, o)                             # g_abij
    C = aibj_to_mat(g_vvoo.transpose(1, 2, 0, 3))              # C(b,i,a,j) = g_abij
    del g_vvoo
    g_voov = eri(L_t1, v, o, o, v)                             # g_aijb
    D = -C + 2.0 * aibj_to_mat(g_voov.transpose(3, 2, 0, 1))   # D(c,k,a,i) += 2 g_aikc
    del g_voov
    # G[(c,k),(d,l)] = g_lckd  (sort_1234_to_2341 of g_ovov)   (:643-644)
    g_ovov = eri(L_t1, o, v, o, v)
    G = aibj_to_mat(g_ovov.transpose(1, 2, 3, 0))              # G(c,l,d,k) = g_kcld
    del g_ovov
    C += symmetric_product(-0.5, G, T)                         # dsymm R,U (:648)
    x = mat_to_aibj(G, n_v, n_o)
    Gc = aibj_to_mat(2.0 * x.transpose(0, 3, 2, 1) - x)       # contra, psrq (:661)
    del x, G
    D += symmetric_product(0.5, Gc, U)                         # dsymm R,L (:663)
    del Gc
    W = symmetric_product(-1.5, T, C)                          # dsymm L,U (:675)
    del C, T
    out = symmetrize_and_pack(W, n_v, n_o, contra=True, psrq=True)
    del W
    V = symmetric_product(0.5, U, D)                           # dsymm L,L (:689)
    del U, D
    out += symmetrize_and_pack(V, n_v, n_o)
    return out


def pair_index_plus(n):
    """(p, r) with p <= r -> r(r+1)/2 + p  (eT pr_p, packed_array_utilities_r.F90:174)."""
    return packed_index(n)


def omega_a2_plus_minus(L_t1, t2, n_o, n_v):
    """omega_ccsd_a2 (omega_ccsd.F90:113-480), single batch, right=.true.

    t+_{ij,cd} = (1-d_cd/2)(1-d_ij/2)(t_cidj + t_cjdi)      (c<=d, i<=j)
    t-_{ij,cd} = t_cidj - t_cjdi                             (c<d,  i<j)
    g+_{cd,ab} = g_acbd + g_adbc,  g-_{cd,ab} = g_acbd - g_adbc
    omega+-_{ij,ab} = 1/2 t+- g+- ; omega+ diagonal (a=b) blocks * 1/2
    then add_double_packed_plus_minus (packed_array_utilities_r.F90:335-399).
    Returns the packed contribution.
    """
    v = slice(n_o, n_o + n_v)
    t = unpack_t2(t2, n_v, n_o)                                # t[c,i,d,j]
    ci, cj = np.triu_indices(n_v)                              # c <= d pairs, ordered by packed index
    order_p = np.argsort(cj * (cj + 1) // 2 + ci)
    c_p, d_p = ci[order_p], cj[order_p]
    oi, oj = np.triu_indices(n_o)
    order_o = np.argsort(oj * (oj + 1) // 2 + oi)
    i_p, j_p = oi[order_o], oj[order_o]
    cm, dm = np.triu_indices(n_v, 1)
    order_m = np.argsort(dm * (dm - 1) // 2 + cm)
    c_m, d_m = cm[order_m], dm[order_m]
    im, jm = np.triu_indices(n_o, 1)
    order_om = np.argsort(jm * (jm - 1) // 2 + im)
    i_m, j_m = im[order_om], jm[order_om]

    # t+- (construct_plus_minus_2413_from_packed + scale_double_packed_blocks)
    tp = (t[c_p[None, :], i_p[:, None], d_p[None, :], j_p[:, None]]
          + t[c_p[None, :], j_p[:, None], d_p[None, :], i_p[:, None]])
    tp *= np.where(c_p == d_p, 0.5, 1.0)[None, :]
    tp *= np.where(i_p == j_p, 0.5, 1.0)[:, None]
    tm = (t[c_m[None, :], i_m[:, None], d_m[None, :], j_m[:, None]]
          - t[c_m[None, :], j_m[:, None], d_m[None, :], i_m[:, None]])
    # g+- (get_packed('vv') + construct_plus_minus_1324_from_RFP)
    g = eri(L_t1, v, v, v, v)                                  # g[a,c,b,d]
    gA = g.transpose(1, 3, 0, 2)                             # gA[c,d,a,b] = g_acbd
    gp = (gA[c_p[:, None], d_p[:, None], c_p[None, :], d_p[None, :]]
          + gA[d_p[:, None], c_p[:, None], c_p[None, :], d_p[None, :]])   # rows cd, cols ab
    gm = (gA[c_m[:, None], d_m[:, None], c_m[None, :], d_m[None, :]]
          - gA[d_m[:, None], c_m[:, None], c_m[None, :], d_m[None, :]])
    wp = 0.5 * tp @ gp                                         # omega+_{ij,ab}
    wm = 0.5 * tm @ gm if tm.size and gm.size else np.zeros((len(i_m), len(c_m)), dtype=tp.dtype)
    wp *= np.where(c_p == d_p, 0.5, 1.0)[None, :]              # a = b columns
    # add_double_packed_plus_minus: (a<=b, i<=j); omega(ai,bj) += w+ + w-, omega(aj,bi) += w+ - w-
    out = np.zeros(n_t2(n_o, n_v), dtype=wp.dtype)
    ov = n_o * n_v
    P = packed_index(ov)
    a_, b_ = c_p[None, :], d_p[None, :]
    i_, j_ = i_p[:, None], j_p[:, None]
    ai, bj = a_ + n_v * i_, b_ + n_v * j_
    aj, bi = a_ + n_v * j_, b_ + n_v * i_
    wm_full = np.zeros_like(wp)
    rows_m = np.searchsorted(j_p * (j_p + 1) // 2 + i_p, j_m * (j_m + 1) // 2 + i_m)
    cols_m = np.searchsorted(d_p * (d_p + 1) // 2 + c_p, d_m * (d_m + 1) // 2 + c_m)
    wm_full[np.ix_(rows_m, cols_m)] = wm
    np.add.at(out, P[ai, bj], wp + wm_full)
    np.add.at(out, P[aj, bi], wp - wm_full)
    return out


_ladder_lib = None


def _ladder_c():
    """The C helper of the batched ladder (oracle/ladder_pm.c, built by oracle/build.py)."""
    global _ladder_lib
    if _ladder_lib is None:
        import ctypes
        import os
        from . import build as _build
        if not os.path.exists(_build.LIB):
            _build.build()
        lib = ctypes.CDLL(_build.LIB)
        i64, dp = ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        lib.ladder_gather_pm.argtypes = [dp, i64, i64, i64, i64, i64, i64, ip, ip, dp, dp]
        lib.ladder_gather_pm.restype = None
        lib.ladder_scatter_pm.argtypes = [dp, dp, i64, i64, dp]
        lib.ladder_scatter_pm.restype = None
        lib.ladder_scatter_block.argtypes = [dp, dp, i64, ip, ip, i64, i64, dp]
        lib.ladder_scatter_block.restype = None
        lib.permute4_copy.argtypes = [dp, ip, ip, dp]
        lib.permute4_copy.restype = None
        lib.sym_unpack.argtypes = [dp, i64, dp]
        lib.sym_unpack.restype = None
        lib.sym_pack_lower.argtypes = [dp, i64, dp]
        lib.sym_pack_lower.restype = None
        _ladder_lib = lib
    return _ladder_lib


def ladder_t_plus_minus(t2, n_o, n_v):
    """construct_plus_minus_2413_from_packed + scale_double_packed_blocks
    (packed_array_utilities_r.F90:140-332): t+ (o+ x v+) and t- (o- x v-) with rows
    ij (i <= j, resp. i < j) and columns cd (c <= d, resp. c < d), each in the packed
    order j(j+1)/2 + i (resp. j(j-1)/2 + i)."""
    t = unpack_t2(t2, n_v, n_o)                                # t[c,i,d,j]
    oi, oj = np.triu_indices(n_o)
    order = np.argsort(oj * (oj + 1) // 2 + oi)
    i_p, j_p = oi[order], oj[order]
    ci, cj = np.triu_indices(n_v)
    order = np.argsort(cj * (cj + 1) // 2 + ci)
    c_p, d_p = ci[order], cj[order]
    im, jm = np.triu_indices(n_o, 1)
    order = np.argsort(jm * (jm - 1) // 2 + im)
    i_m, j_m = im[order], jm[order]
    cm, dm = np.triu_indices(n_v, 1)
    order = np.argsort(dm * (dm - 1) // 2 + cm)
    c_m, d_m = cm[order], dm[order]
    tp = np.empty((len(i_p), len(c_p)))
    for r, (i, j) in enumerate(zip(i_p, j_p)):           # one row of t+ per occupied pair
        tp[r] = t[c_p, i, d_p, j] + t[c_p, j, d_p, i]
    tp *= np.where(c_p == d_p, 0.5, 1.0)[None, :]
    tp *= np.where(i_p == j_p, 0.5, 1.0)[:, None]
    tm = np.empty((len(i_m), len(c_m)))
    for r, (i, j) in enumerate(zip(i_m, j_m)):
        tm[r] = t[c_m, i, d_m, j] - t[c_m, j, d_m, i]
    return tp, tm


def omega_a2_batched(L_t1, t2, n_o, n_v, batch):
    """omega_ccsd_a2 (omega_ccsd.F90:113-480) with eT's batching over virtual blocks:
    for each (a-batch A, b-batch B >= A) the block g_acbd (a in A, b in B) is built from
    the dressed vv Cholesky block (one product, K = n_J), its +/- combinations for the
    pairs a <= b of the block are gathered (ladder_pm.c), multiplied with t+/- (two
    products) and the results scattered into the packed residual.  Equals
    omega_a2_plus_minus for any batch length.  Returns the packed contribution."""
    lib = _ladder_c()
    import ctypes
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int64)
    v = slice(n_o, n_o + n_v)
    n_J = L_t1.shape[0]
    vp, vm = n_v * (n_v + 1) // 2, n_v * (n_v - 1) // 2
    tp, tm = ladder_t_plus_minus(t2, n_o, n_v)
    out = np.zeros(n_t2(n_o, n_v))
    Lvv = np.ascontiguousarray(L_t1[:, v, v])                  # Lvv[J, a, c]
    starts = list(range(0, n_v, batch))
    for a0 in starts:
        la = min(batch, n_v - a0)
        LA = np.ascontiguousarray(Lvv[:, a0:a0 + la, :]).reshape(n_J, la * n_v)
        for b0 in starts:
            if b0 < a0:
                continue
            lb = min(batch, n_v - b0)
            LB = np.ascontiguousarray(Lvv[:, b0:b0 + lb, :]).reshape(n_J, lb * n_v)
            G = LA.T @ LB                                      # G[(a, c), (b, d)] = g_acbd
            aa, bb = np.meshgrid(np.arange(a0, a0 + la), np.arange(b0, b0 + lb), indexing='ij')
            keep = aa <= bb
            pa = np.ascontiguousarray(aa[keep], dtype=np.int64)
            pb = np.ascontiguousarray(bb[keep], dtype=np.int64)
            npairs = pa.size
            gp = np.empty((npairs, vp))
            gm = np.empty((npairs, max(vm, 1)))
            lib.ladder_gather_pm(G.ctypes.data_as(dp), n_v, a0, la, b0, lb, npairs, pa.ctypes.data_as(ip),
                                 pb.ctypes.data_as(ip), gp.ctypes.data_as(dp), gm.ctypes.data_as(dp))
            del G
            wp = np.ascontiguousarray(0.5 * (tp @ gp.T))       # (o+ x pairs)
            wm = np.ascontiguousarray(0.5 * (tm @ gm.T)) if vm and tm.shape[0] else np.zeros((1, npairs))
            lib.ladder_scatter_block(wp.ctypes.data_as(dp), wm.ctypes.data_as(dp), npairs, pa.ctypes.data_as(ip),
                                     pb.ctypes.data_as(ip), n_o, n_v, out.ctypes.data_as(dp))
    return out


def omega_a2_naive(L_t1, t2, n_o, n_v):
    """Ladder sum_cd g_acbd t_cidj with the (1 - delta_{ai,bj}/2) packing factor (test helper)."""
    v = slice(n_o, n_o + n_v)
    t = unpack_t2(t2, n_v, n_o)
    g = eri(L_t1, v, v, v, v)                                  # g_acbd
    W = np.einsum('acbd,cidj->aibj', g, t, optimize=True)
    Wm = aibj_to_mat(W)
    d = np.arange(Wm.shape[0])
    Wm[d, d] *= 0.5
    return packin(Wm)


def omega(L_mo, L_t1, F, t1, t2, n_o, ladder_batch=None):
    """construct_omega_ccsd (omega_ccsd.F90:39-110).  Returns omega(n_t1 + n_t2).

    ladder_batch: virtual batch length of the ladder term (eT's batched form,
    omega_a2_batched); None forms g_vvvv whole (omega_a2_plus_minus)."""
    n_v = L_mo.shape[1] - n_o
    ov = n_o * n_v
    t = unpack_t2(t2, n_v, n_o)
    u = construct_u(t)
    om1 = omega_singles(L_mo, L_t1, F, u, n_o)
    # doubles: E2, symmetric sum, B2, 1/2 diagonal, pack   (:78-102)
    e2 = omega_e2(F, L_t1, t, n_o)
    m = aibj_to_mat(e2)
    m = m + m.T
    m += aibj_to_mat(omega_b2(L_t1, t, n_o))
    d = np.arange(ov)
    m[d, d] *= 0.5
    om2 = packin(m)
    om2 += omega_c2_d2(L_t1, t2, n_o, n_v)                     # (:105)
    if ladder_batch is None:
        om2 += omega_a2_plus_minus(L_t1, t2, n_o, n_v)         # (:106)
    else:
        om2 += omega_a2_batched(L_t1, t2, n_o, n_v, ladder_batch)
    return np.concatenate([om1.ravel(order='F'), om2])


# ----------------------------------------------------------------------------
# ground-state solver restatement (diis_cc_gs_class.F90:209-315, diis_tool_class.F90,
# quasi_newton_updater_class.F90:81-118, amplitude_updater_class.F90:86-120)
# ----------------------------------------------------------------------------

class DIIS:
    """diis_tool (diis_tool_class.F90:302-760), records in memory, no CROP."""

    def __init__(self, dimension=8):
        self.dimension = dimension
        self.iteration = 1
        self.e = []
        self.x = []
        self.G = np.zeros((0, 0))

    def update(self, e, x):
        if len(self.e) == self.dimension:
            self.e.pop(0)
            self.x.pop(0)
            self.G = self.G[1:, 1:]
        self.e.append(e.copy())
        self.x.append(x.copy())
        n = len(self.e)
        G = np.zeros((n, n))
        G[:n - 1, :n - 1] = self.G
        for k in range(n):
            G[n - 1, k] = G[k, n - 1] = float(self.e[k] @ e)
        self.G = G
        A = np.zeros((n + 1, n + 1))
        A[:n, :n] = G
        A[n, :n] = -1.0
        A[:n, n] = -1.0
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        w = np.linalg.solve(A, rhs)[:n]
        self.iteration += 1
        return sum(wk * xk for wk, xk in zip(w, self.x))


def quasi_newton_step(amplitudes, residual, eps, n_o, n_t1):
    """calculate_update_quasi_newton_updater + update (scale amplitudes and residual).

    Returns (new_amplitudes, preconditioned_scaled_residual).
    """
    dt = -residual / orbital_differences(eps, n_o)
    r = scale_doubles_diagonal(dt, n_t1)
    return amplitudes + r, r


def solve_ground_state(L_mo, h_mo, eps, n_o, residual_threshold=1e-10, max_iterations=100,
                       diis_dimension=8, verbose=False):
    """run_diis_cc_gs restated: Fock -> energy -> Omega -> quasi-Newton -> DIIS -> T1 re-dress."""
    n_v = L_mo.shape[1] - n_o
    n_t1 = n_o * n_v
    t1 = np.zeros((n_v, n_o))
    L_t1 = t1_cholesky(L_mo, t1, n_o)
    t2 = mp2_guess(L_t1, eps, n_o)
    diis = DIIS(diis_dimension)
    history = []
    for it in range(1, max_iterations + 1):
        F = fock_t1(h_mo, L_t1, t1, n_o)
        e = correlation_energy(L_t1, t1, t2, n_o)
        om = omega(L_mo, L_t1, F, t1, t2, n_o)
        norm = float(np.linalg.norm(om))
        history.append((e, norm))
        if verbose:
            print(it, e, norm)
        if norm < residual_threshold:
            return t1, t2, e, history
        amps = np.concatenate([t1.ravel(order='F'), t2])
        amps, r = quasi_newton_step(amps, om, eps, n_o, n_t1)
        amps = diis.update(r, amps)
        t1 = amps[:n_t1].reshape(n_v, n_o, order='F')
        t2 = amps[n_t1:]
        L_t1 = t1_cholesky(L_mo, t1, n_o)
    raise RuntimeError('Did not converge in the max number of iterations.')


# ----------------------------------------------------------------------------
# Jacobian transformations (jacobian_ccsd.F90:74-178, jacobian_transpose_ccsd.F90:129-236)
#
#   A_mu,nu = < mu~ | exp(-T) [H, tau_nu] exp(T) | R >,  tau_aibj = E_ai E_bj,
#
# so A = J D2 with J = d Omega / d t in eT's packed amplitude basis (Omega with
# its 1/2 doubles diagonal, t2 packed with T2 = 1/2 sum t_aibj E_ai E_bj) and
# D2 = diag(1, ..., 2 on the doubles diagonal) (eT scales c_aiai by two,
# jacobian_ccsd.F90:123).  The right transform is rho = A c, the left sigma^T = b^T A.
# Omega is a polynomial in t, so the complex step i h d/dt is exact to rounding.
# ----------------------------------------------------------------------------

def residual(L_mo, h_mo, t1, t2, n_o):
    """Omega(t1, t2) including the T1 dressing and Fock matrix of the current t1."""
    Lt = t1_cholesky(L_mo, t1, n_o)
    F = fock_t1(h_mo, Lt, t1, n_o)
    return omega(L_mo, Lt, F, t1, t2, n_o)


def jacobian_transform(L_mo, h_mo, t1, t2, n_o, c, step=1e-30):
    """rho = A c by complex-step differentiation of the residual along D2 c."""
    n_v = L_mo.shape[1] - n_o
    n_t1 = n_o * n_v
    c = np.asarray(c, dtype=float).copy()
    c[doubles_diagonal(n_t1)] *= 2.0
    t1c = t1 + 1j * step * c[:n_t1].reshape(n_v, n_o, order='F')
    t2c = t2 + 1j * step * c[n_t1:]
    return residual(L_mo, h_mo, t1c, t2c, n_o).imag / step


# ----------------------------------------------------------------------------
# Multiplier (Lambda) equations: t-bar^T A + eta = 0
# (multiplier_equation_ccs.F90:58-94, multiplier_equation_ccsd.F90:41-110)
# ----------------------------------------------------------------------------

def construct_eta(L_mo, h_mo, t1, n_o):
    """construct_eta_ccsd (multiplier_equation_ccsd.F90:56-110).

    eta_ai = 2 F_ia (T1 Fock), eta_aibj = 4 g_iajb - 2 g_ibja packed; g_iajb from
    the ov Cholesky block, which the T1 transformation leaves unchanged.
    """
    n_v = L_mo.shape[1] - n_o
    Lt = t1_cholesky(L_mo, t1, n_o)
    F = fock_t1(h_mo, Lt, t1, n_o)
    eta1 = 2.0 * F[:n_o, n_o:].T                       # (v, o): 2 F_ia
    Lov = L_mo[:, :n_o, n_o:]                          # (J, i, a)
    g = np.einsum('Jia,Jjb->iajb', Lov, Lov)
    e2 = 4.0 * g - 2.0 * g.transpose(0, 3, 2, 1)      # (i, a, j, b)
    eta2 = pack_aibj(e2.transpose(1, 0, 3, 2))         # (a, i, b, j) packed
    return np.concatenate([eta1.ravel(order='F'), eta2])


def initial_multipliers(t1, t2, n_o):
    """set_initial_multipliers_guess_ccsd (ccsd_class.F90:423-475): t-bar1 = t1,
    t-bar2 = 2 t_aibj - t_ajbi (construct_packed_contravariant, reordering.F90:4542)."""
    n_v = t1.shape[0]
    T = unpack_t2(t2, n_v, n_o)                       # (a, i, b, j)
    return np.concatenate([t1.ravel(order='F'), pack_aibj(2.0 * T - T.transpose(0, 3, 2, 1))])


def multipliers_dense(L_mo, h_mo, t1, t2, n_o):
    """Solution of A^T t-bar = -eta with the full A (tiny systems only)."""
    A = jacobian_matrix(L_mo, h_mo, t1, t2, n_o)
    return np.linalg.solve(A.T, -construct_eta(L_mo, h_mo, t1, n_o))


def one_electron_expectation(L_mo, h_mo, t1, t2, tbar, n_o, X, step=1e-30):
    """CC mean value of a one-electron operator X (MO basis) with the Lagrangian density.

    eT's ground-state density (construct_gs_density, mean_value_ccs.F90:94-122 and
    mean_value_doubles.F90) gives <X> = <HF|(1 + Lambda) e^-T X e^T|HF>, which is
    dL/de at e = 0 for h -> h + e X with L = E + tbar . Omega at fixed t and tbar:
      dE/de = 2 sum_i X_ii + 2 sum_ai X_ia t_ai   (E = E_ref + 2 F_ia t_ai + ..., F_ia(e) = e X_ia)
      tbar . dOmega/de by the complex step of the residual along X.
    """
    o = slice(0, n_o)
    v = slice(n_o, X.shape[0])
    dE = 2.0 * np.trace(X[o, o]) + 2.0 * np.sum(X[o, v] * t1.T)
    dOm = residual(L_mo, h_mo + 1j * step * X, t1, t2, n_o).imag / step
    return float(dE + np.dot(tbar, dOm))


def jacobian_matrix(L_mo, h_mo, t1, t2, n_o):
    """Full A (n_t1 + n_t2 square) column by column (tiny systems only)."""
    n_v = L_mo.shape[1] - n_o
    n = n_o * n_v + n_t2(n_o, n_v)
    A = np.empty((n, n))
    e = np.zeros(n)
    for k in range(n):
        e[k] = 1.0
        A[:, k] = jacobian_transform(L_mo, h_mo, t1, t2, n_o, e)
        e[k] = 0.0
    return A


def davidson_lowest(transform, diagonal, n_roots, tol=1e-10, max_iter=100):
    """Lowest eigenvalues of a non-symmetric operator by a Davidson iteration
    (solver_tools/davidson: trial vectors orthonormalised, reduced problem by
    a dense eigensolver, residuals preconditioned by the orbital differences)."""
    n = diagonal.size
    start = np.argsort(diagonal)[:n_roots]
    V = np.zeros((n, 0))
    AV = np.zeros((n, 0))
    new = []
    for k in start:
        e = np.zeros(n)
        e[k] = 1.0
        new.append(e)
    for it in range(max_iter):
        for x in new:
            for _ in range(2):
                x = x - V @ (V.T @ x)
            nx = np.linalg.norm(x)
            if nx < 1e-10:
                continue
            x = x / nx
            V = np.column_stack([V, x])
            AV = np.column_stack([AV, transform(x)])
        M = V.T @ AV
        w, S = np.linalg.eig(M)
        order = np.argsort(w.real)[:n_roots]
        w, S = w[order].real, S[:, order].real
        new = []
        done = True
        for r in range(n_roots):
            x = V @ S[:, r]
            res = AV @ S[:, r] - w[r] * x
            if np.linalg.norm(res) > tol:
                done = False
                new.append(res / (w[r] - diagonal + 1e-12))
        if done:
            return w
    raise RuntimeError('Davidson did not converge')
