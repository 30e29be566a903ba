"""Python mirror of eT's CCSD wavefunction interface over libetgpu.so.

The method names and argument meanings follow the type-bound procedures of
eT's ``ccsd`` class that the ground-state solvers call
(src/wavefunctions/ccsd/ccsd_class.F90:101-160, 232-292); errors raise
``ETGError`` where eT would stop with ``output%error_msg``.  Arrays cross the
boundary in eT's Fortran layouts (see include/etgpu.h).  There is no CPU
fallback: constructing a wavefunction without the built library raises.
"""
import ctypes
import os

import numpy as np

from . import restart as _restart

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libetgpu.so')

FP64 = 0
FP32 = 1
BF16X3 = 2
BF16X3_ANY_SHAPE = 3    # etg_test_gemm / etg_bench_gemm: the split-bf16 kernel for every shape
PRECISIONS = {'fp64': FP64, 'fp32': FP32, 'bf16x3': BF16X3}


class ETGError(RuntimeError):
    pass


_lib = None

_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64

# name -> argtypes (restype is int status unless listed in _RESTYPE)
_SIGNATURES = {
    'etg_create': [_i64, _i64, _i64, ctypes.POINTER(ctypes.c_void_p)],
    'etg_destroy': [ctypes.c_void_p],
    'etg_last_error': [],
    'etg_set_cholesky': [ctypes.c_void_p, _c_dbl_p],
    'etg_set_oei': [ctypes.c_void_p, _c_dbl_p],
    'etg_set_orbital_energies': [ctypes.c_void_p, _c_dbl_p],
    'etg_set_amplitudes': [ctypes.c_void_p, _c_dbl_p],
    'etg_get_amplitudes': [ctypes.c_void_p, _c_dbl_p],
    'etg_set_t2_to_mp2_guess': [ctypes.c_void_p],
    'etg_construct_t1_cholesky': [ctypes.c_void_p],
    'etg_construct_fock': [ctypes.c_void_p, _c_dbl_p],
    'etg_calculate_energy': [ctypes.c_void_p, _c_dbl_p],
    'etg_construct_omega': [ctypes.c_void_p, _c_dbl_p, ctypes.c_int],
    'etg_get_omega': [ctypes.c_void_p, _c_dbl_p],
    'etg_omega_norm': [ctypes.c_void_p, _c_dbl_p],
    'etg_get_orbital_differences': [ctypes.c_void_p, _c_dbl_p, _i64],
    'etg_set_ladder_batch': [ctypes.c_void_p, _i64],
    'etg_synchronize': [ctypes.c_void_p],
    'etg_last_omega_flops': [ctypes.c_void_p, _c_dbl_p],
    'etg_prepare_for_jacobian': [ctypes.c_void_p],
    'etg_jacobian_transformation': [ctypes.c_void_p, _c_dbl_p, _c_dbl_p],
    'etg_jacobian_transpose_transformation': [ctypes.c_void_p, _c_dbl_p, _c_dbl_p],
    'etg_construct_omega_resident': [ctypes.c_void_p, ctypes.c_int],
    'etg_jacobian_transformations': [ctypes.c_void_p, ctypes.c_int, _i64, _c_dbl_p, _c_dbl_p,
                                     ctypes.c_int],
    'etg_davidson': [ctypes.c_void_p, ctypes.c_int, _i64, ctypes.c_double, _i64, _i64, ctypes.c_int,
                     _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.POINTER(_i64)],
    'etg_linear_davidson': [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, ctypes.c_double, _i64, _i64,
                            ctypes.c_int, _c_dbl_p, _c_dbl_p, ctypes.POINTER(_i64)],
    'etg_set_cvs': [ctypes.c_void_p, _i64, ctypes.POINTER(_i64)],
    'etg_set_remove_core': [ctypes.c_void_p, _i64, ctypes.POINTER(_i64)],
    'etg_set_ionization': [ctypes.c_void_p, ctypes.c_int],
    'etg_newton_raphson_update': [ctypes.c_void_p, ctypes.c_double, _i64, ctypes.c_int, _c_dbl_p,
                                  ctypes.POINTER(_i64)],
    'etg_diis_excited_states': [ctypes.c_void_p, ctypes.c_int, _i64, ctypes.c_double, ctypes.c_double, _i64,
                                _i64, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                ctypes.POINTER(_i64)],
    'etg_test_reduced_eigen': [_i64, _c_dbl_p, _i64, _c_dbl_p, _c_dbl_p, _c_dbl_p],
    'etg_reduced_eigen_noalloc': [_i64, ctypes.c_void_p, _i64, _i64, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, _i64],
    'etg_set_davidson_start_precision': [ctypes.c_void_p, ctypes.c_int],
    'etg_set_davidson_space': [ctypes.c_void_p, ctypes.c_int],
    'etg_diis_finalize': [ctypes.c_void_p],
    'etg_davidson_timing': [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, ctypes.POINTER(_i64)],
    'etg_diis_initialize': [ctypes.c_void_p, _i64],
    'etg_quasi_newton_update': [ctypes.c_void_p],
    'etg_diis_update': [ctypes.c_void_p],
    'etg_diis_initialize_crop': [ctypes.c_void_p, _i64],
    'etg_prepare_for_multiplier_equation': [ctypes.c_void_p],
    'etg_set_initial_multipliers_guess': [ctypes.c_void_p],
    'etg_set_multipliers': [ctypes.c_void_p, _c_dbl_p],
    'etg_get_multipliers': [ctypes.c_void_p, _c_dbl_p],
    'etg_construct_eta': [ctypes.c_void_p, _c_dbl_p],
    'etg_construct_multiplier_equation': [ctypes.c_void_p, _c_dbl_p, ctypes.c_int],
    'etg_multiplier_residual_norm': [ctypes.c_void_p, _c_dbl_p],
    'etg_multiplier_quasi_newton_update': [ctypes.c_void_p],
    'etg_multiplier_diis_update': [ctypes.c_void_p],
    'etg_set_transform_graphs': [ctypes.c_void_p, ctypes.c_int],
    'etg_one_electron_expectation': [ctypes.c_void_p, _c_dbl_p, _c_dbl_p],
    'etg_one_electron_density': [ctypes.c_void_p, _c_dbl_p],
    'etg_molecule_dipole_integrals': [ctypes.c_void_p, _c_dbl_p, _c_dbl_p],
    'etg_set_transform_batch': [ctypes.c_void_p, _i64],
    'etg_set_kernel_ledger': [ctypes.c_void_p, ctypes.c_int],
    'etg_jacobian_memory': [ctypes.c_void_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    'etg_kernel_ledger': [ctypes.c_void_p, _i64, ctypes.c_char_p, _i64, ctypes.POINTER(_i64),
                          _c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.POINTER(_i64)],
    'etg_test_gemm': [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_double, _c_dbl_p,
                      _i64, _c_dbl_p, _i64, ctypes.c_double, _c_dbl_p, _i64, ctypes.c_int],
    'etg_set_gemm_autotune': [ctypes.c_int, ctypes.POINTER(_i64)],
    'etg_molecule_create': [_i64, _c_dbl_p, _c_dbl_p, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                            ctypes.POINTER(_i64), _c_dbl_p, _c_dbl_p, _i64, ctypes.POINTER(ctypes.c_void_p)],
    'etg_molecule_run': [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _i64],
    'etg_molecule_set_orbitals': [ctypes.c_void_p, _i64, _i64, _c_dbl_p, _c_dbl_p],
    'etg_molecule_set_frozen_core': [ctypes.c_void_p, _i64],
    'etg_molecule_set_cartesian': [ctypes.c_void_p, ctypes.c_int],
    'etg_molecule_dims': [ctypes.c_void_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                          ctypes.POINTER(_i64)],
    'etg_molecule_results': [ctypes.c_void_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p],
    'etg_molecule_one_electron': [ctypes.c_void_p, _c_dbl_p, _c_dbl_p, _c_dbl_p],
    'etg_molecule_ao_eri': [ctypes.c_void_p, _c_dbl_p],
    'etg_molecule_destroy': [ctypes.c_void_p],
    'etg_bench_gemm': [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int,
                       _c_dbl_p],
}
_RESTYPE = {'etg_last_error': ctypes.c_char_p}


def load_library(path=LIB_PATH):
    """Loads libetgpu.so (raises ETGError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ETGError('libetgpu.so not found at %s: run `python -m et_amd.build`' % path)
    lib = ctypes.CDLL(path)
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, ctypes.c_int)
    _lib = lib
    return lib


def _check(status):
    if status != 0:
        raise ETGError(_lib.etg_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_c_dbl_p)


def device_gemm(ta, tb, alpha, A, B, beta, C, precision=FP64):
    """C <- alpha op(A) op(B) + beta C on the matrix-core kernel (diagnostic entry)."""
    lib = load_library()
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    C = np.array(C, dtype=np.float64, order='F', copy=True)
    m, n = C.shape
    k = A.shape[0] if ta == 'T' else A.shape[1]
    _check(lib.etg_test_gemm(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), max(1, A.shape[0]),
                             _ptr(B), max(1, B.shape[0]), beta, _ptr(C), max(1, m), precision))
    return C


def reduced_eigen(G, n_want):
    """Lowest n_want eigenpairs of a small real matrix with the device reduced-space solver."""
    lib = load_library()
    G = np.asfortranarray(G, dtype=np.float64)
    k = G.shape[0]
    wr, wi = np.empty(k), np.empty(k)
    X = np.empty((k, max(n_want, 1)), order='F')
    _check(lib.etg_test_reduced_eigen(k, _ptr(G), n_want, _ptr(wr), _ptr(wi), _ptr(X)))
    return wr, wi, X[:, :n_want]


def set_gemm_autotune(on):
    """Measured tile choice of the FP64/FP32 products on or off (process-wide);
    returns the number of product shapes measured so far."""
    n = _i64()
    _check(load_library().etg_set_gemm_autotune(1 if on else 0, ctypes.byref(n)))
    return n.value


def bench_gemm(m, n, k, precision=FP64, reps=5, ta='N', tb='N'):
    """Seconds per m x n x k product on device-resident operands."""
    lib = load_library()
    out = ctypes.c_double()
    _check(lib.etg_bench_gemm(ta.encode(), tb.encode(), m, n, k, precision, reps, ctypes.byref(out)))
    return out.value


class CCSD:
    """Device-resident CCSD wavefunction (eT ``ccsd`` type-bound procedures).

    Sizes: n_t1 = n_o n_v, n_t2 = n_t1 (n_t1 + 1) / 2.
    """

    def __init__(self, n_o, n_v, n_J):
        self._lib = load_library()
        self.n_o, self.n_v, self.n_J = int(n_o), int(n_v), int(n_J)
        self.n_mo = self.n_o + self.n_v
        self.n_t1 = self.n_o * self.n_v
        self.n_t2 = self.n_t1 * (self.n_t1 + 1) // 2
        self.n_gs_amplitudes = self.n_t1 + self.n_t2
        h = ctypes.c_void_p()
        _check(self._lib.etg_create(self.n_o, self.n_v, self.n_J, ctypes.byref(h)))
        self._h = h
        self.hf_energy = 0.0
        self.correlation_energy = 0.0
        self.energy = 0.0
        self.davidson_stalled = False
        self.files = None

    def close(self):
        if getattr(self, '_h', None) is not None and self._h.value:
            _check(self._lib.etg_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- inputs ---------------------------------------------------------
    def set_cholesky(self, L):
        L = np.asfortranarray(L, dtype=np.float64)
        if L.shape != (self.n_J, self.n_mo, self.n_mo):
            raise ETGError('L must have shape (n_J, n_mo, n_mo)')
        _check(self._lib.etg_set_cholesky(self._h, _ptr(L)))

    def set_oei(self, h):
        h = np.asfortranarray(h, dtype=np.float64)
        if h.shape != (self.n_mo, self.n_mo):
            raise ETGError('h must have shape (n_mo, n_mo)')
        _check(self._lib.etg_set_oei(self._h, _ptr(h)))

    def set_orbital_energies(self, eps):
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        if eps.shape != (self.n_mo,):
            raise ETGError('orbital energies must have length n_mo')
        _check(self._lib.etg_set_orbital_energies(self._h, _ptr(eps)))

    def set_ladder_batch(self, batch):
        _check(self._lib.etg_set_ladder_batch(self._h, int(batch)))

    # ---- amplitudes -----------------------------------------------------
    def set_amplitudes(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        if t.shape != (self.n_gs_amplitudes,):
            raise ETGError('amplitudes must have length n_t1 + n_t2')
        _check(self._lib.etg_set_amplitudes(self._h, _ptr(t)))

    def get_amplitudes(self):
        t = np.empty(self.n_gs_amplitudes)
        _check(self._lib.etg_get_amplitudes(self._h, _ptr(t)))
        return t

    def set_t2_to_cc2_guess(self):
        _check(self._lib.etg_set_t2_to_mp2_guess(self._h))

    def set_initial_amplitudes_guess(self, restart=False):
        """set_initial_amplitudes_guess_ccsd (ccsd_class.F90:357-420).

        Without restart (or with no t file): t1 = 0, L_t1 = L_mo, t2 = MP2 guess.  With
        restart, the t file of initialize_files' directory: a whole vector is taken as
        is; a singles-only one (a CCS / CC2 file) is dropped for the no-restart guess,
        as eT does; any other length is an error.
        """
        if restart and self.files is not None and self.files.t.exists():
            n_read = self.read_amplitudes()
            if n_read == self.n_gs_amplitudes:
                self.construct_t1_cholesky()
                return
            if n_read != self.n_t1:
                raise ETGError('Did not recognize number of t-amplitudes on file expected %d or %d '
                               'found %d amplitudes.' % (self.n_gs_amplitudes, self.n_t1, n_read))
        t = np.zeros(self.n_gs_amplitudes)
        self.set_amplitudes(t)
        self.construct_t1_cholesky()
        self.set_t2_to_cc2_guess()

    # ---- restart files (file_handling_ccs.F90) ------------------------------
    def initialize_files(self, directory):
        """Restart files of this wavefunction in `directory` (file_handling_ccs.F90:36-99):
        t, tbar, r_00k, l_00k in eT's amplitude file format."""
        self.files = _restart.CCSDFiles(directory, self.n_o, self.n_v)
        return self.files

    def _require_files(self):
        if self.files is None:
            raise ETGError('restart files not initialised (initialize_files)')
        return self.files

    def save_amplitudes(self):
        """save_amplitudes_ccs (file_handling_ccs.F90:147-167): t and the energy."""
        f = self._require_files()
        f.make_directory()
        f.t.save(self.get_amplitudes(), self.energy)

    def read_amplitudes(self):
        """read_amplitudes_ccs (file_handling_ccs.F90:170-195); returns the number read."""
        v, _, n_read = self._require_files().t.read()
        self.set_amplitudes(v)
        return n_read

    def save_multipliers(self):
        """save_multipliers_ccs (file_handling_ccs.F90:198-218): tbar and the energy."""
        f = self._require_files()
        f.make_directory()
        f.tbar.save(self.get_multipliers(), self.energy)

    def read_multipliers(self):
        """read_multipliers_ccs (file_handling_ccs.F90:221-246); returns the number read."""
        v, _, n_read = self._require_files().tbar.read()
        self.set_multipliers(v)
        return n_read

    def save_excited_state(self, X, first, last, side, energies):
        """save_excited_state_ccs (file_handling_ccs.F90:249-286): columns of X are the
        states first..last (1-based) of `side`, each file with its excitation energy."""
        f = self._require_files()
        f.make_directory()
        X = np.asarray(X, dtype=np.float64).reshape(self.n_gs_amplitudes, -1)
        if X.shape[1] != last - first + 1 or len(energies) != X.shape[1]:
            raise ETGError('save_excited_state: states %d..%d need as many vectors and energies'
                           % (first, last))
        for c, k in enumerate(range(first, last + 1)):
            f.side(side, k).save(X[:, c], energies[c])

    def read_excited_state(self, first, last, side):
        """read_excited_state_ccs (file_handling_ccs.F90:329-373): (X, energies)."""
        f = self._require_files()
        X = np.empty((self.n_gs_amplitudes, last - first + 1), order='F')
        e = np.empty(last - first + 1)
        for c, k in enumerate(range(first, last + 1)):
            X[:, c], e[c], _ = f.side(side, k).read()
        return X, e

    def restart_start_vectors(self, n_states, side='right'):
        """Davidson start vectors on restart (es_start_vector_tool_class.F90:60-111).

        State k comes from its own file, or else from the other side's through
        get_restart_vector_doubles, normalised; a state with neither file gets the unit
        vector on its orbital difference (the k-th lowest).  Returns (start, energies,
        found) with start (n_t1 + n_t2) x n_states.
        """
        f = self._require_files()
        S = np.zeros((self.n_gs_amplitudes, n_states), order='F')
        e = np.zeros(n_states)
        found = np.zeros(n_states, dtype=bool)
        lowest = None
        for k in range(n_states):
            got = f.check_and_get_restart_vector(k + 1, side)
            if got is not None:
                v, e[k] = got
                S[:, k] = v / np.linalg.norm(v)
                found[k] = True
            else:
                if lowest is None:
                    lowest = np.argsort(self.get_orbital_differences(), kind='stable')
                S[lowest[k], k] = 1.0
        return S, e, found

    # ---- residual path --------------------------------------------------
    def construct_t1_cholesky(self):
        _check(self._lib.etg_construct_t1_cholesky(self._h))

    def construct_fock(self, task=None):
        """All four T1-Fock blocks (CCSD ignores task, fock_ccsd.F90:34-79); returns F."""
        F = np.empty((self.n_mo, self.n_mo), order='F')
        _check(self._lib.etg_construct_fock(self._h, _ptr(F)))
        return F

    def calculate_energy(self):
        e = ctypes.c_double()
        _check(self._lib.etg_calculate_energy(self._h, ctypes.byref(e)))
        self.correlation_energy = e.value
        self.energy = self.hf_energy + e.value
        return e.value

    def construct_omega(self, precision=FP64, copy_out=True):
        if copy_out:
            om = np.empty(self.n_gs_amplitudes)
            _check(self._lib.etg_construct_omega(self._h, _ptr(om), int(precision)))
            return om
        _check(self._lib.etg_construct_omega(self._h, None, int(precision)))
        return None

    def get_omega(self):
        om = np.empty(self.n_gs_amplitudes)
        _check(self._lib.etg_get_omega(self._h, _ptr(om)))
        return om

    def omega_norm(self):
        x = ctypes.c_double()
        _check(self._lib.etg_omega_norm(self._h, ctypes.byref(x)))
        return x.value

    def get_orbital_differences(self, n=None):
        n = self.n_gs_amplitudes if n is None else int(n)
        x = np.empty(n)
        _check(self._lib.etg_get_orbital_differences(self._h, _ptr(x), n))
        return x

    def synchronize(self):
        _check(self._lib.etg_synchronize(self._h))

    # ---- Jacobian transformations -------------------------------------------
    def prepare_for_jacobian(self, side='right'):
        """prepare_for_jacobian(_transpose) (jacobian_ccsd.F90:46-71): records the residual program."""
        _check(self._lib.etg_prepare_for_jacobian(self._h))

    def jacobian_transformation(self, c):
        """rho = A c (jacobian_ccsd.F90:74-178)."""
        c = np.ascontiguousarray(c, dtype=np.float64)
        if c.shape != (self.n_gs_amplitudes,):
            raise ETGError('trial vector must have length n_t1 + n_t2')
        rho = np.empty(self.n_gs_amplitudes)
        _check(self._lib.etg_jacobian_transformation(self._h, _ptr(c), _ptr(rho)))
        return rho

    def jacobian_transpose_transformation(self, b):
        """sigma^T = b^T A (jacobian_transpose_ccsd.F90:129-236)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (self.n_gs_amplitudes,):
            raise ETGError('trial vector must have length n_t1 + n_t2')
        sigma = np.empty(self.n_gs_amplitudes)
        _check(self._lib.etg_jacobian_transpose_transformation(self._h, _ptr(b), _ptr(sigma)))
        return sigma

    def jacobian_transformations(self, X, side='right', precision=FP64):
        """Columns of X transformed by A (right) or A^T (left) in one call."""
        X = np.asfortranarray(X, dtype=np.float64)
        if X.ndim != 2 or X.shape[0] != self.n_gs_amplitudes:
            raise ETGError('trial vectors must be an (n_t1 + n_t2) x n_vec array')
        R = np.empty_like(X, order='F')
        _check(self._lib.etg_jacobian_transformations(self._h, 0 if side == 'right' else 1, X.shape[1],
                                                      _ptr(X), _ptr(R), int(precision)))
        return R

    def jacobian_transformations_at(self, x_addr, r_addr, n_vec, side='right', precision=FP64):
        """The same on caller-owned (n_t1 + n_t2) x n_vec column-major arrays at raw
        addresses, host or device (device-resident vectors skip the PCIe copies)."""
        _check(self._lib.etg_jacobian_transformations(self._h, 0 if side == 'right' else 1, n_vec,
                                                      ctypes.cast(x_addr, _c_dbl_p),
                                                      ctypes.cast(r_addr, _c_dbl_p), int(precision)))

    def davidson(self, n_states, side='right', residual_threshold=1e-6, max_iterations=100,
                 max_dim=0, precision=FP64, vectors=False, start=None):
        """Device Davidson (eigen_davidson_solver_class.F90:122-318): lowest roots of A or A^T.

        start: optional (n_t1 + n_t2) x n_states start vectors.
        Returns (energies, residual-norm history per iteration[, eigenvectors]).  With
        FP32/BF16X3 transformations the solve may end at their rounding floor before
        the threshold; davidson_stalled then reads True."""
        e = np.empty(n_states)
        hist = np.full((max_iterations, n_states), np.nan)
        X = np.empty((self.n_gs_amplitudes, n_states), order='F') if vectors else None
        S = None if start is None else np.asfortranarray(start, dtype=np.float64)
        if S is not None and S.shape != (self.n_gs_amplitudes, n_states):
            raise ETGError('start vectors must be (n_t1 + n_t2) x n_states')
        its = _i64()
        _check(self._lib.etg_davidson(self._h, 0 if side == 'right' else 1, n_states, residual_threshold,
                                      max_iterations, max_dim, int(precision),
                                      _ptr(S) if S is not None else None, _ptr(e),
                                      _ptr(X) if vectors else None, _ptr(hist), ctypes.byref(its)))
        self.davidson_stalled = its.value < 0
        hist = hist[:abs(its.value)]
        return (e, hist, X) if vectors else (e, hist)

    def set_davidson_start_precision(self, precision=None):
        """Precision of the start block's transformation in davidson() (None: the solve's)."""
        _check(self._lib.etg_set_davidson_start_precision(self._h, -1 if precision is None else int(precision)))

    def set_davidson_space(self, where='auto'):
        """Placement of davidson()'s V and A V: 'auto', 'hbm' or 'host' (pinned host memory)."""
        _check(self._lib.etg_set_davidson_space(self._h, {'auto': 0, 'hbm': 1, 'host': 2}[where]))

    def davidson_timing(self, reset=True):
        """(seconds in transformations, seconds in davidson() in total, vectors transformed)
        summed over the solves since the last reset."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), _i64()
        _check(self._lib.etg_davidson_timing(self._h, 1 if reset else 0, ctypes.byref(a), ctypes.byref(b),
                                             ctypes.byref(c)))
        return a.value, b.value, c.value

    def linear_davidson(self, rhs, side='right', start=None, residual_threshold=1e-6, max_iterations=100,
                        max_dim=50, precision=FP64):
        """Solves A X = rhs (right) or A^T X = rhs (left) with eT's linear Davidson solver on
        the device; returns (X, residual-norm history)."""
        F = np.ascontiguousarray(rhs, dtype=np.float64)
        if F.shape != (self.n_gs_amplitudes,):
            raise ETGError('rhs must have length n_t1 + n_t2')
        S = None if start is None else np.ascontiguousarray(start, dtype=np.float64)
        if S is not None and S.shape != F.shape:
            raise ETGError('start vector must have length n_t1 + n_t2')
        X = np.empty(self.n_gs_amplitudes)
        hist = np.full(max_iterations, np.nan)
        its = _i64()
        _check(self._lib.etg_linear_davidson(self._h, 0 if side == 'right' else 1, _ptr(F),
                                             _ptr(S) if S is not None else None, residual_threshold,
                                             max_iterations, max_dim, int(precision), _ptr(X), _ptr(hist),
                                             ctypes.byref(its)))
        return X, hist[:its.value]

    def diis_excited_states(self, n_states, side='right', residual_threshold=1e-6, energy_threshold=0.0,
                            max_iterations=100, diis_dimension=20, precision=FP64, start=None):
        """eT's DIIS excited-state solver on the device (diis_cc_es_class.F90:260-492).

        Returns (energies sorted ascending, residual-norm history per iteration,
        eigenvectors in the same order); the per-iteration energies (start order) are
        left in diis_energy_history."""
        e = np.empty(n_states)
        hist = np.full((max_iterations, n_states), np.nan)
        ehist = np.full((max_iterations, n_states), np.nan)
        X = np.empty((self.n_gs_amplitudes, n_states), order='F')
        S = None if start is None else np.asfortranarray(start, dtype=np.float64)
        if S is not None and S.shape != (self.n_gs_amplitudes, n_states):
            raise ETGError('start vectors must be (n_t1 + n_t2) x n_states')
        its = _i64()
        _check(self._lib.etg_diis_excited_states(self._h, 0 if side == 'right' else 1, n_states,
                                                 residual_threshold, energy_threshold, max_iterations,
                                                 diis_dimension, int(precision),
                                                 _ptr(S) if S is not None else None, _ptr(e), _ptr(X),
                                                 _ptr(hist), _ptr(ehist), ctypes.byref(its)))
        self.diis_energy_history = ehist[:its.value]
        return e, hist[:its.value], X

    def diis_initialize(self, dimension=8, crop=False):
        """HBM records for DIIS (diis_tool_class.F90:249-277); crop: eT's CROP variant
        (diis_tool_class.F90:763-805)."""
        if crop:
            _check(self._lib.etg_diis_initialize_crop(self._h, int(dimension)))
        else:
            _check(self._lib.etg_diis_initialize(self._h, int(dimension)))

    def diis_finalize(self):
        """Releases the DIIS records (diis%finalize, diis_tool_class.F90:280-299)."""
        _check(self._lib.etg_diis_finalize(self._h))

    def quasi_newton_update(self):
        """Device quasi-Newton step t += -omega/eps (quasi_newton_updater_class.F90:81-118)."""
        _check(self._lib.etg_quasi_newton_update(self._h))

    def set_cvs(self, core_mos=()):
        """Core-valence separation for the excited-state solvers: project every transformed
        vector onto excitations out of the occupied MOs core_mos (0-based); () switches it off."""
        core = np.ascontiguousarray(core_mos, dtype=np.int64)
        _check(self._lib.etg_set_cvs(self._h, len(core), core.ctypes.data_as(ctypes.POINTER(_i64))))

    def set_remove_core(self, core_mos=()):
        """eT's `remove core`: project every transformed vector of the excited-state solvers
        onto excitations with no occupied index among core_mos (0-based); () switches it off."""
        core = np.ascontiguousarray(core_mos, dtype=np.int64)
        _check(self._lib.etg_set_remove_core(self._h, len(core), core.ctypes.data_as(ctypes.POINTER(_i64))))

    def set_ionization(self, on=True):
        """eT's `ionization`: the excited-state solvers keep only excitations into the last
        virtual, the bath orbital (et_amd.molecule.with_bath_orbital appends it)."""
        _check(self._lib.etg_set_ionization(self._h, 1 if on else 0))

    def ip_start_vectors(self, n_states):
        """eT's ionisation start vectors (set_ip_start_indices_ccs, ccs_class.F90:1585-1608):
        unit singles (bath, i) for i from the highest occupied down."""
        n = self.n_t1 + self.n_t2
        X = np.zeros((n, n_states), order='F')
        for k in range(n_states):
            X[(self.n_o - 1 - k) * self.n_v + self.n_v - 1, k] = 1.0
        return X

    def cvs_start_vectors(self, n_states, core_mos):
        """eT's CVS start vectors (set_cvs_start_indices_ccs, ccs_class.F90:1393-1438): unit
        singles (a, i) for a = 0, 1, ... and, for each a, the core MOs in the order given."""
        n = self.n_t1 + self.n_t2
        X = np.zeros((n, n_states), order='F')
        for k in range(n_states):
            a, c = divmod(k, len(core_mos))
            X[a + self.n_v * int(core_mos[c]), k] = 1.0
        return X

    def newton_raphson_update(self, rel_threshold=1e-2, max_iterations=100, precision=FP64):
        """eT's Newton-Raphson step t += u with A u = -Omega (linear Davidson); the next
        diis_update records (Omega, t).  Returns the linear iterations' residual norms."""
        hist = np.full(max_iterations, np.nan)
        its = _i64()
        _check(self._lib.etg_newton_raphson_update(self._h, rel_threshold, max_iterations, int(precision),
                                                   _ptr(hist), ctypes.byref(its)))
        return hist[:its.value]

    def diis_update(self):
        """Device DIIS extrapolation of the amplitudes (diis_tool_class.F90:302-347)."""
        _check(self._lib.etg_diis_update(self._h))

    # ---- multiplier equations (multiplier_equation_ccs.F90, multiplier_equation_ccsd.F90) ----
    def prepare_for_multiplier_equation(self):
        _check(self._lib.etg_prepare_for_multiplier_equation(self._h))

    def set_initial_multipliers_guess(self, restart=False):
        """set_initial_multipliers_guess_ccsd (ccsd_class.F90:423-470).

        Without restart (or with no tbar file): tbar = (t1, 2 t_aibj - t_ajbi) on the
        device.  With restart, the tbar file: a whole vector is taken as is; a
        singles-only one keeps its singles and gets the doubles of the guess.
        """
        if restart and self.files is not None and self.files.tbar.exists():
            v, _, n_read = self.files.tbar.read()
            if n_read == self.n_gs_amplitudes:
                self.set_multipliers(v)
                return
            if n_read != self.n_t1:
                raise ETGError('Did not recognize number of multipliers on file expected %d or %d '
                               'found %d' % (self.n_gs_amplitudes, self.n_t1, n_read))
            _check(self._lib.etg_set_initial_multipliers_guess(self._h))
            guess = self.get_multipliers()
            guess[:self.n_t1] = v[:self.n_t1]
            self.set_multipliers(guess)
            return
        _check(self._lib.etg_set_initial_multipliers_guess(self._h))

    def set_multipliers(self, tbar):
        tbar = np.ascontiguousarray(tbar, dtype=np.float64)
        if tbar.shape != (self.n_gs_amplitudes,):
            raise ETGError('multipliers must have length n_t1 + n_t2')
        _check(self._lib.etg_set_multipliers(self._h, _ptr(tbar)))

    def get_multipliers(self):
        out = np.empty(self.n_gs_amplitudes)
        _check(self._lib.etg_get_multipliers(self._h, _ptr(out)))
        return out

    def construct_eta(self, copy_out=True):
        out = np.empty(self.n_gs_amplitudes) if copy_out else None
        _check(self._lib.etg_construct_eta(self._h, _ptr(out) if copy_out else None))
        return out

    def construct_multiplier_equation(self, copy_out=True, precision=FP64):
        """R = A^T tbar + eta, the left transformation in `precision`."""
        out = np.empty(self.n_gs_amplitudes) if copy_out else None
        _check(self._lib.etg_construct_multiplier_equation(self._h, _ptr(out) if copy_out else None,
                                                           int(precision)))
        return out

    def multiplier_residual_norm(self):
        x = ctypes.c_double()
        _check(self._lib.etg_multiplier_residual_norm(self._h, ctypes.byref(x)))
        return x.value

    def multiplier_quasi_newton_update(self):
        _check(self._lib.etg_multiplier_quasi_newton_update(self._h))

    def multiplier_diis_update(self):
        _check(self._lib.etg_multiplier_diis_update(self._h))

    def one_electron_expectation(self, X):
        """CC mean value of the one-electron operator X (n_mo x n_mo, MO basis) with the
        Lagrangian density (needs prepare_for_jacobian at the converged amplitudes and
        the multipliers)."""
        X = np.asfortranarray(X, dtype=np.float64)
        if X.shape != (self.n_mo, self.n_mo):
            raise ETGError('operator must be n_mo x n_mo')
        v = ctypes.c_double()
        _check(self._lib.etg_one_electron_expectation(self._h, _ptr(X), ctypes.byref(v)))
        return v.value

    def one_electron_density(self):
        """CC one-electron density D (n_mo x n_mo, MO basis): <X> = sum(D * X) for any
        one-electron X (one backward replay; same preconditions as
        one_electron_expectation)."""
        D = np.empty((self.n_mo, self.n_mo), order='F')
        _check(self._lib.etg_one_electron_density(self._h, _ptr(D)))
        return D

    def set_transform_graphs(self, on):
        """Replay captured HIP graphs of the transformation passes (default) or run them eagerly."""
        _check(self._lib.etg_set_transform_graphs(self._h, 1 if on else 0))

    def set_transform_batch(self, max_vectors):
        """At most max_vectors trial vectors per transformation pass (0: automatic)."""
        _check(self._lib.etg_set_transform_batch(self._h, int(max_vectors)))

    def jacobian_memory(self):
        """Bytes of the Jacobian record and per trial vector of a pass (after prepare_for_jacobian)."""
        v = [_i64(), _i64(), _i64()]
        _check(self._lib.etg_jacobian_memory(self._h, *[ctypes.byref(x) for x in v]))
        return dict(record=v[0].value, per_vector=v[1].value, per_vector_unshared=v[2].value)

    def set_kernel_ledger(self, on):
        _check(self._lib.etg_set_kernel_ledger(self._h, 1 if on else 0))

    def kernel_ledger(self):
        """[(kernel, launches, seconds, flops, algorithmic bytes)] per kernel instantiation."""
        count = _i64()
        _check(self._lib.etg_kernel_ledger(self._h, -1, None, 0, None, None, None, None,
                                           ctypes.byref(count)))
        out = []
        for i in range(count.value):
            name = ctypes.create_string_buffer(128)
            n, sec, fl, by = _i64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            _check(self._lib.etg_kernel_ledger(self._h, i, name, 128, ctypes.byref(n), ctypes.byref(sec),
                                               ctypes.byref(fl), ctypes.byref(by), None))
            out.append((name.value.decode(), n.value, sec.value, fl.value, by.value))
        return out

    def last_omega_flops(self):
        x = ctypes.c_double()
        _check(self._lib.etg_last_omega_flops(self._h, ctypes.byref(x)))
        return x.value
