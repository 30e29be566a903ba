"""Molecular producer of the CCSD inputs over libetgpu.so: AO integrals -> RHF
-> pivoted Cholesky -> MO Cholesky vectors, all on the device.

Replaces the producers eT runs before its CC solvers: the Libint integral
interface (src/libint/eri.cpp:33-89), the RHF solver
(src/solvers/hf/scf_solver_class.F90) and the Cholesky decomposition of the
AO integrals with the AO->MO transformation
(src/solvers/cholesky/eri_cd_class.F90:241, :3132).  ``Molecule.run`` returns
the arrays ``ccsd.CCSD`` takes (L_mo in eT's (J, p, q) layout, h_mo, orbital
energies) together with E_nuc, E_HF and the MO coefficients.

Also here: the cc-pVDZ basis (et_amd/basis/cc-pvdz.json, from the reference's
ao_basis/cc-pvdz.g94) and the seeded water lattices of the (H2O)_n systems.
"""
import ctypes
import json
import math
import os

import numpy as np

from . import ccsd as _ccsd
from . import restart as _restart

# eT's Bohr radius in Angstrom (CODATA 2014, src/various/parameters.F90:105)
BOHR_TO_ANGSTROM = 0.52917721067
Z_OF = {'H': 1, 'He': 2, 'Li': 3, 'Be': 4, 'B': 5, 'C': 6, 'N': 7, 'O': 8, 'F': 9, 'Ne': 10, 'Xe': 54}
_BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'basis')

# the water molecule of the (H2O)_n systems (Angstrom): O on the z axis, H in the yz plane
WATER = [('O', (0.0, 0.0, 0.1173)), ('H', (0.0, 0.7572, -0.4692)), ('H', (0.0, -0.7572, -0.4692))]


def load_basis(name='cc-pvdz'):
    with open(os.path.join(_BASIS_DIR, name + '.json')) as f:
        return json.load(f)['basis']


def _splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


class _Uniform:
    """Uniform doubles in [0, 1) from a splitmix64 stream of `seed` (53-bit mantissas)."""

    def __init__(self, seed):
        self.state = int(seed) & 0xFFFFFFFFFFFFFFFF

    def __call__(self):
        self.state, z = _splitmix64(self.state)
        return (z >> 11) * (1.0 / (1 << 53))


def random_rotation(u):
    """Uniformly distributed rotation from three uniforms (unit quaternion of
    K. Shoemake, "Uniform random rotations", Graphics Gems III, 1992)."""
    u1, u2, u3 = u(), u(), u()
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    x, y = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
    z, w = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_lattice(shape=(2, 2, 2), spacing=2.9, seed=0, min_distance=1.5, max_draws=1000):
    """(H2O)_n on a cubic lattice: each water (geometry WATER, rotated about its
    origin) is placed on a lattice point, x fastest, with a uniform random
    rotation drawn from the seed; a rotation that brings any atom closer than
    `min_distance` Angstrom to an atom of an earlier molecule is redrawn.
    Returns [(symbol, (x, y, z) Angstrom)]."""
    u = _Uniform(seed)
    atoms, placed = [], []
    for iz in range(shape[2]):
        for iy in range(shape[1]):
            for ix in range(shape[0]):
                origin = spacing * np.array([ix, iy, iz], dtype=float)
                for _ in range(max_draws):
                    R = random_rotation(u)
                    mol = [(s, origin + R @ np.array(x)) for s, x in WATER]
                    if all(np.linalg.norm(p - q) >= min_distance for _, p in mol for q in placed):
                        break
                else:
                    raise RuntimeError('water_lattice: no admissible rotation')
                placed.extend(p for _, p in mol)
                atoms.extend((s, tuple(p)) for s, p in mol)
    return atoms


# basis sets eT uses with Cartesian d functions unless told otherwise
# (set_cartesian_atomic_center, atomic_center_class.F90:706-756)
CARTESIAN_BY_DEFAULT = ('3-21g', '6-31g', '6-31+g', '6-31++g', '6-31g*', '6-31g**', '6-31+g*', '6-31+g**',
                        '6-31++g**', '6-31g(d,p)', '6-31g(2df,p)', '6-31g(3df,3pd)')


def cartesian_by_default(basis_name):
    """Whether eT builds this basis from Cartesian Gaussians when the input says neither
    `cartesian gaussians` nor `pure gaussians` (the Pople sets)."""
    return basis_name.lower() in CARTESIAN_BY_DEFAULT


def with_bath_orbital(r):
    """The CCSD inputs of Molecule.run with eT's bath orbital appended as the last virtual
    (make_bath_orbital_ccs, ccs_class.F90:1494-1582): zero MO coefficients, so zero
    Cholesky vectors and h, and orbital energy 0.  For ionisation energies."""
    L, h, eps, C = r['L_mo'], r['h_mo'], r['eps'], r['C']
    n = h.shape[0]
    out = dict(r)
    out['L_mo'] = np.zeros((L.shape[0], n + 1, n + 1), order='F')
    out['L_mo'][:, :n, :n] = L
    out['h_mo'] = np.zeros((n + 1, n + 1), order='F')
    out['h_mo'][:n, :n] = h
    out['eps'] = np.append(eps, 0.0)
    out['C'] = np.zeros((C.shape[0], n + 1), order='F')
    out['C'][:, :n] = C
    return out


def frozen_core_orbitals(Z):
    """eT's frozen-core count per atom (get_frozen_cores_and_n_frozen_orbitals,
    ao_tool_class.F90:2425-2507): 1 for 5 <= Z <= 12, 5 for 13..30, 9 for 31..38, 14 for
    39..48, 18 for 49..70, 23 for 71..80, 34 for 81..103, 50 for 104..105."""
    for lo, hi, k in ((1, 4, 0), (5, 12, 1), (13, 30, 5), (31, 38, 9), (39, 48, 14), (49, 70, 18), (71, 80, 23),
                      (81, 103, 34), (104, 105, 50)):
        if lo <= Z <= hi:
            return k
    raise _ccsd.ETGError('No support for frozen core for Z > 105.')


class Molecule:
    """A closed-shell molecule in a Gaussian basis on the device (libetgpu.so)."""

    def __init__(self, atoms, basis=None, bohr_to_angstrom=BOHR_TO_ANGSTROM, charge=0, cartesian=False):
        """cartesian: eT's `cartesian gaussians` (six Cartesian d functions per d shell)."""
        basis = load_basis() if basis is None else basis
        self._lib = _ccsd.load_library()
        Z, xyz, sh_atom, sh_l, sh_np, exps, coefs = [], [], [], [], [], [], []
        for ia, (sym, x) in enumerate(atoms):
            Z.append(float(Z_OF[sym]))
            xyz.extend(float(c) / bohr_to_angstrom for c in x)
            for sh in basis[sym]:
                sh_atom.append(ia)
                sh_l.append(int(sh['l']))
                sh_np.append(len(sh['exps']))
                exps.extend(sh['exps'])
                coefs.extend(sh['coefs'])
        self.charges = np.array(Z)
        self._shell_atom, self._shell_l, self._cartesian = list(sh_atom), list(sh_l), bool(cartesian)
        self.xyz_bohr = np.array(xyz).reshape(-1, 3)
        n_el = int(round(sum(Z))) - int(charge)
        arr = lambda v, t: np.ascontiguousarray(v, dtype=t)
        self._keep = [arr(Z, np.float64), arr(xyz, np.float64), arr(sh_atom, np.int64), arr(sh_l, np.int64),
                      arr(sh_np, np.int64), arr(exps, np.float64), arr(coefs, np.float64)]
        i64p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        h = ctypes.c_void_p()
        k = self._keep
        _ccsd._check(self._lib.etg_molecule_create(len(atoms), _ccsd._ptr(k[0]), _ccsd._ptr(k[1]), len(sh_l),
                                                   i64p(k[2]), i64p(k[3]), i64p(k[4]), _ccsd._ptr(k[5]),
                                                   _ccsd._ptr(k[6]), n_el, ctypes.byref(h)))
        self._h = h
        if cartesian:
            _ccsd._check(self._lib.etg_molecule_set_cartesian(self._h, 1))

    def close(self):
        if getattr(self, '_h', None) is not None and self._h.value:
            self._lib.etg_molecule_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def dims(self):
        v = [ctypes.c_int64() for _ in range(4)]
        _ccsd._check(self._lib.etg_molecule_dims(self._h, *[ctypes.byref(x) for x in v]))
        return dict(n_ao=v[0].value, n_o=v[1].value, n_J=v[2].value, scf_iterations=v[3].value)

    def one_electron(self):
        n = self.dims()['n_ao']
        S, T, V = (np.empty((n, n), order='F') for _ in range(3))
        _ccsd._check(self._lib.etg_molecule_one_electron(self._h, _ccsd._ptr(S), _ccsd._ptr(T), _ccsd._ptr(V)))
        return S, T, V

    def dipole_integrals(self, origin=(0.0, 0.0, 0.0)):
        """<mu| r_c - origin_c |nu> (c = x, y, z) in the AO basis, shape (3, n, n)."""
        n = self.dims()['n_ao']
        D = np.empty((n, n, 3), order='F')
        o = np.asarray(origin, dtype=np.float64)
        _ccsd._check(self._lib.etg_molecule_dipole_integrals(self._h, _ccsd._ptr(o), _ccsd._ptr(D)))
        return np.moveaxis(D, 2, 0)

    def ao_eri(self):
        """(pq|rs) as an n^2 x n^2 matrix (small molecules)."""
        n = self.dims()['n_ao']
        G = np.empty((n * n, n * n), order='F')
        _ccsd._check(self._lib.etg_molecule_ao_eri(self._h, _ccsd._ptr(G)))
        return G

    def set_orbitals(self, C, eps):
        """Restart orbitals for run(): C (n_ao, n_mo) and their energies are used as they
        are, the RHF iterations skipped (eT's restarted SCF with `skip`,
        hf_class.F90:1525-1537); run() then checks they are converged."""
        C = np.asfortranarray(C, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        if C.ndim != 2 or eps.shape != (C.shape[1],):
            raise _ccsd.ETGError('orbitals: C must be (n_ao, n_mo) with n_mo orbital energies')
        _ccsd._check(self._lib.etg_molecule_set_orbitals(self._h, C.shape[0], C.shape[1], _ccsd._ptr(C),
                                                         _ccsd._ptr(eps)))

    def read_orbitals(self, path):
        """set_orbitals from an eT orbital_coefficients file (file_handling_hf.F90:55-81)."""
        n = self.dims()['n_ao']
        C, eps = _restart.OrbitalFile(path).read(n_ao=n, n_mo=n)
        self.set_orbitals(C, eps)

    def run(self, cholesky_threshold=1e-8, scf_threshold=1e-10, max_scf_iterations=200, frozen_core=False):
        """Integrals, decomposition, RHF, MO vectors; returns the CCSD inputs.

        frozen_core: eT's `frozen orbitals: core` (remove_core_orbitals_hf,
        frozen_orbital_hf.F90:81-205): the lowest MOs (frozen_core_orbitals per atom) leave
        the correlated space and their Coulomb and exchange go into h_mo.  The returned
        n_o, eps, h_mo, L_mo and C then cover the active MOs only ('n_frozen_core' says how
        many were removed, 'C_frozen' holds them; E_HF is that of all electrons).
        """
        n_fc = int(sum(frozen_core_orbitals(int(round(z))) for z in self.charges)) if frozen_core else 0
        _ccsd._check(self._lib.etg_molecule_set_frozen_core(self._h, n_fc))
        _ccsd._check(self._lib.etg_molecule_run(self._h, cholesky_threshold, scf_threshold, max_scf_iterations))
        d = self.dims()
        n, nJ = d['n_ao'], d['n_J']
        e_nuc, e_hf = ctypes.c_double(), ctypes.c_double()
        eps = np.empty(n)
        h_mo = np.empty((n, n), order='F')
        L_mo = np.empty((nJ, n, n), order='F')
        C = np.empty((n, n), order='F')
        _ccsd._check(self._lib.etg_molecule_results(self._h, ctypes.byref(e_nuc), ctypes.byref(e_hf), _ccsd._ptr(eps),
                                                    _ccsd._ptr(h_mo), _ccsd._ptr(L_mo), _ccsd._ptr(C)))
        C_frozen = np.asfortranarray(C[:, :n_fc])
        if n_fc:
            self._check_frozen_core(C, n_fc)
            a = slice(n_fc, n)
            eps, h_mo, C = eps[a].copy(), np.asfortranarray(h_mo[a, a]), np.asfortranarray(C[:, a])
            L_mo = np.asfortranarray(L_mo[:, a, a])
        return dict(n_o=d['n_o'] - n_fc, n_J=nJ, e_nuc=e_nuc.value, e_hf=e_hf.value, eps=eps, h_mo=h_mo, L_mo=L_mo,
                    C=C, scf_iterations=d['scf_iterations'], n_frozen_core=n_fc, C_frozen=C_frozen)

    def _check_frozen_core(self, C, n_fc):
        """eT's crossover check (frozen_orbital_hf.F90:150-165): the largest AO weight of each
        frozen MO must sit on an atom that is frozen."""
        n = C.shape[0]
        ao_atom = []
        for size in (lambda l: (l + 1) * (l + 2) // 2 if l > 1 and self._cartesian else 2 * l + 1,
                     lambda l: (l + 1) * (l + 2) // 2):
            ao_atom = [a for a, l in zip(self._shell_atom, self._shell_l) for _ in range(size(l))]
            if len(ao_atom) == n:
                break
        frozen = [frozen_core_orbitals(int(round(z))) > 0 for z in self.charges]
        for k in range(n_fc):
            if not frozen[ao_atom[int(np.argmax(np.abs(C[:, k])))]]:
                raise _ccsd.ETGError('Frozen MO has large contribution from a non-frozen atom.')
