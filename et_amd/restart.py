"""eT's restart files: CC vectors and Hartree-Fock orbitals, read and written on the host.

Formats (unformatted stream files in native byte order, no record markers):

* amplitude files, ``amplitude_file_storer`` (src/io/amplitude_file_storer_class.F90:99-190):
  the energy (float64), then for each block its length n (int64) and n float64.  CCSD
  ``t`` and ``tbar`` hold blocks [n_t1, n_t2] and the ground-state energy; ``r_00k`` and
  ``l_00k`` (file_handling_ccs.F90:56-99) hold the k-th right / left excited state with
  its excitation energy.  A file that ends at a block header is a partial vector (a CCS
  or CC2 file read by CCSD): reading stops there and reports how many elements it read.
* ``orbital_coefficients`` (src/wavefunctions/hf/file_handling_hf.F90:36-81): n_ao and
  n_mo (int64), the orbital energies (n_mo), then C (n_ao, n_mo) column-major.
* eT v1.0 files are Fortran sequential records (4-byte length markers around each
  record); ``convert_v1_0_directory`` rewrites a v1.0 restart directory in the formats above
  (the directory layout is described where the converter is defined).

Vectors are in eT's CC layout: singles (v, o) column-major, then the doubles packed over
the compound index ai = a + n_v i (element (ai, bj), ai <= bj, at bj (bj + 1) / 2 + ai).
The packed (co/contra)variant maps used when a right vector restarts a left one or the
reverse (get_restart_vector_doubles, src/wavefunctions/doubles/file_handling_doubles.F90:36-119)
are the permutation ai,bj -> bi,aj applied to that packed storage.
"""
import os

import numpy as np


class RestartError(RuntimeError):
    """A restart file that does not fit the calculation (eT stops with an error message)."""


def _read_exact(f, n, what):
    data = f.read(n)
    if len(data) != n:
        raise RestartError('%s: file ends inside %s' % (f.name, what))
    return data


class AmplitudeFile:
    """One amplitude file: ``amplitude_file_storer`` (amplitude_file_storer_class.F90:66-190)."""

    def __init__(self, path, block_sizes):
        self.path = os.fspath(path)
        self.block_sizes = [int(n) for n in block_sizes]
        self.n_amplitudes = sum(self.block_sizes)

    def exists(self):
        return os.path.isfile(self.path)

    def save(self, vector, energy=0.0):
        """save_ (amplitude_file_storer_class.F90:99-137): energy, then (n, block) per block."""
        v = np.ascontiguousarray(vector, dtype=np.float64).reshape(-1)
        if v.size != self.n_amplitudes:
            raise RestartError('%s: vector of %d elements, the file holds %d'
                               % (self.path, v.size, self.n_amplitudes))
        with open(self.path, 'wb') as f:
            f.write(np.float64(energy).tobytes())
            start = 0
            for n in self.block_sizes:
                f.write(np.int64(n).tobytes())
                f.write(v[start:start + n].tobytes())
                start += n

    def read(self):
        """read_ (amplitude_file_storer_class.F90:140-183): (vector, energy, n_read).

        A block whose header is missing (end of file) ends the read; elements past
        n_read are zero.  A block of another length is an error.
        """
        v = np.zeros(self.n_amplitudes)
        with open(self.path, 'rb') as f:
            energy = float(np.frombuffer(_read_exact(f, 8, 'the energy'), dtype=np.float64)[0])
            start = 0
            for n in self.block_sizes:
                head = f.read(8)
                if not head:
                    break
                if len(head) != 8:
                    raise RestartError('%s: file ends inside a block header' % self.path)
                found = int(np.frombuffer(head, dtype=np.int64)[0])
                if found != n:
                    raise RestartError('Wrong number of amplitudes in %s. Expected %d, found %d'
                                       % (os.path.basename(self.path), n, found))
                v[start:start + n] = np.frombuffer(_read_exact(f, 8 * n, 'a block'), dtype=np.float64)
                start += n
        return v, energy, start


class OrbitalFile:
    """Hartree-Fock orbitals: save_orbitals_hf / read_orbitals_hf (file_handling_hf.F90:36-81)."""

    def __init__(self, path):
        self.path = os.fspath(path)

    def exists(self):
        return os.path.isfile(self.path)

    def save(self, C, eps):
        C = np.asarray(C, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        n_ao, n_mo = C.shape
        if eps.shape != (n_mo,):
            raise RestartError('orbital energies must have length n_mo')
        with open(self.path, 'wb') as f:
            f.write(np.int64(n_ao).tobytes())
            f.write(np.int64(n_mo).tobytes())
            f.write(eps.tobytes())
            f.write(np.asfortranarray(C).tobytes(order='F'))

    def read(self, n_ao=None, n_mo=None):
        """(C (n_ao, n_mo), eps); n_ao / n_mo, when given, must match the file."""
        with open(self.path, 'rb') as f:
            a, m = np.frombuffer(_read_exact(f, 16, 'the dimensions'), dtype=np.int64)
            a, m = int(a), int(m)
            if n_ao is not None and a != n_ao:
                raise RestartError('Number of AOs does not match')
            if n_mo is not None and m != n_mo:
                raise RestartError('Number of MOs does not match')
            eps = np.frombuffer(_read_exact(f, 8 * m, 'the orbital energies'), dtype=np.float64).copy()
            C = np.frombuffer(_read_exact(f, 8 * a * m, 'the coefficients'), dtype=np.float64)
        return C.reshape(m, a).T.copy(), eps


# ---- packed doubles: the ai,bj -> bi,aj permutation ---------------------------------------

def _swap_index(n_v, n_o, lo, hi):
    """Packed positions of element (bi, aj) for the packed positions lo..hi-1 of (ai, bj)."""
    k = np.arange(lo, hi, dtype=np.int64)
    col = ((np.sqrt(8.0 * k + 1.0) - 1.0) / 2.0).astype(np.int64)   # bj, with ai <= bj
    col -= (col * (col + 1) // 2 > k)
    col += ((col + 1) * (col + 2) // 2 <= k)
    row = k - col * (col + 1) // 2                                   # ai
    a, i = row % n_v, row // n_v
    b, j = col % n_v, col // n_v
    x = b + n_v * i
    y = a + n_v * j
    hi_, lo_ = np.maximum(x, y), np.minimum(x, y)
    return hi_ * (hi_ + 1) // 2 + lo_


def _exchange_doubles(x2, n_v, n_o, chunk=1 << 22):
    """y[(ai, bj)] = x[(bi, aj)] over eT's packed doubles."""
    n = x2.size
    if n != (n_v * n_o) * (n_v * n_o + 1) // 2:
        raise RestartError('packed doubles of length %d do not fit n_v = %d, n_o = %d' % (n, n_v, n_o))
    y = np.empty_like(x2)
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        y[lo:hi] = x2[_swap_index(n_v, n_o, lo, hi)]
    return y


def packed_contravariant(x2, n_v, n_o):
    """construct_packed_contravariant (src/tools/reordering.F90:4542-4591): 2 X_aibj - X_biaj."""
    x2 = np.asarray(x2, dtype=np.float64)
    return 2.0 * x2 - _exchange_doubles(x2, n_v, n_o)


def packed_covariant(x2, n_v, n_o):
    """construct_packed_covariant (src/tools/reordering.F90:4594-4645): (2 X_aibj + X_biaj) / 3."""
    x2 = np.asarray(x2, dtype=np.float64)
    return (2.0 * x2 + _exchange_doubles(x2, n_v, n_o)) / 3.0


def restart_vector(vector, n_read, n_o, n_v, restart_from, restart_to, name=''):
    """get_restart_vector_doubles (file_handling_doubles.F90:36-119) on a vector read from file.

    A singles-only vector gets zero doubles.  A right vector restarting a left one
    becomes 2 (R1, 2 R2 - R2^T); a left one restarting a right one (L1, (2 L2 + L2^T) / 3) / 2.
    """
    n_t1 = n_o * n_v
    v = np.array(vector, dtype=np.float64)
    if n_read == n_t1:
        v[n_t1:] = 0.0
        return v
    if n_read != v.size:
        raise RestartError('Did not recognize number of amplitudes in %s expected %d or %d found %d '
                           'amplitudes.' % (name, v.size, n_t1, n_read))
    if restart_from == restart_to:
        return v
    if restart_from == 'right' and restart_to == 'left':
        return 2.0 * np.concatenate([v[:n_t1], packed_contravariant(v[n_t1:], n_v, n_o)])
    if restart_from == 'left' and restart_to == 'right':
        return 0.5 * np.concatenate([v[:n_t1], packed_covariant(v[n_t1:], n_v, n_o)])
    raise RestartError('restart side not recognized: %r -> %r' % (restart_from, restart_to))


# ---- the files of one CCSD calculation -----------------------------------------------------

class CCSDFiles:
    """The restart directory of a CCSD calculation (file_handling_ccs.F90:36-99).

    ``t`` and ``tbar`` hold the ground-state amplitudes and multipliers, ``right(k)`` /
    ``left(k)`` the k-th excited state (k = 1, 2, ...), ``orbitals`` the HF orbitals.
    """

    def __init__(self, directory, n_o, n_v):
        self.directory = os.fspath(directory)
        self.n_o, self.n_v = int(n_o), int(n_v)
        self.n_t1 = self.n_o * self.n_v
        self.n_t2 = self.n_t1 * (self.n_t1 + 1) // 2
        self.block_sizes = [self.n_t1, self.n_t2]
        self.t = AmplitudeFile(self._path('t'), self.block_sizes)
        self.tbar = AmplitudeFile(self._path('tbar'), self.block_sizes)
        self.orbitals = OrbitalFile(self._path('orbital_coefficients'))

    def _path(self, name):
        return os.path.join(self.directory, name)

    def right(self, k):
        return AmplitudeFile(self._path('r_%03d' % k), self.block_sizes)

    def left(self, k):
        return AmplitudeFile(self._path('l_%03d' % k), self.block_sizes)

    def side(self, side, k):
        if side == 'right':
            return self.right(k)
        if side == 'left':
            return self.left(k)
        raise RestartError('Trying to restart an excited state, but argument side not recognized: '
                           + str(side))

    def make_directory(self):
        os.makedirs(self.directory, exist_ok=True)

    def check_and_get_restart_vector(self, k, side):
        """check_and_get_restart_vector_ccs (file_handling_ccs.F90:430-503).

        (vector, energy) of state k for `side`, from its own file or else from the
        other side's (transformed), or None when neither exists.
        """
        other = 'left' if side == 'right' else 'right'
        for source in (side, other):
            f = self.side(source, k)
            if f.exists():
                v, energy, n_read = f.read()
                return restart_vector(v, n_read, self.n_o, self.n_v, source, side,
                                      os.path.basename(f.path)), energy
        return None


# ---- eT v1.0 restart directories ----------------------------------------------------------
#
# Layout of a v1.0 directory (every file Fortran sequential unformatted: each record is
# a 4-byte native-order length L, L bytes, then L again):
#
#   cc_restart_file      four records of one int64: n_o, n_v, and the lengths of the
#                        ground-state and excited-state vectors (n_t1 only for CCS / CC2
#                        ground states, n_t1 + n_t2 for CCSD)
#   t, tbar              the ground-state vector, one record per block (singles, doubles)
#   r_00k, l_00k         the k-th excited state, the whole vector in one record
#   excitation_energies  a record with the int64 state count, then one with the energies
#   orbital_energies     one record of n_mo energies per spin (one for RHF)
#   orbital_coefficients one record of the n_ao x n_mo coefficients per spin
#   scf_restart_file     present when the Hartree-Fock files are
#
# The conversion re-reads every vector as a list of records, checks its length against
# the dimensions of cc_restart_file, and writes it through AmplitudeFile / the orbital
# layout above.  n_mo is the length of an orbital-energy record and n_ao that of a
# coefficient record divided by n_mo.  v1.0 ground-state files carry no energy: 0 is
# written; excited states take theirs from excitation_energies.

def fortran_records(path):
    """Every record of a Fortran sequential unformatted file, as bytes."""
    out = []
    with open(path, 'rb') as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise RestartError('%s: %d stray bytes after the last record' % (path, len(data) - pos))
        n = int(np.frombuffer(data, dtype='=i4', count=1, offset=pos)[0])
        end = pos + 4 + n
        if n < 0 or end + 4 > len(data) or int(np.frombuffer(data, dtype='=i4', count=1, offset=end)[0]) != n:
            raise RestartError('%s: not a Fortran sequential file (record at byte %d)' % (path, pos))
        out.append(data[pos + 4:end])
        pos = end + 4
    return out


def _int_records(path):
    recs = fortran_records(path)
    for r in recs:
        if len(r) != 8:
            raise RestartError('%s: expected records of one 8-byte integer' % path)
    return [int(np.frombuffer(r, dtype='=i8')[0]) for r in recs]


def _doubles(record):
    if len(record) % 8:
        raise RestartError('a record of %d bytes does not hold float64 values' % len(record))
    return np.frombuffer(record, dtype='=f8')


def _vector_records(path, lengths):
    """The float64 records of a v1.0 vector file; their lengths must be `lengths`."""
    recs = [_doubles(r) for r in fortran_records(path)]
    found = [r.size for r in recs]
    if found != list(lengths):
        raise RestartError('%s: records of %s amplitudes, the dimensions in cc_restart_file give %s'
                           % (path, found, list(lengths)))
    return recs


def _convert_orbital_files(old_dir, new_dir):
    energies = [_doubles(r) for r in fortran_records(os.path.join(old_dir, 'orbital_energies'))]
    coefficients = [_doubles(r) for r in fortran_records(os.path.join(old_dir, 'orbital_coefficients'))]
    if not energies or len(energies) != len(coefficients):
        raise RestartError('%s: orbital_energies and orbital_coefficients hold %d and %d records'
                           % (old_dir, len(energies), len(coefficients)))
    n_mo = energies[0].size
    n_ao = coefficients[0].size // max(n_mo, 1)
    for e, c in zip(energies, coefficients):
        if e.size != n_mo or c.size != n_ao * n_mo or n_ao * n_mo == 0:
            raise RestartError('%s: orbital records of inconsistent lengths' % old_dir)
    with open(os.path.join(new_dir, 'orbital_coefficients'), 'wb') as f:
        f.write(np.array([n_ao, n_mo], dtype=np.int64).tobytes())
        for e, c in zip(energies, coefficients):
            f.write(e.astype(np.float64).tobytes())
            f.write(c.astype(np.float64).tobytes())
    return n_ao, n_mo


def convert_v1_0_directory(old_dir, new_dir):
    """Rewrites an eT v1.0 restart directory (layout above) in the current formats:
    amplitude files t, tbar, r_00k, l_00k and the orbital file.  Returns a dict of
    what was written ('orbitals': (n_ao, n_mo) or None, 'ground_state': names,
    'excited_states': number of states)."""
    old_dir, new_dir = os.path.realpath(old_dir), os.path.realpath(new_dir)
    if not os.path.isdir(old_dir):
        raise NotADirectoryError(old_dir)
    if old_dir == new_dir:
        raise RestartError('the converted files would overwrite the v1.0 files in %s' % old_dir)
    os.makedirs(new_dir, exist_ok=True)
    done = {'orbitals': None, 'ground_state': [], 'excited_states': 0}
    have = lambda name: os.path.isfile(os.path.join(old_dir, name))
    if have('scf_restart_file') and have('orbital_energies') and have('orbital_coefficients'):
        done['orbitals'] = _convert_orbital_files(old_dir, new_dir)
    if not have('cc_restart_file'):
        return done
    n_o, n_v, n_gs, n_es = _int_records(os.path.join(old_dir, 'cc_restart_file'))[:4]
    n_t1 = n_o * n_v
    n_t2 = n_t1 * (n_t1 + 1) // 2
    gs_blocks = [n_t1] if n_gs == n_t1 else [n_t1, n_t2]
    es_blocks = [n_t1] if n_es == n_t1 else [n_t1, n_t2]
    for name in ('t', 'tbar'):
        if have(name):
            x = np.concatenate(_vector_records(os.path.join(old_dir, name), gs_blocks))
            AmplitudeFile(os.path.join(new_dir, name), gs_blocks).save(x, 0.0)
            done['ground_state'].append(name)
    if have('excitation_energies'):
        recs = fortran_records(os.path.join(old_dir, 'excitation_energies'))
        if len(recs) < 2 or len(recs[0]) != 8:
            raise RestartError('%s: expected the state count and the energies' % old_dir)
        n_states = int(np.frombuffer(recs[0], dtype='=i8')[0])
        omega = _doubles(recs[1])
        if omega.size < n_states:
            raise RestartError('%s: %d excitation energies for %d states' % (old_dir, omega.size, n_states))
        for k in range(1, n_states + 1):
            for prefix in ('r', 'l'):
                name = '%s_%03d' % (prefix, k)
                if have(name):
                    x = _vector_records(os.path.join(old_dir, name), [sum(es_blocks)])[0]
                    AmplitudeFile(os.path.join(new_dir, name), es_blocks).save(x, float(omega[k - 1]))
        done['excited_states'] = n_states
    return done


def main(argv=None):
    """python -m et_amd.restart V1_0_DIR NEW_DIR: converts an eT v1.0 restart directory."""
    import argparse
    p = argparse.ArgumentParser(description='Rewrite an eT v1.0 restart directory in the current formats.')
    p.add_argument('v1_0_dir')
    p.add_argument('new_dir')
    a = p.parse_args(argv)
    print(convert_v1_0_directory(a.v1_0_dir, a.new_dir))


if __name__ == '__main__':
    main()
