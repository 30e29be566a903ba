#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_v1_0_vectors.npz from eT's v1.0 restart test files.

eT's tests/restart_ccsd_v1.0 (same H2O-He / cc-pVDZ system, geometry and
Cholesky threshold as tests/ccsd_gs_energy) ships converged eT vectors as
Fortran sequential unformatted files, little endian, 4-byte record markers
(layout as eT's dev_tools/convert_v1_0_restart_files.py:83-97,147-215 reads
them).  Read here as data, record by record:
  t, tbar            two records: singles (v, o) and packed doubles
  r_00k, l_00k       one record each: right / left EOM vectors, k = 1..3
  orbital_energies   one record (n_mo)
  excitation_energies  n_states record, then the energies in a second record
Usage: python tests/golden/make_h2o_he_v1_0_vectors.py [/root/reference]
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
SRC = os.path.join(REF, 'tests', 'restart_ccsd_v1.0', 'old_restart_files')
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_v1_0_vectors.npz')


def records(path):
    data = open(path, 'rb').read()
    out, pos = [], 0
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], 'little')
        out.append(data[pos + 4:pos + 4 + n])
        tail = int.from_bytes(data[pos + 4 + n:pos + 8 + n], 'little')
        if tail != n:
            raise ValueError('bad record marker in %s' % path)
        pos += 8 + n
    return out


def doubles(rec):
    return np.frombuffer(rec, dtype='<f8').copy()


def main():
    out = {}
    for name in ('t', 'tbar'):
        r = records(os.path.join(SRC, name))
        out[name] = np.concatenate([doubles(r[0]), doubles(r[1])])
    for side in 'rl':
        for k in (1, 2, 3):
            r = records(os.path.join(SRC, '%s_%03d' % (side, k)))
            out['%s%d' % (side, k)] = np.concatenate([doubles(x) for x in r])
    out['orbital_energies'] = doubles(records(os.path.join(SRC, 'orbital_energies'))[0])
    raw = open(os.path.join(SRC, 'excitation_energies'), 'rb').read()
    first = records(os.path.join(SRC, 'excitation_energies'))[0]
    n_states = int.from_bytes(first, 'little')
    start = 8 + len(first) + 4                     # past the first record and the next marker
    out['excitation_energies'] = np.frombuffer(raw[start:start + 8 * n_states], dtype='<f8').copy()
    np.savez_compressed(OUT, **out)
    for k, v in out.items():
        print(k, v.shape, v[:3])


if __name__ == '__main__':
    main()
