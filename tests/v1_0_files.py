"""Rebuilds eT's v1.0 restart directories (tests/restart_ccsd_v1.0 and
tests/restart_ccsd_from_cc2_v1.0, H2O-He / cc-pVDZ) from the committed fixtures.

golden/h2o_he_v1_0_restart_files.json holds each file's size and SHA-256 and the bytes
of the small ones; the CCSD t / tbar / r_00k / l_00k vectors come from
golden/h2o_he_v1_0_vectors.npz, written back as Fortran sequential records (4-byte
little-endian length markers): t and tbar as two records (singles, doubles), the
excited states as one.  Every rebuilt file is checked against eT's size and hash.
"""
import functools
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = os.path.join(HERE, 'golden', 'h2o_he_v1_0_restart_files.json')
VECTORS = os.path.join(HERE, 'golden', 'h2o_he_v1_0_vectors.npz')
N_O, N_V = 6, 23
N_T1 = N_O * N_V


@functools.lru_cache(maxsize=1)
def fixture():
    return json.load(open(FILES))


@functools.lru_cache(maxsize=1)
def vectors():
    with np.load(VECTORS) as v:
        return {k: v[k] for k in v.files}


def record(*arrays):
    body = b''.join(np.ascontiguousarray(a, dtype='<f8').tobytes() for a in arrays)
    n = len(body).to_bytes(4, 'little')
    return n + body + n


def build(directory, which='ccsd'):
    """Writes the v1.0 directory `which` ('ccsd' or 'cc2') into `directory`; returns its path."""
    os.makedirs(directory, exist_ok=True)
    files = fixture()['sets'][which]
    v = vectors()
    for name, entry in files.items():
        if 'hex' in entry:
            data = bytes.fromhex(entry['hex'])
        elif which == 'ccsd' and name in ('t', 'tbar'):
            data = record(v[name][:N_T1]) + record(v[name][N_T1:])
        elif which == 'ccsd':
            data = record(v[name[0] + str(int(name[2:]))])
        else:
            continue                    # the CC2 excited states are not part of the fixtures
        if len(data) != entry['size'] or hashlib.sha256(data).hexdigest() != entry['sha256']:
            raise AssertionError('rebuilt %s/%s differs from eT\'s file' % (which, name))
        with open(os.path.join(directory, name), 'wb') as f:
            f.write(data)
    return directory
