#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_v1_0_restart_files.json: eT's v1.0 restart directories as data.

eT's tests/restart_ccsd_v1.0 and tests/restart_ccsd_from_cc2_v1.0 (H2O-He / cc-pVDZ)
ship v1.0 restart directories (Fortran sequential files, little endian).  Stored here
for each file: its size and SHA-256; the bytes themselves (hex) for the files other
than the CCSD t / tbar / r_00k / l_00k vectors, whose values are already in
h2o_he_v1_0_vectors.npz.  A test rebuilds the directories from these and checks every
rebuilt file against the size and hash of eT's.
Usage: python tests/golden/make_h2o_he_v1_0_restart_files.py [/root/reference]
"""
import hashlib
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_v1_0_restart_files.json')
SETS = {
    'ccsd': 'tests/restart_ccsd_v1.0/old_restart_files',
    'cc2': 'tests/restart_ccsd_from_cc2_v1.0/old_restart_files',
}
VECTOR_FILES = {'t', 'tbar', 'r_001', 'r_002', 'r_003', 'l_001', 'l_002', 'l_003'}


def main():
    out = {'source': 'eT test data: ' + ', '.join(SETS.values()), 'sets': {}}
    for key, rel in SETS.items():
        src = os.path.join(REF, rel)
        files = {}
        for name in sorted(os.listdir(src)):
            data = open(os.path.join(src, name), 'rb').read()
            entry = {'size': len(data), 'sha256': hashlib.sha256(data).hexdigest()}
            if not (key == 'ccsd' and name in VECTOR_FILES) and not (key == 'cc2' and name.startswith('r_')):
                entry['hex'] = data.hex()
            files[name] = entry
        out['sets'][key] = files
    json.dump(out, open(OUT, 'w'), indent=1, sort_keys=True)
    for key, files in out['sets'].items():
        print(key, {k: (v['size'], 'hex' in v) for k, v in files.items()})


if __name__ == '__main__':
    main()
