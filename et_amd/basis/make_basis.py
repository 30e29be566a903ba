"""Writes cc-pvdz.json, aug-cc-pvdz.json, sto-3g.json, 6-31g.json, ccx-dz.json (H-Ne) and 3-21g.json
(H-Ne, Xe)
from the reference's
Gaussian-94 basis files.

Run in a checkout that has the reference next to it; the JSON files are
committed and are what the package reads.  Data sources: ao_basis/cc-pvdz.g94,
ao_basis/augmentation-cc-pvdz.g94, ao_basis/sto-3g.g94, ao_basis/6-31g.g94 and ao_basis/3-21g.g94 and ao_basis/ccx-dz.g94
(eT's basis library).
Each element maps to a list of shells {"l": 0|1|2, "exps": [...], "coefs": [...]}
in file order; an "SP" shell (shared exponents) becomes an s shell followed by a
p shell.  aug-cc-pVDZ is cc-pVDZ followed by the augmentation shells, as eT reads
an "aug" basis (src/molecule/atomic_center_class.F90:460-497).
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
ELEMENTS = ('H', 'He', 'Li', 'Be', 'B', 'C', 'N', 'O', 'F', 'Ne')


def parse(path, elements=ELEMENTS):
    out = {}
    for block in open(path).read().split('****'):
        rows = [r for r in block.splitlines() if r.strip() and not r.lstrip().startswith('!')]
        if not rows or rows[0].split()[0] not in elements or len(rows[0].split()) != 2:
            continue
        element, shells, i = rows[0].split()[0], [], 1
        while i < len(rows):
            kind, count = rows[i].split()[0], int(rows[i].split()[1])
            cols = [r.replace('D', 'E').split() for r in rows[i + 1:i + 1 + count]]
            exps = [float(c[0]) for c in cols]
            if kind == 'SP':
                shells.append({'l': 0, 'exps': exps, 'coefs': [float(c[1]) for c in cols]})
                shells.append({'l': 1, 'exps': exps, 'coefs': [float(c[2]) for c in cols]})
            else:
                shells.append({'l': 'SPD'.index(kind), 'exps': exps, 'coefs': [float(c[1]) for c in cols]})
            i += 1 + count
        out[element] = shells
    return out


def write(name, data):
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), name + '.json'), 'w') as f:
        json.dump(data, f, indent=0)


if __name__ == '__main__':
    for name in ('cc-pvdz', 'sto-3g', '6-31g', 'ccx-dz'):
        write(name, {'source': 'ao_basis/%s.g94' % name,
                     'basis': parse(os.path.join(REF, 'ao_basis', name + '.g94'))})
    write('3-21g', {'source': 'ao_basis/3-21g.g94',
                    'basis': parse(os.path.join(REF, 'ao_basis', '3-21g.g94'), ELEMENTS + ('Xe',))})
    base = parse(os.path.join(REF, 'ao_basis', 'cc-pvdz.g94'))
    aug = parse(os.path.join(REF, 'ao_basis', 'augmentation-cc-pvdz.g94'))
    write('aug-cc-pvdz', {'source': 'ao_basis/cc-pvdz.g94 + ao_basis/augmentation-cc-pvdz.g94',
                          'basis': {e: base[e] + aug.get(e, []) for e in base}})
