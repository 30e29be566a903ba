"""Writes cc-pvdz.json (H-Ne) from the reference's Gaussian-94 basis file.

Run in a checkout that has the reference next to it; the JSON is committed and
is what the package reads.  Data source: ao_basis/cc-pvdz.g94 (eT's basis
library, the cc-pVDZ sets of Dunning).  Each element maps to a list of shells
{"l": 0|1|2, "exps": [...], "coefs": [...]} in file order, which is eT's AO
shell order for the element.
"""
import json
import os
import sys

SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else '/root/reference', 'ao_basis', 'cc-pvdz.g94')
ELEMENTS = ('H', 'He', 'Li', 'Be', 'B', 'C', 'N', 'O', 'F', 'Ne')


def parse(path):
    out = {}
    for block in open(path).read().split('****'):
        rows = [r for r in block.splitlines() if r.strip() and not r.lstrip().startswith('!')]
        if not rows or rows[0].split()[0] not in ELEMENTS or len(rows[0].split()) != 2:
            continue
        element, shells, i = rows[0].split()[0], [], 1
        while i < len(rows):
            kind, count = rows[i].split()[0], int(rows[i].split()[1])
            pairs = [r.replace('D', 'E').split()[:2] for r in rows[i + 1:i + 1 + count]]
            shells.append({'l': 'SPD'.index(kind), 'exps': [float(a) for a, _ in pairs],
                           'coefs': [float(b) for _, b in pairs]})
            i += 1 + count
        out[element] = shells
    return out


if __name__ == '__main__':
    data = {'source': 'ao_basis/cc-pvdz.g94', 'basis': parse(SRC)}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'cc-pvdz.json'), 'w') as f:
        json.dump(data, f, indent=0)
