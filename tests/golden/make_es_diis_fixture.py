#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_es_diis.json from eT's tests/ccsd_right_es_diis and
tests/ccsd_left_es_diis (H2O-He / cc-pVDZ, the system of h2o_he_ccpvdz.json).

Both run eT's DIIS excited-state solver (diis_cc_es_class.F90) for two states at residual
and energy thresholds 1e-10.  Read as text from result/*.out: every iteration's table
(root, energy, residual norm), the number of iterations and the final excitation energies.

Usage: python tests/golden/make_es_diis_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_es_diis.json')


def main():
    out = {}
    for side in ('right', 'left'):
        rel = 'tests/ccsd_%s_es_diis/result/ccsd_%s_es_diis.out' % (side, side)
        lines = open(os.path.join(REF, rel)).read().splitlines()
        start = next(i for i, ln in enumerate(lines) if 'Root     Eigenvalue (Re)     Residual norm' in ln)
        rows, it = [], 0
        for ln in lines[start - 4:]:
            m = re.match(r'\s+Iteration:\s+(\d+)', ln)
            if m:
                it = int(m.group(1))
                continue
            m = re.match(r'\s+(\d)\s+(0\.\d+)\s+(\d\.\d+E[-+]\d+)\s*$', ln)
            if m:
                rows.append([it, int(m.group(1)), float(m.group(2)), float(m.group(3))])
            if 'Convergence criterion met' in ln:
                n_it = int(re.search(r'in (\d+) iterations', ln).group(1))
                break
        k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
        final = [float(lines[k + 2 + q].split()[1]) for q in range(2)]
        out[side] = {'iterations': n_it, 'table': rows, 'excitation_energies': final,
                     'citation': '%s:%d-%d' % (rel, start, k + 4)}
    json.dump(out, open(OUT, 'w'), indent=0)
    for side, v in out.items():
        print(side, v['iterations'], v['excitation_energies'], len(v['table']), v['table'][:2])


if __name__ == '__main__':
    main()
