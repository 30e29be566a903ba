#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_es_diis.json from eT's tests/ccsd_right_es_diis and
tests/ccsd_left_es_diis (H2O-He / cc-pVDZ, the system of h2o_he_ccpvdz.json).

Both run eT's DIIS excited-state solver (diis_cc_es_class.F90) for two states at residual
and energy thresholds 1e-10.  Read as text from result/*.out: every iteration's table
(root, energy, residual norm), the number of iterations and the final excitation energies; likewise from tests/ccsd_es_remove_core (eigen
Davidson with `remove core: {1}`, key right_remove_core_davidson).
The same for tests/ccsd_right_cvs_es_energies_diis (key right_cvs: core-valence separation,
`core excitation: {1}`, the oxygen 1s), and from tests/ccsd_right_cvs_es_energies (the
eigen Davidson solver with the same CVS settings, key right_cvs_davidson) the number of
iterations and the final excitation energies.

Usage: python tests/golden/make_es_diis_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_es_diis.json')


def final_energies(lines, n=2):
    k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
    return k, [float(lines[k + 2 + q].split()[1]) for q in range(n)]


def main():
    out = {}
    for side, test in (('right', 'ccsd_right_es_diis'), ('left', 'ccsd_left_es_diis'),
                       ('right_cvs', 'ccsd_right_cvs_es_energies_diis')):
        rel = 'tests/%s/result/%s.out' % (test, test)
        lines = open(os.path.join(REF, rel)).read().splitlines()
        start = next(i for i, ln in enumerate(lines) if 'Root     Eigenvalue (Re)     Residual norm' in ln)
        rows, it = [], 0
        for ln in lines[start - 4:]:
            m = re.match(r'\s+Iteration:\s+(\d+)', ln)
            if m:
                it = int(m.group(1))
                continue
            m = re.match(r'\s+(\d)\s+(\d+\.\d+)\s+(\d\.\d+E[-+]\d+)\s*$', ln)
            if m:
                rows.append([it, int(m.group(1)), float(m.group(2)), float(m.group(3))])
            if 'Convergence criterion met' in ln:
                n_it = int(re.search(r'in (\d+) iterations', ln).group(1))
                break
        k, final = final_energies(lines)
        out[side] = {'iterations': n_it, 'table': rows, 'excitation_energies': final,
                     'citation': '%s:%d-%d' % (rel, start, k + 4)}
    for key, test in (('right_cvs_davidson', 'ccsd_right_cvs_es_energies'),
                      ('right_remove_core_davidson', 'ccsd_es_remove_core')):
        rel = 'tests/%s/result/%s.out' % (test, test)
        lines = open(os.path.join(REF, rel)).read().splitlines()
        k, final = final_energies(lines)
        conv = [ln for ln in lines if 'Convergence criterion met' in ln][-1]     # the ground state's comes first
        out[key] = {'iterations': int(re.search(r'in (\d+) iterations', conv).group(1)),
                    'excitation_energies': final, 'citation': '%s:%d-%d' % (rel, k + 3, k + 4)}
    json.dump(out, open(OUT, 'w'), indent=0)
    for side, v in out.items():
        print(side, v['iterations'], v['excitation_energies'], len(v.get('table', [])), v.get('table', [])[:2])


if __name__ == '__main__':
    main()
