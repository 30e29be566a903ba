#!/usr/bin/env python3
"""Writes tests/golden/cc_pvdz_molecules.json: two more of eT's CCSD test molecules in cc-pVDZ.

Reads (text only, nothing is imported or executed from the reference):
  tests/ccsd_lr_polarizability/ccsd_lr_polarizability.inp and result/*.out
      H2O, Cholesky threshold 1e-11: E_nuc, E_HF, n_J, the DIIS energies, E_CCSD
  tests/ccsd_gs_energy_nr_crop/ccsd_gs_energy_nr_crop.inp and result/*.out
      O3, Cholesky threshold 1e-10: E_nuc, E_HF, n_J, E_CCSD (Newton-Raphson, omega 1e-8)
Each value keeps the output line it came from.  Bohr/Angstrom is eT's
(src/various/parameters.F90:105).

Usage: python tests/golden/make_cc_pvdz_molecules.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'cc_pvdz_molecules.json')
TESTS = {
    'h2o': 'tests/ccsd_lr_polarizability/ccsd_lr_polarizability',
    'o3': 'tests/ccsd_gs_energy_nr_crop/ccsd_gs_energy_nr_crop',
}


def geometry(inp):
    atoms, inside = [], False
    for ln in open(inp):
        s = ln.split()
        if ln.startswith('geometry'):
            inside = True
        elif ln.startswith('end geometry'):
            inside = False
        elif inside and len(s) == 4:
            atoms.append([s[0], float(s[1]), float(s[2]), float(s[3])])
    return atoms


def keyword(inp, section, key):
    text = open(inp).read()
    block = re.search(r'solver %s\n(.*?)end solver %s' % (section, section), text, re.S).group(1)
    return float(re.search(r'%s:\s*([0-9.dDeE+-]+)' % key, block).group(1).replace('d', 'e').replace('D', 'e'))


def value(lines, pattern, rel):
    for i, ln in enumerate(lines):
        m = re.search(pattern, ln)
        if m:
            return m.group(1), '%s:%d' % (rel, i + 1)
    raise RuntimeError('not found: ' + pattern)


def main():
    out = {'bohr_to_angstrom': 0.52917721092, 'molecules': {}}
    for name, stem in TESTS.items():
        inp = os.path.join(REF, stem + '.inp')
        res_rel = os.path.join(os.path.dirname(stem), 'result', os.path.basename(stem) + '.out')
        lines = open(os.path.join(REF, res_rel)).read().splitlines()
        g, cite = {}, {}
        for key, pat, conv in (('nuclear_repulsion', r'Nuclear repulsion energy:\s+(-?[0-9.]+)', float),
                               ('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float),
                               ('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int),
                               ('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', float),
                               ('n_o', r'Number of occupied orbitals:\s+([0-9]+)', int)):
            v, c = value(lines, pat, res_rel)
            g[key], cite[key] = conv(v), c
        # the DIIS table of the ground state (energy per iteration); O3 was solved by
        # Newton-Raphson, whose table is not a DIIS trajectory
        start = next((i for i, ln in enumerate(lines) if 'Iteration    Energy (a.u.)        |omega|' in ln), None)
        if start is not None and name == 'h2o':
            rows = []
            for ln in lines[start + 2:]:
                s = ln.split()
                if len(s) != 4 or not s[0].isdigit():
                    break
                rows.append([int(s[0]), float(s[1]), float(s[2])])
            g['ccsd_iterations'] = rows
            cite['ccsd_iterations'] = '%s:%d-%d' % (res_rel, start + 3, start + 2 + len(rows))
        out['molecules'][name] = {
            'source': stem + '.inp',
            'geometry_angstrom': geometry(inp),
            'cholesky_threshold': keyword(inp, 'cholesky', 'threshold'),
            'golden': g,
            'citations': cite,
        }
    json.dump(out, open(OUT, 'w'), indent=1)
    for k, v in out['molecules'].items():
        print(k, v['cholesky_threshold'], {a: b for a, b in v['golden'].items() if a != 'ccsd_iterations'},
              len(v['golden'].get('ccsd_iterations', [])))


if __name__ == '__main__':
    main()
