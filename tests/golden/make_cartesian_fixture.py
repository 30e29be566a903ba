#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_cartesian.json from eT's tests/ccsd_right_es_energies_cartgauss.

H2O-He / cc-pVDZ with `cartesian gaussians` (six Cartesian d functions on O: 30 AOs),
Cholesky threshold 1e-12, ground state and two right EOM-CCSD roots to 1e-10.  Read as
text from the input and result/*.out: geometry (Angstrom), the number of AOs, E_nuc, E_HF,
the number of Cholesky vectors, E_CCSD and the excitation energies, each with its line.

Usage: python tests/golden/make_cartesian_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
STEM = 'tests/ccsd_right_es_energies_cartgauss/ccsd_right_es_energies_cartgauss'
REL = 'tests/ccsd_right_es_energies_cartgauss/result/ccsd_right_es_energies_cartgauss.out'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_cartesian.json')


def main():
    inp = open(os.path.join(REF, STEM + '.inp')).read()
    lines = open(os.path.join(REF, REL)).read().splitlines()
    geom = [[s[0], float(s[1]), float(s[2]), float(s[3])]
            for s in (ln.split() for ln in inp.split('basis: cc-pVDZ')[1].split('end geometry')[0].splitlines())
            if len(s) == 4]
    g, cite = {}, {}
    for key, pattern, conv in (('n_ao', r'Number of atomic orbitals:\s+([0-9]+)', int),
                               ('nuclear_repulsion', r'Nuclear repulsion energy:\s+(-?[0-9.]+)', float),
                               ('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float),
                               ('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int),
                               ('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', float)):
        i = next(i for i, ln in enumerate(lines) if re.search(pattern, ln))
        g[key], cite[key] = conv(re.search(pattern, lines[i]).group(1)), '%s:%d' % (REL, i + 1)
    k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
    g['excitation_energies'] = [float(lines[k + 2 + q].split()[1]) for q in range(2)]
    cite['excitation_energies'] = '%s:%d-%d' % (REL, k + 3, k + 4)
    out = {'source': STEM + '.inp', 'basis': 'cc-pVDZ', 'cartesian': True, 'bohr_to_angstrom': 0.52917721092,
           'geometry_angstrom': geom, 'cholesky_threshold': 1e-12, 'golden': g, 'citations': cite}
    json.dump(out, open(OUT, 'w'), indent=1)
    print(json.dumps(g, indent=1))


if __name__ == '__main__':
    main()
