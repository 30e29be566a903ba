#!/usr/bin/env python3
"""Writes tests/golden/n2_cvs_degeneracy.json from eT's tests/ccsd_cvs_eom_degeneracy.

N2 / cc-pVDZ, Cholesky threshold 1e-12, ground state to 1e-11, then core-valence separated
EOM-CCSD from both 1s MOs (`core excitation: {1,2}`, nearly degenerate), four right and four
left roots (two degenerate pairs) by the eigen Davidson solver at residual 1e-11.  Read as text
from the input and result/*.out: geometry (Angstrom), E_nuc, E_HF, the number of Cholesky
vectors, E_CCSD, and for each side the number of Davidson iterations and the final roots.

Also writes tests/golden/h2o_ccx_cvs.json from eT's tests/ccsd_right_cvs_ccX-dz: H2O with the
core-valence basis ccX-DZ on O and cc-pVDZ on H (48 AOs), Cholesky threshold 1e-10, one right
CVS root out of the O 1s at residual 1e-5: the AO count, E_HF, the number of Cholesky vectors,
E_CCSD and the root.

Usage: python tests/golden/make_cvs_degeneracy_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
STEM = 'tests/ccsd_cvs_eom_degeneracy/ccsd_cvs_eom_degeneracy'
REL = 'tests/ccsd_cvs_eom_degeneracy/result/ccsd_cvs_eom_degeneracy.out'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'n2_cvs_degeneracy.json')


def main():
    inp = open(os.path.join(REF, STEM + '.inp')).read()
    lines = open(os.path.join(REF, REL)).read().splitlines()
    geom = [[s[0], float(s[1]), float(s[2]), float(s[3])]
            for s in (ln.split() for ln in inp.split('basis: cc-pVDZ')[1].split('end geometry')[0].splitlines())
            if len(s) == 4]
    g, cite = {}, {}
    for key, pattern, conv in (('nuclear_repulsion', r'Nuclear repulsion energy:\s+(-?[0-9.]+)', float),
                               ('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float),
                               ('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int),
                               ('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', float)):
        i = next(i for i, ln in enumerate(lines) if re.search(pattern, ln))
        g[key], cite[key] = conv(re.search(pattern, lines[i]).group(1)), '%s:%d' % (REL, i + 1)
    # the two excited-state sections: right, then left
    sections = [i for i, ln in enumerate(lines) if 'Determining CC excited state amplitudes' in ln]
    for side, start in zip(('right', 'left'), sections):
        conv = next(i for i in range(start, len(lines)) if 'Convergence criterion met' in lines[i])
        k = next(i for i in range(start, len(lines)) if 'State                (Hartree)             (eV)' in lines[i])
        g[side] = {'iterations': int(re.search(r'in (\d+) iterations', lines[conv]).group(1)),
                   'excitation_energies': [float(lines[k + 2 + q].split()[1]) for q in range(4)]}
        cite[side] = '%s:%d, %d-%d' % (REL, conv + 1, k + 3, k + 6)
    out = {'source': STEM + '.inp', 'basis': 'cc-pVDZ', 'bohr_to_angstrom': 0.52917721092,
           'geometry_angstrom': geom, 'cholesky_threshold': 1e-12, 'core_mos': [0, 1], 'n_states': 4,
           'residual_threshold': 1e-11, 'golden': g, 'citations': cite}
    json.dump(out, open(OUT, 'w'), indent=1)
    print(json.dumps(g, indent=1))


def ccx():
    stem = 'tests/ccsd_right_cvs_ccX-dz/ccsd_right_cvs_ccX-dz'
    rel = 'tests/ccsd_right_cvs_ccX-dz/result/ccsd_right_cvs_ccX-dz.out'
    inp = open(os.path.join(REF, stem + '.inp')).read()
    lines = open(os.path.join(REF, rel)).read().splitlines()
    geom, basis = [], None
    for ln in inp.split('geometry')[1].split('end geometry')[0].splitlines():
        s = ln.split()
        if ln.strip().startswith('basis:'):
            basis = s[1]
        elif len(s) == 4:
            geom.append([s[0], float(s[1]), float(s[2]), float(s[3]), basis])
    g, cite = {}, {}
    for key, pattern, conv in (('n_ao', r'Number of atomic orbitals:\s+([0-9]+)', int),
                               ('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float),
                               ('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int),
                               ('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', float)):
        i = next(i for i, ln in enumerate(lines) if re.search(pattern, ln))
        g[key], cite[key] = conv(re.search(pattern, lines[i]).group(1)), '%s:%d' % (rel, i + 1)
    k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
    g['excitation_energies'] = [float(lines[k + 2].split()[1])]
    cite['excitation_energies'] = '%s:%d' % (rel, k + 3)
    out = {'source': stem + '.inp', 'bohr_to_angstrom': 0.52917721092, 'geometry_angstrom_basis': geom,
           'cholesky_threshold': 1e-10, 'core_mos': [0], 'n_states': 1, 'golden': g, 'citations': cite}
    json.dump(out, open(os.path.join(os.path.dirname(OUT), 'h2o_ccx_cvs.json'), 'w'), indent=1)
    print(json.dumps(g, indent=1))


if __name__ == '__main__':
    main()
    ccx()
