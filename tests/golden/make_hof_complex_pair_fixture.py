#!/usr/bin/env python3
"""Writes tests/golden/hof_complex_pair.json from eT's tests/ccsd_right_es_energies_imaginary.

HOF / aug-cc-pVDZ, Cholesky threshold 1e-6, four right EOM-CCSD roots, of which the
third and fourth are a complex-conjugate pair.  Read as text from the input and
result/ccsd_right_es_energies_imaginary.out:
  geometry (Angstrom), thresholds;
  E_nuc, E_HF, number of Cholesky vectors, E_CCSD;
  the final excitation energies (real parts);
  the last Davidson iteration: Re and Im of each root.
eT's own check of this test compares these to 1e-4 Eh (its file `test`).

Also writes tests/golden/hof_false_positive.json from eT's tests/ccsd_es_false_positive_defect
(HOF at another geometry, aug-cc-pVDZ, Cholesky 1e-6, three right roots at residual 1e-4,
the third next to a complex pair): geometry, E_HF, the number of Cholesky vectors, E_CCSD,
the number of Davidson iterations, the final roots and the last iteration's residual norms.

Usage: python tests/golden/make_hof_complex_pair_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
STEM = 'tests/ccsd_right_es_energies_imaginary/ccsd_right_es_energies_imaginary'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'hof_complex_pair.json')


def main():
    inp = open(os.path.join(REF, STEM + '.inp')).read()
    rel = 'tests/ccsd_right_es_energies_imaginary/result/ccsd_right_es_energies_imaginary.out'
    lines = open(os.path.join(REF, rel)).read().splitlines()
    geom = [[s[0], float(s[1]), float(s[2]), float(s[3])]
            for s in (ln.split() for ln in inp.split('basis: aug-cc-pVDZ')[1].split('end geometry')[0].splitlines())
            if len(s) == 4]
    g, cite = {}, {}

    def grab(key, pattern, conv):
        for i, ln in enumerate(lines):
            m = re.search(pattern, ln)
            if m:
                g[key], cite[key] = conv(m.group(1)), '%s:%d' % (rel, i + 1)
                return
        raise RuntimeError(pattern)

    grab('nuclear_repulsion', r'Nuclear repulsion energy:\s+(-?[0-9.]+)', float)
    grab('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float)
    grab('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int)
    grab('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', float)
    grab('n_o', r'Number of occupied orbitals:\s+([0-9]+)', int)
    # final table: "State (Hartree) (eV)" then a rule, then one row per state
    k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
    rows = [lines[k + 2 + q].split() for q in range(4)]
    g['excitation_energies'] = [float(r[1]) for r in rows]
    cite['excitation_energies'] = '%s:%d-%d' % (rel, k + 3, k + 6)
    # the last Davidson iteration before convergence
    conv = next(i for i, ln in enumerate(lines) if 'Convergence criterion met' in ln and i > k - 200)
    head = max(i for i, ln in enumerate(lines[:conv]) if 'Root  Eigenvalue (Re)' in ln)
    last = [lines[head + 2 + q].split() for q in range(4)]
    g['last_iteration_re'] = [float(r[1]) for r in last]
    g['last_iteration_im'] = [float(r[2]) for r in last]
    g['last_iteration_residual'] = [float(r[3]) for r in last]
    cite['last_iteration'] = '%s:%d-%d' % (rel, head + 3, head + 6)
    out = {'source': STEM + '.inp', 'basis': 'aug-cc-pVDZ', 'bohr_to_angstrom': 0.52917721092,
           'geometry_angstrom': geom, 'cholesky_threshold': 1e-6, 'n_states': 4,
           'golden': g, 'citations': cite}
    json.dump(out, open(OUT, 'w'), indent=1)
    print(json.dumps(g, indent=1))


def false_positive():
    stem = 'tests/ccsd_es_false_positive_defect/ccsd_es_false_positive_defect'
    rel = 'tests/ccsd_es_false_positive_defect/result/ccsd_es_false_positive_defect.out'
    inp = open(os.path.join(REF, stem + '.inp')).read()
    lines = open(os.path.join(REF, rel)).read().splitlines()
    geom = [[s[0], float(s[1]), float(s[2]), float(s[3])]
            for s in (ln.split() for ln in inp.split('basis: aug-cc-pVDZ')[1].split('end geometry')[0].splitlines())
            if len(s) == 4]
    g, cite = {}, {}
    for key, pattern, conv in (('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float),
                               ('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int),
                               ('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', float)):
        i = next(i for i, ln in enumerate(lines) if re.search(pattern, ln))
        g[key], cite[key] = conv(re.search(pattern, lines[i]).group(1)), '%s:%d' % (rel, i + 1)
    conv = max(i for i, ln in enumerate(lines) if 'Convergence criterion met' in ln)
    g['iterations'] = int(re.search(r'in (\d+) iterations', lines[conv]).group(1))
    head = max(i for i, ln in enumerate(lines[:conv]) if 'Root  Eigenvalue (Re)' in ln)
    g['last_iteration_residual'] = [float(lines[head + 2 + q].split()[3]) for q in range(3)]
    cite['last_iteration_residual'] = '%s:%d-%d' % (rel, head + 3, head + 5)
    k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
    g['excitation_energies'] = [float(lines[k + 2 + q].split()[1]) for q in range(3)]
    cite['excitation_energies'] = '%s:%d-%d' % (rel, k + 3, k + 5)
    out = {'source': stem + '.inp', 'basis': 'aug-cc-pVDZ', 'bohr_to_angstrom': 0.52917721092,
           'geometry_angstrom': geom, 'cholesky_threshold': 1e-6, 'n_states': 3, 'residual_threshold': 1e-4,
           'golden': g, 'citations': cite}
    json.dump(out, open(os.path.join(os.path.dirname(OUT), 'hof_false_positive.json'), 'w'), indent=1)
    print(json.dumps(g, indent=1))


if __name__ == '__main__':
    main()
    false_positive()
