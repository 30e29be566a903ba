#!/usr/bin/env python3
"""Writes the Newton-Raphson fixtures from eT's tests:
  tests/golden/h2o_he_newton_raphson.json  tests/ccsd_gs_energy_nr: H2O-He / cc-pVDZ,
      Cholesky threshold 1e-10, relative micro threshold 1e-2, residual and energy
      thresholds 1e-10, DIIS dimension 8;
  tests/golden/o3_newton_raphson_crop.json  tests/ccsd_gs_energy_nr_crop: ozone / cc-pVDZ,
      Cholesky threshold 1e-10, residual threshold 1e-8, CROP with DIIS dimension 3.
Read as text from each input and its result/*.out:
  geometry (Angstrom), thresholds;
  E_nuc, E_HF, number of Cholesky vectors, number of occupied orbitals;
  the macro-iteration table (energy, |omega|, delta E) and, after each macro iteration,
  the residual norms of its linear (micro) iterations;
  the final ground-state energy.

Usage: python tests/golden/make_newton_raphson_fixtures.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = [
    ('ccsd_gs_energy_nr', 'h2o_he_newton_raphson.json',
     {'cholesky_threshold': 1e-10, 'relative_micro_threshold': 1e-2, 'residual_threshold': 1e-10,
      'energy_threshold': 1e-10, 'diis_dimension': 8, 'crop': False}),
    ('ccsd_gs_energy_nr_crop', 'o3_newton_raphson_crop.json',
     {'cholesky_threshold': 1e-10, 'relative_micro_threshold': 1e-2, 'residual_threshold': 1e-8,
      'energy_threshold': None, 'diis_dimension': 3, 'crop': True}),
]


def write(name, out_name, settings):
    STEM = 'tests/%s/%s' % (name, name)
    REL = 'tests/%s/result/%s.out' % (name, name)
    inp = open(os.path.join(REF, STEM + '.inp')).read()
    lines = open(os.path.join(REF, REL)).read().splitlines()
    geom = [[s[0], float(s[1]), float(s[2]), float(s[3])]
            for s in (ln.split() for ln in inp.split('basis: cc-pVDZ')[1].split('end geometry')[0].splitlines())
            if len(s) == 4]
    g, cite = {}, {}

    def grab(key, pattern, conv):
        for i, ln in enumerate(lines):
            m = re.search(pattern, ln)
            if m:
                g[key], cite[key] = conv(m.group(1)), '%s:%d' % (REL, i + 1)
                return
        raise RuntimeError(pattern)

    grab('nuclear_repulsion', r'Nuclear repulsion energy:\s+(-?[0-9.]+)', float)
    grab('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float)
    grab('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int)
    grab('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', float)
    grab('n_o', r'Number of occupied orbitals:\s+([0-9]+)', int)
    head = next(i for i, ln in enumerate(lines) if 'Iteration    Energy (a.u.)        |omega|' in ln)
    end = next(i for i in range(head, len(lines)) if 'Convergence criterion met' in lines[i])
    macro, micro, block = [], [], None
    for ln in lines[head + 2:end]:
        s = ln.split()
        if len(s) == 4 and s[0].isdigit():
            macro.append([int(s[0]), float(s[1]), float(s[2]), float(s[3])])
        elif 'Iteration     Residual norm' in ln:
            block = []
            micro.append(block)
        elif len(s) == 2 and s[0].isdigit() and block is not None:
            block.append(float(s[1]))
    g['macro_iterations'] = macro             # [iteration, energy, |omega|, delta E]
    g['micro_residual_norms'] = micro         # one list per Newton-Raphson update
    cite['macro_iterations'] = '%s:%d-%d' % (REL, head + 1, end + 1)
    out = {'source': STEM + '.inp', 'basis': 'cc-pVDZ', 'bohr_to_angstrom': 0.52917721092,
           'geometry_angstrom': geom, **settings, 'golden': g, 'citations': cite}
    json.dump(out, open(os.path.join(HERE, out_name), 'w'), indent=1)
    print(out_name, len(g['macro_iterations']), 'macro iterations,',
          [len(m) for m in g['micro_residual_norms']], 'linear iterations')


def main():
    for case in CASES:
        write(*case)


if __name__ == '__main__':
    main()
