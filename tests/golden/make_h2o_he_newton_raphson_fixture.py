#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_newton_raphson.json from eT's tests/ccsd_gs_energy_nr.

H2O-He / cc-pVDZ, Cholesky threshold 1e-10, the CCSD ground state by eT's
Newton-Raphson solver (relative micro threshold 1e-2, residual and energy thresholds
1e-10).  Read as text from the input and result/ccsd_gs_energy_nr.out:
  geometry (Angstrom), thresholds;
  E_nuc, E_HF, number of Cholesky vectors, number of occupied orbitals;
  the macro-iteration table (energy, |omega|, delta E) and, after each macro iteration,
  the residual norms of its linear (micro) iterations;
  the final ground-state energy.

Usage: python tests/golden/make_h2o_he_newton_raphson_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
STEM = 'tests/ccsd_gs_energy_nr/ccsd_gs_energy_nr'
REL = 'tests/ccsd_gs_energy_nr/result/ccsd_gs_energy_nr.out'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_newton_raphson.json')


def main():
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
           'geometry_angstrom': geom, 'cholesky_threshold': 1e-10, 'relative_micro_threshold': 1e-2,
           'residual_threshold': 1e-10, 'energy_threshold': 1e-10, 'golden': g, 'citations': cite}
    json.dump(out, open(OUT, 'w'), indent=1)
    print(json.dumps(g, indent=1))


if __name__ == '__main__':
    main()
