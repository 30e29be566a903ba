#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_ip.json from eT's ionisation tests: tests/ccsd_right_ip_energies,
ccsd_left_ip_energies (eigen Davidson), ccsd_right_ip_diis, ccsd_left_ip_diis (DIIS).

H2O-He / cc-pVDZ with eT's bath orbital, Cholesky threshold 1e-11, two ionised states at
residual threshold 1e-11.  Read as text from the first input (geometry, Angstrom) and each
result/*.out: E_nuc, E_HF, the number of Cholesky vectors, E_CCSD, the excited-state
solver's number of iterations and the final ionisation energies.

Usage: python tests/golden/make_ip_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_ip.json')
TESTS = {'right_davidson': 'ccsd_right_ip_energies', 'left_davidson': 'ccsd_left_ip_energies',
         'right_diis': 'ccsd_right_ip_diis', 'left_diis': 'ccsd_left_ip_diis'}


def main():
    inp = open(os.path.join(REF, 'tests/ccsd_right_ip_energies/ccsd_right_ip_energies.inp')).read()
    geom = [[s[0], float(s[1]), float(s[2]), float(s[3])]
            for s in (ln.split() for ln in inp.split('basis: cc-pVDZ')[1].split('end geometry')[0].splitlines())
            if len(s) == 4]
    out = {'source': 'tests/ccsd_right_ip_energies/ccsd_right_ip_energies.inp', 'basis': 'cc-pVDZ',
           'bohr_to_angstrom': 0.52917721092, 'geometry_angstrom': geom, 'cholesky_threshold': 1e-11,
           'residual_threshold': 1e-11, 'n_states': 2, 'runs': {}}
    for key, test in TESTS.items():
        rel = 'tests/%s/result/%s.out' % (test, test)
        lines = open(os.path.join(REF, rel)).read().splitlines()
        text = '\n'.join(lines)
        run = {'nuclear_repulsion': float(re.search(r'Nuclear repulsion energy:\s+(-?[0-9.]+)', text).group(1)),
               'hf_energy': float(re.search(r'Total energy:\s+(-?[0-9.]+)', text).group(1)),
               'n_cholesky': int(re.search(r'Final number of Cholesky vectors:\s+([0-9]+)', text).group(1)),
               'ccsd_energy': float(re.search(r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', text).group(1))}
        conv = [ln for ln in lines if 'Convergence criterion met' in ln][-1]     # the excited states' is last
        run['iterations'] = int(re.search(r'in (\d+) iterations', conv).group(1))
        k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
        run['ionisation_energies'] = [float(lines[k + 2 + q].split()[1]) for q in range(2)]
        run['citation'] = '%s:%d-%d' % (rel, k + 3, k + 4)
        out['runs'][key] = run
    json.dump(out, open(OUT, 'w'), indent=1)
    for key, run in out['runs'].items():
        print(key, run['iterations'], run['ionisation_energies'], run['n_cholesky'], run['ccsd_energy'])


if __name__ == '__main__':
    main()
