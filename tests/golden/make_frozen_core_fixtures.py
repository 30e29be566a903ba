#!/usr/bin/env python3
"""Writes the frozen-core fixtures from eT's tests/ccsd_frozen_core and tests/ccsd_dipole_fc.

tests/golden/hof_he_frozen_core.json, from tests/ccsd_frozen_core:
HOF-He / cc-pVDZ with `frozen orbitals: core` (one frozen orbital each on O and F),
Cholesky threshold 1e-11, ground state to 1e-11 (residual and energy) by DIIS, two right
EOM-CCSD roots to 1e-11.  Read as text from the input and result/ccsd_frozen_core.out:
geometry (Angstrom), E_nuc, E_HF, the number of Cholesky vectors, the number of occupied
orbitals before freezing and the number frozen, the DIIS table (energy, |omega|), E_CCSD
and the excitation energies.

tests/golden/h2o_631g_dipole_fc.json, from tests/ccsd_dipole_fc: H2O / 6-31G (geometry in
bohr), frozen core (O 1s), Cholesky threshold 1e-11, ground state and multipliers to 1e-11,
the CC mean-value dipole: E_nuc, E_HF, the number of Cholesky vectors, the number frozen,
E_CCSD and the dipole table in Debye (electronic, nuclear, total per component).

tests/golden/xe_321g_frozen_core.json, from tests/ccsd_frozen_core_xenon: Xe / 3-21G (Cartesian d,
eT's default for Pople sets: 33 AOs), 18 frozen core orbitals, Cholesky threshold 1e-12, ground
state to 1e-10, three right roots (a degenerate triple) to 1e-9: the AO count, E_HF, the number
of Cholesky vectors, the number frozen, E_CCSD and the roots.

Each value keeps the output line it came from.

Usage: python tests/golden/make_frozen_core_fixtures.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
STEM = 'tests/ccsd_frozen_core/ccsd_frozen_core'
REL = 'tests/ccsd_frozen_core/result/ccsd_frozen_core.out'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'hof_he_frozen_core.json')


def grabber(lines, rel, g, cite):
    def grab(key, pattern, conv):
        for i, ln in enumerate(lines):
            m = re.search(pattern, ln)
            if m:
                g[key], cite[key] = conv(m.group(1)), '%s:%d' % (rel, i + 1)
                return
        raise RuntimeError(pattern)
    return grab


def dipole_fc():
    stem = 'tests/ccsd_dipole_fc/ccsd_dipole_fc'
    rel = 'tests/ccsd_dipole_fc/result/ccsd_dipole_fc.out'
    inp = open(os.path.join(REF, stem + '.inp')).read()
    lines = open(os.path.join(REF, rel)).read().splitlines()
    geom = [[s[0], float(s[1]), float(s[2]), float(s[3])]
            for s in (ln.split() for ln in inp.split('basis: 6-31g')[1].split('end geometry')[0].splitlines())
            if len(s) == 4]
    g, cite = {}, {}
    grab = grabber(lines, rel, g, cite)
    grab('nuclear_repulsion', r'Nuclear repulsion energy:\s+(-?[0-9.]+)', float)
    grab('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float)
    grab('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int)
    grab('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s+(-?[0-9.]+)', float)
    grab('n_frozen_core', r'There are ([0-9]+) frozen core orbitals', int)
    k = next(i for i, ln in enumerate(lines) if 'Dipole moment in [Debye]' in ln)
    h = next(i for i in range(k, len(lines)) if 'Comp.         Electronic' in lines[i])
    g['dipole_debye'] = {}
    for q in range(3):
        s = lines[h + 2 + q].split()
        g['dipole_debye'][s[0]] = {'electronic': float(s[1]), 'nuclear': float(s[2]), 'total': float(s[3])}
    cite['dipole_debye'] = '%s:%d-%d' % (rel, h + 3, h + 5)
    out = {'source': stem + '.inp', 'basis': '6-31G', 'units': 'bohr', 'geometry_bohr': geom,
           'cholesky_threshold': 1e-11, 'residual_threshold': 1e-11, 'golden': g, 'citations': cite}
    json.dump(out, open(os.path.join(os.path.dirname(OUT), 'h2o_631g_dipole_fc.json'), 'w'), indent=1)
    print(json.dumps(g, indent=1))


def xenon():
    stem = 'tests/ccsd_frozen_core_xenon/ccsd_frozen_core_xenon'
    rel = 'tests/ccsd_frozen_core_xenon/result/ccsd_frozen_core_xenon.out'
    inp = open(os.path.join(REF, stem + '.inp')).read()
    lines = open(os.path.join(REF, rel)).read().splitlines()
    geom = [[s[0], float(s[1]), float(s[2]), float(s[3])]
            for s in (ln.split() for ln in inp.split('basis: 3-21G')[1].split('end geometry')[0].splitlines())
            if len(s) == 4]
    g, cite = {}, {}
    grab = grabber(lines, rel, g, cite)
    grab('n_ao', r'Number of atomic orbitals:\s+([0-9]+)', int)
    grab('hf_energy', r'Total energy:\s+(-?[0-9.]+)', float)
    grab('n_cholesky', r'Final number of Cholesky vectors:\s+([0-9]+)', int)
    grab('ccsd_energy', r'Final ground state energy \(a\.u\.\):\s*(-?[0-9.]+)', float)
    grab('n_frozen_core', r'There are ([0-9]+) frozen core orbitals', int)
    k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
    g['excitation_energies'] = [float(lines[k + 2 + q].split()[1]) for q in range(3)]
    cite['excitation_energies'] = '%s:%d-%d' % (rel, k + 3, k + 5)
    out = {'source': stem + '.inp', 'basis': '3-21G', 'cartesian': True, 'bohr_to_angstrom': 0.52917721092,
           'geometry_angstrom': geom, 'cholesky_threshold': 1e-12, 'residual_threshold': 1e-10,
           'es_residual_threshold': 1e-9, 'n_states': 3, 'golden': g, 'citations': cite}
    json.dump(out, open(os.path.join(os.path.dirname(OUT), 'xe_321g_frozen_core.json'), 'w'), indent=1)
    print(json.dumps(g, indent=1))


def main():
    xenon()
    dipole_fc()
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
    grab('n_o_all_electron', r'Number of occupied orbitals:\s+([0-9]+)', int)
    grab('n_frozen_core', r'There are ([0-9]+) frozen core orbitals', int)
    head = next(i for i, ln in enumerate(lines) if 'Iteration    Energy (a.u.)        |omega|' in ln)
    end = next(i for i in range(head, len(lines)) if 'Convergence criterion met' in lines[i])
    g['ccsd_iterations'] = [[int(s[0]), float(s[1]), float(s[2])]
                            for s in (ln.split() for ln in lines[head + 2:end - 1]) if len(s) == 4]
    cite['ccsd_iterations'] = '%s:%d-%d' % (REL, head + 3, end - 1)
    k = next(i for i, ln in enumerate(lines) if 'State                (Hartree)             (eV)' in ln)
    g['excitation_energies'] = [float(lines[k + 2 + q].split()[1]) for q in range(2)]
    cite['excitation_energies'] = '%s:%d-%d' % (REL, k + 3, k + 4)
    out = {'source': STEM + '.inp', 'basis': 'cc-pVDZ', 'bohr_to_angstrom': 0.52917721092,
           'geometry_angstrom': geom, 'cholesky_threshold': 1e-11, 'residual_threshold': 1e-11,
           'es_residual_threshold': 1e-11, 'n_states': 2, 'golden': g, 'citations': cite}
    json.dump(out, open(OUT, 'w'), indent=1)
    print(json.dumps({k: v for k, v in g.items() if k != 'ccsd_iterations'}, indent=1),
          len(g['ccsd_iterations']), 'DIIS iterations')


if __name__ == '__main__':
    main()
