#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_ccpvdz.json from eT's own test data files.

Reads (text only, nothing is imported or executed from the reference):
  ao_basis/cc-pvdz.g94                         basis set data for H, He, O
  tests/ccsd_gs_energy/ccsd_gs_energy.inp      geometry (Angstrom) and thresholds
  tests/ccsd_gs_energy/result/ccsd_gs_energy.out          golden energies
  tests/ccsd_right_es_energies/result/ccsd_right_es_energies.out  excitation energies
  tests/ccsd_left_es_energies/result/ccsd_left_es_energies.out    excitation energies
and the Bohr/Angstrom factor eT uses (src/various/parameters.F90:105).

Usage: python tests/golden/make_h2o_he_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_ccpvdz.json')


def read_basis(path, elements):
    text = open(path).read().split('****')
    basis = {}
    for block in text:
        lines = [ln for ln in block.strip().splitlines() if ln.strip() and not ln.startswith('!')]
        if not lines:
            continue
        head = lines[0].split()
        if len(head) != 2 or head[0] not in elements:
            continue
        shells = []
        k = 1
        while k < len(lines):
            kind, nprim, _ = lines[k].split()
            nprim = int(nprim)
            exps, coefs = [], []
            for q in range(nprim):
                e, c = lines[k + 1 + q].replace('D', 'E').split()[:2]
                exps.append(float(e))
                coefs.append(float(c))
            shells.append({'l': 'SPDFG'.index(kind), 'exps': exps, 'coefs': coefs})
            k += 1 + nprim
        basis[head[0]] = shells
    return basis


def grab(path, pattern, group=1):
    m = re.search(pattern, open(path).read())
    if not m:
        raise RuntimeError('pattern not found in %s: %s' % (path, pattern))
    return m.group(group)


def main():
    inp = os.path.join(REF, 'tests/ccsd_gs_energy/ccsd_gs_energy.inp')
    out = os.path.join(REF, 'tests/ccsd_gs_energy/result/ccsd_gs_energy.out')
    geom = []
    in_geom = False
    for ln in open(inp):
        if ln.strip().startswith('geometry'):
            in_geom = True
            continue
        if ln.strip().startswith('end geometry'):
            break
        if in_geom and not ln.strip().startswith('basis'):
            sym, x, y, z = ln.split()
            geom.append([sym, float(x), float(y), float(z)])
    bohr = float(grab(os.path.join(REF, 'src/various/parameters.F90'),
                      r'bohr_to_angstrom\s*=\s*([0-9.]+)D0'))
    basis = read_basis(os.path.join(REF, 'ao_basis/cc-pvdz.g94'), {'H', 'He', 'O'})
    iters = re.findall(r'^\s+(\d+)\s+(-\d+\.\d+)\s+(\d\.\d+E[-+]\d+)\s+(\d\.\d+E[-+]\d+)\s*$',
                       open(out).read().split(':: CCSD wavefunction')[1], flags=re.M)
    right = os.path.join(REF, 'tests/ccsd_right_es_energies/result/ccsd_right_es_energies.out')
    left = os.path.join(REF, 'tests/ccsd_left_es_energies/result/ccsd_left_es_energies.out')

    def es(path):
        block = open(path).read().split('State                (Hartree)             (eV)')[-1]
        return [float(v) for v in re.findall(r'^\s+\d+\s+(\d+\.\d+)\s+\d+\.\d+\s*$', block, flags=re.M)]

    fixture = {
        'source': 'eT 1.9 test data: tests/ccsd_gs_energy (input, result/ccsd_gs_energy.out), '
                  'ao_basis/cc-pvdz.g94, src/various/parameters.F90:105',
        'bohr_to_angstrom': bohr,
        'geometry_angstrom': geom,
        'basis': basis,
        'cholesky_threshold': 1e-12,
        'n_o': 6, 'n_v': 23,
        'golden': {
            'nuclear_repulsion': float(grab(out, r'Nuclear repulsion energy:\s+(-?\d+\.\d+)')),
            'hf_energy': float(grab(out, r'Total energy:\s+(-?\d+\.\d+)')),
            'n_cholesky': int(grab(out, r'Final number of Cholesky vectors:\s+(\d+)')),
            'ccsd_energy': float(grab(out, r'Final ground state energy \(a\.u\.\):\s+(-?\d+\.\d+)')),
            'ccsd_correlation_energy': float(grab(out, r'Correlation energy \(a\.u\.\):\s+(-?\d+\.\d+)')),
            'ccsd_diis_iterations': [[int(a), float(b), float(c), float(d)] for a, b, c, d in iters],
            'right_excitation_energies': es(right),
            'left_excitation_energies': es(left),
        },
        'citations': {
            'nuclear_repulsion': 'tests/ccsd_gs_energy/result/ccsd_gs_energy.out:185',
            'hf_energy': 'tests/ccsd_gs_energy/result/ccsd_gs_energy.out (RHF summary)',
            'ccsd_energy': 'tests/ccsd_gs_energy/result/ccsd_gs_energy.out:325-327',
            'ccsd_diis_iterations': 'tests/ccsd_gs_energy/result/ccsd_gs_energy.out:299-321',
            'right_excitation_energies': 'tests/ccsd_right_es_energies/result/ccsd_right_es_energies.out:749-750',
            'left_excitation_energies': 'tests/ccsd_left_es_energies/result/ccsd_left_es_energies.out:758-759',
        },
    }
    with open(OUT, 'w') as f:
        json.dump(fixture, f, indent=1)
    print('wrote', OUT)


if __name__ == '__main__':
    main()
