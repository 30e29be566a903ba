"""Writes h2o_he_sto3g_dipole.json from eT's tests/ccsd_dipole (input and result).

Data only: the geometry (Angstrom), the STO-3G shells of H, O and He from
ao_basis/sto-3g.g94 (an SP shell split into its s and p shells), the thresholds
of the input and the printed reference values (E_nuc, E_HF, number of Cholesky
vectors, E_CCSD, E_corr, the CC mean-value dipole table in Debye and its
Debye <-> a.u. factor).  Run where the reference is at /root/reference (or give
its path as the first argument).
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2o_he_sto3g_dipole.json')
TEST = os.path.join(REF, 'tests', 'ccsd_dipole')


def read_basis(path, elements):
    basis = {}
    for block in open(path).read().split('****'):
        rows = [r for r in block.splitlines() if r.strip() and not r.lstrip().startswith('!')]
        if not rows or len(rows[0].split()) != 2 or rows[0].split()[0] not in elements:
            continue
        shells, i = [], 1
        while i < len(rows):
            kind, count = rows[i].split()[0], int(rows[i].split()[1])
            cols = [r.replace('D', 'E').split() for r in rows[i + 1:i + 1 + count]]
            exps = [float(c[0]) for c in cols]
            if kind == 'SP':
                shells.append({'l': 0, 'exps': exps, 'coefs': [float(c[1]) for c in cols]})
                shells.append({'l': 1, 'exps': exps, 'coefs': [float(c[2]) for c in cols]})
            else:
                shells.append({'l': 'SPD'.index(kind), 'exps': exps, 'coefs': [float(c[1]) for c in cols]})
            i += 1 + count
        basis[rows[0].split()[0]] = shells
    return basis


def grab(text, pattern):
    m = re.search(pattern, text)
    if not m:
        raise RuntimeError('pattern not found: ' + pattern)
    return m.group(1)


def main():
    inp = open(os.path.join(TEST, 'ccsd_dipole.inp')).read()
    out = open(os.path.join(TEST, 'result', 'ccsd_dipole.out')).read()
    geom_block = inp.split('geometry')[1]
    geom = []
    for ln in geom_block.splitlines():
        f = ln.split()
        if len(f) == 4 and f[0].isalpha():
            geom.append([f[0], float(f[1]), float(f[2]), float(f[3])])
    params = open(os.path.join(REF, 'src', 'various', 'parameters.F90')).read()
    bohr = float(grab(params, r'bohr_to_angstrom\s*=\s*([0-9.]+)D0'))
    table = out.split('Dipole moment in [Debye]:')[1]
    comps = {c: [float(x) for x in re.search(r'^\s+%s\s+(-?\d+\.\d+)\s+(-?\d+\.\d+)\s+(-?\d+\.\d+)\s*$' % c,
                                             table, flags=re.M).groups()] for c in 'xyz'}
    # eT's default CCSD multiplier solver (linear Davidson): its residual-norm table
    mult = out.split('Determining CC multipliers')[1].split('Convergence criterion met')[0]
    mult_rows = [float(m) for m in re.findall(r'^\s+\d+\s+(\d\.\d+E[-+]\d+)\s*$', mult, flags=re.M)]
    fixture = {
        'source': 'eT test data: tests/ccsd_dipole (ccsd_dipole.inp, result/ccsd_dipole.out), '
                  'ao_basis/sto-3g.g94, src/various/parameters.F90',
        'bohr_to_angstrom': bohr,
        'geometry_angstrom': geom,
        'basis': read_basis(os.path.join(REF, 'ao_basis', 'sto-3g.g94'), {'H', 'He', 'O'}),
        'cholesky_threshold': float(grab(inp, r'solver cholesky\s+threshold:\s*([0-9.dDeE+-]+)').replace('d', 'e')),
        'golden': {
            'nuclear_repulsion': float(grab(out, r'Nuclear repulsion energy:\s+(-?\d+\.\d+)')),
            'hf_energy': float(grab(out, r'Total energy:\s+(-?\d+\.\d+)')),
            'n_cholesky': int(grab(out, r'Final number of Cholesky vectors:\s+(\d+)')),
            'ccsd_energy': float(grab(out, r'Final ground state energy \(a\.u\.\):\s+(-?\d+\.\d+)')),
            'ccsd_correlation_energy': float(grab(out, r'Correlation energy \(a\.u\.\):\s+(-?\d+\.\d+)')),
            'debye_to_au': float(grab(table, r'Conversion factor from Debye a\.u\.:\s+(\d+\.\d+)')),
            'au_to_debye': float(grab(params, r'au_to_debye\s*=\s*([0-9.]+)D0')),
            'dipole_debye': {c: dict(zip(('electronic', 'nuclear', 'total'), comps[c])) for c in 'xyz'},
            'multiplier_davidson_residuals': mult_rows,
        },
        'citations': {
            'nuclear_repulsion': 'tests/ccsd_dipole/result/ccsd_dipole.out:186',
            'hf_energy': 'tests/ccsd_dipole/result/ccsd_dipole.out:188',
            'n_cholesky': 'tests/ccsd_dipole/result/ccsd_dipole.out:252',
            'ccsd_energy': 'tests/ccsd_dipole/result/ccsd_dipole.out:318',
            'dipole_debye': 'tests/ccsd_dipole/result/ccsd_dipole.out:434-445',
            'multiplier_davidson_residuals': 'tests/ccsd_dipole/result/ccsd_dipole.out:360-395',
            'au_to_debye': 'src/various/parameters.F90:110',
        },
    }
    with open(OUT, 'w') as f:
        json.dump(fixture, f, indent=1)


if __name__ == '__main__':
    main()
