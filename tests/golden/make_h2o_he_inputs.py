#!/usr/bin/env python3
"""Writes tests/golden/h2o_he_ccpvdz_inputs.npz: the molecular inputs of H2O-He / cc-pVDZ.

Runs the oracle's front end (oracle/frontend.py: integrals -> RHF -> pivoted
Cholesky at eT's threshold) on the basis and geometry in h2o_he_ccpvdz.json and
stores what a CCSD calculation consumes, so that GPU runs need no front end:
  L_tri  (n_J, n_mo (n_mo + 1) / 2)  MO Cholesky vectors, upper triangle (p <= q), column-major pairs
  h_mo   (n_mo, n_mo)                MO one-electron Hamiltonian
  eps    (n_mo,)                     RHF orbital energies
  e_hf, e_nuc, n_o
The CPU test tests/test_golden_h2o_he.py checks the file against the front end.

Usage: python tests/golden/make_h2o_he_inputs.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import frontend as fe  # noqa: E402

OUT = os.path.join(HERE, 'h2o_he_ccpvdz_inputs.npz')


def pack_upper(L):
    n = L.shape[1]
    iu = np.triu_indices(n)
    # column-major pair order: (p, q) with p <= q, q slowest
    order = np.lexsort((iu[0], iu[1]))
    return L[:, iu[0][order], iu[1][order]]


def main():
    m = fe.molecular_inputs(json.load(open(os.path.join(HERE, 'h2o_he_ccpvdz.json'))))
    np.savez_compressed(OUT, L_tri=pack_upper(m['L_mo']), h_mo=m['h_mo'], eps=m['eps'],
                        e_hf=m['e_hf'], e_nuc=m['e_nuc'], n_o=m['n_o'])
    print('wrote', OUT)


if __name__ == '__main__':
    main()
