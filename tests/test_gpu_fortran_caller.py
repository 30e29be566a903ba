"""eT's ground-state loop as a Fortran program: tests/fortran/run_diis_cc_gs.F90,
compiled with AMD flang against the library's ISO_C_BINDING module
(et_amd/fortran/etgpu_interface.F90) and linked to libetgpu.so, runs
run_diis_cc_gs (diis_cc_gs_class.F90:209-315) on the committed H2O-He /
cc-pVDZ inputs of eT's tests/ccsd_gs_energy.  It must reproduce eT's
19-iteration DIIS table and E_CCSD = -79.100383481554 Eh
(ccsd_gs_energy.out:299-327) to 1e-10."""
import json
import os
import subprocess

import numpy as np
import pytest

import fortran_driver
from et_amd import workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2o_he_ccpvdz.json')


@pytest.mark.skipif(fortran_driver.compiler() is None, reason='no Fortran compiler')
def test_fortran_diis_loop_reproduces_et(tmp_path):
    g = json.load(open(GOLDEN))['golden']
    m = workloads.h2o_he_inputs()
    n_o = m['n_o']
    n_mo = m['L_mo'].shape[1]
    n_J = m['L_mo'].shape[0]
    with open(tmp_path / 'dims', 'wb') as f:
        f.write(np.array([n_o, n_mo - n_o, n_J], dtype=np.int64).tobytes())
        f.write(np.array([m['e_hf'], 1e-10], dtype=np.float64).tobytes())
    np.asfortranarray(m['L_mo']).ravel(order='F').tofile(str(tmp_path / 'L'))
    np.asfortranarray(m['h_mo']).ravel(order='F').tofile(str(tmp_path / 'h'))
    m['eps'].astype(np.float64).tofile(str(tmp_path / 'eps'))
    exe = fortran_driver.build(str(tmp_path))
    print('running', exe, flush=True)
    res = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(res.stdout[-3000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    rows = [ln.split() for ln in res.stdout.splitlines() if ln.startswith('iteration')]
    table = g['ccsd_diis_iterations']
    assert len(rows) == len(table)
    for row, (it, e_ref, norm_ref, _) in zip(rows, table):
        assert int(row[1]) == it
        assert abs(float(row[2]) - e_ref) < 1e-10, it
    e = float([ln for ln in res.stdout.splitlines() if ln.startswith('E_CCSD')][0].split()[1])
    assert abs(e - g['ccsd_energy']) < 1e-10
