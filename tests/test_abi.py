"""The C ABI library loads and exports every function include/etgpu.h declares.

Runs without a GPU: loading libetgpu.so does not touch the device, and no
compute entry point is called here.
"""
import ctypes
import os
import re

import pytest

from et_amd import ccsd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, 'include', 'etgpu.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(?:int|const char\*)\s+(etg_\w+)\s*\(', text)))


def test_header_declares_the_reference_interface():
    names = declared_functions()
    for required in ['etg_create', 'etg_destroy', 'etg_last_error', 'etg_set_cholesky',
                     'etg_set_oei', 'etg_set_orbital_energies', 'etg_set_amplitudes',
                     'etg_get_amplitudes', 'etg_construct_t1_cholesky', 'etg_construct_fock',
                     'etg_calculate_energy', 'etg_construct_omega', 'etg_get_orbital_differences']:
        assert required in names


@pytest.mark.skipif(not os.path.exists(ccsd.LIB_PATH), reason='library not built')
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ccsd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(ccsd.LIB_PATH), reason='library not built')
def test_python_signatures_cover_the_header():
    assert set(declared_functions()) == set(ccsd._SIGNATURES)


@pytest.mark.skipif(not os.path.exists(ccsd.LIB_PATH), reason='library not built')
def test_last_error_is_empty_initially():
    lib = ccsd.load_library()
    assert lib.etg_last_error() == b''


def test_fortran_module_binds_every_declared_function():
    """et_amd/fortran/etgpu_interface.F90 has a bind(C) interface for every function
    include/etgpu.h declares (the Fortran caller reaches the whole boundary)."""
    text = open(os.path.join(ROOT, 'et_amd', 'fortran', 'etgpu_interface.F90')).read()
    bound = set(re.findall(r"name='(etg_\w+)'", text))
    missing = sorted(set(declared_functions()) - bound)
    assert not missing, missing
