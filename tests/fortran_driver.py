"""Builds tests/fortran/run_diis_cc_gs.F90 against the library's Fortran module
(et_amd/fortran/etgpu_interface.F90) with AMD flang and links libetgpu.so."""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE = os.path.join(ROOT, 'et_amd', 'fortran', 'etgpu_interface.F90')
DRIVER = os.path.join(ROOT, 'tests', 'fortran', 'run_diis_cc_gs.F90')
LIBDIR = os.path.join(ROOT, 'et_amd')


def compiler():
    for c in (shutil.which('amdflang'), '/opt/rocm/bin/amdflang', shutil.which('flang'),
              '/opt/rocm/lib/llvm/bin/flang'):
        if c and os.path.exists(c):
            return c
    return None


def build(workdir):
    """Compiles the module and the driver into `workdir`; returns the executable."""
    fc = compiler()
    if fc is None:
        raise RuntimeError('no Fortran compiler (amdflang) found')
    exe = os.path.join(workdir, 'run_diis_cc_gs')
    subprocess.run([fc, '-c', MODULE, '-J', workdir, '-o', os.path.join(workdir, 'etgpu_interface.o')],
                   check=True, capture_output=True, text=True)
    subprocess.run([fc, DRIVER, os.path.join(workdir, 'etgpu_interface.o'), '-I', workdir, '-J', workdir,
                    '-L' + LIBDIR, '-letgpu', '-Wl,-rpath,' + LIBDIR, '-o', exe],
                   check=True, capture_output=True, text=True)
    return exe
