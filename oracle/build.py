"""Builds the oracle's C helpers (TEST INFRASTRUCTURE ONLY):

    python -m oracle.build      # oracle/_build/libladder_pm.so

gcc -O3 -fopenmp, one shared library per C file; called by __graft_entry__.build().
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, '_build')
LIB = os.path.join(OUT, 'libladder_pm.so')


def build():
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(HERE, 'ladder_pm.c')
    tmp = LIB + '.tmp'
    subprocess.run(['gcc', '-O3', '-march=x86-64-v2', '-fopenmp', '-shared', '-fPIC', '-std=c99', src, '-o', tmp],
                   check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build())
