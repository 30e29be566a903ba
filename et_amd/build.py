"""Builds libetgpu.so (HIP kernels + C ABI) for gfx950, in place.

    python -m et_amd.build            # compile et_amd/libetgpu.so

Compiles each translation unit of et_amd/csrc with hipcc in parallel and
links them into one shared library next to this file.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libetgpu.so')
OBJDIR = os.path.join(HERE, '_build')
ARCH = os.environ.get('ETG_OFFLOAD_ARCH', 'gfx950')

# packed-FP32 instructions are disabled for all device code.  Matrix-core
# accumulators are kept in architectural VGPRs (the VGPR form of MFMA): with
# AGPR accumulators the compiler copies them out and back on every k step.
HIPFLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-mllvm', '-amdgpu-mfma-vgpr-form=1',
            '-Wall', '-Wno-unused-result']


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found')


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


# The split-bf16 product keeps its 256 FP32 accumulators per lane in the
# accumulation registers (the default MFMA form): its fragments fill the VGPRs.
FILE_FLAGS = {'gemm_bf16.hip': ['-mllvm', '-amdgpu-mfma-vgpr-form=0']}


def _flags(src):
    extra = FILE_FLAGS.get(os.path.basename(src))
    if not extra:
        return HIPFLAGS
    flags = list(HIPFLAGS)
    i = flags.index('-amdgpu-mfma-vgpr-form=1')
    del flags[i - 1:i + 1]
    return flags + extra


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + '.o')
    cmd = [hipcc(), *_flags(src), '-c', src, '-o', obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [ln for ln in (res.stdout + res.stderr).splitlines()
            if 'packed-fp32-ops' not in ln]
    if res.returncode != 0:
        raise RuntimeError('compile failed: %s\n%s' % (' '.join(cmd), '\n'.join(msgs)))
    return obj, msgs


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for _, msgs in results:
            for m in msgs:
                print(m, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + '.tmp'
    cmd = [hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp, *objs,
           '-L/opt/rocm/lib', '-lrocsolver', '-lrocblas', '-Wl,-rpath,/opt/rocm/lib']
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('link failed: %s\n%s%s' % (' '.join(cmd), res.stdout, res.stderr))
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build())
