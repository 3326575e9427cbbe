"""Compile the HIP sources in geepee_amd/csrc into geepee_amd/libgpx.so for gfx950.

Run as ``python -m geepee_amd.build_lib``; __graft_entry__.build() calls ``build()``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'libgpx.so')
BUILD = os.path.join(HERE, '_build')
ARCH = os.environ.get('GPX_ARCH', 'gfx950')
# -amdgpu-mfma-vgpr-form: matrix-core accumulators in the unified vector registers
# (gfx90a and later), so the compiler does not copy them to and from the accumulator
# registers around every MFMA group (measured equal or faster on every fused kernel)
FLAGS = ['-O3', '-fPIC', '-std=c++17', '--offload-arch=' + ARCH,
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-mllvm', '-amdgpu-mfma-vgpr-form']


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src).replace('.hip', '.o'))
    cmd = ['hipcc'] + FLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = '\n'.join(l for l in (r.stdout + r.stderr).splitlines()
                     if 'packed-fp32-ops' not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, errs))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + '.tmp'
    cmd = ['hipcc', '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT)
    return OUT


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
