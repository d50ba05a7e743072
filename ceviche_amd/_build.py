"""Builds the HIP library ``ceviche_amd/libceviche_amd.so`` for gfx950.

Each ``csrc/*.hip`` file is compiled to an object with hipcc in parallel,
then the objects are linked into one shared library.  Objects go to
``_build/`` at the repository root (git-ignored); the library is written
in-tree next to this file so that it travels with the source tree.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'csrc')
INC = os.path.join(ROOT, 'include')
OBJ = os.path.join(ROOT, '_build')
LIB = os.path.join(HERE, 'libceviche_amd.so')

ARCH = os.environ.get('CEVICHE_AMD_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = [
    '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC',
    # packed fp32 ops stay off: hand-free codegen only, see the build notes
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I' + INC, '-I' + SRC,
    '-Wno-unused-result', '-Wno-unused-command-line-argument',
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, ' '.join(cmd), p.stderr))
    return obj


def build(verbose=False):
    """Compile every HIP source and link the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(SRC, f) for f in os.listdir(SRC) if f.endswith('.hip'))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    # rocBLAS / rocSOLVER: batched dense LU and products of the nested-dissection factor
    cmd = [HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs + ['-lrocsolver', '-lrocblas']
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (' '.join(cmd), p.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
