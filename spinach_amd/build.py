"""Build the HIP library in-tree: spinach_amd/libspinach_powder.so (gfx950)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, 'libspinach_powder.so')
SOURCES = [os.path.join(HERE, 'csrc', 'spinach_powder.hip'), os.path.join(HERE, 'csrc', 'surrogate.hip')]
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '-shared', '--offload-arch=gfx950',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function', '-Wno-unused-result',
    '-I', os.path.join(ROOT, 'include'),
]


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + ['-o', LIB + '.tmp'] + SOURCES
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed to build libspinach_powder.so')
    os.replace(LIB + '.tmp', LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
