"""Builds the gfx950 shared library of this package in-tree with hipcc.

The library is a single C-ABI shared object, ``hcipy_amd/libhcipy_amd.so``, compiled from
``hcipy_amd/csrc/*.hip``.  Run ``python -m hcipy_amd.build`` (or ``__graft_entry__.build()``).
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB_NAME = 'libhcipy_amd.so'
LIB_PATH = os.path.join(HERE, LIB_NAME)

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '-shared',
    '--offload-arch=gfx950',
    # Packed FP32 instructions are kept out of the device code.
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function', '-Wno-unknown-pragmas',
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, '*.hip')))


def build(verbose=True, jobs=8):
    """Compiles every source to an object in parallel, then links the shared library."""
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + '.o')
        objs.append(obj)
        cmd = [HIPCC] + [f for f in FLAGS if f != '-shared'] + ['-c', src, '-o', obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors='replace')
        text = '\n'.join(l for l in text.splitlines() if 'packed-fp32-ops' not in l)
        if p.returncode != 0:
            failed = True
            sys.stderr.write(' '.join(cmd) + '\n' + text + '\n')
        elif verbose and text.strip():
            sys.stderr.write(text + '\n')
    if failed:
        raise RuntimeError('hipcc compilation failed')
    tmp = LIB_PATH + '.tmp'
    cmd = [HIPCC] + FLAGS + objs + ['-o', tmp]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if res.returncode != 0:
        sys.stderr.write(res.stdout.decode(errors='replace'))
        raise RuntimeError('hipcc link failed')
    os.replace(tmp, LIB_PATH)
    if verbose:
        print('built', LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    build()
