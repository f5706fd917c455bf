"""Build the native parts of quicktex_amd in-tree.

    python -m quicktex_amd.build            # libqtx.so + _quicktex extension
    python -m quicktex_amd.build --verbose

Outputs (next to this file):
  libqtx.so          the C ABI of include/qtx.h: HIP kernels for gfx950 + host runtime
  _quicktex*.so      the pybind11 module behind quicktex_amd's Python API (links libqtx.so)

Every HIP translation unit is compiled with -ffp-contract=off (the encoder's
FP32 arithmetic must round like the reference's) and without packed FP32 ops.
"""
import argparse
import os
import shlex
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('QTX_OFFLOAD_ARCH', 'gfx950')

HIP_FLAGS = [
    '-std=c++17', '-O3', '-fPIC', f'--offload-arch={ARCH}', '-ffp-contract=off',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function', '-Wno-unknown-pragmas',
    '-I' + os.path.join(ROOT, 'include'), '-I' + CSRC,
]
HIP_SOURCES = ['kernels.hip', 'bc1_ideal.hip', 'bc1_ideal_round.hip', 'bc1_nvidia.hip', 'bc1_amd.hip', 'resample.hip', 'capi.cpp', 'qtx_tables.cpp', 'batch.cpp', 'server.cpp', 'mips.cpp', 'staging.cpp']


def _run(cmd, verbose):
    if verbose:
        print(' '.join(shlex.quote(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature name; that warning is noise
    err = '\n'.join(l for l in r.stderr.splitlines() if 'packed-fp32-ops' not in l)
    if r.returncode != 0:
        raise RuntimeError('command failed: %s\n%s%s' % (' '.join(cmd), r.stdout, err))
    if verbose and err.strip():
        print(err)


def _obj(src):
    return os.path.join(BUILD, os.path.splitext(src)[0] + '.o')


def build_libqtx(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    for src in HIP_SOURCES:
        lang = [] if src.endswith('.hip') else ['-x', 'hip']
        cmds.append([HIPCC] + HIP_FLAGS + lang + ['-c', os.path.join(CSRC, src), '-o', _obj(src)])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    out = os.path.join(HERE, 'libqtx.so')
    _run([HIPCC, '-shared', '-fPIC', f'--offload-arch={ARCH}', '-o', out] + [_obj(s) for s in HIP_SOURCES], verbose)
    return out


def build_bindings(verbose=False):
    import pybind11
    suffix = sysconfig.get_config_var('EXT_SUFFIX')
    out = os.path.join(HERE, '_quicktex' + suffix)
    cmd = ['g++', '-std=c++17', '-O2', '-fPIC', '-shared', '-fvisibility=hidden',
           '-I' + pybind11.get_include(), '-I' + sysconfig.get_paths()['include'],
           '-I' + os.path.join(ROOT, 'include'),
           os.path.join(CSRC, '_bindings.cpp'), '-o', out,
           '-L' + HERE, '-lqtx', '-Wl,-rpath,$ORIGIN']
    _run(cmd, verbose)
    return out


def build(verbose=False):
    lib = build_libqtx(verbose)
    ext = build_bindings(verbose) if os.path.exists(os.path.join(CSRC, '_bindings.cpp')) else None
    return lib, ext


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--verbose', action='store_true')
    ap.add_argument('--lib-only', action='store_true')
    a = ap.parse_args()
    if a.lib_only:
        print(build_libqtx(a.verbose))
    else:
        print(build(a.verbose))


if __name__ == '__main__':
    sys.exit(main())
