"""Builds libpirate_amd.so in-tree: host C++ (config, plan, orchestrator, C-ABI) and the
gfx950 HIP kernels, compiled with hipcc and linked into one shared library.

    python -m pirate_amd.build          (or: python pirate_amd/build.py)

All code is built with -ffp-contract=off (separately rounded multiplies and adds, which the
float32 parity with the CPU reference relies on) and with packed-FP32 instructions disabled.
"""

import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, 'build')
LIB = os.path.join(HERE, 'libpirate_amd.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('PIRATE_AMD_ARCH', 'gfx950')

COMMON = ['-O3', '-std=c++17', '-fPIC', '-ffp-contract=off', '-Wall', '-Wno-unused-function',
          '-Wno-unused-variable', '-Wno-unused-result']
HIPFLAGS = ['--offload-arch=' + ARCH, '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def sources():
    return sorted(glob.glob(os.path.join(CSRC, 'host', '*.cpp')) + glob.glob(os.path.join(CSRC, 'gpu', '*.cpp'))
                  + glob.glob(os.path.join(CSRC, 'gpu', '*.hip')) + [os.path.join(CSRC, 'capi.cpp')])


def _compile(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, '_')
    obj = os.path.join(BUILD, rel + '.o')
    cmd = [HIPCC] + COMMON + HIPFLAGS + ['-I', os.path.join(HERE, '..', 'include'), '-c', src, '-o', obj]
    if src.endswith('.cpp'):
        cmd = [HIPCC] + COMMON + ['-I', os.path.join(HERE, '..', 'include'), '-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = '\n'.join(l for l in (r.stdout + r.stderr).splitlines() if 'packed-fp32-ops' not in l)
    return src, obj, r.returncode, out


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    objs = []
    jobs = jobs or min(8, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for src, obj, rc, out in ex.map(_compile, srcs):
            if out.strip() and (verbose or rc != 0):
                print(out, file=sys.stderr)
            if rc != 0:
                raise RuntimeError(f'compilation failed: {src}')
            objs.append(obj)
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError('link failed')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
