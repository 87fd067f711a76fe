"""Build libpyfds_amd.so (HIP, gfx950) in-tree.

    python -m pyfds_amd.build          # or pyfds_amd.build.build()

Compiles every ``pyfds_amd/csrc/*.hip`` translation unit in parallel with hipcc and links
them into ``pyfds_amd/_lib/libpyfds_amd.so``.  The library is compiled with
``-ffp-contract=off`` so the float64 kernels round every product and sum exactly where the
reference's numpy/scipy code does, and with packed-FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, '_lib')
LIB = os.path.join(LIBDIR, 'libpyfds_amd.so')
ARCH = os.environ.get('PYFDS_AMD_ARCH', 'gfx950')

FLAGS = ['-O3', '-std=c++17', '--offload-arch=' + ARCH, '-fPIC', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument']


def _hipcc():
    for cand in (os.environ.get('HIPCC'), shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found: the pyfds_amd HIP library cannot be built')


def _compile(src, obj):
    cmd = [_hipcc()] + FLAGS + ['-c', src, '-o', obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (os.path.basename(src), res.stderr))
    return obj


def build(verbose=True):
    """Compile all HIP sources and link the shared library (always rebuilds)."""
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith('.hip'))
    jobs = {}
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as pool:
        for f in srcs:
            obj = os.path.join(objdir, f[:-4] + '.o')
            jobs[pool.submit(_compile, os.path.join(CSRC, f), obj)] = f
        objs = []
        for fut in cf.as_completed(jobs):
            objs.append(fut.result())
            if verbose:
                print('compiled', jobs[fut], flush=True)
    tmp = LIB + '.tmp'
    cmd = [_hipcc(), '-shared', '--offload-arch=' + ARCH, '-o', tmp] + sorted(objs)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('link failed:\n' + res.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print('built', os.path.relpath(LIB, os.path.dirname(HERE)), flush=True)
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
