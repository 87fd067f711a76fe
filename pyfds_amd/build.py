"""Build libpyfds_amd.so (HIP, gfx950) in-tree.

    python -m pyfds_amd.build          # or pyfds_amd.build.build()

Compiles every ``pyfds_amd/csrc/*.hip`` translation unit in parallel with hipcc (the
streaming kernels' ``fds_stream_tu.hip`` once per numeric form and streaming shape) and links
them into ``pyfds_amd/_lib/libpyfds_amd.so``.  The library is compiled with
``-ffp-contract=off`` so the float64 kernels round every product and sum exactly where the
reference's numpy/scipy code does, and with packed-FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import re
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


STREAM_TU = 'fds_stream_tu.hip'


def _num_stream_shapes():
    with open(os.path.join(CSRC, 'fds_kernels.h')) as f:
        m = re.search(r'kNumStream\s*=\s*(\d+)', f.read())
    if not m:
        raise RuntimeError('kNumStream not found in fds_kernels.h')
    return int(m.group(1))


def _hipcc():
    for cand in (os.environ.get('HIPCC'), shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found: the pyfds_amd HIP library cannot be built')


def _compile(src, obj, defines=()):
    cmd = [_hipcc()] + FLAGS + ['-D' + d for d in defines] + ['-c', src, '-o', obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (os.path.basename(src), res.stderr))
    return obj


def build(verbose=True):
    """Compile all HIP sources and link the shared library (always rebuilds)."""
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith('.hip') and f != STREAM_TU)
    # (source, object, defines): the streaming kernels once per form and shape
    units = [(f, f[:-4] + '.o', ()) for f in srcs]
    for form in range(3):
        for shape in range(_num_stream_shapes()):
            units.append((STREAM_TU, 'fds_stream_f%d_s%d.o' % (form, shape),
                          ('FDS_FORM=%d' % form, 'FDS_SHAPE=%d' % shape)))
    # longest first: the streaming units dominate
    units.sort(key=lambda u: u[0] != STREAM_TU)
    jobs = {}
    with cf.ThreadPoolExecutor(max_workers=min(os.cpu_count() or 8, 16, len(units))) as pool:
        for f, o, d in units:
            obj = os.path.join(objdir, o)
            jobs[pool.submit(_compile, os.path.join(CSRC, f), obj, d)] = o
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
