"""ctypes binding of libpyfds_amd.so (declared in include/pyfds_amd.h).

The library is the only compute path of pyfds_amd: if it is missing or cannot be loaded,
every simulation call raises :class:`NativeError` -- there is no CPU fallback.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'libpyfds_amd.so')

# models / forms / codes, mirroring include/pyfds_amd.h
ACOUSTIC1D, ACOUSTIC2D, ACOUSTIC_AXI = 1, 2, 3
THERMAL1D, THERMAL2D, THERMAL_AXI = 4, 5, 6
IDEALGAS1D, ACOUSTIC2ND1D = 7, 8
FORMS = {'fp64': 0, 'fp32': 1, 'fp32_kahan': 2}
E_ARG, E_HIP, E_INDEX, E_STATE, E_NONFINITE = -1, -2, -3, -4, -5
NUM_COEFF, NUM_SCALARS = 8, 8

EXPORTS = ('fds_create', 'fds_destroy', 'fds_set_scalars', 'fds_set_coeff', 'fds_commit',
           'fds_set_state', 'fds_get_state', 'fds_set_boundaries', 'fds_set_outputs',
           'fds_run', 'fds_run_async', 'fds_run_take', 'fds_prepare_run', 'fds_read_outputs', 'fds_set_tile', 'fds_tile_count', 'fds_tile_shape',
           'fds_tile_engine', 'fds_tile_traits', 'fds_set_strip_height', 'fds_engine', 'fds_steps_per_launch', 'fds_set_line_shape', 'fds_set_flow',
           'fds_couple_loss_run',
           'fds_last_run_ms', 'fds_stream_copy_bench', 'fds_stream_pattern_bench', 'fds_batch_create', 'fds_batch_run',
           'fds_batch_run_async', 'fds_batch_take', 'fds_batch_steps_per_launch',
           'fds_batch_last_run_ms', 'fds_batch_destroy', 'fds_batch_set_schedule',
           'fds_set_learned_boundary',
           'fds_lbc_conv_fwd', 'fds_lbc_conv_bwd_data', 'fds_lbc_conv_bwd_weight',
           'fds_lbc_reduce_weight', 'fds_lbc_mse', 'fds_lbc_adam', 'fds_lbc_pack',
           'fds_lbc_gather', 'fds_lbc_band_rms', 'fds_window_sq_diff', 'fds_window_sq_diff_static',
           'fds_output_device_ptr',
           'fds_set_nonfinite_check', 'fds_coupling_create', 'fds_coupling_set_material',
           'fds_coupling_set_scalars', 'fds_coupling_state', 'fds_coupling_run',
           'fds_coupling_destroy', 'fds_get_coeff', 'fds_last_error', 'fds_version')


class NativeError(RuntimeError):
    """The HIP library is missing, failed to load, or a device call failed."""


def _preload_hip_runtime():
    """Make the HIP runtime that PyTorch-ROCm ships (when it is installed) the process's
    libamdhip64 before this library pulls in the system one, without importing torch.

    Both carry the soname libamdhip64.so.7; whichever is loaded first serves the whole process,
    and torch reports no GPU when the other one is already loaded.  Loading torch's file by
    path first keeps a later ``import torch`` (the learned boundary, user code) working in
    the same process; without torch the system runtime is used."""
    import importlib.util
    import sys
    if 'torch' in sys.modules:
        return
    try:
        spec = importlib.util.find_spec('torch')
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], 'lib', 'libamdhip64.so')
    if os.path.exists(path):
        ct.CDLL(path, mode=ct.RTLD_GLOBAL)


_lib = None

_p = ct.c_void_p
_i64 = ct.c_int64
_i32 = ct.c_int32
_dp = ct.POINTER(ct.c_double)
_lp = ct.POINTER(ct.c_int64)
_ip = ct.POINTER(ct.c_int32)


def load():
    """Load and type the library once; raise NativeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError('pyfds_amd HIP library not built (%s missing); run '
                          '`python -m pyfds_amd.build`' % LIB_PATH)
    try:
        _preload_hip_runtime()
        lib = ct.CDLL(LIB_PATH)
    except OSError as exc:
        raise NativeError('cannot load %s: %s' % (LIB_PATH, exc)) from exc
    sig = {
        'fds_create': (ct.c_int, [ct.POINTER(_p), ct.c_int, ct.c_int, _i64, _i64, ct.c_int]),
        'fds_destroy': (ct.c_int, [_p]),
        'fds_set_scalars': (ct.c_int, [_p, _dp, ct.c_int]),
        'fds_set_coeff': (ct.c_int, [_p, ct.c_int, _dp, _i64]),
        'fds_commit': (ct.c_int, [_p]),
        'fds_set_state': (ct.c_int, [_p, ct.c_int, _dp, _i64]),
        'fds_get_state': (ct.c_int, [_p, ct.c_int, _dp, _i64]),
        'fds_set_boundaries': (ct.c_int, [_p, ct.c_int, _i64, _lp, _ip, _i64, _dp, _lp, _lp,
                                          _lp, _i64, _dp]),
        'fds_set_outputs': (ct.c_int, [_p, ct.c_int, _i64, _lp]),
        'fds_run': (ct.c_int, [_p, _i64, _i64, ct.c_int]),
        'fds_run_async': (ct.c_int, [_p, _i64, _i64, ct.c_int]),
        'fds_run_take': (ct.c_int, [_p, _dp, _dp, _dp]),
        'fds_prepare_run': (ct.c_int, [_p, _i64, _i64, ct.c_int]),
        'fds_read_outputs': (ct.c_int, [_p, ct.c_int, _i64, _i64, _dp]),
        'fds_set_tile': (ct.c_int, [_p, ct.c_int]),
        'fds_last_run_ms': (ct.c_int, [_p, _dp]),
        'fds_stream_copy_bench': (ct.c_int, [_i64, ct.c_int, ct.c_int, _dp, _dp]),
        'fds_stream_pattern_bench': (ct.c_int, [ct.c_int] * 8 + [_dp, _dp]),
        'fds_batch_create': (ct.c_int, [ct.POINTER(_p), ct.POINTER(_p), ct.c_int, ct.c_int]),
        'fds_batch_run': (ct.c_int, [_p, _i64, _i64, ct.c_int]),
        'fds_batch_run_async': (ct.c_int, [_p, _i64, _i64, ct.c_int]),
        'fds_batch_take': (ct.c_int, [_p, ct.POINTER(_p)]),
        'fds_batch_last_run_ms': (ct.c_int, [_p, _dp]),
        'fds_batch_destroy': (ct.c_int, [_p]),
        'fds_batch_set_schedule': (ct.c_int, [_p, ct.c_int]),
        'fds_set_learned_boundary': (ct.c_int, [_p, _p, _p, ct.c_int, _p]),
        'fds_lbc_conv_fwd': (ct.c_int, [_p, _p, _p, _p, _p] + [ct.c_int] * 5 + [_p]),
        'fds_lbc_conv_bwd_data': (ct.c_int, [_p, _p, _p, _p, _p] + [ct.c_int] * 5 + [_p]),
        'fds_lbc_conv_bwd_weight': (ct.c_int, [_p, _p, _p] + [ct.c_int] * 5 + [_p]),
        'fds_lbc_reduce_weight': (ct.c_int, [_p, _p, _p] + [ct.c_int] * 3 + [_p]),
        'fds_lbc_mse': (ct.c_int, [_p] * 7 + [ct.c_int] * 4 + [ct.c_float, _p]),
        'fds_lbc_adam': (ct.c_int, [_p] * 4 + [ct.c_int] + [ct.c_float] * 4 + [ct.c_int, _p]),
        'fds_lbc_pack': (ct.c_int, [_p] * 3 + [ct.c_int] * 4 + [_p]),
        'fds_lbc_gather': (ct.c_int, [_p] * 7 + [ct.c_int] * 5 + [_p, _p]),
        'fds_lbc_band_rms': (ct.c_int, [_p, _p] + [ct.c_int] * 4 + [_p]),
        'fds_window_sq_diff': (ct.c_int, [_p, ct.c_int, _p, ct.c_int] + [ct.c_int] * 6
                               + [_dp, _dp]),
        'fds_window_sq_diff_static': (ct.c_int, [_p, ct.c_int, _p, ct.c_int] + [ct.c_int] * 6
                                      + [_dp, _dp]),
        'fds_output_device_ptr': (ct.c_int, [_p, ct.c_int, ct.POINTER(_p), ct.POINTER(_i64),
                                             ct.POINTER(_i64)]),
        'fds_tile_count': (ct.c_int, []),
        'fds_tile_engine': (ct.c_int, [ct.c_int]),
        'fds_tile_traits': (ct.c_int, [ct.c_int, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                                       ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]),
        'fds_set_strip_height': (ct.c_int, [_p, ct.c_int]),
        'fds_engine': (ct.c_int, [_p]),
        'fds_steps_per_launch': (ct.c_int, [_p]),
        'fds_batch_steps_per_launch': (ct.c_int, [_p]),
        'fds_set_line_shape': (ct.c_int, [_p, ct.c_int, ct.c_int, ct.c_int]),
        'fds_couple_loss_run': (ct.c_int, [_p, _p, _dp, _dp, ct.c_double, ct.c_int64, ct.c_int,
                                           _dp, ct.c_int64, ct.c_int64]),
        'fds_set_flow': (ct.c_int, [_p, _lp, ct.c_int64]),
        'fds_set_nonfinite_check': (ct.c_int, [_p, ct.c_int]),
        'fds_coupling_create': (ct.c_int, [ct.POINTER(_p), ct.POINTER(_p), ct.c_int, _p,
                                           ct.c_int]),
        'fds_coupling_set_material': (ct.c_int, [_p, ct.c_int, ct.c_int, _dp, _i64]),
        'fds_coupling_set_scalars': (ct.c_int, [_p, ct.c_int, _dp, ct.c_int]),
        'fds_coupling_state': (ct.c_int, [_p, ct.c_int, ct.c_int, _dp, _i64]),
        'fds_coupling_run': (ct.c_int, [_p, _i64, _i64, _dp, _i64, ct.POINTER(_i64)]),
        'fds_coupling_destroy': (ct.c_int, [_p]),
        'fds_get_coeff': (ct.c_int, [_p, ct.c_int, _dp, _i64]),
        'fds_tile_shape': (ct.c_int, [ct.c_int, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]),
        'fds_last_error': (ct.c_char_p, []),
        'fds_version': (ct.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def error_message():
    """The library's last error message on this thread."""
    return load().fds_last_error().decode(errors='replace')


def check(rc):
    """Turn a negative return code into the reference's exception types."""
    if rc >= 0:
        return rc
    msg = load().fds_last_error().decode(errors='replace')
    if rc == E_INDEX:
        raise IndexError(msg)
    if rc == E_ARG:
        raise ValueError(msg)
    if rc == E_NONFINITE:
        raise FloatingPointError(msg)
    raise NativeError('pyfds_amd device call failed (%d): %s' % (rc, msg))


def dptr(a):
    return a.ctypes.data_as(_dp)


def lptr(a):
    return a.ctypes.data_as(_lp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def tile_shapes():
    lib = load()
    out = []
    for i in range(lib.fds_tile_count()):
        tx, ty = ct.c_int(), ct.c_int()
        check(lib.fds_tile_shape(i, ct.byref(tx), ct.byref(ty)))
        out.append((tx.value, ty.value))
    return out


def tile_table():
    """Every shape of the step kernels: dicts with index, engine, tx, ty, steps (per launch),
    nontemporal, prefetch (rows loaded ahead by a streaming workgroup) and lds_ring (those
    rows go straight to LDS); ``single_only``: the shape serves single lossless fields (no
    batches, no lossy fields)."""
    lib = load()
    out = []
    for i, ((tx, ty), eng) in enumerate(zip(tile_shapes(), tile_engines())):
        st, nt, pf, rg = ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int()
        check(lib.fds_tile_traits(i, ct.byref(st), ct.byref(nt), ct.byref(pf), ct.byref(rg)))
        out.append(dict(index=i, engine=eng, tx=tx, ty=ty, steps=st.value,
                        nontemporal=bool(nt.value), prefetch=pf.value, lds_ring=bool(rg.value),
                        single_only=pf.value > 1))
    return out


ENGINES = ('tile', 'stream', 'line', 'nonlinear')


def engine_of(ctx):
    """Engine that runs a context's steps: 'tile', 'stream', 'line' (resident 1D) or
    'nonlinear'."""
    code = load().fds_engine(ctx)
    check(code if code < 0 else 0)
    return ENGINES[code]


def tile_engines():
    """Per shape index: 'tile' (LDS tiles) or 'stream' (row streaming, lossless 2D only)."""
    lib = load()
    return ['stream' if lib.fds_tile_engine(i) == 1 else 'tile'
            for i in range(lib.fds_tile_count())]
