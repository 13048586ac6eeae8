"""ctypes binding of ``libpyrecon_amd.so`` (the HIP library, built in-tree).

The library is REQUIRED: there is no CPU fallback.  Loading fails with an
:class:`ImportError` that says how to build it.  Every C call returns an int
status; :func:`check` turns a negative status into the matching exception
(``MeshError`` for the errors the reference raises as MeshError,
pyrecon/mesh.py:22-24).
"""
import ctypes
import os

import numpy as np

# PYRECON_AMD_LIBRARY may name another build of the same library (e.g. a variant built with
# different kernel options for a comparison); by default the in-tree build is loaded.
LIB_PATH = os.environ.get('PYRECON_AMD_LIBRARY') or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                    'libpyrecon_amd.so')

OK = 0
ERR_ARG = -1
ERR_RANGE = -2
ERR_NONFINITE = -3
ERR_HIP = -4
ERR_NOMEM = -5
ERR_FFT = -6
ERR_CAPACITY = -7
ERR_STATE = -8

F32, F64 = 4, 8
FIELD_DISP, FIELD_RSD, FIELD_DISP_RSD, FIELD_VALUE = 0, 1, 2, 3
TIMER_JACOBI_FINEST, TIMER_PAINT_GATHER, TIMER_READ_SHIFTS, TIMER_PAIR_FINEST, TIMER_FFT_PASS = 0, 1, 2, 3, 4
TUNE_PAIRED_SWEEPS = 0
FFT_COMPLEX, FFT_REAL_IN, FFT_REAL_OUT = 0, 1, 2
SPEC_PLANE_PARALLEL, SPEC_INV_K2, SPEC_LOS_SQ, SPEC_KA_KB, SPEC_GRAD, SPEC_I_KA = 0, 1, 2, 3, 4, 5

WS_MULTIGRID, WS_PAINT, WS_FFT, WS_MIXED = 1, 2, 4, 8
OP_MUL, OP_DIV, OP_RDIV = 0, 1, 2
WS_ALL = 15


class MeshError(Exception):
    """Exception raised when issue with mesh (same role as the reference's MeshError)."""


class DeviceError(RuntimeError):
    """HIP or allocation failure reported by the library."""


def dtype_code(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32: return F32
    if dtype == np.float64: return F64
    raise TypeError('only float32 and float64 are supported, got {}'.format(dtype))


_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i = ctypes.c_int
_d = ctypes.c_double
_f = ctypes.c_float
_szp = ctypes.POINTER(ctypes.c_size_t)


def _signatures():
    sig = {
        'prh_last_error': (ctypes.c_char_p, []),
        'prh_version': (_i, []),
        'prh_device_count': (_i, [ctypes.POINTER(_i)]),
        'prh_set_device': (_i, [_i]),
        'prh_get_device_name': (_i, [ctypes.c_char_p, _sz]),
        'prh_synchronize': (_i, [_vp]),
        'prh_stream_create': (_i, [ctypes.POINTER(_vp)]),
        'prh_stream_destroy': (_i, [_vp]),
        'prh_malloc': (_i, [ctypes.POINTER(_vp), _sz]),
        'prh_free': (_i, [_vp]),
        'prh_host_alloc': (_i, [ctypes.POINTER(_vp), _sz]),
        'prh_host_free': (_i, [_vp]),
        'prh_memcpy_htod': (_i, [_vp, _vp, _sz, _vp]),
        'prh_memcpy_dtoh': (_i, [_vp, _vp, _sz, _vp]),
        'prh_memcpy_htod_staged': (_i, [_vp, _vp, _sz, _vp]),
        'prh_memcpy_dtoh_staged': (_i, [_vp, _vp, _sz, _vp]),
        'prh_memcpy_dtod': (_i, [_vp, _vp, _sz, _vp]),
        'prh_memset': (_i, [_vp, _i, _sz, _vp]),
        'prh_mem_info': (_i, [_szp, _szp]),
        'prh_event_create': (_i, [ctypes.POINTER(_vp)]),
        'prh_event_destroy': (_i, [_vp]),
        'prh_event_record': (_i, [_vp, _vp]),
        'prh_event_synchronize': (_i, [_vp]),
        'prh_stream_wait_event': (_i, [_vp, _vp]),
        'prh_event_elapsed_ms': (_i, [_vp, _vp, ctypes.POINTER(_f)]),
        'prh_workspace_create': (_i, [ctypes.POINTER(_vp), _szp, _i, _sz, ctypes.c_uint]),
        'prh_workspace_destroy': (_i, [_vp]),
        'prh_workspace_bytes': (_i, [_vp, _szp]),
        'prh_workspace_footprint': (_i, [_szp, _i, _sz, ctypes.c_uint, _szp]),
        'prh_workspace_reserve_paint': (_i, [_vp, _sz]),
        'prh_shift_algebra_f32': (_i, [_vp, _i, _vp, _vp, _vp, _sz, _d, _vp, _i, _i, _vp, _i, _vp]),
        'prh_shift_algebra_f64': (_i, [_vp, _i, _vp, _vp, _vp, _sz, _d, _vp, _i, _i, _vp, _i, _vp]),
        'prh_timer_enable': (_i, [_i, _i]),
        'prh_set_tuning': (_i, [_i, _i]),
        'prh_jacobi_fast_f32': (_i, [_vp, _vp, _vp, _szp, _vp, _vp, _f, _f, _i, _vp, _i, _vp]),
        'prh_timer_reset': (_i, [_i]),
        'prh_timer_read': (_i, [_i, ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_d)]),
        'prh_copy_kernel': (_i, [_vp, _vp, _sz, _vp]),
        'prh_fmg_mixed': (_i, [_vp, _vp, _vp, _szp, _vp, _vp, _d, _d, _i, _i, _vp, _i, _i, _vp, _vp]),
        'prh_synth_points': (_i, [_vp, _vp, _i, _sz, ctypes.c_uint64, _vp, _vp, _vp, _i, _d, _d, _d, _vp]),
        'prh_fmg_mixed_group': (_i, [_vp, _i, _vp, _vp, _szp, _vp, _vp, _vp, _d, _i, _i, _vp, _i, _i, _vp]),
        'prh_residual_norm_f64': (_i, [_vp, _vp, _vp, _szp, _vp, _vp, _d, _vp, ctypes.POINTER(_d), _vp]),
        'prh_paint_mixed_f64': (_i, [_vp, _vp, _szp, _vp, _vp, _vp, _i, _vp, _i, _sz, _i, _vp]),
        'prh_fmg_noalloc_f64': (_i, [_vp, _vp, _szp, _vp, _vp, _d, _d, _i, _i, _vp, _vp]),
    }
    for suffix, real in [('f32', _f), ('f64', _d)]:
        sig.update({
            'prh_assign_cic_' + suffix: (_i, [_vp, _vp, _szp, _vp, _vp, _sz, _vp]),
            'prh_read_cic_' + suffix: (_i, [_vp, _szp, _vp, _vp, _sz, _vp]),
            'prh_read_finite_difference_cic_' + suffix: (_i, [_vp, _szp, _vp, _vp, _vp, _sz, _vp]),
            'prh_jacobi_' + suffix: (_i, [_vp, _vp, _vp, _szp, _vp, _vp, real, real, _i, _vp, _vp]),
            'prh_residual_' + suffix: (_i, [_vp, _vp, _vp, _szp, _vp, _vp, real, _vp, _vp]),
            'prh_prolong_' + suffix: (_i, [_vp, _vp, _szp, _vp]),
            'prh_reduce_' + suffix: (_i, [_vp, _vp, _szp, _vp]),
            'prh_vcycle_' + suffix: (_i, [_vp, _vp, _vp, _szp, _vp, _vp, real, real, _i, _vp, _vp]),
            'prh_fmg_' + suffix: (_i, [_vp, _vp, _vp, _szp, _vp, _vp, real, real, _i, _i, _vp, _vp]),
            'prh_paint_' + suffix: (_i, [_vp, _vp, _szp, _vp, _vp, _vp, _i, _vp, _i, _sz, _i, _vp]),
            'prh_smooth_gaussian_' + suffix: (_i, [_vp, _vp, _szp, _vp, _vp, _vp]),
            'prh_mesh_sums_' + suffix: (_i, [_vp, _vp, _vp, _sz, _vp, _vp]),
            'prh_density_contrast_' + suffix: (_i, [_vp, _vp, _vp, _sz, real, real, real, real, _vp]),
            'prh_fmg_group_' + suffix: (_i, [_vp, _i, _vp, _vp, _szp, _vp, _vp, _vp, real, _i, _i, _vp, _vp]),
            'prh_martin_sums_' + suffix: (_i, [_vp, _vp, _vp, _sz, real, _vp, _vp]),
            'prh_martin_contrast_' + suffix: (_i, [_vp, _vp, _vp, _vp, _sz, real, real, real, _vp, _vp]),
            'prh_subtract_nonzero_mean_' + suffix: (_i, [_vp, _sz, real, _vp]),
            'prh_mesh_scalar_op_' + suffix: (_i, [_vp, _vp, _sz, _i, _d, _vp]),
            'prh_fft3d_' + suffix: (_i, [_vp, _vp, _vp, _szp, _i, _i, _vp]),
            'prh_spectral_' + suffix: (_i, [_vp, _i, _vp, _vp, _szp, _vp, _vp, _d, _d, _i, _i, _vp]),
            'prh_rsd_update_' + suffix: (_i, [_vp, _vp, _szp, _vp, _vp, _d, _i, _i, _vp]),
            'prh_read_shifts_' + suffix: (_i, [_vp, _vp, _szp, _vp, _vp, _vp, _i, _sz, _d, _vp, _i, _i, _i,
                                                _vp, _i, ctypes.POINTER(_i), _vp]),
        })
    return sig


SIGNATURES = _signatures()


def _load():
    if not os.path.isfile(LIB_PATH):
        raise ImportError('pyrecon_amd: the HIP library {} is not built. Build it with '
                          '"make -C pyrecon_amd/csrc" (hipcc, gfx950) or __graft_entry__.build().'.format(LIB_PATH))
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


lib = _load()


def last_error():
    msg = lib.prh_last_error()
    return msg.decode() if msg else ''


def check(status, what=''):
    """Raise the exception matching a negative status code."""
    if status == OK:
        return
    msg = '{}: {}'.format(what, last_error()) if what else last_error()
    if status in (ERR_RANGE, ERR_NONFINITE):
        raise MeshError(msg)
    if status == ERR_NOMEM:
        raise MemoryError(msg)
    if status in (ERR_ARG, ERR_CAPACITY, ERR_STATE):
        raise ValueError(msg)
    raise DeviceError('status {}: {}'.format(status, msg))


def size3(values):
    """(3,) size_t array and the pointer to pass for a ``const size_t*`` argument."""
    arr = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.uintp), (3,)))
    return arr, arr.ctypes.data_as(_szp)


def vec3(values, dtype):
    arr = np.empty(3, dtype=dtype)
    arr[...] = values
    return arr, arr.ctypes.data_as(_vp)


def device_count():
    n = ctypes.c_int(0)
    status = lib.prh_device_count(ctypes.byref(n))
    if status != OK: return 0
    return n.value


class KernelTimer(object):
    """Device time of one kernel class (``TIMER_*``) over a region, with HIP events."""

    def __init__(self, which):
        self.which = int(which)

    def __enter__(self):
        check(lib.prh_timer_enable(self.which, 1), 'timer enable')
        check(lib.prh_timer_reset(self.which), 'timer reset')
        return self

    def __exit__(self, *exc):
        check(lib.prh_timer_enable(self.which, 0), 'timer disable')

    def read(self):
        """(total device ms, number of launches, algorithmic bytes); synchronizes."""
        ms, n, nbytes = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        check(lib.prh_timer_read(self.which, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(nbytes)), 'timer read')
        return ms.value, n.value, nbytes.value
