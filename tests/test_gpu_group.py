"""Grouped multigrid solves (several independent problems per launch) give each
problem the same bits as solving it alone.

The grouped entry points have no counterpart in the reference; their contract
is equality with the single-problem entry points, which the other GPU tests
tie bit-for-bit to the reference's C (src/multigrid.c:303-328).  Shapes cover
a mesh whose levels are all handled by grouped launches (32^3, 24x32x40) and
one whose finest level runs the blocked kernel per job (64^3); the jobs differ
in box, centre, beta and line of sight (local and global in one group).
"""
import ctypes

import numpy as np
import pytest

from pyrecon_amd import _lib
from pyrecon_amd._lib import lib, check
from pyrecon_amd.device import DeviceArray
from pyrecon_amd.mesh import Workspace

pytestmark = pytest.mark.gpu

JOBS = [
    dict(boxsize=(1000., 1000., 1000.), boxcenter=(0., 0., 1500.), beta=0.4, los=None),
    dict(boxsize=(800., 900., 1000.), boxcenter=(100., -50., 2000.), beta=0.35, los=(1., 0., 0.)),
    dict(boxsize=(1500., 1500., 1500.), boxcenter=(0., 0., 3000.), beta=0.5, los=None),
]


def bits(a):
    return np.ascontiguousarray(a).view('u{:d}'.format(a.dtype.itemsize))


def sources(shape, dtype, count, seed=5):
    rng = np.random.RandomState(seed)
    return [rng.normal(size=shape).astype(dtype) for _ in range(count)]


def solve_alone(f, job, dtype, mixed=False):
    flags = _lib.WS_MULTIGRID | (_lib.WS_MIXED if mixed else 0)
    ws = Workspace(f.shape, dtype, paint_chunk=1, flags=flags)
    df, dv = DeviceArray.from_numpy(f), DeviceArray.zeros(f.shape, dtype)
    nm, nm_ptr = _lib.size3(f.shape)
    bs, bs_ptr = _lib.vec3(job['boxsize'], dtype)
    bc, bc_ptr = _lib.vec3(job['boxcenter'], dtype)
    los_ptr = None
    if job['los'] is not None:
        la, los_ptr = _lib.vec3(job['los'], dtype)
    if mixed:
        check(lib.prh_fmg_mixed(ws.handle, df.ptr, dv.ptr, nm_ptr, bs_ptr, bc_ptr, job['beta'], 0.4, 5, 6, los_ptr,
                                0, 1, None, None), 'fmg_mixed')
    else:
        fn = getattr(lib, 'prh_fmg_f32' if np.dtype(dtype) == np.float32 else 'prh_fmg_f64')
        check(fn(ws.handle, df.ptr, dv.ptr, nm_ptr, bs_ptr, bc_ptr, job['beta'], 0.4, 5, 6, los_ptr, None), 'fmg')
    return dv.numpy()


def solve_group(fs, jobs, dtype, mixed=False):
    count = len(fs)
    flags = _lib.WS_MULTIGRID | (_lib.WS_MIXED if mixed else 0)
    wss = [Workspace(fs[0].shape, dtype, paint_chunk=1, flags=flags) for _ in range(count)]
    dfs = [DeviceArray.from_numpy(f) for f in fs]
    dvs = [DeviceArray.zeros(fs[0].shape, dtype) for _ in range(count)]
    real = ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double
    ws_arr = (ctypes.c_void_p * count)(*[w.handle for w in wss])
    f_arr = (ctypes.c_void_p * count)(*[d.ptr for d in dfs])
    v_arr = (ctypes.c_void_p * count)(*[d.ptr for d in dvs])
    bs = (real * (3 * count))(*[x for j in jobs for x in j['boxsize']])
    bc = (real * (3 * count))(*[x for j in jobs for x in j['boxcenter']])
    beta = (real * count)(*[j['beta'] for j in jobs])
    los_keep = [None if j['los'] is None else (real * 3)(*j['los']) for j in jobs]
    los_arr = (ctypes.c_void_p * count)(*[None if l is None else ctypes.cast(l, ctypes.c_void_p).value for l in los_keep])
    nm, nm_ptr = _lib.size3(fs[0].shape)
    if mixed:
        check(lib.prh_fmg_mixed_group(ws_arr, count, f_arr, v_arr, nm_ptr, bs, bc, beta, 0.4, 5, 6, los_arr, 0, 1,
                                      None), 'fmg_mixed_group')
    else:
        fn = getattr(lib, 'prh_fmg_group_f32' if np.dtype(dtype) == np.float32 else 'prh_fmg_group_f64')
        check(fn(ws_arr, count, f_arr, v_arr, nm_ptr, bs, bc, beta, 0.4, 5, 6, los_arr, None), 'fmg_group')
    return [d.numpy() for d in dvs]


@pytest.mark.parametrize('shape', [(32, 32, 32), (24, 32, 40), (64, 64, 64)])
@pytest.mark.parametrize('dtype', ['f4', 'f8'])
def test_group_fmg_equals_alone(shape, dtype):
    fs = sources(shape, dtype, len(JOBS))
    group = solve_group(fs, JOBS, dtype)
    for f, job, g in zip(fs, JOBS, group):
        assert np.array_equal(bits(g), bits(solve_alone(f, job, dtype)))


@pytest.mark.parametrize('shape', [(32, 32, 32), (64, 64, 64)])
def test_group_mixed_equals_alone(shape):
    fs = sources(shape, 'f8', len(JOBS), seed=6)
    group = solve_group(fs, JOBS, 'f8', mixed=True)
    for f, job, g in zip(fs, JOBS, group):
        assert np.array_equal(bits(g), bits(solve_alone(f, job, 'f8', mixed=True)))


def test_group_of_one_and_errors():
    fs = sources((32, 32, 32), 'f8', 1, seed=7)
    assert np.array_equal(bits(solve_group(fs, JOBS[:1], 'f8')[0]), bits(solve_alone(fs[0], JOBS[0], 'f8')))
    # a workspace may appear once per group
    ws = Workspace((32, 32, 32), 'f8', paint_chunk=1, flags=_lib.WS_MULTIGRID)
    d = [DeviceArray.zeros((32, 32, 32), 'f8') for _ in range(4)]
    ws_arr = (ctypes.c_void_p * 2)(ws.handle, ws.handle)
    f_arr = (ctypes.c_void_p * 2)(d[0].ptr, d[1].ptr)
    v_arr = (ctypes.c_void_p * 2)(d[2].ptr, d[3].ptr)
    bs = (ctypes.c_double * 6)(*([1000.] * 6))
    bc = (ctypes.c_double * 6)(*([0.] * 6))
    beta = (ctypes.c_double * 2)(0.4, 0.4)
    nm, nm_ptr = _lib.size3((32, 32, 32))
    assert lib.prh_fmg_group_f64(ws_arr, 2, f_arr, v_arr, nm_ptr, bs, bc, beta, 0.4, 5, 6, None, None) == _lib.ERR_ARG
    assert 'share a workspace' in _lib.last_error()
    assert lib.prh_fmg_group_f64(ws_arr, 9, f_arr, v_arr, nm_ptr, bs, bc, beta, 0.4, 5, 6, None, None) == _lib.ERR_ARG
