"""The allocation-free full multigrid (prh_fmg_noalloc_f64) against the workspace one
(prh_fmg_f64, the reference's fmg bit for bit): identical bits on two calls in a row, the second
on a smaller mesh (the shared scratch keeps the first call's contents), for a local and a global
line of sight; meshes above 64^3 cells are refused before anything runs.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view('u8')


def _solve(noalloc, f, boxsize, boxcenter, los):
    from pyrecon_amd import _lib, DeviceArray, Workspace
    dv, df = DeviceArray.zeros(f.shape, 'f8'), DeviceArray.from_numpy(f)
    nm, nm_ptr = _lib.size3(f.shape)
    bs, bs_ptr = _lib.vec3(boxsize, 'f8')
    bc, bc_ptr = _lib.vec3(boxcenter, 'f8')
    lp = None
    if los is not None:
        la, lp = _lib.vec3(los, 'f8')
    if noalloc:
        status = _lib.lib.prh_fmg_noalloc_f64(df.ptr, dv.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, 5, 6, lp, None)
    else:
        ws = Workspace(f.shape, 'f8', flags=_lib.WS_MULTIGRID)
        status = _lib.lib.prh_fmg_f64(ws.handle, df.ptr, dv.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, 5, 6, lp, None)
    _lib.check(status, 'fmg')
    return dv.numpy()


@pytest.mark.parametrize('los', [None, (0., 0., 1.)])
def test_noalloc_fmg_matches_workspace_fmg(los):
    rng = np.random.RandomState(5)
    for shape, boxsize, boxcenter in [((64, 64, 64), 500., (0., 0., 800.)), ((16, 24, 32), (200., 300., 400.), (10., 0., 600.))]:
        f = rng.normal(size=shape)
        expect = _solve(False, f, boxsize, boxcenter, los)
        got = _solve(True, f, boxsize, boxcenter, los)
        assert np.array_equal(bits(got), bits(expect)), shape


def test_noalloc_fmg_refuses_large_meshes():
    from pyrecon_amd import _lib, DeviceArray
    shape = (128, 64, 64)  # 2^19 cells: checked on the host before any launch
    nm, nm_ptr = _lib.size3(shape)
    bs, bs_ptr = _lib.vec3(500., 'f8')
    bc, bc_ptr = _lib.vec3(0., 'f8')
    d = DeviceArray.zeros((8,), 'f8')
    status = _lib.lib.prh_fmg_noalloc_f64(d.ptr, d.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, 5, 6, None, None)
    assert status == _lib.ERR_CAPACITY
