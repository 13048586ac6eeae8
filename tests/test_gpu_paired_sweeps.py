"""Two Jacobi sweeps per launch (k_stencil_pair, PRH_TUNE_PAIRED_SWEEPS = 1: the mixed-precision
solve on levels with whole rows of 256 or 512 cells) against one sweep per launch: the same bits,
for the FP32 smoother alone (prh_jacobi_fast_f32) and for the whole mixed solve.

The shapes cover one x plane per block with the periodic wrap in x (n0 = 6), several planes per
block with a partial last run (n0 = 1100: runs of 3), rows of 256 and 512 cells, a level below
that is left to the single sweeps, an odd number of sweeps (the last one alone) and an even one
(the last pair adds into the FP64 iterate), zero starts, and a local and a global line of sight.
Also: a mixed solve whose hierarchy has two levels starts its coarse stage from zero (repeatable,
and within 1e-5 of the FP64 full multigrid); and the first two sweeps from a zero iterate, one
launch on the whole-row levels, give the bits of a zero-start sweep followed by a sweep.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view('u{:d}'.format(a.dtype.itemsize))


def _mixed(f, boxsize, boxcenter, los, niter, nrefine, vpr):
    from pyrecon_amd import _lib, DeviceArray, Workspace
    ws = Workspace(f.shape, 'f8', flags=_lib.WS_MULTIGRID | _lib.WS_MIXED)
    df, dv = DeviceArray.from_numpy(f), DeviceArray.zeros(f.shape, 'f8')
    nm, nm_ptr = _lib.size3(f.shape)
    bs, bs_ptr = _lib.vec3(boxsize, 'f8')
    bc, bc_ptr = _lib.vec3(boxcenter, 'f8')
    lp = None
    if los is not None:
        la, lp = _lib.vec3(los, 'f8')
    _lib.check(_lib.lib.prh_fmg_mixed(ws.handle, df.ptr, dv.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, niter, 6, lp,
                                      nrefine, vpr, None, None), 'fmg_mixed')
    out = dv.numpy()
    ws.close()
    return out


@pytest.mark.parametrize('shape, boxsize', [((6, 8, 512), (60., 80., 2000.)), ((1100, 8, 256), (3000., 40., 900.)),
                                            ((32, 16, 256), 500.), ((16, 32, 512), (300., 400., 1000.))])
@pytest.mark.parametrize('los', [None, (0., 0., 1.)])
@pytest.mark.parametrize('niter, nrefine, vpr', [(5, 0, 1), (4, 1, 2)])
def test_paired_sweeps_same_bits(shape, boxsize, los, niter, nrefine, vpr):
    from pyrecon_amd import _lib
    rng = np.random.RandomState(sum(shape) + niter)
    f = rng.normal(size=shape)
    f -= f.mean()
    boxcenter = (10., -20., 1500.)
    got = {}
    for paired in (1, 0):
        prev = _lib.lib.prh_set_tuning(_lib.TUNE_PAIRED_SWEEPS, paired)
        try:
            got[paired] = _mixed(f, boxsize, boxcenter, los, niter, nrefine, vpr)
        finally:
            _lib.lib.prh_set_tuning(_lib.TUNE_PAIRED_SWEEPS, prev)
    assert np.all(np.isfinite(got[1]))
    assert np.array_equal(bits(got[1]), bits(got[0]))


def _plain(f, boxsize, boxcenter, los, niter):
    from pyrecon_amd import _lib, DeviceArray, Workspace
    ws = Workspace(f.shape, 'f8', flags=_lib.WS_MULTIGRID)
    df, dv = DeviceArray.from_numpy(f), DeviceArray.zeros(f.shape, 'f8')
    nm, nm_ptr = _lib.size3(f.shape)
    bs, bs_ptr = _lib.vec3(boxsize, 'f8')
    bc, bc_ptr = _lib.vec3(boxcenter, 'f8')
    lp = None
    if los is not None:
        la, lp = _lib.vec3(los, 'f8')
    _lib.check(_lib.lib.prh_fmg_f64(ws.handle, df.ptr, dv.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, niter, 6, lp, None),
               'fmg_f64')
    out = dv.numpy()
    ws.close()
    return out


@pytest.mark.parametrize('shape', [(6, 8, 512), (10, 10, 10)])
def test_two_level_mixed_solve(shape):
    # a hierarchy of two levels (level 1 cannot be coarsened): the coarse stage starts from zero,
    # as the full multigrid's coarsest level does, so repeated solves give the same bits and track
    # the FP64 full multigrid
    rng = np.random.RandomState(7)
    f = rng.normal(size=shape)
    f -= f.mean()
    args = ((300., 400., 1000.), (10., -20., 1500.), None, 5)
    a = _mixed(f, *args, 0, 1)
    b = _mixed(f, *args, 0, 1)
    assert np.array_equal(bits(a), bits(b))
    p = _plain(f, *args)
    a, p = a - a.mean(), p - p.mean()
    assert np.sqrt(np.sum((a - p)**2) / np.sum(p**2)) < 1e-5


def _smooth(v, f, boxsize, boxcenter, los, niter, zero):
    from pyrecon_amd import _lib, DeviceArray, Workspace
    ws = Workspace(f.shape, 'f4', flags=_lib.WS_MULTIGRID)
    dv, df = DeviceArray.from_numpy(v), DeviceArray.from_numpy(f)
    nm, nm_ptr = _lib.size3(f.shape)
    bs, bs_ptr = _lib.vec3(boxsize, 'f4')
    bc, bc_ptr = _lib.vec3(boxcenter, 'f4')
    lp = None
    if los is not None:
        la, lp = _lib.vec3(los, 'f4')
    _lib.check(_lib.lib.prh_jacobi_fast_f32(ws.handle, dv.ptr, df.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, niter, lp,
                                            int(zero), None), 'jacobi_fast')
    out = dv.numpy()
    ws.close()
    return out


@pytest.mark.parametrize('shape', [(6, 8, 512), (1100, 8, 256), (32, 16, 256)])
@pytest.mark.parametrize('los', [None, (0., 0., 1.)])
@pytest.mark.parametrize('niter, zero', [(2, False), (2, True), (3, False), (4, True)])
def test_paired_smoother_same_bits(shape, los, niter, zero):
    # the FP32 smoother alone: one launch per two sweeps against one per sweep
    from pyrecon_amd import _lib
    rng = np.random.RandomState(niter + sum(shape))
    v = rng.normal(size=shape).astype('f4')
    f = rng.normal(size=shape).astype('f4')
    got = {}
    for paired in (1, 0):
        prev = _lib.lib.prh_set_tuning(_lib.TUNE_PAIRED_SWEEPS, paired)
        try:
            got[paired] = _smooth(v, f, (300., 400., 1000.), (10., -20., 1500.), los, niter, zero)
        finally:
            _lib.lib.prh_set_tuning(_lib.TUNE_PAIRED_SWEEPS, prev)
    diff = got[1] != got[0]
    where = np.argwhere(diff)
    assert not diff.any(), (int(diff.sum()), where[:5].tolist(), where[-5:].tolist())


def test_tuning_switch():
    from pyrecon_amd import _lib
    prev = _lib.lib.prh_set_tuning(_lib.TUNE_PAIRED_SWEEPS, 0)
    assert _lib.lib.prh_set_tuning(_lib.TUNE_PAIRED_SWEEPS, prev) == 0
    assert _lib.lib.prh_set_tuning(12345, 1) == _lib.ERR_ARG


@pytest.mark.parametrize('shape', [(6, 8, 512), (16, 32, 512), (32, 16, 256), (5, 8, 1024), (3, 4, 1024),
                                   (9, 16, 128), (7, 8, 64)])
@pytest.mark.parametrize('los', [None, (0., 0., 1.)])
def test_two_sweeps_from_zero_same_bits(shape, los):
    # the first two sweeps from a zero iterate run as one launch on the whole-row levels
    # (kOpJacobiAfterZero): the same bits as a zero-start sweep followed by a sweep
    rng = np.random.RandomState(sum(shape))
    v = rng.normal(size=shape).astype('f4')  # ignored by the zero start
    f = rng.normal(size=shape).astype('f4')
    args = ((300., 400., 1000.), (10., -20., 1500.), los)
    together = _smooth(v, f, *args, 2, True)
    first = _smooth(v, f, *args, 1, True)
    both = _smooth(first, f, *args, 1, False)
    assert np.array_equal(bits(together), bits(both))
    # and three sweeps: the pair from zero, then one more
    three = _smooth(v, f, *args, 3, True)
    assert np.array_equal(bits(three), bits(_smooth(both, f, *args, 1, False)))
