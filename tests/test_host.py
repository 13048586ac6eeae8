"""Host-side logic of the drop-in classes (no GPU needed).

MeshInfo geometry and nmesh selection follow the reference (pyrecon/mesh.py:342-479,
pyrecon/multigrid.py:23-36; its test_nmesh, pyrecon/tests/test_multigrid.py:97-104);
the output dtypes of the shift reads follow numpy's promotion as the reference's
numpy code produces them.
"""
import warnings

import numpy as np
import pytest

from pyrecon_amd.mesh import MeshInfo
from pyrecon_amd._lib import MeshError
from pyrecon_amd.multigrid import OriginalMultiGridReconstruction, MultiGridReconstruction
from pyrecon_amd.recon import BaseReconstruction


def test_meshinfo_storage():
    info = MeshInfo(boxsize=1000., boxcenter=0., nmesh=64, dtype='f4')
    assert info.nmesh.dtype == np.intc and np.all(info.nmesh == 64)
    assert info.boxsize.dtype == np.float32 and info.boxcenter.dtype == np.float32
    info.boxsize = 2000.
    assert np.all(info.boxsize == 2000.)
    assert np.allclose(info.offset, -1000.) and np.allclose(info.cellsize, 2000. / 64)
    assert info.shape == (64, 64, 64) and info.size == 64**3


def test_meshinfo_from_positions():
    rng = np.random.RandomState(0)
    pos = rng.uniform(-100., 300., (1000, 3))
    info = MeshInfo(positions=pos, nmesh=32)
    delta = pos.max(axis=0) - pos.min(axis=0)
    assert np.allclose(info.boxsize, 1.5 * delta)
    assert np.allclose(info.boxcenter, 0.5 * (pos.min(axis=0) + pos.max(axis=0)))
    with pytest.raises(MeshError):
        MeshInfo(boxsize=1000.)
    with pytest.raises(MeshError):
        MeshInfo(positions=pos, boxsize=10., nmesh=8)
    with pytest.raises(MeshError):
        MeshInfo(boxsize=1000., boxcenter=0.)


def test_select_nmesh():
    select = OriginalMultiGridReconstruction._select_nmesh
    assert list(select(np.array([12, 13, 17]))) == [12, 14, 20]
    assert list(select(np.array([64, 65, 100]))) == [64, 80, 112]
    for n in range(4, 600):
        m = select(np.array([n]))[0]
        assert m >= n
        k = m
        while k % 2 == 0: k //= 2
        assert k in (1, 3, 5, 7)
    rng = np.random.RandomState(81)
    pos = rng.uniform(0., 1000., (1000, 3))
    info = MeshInfo(positions=pos, cellsize=[10, 8, 9], select_nmesh=select)
    assert np.all(info.nmesh % 2 == 0)
    with pytest.warns(UserWarning):
        MeshInfo(positions=pos, nmesh=[12, 13, 17], select_nmesh=select)
    with warnings.catch_warnings():
        warnings.simplefilter('error')
        MeshInfo(positions=pos, nmesh=[12, 14, 20], select_nmesh=select)


def bare(dtype, los, wrap):
    r = object.__new__(MultiGridReconstruction)
    r.info = MeshInfo(boxsize=1000., boxcenter=0., nmesh=8, dtype=dtype)
    r.wrap = wrap
    r.f, r.bias = 0.8, 2.
    r.set_los(los)
    return r


@pytest.mark.parametrize('mesh', ['f4', 'f8'])
@pytest.mark.parametrize('pos', ['f4', 'f8'])
@pytest.mark.parametrize('los', [None, 'x'])
@pytest.mark.parametrize('wrap', [False, True])
def test_shift_dtypes_follow_numpy(mesh, pos, los, wrap):
    # what the reference's numpy code returns (recon.py:288-325, multigrid.py:137-158)
    r = bare(mesh, los, wrap)
    shifts = np.zeros((2, 3), dtype=pos)  # read_finite_difference_cic casts to positions.dtype
    x = np.ones((2, 3), dtype=pos)
    lo = x / np.sqrt((x**2).sum(axis=-1))[:, None] if los is None else r.los
    rsd = r.f * np.sum(shifts * lo, axis=-1)[:, None] * lo
    expect = {'disp': shifts.dtype, 'rsd': rsd.dtype, 'disp+rsd': shifts.dtype}
    for field, dt in expect.items():
        assert r._shift_out_dtype(np.dtype(pos), field, shifted=False) == dt
        out = x - (rsd if field == 'rsd' else shifts)
        if wrap: out = r.info.wrap(out)
        assert r._shift_out_dtype(np.dtype(pos), field, shifted=True) == out.dtype


def test_set_los():
    r = bare('f8', 'y', False)
    assert np.allclose(r.los, [0., 1., 0.])
    r.set_los([3., 4., 0.])
    assert np.allclose(r.los, [0.6, 0.8, 0.]) and r.los.dtype == np.float64
    r.set_los('local')
    assert r.los is None
    r = bare('f4', 2, False)
    assert r.los.dtype == np.float32 and np.allclose(r.los, [0., 0., 1.])


def test_cosmo():
    r = bare('f8', None, False)
    assert np.isclose(r.beta, 0.4)
    r.set_cosmo(beta=0.5)
    assert np.isclose(r.f, 1.0) and np.isclose(r.bias, 2.)
    assert issubclass(MultiGridReconstruction, BaseReconstruction)


def test_serving_footprint_and_modes():
    # admission sizes a job by its workspace (hierarchy, paint scratch, FFT) plus meshes, catalogs and outputs
    from pyrecon_amd.serving import ReconJob, zeldovich_modes
    from pyrecon_amd.mesh import Workspace
    job = ReconJob((np.zeros((1000, 3)), None), (np.zeros((10000, 3)), None), nmesh=64, boxsize=500., boxcenter=0.,
                   shifted=('data', 'randoms'))
    ws, other = job.footprint(1 << 16)
    assert ws == Workspace.footprint((64, 64, 64), 'f8', paint_chunk=1 << 16, flags=job.flags())
    assert other == 3 * 64**3 * 8 + 11000 * 32 + 11000 * 24
    mixed = ReconJob((np.zeros((10, 3)), None), nmesh=64, boxsize=500., boxcenter=0., precision='mixed')
    assert mixed.footprint(1 << 16)[0] > ws  # the FP32 hierarchy of the mixed solve
    with pytest.raises(ValueError):
        ReconJob(lambda s: None, nmesh=64, boxsize=500., boxcenter=0.)
    # random Fourier modes: every mode carries an equal share of the displacement variance
    # sigma^2 = int P(k) dk / (2 pi^2); the per-component variance of psi is sigma^2 / 3
    modes = zeldovich_modes([1000.] * 3, [128] * 3, seed=3, nmodes=512)
    k = np.sqrt(np.sum(modes[:, :3]**2, axis=-1))
    assert np.all(k >= 2 * np.pi / 1000. - 1e-12) and np.all(k <= np.pi * 128 / 1000. + 1e-12)
    grid = np.linspace(2 * np.pi / 1000., np.pi * 128 / 1000., 200001)
    pk = 2e4 * (grid / 0.1)**-1.5 * np.exp(-(grid / 0.3)**2)
    sigma2 = np.trapezoid(pk, grid) / (2 * np.pi**2)
    var = np.sum((modes[:, 3] * k)**2 / 2.)
    assert np.isclose(var, sigma2, rtol=1e-3)
