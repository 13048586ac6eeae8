"""Pins the numpy half of the oracle (oracle/recon_numpy.py) with closed forms.

- Gaussian smoothing: a single Fourier mode is scaled by exactly
  exp(-(R k)^2 / 2); the sum is conserved (the reference's test_smoothing,
  pyrecon/tests/test_mesh.py:83-100); the real-to-complex route gives the same
  field (test_hermitian, test_mesh.py:162-170).
- Density contrast: the rule of pyrecon/recon.py:227-247 on hand-built meshes
  (mask below threshold -> 0, ratio above), and mean zero without randoms
  (test_multigrid.py:33).
- Martin White's density contrast (pyrecon/multigrid.py:69-97) on a hand-built
  mesh whose answer is worked out in fractions: the cell with 0 < randoms <
  ran_min loses its data, the empty-randoms cell keeps its data in alpha, the
  non-zero cells end with mean zero.
- Mesh-unit transform and wrap: numpy's promotion for f4/f8 and Python's
  modulo sign convention.
- Shift algebra: 'disp+rsd' = 'disp' + 'rsd', rsd parallel to the line of
  sight with the growth-rate factor, shifted positions = positions - shifts.
- Seeded catalogs regenerate identically (SHA-256 of configuration 1).
"""
import numpy as np
import pytest

from oracle.recon_numpy import OracleMultiGridReconstruction, OracleOriginalMultiGridReconstruction, OracleError, distance


def make(dtype='f8', los=None, nmesh=16, boxsize=800., boxcenter=(0., 0., 1000.), wrap=False, f=0.8, bias=2.):
    return OracleMultiGridReconstruction(f=f, bias=bias, los=los, nmesh=nmesh, boxsize=boxsize, boxcenter=boxcenter,
                                         dtype=dtype, wrap=wrap)


@pytest.mark.parametrize('m', [(1, 0, 0), (0, 3, 0), (2, 1, 5)])
def test_smoothing_single_mode(m):
    r = make(nmesh=(16, 12, 20), boxsize=(800., 600., 1000.))
    grids = np.meshgrid(*[np.arange(n) for n in r.shape], indexing='ij')
    field = np.cos(sum(2 * np.pi * mi * g / n for mi, g, n in zip(m, grids, r.shape)))
    out = r.smooth_gaussian(field, 15.)
    k2 = sum((2 * np.pi * mi / L)**2 for mi, L in zip(m, r.boxsize))
    assert np.allclose(out, np.exp(-0.5 * 15.**2 * k2) * field, atol=1e-12)


def test_smoothing_sum_and_hermitian():
    r = make(nmesh=32)
    rng = np.random.RandomState(0)
    value = rng.uniform(0., 1., r.shape)
    out = r.smooth_gaussian(value, 10.)
    assert np.isclose(out.sum(), value.sum(), rtol=1e-12)
    # real-to-complex route
    k = [2 * np.pi * np.fft.fftfreq(n, d=L / n) for n, L in zip(r.shape[:2], r.boxsize[:2])]
    kz = 2 * np.pi * np.fft.rfftfreq(r.shape[2], d=r.boxsize[2] / r.shape[2])
    gauss = np.exp(-0.5 * 100. * (k[0][:, None, None]**2 + k[1][None, :, None]**2 + kz[None, None, :]**2))
    herm = np.fft.irfftn(np.fft.rfftn(value) * gauss, s=r.shape, axes=(0, 1, 2))
    assert np.allclose(herm, out, atol=1e-13)


def test_density_rule():
    r = make(nmesh=4)
    data = np.full(r.shape, 2.)
    randoms = np.full(r.shape, 4.)
    randoms[0, 0, 0] = 0.  # empty cell: below the threshold
    r.mesh_data, r.mesh_randoms, r._size_randoms = data, randoms, 100
    r.smooth_gaussian = lambda value, radius: value  # isolate the rule from the smoothing
    r.set_density_contrast(ran_min=0.01)
    alpha = data.sum() / randoms.sum()
    assert np.isclose(r.alpha, alpha)
    assert np.isclose(r.threshold, 0.01 * randoms.sum() / 100)
    expect = (2. - alpha * 4.) / (2. * alpha * 4.)
    assert r.mesh_delta[0, 0, 0] == 0.
    assert np.allclose(r.mesh_delta.ravel()[1:], expect)
    assert r.mask_count == randoms.size - 1


def test_martin_density_rule():
    r = OracleOriginalMultiGridReconstruction(f=0.8, bias=2., nmesh=(2, 2, 2), boxsize=800., boxcenter=0.)
    randoms = np.array([1., 2., 4., 8., 0.5, 0., 0.5, 0.]).reshape(r.shape)
    data = np.array([1., 1., 1., 1., 3., 5., 0., 0.]).reshape(r.shape)
    r.mesh_data, r.mesh_randoms, r._size_randoms = data, randoms, 100
    r.smooth_gaussian = lambda value, radius: value  # isolate the rule from the smoothing
    r.set_density_contrast()
    # alpha = (1 + 1 + 1 + 1 + 5) / (1 + 2 + 4 + 8): the r = 0.5 cell's data is zeroed, the r = 0 cell's is kept
    assert np.isclose(r.alpha, 0.6)
    expect = np.array([85., 5., -35., -55., 0., 0., 0., 0.]) / 192.
    assert np.allclose(r.mesh_delta.ravel(), expect, rtol=0, atol=1e-15)
    assert r.mask_count == 4


def test_martin_density_no_randoms():
    r = OracleOriginalMultiGridReconstruction(f=0.8, bias=2., nmesh=8, boxsize=800., boxcenter=0.)
    rng = np.random.RandomState(3)
    data = rng.uniform(0.5, 1.5, r.shape)
    r.mesh_data = data.copy()
    r.smooth_gaussian = lambda value, radius: value
    r.set_density_contrast()
    # 1 + bias delta is the data in units of its mean: ratios of cells are ratios of data
    ratio = (1. + 2. * r.mesh_delta) / (1. + 2. * r.mesh_delta.ravel()[0])
    assert np.allclose(ratio, data / data.ravel()[0], rtol=1e-14)
    assert np.isclose(np.mean(r.mesh_delta), 0., atol=1e-15)


def test_density_no_randoms():
    r = make(nmesh=8)
    rng = np.random.RandomState(1)
    r.mesh_data = rng.uniform(0.5, 1.5, r.shape)
    r.set_density_contrast(smoothing_radius=15.)
    assert np.isclose(np.mean(r.mesh_delta), 0., atol=1e-14)


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
def test_mesh_units_and_wrap(dtype):
    r = make(dtype=dtype, nmesh=16, boxsize=800., boxcenter=(0., 0., 1000.))
    low = r.boxcenter - r.boxsize / 2.
    pos = np.array([low, low + r.boxsize / 2., low + r.boxsize], dtype='f8')
    mu = r.mesh_units(pos)
    assert mu.dtype == np.dtype(dtype)
    assert np.allclose(mu, [[0., 0., 0.], [8., 8., 8.], [16., 16., 16.]])
    out = r.wrap_positions(np.array([[low[0] - 1., low[1] + 801., 1000.]]))
    assert np.allclose(out, [[low[0] + 799., low[1] + 1., 1000.]])
    with pytest.raises(OracleError):
        r.assign_data(np.array([[0., 0., 2000.]]))
    with pytest.raises(OracleError):
        r.assign_data(np.array([[0., np.nan, 1000.]]))


@pytest.mark.parametrize('los', [None, 'x', [1., 1., 0.]])
def test_shift_algebra(los):
    r = make(los=los, nmesh=16)
    rng = np.random.RandomState(2)
    r.mesh_phi = rng.normal(size=r.shape)
    pos = rng.uniform(-350., 350., (500, 3)) + r.boxcenter
    disp = r.read_shifts(pos, field='disp')
    rsd = r.read_shifts(pos, field='rsd')
    both = r.read_shifts(pos, field='disp+rsd')
    assert np.allclose(both, disp + rsd, rtol=1e-14, atol=1e-14)
    rhat = pos / distance(pos)[:, None] if los is None else np.broadcast_to(r.los, pos.shape)
    assert np.allclose(rsd, r.f * np.sum(disp * rhat, axis=-1)[:, None] * rhat)
    assert np.allclose(np.cross(rsd, rhat), 0., atol=1e-12)
    assert np.allclose(r.read_shifted_positions(pos), pos - both)
    with pytest.raises(OracleError):
        r.read_shifts(pos, field='rsd+disp')


def test_catalog_regeneration_is_frozen():
    from catalogs import make_catalogs, sha256
    import json, os
    fixture = json.load(open(os.path.join(os.path.dirname(__file__), 'golden', 'catalogs_sha256.json')))
    for name in ('c1', 'c2'):
        cat = make_catalogs(name)
        assert sha256(cat['data'], cat['data_weights']) == fixture[name]['data'], name
        assert sha256(cat['randoms'], cat['randoms_weights']) == fixture[name]['randoms'], name


def test_threaded_smoothing_matches_numpy():
    # the oracle's scipy.fft route (several threads, for 512^3 meshes) against numpy's engine
    r = make(nmesh=(24, 16, 20), boxsize=(800., 600., 1000.))
    rng = np.random.RandomState(3)
    value = rng.uniform(0., 1., r.shape)
    ref = r.smooth_gaussian(value, 15.)
    r.fft_workers = 4
    out = r.smooth_gaussian(value, 15.)
    assert np.max(np.abs(out - ref)) < 1e-14 * np.max(np.abs(ref))
    grids = np.meshgrid(*[np.arange(n) for n in r.shape], indexing='ij')
    field = np.cos(2 * np.pi * (2 * grids[0] / 24 + grids[2] / 20))
    k2 = (2 * np.pi * 2 / 800.)**2 + (2 * np.pi / 1000.)**2
    assert np.allclose(r.smooth_gaussian(field, 15.), np.exp(-0.5 * 15.**2 * k2) * field, atol=1e-12)
