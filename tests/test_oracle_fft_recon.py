"""Pins the oracle of the FFT reconstructions (oracle/recon_numpy.py) with closed forms.

- PlaneParallelFFTReconstruction (pyrecon/plane_parallel_fft.py:41-55) on a
  single Fourier mode delta = cos(k0 . x): psi = -k0 sin(k0 . x) / (k0^2 (1 +
  beta mu0^2)) at the mesh nodes, mu0 = k0 . los / |k0|.
- IterativeFFTReconstruction (pyrecon/iterative_fft.py:12-82) with a global
  line of sight on the same mode: each iteration maps the amplitude a of the
  real-space density to 1 - factor mu0^2 a (factor = beta / (1 + beta) first,
  then beta), starting from a = 1; psi = -k0 a sin(k0 . x) / k0^2.  With no
  iteration, psi is the gradient of the inverse Laplacian for either line of
  sight.
- Local against global line of sight for a box far along z (the reference's
  own local ~ global check for the multigrid, test_multigrid.py:124-145).
- Base-class shifts (pyrecon/recon.py:255-325): constant displacement meshes
  read back as the constants; 'disp+rsd' = 'disp' + f (s . los) los.
- OriginalIterativeFFTParticleReconstruction (iterative_fft_particle.py:9-239):
  after one iteration the data move by -f / (1 + beta) (psi . rhat) rhat, psi
  being the field-level oracle's displacement of the same density contrast read
  at the data (pins the first-iteration factor, the sign and the read
  positions); with f = 0 the repainted data stay put and every iteration gives
  the same field; the 'data' shifts split as rsd = x - x_rec and
  x - shifts = x_rec - disp.
"""
import numpy as np
import pytest

from oracle.recon_numpy import OraclePlaneParallelFFTReconstruction, OracleIterativeFFTReconstruction, OracleError

SHAPE = (16, 12, 20)
BOX = np.array([800., 600., 1000.])


def mode(r, m):
    """cos(k0 . x) on the nodes of ``r``'s mesh, with k0 = 2 pi m / L; returns (field, phase, k0)."""
    grids = np.meshgrid(*[np.arange(n) for n in r.shape], indexing='ij')
    phase = sum(2 * np.pi * mi * g / n for mi, g, n in zip(m, grids, r.shape))
    k0 = 2 * np.pi * np.array(m) / np.asarray(r.boxsize, dtype='f8')
    return np.cos(phase), phase, k0


def make(cls, los, boxcenter=(0., 0., 1500.), f=0.8, bias=2.):
    return cls(f=f, bias=bias, los=los, nmesh=SHAPE, boxsize=BOX, boxcenter=boxcenter, dtype='f8')


@pytest.mark.parametrize('m', [(1, 0, 0), (0, 2, 3), (2, 1, 5)])
@pytest.mark.parametrize('los', ['z', [0.3, 0.4, np.sqrt(0.75)]])
def test_plane_parallel_single_mode(m, los):
    r = make(OraclePlaneParallelFFTReconstruction, los)
    field, phase, k0 = mode(r, m)
    r.mesh_delta = field
    r.run()
    k2 = np.sum(k0**2)
    mu2 = np.dot(k0, r.los)**2 / k2
    for a in range(3):
        expect = -k0[a] * np.sin(phase) / (k2 * (1. + r.beta * mu2))
        assert np.allclose(r.mesh_psi[a], expect, atol=1e-12 * np.max(np.abs(k0)) / k2)


def test_plane_parallel_needs_global_los():
    r = make(OraclePlaneParallelFFTReconstruction, None)
    r.mesh_delta = np.zeros(SHAPE)
    with pytest.raises(OracleError):
        r.run()


@pytest.mark.parametrize('niterations', [0, 1, 3])
@pytest.mark.parametrize('m', [(1, 0, 0), (0, 2, 3), (2, 1, 5)])
def test_iterative_global_single_mode(niterations, m):
    los = [0.3, 0.4, np.sqrt(0.75)]
    r = make(OracleIterativeFFTReconstruction, los)
    field, phase, k0 = mode(r, m)
    r.mesh_delta = field
    r.run(niterations=niterations)
    k2 = np.sum(k0**2)
    mu2 = np.dot(k0, r.los)**2 / k2
    a = 1.
    for it in range(niterations):
        factor = r.beta / (1. + r.beta) if it == 0 else r.beta
        a = 1. - factor * mu2 * a
    for ax in range(3):
        expect = -k0[ax] * a * np.sin(phase) / k2
        assert np.allclose(r.mesh_psi[ax], expect, atol=1e-12 * np.max(np.abs(k0)) / k2)


def test_iterative_no_iteration_is_inverse_laplacian_gradient():
    r = make(OracleIterativeFFTReconstruction, None)
    field, phase, k0 = mode(r, (2, 1, 5))
    r.mesh_delta = field
    r.run(niterations=0)
    k2 = np.sum(k0**2)
    for ax in range(3):
        assert np.allclose(r.mesh_psi[ax], -k0[ax] * np.sin(phase) / k2, atol=1e-12 / np.sqrt(k2))


def test_iterative_local_far_box_matches_global():
    # every cell of a box 1e6 Mpc/h away along z sees x / |x| within 5e-4 of z
    far = (0., 0., 1e6)
    rng = np.random.RandomState(5)
    field = rng.normal(size=SHAPE)
    out = []
    for los in [None, 'z']:
        r = make(OracleIterativeFFTReconstruction, los, boxcenter=far)
        r.mesh_delta = field.copy()
        r.run(niterations=3)
        out.append(np.array(r.mesh_psi))
    scale = np.max(np.abs(out[1]))
    assert np.max(np.abs(out[0] - out[1])) < 2e-3 * scale
    # and the local line of sight does differ from no RSD correction at all
    r = make(OracleIterativeFFTReconstruction, 'z', boxcenter=far, f=0.)
    r.mesh_delta = field.copy()
    r.run(niterations=3)
    assert np.max(np.abs(np.array(r.mesh_psi) - out[1])) > 1e-2 * scale


@pytest.mark.parametrize('los', [None, 'x'])
def test_base_class_shifts_of_constant_psi(los):
    r = make(OraclePlaneParallelFFTReconstruction, 'x')
    r.los = None if los is None else r.los
    c = np.array([3., -2., 5.])
    r.mesh_psi = [np.full(SHAPE, ci) for ci in c]
    rng = np.random.RandomState(3)
    pos = rng.uniform(-300., 300., (200, 3)) + np.array([0., 0., 1500.])
    disp = r.read_shifts(pos, field='disp')
    assert np.allclose(disp, c[None, :], rtol=0, atol=1e-12)
    rhat = pos / np.sqrt(np.sum(pos**2, axis=-1))[:, None] if los is None else np.broadcast_to(r.los, pos.shape)
    rsd = r.read_shifts(pos, field='rsd')
    assert np.allclose(rsd, r.f * np.sum(c * rhat, axis=-1)[:, None] * rhat, atol=1e-12)
    assert np.allclose(r.read_shifts(pos), disp + rsd, atol=1e-12)
    assert np.allclose(r.read_shifted_positions(pos), pos - disp - rsd, atol=1e-12)


# ---------------------------------------------------------------------------
# IterativeFFTParticleReconstruction (pyrecon/iterative_fft_particle.py:9-239)
# ---------------------------------------------------------------------------
def _small_catalogs():
    from catalogs import zeldovich_catalog, uniform_box
    boxsize, boxcenter = 500., (0., 0., 800.)
    data = zeldovich_catalog(24, boxsize, boxcenter, f=0.8, workers=2)
    randoms = uniform_box(200000, boxsize, boxcenter, seed=7)
    return dict(boxsize=boxsize, boxcenter=boxcenter, data=data, randoms=randoms)


@pytest.fixture(scope='module')
def small():
    return _small_catalogs()


def _particle(small, los, niterations, f=0.8):
    from oracle.recon_numpy import OracleOriginalIterativeFFTParticleReconstruction
    r = OracleOriginalIterativeFFTParticleReconstruction(f=f, bias=2., los=los, nmesh=32, boxsize=small['boxsize'],
                                                         boxcenter=small['boxcenter'], dtype='f8')
    r.assign_data(small['data'])
    r.assign_randoms(small['randoms'])
    r.set_density_contrast(smoothing_radius=15.)
    r.run(niterations=niterations)
    return r


def _field_psi(small, los):
    """The displacement of the smoothed density contrast, from the field-level oracle (no iteration)."""
    r = OracleIterativeFFTReconstruction(f=0.8, bias=2., los=los, nmesh=32, boxsize=small['boxsize'],
                                         boxcenter=small['boxcenter'], dtype='f8')
    r.assign_data(small['data'])
    r.assign_randoms(small['randoms'])
    r.set_density_contrast(smoothing_radius=15.)
    r.run(niterations=0)
    return r


@pytest.mark.parametrize('los', ['z', None])
def test_particle_first_iteration(small, los):
    # one iteration: x_rec = x - f / (1 + beta) (psi(x) . rhat) rhat with psi the displacement of the
    # smoothed density contrast read at the data, rhat the line of sight
    r = _particle(small, los, 1)
    field = _field_psi(small, los)
    x = small['data']
    psi = np.stack([field.read_cic_nowrap(p, x) for p in field.mesh_psi], axis=-1)
    rhat = x / np.sqrt(np.sum(x**2, axis=-1))[:, None] if los is None else np.broadcast_to(r.los, x.shape)
    expect = x - r.f / (1. + r.beta) * np.sum(psi * rhat, axis=-1)[:, None] * rhat
    scale = np.max(np.abs(psi))
    assert np.max(np.abs(r._positions_rec_data - expect)) < 1e-10 * scale
    for a in range(3):
        assert np.allclose(r.mesh_psi[a], field.mesh_psi[a], rtol=0, atol=1e-12 * scale)


def test_particle_without_rsd_is_stationary(small):
    # f = 0: the data are repainted where they were, so every iteration gives the same field
    one, three = _particle(small, 'z', 1, f=0.), _particle(small, 'z', 3, f=0.)
    assert np.array_equal(three._positions_rec_data, small['data'])
    scale = np.max(np.abs(one.mesh_psi[0]))
    for a in range(3):
        assert np.allclose(three.mesh_psi[a], one.mesh_psi[a], rtol=0, atol=1e-12 * scale)


def test_particle_data_shifts(small):
    r = _particle(small, None, 2)
    x = small['data']
    rsd = r.read_shifts('data', field='rsd')
    assert np.array_equal(rsd, x - r._positions_rec_data)
    disp = r.read_shifts('data', field='disp')
    # the reconstructed data are the real-space estimate minus the displacement read there
    assert np.allclose(r.read_shifted_positions('data'), r._positions_rec_data - disp, rtol=0, atol=1e-9)
