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
