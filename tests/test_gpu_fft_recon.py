"""Parity of the FFT reconstructions on the device with the CPU oracle.

- 3-D transforms (prh_fft3d_*) against numpy.fft.fftn / ifftn for shapes that
  cover every radix and 512-cell lines along each axis: relative L2 <= 1e-13
  (FP64), 2e-6 (FP32).
- Fourier-space operators (prh_spectral_*) against the oracle's numpy
  expressions of plane_parallel_fft.py:45-53, iterative_fft.py:37-81 and
  iterative_fft_particle.py:112.
- PlaneParallelFFTReconstruction, IterativeFFTReconstruction and
  IterativeFFTParticleReconstruction (global and local line of sight) end to end on the configuration-1 catalogs against
  their oracle classes in oracle/recon_numpy.py:
  displacement meshes to relative L2 <= 1e-10 (FP64; the smoothing and the
  transforms round differently from numpy's) and shifted positions to
  1e-8 Mpc/h; FP32 meshes to the reference's own f4-vs-f8 tolerance
  (pyrecon/tests/test_multigrid.py:94, 1e-2).
"""
import ctypes

import numpy as np
import pytest

from catalogs import make_catalogs, RECON_KWARGS, SMOOTHING_RADIUS

pytestmark = pytest.mark.gpu

SHAPES = [(8, 12, 20), (16, 16, 16), (6, 10, 14), (15, 9, 25), (4, 6, 512), (512, 2, 16), (2, 512, 8),
          (11, 13, 6), (22, 4, 17)]  # the last two: direct DFT along axes with prime factors above 7


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.sqrt(np.sum(np.abs(a - b)**2) / np.sum(np.abs(b)**2))


def _dev_complex(value, dtype):
    from pyrecon_amd import DeviceArray
    inter = np.empty(value.shape + (2,), dtype=dtype)
    inter[..., 0], inter[..., 1] = value.real, value.imag
    return DeviceArray.from_numpy(inter)


def _host_complex(arr):
    v = arr.numpy()
    return v[..., 0].astype('f8') + 1j * v[..., 1].astype('f8')


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', SHAPES)
def test_fft3d(dtype, shape):
    from pyrecon_amd import DeviceArray, Workspace, _lib
    rng = np.random.RandomState(sum(shape))
    ws = Workspace(shape, dtype, flags=_lib.WS_FFT)
    fn = getattr(_lib.lib, 'prh_fft3d_' + ('f32' if dtype == 'f4' else 'f64'))
    nm, nm_ptr = _lib.size3(shape)
    tol = 1e-13 if dtype == 'f8' else 2e-6
    real = rng.normal(size=shape).astype(dtype)
    cplx = (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    cplx = cplx.astype('c16' if dtype == 'f8' else 'c8')
    # forward, real in
    src = DeviceArray.from_numpy(real)
    dst = DeviceArray(shape + (2,), dtype)
    _lib.check(fn(ws.handle, src.ptr, dst.ptr, nm_ptr, 0, _lib.FFT_REAL_IN, None))
    assert rel_l2(_host_complex(dst), np.fft.fftn(real.astype('f8'))) < tol
    # forward, complex in place
    buf = _dev_complex(cplx, dtype)
    _lib.check(fn(ws.handle, buf.ptr, buf.ptr, nm_ptr, 0, _lib.FFT_COMPLEX, None))
    assert rel_l2(_host_complex(buf), np.fft.fftn(cplx.astype('c16'))) < tol
    # inverse (no 1 / N), complex out and real part out
    buf = _dev_complex(cplx, dtype)
    out = DeviceArray(shape + (2,), dtype)
    _lib.check(fn(ws.handle, buf.ptr, out.ptr, nm_ptr, 1, _lib.FFT_COMPLEX, None))
    expect = np.fft.ifftn(cplx.astype('c16')) * np.prod(shape)
    assert rel_l2(_host_complex(out), expect) < tol
    buf = _dev_complex(cplx, dtype)
    outr = DeviceArray(shape, dtype)
    _lib.check(fn(ws.handle, buf.ptr, outr.ptr, nm_ptr, 1, _lib.FFT_REAL_OUT, None))
    assert rel_l2(outr.numpy(), expect.real) < tol
    # a real input with the inverse is refused
    assert fn(ws.handle, src.ptr, dst.ptr, nm_ptr, 1, _lib.FFT_REAL_IN, None) == _lib.ERR_ARG


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
def test_spectral_operators(dtype):
    from pyrecon_amd import DeviceArray, Workspace, _lib
    from oracle.recon_numpy import OracleIterativeFFTReconstruction
    shape, boxsize = (8, 12, 10), np.array([800., 600., 1000.])
    r = OracleIterativeFFTReconstruction(f=0.8, bias=2., los='x', nmesh=shape, boxsize=boxsize,
                                         boxcenter=(0., 0., 1500.), dtype=dtype)
    k = r.kcoords()
    rng = np.random.RandomState(2)
    spec = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(r.ctype)
    ws = Workspace(shape, dtype, flags=_lib.WS_FFT)
    fn = getattr(_lib.lib, 'prh_spectral_' + ('f32' if dtype == 'f4' else 'f64'))
    nm, nm_ptr = _lib.size3(shape)
    bs, bs_ptr = _lib.vec3(boxsize, dtype)
    los = np.array([0.3, 0.4, np.sqrt(0.75)])
    los_c = (ctypes.c_double * 3)(*los)
    beta, norm = 0.4, 0.25
    tol = 1e-14 if dtype == 'f8' else 1e-6

    def run(op, a=0, b=0, use_los=False):
        src = _dev_complex(spec, dtype)
        dst = DeviceArray(shape + (2,), dtype)
        _lib.check(fn(ws.handle, op, src.ptr, dst.ptr, nm_ptr, bs_ptr, los_c if use_los else None, beta, norm, a, b,
                      None))
        return _host_complex(dst)

    k2 = sum(kk**2 for kk in k)
    k2[0, 0, 0] = 1.
    d = spec / k2
    mu2 = sum(kk * ll for ll, kk in zip(k, los))**2 / k2
    for a in range(3):
        assert rel_l2(run(_lib.SPEC_PLANE_PARALLEL, a=a, use_los=True),
                      norm * (1j * k[a] * d / (1. + beta * mu2))) < tol
        g = 1j * k[a] * d
        g[0, 0, 0] = 0.
        assert rel_l2(run(_lib.SPEC_GRAD, a=a), norm * g) < tol
        assert rel_l2(run(_lib.SPEC_I_KA, a=a), norm * (spec * 1j) * k[a]) < tol
        for b in range(a, 3):
            assert rel_l2(run(_lib.SPEC_KA_KB, a=a, b=b), norm * k[a] * k[b] * spec) < tol
    assert rel_l2(run(_lib.SPEC_LOS_SQ, use_los=True), norm * sum(kk * ll for ll, kk in zip(k, los))**2 * spec) < tol
    kk2 = (np.asarray(k[1]**2, dtype) + np.asarray(k[0]**2, dtype)) + np.asarray(k[2]**2, dtype)
    with np.errstate(divide='ignore', invalid='ignore'):
        inv = spec / kk2
    inv[0, 0, 0] = 0.
    assert rel_l2(run(_lib.SPEC_INV_K2), inv) < tol


@pytest.fixture(scope='module')
def c1():
    return make_catalogs('c1')


def _pair(cat, name, dtype, los):
    import pyrecon_amd
    from oracle import recon_numpy
    cfg = cat['config']
    geo = dict(nmesh=cfg['nmesh'], boxsize=cfg['boxsize'], boxcenter=cfg['boxcenter'], dtype=dtype)
    rkw = dict(RECON_KWARGS, los=los)
    gpu = getattr(pyrecon_amd, name)(**rkw, **geo)
    cpu = getattr(recon_numpy, 'Oracle' + name)(**rkw, **geo)
    for r in (gpu, cpu):
        r.assign_data(cat['data'], cat['data_weights'])
        r.assign_randoms(cat['randoms'], cat['randoms_weights'])
        r.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    return gpu, cpu


@pytest.mark.parametrize('dtype', ['f8', 'f4'])
@pytest.mark.parametrize('case', [('PlaneParallelFFTReconstruction', 'x'), ('IterativeFFTReconstruction', None),
                                  ('IterativeFFTReconstruction', 'z')])
def test_fft_reconstruction_vs_oracle(c1, dtype, case):
    name, los = case
    gpu, cpu = _pair(c1, name, dtype, los)
    gpu.run()
    cpu.run()
    for a in range(3):
        err = rel_l2(gpu.mesh_psi[a].value.astype('f8'), np.asarray(cpu.mesh_psi[a], dtype='f8'))
        assert err < (1e-10 if dtype == 'f8' else 1e-4), (a, err)
    for field in ['disp', 'rsd', 'disp+rsd']:
        xg = gpu.read_shifted_positions(c1['data'], field=field)
        xc = cpu.read_shifted_positions(c1['data'], field=field)
        assert xg.dtype == xc.dtype
        if dtype == 'f8':
            assert np.max(np.abs(xg - xc)) < 1e-8, field
        else:
            assert np.allclose(xg, xc, atol=1e-2, rtol=1e-2), field


def test_plane_parallel_needs_global_los():
    import pyrecon_amd
    with pytest.raises(pyrecon_amd.ReconstructionError):
        pyrecon_amd.PlaneParallelFFTReconstruction(f=0.8, bias=2., los=None, nmesh=16, boxsize=500., boxcenter=0.)


@pytest.mark.parametrize('dtype', ['f8', 'f4'])
@pytest.mark.parametrize('los', [None, 'x'])
def test_iterative_particle_vs_oracle(c1, dtype, los):
    # OriginalIterativeFFTParticleReconstruction (iterative_fft_particle.py:9-239): the data are repainted
    # at their reconstructed positions each iteration; positions, fields and shifts against the oracle
    import pyrecon_amd
    from oracle import recon_numpy
    cfg = c1['config']
    geo = dict(nmesh=cfg['nmesh'], boxsize=cfg['boxsize'], boxcenter=cfg['boxcenter'], dtype=dtype)
    rkw = dict(RECON_KWARGS, los=los)
    gpu = pyrecon_amd.IterativeFFTParticleReconstruction(**rkw, **geo)
    cpu = recon_numpy.OracleIterativeFFTParticleReconstruction(**rkw, **geo)
    for r in (gpu, cpu):
        r.assign_data(c1['data'], c1['data_weights'])
        r.assign_randoms(c1['randoms'], c1['randoms_weights'])
        r.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
        r.run(niterations=3)
    tol_x = 1e-8 if dtype == 'f8' else 1e-2
    assert np.max(np.abs(gpu._positions_rec_data - cpu._positions_rec_data)) < tol_x
    for a in range(3):
        err = rel_l2(gpu.mesh_psi[a].value.astype('f8'), np.asarray(cpu.mesh_psi[a], dtype='f8'))
        assert err < (1e-10 if dtype == 'f8' else 1e-4), (a, err)
    for positions in ['data', c1['randoms'][:20000]]:
        for field in ['disp', 'rsd', 'disp+rsd']:
            xg = gpu.read_shifted_positions(positions, field=field)
            xc = cpu.read_shifted_positions(positions, field=field)
            if dtype == 'f8':
                assert np.max(np.abs(xg - xc)) < tol_x, field
            else:
                assert np.allclose(xg, xc, atol=1e-2, rtol=1e-2), field
