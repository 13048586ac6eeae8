"""Configuration 2 end to end: nmesh=256, boxsize 2000 Mpc/h, boxcenter (0, 0, 2000), 2e6 Zel'dovich
data with weights uniform in [0.5, 1.5] and 2e7 uniform randoms, local line of sight.

The plain FP64 and FP32 GPU paths, and the mixed path, are each checked against the oracle's FP64
run on identical inputs (oracle/recon_numpy.py, whose C kernels are the reference's bit for bit):
- CIC meshes: the FP64 paint is bit-exact to the oracle, and two GPU paints are bitwise identical;
- FP64: potential (mean removed) relative L2 <= 1e-10, shifted positions <= 1e-8 Mpc/h;
- mixed: shifted positions <= 1e-3 Mpc/h;
- FP32: shifted positions within the reference's own f4-vs-f8 tolerance
  (pyrecon/tests/test_multigrid.py:94, atol = rtol = 1e-2) of the FP64 run, CIC meshes bit-exact
  to the oracle's FP32 paint, and shifted positions close to the oracle's FP32 run.
"""
import numpy as np
import pytest

from catalogs import make_catalogs, RECON_KWARGS, SMOOTHING_RADIUS

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view('u{:d}'.format(a.dtype.itemsize))


def rel_l2(a, b):
    a, b = np.asarray(a, dtype='f8'), np.asarray(b, dtype='f8')
    a, b = a - a.mean(), b - b.mean()
    return np.sqrt(np.sum((a - b)**2) / np.sum(b**2))


@pytest.fixture(scope='module')
def c2():
    cat = make_catalogs('c2', workers=16)
    cfg = cat['config']
    geo = dict(nmesh=cfg['nmesh'], boxsize=cfg['boxsize'], boxcenter=cfg['boxcenter'])
    from oracle.recon_numpy import OracleMultiGridReconstruction
    cpu = OracleMultiGridReconstruction(**RECON_KWARGS, **geo, dtype='f8')
    cpu.assign_data(cat['data'], cat['data_weights'])
    cpu.assign_randoms(cat['randoms'], cat['randoms_weights'])
    meshes = (cpu.mesh_data.copy(), cpu.mesh_randoms.copy())
    cpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    cpu.run()
    sample = cat['randoms'][::10]
    ref = dict(meshes=meshes, phi=cpu.mesh_phi, data=cpu.read_shifted_positions(cat['data']),
               randoms=cpu.read_shifted_positions(sample))
    return cat, geo, sample, ref


def _gpu(cat, geo, dtype, precision='plain'):
    import pyrecon_amd
    gpu = pyrecon_amd.MultiGridReconstruction(**RECON_KWARGS, **geo, dtype=dtype, precision=precision)
    gpu.assign_data(cat['data'], cat['data_weights'])
    gpu.assign_randoms(cat['randoms'], cat['randoms_weights'])
    return gpu


def test_config2_fp64(c2):
    cat, geo, sample, ref = c2
    gpu = _gpu(cat, geo, 'f8')
    meshes = (gpu.mesh_data.value, gpu.mesh_randoms.value)
    for a, b in zip(meshes, ref['meshes']):
        assert np.array_equal(bits(a), bits(b)), 'CIC mesh not bit-exact to the oracle'
    again = _gpu(cat, geo, 'f8')
    for a, b in zip(meshes, (again.mesh_data.value, again.mesh_randoms.value)):
        assert np.array_equal(bits(a), bits(b)), 'CIC mesh differs between two GPU runs'
    gpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    gpu.run()
    err_phi = rel_l2(gpu.mesh_phi.value, ref['phi'])
    dx = max(np.max(np.abs(gpu.read_shifted_positions(cat['data']) - ref['data'])),
             np.max(np.abs(gpu.read_shifted_positions(sample) - ref['randoms'])))
    print('configuration 2, FP64: phi rel L2 {:.3e}, max |dx| {:.3e} Mpc/h'.format(err_phi, dx))
    assert err_phi < 1e-10
    assert dx < 1e-8


def test_config2_mixed(c2):
    cat, geo, sample, ref = c2
    gpu = _gpu(cat, geo, 'f8', precision='mixed')
    gpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    gpu.run()
    err_phi = rel_l2(gpu.mesh_phi.value, ref['phi'])
    dx = max(np.max(np.abs(gpu.read_shifted_positions(cat['data']) - ref['data'])),
             np.max(np.abs(gpu.read_shifted_positions(sample) - ref['randoms'])))
    print('configuration 2, mixed: phi rel L2 {:.3e}, max |dx| {:.3e} Mpc/h'.format(err_phi, dx))
    assert err_phi < 1e-5
    assert dx <= 1e-3


def test_config2_fp32(c2):
    cat, geo, sample, ref = c2
    gpu = _gpu(cat, geo, 'f4')
    meshes = (gpu.mesh_data.value, gpu.mesh_randoms.value)
    again = _gpu(cat, geo, 'f4')
    for a, b in zip(meshes, (again.mesh_data.value, again.mesh_randoms.value)):
        assert np.array_equal(bits(a), bits(b)), 'FP32 CIC mesh differs between two GPU runs'
    gpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    gpu.run()
    err_phi = rel_l2(gpu.mesh_phi.value, ref['phi'])
    xd, xr = gpu.read_shifted_positions(cat['data']), gpu.read_shifted_positions(sample)
    dx = max(np.max(np.abs(xd - ref['data'])), np.max(np.abs(xr - ref['randoms'])))
    rms = np.sqrt(np.mean(np.concatenate([(xd - ref['data']).ravel(), (xr - ref['randoms']).ravel()])**2))
    print('configuration 2, FP32 vs FP64: phi rel L2 {:.3e}, max |dx| {:.3e} Mpc/h, rms {:.3e}'.format(err_phi, dx, rms))
    assert np.allclose(xd, ref['data'], atol=1e-2, rtol=1e-2)
    assert np.allclose(xr, ref['randoms'], atol=1e-2, rtol=1e-2)
    # against the oracle's own FP32 run: the same algorithm (the paint and the multigrid are the
    # reference's bit for bit, including the upper-edge weights of the two randoms whose FP32 mesh
    # coordinate rounds to n); only the Fourier smoothing rounds differently
    from oracle.recon_numpy import OracleMultiGridReconstruction
    cpu = OracleMultiGridReconstruction(**RECON_KWARGS, **geo, dtype='f4')
    cpu.assign_data(cat['data'], cat['data_weights'])
    cpu.assign_randoms(cat['randoms'], cat['randoms_weights'])
    assert np.array_equal(bits(meshes[0]), bits(cpu.mesh_data)) and np.array_equal(bits(meshes[1]), bits(cpu.mesh_randoms))
    cpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    cpu.run()
    d32 = np.concatenate([(xd - cpu.read_shifted_positions(cat['data'])).ravel(),
                          (xr - cpu.read_shifted_positions(sample)).ravel()])
    dx32, rms32 = np.max(np.abs(d32)), np.sqrt(np.mean(d32**2))
    print('configuration 2, FP32 vs the FP32 oracle: max |dx| {:.3e} Mpc/h, rms {:.3e}'.format(dx32, rms32))
    assert rms32 < 5e-4 and dx32 < 1e-2
