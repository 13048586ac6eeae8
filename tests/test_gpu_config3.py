"""Configuration 3 end to end, the headline workload: nmesh=512, boxsize 4000 Mpc/h, boxcenter
(0, 0, 3000), 2e7 Zel'dovich data and 1e8 randoms in the shell 1000 < r < 3000 Mpc/h with FKP
weights, local line of sight, f = 0.8, bias = 2, smoothing radius 15, the reference's solver
defaults.

The oracle (oracle/recon_numpy.py; its C kernels are the reference's bit for bit) runs the whole
FP64 reconstruction on the CPU from the same seeded catalogs, as the reference's integration
check does on full catalogs (pyrecon/tests/test_multigrid.py:183-239).  Against it:
- CIC meshes of data and randoms: bit-exact, and bit-identical between two GPU paints;
- density contrast: relative L2 <= 1e-12 (separable device smoothing vs pocketfft c2c);
- potential (mean removed): relative L2 <= 1e-10; shifted positions <= 1e-8 Mpc/h;
- the GPU multigrid given the oracle's density contrast: the oracle's potential bit for bit;
- the GPU read given the oracle's potential: the oracle's shifted positions bit for bit;
- the mixed path (CIC weights summed in fixed point per tile, FP32 V-cycles inside FP64 defect
  correction): paint meshes identical between two runs and within 1e-6 (relative to the largest
  cell) of the FP64 paint; shifted positions within 1e-3 Mpc/h of the oracle,
  over all data and a 1e6-point randoms sample.
"""
import json
import os
import time

import numpy as np
import pytest

from catalogs import make_catalogs, sha256, RECON_KWARGS, SMOOTHING_RADIUS

pytestmark = pytest.mark.gpu

T0 = time.time()


def say(msg):
    print('[config3 {:6.1f}s] {}'.format(time.time() - T0, msg), flush=True)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view('u{:d}'.format(a.dtype.itemsize))


def rel_l2(a, b, remove_mean=False):
    a, b = np.asarray(a, dtype='f8'), np.asarray(b, dtype='f8')
    if remove_mean:
        a, b = a - a.mean(), b - b.mean()
    return np.sqrt(np.sum((a - b)**2) / np.sum(b**2))


def _workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope='module')
def c3():
    from oracle.recon_numpy import OracleMultiGridReconstruction
    workers = _workers()
    say('generating the catalogs ({:d} threads)'.format(workers))
    cat = make_catalogs('c3', workers=workers, progress=say)
    fixture = json.load(open(os.path.join(os.path.dirname(__file__), 'golden', 'catalogs_sha256.json')))['c3']
    assert sha256(cat['data'], cat['data_weights']) == fixture['data'], 'configuration-3 data catalog changed'
    assert sha256(cat['randoms'], cat['randoms_weights']) == fixture['randoms'], 'configuration-3 randoms changed'
    say('catalogs match their SHA-256 pins')
    cfg = cat['config']
    geo = dict(nmesh=cfg['nmesh'], boxsize=cfg['boxsize'], boxcenter=cfg['boxcenter'])
    sample = cat['randoms'][::100]  # 1e6 randoms
    cpu = OracleMultiGridReconstruction(**RECON_KWARGS, **geo, dtype='f8', fft_workers=workers)
    cpu.assign_data(cat['data'], cat['data_weights'])
    cpu.assign_randoms(cat['randoms'], cat['randoms_weights'])
    say('oracle paint')
    meshes = (cpu.mesh_data.copy(), cpu.mesh_randoms.copy())
    cpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    delta = cpu.mesh_delta.copy()
    say('oracle density contrast')
    cpu.run()
    say('oracle multigrid')
    ref = dict(meshes=meshes, delta=delta, phi=cpu.mesh_phi, data=cpu.read_shifted_positions(cat['data']),
               randoms=cpu.read_shifted_positions(sample), alpha=cpu.alpha, mask_count=cpu.mask_count)
    say('oracle shifts')
    return cat, geo, sample, ref


@pytest.fixture(scope='module')
def device_catalogs(c3):
    from pyrecon_amd.device import DeviceArray
    cat = c3[0]
    return {k: DeviceArray.from_numpy(cat[k]) for k in ('data', 'data_weights', 'randoms', 'randoms_weights')}


def _painted(geo, dev, precision='plain'):
    import pyrecon_amd
    gpu = pyrecon_amd.MultiGridReconstruction(**RECON_KWARGS, **geo, dtype='f8', precision=precision)
    gpu.assign_data(dev['data'], dev['data_weights'])
    gpu.assign_randoms(dev['randoms'], dev['randoms_weights'])
    return gpu


def test_config3_paint_and_density(c3, device_catalogs):
    cat, geo, sample, ref = c3
    gpu = _painted(geo, device_catalogs)
    meshes = (gpu.mesh_data.value, gpu.mesh_randoms.value)
    for name, a, b in zip(('data', 'randoms'), meshes, ref['meshes']):
        assert np.array_equal(bits(a), bits(b)), 'configuration 3: {} mesh not bit-exact to the oracle'.format(name)
    again = _painted(geo, device_catalogs)
    for name, a in zip(('data', 'randoms'), (again.mesh_data.value, again.mesh_randoms.value)):
        b = meshes[0] if name == 'data' else meshes[1]
        assert np.array_equal(bits(a), bits(b)), 'configuration 3: {} mesh differs between two paints'.format(name)
    del again
    say('paint: bit-exact to the oracle and between runs')
    gpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    err = rel_l2(gpu.mesh_delta.value, ref['delta'])
    say('density contrast: rel L2 {:.3e}; alpha {!r} vs {!r}'.format(err, gpu.alpha, ref['alpha']))
    assert err < 1e-12
    assert abs(gpu.alpha / ref['alpha'] - 1.) < 1e-13


def test_config3_fp64(c3, device_catalogs):
    cat, geo, sample, ref = c3
    gpu = _painted(geo, device_catalogs)
    gpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    gpu.run()
    err_phi = rel_l2(gpu.mesh_phi.value, ref['phi'], remove_mean=True)
    xd = gpu.read_shifted_positions(device_catalogs['data']).numpy()
    xr = gpu.read_shifted_positions(sample)
    dx = max(np.max(np.abs(xd - ref['data'])), np.max(np.abs(xr - ref['randoms'])))
    say('FP64 against the oracle: phi rel L2 {:.3e}, max |dx| {:.3e} Mpc/h'.format(err_phi, dx))
    assert err_phi < 1e-10
    assert dx < 1e-8


def test_config3_fmg_and_read_bitexact(c3):
    """Given the oracle's density contrast, the GPU multigrid returns the oracle's potential bit for
    bit; given that potential, the GPU read returns the oracle's shifted positions bit for bit."""
    import pyrecon_amd
    cat, geo, sample, ref = c3
    gpu = pyrecon_amd.MultiGridReconstruction(**RECON_KWARGS, **geo, dtype='f8')
    gpu.mesh_delta = pyrecon_amd.RealMesh(ref['delta'], info=gpu.info, workspace=gpu.workspace)
    gpu.run()
    assert np.array_equal(bits(gpu.mesh_phi.value), bits(ref['phi'])), 'configuration 3: potential not bit-exact'
    xd = gpu.read_shifted_positions(cat['data'])
    xr = gpu.read_shifted_positions(sample)
    assert np.array_equal(bits(xd), bits(ref['data']))
    assert np.array_equal(bits(xr), bits(ref['randoms']))
    say('multigrid and read: bit-exact to the oracle')


def test_config3_mixed(c3, device_catalogs):
    cat, geo, sample, ref = c3
    gpu = _painted(geo, device_catalogs, 'mixed')
    # the tile-sum paint: identical bits on a repeated run, within the FP64 rounding of the serial paint
    again = _painted(geo, device_catalogs, 'mixed')
    for name, exact in zip(('data', 'randoms'), ref['meshes']):
        a, b = getattr(gpu, 'mesh_' + name).value, getattr(again, 'mesh_' + name).value
        assert np.array_equal(bits(a), bits(b)), 'configuration 3: mixed {} mesh differs between two paints'.format(name)
        err = np.max(np.abs(a - exact)) / np.max(np.abs(exact))
        say('mixed paint, {}: max |mesh - FP64 mesh| / max |mesh| = {:.3e}'.format(name, err))
        assert err < 1e-13
    del again
    gpu.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    # the density rule's mask (randoms above the threshold) is the oracle's, cell for cell
    flips = int(np.count_nonzero((gpu.mesh_delta.value == 0.) != (ref['delta'] == 0.)))
    say('mixed density contrast: {:d} cells with the other mask value'.format(flips))
    assert flips == 0
    gpu.run()
    err_phi = rel_l2(gpu.mesh_phi.value, ref['phi'], remove_mean=True)
    xd = gpu.read_shifted_positions(device_catalogs['data']).numpy()
    xr = gpu.read_shifted_positions(sample)
    d = np.concatenate([(xd - ref['data']).ravel(), (xr - ref['randoms']).ravel()])
    dx, rms = np.max(np.abs(d)), np.sqrt(np.mean(d**2))
    say('mixed against the FP64 oracle: phi rel L2 {:.3e}, max |dx| {:.3e} Mpc/h, rms {:.3e}'.format(err_phi, dx, rms))
    assert err_phi < 1e-5
    assert dx <= 1e-3
