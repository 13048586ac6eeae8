"""Configuration 4, the large mesh: nmesh=1024, boxsize 4000 Mpc/h, the configuration-3 shell
geometry, 5e7 data and 2.5e8 randoms, FKP weights, local line of sight.

At this size the FP64 oracle is the plain FP64 GPU path: it is the reference's algorithm bit for bit
(the paint, the multigrid and the read are checked bit-exact against the reference's C at
configurations 1 to 3), and a CPU run does not fit the test's time.  Against it:
- the mixed path (FP32 V-cycles inside FP64 defect correction): potential (mean removed)
  relative L2 <= 1e-5, shifted positions within 1e-3 Mpc/h over all data
  and a 2.5e6-point randoms sample;
- the CIC meshes of two paints of the same catalogs are bit-identical.
"""
import os
import time

import numpy as np
import pytest

from catalogs import make_catalogs, RECON_KWARGS, SMOOTHING_RADIUS

pytestmark = pytest.mark.gpu

T0 = time.time()


def say(msg):
    print('[config4 {:6.1f}s] {}'.format(time.time() - T0, msg), flush=True)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view('u{:d}'.format(a.dtype.itemsize))


@pytest.fixture(scope='module')
def c4():
    import pyrecon_amd as P
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    say('generating the catalogs ({:d} threads)'.format(workers))
    cat = make_catalogs('c4', workers=workers, progress=say)
    cfg = cat['config']
    geo = dict(nmesh=cfg['nmesh'], boxsize=cfg['boxsize'], boxcenter=cfg['boxcenter'], dtype='f8')
    dev = {k: P.DeviceArray.from_numpy(cat[k]) for k in ('data', 'data_weights', 'randoms', 'randoms_weights')}
    sample = P.DeviceArray.from_numpy(np.ascontiguousarray(cat['randoms'][::100]))
    del cat
    ws = P.Workspace((cfg['nmesh'],) * 3, 'f8', paint_chunk=1 << 28)
    say('catalogs on the device; workspace {:.1f} GB'.format(ws.nbytes / 1e9))
    yield geo, dev, sample, ws
    ws.close()
    P.empty_cache()


def _painted(c4, precision='plain'):
    import pyrecon_amd as P
    geo, dev, sample, ws = c4
    recon = P.MultiGridReconstruction(**RECON_KWARGS, **geo, workspace=ws, precision=precision)
    recon.assign_data(dev['data'], dev['data_weights'])
    recon.assign_randoms(dev['randoms'], dev['randoms_weights'])
    return recon


def _solved(c4, precision):
    geo, dev, sample, ws = c4
    recon = _painted(c4, precision)
    recon.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    recon.run()
    phi = recon.mesh_phi.value
    phi -= phi.mean()
    xd = recon.read_shifted_positions(dev['data']).numpy()
    xr = recon.read_shifted_positions(sample).numpy()
    return phi, xd, xr


@pytest.mark.parametrize('precision', ['plain', 'mixed'])
def test_config4_paint_repeatable(c4, precision):
    first = _painted(c4, precision)
    second = _painted(c4, precision)
    for name in ('mesh_data', 'mesh_randoms'):
        a = getattr(first, name).value
        b = getattr(second, name).value
        assert np.array_equal(bits(a), bits(b)), 'configuration 4: {} differs between two paints'.format(name)
        del a, b
    say('{} paint: bit-identical between two runs'.format(precision))


def test_config4_mixed_vs_fp64(c4):
    phi_p, xd_p, xr_p = _solved(c4, 'plain')
    say('plain FP64 solved')
    phi_m, xd_m, xr_m = _solved(c4, 'mixed')
    say('mixed solved')
    err_phi = np.sqrt(np.sum((phi_m - phi_p)**2) / np.sum(phi_p**2))
    del phi_m, phi_p
    d = np.concatenate([(xd_m - xd_p).ravel(), (xr_m - xr_p).ravel()])
    dx, rms = np.max(np.abs(d)), np.sqrt(np.mean(d**2))
    say('configuration 4, mixed against plain FP64: phi rel L2 {:.3e}, max |dx| {:.3e} Mpc/h, rms {:.3e}'.format(
        err_phi, dx, rms))
    assert err_phi < 1e-5
    assert dx <= 1e-3
