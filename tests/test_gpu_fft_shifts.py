"""The FFT reconstructions' shifts without wrap run their RSD algebra on the device
(prh_shift_algebra_*, pyrecon/recon.py:255-325): the same bits as the reference's numpy algebra on the
host (the class's own host form), for every field, shifted and not, the local and a global line of
sight, FP32 and FP64 positions and meshes; device positions give device results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(np.asarray(a))
    return a.view('u{:d}'.format(a.dtype.itemsize))


def _recon(cls, dtype, los):
    rng = np.random.RandomState(11)
    center = np.array([0., 0., 900.])
    data = rng.uniform(-300., 300., size=(4000, 3)) + center
    randoms = rng.uniform(-300., 300., size=(16000, 3)) + center
    rec = cls(f=0.8, bias=2., los=los, nmesh=32, boxsize=700., boxcenter=center, dtype=dtype)
    rec.assign_data(data, rng.uniform(0.5, 1.5, size=len(data)))
    rec.assign_randoms(randoms)
    rec.set_density_contrast(smoothing_radius=20.)
    rec.run()
    return rec, randoms


@pytest.mark.parametrize('algorithm, los', [('iterative', None), ('iterative', 'x'), ('plane', 'z')])
@pytest.mark.parametrize('dtype', ['f8', 'f4'])
@pytest.mark.parametrize('pdtype', ['f8', 'f4'])
def test_device_algebra_same_bits(algorithm, los, dtype, pdtype):
    from pyrecon_amd import IterativeFFTReconstruction, PlaneParallelFFTReconstruction, DeviceArray
    cls = IterativeFFTReconstruction if algorithm == 'iterative' else PlaneParallelFFTReconstruction
    rec, randoms = _recon(cls, dtype, los)
    pos = randoms[:5000].astype(pdtype)
    for field in ['disp', 'rsd', 'disp+rsd']:
        host = rec._host_shifts(pos, field)
        dev = rec.read_shifts(pos, field=field)
        assert isinstance(dev, np.ndarray) and dev.dtype == host.dtype, (field, dev.dtype, host.dtype)
        assert np.array_equal(bits(dev), bits(host)), field
        shifted = rec.read_shifted_positions(pos, field=field)
        expect = pos - host
        assert shifted.dtype == expect.dtype and np.array_equal(bits(shifted), bits(expect)), field
    # device positions: device results, the same values
    out = rec.read_shifted_positions(DeviceArray.from_numpy(pos))
    assert isinstance(out, DeviceArray)
    assert np.array_equal(bits(out.numpy()), bits(pos - rec._host_shifts(pos, 'disp+rsd')))
