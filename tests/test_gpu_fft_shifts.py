"""The FFT reconstructions' shifts without wrap run their RSD algebra on the device
(prh_shift_algebra_*, pyrecon/recon.py:255-325): the same bits as the reference's numpy algebra on the
host (the class's own host form), for every field, shifted and not, the local and a global line of
sight, FP32 and FP64 positions and meshes; device positions give device results.  Repeated reads of
the same positions share one binning of the points, which a paint in between invalidates."""
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


def test_read_bins_reused_and_invalidated():
    # the three component reads of one set of positions share one binning; a paint in between
    # overwrites the memory that holds it, and the next read bins again: the same values throughout
    from pyrecon_amd import IterativeFFTReconstruction, DeviceArray
    rec, randoms = _recon(IterativeFFTReconstruction, 'f8', None)
    pos = DeviceArray.from_numpy(randoms[:7000])
    first = [psi.read_cic(pos).numpy() for psi in rec.mesh_psi]
    again = [psi.read_cic(pos).numpy() for psi in rec.mesh_psi]
    from pyrecon_amd.mesh import RealMesh
    scratch = RealMesh(info=rec.info, workspace=rec.workspace)
    scratch.assign_cic(randoms[3000:9000])  # a paint through the same workspace
    after = [psi.read_cic(pos).numpy() for psi in rec.mesh_psi]
    fresh = [psi.read_cic(randoms[:7000]) for psi in rec.mesh_psi]
    for a, b, c, d in zip(first, again, after, fresh):
        assert np.array_equal(bits(a), bits(b)) and np.array_equal(bits(a), bits(c)) and np.array_equal(bits(a), bits(d))
