"""Workspaces of the drop-in classes.

Growth of the paint scratch (Workspace.reserve, prh_workspace_reserve_paint): a workspace that
starts with chunks of 1000 points and grows when the randoms (4000 points) are painted gives the
bits of one sized for both catalogs from the start -- meshes, potential and the shifted positions of
both catalogs, the data's read walking the paint order (and, mixed, the tile-order copy) that the
growth carried over.  A workspace with a fixed chunk does not grow and paints in several passes.
A released reconstruction's workspace is taken by the next one of the same kind, with the results
of a fresh workspace."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(np.asarray(a))
    return a.view('u{:d}'.format(a.dtype.itemsize))


def _catalogs():
    rng = np.random.RandomState(4)
    data = rng.uniform(-400., 400., size=(900, 3)) + np.array([0., 0., 1000.])
    wdata = rng.uniform(0.5, 1.5, size=900)
    randoms = rng.uniform(-400., 400., size=(4000, 3)) + np.array([0., 0., 1000.])
    return data, wdata, randoms


@pytest.mark.parametrize('precision', ['plain', 'mixed'])
def test_paint_scratch_grows(precision):
    from pyrecon_amd import MultiGridReconstruction
    from pyrecon_amd.mesh import Workspace
    data, wdata, randoms = _catalogs()
    out = {}
    for name, kwargs in (('sized', dict(paint_chunk=8192)), ('grown', dict(paint_chunk=1000, grow=True))):
        ws = Workspace((32, 32, 32), 'f8', **kwargs)
        rec = MultiGridReconstruction(f=0.8, bias=2., nmesh=32, boxsize=900., boxcenter=(0., 0., 1000.),
                                      precision=precision, workspace=ws)
        rec.assign_data(data, wdata)
        assert ws.paint_chunk == kwargs['paint_chunk']
        rec.assign_randoms(randoms)
        if name == 'grown':
            assert ws.paint_chunk >= len(randoms)
        meshes = (np.array(rec.mesh_data.value), np.array(rec.mesh_randoms.value))
        rec.set_density_contrast(smoothing_radius=15.)
        rec.run()
        out[name] = meshes + (np.array(rec.mesh_phi.value), rec.read_shifted_positions(data),
                              rec.read_shifted_positions(randoms))
    for a, b in zip(out['sized'], out['grown']):
        assert np.all(np.isfinite(a))
        assert np.array_equal(bits(a), bits(b))


def test_fixed_chunk_does_not_grow():
    from pyrecon_amd import MultiGridReconstruction
    from pyrecon_amd.mesh import Workspace
    data, wdata, randoms = _catalogs()
    ws = Workspace((32, 32, 32), 'f8', paint_chunk=1000)
    assert not ws.grow and not ws.reserve(len(randoms)) and ws.paint_chunk == 1000
    rec = MultiGridReconstruction(f=0.8, bias=2., nmesh=32, boxsize=900., boxcenter=(0., 0., 1000.), workspace=ws)
    rec.assign_randoms(randoms)  # four passes
    ref = MultiGridReconstruction(f=0.8, bias=2., nmesh=32, boxsize=900., boxcenter=(0., 0., 1000.))
    assert ref.workspace.grow
    ref.assign_randoms(randoms)
    assert ws.paint_chunk == 1000
    assert np.array_equal(bits(rec.mesh_randoms.value), bits(ref.mesh_randoms.value))


def test_released_workspace_reused():
    # a reconstruction on the default stream that is no longer referenced leaves its workspace for
    # the next one of the same kind (scratch only: the results are those of a fresh workspace)
    import gc
    from pyrecon_amd import MultiGridReconstruction, empty_cache
    from pyrecon_amd.mesh import Workspace
    data, wdata, randoms = _catalogs()
    kwargs = dict(f=0.8, bias=2., nmesh=32, boxsize=900., boxcenter=(0., 0., 1000.), precision='mixed')

    def reconstruct(rec):
        rec.assign_data(data, wdata)
        rec.assign_randoms(randoms)
        rec.set_density_contrast(smoothing_radius=15.)
        rec.run()
        return rec.read_shifted_positions(data), rec.read_shifted_positions(randoms)

    empty_cache()
    first = MultiGridReconstruction(**kwargs)
    handle = first.workspace.handle
    out_first = reconstruct(first)
    del first
    gc.collect()
    second = MultiGridReconstruction(**kwargs)
    assert second.workspace.handle == handle
    out_second = reconstruct(second)
    fresh = MultiGridReconstruction(workspace=Workspace((32, 32, 32), 'f8'), **kwargs)
    out_fresh = reconstruct(fresh)
    for a, b, c in zip(out_first, out_second, out_fresh):
        assert np.array_equal(bits(a), bits(b)) and np.array_equal(bits(a), bits(c))
    del second
    gc.collect()
    empty_cache()  # destroys the kept workspace
    third = MultiGridReconstruction(**kwargs)
    assert third.workspace.handle is not None
