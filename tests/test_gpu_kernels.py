"""Parity of the reference-shaped HIP kernels with the CPU oracle.

The oracle (oracle/orc_kernels.c) is itself pinned bit-for-bit to the
reference's C sources (tests/test_oracle_kernels.py), so equality here means
equality with the reference's src/mesh.c and src/multigrid.c.  The bar is
bit-exact: the HIP kernels evaluate every expression in the reference's order
without fused multiply-adds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(8, 8, 8), (8, 12, 16), (16, 16, 16), (6, 10, 12), (32, 32, 32), (24, 40, 20),
          # shapes taking the blocked kernels: 64-wide tiles (n2 % 64 == 0, n1 % 8 == 0, incl. a partial
          # x run) and, in FP32, whole rows of 64, 128, 256 and 512 cells (n1 % 8 == 0) and of 1024 cells
          # (n1 % 4 == 0)
          (6, 16, 64), (10, 32, 128), (64, 64, 64), (6, 16, 256), (4, 8, 512), (9, 24, 512),
          (4, 8, 1024), (3, 4, 1024), (5, 12, 1024)]
LOS = [None, [0., 0., 1.], [0.3, 0.4, np.sqrt(1 - 0.25)]]
BOX = dict(boxsize=np.array([1000., 1200., 900.]), boxcenter=np.array([0., 100., 1500.]))


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view('u{:d}'.format(a.dtype.itemsize))


def assert_bitexact(a, b, what=''):
    assert a.shape == b.shape and a.dtype == b.dtype, (what, a.shape, b.shape, a.dtype, b.dtype)
    ne = np.count_nonzero(bits(a) != bits(b))
    assert ne == 0, '{}: {} of {} values differ, max |diff| {}'.format(what, ne, a.size, np.max(np.abs(a - b)))


def _case(rng, shape, dtype):
    f = rng.normal(size=shape).astype(dtype)
    v = rng.normal(size=shape).astype(dtype)
    return f, v


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', SHAPES)
@pytest.mark.parametrize('ilos', range(len(LOS)))
def test_jacobi_residual(gpu, oracle, dtype, shape, ilos):
    rng = np.random.RandomState(1 + ilos)
    f, v = _case(rng, shape, dtype)
    los = LOS[ilos]
    a, b = v.copy(), v.copy()
    oracle.jacobi(a, f, BOX['boxsize'], BOX['boxcenter'], 0.4, 0.4, 3, los)
    gpu.jacobi(b, f, BOX['boxsize'], BOX['boxcenter'], 0.4, 0.4, 3, los)
    assert_bitexact(a, b, 'jacobi')
    assert_bitexact(oracle.residual(v, f, BOX['boxsize'], BOX['boxcenter'], 0.4, los),
                    gpu.residual(v, f, BOX['boxsize'], BOX['boxcenter'], 0.4, los), 'residual')


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', [s for s in SHAPES if all(n % 2 == 0 for n in s)])
def test_reduce_prolong(gpu, oracle, dtype, shape):
    rng = np.random.RandomState(3)
    fine = rng.normal(size=shape).astype(dtype)
    assert_bitexact(oracle.reduce(fine), gpu.reduce(fine), 'reduce')
    coarse = rng.normal(size=tuple(n // 2 for n in shape)).astype(dtype)
    assert_bitexact(oracle.prolong(coarse), gpu.prolong(coarse), 'prolong')


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', [(8, 8, 8), (16, 16, 16), (8, 12, 16), (24, 40, 20), (32, 32, 32)])
@pytest.mark.parametrize('ilos', [0, 2])
def test_vcycle_fmg(gpu, oracle, dtype, shape, ilos):
    rng = np.random.RandomState(5)
    f, v = _case(rng, shape, dtype)
    los = LOS[ilos]
    a, b = v.copy(), v.copy()
    oracle.vcycle(a, f, BOX['boxsize'], BOX['boxcenter'], 0.4, 0.4, 5, los)
    gpu.vcycle(b, f, BOX['boxsize'], BOX['boxcenter'], 0.4, 0.4, 5, los)
    assert_bitexact(a, b, 'vcycle')
    for niter, ncycle in [(5, 6), (2, 1), (3, 2)]:
        va = oracle.fmg(f, BOX['boxsize'], BOX['boxcenter'], 0.4, 0.4, niter, ncycle, los)
        vb = gpu.fmg(f, BOX['boxsize'], BOX['boxcenter'], 0.4, 0.4, niter, ncycle, los)
        assert_bitexact(va, vb, 'fmg niter={} ncycle={}'.format(niter, ncycle))


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
def test_fmg_64(gpu, oracle, dtype):
    rng = np.random.RandomState(7)
    shape = (64, 64, 64)
    f = rng.normal(size=shape).astype(dtype)
    for los in [None, [0., 0., 1.]]:
        va = oracle.fmg(f, 1000., [0., 0., 1500.], 0.4, los=los)
        vb = gpu.fmg(f, 1000., [0., 0., 1500.], 0.4, los=los)
        assert_bitexact(va, vb, 'fmg 64^3')


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
def test_fmg_128(gpu, oracle, dtype):
    # blocked kernels on two levels (128^3 and 64^3), zero-start sweeps on a blocked level, and the
    # row-wise restriction and prolongation at full size
    rng = np.random.RandomState(8)
    shape = (128, 128, 128)
    f = rng.normal(size=shape).astype(dtype)
    for los in [None, [0., 0., 1.]]:
        va = oracle.fmg(f, 1000., [0., 0., 1500.], 0.4, los=los)
        vb = gpu.fmg(f, 1000., [0., 0., 1500.], 0.4, los=los)
        assert_bitexact(va, vb, 'fmg 128^3')


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', [(8, 8, 8), (16, 12, 10), (2, 2, 2), (64, 64, 64), (2, 4, 1024)])
@pytest.mark.parametrize('chunk', [1 << 16, 1000])
def test_assign_cic(gpu, oracle, dtype, shape, chunk):
    rng = np.random.RandomState(11)
    n = 20000
    pos = (rng.uniform(0., 1., size=(n, 3)) * np.array(shape)).astype(dtype)
    pos[:4] = np.array(shape)  # upper edge: d = n, the reference's quirk
    pos[4:8] = 0.
    w = rng.uniform(0.5, 1.5, n).astype(dtype)
    mesh0 = rng.normal(size=shape).astype(dtype)  # accumulation into an existing mesh
    a, b = mesh0.copy(), mesh0.copy()
    oracle.assign_cic(a, pos, w)
    gpu.assign_cic(b, pos, w, chunk=chunk)
    assert_bitexact(a, b, 'assign_cic')


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
def test_assign_cic_dense_cells(gpu, oracle, dtype):
    # many points per cell: exercises the long-group ordering paths (LDS sort and rank placement)
    # and the gather's staging: whole 9^3 region, two x-halves, and the global-memory merge
    rng = np.random.RandomState(12)
    cases = [((4, 4, 4), 300), ((4, 4, 4), 5000), ((4, 4, 4), 40000),
             ((16, 16, 16), 14000), ((16, 16, 16), 20000), ((16, 16, 16), 30000)]
    for shape, n in cases:
        pos = (rng.uniform(0., 1., size=(n, 3)) * np.array(shape)).astype(dtype)
        w = rng.uniform(0.5, 1.5, n).astype(dtype)
        a, b = np.zeros(shape, dtype), np.zeros(shape, dtype)
        oracle.assign_cic(a, pos, w)
        gpu.assign_cic(b, pos, w, chunk=1 << 17)
        assert_bitexact(a, b, 'assign_cic dense n={}'.format(n))


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', [(8, 8, 8), (16, 12, 10), (64, 64, 64), (4, 2, 1024)])
def test_reads(gpu, oracle, dtype, shape):
    rng = np.random.RandomState(13)
    n = 50000
    mesh = rng.normal(size=shape).astype(dtype)
    pos = (rng.uniform(0., 1., size=(n, 3)) * np.array(shape)).astype(dtype)
    pos[:3] = np.array(shape)
    assert_bitexact(oracle.read_cic(mesh, pos), gpu.read_cic(mesh, pos), 'read_cic')
    boxsize = np.array([1000., 800., 900.])
    assert_bitexact(oracle.read_finite_difference_cic(mesh, boxsize, pos),
                    gpu.read_finite_difference_cic(mesh, boxsize, pos), 'read_finite_difference_cic')


SMOOTH_SHAPES = [(8, 8, 8), (8, 12, 20), (6, 10, 14), (15, 9, 25), (10, 32, 128), (64, 64, 64), (96, 80, 112),
                 # 512-cell lines: the register-resident transforms along z (an even and an odd number
                 # of rows), x and y (16 adjacent rows per tile), and x with a z extent that is not a
                 # multiple of 16 (the LDS transform)
                 (4, 6, 512), (3, 5, 512), (512, 2, 16), (2, 512, 32), (512, 3, 5),
                 # 1024-cell lines: in registers along z (an even and an odd number of rows) and
                 # along x and y (16 adjacent rows per tile), and the LDS transform along x and y
                 # with a z extent that is not a multiple of 16
                 (2, 4, 1024), (3, 5, 1024), (1024, 2, 16), (2, 1024, 32), (1024, 2, 8), (4, 1024, 2),
                 # prime factors above 7: the direct DFT along those axes (11, 13, 17, 2 x 11 x 13,
                 # 131), next to radix axes
                 (11, 13, 17), (286, 4, 6), (8, 131, 10)]


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', SMOOTH_SHAPES)
def test_smooth_gaussian(dtype, shape):
    # separable Fourier-space smoothing (mixed-radix FFTs along each axis) against the oracle's
    # numpy c2c 3-D transform (pyrecon/mesh.py:655-680); shapes cover radices 2, 3, 4, 5, 7
    from pyrecon_amd import RealMesh, MeshInfo
    from oracle.recon_numpy import OracleMultiGridReconstruction
    rng = np.random.RandomState(sum(shape))
    boxsize, boxcenter = np.array([800., 700., 1100.]), np.array([0., 0., 1500.])
    value = rng.uniform(0., 3., shape).astype(dtype)
    cpu = OracleMultiGridReconstruction(nmesh=shape, boxsize=boxsize, boxcenter=boxcenter, dtype=dtype)
    expect = np.asarray(cpu.smooth_gaussian(value.astype('f8') if dtype == 'f8' else value, 30.), dtype='f8')
    mesh = RealMesh(info=MeshInfo(nmesh=shape, boxsize=boxsize, boxcenter=boxcenter, dtype=dtype))
    mesh.value = value
    mesh.smooth_gaussian(30.)
    got = mesh.value.astype('f8')
    err = np.sqrt(np.sum((got - expect)**2) / np.sum(expect**2))
    assert err < (1e-13 if dtype == 'f8' else 2e-6), err
    # the k = 0 mode has gain 1: the sum is conserved
    assert np.isclose(got.sum(), value.astype('f8').sum(), rtol=1e-12 if dtype == 'f8' else 1e-5)


def test_smooth_gaussian_unsupported_size():
    # a dimension longer than the direct DFT's LDS line (4096) with a prime factor above 7: the
    # call fails cleanly
    from pyrecon_amd import RealMesh, MeshInfo
    mesh = RealMesh(info=MeshInfo(nmesh=(4099, 2, 2), boxsize=100., boxcenter=0., dtype='f8'))
    mesh.value = np.ones((4099, 2, 2))
    with pytest.raises(ValueError):
        mesh.smooth_gaussian(10.)
