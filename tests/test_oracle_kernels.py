"""Pins the CPU oracle (test infrastructure) before it is trusted as the checker.

1. Bit-for-bit against the reference's own C (src/mesh.c, src/multigrid.c built
   into oracle/_ref by oracle/Makefile), on random inputs, f4 and f8, local and
   global line of sight, cubic and non-cubic meshes, including the upper-edge
   CIC quirk (position exactly n).
2. Against what the mathematics fixes, independently of the reference:
   - CIC paint equals a brute-force np.add.at deposit and conserves the weight
     sum (the reference's own test_cic, pyrecon/tests/test_mesh.py:17-66);
   - CIC read reproduces a constant and a brute-force interpolation;
   - finite-difference read of a constant is 0 (test_mesh.py:69-80) and of a
     single Fourier mode is the closed-form central difference;
   - global line of sight: a Fourier mode is an eigenvector of the discrete
     operator with the closed-form eigenvalue
     lambda(k) = sum_i (1 + beta l_i^2)(2 - 2 cos k_i h_i)/h_i^2
                 + 2 beta sum_{i<j} l_i l_j sin(k_i h_i) sin(k_j h_j)/(h_i h_j),
     so residual(v=mode, f=0) = -lambda v and one damped Jacobi sweep scales it
     by 1 - omega lambda / D;
   - restriction preserves constants and annihilates the Nyquist checkerboard;
     prolongation preserves constants;
   - the full multigrid converges to the discrete solution (solved exactly with
     FFTs through lambda(k)) for a global line of sight.
Each check is chosen so that a plausible slip (a dropped term, a sign, a
transposed index) fails it.
"""
import numpy as np
import pytest

from oracle.cpu import OracleKernels, RefKernels

SHAPES = [(8, 8, 8), (8, 12, 16), (6, 10, 12), (16, 16, 16)]
LOS = [None, [0., 0., 1.], [0.3, 0.4, np.sqrt(1 - 0.25)]]
BOXSIZE = np.array([1000., 1200., 900.])
BOXCENTER = np.array([0., 100., 1500.])


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view('u{:d}'.format(a.dtype.itemsize))


def same(a, b):
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(bits(a), bits(b))


@pytest.fixture(scope='module')
def orc():
    return OracleKernels()


# ---------------------------------------------------------------------------
# 1. bit-exact against the reference's C
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', SHAPES)
@pytest.mark.parametrize('ilos', range(len(LOS)))
def test_multigrid_vs_reference(orc, ref, dtype, shape, ilos):
    rng = np.random.RandomState(10 + ilos)
    los = LOS[ilos]
    f = rng.normal(size=shape).astype(dtype)
    v0 = rng.normal(size=shape).astype(dtype)
    a, b = v0.copy(), v0.copy()
    orc.jacobi(a, f, BOXSIZE, BOXCENTER, 0.4, 0.4, 3, los)
    ref.jacobi(b, f, BOXSIZE, BOXCENTER, 0.4, 0.4, 3, los)
    assert same(a, b)
    assert same(orc.residual(v0, f, BOXSIZE, BOXCENTER, 0.4, los), ref.residual(v0, f, BOXSIZE, BOXCENTER, 0.4, los))
    if all(n % 2 == 0 for n in shape):
        assert same(orc.reduce(f), ref.reduce(f))
        c = rng.normal(size=tuple(n // 2 for n in shape)).astype(dtype)
        assert same(orc.prolong(c), ref.prolong(c))
    a, b = v0.copy(), v0.copy()
    orc.vcycle(a, f, BOXSIZE, BOXCENTER, 0.4, 0.4, 2, los)
    ref.vcycle(b, f, BOXSIZE, BOXCENTER, 0.4, 0.4, 2, los)
    assert same(a, b)
    for niter, ncycle in [(5, 6), (2, 1)]:
        assert same(orc.fmg(f, BOXSIZE, BOXCENTER, 0.4, 0.4, niter, ncycle, los),
                    ref.fmg(f, BOXSIZE, BOXCENTER, 0.4, 0.4, niter, ncycle, los))


def test_fmg_initial_guess_vs_reference(orc, ref):
    # an odd dimension stops the recursion at the top: the caller's iterate is the start
    rng = np.random.RandomState(3)
    f = rng.normal(size=(9, 8, 8))
    v0 = rng.normal(size=(9, 8, 8))
    a = orc.fmg(f, BOXSIZE, BOXCENTER, 0.4, 0.4, 3, 2, None, v=v0.copy())
    b = ref.fmg(f, BOXSIZE, BOXCENTER, 0.4, 0.4, 3, 2, None, v=v0.copy())
    assert same(a, b)


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
@pytest.mark.parametrize('shape', [(8, 8, 8), (16, 12, 10), (2, 3, 5)])
def test_cic_vs_reference(orc, ref, dtype, shape):
    rng = np.random.RandomState(11)
    n = 20000
    pos = (rng.uniform(0., 1., size=(n, 3)) * np.array(shape)).astype(dtype)
    pos[:3] = np.array(shape)  # upper edge: the reference's d = n quirk
    w = rng.uniform(0.5, 1.5, n).astype(dtype)
    mesh0 = rng.normal(size=shape).astype(dtype)
    a, b = mesh0.copy(), mesh0.copy()
    orc.assign_cic(a, pos, w)
    ref.assign_cic(b, pos, w)
    assert same(a, b)
    assert same(orc.read_cic(a, pos), ref.read_cic(a, pos))
    assert same(orc.read_finite_difference_cic(a, BOXSIZE, pos), ref.read_finite_difference_cic(a, BOXSIZE, pos))


# ---------------------------------------------------------------------------
# 2. closed forms (independent of the reference)
# ---------------------------------------------------------------------------

def brute_cic(shape, pos, w):
    mesh = np.zeros(shape)
    index = pos.astype(int)
    frac = pos - index
    for corner in np.array(np.meshgrid([0, 1], [0, 1], [0, 1], indexing='ij')).reshape(3, -1).T:
        weight = np.prod(np.where(corner, frac, 1 - frac), axis=-1) * w
        np.add.at(mesh, tuple(((index + corner) % np.array(shape)).T), weight)
    return mesh


def brute_read(mesh, pos):
    index = pos.astype(int)
    frac = pos - index
    out = np.zeros(len(pos))
    for corner in np.array(np.meshgrid([0, 1], [0, 1], [0, 1], indexing='ij')).reshape(3, -1).T:
        weight = np.prod(np.where(corner, frac, 1 - frac), axis=-1)
        out += mesh[tuple(((index + corner) % np.array(mesh.shape)).T)] * weight
    return out


def test_cic_closed_forms(orc):
    rng = np.random.RandomState(42)
    shape = (16, 12, 10)
    pos = rng.uniform(0., 1., size=(50000, 3)) * (np.array(shape) - 1e-9)
    w = rng.uniform(0.5, 1., 50000)
    mesh = np.zeros(shape)
    orc.assign_cic(mesh, pos, w)
    assert np.allclose(mesh, brute_cic(shape, pos, w), rtol=1e-12, atol=1e-12)
    assert np.isclose(mesh.sum(), w.sum(), rtol=1e-12)
    # a single point: the 8 corner weights are products of (1-d, d)
    one = np.zeros(shape)
    orc.assign_cic(one, np.array([[3.25, 4.5, 9.75]]), np.array([2.]))
    assert np.isclose(one[3, 4, 9], 2 * 0.75 * 0.5 * 0.25) and np.isclose(one[4, 5, 0], 2 * 0.25 * 0.5 * 0.75)
    assert np.count_nonzero(one) == 8
    values = orc.read_cic(mesh, pos)
    assert np.allclose(values, brute_read(mesh, pos), rtol=1e-12, atol=1e-12)
    assert np.allclose(orc.read_cic(np.full(shape, 3.5), pos), 3.5, rtol=1e-14)


def test_finite_difference_closed_forms(orc):
    rng = np.random.RandomState(1)
    shape = (16, 12, 20)
    boxsize = np.array([800., 600., 1000.])
    pos = rng.uniform(0., 1., size=(2000, 3)) * np.array(shape)
    assert np.allclose(orc.read_finite_difference_cic(np.full(shape, 7.), boxsize, pos), 0.)
    # single mode along y: the gradient is the CIC interpolation of the central difference
    m, n1 = 2, shape[1]
    h = boxsize[1] / n1
    iy = np.arange(n1)
    phi = np.broadcast_to(np.sin(2 * np.pi * m * iy / n1)[None, :, None], shape).copy()
    grad = np.broadcast_to((np.sin(2 * np.pi * m * (iy + 1) / n1) - np.sin(2 * np.pi * m * (iy - 1) / n1))
                           [None, :, None] / (2 * h), shape).copy()
    out = orc.read_finite_difference_cic(phi, boxsize, pos)
    assert np.allclose(out[:, 1], brute_read(grad, pos), atol=1e-12)
    assert np.allclose(out[:, [0, 2]], 0., atol=1e-12)


def eigenvalue(shape, boxsize, beta, los, m):
    h = np.asarray(boxsize, dtype='f8') / np.asarray(shape)
    theta = [2 * np.pi * np.asarray(mi) / n for mi, n in zip(m, shape)]
    lam = sum((1 + beta * los[i]**2) * (2 - 2 * np.cos(theta[i])) / h[i]**2 for i in range(3))
    for i in range(3):
        for j in range(i + 1, 3):
            lam += 2 * beta * los[i] * los[j] * np.sin(theta[i]) * np.sin(theta[j]) / (h[i] * h[j])
    return lam


def mode(shape, m):
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing='ij')
    return np.cos(sum(2 * np.pi * mi * g / n for mi, g, n in zip(m, grids, shape)))


@pytest.mark.parametrize('m', [(1, 0, 0), (1, 2, 3), (3, 1, 2), (0, 0, 1)])
def test_global_los_eigenvalue(orc, m):
    shape, beta = (16, 12, 20), 0.4
    los = np.array([0.3, 0.4, np.sqrt(1 - 0.25)])
    v = mode(shape, m)
    lam = eigenvalue(shape, BOXSIZE, beta, los, m)
    r = orc.residual(v, np.zeros(shape), BOXSIZE, BOXCENTER, beta, los)
    assert np.allclose(r, -lam * v, atol=1e-10 * lam)
    # one damped Jacobi sweep: v -> (1 - omega lambda / D) v with D = 2 sum_i (1/h_i^2 + beta l_i^2/h_i^2)
    h = BOXSIZE / np.array(shape)
    d = 2 * np.sum(1. / h**2 + beta * los**2 / h**2)
    w = v.copy()
    orc.jacobi(w, np.zeros(shape), BOXSIZE, BOXCENTER, beta, 0.4, 1, los)
    assert np.allclose(w, (1 - 0.4 * lam / d) * v, atol=1e-12)


def test_transfer_closed_forms(orc):
    shape = (8, 12, 16)
    assert np.allclose(orc.reduce(np.full(shape, 2.5)), 2.5, rtol=1e-15)
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing='ij')
    checker = (-1.)**(grids[0] + grids[1] + grids[2])
    assert np.allclose(orc.reduce(checker), 0., atol=1e-15)
    assert np.allclose(orc.prolong(np.full((4, 6, 8), -1.25)), -1.25, rtol=1e-15)
    # prolongation interpolates linearly between coarse nodes (interior)
    coarse = np.broadcast_to(np.arange(6.)[None, :, None], (4, 6, 8)).copy()
    fine = orc.prolong(coarse)
    assert np.allclose(fine[:, 1:-2, :], np.broadcast_to((np.arange(1, 10) / 2.)[None, :, None], (8, 9, 16)))


def test_fmg_converges_to_discrete_solution(orc):
    # global line of sight: A is diagonal in Fourier space with eigenvalue lambda(k);
    # the exact discrete solution of A v = f (zero mean) is ifft(fft(f) / lambda)
    shape, beta = (32, 32, 32), 0.4
    boxsize = np.array([1000.] * 3)
    los = np.array([0., 0., 1.])
    rng = np.random.RandomState(5)
    f = rng.normal(size=shape)
    f -= f.mean()
    k = [np.fft.fftfreq(n) * n for n in shape]
    m = np.meshgrid(*k, indexing='ij')
    lam = eigenvalue(shape, boxsize, beta, los, m)
    lam[0, 0, 0] = 1.
    fk = np.fft.fftn(f)
    fk[0, 0, 0] = 0.
    exact = np.fft.ifftn(fk / lam).real
    v = orc.fmg(f, boxsize, [0., 0., 0.], beta, los=los)
    v -= v.mean()
    err = np.sqrt(np.sum((v - exact)**2) / np.sum(exact**2))
    assert err < 1e-3
    # more V-cycles converge further
    v2 = orc.fmg(f, boxsize, [0., 0., 0.], beta, vcycle_niterations=12, los=los)
    v2 -= v2.mean()
    assert np.sqrt(np.sum((v2 - exact)**2) / np.sum(exact**2)) < err / 10
