"""Arithmetic and indexing of the device RealMesh, as the reference's BaseMesh
(pyrecon/mesh.py:156-226): *, / on either side and in place, numpy ufuncs through
__array_ufunc__ (a RealMesh back for same-shape non-boolean results, arrays otherwise) and
__getitem__ / __setitem__ on the value.  Each result is compared bit for bit with numpy on the
host copy of the value.  With a Python scalar, * and / run on the device (prh_mesh_scalar_op_*).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view('u{:d}'.format(a.dtype.itemsize))


def make(dtype, seed=0):
    from pyrecon_amd import RealMesh
    rng = np.random.RandomState(seed)
    value = rng.uniform(-2., 2., (8, 12, 16)).astype(dtype)
    return RealMesh(value, boxsize=100., boxcenter=0., dtype=dtype), value


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
def test_scalar_arithmetic(dtype):
    from pyrecon_amd import RealMesh
    mesh, value = make(dtype)
    for got, expect in [(mesh * 2.5, value * 2.5), (mesh * 3, value * 3), (mesh / 0.3, value / 0.3),
                        (mesh.__rtruediv__(0.7), 0.7 / value)]:
        assert isinstance(got, RealMesh) and got.dtype == np.dtype(dtype)
        assert np.array_equal(bits(got.value), bits(expect.astype(dtype)))
    assert np.array_equal(bits(mesh.value), bits(value)), 'out-of-place operators must not touch the mesh'
    m2, v2 = make(dtype, 1)
    m2 *= 1.7
    v2 *= np.dtype(dtype).type(1.7)
    m2 /= 0.9
    v2 /= np.dtype(dtype).type(0.9)
    assert np.array_equal(bits(m2.value), bits(v2))
    # an array operand takes the host path with numpy's own semantics
    other = np.linspace(0.5, 1.5, value.size).reshape(value.shape).astype(dtype)
    assert np.array_equal(bits((mesh * other).value), bits(value * other))


@pytest.mark.parametrize('dtype', ['f4', 'f8'])
def test_ufuncs_and_indexing(dtype):
    from pyrecon_amd import RealMesh
    mesh, value = make(dtype)
    other, ovalue = make(dtype, 2)
    r = np.add(mesh, other)
    assert isinstance(r, RealMesh) and np.array_equal(bits(r.value), bits(value + ovalue))
    r = np.sqrt(np.abs(mesh))
    assert isinstance(r, RealMesh) and np.array_equal(bits(r.value), bits(np.sqrt(np.abs(value))))
    mask = np.greater(mesh, 0.)
    assert isinstance(mask, np.ndarray) and mask.dtype == bool and np.array_equal(mask, value > 0.)
    s = np.sum(mesh)  # reduction: a scalar, not a mesh
    assert np.isclose(s, value.sum())
    np.multiply(mesh, 2., out=(mesh,))
    assert np.array_equal(bits(mesh.value), bits(value * np.dtype(dtype).type(2.)))
    mesh[mask] = 0.
    expect = value * np.dtype(dtype).type(2.)
    expect[value > 0.] = 0.
    assert np.array_equal(bits(mesh.value), bits(expect))
    assert np.array_equal(mesh[1:3, 2], expect[1:3, 2])
    with pytest.raises(TypeError):
        mesh + 'text'
