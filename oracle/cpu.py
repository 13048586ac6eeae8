"""ctypes bindings of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Two interchangeable back ends with one interface:

- :class:`OracleKernels` binds ``oracle/liboracle.so``, this project's C
  restatement (``oracle/orc_kernels.c``) of the reference's native kernels.
- :class:`RefKernels` binds ``oracle/_ref/{mesh,multigrid}_{float,double}.so``,
  the reference's own ``src/mesh.c`` and ``src/multigrid.c`` compiled by
  ``oracle/Makefile`` (present only when it could be built).

Every method takes numpy arrays in the reference's layout: meshes are C-order
``(n0, n1, n2)`` (x slowest), positions are ``(N, 3)`` already in mesh units
(``((x - boxcenter) / boxsize + 0.5) * nmesh``, pyrecon/mesh.py:554).  The
argument meaning follows the reference C signatures (src/mesh.c:26, 187, 269;
src/multigrid.c:44, 107, 167, 212, 270, 303).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may use
this module; the shipped package never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, 'liboracle.so')
REF_DIR = os.path.join(HERE, '_ref')

_c_size3 = np.ctypeslib.ndpointer(dtype=np.uintp, shape=3, flags='C')


def _ctype(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def _ptr(array, dtype):
    if not (isinstance(array, np.ndarray) and array.dtype == dtype and array.flags.c_contiguous):
        raise TypeError('expected a C-contiguous {} array'.format(np.dtype(dtype)))
    return array.ctypes.data_as(ctypes.c_void_p)


def _vec3(values, dtype):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=dtype), (3,)))


class _Kernels(object):
    """Shared argument handling; subclasses provide ``_fn(name, dtype)``."""

    def _fn(self, name, dtype):
        raise NotImplementedError

    def assign_cic(self, mesh, positions, weights):
        dtype = mesh.dtype
        fn = self._fn('assign_cic', dtype)
        fn.restype = ctypes.c_int
        nmesh = np.array(mesh.shape, dtype=np.uintp)
        positions = np.ascontiguousarray(positions, dtype=dtype)
        weights = np.ascontiguousarray(weights, dtype=dtype)
        flag = fn(_ptr(mesh, dtype), nmesh.ctypes.data_as(ctypes.c_void_p), _ptr(positions, dtype),
                  _ptr(weights, dtype), ctypes.c_size_t(len(weights)))
        if flag != 0: raise RuntimeError('assign_cic failed')

    def read_cic(self, mesh, positions):
        dtype = mesh.dtype
        fn = self._fn('read_cic', dtype)
        fn.restype = ctypes.c_int
        nmesh = np.array(mesh.shape, dtype=np.uintp)
        positions = np.ascontiguousarray(positions, dtype=dtype)
        out = np.empty(len(positions), dtype=dtype)
        flag = fn(_ptr(mesh, dtype), nmesh.ctypes.data_as(ctypes.c_void_p), _ptr(positions, dtype),
                  _ptr(out, dtype), ctypes.c_size_t(len(out)))
        if flag != 0: raise RuntimeError('read_cic failed')
        return out

    def read_finite_difference_cic(self, mesh, boxsize, positions):
        dtype = mesh.dtype
        fn = self._fn('read_finite_difference_cic', dtype)
        fn.restype = ctypes.c_int
        nmesh = np.array(mesh.shape, dtype=np.uintp)
        boxsize = _vec3(boxsize, dtype)
        positions = np.ascontiguousarray(positions, dtype=dtype)
        out = np.empty((len(positions), 3), dtype=dtype)
        flag = fn(_ptr(mesh, dtype), nmesh.ctypes.data_as(ctypes.c_void_p), _ptr(boxsize, dtype),
                  _ptr(positions, dtype), _ptr(out, dtype), ctypes.c_size_t(len(out)))
        if flag != 0: raise RuntimeError('read_finite_difference_cic failed')
        return out

    def _geom(self, shape, boxsize, boxcenter, los, dtype):
        nmesh = np.array(shape, dtype=np.uintp)
        boxsize = _vec3(boxsize, dtype)
        boxcenter = _vec3(boxcenter, dtype)
        los_arr = None if los is None else _vec3(los, dtype)
        los_ptr = ctypes.c_void_p(None) if los_arr is None else _ptr(los_arr, dtype)
        return nmesh, boxsize, boxcenter, los_arr, los_ptr

    def jacobi(self, v, f, boxsize, boxcenter, beta, damping_factor, niterations, los=None):
        dtype = v.dtype
        ct = _ctype(dtype)
        fn = self._fn('jacobi', dtype)
        fn.restype = None
        nmesh, bs, bc, los_arr, los_ptr = self._geom(v.shape, boxsize, boxcenter, los, dtype)
        fn(_ptr(v, dtype), _ptr(f, dtype), nmesh.ctypes.data_as(ctypes.c_void_p), _ptr(bs, dtype), _ptr(bc, dtype),
           ct(beta), ct(damping_factor), ctypes.c_int(niterations), los_ptr)

    def residual(self, v, f, boxsize, boxcenter, beta, los=None):
        dtype = v.dtype
        ct = _ctype(dtype)
        fn = self._fn('residual', dtype)
        fn.restype = None
        r = np.empty_like(v)
        nmesh, bs, bc, los_arr, los_ptr = self._geom(v.shape, boxsize, boxcenter, los, dtype)
        fn(_ptr(v, dtype), _ptr(f, dtype), _ptr(r, dtype), nmesh.ctypes.data_as(ctypes.c_void_p), _ptr(bs, dtype),
           _ptr(bc, dtype), ct(beta), los_ptr)
        return r

    def reduce(self, fine):
        dtype = fine.dtype
        fn = self._fn('reduce', dtype)
        fn.restype = None
        nmesh = np.array(fine.shape, dtype=np.uintp)
        coarse = np.empty(tuple(n // 2 for n in fine.shape), dtype=dtype)
        fn(_ptr(fine, dtype), _ptr(coarse, dtype), nmesh.ctypes.data_as(ctypes.c_void_p))
        return coarse

    def prolong(self, coarse):
        dtype = coarse.dtype
        fn = self._fn('prolong', dtype)
        fn.restype = None
        nmesh = np.array(coarse.shape, dtype=np.uintp)
        fine = np.empty(tuple(2 * n for n in coarse.shape), dtype=dtype)
        fn(_ptr(coarse, dtype), _ptr(fine, dtype), nmesh.ctypes.data_as(ctypes.c_void_p))
        return fine

    def vcycle(self, v, f, boxsize, boxcenter, beta, damping_factor, niterations, los=None):
        dtype = v.dtype
        ct = _ctype(dtype)
        fn = self._fn('vcycle', dtype)
        fn.restype = None
        nmesh, bs, bc, los_arr, los_ptr = self._geom(v.shape, boxsize, boxcenter, los, dtype)
        fn(_ptr(v, dtype), _ptr(f, dtype), nmesh.ctypes.data_as(ctypes.c_void_p), _ptr(bs, dtype), _ptr(bc, dtype),
           ct(beta), ct(damping_factor), ctypes.c_int(niterations), los_ptr)

    def fmg(self, f, boxsize, boxcenter, beta, damping_factor=0.4, niterations=5, vcycle_niterations=6, los=None, v=None):
        """Solve with the full multigrid; ``v`` (default zeros) is the initial iterate, updated in place."""
        dtype = f.dtype
        ct = _ctype(dtype)
        fn = self._fn('fmg', dtype)
        fn.restype = ctypes.c_void_p
        if v is None: v = np.zeros_like(f)
        nmesh, bs, bc, los_arr, los_ptr = self._geom(f.shape, boxsize, boxcenter, los, dtype)
        fn(_ptr(f, dtype), _ptr(v, dtype), nmesh.ctypes.data_as(ctypes.c_void_p), _ptr(bs, dtype), _ptr(bc, dtype),
           ct(beta), ct(damping_factor), ctypes.c_int(niterations), ctypes.c_int(vcycle_niterations), los_ptr)
        return v


class OracleKernels(_Kernels):
    """This project's C restatement (``oracle/orc_kernels.c``)."""

    def __init__(self, path=ORACLE_LIB):
        if not os.path.isfile(path):
            raise OSError('oracle library {} is not built; run "make -C oracle"'.format(path))
        self._lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)

    def _fn(self, name, dtype):
        suffix = '_f32' if np.dtype(dtype) == np.float32 else '_f64'
        return getattr(self._lib, 'orc_{}{}'.format(name, suffix))


class RefKernels(_Kernels):
    """The reference's own C sources, compiled into ``oracle/_ref``."""

    _which = {'assign_cic': 'mesh', 'read_cic': 'mesh', 'read_finite_difference_cic': 'mesh',
              'jacobi': 'multigrid', 'residual': 'multigrid', 'reduce': 'multigrid', 'prolong': 'multigrid',
              'vcycle': 'multigrid', 'fmg': 'multigrid'}

    def __init__(self, directory=REF_DIR):
        self._libs = {}
        for lib in ['mesh', 'multigrid']:
            for precision in ['float', 'double']:
                path = os.path.join(directory, '{}_{}.so'.format(lib, precision))
                if not os.path.isfile(path):
                    raise OSError('reference build {} is missing; run "make -C oracle ref"'.format(path))
                self._libs[lib, precision] = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)

    @staticmethod
    def available(directory=REF_DIR):
        return all(os.path.isfile(os.path.join(directory, '{}_{}.so'.format(lib, p)))
                   for lib in ['mesh', 'multigrid'] for p in ['float', 'double'])

    def _fn(self, name, dtype):
        precision = 'float' if np.dtype(dtype) == np.float32 else 'double'
        return getattr(self._libs[self._which[name], precision], name)

    def set_num_threads(self, nthreads):
        for precision in ['float', 'double']:
            fn = self._libs['mesh', precision].set_num_threads
            fn.argtypes = (ctypes.c_int,)
            fn(int(nthreads))
