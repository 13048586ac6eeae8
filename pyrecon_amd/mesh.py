"""Mesh geometry and device-resident meshes.

:class:`MeshInfo` restates the reference's geometry rules (pyrecon/mesh.py:322-483):
``nmesh`` stored as C int, ``boxsize``/``boxcenter`` stored in the mesh dtype,
box from ``positions`` with ``boxpad`` when not given, ``cellsize`` -> ``nmesh``,
the recommended-size warning, ``offset``, ``cellsize`` and ``wrap``.

:class:`RealMesh` holds its values on the device; ``.value`` downloads a numpy
copy (and assigning ``.value`` uploads), so code written against the
reference's numpy-backed RealMesh keeps working.
"""
import numbers
import threading
import warnings

import numpy as np

from . import _lib
from ._lib import lib, check, MeshError
from . import device as _device
from .device import DeviceArray, as_device, stream_handle, mem_info, empty_cache

DEFAULT_PAINT_CHUNK = 1 << 24
# a workspace created with paint_chunk=None grows its paint scratch up to this many points
MAX_AUTO_PAINT_CHUNK = 1 << 28


def _make_array(value, shape, dtype='f8'):
    out = np.empty(shape, dtype=dtype)
    out[...] = value
    return out


class MeshInfo(object):
    """Mesh geometry (pyrecon/mesh.py:322-483)."""

    _attrs = ['dtype', 'nmesh', 'boxsize', 'boxcenter']

    def __init__(self, nmesh=None, boxsize=None, boxcenter=None, cellsize=None, value=None, positions=None,
                 boxpad=1.5, select_nmesh=None, dtype=None):
        if value is not None:
            value = np.asarray(value)
            if dtype is None: dtype = value.dtype
            if nmesh is None: nmesh = value.shape
        dtype = np.dtype(dtype if dtype is not None else 'f8')
        if dtype not in (np.dtype('f4'), np.dtype('f8')):
            raise MeshError('Provide float dtype (f4 or f8)')
        if boxsize is None or boxcenter is None:
            if positions is None:
                raise MeshError('positions must be provided if boxsize or boxcenter are not specified')
            if isinstance(positions, DeviceArray): positions = positions.numpy()
            positions = np.asarray(positions)
            pos_min, pos_max = positions.min(axis=0), positions.max(axis=0)
            delta = np.abs(pos_max - pos_min)
            if boxcenter is None: boxcenter = 0.5 * (pos_min + pos_max)
            if boxsize is None:
                if cellsize is not None and nmesh is not None:
                    boxsize = nmesh * cellsize
                else:
                    boxsize = delta * boxpad
            if (boxsize < delta).any(): raise MeshError('boxsize too small to contain all data')
        boxsize = _make_array(boxsize, 3, dtype='f8')
        if nmesh is None:
            if cellsize is not None:
                cellsize = _make_array(cellsize, 3, dtype='f8')
                nmesh = boxsize / cellsize
                nmesh = np.ceil(nmesh).astype('i8')
                if select_nmesh is not None: nmesh = select_nmesh(nmesh)
                boxsize = nmesh * cellsize
            else:
                raise MeshError('nmesh (or cellsize) must be specified')
        self.dtype = dtype
        self.boxsize = boxsize
        self.boxcenter = boxcenter
        self.nmesh = nmesh
        if select_nmesh is not None:
            recommended = select_nmesh(self.nmesh)
            if not np.all(recommended == self.nmesh):
                warnings.warn('Recommended nmesh is {}, provided nmesh is {}'.format(recommended, self.nmesh))

    @property
    def dtype(self):
        return self._dtype

    @dtype.setter
    def dtype(self, dtype):
        self._dtype = np.dtype(dtype)
        if hasattr(self, '_boxsize'):
            self.boxsize = self._boxsize
            self.boxcenter = self._boxcenter

    @property
    def boxsize(self):
        return self._boxsize

    @boxsize.setter
    def boxsize(self, value):
        self._boxsize = _make_array(value, 3, dtype=self.dtype)

    @property
    def boxcenter(self):
        return self._boxcenter

    @boxcenter.setter
    def boxcenter(self, value):
        self._boxcenter = _make_array(value, 3, dtype=self.dtype)

    @property
    def nmesh(self):
        return self._nmesh

    @nmesh.setter
    def nmesh(self, value):
        self._nmesh = _make_array(value, 3, dtype=np.intc)

    @property
    def offset(self):
        """Coordinates of the (0, 0, 0) corner of the box."""
        return self.boxcenter - self.boxsize / 2.

    @property
    def cellsize(self):
        return self.boxsize / self.nmesh

    @property
    def ndim(self):
        return 3

    @property
    def shape(self):
        return tuple(int(n) for n in self.nmesh)

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def wrap(self, positions):
        """Wrap positions into the box (host arrays)."""
        return ((positions - self.offset) % self.boxsize) + self.offset

    def clone(self, **kwargs):
        for name in self._attrs: kwargs.setdefault(name, getattr(self, name))
        return self.__class__(**kwargs)

    def __repr__(self):
        return 'MeshInfo(dtype={}, nmesh={}, boxsize={}, boxcenter={})'.format(self.dtype, self.nmesh, self.boxsize,
                                                                               self.boxcenter)


# Workspaces marked ``reuse`` that no object references any more, by kind (one each, the last
# _MAX_IDLE kinds): a new workspace of the same kind takes one instead of allocating and setting
# up its own.
_idle = {}
_idle_lock = threading.Lock()
_MAX_IDLE = 2


def _drop_idle():
    with _idle_lock:
        handles = [entry[0] for entry in _idle.values()]
        _idle.clear()
    for handle in handles: lib.prh_workspace_destroy(handle)


_device._release_hooks.append(_drop_idle)


class Workspace(object):
    """Owner of a ``prh_workspace``: multigrid hierarchy, paint scratch, FFT plans.

    ``paint_chunk`` is the number of points painted per pass.  ``None`` starts at
    ``DEFAULT_PAINT_CHUNK`` and grows the scratch for larger catalogs (``reserve``), up to
    ``MAX_AUTO_PAINT_CHUNK`` points and while device memory allows, so that a catalog is painted
    in one pass (and its read can walk the order that pass left); an integer fixes it.
    ``grow`` overrides that choice.

    ``reuse``: for a workspace used on the default stream only.  When no object references it
    any more it is kept (one per kind) for the next workspace of the same kind, instead of being
    destroyed; ``empty_cache()`` destroys the kept ones.
    """

    def __init__(self, nmesh, dtype, paint_chunk=None, flags=_lib.WS_ALL, grow=None, reuse=False):
        import ctypes
        self.nmesh = tuple(int(n) for n in nmesh)
        self.dtype = np.dtype(dtype)
        self.grow = (paint_chunk is None) if grow is None else bool(grow)
        self.paint_chunk = int(DEFAULT_PAINT_CHUNK if paint_chunk is None else paint_chunk)
        self.flags = int(flags)
        if self.dtype == np.float32: self.flags &= ~_lib.WS_MIXED
        self._reuse_key = (self.nmesh, self.dtype.str, self.flags, paint_chunk, self.grow) if reuse else None
        if self._reuse_key is not None:
            with _idle_lock:
                entry = _idle.pop(self._reuse_key, None)
            if entry is not None:
                self.handle, self.nbytes, self.paint_chunk, self.grow = entry
                return
        handle = ctypes.c_void_p()
        nm, nm_ptr = _lib.size3(self.nmesh)
        create = lambda: lib.prh_workspace_create(ctypes.byref(handle), nm_ptr, _lib.dtype_code(self.dtype),
                                                  self.paint_chunk, self.flags)
        status = create()
        if status == _lib.ERR_NOMEM:  # cached blocks or kept workspaces may hold the memory
            empty_cache()
            status = create()
        check(status, 'workspace create')
        self.handle = handle.value
        nbytes = ctypes.c_size_t()
        check(lib.prh_workspace_bytes(self.handle, ctypes.byref(nbytes)), 'workspace bytes')
        self.nbytes = nbytes.value

    def _update_bytes(self):
        import ctypes
        nbytes = ctypes.c_size_t()
        check(lib.prh_workspace_bytes(self.handle, ctypes.byref(nbytes)), 'workspace bytes')
        self.nbytes = nbytes.value

    def reserve(self, npoints):
        """Grow the paint scratch (if this workspace grows) so that ``npoints`` points are painted
        in one pass.  Returns whether they are; if not, the paint takes several passes."""
        npoints = int(npoints)
        if npoints <= self.paint_chunk: return True
        if not self.grow or not (self.flags & _lib.WS_PAINT) or npoints > MAX_AUTO_PAINT_CHUNK: return False
        step = 1 << 22
        chunk = min(-(-npoints // step) * step, MAX_AUTO_PAINT_CHUNK)
        # the new arrays are allocated before the old ones are freed; keep a margin for the meshes
        need = Workspace.footprint(self.nmesh, self.dtype, chunk, self.flags) - \
            Workspace.footprint(self.nmesh, self.dtype, 1, self.flags)
        free, total = mem_info()
        margin = max(2 << 30, total // 20)
        if free < need + margin:
            empty_cache()
            free, total = mem_info()
        if free < need + margin:
            self.grow = False
            return False
        status = lib.prh_workspace_reserve_paint(self.handle, chunk)
        if status == _lib.ERR_NOMEM:
            self.grow = False
            return False
        check(status, 'paint scratch reserve')
        self.paint_chunk = chunk
        self._update_bytes()
        return True

    @staticmethod
    def footprint(nmesh, dtype, paint_chunk=DEFAULT_PAINT_CHUNK, flags=_lib.WS_ALL):
        import ctypes
        dtype = np.dtype(dtype)
        if paint_chunk is None: paint_chunk = DEFAULT_PAINT_CHUNK
        if dtype == np.float32: flags &= ~_lib.WS_MIXED
        nm, nm_ptr = _lib.size3(nmesh)
        out = ctypes.c_size_t()
        check(lib.prh_workspace_footprint(nm_ptr, _lib.dtype_code(dtype), int(paint_chunk), int(flags),
                                          ctypes.byref(out)), 'workspace footprint')
        return out.value

    def close(self):
        handle = getattr(self, 'handle', None)
        if handle:
            lib.prh_workspace_destroy(handle)
            self.handle = None

    def __del__(self):
        try:
            handle = getattr(self, 'handle', None)
            if handle and getattr(self, '_reuse_key', None) is not None:
                lib.prh_synchronize(None)  # its work on the default stream has finished
                evicted = []
                with _idle_lock:
                    if self._reuse_key not in _idle:
                        _idle[self._reuse_key] = (handle, self.nbytes, self.paint_chunk, self.grow)
                        self.handle = None
                        while len(_idle) > _MAX_IDLE:
                            evicted.append(_idle.pop(next(iter(_idle)))[0])
                for old in evicted: lib.prh_workspace_destroy(old)
                if self.handle is None: return
            self.close()
        except Exception:
            pass


_workspace_cache = {}


def shared_workspace(nmesh, dtype, flags=_lib.WS_ALL, paint_chunk=None):
    key = (tuple(int(n) for n in nmesh), np.dtype(dtype).str, int(flags), paint_chunk)
    ws = _workspace_cache.get(key)
    if ws is None:
        ws = _workspace_cache[key] = Workspace(nmesh, dtype, paint_chunk=paint_chunk, flags=flags)
    return ws


def _suffix(dtype):
    return 'f32' if np.dtype(dtype) == np.float32 else 'f64'


class RealMesh(object):
    """Real mesh whose values live on the device (pyrecon/mesh.py:486-736)."""

    def __init__(self, value=None, info=None, dtype=None, workspace=None, stream=None, **kwargs):
        if info is None:
            if dtype is None and (value is None or np.ndim(value) == 0): dtype = 'f8'
            info = MeshInfo(value=None if isinstance(value, DeviceArray) or np.ndim(value) == 0 else value,
                            dtype=dtype, **kwargs)
        self.info = info
        self.stream = stream
        self._workspace = workspace
        self.buffer = None
        if value is not None: self.value = value

    # geometry passthrough
    dtype = property(lambda self: self.info.dtype)
    nmesh = property(lambda self: self.info.nmesh)
    boxsize = property(lambda self: self.info.boxsize)
    boxcenter = property(lambda self: self.info.boxcenter)
    offset = property(lambda self: self.info.offset)
    cellsize = property(lambda self: self.info.cellsize)
    shape = property(lambda self: self.info.shape)
    size = property(lambda self: self.info.size)
    ndim = property(lambda self: 3)

    @property
    def workspace(self):
        if self._workspace is None:
            self._workspace = shared_workspace(self.shape, self.dtype)
        return self._workspace

    def _allocate(self, zero=True):
        if self.buffer is None:
            self.buffer = DeviceArray(self.shape, self.dtype, stream=self.stream)
            if zero: self.buffer.fill_zero(stream=self.stream)
        return self.buffer

    @property
    def value(self):
        if self.buffer is None: return None
        return self.buffer.numpy(stream=self.stream)

    @value.setter
    def value(self, value):
        if value is None:
            self.buffer = None
            return
        if isinstance(value, DeviceArray):
            if value.size != self.size or value.dtype != self.dtype:
                raise MeshError('device value must have size {} and dtype {}'.format(self.size, self.dtype))
            self.buffer = value.reshape(self.shape) if value.shape != self.shape else value
            return
        array = np.empty(self.shape, dtype=self.dtype)
        array[...] = value
        self._allocate(zero=False).upload(array, stream=self.stream)

    def __array__(self, dtype=None, copy=None):
        value = self.value
        if dtype is not None: value = value.astype(dtype)
        return value

    # ------------------------------------------------------------------ arithmetic and indexing
    # The reference's BaseMesh operators (pyrecon/mesh.py:156-226): *, / (both sides) and their
    # in-place forms, numpy ufuncs through __array_ufunc__, and indexing of the value.  With a
    # Python scalar, * and / run on the device; other operands take numpy's path on the host.
    _HANDLED_TYPES = (np.ndarray, numbers.Number)

    def _scalar_op(self, op, other, out):
        """out (a RealMesh, possibly self) = self op other on the device; False if not applicable."""
        if self.buffer is None or isinstance(other, (bool, np.bool_)) or not isinstance(other, (int, float)):
            return False
        fn = getattr(lib, 'prh_mesh_scalar_op_' + _suffix(self.dtype))
        out._allocate(zero=False)
        check(fn(self.buffer.ptr, out.buffer.ptr, self.size, op, float(other), stream_handle(self.stream)),
              'mesh arithmetic')
        return True

    def _binary(self, op, other, reverse=False):
        new = self.deepcopy(copy_value=False)
        code = _lib.OP_RDIV if reverse else op
        if not self._scalar_op(code, other, new):
            value = self.value
            if op == _lib.OP_MUL: result = value * other
            elif reverse: result = other / value
            else: result = value / other
            new.value = result
        return new

    def _inplace(self, op, other):
        if not self._scalar_op(op, other, self):
            value = self.value
            if op == _lib.OP_MUL: value *= other
            else: value /= other
            self.value = value
        return self

    def __mul__(self, other):
        return self._binary(_lib.OP_MUL, other)

    def __imul__(self, other):
        return self._inplace(_lib.OP_MUL, other)

    def __truediv__(self, other):
        return self._binary(_lib.OP_DIV, other)

    def __rtruediv__(self, other):
        return self._binary(_lib.OP_DIV, other, reverse=True)

    def __itruediv__(self, other):
        return self._inplace(_lib.OP_DIV, other)

    __div__, __rdiv__, __idiv__ = __truediv__, __rtruediv__, __itruediv__

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        out = kwargs.get('out', ())
        for x in inputs + out:
            if not isinstance(x, self._HANDLED_TYPES + (RealMesh,)):
                return NotImplemented
        values = tuple(x.value if isinstance(x, RealMesh) else x for x in inputs)
        if out:
            kwargs['out'] = tuple(x.value if isinstance(x, RealMesh) else x for x in out)
        result = getattr(ufunc, method)(*values, **kwargs)
        for target, host in zip(out, kwargs.get('out', ())):
            if isinstance(target, RealMesh):
                target.value = host

        def cast(res):
            # booleans and other shapes are no meshes: plain arrays
            if res.dtype == '?' or res.shape != self.shape:
                return res
            new = self.deepcopy(copy_value=False)
            new.value = res
            return new

        if method == 'at':
            return None
        if out:
            return out[0] if len(out) == 1 else out
        if type(result) is tuple:
            return tuple(cast(r) for r in result)
        return cast(result)

    def __getitem__(self, index):
        return self.value[index]

    def __setitem__(self, index, value):
        array = self.value
        array[index] = value
        self.value = array

    def __repr__(self):
        return 'RealMesh(dtype={}, nmesh={}, boxsize={}, boxcenter={})'.format(self.dtype, self.nmesh, self.boxsize,
                                                                               self.boxcenter)

    def zeros_like(self):
        new = RealMesh(info=self.info, workspace=self._workspace, stream=self.stream)
        new._allocate(zero=True)
        return new

    def deepcopy(self, copy_value=True):
        new = RealMesh(info=self.info.clone(), workspace=self._workspace, stream=self.stream)
        if copy_value and self.buffer is not None:
            new._allocate(zero=False).copy_(self.buffer, stream=self.stream)
        return new

    def coords(self):
        return tuple(o + d * np.arange(n) for n, o, d in zip(self.nmesh, self.offset, self.boxsize / self.nmesh))

    # ------------------------------------------------------------------ kernels
    def assign_cic(self, positions, weights=None, wrap=False, tile_sums=False):
        """Paint positions with CIC (pyrecon/mesh.py:529-568); accumulates into the mesh.

        ``tile_sums`` (FP64 meshes): the mixed-precision paint, with the FP64 corner weights summed
        in fixed point per 8^3-cell tile and in FP64 across tiles (deterministic; the reference's
        sums to within their own rounding, not its bits).
        """
        pos, _ = as_device(positions, stream=self.stream)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError('positions must have shape (N, 3)')
        wdev, wcode = None, 0
        if weights is not None:
            wdev, _ = as_device(weights, stream=self.stream)
            if wdev.size != len(pos):
                raise ValueError('weights must have shape (N,)')
            wcode = _lib.dtype_code(wdev.dtype)
        self._allocate(zero=True)
        self.workspace.reserve(len(pos))  # one pass for the whole catalog, where the workspace grows
        nm, nm_ptr = _lib.size3(self.shape)
        bs, bs_ptr = _lib.vec3(self.boxsize, self.dtype)
        bc, bc_ptr = _lib.vec3(self.boxcenter, self.dtype)
        if tile_sums and self.dtype != np.float64:
            raise ValueError('the tile-sum paint writes FP64 meshes (dtype="f8")')
        fn = lib.prh_paint_mixed_f64 if tile_sums else getattr(lib, 'prh_paint_' + _suffix(self.dtype))
        check(fn(self.workspace.handle, self.buffer.ptr, nm_ptr, bs_ptr, bc_ptr, pos.ptr,
                 _lib.dtype_code(pos.dtype), wdev.ptr if wdev is not None else None, wcode, len(pos), int(bool(wrap)),
                 stream_handle(self.stream)), 'assign_cic')

    def _read(self, positions, field, wrap):
        import ctypes
        import warnings
        if self.buffer is None: raise MeshError('mesh has no value')
        pos, is_host = as_device(positions, stream=self.stream)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError('positions must have shape (N, 3)')
        if pos.dtype not in (np.dtype('f4'), np.dtype('f8')):
            pos, is_host = as_device(np.asarray(positions, dtype='f8'), stream=self.stream)
        # values come back in the positions dtype (mesh.py:610, 653)
        out = DeviceArray((len(pos),) if field == _lib.FIELD_VALUE else (len(pos), 3), pos.dtype, stream=self.stream)
        nm, nm_ptr = _lib.size3(self.shape)
        bs, bs_ptr = _lib.vec3(self.boxsize, self.dtype)
        bc, bc_ptr = _lib.vec3(self.boxcenter, self.dtype)
        nonfinite = ctypes.c_int(0)
        fn = getattr(lib, 'prh_read_shifts_' + _suffix(self.dtype))
        status = fn(self.workspace.handle, self.buffer.ptr, nm_ptr, bs_ptr, bc_ptr, pos.ptr, _lib.dtype_code(pos.dtype),
                    len(pos), 0., None, field, int(bool(wrap)), 0, out.ptr, _lib.dtype_code(pos.dtype),
                    ctypes.byref(nonfinite), stream_handle(self.stream))
        if status == _lib.ERR_RANGE:
            low, high = self.boxcenter - self.boxsize / 2., self.boxcenter + self.boxsize / 2.
            raise MeshError('positions not in box range {} - {}'.format(low, high))
        check(status, 'read')
        if nonfinite.value:
            warnings.warn('Some positions are NaN/inf')
        return out.numpy(stream=self.stream) if is_host else out

    def read_cic(self, positions, wrap=False):
        """Mesh values at ``positions`` by CIC interpolation (pyrecon/mesh.py:570-610), in the positions
        dtype; numpy in -> numpy out, device in -> device out."""
        return self._read(positions, _lib.FIELD_VALUE, wrap)

    def read_finite_difference_cic(self, positions, wrap=False):
        """Finite-difference gradient of the mesh at ``positions`` (pyrecon/mesh.py:612-653), shape (N, 3),
        in the positions dtype."""
        return self._read(positions, _lib.FIELD_DISP, wrap)

    def smooth_gaussian(self, radius, method='fft', nsigmas=2.5, **kwargs):
        """Gaussian smoothing in Fourier space (pyrecon/mesh.py:655-680)."""
        if method != 'fft':
            raise NotImplementedError('only method="fft" is implemented on the device')
        if self.buffer is None: raise MeshError('mesh has no value')
        radius_ = np.empty(3, dtype='f8')
        radius_[:] = radius
        nm, nm_ptr = _lib.size3(self.shape)
        bs, bs_ptr = _lib.vec3(self.boxsize, self.dtype)
        fn = getattr(lib, 'prh_smooth_gaussian_' + _suffix(self.dtype))
        check(fn(self.workspace.handle, self.buffer.ptr, nm_ptr, bs_ptr, radius_.ctypes.data,
                 stream_handle(self.stream)), 'smooth_gaussian')

    def sum(self, *args, **kwargs):
        """numpy's sum of the value (what np.sum does with the reference's mesh)."""
        return np.sum(self.value, *args, **kwargs)

    def device_sum(self):
        """Deterministic FP64 sum of the cells on the device (one number back to the host)."""
        import ctypes
        sums = (ctypes.c_double * 2)()
        fn = getattr(lib, 'prh_mesh_sums_' + _suffix(self.dtype))
        check(fn(self.workspace.handle, self.buffer.ptr, None, self.size, sums, stream_handle(self.stream)), 'sum')
        return sums[0]
