"""Device buffers and streams for the HIP library.

:class:`DeviceArray` is a typed, C-contiguous device allocation owned by the
library (``prh_malloc``); it is the device-resident stand-in for the numpy
arrays the reference passes to its C functions.  Objects exposing a device
pointer through ``data_ptr()`` (e.g. PyTorch-ROCm tensors) can be wrapped
without a copy with :func:`as_device`.
"""
import ctypes
import threading

import numpy as np

from . import _lib
from ._lib import lib, check

# host copies of at least this many bytes go through the library's pinned
# staging buffers (prh_memcpy_*_staged): host threads fill one chunk while the
# previous one is in flight; smaller ones are left to HIP's own staging
STAGED_COPY_BYTES = 8 << 20


class Stream(object):
    """A non-blocking HIP stream (``None`` means the default stream)."""

    def __init__(self):
        ptr = ctypes.c_void_p()
        check(lib.prh_stream_create(ctypes.byref(ptr)), 'stream create')
        self.handle = ptr.value

    def synchronize(self):
        check(lib.prh_synchronize(self.handle), 'stream synchronize')

    def __del__(self):
        handle = getattr(self, 'handle', None)
        if handle:
            lib.prh_stream_destroy(handle)
            self.handle = None


def stream_handle(stream):
    if stream is None: return None
    if isinstance(stream, Stream): return stream.handle
    if isinstance(stream, int): return stream
    handle = getattr(stream, 'cuda_stream', None)  # torch.cuda.Stream
    if handle is not None: return handle
    raise TypeError('unsupported stream object {!r}'.format(stream))


def synchronize(stream=None):
    check(lib.prh_synchronize(stream_handle(stream)), 'synchronize')


def block_bytes(nbytes):
    """Allocation size of an array of ``nbytes``: up to 1 MiB a multiple of 4 KiB, above it
    rounded up to one of eight sizes per power of two (at most 12.5% more), so that arrays
    of similar sizes (catalogs of different lengths) share cached blocks."""
    nbytes = int(nbytes)
    if nbytes <= 1 << 20:
        step = 4096
    else:
        step = 1 << (nbytes.bit_length() - 4)
    return -(-nbytes // step) * step


class _BlockCache(object):
    """Stream-ordered cache of freed device blocks, by size class (:func:`block_bytes`), so
    that repeated reconstructions of the same or similar shapes reuse their allocations
    instead of calling hipMalloc/hipFree (which waits for the whole device) each time.

    A block is returned to the cache of the stream its array was allocated for and is
    handed out again only for that stream: work queued on the stream before the release
    completes before any work queued after the reuse, so an array may be released while
    kernels that read it are still queued on its own stream.  Work on other streams that
    uses an array must be ordered before its release by the caller (events)."""

    def __init__(self, limit_bytes=64 << 30):
        self.blocks = {}
        self.cached = 0
        self.limit = limit_bytes
        self.lock = threading.Lock()  # arrays are created and released from several threads when serving

    def take(self, nbytes, stream=None):
        with self.lock:
            stack = self.blocks.get((nbytes, stream))
            if stack:
                self.cached -= nbytes
                return stack.pop()
        return None

    def give(self, ptr, nbytes, stream=None):
        with self.lock:
            if self.cached + nbytes <= self.limit:
                self.blocks.setdefault((nbytes, stream), []).append(ptr)
                self.cached += nbytes
                return
        lib.prh_free(ptr)

    def empty(self):
        with self.lock:
            ptrs = [ptr for stack in self.blocks.values() for ptr in stack]
            self.blocks.clear()
            self.cached = 0
        for ptr in ptrs: lib.prh_free(ptr)


_cache = _BlockCache()
# other holders of idle device memory (idle workspaces), released by empty_cache()
_release_hooks = []


def empty_cache():
    """Return every cached device block, and every idle workspace, to the HIP runtime."""
    synchronize()
    _cache.empty()
    for hook in _release_hooks: hook()


class DeviceArray(object):
    """C-contiguous device array.

    Parameters
    ----------
    shape : tuple
    dtype : numpy dtype
    ptr : int, optional
        Existing device address (not owned unless ``owner``).
    stream : Stream, optional
        Stream whose work uses the array: a released block is reused for the same stream only.
    """

    def __init__(self, shape, dtype, ptr=None, owner=None, base=None, stream=None):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.base = base
        self.stream = stream_handle(stream)
        if ptr is None:
            self._block = block_bytes(self.nbytes)
            cached = _cache.take(self._block, self.stream) if self.nbytes else None
            if cached is None:
                p = ctypes.c_void_p()
                status = lib.prh_malloc(ctypes.byref(p), self._block)
                if status == _lib.ERR_NOMEM:
                    empty_cache()
                    status = lib.prh_malloc(ctypes.byref(p), self._block)
                check(status, 'device allocation of {} bytes'.format(self._block))
                cached = p.value or 0
            self.ptr = cached
            self.owner = True if owner is None else owner
        else:
            self.ptr = int(ptr)
            self.owner = bool(owner)
            self._block = None  # an adopted block is cached under its exact size

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    @property
    def ndim(self):
        return len(self.shape)

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        return 'DeviceArray(shape={}, dtype={}, ptr=0x{:x})'.format(self.shape, self.dtype, self.ptr or 0)

    @classmethod
    def empty(cls, shape, dtype, stream=None):
        return cls(shape, dtype, stream=stream)

    @classmethod
    def zeros(cls, shape, dtype, stream=None):
        arr = cls(shape, dtype, stream=stream)
        arr.fill_zero(stream=stream)
        return arr

    @classmethod
    def from_numpy(cls, array, stream=None):
        array = np.ascontiguousarray(array)
        arr = cls(array.shape, array.dtype, stream=stream)
        arr.upload(array, stream=stream)
        return arr

    def upload(self, array, stream=None):
        array = np.ascontiguousarray(array, dtype=self.dtype)
        if array.size != self.size:
            raise ValueError('size mismatch: {} vs {}'.format(array.size, self.size))
        copy = lib.prh_memcpy_htod_staged if self.nbytes >= STAGED_COPY_BYTES else lib.prh_memcpy_htod
        check(copy(self.ptr, array.ctypes.data, self.nbytes, stream_handle(stream)), 'upload')
        check(lib.prh_synchronize(stream_handle(stream)), 'upload')

    def numpy(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        copy = lib.prh_memcpy_dtoh_staged if self.nbytes >= STAGED_COPY_BYTES else lib.prh_memcpy_dtoh
        check(copy(out.ctypes.data, self.ptr, self.nbytes, stream_handle(stream)), 'download')
        check(lib.prh_synchronize(stream_handle(stream)), 'download')
        return out

    get = numpy

    def fill_zero(self, stream=None):
        check(lib.prh_memset(self.ptr, 0, self.nbytes, stream_handle(stream)), 'memset')

    def copy_(self, other, stream=None):
        if other.nbytes != self.nbytes:
            raise ValueError('size mismatch')
        check(lib.prh_memcpy_dtod(self.ptr, other.ptr, self.nbytes, stream_handle(stream)), 'device copy')
        return self

    def rows(self, start, stop):
        """View of rows [start, stop) along the first axis (no copy)."""
        start, stop = int(start), int(stop)
        rowbytes = self.nbytes // max(self.shape[0], 1)
        return DeviceArray((stop - start,) + self.shape[1:], self.dtype, ptr=self.ptr + start * rowbytes,
                           owner=False, base=self)

    def reshape(self, *shape):
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)): shape = tuple(shape[0])
        shape = list(shape)
        if -1 in shape:
            known = int(np.prod([s for s in shape if s != -1]))
            shape[shape.index(-1)] = self.size // known
        if int(np.prod(shape)) != self.size:
            raise ValueError('cannot reshape {} to {}'.format(self.shape, shape))
        return DeviceArray(tuple(shape), self.dtype, ptr=self.ptr, owner=False, base=self)

    def free(self):
        if self.owner and self.ptr:
            _cache.give(self.ptr, getattr(self, '_block', None) or self.nbytes, self.stream)
        self.ptr = 0
        self.owner = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def as_device(obj, dtype=None, stream=None):
    """Return ``(device_array, is_host)``: wrap a device object or upload a host array."""
    if isinstance(obj, DeviceArray):
        return obj, False
    data_ptr = getattr(obj, 'data_ptr', None)
    if data_ptr is not None and getattr(obj, 'is_cuda', False):
        if not obj.is_contiguous():
            raise ValueError('device tensors must be contiguous')
        np_dtype = np.dtype(str(obj.dtype).replace('torch.', ''))
        return DeviceArray(tuple(obj.shape), np_dtype, ptr=data_ptr(), owner=False, base=obj), False
    array = np.asarray(obj)
    if dtype is not None: array = array.astype(dtype, copy=False)
    return DeviceArray.from_numpy(array, stream=stream), True


class Event(object):
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib.prh_event_create(ctypes.byref(p)), 'event create')
        self.handle = p.value

    def record(self, stream=None):
        check(lib.prh_event_record(self.handle, stream_handle(stream)), 'event record')

    def synchronize(self):
        check(lib.prh_event_synchronize(self.handle), 'event synchronize')

    def wait(self, stream=None):
        """Make later work on ``stream`` wait until this event has completed (no host wait)."""
        check(lib.prh_stream_wait_event(stream_handle(stream), self.handle), 'stream wait event')

    def elapsed_ms(self, stop):
        ms = ctypes.c_float()
        check(lib.prh_event_elapsed_ms(self.handle, stop.handle, ctypes.byref(ms)), 'event elapsed')
        return ms.value

    def __del__(self):
        handle = getattr(self, 'handle', None)
        if handle:
            lib.prh_event_destroy(handle)
            self.handle = None


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib.prh_get_device_name(buf, 256), 'device name')
    return buf.value.decode()


def mem_info():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib.prh_mem_info(ctypes.byref(free), ctypes.byref(total)), 'mem info')
    return free.value, total.value
