"""Host copies through the library's pinned staging buffers (prh_memcpy_htod_staged,
prh_memcpy_dtoh_staged) against HIP's own copies: the same bytes for sizes below one chunk, at
chunk edges and over more chunks than there are buffers, from two threads on two streams at once,
and through DeviceArray (which takes the staged path from 8 MB on)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK = 32 << 20


def _pattern(n, seed):
    return np.random.RandomState(seed).randint(0, 256, size=n).astype(np.uint8)


@pytest.mark.parametrize('n', [1, 4097, CHUNK - 8, CHUNK, CHUNK + 24, 4 * CHUNK + 12])
def test_staged_round_trip(n):
    from pyrecon_amd import _lib, DeviceArray, synchronize
    lib = _lib.lib
    host = _pattern(n, n % 1000)
    dev = DeviceArray.empty((n,), 'u1')
    # staged up, HIP's copy down
    _lib.check(lib.prh_memcpy_htod_staged(dev.ptr, host.ctypes.data, n, None), 'staged upload')
    synchronize()
    back = np.zeros(n, np.uint8)
    _lib.check(lib.prh_memcpy_dtoh(back.ctypes.data, dev.ptr, n, None), 'download')
    synchronize()
    assert np.array_equal(back, host)
    # HIP's copy up, staged down
    host2 = _pattern(n, n % 1000 + 1)
    _lib.check(lib.prh_memcpy_htod(dev.ptr, host2.ctypes.data, n, None), 'upload')
    synchronize()
    back2 = np.zeros(n, np.uint8)
    _lib.check(lib.prh_memcpy_dtoh_staged(back2.ctypes.data, dev.ptr, n, None), 'staged download')
    assert np.array_equal(back2, host2)


def test_staged_two_streams():
    from pyrecon_amd import _lib, DeviceArray, Stream
    lib = _lib.lib
    n = 3 * CHUNK + 4096 + 8
    errors = []

    def work(seed):
        try:
            stream = Stream()
            dev = DeviceArray.empty((n,), 'u1', stream=stream)
            for rep in range(3):
                host = _pattern(n, seed + rep)
                _lib.check(lib.prh_memcpy_htod_staged(dev.ptr, host.ctypes.data, n, stream.handle), 'up')
                back = np.zeros(n, np.uint8)
                _lib.check(lib.prh_memcpy_dtoh_staged(back.ctypes.data, dev.ptr, n, stream.handle), 'down')
                if not np.array_equal(back, host):
                    errors.append((seed, rep))
            stream.synchronize()
            dev.free()
        except Exception as exc:  # reported below
            errors.append(repr(exc))

    threads = [threading.Thread(target=work, args=(10 * t,)) for t in range(2)]
    for t in threads: t.start()
    for t in threads: t.join()
    assert not errors, errors


def test_device_array_staged():
    from pyrecon_amd import DeviceArray
    from pyrecon_amd.device import STAGED_COPY_BYTES
    a = np.random.RandomState(5).normal(size=(1 << 20, 3))  # 24 MB: the staged path
    assert a.nbytes >= STAGED_COPY_BYTES
    d = DeviceArray.from_numpy(a)
    assert np.array_equal(d.numpy(), a)
    b = np.random.RandomState(6).normal(size=1000)  # 8 KB: HIP's own copy
    assert np.array_equal(DeviceArray.from_numpy(b).numpy(), b)
