"""numpy-in / numpy-out drivers of the reference-shaped C-ABI entry points.

Each method uploads its inputs, calls one ``prh_*`` function of
libpyrecon_amd.so through ctypes, and downloads the result, so GPU tests read
like the oracle's (``oracle.cpu``) calls.
"""
import ctypes

import numpy as np

from pyrecon_amd import _lib
from pyrecon_amd._lib import lib, check
from pyrecon_amd.device import DeviceArray, synchronize
from pyrecon_amd.mesh import Workspace


def _sfx(dtype):
    return 'f32' if np.dtype(dtype) == np.float32 else 'f64'


def _real(dtype):
    return float


class GPU(object):

    def __init__(self):
        n = _lib.device_count()
        if n < 1:
            raise RuntimeError('no HIP device visible: GPU tests need an MI355X')
        self._ws = {}

    def workspace(self, shape, dtype, flags=_lib.WS_MULTIGRID, chunk=1 << 16):
        key = (tuple(shape), np.dtype(dtype).str, flags, chunk)
        if key not in self._ws:
            self._ws[key] = Workspace(shape, dtype, paint_chunk=chunk, flags=flags)
        return self._ws[key]

    def _geom(self, shape, boxsize, boxcenter, los, dtype):
        nm, nm_ptr = _lib.size3(shape)
        bs, bs_ptr = _lib.vec3(boxsize, dtype)
        bc, bc_ptr = _lib.vec3(boxcenter, dtype)
        keep = [nm, bs, bc]
        los_ptr = None
        if los is not None:
            la, los_ptr = _lib.vec3(los, dtype)
            keep.append(la)
        return keep, nm_ptr, bs_ptr, bc_ptr, los_ptr

    def jacobi(self, v, f, boxsize, boxcenter, beta, damping_factor, niterations, los=None):
        dtype = v.dtype
        ws = self.workspace(v.shape, dtype)
        dv, df = DeviceArray.from_numpy(v), DeviceArray.from_numpy(f)
        keep, nm, bs, bc, lp = self._geom(v.shape, boxsize, boxcenter, los, dtype)
        fn = getattr(lib, 'prh_jacobi_' + _sfx(dtype))
        check(fn(ws.handle, dv.ptr, df.ptr, nm, bs, bc, beta, damping_factor, int(niterations), lp, None), 'jacobi')
        v[...] = dv.numpy()

    def residual(self, v, f, boxsize, boxcenter, beta, los=None):
        dtype = v.dtype
        dv, df, dr = DeviceArray.from_numpy(v), DeviceArray.from_numpy(f), DeviceArray(v.shape, dtype)
        keep, nm, bs, bc, lp = self._geom(v.shape, boxsize, boxcenter, los, dtype)
        fn = getattr(lib, 'prh_residual_' + _sfx(dtype))
        check(fn(dv.ptr, df.ptr, dr.ptr, nm, bs, bc, beta, lp, None), 'residual')
        return dr.numpy()

    def reduce(self, fine):
        dtype = fine.dtype
        d1 = DeviceArray.from_numpy(fine)
        d2 = DeviceArray(tuple(n // 2 for n in fine.shape), dtype)
        nm, nm_ptr = _lib.size3(fine.shape)
        check(getattr(lib, 'prh_reduce_' + _sfx(dtype))(d1.ptr, d2.ptr, nm_ptr, None), 'reduce')
        return d2.numpy()

    def prolong(self, coarse):
        dtype = coarse.dtype
        d2 = DeviceArray.from_numpy(coarse)
        d1 = DeviceArray(tuple(2 * n for n in coarse.shape), dtype)
        nm, nm_ptr = _lib.size3(coarse.shape)
        check(getattr(lib, 'prh_prolong_' + _sfx(dtype))(d2.ptr, d1.ptr, nm_ptr, None), 'prolong')
        return d1.numpy()

    def vcycle(self, v, f, boxsize, boxcenter, beta, damping_factor, niterations, los=None):
        dtype = v.dtype
        ws = self.workspace(v.shape, dtype)
        dv, df = DeviceArray.from_numpy(v), DeviceArray.from_numpy(f)
        keep, nm, bs, bc, lp = self._geom(v.shape, boxsize, boxcenter, los, dtype)
        fn = getattr(lib, 'prh_vcycle_' + _sfx(dtype))
        check(fn(ws.handle, dv.ptr, df.ptr, nm, bs, bc, beta, damping_factor, int(niterations), lp, None), 'vcycle')
        v[...] = dv.numpy()

    def fmg(self, f, boxsize, boxcenter, beta, damping_factor=0.4, niterations=5, vcycle_niterations=6, los=None,
            v=None):
        dtype = f.dtype
        ws = self.workspace(f.shape, dtype)
        if v is None: v = np.zeros_like(f)
        dv, df = DeviceArray.from_numpy(v), DeviceArray.from_numpy(f)
        keep, nm, bs, bc, lp = self._geom(f.shape, boxsize, boxcenter, los, dtype)
        fn = getattr(lib, 'prh_fmg_' + _sfx(dtype))
        check(fn(ws.handle, df.ptr, dv.ptr, nm, bs, bc, beta, damping_factor, int(niterations),
                 int(vcycle_niterations), lp, None), 'fmg')
        v[...] = dv.numpy()
        return v

    def fmg_mixed(self, f, boxsize, boxcenter, beta, damping_factor=0.4, niterations=5, vcycle_niterations=6,
                  los=None, nrefine=0, vcycles_per_refine=1):
        ws = self.workspace(f.shape, np.float64, flags=_lib.WS_MULTIGRID | _lib.WS_MIXED)
        dv, df = DeviceArray.zeros(f.shape, np.float64), DeviceArray.from_numpy(f)
        keep, nm, bs, bc, lp = self._geom(f.shape, boxsize, boxcenter, los, np.float64)
        norms = (ctypes.c_double * (vcycle_niterations + nrefine + 1))()
        check(lib.prh_fmg_mixed(ws.handle, df.ptr, dv.ptr, nm, bs, bc, beta, damping_factor, int(niterations),
                                int(vcycle_niterations), lp, int(nrefine), int(vcycles_per_refine), norms, None),
              'fmg_mixed')
        return dv.numpy(), np.array(norms[:])

    def residual_norm(self, v, f, boxsize, boxcenter, beta, los=None):
        ws = self.workspace(f.shape, np.float64)
        dv, df = DeviceArray.from_numpy(v), DeviceArray.from_numpy(f)
        keep, nm, bs, bc, lp = self._geom(f.shape, boxsize, boxcenter, los, np.float64)
        rms = ctypes.c_double()
        check(lib.prh_residual_norm_f64(ws.handle, dv.ptr, df.ptr, nm, bs, bc, beta, lp, ctypes.byref(rms), None),
              'residual_norm')
        return rms.value

    def assign_cic(self, mesh, positions, weights, chunk=1 << 16):
        dtype = mesh.dtype
        ws = self.workspace(mesh.shape, dtype, flags=_lib.WS_PAINT, chunk=chunk)
        dm = DeviceArray.from_numpy(mesh)
        dp = DeviceArray.from_numpy(np.ascontiguousarray(positions, dtype=dtype))
        dw = DeviceArray.from_numpy(np.ascontiguousarray(weights, dtype=dtype))
        nm, nm_ptr = _lib.size3(mesh.shape)
        fn = getattr(lib, 'prh_assign_cic_' + _sfx(dtype))
        check(fn(ws.handle, dm.ptr, nm_ptr, dp.ptr, dw.ptr, len(dw), None), 'assign_cic')
        mesh[...] = dm.numpy()

    def read_cic(self, mesh, positions):
        dtype = mesh.dtype
        dm = DeviceArray.from_numpy(mesh)
        dp = DeviceArray.from_numpy(np.ascontiguousarray(positions, dtype=dtype))
        out = DeviceArray((len(dp),), dtype)
        nm, nm_ptr = _lib.size3(mesh.shape)
        check(getattr(lib, 'prh_read_cic_' + _sfx(dtype))(dm.ptr, nm_ptr, dp.ptr, out.ptr, len(dp), None), 'read_cic')
        return out.numpy()

    def read_finite_difference_cic(self, mesh, boxsize, positions):
        dtype = mesh.dtype
        dm = DeviceArray.from_numpy(mesh)
        dp = DeviceArray.from_numpy(np.ascontiguousarray(positions, dtype=dtype))
        out = DeviceArray((len(dp), 3), dtype)
        nm, nm_ptr = _lib.size3(mesh.shape)
        bs, bs_ptr = _lib.vec3(boxsize, dtype)
        fn = getattr(lib, 'prh_read_finite_difference_cic_' + _sfx(dtype))
        check(fn(dm.ptr, nm_ptr, bs_ptr, dp.ptr, out.ptr, len(dp), None), 'read_fd')
        return out.numpy()
