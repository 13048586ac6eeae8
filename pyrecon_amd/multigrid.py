"""Drop-in ``MultiGridReconstruction`` on MI355X (pyrecon/multigrid.py:161-167).

The constructor and methods are the reference's; the work runs in the HIP
library: deterministic CIC paint, Fourier smoothing and density contrast, the
full-multigrid solve (bit-identical to the reference's C in ``precision='plain'``)
and the fused finite-difference read of the shifts.

``run(precision='mixed')`` (FP64 meshes only) solves with FP32 V-cycles inside
an FP64 defect-correction loop instead of the plain FP64 multigrid.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import lib, check
from .device import DeviceArray, stream_handle
from .mesh import RealMesh
from .recon import BaseReconstruction, ReconstructionError, _suffix, numpy_mean


_OWN = object()  # "this reconstruction's own stream"


class OriginalMultiGridReconstruction(BaseReconstruction):
    """Multigrid solve of Martin J. White's reconstruction (pyrecon/multigrid.py:13-158)."""

    @staticmethod
    def _select_nmesh(nmesh):
        # the closest size >= nmesh of the form 2^n, 3 2^(n-1), 5 2^(n-2), 7 2^(n-2) (multigrid.py:23-36)
        toret = []
        for n in nmesh:
            nbits = int(n).bit_length() - 1
            ntries = [2**nbits, 3 * 2**(nbits - 1), 5 * 2**(nbits - 2), 7 * 2**(nbits - 2), 2**(nbits + 1)]
            mindiff, iclosest = n, None
            for itry, ntry in enumerate(ntries):
                diff = ntry - n
                if diff >= 0 and diff < mindiff:
                    mindiff, iclosest = diff, itry
            toret.append(ntries[iclosest])
        return np.array(toret, dtype='i8')

    def __init__(self, *args, **kwargs):
        kwargs.setdefault('fft_plan', 'estimate')
        super(OriginalMultiGridReconstruction, self).__init__(*args, **kwargs)

    def set_density_contrast(self, ran_min=0.75, smoothing_radius=15., **kwargs):
        """Martin White's density contrast (pyrecon/multigrid.py:49-97).

        With randoms: cells with 0 < randoms < ran_min lose their data, alpha is the data sum over
        the randoms sum where randoms >= ran_min, delta = (data / (alpha randoms) - 1) / bias inside
        that mask and 0 outside, the mean of the non-zero cells is removed, then delta is smoothed.
        Without randoms: delta = (data / mean - 1) / bias, then smoothed.  The sums are
        deterministic FP64 (numpy sums in the mesh precision); the scalar algebra uses numpy's types.
        """
        dtype = self.info.dtype
        rtype = dtype.type
        sfx = _suffix(dtype)
        data = self.mesh_data
        if data.buffer is None:
            raise ReconstructionError('assign data before setting the density contrast')
        stream = stream_handle(self.stream)
        size = self.info.size
        if not self.has_randoms:
            s_data, _ = self._mesh_sums(data.buffer)
            mean = numpy_mean(rtype, s_data, size)
            check(getattr(lib, 'prh_density_contrast_' + sfx)(data.buffer.ptr, None, data.buffer.ptr, size, float(mean),
                                                             0., 0., float(rtype(self.bias)), stream), 'density contrast')
        else:
            randoms = self.mesh_randoms
            sums = (ctypes.c_double * 2)()
            check(getattr(lib, 'prh_martin_sums_' + sfx)(self.workspace.handle, data.buffer.ptr, randoms.buffer.ptr, size,
                                                        float(rtype(ran_min)), sums, stream), 'density sums')
            alpha = rtype(sums[0]) / rtype(sums[1])
            check(getattr(lib, 'prh_martin_contrast_' + sfx)(self.workspace.handle, data.buffer.ptr, randoms.buffer.ptr,
                                                            data.buffer.ptr, size, float(rtype(ran_min)), float(alpha),
                                                            float(rtype(self.bias)), sums, stream), 'density contrast')
            count = int(round(sums[1]))
            if count:
                mean = numpy_mean(rtype, sums[0], count)
                check(getattr(lib, 'prh_subtract_nonzero_mean_' + sfx)(data.buffer.ptr, size, float(mean), stream),
                      'mean removal')
            self.alpha = alpha
        self.mesh_delta = data
        self.mesh_data = None
        self.mesh_randoms = None
        self.mesh_delta.smooth_gaussian(smoothing_radius, **kwargs)

    def run(self, jacobi_damping_factor=0.4, jacobi_niterations=5, vcycle_niterations=6, precision=None,
            nrefine=0, vcycles_per_refine=1, residual_norms=False):
        """Solve for the displacement potential :attr:`mesh_phi` (pyrecon/multigrid.py:99-135).

        Parameters
        ----------
        jacobi_damping_factor, jacobi_niterations, vcycle_niterations : as the reference.
        precision : 'plain' (the reference's algorithm in the mesh dtype) or 'mixed'
            (FP64 meshes: the coarse full-multigrid stages in FP32, then the
            ``vcycle_niterations`` finest-level V-cycles as FP64 defect corrections with FP32
            V-cycles, tracking the plain FP64 iterate; ``nrefine`` extra corrections may follow).
            None (default): the constructor's ``precision``.
        residual_norms : if True (mixed only), store the FP64 RMS residual before each finest-level
            correction and at the end in :attr:`residual_norms` (one synchronisation each).
        """
        dtype = self.info.dtype
        if precision is None: precision = self.precision
        delta = self._solve_source()
        phi = self._new_phi()
        nm, nm_ptr = _lib.size3(self.info.shape)
        bs, bs_ptr = _lib.vec3(self.info.boxsize, dtype)
        bc, bc_ptr = _lib.vec3(self.info.boxcenter, dtype)
        los_ptr = None
        if self.los is not None:
            los_arr, los_ptr = _lib.vec3(self.los, dtype)
        self.logger.info('Computing displacement potential.')
        stream = stream_handle(self.stream)
        if precision == 'plain':
            fn = getattr(lib, 'prh_fmg_' + _suffix(dtype))
            check(fn(self.workspace.handle, delta.ptr, phi.buffer.ptr, nm_ptr, bs_ptr, bc_ptr, self.beta,
                     jacobi_damping_factor, int(jacobi_niterations), int(vcycle_niterations), los_ptr, stream), 'fmg')
        elif precision == 'mixed':
            if dtype != np.float64:
                raise ValueError('precision="mixed" needs an FP64 mesh (dtype="f8")')
            norms = (ctypes.c_double * (int(vcycle_niterations) + int(nrefine) + 1))() if residual_norms else None
            check(lib.prh_fmg_mixed(self.workspace.handle, delta.ptr, phi.buffer.ptr, nm_ptr, bs_ptr, bc_ptr,
                                    float(self.beta), float(jacobi_damping_factor), int(jacobi_niterations),
                                    int(vcycle_niterations), los_ptr, int(nrefine), int(vcycles_per_refine),
                                    norms, stream), 'fmg_mixed')
            if residual_norms: self.residual_norms = np.array(norms[:])
        else:
            raise ValueError('precision must be "plain" or "mixed"')
        self.mesh_phi = phi
        self.mesh_delta = None


    def _solve_source(self):
        if getattr(self, 'mesh_delta', None) is None or self.mesh_delta.buffer is None:
            raise ReconstructionError('set the density contrast before running')
        return self.mesh_delta.buffer

    def _new_phi(self, zero_stream=_OWN):
        """Potential mesh for the solve, zero-filled where the solve starts from it: the full
        multigrid overwrites the iterate with the prolongation from the coarser level whenever
        the mesh can be coarsened (every extent even and above 4, src/multigrid.c:303-328), and
        only starts from the given iterate otherwise.  The fill is queued on ``zero_stream``
        (default: this reconstruction's stream), the stream of the solve that writes it."""
        phi = RealMesh(info=self.info, workspace=self.workspace, stream=self.stream)
        phi._allocate(zero=False)
        if not all(n > 4 and n % 2 == 0 for n in self.info.shape):
            phi.buffer.fill_zero(stream=self.stream if zero_stream is _OWN else zero_stream)
        return phi


MAX_GROUP = 8


def run_group(recons, jacobi_damping_factor=0.4, jacobi_niterations=5, vcycle_niterations=6, precision='plain',
              nrefine=0, vcycles_per_refine=1, stream=None):
    """Run the multigrid solve of several reconstructions at once.

    All reconstructions share the mesh shape and dtype and have their own workspace; the
    coarse levels of all of them are relaxed by the same launches.  Each gets the potential
    :meth:`OriginalMultiGridReconstruction.run` would give it alone, bit for bit.  The solve
    is queued on ``stream``; work on the reconstructions' own streams that produced the
    density contrasts must be complete or ordered before it (events), and later work on
    them ordered after it.
    """
    recons = list(recons)
    count = len(recons)
    if not 1 <= count <= MAX_GROUP:
        raise ValueError('a group holds 1 to {:d} reconstructions'.format(MAX_GROUP))
    first = recons[0]
    dtype, shape = first.info.dtype, tuple(first.info.shape)
    for r in recons:
        if tuple(r.info.shape) != shape or r.info.dtype != dtype:
            raise ValueError('grouped reconstructions must share mesh shape and dtype')
    if len(set(id(r.workspace) for r in recons)) != count:
        raise ValueError('grouped reconstructions need distinct workspaces')
    if precision == 'mixed' and dtype != np.float64:
        raise ValueError('precision="mixed" needs an FP64 mesh (dtype="f8")')
    if precision not in ('plain', 'mixed'):
        raise ValueError('precision must be "plain" or "mixed"')
    deltas = [r._solve_source() for r in recons]
    phis = [r._new_phi(zero_stream=stream) for r in recons]
    real = ctypes.c_float if dtype == np.float32 else ctypes.c_double
    ws = (ctypes.c_void_p * count)(*[r.workspace.handle for r in recons])
    f = (ctypes.c_void_p * count)(*[d.ptr for d in deltas])
    v = (ctypes.c_void_p * count)(*[p.buffer.ptr for p in phis])
    bs = (real * (3 * count))(*[x for r in recons for x in r.info.boxsize])
    bc = (real * (3 * count))(*[x for r in recons for x in r.info.boxcenter])
    beta = (real * count)(*[r.beta for r in recons])
    los_keep = [None if r.los is None else (real * 3)(*r.los) for r in recons]
    los = (ctypes.c_void_p * count)(*[None if l is None else ctypes.addressof(l) for l in los_keep])
    nm, nm_ptr = _lib.size3(shape)
    handle = stream_handle(stream)
    if precision == 'plain':
        fn = getattr(lib, 'prh_fmg_group_' + _suffix(dtype))
        check(fn(ws, count, f, v, nm_ptr, bs, bc, beta, jacobi_damping_factor, int(jacobi_niterations),
                 int(vcycle_niterations), los, handle), 'fmg_group')
    else:
        check(lib.prh_fmg_mixed_group(ws, count, f, v, nm_ptr, bs, bc, beta, float(jacobi_damping_factor),
                                      int(jacobi_niterations), int(vcycle_niterations), los, int(nrefine),
                                      int(vcycles_per_refine), handle), 'fmg_mixed_group')
    for r, phi in zip(recons, phis):
        r.mesh_phi = phi
    return deltas


class MultiGridReconstruction(OriginalMultiGridReconstruction):
    """Multigrid reconstruction with the base-class density contrast (pyrecon/multigrid.py:161-167)."""

    def set_density_contrast(self, *args, **kwargs):
        BaseReconstruction.set_density_contrast(self, *args, **kwargs)
