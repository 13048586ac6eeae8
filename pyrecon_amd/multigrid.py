"""Drop-in ``MultiGridReconstruction`` on MI355X (pyrecon/multigrid.py:161-167).

The constructor and methods are the reference's; the work runs in the HIP
library: deterministic CIC paint, hipFFT smoothing and density contrast, the
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
from .recon import BaseReconstruction, ReconstructionError, _suffix


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

    def run(self, jacobi_damping_factor=0.4, jacobi_niterations=5, vcycle_niterations=6, precision='plain',
            nrefine=0, vcycles_per_refine=1, residual_norms=False):
        """Solve for the displacement potential :attr:`mesh_phi` (pyrecon/multigrid.py:99-135).

        Parameters
        ----------
        jacobi_damping_factor, jacobi_niterations, vcycle_niterations : as the reference.
        precision : 'plain' (the reference's algorithm in the mesh dtype) or 'mixed'
            (FP64 meshes: the coarse full-multigrid stages in FP32, then the
            ``vcycle_niterations`` finest-level V-cycles as FP64 defect corrections with FP32
            V-cycles, tracking the plain FP64 iterate; ``nrefine`` extra corrections may follow).
        residual_norms : if True (mixed only), store the FP64 RMS residual before each finest-level
            correction and at the end in :attr:`residual_norms` (one synchronisation each).
        """
        if getattr(self, 'mesh_delta', None) is None:
            raise ReconstructionError('set the density contrast before running')
        dtype = self.info.dtype
        delta = self.mesh_delta.buffer
        phi = RealMesh(info=self.info, workspace=self.workspace, stream=self.stream)
        phi._allocate(zero=True)
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


class MultiGridReconstruction(OriginalMultiGridReconstruction):
    """Multigrid reconstruction with the base-class density contrast (pyrecon/multigrid.py:161-167)."""

    def set_density_contrast(self, *args, **kwargs):
        BaseReconstruction.set_density_contrast(self, *args, **kwargs)
