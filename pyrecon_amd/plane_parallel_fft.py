"""Drop-in ``PlaneParallelFFTReconstruction`` on MI355X (pyrecon/plane_parallel_fft.py:7-55).

Eisenstein et al. 2007 with a global line of sight: psi_k = i k delta_k / (k^2 (1 + beta mu^2)).
The forward transform, the Fourier-space operator (one per axis, with the 1 / N of the inverse
folded in) and the inverse transforms run in the HIP library; the paint, smoothing, density
contrast and CIC reads are the multigrid path's.
"""
from . import _lib
from .recon import SpectralReconstruction, ReconstructionError


class PlaneParallelFFTReconstruction(SpectralReconstruction):
    """Implementation of Eisenstein et al. 2007 (pyrecon/plane_parallel_fft.py:7-55)."""

    def __init__(self, los=None, **kwargs):
        super(PlaneParallelFFTReconstruction, self).__init__(los=los, **kwargs)

    def set_los(self, los=None):
        """Set a global line of sight: 'x', 'y', 'z' or a 3-vector (plane_parallel_fft.py:27-39)."""
        super(PlaneParallelFFTReconstruction, self).set_los(los=los)
        if self.los is None:
            raise ReconstructionError('A (global) line-of-sight must be provided')

    def run(self):
        """Compute the Zel'dovich displacement meshes :attr:`mesh_psi` (plane_parallel_fft.py:41-55)."""
        if getattr(self, 'mesh_delta', None) is None or self.mesh_delta.buffer is None:
            raise ReconstructionError('set the density contrast before running')
        delta_k = self._spectrum(self.mesh_delta)
        self.mesh_delta = None
        norm = 1. / self.info.size
        work = self._complex()
        psis = []
        for iaxis in range(3):
            self._spectral(_lib.SPEC_PLANE_PARALLEL, delta_k, work, a=iaxis, norm=norm, los=self.los, beta=self.beta)
            psis.append(self._to_real(work))
        self.mesh_psi = psis
