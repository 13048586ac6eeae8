"""Drop-in ``IterativeFFTReconstruction`` on MI355X (pyrecon/iterative_fft.py:7-82).

Burden et al. 2015, field-level iterations (eq. 22 and 24 of arXiv:1504.02591): each iteration
transforms the current real-space density, divides by k^2, and removes the RSD term from the
redshift-space density, with a global line of sight as beta IFFT((k.los)^2 delta_k / k^2) and with
a local one as beta sum_ij x_i x_j / |x|^2 IFFT(k_i k_j delta_k / k^2) (6 inverse transforms).
Transforms, Fourier-space operators and the real-space updates run in the HIP library.
"""
from . import _lib
from .device import stream_handle
from .recon import SpectralReconstruction, ReconstructionError, _suffix


class IterativeFFTReconstruction(SpectralReconstruction):
    """Implementation of Burden et al. 2015, field-level algorithm (pyrecon/iterative_fft.py:7-82)."""

    def run(self, niterations=3):
        """Compute the Zel'dovich displacement meshes :attr:`mesh_psi` (iterative_fft.py:12-28)."""
        if getattr(self, 'mesh_delta', None) is None or self.mesh_delta.buffer is None:
            raise ReconstructionError('set the density contrast before running')
        self._iter = 0
        self.mesh_delta_real = self.mesh_delta.deepcopy()
        for _ in range(niterations):
            self._iterate()
        self.mesh_delta = None
        self.mesh_psi = self._compute_psi()
        self.mesh_delta_real = None

    def _rsd_update(self, t, factor, a=-1, b=-1):
        nm, nm_ptr = _lib.size3(self.info.shape)
        bs, bs_ptr = _lib.vec3(self.info.boxsize, self.info.dtype)
        bc, bc_ptr = _lib.vec3(self.info.boxcenter, self.info.dtype)
        fn = getattr(_lib.lib, 'prh_rsd_update_' + _suffix(self.info.dtype))
        _lib.check(fn(self.mesh_delta_real.buffer.ptr, t.buffer.ptr, nm_ptr, bs_ptr, bc_ptr, float(factor), int(a),
                      int(b), stream_handle(self.stream)), 'rsd update')

    def _iterate(self):
        """One iteration (iterative_fft.py:30-67)."""
        self.logger.info('Running iteration {:d}.'.format(self._iter))
        delta_k = self._spectrum(self.mesh_delta_real)
        self._spectral(_lib.SPEC_INV_K2, delta_k, delta_k)
        self.mesh_delta_real.buffer.copy_(self.mesh_delta.buffer, stream=self.stream)
        norm = 1. / self.info.size
        work = self._complex()
        t = self._new_mesh()
        if self.los is not None:
            self._spectral(_lib.SPEC_LOS_SQ, delta_k, work, norm=norm, los=self.los)
            self._to_real(work, out=t)
            factor = self.beta
            if self._iter == 0:
                # Burden et al. 2015, eq. 12 (flat-sky approximation)
                factor /= (1. + self.beta)
            self._rsd_update(t, factor)
        else:
            for i in range(3):
                for j in range(i, 3):
                    self._spectral(_lib.SPEC_KA_KB, delta_k, work, a=i, b=j, norm=norm)
                    self._to_real(work, out=t)
                    # j > i counted twice for j < i
                    factor = (1. + (i != j)) * self.beta
                    if self._iter == 0:
                        factor /= (1. + self.beta)
                    self._rsd_update(t, factor, a=i, b=j)
        self._iter += 1

    def _compute_psi(self):
        """Zel'dovich displacements of the real-space density (iterative_fft.py:69-82)."""
        delta_k = self._spectrum(self.mesh_delta_real)
        norm = 1. / self.info.size
        work = self._complex()
        psis = []
        for iaxis in range(3):
            self._spectral(_lib.SPEC_GRAD, delta_k, work, a=iaxis, norm=norm)
            psis.append(self._to_real(work))
        return psis

