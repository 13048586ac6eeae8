"""Drop-in ``OriginalIterativeFFTParticleReconstruction`` / ``IterativeFFTParticleReconstruction`` on MI355X
(pyrecon/iterative_fft_particle.py:9-245; Bautista et al. 2018).

Each iteration repaints the data at their current real-space estimate (wrapped), smooths, forms the
density contrast, transforms, divides by k^2 and reads the displacement at the data; the RSD part
of the displacement (reduced by beta / (1 + beta) in the first iteration) is then removed from the
original positions.  Paint, smoothing, density contrast, transforms, Fourier-space operators and
CIC reads run in the HIP library; the per-point RSD algebra is the reference's numpy.
"""
import numpy as np

from . import _lib
from ._lib import lib, check
from .device import DeviceArray, as_device, stream_handle
from .mesh import RealMesh
from .recon import BaseReconstruction, SpectralReconstruction, ReconstructionError, _suffix, distance, numpy_mean


class OriginalIterativeFFTParticleReconstruction(SpectralReconstruction):
    """Bautista et al. 2018 (pyrecon/iterative_fft_particle.py:9-239)."""

    def assign_data(self, positions, weights=None):
        """Paint the data and keep their positions (for :meth:`run`) and weights (iterative_fft_particle.py:16-31)."""
        if isinstance(positions, DeviceArray): positions = positions.numpy(stream=self.stream)
        if isinstance(weights, DeviceArray): weights = weights.numpy(stream=self.stream)
        positions = np.asarray(positions)
        if weights is None:
            weights = np.ones_like(positions, shape=(len(positions),))
        if self.wrap: positions = self.info.wrap(positions)
        if self.mesh_data.buffer is None:
            self._positions_data = positions
            self._weights_data = weights
        else:
            self._positions_data = np.concatenate([self._positions_data, positions], axis=0)
            self._weights_data = np.concatenate([self._weights_data, weights], axis=0)
        self.mesh_data.assign_cic(positions, weights=weights, wrap=self.wrap)

    def set_density_contrast(self, ran_min=0.01, smoothing_radius=15.):
        """Density contrast of the (already smoothed) meshes (iterative_fft_particle.py:33-59); the data and
        randoms meshes are kept.  Sums are deterministic FP64."""
        self.ran_min = ran_min
        self.smoothing_radius = smoothing_radius
        dtype = self.info.dtype
        rtype = dtype.type
        delta = RealMesh(info=self.info, workspace=self.workspace, stream=self.stream)
        delta._allocate(zero=False)
        fn = getattr(lib, 'prh_density_contrast_' + _suffix(dtype))
        data = self.mesh_data.buffer
        if self.has_randoms:
            randoms = self.mesh_randoms.buffer
            s_data, s_randoms = self._mesh_sums(data, randoms)
            sum_data, sum_randoms = rtype(s_data), rtype(s_randoms)
            alpha = sum_data * 1. / sum_randoms
            threshold = ran_min * sum_randoms / self._size_randoms
            status = fn(data.ptr, randoms.ptr, delta.buffer.ptr, self.info.size, float(alpha), float(threshold),
                        float(self.bias * alpha), float(rtype(self.bias)), stream_handle(self.stream))
        else:
            s_data, _ = self._mesh_sums(data)
            mean = numpy_mean(rtype, s_data, self.info.size)
            status = fn(data.ptr, None, delta.buffer.ptr, self.info.size, float(mean), 0., 0., float(rtype(self.bias)),
                        stream_handle(self.stream))
        check(status, 'density contrast')
        self.mesh_delta = delta

    def run(self, niterations=3):
        """Reconstructed real-space data positions and the displacement meshes :attr:`mesh_psi`
        (iterative_fft_particle.py:61-78)."""
        if self.mesh_data.buffer is None or not hasattr(self, 'smoothing_radius'):
            raise ReconstructionError('assign data and set the density contrast before running')
        self._iter = 0
        self.mesh_data.smooth_gaussian(self.smoothing_radius)
        if self.has_randoms: self.mesh_randoms.smooth_gaussian(self.smoothing_radius)
        self._positions_rec_data = self._positions_data.copy()
        for it in range(niterations):
            self.mesh_psi = self._iterate(return_psi=it == niterations - 1)
        self.mesh_randoms = None

    def _iterate(self, return_psi=False):
        """One iteration (iterative_fft_particle.py:80-139)."""
        self.logger.info('Running iteration {:d}.'.format(self._iter))
        if self._iter > 0:
            self.mesh_data = RealMesh(info=self.info, workspace=self.workspace, stream=self.stream)
            # paint the reconstructed real-space data positions, wrapped
            wrap = self.wrap
            self.wrap = True
            BaseReconstruction.assign_data(self, self._positions_rec_data, weights=self._weights_data)
            self.wrap = wrap
            self.mesh_data.smooth_gaussian(self.smoothing_radius)
        self.set_density_contrast(ran_min=self.ran_min, smoothing_radius=self.smoothing_radius)
        self.mesh_data = None
        delta_k = self._spectrum(self.mesh_delta)
        self.mesh_delta = None
        self._spectral(_lib.SPEC_INV_K2, delta_k, delta_k)
        shifts = np.empty_like(self._positions_rec_data)
        pos, _ = as_device(self._positions_rec_data, stream=self.stream)
        norm = 1. / self.info.size
        work = self._complex()
        psis = []
        for iaxis in range(3):
            # no displacement is needed along an axis the line of sight does not touch
            if not return_psi and self.los is not None and self.los[iaxis] == 0:
                shifts[:, iaxis] = 0.
                continue
            self._spectral(_lib.SPEC_I_KA, delta_k, work, a=iaxis, norm=norm)
            psi = self._to_real(work)
            shifts[:, iaxis] = psi.read_cic(pos, wrap=True).numpy(stream=self.stream)
            if return_psi: psis.append(psi)
        if self.los is None:
            los = self._positions_data / distance(self._positions_data)[:, None]
        else:
            los = self.los
        if self._iter == 0:
            # first iteration: remove the RSD part approximately (Burden et al. 2015, eq. 12)
            shifts -= self.beta / (1 + self.beta) * np.sum(shifts * los, axis=-1)[:, None] * los
        # the RSD is removed from the original positions; the next iteration paints these
        self._positions_rec_data = self._positions_data - self.f * np.sum(shifts * los, axis=-1)[:, None] * los
        diff = self._positions_rec_data - self.info.offset
        if (not self.wrap) and np.any((diff < 0) | (diff > self.info.boxsize - self.info.cellsize)):
            self.logger.warning('Some particles are out-of-bounds.')
        self._iter += 1
        if return_psi:
            return psis

    def _read_psi(self, positions, wrap=False):
        shifts = np.empty_like(positions)
        pos, _ = as_device(positions, stream=self.stream)
        for iaxis, psi in enumerate(self.mesh_psi):
            shifts[:, iaxis] = psi.read_cic(pos, wrap=wrap).numpy(stream=self.stream)
        return shifts

    def read_shifts(self, positions, field='disp+rsd'):
        """Displacements at ``positions``, or with ``positions='data'`` at the reconstructed data
        (iterative_fft_particle.py:141-212)."""
        field = field.lower()
        allowed_fields = ['disp', 'rsd', 'disp+rsd']
        if field not in allowed_fields:
            raise ReconstructionError('Unknown field {}. Choices are {}'.format(field, allowed_fields))
        if getattr(self, 'mesh_psi', None) is None:
            raise ReconstructionError('run the reconstruction before reading shifts')
        if isinstance(positions, str) and positions == 'data':
            shifts = self._read_psi(self._positions_rec_data, wrap=True)
            if field == 'disp':
                return shifts
            rsd = self._positions_data - self._positions_rec_data
            if field == 'rsd':
                return rsd
            shifts += rsd
            return shifts
        if isinstance(positions, DeviceArray): positions = positions.numpy(stream=self.stream)
        positions = np.asarray(positions)
        if self.wrap: positions = self.info.wrap(positions)
        shifts = self._read_psi(positions, wrap=False)
        if field == 'disp':
            return shifts
        if self.los is None:
            los = positions / distance(positions)[:, None]
        else:
            los = self.los
        rsd = self.f * np.sum(shifts * los, axis=-1)[:, None] * los
        if field == 'rsd':
            return rsd
        # the original algorithm's convention: RSD removed first, then the displacement read there
        real_positions = positions - rsd
        diff = real_positions - self.info.offset
        if (not self.wrap) and np.any((diff < 0) | (diff > self.info.boxsize - self.info.cellsize)):
            self.logger.warning('Some particles are out-of-bounds.')
        shifts = self._read_psi(real_positions, wrap=True)
        return shifts + rsd

    def read_shifted_positions(self, positions, field='disp+rsd'):
        """``positions - read_shifts(positions)`` (``'data'``: the data given to :meth:`assign_data`), wrapped
        if :attr:`wrap` (iterative_fft_particle.py:214-239)."""
        if isinstance(positions, DeviceArray): positions = positions.numpy(stream=self.stream)
        shifts = self.read_shifts(positions, field=field)
        if isinstance(positions, str) and positions == 'data':
            positions = self._positions_data
        positions = positions - shifts
        if self.wrap: positions = self.info.wrap(positions)
        return positions


class IterativeFFTParticleReconstruction(OriginalIterativeFFTParticleReconstruction):
    """Any update upon the original algorithm (iterative_fft_particle.py:242-245)."""
