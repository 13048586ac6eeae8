"""End-to-end CPU oracle of ``MultiGridReconstruction`` -- TEST INFRASTRUCTURE ONLY.

A restatement of the reference's Python layer with numpy, calling the C oracle
(:mod:`oracle.cpu`) for the native kernels.  It follows, step for step:

- mesh geometry and the mesh-unit transform: pyrecon/mesh.py:462-479, 529-568,
  612-653 (``offset``, ``wrap``, range/finite checks, ``((x - c) / L + 0.5) * n``);
- FFT Gaussian smoothing with numpy's complex-to-complex transform, as the
  reference does with ``hermitian=False``: pyrecon/mesh.py:655-680, 697-705,
  812-820, 827-844, 974-988, pyrecon/recon.py:117;
- the density contrast: pyrecon/recon.py:198-249 (used by
  MultiGridReconstruction, pyrecon/multigrid.py:165-167);
- Martin White's density contrast of OriginalMultiGridReconstruction:
  pyrecon/multigrid.py:69-97;
- the solve: pyrecon/multigrid.py:99-135 (``beta = f / bias``, recon.py:122-125);
- the shifts: pyrecon/multigrid.py:137-158 and pyrecon/recon.py:304-325;
- the FFT reconstructions: pyrecon/plane_parallel_fft.py:41-55 and
  pyrecon/iterative_fft.py:12-82, with the base-class shifts of
  pyrecon/recon.py:255-325, and the particle-level iterations of
  pyrecon/iterative_fft_particle.py:9-239.

The numpy operations (and therefore the dtypes of every intermediate) are the
reference's, so with the same native kernels this reproduces the reference's
output bit for bit.  Nothing in the shipped package imports this module.
"""
import numpy as np

from .cpu import OracleKernels


class OracleError(Exception):
    """Raised where the reference raises MeshError / ReconstructionError."""


def distance(positions):
    """Cartesian norm along the last axis (pyrecon/utils.py:146-148)."""
    return np.sqrt((positions**2).sum(axis=-1))


def _vec3(value, dtype):
    out = np.empty(3, dtype=dtype)
    out[...] = value
    return out


class OracleMultiGridReconstruction(object):
    """Same constructor arguments and methods as the reference's class, for explicit
    ``nmesh``, ``boxsize`` and ``boxcenter``."""

    def __init__(self, f=0., bias=1., los=None, nmesh=None, boxsize=None, boxcenter=None, dtype='f8', wrap=False,
                 kernels=None, fft_workers=None):
        self.f = f
        # fft_workers: None -> numpy.fft (the reference's engine); an int -> scipy.fft with that many
        # threads, the same pocketfft algorithm, for meshes where numpy's single thread is too slow
        self.fft_workers = fft_workers
        self.bias = bias
        self.wrap = wrap
        self.dtype = np.dtype(dtype)
        # MeshInfo stores nmesh as c_int and boxsize / boxcenter in the mesh dtype (mesh.py:447-460)
        self.nmesh = _vec3(nmesh, np.intc)
        self.boxsize = _vec3(boxsize, self.dtype)
        self.boxcenter = _vec3(boxcenter, self.dtype)
        self.kernels = kernels if kernels is not None else OracleKernels()
        if los in [None, 'local']:
            self.los = None
        else:
            if isinstance(los, str):
                los = 'xyz'.index(los)
            if np.ndim(los) == 0:
                ilos = los
                los = np.zeros(3, dtype=self.dtype)
                los[ilos] = 1.
            los = np.array(los, dtype=self.dtype)
            self.los = los / distance(los)
        self.mesh_data = None
        self.mesh_randoms = None
        self._size_randoms = 0

    @property
    def beta(self):
        return self.f / self.bias

    @property
    def offset(self):
        return self.boxcenter - self.boxsize / 2.

    @property
    def shape(self):
        return tuple(int(n) for n in self.nmesh)

    def wrap_positions(self, positions):
        return ((positions - self.offset) % self.boxsize) + self.offset

    def _check_range(self, positions):
        low, high = self.boxcenter - self.boxsize / 2., self.boxcenter + self.boxsize / 2.
        if not np.all((positions >= low) & (positions <= high)):
            raise OracleError('positions not in box range {} - {}'.format(low, high))

    def mesh_units(self, positions):
        """Transform to mesh units and cast to the mesh dtype (mesh.py:554-555)."""
        positions = ((positions - self.boxcenter) / self.boxsize + 0.5) * self.nmesh
        return positions.astype(self.dtype, copy=False)

    def _paint(self, mesh, positions, weights):
        size = len(positions)
        if weights is None: weights = np.ones_like(positions, shape=size, dtype=self.dtype)
        if self.wrap:
            positions = self.wrap_positions(positions)
        else:
            self._check_range(positions)
        if not (np.isfinite(positions).all() and np.isfinite(weights).all()):
            raise OracleError('Some positions or weights are NaN/inf')
        positions = self.mesh_units(positions)
        weights = weights.astype(self.dtype, copy=False)
        if mesh is None: mesh = np.zeros(self.shape, dtype=self.dtype)
        self.kernels.assign_cic(mesh, positions, weights)
        return mesh

    def assign_data(self, positions, weights=None):
        self.mesh_data = self._paint(self.mesh_data, positions, weights)

    def assign_randoms(self, positions, weights=None):
        if self.mesh_randoms is None: self._size_randoms = 0
        self.mesh_randoms = self._paint(self.mesh_randoms, positions, weights)
        self._size_randoms += len(positions)

    @property
    def has_randoms(self):
        return self.mesh_randoms is not None

    def smooth_gaussian(self, value, radius):
        """Complex-to-complex FFT smoothing (hermitian=False), numpy engine."""
        radius_ = np.empty_like(self.boxsize, order='C')
        radius_[:] = radius
        ctype = np.dtype('c{:d}'.format(2 * self.dtype.itemsize))
        valuek = self._fftn(value).astype(ctype, copy=False)
        coords = []
        for n, d in zip(self.nmesh, self.boxsize / self.nmesh):
            coords.append(2 * np.pi * np.fft.fftfreq(n, d=d))
        grids = []
        for iaxis, k in enumerate(coords):
            sl = [None] * 3
            sl[iaxis] = slice(None)
            grids.append(k[tuple(sl)])
        k2 = sum(-0.5 * (r * k)**2 for r, k in zip(radius_, grids))
        valuek *= np.exp(k2)
        real = self._fftn(valuek, inverse=True).astype(ctype, copy=False).real
        return real.astype(self.dtype, copy=False)

    def _fftn(self, value, inverse=False):
        workers = getattr(self, 'fft_workers', None)
        if workers is None:
            return (np.fft.ifftn if inverse else np.fft.fftn)(value)
        import scipy.fft
        return (scipy.fft.ifftn if inverse else scipy.fft.fftn)(value, workers=int(workers))

    def set_density_contrast(self, ran_min=0.01, smoothing_radius=15.):
        data = self.smooth_gaussian(self.mesh_data, smoothing_radius)
        if self.has_randoms:
            randoms = self.smooth_gaussian(self.mesh_randoms, smoothing_radius)
            sum_data, sum_randoms = np.sum(data), np.sum(randoms)
            alpha = sum_data * 1. / sum_randoms
            delta = data - alpha * randoms
            threshold = ran_min * sum_randoms / self._size_randoms
            mask = randoms > threshold
            delta[mask] /= (self.bias * alpha * randoms[mask])
            delta[~mask] = 0.
            self.alpha, self.threshold, self.mask_count = alpha, threshold, int(mask.sum())
        else:
            delta = data / np.mean(data) - 1.
            delta /= self.bias
        self.mesh_delta = delta
        self.mesh_data = self.mesh_randoms = None

    def run(self, jacobi_damping_factor=0.4, jacobi_niterations=5, vcycle_niterations=6):
        los = None if self.los is None else self.los.astype(self.dtype, copy=False)
        self.mesh_phi = self.kernels.fmg(self.mesh_delta, self.boxsize, self.boxcenter, self.beta,
                                         jacobi_damping_factor, jacobi_niterations, vcycle_niterations, los=los)
        del self.mesh_delta

    def read_finite_difference_cic(self, positions):
        dtype = positions.dtype
        if self.wrap:
            positions = self.wrap_positions(positions)
        else:
            self._check_range(positions)
        positions = self.mesh_units(positions)
        values = self.kernels.read_finite_difference_cic(self.mesh_phi, self.boxsize, positions)
        return values.astype(dtype=dtype, copy=False)

    def read_cic(self, mesh, positions):
        dtype = positions.dtype
        if self.wrap:
            positions = self.wrap_positions(positions)
        else:
            self._check_range(positions)
        positions = self.mesh_units(positions)
        return self.kernels.read_cic(mesh, positions).astype(dtype=dtype, copy=False)

    def read_shifts(self, positions, field='disp+rsd'):
        field = field.lower()
        if field not in ['disp', 'rsd', 'disp+rsd']:
            raise OracleError('Unknown field {}'.format(field))
        shifts = self.read_finite_difference_cic(positions)
        if field == 'disp':
            return shifts
        if self.los is None:
            los = positions / distance(positions)[:, None]
        else:
            los = self.los
        rsd = self.f * np.sum(shifts * los, axis=-1)[:, None] * los
        if field == 'rsd':
            return rsd
        shifts += rsd
        return shifts

    def read_shifted_positions(self, positions, field='disp+rsd'):
        shifts = self.read_shifts(positions, field=field)
        positions = positions - shifts
        if self.wrap: positions = self.wrap_positions(positions)
        return positions


class OracleOriginalMultiGridReconstruction(OracleMultiGridReconstruction):
    """The reference's OriginalMultiGridReconstruction: Martin White's density contrast."""

    def set_density_contrast(self, ran_min=0.75, smoothing_radius=15.):
        # pyrecon/multigrid.py:69-97
        data = np.array(self.mesh_data, copy=True)
        if not self.has_randoms:
            delta = data / np.mean(data) - 1.
            delta /= self.bias
        else:
            randoms = self.mesh_randoms
            mask = randoms >= ran_min
            data[(randoms > 0) & (randoms < ran_min)] = 0.
            alpha = np.sum(data) / np.sum(randoms[mask])
            data[mask] /= alpha * randoms[mask]
            delta = data
            delta -= 1
            delta[~mask] = 0.
            delta /= self.bias
            nonzero = delta != 0.
            delta[nonzero] -= np.mean(delta[nonzero])
            self.alpha, self.mask_count = alpha, int(mask.sum())
        self.mesh_delta = np.ascontiguousarray(self.smooth_gaussian(delta, smoothing_radius))
        self.mesh_data = self.mesh_randoms = None



class _OracleFFTReconstruction(OracleMultiGridReconstruction):
    """Shared parts of the FFT reconstructions: numpy's c2c engine (pyrecon/mesh.py:974-988, through
    RealMesh.to_complex and ComplexMesh.to_real, 697-705 and 827-844), the wavenumbers and mesh
    coordinates (ComplexMesh.coords 812-820, RealMesh.coords 522-527), and the base-class shifts,
    which read each of the three displacement meshes by CIC (pyrecon/recon.py:255-325)."""

    @property
    def ctype(self):
        return np.dtype('c{:d}'.format(2 * self.dtype.itemsize))

    def to_complex(self, value):
        return np.fft.fftn(value).astype(self.ctype, copy=False)

    def to_real(self, valuek):
        return np.fft.ifftn(valuek).astype(self.ctype, copy=False).real

    def kcoords(self):
        k = [2 * np.pi * np.fft.fftfreq(n, d=d) for n, d in zip(self.nmesh, self.boxsize / self.nmesh)]
        return _broadcast(k)

    def xcoords(self):
        x = [o + d * np.arange(n) for n, o, d in zip(self.nmesh, self.offset, self.boxsize / self.nmesh)]
        return _broadcast(x)

    def read_cic_nowrap(self, mesh, positions):
        """RealMesh.read_cic(positions, wrap=False) (mesh.py:570-610)."""
        dtype = positions.dtype
        self._check_range(positions)
        positions = self.mesh_units(positions)
        return self.kernels.read_cic(np.ascontiguousarray(mesh, dtype=self.dtype), positions).astype(dtype, copy=False)

    def read_shifts(self, positions, field='disp+rsd'):
        field = field.lower()
        if field not in ['disp', 'rsd', 'disp+rsd']:
            raise OracleError('Unknown field {}'.format(field))
        shifts = np.empty_like(positions)
        if self.wrap: positions = self.wrap_positions(positions)
        for iaxis, psi in enumerate(self.mesh_psi):
            shifts[:, iaxis] = self.read_cic_nowrap(psi, positions)
        if field == 'disp':
            return shifts
        if self.los is None:
            los = positions / distance(positions)[:, None]
        else:
            los = self.los
        rsd = self.f * np.sum(shifts * los, axis=-1)[:, None] * los
        if field == 'rsd':
            return rsd
        shifts += rsd
        return shifts


def _broadcast(arrays):
    """pyrecon/utils.py:125-143: 1D arrays to broadcastable 3D arrays."""
    out = []
    for iaxis, a in enumerate(arrays):
        sl = [None] * len(arrays)
        sl[iaxis] = slice(None)
        out.append(a[tuple(sl)])
    return out


class OraclePlaneParallelFFTReconstruction(_OracleFFTReconstruction):
    """Eisenstein et al. 2007 with a global line of sight (pyrecon/plane_parallel_fft.py:41-55)."""

    def run(self):
        if self.los is None:
            raise OracleError('A (global) line-of-sight must be provided')
        delta_k = self.to_complex(self.mesh_delta)
        del self.mesh_delta
        k = self.kcoords()
        k2 = sum(kk**2 for kk in k)
        k2[0, 0, 0] = 1.
        delta_k /= k2
        mu2 = sum(kk * ll for ll, kk in zip(k, self.los))**2 / k2
        psis = []
        for iaxis in range(3):
            tmp = 1j * k[iaxis] * delta_k / (1. + self.beta * mu2)
            psis.append(self.to_real(tmp))
        self.mesh_psi = psis


class OracleIterativeFFTReconstruction(_OracleFFTReconstruction):
    """Burden et al. 2015, field-level iterations (pyrecon/iterative_fft.py:12-82)."""

    def run(self, niterations=3):
        self._iter = 0
        self.mesh_delta_real = np.array(self.mesh_delta, copy=True)
        for _ in range(niterations):
            self._iterate()
        del self.mesh_delta
        self.mesh_psi = self._compute_psi()
        del self.mesh_delta_real

    def _iterate(self):
        delta_k = self.to_complex(self.mesh_delta_real)
        k = self.kcoords()
        # ComplexMesh.prod_sum(exp=-1) (src/mesh.c:327-365): squares cast to the mesh dtype and
        # summed as (y + x) + z in it; real and imaginary parts divided separately
        k2 = [np.asarray(kk**2, dtype=self.dtype) for kk in k]
        den = (k2[1] + k2[0]) + k2[2]
        with np.errstate(divide='ignore', invalid='ignore'):
            re = delta_k.real / den
            im = delta_k.imag / den
        delta_k = (re + 1j * im).astype(self.ctype, copy=False)
        delta_k[0, 0, 0] = 0.
        self.mesh_delta_real = np.array(self.mesh_delta, copy=True)
        if self.los is not None:
            disp_deriv_k = sum(kk * ll for ll, kk in zip(k, self.los))**2 * delta_k
            delta_rsd = self.to_real(disp_deriv_k)
            factor = self.beta
            if self._iter == 0:
                factor /= (1. + self.beta)
            self.mesh_delta_real -= factor * delta_rsd
        else:
            x = self.xcoords()
            x2 = sum(x_**2 for x_ in x)
            for i in range(3):
                for j in range(i, 3):
                    disp_deriv_k = k[i] * k[j] * delta_k
                    delta_rsd = x[i] * x[j] * self.to_real(disp_deriv_k) / x2
                    factor = (1. + (i != j)) * self.beta
                    if self._iter == 0:
                        factor /= (1. + self.beta)
                    self.mesh_delta_real -= factor * delta_rsd
        self._iter += 1

    def _compute_psi(self):
        delta_k = self.to_complex(self.mesh_delta_real)
        k = self.kcoords()
        k2 = sum(kk**2 for kk in k)
        k2[0, 0, 0] = 1.
        delta_k /= k2
        delta_k[0, 0, 0] = 0.
        return [self.to_real(1j * k[iaxis] * delta_k) for iaxis in range(3)]


class OracleOriginalIterativeFFTParticleReconstruction(_OracleFFTReconstruction):
    """Bautista et al. 2018, particle-level iterations (pyrecon/iterative_fft_particle.py:9-239)."""

    def assign_data(self, positions, weights=None):
        if weights is None:
            weights = np.ones_like(positions, shape=(len(positions),))
        if self.wrap: positions = self.wrap_positions(positions)
        if self.mesh_data is None:
            self._positions_data = positions
            self._weights_data = weights
        else:
            self._positions_data = np.concatenate([self._positions_data, positions], axis=0)
            self._weights_data = np.concatenate([self._weights_data, weights], axis=0)
        self.mesh_data = self._paint(self.mesh_data, positions, weights)

    def set_density_contrast(self, ran_min=0.01, smoothing_radius=15.):
        # the meshes are taken as already smoothed and are kept (iterative_fft_particle.py:33-59)
        self.ran_min = ran_min
        self.smoothing_radius = smoothing_radius
        if self.has_randoms:
            sum_data, sum_randoms = np.sum(self.mesh_data), np.sum(self.mesh_randoms)
            alpha = sum_data * 1. / sum_randoms
            delta = self.mesh_data - alpha * self.mesh_randoms
            threshold = ran_min * sum_randoms / self._size_randoms
            mask = self.mesh_randoms > threshold
            delta[mask] /= (self.bias * alpha * self.mesh_randoms[mask])
            delta[~mask] = 0.
        else:
            delta = self.mesh_data / np.mean(self.mesh_data) - 1.
            delta /= self.bias
        self.mesh_delta = delta

    def read_cic_wrap(self, mesh, positions):
        """RealMesh.read_cic(positions, wrap=True) (mesh.py:570-610)."""
        dtype = positions.dtype
        positions = self.mesh_units(self.wrap_positions(positions))
        return self.kernels.read_cic(np.ascontiguousarray(mesh, dtype=self.dtype), positions).astype(dtype, copy=False)

    def run(self, niterations=3):
        self._iter = 0
        self.mesh_data = self.smooth_gaussian(self.mesh_data, self.smoothing_radius)
        if self.has_randoms: self.mesh_randoms = self.smooth_gaussian(self.mesh_randoms, self.smoothing_radius)
        self._positions_rec_data = self._positions_data.copy()
        for it in range(niterations):
            self.mesh_psi = self._iterate(return_psi=it == niterations - 1)
        self.mesh_randoms = None

    def _iterate(self, return_psi=False):
        if self._iter > 0:
            wrap = self.wrap
            self.wrap = True
            self.mesh_data = self._paint(None, self._positions_rec_data, self._weights_data)
            self.wrap = wrap
            self.mesh_data = self.smooth_gaussian(self.mesh_data, self.smoothing_radius)
        self.set_density_contrast(ran_min=self.ran_min, smoothing_radius=self.smoothing_radius)
        self.mesh_data = None
        delta_k = self.to_complex(self.mesh_delta)
        del self.mesh_delta
        k = self.kcoords()
        # ComplexMesh.prod_sum(exp=-1), as in OracleIterativeFFTReconstruction._iterate
        k2 = [np.asarray(kk**2, dtype=self.dtype) for kk in k]
        den = (k2[1] + k2[0]) + k2[2]
        with np.errstate(divide='ignore', invalid='ignore'):
            delta_k = (delta_k.real / den + 1j * (delta_k.imag / den)).astype(self.ctype, copy=False)
        delta_k[0, 0, 0] = 0.
        shifts = np.empty_like(self._positions_rec_data)
        psis = []
        for iaxis in range(3):
            if not return_psi and self.los is not None and self.los[iaxis] == 0:
                shifts[:, iaxis] = 0.
                continue
            psi = self.to_real(delta_k * 1j * k[iaxis])
            shifts[:, iaxis] = self.read_cic_wrap(psi, self._positions_rec_data)
            if return_psi: psis.append(psi)
        if self.los is None:
            los = self._positions_data / distance(self._positions_data)[:, None]
        else:
            los = self.los
        if self._iter == 0:
            shifts -= self.beta / (1 + self.beta) * np.sum(shifts * los, axis=-1)[:, None] * los
        self._positions_rec_data = self._positions_data - self.f * np.sum(shifts * los, axis=-1)[:, None] * los
        self._iter += 1
        if return_psi:
            return psis

    def read_shifts(self, positions, field='disp+rsd'):
        field = field.lower()
        if field not in ['disp', 'rsd', 'disp+rsd']:
            raise OracleError('Unknown field {}'.format(field))

        def read_cic(positions, wrap=False):
            shifts = np.empty_like(positions)
            for iaxis, psi in enumerate(self.mesh_psi):
                shifts[:, iaxis] = self.read_cic_wrap(psi, positions) if wrap else self.read_cic_nowrap(psi, positions)
            return shifts

        if isinstance(positions, str) and positions == 'data':
            shifts = read_cic(self._positions_rec_data, wrap=True)
            if field == 'disp':
                return shifts
            rsd = self._positions_data - self._positions_rec_data
            if field == 'rsd':
                return rsd
            shifts += rsd
            return shifts
        if self.wrap: positions = self.wrap_positions(positions)
        shifts = read_cic(positions, wrap=False)
        if field == 'disp':
            return shifts
        if self.los is None:
            los = positions / distance(positions)[:, None]
        else:
            los = self.los
        rsd = self.f * np.sum(shifts * los, axis=-1)[:, None] * los
        if field == 'rsd':
            return rsd
        real_positions = positions - rsd
        shifts = read_cic(real_positions, wrap=True)
        return shifts + rsd

    def read_shifted_positions(self, positions, field='disp+rsd'):
        shifts = self.read_shifts(positions, field=field)
        if isinstance(positions, str) and positions == 'data':
            positions = self._positions_data
        positions = positions - shifts
        if self.wrap: positions = self.wrap_positions(positions)
        return positions


OracleIterativeFFTParticleReconstruction = OracleOriginalIterativeFFTParticleReconstruction
