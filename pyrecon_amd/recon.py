"""Base reconstruction class: the reference's BaseReconstruction API
(pyrecon/recon.py:15-338) over device-resident meshes.

Positions and weights may be numpy arrays (uploaded for the call) or device
arrays (:class:`~pyrecon_amd.device.DeviceArray`, or a tensor exposing
``data_ptr()``), in which case nothing crosses PCIe.
"""
import ctypes
import logging

import numpy as np

from . import _lib
from ._lib import lib, check, MeshError
from .device import DeviceArray, as_device, stream_handle
from .mesh import MeshInfo, RealMesh, Workspace


def numpy_mean(rtype, total, count):
    """np.mean's arithmetic (numpy/_core/_methods.py): the sum in the array's type divided by an
    intp count, cast back to the array's type."""
    return rtype(rtype(total) / np.intp(count))


class ReconstructionError(Exception):
    """Error raised when issue with reconstruction (pyrecon/recon.py:10-12)."""


def distance(position):
    """Cartesian distance along the last axis (pyrecon/utils.py:146-148)."""
    return np.sqrt((position**2).sum(axis=-1))


def _suffix(dtype):
    return 'f32' if np.dtype(dtype) == np.float32 else 'f64'


class BaseReconstruction(object):
    """Base template reconstruction class; see the reference for the full docs.

    Parameters (as the reference, pyrecon/recon.py:66)
    ----------
    f, bias : float
    los : None, 'local', 'x', 'y', 'z', int or 3-vector
    fft_engine, fft_wisdom, save_fft_wisdom, fft_plan : accepted for compatibility; the device
        always uses its own transforms.
    wrap : bool
    precision : 'plain' (default: the reference's algorithm in the mesh dtype, bit for bit where the
        reference is deterministic C) or 'mixed' (FP64 meshes: CIC weights summed in fixed point per tile
        and in FP64 across tiles, and the mixed-precision solve of ``run``).
    kwargs : MeshInfo arguments (nmesh, boxsize, boxcenter, cellsize, positions, boxpad, dtype);
        ``nthreads`` is accepted and ignored.  Device-specific: ``stream``, ``paint_chunk`` (points
        painted per pass; ``None`` grows the scratch so that each catalog is painted in one pass,
        see :class:`~pyrecon_amd.mesh.Workspace`), ``workspace`` (a
        :class:`~pyrecon_amd.mesh.Workspace` to share).
    """

    @staticmethod
    def _select_nmesh(nmesh):
        return nmesh.copy()

    def __init__(self, f=0., bias=1., los=None, fft_engine='numpy', fft_wisdom=None, save_fft_wisdom=None,
                 fft_plan='measure', wrap=False, stream=None, paint_chunk=None, workspace=None,
                 precision='plain', **kwargs):
        self.logger = logging.getLogger(self.__class__.__name__)
        self.set_cosmo(f=f, bias=bias)
        self.wrap = wrap
        self.stream = stream
        kwargs.pop('nthreads', None)
        self.info = MeshInfo(select_nmesh=self._select_nmesh, **kwargs)
        if precision not in ('plain', 'mixed'):
            raise ValueError('precision must be "plain" or "mixed"')
        if precision == 'mixed' and self.info.dtype != np.float64:
            raise ValueError('precision="mixed" needs an FP64 mesh (dtype="f8")')
        self.precision = precision
        if workspace is None:
            # on the default stream, the workspace of a released reconstruction of the same kind
            workspace = Workspace(self.info.shape, self.info.dtype, paint_chunk=paint_chunk, reuse=stream is None)
        elif tuple(workspace.nmesh) != self.info.shape or workspace.dtype != self.info.dtype:
            raise ValueError('workspace does not match the mesh')
        self.workspace = workspace
        self.mesh_data = RealMesh(info=self.info, workspace=workspace, stream=stream)
        self.mesh_randoms = RealMesh(info=self.info, workspace=workspace, stream=stream)
        self._size_randoms = 0
        self.set_los(los)
        self.logger.info('Using mesh {}.'.format(self.mesh_data))

    @property
    def beta(self):
        return self.f / self.bias

    def set_cosmo(self, f=None, bias=None, beta=None):
        if bias is None: bias = self.bias
        if beta is not None: f = beta * bias
        if f is None: f = self.f
        self.f = f
        self.bias = bias

    def set_los(self, los=None):
        if los is None or (isinstance(los, str) and los == 'local'):
            self.los = None
        else:
            if isinstance(los, str):
                los = 'xyz'.index(los)
            if np.ndim(los) == 0:
                ilos = los
                los = np.zeros(3, dtype=self.info.dtype)
                los[ilos] = 1.
            los = np.array(los, dtype=self.info.dtype)
            self.los = los / distance(los)

    def assign_data(self, positions, weights=None):
        self.mesh_data.assign_cic(positions, weights=weights, wrap=self.wrap, tile_sums=self.precision == 'mixed')

    def assign_randoms(self, positions, weights=None):
        if self.mesh_randoms.buffer is None:
            self._size_randoms = 0
        self.mesh_randoms.assign_cic(positions, weights=weights, wrap=self.wrap,
                                     tile_sums=self.precision == 'mixed')
        self._size_randoms += len(positions)

    @property
    def has_randoms(self):
        return self.mesh_randoms is not None and self.mesh_randoms.buffer is not None

    def _mesh_sums(self, a, b=None):
        sums = (ctypes.c_double * 2)()
        fn = getattr(lib, 'prh_mesh_sums_' + _suffix(self.info.dtype))
        check(fn(self.workspace.handle, a.ptr, b.ptr if b is not None else None, self.info.size, sums,
                 stream_handle(self.stream)), 'mesh sums')
        return sums[0], sums[1]

    def set_density_contrast(self, ran_min=0.01, smoothing_radius=15., check=False, **kwargs):
        """Smoothed density contrast (pyrecon/recon.py:198-249)."""
        dtype = self.info.dtype
        rtype = dtype.type
        data = self.mesh_data
        if data.buffer is None:
            raise ReconstructionError('assign data before setting the density contrast')
        if check and self.has_randoms:
            randoms_value = self.mesh_randoms.value
            mask_nonzero = randoms_value > 0.
            nnonzero = mask_nonzero.sum()
            if nnonzero < 2: raise ValueError('Very few randoms!')
        data.smooth_gaussian(smoothing_radius, **kwargs)
        fn = getattr(lib, 'prh_density_contrast_' + _suffix(dtype))
        if self.has_randoms:
            randoms = self.mesh_randoms
            randoms.smooth_gaussian(smoothing_radius, **kwargs)
            s_data, s_randoms = self._mesh_sums(data.buffer, randoms.buffer)
            # scalar algebra with numpy's types, as recon.py:227-243 evaluates it
            sum_data, sum_randoms = rtype(s_data), rtype(s_randoms)
            alpha = sum_data * 1. / sum_randoms
            threshold = ran_min * sum_randoms / self._size_randoms
            bias_alpha = self.bias * alpha
            if check:
                smoothed = randoms.value
                mean_nran = smoothed[mask_nonzero].mean()
                std_nran = smoothed[mask_nonzero].std(ddof=1)
                self.logger.info('Mean smoothed random density in non-empty cells is {:.4f} (std = {:.4f}), '
                                 'threshold is (ran_min * mean weight) = {:.4f}.'.format(mean_nran, std_nran, threshold))
                frac = 1. - (smoothed > threshold).sum() / nnonzero
                if frac > 0.1:
                    self.logger.warning('Masking a large fraction {:.4f} of non-empty cells. You should probably '
                                        'increase the number of randoms.'.format(frac))
                else:
                    self.logger.info('Masking a fraction {:.4f} of non-empty cells.'.format(frac))
            check_status = fn(data.buffer.ptr, randoms.buffer.ptr, data.buffer.ptr, self.info.size, float(alpha),
                              float(threshold), float(bias_alpha), float(rtype(self.bias)), stream_handle(self.stream))
            self.alpha, self.threshold = alpha, threshold
        else:
            s_data, _ = self._mesh_sums(data.buffer)
            mean = numpy_mean(rtype, s_data, self.info.size)
            check_status = fn(data.buffer.ptr, None, data.buffer.ptr, self.info.size, float(mean), 0., 0.,
                              float(rtype(self.bias)), stream_handle(self.stream))
        _lib.check(check_status, 'density contrast')
        self.mesh_delta = data
        self.mesh_data = None
        self.mesh_randoms = None

    def run(self, *args, **kwargs):
        raise NotImplementedError('Implement method "run" in your "{}"-inherited algorithm'.format(self.__class__.__name__))

    # ------------------------------------------------------------------ shifts
    def _shift_out_dtype(self, pos_dtype, field, shifted):
        """Output dtype of read_shifts / read_shifted_positions under numpy's promotion."""
        pos_dtype = np.dtype(pos_dtype)
        out = pos_dtype
        if field == 'rsd' and self.los is not None:
            out = np.result_type(pos_dtype, self.info.dtype)
        if shifted and self.wrap:
            out = np.result_type(out, self.info.dtype)
        return np.dtype(out)

    def _read(self, positions, field, shifted, out=None):
        field = field.lower()
        allowed_fields = ['disp', 'rsd', 'disp+rsd']
        if field not in allowed_fields:
            raise ReconstructionError('Unknown field {}. Choices are {}'.format(field, allowed_fields))
        if getattr(self, 'mesh_phi', None) is None or self.mesh_phi.buffer is None:
            raise ReconstructionError('run the reconstruction before reading shifts')
        pos, is_host = as_device(positions, stream=self.stream)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError('positions must have shape (N, 3)')
        pdtype = np.dtype(pos.dtype)
        if pdtype not in (np.dtype('f4'), np.dtype('f8')):
            pos, is_host = as_device(np.asarray(positions, dtype='f8'), stream=self.stream)
            pdtype = np.dtype('f8')
        odtype = self._shift_out_dtype(pdtype, field, shifted)
        if out is None:
            out = DeviceArray((len(pos), 3), odtype, stream=self.stream)
        elif not (isinstance(out, DeviceArray) and out.shape == (len(pos), 3) and out.dtype == odtype):
            raise ValueError('out must be a DeviceArray of shape ({:d}, 3) and dtype {}'.format(len(pos), odtype))
        code = {'disp': _lib.FIELD_DISP, 'rsd': _lib.FIELD_RSD, 'disp+rsd': _lib.FIELD_DISP_RSD}[field]
        dtype = self.info.dtype
        nm, nm_ptr = _lib.size3(self.info.shape)
        bs, bs_ptr = _lib.vec3(self.info.boxsize, dtype)
        bc, bc_ptr = _lib.vec3(self.info.boxcenter, dtype)
        los_ptr = None
        if self.los is not None:
            los_arr, los_ptr = _lib.vec3(self.los, dtype)
        nonfinite = ctypes.c_int(0)
        fn = getattr(lib, 'prh_read_shifts_' + _suffix(dtype))
        check(fn(self.workspace.handle, self.mesh_phi.buffer.ptr, nm_ptr, bs_ptr, bc_ptr, pos.ptr,
                 _lib.dtype_code(pdtype), len(pos), float(self.f), los_ptr, code, int(bool(self.wrap)),
                 int(bool(shifted)), out.ptr, _lib.dtype_code(odtype), ctypes.byref(nonfinite),
                 stream_handle(self.stream)), 'read_shifts')
        if nonfinite.value:
            import warnings
            warnings.warn('Some positions are NaN/inf')
        return out.numpy(stream=self.stream) if is_host else out

    def read_shifts(self, positions, field='disp+rsd', out=None):
        """Displacements at ``positions`` (multigrid.py:137-158); numpy in -> numpy out, device in -> device
        out (written into ``out`` if given)."""
        return self._read(positions, field, shifted=False, out=out)

    def read_shifted_positions(self, positions, field='disp+rsd', out=None):
        """``positions - read_shifts(positions)``, wrapped if :attr:`wrap` (recon.py:304-325)."""
        return self._read(positions, field, shifted=True, out=out)


def _make_property(name):

    @property
    def func(self):
        return getattr(self.info, name)

    return func


for _name in ['boxsize', 'boxcenter', 'nmesh', 'offset', 'cellsize']:
    setattr(BaseReconstruction, _name, _make_property(_name))


class SpectralReconstruction(BaseReconstruction):
    """Device parts shared by the FFT reconstructions: 3-D transforms of the mesh, Fourier-space
    operators, and the base-class shifts, read by CIC from the three displacement meshes
    :attr:`mesh_psi` (pyrecon/recon.py:255-325)."""

    def _complex(self):
        """An interleaved complex mesh (n0, n1, n2, 2) in the mesh precision."""
        return DeviceArray(tuple(self.info.shape) + (2,), self.info.dtype, stream=self.stream)

    def _fft(self, src, dst, inverse, io):
        nm, nm_ptr = _lib.size3(self.info.shape)
        fn = getattr(lib, 'prh_fft3d_' + _suffix(self.info.dtype))
        check(fn(self.workspace.handle, src.ptr, dst.ptr, nm_ptr, int(inverse), int(io), stream_handle(self.stream)),
              'fft3d')

    def _spectrum(self, mesh):
        """numpy.fft.fftn of a real mesh (RealMesh.to_complex, pyrecon/mesh.py:697-705)."""
        out = self._complex()
        self._fft(mesh.buffer, out, False, _lib.FFT_REAL_IN)
        return out

    def _spectral(self, op, src, dst, a=0, b=0, norm=1., los=None, beta=0.):
        nm, nm_ptr = _lib.size3(self.info.shape)
        bs, bs_ptr = _lib.vec3(self.info.boxsize, self.info.dtype)
        los_arr = None if los is None else (ctypes.c_double * 3)(*[float(x) for x in los])
        fn = getattr(lib, 'prh_spectral_' + _suffix(self.info.dtype))
        check(fn(self.workspace.handle, int(op), src.ptr, dst.ptr, nm_ptr, bs_ptr, los_arr, float(beta), float(norm),
                 int(a), int(b), stream_handle(self.stream)), 'spectral operator')

    def _new_mesh(self):
        mesh = RealMesh(info=self.info, workspace=self.workspace, stream=self.stream)
        mesh._allocate(zero=False)
        return mesh

    def _to_real(self, spectrum, out=None):
        """Real part of numpy.fft.ifftn times N (ComplexMesh.to_real, pyrecon/mesh.py:827-844; the 1 / N is
        applied by the Fourier-space operator that made ``spectrum``); ``spectrum`` is overwritten."""
        if out is None: out = self._new_mesh()
        self._fft(spectrum, out.buffer, True, _lib.FFT_REAL_OUT)
        return out

    _FIELDS = {'disp': _lib.FIELD_DISP, 'rsd': _lib.FIELD_RSD, 'disp+rsd': _lib.FIELD_DISP_RSD}

    def _check_field(self, field):
        field = field.lower()
        if field not in self._FIELDS:
            raise ReconstructionError('Unknown field {}. Choices are {}'.format(field, list(self._FIELDS)))
        if getattr(self, 'mesh_psi', None) is None:
            raise ReconstructionError('run the reconstruction before reading shifts')
        return field

    def _device_shifts(self, positions, field, shifted):
        """Shifts (or, ``shifted``, shifted positions) without wrap, on the device: the three CIC reads
        of :attr:`mesh_psi`, then the RSD algebra of pyrecon/recon.py:255-325 in numpy's dtypes and
        order (``prh_shift_algebra_*``).  Host in -> host out, device in -> device out."""
        pos, is_host = as_device(positions, stream=self.stream)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError('positions must have shape (N, 3)')
        if pos.dtype not in (np.dtype('f4'), np.dtype('f8')):
            pos, is_host = as_device(np.asarray(positions, dtype='f8'), stream=self.stream)
        comps = [psi.read_cic(pos, wrap=False) for psi in self.mesh_psi]
        los = None if self.los is None else np.ascontiguousarray(self.los, dtype=self.info.dtype)
        odt = pos.dtype
        if field == 'rsd' and los is not None: odt = np.promote_types(pos.dtype, self.info.dtype)
        out = DeviceArray((len(pos), 3), odt, stream=self.stream)
        fn = getattr(lib, 'prh_shift_algebra_' + _suffix(self.info.dtype))
        check(fn(pos.ptr, _lib.dtype_code(pos.dtype), comps[0].ptr, comps[1].ptr, comps[2].ptr, len(pos),
                 float(self.f), los.ctypes.data if los is not None else None, self._FIELDS[field], int(bool(shifted)),
                 out.ptr, _lib.dtype_code(odt), stream_handle(self.stream)), 'shift algebra')
        return out.numpy(stream=self.stream) if is_host else out

    def read_shifts(self, positions, field='disp+rsd'):
        """Displacements at ``positions`` (pyrecon/recon.py:255-302): each component read by CIC from
        :attr:`mesh_psi` on the device.  Without :attr:`wrap` the RSD algebra runs on the device
        (host in -> host out, device in -> device out); with it, positions are wrapped first and the
        algebra is the reference's numpy on the host (host arrays out)."""
        field = self._check_field(field)
        if not self.wrap: return self._device_shifts(positions, field, False)
        return self._host_shifts(positions, field)

    def _host_shifts(self, positions, field):
        """The shifts with the reference's numpy RSD algebra on the host, positions wrapped first if
        :attr:`wrap` (host arrays out)."""
        if isinstance(positions, DeviceArray): positions = positions.numpy(stream=self.stream)
        positions = np.asarray(positions)
        shifts = np.empty_like(positions)
        if self.wrap: positions = self.info.wrap(positions)
        pos, _ = as_device(positions, stream=self.stream)
        for iaxis, psi in enumerate(self.mesh_psi):
            shifts[:, iaxis] = psi.read_cic(pos, wrap=False).numpy(stream=self.stream)
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
        """``positions - read_shifts(positions)``, wrapped if :attr:`wrap` (pyrecon/recon.py:304-325)."""
        field = self._check_field(field)
        if not self.wrap: return self._device_shifts(positions, field, True)
        if isinstance(positions, DeviceArray): positions = positions.numpy(stream=self.stream)
        shifts = self._host_shifts(positions, field)
        positions = positions - shifts
        if self.wrap: positions = self.info.wrap(positions)
        return positions
