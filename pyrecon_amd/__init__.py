"""pyrecon_amd: MI355X-native MultiGridReconstruction (drop-in for pyrecon's hot path), and the
FFT reconstructions (plane-parallel, iterative, iterative particle) on the same kernels.

Usage is the reference's::

    from pyrecon_amd import MultiGridReconstruction
    recon = MultiGridReconstruction(f=0.8, bias=2.0, los=None, nmesh=512, boxsize=1000., boxcenter=2000.)
    recon.assign_data(data_positions, data_weights)
    recon.assign_randoms(randoms_positions, randoms_weights)
    recon.set_density_contrast(smoothing_radius=15.)
    recon.run()
    data_positions_rec = recon.read_shifted_positions(data_positions)

Importing the package loads ``libpyrecon_amd.so`` (HIP, gfx950); it raises
ImportError if the library has not been built.  There is no CPU fallback.
"""
__version__ = '0.1.0'

from ._lib import MeshError, DeviceError
from .device import DeviceArray, Stream, Event, synchronize, empty_cache
from .mesh import MeshInfo, RealMesh, Workspace
from .recon import BaseReconstruction, ReconstructionError
from .multigrid import MultiGridReconstruction, OriginalMultiGridReconstruction, run_group
from .plane_parallel_fft import PlaneParallelFFTReconstruction
from .iterative_fft import IterativeFFTReconstruction
from .iterative_fft_particle import OriginalIterativeFFTParticleReconstruction, IterativeFFTParticleReconstruction
from .serving import ReconServer, ReconJob, AdmissionError

__all__ = ['MultiGridReconstruction', 'OriginalMultiGridReconstruction', 'PlaneParallelFFTReconstruction',
           'IterativeFFTReconstruction', 'OriginalIterativeFFTParticleReconstruction',
           'IterativeFFTParticleReconstruction', 'BaseReconstruction', 'ReconstructionError',
           'MeshError', 'DeviceError', 'MeshInfo', 'RealMesh', 'Workspace', 'DeviceArray', 'Stream', 'Event',
           'synchronize', 'empty_cache', 'run_group', 'ReconServer', 'ReconJob', 'AdmissionError']
