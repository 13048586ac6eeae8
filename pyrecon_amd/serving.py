"""Serving many reconstructions on one device: continuous batching of uneven jobs.

The reference reconstructs one catalog per process call.  A :class:`ReconServer` takes
reconstruction jobs as they arrive and keeps several on the device at once:

- **Admission by device footprint.**  A job is admitted (FIFO) when its bytes (workspace
  for its mesh, the four meshes, device catalogs and outputs) fit in the remaining budget
  and a slot is free; memory returns to the budget when a job ends.  A job larger than the
  whole budget fails alone with :class:`AdmissionError`.
- **Overlap.**  Every admitted job runs on its own HIP stream, driven by its own host
  thread (the library's calls release the GIL), so painting, smoothing and reading of
  different jobs overlap on the device.
- **Grouped solves.**  A job whose density contrast is ready hands its solve to the solver
  thread.  Jobs waiting with the same mesh shape, dtype and solver settings are solved
  together (:func:`pyrecon_amd.multigrid.run_group`): one launch relaxes the coarse levels
  of all of them.  One group per mesh shape is in flight at a time, so jobs that become
  ready meanwhile gather into the next group.
- **Determinism.**  Every kernel of a job depends on that job's data only, so a job's
  result is bit-identical alone and in any mix.

Workspaces (hierarchies, paint scratch, FFT plans) are pooled per mesh shape and reused
across jobs; idle ones count against the budget and are released when room is needed.
"""
import collections
import ctypes
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _lib
from ._lib import lib, check
from .device import DeviceArray, Event, Stream, as_device, stream_handle
from .mesh import Workspace
from .multigrid import MultiGridReconstruction, OriginalMultiGridReconstruction, run_group, MAX_GROUP


class AdmissionError(RuntimeError):
    """A job cannot be admitted (larger than the whole device budget)."""


def synthetic_catalog(size, boxsize, boxcenter, seed, modes=None, f=0., weights=(0.5, 1.5), dtype='f8',
                      stream=None):
    """Generate ``(positions, weights)`` on the device from a seed (see ``prh_synth_points``).

    ``modes`` is None for uniform points, or an (M, 6) array of Fourier modes
    (kx, ky, kz, amplitude / |k|, phase, 0) for Zel'dovich-displaced points, with ``f`` the
    growth rate of the line-of-sight shift.
    """
    pos = DeviceArray((int(size), 3), dtype, stream=stream)
    w = DeviceArray((int(size),), dtype, stream=stream)
    bs = np.empty(3, dtype='f8')
    bs[:] = boxsize
    bc = np.empty(3, dtype='f8')
    bc[:] = boxcenter
    dmodes, nmodes = None, 0
    if modes is not None and len(modes):
        dmodes = DeviceArray.from_numpy(np.ascontiguousarray(modes, dtype='f8'), stream=stream)
        nmodes = len(modes)
    check(lib.prh_synth_points(pos.ptr, w.ptr, _lib.dtype_code(dtype), int(size), int(seed) & ((1 << 64) - 1),
                               bs.ctypes.data, bc.ctypes.data, dmodes.ptr if dmodes is not None else None, nmodes,
                               float(f), float(weights[0]), float(weights[1]), stream_handle(stream)), 'synth')
    if dmodes is not None:
        check(lib.prh_synchronize(stream_handle(stream)), 'synth')
    return pos, w


def zeldovich_modes(boxsize, nmesh, seed, nmodes=256, amplitude=2e4, kpivot=0.1, slope=-1.5, kcut=0.3):
    """Random Fourier modes of the displacement of a Gaussian field with
    P(k) = amplitude (k / kpivot)^slope exp(-(k / kcut)^2), for :func:`synthetic_catalog`.

    Kraichnan's randomization: |k| is drawn with density proportional to P(k) between the
    fundamental and the Nyquist wavenumber of the mesh, the direction isotropically, the
    phase uniformly; every mode carries the same share of the displacement variance
    sigma^2 = int P(k) dk / (2 pi^2), so the field has that spectrum on average.
    """
    rng = np.random.RandomState(seed)
    kmin = 2. * np.pi / float(np.min(boxsize))
    kmax = np.pi * float(np.max(nmesh)) / float(np.max(boxsize))
    grid = np.linspace(kmin, kmax, 4096)
    pk = amplitude * (grid / kpivot)**slope * np.exp(-(grid / kcut)**2)
    cdf = np.concatenate([[0.], np.cumsum(0.5 * (pk[1:] + pk[:-1]) * np.diff(grid))])
    sigma2 = cdf[-1] / (2. * np.pi**2)
    k = np.interp(rng.uniform(0., cdf[-1], nmodes), cdf, grid)
    mu = rng.uniform(-1., 1., nmodes)
    phi = rng.uniform(0., 2. * np.pi, nmodes)
    s = np.sqrt(1. - mu**2)
    kvec = k[:, None] * np.stack([s * np.cos(phi), s * np.sin(phi), mu], axis=-1)
    amp = np.sqrt(2. * sigma2 / nmodes) / k
    modes = np.zeros((nmodes, 6))
    modes[:, :3] = kvec
    modes[:, 3] = amp
    modes[:, 4] = rng.uniform(0., 2. * np.pi, nmodes)
    return modes


class ReconJob(object):
    """One reconstruction request.

    Parameters
    ----------
    data, randoms : tuple (positions, weights), or callable ``fn(stream) -> (positions, weights)``
        Catalogs as numpy arrays or device arrays, or a function producing device arrays on
        the job's stream once the job is admitted (``randoms`` may be None).
    ndata, nrandoms : int, optional
        Catalog sizes, needed for the footprint when the catalogs are callables.
    nmesh, boxsize, boxcenter, f, bias, los, dtype, wrap : as :class:`MultiGridReconstruction`.
    smoothing_radius : float
    precision : 'plain' or 'mixed' (see :meth:`OriginalMultiGridReconstruction.run`).
    shifted : tuple of 'data', 'randoms'
        Which catalogs' shifted positions to return.
    field, randoms_field : 'disp', 'rsd' or 'disp+rsd'
        Fields of the data and randoms shifts; the defaults are the RecIso convention of
        ``bin/pyrecon:348-361`` (RecSym moves the randoms by 'disp+rsd').
    original : bool
        Use :class:`OriginalMultiGridReconstruction` (Martin White's density rule).
    to_host : bool
        Return numpy arrays (True) or device arrays (False, default).
    """

    def __init__(self, data, randoms=None, ndata=None, nrandoms=None, nmesh=None, boxsize=None, boxcenter=None,
                 f=0., bias=1., los=None, dtype='f8', wrap=False, smoothing_radius=15., precision='plain',
                 shifted=('data',), field='disp+rsd', randoms_field='disp', original=False, to_host=False, run_kwargs=None, name=None):
        self.data, self.randoms = data, randoms
        self.ndata = ndata if ndata is not None else (len(data[0]) if not callable(data) else None)
        if randoms is None:
            self.nrandoms = 0
        else:
            self.nrandoms = nrandoms if nrandoms is not None else (len(randoms[0]) if not callable(randoms) else None)
        if self.ndata is None or self.nrandoms is None:
            raise ValueError('give ndata / nrandoms for catalogs produced by callables')
        self.nmesh = tuple(int(n) for n in np.broadcast_to(nmesh, 3))
        self.boxsize, self.boxcenter = boxsize, boxcenter
        self.f, self.bias, self.los, self.dtype, self.wrap = f, bias, los, np.dtype(dtype), wrap
        self.smoothing_radius, self.precision = smoothing_radius, precision
        self.field, self.randoms_field = field, randoms_field
        self.shifted, self.original, self.to_host = tuple(shifted), original, to_host
        self.run_kwargs = dict(run_kwargs or {})
        self.name = name

    def footprint(self, paint_chunk):
        """Device bytes of the job: ``(workspace, meshes + catalogs + outputs)``."""
        W = self.dtype.itemsize
        cells = int(np.prod(self.nmesh))
        ws = Workspace.footprint(self.nmesh, self.dtype, paint_chunk=paint_chunk, flags=self.flags())
        meshes = 3 * cells * W  # data (then delta), randoms, phi
        points = self.ndata + self.nrandoms
        catalogs = points * (3 * 8 + 8)
        outputs = sum(n * 3 * 8 for which, n in (('data', self.ndata), ('randoms', self.nrandoms))
                      if which in self.shifted)
        return ws, meshes + catalogs + outputs

    def flags(self):
        return _lib.WS_ALL if self.precision == 'mixed' else (_lib.WS_ALL & ~_lib.WS_MIXED)

    def solve_key(self):
        kw = tuple(sorted(self.run_kwargs.items()))
        return (self.nmesh, self.dtype.str, self.precision, kw)


class JobResult(object):
    """Outputs and timeline of a finished job (times in seconds, ``time.perf_counter``)."""

    def __init__(self, job):
        self.job = job
        self.shifted_data = None
        self.shifted_randoms = None
        self.times = {}
        self.group_size = 1

    @property
    def latency(self):
        return self.times['done'] - self.times['submit']


class _Active(object):
    def __init__(self, job, future, ws_bytes, other_bytes, workspace, stream):
        self.job, self.future = job, future
        self.ws_bytes, self.other_bytes = ws_bytes, other_bytes
        self.workspace, self.stream = workspace, stream
        self.result = JobResult(job)
        self.result.times['submit'] = future._submit_time


class ReconServer(object):
    """Continuous-batching reconstruction server on the current device.

    Parameters
    ----------
    budget_bytes : int
        Device bytes the jobs may hold together (workspaces included).
    slots : int
        Jobs on the device at once (one stream and host thread each).
    paint_chunk : int
        Points painted per pass (sizes the paint scratch of the pooled workspaces).
    group_solves : bool
        Solve jobs of the same mesh shape together (one launch per coarse-level operation).
    """

    def __init__(self, budget_bytes=200 * 10**9, slots=4, paint_chunk=1 << 23, group_solves=True, max_group=MAX_GROUP):
        self.budget = int(budget_bytes)
        self.slots = int(slots)
        self.paint_chunk = int(paint_chunk)
        self.group_solves = bool(group_solves)
        self.max_group = int(min(max_group, MAX_GROUP))
        self._cv = threading.Condition()
        self._pending = collections.deque()
        self._active = 0
        self._used = 0
        self._idle_ws = collections.OrderedDict()  # id -> (key, workspace, bytes), oldest first
        self._streams = [Stream() for _ in range(self.slots)]
        self._free_streams = list(self._streams)
        self._solve_queue = []
        self._solve_busy = set()
        self._solve_streams = {}
        self._closed = False
        self.results = []
        self._threads = [threading.Thread(target=self._admit_loop, name='recon-admit', daemon=True)]
        if self.group_solves:
            self._threads.append(threading.Thread(target=self._solve_loop, name='recon-solve', daemon=True))
        for t in self._threads: t.start()

    # ------------------------------------------------------------------ client API
    def submit(self, job):
        """Queue a job; returns a :class:`concurrent.futures.Future` of its :class:`JobResult`."""
        fut = Future()
        fut._submit_time = time.perf_counter()
        with self._cv:
            if self._closed:
                raise RuntimeError('server is closed')
            self._pending.append((job, fut))
            self._cv.notify_all()
        return fut

    def close(self, wait=True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            for t in self._threads: t.join()
        with self._cv:
            for _, (key, ws, nbytes) in self._idle_ws.items():
                ws.close()
            self._idle_ws.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def report(self, results=None):
        """Latency percentiles (s) and throughput of the finished jobs."""
        results = self.results if results is None else results
        if not results: return {}
        lat = np.array([r.latency for r in results])
        t0 = min(r.times['submit'] for r in results)
        t1 = max(r.times['done'] for r in results)
        return {'jobs': len(results), 'p50_s': float(np.percentile(lat, 50)), 'p95_s': float(np.percentile(lat, 95)),
                'p99_s': float(np.percentile(lat, 99)), 'mean_s': float(lat.mean()),
                'jobs_per_s': len(results) / (t1 - t0) if t1 > t0 else float('nan'),
                'mean_group': float(np.mean([r.group_size for r in results]))}

    # ------------------------------------------------------------------ admission
    # Accounting (under the lock): `_used` holds the bytes of the admitted jobs, their
    # workspaces included; `_idle_ws` the pooled workspaces of finished jobs.  Both count
    # against the budget; idle workspaces are destroyed, oldest first, to make room.
    def _admit_loop(self):
        while True:
            with self._cv:
                while True:
                    if self._closed and not self._pending and self._active == 0:
                        return
                    act = self._try_admit() if (self._pending and self._active < self.slots) else None
                    if act is not None:
                        break
                    self._cv.wait()
            act.result.times['admit'] = time.perf_counter()
            threading.Thread(target=self._run_job, args=(act,), daemon=True).start()

    def _try_admit(self):
        while self._pending:
            job, fut = self._pending[0]
            try:
                ws_bytes, other = job.footprint(self.paint_chunk)
            except Exception as exc:
                self._pending.popleft()
                fut.set_exception(exc)
                continue
            if ws_bytes + other > self.budget:
                self._pending.popleft()
                fut.set_exception(AdmissionError('job needs {:d} bytes, the budget is {:d}'.format(
                    ws_bytes + other, self.budget)))
                continue
            key = self._ws_key(job)
            ws = None
            for i, (k, w, nbytes) in self._idle_ws.items():
                if k == key:
                    ws, ws_bytes = w, nbytes
                    del self._idle_ws[i]
                    break
            need = other + ws_bytes
            while self._used + self._idle_bytes() + need > self.budget and self._idle_ws:
                _, (_, victim, _) = self._idle_ws.popitem(last=False)
                victim.close()
            if self._used + self._idle_bytes() + need > self.budget:
                if ws is not None:
                    self._idle_ws[id(ws)] = (key, ws, ws_bytes)
                return None  # FIFO: wait for running jobs to free memory
            self._pending.popleft()
            self._active += 1
            self._used += need
            return _Active(job, fut, ws_bytes, other, ws, self._free_streams.pop())
        return None

    def _idle_bytes(self):
        return sum(b for _, _, b in self._idle_ws.values())

    def _ws_key(self, job):
        return (job.nmesh, job.dtype.str, job.flags())

    def _release(self, act):
        with self._cv:
            self._used -= act.other_bytes + act.ws_bytes
            if act.workspace is not None and not self._closed:
                self._idle_ws[id(act.workspace)] = (self._ws_key(act.job), act.workspace, act.ws_bytes)
            elif act.workspace is not None:
                act.workspace.close()
            self._active -= 1
            self._free_streams.append(act.stream)
            self._cv.notify_all()

    # ------------------------------------------------------------------ one job
    def _run_job(self, act):
        job, res, stream = act.job, act.result, act.stream
        try:
            if act.workspace is None:
                act.workspace = Workspace(job.nmesh, job.dtype, paint_chunk=self.paint_chunk, flags=job.flags())
            res.times['start'] = time.perf_counter()
            cls = OriginalMultiGridReconstruction if job.original else MultiGridReconstruction
            recon = cls(f=job.f, bias=job.bias, los=job.los, nmesh=job.nmesh, boxsize=job.boxsize,
                        boxcenter=job.boxcenter, dtype=job.dtype, wrap=job.wrap, workspace=act.workspace,
                        stream=stream, precision=job.precision)
            data_pos, data_w = self._catalog(job.data, stream)
            recon.assign_data(data_pos, data_w)
            rand_pos = None
            if job.randoms is not None:
                rand_pos, rand_w = self._catalog(job.randoms, stream)
                recon.assign_randoms(rand_pos, rand_w)
                del rand_w
            del data_w
            recon.set_density_contrast(smoothing_radius=job.smoothing_radius)
            res.times['density'] = time.perf_counter()
            if self.group_solves:
                res.group_size = self._solve_grouped(recon, job, stream)
            else:
                recon.run(precision=job.precision, **job.run_kwargs)
            res.times['solved'] = time.perf_counter()
            if 'data' in job.shifted:
                out = recon.read_shifted_positions(data_pos, field=job.field)
                res.shifted_data = out.numpy(stream=stream) if (job.to_host and isinstance(out, DeviceArray)) else out
            if 'randoms' in job.shifted and rand_pos is not None:
                out = recon.read_shifted_positions(rand_pos, field=job.randoms_field)
                res.shifted_randoms = out.numpy(stream=stream) if (job.to_host and isinstance(out, DeviceArray)) else out
            check(lib.prh_synchronize(stream_handle(stream)), 'job')
            del recon, data_pos, rand_pos
            res.times['done'] = time.perf_counter()
            with self._cv:
                self.results.append(res)
            act.future.set_result(res)
        except BaseException as exc:  # a failed job must not take the server down
            lib.prh_synchronize(stream_handle(stream))
            res.times['done'] = time.perf_counter()
            act.future.set_exception(exc)
        finally:
            self._release(act)

    @staticmethod
    def _catalog(spec, stream):
        if callable(spec):
            pos, w = spec(stream)
        else:
            pos, w = spec
            pos, _ = as_device(pos, stream=stream)
            if w is not None:
                w, _ = as_device(w, stream=stream)
        return pos, w

    # ------------------------------------------------------------------ grouped solves
    def _solve_grouped(self, recon, job, stream):
        ready = Event()
        ready.record(stream)
        req = {'recon': recon, 'job': job, 'ready': ready, 'launched': threading.Event(), 'done': None,
               'error': None, 'count': 1, 'stream': stream}
        with self._cv:
            self._solve_queue.append(req)
            self._cv.notify_all()
        req['launched'].wait()
        if req['error'] is not None:
            raise req['error']
        req['done'].wait(stream)  # later work of this job waits for the group's solve
        recon.mesh_delta = None
        return req['count']

    def _solve_loop(self):
        while True:
            with self._cv:
                while True:
                    if self._closed and not self._solve_queue and self._active == 0:
                        return
                    batch = self._next_group()
                    if batch is not None:
                        key = batch[0]['job'].solve_key()
                        self._solve_busy.add(key)
                        break
                    self._cv.wait()
            self._launch_group(key, batch)

    def _next_group(self):
        # with the lock held: the oldest waiting request whose key has no group in flight,
        # with every other waiting request of that key (up to max_group)
        for i, req in enumerate(self._solve_queue):
            key = req['job'].solve_key()
            if key in self._solve_busy:
                continue
            batch = [r for r in self._solve_queue if r['job'].solve_key() == key][:self.max_group]
            for r in batch:
                self._solve_queue.remove(r)
            return batch
        return None

    def _launch_group(self, key, batch):
        stream = self._solve_streams.get(key)
        if stream is None:
            stream = self._solve_streams[key] = Stream()
        done = Event()
        try:
            for r in batch:
                r['ready'].wait(stream)
            job = batch[0]['job']
            run_group([r['recon'] for r in batch], precision=job.precision, stream=stream, **job.run_kwargs)
            done.record(stream)
        except BaseException as exc:
            stream.synchronize()  # nothing of a failed group may still run when its jobs release memory
            for r in batch:
                r['error'] = exc
        for r in batch:
            r['done'] = done
            r['count'] = len(batch)
            r['launched'].set()

        def finish():
            try:
                done.synchronize()
            except Exception:
                pass
            with self._cv:
                self._solve_busy.discard(key)
                self._cv.notify_all()

        threading.Thread(target=finish, daemon=True).start()
