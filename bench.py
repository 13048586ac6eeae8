#!/usr/bin/env python
"""Benchmark of MultiGridReconstruction on one MI355X.

A step is one full reconstruction of the nmesh=512 survey-like workload
(configuration 3: boxsize 4000 Mpc/h, boxcenter (0, 0, 3000), 2e7 data and
1e8 randoms in the shell 1000 < r < 3000 Mpc/h with FKP weights, local line of
sight, f = 0.8, bias = 2, smoothing radius 15, reference defaults for the
solver): paint data and randoms, smooth, density contrast, solve, and the
shifted positions of data and randoms (RecSym).  Catalogs are generated once
(seeded, synthetic) and uploaded before the timed region, so inputs are
resident in HBM.  ``value`` is recon/s of the mixed-precision path (FP32
V-cycles inside FP64 defect correction, FP64 meshes); the plain FP64 path is
timed in the same run, and the mixed result is compared with it (the plain
FP64 solve is bit-identical to the reference's C multigrid).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))

METRIC = 'recon/s at nmesh=512 (FP32+FP64-refined vs FP64, 1 GPU); p50/p99 job latency'
HBM_PEAK_GBS = 8000.


def log(*args):
    print('[bench {:7.1f}s]'.format(time.time() - T0), *args, file=sys.stderr, flush=True)


T0 = time.time()


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=10)
    p.add_argument('--warmup', type=int, default=2)
    p.add_argument('--config', default='c3')
    p.add_argument('--scale', type=float, default=1.0, help='fraction of the catalog sizes (tests only)')
    p.add_argument('--fp64-steps', type=int, default=None, help='timed steps of the plain FP64 path')
    p.add_argument('--nrefine', type=int, default=0, help='extra FP64 corrections of the mixed path')
    p.add_argument('--no-cpu-baseline', action='store_true')
    p.add_argument('--no-accuracy', action='store_true')
    p.add_argument('--kernel-only', default=None, choices=['jacobi32', 'jacobi64', 'copy', 'fmg_mixed', 'fmg64'],
                   help='run only one kernel repeatedly (for rocprofv3 counter runs)')
    p.add_argument('--workers', type=int, default=len(os.sched_getaffinity(0)),
                   help='host threads (catalog generation, CPU baseline); default: the cores this process may use')
    p.add_argument('--cpu-baseline-full', action='store_true',
                   help='time one full-size CPU reconstruction (minutes) and print it instead of the GPU bench')
    p.add_argument('--serving-jobs', type=int, default=256, help='jobs of the serving stream (0: skip it)')
    p.add_argument('--serving-rate', type=float, default=4.0, help='Poisson arrival rate of the serving stream (jobs/s)')
    p.add_argument('--serving-only', action='store_true', help='run only the serving stream')
    p.add_argument('--host-io-steps', type=int, default=2,
                   help='timed mixed steps with numpy inputs and outputs (PCIe-inclusive rate; 0: skip)')
    p.add_argument('--fft-recon', action='store_true',
                   help='the FFT reconstructions (iterative, plane-parallel, iterative particle) on the workload')
    p.add_argument('--catalog-cache', default=None,
                   help='directory where the generated catalogs are saved and, on later runs, loaded from')
    return p.parse_args()


def serving_bench(args, precision, stream_spec, saturate=False, group=True, keep=False, reference=None):
    """Run the serving stream through one ReconServer.  ``saturate``: every job submitted at once
    (instead of at its Poisson arrival time).  ``keep``: return every job's shifted data positions
    (host); with ``reference`` (such a dict of another run), the per-job errors against it."""
    from pyrecon_amd.serving import ReconServer
    from catalogs import serving_job
    kept = {}

    def done(i, fut):
        res = fut.result()
        if keep or reference is not None:
            x = res.shifted_data.numpy()
            if reference is not None:
                d = np.abs(x - reference[i])
                kept[i] = (float(d.max()), float(np.sum(d * d)), d.size)
            else:
                kept[i] = x
        res.shifted_data = res.shifted_randoms = None

    server = ReconServer(budget_bytes=200 * 10**9, slots=4, paint_chunk=1 << 23, group_solves=group)
    futures = []
    t0 = time.perf_counter()
    for i, spec in enumerate(stream_spec):
        if not saturate:
            delay = t0 + spec['arrival'] - time.perf_counter()
            if delay > 0: time.sleep(delay)
        fut = server.submit(serving_job(spec, precision, name=str(i)))
        fut.add_done_callback(lambda f, i=i: done(i, f))
        futures.append(fut)
        if (i + 1) % 64 == 0:
            log('serving {}: submitted {:d} jobs'.format(precision, i + 1))
    for fut in futures: fut.result()
    report = server.report()
    per_mesh = {}
    for res in server.results:
        per_mesh.setdefault(res.job.nmesh[0], []).append(res.latency)
    report['p50_s_by_nmesh'] = {str(k): round(float(np.percentile(v, 50)), 4) for k, v in sorted(per_mesh.items())}
    report['max_group'] = int(max(r.group_size for r in server.results))
    server.close()
    if reference is not None:
        worst = max(kept, key=lambda i: kept[i][0])
        report['shift_max_mpc_h'] = kept[worst][0]
        report['shift_rms_mpc_h'] = float(np.sqrt(sum(v[1] for v in kept.values()) / sum(v[2] for v in kept.values())))
        report['error_jobs'] = len(kept)
        report['worst_job'] = dict(index=worst, nmesh=stream_spec[worst]['nmesh'], ndata=stream_spec[worst]['ndata'],
                                   shift_max_mpc_h=kept[worst][0])
        return report, None
    return report, kept


def run_serving(args):
    """Configuration 5.  At its Poisson arrival rate (4 jobs/s): latency percentiles, plain and mixed.
    Saturated (all 256 jobs submitted at once): throughput with grouped solves and without, and the
    mixed path's shifted-position error against plain FP64 over every job."""
    from catalogs import serving_stream
    spec = serving_stream(args.serving_jobs, args.serving_rate)
    log('serving stream: {:d} jobs at {:.1f}/s'.format(len(spec), args.serving_rate))

    def fmt(rep):
        return {k: (float('{:.5g}'.format(v)) if isinstance(v, float) else v) for k, v in rep.items()}

    plain, _ = serving_bench(args, 'plain', spec)
    log('serving plain', plain)
    mixed, _ = serving_bench(args, 'mixed', spec)
    log('serving mixed', mixed)
    sat_plain, ref = serving_bench(args, 'plain', spec, saturate=True, keep=True)
    log('saturated plain', sat_plain)
    sat_mixed, _ = serving_bench(args, 'mixed', spec, saturate=True, reference=ref)
    del ref
    log('saturated mixed', sat_mixed)
    sat_mixed_nogroup, _ = serving_bench(args, 'mixed', spec, saturate=True, group=False)
    log('saturated mixed, ungrouped', sat_mixed_nogroup)
    return dict(workload='configuration 5: {:d} jobs, Poisson {:.1f}/s, nmesh {{128,256,512}} p={{.5,.35,.15}}, '
                         'boxsize 7.8 nmesh, data lognormal median 1e6 [1e5, 2e7], randoms 10x, '
                         '200 GB budget, 4 slots'.format(len(spec), args.serving_rate),
                catalogs='generated on the device per job from its seed (uniform randoms; data Zel\'dovich-displaced '
                         'by 256 random Fourier modes of P(k)), inside the job latency',
                mixed=fmt(mixed), plain=fmt(plain),
                saturated=dict(note='all jobs submitted at once',
                               plain=fmt(sat_plain), mixed=fmt(sat_mixed), mixed_ungrouped=fmt(sat_mixed_nogroup)),
                error_against='plain FP64 run of the same jobs (saturated), every job\'s shifted data positions')


def kernel_only(args):
    """A short run of one kernel at the workload's finest-level shape (for --pmc profiles)."""
    import pyrecon_amd as P
    from catalogs import CONFIGS
    from pyrecon_amd import _lib
    cfg = CONFIGS[args.config]
    n = cfg['nmesh']
    shape = (n, n, n)
    if args.kernel_only == 'copy':
        nbytes = n**3 * 4
        a, b = P.DeviceArray((nbytes,), 'u1'), P.DeviceArray((nbytes,), 'u1')
        a.fill_zero()
        for _ in range(args.steps):
            _lib.check(_lib.lib.prh_copy_kernel(b.ptr, a.ptr, nbytes, None))
        P.synchronize()
        log('copy kernel: {} launches of {} bytes'.format(args.steps, nbytes))
        return
    if args.kernel_only in ('fmg_mixed', 'fmg64'):
        # the multigrid solve alone on a random source (finest-level kernels at the workload's shape)
        mixed = args.kernel_only == 'fmg_mixed'
        ws = P.Workspace(shape, 'f8', flags=_lib.WS_MULTIGRID | (_lib.WS_MIXED if mixed else 0))
        rng = np.random.default_rng(0)
        f = P.DeviceArray.from_numpy(rng.standard_normal(shape))
        v = P.DeviceArray.zeros(shape, 'f8')
        nm, nm_ptr = _lib.size3(shape)
        bs, bs_ptr = _lib.vec3(cfg['boxsize'], 'f8')
        bc, bc_ptr = _lib.vec3(cfg['boxcenter'], 'f8')
        with _lib.KernelTimer(_lib.TIMER_JACOBI_FINEST) as kt:
            t0 = time.perf_counter()
            for _ in range(args.steps):
                if mixed:
                    _lib.check(_lib.lib.prh_fmg_mixed(ws.handle, f.ptr, v.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, 5, 6,
                                                      None, 0, 1, None, None))
                else:
                    _lib.check(_lib.lib.prh_fmg_f64(ws.handle, f.ptr, v.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, 5, 6,
                                                    None, None))
            P.synchronize()
            elapsed = time.perf_counter() - t0
            ms, launches, nbytes = kt.read()
        log('{}: {:.2f} ms per solve; finest-level sweeps {:.4f} ms/launch, {:.1f} GB/s algorithmic'.format(
            args.kernel_only, 1e3 * elapsed / args.steps, ms / launches, nbytes / (ms * 1e-3) / 1e9))
        return
    dtype = 'f4' if args.kernel_only == 'jacobi32' else 'f8'
    ws = P.Workspace(shape, dtype, flags=_lib.WS_MULTIGRID)
    rng = np.random.default_rng(0)
    v = P.DeviceArray.from_numpy(rng.standard_normal(shape).astype(dtype))
    f = P.DeviceArray.from_numpy(rng.standard_normal(shape).astype(dtype))
    nm, nm_ptr = _lib.size3(shape)
    bs, bs_ptr = _lib.vec3(cfg['boxsize'], dtype)
    bc, bc_ptr = _lib.vec3(cfg['boxcenter'], dtype)
    fn = getattr(_lib.lib, 'prh_jacobi_' + ('f32' if dtype == 'f4' else 'f64'))
    with _lib.KernelTimer(_lib.TIMER_JACOBI_FINEST) as kt:
        _lib.check(fn(ws.handle, v.ptr, f.ptr, nm_ptr, bs_ptr, bc_ptr, 0.4, 0.4, args.steps, None, None))
        ms, launches, nbytes = kt.read()
    log('jacobi {}: {} launches, {:.4f} ms/launch, {:.1f} GB/s algorithmic'.format(
        dtype, launches, ms / launches, nbytes / (ms * 1e-3) / 1e9))


def recon_step(P, ws, geo, catalogs, outs, precision, nrefine=0, stages=None):
    """One reconstruction: paint data and randoms, smoothing and density contrast, the solve in
    ``precision``, and the shifted positions of data and randoms (RecSym) into ``outs``.  With
    ``stages`` (a dict), each stage's wall time is added to it after a synchronisation."""
    from catalogs import RECON_KWARGS, SMOOTHING_RADIUS
    d, dw, r, rw = catalogs
    out_d, out_r = outs

    def mark(name):
        if stages is not None:
            P.synchronize()
            stages[name] = stages.get(name, 0.) + time.perf_counter() - mark.t
            mark.t = time.perf_counter()
    mark.t = time.perf_counter()
    recon = P.MultiGridReconstruction(**RECON_KWARGS, **geo, workspace=ws, precision=precision)
    recon.assign_data(d, dw)
    recon.assign_randoms(r, rw)
    mark('paint')
    recon.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
    mark('smoothing+density')
    recon.run(precision=precision, nrefine=nrefine)
    mark('multigrid')
    recon.read_shifted_positions(d, out=out_d)
    recon.read_shifted_positions(r, out=out_r)
    mark('shifts')
    return recon


def run_large(args):
    """Configuration 4: nmesh=1024, boxsize 4000 Mpc/h, 5e7 data and 2.5e8 randoms in the shell.

    Three paths on the same device-resident catalogs: plain FP64 (FP64 meshes, the reference's
    algorithm; the oracle of the other two), mixed (FP64 meshes, FP32 V-cycles inside FP64 defect
    correction) and FP32 storage (FP32 meshes, about 4.3 GB each: the reference's FLOAT32
    algorithm).  For each: recon/s, per-stage times, the finest-level Jacobi sweep against HBM
    peak, and the errors of the potential and of the shifted positions against plain FP64."""
    import pyrecon_amd as P
    from pyrecon_amd import _lib
    from pyrecon_amd.device import device_name
    from catalogs import make_catalogs
    log('device', device_name())
    cat = load_catalogs(args, make_catalogs)
    cfg = cat['config']
    nd, nr = len(cat['data']), len(cat['randoms'])
    log('catalogs: {:d} data, {:d} randoms'.format(nd, nr))
    dev = tuple(P.DeviceArray.from_numpy(cat[k]) for k in ('data', 'data_weights', 'randoms', 'randoms_weights'))
    del cat['data'], cat['data_weights'], cat['randoms'], cat['randoms_weights']
    out_d, out_r = P.DeviceArray((nd, 3), 'f8'), P.DeviceArray((nr, 3), 'f8')
    shape = (cfg['nmesh'],) * 3
    # one paint chunk holds the 2.5e8 randoms (no separate validation pass)
    chunk = 1 << (max(nd, nr) - 1).bit_length()
    paths = [('fp64', 'f8', 'plain'), ('mixed', 'f8', 'mixed'), ('fp32_storage', 'f4', 'plain')]
    report, ref = {}, None
    for name, dtype, precision in paths:
        ws = P.Workspace(shape, dtype, paint_chunk=chunk)
        log('{}: workspace {:.2f} GB'.format(name, ws.nbytes / 1e9))
        geo = dict(nmesh=cfg['nmesh'], boxsize=cfg['boxsize'], boxcenter=cfg['boxcenter'], dtype=dtype)
        for _ in range(args.warmup):
            recon_step(P, ws, geo, dev, (out_d, out_r), precision, args.nrefine)
        P.synchronize()
        with _lib.KernelTimer(_lib.TIMER_JACOBI_FINEST) as kt:
            t0 = time.perf_counter()
            for _ in range(args.steps):
                recon_step(P, ws, geo, dev, (out_d, out_r), precision, args.nrefine)
            P.synchronize()
            elapsed = time.perf_counter() - t0
            ms_j, nj, bytes_j = kt.read()
        stages = {}
        recon = recon_step(P, ws, geo, dev, (out_d, out_r), precision, args.nrefine, stages=stages)
        phi = recon.mesh_phi.value.astype('f8')
        phi -= phi.mean()
        xd, xr = out_d.numpy(), out_r.numpy()
        del recon
        entry = dict(value=round(args.steps / elapsed, 4), unit='recon/s',
                     ms_per_step=round(1e3 * elapsed / args.steps, 2), mesh_dtype=dtype, solve=precision,
                     stages_ms={k: round(1e3 * v, 2) for k, v in stages.items()},
                     jacobi_finest=dict(launches=nj, avg_launch_ms=round(ms_j / max(nj, 1), 5),
                                        achieved_gbs=round(bytes_j / (ms_j * 1e-3) / 1e9, 1) if ms_j else None,
                                        frac_of_hbm_peak=round(bytes_j / (ms_j * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                        if ms_j else None))
        if ref is None:
            ref = (phi, xd, xr)
        else:
            dx = np.concatenate([np.abs(xd - ref[1]).ravel(), np.abs(xr - ref[2]).ravel()])
            entry['accuracy'] = dict(phi_rel_l2=float(np.sqrt(np.sum((phi - ref[0])**2) / np.sum(ref[0]**2))),
                                     shift_max_mpc_h=float(dx.max()), shift_rms_mpc_h=float(np.sqrt(np.mean(dx**2))),
                                     against='plain FP64 GPU path (potentials compared after removing their means)')
            del dx
        del phi, xd, xr
        log(name, entry)
        report[name] = entry
        ws.close()
        P.empty_cache()
    line = dict(metric='recon/s at nmesh=1024 (configuration 4), FP32 storage and mixed vs FP64, 1 GPU',
                config=dict(workload='configuration 4: nmesh=1024, boxsize=4000 Mpc/h, boxcenter=(0,0,3000), '
                                     '{:d} data + {:d} randoms in the 1000<r<3000 shell, FKP weights, local LOS, '
                                     'f=0.8, bias=2, R=15, 5 sweeps x 6 V-cycles'.format(nd, nr),
                            nmesh=cfg['nmesh'], ndata=nd, nrandoms=nr, inputs='device-resident'),
                n_gpus=1, steps=args.steps, warmup=args.warmup, data='synthetic', paths=report)
    print(json.dumps(line), flush=True)


def host_cores():
    """(usable, physical): the cores this process may run on, and the physical cores lscpu reports."""
    usable = len(os.sched_getaffinity(0))
    physical = None
    try:
        import subprocess
        out = subprocess.run(['lscpu', '-p=CORE,SOCKET'], capture_output=True, text=True, timeout=10).stdout
        pairs = {tuple(line.split(',')) for line in out.splitlines() if line and not line.startswith('#')}
        physical = len(pairs) or None
    except Exception:
        pass
    return usable, physical


def cpu_baseline(cat, threads, full=False):
    """The reference's C kernels (oracle/_ref, built from the reference's own src/*.c) plus the
    reference's numpy glue, timed stage by stage for one reconstruction of the workload.

    Default: bounded samples of each stage scaled to the workload (paint and shifts of 4e6 randoms,
    one c2c FFT x4, the density on a 1/8 slab x8, one finest-level V-cycle x 6 x 8/7), with
    ``threads`` OpenMP threads, plus the multigrid and the read at one thread (a 1/8-size
    multigrid problem x8).  ``full``: every stage at full size, once (minutes of CPU time)."""
    from oracle.cpu import RefKernels, OracleKernels
    from oracle.recon_numpy import OracleMultiGridReconstruction
    kind = 'reference'
    try:
        kern = RefKernels()
    except OSError:
        kind, kern = 'port', OracleKernels()
    cfg = cat['config']
    n = cfg['nmesh']

    def use_threads(k):
        os.environ['OMP_NUM_THREADS'] = str(k)
        if kind == 'reference': kern.set_num_threads(k)

    use_threads(threads)
    recon = OracleMultiGridReconstruction(f=0.8, bias=2., los=None, nmesh=n, boxsize=cfg['boxsize'],
                                          boxcenter=cfg['boxcenter'], dtype='f8', kernels=kern)
    ntot = len(cat['data']) + len(cat['randoms'])
    stages, one = {}, {}
    ns = ntot if full else min(4000000, len(cat['randoms']))
    # 1. paint: the reference's serial C loop (one thread whatever OMP_NUM_THREADS) + numpy glue
    t = time.perf_counter()
    if full:
        recon.assign_data(cat['data'], cat['data_weights'])
        recon.assign_randoms(cat['randoms'], cat['randoms_weights'])
    else:
        recon.assign_randoms(cat['randoms'][:ns], cat['randoms_weights'][:ns])
    stages['paint'] = (time.perf_counter() - t) * ntot / ns
    log('cpu baseline: paint {:.2f} s'.format(stages['paint']))
    # 2.-3. smoothing (numpy c2c, single-threaded) and the density contrast
    if full:
        t = time.perf_counter()
        recon.set_density_contrast(smoothing_radius=15.)
        stages['smoothing+density'] = time.perf_counter() - t
        delta = recon.mesh_delta
    else:
        mesh = recon.mesh_randoms
        t = time.perf_counter()
        spec = np.fft.fftn(mesh)
        t_fft = time.perf_counter() - t
        del spec
        slab = mesh[:n // 8]
        t = time.perf_counter()
        s_d, s_r = np.sum(slab), np.sum(slab)
        alpha = s_d * 1. / s_r
        dslab = slab - alpha * slab
        mask = slab > 0.01 * s_r / ns
        dslab[mask] /= (2. * alpha * slab[mask])
        dslab[~mask] = 0.
        stages['smoothing+density'] = 4 * t_fft + (time.perf_counter() - t) * 8
        delta = np.random.default_rng(1).standard_normal((n, n, n))
    log('cpu baseline: smoothing+density {:.2f} s'.format(stages['smoothing+density']))
    # 4. multigrid
    t = time.perf_counter()
    if full:
        phi = kern.fmg(delta, cfg['boxsize'], cfg['boxcenter'], 0.4, 0.4, 5, 6, None)
        stages['multigrid'] = time.perf_counter() - t
    else:
        v = np.zeros_like(delta)
        kern.vcycle(v, delta, cfg['boxsize'], cfg['boxcenter'], 0.4, 0.4, 5, None)
        stages['multigrid'] = (time.perf_counter() - t) * 6 * 8. / 7.
        phi = v
        # one thread: the same V-cycle on a 1/8-size problem, x 8
        use_threads(1)
        h = n // 2
        small = np.ascontiguousarray(delta[:h, :h, :h])
        vs = np.zeros_like(small)
        t = time.perf_counter()
        kern.vcycle(vs, small, [b / 2. for b in np.broadcast_to(cfg['boxsize'], 3)], cfg['boxcenter'], 0.4, 0.4, 5,
                    None)
        one['multigrid'] = (time.perf_counter() - t) * 8 * 6 * 8. / 7.
        del vs, small
        use_threads(threads)
    log('cpu baseline: multigrid {:.2f} s'.format(stages['multigrid']))
    # 5. shifts: the finite-difference read (OpenMP C) + numpy RSD algebra
    recon.mesh_phi = phi
    t = time.perf_counter()
    if full:
        recon.read_shifted_positions(cat['data'])
        recon.read_shifted_positions(cat['randoms'])
    else:
        recon.read_shifted_positions(cat['randoms'][:ns])
    stages['shifts'] = (time.perf_counter() - t) * ntot / ns
    if not full:
        use_threads(1)
        t = time.perf_counter()
        recon.read_shifted_positions(cat['randoms'][:ns // 4])
        one['shifts'] = (time.perf_counter() - t) * 4 * ntot / ns
        use_threads(threads)
    log('cpu baseline: shifts {:.2f} s'.format(stages['shifts']))
    total = sum(stages.values())
    usable, physical = host_cores()
    out = dict(value=1. / total, unit='recon/s', cores=threads, kind=kind,
               host_cores=dict(usable=usable, physical=physical),
               stages_s={k: round(v, 3) for k, v in stages.items()})
    if full:
        out['sample'] = ('one full reconstruction of the workload ({:d} points, {n}^3): the reference\'s serial paint, '
                         'numpy c2c smoothing and density, full FMG, reads of all data and randoms; reference '
                         'src/*.c via oracle/_ref with {:d} OpenMP threads').format(ntot, threads, n=n)
    else:
        one_total = stages['paint'] + stages['smoothing+density'] + one['multigrid'] + one['shifts']
        out['one_thread'] = dict(value=1. / one_total, unit='recon/s',
                                 stages_s={k: round(v, 3) for k, v in dict(stages, **one).items()})
        out['sample'] = ('one reconstruction of the same workload timed stage by stage on samples: paint and shifts '
                         'of {:d} randoms scaled to {:d} points, one {n}^3 c2c FFT x4, density on a 1/8 slab x8, '
                         'one {n}^3 V-cycle x6x8/7; reference src/*.c via oracle/_ref with {:d} OpenMP threads '
                         '(the paint loop is serial in the reference), numpy glue single-threaded; one_thread: '
                         'the multigrid and the read at one thread ({h}^3 V-cycle x8, read of {:d} randoms '
                         'scaled)').format(ns, ntot, threads, ns // 4, n=n, h=n // 2)
    return out


FFT_METRIC = 'recon/s at nmesh=512, IterativeFFTReconstruction (FP64, 1 GPU)'


def fft_cpu_baseline(cat, threads, transforms):
    """The reference's path for the FFT reconstructions on bounded samples: its C paint (oracle/_ref)
    of 4e6 randoms scaled to every point, numpy's c2c transform of the 512^3 mesh (one, timed) times
    the transforms of one step (the smoothing's four and those of the algorithm), and its shifts (the
    three CIC reads of oracle/_ref plus the numpy algebra) of 4e6 randoms scaled."""
    from oracle.cpu import RefKernels, OracleKernels
    from oracle.recon_numpy import OracleIterativeFFTReconstruction
    kind = 'reference'
    try:
        kern = RefKernels()
    except OSError:
        kind, kern = 'port', OracleKernels()
    os.environ['OMP_NUM_THREADS'] = str(threads)
    if kind == 'reference': kern.set_num_threads(threads)
    cfg = cat['config']
    n = cfg['nmesh']
    recon = OracleIterativeFFTReconstruction(f=0.8, bias=2., los=None, nmesh=n, boxsize=cfg['boxsize'],
                                             boxcenter=cfg['boxcenter'], dtype='f8', kernels=kern)
    ntot = len(cat['data']) + len(cat['randoms'])
    ns = min(4000000, len(cat['randoms']))
    stages = {}
    t = time.perf_counter()
    recon.assign_randoms(cat['randoms'][:ns], cat['randoms_weights'][:ns])
    stages['paint'] = (time.perf_counter() - t) * ntot / ns
    t = time.perf_counter()
    spec = np.fft.fftn(recon.mesh_randoms)
    t_fft = time.perf_counter() - t
    del spec
    stages['transforms'] = (4 + transforms) * t_fft
    rng = np.random.default_rng(1)
    recon.mesh_psi = [rng.standard_normal((n, n, n)) for _ in range(3)]
    t = time.perf_counter()
    recon.read_shifted_positions(cat['randoms'][:ns])
    stages['shifts'] = (time.perf_counter() - t) * ntot / ns
    usable, physical = host_cores()
    return dict(value=1. / sum(stages.values()), unit='recon/s', cores=threads, kind=kind,
                host_cores=dict(usable=usable, physical=physical), stages_s={k: round(v, 3) for k, v in stages.items()},
                sample=('one IterativeFFTReconstruction step of the same workload on samples: paint and shifts (three '
                        'CIC reads + numpy algebra) of {:d} randoms scaled to {:d} points, one {n}^3 numpy c2c '
                        'transform x {:d} (4 for the smoothing, {:d} for the algorithm); reference src/*.c via '
                        'oracle/_ref with {:d} OpenMP threads, numpy single-threaded').format(
                            ns, ntot, 4 + transforms, transforms, threads, n=n))


def run_fft(args):
    """The FFT reconstructions on the workload's catalogs (FP64 meshes): IterativeFFTReconstruction
    (niterations=3, the value), PlaneParallelFFTReconstruction (global line of sight along z) and
    IterativeFFTParticleReconstruction (niterations=3).  A step: paint, smoothing and density contrast,
    run, and the shifted positions of data and randoms.  Their shifts take and return host arrays
    (the particle variant also keeps the data on the host), so a step includes those transfers; the
    paint reads the device-resident catalogs.  The roofline is that of the 3-D transform passes."""
    import pyrecon_amd as P
    from pyrecon_amd import _lib
    from pyrecon_amd.device import device_name
    from catalogs import make_catalogs
    log('device', device_name())
    cat = load_catalogs(args, make_catalogs)
    cfg = cat['config']
    nd, nr = len(cat['data']), len(cat['randoms'])
    d, dw = P.DeviceArray.from_numpy(cat['data']), P.DeviceArray.from_numpy(cat['data_weights'])
    r, rw = P.DeviceArray.from_numpy(cat['randoms']), P.DeviceArray.from_numpy(cat['randoms_weights'])
    geo = dict(nmesh=cfg['nmesh'], boxsize=cfg['boxsize'], boxcenter=cfg['boxcenter'], dtype='f8')
    steps = max(1, min(args.steps, 3))
    algorithms = [('IterativeFFTReconstruction', P.IterativeFFTReconstruction, {}, dict(niterations=3), True),
                  ('PlaneParallelFFTReconstruction', P.PlaneParallelFFTReconstruction, dict(los='z'), {}, True),
                  ('IterativeFFTParticleReconstruction', P.IterativeFFTParticleReconstruction, {},
                   dict(niterations=3), False)]
    results = {}
    for name, cls, kwargs, run_kwargs, device_data in algorithms:
        def step():
            rec = cls(f=0.8, bias=2., **geo, **kwargs)
            if device_data:
                rec.assign_data(d, dw)
            else:
                rec.assign_data(cat['data'], cat['data_weights'])
            rec.assign_randoms(r, rw)
            rec.set_density_contrast(smoothing_radius=15.)
            rec.run(**run_kwargs)
            return rec.read_shifted_positions(cat['data']), rec.read_shifted_positions(cat['randoms'])

        step()
        P.synchronize()
        with _lib.KernelTimer(_lib.TIMER_FFT_PASS) as kt:
            t0 = time.perf_counter()
            for _ in range(steps): step()
            P.synchronize()
            elapsed = time.perf_counter() - t0
            ms, nl, nb = kt.read()
        results[name] = dict(value=round(steps / elapsed, 4), unit='recon/s', ms_per_step=round(1e3 * elapsed / steps, 2),
                             fft_passes_per_step=nl // steps, fft_ms_per_step=round(ms / steps, 2),
                             fft_pass_avg_ms=round(ms / max(nl, 1), 4),
                             fft_pass_gbs=round(nb / (ms * 1e-3) / 1e9, 1) if ms else None)
        log(name, results[name])
    main = results['IterativeFFTReconstruction']
    achieved = main['fft_pass_gbs']
    roofline = dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s',
                    frac=round(achieved / HBM_PEAK_GBS, 4) if achieved else None, traffic=None,
                    kernel='k_fft_axis: one axis pass of a 3-D c2c transform, 512^3 FP64, LDS Stockham passes; '
                           'achieved = algorithmic bytes (4 n^3 x 8 B per complex pass, 3 n^3 x 8 B for the real-in '
                           'and real-out passes) / average pass time (HIP events) over the IterativeFFTReconstruction '
                           'steps', avg_launch_ms=main['fft_pass_avg_ms'])
    cpu = None
    if not args.no_cpu_baseline:
        ntrans = main['fft_passes_per_step'] // 3
        log('cpu baseline ({:d} transforms per step)'.format(ntrans))
        cpu = fft_cpu_baseline(cat, threads=args.workers, transforms=ntrans)
        log('cpu baseline', cpu)
    line = dict(metric=FFT_METRIC, value=main['value'], unit='recon/s', n_gpus=1, steps=steps, warmup=1,
                ms_per_step=main['ms_per_step'], higher_is_better=True, scaling='weak', vs_baseline=None,
                dtype='f64', data='synthetic',
                config=dict(workload='configuration 3 catalogs ({:d} data + {:d} randoms, nmesh=512, boxsize=4000, '
                                     'boxcenter=(0,0,3000)), f=0.8, bias=2, R=15; IterativeFFTReconstruction '
                                     'niterations=3, local LOS'.format(nd, nr),
                            nmesh=cfg['nmesh'], ndata=nd, nrandoms=nr, inputs='catalogs device-resident for the '
                            'paint; shifts from and to host arrays'),
                roofline=roofline, cpu_baseline=cpu, algorithms=results)
    print(json.dumps(line), flush=True)


def load_catalogs(args, make_catalogs):
    """The seeded catalogs of the configuration, generated or read back from --catalog-cache."""
    keys = ['data', 'data_weights', 'randoms', 'randoms_weights']
    path = None
    if args.catalog_cache:
        path = os.path.join(args.catalog_cache, '{}_{:g}'.format(args.config, args.scale))
        if all(os.path.isfile(os.path.join(path, k + '.npy')) for k in keys):
            log('loading catalogs from', path)
            from catalogs import CONFIGS
            cat = {k: np.load(os.path.join(path, k + '.npy')) for k in keys}
            cat['config'] = dict(CONFIGS[args.config])
            return cat
    log('generating catalogs for', args.config)
    cat = make_catalogs(args.config, scale=args.scale, workers=args.workers, progress=log)
    if path:
        os.makedirs(path, exist_ok=True)
        for k in keys: np.save(os.path.join(path, k + '.npy'), cat[k])
    return cat


def main():
    args = parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: every GPU box here has one GPU, so --gpus {} is not '
              'supported (multi-GPU sharding is a design note, not a measured path).'.format(args.gpus),
              file=sys.stderr)
        sys.exit(2)
    if args.kernel_only:
        return kernel_only(args)
    if args.serving_only:
        print(json.dumps(dict(metric=METRIC, serving=run_serving(args))), flush=True)
        return
    if args.cpu_baseline_full:
        from catalogs import make_catalogs
        cat = load_catalogs(args, make_catalogs)
        print(json.dumps(dict(metric=METRIC, config=args.config,
                              cpu_baseline=cpu_baseline(cat, threads=args.workers, full=True))), flush=True)
        return
    if args.config == 'c4':
        return run_large(args)
    if args.fft_recon:
        return run_fft(args)
    import pyrecon_amd as P
    from pyrecon_amd import _lib
    from pyrecon_amd.device import device_name
    from catalogs import make_catalogs
    log('device', device_name())
    cat = load_catalogs(args, make_catalogs)
    cfg = cat['config']
    nd, nr = len(cat['data']), len(cat['randoms'])
    log('catalogs: {:d} data, {:d} randoms'.format(nd, nr))
    d, dw = P.DeviceArray.from_numpy(cat['data']), P.DeviceArray.from_numpy(cat['data_weights'])
    r, rw = P.DeviceArray.from_numpy(cat['randoms']), P.DeviceArray.from_numpy(cat['randoms_weights'])
    shape = (cfg['nmesh'],) * 3
    ws = P.Workspace(shape, "f8", paint_chunk=1 << 27)  # one chunk for 1.2e8 points: no validation pass
    log('workspace {:.2f} GB'.format(ws.nbytes / 1e9))
    out_d, out_r = P.DeviceArray((nd, 3), 'f8'), P.DeviceArray((nr, 3), 'f8')
    geo = dict(nmesh=cfg['nmesh'], boxsize=cfg['boxsize'], boxcenter=cfg['boxcenter'], dtype='f8')

    def step(precision, stages=None):
        return recon_step(P, ws, geo, (d, dw, r, rw), (out_d, out_r), precision, args.nrefine, stages)

    def timed(precision, steps, warmup):
        for _ in range(warmup): step(precision)
        P.synchronize()
        with _lib.KernelTimer(_lib.TIMER_JACOBI_FINEST) as kt, _lib.KernelTimer(_lib.TIMER_PAIR_FINEST) as kp:
            t0 = time.perf_counter()
            for _ in range(steps): step(precision)
            P.synchronize()
            elapsed = time.perf_counter() - t0
            jac, pair = kt.read(), kp.read()
        return elapsed, jac, pair

    log('mixed: warmup {:d}, timed {:d}'.format(args.warmup, args.steps))
    elapsed, jac, pair = timed('mixed', args.steps, args.warmup)
    value = args.steps / elapsed
    log('mixed: {:.3f} recon/s ({:.1f} ms/step)'.format(value, 1e3 * elapsed / args.steps))
    stages_mixed = {}
    step('mixed', stages=stages_mixed)
    nfp64 = args.fp64_steps if args.fp64_steps is not None else max(1, args.steps // 2)
    log('plain fp64: timed {:d}'.format(nfp64))
    elapsed64, jac64, _ = timed('plain', nfp64, 1)
    value64 = nfp64 / elapsed64
    log('plain fp64: {:.3f} recon/s ({:.1f} ms/step)'.format(value64, 1e3 * elapsed64 / nfp64))
    stages_plain = {}
    step('plain', stages=stages_plain)

    host_io = None
    if args.host_io_steps > 0:
        # end to end from numpy, as a drop-in user calls it: a new reconstruction object per step with
        # the default workspace settings, catalogs uploaded and shifted positions returned inside each step
        from catalogs import RECON_KWARGS, SMOOTHING_RADIUS

        def host_step():
            recon = P.MultiGridReconstruction(**RECON_KWARGS, **geo, precision='mixed')
            recon.assign_data(cat['data'], cat['data_weights'])
            recon.assign_randoms(cat['randoms'], cat['randoms_weights'])
            recon.set_density_contrast(smoothing_radius=SMOOTHING_RADIUS)
            recon.run(precision='mixed', nrefine=args.nrefine)
            return recon.read_shifted_positions(cat['data']), recon.read_shifted_positions(cat['randoms'])

        host_step()
        P.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.host_io_steps): host_step()
        P.synchronize()
        th = (time.perf_counter() - t0) / args.host_io_steps
        host_io = dict(value=round(1. / th, 4), unit='recon/s', ms_per_step=round(1e3 * th, 2),
                       steps=args.host_io_steps, precision='mixed',
                       note='a new MultiGridReconstruction per step with default settings (its workspace kept '
                            'between steps), catalogs uploaded from numpy (pageable host memory, copied through the '
                            'library\'s pinned staging buffers) and shifted positions of data and randoms returned '
                            'to numpy inside each step: PCIe-inclusive, not the headline value')
        log('host io', host_io)

    accuracy = None
    if not args.no_accuracy:
        # mixed vs plain FP64 (the plain FP64 solve is bit-identical to the reference's C multigrid)
        rp = step('plain')
        phi_p, xd_p, xr_p = rp.mesh_phi.value, out_d.numpy(), out_r.numpy()
        rm = step('mixed')
        phi_m, xd_m, xr_m = rm.mesh_phi.value, out_d.numpy(), out_r.numpy()
        a, b = phi_m - phi_m.mean(), phi_p - phi_p.mean()
        dx = np.concatenate([np.abs(xd_m - xd_p).ravel(), np.abs(xr_m - xr_p).ravel()])
        accuracy = dict(phi_rel_l2=float(np.sqrt(np.sum((a - b)**2) / np.sum(b**2))),
                        shift_max_mpc_h=float(dx.max()), shift_rms_mpc_h=float(np.sqrt(np.mean(dx**2))),
                        against='plain FP64 GPU path (bit-identical to the reference FP64 multigrid)')
        log('accuracy', accuracy)
        del phi_p, xd_p, xr_p, phi_m, xd_m, xr_m, a, b, dx

    # the dominant kernel: the finest-level sweeps of the mixed path, two per launch where paired
    ms_p, np_, bytes_p = pair
    ms_j, nj, bytes_j = jac
    single = dict(launches=nj, avg_launch_ms=round(ms_j / max(nj, 1), 5),
                  achieved_gbs=round(bytes_j / (ms_j * 1e-3) / 1e9, 1) if ms_j > 0 else None,
                  kernel='k_stencil_rows<float, fast> (one sweep per launch)')
    paired = ms_p > ms_j
    ms_d, nd_, bytes_d = pair if paired else jac
    achieved = bytes_d / (ms_d * 1e-3) / 1e9 if ms_d > 0 else 0.
    traffic = None
    tfile = os.path.join(ROOT, 'profiles', 'stencil_pair_f32_512_traffic.json' if paired else 'jacobi_f32_512_traffic.json')
    if os.path.isfile(tfile):
        with open(tfile) as fh: traffic = json.load(fh).get('hbm_bytes_per_launch')
    if paired:
        kernel = ('k_stencil_pair<float, fast>: two damped-Jacobi sweeps per launch (blocks of 8 whole 512-cell rows, '
                  'the intermediate iterate in LDS), finest level 512^3 (mixed path); achieved = algorithmic bytes '
                  '(3 x 512^3 x 4 B per pair, 2 x 512^3 x 4 B from a zero start) / average launch time (HIP events); '
                  'traffic = FETCH_SIZE + WRITE_SIZE per launch, calibrated on a copy kernel')
    else:
        kernel = ('k_stencil_rows<float, fast> damped-Jacobi sweep (blocks of 8 whole 512-cell rows), '
                  'finest level 512^3 (mixed path); achieved = algorithmic bytes (3 x 512^3 x 4 B per '
                  'sweep) / average launch time (HIP events); the last sweep of each correction (added into the '
                  'FP64 iterate) and the first two sweeps from zero (one launch) are other kernels, not timed; '
                  'traffic = FETCH_SIZE + WRITE_SIZE per launch, calibrated on a copy kernel')
    roofline = dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=kernel,
                    launches=nd_, avg_launch_ms=round(ms_d / max(nd_, 1), 5),
                    bytes_per_launch=(bytes_d / nd_) if nd_ else None)
    if paired:
        roofline['single_sweeps'] = single
    cpu = None
    if not args.no_cpu_baseline:
        log('cpu baseline')
        cpu = cpu_baseline(cat, threads=args.workers)
        log('cpu baseline', cpu)
    serving = None
    if args.serving_jobs > 0:
        del d, dw, r, rw, out_d, out_r
        ws.close()
        P.empty_cache()
        serving = run_serving(args)
    ms_j64, nj64, bytes_j64 = jac64
    line = dict(metric=METRIC, value=round(value, 4), unit='recon/s', n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=round(1e3 * elapsed / args.steps, 3), higher_is_better=True,
                scaling='weak', vs_baseline=None, dtype='f32+f64', data='synthetic',
                config=dict(workload='configuration 3: nmesh=512, boxsize=4000 Mpc/h, boxcenter=(0,0,3000), '
                                     '{:d} data + {:d} randoms in the 1000<r<3000 shell, FKP weights, local LOS, '
                                     'f=0.8, bias=2, R=15, 5 sweeps x 6 V-cycles'.format(nd, nr),
                            nmesh=cfg['nmesh'], ndata=nd, nrandoms=nr, precision='mixed', nrefine=args.nrefine,
                            inputs='device-resident'),
                roofline=roofline, cpu_baseline=cpu,
                fp64_plain=dict(value=round(value64, 4), unit='recon/s', ms_per_step=round(1e3 * elapsed64 / nfp64, 3),
                                jacobi_avg_launch_ms=round(ms_j64 / max(nj64, 1), 5),
                                jacobi_achieved_gbs=round(bytes_j64 / (ms_j64 * 1e-3) / 1e9, 1) if ms_j64 else None),
                speedup_mixed_vs_fp64=round(value / value64, 3), accuracy=accuracy, host_io=host_io,
                stages_ms=dict(mixed={k: round(1e3 * v, 2) for k, v in stages_mixed.items()},
                               plain={k: round(1e3 * v, 2) for k, v in stages_plain.items()}),
                serving=serving)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
