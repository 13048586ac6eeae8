"""Seeded synthetic catalogs for the benchmark configurations.

Data: a uniform cubic lattice displaced by the Zel'dovich field
psi = -grad(inverse Laplacian(delta)), i.e. psi_k = i k delta_k / k^2, of a
Gaussian field with P(k) = 2e4 (k/0.1)^-1.5 exp(-(k/0.3)^2) (Mpc/h)^3 drawn on
the lattice grid (k = 0 mode removed), plus the line-of-sight RSD shift
f (psi . rhat) rhat with rhat = x / |x| (observer at the origin), wrapped into
the box.  Randoms: uniform in the box, or in the shell 1000 < r < 3000 Mpc/h
(intersected with the box) with the radial profile
n(r) = 3e-4 exp(-((r - 2000) / 600)^2).

Seeds follow the reference tests' convention (tests/utils.py:38, 43, 68 of
pyrecon): 42 for the field, 43 for sampling, 44 for the randoms.  FFTs use
scipy.fft with several workers (deterministic: work is split over
independent 1D transforms).
"""
import hashlib

import numpy as np

FIELD_SEED, SAMPLE_SEED, RANDOMS_SEED = 42, 43, 44


def power_spectrum(k):
    with np.errstate(divide='ignore'):
        pk = 2e4 * (k / 0.1)**-1.5 * np.exp(-(k / 0.3)**2)
    return np.where(k > 0, pk, 0.)


def zeldovich_displacements(nlattice, boxsize, seed=FIELD_SEED, workers=8):
    """Displacement field (3, n, n, n) float64 on an n^3 lattice of a periodic box."""
    import scipy.fft
    n = int(nlattice)
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n), dtype=np.float64)
    wk = scipy.fft.rfftn(white, workers=workers)
    del white
    cell = boxsize / n
    kx = 2 * np.pi * np.fft.fftfreq(n, d=cell)
    kz = 2 * np.pi * np.fft.rfftfreq(n, d=cell)
    k2 = kx[:, None, None]**2 + kx[None, :, None]**2 + kz[None, None, :]**2
    amp = np.sqrt(power_spectrum(np.sqrt(k2)) / cell**3)
    deltak = wk * amp
    del wk, amp
    with np.errstate(divide='ignore', invalid='ignore'):
        inv = np.where(k2 > 0, 1. / k2, 0.)
    psi = np.empty((3, n, n, n), dtype=np.float64)
    for axis, kk in enumerate([kx[:, None, None], kx[None, :, None], kz[None, None, :]]):
        psi[axis] = scipy.fft.irfftn(1j * kk * inv * deltak, s=(n, n, n), workers=workers)
    return psi


def lattice_positions(nlattice, boxsize, boxcenter):
    n = int(nlattice)
    offset = np.asarray(boxcenter, dtype='f8') - boxsize / 2.
    q = (np.arange(n) + 0.5) * (boxsize / n)
    grid = np.meshgrid(q, q, q, indexing='ij')
    return np.stack([g.ravel() for g in grid], axis=-1) + offset


def zeldovich_catalog(nlattice, boxsize, boxcenter, f=0.8, seed=FIELD_SEED, workers=8):
    """Displaced lattice with RSD along x/|x|, wrapped into the box; (n^3, 3) float64."""
    psi = zeldovich_displacements(nlattice, boxsize, seed=seed, workers=workers)
    psi = psi.reshape(3, -1).T
    x = lattice_positions(nlattice, boxsize, boxcenter) + psi
    rhat = x / np.sqrt(np.sum(x**2, axis=-1))[:, None]
    x += f * np.sum(psi * rhat, axis=-1)[:, None] * rhat
    offset = np.asarray(boxcenter, dtype='f8') - boxsize / 2.
    x = (x - offset) % boxsize + offset
    return x


def uniform_box(size, boxsize, boxcenter, seed=RANDOMS_SEED):
    rng = np.random.default_rng(seed)
    offset = np.asarray(boxcenter, dtype='f8') - boxsize / 2.
    return rng.random((int(size), 3)) * boxsize + offset


def shell_nbar(r):
    return 3e-4 * np.exp(-((r - 2000.) / 600.)**2)


def fkp_weights(r):
    return 1. / (1. + 1e4 * shell_nbar(r))


def in_shell_box(x, boxsize, boxcenter, rmin=1000., rmax=3000.):
    r = np.sqrt(np.sum(x**2, axis=-1))
    lo = np.asarray(boxcenter) - boxsize / 2.
    hi = np.asarray(boxcenter) + boxsize / 2.
    return (r > rmin) & (r < rmax) & np.all((x >= lo) & (x < hi), axis=-1), r


def shell_randoms(size, boxsize, boxcenter, seed=RANDOMS_SEED, batch=1 << 22, workers=8):
    """``size`` points uniform in the shell intersected with the box, thinned by n(r)/max n.

    Candidates are drawn in batches, each with its own generator spawned from ``seed``,
    and accepted points are concatenated in batch order: the result does not depend on
    the number of worker threads."""
    from concurrent.futures import ThreadPoolExecutor
    offset = np.asarray(boxcenter, dtype='f8') - boxsize / 2.
    seeds = np.random.SeedSequence(seed)

    def one(child):
        rng = np.random.default_rng(child)
        x = rng.random((batch, 3)) * boxsize + offset
        keep, r = in_shell_box(x, boxsize, boxcenter)
        keep &= rng.random(batch) < shell_nbar(r) / 3e-4
        return x[keep]

    out = np.empty((int(size), 3), dtype='f8')
    filled = 0
    with ThreadPoolExecutor(max(int(workers), 1)) as pool:
        while filled < size:
            for x in pool.map(one, seeds.spawn(max(int(workers), 1))):
                take = min(len(x), size - filled)
                out[filled:filled + take] = x[:take]
                filled += take
                if filled >= size: break
    return out


CONFIGS = {
    # name: (nmesh, boxsize, boxcenter, ndata, nrandoms, nlattice, geometry, data weights)
    'c1': dict(nmesh=64, boxsize=1000., boxcenter=(0., 0., 1500.), ndata=100000, nrandoms=1000000, nlattice=47,
               geometry='box', weights='unit'),
    'c2': dict(nmesh=256, boxsize=2000., boxcenter=(0., 0., 2000.), ndata=2000000, nrandoms=20000000, nlattice=126,
               geometry='box', weights='uniform'),
    'c3': dict(nmesh=512, boxsize=4000., boxcenter=(0., 0., 3000.), ndata=20000000, nrandoms=100000000,
               nlattice=512, geometry='shell', weights='fkp'),
    # about 20% of the lattice lies in the shell after thinning: 672^3 gives 6.2e7 >= 5e7 points
    'c4': dict(nmesh=1024, boxsize=4000., boxcenter=(0., 0., 3000.), ndata=50000000, nrandoms=250000000,
               nlattice=672, geometry='shell', weights='fkp'),
}
RECON_KWARGS = dict(f=0.8, bias=2.0, los=None)
SMOOTHING_RADIUS = 15.


def make_catalogs(name, scale=1., workers=8, with_randoms=True, progress=None):
    """Data/randoms positions and weights of a configuration.

    ``scale`` < 1 shrinks the point counts (not the mesh) for quick tests; ``progress`` is
    called with a short message after each generation step.
    """
    progress = progress or (lambda msg: None)
    cfg = dict(CONFIGS[name])
    ndata = int(cfg['ndata'] * scale)
    nrandoms = int(cfg['nrandoms'] * scale)
    boxsize, boxcenter = cfg['boxsize'], cfg['boxcenter']
    x = zeldovich_catalog(cfg['nlattice'], boxsize, boxcenter, f=RECON_KWARGS['f'], workers=workers)
    progress('displaced lattice of {:d} points'.format(len(x)))
    rng = np.random.default_rng(SAMPLE_SEED)
    if cfg['geometry'] == 'shell':
        keep, r = in_shell_box(x, boxsize, boxcenter)
        keep &= rng.random(len(x)) < shell_nbar(r) / 3e-4
        x = x[keep]
    if len(x) < ndata:
        raise ValueError('lattice too small: {:d} < {:d} points'.format(len(x), ndata))
    index = np.sort(rng.choice(len(x), size=ndata, replace=False))
    data = np.ascontiguousarray(x[index])
    del x
    if cfg['weights'] == 'unit':
        data_weights = np.ones(ndata, dtype='f8')
    elif cfg['weights'] == 'uniform':
        data_weights = rng.uniform(0.5, 1.5, ndata)
    else:
        data_weights = fkp_weights(np.sqrt(np.sum(data**2, axis=-1)))
    out = dict(config=cfg, data=data, data_weights=data_weights)
    progress('data catalog')
    if with_randoms:
        if cfg['geometry'] == 'shell':
            randoms = shell_randoms(nrandoms, boxsize, boxcenter, workers=workers)
            randoms_weights = fkp_weights(np.sqrt(np.sum(randoms**2, axis=-1)))
        else:
            randoms = uniform_box(nrandoms, boxsize, boxcenter)
            randoms_weights = np.ones(nrandoms, dtype='f8')
        out.update(randoms=randoms, randoms_weights=randoms_weights)
    return out


def sha256(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def serving_stream(njobs, rate, seed=2024):
    """Configuration 5: Poisson arrivals at ``rate`` jobs/s; nmesh in {128, 256, 512} with
    probabilities {0.5, 0.35, 0.15}; boxsize = 7.8 nmesh Mpc/h centred at (0, 0, boxsize); data
    count lognormal with median 1e6 and sigma(ln N) = 1, clipped to [1e5, 2e7]; randoms 10x."""
    rng = np.random.RandomState(seed)
    arrivals = np.cumsum(rng.exponential(1. / rate, njobs))
    nmesh = rng.choice([128, 256, 512], size=njobs, p=[0.5, 0.35, 0.15])
    ndata = np.clip(np.exp(np.log(1e6) + rng.normal(size=njobs)), 1e5, 2e7).astype(np.int64)
    return [dict(arrival=float(a), nmesh=int(n), ndata=int(d), seed=int(seed * 1000 + i))
            for i, (a, n, d) in enumerate(zip(arrivals, nmesh, ndata))]


def serving_catalogs(spec):
    """Geometry and device catalog makers ``fn(stream) -> (positions, weights)`` of one serving job:
    data Zel'dovich-displaced by 256 random Fourier modes of P(k) with the RSD shift and weights
    uniform in [0.5, 1.5] (configuration 2's generator, on the device), randoms uniform with unit
    weights, each from the job's seed."""
    from pyrecon_amd.serving import synthetic_catalog, zeldovich_modes
    n, nd = spec['nmesh'], spec['ndata']
    boxsize = 7.8 * n
    boxcenter = (0., 0., boxsize)
    modes = zeldovich_modes([boxsize] * 3, [n] * 3, spec['seed'])

    def data(stream):
        return synthetic_catalog(nd, boxsize, boxcenter, 2 * spec['seed'], modes=modes, f=RECON_KWARGS['f'],
                                 weights=(0.5, 1.5), stream=stream)

    def randoms(stream):
        return synthetic_catalog(10 * nd, boxsize, boxcenter, 2 * spec['seed'] + 1, weights=(1., 1.), stream=stream)

    return boxsize, boxcenter, data, randoms


def serving_job(spec, precision, name=None, **kwargs):
    """The :class:`pyrecon_amd.serving.ReconJob` of one configuration-5 job (RecSym shifts of data and
    randoms, device outputs)."""
    from pyrecon_amd.serving import ReconJob
    n, nd = spec['nmesh'], spec['ndata']
    boxsize, boxcenter, data, randoms = serving_catalogs(spec)
    kw = dict(shifted=('data', 'randoms'), field='disp+rsd', randoms_field='disp+rsd', to_host=False)
    kw.update(kwargs)
    return ReconJob(data, randoms, ndata=nd, nrandoms=10 * nd, nmesh=n, boxsize=boxsize, boxcenter=boxcenter,
                    f=RECON_KWARGS['f'], bias=RECON_KWARGS['bias'], los=RECON_KWARGS['los'], dtype='f8',
                    smoothing_radius=SMOOTHING_RADIUS, precision=precision, name=name, **kw)
