"""Serving: a job's result is the same alone and in a mix of concurrent jobs.

Jobs of uneven mesh size, catalog size, precision and density rule are submitted to
one :class:`ReconServer` at once (several streams, grouped solves); every job's
shifted positions must equal, bit for bit, those of the same reconstruction run
alone through the drop-in class.  Admission: a job larger than the budget fails
alone and the other jobs still complete.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view('u{:d}'.format(a.dtype.itemsize))


def make_specs():
    # (nmesh, ndata, precision, original, los); boxes of 7.8 Mpc/h cells as in the serving stream
    return [(32, 20000, 'plain', False, None), (64, 60000, 'mixed', False, None), (32, 15000, 'plain', False, 'x'),
            (64, 40000, 'plain', True, None), (32, 25000, 'mixed', False, None), (64, 50000, 'plain', False, None),
            (32, 10000, 'plain', False, None)]


def catalogs(spec, seed):
    from pyrecon_amd.serving import synthetic_catalog, zeldovich_modes
    nmesh, ndata = spec[0], spec[1]
    boxsize = 7.8 * nmesh
    boxcenter = (0., 0., 2. * boxsize)
    modes = zeldovich_modes([boxsize] * 3, [nmesh] * 3, seed, nmodes=64)

    def data(stream):
        return synthetic_catalog(ndata, boxsize, boxcenter, seed=2 * seed, modes=modes, f=0.8, stream=stream)

    def randoms(stream):
        return synthetic_catalog(10 * ndata, boxsize, boxcenter, seed=2 * seed + 1, stream=stream)

    return boxsize, boxcenter, data, randoms


def job_for(spec, seed):
    from pyrecon_amd.serving import ReconJob
    nmesh, ndata, precision, original, los = spec
    boxsize, boxcenter, data, randoms = catalogs(spec, seed)
    return ReconJob(data, randoms, ndata=ndata, nrandoms=10 * ndata, nmesh=nmesh, boxsize=boxsize,
                    boxcenter=boxcenter, f=0.8, bias=2., los=los, precision=precision, original=original,
                    shifted=('data', 'randoms'), to_host=True, name='job{:d}'.format(seed))


def alone(spec, seed):
    import pyrecon_amd
    nmesh, ndata, precision, original, los = spec
    boxsize, boxcenter, data, randoms = catalogs(spec, seed)
    cls = pyrecon_amd.OriginalMultiGridReconstruction if original else pyrecon_amd.MultiGridReconstruction
    recon = cls(f=0.8, bias=2., los=los, nmesh=nmesh, boxsize=boxsize, boxcenter=boxcenter, dtype='f8',
                precision=precision)
    dpos, dw = data(None)
    rpos, rw = randoms(None)
    recon.assign_data(dpos, dw)
    recon.assign_randoms(rpos, rw)
    recon.set_density_contrast(smoothing_radius=15.)
    recon.run(precision=precision)
    return recon.read_shifted_positions(dpos).numpy(), recon.read_shifted_positions(rpos, field='disp').numpy()


def test_alone_equals_mixed():
    from pyrecon_amd.serving import ReconServer
    specs = make_specs()
    expect = [alone(spec, seed) for seed, spec in enumerate(specs, start=1)]
    with ReconServer(budget_bytes=40 * 10**9, slots=4, paint_chunk=1 << 20) as server:
        futures = [server.submit(job_for(spec, seed)) for seed, spec in enumerate(specs, start=1)]
        results = [f.result(timeout=300) for f in futures]
        report = server.report()
    for (d, r), res in zip(expect, results):
        assert np.array_equal(bits(res.shifted_data), bits(d)), res.job.name
        assert np.array_equal(bits(res.shifted_randoms), bits(r)), res.job.name
    assert report['jobs'] == len(specs)
    assert report['p50_s'] <= report['p95_s'] <= report['p99_s']
    print('serving mix:', report)


def test_admission_failure_is_isolated():
    from pyrecon_amd.serving import ReconServer, AdmissionError
    specs = make_specs()
    big = job_for((64, 10**7, 'plain', False, None), 99)  # needs far more than the budget below
    small = job_for(specs[0], 1)
    with ReconServer(budget_bytes=3 * 10**9, slots=2, paint_chunk=1 << 20) as server:
        f_big = server.submit(big)
        f_small = server.submit(small)
        with pytest.raises(AdmissionError):
            f_big.result(timeout=120)
        res = f_small.result(timeout=120)
    d, r = alone(specs[0], 1)
    assert np.array_equal(bits(res.shifted_data), bits(d))
