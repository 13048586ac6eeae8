"""Host side of the command-line reconstruction (pyrecon_amd/cli.py, after bin/pyrecon:70-423) and
the coordinate utilities (pyrecon_amd/utils.py, after pyrecon/utils.py:146-282).

- sky <-> Cartesian conversions are inverse to each other and give the textbook values on the axes;
- the flat LCDM comoving distance reduces to the Einstein-de Sitter closed form
  D = 2 (c / 100) (1 - 1 / sqrt(1 + z)) Mpc/h at Omega_m = 1, and matches
  D(0.5) = 1322.04 Mpc/h, D(1) = 2312.68 Mpc/h at Omega_m = 0.3 (1888.63 and 3303.83 Mpc for h = 0.7);
- DistanceToRedshift inverts it;
- catalogs are read with the reference's column, mask, weight and format rules.
"""
import os

import numpy as np
import pytest

from pyrecon_amd import utils, cli


def test_sky_cartesian_round_trip():
    rng = np.random.RandomState(1)
    pos = rng.normal(size=(1000, 3)) * 1000.
    dist, ra, dec = utils.cartesian_to_sky(pos)
    assert np.all((ra >= 0.) & (ra < 360.)) and np.all(np.abs(dec) <= 90.)
    assert np.allclose(utils.sky_to_cartesian(dist, ra, dec), pos, rtol=0, atol=1e-9)
    d, r, de = utils.cartesian_to_sky(np.array([[2., 0., 0.], [0., 3., 0.], [0., 0., 4.], [0., -1., 0.]]))
    assert np.allclose(d, [2., 3., 4., 1.]) and np.allclose(r[[0, 1, 3]], [0., 90., 270.]) and np.allclose(de[2], 90.)


def test_flat_lcdm_distance():
    z = np.array([0., 0.01, 0.5, 1., 3., 10.])
    eds = utils.FlatLCDMDistance(1.)
    assert np.allclose(eds(z), 2. * utils.C_KMS / 100. * (1. - 1. / np.sqrt(1. + z)), rtol=1e-10, atol=1e-9)
    lcdm = utils.FlatLCDMDistance(0.3)
    assert np.allclose(lcdm([0.5, 1.]), [1888.63 * 0.7, 3303.83 * 0.7], rtol=2e-5)
    inv = utils.DistanceToRedshift(lcdm)
    zz = np.linspace(0.01, 3., 50)
    assert np.allclose(inv(lcdm(zz)), zz, rtol=1e-6)


def test_eval_strings_and_paths():
    assert cli.decode_eval_str('(${RA}>0.) & (${DEC}<${RA})') == ('(RA>0.) & (DEC<RA)', ['RA', 'DEC'])
    assert cli.decode_eval_str(None) == ('', [])
    assert cli.sepjoin('/galaxies', 'RA') == '/galaxies/RA' and cli.sepjoin('/a/', '/b') == '/b'
    assert cli.make_list('A') == ['A'] and cli.remove_duplicates(['A', 'B', 'A']) == ['A', 'B']


def test_yaml_loader(tmp_path):
    import yaml
    fn = tmp_path / 'c.yaml'
    fn.write_text('a: 1e-5\nb: None\nc: 3\n')
    out = yaml.load(open(fn), Loader=cli._yaml_loader())
    assert out == {'a': 1e-5, 'b': None, 'c': 3}


def _write(tmp_path, n=500, seed=0):
    rng = np.random.RandomState(seed)
    cat = dict(RA=rng.uniform(0., 60., n), DEC=rng.uniform(-10., 40., n), Z=rng.uniform(0.4, 0.8, n),
               WEIGHT=rng.uniform(0.5, 1.5, n), NZ=rng.uniform(1e-4, 3e-4, n))
    np.savez(tmp_path / 'data.npz', **cat)
    pos = rng.uniform(-500., 500., (n, 3))
    np.savez(tmp_path / 'randoms.npz', POSITION=pos, X=pos[:, 0], Y=pos[:, 1], Z=pos[:, 2])
    arr = np.empty(n, dtype=[('RA', 'f8'), ('DEC', 'f8'), ('Z', 'f8')])
    for col in arr.dtype.names: arr[col] = cat[col]
    np.save(tmp_path / 'data.npy', arr)
    return cat, pos


def test_load_inputs(tmp_path):
    cat, pos = _write(tmp_path)
    config = cli.default_config()
    cli.update_config(config, {'input': {'dir': str(tmp_path), 'data_fn': 'data.npz', 'randoms_fn': 'randoms.npz',
                                         'rdz': ['RA', 'DEC', 'Z'], 'xyz_randoms': 'POSITION',
                                         'weights': '${WEIGHT} * 2', 'mask': '(${RA} > 10.) & (${DEC} < 30.)',
                                         'weights_randoms': '', 'mask_randoms': '${X} > 0.'},
                               'output': {'data_fn': 'out.npz', 'columns': ['NZ'], 'columns_randoms': ''},
                               'cosmology': {'Omega_m': 0.3}})
    sel, ifn, ofn, catalogs, positions, weights, keep = cli.load_inputs(config)
    assert sel == 'randoms' and set(ifn) == {'data', 'randoms'} and set(ofn) == {'data'}
    mask = (cat['RA'] > 10.) & (cat['DEC'] < 30.)
    dist = utils.FlatLCDMDistance(0.3)(cat['Z'][mask])
    assert np.allclose(positions['data'], utils.sky_to_cartesian(dist, cat['RA'][mask], cat['DEC'][mask]))
    assert np.array_equal(weights['data'], cat['WEIGHT'][mask] * 2)
    assert np.array_equal(catalogs['data']['NZ'], cat['NZ'][mask]) and list(catalogs['data']) == ['NZ']
    assert np.array_equal(positions['randoms'], pos[pos[:, 0] > 0.]) and weights['randoms'] is None
    # three xyz columns, structured .npy input, no randoms -> uniform selection
    config = cli.default_config()
    cli.update_config(config, {'input': {'dir': str(tmp_path), 'data_fn': 'data.npy', 'xyz': ['RA', 'DEC', 'Z']}})
    sel, ifn, ofn, catalogs, positions, weights, keep = cli.load_inputs(config)
    assert sel == 'uniform' and np.array_equal(positions['data'], np.array([cat['RA'], cat['DEC'], cat['Z']]).T)


@pytest.mark.parametrize('bad, match', [
    ({'input': {'data_fn': 'data.npz'}}, 'Unknown input format'),
    ({'input': {'data_fn': 'data.npz', 'rdz': ['RA', 'DEC', 'Z'], 'xyz': 'P'}}, 'two different input formats'),
    ({'input': {'data_fn': 'data.npz', 'rdz': ['RA', 'DEC', 'Z']}}, 'Omega_m'),
    ({'input': {'data_fn': 'data.fits', 'xyz': 'P'}}, 'fitsio'),
    ({'input': {'data_fn': 'data.txt', 'xyz': 'P'}}, 'not recognized'),
    ({'input': {'data_fn': 'data.npz', 'xyz': 'P'}, 'delta': {'selection_function': 'randoms'}}, 'randoms catalog'),
    ({'input': {'xyz': 'P'}}, 'data catalog'),
    ({'input': {'data_fn': 'data.npz', 'xyz': 'P'}, 'output': {'randoms_fn': 'r.npz'}}, 'Cannot output randoms'),
])
def test_config_errors(tmp_path, bad, match):
    _write(tmp_path)
    config = cli.default_config()
    bad = {k: dict(v) for k, v in bad.items()}
    bad.setdefault('input', {})['dir'] = str(tmp_path)
    cli.update_config(config, bad)
    with pytest.raises(cli.ConfigError, match=match):
        cli.load_inputs(config)


def test_unknown_convention(tmp_path):
    with pytest.raises(cli.ConfigError, match='convention'):
        cli.main([], input={'data_fn': str(tmp_path / 'x.npz'), 'xyz': 'P'}, algorithm={'convention': 'nope'})
