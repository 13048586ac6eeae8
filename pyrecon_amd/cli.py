"""Command-line reconstruction of catalogs (pyrecon's bin/pyrecon:70-423).

    python -m pyrecon_amd.cli config.yaml [--data-fn F] [--randoms-fn F] [--output-data-fn F]
                                          [--output-randoms-fn F] [--growth-rate f] [--galaxy-bias b]

The configuration file has the reference's sections and keys (``input``, ``output``,
``algorithm``, ``delta``, ``cosmology``, ``mesh``; bin/config_example.yaml):
- positions come from ``xyz`` (1 or 3 columns) or ``rdz`` (RA, Dec, redshift) columns, with
  ``_data`` / ``_randoms`` variants;
- ``weights`` and ``mask`` are expressions over ``${column}`` names, evaluated with numpy;
- the ``convention`` is RecSym (data and randoms shifted by 'disp+rsd'), RecIso (randoms by
  'disp') or RSD (data by 'rsd', randoms unchanged);
- outputs keep the requested input columns and add the input and reconstructed positions
  under the ``xyz``/``rdz`` and ``xyz_rec``/``rdz_rec`` names.

The mask and weight expressions are evaluated with Python's ``eval`` over numpy and the catalog
columns, as the reference does: a configuration file is trusted input.

Catalog files are numpy ``.npz`` (one array per column) or ``.npy`` (a structured array); FITS and
HDF5 files are read and written when fitsio or h5py is importable.  Redshift <-> distance uses a
flat LCDM cosmology with ``cosmology.Omega_m`` (the reference uses astropy's FlatLambdaCDM).
"""
import argparse
import logging
import os
import re
from numbers import Number

import numpy as np

from . import utils

logger = logging.getLogger('pyrecon_amd.cli')

EXTENSIONS = {'npz': ['.npz'], 'npy': ['.npy'], 'fits': ['.fits'],
              'hdf5': ['.hdf', '.h4', '.hdf4', '.he2', '.h5', '.hdf5', '.he5', '.h5py']}
ALLOWED_FORMATS = ['rdz', 'xyz']
ALLOWED_CONVENTIONS = ['recsym', 'reciso', 'rsd']
ALLOWED_SELECTION_FUNCTIONS = ['uniform', 'randoms', '']


class ConfigError(Exception):
    """Raised for an invalid configuration (bin/pyrecon:60-61)."""


def _yaml_loader():
    """yaml.SafeLoader that reads 1e-5-style floats and 'None' (bin/pyrecon:28-50)."""
    import yaml

    class Loader(yaml.SafeLoader):
        pass

    Loader.add_implicit_resolver(u'tag:yaml.org,2002:float', re.compile(u'''^(?:
        [-+]?(?:[0-9][0-9_]*)\\.[0-9_]*(?:[eE][-+]?[0-9]+)?
        |[-+]?(?:[0-9][0-9_]*)(?:[eE][-+]?[0-9]+)
        |\\.[0-9_]+(?:[eE][-+][0-9]+)?
        |[-+]?[0-9][0-9_]*(?::[0-5]?[0-9])+\\.[0-9_]*
        |[-+]?\\.(?:inf|Inf|INF)
        |\\.(?:nan|NaN|NAN))$''', re.X), list(u'-+0123456789.'))
    Loader.add_implicit_resolver('!none', re.compile('None$'), first='None')
    Loader.add_constructor('!none', lambda loader, node: None)
    return Loader


def default_config():
    return {'input': {}, 'output': {}, 'algorithm': {'name': 'MultiGridReconstruction', 'convention': 'RecSym'},
            'delta': {'smoothing_radius': 15.}, 'cosmology': {}, 'mesh': {'nmesh': 512, 'dtype': 'f4'}}


def update_config(c1, c2):
    """Update c1 in place with c2, section by section; empty values become None."""
    for section, value in c2.items():
        c1.setdefault(section, {})
        c1[section].update(value or {})
    for section in c1:
        for name, value in c1[section].items():
            if not isinstance(value, Number) and not isinstance(value, np.ndarray) and not value:
                c1[section][name] = None


def make_list(cols):
    if cols is None: return []
    if isinstance(cols, str): cols = [cols]
    return list(cols)


def remove_duplicates(cols):
    toret = []
    for col in cols:
        if col not in toret: toret.append(col)
    return toret


def sepjoin(*paths):
    """Join HDF5 dataset paths with single '/' separators."""
    sep = '/'
    toret = ''
    for path in paths:
        if path is None: continue
        if toret.endswith(sep) and path.startswith(sep):
            toret = path
        elif (not toret.endswith(sep)) and (not path.startswith(sep)):
            toret += sep + path
        else:
            toret += path
    return toret


def decode_eval_str(s):
    """'${A} > 0' -> ('A > 0', ['A'])."""
    if s is None:
        return '', []
    toret = str(s)
    columns = []
    for replace in re.finditer(r'(\${.*?})', toret):
        value = replace.group(1)
        col = value[2:-1]
        toret = toret.replace(value, col)
        if col not in columns: columns.append(col)
    return toret, columns


def _file_format(fn):
    for file_format, exts in EXTENSIONS.items():
        if any(fn.endswith(ext) for ext in exts):
            return file_format
    allowed = ', '.join(['{} ({})'.format(f, ', '.join(e)) for f, e in EXTENSIONS.items()])
    raise ConfigError('File format of {} is not recognized. Formats {} only are handled.'.format(fn, allowed))


def read_catalog(fn, columns, hdf5_prefix=''):
    file_format = _file_format(fn)
    catalog = {}
    if file_format == 'npz':
        with np.load(fn, allow_pickle=False) as file:
            for col in columns: catalog[col] = file[col]
    elif file_format == 'npy':
        array = np.load(fn, allow_pickle=False)
        for col in columns: catalog[col] = array[col]
    elif file_format == 'fits':
        try:
            import fitsio
        except ImportError as exc:
            raise ConfigError('reading {} needs fitsio, which is not installed'.format(fn)) from exc
        tmp = fitsio.read(fn, columns=columns)
        for col in columns: catalog[col] = tmp[col]
    else:
        try:
            import h5py
        except ImportError as exc:
            raise ConfigError('reading {} needs h5py, which is not installed'.format(fn)) from exc
        with h5py.File(fn, 'r') as file:
            for col in columns: catalog[col] = file[sepjoin(hdf5_prefix, col)][:]
    return catalog


def write_catalog(fn, catalog, columns, hdf5_prefix=''):
    file_format = _file_format(fn)
    if file_format == 'npz':
        np.savez(fn, **{col: catalog[col] for col in columns})
    elif file_format == 'npy':
        array = np.empty(len(catalog[columns[0]]), dtype=[(col, catalog[col].dtype, catalog[col].shape[1:])
                                                           for col in columns])
        for col in columns: array[col] = catalog[col]
        np.save(fn, array)
    elif file_format == 'fits':
        try:
            import fitsio
        except ImportError as exc:
            raise ConfigError('writing {} needs fitsio, which is not installed'.format(fn)) from exc
        array = np.empty(len(catalog[columns[0]]), dtype=[(col, catalog[col].dtype, catalog[col].shape[1:])
                                                           for col in columns])
        for col in columns: array[col] = catalog[col]
        fitsio.write(fn, array, clobber=True)
    else:
        try:
            import h5py
        except ImportError as exc:
            raise ConfigError('writing {} needs h5py, which is not installed'.format(fn)) from exc
        with h5py.File(fn, 'w') as file:
            for col in columns: file[sepjoin(hdf5_prefix, col)] = catalog[col]


def comoving_distance(config):
    Omega_m = config['cosmology'].get('Omega_m', None)
    if Omega_m is None:
        raise ConfigError('Provide Omega_m for redshift -> distance conversion.')
    return utils.FlatLCDMDistance(Omega_m)


def load_inputs(config):
    """Read the input catalogs of ``config`` (bin/pyrecon:166-337).

    Returns (selection_function, input_fns, output_fns, catalogs, positions, weights,
    keep_input_columns)."""
    selection_function = (config['delta'].pop('selection_function', '') or '').lower()
    if selection_function not in ALLOWED_SELECTION_FUNCTIONS:
        raise ConfigError('Unknown input selection function {}. Choices are {}'.format(
            selection_function, ALLOWED_SELECTION_FUNCTIONS))
    input_fns, output_fns = {}, {}
    for name in ['data', 'randoms']:
        tmp_fn = config['input'].get('{}_fn'.format(name), None)
        if tmp_fn is None:
            if name == 'randoms':
                if selection_function == 'randoms':
                    raise ConfigError('Please provide randoms catalog.')
                if not selection_function:
                    logger.info('No randoms provided.')
                    selection_function = 'uniform'
            else:
                raise ConfigError('Please provide data catalog.')
        else:
            input_fns[name] = tmp_fn
        tmp_fn = config['output'].get('{}_fn'.format(name), None)
        if tmp_fn is not None:
            if name not in input_fns:
                raise ConfigError('Cannot output {} catalog if not provided as input.'.format(name))
            output_fns[name] = tmp_fn
    if not selection_function:
        selection_function = 'randoms'
    logger.info('Using {} selection function.'.format(selection_function))
    catalogs, positions, weights, keep_input_columns = {}, {}, {}, {}
    for name in input_fns:
        fmt, position_columns = None, None
        for key in ['{}_' + name, '{}']:
            for format in ALLOWED_FORMATS:
                cols = config['input'].get(key.format(format), None)
                if cols is not None:
                    if fmt is not None:
                        raise ConfigError('Cannot use two different input formats')
                    fmt, position_columns = format, cols
            if fmt is not None: break
        if fmt is None:
            raise ConfigError('Unknown input format. Choices are {}'.format(ALLOWED_FORMATS))
        position_columns = make_list(position_columns)
        fn = os.path.join(config['input'].get('dir', None) or '', input_fns[name])
        mask_str, mask_columns = decode_eval_str(config['input'].get('mask_{}'.format(name),
                                                                     config['input'].get('mask', None)))
        weight_str, weight_columns = decode_eval_str(config['input'].get('weights_{}'.format(name),
                                                                         config['input'].get('weights', None)))
        keep_input_columns[name] = make_list(config['output'].get('columns_{}'.format(name),
                                                                  config['output'].get('columns', None)))
        columns = remove_duplicates(position_columns + weight_columns + mask_columns + keep_input_columns[name])
        logger.info('Loading {} catalog {}.'.format(name, fn))
        catalog = read_catalog(fn, columns, hdf5_prefix=config['input'].get('hdf5_prefix', None) or '')
        if mask_str:
            mask = eval(mask_str, dict(np=np, **catalog), {})
            for col in columns: catalog[col] = catalog[col][mask]
            logger.info('Size of catalog {} after masking is {:d}.'.format(name, len(catalog[columns[0]])))
        if fmt == 'rdz':
            if len(position_columns) != 3:
                raise ConfigError('Format rdz requires 3 position columns')
            dist = comoving_distance(config)(catalog[position_columns[2]])
            positions[name] = utils.sky_to_cartesian(dist, catalog[position_columns[0]], catalog[position_columns[1]])
        else:
            if len(position_columns) == 3:
                positions[name] = np.array([catalog[col] for col in position_columns]).T
            elif len(position_columns) == 1:
                positions[name] = catalog[position_columns[0]]
            else:
                raise ConfigError('Format xyz requires 1 or 3 position columns')
        weights[name] = eval(weight_str, dict(np=np, **catalog), {}) if weight_str else None
        for col in columns:
            if col not in keep_input_columns[name]:
                del catalog[col]
        catalogs[name] = catalog
    return selection_function, input_fns, output_fns, catalogs, positions, weights, keep_input_columns


def main(args=None, **input_config):
    """Run a reconstruction as configured; returns the reconstructed positions by catalog name."""
    import pyrecon_amd
    parser = argparse.ArgumentParser(formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument('config_fn', action='store', type=str, nargs='?', default=None, help='Name of configuration file')
    parser.add_argument('--data-fn', nargs='?', help='Path to input data file (overrides configuration file)')
    parser.add_argument('--randoms-fn', nargs='?', help='Path to input randoms file (overrides configuration file)')
    parser.add_argument('--output-data-fn', nargs='?', help='Path to output data file (overrides configuration file)')
    parser.add_argument('--output-randoms-fn', nargs='?', help='Path to output randoms file (overrides configuration file)')
    parser.add_argument('--growth-rate', nargs='?', type=float, help='Logarithmic growth rate (overrides configuration file)')
    parser.add_argument('--galaxy-bias', nargs='?', type=float, help='Linear galaxy bias (overrides configuration file)')
    opt = parser.parse_args(args=args)
    logging.basicConfig(level=logging.INFO, format='[%(asctime)s] %(name)s %(message)s')
    config = default_config()
    update_config(config, input_config)
    if opt.config_fn is not None:
        logger.info('Loading config file {}.'.format(opt.config_fn))
        with open(opt.config_fn, 'r') as file:
            import yaml
            update_config(config, yaml.load(file, Loader=_yaml_loader()))
    if opt.data_fn: config['input']['data_fn'] = opt.data_fn
    if opt.randoms_fn: config['input']['randoms_fn'] = opt.randoms_fn
    if opt.output_data_fn: config['output']['data_fn'] = opt.output_data_fn
    if opt.output_randoms_fn: config['output']['randoms_fn'] = opt.output_randoms_fn
    if opt.growth_rate: config['cosmology']['f'] = opt.growth_rate
    if opt.galaxy_bias: config['cosmology']['bias'] = opt.galaxy_bias
    algorithm = getattr(pyrecon_amd, config['algorithm'].pop('name'))
    config_cosmo = {name: value for name, value in config['cosmology'].items() if name in ['f', 'bias']}
    config_cosmo['los'] = config['algorithm'].pop('los', None)
    convention = (config['algorithm'].pop('convention') or '').lower()
    config['algorithm'].pop('nthreads', None)
    if convention not in ALLOWED_CONVENTIONS:
        raise ConfigError('Unknown convention {}. Choices are {}'.format(convention, ALLOWED_CONVENTIONS))
    logger.info('Convention is {}.'.format(convention))
    (selection_function, input_fns, output_fns, catalogs, positions, weights,
     keep_input_columns) = load_inputs(config)
    mesh = {name: value for name, value in config['mesh'].items() if value is not None}
    recon = algorithm(**config_cosmo, **mesh,
                      positions=positions['randoms'] if selection_function == 'randoms' else None)
    recon.assign_data(positions['data'], weights['data'])
    if selection_function == 'randoms': recon.assign_randoms(positions['randoms'], weights['randoms'])
    recon.set_density_contrast(**{k: v for k, v in config['delta'].items() if v is not None})
    recon.run(**{k: v for k, v in config['algorithm'].items() if v is not None})
    positions_rec = {}
    field = 'rsd' if convention == 'rsd' else 'disp+rsd'
    if isinstance(recon, pyrecon_amd.OriginalIterativeFFTParticleReconstruction):
        positions_rec['data'] = recon.read_shifted_positions('data', field=field)
    else:
        positions_rec['data'] = recon.read_shifted_positions(positions['data'], field=field)
    if 'randoms' in output_fns:
        if convention == 'rsd':
            positions_rec['randoms'] = positions['randoms']
        else:
            field = 'disp+rsd' if convention == 'recsym' else 'disp'
            positions_rec['randoms'] = recon.read_shifted_positions(positions['randoms'], field=field)
    for name in output_fns:
        catalog = catalogs[name]
        columns = list(keep_input_columns[name])
        for rec, positions_ in zip(['', '_rec'], [positions[name], positions_rec[name]]):
            for format in ALLOWED_FORMATS:
                position_columns = config['output'].get('{}{}_{}'.format(format, rec, name), None)
                if position_columns is None:
                    position_columns = config['output'].get('{}{}'.format(format, rec), None)
                if position_columns is None: continue
                position_columns = make_list(position_columns)
                if format == 'rdz':
                    if len(position_columns) != 3:
                        raise ConfigError('Format rdz requires 3 position columns')
                    dist, ra, dec = utils.cartesian_to_sky(positions_)
                    z = utils.DistanceToRedshift(comoving_distance(config))(dist)
                    for col, value in zip(position_columns, [ra, dec, z]):
                        catalog[col] = value
                else:
                    if len(position_columns) == 3:
                        for icol, col in enumerate(position_columns):
                            catalog[col] = positions_[:, icol]
                    elif len(position_columns) == 1:
                        catalog[position_columns[0]] = positions_
                    else:
                        raise ConfigError('Format xyz requires 1 or 3 position columns')
                columns += position_columns
        columns = remove_duplicates(columns)
        fn = os.path.join(config['output'].get('dir', None) or '', output_fns[name])
        logger.info('Saving {} catalog to {}.'.format(name, fn))
        write_catalog(fn, catalog, columns, hdf5_prefix=config['output'].get('hdf5_prefix', None) or '')
    return positions_rec


if __name__ == '__main__':
    main()
