"""Command-line reconstruction of catalog files.

    python -m pyrecon_amd.cli [config.yaml] [--data-fn F] [--randoms-fn F] [--output-data-fn F]
                              [--output-randoms-fn F] [--growth-rate f] [--galaxy-bias b]

The configuration file has the sections and keys of pyrecon's command line (bin/pyrecon:70-423,
bin/config_example.yaml), so its configuration files run unchanged:

- ``input``: ``data_fn``, ``randoms_fn`` and ``dir``; positions as ``xyz`` (one (N, 3) column or
  three columns) or ``rdz`` (RA, Dec, redshift) columns, each key optionally suffixed ``_data`` /
  ``_randoms``; ``weights`` and ``mask`` expressions over ``${column}`` names (suffixable too);
  ``hdf5_prefix``.
- ``output``: ``data_fn``, ``randoms_fn``, ``dir``; ``columns`` (suffixable) to carry over from the
  input; ``xyz`` / ``rdz`` columns for the input positions and ``xyz_rec`` / ``rdz_rec`` for the
  reconstructed ones (suffixable); ``hdf5_prefix``.
- ``algorithm``: ``name`` (a reconstruction class of this package), ``convention`` (RecSym: data
  and randoms shifted by 'disp+rsd'; RecIso: randoms by 'disp'; RSD: data by 'rsd', randoms
  unchanged), ``los``, ``nthreads``; any other key is passed to ``run``.
- ``delta``: ``selection_function`` ('randoms' or 'uniform'; by default 'randoms' when a randoms
  file is given), the rest is passed to ``set_density_contrast``.
- ``cosmology``: ``f``, ``bias``, and ``Omega_m`` for redshift <-> distance.
- ``mesh``: constructor keywords (``nmesh``, ``boxsize``, ``boxcenter``, ``dtype``, ``wrap``, ...).

Settings are layered: built-in defaults, then keyword arguments of :func:`main`, then the file,
then the command-line options.  An empty value (empty string, list or mapping) means "not set".
Where a suffixed key is present it replaces the plain one, even when it is empty (an empty
``weights_randoms`` gives unit weights to the randoms whatever ``weights`` says).

Mask and weight expressions are Python expressions over numpy (``np``) and the catalog columns,
evaluated with ``eval``: a configuration file is trusted input, as it is for the reference.
Catalog files are read and written by :mod:`pyrecon_amd.catalog_io`.
"""
import argparse
import copy
import logging
import os
import re

import numpy as np

from . import catalog_io, utils

logger = logging.getLogger('pyrecon_amd.cli')

ROLES = ('data', 'randoms')
POSITION_KINDS = ('rdz', 'xyz')
# convention -> (field of the data shifts, field of the randoms shifts; None: randoms not moved)
CONVENTIONS = {'recsym': ('disp+rsd', 'disp+rsd'), 'reciso': ('disp+rsd', 'disp'), 'rsd': ('rsd', None)}
SELECTION_FUNCTIONS = ('randoms', 'uniform')
DEFAULTS = {'input': {}, 'output': {},
            'algorithm': {'name': 'MultiGridReconstruction', 'convention': 'RecSym'},
            'delta': {'smoothing_radius': 15.}, 'cosmology': {}, 'mesh': {'nmesh': 512, 'dtype': 'f4'}}


class ConfigError(Exception):
    """The configuration cannot be run (the reference raises ``ConfigError`` for the same cases)."""


_FLOAT_TEXT = re.compile(r'^[-+]?(\d[\d_]*\.?[\d_]*|\.\d[\d_]*)([eE][-+]?\d+)?$')


def normalize(value):
    """A configuration value as the run uses it: empty -> None; text spelling a number (YAML 1.1
    reads '1e-5' as text) -> float; 'None' / 'none' -> None."""
    if isinstance(value, str):
        text = value.strip()
        if text in ('', 'None', 'none'):
            return None
        if _FLOAT_TEXT.match(text) and ('.' in text or 'e' in text.lower()):
            return float(text.replace('_', ''))
        return value
    if isinstance(value, (list, tuple, dict)) and len(value) == 0:
        return None
    return value


class Settings(object):
    """Configuration sections ``{section: {key: value}}`` merged from successive layers."""

    def __init__(self, *layers):
        self.sections = copy.deepcopy(DEFAULTS)
        for layer in layers:
            self.merge(layer)

    def merge(self, layer):
        for section, entries in (layer or {}).items():
            target = self.sections.setdefault(section, {})
            for key, value in (entries or {}).items():
                target[key] = normalize(value)
        return self

    def section(self, name):
        return self.sections.setdefault(name, {})

    def get(self, section, key, default=None):
        """The value of ``key`` if set (not None), else ``default``."""
        value = self.section(section).get(key)
        return default if value is None else value

    def suffixed(self, section, key, role):
        """``key_role`` if that key is present (even empty), else ``key``."""
        entries = self.section(section)
        specific = '{}_{}'.format(key, role)
        return entries[specific] if specific in entries else entries.get(key)

    def pop(self, section, key, default=None):
        value = self.section(section).pop(key, None)
        return default if value is None else value


def as_columns(value):
    """Column names as a list: None -> [], a name -> [name]."""
    if value is None:
        return []
    if isinstance(value, str):
        return [value]
    return [str(v) for v in value]


def unique(names):
    return list(dict.fromkeys(names))


class ColumnExpression(object):
    """A mask or weight expression over ``${column}`` names.

    Each ``${...}`` is replaced by a private variable bound to that column when evaluated, so any
    column name works, including ones that are not Python identifiers.
    """

    _REF = re.compile(r'\$\{([^}]*)\}')

    def __init__(self, text):
        self.text = '' if text is None else str(text)
        self.columns = unique(m.group(1) for m in self._REF.finditer(self.text))
        slot = {name: '_col{:d}'.format(i) for i, name in enumerate(self.columns)}
        self.code = self._REF.sub(lambda m: slot[m.group(1)], self.text)
        self._slots = slot

    def __bool__(self):
        return bool(self.text.strip())

    def evaluate(self, catalog):
        namespace = {'np': np}
        namespace.update({slot: catalog[name] for name, slot in self._slots.items()})
        return eval(self.code, namespace, {})


class Cosmology(object):
    """Redshift <-> comoving distance (Mpc/h) from ``cosmology.Omega_m``, built when first needed."""

    def __init__(self, Omega_m):
        self.Omega_m = Omega_m
        self._distance = self._redshift = None

    def distance(self, z):
        if self._distance is None:
            if self.Omega_m is None:
                raise ConfigError('rdz positions need cosmology.Omega_m (redshift <-> distance)')
            self._distance = utils.FlatLCDMDistance(self.Omega_m)
        return self._distance(z)

    def redshift(self, dist):
        if self._redshift is None:
            self.distance(0.)
            self._redshift = utils.DistanceToRedshift(self._distance)
        return self._redshift(dist)


class PositionColumns(object):
    """Where a catalog keeps its positions: ``kind`` 'xyz' (one (N, 3) column or three columns) or
    'rdz' (RA and Dec in degrees, redshift)."""

    def __init__(self, kind, columns):
        self.kind, self.columns = kind, as_columns(columns)
        if kind == 'rdz' and len(self.columns) != 3:
            raise ConfigError('rdz positions take 3 columns (RA, Dec, redshift), got {}'.format(self.columns))
        if kind == 'xyz' and len(self.columns) not in (1, 3):
            raise ConfigError('xyz positions take 1 column of shape (N, 3) or 3 columns, got {}'.format(self.columns))

    def to_cartesian(self, catalog, cosmology):
        cols = [catalog[c] for c in self.columns]
        if self.kind == 'rdz':
            return utils.sky_to_cartesian(cosmology.distance(cols[2]), cols[0], cols[1])
        if len(cols) == 1:
            return np.asarray(cols[0])
        return np.stack(cols, axis=-1)

    def from_cartesian(self, positions, cosmology):
        if self.kind == 'rdz':
            dist, ra, dec = utils.cartesian_to_sky(positions)
            return dict(zip(self.columns, (ra, dec, cosmology.redshift(dist))))
        if len(self.columns) == 1:
            return {self.columns[0]: positions}
        return {c: positions[:, i] for i, c in enumerate(self.columns)}


def input_positions(settings, role):
    """The input :class:`PositionColumns` of ``role``: a role-suffixed key wins over the plain ones."""
    for keys in (['{}_{}'.format(kind, role) for kind in POSITION_KINDS], list(POSITION_KINDS)):
        found = [(key.split('_')[0], settings.get('input', key)) for key in keys]
        found = [(kind, cols) for kind, cols in found if cols is not None]
        if len(found) > 1:
            raise ConfigError('{} positions are given in two formats ({})'.format(role, ', '.join(k for k, _ in found)))
        if found:
            return PositionColumns(*found[0])
    raise ConfigError('no position columns for the {}: give one of {}'.format(
        role, ', '.join('{0} or {0}_{1}'.format(k, role) for k in POSITION_KINDS)))


def output_positions(settings, role, reconstructed):
    """The :class:`PositionColumns` under which ``role``'s input (or reconstructed) positions are written."""
    tag = '_rec' if reconstructed else ''
    out = []
    for kind in POSITION_KINDS:
        cols = settings.get('output', '{}{}_{}'.format(kind, tag, role))
        if cols is None:
            cols = settings.get('output', kind + tag)
        if cols is not None:
            out.append(PositionColumns(kind, cols))
    return out


class CatalogPlan(object):
    """How one catalog (data or randoms) is read: file, positions, mask, weights, kept columns."""

    def __init__(self, settings, role, path):
        self.role, self.path = role, path
        self.positions = input_positions(settings, role)
        self.mask = ColumnExpression(settings.suffixed('input', 'mask', role))
        self.weights = ColumnExpression(settings.suffixed('input', 'weights', role))
        self.keep = as_columns(settings.suffixed('output', 'columns', role))
        self.prefix = settings.get('input', 'hdf5_prefix', '')

    def columns(self):
        return unique(self.positions.columns + self.weights.columns + self.mask.columns + self.keep)

    def load(self, cosmology):
        """``(kept columns, positions (N, 3), weights or None)`` after the mask."""
        names = self.columns()
        try:
            catalog = catalog_io.read_catalog(self.path, names, prefix=self.prefix)
        except catalog_io.CatalogFormatError as exc:
            raise ConfigError(str(exc)) from exc
        logger.info('%s catalog %s: %d rows', self.role, self.path, len(catalog[names[0]]))
        if self.mask:
            selected = np.asarray(self.mask.evaluate(catalog))
            catalog = {name: value[selected] for name, value in catalog.items()}
            logger.info('%s catalog after the mask: %d rows', self.role, len(catalog[names[0]]))
        positions = self.positions.to_cartesian(catalog, cosmology)
        weights = self.weights.evaluate(catalog) if self.weights else None
        kept = {name: catalog[name] for name in self.keep}
        return kept, positions, weights


def settings_from(args=None, **input_config):
    """Parse the command line and layer defaults, ``input_config``, the file and the options."""
    parser = argparse.ArgumentParser(prog='python -m pyrecon_amd.cli',
                                     description='Reconstruct catalog files on the GPU.',
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument('config_fn', nargs='?', default=None, help='YAML configuration file')
    parser.add_argument('--data-fn', help='input data file (overrides the configuration)')
    parser.add_argument('--randoms-fn', help='input randoms file (overrides the configuration)')
    parser.add_argument('--output-data-fn', help='output data file (overrides the configuration)')
    parser.add_argument('--output-randoms-fn', help='output randoms file (overrides the configuration)')
    parser.add_argument('--growth-rate', type=float, help='logarithmic growth rate f (overrides the configuration)')
    parser.add_argument('--galaxy-bias', type=float, help='linear galaxy bias (overrides the configuration)')
    opt = parser.parse_args(args=args)
    settings = Settings(input_config)
    if opt.config_fn is not None:
        import yaml
        logger.info('Reading configuration %s', opt.config_fn)
        with open(opt.config_fn, 'r') as fh:
            settings.merge(yaml.safe_load(fh))
    overrides = [('input', 'data_fn', opt.data_fn), ('input', 'randoms_fn', opt.randoms_fn),
                 ('output', 'data_fn', opt.output_data_fn), ('output', 'randoms_fn', opt.output_randoms_fn),
                 ('cosmology', 'f', opt.growth_rate), ('cosmology', 'bias', opt.galaxy_bias)]
    for section, key, value in overrides:
        if value is not None:
            settings.section(section)[key] = value
    return settings


def plan(settings):
    """Resolve the catalogs of a run: ``(selection, {role: CatalogPlan}, {role: output path})``."""
    selection = str(settings.pop('delta', 'selection_function', '')).lower()
    if selection and selection not in SELECTION_FUNCTIONS:
        raise ConfigError('selection_function must be one of {}, not "{}"'.format(SELECTION_FUNCTIONS, selection))
    paths = {role: settings.get('input', role + '_fn') for role in ROLES}
    if paths['data'] is None:
        raise ConfigError('no data catalog: set input.data_fn')
    if paths['randoms'] is None:
        if selection == 'randoms':
            raise ConfigError('selection_function "randoms" needs a randoms catalog (input.randoms_fn)')
        selection = 'uniform'
    elif not selection:
        selection = 'randoms'
    outputs = {}
    for role in ROLES:
        path = settings.get('output', role + '_fn')
        if path is None:
            continue
        if paths[role] is None:
            raise ConfigError('cannot write a {0} catalog without an input {0} catalog'.format(role))
        outputs[role] = os.path.join(settings.get('output', 'dir', ''), path)
    in_dir = settings.get('input', 'dir', '')
    plans = {role: CatalogPlan(settings, role, os.path.join(in_dir, path)) for role, path in paths.items()
             if path is not None}
    return selection, plans, outputs


def reconstruct(settings, selection, inputs):
    """Run the configured algorithm on the loaded ``inputs`` ``{role: (kept, positions, weights)}``;
    returns ``{role: reconstructed positions}`` of the roles the convention moves or keeps."""
    import pyrecon_amd
    name = settings.pop('algorithm', 'name')
    cls = getattr(pyrecon_amd, str(name), None)
    if not (isinstance(cls, type) and hasattr(cls, 'assign_data') and hasattr(cls, 'run')):
        raise ConfigError('unknown reconstruction algorithm "{}"'.format(name))
    convention = str(settings.pop('algorithm', 'convention', '')).lower()
    if convention not in CONVENTIONS:
        raise ConfigError('convention must be one of RecSym, RecIso, RSD, not "{}"'.format(convention))
    los = settings.pop('algorithm', 'los')
    nthreads = settings.pop('algorithm', 'nthreads')
    cosmo = {k: settings.get('cosmology', k) for k in ('f', 'bias') if settings.get('cosmology', k) is not None}
    mesh = {k: v for k, v in settings.section('mesh').items() if v is not None}
    use_randoms = selection == 'randoms'
    box_from = inputs['randoms'][1] if use_randoms else None
    recon = cls(los=los, positions=box_from, nthreads=nthreads, **cosmo, **mesh)
    recon.assign_data(inputs['data'][1], inputs['data'][2])
    if use_randoms:
        recon.assign_randoms(inputs['randoms'][1], inputs['randoms'][2])
    recon.set_density_contrast(**{k: v for k, v in settings.section('delta').items() if v is not None})
    recon.run(**{k: v for k, v in settings.section('algorithm').items() if v is not None})
    data_field, randoms_field = CONVENTIONS[convention]
    # the particle-level iterative algorithm returns its own iterated data positions for 'data'
    particle = type(recon) is getattr(pyrecon_amd, 'IterativeFFTParticleReconstruction', None)
    out = {'data': recon.read_shifted_positions('data' if particle else inputs['data'][1], field=data_field)}
    if 'randoms' in inputs:
        if randoms_field is None:
            out['randoms'] = inputs['randoms'][1]
        else:
            out['randoms'] = recon.read_shifted_positions(inputs['randoms'][1], field=randoms_field)
    return {role: np.asarray(x) for role, x in out.items()}


def main(args=None, **input_config):
    """Run the command line; ``input_config`` (``{section: {key: value}}``) is layered under the file."""
    settings = settings_from(args, **input_config)
    selection, plans, outputs = plan(settings)
    logger.info('Selection function: %s', selection)
    cosmology = Cosmology(settings.get('cosmology', 'Omega_m'))
    inputs = {role: p.load(cosmology) for role, p in plans.items()}
    rec = reconstruct(settings, selection, inputs)
    prefix = settings.get('output', 'hdf5_prefix', '')
    for role, path in outputs.items():
        catalog = dict(inputs[role][0])
        for reconstructed, positions in ((False, inputs[role][1]), (True, rec[role])):
            for spec in output_positions(settings, role, reconstructed):
                catalog.update(spec.from_cartesian(positions, cosmology))
        logger.info('Writing %s catalog %s', role, path)
        try:
            catalog_io.write_catalog(path, catalog, prefix=prefix)
        except catalog_io.CatalogFormatError as exc:
            raise ConfigError(str(exc)) from exc
    return rec


if __name__ == '__main__':
    logging.basicConfig(level=logging.INFO, format='%(asctime)s %(name)s %(levelname)s %(message)s')
    main()
