"""Column-oriented catalog files for the command line.

A catalog is a dict ``{column name: numpy array}`` whose arrays share their first dimension.  Each
file type is a :class:`CatalogFormat` with a reader of a list of columns and a writer of a catalog,
chosen by the file's extension:

- ``.npz``: one array per column (numpy);
- ``.npy``: one structured array, a column per field (numpy);
- FITS (``.fits``) through fitsio, and HDF5 (``.h5``, ``.hdf5`` and the other HDF extensions of
  bin/pyrecon:56-58) through h5py, with an optional group prefix; these two modules are optional
  and only imported for such files.
"""
import os

import numpy as np


class CatalogFormatError(ValueError):
    """The file type is unknown or cannot be handled here."""


class CatalogFormat(object):
    """One file type: ``read(path, columns, prefix)`` and ``write(path, catalog, prefix)``."""

    def __init__(self, name, extensions, read, write, module=None):
        self.name, self.extensions = name, tuple(extensions)
        self._read, self._write, self.module = read, write, module

    def matches(self, path):
        return str(path).lower().endswith(self.extensions)

    def _check_module(self):
        if self.module is not None:
            try:
                __import__(self.module)
            except ImportError as exc:
                raise CatalogFormatError('{} files need the "{}" module, which is not installed'.format(
                    self.name, self.module)) from exc

    def read(self, path, columns, prefix=''):
        self._check_module()
        catalog = self._read(path, list(columns), prefix)
        missing = [c for c in columns if c not in catalog]
        if missing:
            raise KeyError('columns {} not found in {}'.format(missing, path))
        sizes = {len(catalog[c]) for c in columns}
        if len(sizes) > 1:
            raise ValueError('columns of {} differ in length'.format(path))
        return catalog

    def write(self, path, catalog, prefix=''):
        self._check_module()
        self._write(path, catalog, prefix)


def hdf5_path(prefix, column):
    """Dataset path of ``column`` under the group ``prefix``; a column that starts with '/' is absolute."""
    column = str(column)
    if column.startswith('/'):
        return column
    parts = [p for p in str(prefix or '').split('/') if p] + [p for p in column.split('/') if p]
    return '/' + '/'.join(parts)


def _read_npz(path, columns, prefix):
    with np.load(path, allow_pickle=False) as archive:
        return {c: archive[c] for c in columns if c in archive.files}


def _write_npz(path, catalog, prefix):
    with open(path, 'wb') as fh:
        np.savez(fh, **catalog)


def _read_npy(path, columns, prefix):
    array = np.load(path, mmap_mode='r', allow_pickle=False)
    if array.dtype.names is None:
        raise CatalogFormatError('{}: a .npy catalog holds a structured array'.format(path))
    return {c: np.array(array[c]) for c in columns if c in array.dtype.names}


def _structured(catalog):
    names = list(catalog)
    size = len(catalog[names[0]]) if names else 0
    out = np.empty(size, dtype=[(n, catalog[n].dtype, catalog[n].shape[1:]) for n in names])
    for n in names:
        out[n] = catalog[n]
    return out


def _write_npy(path, catalog, prefix):
    np.save(path, _structured(catalog), allow_pickle=False)


def _read_fits(path, columns, prefix):
    import fitsio
    table = fitsio.read(path, columns=columns)
    return {c: np.asarray(table[c]) for c in columns}


def _write_fits(path, catalog, prefix):
    import fitsio
    fitsio.write(path, _structured(catalog), clobber=True)


def _read_hdf5(path, columns, prefix):
    import h5py
    with h5py.File(path, 'r') as fh:
        return {c: fh[hdf5_path(prefix, c)][...] for c in columns}


def _write_hdf5(path, catalog, prefix):
    import h5py
    with h5py.File(path, 'w') as fh:
        for c, value in catalog.items():
            fh[hdf5_path(prefix, c)] = value


FORMATS = [
    CatalogFormat('numpy archive', ['.npz'], _read_npz, _write_npz),
    CatalogFormat('numpy structured array', ['.npy'], _read_npy, _write_npy),
    CatalogFormat('FITS', ['.fits'], _read_fits, _write_fits, module='fitsio'),
    CatalogFormat('HDF5', ['.hdf', '.h4', '.hdf4', '.he2', '.h5', '.hdf5', '.he5', '.h5py'], _read_hdf5, _write_hdf5,
                  module='h5py'),
]


def format_of(path):
    for fmt in FORMATS:
        if fmt.matches(path):
            return fmt
    known = '; '.join('{} ({})'.format(f.name, ', '.join(f.extensions)) for f in FORMATS)
    raise CatalogFormatError('{}: file type not recognized; known types are {}'.format(os.path.basename(str(path)),
                                                                                      known))


def read_catalog(path, columns, prefix=''):
    """``{column: array}`` of the requested columns of a catalog file."""
    return format_of(path).read(path, columns, prefix)


def write_catalog(path, catalog, prefix=''):
    """Write ``catalog`` (columns in insertion order) to ``path``, typed by its extension."""
    format_of(path).write(path, catalog, prefix)
