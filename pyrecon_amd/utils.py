"""Coordinate and distance utilities of the catalog formats (pyrecon/utils.py:125-282).

Host-side numpy: the conversions between (distance, RA, Dec) and Cartesian positions, the
distance -> redshift inversion by interpolation, and the comoving distance of a flat LCDM
cosmology without radiation in Mpc/h, which the reference's command line takes from
astropy's ``FlatLambdaCDM(H0=70, Om0=Omega_m)`` (bin/pyrecon:124-135).
"""
import numpy as np

# speed of light in km/s; c / (100 h km/s/Mpc) is the Hubble distance in Mpc/h
C_KMS = 299792.458


def distance(position):
    """Cartesian distance along the last axis (pyrecon/utils.py:146-148)."""
    return np.sqrt((position**2).sum(axis=-1))


def cartesian_to_sky(position, wrap=True, degree=True):
    """Distance, RA, Dec of Cartesian positions (pyrecon/utils.py:151-182)."""
    dist = distance(position)
    ra = np.arctan2(position[:, 1], position[:, 0])
    if wrap: ra %= 2. * np.pi
    dec = np.arcsin(position[:, 2] / dist)
    conversion = np.pi / 180. if degree else 1.
    return dist, ra / conversion, dec / conversion


def sky_to_cartesian(dist, ra, dec, degree=True, dtype=None):
    """Cartesian positions (N, 3) of distance, RA, Dec (pyrecon/utils.py:185-217)."""
    conversion = np.pi / 180. if degree else 1.
    position = [None] * 3
    cos_dec = np.cos(dec * conversion)
    position[0] = cos_dec * np.cos(ra * conversion)
    position[1] = cos_dec * np.sin(ra * conversion)
    position[2] = np.sin(dec * conversion)
    return (dist * np.asarray(position, dtype=dtype)).T


class DistanceToRedshift(object):
    """Distance -> redshift by interpolating a redshift -> distance table on a logarithmic grid
    (pyrecon/utils.py:220-282)."""

    def __init__(self, distance, zmax=100., nz=2048, interp_order=3):
        from scipy import interpolate
        self.distance = distance
        self.zmax = zmax
        self.nz = nz
        zgrid = np.logspace(-8, np.log10(self.zmax), self.nz)
        self.zgrid = np.concatenate([[0.], zgrid])
        self.rgrid = self.distance(self.zgrid)
        self.interp = interpolate.UnivariateSpline(self.rgrid, self.zgrid, k=interp_order, s=0, ext='raise')

    def __call__(self, distance):
        return self.interp(distance)


class FlatLCDMDistance(object):
    """Comoving distance (Mpc/h) of a flat LCDM cosmology without radiation:
    D(z) = (c / 100) int_0^z dz' / sqrt(Omega_m (1 + z')^3 + 1 - Omega_m), by Gauss-Legendre quadrature
    (64 nodes) on a fine grid of z, cumulated and interpolated by cubic splines."""

    def __init__(self, Omega_m, zmax=100., nz=4096):
        from scipy import interpolate
        self.Omega_m = float(Omega_m)
        edges = np.concatenate([[0.], np.logspace(-6, np.log10(zmax), nz)])
        x, w = np.polynomial.legendre.leggauss(64)
        lo, hi = edges[:-1], edges[1:]
        mid, half = (lo + hi) / 2., (hi - lo) / 2.
        z = mid[:, None] + half[:, None] * x[None, :]
        seg = half * np.sum(w[None, :] / self.efunc(z), axis=-1)
        self.zgrid = edges
        self.dgrid = (C_KMS / 100.) * np.concatenate([[0.], np.cumsum(seg)])
        self._interp = interpolate.CubicSpline(self.zgrid, self.dgrid)

    def efunc(self, z):
        return np.sqrt(self.Omega_m * (1. + z)**3 + (1. - self.Omega_m))

    def __call__(self, z):
        return self._interp(np.asarray(z, dtype='f8'))
