"""Coordinates and distances for the catalog formats of the command line.

Host-side numpy.  The public names and arguments are those of pyrecon's utilities
(``distance``, ``cartesian_to_sky``, ``sky_to_cartesian``, ``DistanceToRedshift``;
pyrecon/utils.py:146-282), so scripts written against them keep working:

- spherical <-> Cartesian conversion of (distance, right ascension, declination), angles in
  degrees by default, RA folded into [0, 360);
- the comoving distance of a flat LCDM cosmology without radiation in Mpc/h
  (:class:`FlatLCDMDistance`), which the reference's command line takes from astropy's
  ``FlatLambdaCDM(H0=70, Om0=Omega_m)`` (bin/pyrecon:136-147); astropy is not a dependency here;
- the inverse of any increasing distance-redshift relation (:class:`DistanceToRedshift`).
"""
import numpy as np

# speed of light in km/s; c / (100 km/s/Mpc) is the Hubble distance in Mpc/h
C_KMS = 299792.458


def distance(position):
    """Euclidean norm along the last axis."""
    return np.sqrt(np.sum(np.square(position), axis=-1))


def cartesian_to_sky(position, wrap=True, degree=True):
    """``(distance, ra, dec)`` of Cartesian positions of shape (N, 3).

    ``ra`` is the azimuth of (x, y) measured from the x axis towards y, folded into [0, 2 pi)
    when ``wrap``; ``dec`` is the elevation above the (x, y) plane.  Angles are in degrees when
    ``degree``, radians otherwise.
    """
    position = np.asarray(position)
    x, y, z = position[..., 0], position[..., 1], position[..., 2]
    rho2 = x * x + y * y
    dist = np.sqrt(rho2 + z * z)
    ra = np.arctan2(y, x)
    if wrap:
        ra = np.where(ra < 0., ra + 2. * np.pi, ra)
        ra = np.where(ra >= 2. * np.pi, ra - 2. * np.pi, ra)  # -0. and rounding at 2 pi
    dec = np.arctan2(z, np.sqrt(rho2))
    if degree:
        ra, dec = np.degrees(ra), np.degrees(dec)
    return dist, ra, dec


def sky_to_cartesian(dist, ra, dec, degree=True, dtype=None):
    """Cartesian positions (N, 3) of distances and sky angles (degrees when ``degree``)."""
    ra, dec = np.asarray(ra, dtype='f8'), np.asarray(dec, dtype='f8')
    if degree:
        ra, dec = np.radians(ra), np.radians(dec)
    cos_dec = np.cos(dec)
    unit = np.stack([cos_dec * np.cos(ra), cos_dec * np.sin(ra), np.sin(dec)], axis=-1)
    out = np.asarray(dist)[..., None] * unit
    return out.astype(dtype, copy=False) if dtype is not None else out


class FlatLCDMDistance(object):
    """Comoving distance (Mpc/h) of a flat LCDM cosmology without radiation:
    D(z) = (c / 100) int_0^z dz' / E(z'), E(z) = sqrt(Omega_m (1 + z)^3 + 1 - Omega_m).

    The integral is tabulated once by 64-point Gauss-Legendre quadrature on each interval of a grid
    uniform in log(1 + z) up to ``zmax``, cumulated, and interpolated by a cubic spline.
    """

    def __init__(self, Omega_m, zmax=100., nz=4096):
        from scipy import interpolate
        self.Omega_m = float(Omega_m)
        self.zmax = float(zmax)
        edges = np.expm1(np.linspace(0., np.log1p(self.zmax), int(nz) + 1))
        edges[-1] = self.zmax  # expm1(log1p(zmax)) may round above zmax
        nodes, wts = np.polynomial.legendre.leggauss(64)
        lo, hi = edges[:-1], edges[1:]
        z = 0.5 * (lo + hi)[:, None] + 0.5 * (hi - lo)[:, None] * nodes[None, :]
        pieces = 0.5 * (hi - lo) * (wts[None, :] / self.efunc(z)).sum(axis=-1)
        self.zgrid = edges
        self.dgrid = (C_KMS / 100.) * np.concatenate([[0.], np.cumsum(pieces)])
        self._spline = interpolate.CubicSpline(self.zgrid, self.dgrid)

    def efunc(self, z):
        return np.sqrt(self.Omega_m * (1. + z)**3 + (1. - self.Omega_m))

    def derivative(self, z):
        """dD/dz = (c / 100) / E(z)."""
        return (C_KMS / 100.) / self.efunc(np.asarray(z, dtype='f8'))

    def __call__(self, z):
        z = np.asarray(z, dtype='f8')
        if np.any(z < 0.) or np.any(z > self.zmax):
            raise ValueError('redshift outside [0, {:g}]'.format(self.zmax))
        return self._spline(z)


class DistanceToRedshift(object):
    """Inverse of an increasing distance-redshift relation ``distance(z)``.

    The relation is sampled on a grid uniform in log(1 + z) over [0, ``zmax``] with ``nz``
    intervals, and the redshift of a distance is read from a monotone (PCHIP) interpolant of the
    table; when the relation has a ``derivative`` method, two Newton steps on distance(z) = d
    then polish it.  Distances outside the table raise ``ValueError``.
    """

    def __init__(self, distance, zmax=100., nz=2048, interp_order=3):
        from scipy import interpolate
        self.distance = distance
        self.zmax = float(zmax)
        self.nz = int(nz)
        self.zgrid = np.expm1(np.linspace(0., np.log1p(self.zmax), self.nz + 1))
        self.zgrid[-1] = self.zmax
        self.rgrid = np.asarray(distance(self.zgrid), dtype='f8')
        if not np.all(np.diff(self.rgrid) > 0.):
            raise ValueError('distance(z) must increase with z')
        self._inverse = interpolate.PchipInterpolator(self.rgrid, self.zgrid, extrapolate=False)

    def __call__(self, distance):
        d = np.asarray(distance, dtype='f8')
        if np.any(d < self.rgrid[0]) or np.any(d > self.rgrid[-1]):
            raise ValueError('distance outside [{:g}, {:g}]'.format(self.rgrid[0], self.rgrid[-1]))
        z = self._inverse(d)
        derivative = getattr(self.distance, 'derivative', None)
        if derivative is not None:
            for _ in range(2):
                z = np.clip(z - (self.distance(z) - d) / derivative(z), 0., self.zmax)
        return z
