"""Geostatistical functions (GeoStatsFunctions 0.15 restated, un-vendored):
spherical, exponential and gaussian variograms and covariances with nugget and
axis ranges (optionally rotated). C(h) = sill - gamma(h), gamma(0) = 0 and
gamma(u) = (sill - nugget) * shape(u) + nugget for u > 0 (practical ranges).
Gaussian models add a stabilising nugget of 1e-6 (sill units) to their own, as
the backend's covariance does: the reference's LUSIM testset factors Gaussian
covariance matrices that are singular in FP64 without it (test/field.jl:29-56)."""
from __future__ import annotations

import numpy as np

from . import _lib


class GeoStatsFunction:
    family = None
    kind = _lib.VARIOGRAM
    stabilising_nugget = 0.0  # added to the nugget inside gamma (gaussian: 1e-6)
    __array_ufunc__ = None  # `A * f` with a numpy matrix A goes to __rmul__

    def __init__(self, sill=1.0, range=1.0, nugget=0.0, ranges=None, rotation=None):
        self._sill = float(sill)
        self._nugget = float(nugget)
        self._ranges = tuple(float(r) for r in ranges) if ranges is not None else (float(range),)
        self.rotation = None if rotation is None else np.asarray(rotation, dtype=np.float64).reshape(3, 3)

    # trait functions (GeoStatsFunctions: isbanded, isstationary, issymmetric)
    def isbanded(self):
        return self.kind == _lib.COVARIANCE

    def isstationary(self):
        return True

    def issymmetric(self):
        return True

    def nvariables(self):
        return 1

    def sill(self):
        return self._sill

    def nugget(self):
        return self._nugget

    def ranges(self, ndim=None):
        r = self._ranges
        if ndim is None:
            return r
        if len(r) == 1:
            return r * ndim
        if len(r) != ndim:
            raise ValueError(f"function has {len(r)} ranges, domain has {ndim} dimensions")
        return r

    def range(self):
        """range(f): the largest range."""
        return max(self._ranges)

    def shape(self, u):
        raise NotImplementedError

    def gamma(self, h):
        """Variogram value at normalised lag(s) u = h / range."""
        u = np.asarray(h, dtype=np.float64) / self._ranges[0]
        n = self._nugget + self.stabilising_nugget
        g = (self._sill - n) * self.shape(u) + n
        return np.where(u > 0, g, 0.0)

    def cov(self, h):
        return self._sill - self.gamma(h)

    def struct(self, ndim) -> _lib.GspModel:
        m = _lib.GspModel()
        m.family = self.family
        m.kind = self.kind
        m.sill = self._sill
        m.nugget = self._nugget
        r = self.ranges(ndim)
        for d in range(3):
            m.ranges[d] = r[d] if d < ndim else 1.0
        if self.rotation is not None:
            for q in range(9):
                m.rotation[q] = self.rotation.flat[q]
            m.has_rotation = 1
        return m

    def __repr__(self):
        return (f"{type(self).__name__}(sill={self._sill}, nugget={self._nugget}, "
                f"ranges={self._ranges})")

    def __rmul__(self, A):
        """A * f with a symmetric matrix A: the multivariate function
        C_ij(h) = A_ij c(h) (intrinsic coregionalization)."""
        return CoregionalizedFunction(A, self)


class CoregionalizedFunction:
    """A * f: n variables sharing the spatial structure of f (GeoStatsFunctions
    composite of a matrix and a function)."""

    def __init__(self, A, base):
        A = np.atleast_2d(np.asarray(A, dtype=np.float64))
        if A.shape[0] != A.shape[1] or not np.allclose(A, A.T):
            raise ValueError("the coefficient matrix must be square and symmetric")
        self.A = A
        self.base = base

    def nvariables(self):
        return self.A.shape[0]

    def isstationary(self):
        return self.base.isstationary()

    def issymmetric(self):
        return True

    def isbanded(self):
        return self.base.isbanded()

    def range(self):
        return self.base.range()

    def ranges(self, ndim=None):
        return self.base.ranges(ndim)

    def sill(self):
        return self.A * self.base.sill()

    def __repr__(self):
        return f"{self.A.tolist()} * {self.base!r}"


class _Spherical:
    family = _lib.SPHERICAL

    def shape(self, u):
        u = np.asarray(u, dtype=np.float64)
        return np.where(u < 1.0, 1.5 * u - 0.5 * u ** 3, 1.0)


class _Exponential:
    family = _lib.EXPONENTIAL

    def shape(self, u):
        return 1.0 - np.exp(-3.0 * np.asarray(u, dtype=np.float64))


class _Gaussian:
    family = _lib.GAUSSIAN
    stabilising_nugget = 1e-6

    def shape(self, u):
        u = np.asarray(u, dtype=np.float64)
        return 1.0 - np.exp(-3.0 * u * u)


class SphericalVariogram(_Spherical, GeoStatsFunction):
    kind = _lib.VARIOGRAM


class ExponentialVariogram(_Exponential, GeoStatsFunction):
    kind = _lib.VARIOGRAM


class GaussianVariogram(_Gaussian, GeoStatsFunction):
    kind = _lib.VARIOGRAM


class SphericalCovariance(_Spherical, GeoStatsFunction):
    kind = _lib.COVARIANCE


class ExponentialCovariance(_Exponential, GeoStatsFunction):
    kind = _lib.COVARIANCE


class GaussianCovariance(_Gaussian, GeoStatsFunction):
    kind = _lib.COVARIANCE


class Transiogram:
    """Shape transiograms of GeoStatsFunctions 0.15 (un-vendored; the form is
    not fixed by any reference test): transition probabilities
    T_ij(h) = (1 - v(h)) delta_ij + v(h) p_j, where v is the family's unit-sill
    variogram shape at the normalised lag (0 at h = 0, 1 at and beyond the
    range) and p the proportions. The indicator cross-covariances are then
    p_i (T_ij(h) - p_j) = c(h) (diag(p) - p p^T)_ij with c = 1 - v: an intrinsic
    coregionalization of the unit-sill covariance c of the same family and
    range, which is how SEQSIM simulates an IndicatorProcess of a transiogram
    here (its simple-kriging weights are those of c)."""
    covariance_type = None
    __array_ufunc__ = None

    def __init__(self, range=1.0, proportions=(0.5, 0.5), ranges=None, rotation=None):
        self._ranges = tuple(float(r) for r in ranges) if ranges is not None else (float(range),)
        prop = np.asarray(proportions, dtype=np.float64)
        if prop.ndim != 1 or len(prop) < 2 or np.any(prop < 0):
            raise ValueError("proportions must be two or more non-negative values")
        self._prop = prop / prop.sum()
        self.rotation = rotation

    def proportions(self):
        return self._prop.copy()

    def nvariables(self):
        return len(self._prop)

    def isstationary(self):
        return True

    def issymmetric(self):
        return False

    def isbanded(self):
        return False

    def range(self):
        return max(self._ranges)

    def ranges(self, ndim=None):
        return self.base_covariance().ranges(ndim)

    def base_covariance(self):
        """The unit-sill covariance c = 1 - v of the same family and range."""
        return self.covariance_type(sill=1.0, nugget=0.0, ranges=self._ranges,
                                    rotation=self.rotation)

    def coregionalization(self):
        """diag(p) - p p^T: the indicator covariances are c(h) times this."""
        p = self._prop
        return np.diag(p) - np.outer(p, p)

    def transition(self, h):
        """T(h) for a scalar lag (in the units of the range)."""
        v = 1.0 - float(self.base_covariance().cov(np.asarray([h], dtype=np.float64))[0])
        p = self._prop
        return (1.0 - v) * np.eye(len(p)) + v * np.broadcast_to(p, (len(p), len(p)))

    def __repr__(self):
        return (f"{type(self).__name__}(ranges={self._ranges}, "
                f"proportions={tuple(float(x) for x in self._prop)})")


class SphericalTransiogram(Transiogram):
    covariance_type = SphericalCovariance


class ExponentialTransiogram(Transiogram):
    covariance_type = ExponentialCovariance


class GaussianTransiogram(Transiogram):
    covariance_type = GaussianCovariance
