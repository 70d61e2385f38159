"""Domains the backend simulates on: Meshes' CartesianGrid (and PointSet for
initialization only). Linear indices are column-major (x fastest), like
`LinearIndices(size(grid))` in the reference's tests (test/field.jl:85)."""
from __future__ import annotations

import numpy as np


class CartesianGrid:
    """CartesianGrid(dims...) / CartesianGrid(start, finish, dims=...) /
    CartesianGrid(dims, origin=..., spacing=...).

    Cell centroid of 0-based index (i, j, k): origin + ((i, j, k) + 1/2) * spacing.
    """

    def __init__(self, *args, dims=None, origin=None, spacing=None):
        if len(args) == 2 and dims is not None:
            start = np.asarray(args[0], dtype=np.float64).ravel()
            finish = np.asarray(args[1], dtype=np.float64).ravel()
            dims = tuple(int(d) for d in np.atleast_1d(dims))
            origin = start
            spacing = (finish - start) / np.asarray(dims, dtype=np.float64)
        else:
            if dims is None:
                if len(args) == 1 and np.ndim(args[0]) > 0:
                    dims = tuple(int(d) for d in args[0])
                else:
                    dims = tuple(int(d) for d in args)
            dims = tuple(int(d) for d in np.atleast_1d(dims))
        nd = len(dims)
        if not 1 <= nd <= 3:
            raise ValueError("grids of 1 to 3 dimensions are supported")
        self.dims = dims
        self.origin = np.zeros(nd) if origin is None else np.asarray(origin, dtype=np.float64).reshape(nd)
        self.spacing = np.ones(nd) if spacing is None else np.asarray(spacing, dtype=np.float64).reshape(nd)

    @property
    def ndim(self):
        return len(self.dims)

    def nelements(self):
        return int(np.prod(self.dims))

    def __len__(self):
        return self.nelements()

    def size(self):
        return self.dims

    def linear_index(self, *ijk):
        """0-based column-major linear index of 0-based (i, j, k)."""
        lin, stride = 0, 1
        for d, i in enumerate(ijk):
            lin += int(i) * stride
            stride *= self.dims[d]
        return lin

    def centroids(self) -> np.ndarray:
        idx = np.indices(self.dims[::-1]).reshape(self.ndim, -1)[::-1].T
        return self.origin + (idx + 0.5) * self.spacing

    def centroid(self, i: int) -> np.ndarray:
        ijk = np.unravel_index(int(i), self.dims, order="F")
        return self.origin + (np.asarray(ijk) + 0.5) * self.spacing

    def boundingbox(self):
        return self.origin.copy(), self.origin + np.asarray(self.dims) * self.spacing

    def __eq__(self, other):
        return (isinstance(other, CartesianGrid) and self.dims == other.dims
                and np.array_equal(self.origin, other.origin)
                and np.array_equal(self.spacing, other.spacing))

    def __repr__(self):
        return f"CartesianGrid(dims={self.dims}, origin={tuple(self.origin)}, spacing={tuple(self.spacing)})"


class PointSet:
    def __init__(self, points):
        self.points = np.atleast_2d(np.asarray(points, dtype=np.float64))

    @property
    def ndim(self):
        return self.points.shape[1]

    def nelements(self):
        return len(self.points)

    def __len__(self):
        return self.nelements()

    def centroids(self) -> np.ndarray:
        return self.points

    def centroid(self, i: int) -> np.ndarray:
        return self.points[int(i)]


class GeoTable:
    """Minimal geotable: named value columns over a domain (georef)."""

    def __init__(self, values: dict, domain):
        self.values = {k: np.asarray(v) for k, v in values.items()}
        self.domain = domain
        for k, v in self.values.items():
            if len(v) != domain.nelements():
                raise ValueError(f"column {k} has {len(v)} rows, domain has {domain.nelements()}")

    def __getattr__(self, name):
        vals = self.__dict__.get("values", {})
        if name in vals:
            return vals[name]
        raise AttributeError(name)

    def __getitem__(self, name):
        return self.values[name]

    def nrow(self):
        return self.domain.nelements()

    def names(self):
        return tuple(self.values)


def georef(values: dict, domain) -> GeoTable:
    """georef(table, domain); a list of coordinate tuples becomes a PointSet."""
    if not isinstance(domain, (CartesianGrid, PointSet)):
        domain = PointSet(domain)
    return GeoTable(values, domain)
