"""Placement of conditioning data into realization buffers.

NearestInit (src/initialization/nearest.jl:12-34): each datum goes to the
nearest element of the domain; a later datum overwrites an earlier one.
ExplicitInit (src/initialization/explicit.jl:12-46): datum orig[i] goes to
element dest[i]. initialize (src/processes/field.jl:43-58) allocates the
realization and the simulated mask and applies the init method.
"""
from __future__ import annotations

import numpy as np

from .geometry import CartesianGrid


class InitMethod:
    pass


class NearestInit(InitMethod):
    def destinations(self, domain, data) -> np.ndarray:
        pts = np.asarray(data.domain.centroids(), dtype=np.float64)
        if isinstance(domain, CartesianGrid):
            # nearest cell centre: round per axis (ties to the lower index)
            rel = (pts - domain.origin) / domain.spacing - 0.5
            idx = np.ceil(rel - 0.5).astype(np.int64)
            idx = np.clip(idx, 0, np.asarray(domain.dims) - 1)
            mult = np.cumprod((1,) + tuple(domain.dims[:-1]))
            return (idx * mult).sum(axis=1)
        cents = np.asarray(domain.centroids(), dtype=np.float64)
        out = np.empty(len(pts), dtype=np.int64)
        for i, x in enumerate(pts):
            out[i] = int(np.argmin(np.sum((cents - x) ** 2, axis=1)))
        return out


class ExplicitInit(InitMethod):
    """ExplicitInit([orig], dest) with 0-based indices."""

    def __init__(self, *args):
        if len(args) == 1:
            self.orig, self.dest = None, np.asarray(args[0], dtype=np.int64)
        elif len(args) == 2:
            self.orig = np.asarray(args[0], dtype=np.int64)
            self.dest = np.asarray(args[1], dtype=np.int64)
        else:
            raise TypeError("ExplicitInit([orig], dest)")

    def destinations(self, domain, data) -> np.ndarray:
        orig = np.arange(data.domain.nelements()) if self.orig is None else self.orig
        if len(orig) != len(self.dest):
            raise AssertionError("invalid explicit initialization")
        return orig, self.dest


def initialize(process, domain, data, init):
    """Returns (real, mask): dicts of column -> buffer / bool mask."""
    n = domain.nelements()
    if data is None:
        names = process.defaultnames()
        real = {k: np.full(n, np.nan) for k in names}
        mask = {k: np.zeros(n, dtype=bool) for k in names}
        return real, mask
    names = data.names()
    real = {k: np.full(n, np.nan) for k in names}
    mask = {k: np.zeros(n, dtype=bool) for k in names}
    if isinstance(init, ExplicitInit):
        orig, dest = init.destinations(domain, data)
    else:
        dest = init.destinations(domain, data)
        orig = np.arange(len(dest))
    for k in names:
        vals = np.asarray(data[k], dtype=np.float64)
        for i, j in zip(orig, dest):
            v = vals[i]
            if not np.isnan(v):
                real[k][j] = v
                mask[k][j] = True
    return real, mask
