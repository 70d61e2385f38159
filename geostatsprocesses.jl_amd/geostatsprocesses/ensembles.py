"""Ensemble of realizations (src/ensembles.jl:10-85): indexing returns a geotable,
reductions act per element across realizations."""
from __future__ import annotations

import numpy as np

from .geometry import GeoTable


class Ensemble:
    def __init__(self, domain, reals: dict):
        """reals: column name -> (R, N) array (realization-major)."""
        self.domain = domain
        self.reals = reals

    def __len__(self):
        return next(iter(self.reals.values())).shape[0]

    def __getitem__(self, i):
        if isinstance(i, (list, tuple, np.ndarray)):
            return [self[j] for j in i]
        return GeoTable({k: v[i] for k, v in self.reals.items()}, self.domain)

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]

    def _reduce(self, f):
        return GeoTable({k: f(v) for k, v in self.reals.items()}, self.domain)

    def mean(self):
        return self._reduce(lambda v: v.mean(axis=0))

    def var(self):
        return self._reduce(lambda v: v.var(axis=0, ddof=1))

    def cdf(self, x):
        return self._reduce(lambda v: (v <= x).mean(axis=0))

    def ccdf(self, x):
        return self._reduce(lambda v: (v > x).mean(axis=0))

    def quantile(self, p):
        if np.ndim(p):
            return [self.quantile(q) for q in p]
        return self._reduce(lambda v: np.quantile(v, p, axis=0))

    def __repr__(self):
        return f"{self.domain.ndim}D Ensemble"
