"""Ensemble of realizations (src/ensembles.jl:10-85): indexing returns a geotable,
reductions act per element across realizations (ereduce, ensembles.jl:76-85).

The reductions run on the GPU (gsp_ensemble_stats): mean, var, cdf, ccdf and
quantiles of every cell in one pass over the realizations; `stats` asks for
several at once.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .geometry import GeoTable


class Ensemble:
    def __init__(self, domain, reals: dict, ctx=None):
        """reals: column name -> (R, N) array (realization-major)."""
        self.domain = domain
        self.reals = reals
        self.ctx = ctx

    def stats(self, *requests):
        """Several statistics in one pass: requests like ("mean",), ("var",),
        ("cdf", x), ("ccdf", x), ("quantile", p). Returns one GeoTable per
        request."""
        kinds = {"mean": _lib.STAT_MEAN, "var": _lib.STAT_VAR, "cdf": _lib.STAT_CDF,
                 "ccdf": _lib.STAT_CCDF, "quantile": _lib.STAT_QUANTILE}
        stats = [(kinds[r[0]], r[1] if len(r) > 1 else 0.0) for r in requests]
        ctx = _lib.default_context() if self.ctx is None else self.ctx
        cols = {k: _lib.ensemble_stats(ctx, v, stats) for k, v in self.reals.items()}
        return [GeoTable({k: c[q] for k, c in cols.items()}, self.domain)
                for q in range(len(stats))]

    def __len__(self):
        return next(iter(self.reals.values())).shape[0]

    def __getitem__(self, i):
        if isinstance(i, (list, tuple, np.ndarray)):
            return [self[j] for j in i]
        return GeoTable({k: v[i] for k, v in self.reals.items()}, self.domain)

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]

    def mean(self):
        return self.stats(("mean",))[0]

    def var(self):
        return self.stats(("var",))[0]

    def cdf(self, x):
        return self.stats(("cdf", x))[0]

    def ccdf(self, x):
        return self.stats(("ccdf", x))[0]

    def quantile(self, p):
        if np.ndim(p):
            return self.stats(*[("quantile", q) for q in p])
        return self.stats(("quantile", p))[0]

    def __repr__(self):
        return f"{self.domain.ndim}D Ensemble"
