"""Expectation of Gaussian fields (src/expectation/field.jl:5-36,
src/expectation/field/gaussian.jl:5-25):

  expectedvalue(process, domain; data, init, kwargs...)
  mean(process, domain; data, init, kwargs...)   prior mean without data,
                                                 posterior mean with data
  priormean(process, domain)                     the process mean everywhere
  posteriormean(process, domain, data; kwargs...)

The posterior mean of a GaussianProcess is simple kriging with the process
function and mean over the data (gaussian.jl:21-25, GeoStatsModels
`fitpredict(Kriging(f, mu), data, domain; kwargs...)`). It runs on the GPU as a
SEQSIM plan in data-only search mode: every node takes its k nearest data
(kernel (a)) and solves its simple kriging system (kernel (b)); one
realization swept with zero normals then gives mu + sum_j lambda_j (z_j - mu)
at every node, and the plan's sigma the kriging standard deviation.

The neighbourhood keywords follow GeoStatsModels 0.13 `fitpredict`
(minneighbors=1, maxneighbors=10, neighborhood=nothing: the k nearest data;
a MetricBall or radius: the k nearest within the ball; neighbors=false: all
data). A node with fewer than minneighbors data, or whose system is not
positive definite, is missing (NaN), like the reference's `missing`.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .geometry import CartesianGrid, GeoTable
from .processes import FieldProcess, GaussianProcess
from .simulation import MetricBall, _grid_struct


def expectedvalue(process, domain, *, data=None, init=None, ctx=None, **kwargs):
    """field.jl:16-22: the mean of continuous analytical processes."""
    if isinstance(process, FieldProcess) and process.iscontinuous() and process.isanalytical():
        return mean(process, domain, data=data, init=init, ctx=ctx, **kwargs)
    raise NotImplementedError("not implemented")


def mean(process, domain, *, data=None, init=None, ctx=None, **kwargs):
    """field.jl:35-36."""
    if data is None:
        return priormean(process, domain, **kwargs)
    return posteriormean(process, domain, data, init=init, ctx=ctx, **kwargs)


def priormean(process, domain, **kwargs):
    """gaussian.jl:5-19: each variable filled with its process mean."""
    if not isinstance(process, GaussianProcess):
        raise _lib.UnsupportedError("only GaussianProcess is provided by this backend")
    names = process.defaultnames()
    mu = np.atleast_1d(np.asarray(process.mean, dtype=np.float64))
    n = domain.nelements()
    return GeoTable({v: np.full(n, mu[i]) for i, v in enumerate(names)}, domain)


def _radius(neighborhood, f):
    if neighborhood is None:
        return 0.0  # KNearestSearch
    if isinstance(neighborhood, MetricBall):
        return neighborhood.radius
    if isinstance(neighborhood, str) and neighborhood == ":range":
        return f.range()
    return float(neighborhood)


class KrigingPlan:
    """Simple kriging weights of every node of a grid over a fixed set of data
    locations (a data-only SEQSIM plan). `predict(values)` maps one column of
    data values to (mean, variance) at every node; the weights are shared by
    every column, as they depend on the locations only."""

    def __init__(self, process, domain, coords, *, minneighbors=1, maxneighbors=10,
                 neighborhood=None, precision=_lib.FP64, ctx=None):
        ctx = _lib.default_context() if ctx is None else ctx
        f = process.func
        if f.nvariables() != 1:
            raise _lib.UnsupportedError("multivariate kriging is not supported by this backend")
        nd = domain.ndim
        coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, nd)
        if len(coords) == 0:
            raise _lib.ArgumentError("the posterior mean needs at least one datum")
        N = domain.nelements()
        path = np.arange(N, dtype=np.int64)
        p = _lib.GspSeqsimProblem()
        p.grid = _grid_struct(domain)
        p.model = f.struct(nd)
        p.mean = float(process.mean)
        p.n_data = len(coords)
        p.data_coords = coords.ctypes.data
        p.n_path = N
        p.path = path.ctypes.data
        p.minneighbors = int(minneighbors)
        p.maxneighbors = min(int(maxneighbors), len(coords))
        p.radius = _radius(neighborhood, f)
        p.search_mode = _lib.DATA_ONLY
        p.precision = precision
        p.device_pointers = 0
        self.plan = _lib.SeqsimPlan(ctx, p, keep=(coords, path))
        self.n_nodes = N
        out = self.plan.export(fields=("sigma", "flags"))
        self.variance = out["sigma"] ** 2
        self.missing = (out["flags"] & (_lib.FLAG_FEW | _lib.FLAG_FAIL)) != 0
        self.variance[self.missing] = np.nan

    def predict(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if np.isnan(v).any():
            raise _lib.UnsupportedError("missing data values are not supported by this backend")
        m = self.plan.sweep(np.zeros((1, self.n_nodes)), v)[0]
        m[self.missing] = np.nan
        return m, self.variance


def posteriormean(process, domain, data, *, init=None, path=None, point=True, prob=False,
                  neighbors=True, minneighbors=1, maxneighbors=10, distance="euclidean",
                  neighborhood=None, precision=_lib.FP64, ctx=None):
    """gaussian.jl:21-25: simple kriging of every data column over the domain.

    With prob=True each column `v` comes with `v_variance`, the simple kriging
    variance (the Normal(mean, sqrt(variance)) of GeoStatsModels' probabilistic
    prediction). `init` is accepted and unused, as in the reference; `path`
    does not change the result (every node is predicted independently)."""
    if not isinstance(process, GaussianProcess):
        raise _lib.UnsupportedError("only GaussianProcess is provided by this backend")
    if not isinstance(domain, CartesianGrid):
        raise _lib.UnsupportedError("this backend predicts on CartesianGrid domains")
    if not point:
        raise _lib.UnsupportedError("block kriging (point=false) is not supported by this backend")
    if distance not in ("euclidean", None):
        raise _lib.UnsupportedError("only the Euclidean distance is supported")
    coords = data.domain.centroids()
    if not neighbors:
        minneighbors, maxneighbors, neighborhood = 1, len(coords), None
    kp = KrigingPlan(process, domain, coords, minneighbors=minneighbors,
                     maxneighbors=maxneighbors, neighborhood=neighborhood, precision=precision,
                     ctx=ctx)
    cols = {}
    for name in data.names():
        m, var = kp.predict(data[name])
        cols[name] = m
        if prob:
            cols[f"{name}_variance"] = var.copy()
    return GeoTable(cols, domain)
