"""Field simulation API mirroring the reference's plugin contract:

  rand(process, domain, nreals; data, method, init, ...)   field.jl:47-124
  preprocess(rng, process, method, init, domain, data)    called once
  randsingle(rng, process, method, domain, data, preproc) called per realization

for (GaussianProcess, SEQSIM), (GaussianProcess, LUSIM) and (GaussianProcess,
FFTSIM). All arithmetic runs in libgsp_hip.so on the GPU; this module only
shapes arguments. The ensemble form of `rand` draws every realization's random
numbers in the reference's order (realization after realization, one per path
node / simulated cell / grid cell) and runs them in one device call.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .ensembles import Ensemble
from .geometry import CartesianGrid, GeoTable, PointSet
from .initialization import ExplicitInit, NearestInit, initialize
from .functions import CoregionalizedFunction, Transiogram
from .processes import GaussianProcess, IndicatorProcess, LinearPath, SourcePath


class FieldSimulationMethod:
    pass


class MetricBall:
    """Isotropic search ball of a given radius (Meshes MetricBall)."""

    def __init__(self, radius):
        self.radius = float(radius)


class SEQSIM(FieldSimulationMethod):
    """SEQSIM(path, minneighbors=1, maxneighbors=26, neighborhood=':range',
    distance='euclidean') — seqsim.jl:42-48, plus the backend options
    search_mode (k nearest then mask | masked k nearest) and precision
    (FP64 | FP32 with one FP64-residual refinement step)."""

    def __init__(self, path=None, minneighbors=1, maxneighbors=26, neighborhood=":range",
                 distance="euclidean", search_mode=_lib.KNN_THEN_MASK, precision=_lib.FP64):
        self.path = LinearPath() if path is None else path
        self.minneighbors = int(minneighbors)
        self.maxneighbors = int(maxneighbors)
        self.neighborhood = neighborhood
        if distance not in ("euclidean", None):
            raise _lib.UnsupportedError("only the Euclidean distance is supported")
        self.distance = distance
        self.search_mode = search_mode
        self.precision = precision


class LUSIM(FieldSimulationMethod):
    """LUSIM() — lusim.jl:36; precision LU_FP64 or LU_MIXED (FP32 MFMA trailing
    updates, FP64 panels and conditioning)."""

    def __init__(self, precision=_lib.LU_FP64):
        self.precision = precision


class FFTSIM(FieldSimulationMethod):
    """FFTSIM(minneighbors=1, maxneighbors=26, neighborhood=None,
    distance='euclidean') — fftsim.jl:49-53. The options are those of the
    kriging that conditions the realizations on data."""

    def __init__(self, minneighbors=1, maxneighbors=26, neighborhood=None, distance="euclidean"):
        self.minneighbors = int(minneighbors)
        self.maxneighbors = int(maxneighbors)
        self.neighborhood = neighborhood
        if distance not in ("euclidean", None):
            raise _lib.UnsupportedError("only the Euclidean distance is supported")
        self.distance = distance


def _grid_struct(domain) -> _lib.GspGrid:
    if not isinstance(domain, CartesianGrid):
        raise _lib.UnsupportedError("this backend simulates on CartesianGrid domains")
    g = _lib.GspGrid()
    g.ndim = domain.ndim
    for d in range(3):
        g.dims[d] = domain.dims[d] if d < domain.ndim else 1
        g.origin[d] = float(domain.origin[d]) if d < domain.ndim else 0.0
        g.spacing[d] = float(domain.spacing[d]) if d < domain.ndim else 1.0
    return g


def _sides(domain):
    lo, hi = domain.boundingbox() if isinstance(domain, CartesianGrid) else (
        domain.centroids().min(axis=0), domain.centroids().max(axis=0))
    return hi - lo


def defaultsimulation(process, domain, data=None):
    """field.jl:152-180."""
    if isinstance(process, IndicatorProcess):
        # SEQSIM visiting the data's nearest cells first (field.jl:168-180)
        if data is None:
            return SEQSIM(path=LinearPath())
        dest = NearestInit().destinations(domain, data)
        return SEQSIM(path=SourcePath(_unique_in_order(dest)))
    if not isinstance(process, GaussianProcess):
        return None
    f = process.func
    if (isinstance(domain, CartesianGrid) and f.isstationary() and f.nvariables() == 1
            and f.range() <= np.min(_sides(domain)) / 3 and data is None):
        return FFTSIM()
    if domain.nelements() < 100 * 100 and f.isstationary() and f.issymmetric() and f.isbanded():
        return LUSIM()
    return SEQSIM()


def _unique_in_order(x):
    """unique(x) keeping first occurrences in order (Julia's `unique`)."""
    _, first = np.unique(x, return_index=True)
    return np.asarray(x)[np.sort(first)]


# ---------------------------------------------------------------- SEQSIM ----
class SeqsimPreproc:
    """The plan and the values the sweeps need; `A` and `means` for a
    multivariate process of A * c(h) (data_vals then one row per variable)."""

    def __init__(self, plan, names, data_vals, path, A=None, means=None):
        self.plan = plan
        self.names = names
        self.data_vals = data_vals
        self.path = path
        self.A = A
        self.means = means


def _seqsim_problem(process, method, domain, data, ctx, func=None, data_vals=None):
    """The SEQSIM plan (seqsim.jl:50-70). `func` replaces the process function
    for the search and the kriging weights (the spatial part of an indicator
    process); `data_vals` replaces the data values."""
    f = process.func if func is None else func
    A = means = None
    if f.nvariables() != 1:
        # an intrinsic coregionalization A * c(h): the plan of c(h) serves every
        # variable (simple cokriging weights lambda_j * I), the draws couple them
        if not isinstance(f, CoregionalizedFunction) or not isinstance(process, GaussianProcess):
            raise _lib.UnsupportedError(
                "multivariate SEQSIM needs a coregionalized function A * c (intrinsic model)")
        A = np.ascontiguousarray(f.A, dtype=np.float64)
        means = np.ascontiguousarray(np.atleast_1d(process.mean), dtype=np.float64)
        f = f.base
    grid = _grid_struct(domain)
    nd = domain.ndim
    p = _lib.GspSeqsimProblem()
    p.grid = grid
    p.model = f.struct(nd)
    p.mean = (float(process.mean) if A is None else 0.0) if isinstance(process, GaussianProcess) \
        else 0.0
    if data is not None:
        coords = np.ascontiguousarray(data.domain.centroids(), dtype=np.float64).reshape(-1, nd)
        names = data.names()
        nvar = 1 if A is None else len(A)
        if len(names) != nvar:
            raise _lib.UnsupportedError(f"SEQSIM data must have {nvar} variable(s)")
        if A is None:
            vals = np.ascontiguousarray(data[names[0]] if data_vals is None else data_vals,
                                        dtype=np.float64)
        else:
            vals = np.ascontiguousarray(np.stack([np.asarray(data[n], dtype=np.float64)
                                                  for n in names]))
        if np.isnan(vals).any():
            raise _lib.UnsupportedError("missing data values are not supported by this backend")
    else:
        coords = np.zeros((0, nd))
        names = process.defaultnames()
        vals = np.zeros(0)
    path = np.ascontiguousarray(method.path.traverse(domain), dtype=np.int64)
    p.n_data = len(coords)
    p.data_coords = coords.ctypes.data if len(coords) else None
    p.n_path = len(path)
    p.path = path.ctypes.data
    p.minneighbors = method.minneighbors
    p.maxneighbors = method.maxneighbors
    neigh = method.neighborhood
    if isinstance(neigh, str) and neigh == ":range":
        p.radius = f.range()
    elif isinstance(neigh, MetricBall):
        p.radius = neigh.radius
    elif neigh is None:
        p.radius = 0.0
    else:
        p.radius = float(neigh)
    p.search_mode = method.search_mode
    p.precision = method.precision
    p.device_pointers = 0
    plan = _lib.SeqsimPlan(ctx, p, keep=(coords, path))
    return SeqsimPreproc(plan, names, vals, path, A=A, means=means)


class IndicatorPreproc:
    def __init__(self, plan, names, labels, data_cat, prob):
        self.plan = plan
        self.names = names
        self.labels = labels
        self.data_cat = data_cat
        self.prob = prob


def _indicator_problem(process, method, domain, data, ctx):
    """SEQSIM of an IndicatorProcess (seqsim.jl:247-251,292-310,342-347): the
    data one-hot encoded by their sorted distinct values (OneHot), simple
    cokriging with the a priori probabilities as mean, clamp + normalize, a
    categorical draw. With an intrinsic coregionalization A * c(h) and every
    category observed at every neighbour, the cokriging weights are lambda_j
    times the identity, lambda the simple-kriging weights of c(h): the plan is
    the Gaussian one built with c(h)."""
    f = process.func
    if isinstance(f, Transiogram):
        base = f.base_covariance()  # shape transiogram: A * c(h), A = diag(p) - p p^T
    elif isinstance(f, CoregionalizedFunction):
        base = f.base
    else:
        raise _lib.UnsupportedError(
            "this backend simulates indicator processes of a coregionalized covariance "
            "A * c(h) or of a shape transiogram")
    if data is not None:
        names = data.names()
        if len(names) != 1:
            raise _lib.UnsupportedError("indicator data must have exactly one variable")
        raw = np.asarray(data[names[0]])
        labels, data_cat = np.unique(raw, return_inverse=True)
        if len(labels) != f.nvariables():
            raise AssertionError(f"the data have {len(labels)} categories, the process "
                                 f"{f.nvariables()}")
    else:
        names = process.defaultnames()
        labels = np.arange(1, f.nvariables() + 1)
        data_cat = np.zeros(0, np.int64)
    data_cat = np.ascontiguousarray(data_cat, dtype=np.int32)
    pre = _seqsim_problem(process, method, domain, data, ctx, func=base,
                          data_vals=data_cat.astype(np.float64) if data is not None else None)
    return IndicatorPreproc(pre.plan, names, labels, data_cat, process.prob)


# ----------------------------------------------------------------- LUSIM ----
class LusimPreproc:
    """One factorized plan per variable (lusim.jl:64-107); `rho` couples the
    second variable's normals to the first's (lusim.jl:118-120)."""

    def __init__(self, plans, names, rho=None):
        self.plans = plans
        self.names = names
        self.rho = rho

    @property
    def plan(self):
        return self.plans[0]


def _marginal_struct(f, j, ndim):
    """_marginalize (lusim.jl:132-137) of A * c: the covariance A[j, j] c(h) of
    variable j (its nugget scaled with it)."""
    if not isinstance(f, CoregionalizedFunction):
        raise _lib.UnsupportedError(
            "bivariate LUSIM needs a coregionalized function A * c (intrinsic model)")
    m = f.base.struct(ndim)
    m.sill = float(f.A[j, j]) * m.sill
    m.nugget = float(f.A[j, j]) * m.nugget
    return m


def _rho(f):
    """_rho (lusim.jl:139-143): C12(0) / sqrt(s1 s2), with C(0) = sill for a
    covariance, so that rho = A[0, 1] / sqrt(A[0, 0] A[1, 1])."""
    c12 = f.A[0, 1] * f.base.sill()
    s1, s2 = np.diag(f.sill())
    return float(c12 / np.sqrt(s1 * s2))


def _lusim_problem(process, method, init, domain, data, ctx):
    f = process.func
    grid = _grid_struct(domain)
    real, mask = initialize(process, domain, data, init)
    names = tuple(real)
    assert len(names) == f.nvariables(), "incompatible number of variables for geostatistical function"
    assert len(names) in (1, 2), "LUSIM only supports univariate and bivariate simulation"
    means = np.atleast_1d(np.asarray(process.mean, dtype=np.float64))
    plans = []
    for j, var in enumerate(names):
        dinds = np.ascontiguousarray(np.flatnonzero(mask[var]), dtype=np.int64)
        z1 = np.ascontiguousarray(real[var][dinds], dtype=np.float64)
        p = _lib.GspLusimProblem()
        p.grid = grid
        p.model = (_marginal_struct(f, j, domain.ndim)
                   if len(names) == 2 or isinstance(f, CoregionalizedFunction)
                   else f.struct(domain.ndim))
        p.mean = float(means[j])
        p.n_data = len(dinds)
        p.dinds = dinds.ctypes.data if len(dinds) else None
        p.z1 = z1.ctypes.data if len(dinds) else None
        p.precision = method.precision
        p.device_pointers = 0
        plans.append(_lib.LusimPlan(ctx, p, keep=(dinds, z1)))
    rho = None
    if len(names) == 2:
        rho = _rho(f)
        if plans[0].n_sim != plans[1].n_sim:
            # the reference mixes w1 into the second variable's draw (lusim.jl:162)
            raise _lib.ArgumentError(
                "bivariate LUSIM draws both variables from the same normals: the variables "
                f"need as many simulated cells ({plans[0].n_sim} != {plans[1].n_sim})")
    return LusimPreproc(plans, names, rho)


# ---------------------------------------------------------------- FFTSIM ----
class FftsimPreproc:
    def __init__(self, plan, names):
        self.plan = plan
        self.names = names


def _fftsim_problem(process, method, init, domain, data, ctx):
    """fftsim.jl:54-107: the amplitude spectrum and, with data, the kriging
    mean and the cells initialize() placed the data on."""
    f = process.func
    if not f.isstationary():
        raise AssertionError("geostatistical function must be stationary")
    grid = _grid_struct(domain)
    real, mask = initialize(process, domain, data, init)
    names = tuple(real)
    if len(names) != 1:
        raise AssertionError("FFTSIM does not support multivariate simulation")
    var = names[0]
    nd = domain.ndim
    p = _lib.GspFftsimProblem()
    p.grid = grid
    p.model = f.struct(nd)
    p.mean = float(process.mean)
    keep = ()
    if data is not None:
        coords = np.ascontiguousarray(data.domain.centroids(), dtype=np.float64).reshape(-1, nd)
        vals = np.ascontiguousarray(data[var], dtype=np.float64)
        if np.isnan(vals).any():
            raise _lib.UnsupportedError("missing data values are not supported by this backend")
        cells = np.ascontiguousarray(np.flatnonzero(mask[var]), dtype=np.int64)
        p.n_data = len(coords)
        p.data_coords = coords.ctypes.data
        p.data_vals = vals.ctypes.data
        p.n_cells = len(cells)
        p.cells = cells.ctypes.data
        keep = (coords, vals, cells)
    p.minneighbors = method.minneighbors
    p.maxneighbors = method.maxneighbors
    neigh = method.neighborhood
    if neigh is None:
        p.radius = 0.0
    elif isinstance(neigh, MetricBall):
        p.radius = neigh.radius
    else:
        p.radius = float(neigh)
    p.device_pointers = 0
    plan = _lib.FftsimPlan(ctx, p, keep=keep)
    return FftsimPreproc(plan, names)


# -------------------------------------------------------- plugin contract ----
def preprocess(rng, process, method, init, domain, data, ctx=None):
    ctx = _lib.default_context() if ctx is None else ctx
    if isinstance(process, IndicatorProcess) and isinstance(method, SEQSIM):
        return _indicator_problem(process, method, domain, data, ctx)
    if not isinstance(process, GaussianProcess):
        raise _lib.UnsupportedError("only GaussianProcess and IndicatorProcess (SEQSIM) are "
                                    "accelerated by this backend")
    if isinstance(method, SEQSIM):
        return _seqsim_problem(process, method, domain, data, ctx)
    if isinstance(method, LUSIM):
        return _lusim_problem(process, method, init, domain, data, ctx)
    if isinstance(method, FFTSIM):
        return _fftsim_problem(process, method, init, domain, data, ctx)
    raise _lib.UnsupportedError(f"{type(method).__name__} is not provided by this backend")


def _draw(rng, preproc, R):
    """R realizations of every variable: {name: (R, N) array}."""
    name = preproc.names[0]
    if isinstance(preproc, IndicatorPreproc):
        # one uniform per path node for the categorical draw (seqsim.jl:342-347)
        u = rng.random((R, preproc.plan.n_nodes))
        cats = preproc.plan.sweep_indicator(u, preproc.prob, preproc.data_cat)
        return {name: preproc.labels[cats]}
    if isinstance(preproc, SeqsimPreproc) and preproc.A is not None:
        # nvars normals per path node (the MvNormal draws of seqsim.jl:335-340)
        nv = len(preproc.A)
        z = rng.standard_normal((R, preproc.plan.n_nodes, nv))
        out = preproc.plan.sweep_multivariate(z, preproc.A, preproc.means,
                                              preproc.data_vals if preproc.plan.info.n_data
                                              else None)
        return {n: out[v] for v, n in enumerate(preproc.names)}
    if isinstance(preproc, SeqsimPreproc):
        z = rng.standard_normal((R, preproc.plan.n_nodes))
        return {name: preproc.plan.sweep(z, preproc.data_vals)}
    if isinstance(preproc, FftsimPreproc):
        # uniform white noise, one value per cell (fftsim.jl:124)
        w = rng.random((R, preproc.plan.n_nodes))
        return {name: preproc.plan.sample(w)}
    # LUSIM: per realization the first variable's normals, then the second's
    # (lusim.jl:112-126, 160)
    ns = [pl.n_sim for pl in preproc.plans]
    w = rng.standard_normal((R, sum(ns)))
    w1 = w[:, :ns[0]]
    out = {name: preproc.plans[0].sample(w1)}
    if len(preproc.plans) == 2:
        w2 = w[:, ns[0]:]
        rho = preproc.rho
        out[preproc.names[1]] = preproc.plans[1].sample(rho * w1 + np.sqrt(1.0 - rho ** 2) * w2)
    return out


def randsingle(rng, process, method, domain, data, preproc):
    return {k: v[0] for k, v in _draw(rng, preproc, 1).items()}


def rand(process, domain, nreals=None, *, rng=None, data=None, method=None, init=None,
         workers=None, async_=False, showprogress=False, ctx=None):
    """rand([rng], process, domain, [nreals]; data, method, init, ...)."""
    if rng is None:
        rng = np.random.default_rng()
    init = NearestInit() if init is None else init
    if async_ and workers is not None and 0 in list(workers):
        raise _lib.ArgumentError(
            "the `async` option is not allowed when the master process is in the `workers`")
    smethod = defaultsimulation(process, domain, data) if method is None else method
    preproc = preprocess(rng, process, smethod, init, domain, data, ctx=ctx)
    if nreals is None:
        return GeoTable(randsingle(rng, process, smethod, domain, data, preproc), domain)
    return Ensemble(domain, _draw(rng, preproc, int(nreals)), ctx=ctx)
