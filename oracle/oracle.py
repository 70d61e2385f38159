"""CPU oracle bindings — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker (or as the timed CPU restatement). The product
library never loads it.

* SEQSIM: ctypes bindings to oracle/libgsp_oracle.so (gsp_oracle.c restates
  src/simulation/field/seqsim.jl:50-156 and the dependency semantics it calls).
* LUSIM: a numpy/scipy restatement of src/simulation/field/lusim.jl:38-175 and
  src/utils.jl:50-62 (LAPACK potrf/trsm through scipy).
* NearestInit / ExplicitInit: src/initialization/nearest.jl:12-34 and
  src/initialization/explicit.jl:19-46.
* Learned kriging weights (no counterpart in the reference): a numpy
  restatement of the surrogate's forward pass (BF16 rounding of weights and
  activations, FP32 accumulation, tanh-form GELU) and the experimental
  variogram used to compare learned and exact realizations.
"""
from __future__ import annotations

import ctypes
import os
import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SPHERICAL, EXPONENTIAL, GAUSSIAN = 0, 1, 2
# Gaussian models carry this nugget on top of their own (unpinned value; the
# reference's LUSIM testset needs a positive one, test/field.jl:29-56)
GAUSSIAN_NUGGET = 1e-6
KNN_THEN_MASK, MASKED_KNN, DATA_ONLY = 0, 1, 2
FP64, FP32_REFINE = 0, 1
FLAG_FEW, FLAG_FAIL, FLAG_FP64_FALLBACK, FLAG_LEARNED = 1, 2, 4, 8


class OrModel(ctypes.Structure):
    _fields_ = [("family", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("sill", ctypes.c_double), ("nugget", ctypes.c_double),
                ("ranges", ctypes.c_double * 3), ("rotation", ctypes.c_double * 9),
                ("has_rotation", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class OrGrid(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("dims", ctypes.c_int64 * 3), ("origin", ctypes.c_double * 3),
                ("spacing", ctypes.c_double * 3)]


class OrProblem(ctypes.Structure):
    _fields_ = [("grid", OrGrid), ("model", OrModel), ("mean", ctypes.c_double),
                ("n_data", ctypes.c_int64), ("data_coords", ctypes.POINTER(ctypes.c_double)),
                ("n_path", ctypes.c_int64), ("path", ctypes.POINTER(ctypes.c_int64)),
                ("minneighbors", ctypes.c_int32), ("maxneighbors", ctypes.c_int32),
                ("radius", ctypes.c_double), ("search_mode", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("device_pointers", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]


@dataclass
class SeqsimCase:
    """A SEQSIM problem in physical units (0-based grid ids, column-major grid)."""
    dims: tuple
    origin: tuple = None
    spacing: tuple = None
    family: int = SPHERICAL
    sill: float = 1.0
    nugget: float = 0.0
    ranges: tuple = (1.0,)
    rotation: np.ndarray | None = None
    mean: float = 0.0
    data_coords: np.ndarray | None = None
    data_vals: np.ndarray | None = None
    path: np.ndarray | None = None
    kmin: int = 1
    kmax: int = 26
    radius: float = 0.0
    mode: int = KNN_THEN_MASK
    precision: int = FP64
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        nd = len(self.dims)
        if self.origin is None:
            self.origin = (0.0,) * nd
        if self.spacing is None:
            self.spacing = (1.0,) * nd
        if len(self.ranges) == 1 and nd > 1:
            self.ranges = tuple(self.ranges) * nd
        if self.data_coords is None:
            self.data_coords = np.zeros((0, nd))
        self.data_coords = np.ascontiguousarray(self.data_coords, dtype=np.float64).reshape(-1, nd)
        if self.data_vals is None:
            self.data_vals = np.zeros(len(self.data_coords))
        self.data_vals = np.ascontiguousarray(self.data_vals, dtype=np.float64)
        if self.path is None:
            self.path = np.arange(self.n_nodes, dtype=np.int64)
        self.path = np.ascontiguousarray(self.path, dtype=np.int64)

    @property
    def ndim(self):
        return len(self.dims)

    @property
    def n_nodes(self):
        return int(np.prod(self.dims))

    @property
    def n_data(self):
        return len(self.data_coords)

    def fill(self, st):
        """Fill a ctypes problem struct (shared field layout) from this case."""
        nd = self.ndim
        st.grid.ndim = nd
        for d in range(3):
            st.grid.dims[d] = int(self.dims[d]) if d < nd else 1
            st.grid.origin[d] = float(self.origin[d]) if d < nd else 0.0
            st.grid.spacing[d] = float(self.spacing[d]) if d < nd else 1.0
            st.model.ranges[d] = float(self.ranges[d]) if d < nd else 1.0
        st.model.family = self.family
        st.model.kind = 0
        st.model.sill = self.sill
        st.model.nugget = self.nugget
        if self.rotation is not None:
            rot = np.asarray(self.rotation, dtype=np.float64).reshape(3, 3)
            for q in range(9):
                st.model.rotation[q] = rot.flat[q]
            st.model.has_rotation = 1
        st.mean = self.mean
        st.n_data = self.n_data
        st.data_coords = self.data_coords.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        st.n_path = len(self.path)
        st.path = self.path.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        st.minneighbors = self.kmin
        st.maxneighbors = self.kmax
        st.radius = self.radius
        st.search_mode = self.mode
        st.precision = self.precision
        return st

    def struct(self):
        return self.fill(OrProblem())


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libgsp_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.or_exp.restype = ctypes.c_double
        L.or_exp.argtypes = [ctypes.c_double]
        L.or_cov_u.restype = ctypes.c_double
        L.or_cov_u.argtypes = [P(OrModel), ctypes.c_double]
        L.or_scale_factor.restype = ctypes.c_double
        L.or_scale_factor.argtypes = [P(OrProblem)]
        L.or_search_brute.restype = ctypes.c_int32
        L.or_search_brute.argtypes = [P(OrProblem), ctypes.c_int32, P(ctypes.c_int64),
                                      ctypes.c_int64, P(ctypes.c_int32)]
        L.or_bucket_create.restype = ctypes.c_void_p
        L.or_bucket_create.argtypes = [P(OrProblem)]
        L.or_bucket_free.argtypes = [ctypes.c_void_p]
        L.or_search_bucket.restype = ctypes.c_int32
        L.or_search_bucket.argtypes = [P(OrProblem), ctypes.c_void_p, ctypes.c_int32,
                                       P(ctypes.c_int64), ctypes.c_int64, P(ctypes.c_int32)]
        L.or_krige.restype = ctypes.c_int32
        L.or_krige.argtypes = [P(OrProblem), ctypes.c_double, ctypes.c_int64, P(ctypes.c_int32),
                               ctypes.c_int32, ctypes.c_int32, P(ctypes.c_double),
                               P(ctypes.c_double)]
        L.or_seqsim_plan.restype = ctypes.c_int32
        L.or_seqsim_plan.argtypes = [P(OrProblem), ctypes.c_int32, P(ctypes.c_int32),
                                     P(ctypes.c_int32), P(ctypes.c_double), P(ctypes.c_double),
                                     P(ctypes.c_uint8), P(ctypes.c_int32), P(ctypes.c_int32)]
        L.or_seqsim_sweep.restype = None
        L.or_seqsim_sweep.argtypes = [P(OrProblem), ctypes.c_int32, P(ctypes.c_int32),
                                      P(ctypes.c_int32), P(ctypes.c_double), P(ctypes.c_double),
                                      P(ctypes.c_uint8), ctypes.c_int32, P(ctypes.c_double),
                                      P(ctypes.c_double), P(ctypes.c_double)]
        L.or_seqsim_reference_cost.restype = ctypes.c_int64
        L.or_seqsim_reference_cost.argtypes = [P(OrProblem), ctypes.c_int32, P(ctypes.c_double),
                                               P(ctypes.c_double), P(ctypes.c_double),
                                               ctypes.c_int32, ctypes.c_int64]
        _LIB = L
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


# ------------------------------------------------------------------ primitives
def exp(x: float) -> float:
    return lib().or_exp(float(x))


def cov_u(family, sill, nugget, u) -> float:
    m = OrModel(family=family, kind=0, sill=sill, nugget=nugget)
    return lib().or_cov_u(ctypes.byref(m), float(u))


def scale_factor(case: SeqsimCase) -> float:
    return lib().or_scale_factor(ctypes.byref(case.struct()))


def clamp_neighbors(nelm, nmin, nmax):
    """seqsim.jl:202-208."""
    if nmax > nelm or nmax < 1:
        nmax = nelm
    if nmin > nmax or nmin < 1:
        nmin = 1
    return nmin, nmax


def positions(case: SeqsimCase) -> np.ndarray:
    pos = np.empty(case.n_nodes, dtype=np.int64)
    pos[case.path] = np.arange(len(case.path))
    return pos


def search(case: SeqsimCase, target: int, brute=True, pos=None, bucket=None):
    st = case.struct()
    _, nmax = clamp_neighbors(case.n_nodes, case.kmin, case.kmax)
    pos = positions(case) if pos is None else pos
    out = np.full(nmax, -1, dtype=np.int32)
    if brute:
        n = lib().or_search_brute(ctypes.byref(st), nmax, _p(pos, ctypes.c_int64), int(target),
                                  _p(out, ctypes.c_int32))
    else:
        own = bucket is None
        b = lib().or_bucket_create(ctypes.byref(st)) if own else bucket
        n = lib().or_search_bucket(ctypes.byref(st), b, nmax, _p(pos, ctypes.c_int64),
                                   int(target), _p(out, ctypes.c_int32))
        if own:
            lib().or_bucket_free(b)
    return out[:n].copy()


def krige(case: SeqsimCase, target: int, nbr, precision=None):
    st = case.struct()
    alpha = lib().or_scale_factor(ctypes.byref(st))
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    lam = np.zeros(max(len(nbr), 1))
    sig = np.zeros(1)
    prec = case.precision if precision is None else precision
    f = lib().or_krige(ctypes.byref(st), alpha, int(target), _p(nbr, ctypes.c_int32), len(nbr),
                       prec, _p(lam, ctypes.c_double), _p(sig, ctypes.c_double))
    return lam[:len(nbr)], float(sig[0]), int(f)


@dataclass
class Plan:
    kmax: int
    nfound: np.ndarray
    nbr: np.ndarray
    lam: np.ndarray
    sigma: np.ndarray
    flags: np.ndarray
    level: np.ndarray


def seqsim_plan(case: SeqsimCase, brute=False) -> Plan:
    st = case.struct()
    N = case.n_nodes
    _, nmax = clamp_neighbors(N, case.kmin, case.kmax)
    nfound = np.zeros(N, np.int32)
    nbr = np.zeros((N, nmax), np.int32)
    lam = np.zeros((N, nmax))
    sigma = np.zeros(N)
    flags = np.zeros(N, np.uint8)
    level = np.zeros(N, np.int32)
    kout = ctypes.c_int32(0)
    rc = lib().or_seqsim_plan(ctypes.byref(st), int(brute), _p(nfound, ctypes.c_int32),
                              _p(nbr, ctypes.c_int32), _p(lam, ctypes.c_double),
                              _p(sigma, ctypes.c_double), _p(flags, ctypes.c_uint8),
                              _p(level, ctypes.c_int32), ctypes.byref(kout))
    if rc != 0:
        raise ValueError("oracle rejected the problem")
    return Plan(kout.value, nfound, nbr, lam, sigma, flags, level)


def seqsim_sweep(case: SeqsimCase, plan: Plan, z: np.ndarray) -> np.ndarray:
    """z: (R, n_path) standard normals in path order. Returns (R, N)."""
    st = case.struct()
    z = np.ascontiguousarray(z, dtype=np.float64)
    R = z.shape[0]
    out = np.zeros((R, case.n_nodes))
    lib().or_seqsim_sweep(ctypes.byref(st), plan.kmax, _p(plan.nfound, ctypes.c_int32),
                          _p(plan.nbr, ctypes.c_int32), _p(plan.lam, ctypes.c_double),
                          _p(plan.sigma, ctypes.c_double), _p(plan.flags, ctypes.c_uint8), R,
                          _p(z, ctypes.c_double), _p(case.data_vals, ctypes.c_double),
                          _p(out, ctypes.c_double))
    return out


# -------------------------------------------------------- indicator SEQSIM ----
def mv_factor(A):
    """Cholesky factor of the coregionalization matrix A (lower), with the
    sums taken in ascending order: L_jj = sqrt(A_jj - sum_k L_jk^2),
    L_ij = (A_ij - sum_k L_ik L_jk) / L_jj."""
    A = np.asarray(A, dtype=np.float64)
    nv = A.shape[0]
    L = [[0.0] * nv for _ in range(nv)]
    for j in range(nv):
        for i in range(j, nv):
            acc = float(A[i, j])
            for k in range(j):
                acc = acc - L[i][k] * L[j][k]
            if i == j:
                if not acc > 0.0:
                    raise np.linalg.LinAlgError("coregionalization matrix not positive definite")
                L[j][j] = math.sqrt(acc)
            else:
                L[i][j] = acc / L[j][j]
    return np.array(L)


def seqsim_sweep_multivariate(case: SeqsimCase, plan: Plan, z, A, means, data_vals):
    """seqsim.jl:72-156 for a GaussianProcess of A * c(h) with the MvNormal
    prior (seqsim.jl:287) and draws (seqsim.jl:335-340: nv normals per node).
    Simple cokriging with A * c(h) and every variable at every neighbour has the
    weights lambda_j * I of c(h)'s system and error covariance A sigma^2, so a
    node draws mu_SK + sigma L z, L = chol(A): the normals are mixed per node,
    w_v = sum_{u<=v} L_vu z_u (ascending u), then every variable is the scalar
    sweep of the plan of c(h) with its own mean and data values.
    z: (R, N, nv); data_vals: (nv, n_data). Returns (nv, R, N)."""
    z = np.asarray(z, dtype=np.float64)
    R, N, nv = z.shape
    L = mv_factor(A)
    out = []
    for v in range(nv):
        w = np.zeros((R, N))
        for u in range(v + 1):
            w = w + L[v, u] * z[:, :, u]
        cv = dataclasses.replace(case, mean=float(means[v]),
                                 data_vals=np.asarray(data_vals[v], dtype=np.float64)
                                 if case.n_data else case.data_vals)
        out.append(seqsim_sweep(cv, plan, w))
    return np.stack(out)


def draw_category(q, u):
    """Distributions 0.25 `rand(rng, Categorical(q))` (DiscreteNonParametric,
    not vendored): cp = q[0]; i = 0; while cp <= u and i < n - 1: i += 1,
    cp += q[i]. Vectorised over rows of q (R, n) and u (R,)."""
    q = np.asarray(q, dtype=np.float64)
    n = q.shape[1]
    cat = np.zeros(len(u), np.int64)
    cp = q[:, 0].copy()
    more = cp <= u
    for c in range(1, n):
        cat = np.where(more, c, cat)
        cp = np.where(more, cp + q[:, c], cp)
        more = more & (cp <= u)
    return cat


def indicator_probabilities(q):
    """seqsim.jl:306-310: normalize(clamp.(p, 0, 1), 1) per row, i.e. the
    clamped vector times the inverse of its 1-norm (LinearAlgebra's
    normalize, with its rescaling below 1/floatmax); the norm sums left to
    right."""
    q = np.where(q > 1.0, 1.0, np.where(q < 0.0, 0.0, q))
    s = np.abs(q[:, 0]).copy()
    for c in range(1, q.shape[1]):
        s = s + np.abs(q[:, c])
    delta = 1.0 / np.finfo(np.float64).max
    ed = np.finfo(np.float64).eps / delta
    big = s >= delta
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        out = np.where(big[:, None], q * (1.0 / s)[:, None], (q * ed) * (1.0 / (s * ed))[:, None])
    return out


def indicator_sweep(case: SeqsimCase, plan: Plan, u: np.ndarray, prob, data_cat):
    """SEQSIM of an IndicatorProcess whose function is A * c(h), given the plan
    of c(h) (seqsim.jl:98-138 with _probmodel/_conditional/_draw! of :292-310,
    342-347): in path order, p_c + sum_j lambda_j ([x_j == c] - p_c) per
    category, normalised, drawn with one uniform per node; prior draws use
    `prob`. u: (R, n_path). Returns (R, N) 0-based categories."""
    prob = np.asarray(prob, dtype=np.float64)
    u = np.atleast_2d(np.asarray(u, dtype=np.float64))
    R, N = u.shape[0], case.n_nodes
    X = np.zeros((N + case.n_data, R), np.int64)
    X[N:] = np.asarray(data_cat, dtype=np.int64)[:, None]
    for q, i in enumerate(case.path):
        prior = (plan.flags[i] & 3) != 0
        k = 0 if prior else plan.nfound[i]
        p = np.tile(prob, (R, 1))
        for j in range(k):
            x = X[plan.nbr[i, j]]
            lam = plan.lam[i, j]
            for c in range(len(prob)):
                p[:, c] = p[:, c] + lam * ((x == c).astype(np.float64) - prob[c])
        if not prior:
            p = indicator_probabilities(p)
        X[i] = draw_category(p, u[:, q])
    return X[:N].T.copy()


def simple_kriging_weights(case: SeqsimCase, kmax=None):
    """Simple kriging weights of every grid centroid over the data, restated in
    numpy independently of the C oracle: GeoStatsModels 0.13 `fitpredict(
    Kriging(f, mu), data, domain)` (the posterior mean of expectation/field/
    gaussian.jl:21-25 and the conditioning of fftsim.jl:99,148; GeoStatsModels
    is not vendored, its published algorithm is: the k nearest data of each
    centroid, within the search ball when there is one, then lambda = C^-1 c,
    variance = C(0) - lambda . c). Neighbours are ordered by (distance, data
    index). Returns (ids (N, k) with -1 padding, lambda (N, k), variance,
    ok), ok False where fewer than kmin data were found."""
    kmax = case.kmax if kmax is None else kmax
    X = grid_centroids(case.dims, case.origin, case.spacing)
    D = case.data_coords
    k = min(kmax, len(D))
    r2 = case.radius ** 2 if case.radius > 0 else np.inf
    var = np.full(len(X), np.nan)
    ids = np.full((len(X), k), -1, np.int64)
    lam = np.zeros((len(X), k))
    ok = np.zeros(len(X), bool)
    for i, x in enumerate(X):
        d2 = np.sum((D - x) ** 2, axis=1)
        order = np.lexsort((np.arange(len(D)), d2))[:k]
        order = order[d2[order] <= r2]
        if len(order) < max(case.kmin, 1):
            continue
        ids[i, :len(order)] = order
        xd = D[order]
        C = covariance_np(case.family, case.sill, case.nugget, case.ranges, xd, xd, case.rotation)
        c = covariance_np(case.family, case.sill, case.nugget, case.ranges, xd, x[None, :],
                          case.rotation)[:, 0]
        lam[i, :len(order)] = np.linalg.solve(C, c)
        var[i] = case.sill - lam[i, :len(order)] @ c
        ok[i] = True
    return ids, lam, var, ok


def simple_kriging_apply(ids, lam, ok, vals, mean):
    """mean + sum_j lambda_j (z_j - mean) for every row of `vals` (R, n_data);
    NaN where the node has no estimate (the reference's `missing`)."""
    vals = np.atleast_2d(np.asarray(vals, dtype=np.float64))
    used = ids >= 0
    g = vals[:, np.where(used, ids, 0)] - mean        # (R, N, k)
    out = mean + np.sum(np.where(used, lam, 0.0) * g, axis=2)
    out[:, ~ok] = np.nan
    return out


def simple_kriging_np(case: SeqsimCase, kmax=None):
    """Posterior mean and variance of the data values (simple_kriging_weights).
    Returns (mean, variance, neighbour ids)."""
    ids, lam, var, ok = simple_kriging_weights(case, kmax)
    mean = simple_kriging_apply(ids, lam, ok, case.data_vals, case.mean)[0]
    return mean, var, ids


# ---------------------------------------------------------------- FFTSIM ----
def fftsim_amplitude(dims, family, sill, nugget, ranges, origin=None, spacing=None, rotation=None):
    """fftsim.jl:79-91: C between the centroid of cell dims .÷ 2 (1-based) and
    every centroid, F = sqrt.(abs.(fft(fftshift(C)))), F[1] = 0. The grid is
    column-major, so the numpy array has axes (z, y, x)."""
    nd = len(dims)
    origin = np.zeros(nd) if origin is None else np.asarray(origin, dtype=np.float64)
    spacing = np.ones(nd) if spacing is None else np.asarray(spacing, dtype=np.float64)
    cents = grid_centroids(dims, origin, spacing)
    centre = origin + (np.array([d // 2 - 1 for d in dims]) + 0.5) * spacing
    C = covariance_np(family, sill, nugget, ranges, centre[None, :], cents, rotation)[0]
    F = np.sqrt(np.abs(np.fft.fftn(np.fft.fftshift(C.reshape(tuple(dims)[::-1])))))
    F.flat[0] = 0.0
    return F


def fftsim_unconditional(F, w, sill, mean):
    """fftsim.jl:123-134 for every row of w (R, N): P = F exp(i angle(fft(w))),
    Z = real(ifft(P)), Z = sqrt(sill / var(Z, mean=0)) Z + mean."""
    out = []
    for wr in np.atleast_2d(np.asarray(w, dtype=np.float64)):
        P = F * np.exp(1j * np.angle(np.fft.fftn(wr.reshape(F.shape))))
        Z = np.real(np.fft.ifftn(P)).ravel()
        s2 = np.sum(Z * Z) / (Z.size - 1)  # Statistics.var(Z, mean=0), corrected
        out.append(np.sqrt(sill / s2) * Z + mean)
    return np.array(out)


def fftsim_conditional(zu, zbar, cells, kweights, mean):
    """fftsim.jl:139-153: z = zbar + (zu - krige(zu[cells])), `kweights` the
    simple_kriging_weights of the cell centroids."""
    ids, lam, _, ok = kweights
    zu = np.atleast_2d(zu)
    return zbar + zu - simple_kriging_apply(ids, lam, ok, zu[:, cells], mean)


# ------------------------------------------------------ ensemble statistics ----
# ensembles.jl:42-52 reduce with Julia's Statistics (stdlib of Julia 1.10, not
# vendored): mean(itr) sums left to right and divides by the count; var(itr)
# runs Welford's recurrence M_k = M + (x - M) / k, S += (x - M)(x - M_k) and
# returns S / (n - 1); quantile(itr, p) sorts and, with alpha = beta = 1,
# takes h = fma(n, p, 1 - p), j = clamp(trunc(h), 1, n - 1),
# g = clamp(h - j, 0, 1), a + g (b - a) of the j-th and (j+1)-th values
# (1-based), or (1 - g) a + g b when one is not finite.
def _fma_exact(a, b, c):
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def ensemble_stat(reals, kind, param=0.0):
    """One statistic of every column of `reals` (R, N), in the operation order
    above (vectorised across cells, sequential across realizations)."""
    v = np.asarray(reals, dtype=np.float64)
    R = v.shape[0]
    if kind == "mean":
        acc = v[0].copy()
        for r in range(1, R):
            acc = acc + v[r]
        return acc / float(R)
    if kind == "var":
        M = v[0].copy()
        S = np.zeros_like(M)
        for r in range(1, R):
            nm = M + (v[r] - M) / float(r + 1)
            S = S + (v[r] - M) * (v[r] - nm)
            M = nm
        with np.errstate(invalid="ignore", divide="ignore"):
            return S / float(R - 1)
    if kind == "cdf":
        return np.sum(v <= param, axis=0) / float(R)
    if kind == "ccdf":
        return np.sum(v > param, axis=0) / float(R)
    if kind == "quantile":
        p = float(param)
        m = 1.0 + p * (1.0 - 1.0 - 1.0)
        h = _fma_exact(float(R), p, m)
        srt = np.sort(v, axis=0)
        if R == 1:
            a = b = srt[0]
            g = 0.0
        else:
            j = min(max(int(np.trunc(h)), 1), R - 1)
            g = min(max(h - j, 0.0), 1.0)
            a, b = srt[j - 1], srt[j]
        fin = np.isfinite(a) & np.isfinite(b)
        with np.errstate(invalid="ignore"):
            out = np.where(fin, a + g * (b - a), (1.0 - g) * a + g * b)
        out[np.isnan(v).any(axis=0)] = np.nan
        return out
    raise ValueError(kind)


def seqsim_reference_cost(case: SeqsimCase, z: np.ndarray, threads=1, n_nodes_limit=0):
    """Search + solve per node per realization (seqsim.jl:98-138 cost structure)."""
    st = case.struct()
    z = np.ascontiguousarray(z, dtype=np.float64)
    R = z.shape[0]
    out = np.zeros((R, case.n_nodes))
    n = lib().or_seqsim_reference_cost(ctypes.byref(st), R, _p(z, ctypes.c_double),
                                       _p(case.data_vals, ctypes.c_double),
                                       _p(out, ctypes.c_double), int(threads), int(n_nodes_limit))
    return out, int(n)


# ---------------------------------------------------------------------- LUSIM
def covariance_np(family, sill, nugget, ranges, xa, xb, rotation=None):
    """Pairwise C(x_a, x_b) = sill - gamma(h) (utils.jl:50-62), numpy exp."""
    xa = np.asarray(xa, dtype=np.float64)
    xb = np.asarray(xb, dtype=np.float64)
    delta = xa[:, None, :] - xb[None, :, :]
    if rotation is not None:
        delta = delta @ np.asarray(rotation, dtype=np.float64).reshape(3, 3)[:xa.shape[1], :xa.shape[1]]
    r = np.asarray(ranges, dtype=np.float64)[: xa.shape[1]]
    u = np.sqrt(np.sum((delta / r) ** 2, axis=-1))
    if family == SPHERICAL:
        shape = np.where(u < 1.0, 1.5 * u - 0.5 * u ** 3, 1.0)
    elif family == EXPONENTIAL:
        shape = 1.0 - np.exp(-3.0 * u)
    else:
        # the gaussian model's stabilising nugget (gsp_oracle.c, or_cov_u)
        nugget = nugget + GAUSSIAN_NUGGET
        shape = 1.0 - np.exp(-3.0 * u * u)
    gamma = (sill - nugget) * shape + nugget
    C = sill - gamma
    return np.where(u == 0.0, sill, C)


def grid_centroids(dims, origin=None, spacing=None):
    """Cell centroids of a CartesianGrid, column-major (x fastest), 0-based."""
    nd = len(dims)
    origin = np.zeros(nd) if origin is None else np.asarray(origin, dtype=np.float64)
    spacing = np.ones(nd) if spacing is None else np.asarray(spacing, dtype=np.float64)
    idx = np.indices(dims[::-1]).reshape(nd, -1)[::-1].T  # column-major order
    return origin + (idx + 0.5) * spacing


def nearest_init(centroids, data_coords):
    """NearestInit (nearest.jl:12-34): each datum goes to its nearest element;
    a later datum overwrites an earlier one. Ties go to the lower index."""
    centroids = np.asarray(centroids, dtype=np.float64)
    dest = np.empty(len(data_coords), dtype=np.int64)
    for i, x in enumerate(np.asarray(data_coords, dtype=np.float64)):
        d2 = np.sum((centroids - x) ** 2, axis=1)
        dest[i] = int(np.argmin(d2))  # argmin returns the first (lowest) index on ties
    return dest


def lusim_preprocess(centroids, dinds, z1, family, sill, nugget, ranges):
    """lusim.jl:66-107 for one variable: returns (d2, L22, sinds)."""
    import scipy.linalg as sla
    N = len(centroids)
    dinds = np.asarray(dinds, dtype=np.int64)
    sinds = np.setdiff1d(np.arange(N), dinds)
    sdom = centroids[sinds]
    C22 = covariance_np(family, sill, nugget, ranges, sdom, sdom)
    if len(dinds) == 0:
        L22 = sla.cholesky(C22, lower=True)
        return np.zeros(len(sinds)), L22, sinds
    ddom = centroids[dinds]
    C11 = covariance_np(family, sill, nugget, ranges, ddom, ddom)
    C12 = covariance_np(family, sill, nugget, ranges, ddom, sdom)
    L11 = sla.cholesky(C11, lower=True)
    B12 = sla.solve_triangular(L11, C12, lower=True)
    d2 = B12.T @ sla.solve_triangular(L11, np.asarray(z1, dtype=np.float64), lower=True)
    L22 = sla.cholesky(C22 - B12.T @ B12, lower=True)
    return d2, L22, sinds


def lusim_sample(d2, L22, dinds, sinds, z1, mean, w):
    """lusim.jl:145-175: w is (R, n_s); returns (R, N)."""
    w = np.asarray(w, dtype=np.float64)
    R = w.shape[0]
    N = len(dinds) + len(sinds)
    out = np.empty((R, N))
    z2 = d2[None, :] + w @ L22.T
    out[:, sinds] = z2
    if len(dinds):
        out[:, dinds] = np.asarray(z1)[None, :]
    else:
        out += mean
    return out


# ------------------------------------------------------ learned kriging weights
MLP_IN, MLP_HID, MLP_OUT = 128, 256, 64
MLP_SLOTS = 32
MLP_FEAT_SCALE, MLP_NUG_SCALE = 8.0, 10.0
MLP_DATA_SLOTS = 4                     # data neighbours a node may have
MLP_DFEAT = 3 * MLP_DATA_SLOTS + 1     # their offsets, then the nearest one's length


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def mlp_unpack(params):
    """Flat FP32 parameters -> [(W, b)] per layer (W: out x in)."""
    p = np.asarray(params, dtype=np.float32)
    shapes = [(MLP_HID, MLP_IN), (MLP_HID, MLP_HID), (MLP_HID, MLP_HID), (MLP_OUT, MLP_HID)]
    out, off = [], 0
    for o, i in shapes:
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((W, b))
    assert off == len(p)
    return out


def mlp_request_row(case: SeqsimCase, slot_cells):
    """The input row of a request for a surrogate of len(slot_cells) slots
    (its slot offsets in ranges times the feature scale, then the nugget
    feature), float32 as the device computes it: offset * (spacing / range)
    in float64, rounded once."""
    nd = case.ndim
    S = len(slot_cells)
    sp = list(case.spacing) + [1.0] * (3 - nd)
    rg = list(case.ranges) * nd if len(case.ranges) == 1 else list(case.ranges)
    row = np.zeros(MLP_IN, np.float32)
    for t in range(S):
        for d in range(3):
            c = float(slot_cells[t][d]) * (sp[d] / rg[d]) if d < nd else 0.0
            row[3 * t + d] = np.float32(c * MLP_FEAT_SCALE)
    row[3 * S] = np.float32(case.nugget / case.sill * MLP_NUG_SCALE)
    return row


def mlp_features(row, masks, S, dfeat=None):
    """Inputs of an S-slot surrogate for slot masks (bit s: neighbour in slot
    s) of one request's input row, and the nodes' data features (n,
    MLP_DFEAT) when given."""
    masks = np.asarray(masks, dtype=np.uint32)
    X = np.zeros((len(masks), MLP_IN), np.float32)
    occ = ((masks[:, None] >> np.arange(S, dtype=np.uint32)[None, :]) & 1).astype(np.float32)
    k = 3 * S
    X[:, :k] = (occ[:, :, None] * np.asarray(row[:k], np.float32).reshape(S, 3)[None, :, :]
                ).reshape(len(masks), k)
    X[:, k:] = np.asarray(row[k:], np.float32)[None, :]
    if dfeat is not None:
        X[:, k + 1:k + 1 + MLP_DFEAT] = np.asarray(dfeat, np.float32)
    return X


def gelu_tanh(z):
    z = z.astype(np.float32)
    u = np.float32(0.7978845608028654) * (z + np.float32(0.044715) * z * z * z)
    return z / (np.float32(1.0) + np.exp(np.float32(-2.0) * u))


def mlp_forward(params, row, masks, S, dfeat=None):
    """Outputs (n, S + 1 + MLP_DATA_SLOTS) of an S-slot surrogate: S slot
    weights, log(sigma^2 / sill) and the weights of the data slots."""
    h = bf16(mlp_features(row, masks, S, dfeat))
    layers = mlp_unpack(params)
    for li, (W, b) in enumerate(layers):
        z = h @ bf16(W).T + b
        h = bf16(gelu_tanh(z)) if li < 3 else z
    return h[:, :S + 1 + MLP_DATA_SLOTS]


def grid_slots(case: SeqsimCase, k=MLP_SLOTS):
    """The first k grid offsets (cells) of the search's offset table: squared
    distance, then linear index delta. The zero offset (the node itself, a
    candidate of the k-nearest-then-mask search that the mask always drops)
    comes first."""
    nd = case.ndim
    nx = case.dims[0]
    ny = case.dims[1] if nd > 1 else 1
    r = {1: 2 * k, 2: 5, 3: 3}[nd]
    rng = [range(-r, r + 1) if d < nd else range(0, 1) for d in range(3)]
    offs = [(dx, dy, dz) for dz in rng[2] for dy in rng[1] for dx in rng[0]]
    sp = list(case.spacing) + [1.0] * (3 - nd)
    offs.sort(key=lambda o: (sum((o[d] * sp[d]) ** 2 for d in range(nd)), o[0] + nx * (o[1] + ny * o[2])))
    return np.array(offs[:k])


def node_slot_masks(case: SeqsimCase, plan: Plan, nodes, slot_cells):
    """Slot mask and per-neighbour slot of each node (-1 mask when a neighbour
    is a datum or outside the slot offsets)."""
    masks, slots, _, _ = node_inputs(case, plan, nodes, slot_cells, 0)
    return masks, slots


def node_inputs(case: SeqsimCase, plan: Plan, nodes, slot_cells, data_slots):
    """The surrogate's view of each node: slot mask, the output index of each
    neighbour's weight (its slot, or S + 1 + q for the q-th datum among the
    neighbours), the data features (MLP_DFEAT float32) and C(datum q, node) /
    sill of the data slots. Mask -1 when a grid neighbour lies outside the
    slot offsets or more than `data_slots` neighbours are data. The data
    offsets are taken as the device takes them: in the kriging's scaled
    coordinates (alpha x physical, cell centres alpha (origin + (c + 1/2)
    spacing)) times 1 / (alpha range), scaled by MLP_FEAT_SCALE, in float64
    and rounded once; the distance is the float32 length of the first."""
    nd = case.ndim
    nx = case.dims[0]
    ny = case.dims[1] if nd > 1 else 1
    S = len(slot_cells)
    key = {tuple(o): s for s, o in enumerate(slot_cells)}
    N = case.n_nodes
    alpha = scale_factor(case)
    org = list(case.origin) + [0.0] * (3 - nd)
    sp = list(case.spacing) + [0.0] * (3 - nd)
    rg = list(case.ranges) * nd if len(case.ranges) == 1 else list(case.ranges)
    irs = [1.0 / (alpha * rg[e]) if e < nd else 1.0 for e in range(3)]
    masks = np.zeros(len(nodes), np.int64)
    slots = np.full((len(nodes), plan.kmax), -1, np.int64)
    dfeat = np.zeros((len(nodes), MLP_DFEAT), np.float32)
    cd = np.zeros((len(nodes), MLP_DATA_SLOTS))
    for t, i in enumerate(nodes):
        ci = (i % nx, (i // nx) % ny, i // (nx * ny))
        cen = [alpha * org[e] + (ci[e] + 0.5) * (alpha * sp[e]) if e < nd else 0.0
               for e in range(3)]
        m, q = 0, 0
        for j in range(plan.nfound[i]):
            a = int(plan.nbr[i, j])
            if a >= N:
                if q >= data_slots:
                    m = -1
                    break
                p = case.data_coords[a - N]
                x = [alpha * float(p[e]) if e < nd else 0.0 for e in range(3)]
                u = [(x[e] - cen[e]) * irs[e] for e in range(3)]
                for e in range(3):
                    dfeat[t, 3 * q + e] = np.float32(u[e] * MLP_FEAT_SCALE) if e < nd else 0.0
                h = math.sqrt(sum(v * v for v in u[:nd]))
                cd[t, q] = cov_u(case.family, case.sill, case.nugget, h) / case.sill
                slots[t, j] = S + 1 + q
                q += 1
                continue
            c = (a % nx - ci[0], (a // nx) % ny - ci[1], a // (nx * ny) - ci[2])
            s = key.get(c)
            if s is None:
                m = -1
                break
            slots[t, j] = s
            m |= 1 << s
        masks[t] = m
        if m >= 0 and q > 0:
            f = dfeat[t, :3]
            dfeat[t, 3 * MLP_DATA_SLOTS] = np.sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2],
                                                   dtype=np.float32)
    return masks, slots, dfeat, cd


def axis_variogram(fields, dims, axis, lags):
    """Experimental semivariogram of realizations (R, N) along one grid axis,
    averaged over realizations: gamma(h) = mean (x(i + h e) - x(i))^2 / 2."""
    R = fields.shape[0]
    shape = (R,) + tuple(reversed(dims))          # (R, nz, ny, nx) for column-major ids
    X = fields.reshape(shape)
    ax = len(dims) - axis                          # numpy axis of grid axis `axis`
    Xm = np.moveaxis(X, ax, -1)
    out = []
    for h in lags:
        d = Xm[..., h:] - Xm[..., :-h]
        out.append(0.5 * float(np.mean(d * d, dtype=np.float64)))
    return np.array(out)
