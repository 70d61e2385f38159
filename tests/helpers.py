"""Shared test helpers: build C-ABI problems from oracle cases, synthetic
workloads of the benchmark configurations, comparison metrics."""
import ctypes
import sys

import numpy as np

import oracle as O
from geostatsprocesses import _lib


def gpu_problem(case: O.SeqsimCase, device_pointers=False, data_ptr=None, path_ptr=None):
    p = _lib.GspSeqsimProblem()
    nd = case.ndim
    p.grid.ndim = nd
    for d in range(3):
        p.grid.dims[d] = int(case.dims[d]) if d < nd else 1
        p.grid.origin[d] = float(case.origin[d]) if d < nd else 0.0
        p.grid.spacing[d] = float(case.spacing[d]) if d < nd else 1.0
        p.model.ranges[d] = float(case.ranges[d]) if d < nd else 1.0
    p.model.family = case.family
    p.model.kind = _lib.VARIOGRAM
    p.model.sill = case.sill
    p.model.nugget = case.nugget
    if case.rotation is not None:
        rot = np.asarray(case.rotation, dtype=np.float64).reshape(3, 3)
        for q in range(9):
            p.model.rotation[q] = rot.flat[q]
        p.model.has_rotation = 1
    p.mean = case.mean
    p.n_data = case.n_data
    p.data_coords = data_ptr if data_ptr is not None else (case.data_coords.ctypes.data if case.n_data else None)
    p.n_path = len(case.path)
    p.path = path_ptr if path_ptr is not None else case.path.ctypes.data
    p.minneighbors = case.kmin
    p.maxneighbors = case.kmax
    p.radius = case.radius
    p.search_mode = case.mode
    p.precision = case.precision
    p.device_pointers = int(device_pointers)
    return p


def gpu_plan(ctx, case: O.SeqsimCase):
    return _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,))


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


# ------------------------------------------------- benchmark configurations --
def config1(conditional=True, seed=1001, R=10):
    """100x100 unit grid, spherical sill 1 range 10 nugget 0, maxneighbors 12,
    MetricBall(30), 50 uniform data with N(0,1) values, random path."""
    rng = np.random.default_rng(seed)
    path = rng.permutation(10000)
    nd = 50 if conditional else 0
    data = np.random.default_rng(seed + 1).uniform(0, 100, (nd, 2))
    vals = np.random.default_rng(seed + 2).standard_normal(nd)
    case = O.SeqsimCase(dims=(100, 100), family=O.SPHERICAL, sill=1.0, ranges=(10.0,), nugget=0.0,
                        kmax=12, radius=30.0, data_coords=data, data_vals=vals, path=path)
    z = np.random.default_rng(seed + 3).standard_normal((R, 10000))
    return case, z


def config3(dims=(128, 128, 64), n_data=2000, seed=1003, precision=O.FP32_REFINE):
    """Anisotropic exponential sill 1 ranges (40,40,8) nugget 0.05, maxneighbors 24,
    :range neighbourhood (radius = range(f) = 40), uniform data, random path."""
    rng = np.random.default_rng(seed)
    N = int(np.prod(dims))
    path = rng.permutation(N)
    data = np.random.default_rng(seed + 1).uniform(0, 1, (n_data, 3)) * np.asarray(dims, dtype=float)
    vals = np.random.default_rng(seed + 2).standard_normal(n_data)
    return O.SeqsimCase(dims=dims, family=O.EXPONENTIAL, sill=1.0, ranges=(40.0, 40.0, 8.0),
                        nugget=0.05, kmax=24, radius=40.0, data_coords=data, data_vals=vals,
                        path=path, precision=precision)


def config5(n_req=512, side_lo=32, side_hi=512, seed=1005, precision=O.FP32_REFINE):
    """Mixed batch of independent 2D requests: square grids with side
    log-uniform in [side_lo, side_hi]; spherical/exponential/gaussian
    variograms (sill 1) with range uniform in [5, 60] cells and nugget uniform
    in [0, 0.1]; 0-500 uniform data with N(0,1) values; maxneighbors from
    {8, 16, 32}; :range neighbourhood; random paths."""
    rng = np.random.default_rng(seed)
    cases = []
    for q in range(n_req):
        side = int(round(np.exp(rng.uniform(np.log(side_lo), np.log(side_hi)))))
        fam = int(rng.integers(0, 3))
        rang = float(rng.uniform(5.0, 60.0))
        nug = float(rng.uniform(0.0, 0.1))
        nd = int(rng.integers(0, 501))
        kmax = int(rng.choice([8, 16, 32]))
        data = rng.uniform(0, side, (nd, 2))
        vals = rng.standard_normal(nd)
        path = rng.permutation(side * side)
        cases.append(O.SeqsimCase(dims=(side, side), family=fam, sill=1.0, ranges=(rang,),
                                  nugget=nug, kmax=kmax, radius=rang, data_coords=data,
                                  data_vals=vals, path=path, precision=precision))
    return cases


def batch_arrays(cases, node_offset, z_list=None):
    """Concatenated data values (and normals (R, N_total)) of a batch."""
    vals = np.concatenate([c.data_vals for c in cases]) if cases else np.zeros(0)
    if z_list is None:
        return vals
    R = z_list[0].shape[0]
    z = np.empty((R, int(node_offset[-1])))
    for q, zq in enumerate(z_list):
        z[:, node_offset[q]:node_offset[q + 1]] = zq
    return vals, z


def local_plan(g, q, plan, cases):
    """Request q's rows of an exported batch plan with ids made local
    (grid ids 0..N-1, data N + d), columns cut to its maxneighbors."""
    no, do = plan.node_offset, plan.data_offset
    NT = int(no[-1])
    N = int(no[q + 1] - no[q])
    k = min(cases[q].kmax, N)
    rows = slice(int(no[q]), int(no[q + 1]))
    nbr = g["nbr"][rows].astype(np.int64)
    loc = np.where(nbr < 0, -1, np.where(nbr < NT, nbr - no[q], nbr - NT - do[q] + N))
    return dict(nfound=g["nfound"][rows], nbr=loc[:, :k], nbr_tail=loc[:, k:],
                lam=g["lam"][rows][:, :k], lam_tail=g["lam"][rows][:, k:],
                sigma=g["sigma"][rows], flags=g["flags"][rows], level=g["level"][rows])
