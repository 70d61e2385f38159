"""LUSIM parity on the GPU against the numpy/scipy oracle (lusim.jl:38-175)."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_err, say
import geostatsprocesses as gsp
from geostatsprocesses import _lib

pytestmark = pytest.mark.gpu


def lusim_problem(dims, family, sill, nugget, rng_, dinds, z1, mean=0.0, kind=_lib.COVARIANCE,
                  precision=_lib.LU_FP64, origin=None, spacing=None):
    p = _lib.GspLusimProblem()
    nd = len(dims)
    p.grid.ndim = nd
    for d in range(3):
        p.grid.dims[d] = dims[d] if d < nd else 1
        p.grid.origin[d] = (origin[d] if origin is not None else 0.0) if d < nd else 0.0
        p.grid.spacing[d] = (spacing[d] if spacing is not None else 1.0) if d < nd else 1.0
        p.model.ranges[d] = rng_[d] if d < nd else 1.0
    p.model.family = family
    p.model.kind = kind
    p.model.sill = sill
    p.model.nugget = nugget
    p.mean = mean
    dinds = np.ascontiguousarray(dinds, dtype=np.int64)
    z1 = np.ascontiguousarray(z1, dtype=np.float64)
    p.n_data = len(dinds)
    p.dinds = dinds.ctypes.data if len(dinds) else None
    p.z1 = z1.ctypes.data if len(dinds) else None
    p.precision = precision
    return p, (dinds, z1)


@pytest.mark.parametrize("fp32", [False, True])
def test_mfma_lane_layout(ctx, fp32):
    """16x16x4 MFMA: A[i][k] and B[k][j] sit in lane i + 16k / j + 16k; the
    accumulator register v of lane l holds row 4v + l/16 (FP64) or 4(l/16) + v
    (FP32), column l % 16 — the layout the LUSIM GEMM epilogue uses."""
    out = _lib.selftest_mfma(ctx, fp32)
    lane = np.arange(64)[:, None]
    v = np.arange(4)[None, :]
    row = 4 * (lane // 16) + v if fp32 else 4 * v + lane // 16
    col = np.broadcast_to(lane % 16, (64, 4))
    # row sums of A with A[i][k] = i + 16k + 1, column sums of B likewise
    np.testing.assert_array_equal(out[0], 4 * row + 100)
    np.testing.assert_array_equal(out[1], 4 * col + 100)


def run_case(ctx, dims, family, sill, nugget, ranges, n_data, R, seed, precision=_lib.LU_FP64,
             mean=0.0, tol=1e-10):
    rng = np.random.default_rng(seed)
    N = int(np.prod(dims))
    dinds = np.sort(rng.choice(N, n_data, replace=False)) if n_data else np.zeros(0, np.int64)
    z1 = rng.standard_normal(n_data)
    say(f"lusim dims={dims} fam={family} n_data={n_data} R={R} prec={precision}")
    p, keep = lusim_problem(dims, family, sill, nugget, ranges, dinds, z1, mean=mean,
                            precision=precision)
    plan = _lib.LusimPlan(ctx, p, keep=keep)
    cents = O.grid_centroids(dims)
    d2, L22, sinds = O.lusim_preprocess(cents, dinds, z1, family, sill, nugget, ranges)
    assert plan.n_sim == len(sinds) and plan.n_data == n_data
    gL, gd2 = plan.export()
    w = rng.standard_normal((R, len(sinds)))
    out = plan.sample(w)
    ref = O.lusim_sample(d2, L22, dinds, sinds, z1, mean, w)
    if n_data:
        np.testing.assert_array_equal(out[:, dinds], np.broadcast_to(z1, (R, n_data)))
    return dict(L=rel_err(gL, L22), d2=rel_err(gd2, d2) if n_data else 0.0,
                out=rel_err(out, ref), gL=gL, L22=L22, out_arr=out, ref=ref)


@pytest.mark.parametrize("dims,n_data", [((1,), 0), ((3,), 1), ((70,), 0), ((100,), 5),
                                          ((13, 11), 0), ((13, 11), 17), ((9, 7, 5), 12),
                                          ((40, 33), 64)])
def test_lusim_fp64_parity_small(ctx, dims, n_data):
    r = run_case(ctx, dims, O.EXPONENTIAL, 1.0, 0.01, (6.0, 5.0, 3.0)[:len(dims)], n_data, 5, 3)
    assert r["L"] <= 1e-12 and r["d2"] <= 1e-12 and r["out"] <= 1e-10


@pytest.mark.parametrize("fam", [O.SPHERICAL, O.GAUSSIAN])
def test_lusim_families(ctx, fam):
    r = run_case(ctx, (30, 30), fam, 2.0, 0.1, (8.0, 8.0), 25, 4, 9, mean=1.5)
    assert r["L"] <= 1e-12 and r["out"] <= 1e-10


def test_lusim_gaussian_without_nugget(ctx):
    """GaussianCovariance with nugget 0 (the reference's LUSIM testset, test/field.jl:
    39-56): the matrix is regular only through the stabilising nugget (1e-6), so
    kappa is about 1e7. Unconditional: the factor is checked by its backward error
    |L L' - C| / |C| <= 10 n eps (the bound of a stable Cholesky; the device and
    numpy covariances differ by a few ulp). Conditional: data exact and the
    realizations within kappa-scaled rounding (1e-6 relative) of scipy's."""
    r = run_case(ctx, (20, 20), O.GAUSSIAN, 1.0, 0.0, (10.0, 10.0), 0, 3, 21)
    gL = r["gL"]
    cents = O.grid_centroids((20, 20))
    C = O.covariance_np(O.GAUSSIAN, 1.0, 0.0, (10.0, 10.0), cents, cents)
    back = np.linalg.norm(gL @ gL.T - C) / np.linalg.norm(C)
    say(f"gaussian nugget 0: backward error {back:.3e}, L rel {r['L']:.3e}")
    assert back <= 10 * 400 * np.finfo(np.float64).eps
    assert r["out"] <= 1e-6
    r = run_case(ctx, (20, 20), O.GAUSSIAN, 1.0, 0.0, (10.0, 10.0), 12, 3, 22)
    say(f"gaussian nugget 0 with data: out rel {r['out']:.3e}")
    assert r["out"] <= 1e-6


def test_lusim_unconditional_adds_mean(ctx):
    r = run_case(ctx, (20, 20), O.SPHERICAL, 1.0, 0.0, (6.0, 6.0), 0, 3, 4, mean=0.7)
    assert r["out"] <= 1e-10


def test_config2_fp64_parity(ctx):
    """64x64, exponential covariance sill 1 range 20 nugget 0.01, 200 uniform
    points snapped by NearestInit (later data win), R=256."""
    rng = np.random.default_rng(1002)
    pts = rng.uniform(0, 64, (200, 2))
    vals = np.random.default_rng(1004).standard_normal(200)
    grid = gsp.CartesianGrid(64, 64)
    proc = gsp.GaussianProcess(gsp.ExponentialCovariance(sill=1.0, range=20.0, nugget=0.01))
    data = gsp.georef({"Z": vals}, [tuple(p) for p in pts])
    real, mask = gsp.initialize(proc, grid, data, gsp.NearestInit())
    dinds = np.flatnonzero(mask["Z"])
    z1 = real["Z"][dinds]
    say(f"config2 n_data={len(dinds)}")
    p, keep = lusim_problem((64, 64), O.EXPONENTIAL, 1.0, 0.01, (20.0, 20.0), dinds, z1)
    plan = _lib.LusimPlan(ctx, p, keep=keep)
    d2, L22, sinds = O.lusim_preprocess(grid.centroids(), dinds, z1, O.EXPONENTIAL, 1.0, 0.01,
                                        (20.0, 20.0))
    gL, gd2 = plan.export()
    assert rel_err(gL, L22) <= 1e-11
    assert rel_err(gd2, d2) <= 1e-11
    w = np.random.default_rng(1005).standard_normal((256, len(sinds)))
    out = plan.sample(w)
    ref = O.lusim_sample(d2, L22, dinds, sinds, z1, 0.0, w)
    per_real = np.max(np.abs(out - ref), axis=1) / np.max(np.abs(ref), axis=1)
    say(f"config2 fp64 max per-realization rel err {per_real.max():.3e}")
    assert per_real.max() <= 1e-10


def test_config2_mixed_precision_error(ctx):
    """The mixed form (FP32 MFMA trailing updates, FP64 diagonal tiles, one FP64
    refinement step of the conditioning solve) against the FP64 oracle on the 64 x 64 grid.
    Bounds: d2 comes from the refined solve, x1 = x0 + C11~^-1 (z1 - C11 x0), whose
    error is (kappa(C11) u32)^2 relative (kappa(C11) ~ 1e2 here): <= 1e-11. The
    factor is backward stable in FP32 accumulation: |L~L~' - S| / |S| <= 64 u32
    (K = 64 terms per update). Realizations carry kappa(S) times that backward
    error (kappa(S) = 284 with the 0.01 nugget): <= 1e-5 per realization."""
    rng = np.random.default_rng(1002)
    pts = rng.uniform(0, 64, (200, 2))
    grid = gsp.CartesianGrid(64, 64)
    dest = gsp.NearestInit().destinations(grid, gsp.georef({"Z": np.zeros(200)}, [tuple(p) for p in pts]))
    dinds = np.unique(dest)
    z1 = np.random.default_rng(1004).standard_normal(len(dinds))
    p, keep = lusim_problem((64, 64), O.EXPONENTIAL, 1.0, 0.01, (20.0, 20.0), dinds, z1,
                            precision=_lib.LU_MIXED)
    plan = _lib.LusimPlan(ctx, p, keep=keep)
    d2, L22, sinds = O.lusim_preprocess(grid.centroids(), dinds, z1, O.EXPONENTIAL, 1.0, 0.01,
                                        (20.0, 20.0))
    gL, gd2 = plan.export()
    u32 = 2.0 ** -24
    S = L22 @ L22.T
    back = np.linalg.norm(gL @ gL.T - S) / np.linalg.norm(S)
    w = np.random.default_rng(1005).standard_normal((64, len(sinds)))
    out = plan.sample(w)
    ref = O.lusim_sample(d2, L22, dinds, sinds, z1, 0.0, w)
    per_real = np.max(np.abs(out - ref), axis=1) / np.max(np.abs(ref), axis=1)
    say(f"config2 mixed: d2 rel {rel_err(gd2, d2):.3e}, backward {back:.3e}, "
        f"L rel {rel_err(gL, L22):.3e}, max per-realization {per_real.max():.3e}")
    assert rel_err(gd2, d2) <= 1e-11
    assert back <= 64 * u32
    assert per_real.max() <= 1e-5
    np.testing.assert_array_equal(out[:, dinds], np.broadcast_to(z1, (64, len(dinds))))


def test_lusim_blocked_sizes(ctx):
    """Panel edges of the blocked factorization: sizes around multiples of the
    64-wide panels and of the 16-wide sub-panels, with and without data, FP64
    and mixed."""
    for dims, nd in [((8, 8), 0), ((65,), 0), ((127,), 1), ((12, 16), 64), ((17, 16), 16),
                     ((11, 23), 40)]:
        r = run_case(ctx, dims, O.EXPONENTIAL, 1.0, 0.05, (5.0, 4.0)[:len(dims)], nd, 3, 31)
        assert r["L"] <= 1e-12 and r["d2"] <= 1e-12 and r["out"] <= 1e-10
        r = run_case(ctx, dims, O.EXPONENTIAL, 1.0, 0.05, (5.0, 4.0)[:len(dims)], nd, 3, 31,
                     precision=_lib.LU_MIXED)
        assert r["d2"] <= 1e-11 and r["out"] <= 1e-5


def test_lusim_rejects_variogram(ctx):
    p, keep = lusim_problem((10, 10), O.SPHERICAL, 1.0, 0.0, (4.0, 4.0), [], [],
                            kind=_lib.VARIOGRAM)
    with pytest.raises(_lib.ArgumentError):
        _lib.LusimPlan(ctx, p, keep=keep)


def test_lusim_not_positive_definite(ctx):
    # a gaussian covariance whose sill lies below the model's stabilising nugget
    # (1e-6): its off-diagonal covariances are negative and exceed the diagonal
    p, keep = lusim_problem((40, 40), O.GAUSSIAN, 1e-9, 0.0, (200.0, 200.0), [], [])
    with pytest.raises(_lib.PosDefException):
        _lib.LusimPlan(ctx, p, keep=keep)


def test_lusim_all_cells_are_data(ctx):
    dinds = np.arange(12)
    z1 = np.linspace(-1, 1, 12)
    p, keep = lusim_problem((4, 3), O.EXPONENTIAL, 1.0, 0.1, (2.0, 2.0), dinds, z1)
    plan = _lib.LusimPlan(ctx, p, keep=keep)
    assert plan.n_sim == 0
    out = plan.sample(np.zeros((2, 0)))
    np.testing.assert_array_equal(out, np.broadcast_to(z1, (2, 12)))


def test_lusim_bivariate_matches_oracle(ctx):
    """Bivariate LUSIM (lusim.jl:64-126, 132-143) for A * c with data on both
    variables: each variable is the oracle's univariate LUSIM with the
    covariance A[j, j] c(h), the second one driven by rho w1 + sqrt(1 - rho^2) w2."""
    A = np.array([[1.0, 0.8], [0.8, 2.0]])
    proc = gsp.GaussianProcess(A * gsp.SphericalCovariance(range=9.0, sill=1.5, nugget=0.05),
                               [0.1, -0.3])
    grid = gsp.CartesianGrid(14, 11)
    pts = [(2.5, 3.5), (7.5, 8.5), (11.5, 2.5), (4.5, 9.5)]
    data = gsp.georef({"Cu": [0.3, -0.2, 0.5, 0.0], "Zn": [1.0, 0.4, -0.6, 0.2]}, pts)
    R = 6
    ens = gsp.rand(proc, grid, R, rng=np.random.default_rng(21), method=gsp.LUSIM(), data=data)
    real, mask = gsp.initialize(proc, grid, data, gsp.NearestInit())
    n_s = grid.nelements() - int(mask["Cu"].sum())
    w = np.random.default_rng(21).standard_normal((R, 2 * n_s))
    rho = 0.8 / np.sqrt(2.0)
    for j, (var, wj) in enumerate([("Cu", w[:, :n_s]),
                                   ("Zn", rho * w[:, :n_s] + np.sqrt(1 - rho ** 2) * w[:, n_s:])]):
        dinds = np.flatnonzero(mask[var])
        z1 = real[var][dinds]
        d2, L22, sinds = O.lusim_preprocess(grid.centroids(), dinds, z1, O.SPHERICAL,
                                            1.5 * A[j, j], 0.05 * A[j, j], (9.0, 9.0))
        ref = O.lusim_sample(d2, L22, dinds, sinds, z1, 0.0, wj)
        got = np.stack([ens[r][var] for r in range(R)])
        say(f"bivariate {var}: rel err {rel_err(got, ref):.2e}")
        assert rel_err(got, ref) <= 1e-10
        np.testing.assert_array_equal(got[:, dinds], np.broadcast_to(z1, (R, len(dinds))))


def test_lusim_bivariate_cross_covariance(ctx):
    """The construction reproduces the coregionalization: over many
    unconditional realizations Cov(z1(x), z2(y)) = A12 C(x - y) and
    Var z2 = A22 C(0), which fixes rho and the sqrt(1 - rho^2) factor."""
    A = np.array([[1.0, 0.7], [0.7, 2.0]])
    base = gsp.ExponentialCovariance(range=6.0, sill=1.0)
    proc = gsp.GaussianProcess(A * base, [0.0, 0.0])
    grid = gsp.CartesianGrid(12, 10)
    R = 4000
    ens = gsp.rand(proc, grid, R, rng=np.random.default_rng(8), method=gsp.LUSIM())
    z1 = np.stack([ens[r]["field1"] for r in range(R)])
    z2 = np.stack([ens[r]["field2"] for r in range(R)])
    c0 = np.mean(z1 * z2)  # lag 0, every cell
    # lag 1 along x (cells are x-fastest: row-major (ny, nx) = (10, 12))
    c1 = np.mean(z1.reshape(R, 10, 12)[:, :, :-1] * z2.reshape(R, 10, 12)[:, :, 1:])
    v2 = np.mean(z2 * z2)
    C1 = float(O.cov_u(O.EXPONENTIAL, 1.0, 0.0, 1.0 / 6.0))
    say(f"cross cov lag0 {c0:.3f} (expect {0.7:.3f}), lag1 {c1:.3f} (expect {0.7 * C1:.3f}), "
        f"var z2 {v2:.3f} (expect 2)")
    assert abs(c0 - 0.7) < 0.05 and abs(c1 - 0.7 * C1) < 0.05 and abs(v2 - 2.0) < 0.1
