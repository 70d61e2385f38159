"""Kriging posterior mean on the GPU (expectation/field/gaussian.jl:21-25): the
data-only plan against the oracle bit for bit, the mean and variance against
the independent numpy simple kriging, and the reference's Expectation
testset (test/field.jl:156-178) through the package API."""
import numpy as np
import pytest

import oracle as O
from helpers import gpu_problem, say
from geostatsprocesses import _lib
import geostatsprocesses as gsp

pytestmark = pytest.mark.gpu

JULIA_RTOL = float(np.sqrt(np.finfo(float).eps))  # Base.isapprox default


def data_case(seed, dims=(30, 24, 10), n_data=300, kmax=10, radius=0.0, kmin=1,
              precision=O.FP64, family=O.SPHERICAL, spread=(-0.05, 1.05)):
    rng = np.random.default_rng(seed)
    data = rng.uniform(*spread, (n_data, len(dims))) * np.asarray(dims, float)
    return O.SeqsimCase(dims=dims, family=family, sill=1.3, nugget=0.05,
                        ranges=(9.0, 7.0, 4.0)[:len(dims)], mean=0.25, data_coords=data,
                        data_vals=rng.standard_normal(n_data), kmin=kmin, kmax=kmax,
                        radius=radius, mode=O.DATA_ONLY, precision=precision,
                        path=rng.permutation(int(np.prod(dims))))


@pytest.mark.parametrize("spec", [
    dict(seed=1),
    dict(seed=2, kmax=24, family=O.EXPONENTIAL),
    dict(seed=3, kmax=40, n_data=120, family=O.GAUSSIAN),
    dict(seed=4, kmax=12, radius=3.0, kmin=3),
    dict(seed=5, dims=(64, 48), n_data=500, kmax=16),
    dict(seed=6, kmax=10, precision=O.FP32_REFINE),
    dict(seed=7, kmax=1, n_data=30),
    # rings of hash cells: data clustered in one corner, data far outside the
    # grid, a 1D grid, more data than nodes
    dict(seed=8, n_data=200, spread=(0.0, 0.2)),
    dict(seed=9, n_data=200, kmax=20, spread=(-2.0, 3.0)),
    dict(seed=10, dims=(300,), n_data=17, kmax=6),
    dict(seed=11, dims=(20, 16), n_data=2000, kmax=12),
    # the tile search (kmax <= 16): data far outside the grid, fewer data than
    # kmax (tiles widen until every cell is visited), partial tiles on every axis
    dict(seed=12, n_data=200, kmax=12, spread=(-2.0, 3.0)),
    dict(seed=13, dims=(40, 36, 12), n_data=5, kmax=8),
    dict(seed=14, dims=(17, 9, 5), n_data=400, kmax=16, family=O.EXPONENTIAL),
])
def test_data_only_plan_matches_oracle(ctx, spec):
    case = data_case(**spec)
    say("data-only plan", spec)
    plan = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,))
    g = plan.export()
    o = O.seqsim_plan(case)
    for k in ("nfound", "nbr", "lam", "sigma", "flags", "level"):
        np.testing.assert_array_equal(g[k], getattr(o, k), err_msg=k)
    assert np.all(g["level"] == 0)
    z = np.zeros((1, case.n_nodes))
    np.testing.assert_array_equal(plan.sweep(z, case.data_vals), O.seqsim_sweep(case, o, z))


TILE_SPECS = [dict(seed=1), dict(seed=4, kmax=12, radius=3.0, kmin=3), dict(seed=8, n_data=200, spread=(0.0, 0.2)),
              dict(seed=10, dims=(300,), n_data=17, kmax=6), dict(seed=13, dims=(40, 36, 12), n_data=5, kmax=8)]


@pytest.mark.parametrize("spec", TILE_SPECS)
def test_tile_and_wavefront_search_agree(ctx, spec, monkeypatch):
    """The data-only plans of kmax <= 16 search by tiles of nodes; the
    wavefront search (GSP_TILE_SEARCH=0) must give the same plan, and both the
    oracle's."""
    case = data_case(**spec)
    say("tile vs wavefront search", spec)
    tile = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,)).export()
    monkeypatch.setenv("GSP_TILE_SEARCH", "0")
    wave = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,)).export()
    o = O.seqsim_plan(case)
    for k in ("nfound", "nbr", "lam", "sigma", "flags"):
        np.testing.assert_array_equal(wave[k], getattr(o, k), err_msg=k)
        np.testing.assert_array_equal(tile[k], wave[k], err_msg=k)


def test_data_only_batch(ctx):
    """Several data-only requests in one plan (1D, 2D and 3D grids, each its
    own run of tiles): every request gives what the oracle gives for it alone."""
    from test_batch_gpu import check_batch
    cases = [data_case(21), data_case(22, dims=(64, 48), n_data=500, kmax=16),
             data_case(23, dims=(300,), n_data=17, kmax=6),
             data_case(24, dims=(17, 9, 5), n_data=400, kmax=12, radius=2.5)]
    check_batch(ctx, cases, R=3, seed=2)


def test_posterior_mean_matches_numpy_simple_kriging(ctx):
    case = data_case(11, dims=(40, 30), n_data=200, kmax=10)
    grid = gsp.CartesianGrid(case.dims)
    proc = gsp.GaussianProcess(gsp.SphericalVariogram(sill=1.3, nugget=0.05, ranges=(9.0, 7.0)),
                               0.25)
    data = gsp.georef({"Z": case.data_vals}, case.data_coords)
    out = gsp.mean(proc, grid, data=data, prob=True, ctx=ctx)
    ref, var, _ = O.simple_kriging_np(case)
    err = np.abs(out["Z"] - ref).max()
    say(f"posterior mean vs numpy simple kriging: max abs err {err:.2e}")
    np.testing.assert_allclose(out["Z"], ref, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out["Z_variance"], var, rtol=1e-9, atol=1e-12)


def test_reference_expectation_testset(ctx):
    """test/field.jl:156-178."""
    proc = gsp.GaussianProcess(gsp.SphericalVariogram(range=35.0))
    grid = gsp.CartesianGrid((0.5, 0.5), (100.5, 100.5), dims=(100, 100))
    mval = gsp.mean(proc, grid)
    assert np.all(mval["field"] == 0)
    data = gsp.georef({"Z": [1.0, 0.0, 1.0]}, [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    mval = gsp.mean(proc, grid, data=data, ctx=ctx)
    for (i, j), v in {(25, 25): 1.0, (50, 75): 0.0, (75, 50): 1.0}.items():
        got = mval["Z"][grid.linear_index(i - 1, j - 1)]
        assert got == pytest.approx(v, rel=JULIA_RTOL, abs=0.0)
    e0 = gsp.expectedvalue(proc, grid)
    np.testing.assert_array_equal(e0["field"], gsp.mean(proc, grid)["field"])
    e1 = gsp.expectedvalue(proc, grid, data=data, ctx=ctx)
    np.testing.assert_array_equal(e1["Z"], mval["Z"])


def test_posterior_mean_options(ctx):
    """A search ball leaves far nodes missing; neighbors=false uses every
    datum; several columns share one set of weights."""
    rng = np.random.default_rng(3)
    grid = gsp.CartesianGrid((50, 40))
    coords = rng.uniform(0, 1, (25, 2)) * [50, 40]
    vals = rng.standard_normal((25, 2))
    proc = gsp.GaussianProcess(gsp.ExponentialVariogram(range=12.0), 0.0)
    data = gsp.georef({"a": vals[:, 0], "b": vals[:, 1]}, coords)
    out = gsp.posteriormean(proc, grid, data, neighborhood=gsp.MetricBall(4.0), minneighbors=2,
                            ctx=ctx)
    case = O.SeqsimCase(dims=(50, 40), family=O.EXPONENTIAL, ranges=(12.0,), data_coords=coords,
                        data_vals=vals[:, 0], kmin=2, kmax=10, radius=4.0, mode=O.DATA_ONLY)
    ref, _, _ = O.simple_kriging_np(case)
    miss = np.isnan(ref)
    assert miss.any() and (~miss).any()
    np.testing.assert_array_equal(np.isnan(out["a"]), miss)
    np.testing.assert_allclose(out["a"][~miss], ref[~miss], rtol=1e-9, atol=1e-12)
    case.data_vals = np.ascontiguousarray(vals[:, 1])
    ref_b, _, _ = O.simple_kriging_np(case)
    np.testing.assert_allclose(out["b"][~miss], ref_b[~miss], rtol=1e-9, atol=1e-12)
    full = gsp.posteriormean(proc, grid, data, neighbors=False, ctx=ctx)
    case = O.SeqsimCase(dims=(50, 40), family=O.EXPONENTIAL, ranges=(12.0,), data_coords=coords,
                        data_vals=vals[:, 0], kmax=25, mode=O.DATA_ONLY)
    ref, _, _ = O.simple_kriging_np(case)
    np.testing.assert_allclose(full["a"], ref, rtol=1e-9, atol=1e-12)


def test_posterior_mean_missing_values_rejected(ctx):
    proc = gsp.GaussianProcess(gsp.SphericalVariogram(range=5.0))
    grid = gsp.CartesianGrid((10, 10))
    data = gsp.georef({"Z": [1.0, np.nan]}, [(1.0, 1.0), (3.0, 4.0)])
    with pytest.raises(_lib.UnsupportedError):
        gsp.mean(proc, grid, data=data, ctx=ctx)


def test_masked_search_without_ball_matches_oracle(ctx):
    """Masked k nearest with no search ball: the first nodes of the path find
    fewer than kmax simulated nodes and search the data without a bound."""
    rng = np.random.default_rng(4)
    dims = (18, 14, 6)
    data = rng.uniform(0.0, 0.3, (60, 3)) * np.asarray(dims, float)
    case = O.SeqsimCase(dims=dims, family=O.EXPONENTIAL, ranges=(6.0,), data_coords=data,
                        data_vals=rng.standard_normal(60), kmax=12, mode=O.MASKED_KNN,
                        path=rng.permutation(int(np.prod(dims))))
    g = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,)).export()
    o = O.seqsim_plan(case)
    for k in ("nfound", "nbr", "lam", "sigma", "flags", "level"):
        np.testing.assert_array_equal(g[k], getattr(o, k), err_msg=k)
