"""IndicatorProcess SEQSIM on the GPU (seqsim.jl:247-251,292-310,342-347) for a
coregionalized covariance A * c(h): realizations equal the oracle's indicator
sweep category for category (the same plan, the same operation order per node,
the same uniforms), and the reference's IndicatorProcess testset
(test/field.jl:181-204) through `rand`, with its shape transiograms and with
A * SphericalCovariance."""
import numpy as np
import pytest

import oracle as O
from helpers import gpu_problem, say
from geostatsprocesses import _lib
import geostatsprocesses as gsp

pytestmark = pytest.mark.gpu


def ind_case(seed, dims, n_data, kmax, radius=0.0, n_cat=3, family=O.SPHERICAL):
    rng = np.random.default_rng(seed)
    N = int(np.prod(dims))
    data = rng.uniform(0, 1, (n_data, len(dims))) * np.asarray(dims, float)
    cats = rng.integers(0, n_cat, n_data)
    case = O.SeqsimCase(dims=dims, family=family, sill=1.0, ranges=(8.0,), kmax=kmax,
                        radius=radius, data_coords=data, data_vals=cats.astype(np.float64),
                        path=rng.permutation(N))
    prob = rng.dirichlet(np.ones(n_cat))
    return case, prob, cats.astype(np.int32)


@pytest.mark.parametrize("spec", [
    dict(seed=1, dims=(40, 30), n_data=25, kmax=12, radius=10.0, n_cat=2),
    dict(seed=2, dims=(36, 28), n_data=40, kmax=16, n_cat=3),
    dict(seed=3, dims=(14, 12, 8), n_data=30, kmax=24, n_cat=4, family=O.EXPONENTIAL),
    dict(seed=4, dims=(50, 40), n_data=0, kmax=8, n_cat=8),
])
def test_indicator_sweep_matches_oracle(ctx, spec):
    case, prob, cats = ind_case(**spec)
    say("indicator sweep", spec)
    plan = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,))
    o = O.seqsim_plan(case)
    u = np.random.default_rng(spec["seed"] + 100).random((6, case.n_nodes))
    got = plan.sweep_indicator(u, prob, cats)
    ref = O.indicator_sweep(case, o, u, prob, cats)
    np.testing.assert_array_equal(got, ref)
    say("category frequencies", np.bincount(got.ravel(), minlength=len(prob)) / got.size,
        "prior", np.round(prob, 3))


def test_reference_indicator_testset_transiograms(ctx):
    """test/field.jl:181-204 as written: IndicatorProcess(SphericalTransiogram(
    range=35)) with two categories and with proportions (0.7, 0.2, 0.1), on a
    100 x 100 grid, unconditional and with data (integer and string labels)."""
    rng = np.random.default_rng(123)
    proc = gsp.IndicatorProcess(gsp.SphericalTransiogram(range=35.0))
    grid = gsp.CartesianGrid(100, 100)
    data = gsp.georef({"C": [1, 3, 1]}, [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, ctx=ctx)
    assert np.issubdtype(real["field"].dtype, np.integer)
    assert set(real["field"].tolist()) == {1, 2}
    real = gsp.rand(proc, grid, rng=rng, data=data, ctx=ctx)
    assert np.issubdtype(real["C"].dtype, np.integer)
    assert set(real["C"].tolist()) == {1, 3}
    rng = np.random.default_rng(123)
    proc = gsp.IndicatorProcess(gsp.SphericalTransiogram(range=35.0, proportions=(0.7, 0.2, 0.1)))
    real = gsp.rand(proc, grid, rng=rng, ctx=ctx)
    assert set(real["field"].tolist()) == {1, 2, 3}
    data = gsp.georef({"C": ["a", "b", "c"]}, [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, data=data, ctx=ctx)
    assert set(real["C"].tolist()) == {"a", "b", "c"}
    # the category frequencies follow the proportions
    ens = gsp.rand(proc, grid, 4, rng=rng, ctx=ctx)
    freq = np.bincount(np.concatenate([ens[i]["field"] for i in range(4)]) - 1) / 40000
    say("transiogram category frequencies", np.round(freq, 3))
    assert freq[0] > freq[2]


def test_reference_indicator_testset(ctx):
    """test/field.jl:181-204 with A * SphericalCovariance(range=35)."""
    rng = np.random.default_rng(123)
    A2 = np.array([[1.0, 0.5], [0.5, 1.0]])
    proc = gsp.IndicatorProcess(A2 * gsp.SphericalCovariance(range=35.0), [0.5, 0.5])
    grid = gsp.CartesianGrid(100, 100)
    data = gsp.georef({"C": [1, 3, 1]}, [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, ctx=ctx)
    assert np.issubdtype(real["field"].dtype, np.integer)
    assert set(real["field"].tolist()) == {1, 2}
    real = gsp.rand(proc, grid, rng=rng, data=data, ctx=ctx)
    assert np.issubdtype(real["C"].dtype, np.integer)
    assert set(real["C"].tolist()) == {1, 3}
    # three categories, string labels
    A3 = np.array([[1.0, 0.2, 0.1], [0.2, 1.0, 0.3], [0.1, 0.3, 1.0]])
    proc = gsp.IndicatorProcess(A3 * gsp.SphericalCovariance(range=35.0), [0.7, 0.2, 0.1])
    real = gsp.rand(proc, grid, rng=rng, ctx=ctx)
    assert set(real["field"].tolist()) == {1, 2, 3}
    data = gsp.georef({"C": ["a", "b", "c"]}, [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, data=data, ctx=ctx)
    assert set(real["C"].tolist()) == {"a", "b", "c"}
    ens = gsp.rand(proc, grid, 3, rng=rng, ctx=ctx)
    assert len(ens) == 3


def test_indicator_errors(ctx):
    with pytest.raises(AssertionError):
        gsp.IndicatorProcess(gsp.SphericalCovariance(range=5.0), [1.0])
    A2 = np.eye(2)
    with pytest.raises(AssertionError):
        gsp.IndicatorProcess(A2 * gsp.SphericalCovariance(range=5.0), [0.5, 0.3, 0.2])
    case, prob, cats = ind_case(1, (20, 20), 10, 8, n_cat=2)
    plan = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,))
    u = np.random.default_rng(0).random((1, 400))
    with pytest.raises(AssertionError):
        plan.sweep_indicator(u, [0.6, 0.6], cats)
    with pytest.raises(_lib.UnsupportedError):
        plan.sweep_indicator(u, np.full(9, 1 / 9), cats)
