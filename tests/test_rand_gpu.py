"""The reference's own field-simulation tests (test/field.jl:17-113), run through
the mirrored `rand` API on the GPU. Units are not modelled by this backend, so
the unit-only checks are omitted; eltype and data-honouring checks are kept."""
import numpy as np
import pytest

import oracle as O
import geostatsprocesses as gsp
from helpers import say

pytestmark = pytest.mark.gpu


def test_lusim_reference_cases(ctx):
    """test/field.jl:17-28 and 46-56 (without units)."""
    method = gsp.LUSIM()
    rng = np.random.default_rng(123)
    proc = gsp.GaussianProcess(gsp.SphericalCovariance(range=10.0))
    grid = gsp.CartesianGrid(100)
    data = gsp.georef({"Z": [0.0, 1.0, 0.0, 1.0, 0.0]}, [(0.0,), (25.0,), (50.0,), (75.0,), (100.0,)])
    real = gsp.rand(proc, grid, rng=rng, method=method)
    assert real.field.dtype == np.float64 and np.isfinite(real.field).all()
    real = gsp.rand(proc, grid, rng=rng, method=method, data=data)
    assert real.Z.dtype == np.float64
    # data reproduced at their NearestInit cells
    dest = gsp.NearestInit().destinations(grid, data)
    np.testing.assert_array_equal(real.Z[dest], data.Z)
    proc = gsp.GaussianProcess(gsp.SphericalCovariance(range=35.0), 0.0)
    grid = gsp.CartesianGrid(10, 10)
    data = gsp.georef({"Z": [1.0, 0.0, 1.0]}, [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, method=method, data=data)
    assert real.Z.dtype == np.float64
    # test/field.jl:46-56: GaussianCovariance(range=35) on the 10 x 10 grid, with
    # and without data (the gaussian model's stabilising nugget keeps it regular)
    proc = gsp.GaussianProcess(gsp.GaussianCovariance(range=35.0), 0.0)
    real = gsp.rand(proc, grid, rng=rng, method=method)
    assert real.field.dtype == np.float64 and np.isfinite(real.field).all()
    real = gsp.rand(proc, grid, rng=rng, method=method, data=data)
    assert real.Z.dtype == np.float64 and np.isfinite(real.Z).all()
    dest = np.asarray(gsp.NearestInit().destinations(grid, data))
    for d in np.unique(dest):
        assert real.Z[d] == data.Z[np.flatnonzero(dest == d)[-1]]


def test_lusim_reference_gaussian_2d(ctx):
    """test/field.jl:39-44: GaussianCovariance(range=10) on CartesianGrid(100, 100),
    a 10,000-cell covariance matrix factored whole; the result is a Float64
    field whose variance is about the sill."""
    rng = np.random.default_rng(123)
    proc = gsp.GaussianProcess(gsp.GaussianCovariance(range=10.0))
    grid = gsp.CartesianGrid(100, 100)
    real = gsp.rand(proc, grid, rng=rng, method=gsp.LUSIM())
    f = real.field
    assert f.dtype == np.float64 and f.shape == (10000,) and np.isfinite(f).all()
    say(f"gaussian 100x100 LUSIM: mean {f.mean():.3f} var {f.var():.3f}")
    assert 0.3 < f.var() < 3.0


def test_lusim_bivariate_reference_cases(ctx):
    """test/field.jl:29-37 and 58-70 (without units): cosimulation of A * c,
    unconditional (variables field1, field2) and with data (Cu, Zn)."""
    method = gsp.LUSIM()
    rng = np.random.default_rng(123)
    grid = gsp.CartesianGrid(100)
    # field.jl:31-37: cosimulation of [1 .95; .95 1] * GaussianCovariance(range=10)
    proc = gsp.GaussianProcess(np.array([[1.0, 0.95], [0.95, 1.0]])
                               * gsp.GaussianCovariance(range=10.0), [0.0, 0.0])
    real = gsp.rand(proc, grid, rng=rng, method=method)
    assert real.field1.dtype == np.float64 and real.field2.dtype == np.float64
    assert np.isfinite(real.field1).all() and np.isfinite(real.field2).all()
    proc = gsp.GaussianProcess(np.array([[1.0, 0.95], [0.95, 1.0]])
                               * gsp.SphericalCovariance(range=10.0), [0.0, 0.0])
    real = gsp.rand(proc, grid, rng=rng, method=method)
    assert real.field1.dtype == np.float64 and real.field2.dtype == np.float64
    func = np.array([[1.0, 0.8], [0.8, 1.0]]) * gsp.SphericalCovariance(range=35.0, sill=1.0)
    proc = gsp.GaussianProcess(func, [0.1, 0.2])
    grid = gsp.CartesianGrid(10, 10)
    data = gsp.georef({"Cu": [0.0, 0.1, 0.0], "Zn": [0.1, 0.0, 0.1]},
                      [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, method=method)
    assert real.field1.dtype == np.float64 and real.field2.dtype == np.float64
    real = gsp.rand(proc, grid, rng=rng, method=method, data=data)
    # the points lie outside the 10 x 10 unit grid and snap to one corner cell,
    # where the later datum wins
    dest = np.asarray(gsp.NearestInit().destinations(grid, data))
    for d in np.unique(dest):
        last = np.flatnonzero(dest == d)[-1]
        assert real.Cu[d] == data.Cu[last] and real.Zn[d] == data.Zn[last]
    ens = gsp.rand(proc, grid, 3, rng=rng, method=method, data=data)
    assert len(ens) == 3 and ens[2].Zn.dtype == np.float64


def test_lusim_ensemble_matches_single_draws(ctx):
    """The batched ensemble equals R sequential randsingle calls with the same stream."""
    proc = gsp.GaussianProcess(gsp.ExponentialCovariance(range=6.0, nugget=0.05), 0.4)
    grid = gsp.CartesianGrid(20, 15)
    ens = gsp.rand(proc, grid, 4, rng=np.random.default_rng(5), method=gsp.LUSIM())
    rng = np.random.default_rng(5)
    pre = gsp.preprocess(rng, proc, gsp.LUSIM(), gsp.NearestInit(), grid, None)
    for r in range(4):
        single = gsp.randsingle(rng, proc, gsp.LUSIM(), grid, None, pre)["field"]
        np.testing.assert_allclose(ens[r].field, single, rtol=0, atol=1e-12)


def test_seqsim_reference_case(ctx):
    """test/field.jl:73-86: SEQSIM(maxneighbors=3), data honoured at (25, 25)."""
    method = gsp.SEQSIM(maxneighbors=3)
    rng = np.random.default_rng(2017)
    proc = gsp.GaussianProcess(gsp.SphericalVariogram(range=35.0))
    grid = gsp.CartesianGrid((0.5, 0.5), (100.5, 100.5), dims=(100, 100))
    data = gsp.georef({"Z": [1.0, 0.0, 1.0]}, [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, method=method)
    assert real.field.dtype == np.float64 and np.isfinite(real.field).all()
    real = gsp.rand(proc, grid, rng=rng, method=method, data=data)
    assert real.Z.dtype == np.float64
    assert abs(real.Z[grid.linear_index(24, 24)] - 1.0) <= 1e-4


def test_seqsim_rand_equals_oracle(ctx):
    """rand(process, grid, R; method=SEQSIM(path=RandomPath)) equals the oracle run
    with the same path and normals."""
    proc = gsp.GaussianProcess(gsp.ExponentialVariogram(range=8.0, nugget=0.02), 0.3)
    grid = gsp.CartesianGrid(50, 40)
    rngd = np.random.default_rng(3)
    pts = rngd.uniform(0, 40, (30, 2))
    vals = rngd.standard_normal(30)
    data = gsp.georef({"Z": vals}, [tuple(p) for p in pts])
    method = gsp.SEQSIM(path=gsp.RandomPath(seed=11), maxneighbors=16)
    ens = gsp.rand(proc, grid, 6, rng=np.random.default_rng(4), data=data, method=method)
    path = gsp.RandomPath(seed=11).traverse(grid)
    case = O.SeqsimCase(dims=(50, 40), family=O.EXPONENTIAL, ranges=(8.0,), nugget=0.02, mean=0.3,
                        kmax=16, radius=8.0, data_coords=pts, data_vals=vals, path=path)
    plan = O.seqsim_plan(case)
    z = np.random.default_rng(4).standard_normal((6, 2000))
    ref = O.seqsim_sweep(case, plan, z)
    for r in range(6):
        np.testing.assert_array_equal(ens[r].Z, ref[r])
    say("rand SEQSIM equals oracle")


def test_seqsim_unsupported_paths(ctx):
    proc = gsp.GaussianProcess(gsp.SphericalVariogram(range=5.0))
    with pytest.raises(gsp.UnsupportedError):
        gsp.rand(proc, gsp.PointSet(np.random.default_rng(0).random((20, 2))),
                 method=gsp.SEQSIM())
