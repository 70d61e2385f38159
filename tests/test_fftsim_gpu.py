"""FFTSIM on the GPU (src/simulation/field/fftsim.jl:54-156): the amplitude
spectrum, unconditional and conditional realizations against the numpy oracle
for the same white noise, and the reference's FFTSIM testset
(test/field.jl:115-154) through the package API.

Tolerance: the transforms (rocFFT here, numpy's pocketfft in the oracle)
round differently, and F = sqrt|X| turns an absolute rounding error e of the
spectrum X into about sqrt(e) where the spectrum is near zero (smooth
covariances such as the Gaussian reach the rounding floor). With
e ~ eps max|X| sqrt(N), F and the realizations agree to
fft_tol = 8 sqrt(eps max(F)^2 sqrt(N)) (1e-6 for a Gaussian range 10 on 100
cells, 1e-13 where the spectrum stays well above the floor)."""
import numpy as np
import pytest

import oracle as O
from helpers import say
from geostatsprocesses import _lib
import geostatsprocesses as gsp

pytestmark = pytest.mark.gpu

def fft_problem(dims, family, sill, nugget, ranges, mean=0.0, origin=None, spacing=None):
    nd = len(dims)
    p = _lib.GspFftsimProblem()
    p.grid.ndim = nd
    for d in range(3):
        p.grid.dims[d] = dims[d] if d < nd else 1
        p.grid.origin[d] = origin[d] if origin is not None and d < nd else 0.0
        p.grid.spacing[d] = spacing[d] if spacing is not None and d < nd else 1.0
        p.model.ranges[d] = ranges[d] if d < nd else 1.0
    p.model.family = family
    p.model.kind = _lib.VARIOGRAM
    p.model.sill = sill
    p.model.nugget = nugget
    p.mean = mean
    p.minneighbors = 1
    p.maxneighbors = 26
    return p


def fft_tol(F):
    return 8.0 * np.sqrt(np.finfo(float).eps * float(F.max()) ** 2 * np.sqrt(F.size))


@pytest.mark.parametrize("spec", [
    dict(dims=(100,), family=O.GAUSSIAN, ranges=(10.0,)),
    dict(dims=(101,), family=O.EXPONENTIAL, ranges=(7.0,)),
    dict(dims=(48, 40), family=O.SPHERICAL, ranges=(9.0, 6.0), spacing=(1.0, 1.5)),
    dict(dims=(45, 33), family=O.GAUSSIAN, ranges=(8.0, 8.0), nugget=0.1),
    dict(dims=(16, 12, 10), family=O.EXPONENTIAL, ranges=(5.0, 4.0, 3.0), sill=2.0, mean=1.5),
])
def test_fftsim_unconditional_matches_oracle(ctx, spec):
    spec = dict(spec)
    dims, fam, ranges = spec["dims"], spec["family"], spec["ranges"]
    sill, nug, mean = spec.get("sill", 1.0), spec.get("nugget", 0.0), spec.get("mean", 0.0)
    sp = spec.get("spacing")
    say("fftsim", spec)
    plan = _lib.FftsimPlan(ctx, fft_problem(dims, fam, sill, nug, ranges, mean, spacing=sp))
    F_gpu, zbar = plan.export()
    assert zbar is None
    F = O.fftsim_amplitude(dims, fam, sill, nug, ranges, spacing=sp)
    half = F[..., : dims[0] // 2 + 1]
    np.testing.assert_allclose(F_gpu, half, rtol=0, atol=fft_tol(F))
    N = int(np.prod(dims))
    w = np.random.default_rng(sum(dims)).random((3, N))
    out = plan.sample(w)
    ref = O.fftsim_unconditional(F, w, sill, mean)
    err = np.abs(out - ref).max()
    say(f"max |gpu - oracle| {err:.2e}")
    np.testing.assert_allclose(out, ref, rtol=0, atol=fft_tol(F) * np.sqrt(sill))
    np.testing.assert_allclose(np.sum((out - mean) ** 2, axis=1) / (N - 1), sill, rtol=1e-12)


def test_fftsim_batched_transforms(ctx):
    """More realizations than one batch of transforms holds: the batched and
    the single-realization transform plans give the same realizations."""
    dims = (20, 16)
    plan = _lib.FftsimPlan(ctx, fft_problem(dims, O.GAUSSIAN, 1.0, 0.0, (5.0, 5.0)))
    w = np.random.default_rng(9).random((70, 320))
    out = plan.sample(w)
    F = O.fftsim_amplitude(dims, O.GAUSSIAN, 1.0, 0.0, (5.0, 5.0))
    np.testing.assert_allclose(out, O.fftsim_unconditional(F, w, 1.0, 0.0), rtol=0,
                               atol=fft_tol(F))
    np.testing.assert_allclose(plan.sample(w[65:67]), out[65:67], rtol=0, atol=fft_tol(F))


def cond_setup(dims, family, ranges, coords, vals, kmax, origin=None, mean=0.0, radius=0.0,
               kmin=1):
    cents = O.grid_centroids(dims, origin)
    cells = np.unique(O.nearest_init(cents, coords))
    dcase = O.SeqsimCase(dims=dims, origin=origin, family=family, ranges=ranges, mean=mean,
                         data_coords=coords, data_vals=vals, kmin=kmin, kmax=kmax, radius=radius,
                         mode=O.DATA_ONLY)
    zbar, _, _ = O.simple_kriging_np(dcase)
    ccase = O.SeqsimCase(dims=dims, origin=origin, family=family, ranges=ranges, mean=mean,
                         data_coords=cents[cells], kmin=kmin, kmax=kmax, radius=radius,
                         mode=O.DATA_ONLY)
    return cells, zbar, O.simple_kriging_weights(ccase)


def test_fftsim_conditional_matches_oracle(ctx):
    rng = np.random.default_rng(12)
    dims = (40, 30)
    coords = rng.uniform(0, 1, (25, 2)) * [40, 30]
    vals = rng.standard_normal(25)
    cells, zbar, kw = cond_setup(dims, O.EXPONENTIAL, (8.0, 8.0), coords, vals, 10, mean=0.4)
    p = fft_problem(dims, O.EXPONENTIAL, 1.0, 0.0, (8.0, 8.0), mean=0.4)
    p.n_data = 25
    p.data_coords = coords.ctypes.data
    p.data_vals = vals.ctypes.data
    p.n_cells = len(cells)
    p.cells = cells.ctypes.data
    p.maxneighbors = 10
    plan = _lib.FftsimPlan(ctx, p, keep=(coords, vals, cells))
    _, zbar_gpu = plan.export()
    np.testing.assert_allclose(zbar_gpu, zbar, rtol=1e-9, atol=1e-12)
    w = rng.random((4, 1200))
    out = plan.sample(w)
    F = O.fftsim_amplitude(dims, O.EXPONENTIAL, 1.0, 0.0, (8.0, 8.0))
    zu = O.fftsim_unconditional(F, w, 1.0, 0.4)
    ref = O.fftsim_conditional(zu, zbar, cells, kw, 0.4)
    err = np.abs(out - ref).max()
    say(f"conditional: max |gpu - oracle| {err:.2e}")
    np.testing.assert_allclose(out, ref, rtol=0, atol=fft_tol(F))


def test_fftsim_conditional_missing_estimates(ctx):
    """A search ball that leaves cells without enough data: those cells are
    NaN (the reference's `missing`), the others match the oracle."""
    rng = np.random.default_rng(5)
    dims = (36, 28)
    coords = rng.uniform(0, 1, (12, 2)) * [36, 28]
    vals = rng.standard_normal(12)
    cells, zbar, kw = cond_setup(dims, O.SPHERICAL, (10.0, 10.0), coords, vals, 6, radius=5.0,
                                 kmin=2)
    p = fft_problem(dims, O.SPHERICAL, 1.0, 0.0, (10.0, 10.0))
    p.n_data = 12
    p.data_coords = coords.ctypes.data
    p.data_vals = vals.ctypes.data
    p.n_cells = len(cells)
    p.cells = cells.ctypes.data
    p.minneighbors, p.maxneighbors, p.radius = 2, 6, 5.0
    plan = _lib.FftsimPlan(ctx, p, keep=(coords, vals, cells))
    w = rng.random((2, 1008))
    out = plan.sample(w)
    F = O.fftsim_amplitude(dims, O.SPHERICAL, 1.0, 0.0, (10.0, 10.0))
    ref = O.fftsim_conditional(O.fftsim_unconditional(F, w, 1.0, 0.0), zbar, cells, kw, 0.0)
    miss = np.isnan(ref)
    assert miss.any() and (~miss).any()
    np.testing.assert_array_equal(np.isnan(out), miss)
    np.testing.assert_allclose(out[~miss], ref[~miss], rtol=0, atol=fft_tol(F))


def test_reference_fftsim_testset(ctx):
    """test/field.jl:115-141 (the view-of-grid and units cases have no
    counterpart here): basic and conditional simulation through rand."""
    rng = np.random.default_rng(2019)
    proc = gsp.GaussianProcess(gsp.GaussianVariogram(range=10.0))
    grid = gsp.CartesianGrid(100, 100)
    real = gsp.rand(proc, grid, rng=rng, method=gsp.FFTSIM(), ctx=ctx)
    assert real["field"].dtype == np.float64 and len(real["field"]) == 10000
    # the default method of an unconditional grid (field.jl:159) is FFTSIM
    ens = gsp.rand(proc, grid, 3, rng=np.random.default_rng(1), ctx=ctx)
    assert len(ens) == 3
    proc = gsp.GaussianProcess(gsp.GaussianVariogram(range=35.0))
    grid = gsp.CartesianGrid((0.5, 0.5), (100.5, 100.5), dims=(100, 100))
    data = gsp.georef({"Z": [1.0, 0.0, 1.0]}, [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, method=gsp.FFTSIM(maxneighbors=3), data=data, ctx=ctx)
    assert real.domain == grid and real.nrow() == 10000
    # data on cell centroids are reproduced
    for (i, j), v in {(25, 25): 1.0, (50, 75): 0.0, (75, 50): 1.0}.items():
        assert real["Z"][grid.linear_index(i - 1, j - 1)] == pytest.approx(v, abs=1e-9)


def test_fftsim_errors(ctx):
    with pytest.raises(_lib.ArgumentError):
        _lib.FftsimPlan(ctx, fft_problem((10, 1), O.GAUSSIAN, 1.0, 0.0, (3.0, 3.0)))
    p = fft_problem((10, 10), O.GAUSSIAN, 1.0, 0.0, (3.0, 3.0))
    p.n_data = 2
    with pytest.raises(_lib.ArgumentError):
        _lib.FftsimPlan(ctx, p)
