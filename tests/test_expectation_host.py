"""Expectation API host logic (expectation/field.jl:16-36,
expectation/field/gaussian.jl:5-19): the prior mean and the argument checks
that run before any device work."""
import numpy as np
import pytest

import geostatsprocesses as gsp
from geostatsprocesses import _lib


def test_prior_mean_reference_case():
    """test/field.jl:157-161: the unconditional expectation is zero."""
    proc = gsp.GaussianProcess(gsp.SphericalVariogram(range=35.0))
    grid = gsp.CartesianGrid((0.5, 0.5), (100.5, 100.5), dims=(100, 100))
    mval = gsp.mean(proc, grid)
    assert mval.names() == ("field",) and len(mval["field"]) == 10000
    assert np.all(mval["field"] == 0)
    np.testing.assert_array_equal(gsp.expectedvalue(proc, grid)["field"], mval["field"])


def test_prior_mean_fills_process_mean():
    proc = gsp.GaussianProcess(gsp.GaussianVariogram(range=3.0), 2.5)
    grid = gsp.CartesianGrid((4, 3, 2))
    np.testing.assert_array_equal(gsp.priormean(proc, grid)["field"], np.full(24, 2.5))


def test_posterior_mean_argument_errors():
    proc = gsp.GaussianProcess(gsp.SphericalVariogram(range=5.0))
    grid = gsp.CartesianGrid((10, 10))
    data = gsp.georef({"Z": [1.0, 2.0]}, [(1.0, 1.0), (3.0, 4.0)])
    with pytest.raises(_lib.UnsupportedError):
        gsp.posteriormean(proc, grid, data, point=False)
    with pytest.raises(_lib.UnsupportedError):
        gsp.posteriormean(proc, gsp.PointSet([(1.0, 1.0)]), data)
    with pytest.raises(_lib.UnsupportedError):
        gsp.posteriormean(proc, grid, data, distance="cityblock")
    with pytest.raises(NotImplementedError):
        gsp.expectedvalue(object(), grid)
