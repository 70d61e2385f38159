"""Host side of bivariate LUSIM (lusim.jl:132-143): the marginal covariance of
each variable and the correlation of the normals, without a GPU."""
import numpy as np
import pytest

import geostatsprocesses as gsp
from geostatsprocesses.simulation import _lusim_problem, _marginal_struct, _rho


def test_rho_closed_form():
    # rho = C12(0) / sqrt(s1 s2) = A12 / sqrt(A11 A22), whatever the base sill
    for A, sill in [([[1.0, 0.95], [0.95, 1.0]], 1.0), ([[2.0, 0.5], [0.5, 1.0]], 3.0),
                    ([[4.0, -1.0], [-1.0, 1.0]], 0.5)]:
        A = np.asarray(A)
        f = A * gsp.SphericalCovariance(range=10.0, sill=sill, nugget=0.1 * sill)
        assert _rho(f) == pytest.approx(A[0, 1] / np.sqrt(A[0, 0] * A[1, 1]), rel=1e-15)


def test_marginal_covariance_scales_sill_and_nugget():
    A = np.array([[2.0, 0.5], [0.5, 3.0]])
    base = gsp.ExponentialCovariance(range=7.0, sill=1.5, nugget=0.2)
    for j in range(2):
        m = _marginal_struct(A * base, j, 2)
        b = base.struct(2)
        assert m.family == b.family and m.kind == b.kind
        assert m.sill == pytest.approx(A[j, j] * 1.5) and m.nugget == pytest.approx(A[j, j] * 0.2)
        assert list(m.ranges)[:2] == list(b.ranges)[:2]


def test_lusim_variable_count_checks():
    grid = gsp.CartesianGrid(5, 4)
    A3 = np.eye(3)
    proc = gsp.GaussianProcess(A3 * gsp.SphericalCovariance(range=3.0), np.zeros(3))
    with pytest.raises(AssertionError, match="univariate and bivariate"):
        _lusim_problem(proc, gsp.LUSIM(), gsp.NearestInit(), grid, None, None)
    proc = gsp.GaussianProcess(np.eye(2) * gsp.SphericalCovariance(range=3.0), [0.0, 0.0])
    data = gsp.georef({"Z": [1.0]}, [(1.0, 1.0)])
    with pytest.raises(AssertionError, match="incompatible number of variables"):
        _lusim_problem(proc, gsp.LUSIM(), gsp.NearestInit(), grid, data, None)
