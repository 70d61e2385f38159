"""Multivariate Gaussian SEQSIM of an intrinsic coregionalization A * c(h)
(seqsim.jl:72-156 with the MvNormal prior of seqsim.jl:287) on the GPU: the
realizations of every variable against the oracle, bit for bit, and the
reference's multivariate SEQSIM testset (test/field.jl:102-112) through rand."""
import numpy as np
import pytest

import oracle as O
import geostatsprocesses as gsp
from helpers import gpu_plan, say
from geostatsprocesses import _lib

pytestmark = pytest.mark.gpu

CASES = [
    dict(dims=(30, 25), family=O.SPHERICAL, ranges=(7.0,), kmax=12, radius=9.0, nd=20, nv=2,
         A=[[1.0, 0.8], [0.8, 1.0]], means=[0.1, 0.2], precision=O.FP64),
    dict(dims=(16, 14, 6), family=O.EXPONENTIAL, ranges=(6.0, 5.0, 3.0), kmax=24, radius=6.0,
         nd=30, nv=3, A=[[2.0, 0.5, -0.3], [0.5, 1.0, 0.2], [-0.3, 0.2, 0.7]],
         means=[1.0, -2.0, 0.5], precision=O.FP32_REFINE),
    dict(dims=(40, 31), family=O.GAUSSIAN, ranges=(8.0,), kmax=16, radius=0.0, nd=0, nv=2,
         A=[[1.0, -0.6], [-0.6, 2.5]], means=[0.0, 3.0], precision=O.FP64),
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_multivariate_parity(ctx, i):
    spec = dict(CASES[i])
    nd, nv, A, means = spec.pop("nd"), spec.pop("nv"), spec.pop("A"), spec.pop("means")
    rng = np.random.default_rng(300 + i)
    dims = spec["dims"]
    N = int(np.prod(dims))
    data = rng.uniform(0, 1, (nd, len(dims))) * np.asarray(dims)
    vals = rng.standard_normal((nv, nd))
    case = O.SeqsimCase(**spec, data_coords=data, data_vals=vals[0] if nd else None,
                        path=rng.permutation(N))
    say(f"multivariate dims={dims} nv={nv} nd={nd} prec={spec['precision']}")
    plan = gpu_plan(ctx, case)
    z = rng.standard_normal((4, N, nv))
    got = plan.sweep_multivariate(z, A, means, vals if nd else None)
    ref = O.seqsim_sweep_multivariate(case, O.seqsim_plan(case), z, A, means, vals)
    np.testing.assert_array_equal(got, ref)


def test_multivariate_errors(ctx):
    case = O.SeqsimCase(dims=(10, 10), family=O.SPHERICAL, ranges=(5.0,), kmax=6, radius=5.0,
                        path=np.random.default_rng(1).permutation(100))
    plan = gpu_plan(ctx, case)
    z = np.zeros((1, 100, 2))
    with pytest.raises(_lib.PosDefException):
        plan.sweep_multivariate(z, [[1.0, 2.0], [2.0, 1.0]], [0.0, 0.0])
    with pytest.raises(_lib.ArgumentError):
        plan.sweep_multivariate(z, [[1.0, 0.5], [0.4, 1.0]], [0.0, 0.0])
    with pytest.raises(_lib.UnsupportedError):
        plan.sweep_multivariate(np.zeros((1, 100, 5)), np.eye(5), np.zeros(5))


def test_reference_multivariate_seqsim(ctx):
    """test/field.jl:102-112 without units: [1 .8; .8 1] * SphericalCovariance(range
    35) with means (0.1, 0.2) on a 10 x 10 grid, with and without data (Cu, Zn)."""
    rng = np.random.default_rng(2017)
    method = gsp.SEQSIM(maxneighbors=3)
    func = np.array([[1.0, 0.8], [0.8, 1.0]]) * gsp.SphericalCovariance(range=35.0, sill=1.0)
    proc = gsp.GaussianProcess(func, [0.1, 0.2])
    grid = gsp.CartesianGrid(10, 10)
    real = gsp.rand(proc, grid, rng=rng, method=method)
    assert real.field1.dtype == np.float64 and real.field2.dtype == np.float64
    assert np.isfinite(real.field1).all() and np.isfinite(real.field2).all()
    data = gsp.georef({"Cu": [0.0, 0.1, 0.0], "Zn": [0.1, 0.0, 0.1]},
                      [(25.0, 25.0), (50.0, 75.0), (75.0, 50.0)])
    real = gsp.rand(proc, grid, rng=rng, method=method, data=data)
    assert real.Cu.dtype == np.float64 and real.Zn.dtype == np.float64
    ens = gsp.rand(proc, grid, 3, rng=rng, method=method, data=data)
    assert len(ens) == 3 and ens[1].Zn.dtype == np.float64
