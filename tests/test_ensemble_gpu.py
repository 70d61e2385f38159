"""Ensemble statistics on the GPU (ensembles.jl:42-52,76-85) against the
oracle's restatement of Julia's Statistics, bit for bit: the same operation
order per cell (left-to-right sums, Welford's recurrence, count / R, the
type-7 quantile interpolation with an exact fma)."""
import numpy as np
import pytest

import oracle as O
from helpers import say
from geostatsprocesses import _lib
import geostatsprocesses as gsp

pytestmark = pytest.mark.gpu

KIND = {"mean": _lib.STAT_MEAN, "var": _lib.STAT_VAR, "cdf": _lib.STAT_CDF,
        "ccdf": _lib.STAT_CCDF, "quantile": _lib.STAT_QUANTILE}
REQUESTS = [("mean", 0.0), ("var", 0.0), ("cdf", 0.3), ("ccdf", -0.5), ("quantile", 0.0),
            ("quantile", 0.1), ("quantile", 0.5), ("quantile", 0.9), ("quantile", 1.0),
            ("cdf", 1.0)]


@pytest.mark.parametrize("R,N", [(1, 100), (2, 257), (7, 1000), (16, 4099), (64, 3000),
                                 (128, 500), (129, 300), (200, 130)])
def test_ensemble_stats_match_oracle(ctx, R, N):
    v = np.random.default_rng(R * 31 + N).standard_normal((R, N))
    v[:, :5] = np.round(v[:, :5])  # ties
    say("ensemble stats", R, N)
    out = _lib.ensemble_stats(ctx, v, [(KIND[k], p) for k, p in REQUESTS])
    for q, (k, p) in enumerate(REQUESTS):
        ref = O.ensemble_stat(v, k, p)
        np.testing.assert_array_equal(out[q], ref, err_msg=f"{k}({p})")


def test_ensemble_reference_testset(ctx):
    """test/ensembles.jl:24-59 through the Ensemble API."""
    grid = gsp.CartesianGrid(3, 3)
    reals = np.stack([i * np.ones(9) for i in (1.0, 2.0, 3.0)])
    e = gsp.Ensemble(grid, {"z": reals}, ctx=ctx)
    assert np.all(e.mean().z == 2.0)
    assert np.all(e.var().z == 1.0)
    for i in (1, 2, 3):
        assert np.all(e.cdf(i).z == i / 3)
        np.testing.assert_allclose(e.ccdf(i).z, 1 - e.cdf(i).z)
    q = e.quantile(0.5)
    assert np.all(q.z == 2.0)
    qs = e.quantile([0.0, 0.5, 1.0])
    assert np.array_equal(qs[1].z, q.z)
    assert np.all(qs[0].z == 1.0) and np.all(qs[2].z == 3.0)


def test_ensemble_nan_quantile_and_device_buffers(ctx):
    """A NaN among a cell's values makes its quantiles NaN (the reference
    throws); device-resident realizations are reduced in place."""
    v = np.random.default_rng(3).standard_normal((9, 64))
    v[4, 10] = np.nan
    out = _lib.ensemble_stats(ctx, v, [(_lib.STAT_QUANTILE, 0.5), (_lib.STAT_MEAN, 0.0)])
    assert np.isnan(out[0, 10]) and np.isnan(out[1, 10])
    assert not np.isnan(np.delete(out[0], 10)).any()
    dv = _lib.DeviceBuffer.from_array(ctx, v)
    dout = _lib.DeviceBuffer(ctx, (2, 64))
    _lib.ensemble_stats(ctx, dv, [(_lib.STAT_QUANTILE, 0.5), (_lib.STAT_MEAN, 0.0)], out=dout,
                        device_pointers=True, shape=(9, 64))
    ctx.synchronize()
    np.testing.assert_array_equal(dout.download(), out)


def test_ensemble_stats_errors(ctx):
    v = np.zeros((3, 4))
    with pytest.raises(_lib.ArgumentError):
        _lib.ensemble_stats(ctx, v, [(_lib.STAT_QUANTILE, 1.5)])
    with pytest.raises(_lib.ArgumentError):
        _lib.ensemble_stats(ctx, v, [(9, 0.0)])
    with pytest.raises(_lib.ArgumentError):
        _lib.ensemble_stats(ctx, v, [(_lib.STAT_MEAN, 0.0)] * 33)
