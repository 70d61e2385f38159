"""gsp_seqsim_sweep_noalloc: the SEQSIM sweep with its scratch in the library's
fixed device array, for callers that cannot allocate. Its realizations must
equal gsp_seqsim_sweep's (device pointers) bit for bit, call after call, with
the array keeping the previous call's contents."""
import ctypes

import numpy as np
import pytest

import oracle as O
from helpers import config1, gpu_plan, gpu_problem, say
from geostatsprocesses import _lib

pytestmark = pytest.mark.gpu


def sweep_both(ctx, plan, z, vals):
    """(gsp_seqsim_sweep, gsp_seqsim_sweep_noalloc) outputs for the same inputs."""
    R, N = z.shape
    dz = _lib.DeviceBuffer.from_array(ctx, z)
    dv = _lib.DeviceBuffer.from_array(ctx, vals) if len(vals) else None
    ref = _lib.DeviceBuffer(ctx, (R, N))
    got = _lib.DeviceBuffer(ctx, (R, N))
    plan.sweep(dz, dv, ref, device_pointers=True)
    _lib.check(_lib.load().gsp_seqsim_sweep_noalloc(ctx.handle, plan._h, R, dz.ptr,
                                                    dv.ptr if dv is not None else None,
                                                    got.ptr))
    ctx.synchronize()
    return ref.download(), got.download()


def test_noalloc_sweep_matches_sweep_twice(ctx):
    # first call: the 100 x 100 grid with 50 data, R = 10
    case, z = config1(conditional=True, R=10)
    say("noalloc: 100 x 100 grid, 50 data, R=10")
    plan = gpu_plan(ctx, case)
    ref, got = sweep_both(ctx, plan, z, case.data_vals)
    np.testing.assert_array_equal(got, ref)
    # second call, smaller: 40 x 30 grid without data, R = 3
    rng = np.random.default_rng(7)
    small = O.SeqsimCase(dims=(40, 30), family=O.EXPONENTIAL, ranges=(8.0,), nugget=0.02,
                         kmax=16, radius=12.0, path=rng.permutation(1200))
    say("noalloc: 40x30 unconditional R=3")
    plan2 = gpu_plan(ctx, small)
    z2 = rng.standard_normal((3, 1200))
    ref2, got2 = sweep_both(ctx, plan2, z2, small.data_vals)
    np.testing.assert_array_equal(got2, ref2)
    # and the oracle agrees with both
    np.testing.assert_array_equal(got2, O.seqsim_sweep(small, O.seqsim_plan(small), z2))


def test_noalloc_sweep_in_realization_chunks(ctx):
    """A plan of 1,048,576 nodes leaves room for 3 realizations per pass of the
    8,388,608-word array (2 N + n_data + 32 words each): R = 7 runs as 3 + 3 + 1."""
    rng = np.random.default_rng(11)
    dims = (1024, 1024)
    N = dims[0] * dims[1]
    data = rng.uniform(0, 1024, (100, 2))
    case = O.SeqsimCase(dims=dims, family=O.SPHERICAL, ranges=(12.0,), kmax=8, radius=12.0,
                        data_coords=data, data_vals=rng.standard_normal(100),
                        path=rng.permutation(N))
    say("noalloc: 1024x1024, R=7 in chunks")
    plan = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,))
    z = rng.standard_normal((7, N))
    ref, got = sweep_both(ctx, plan, z, case.data_vals)
    np.testing.assert_array_equal(got, ref)


def test_noalloc_sweep_rejects_oversize_plan(ctx):
    rng = np.random.default_rng(12)
    dims = (2048, 2049)   # 2 N exceeds the array
    N = dims[0] * dims[1]
    case = O.SeqsimCase(dims=dims, family=O.SPHERICAL, ranges=(6.0,), kmax=4, radius=6.0,
                        path=rng.permutation(N))
    say("noalloc: 2048x2049 rejected")
    plan = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,))
    dz = _lib.DeviceBuffer(ctx, (1, N))
    out = _lib.DeviceBuffer(ctx, (1, N))
    with pytest.raises(_lib.ArgumentError, match="too large"):
        _lib.check(_lib.load().gsp_seqsim_sweep_noalloc(ctx.handle, plan._h, 1, dz.ptr, None,
                                                        out.ptr))
