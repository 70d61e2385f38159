"""Sweeps replayed from a captured HIP graph: a plan's repeated sweeps with the
same device buffers are captured once and replayed; every replay must equal
the eager launches (made with the timing events on, which keep launches
individual) bit for bit, with other work and new input values between the
replays, and new buffers must be captured again."""
import numpy as np
import pytest

from helpers import config1, gpu_plan, say
from geostatsprocesses import _lib

pytestmark = pytest.mark.gpu


def eager(ctx, plan, dz, dv, dout):
    ctx.enable_timing(True)
    plan.sweep(dz, dv, dout, device_pointers=True)
    ctx.synchronize()
    ctx.enable_timing(False)
    return dout.download()


def test_graph_replays_equal_eager(ctx):
    case, z = config1(conditional=True, R=10)
    plan = gpu_plan(ctx, case)
    assert plan.info.depth >= 8
    other, z_other = config1(conditional=False, seed=2001, R=3)
    other_plan = gpu_plan(ctx, other)
    dz = _lib.DeviceBuffer.from_array(ctx, z)
    dv = _lib.DeviceBuffer.from_array(ctx, case.data_vals)
    dout = _lib.DeviceBuffer(ctx, z.shape)
    rng = np.random.default_rng(4)
    for rep in range(4):
        zr = rng.standard_normal(z.shape)
        dz.upload(zr)
        ref = eager(ctx, plan, dz, dv, dout)
        say(f"graph replay {rep}")
        plan.sweep(dz, dv, dout, device_pointers=True)    # captured on the first pass
        ctx.synchronize()
        np.testing.assert_array_equal(dout.download(), ref)
        # other work on the stream between replays
        other_plan.sweep(z_other, None)
    # new output buffer: captured again
    dout2 = _lib.DeviceBuffer(ctx, z.shape)
    plan.sweep(dz, dv, dout2, device_pointers=True)
    ctx.synchronize()
    np.testing.assert_array_equal(dout2.download(), eager(ctx, plan, dz, dv, dout))
    # host-pointer sweeps (staged buffers): captured, then replayed
    ctx.enable_timing(True)
    ref = plan.sweep(z, case.data_vals)
    ctx.enable_timing(False)
    for _ in range(3):
        np.testing.assert_array_equal(plan.sweep(z, case.data_vals), ref)
