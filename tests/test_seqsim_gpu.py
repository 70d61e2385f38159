"""SEQSIM parity on the GPU: libgsp_hip.so through its C ABI against the CPU
oracle on the same inputs. Neighbour sets, counts and flags must be identical;
FP64 weights/sigmas bitwise; realizations within 1e-10 relative."""
import numpy as np
import pytest

import oracle as O
from helpers import config1, config3, gpu_plan, gpu_problem, rel_err, say
from geostatsprocesses import _lib

pytestmark = pytest.mark.gpu


def check_plan(ctx, case, brute=False, bitwise=True):
    say(f"plan parity dims={case.dims} k={case.kmax} r={case.radius} mode={case.mode} "
        f"prec={case.precision} nd={case.n_data}")
    gp = gpu_plan(ctx, case)
    g = gp.export()
    o = O.seqsim_plan(case, brute=brute)
    assert gp.kmax == o.kmax
    np.testing.assert_array_equal(g["nfound"], o.nfound)
    np.testing.assert_array_equal(g["nbr"], o.nbr)
    np.testing.assert_array_equal(g["flags"], o.flags)
    np.testing.assert_array_equal(g["level"], o.level)
    if bitwise:
        np.testing.assert_array_equal(g["lam"], o.lam)
        np.testing.assert_array_equal(g["sigma"], o.sigma)
    else:
        assert rel_err(g["lam"], o.lam) <= 1e-12
        assert rel_err(g["sigma"], o.sigma) <= 1e-12
    assert gp.info.depth == o.level.max() + 1
    assert gp.info.n_prior == int(np.count_nonzero(o.flags & 3))
    return gp, o


CASES = [
    dict(dims=(64,), family=O.SPHERICAL, ranges=(6.0,), kmax=5, radius=0.0, nd=4),
    dict(dims=(33, 27), family=O.SPHERICAL, ranges=(8.0,), kmax=12, radius=10.0, nd=15),
    dict(dims=(33, 27), family=O.GAUSSIAN, ranges=(8.0,), nugget=0.1, kmax=16, radius=0.0, nd=0),
    dict(dims=(17, 13, 9), family=O.EXPONENTIAL, ranges=(6.0, 5.0, 2.0), nugget=0.05, kmax=24,
         radius=6.0, nd=40),
    dict(dims=(20, 16), spacing=(1.0, 2.5), origin=(-3.0, 7.0), family=O.EXPONENTIAL,
         ranges=(12.0, 6.0), kmax=32, radius=14.0, nd=60),
    dict(dims=(24, 24), family=O.SPHERICAL, ranges=(10.0,), kmax=48, radius=12.0, nd=20),
    dict(dims=(20, 20), family=O.EXPONENTIAL, ranges=(9.0,), kmax=64, radius=0.0, nd=30),
    dict(dims=(30, 30), family=O.SPHERICAL, ranges=(7.0,), kmax=8, radius=7.0, nd=400),
]


def make_case(i, mode, precision=O.FP64):
    spec = dict(CASES[i])
    nd = spec.pop("nd")
    rng = np.random.default_rng(100 + i)
    dims = spec["dims"]
    N = int(np.prod(dims))
    sp = np.asarray(spec.get("spacing", (1.0,) * len(dims)))
    org = np.asarray(spec.get("origin", (0.0,) * len(dims)))
    data = org + rng.uniform(-0.05, 1.05, (nd, len(dims))) * np.asarray(dims) * sp
    if nd >= 3:
        data[1] = org + (np.array([2] * len(dims)) + 0.5) * sp      # on a node centre
        data[2] = data[1]                                             # duplicated datum
    if mode == O.MASKED_KNN and spec["radius"] == 0.0:
        spec["radius"] = 9.0
    return O.SeqsimCase(**spec, data_coords=data, data_vals=rng.standard_normal(nd),
                        path=rng.permutation(N), mode=mode, precision=precision)


@pytest.mark.parametrize("mode", [O.KNN_THEN_MASK, O.MASKED_KNN])
@pytest.mark.parametrize("i", range(len(CASES)))
def test_plan_parity_fp64(ctx, i, mode):
    case = make_case(i, mode)
    check_plan(ctx, case, brute=case.n_nodes <= 1000)


@pytest.mark.parametrize("mode", [O.KNN_THEN_MASK, O.MASKED_KNN])
def test_plan_parity_large_levels(ctx, mode):
    """34,560 nodes, more than the one-workgroup level pass takes: the levels
    come from the launches per level, and are built again after failed systems
    (a duplicated datum makes every system holding both copies singular)."""
    rng = np.random.default_rng(77)
    dims = (192, 180)
    data = rng.uniform(0, 1, (50, 2)) * np.asarray(dims, float)
    data[1] = [40.5, 60.5]
    data[2] = data[1]
    case = O.SeqsimCase(dims=dims, family=O.SPHERICAL, ranges=(8.0,), kmax=12, radius=10.0,
                        data_coords=data, data_vals=rng.standard_normal(50),
                        path=rng.permutation(192 * 180), mode=mode)
    gp, o = check_plan(ctx, case)
    assert np.count_nonzero(o.flags & O.FLAG_FAIL) > 0
    assert np.count_nonzero(o.flags & O.FLAG_FEW) > 0
    # the sweep visits every node once, after its neighbours
    z = np.random.default_rng(3).standard_normal((3, case.n_nodes))
    plan = O.Plan(case.kmax, o.nfound, o.nbr, o.lam, o.sigma, o.flags, o.level)
    np.testing.assert_array_equal(gp.sweep(z, case.data_vals), O.seqsim_sweep(case, plan, z))


@pytest.mark.parametrize("i", [1, 3, 4, 5])
def test_plan_parity_fp32_refine(ctx, i):
    case = make_case(i, O.KNN_THEN_MASK, precision=O.FP32_REFINE)
    gp, o = check_plan(ctx, case)
    # the refined FP32 weights stay close to the FP64 ones
    ref = O.seqsim_plan(make_case(i, O.KNN_THEN_MASK, precision=O.FP64))
    assert rel_err(gp.export()["lam"], ref.lam) <= 1e-6


@pytest.mark.parametrize("precision", [O.FP64, O.FP32_REFINE])
@pytest.mark.parametrize("i", [0, 1, 2, 3, 4, 7])
def test_plan_parity_lane_kernel(ctx, i, precision, monkeypatch):
    """Systems of at most 8 (FP32 + refinement) or 10 (FP64) neighbours solved
    one per lane (forced on these small plans, which would otherwise stay in
    the four-per-wavefront kernel): bitwise equal to the oracle in both
    precisions, including prior draws, failed systems (duplicated data) and
    FP64 fallbacks."""
    monkeypatch.setenv("GSP_LANE_MIN", "0")
    check_plan(ctx, make_case(i, O.KNN_THEN_MASK, precision=precision))


@pytest.mark.parametrize("mode", [O.KNN_THEN_MASK, O.MASKED_KNN])
@pytest.mark.parametrize("i", [0, 1, 3, 4, 7])
def test_sweep_parity(ctx, i, mode):
    case = make_case(i, mode)
    say(f"sweep parity case={i} mode={mode}")
    gp = gpu_plan(ctx, case)
    o = O.seqsim_plan(case)
    rng = np.random.default_rng(7 + i)
    for R in (1, 10, 64, 70):
        z = rng.standard_normal((R, case.n_nodes))
        got = gp.sweep(z, case.data_vals)
        ref = O.seqsim_sweep(case, o, z)
        assert rel_err(got, ref) <= 1e-10, R
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("chunk", [1, 3, 16])
def test_sweep_realization_chunks(ctx, chunk, monkeypatch):
    """Realizations swept in passes of `chunk` give the same bits."""
    monkeypatch.setenv("GSP_SWEEP_CHUNK", str(chunk))
    case = make_case(3, O.KNN_THEN_MASK)
    gp = gpu_plan(ctx, case)
    o = O.seqsim_plan(case)
    z = np.random.default_rng(31).standard_normal((20, case.n_nodes))
    np.testing.assert_array_equal(gp.sweep(z, case.data_vals), O.seqsim_sweep(case, o, z))


def test_config1_parity(ctx):
    """100x100, spherical range 10, k=12, radius 30, 50 data, R=10 — conditional
    and unconditional; plan bitwise and realizations bitwise against the oracle."""
    for cond in (False, True):
        case, z = config1(conditional=cond)
        say(f"config1 conditional={cond}")
        gp, o = check_plan(ctx, case)
        got = gp.sweep(z, case.data_vals)
        ref = O.seqsim_sweep(case, o, z)
        np.testing.assert_array_equal(got, ref)
        ref2, _ = O.seqsim_reference_cost(case, z, threads=4)
        np.testing.assert_array_equal(got, ref2)


def test_config3_reduced_parity(ctx):
    """The 3D anisotropic configuration at 48x48x24, FP32 + refinement and FP64."""
    for prec in (O.FP32_REFINE, O.FP64):
        case = config3(dims=(48, 48, 24), n_data=110, precision=prec)
        check_plan(ctx, case)


def test_config3_full_size_properties(ctx):
    """128x128x64 with 2000 data: neighbour lists of 3000 sampled nodes equal the
    oracle's brute-force search; weights of those nodes equal the oracle's
    solve; data honoured; level depth consistent; realization statistics."""
    case = config3()
    say("config3 full-size plan")
    gp = gpu_plan(ctx, case)
    g = gp.export()
    pos = O.positions(case)
    rng = np.random.default_rng(5)
    b = O.lib().or_bucket_create(__import__("ctypes").byref(case.struct()))
    try:
        for t in rng.choice(case.n_nodes, 3000, replace=False):
            nb = O.search(case, t, brute=False, pos=pos, bucket=b)
            n = int(g["nfound"][t])
            assert n == len(nb) and np.array_equal(g["nbr"][t, :n], nb), t
            if n >= 1 and not g["flags"][t] & 3:
                lam, sig, f = O.krige(case, t, nb)
                assert f == g["flags"][t]
                np.testing.assert_array_equal(g["lam"][t, :n], lam)
                assert g["sigma"][t] == sig
    finally:
        O.lib().or_bucket_free(b)
    for t in rng.choice(case.n_nodes, 200, replace=False):
        nb = O.search(case, t, brute=True, pos=pos)
        assert np.array_equal(g["nbr"][t, :len(nb)], nb)
    # dependency depth: level(i) = 1 + max(level(grid neighbours))
    k = np.where(g["flags"] & 3, 0, g["nfound"])
    nbr = np.where(np.arange(case.kmax)[None, :] < k[:, None], g["nbr"], -1)
    dep = np.where((nbr >= 0) & (nbr < case.n_nodes),
                   g["level"][np.clip(nbr, 0, case.n_nodes - 1)], -1)
    np.testing.assert_array_equal(g["level"], dep.max(axis=1) + 1)
    # realizations: the GPU sweep equals the oracle's sweep over the same plan
    z = np.random.default_rng(9).standard_normal((8, case.n_nodes))
    out = gp.sweep(z, case.data_vals)
    plan = O.Plan(case.kmax, g["nfound"], g["nbr"], g["lam"], g["sigma"], g["flags"], g["level"])
    np.testing.assert_array_equal(out, O.seqsim_sweep(case, plan, z))
    # k nearest first, then mask: about 3/4 of the sill is reproduced pointwise
    assert abs(out.mean()) < 0.2 and 0.4 < out.var() < 1.2


def test_device_pointer_sweep_matches_host(ctx):
    case, z = config1()
    gp = gpu_plan(ctx, case)
    host = gp.sweep(z, case.data_vals)
    dz = _lib.DeviceBuffer.from_array(ctx, z)
    dv = _lib.DeviceBuffer.from_array(ctx, case.data_vals)
    dout = _lib.DeviceBuffer(ctx, (z.shape[0], case.n_nodes))
    gp.sweep(dz, dv, dout, device_pointers=True)
    ctx.synchronize()
    np.testing.assert_array_equal(dout.download(), host)


def test_device_pointer_plan_matches_host(ctx):
    case, _ = config1()
    dpath = _lib.DeviceBuffer.from_array(ctx, case.path)
    ddata = _lib.DeviceBuffer.from_array(ctx, case.data_coords)
    p = gpu_problem(case, device_pointers=True, data_ptr=ddata.ptr, path_ptr=dpath.ptr)
    gp = _lib.SeqsimPlan(ctx, p)
    ref = gpu_plan(ctx, case).export()
    got = gp.export()
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k])


def test_repeated_sweeps_are_identical(ctx):
    case, z = config1(R=3)
    gp = gpu_plan(ctx, case)
    a = gp.sweep(z, case.data_vals)
    for _ in range(3):
        np.testing.assert_array_equal(gp.sweep(z, case.data_vals), a)


def test_errors(ctx):
    case, _ = config1()
    bad = O.SeqsimCase(dims=(10, 10), kmax=5, radius=3.0, path=np.zeros(100, np.int64))
    with pytest.raises(_lib.UnsupportedError):
        gpu_plan(ctx, bad)
    bad = O.SeqsimCase(dims=(10, 10), kmax=5, radius=3.0, path=np.arange(99))
    with pytest.raises(_lib.UnsupportedError):
        gpu_plan(ctx, bad)
    big = O.SeqsimCase(dims=(20, 20), kmax=65, radius=0.0)
    with pytest.raises(_lib.UnsupportedError):
        gpu_plan(ctx, big)
    neg = O.SeqsimCase(dims=(20, 20), kmax=5, ranges=(-1.0,))
    with pytest.raises(_lib.ArgumentError):
        gpu_plan(ctx, neg)
    gp = gpu_plan(ctx, case)
    with pytest.raises(_lib.ArgumentError):
        gp.sweep(np.zeros((1, 5)), case.data_vals)


@pytest.mark.parametrize("mode,kmin", [(O.KNN_THEN_MASK, 1), (O.MASKED_KNN, 1),
                                       (O.KNN_THEN_MASK, 6)])
def test_grid_lane_search(ctx, mode, kmin, monkeypatch):
    """A 3D grid without data (48 x 40 x 24 = 46,080 nodes, kmax 24, anisotropic
    ranges and spacing): the one-node-per-lane grid search gives the oracle's
    neighbour lists, counts, flags, levels and weights bit for bit, and the
    wavefront search's (GSP_GRID_LANE=0) the same lists and levels. With
    kmin 6, nodes with 1..5 simulated neighbours draw from the prior and have
    no dependencies: the dependency counts their walk added are taken back."""
    rng = np.random.default_rng(91)
    dims = (48, 40, 24)
    N = int(np.prod(dims))
    case = O.SeqsimCase(dims=dims, spacing=(1.0, 1.5, 0.5), family=O.EXPONENTIAL,
                        ranges=(10.0, 7.0, 3.0), nugget=0.02, kmax=24, kmin=kmin,
                        radius=9.0 if mode == O.MASKED_KNN else 0.0,
                        data_coords=np.zeros((0, 3)), data_vals=np.zeros(0),
                        path=rng.permutation(N), mode=mode)
    gp, o = check_plan(ctx, case)
    if kmin > 1:
        few = (o.nfound > 0) & (o.nfound < kmin)
        assert np.count_nonzero(few) > 0
    monkeypatch.setenv("GSP_GRID_LANE", "0")
    w = gpu_plan(ctx, case).export()
    g = gp.export()
    for key in ("nfound", "nbr", "flags", "level", "lam", "sigma"):
        np.testing.assert_array_equal(g[key], w[key], err_msg=key)
