"""Batched SEQSIM on the GPU: many independent requests (grids, variograms,
data, paths and neighbourhoods of their own) in one plan and one sweep. Every
request of a batch must give exactly what the oracle gives for it alone."""
import numpy as np
import pytest

import oracle as O
from helpers import batch_arrays, config5, gpu_problem, local_plan, say
from test_seqsim_gpu import make_case
from geostatsprocesses import _lib

pytestmark = pytest.mark.gpu


def check_batch(ctx, cases, R=6, seed=0, problems=None):
    say(f"batch of {len(cases)}: sizes {[c.n_nodes for c in cases]}")
    probs = problems if problems is not None else [gpu_problem(c) for c in cases]
    plan = _lib.SeqsimPlan(ctx, probs, keep=tuple(cases))
    assert plan.n_requests == len(cases)
    assert list(np.diff(plan.node_offset)) == [c.n_nodes for c in cases]
    assert list(np.diff(plan.data_offset)) == [c.n_data for c in cases]
    g = plan.export()
    rng = np.random.default_rng(seed)
    zs = [rng.standard_normal((R, c.n_nodes)) for c in cases]
    vals, z = batch_arrays(cases, plan.node_offset, zs)
    out = plan.sweep(z, vals)
    depth, prior = 0, 0
    for q, c in enumerate(cases):
        o = O.seqsim_plan(c)
        loc = local_plan(g, q, plan, cases)
        np.testing.assert_array_equal(loc["nfound"], o.nfound, err_msg=f"request {q}")
        np.testing.assert_array_equal(loc["nbr"], o.nbr, err_msg=f"request {q}")
        assert (loc["nbr_tail"] == -1).all() and (loc["lam_tail"] == 0).all()
        np.testing.assert_array_equal(loc["flags"], o.flags, err_msg=f"request {q}")
        np.testing.assert_array_equal(loc["level"], o.level, err_msg=f"request {q}")
        np.testing.assert_array_equal(loc["lam"], o.lam, err_msg=f"request {q}")
        np.testing.assert_array_equal(loc["sigma"], o.sigma, err_msg=f"request {q}")
        ref = O.seqsim_sweep(c, o, zs[q])
        got = out[:, plan.node_offset[q]:plan.node_offset[q + 1]]
        np.testing.assert_array_equal(got, ref, err_msg=f"request {q}")
        depth = max(depth, int(o.level.max()) + 1)
        prior += int(np.count_nonzero(o.flags & 3))
    assert plan.info.depth == depth
    assert plan.info.n_prior == prior
    assert plan.kmax == max(min(c.kmax, c.n_nodes) for c in cases)
    return plan


@pytest.mark.parametrize("precision", [O.FP64, O.FP32_REFINE])
def test_heterogeneous_batch(ctx, precision):
    """1D/2D/3D grids, all families, both search modes, with and without data."""
    cases = [make_case(i, mode, precision=precision)
             for i, mode in [(0, O.KNN_THEN_MASK), (1, O.MASKED_KNN), (2, O.KNN_THEN_MASK),
                             (3, O.KNN_THEN_MASK), (4, O.MASKED_KNN), (5, O.KNN_THEN_MASK),
                             (6, O.KNN_THEN_MASK), (7, O.MASKED_KNN)]]
    check_batch(ctx, cases, R=5, seed=3)


def test_batch_of_one_equals_single_plan(ctx):
    case = make_case(3, O.KNN_THEN_MASK)
    single = _lib.SeqsimPlan(ctx, gpu_problem(case), keep=(case,))
    batch = check_batch(ctx, [case], R=3)
    a, b = single.export(), batch.export()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])


def test_repeated_request_in_batch(ctx):
    """The same request twice gives the same plan rows and, with the same
    normals, the same realizations."""
    case = make_case(1, O.KNN_THEN_MASK)
    plan = check_batch(ctx, [case, case, make_case(0, O.KNN_THEN_MASK)], R=4, seed=5)
    g = plan.export()
    n = case.n_nodes
    np.testing.assert_array_equal(g["lam"][:n], g["lam"][n:2 * n])
    np.testing.assert_array_equal(g["sigma"][:n], g["sigma"][n:2 * n])


def test_config5_small_batch(ctx):
    """The mixed-size workload at reduced sides ([32, 80]): 24 requests."""
    cases = config5(n_req=24, side_lo=32, side_hi=80, seed=77)
    check_batch(ctx, cases, R=8, seed=11)


def test_batch_device_pointers(ctx):
    cases = [make_case(i, O.KNN_THEN_MASK) for i in (1, 3, 7)]
    bufs, probs = [], []
    for c in cases:
        dp = _lib.DeviceBuffer.from_array(ctx, c.path)
        dd = _lib.DeviceBuffer.from_array(ctx, c.data_coords) if c.n_data else None
        bufs += [dp, dd]
        probs.append(gpu_problem(c, device_pointers=True, data_ptr=dd.ptr if dd else None,
                                 path_ptr=dp.ptr))
    plan = check_batch(ctx, cases, R=2, problems=probs)
    # the sweep with device buffers gives the host-pointer result
    rng = np.random.default_rng(9)
    zs = [rng.standard_normal((3, c.n_nodes)) for c in cases]
    vals, z = batch_arrays(cases, plan.node_offset, zs)
    ref = plan.sweep(z, vals)
    dz, dv = _lib.DeviceBuffer.from_array(ctx, z), _lib.DeviceBuffer.from_array(ctx, vals)
    dout = _lib.DeviceBuffer(ctx, z.shape)
    plan.sweep(dz, dv, dout, device_pointers=True)
    np.testing.assert_array_equal(dout.download(), ref)


def test_batch_errors(ctx):
    a, b = make_case(1, O.KNN_THEN_MASK), make_case(0, O.KNN_THEN_MASK, precision=O.FP32_REFINE)
    with pytest.raises(_lib.ArgumentError, match="one precision"):
        _lib.SeqsimPlan(ctx, [gpu_problem(a), gpu_problem(b)], keep=(a, b))
    bad = make_case(0, O.KNN_THEN_MASK)
    bad.path = bad.path.copy()
    bad.path[3] = bad.path[4]
    with pytest.raises(_lib.UnsupportedError, match="permutation"):
        _lib.SeqsimPlan(ctx, [gpu_problem(a), gpu_problem(bad)], keep=(a, bad))
    bad2 = make_case(0, O.KNN_THEN_MASK)
    p = gpu_problem(bad2)
    p.model.sill = -1.0
    with pytest.raises(_lib.ArgumentError, match="request 1"):
        _lib.SeqsimPlan(ctx, [gpu_problem(a), p], keep=(a, bad2))
    with pytest.raises(_lib.ArgumentError):
        _lib.SeqsimPlan(ctx, [], keep=())


def test_config5_full_batch(ctx):
    """The full mixed-size workload: 512 requests, sides log-uniform in [32, 512]
    (about 24M nodes), spherical/exponential/gaussian variograms, 0-500 data,
    maxneighbors in {8, 16, 32}, R = 8, FP32 solves with refinement, in one plan
    and one sweep. Checked at full size: every node's dependency level is 1 + the
    largest level of its grid neighbours (0 for prior draws and nodes without
    grid neighbours: the longest-chain definition) and the plan depth is the
    largest level + 1; every realization value of two realizations satisfies the
    sweep recurrence x_i = (mean + sum_j lambda_ij (x_j - mean)) + sigma_i z_i
    bit for bit, with data neighbours reading the data values. Ten requests of
    every family and maxneighbors (the largest grid among them) are checked
    against the oracle bit for bit: neighbours, flags, levels, weights, sigmas
    and all 8 realizations."""
    R = 8
    cases = config5(n_req=512)
    sides = np.array([c.dims[0] for c in cases])
    say(f"config5: {len(cases)} requests, {int((sides ** 2).sum())} nodes, "
        f"{sum(c.n_data for c in cases)} data")
    probs = [gpu_problem(c) for c in cases]
    plan = _lib.SeqsimPlan(ctx, probs, keep=tuple(cases))
    no, do = plan.node_offset, plan.data_offset
    NT = int(no[-1])
    assert plan.n_requests == 512 and NT == int((sides ** 2).sum())
    # normals and data values in the batch layout, on the device
    rng = np.random.default_rng(12)
    z = rng.standard_normal((R, NT))
    vals = np.concatenate([c.data_vals for c in cases])
    dz, dv = _lib.DeviceBuffer.from_array(ctx, z), _lib.DeviceBuffer.from_array(ctx, vals)
    dout = _lib.DeviceBuffer(ctx, (R, NT))
    plan.sweep(dz, dv, dout, device_pointers=True)
    ctx.synchronize()
    out = dout.download()
    say("config5: plan and sweep done")
    g = plan.export(fields=("nfound", "nbr", "flags", "level"))
    nbr, flags, level = g["nbr"], g["flags"], g["level"]
    prior = (flags & 3) != 0
    # levels: the longest-chain recursion at every node
    grid_nb = (nbr >= 0) & (nbr < NT) & ~prior[:, None]
    nb_level = np.where(grid_nb, level[np.where(grid_nb, nbr, 0)], -1)
    expect = nb_level.max(axis=1) + 1
    np.testing.assert_array_equal(level, expect)
    assert plan.info.depth == int(level.max()) + 1
    del nb_level, expect, grid_nb
    say(f"config5: levels verified (depth {plan.info.depth})")
    # sweep recurrence on two realizations, every node
    ls = plan.export(fields=("lam", "sigma"))
    lam, sigma = ls["lam"], ls["sigma"]
    pos = np.empty(NT, np.int64)
    for q, c in enumerate(cases):
        pos[no[q] + c.path] = no[q] + np.arange(c.n_nodes)
    mean = np.repeat([c.mean for c in cases], np.diff(no)).astype(np.float64)
    k = np.where(prior, 0, g["nfound"])
    for r in (0, R - 1):
        xa = np.concatenate([out[r], vals])
        acc = np.zeros(NT)
        for j in range(plan.kmax):
            use = j < k
            term = lam[:, j] * (xa[np.where(use, nbr[:, j], 0)] - mean)
            acc = np.where(use, acc + term, acc)
        x = (mean + acc) + sigma * z[r, pos]
        np.testing.assert_array_equal(out[r], x, err_msg=f"realization {r}")
    del xa, acc, term, x
    say("config5: sweep recurrence verified")
    # ten requests against the oracle
    pick = [int(np.argmax(sides))]
    for kmax in (8, 16, 32):
        for fam in (O.SPHERICAL, O.EXPONENTIAL, O.GAUSSIAN):
            cand = [q for q, c in enumerate(cases)
                    if c.kmax == kmax and c.family == fam and sides[q] <= 160 and q not in pick]
            if cand:
                pick.append(cand[0])
    assert len(pick) >= 10
    for q in pick[:10]:
        c = cases[q]
        say(f"config5 request {q}: side {c.dims[0]} family {c.family} k {c.kmax} data {c.n_data}")
        o = O.seqsim_plan(c)
        rows = slice(int(no[q]), int(no[q + 1]))
        kq = min(c.kmax, c.n_nodes)
        loc = nbr[rows].astype(np.int64)
        loc = np.where(loc < 0, -1, np.where(loc < NT, loc - no[q], loc - NT - do[q] + c.n_nodes))
        np.testing.assert_array_equal(g["nfound"][rows], o.nfound)
        np.testing.assert_array_equal(loc[:, :kq], o.nbr)
        np.testing.assert_array_equal(flags[rows], o.flags)
        np.testing.assert_array_equal(level[rows], o.level)
        np.testing.assert_array_equal(lam[rows][:, :kq], o.lam)
        np.testing.assert_array_equal(sigma[rows], o.sigma)
        np.testing.assert_array_equal(out[:, rows], O.seqsim_sweep(c, o, z[:, rows]))
