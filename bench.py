#!/usr/bin/env python3
"""SEQSIM throughput on one MI355X: nodes x realizations per second.

Workload "seqsim-3d-128x128x64": a 128 x 128 x 64 unit grid
(1,048,576 nodes), anisotropic exponential variogram sill 1, ranges
(40, 40, 8), nugget 0.05, 2000 conditioning points uniform in the domain with
N(0,1) values, maxneighbors 24 with the :range neighbourhood (radius 40), a
random path, 64 realizations, local solves in FP32 with one FP64-residual
refinement step, realizations in FP64. One step is one full
`rand(process, grid, 64; method=SEQSIM(...))`: plan (neighbour search +
kriging solves) and the 64-realization sweep, with the path, the data and the
64 x N standard normals already resident in HBM.

Workload "seqsim-learned-256x256x128" (default): a 256 x 256 x 128 unit grid
(8,388,608 nodes), exponential variogram sill 1, range 20, nugget 0.01,
maxneighbors 24 with the :range neighbourhood, no conditioning data, a random
path, 16 realizations. The kriging weights and variances come from the learned
surrogate (BF16 MLP 97-256-256-256-33 on the matrix cores) wherever it applies
and its variance check passes, from exact FP32 + refinement solves elsewhere.
The surrogate is trained before timing on 2M kriging systems of exact plans of
the same variogram (three 128 x 128 x 64 grids). One step is one learned plan
plus the 16-realization sweep; the same steps with exact kriging are timed
beside it, and the line reports the RMS weight error against the exact FP64
plan and the paired experimental-variogram error (learned vs exact
realizations from the same normals).

Workload "seqsim-batch-512": 512 independent 2D requests in one plan and one
sweep: square grids with side log-uniform in [32, 512] (about 23M nodes in
total), spherical/exponential/gaussian variograms (sill 1) with range uniform
in [5, 60] cells and nugget uniform in [0, 0.1], 0-500 uniform data each,
maxneighbors from {8, 16, 32}, :range neighbourhood, random paths, 8
realizations each, FP32 solves with refinement. One step is one batched plan
plus the batched sweep.

Workload "seqsim-batch-512-learned": the same 512 requests with learned
weights: one surrogate per variogram family, trained before timing on up to
4M systems (20 epochs) of the exact plans of 3 x 96 other requests of the same
generator; each request
inside its family surrogate's trained spans of range and nugget takes learned
weights where the variance check passes, the rest are solved exactly. The
exact batch is timed beside it; the accuracy (RMS weight error against exact
FP64, paired variograms) is measured on 32 of the requests.

Workload "kriging-mean-256x256x128": the Gaussian posterior mean (simple
kriging, `mean(process, grid; data)`) on a 256 x 256 x 128 unit grid
(8,388,608 nodes) from 100,000 data uniform in the domain, exponential
variogram sill 1, range 20, nugget 0.01, the 10 nearest data of every node
(GeoStatsModels' fitpredict defaults), FP64 solves. One step is one
data-only plan plus its zero-normal sweep (the mean at every node).

Workload "indicator-3d-128x128x64": IndicatorProcess SEQSIM with 3 categories
(probabilities 0.5, 0.3, 0.2, coregionalized A * exponential covariance with
the config-3 ranges (40, 40, 8), nugget 0.05) on the 128 x 128 x 64 grid,
2000 categorical data, maxneighbors 24 with the :range neighbourhood, random
path, 64 realizations. One step is one plan plus the 64-realization
categorical sweep, uniforms resident in HBM.

Workload "fftsim-256x256x128": FFT moving-average simulation (FFTSIM, the
default method of unconditional grids) on a 256 x 256 x 128 unit grid
(8,388,608 cells), Gaussian variogram sill 1, range 20, 16 realizations from
uniform white noise resident in HBM. One step is one plan (covariance,
amplitude spectrum) plus the 16 realizations (forward and inverse transforms,
phase, variance scaling).

Workload "ensemble-stats-32x256x256x128": per-cell statistics of an ensemble
of 32 realizations on a 256 x 256 x 128 grid (2.1 GB, realization-major in
HBM): mean, variance, cdf(0), ccdf(1) and the 5%, 50% and 95% quantiles in
one pass. The same statistics without the quantiles (a pure streaming pass)
are timed beside it.

Workload "lusim-64x64": LUSIM on a 64 x 64 unit grid (4,096 cells),
exponential covariance sill 1, range 20, nugget 0.01, 200 uniform conditioning
points snapped to cells by NearestInit (later data win), 256 realizations.
One step is one factorization (covariance assembly, L11, B12, d2 and the
blocked Cholesky of C22 - B12' B12 on the matrix cores, FP64) plus the
256-realization sample d2 + L22 W and scatter, with the data and the normals
resident in HBM. The mixed form (FP32 MFMA trailing updates, FP64 diagonal
tiles, one FP64 refinement step of the conditioning solve) is timed beside it;
both are compared per realization with the FP64 numpy/scipy restatement,
which is also the CPU baseline (LAPACK at 1 thread and at every core).

Prints one JSON line. `cpu_baseline` times the CPU restatement of the
reference's cost structure (search + solve per node per realization) on a
bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "geostatsprocesses.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "seqsim-3d-128x128x64": dict(dims=(128, 128, 64), family=1, sill=1.0, ranges=(40.0, 40.0, 8.0),
                                 nugget=0.05, n_data=2000, kmax=24, radius=40.0, R=64, seed=1003),
    "seqsim-2d-100x100": dict(dims=(100, 100), family=0, sill=1.0, ranges=(10.0, 10.0), nugget=0.0,
                              n_data=50, kmax=12, radius=30.0, R=10, seed=1001, precision="fp64"),
    "seqsim-batch-512": dict(batch=512, side=(32, 512), R=8, seed=1005),
    "seqsim-batch-512-learned": dict(
        batch=512, side=(32, 512), R=8, seed=1005,
        learned=dict(train_requests=192, train_seed=3005, samples=4_000_000, epochs=20, batch=1024,
                     lr=2e-3, tolerance=0.05, accuracy_requests=32)),
    "seqsim-learned-256x256x128": dict(
        dims=(256, 256, 128), family=1, sill=1.0, ranges=(20.0, 20.0, 20.0), nugget=0.01, n_data=0,
        kmax=24, radius=20.0, R=16, seed=1004,
        learned=dict(train_dims=(128, 128, 64), train_plans=3, samples=2_000_000, epochs=20,
                     batch=1024, lr=2e-3, tolerance=0.05)),
    "seqsim-learned-256x256x128-conditional": dict(
        dims=(256, 256, 128), family=1, sill=1.0, ranges=(20.0, 20.0, 20.0), nugget=0.01,
        n_data=16000, kmax=24, radius=20.0, R=16, seed=1010,
        learned=dict(train_dims=(128, 128, 64), train_plans=3, train_data=2000,
                     samples=2_000_000, epochs=20, batch=1024, lr=2e-3, tolerance=0.05,
                     data_slots=4)),
    "kriging-mean-256x256x128": dict(
        dims=(256, 256, 128), family=1, sill=1.0, ranges=(20.0, 20.0, 20.0), nugget=0.01,
        n_data=100_000, kmax=10, radius=0.0, R=1, seed=1006, mode="data_only", precision="fp64",
        cpu_nodes=1_500_000),
    "ensemble-stats-32x256x256x128": dict(dims=(256, 256, 128), R=32, seed=1008, method="stats",
                                          cpu_cells=1_000_000),
    "indicator-3d-128x128x64": dict(dims=(128, 128, 64), family=1, sill=1.0, ranges=(40.0, 40.0, 8.0),
                                    nugget=0.05, n_data=2000, kmax=24, radius=40.0, R=64, seed=1009,
                                    method="indicator", prob=(0.5, 0.3, 0.2)),
    "fftsim-256x256x128": dict(dims=(256, 256, 128), family=2, sill=1.0, ranges=(20.0, 20.0, 20.0),
                               nugget=0.0, R=16, seed=1007, method="fftsim", cpu_reals=2),
    "lusim-64x64": dict(dims=(64, 64), family=1, sill=1.0, ranges=(20.0, 20.0), nugget=0.01,
                        n_data=200, R=256, seed=1002, method="lusim"),
}
FP64_PEAK_TFLOPS = 78.6


def batch_specs(spec):
    """The requests of a batch workload (same draws as the tests' mixed-batch
    generator): one spec dict per request."""
    rng = np.random.default_rng(spec["seed"])
    lo, hi = spec["side"]
    out = []
    for _ in range(spec["batch"]):
        side = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        fam = int(rng.integers(0, 3))
        rang = float(rng.uniform(5.0, 60.0))
        nug = float(rng.uniform(0.0, 0.1))
        nd = int(rng.integers(0, 501))
        kmax = int(rng.choice([8, 16, 32]))
        data = rng.uniform(0, side, (nd, 2))
        vals = rng.standard_normal(nd)
        path = rng.permutation(side * side).astype(np.int64)
        out.append(dict(dims=(side, side), family=fam, sill=1.0, ranges=(rang, rang), nugget=nug,
                        n_data=nd, kmax=kmax, radius=rang, data=np.ascontiguousarray(data),
                        vals=vals, path=path))
    return out


def make_inputs(spec):
    dims = spec["dims"]
    N = int(np.prod(dims))
    rng = np.random.default_rng(spec["seed"])
    # the posterior mean visits the nodes in order (its result does not depend on it)
    path = np.arange(N, dtype=np.int64) if spec.get("mode") == "data_only" \
        else rng.permutation(N).astype(np.int64)
    data = np.random.default_rng(spec["seed"] + 1).uniform(0, 1, (spec["n_data"], len(dims))) \
        * np.asarray(dims, dtype=np.float64)
    vals = np.random.default_rng(spec["seed"] + 2).standard_normal(spec["n_data"])
    return N, path, np.ascontiguousarray(data), vals


def problem(spec, path_ptr, data_ptr, precision, N):
    from geostatsprocesses import _lib
    dims = spec["dims"]
    nd = len(dims)
    p = _lib.GspSeqsimProblem()
    p.grid.ndim = nd
    for d in range(3):
        p.grid.dims[d] = dims[d] if d < nd else 1
        p.grid.origin[d] = 0.0
        p.grid.spacing[d] = 1.0
        p.model.ranges[d] = spec["ranges"][d] if d < nd else 1.0
    p.model.family = spec["family"]
    p.model.kind = _lib.VARIOGRAM
    p.model.sill = spec["sill"]
    p.model.nugget = spec["nugget"]
    p.mean = 0.0
    p.n_data = spec["n_data"]
    p.data_coords = data_ptr
    p.n_path = N
    p.path = path_ptr
    p.minneighbors = 1
    p.maxneighbors = spec["kmax"]
    p.radius = spec["radius"]
    p.search_mode = _lib.DATA_ONLY if spec.get("mode") == "data_only" else _lib.KNN_THEN_MASK
    p.precision = precision
    p.device_pointers = 1
    return p


# kernels of each bench name in the committed traffic profiles
PROFILE_KERNELS = {"search": ["k_search<false>", "k_search<true>"],
                   "krige": ["k_krige_lane<8, 1>", "k_krige_lane<10, 0>", "k_krige16<1>",
                             "k_krige16<0>", "k_krige<24, 1>", "k_krige<24, 0>", "k_krige<32, 1>",
                             "k_krige<32, 0>", "k_krige32<24, 1>", "k_krige32<24, 0>",
                             "k_krige32<32, 1>", "k_krige32<32, 0>"],
                   "sweep": ["k_sweep_level"], "learned": ["k_mlp_infer"],
                   "stats": ["k_ens_stats<false>"]}


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of the kernel family `kernel` (read + written,
    counter-calibrated) from profiles/<workload>_traffic.md, which rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE runs of this script produced: the sum over the
    family's kernels of each one's average per dispatch (a family launch, such
    as one plan's kriging, runs each of its kernels once); None when no profile
    covers it."""
    path = os.path.join(ROOT, "profiles", workload.replace("-", "_") + "_traffic.md")
    if not os.path.exists(path):
        return None
    total, found = 0.0, False
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] in PROFILE_KERNELS.get(kernel, []):
            try:
                per = (float(cells[2]) + float(cells[3])) * 1e6
            except ValueError:
                continue
            total += per
            found = True
    return int(total) if found else None


def cpu_share():
    """(CPU model, cores visible to this process, threads used): one GPU's
    share of the box's host is 16 cores, whatever the machine has."""
    model, ncores = cpu_info()
    return model, ncores, max(1, min(16, ncores))


def cpu_baseline(reqs, R_batch=None, node_budget=None, node_limit=0):
    """The oracle's restatement of the reference's cost structure: search +
    FP64 Cholesky solve per node per realization (seqsim.jl:98-138), with a
    cell-hash neighbour search (on grids faster than the reference's KD-tree,
    so the figure favours the CPU), OpenMP over realizations like pmap over
    workers (field.jl:113-119). Two legs on the workload's own grid(s): one
    thread (the reference's default of one process) and the GPU's 16-core
    share of the host.

    Single workloads: whole realizations, one per thread, repeated until the
    leg has run for 3 s. Batches: the first requests up to `node_budget` nodes,
    R_batch realizations each. node_limit: the first node_limit path nodes of
    each realization (the posterior mean, whose nodes are independent)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    model, ncores, share = cpu_share()

    def leg(threads):
        n_total, dt_total, used, nodes, reps = 0, 0.0, 0, 0, 0
        for spec in reqs:
            dims = spec["dims"]
            case = O.SeqsimCase(dims=dims, family=spec["family"], sill=spec["sill"],
                                ranges=spec["ranges"][:len(dims)], nugget=spec["nugget"],
                                kmax=spec["kmax"], radius=spec["radius"], data_coords=spec["data"],
                                data_vals=spec["vals"], path=spec["path"], precision=O.FP64,
                                mode=O.DATA_ONLY if spec.get("mode") == "data_only" else O.KNN_THEN_MASK)
            N = case.n_nodes
            R = R_batch if R_batch is not None else threads
            z = np.random.default_rng(7).standard_normal((R, N))
            reps = 0
            while True:
                t0 = time.perf_counter()
                _, n = O.seqsim_reference_cost(case, z, threads=threads, n_nodes_limit=node_limit or N)
                dt_total += time.perf_counter() - t0
                n_total += n
                reps += 1
                if R_batch is not None or dt_total >= 3.0:
                    break
            used += 1
            nodes += N
            if node_budget is None or nodes >= node_budget:
                break
        return n_total / dt_total, dt_total, used, nodes, reps

    v1, t1, used, nodes, reps1 = leg(1)
    vs, ts, _, _, repss = leg(share)
    if node_limit:
        what = f"the first {node_limit} nodes of the path"
        s1, ss = f"1 pass over {what} ({t1:.1f} s)", f"{share} passes over {what} ({ts:.1f} s)"
    elif R_batch is None:
        what = f"full realizations of the workload's {nodes}-node grid"
        s1 = f"{reps1} {what} at 1 thread ({t1:.1f} s)"
        ss = f"{share * repss} {what}, one per thread ({ts:.1f} s)"
    else:
        what = f"the first {used} requests of the batch ({nodes} nodes), {R_batch} full realizations each"
        s1, ss = f"{what} at 1 thread ({t1:.1f} s)", f"{what} over {share} threads ({ts:.1f} s)"
    common = ("search + FP64 kriging solve per node per realization (the reference's cost "
              "structure; cell-hash search)")
    return {"value": vs, "unit": "nodes*realizations/s", "cores": share, "kind": "port",
            "sample": f"{ss}; {common}",
            "one_thread": {"value": v1, "cores": 1, "sample": s1},
            "cpu_model": model, "cores_visible": ncores}


def mlp_flops(slots):
    """Flops of one forward pass of the S-slot surrogate (3S+1 -> 256 -> 256 ->
    256 -> S+1): 312,320 at S = 24, 328,704 at S = 32."""
    return 2 * ((3 * slots + 1) * 256 + 2 * 256 * 256 + 256 * (slots + 1))
BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3


def krige_flops(nfound, flags):
    """Algorithmic flops of the exact kriging solves of a plan: per solved
    system of n neighbours n^3/3 + 2n^2 + 2n (factor and two substitutions)
    + 10 (n(n+1)/2 + n) (covariance entries); prior draws (flags & 3) and
    learned nodes (flags & 8) solve nothing."""
    n = np.where((flags & 11) != 0, 0, nfound).astype(np.float64)
    f = n ** 3 / 3 + 2 * n * n + 2 * n + 10 * (n * (n + 1) / 2 + n)
    return float(f.sum())


def upload_normals(ctx, _lib, R, N, seed):
    z = _lib.DeviceBuffer(ctx, (R, N))
    for r0 in range(R):  # row by row keeps host memory small
        z_row = np.random.default_rng(seed + 3 + r0).standard_normal(N)
        _lib.check(_lib.load().gsp_memcpy_to_device(ctx.handle, z.ptr + r0 * N * 8,
                                                    z_row.ctypes.data, N * 8))
    return z


def timed(ctx, step, steps, warmup):
    """W untimed steps, then K steps between device synchronisations."""
    for _ in range(warmup):
        step().close()
    ctx.synchronize()
    ctx.enable_timing(True)
    ctx.timings(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        step().close()
    ctx.synchronize()
    t1 = time.perf_counter()
    tim = ctx.timings(reset=True)
    ctx.enable_timing(False)
    return (t1 - t0) * 1e3 / steps, tim


def kernel_table(_lib, tim, steps, N, R, K, kbar, n_mlp=0, kflops=None, mlp_fl=0):
    """Average launch time per kernel family and its algorithmic bytes (flops for
    the surrogate) per launch."""
    sweep_bytes = N * R * (8 + 8) + N * (kbar * (4 + 8) + 8)     # z read, x write, plan read
    search_bytes = N * (8 + K * 4 + 4)                           # path, nbr, nfound
    plan_bytes = N * (K * (4 + 8) + 4 + 8 + 1 + 4)               # nbr + nfound in, lambda, sigma, flags, degrees
    kern = {}
    for slot, name, nbytes in ((_lib.K_SEARCH, "search", search_bytes), (_lib.K_PLAN, "krige", plan_bytes),
                               (_lib.K_LEARNED, "learned", None), (_lib.K_SWEEP, "sweep", sweep_bytes),
                               (_lib.K_LEVELS, "levels", None), (_lib.K_TRANSPOSE, "transpose", 2 * 8 * N * R)):
        ms, launches = tim[slot]
        if launches == 0:
            continue
        avg = ms / launches
        kern[name] = {"avg_ms": round(avg, 4), "launches_per_step": launches // steps}
        if nbytes is not None:
            per_launch = nbytes / (launches / steps)
            kern[name]["algorithmic_bytes_per_launch"] = int(per_launch)
            kern[name]["achieved_GBs"] = round(per_launch / (avg * 1e-3) / 1e9, 1)
        if name == "krige" and kflops is not None:
            per_launch = kflops / (launches / steps)
            kern[name]["algorithmic_flops_per_launch"] = int(per_launch)
            kern[name]["achieved_TFLOPs"] = round(per_launch / (avg * 1e-3) / 1e12, 3)
        if name == "learned":
            kern[name]["algorithmic_flops_per_launch"] = int(mlp_fl * n_mlp)
            kern[name]["achieved_TFLOPs"] = round(mlp_fl * n_mlp / (avg * 1e-3) / 1e12, 2)
    return kern


def roofline_of(kern, workload, fp64=False):
    dominant = max(("search", "krige", "learned", "sweep"),
                   key=lambda k: kern.get(k, {}).get("avg_ms", 0) * kern.get(k, {}).get("launches_per_step", 0))
    d = kern[dominant]
    if dominant == "krige" and "achieved_TFLOPs" in d:
        # compute-bound (vector ALU): against the peak of the factor's type
        peak = 78.6 if fp64 else FP32_PEAK_TFLOPS
        return {"bound": "mfma", "kernel": "krige (k_krige16 + k_krige/k_krige32, vector ALU)",
                "achieved": d["achieved_TFLOPs"], "peak": peak, "unit": "TFLOP/s",
                "frac": round(d["achieved_TFLOPs"] / peak, 4),
                "traffic": profiled_traffic(workload, dominant)}
    if dominant == "learned":
        return {"bound": "mfma", "kernel": dominant, "achieved": d["achieved_TFLOPs"],
                "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(d["achieved_TFLOPs"] / BF16_PEAK_TFLOPS, 4),
                "traffic": profiled_traffic(workload, dominant)}
    return {"bound": "hbm", "kernel": dominant, "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(d["achieved_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(workload, dominant)}


def model_variogram(spec, h):
    u = np.asarray(h, float) / spec["ranges"][0]
    shape = 1.0 - np.exp(-3.0 * u)
    return (spec["sill"] - spec["nugget"]) * shape + spec["nugget"]


def axis_variogram(fields, dims, axis, lags):
    """Experimental semivariogram of realizations (R, N) along one grid axis,
    averaged over realizations: gamma(h) = mean (x(i + h e) - x(i))^2 / 2."""
    X = fields.reshape((fields.shape[0],) + tuple(reversed(dims)))  # column-major ids
    Xm = np.moveaxis(X, len(dims) - axis, -1)
    return np.array([0.5 * float(np.mean((Xm[..., h:] - Xm[..., :-h]) ** 2, dtype=np.float64))
                     for h in lags])


def learned_accuracy(ctx, _lib, spec, prob_fp32r, prob_fp64, prob_masked, S, tol, z, dvals, N, R):
    """RMS weight error of a learned plan against the exact FP64 plan, and the
    experimental variograms of learned and exact realizations drawn from the
    same normals (paired), along each axis at lags 1..20; beside them the same
    variogram of exact realizations with the masked k-nearest search (the k
    nearest among simulated points) against the model."""
    lp = _lib.SeqsimPlan(ctx, prob_fp32r, surrogate=S, tolerance=tol)
    ep = _lib.SeqsimPlan(ctx, prob_fp64)
    gl = lp.export(fields=("nbr", "lam", "sigma", "flags"))
    ge = ep.export(fields=("lam", "sigma"))
    learned = (gl["flags"] & _lib.FLAG_LEARNED) != 0
    valid = gl["nbr"] >= 0
    with_data = learned & (gl["nbr"] >= N).any(axis=1)
    d = gl["lam"] - ge["lam"]
    rms_learned = float(np.sqrt(np.mean(d[learned][valid[learned]] ** 2)))
    rms_data = float(np.sqrt(np.mean(d[with_data][valid[with_data]] ** 2))) if with_data.any() \
        else None
    data_nodes = int((gl["nbr"] >= N).any(axis=1).sum())
    rms_all = float(np.sqrt(np.mean(d[valid] ** 2)))
    sig = np.abs(gl["sigma"] - ge["sigma"])[learned] / np.maximum(ge["sigma"][learned], 1e-300)
    del gl, ge, d, valid
    Rv = min(R, 4)
    dims = spec["dims"]
    outs = []
    mp = _lib.SeqsimPlan(ctx, prob_masked)
    for plan in (lp, ep, mp):
        out = _lib.DeviceBuffer(ctx, (R, N))
        plan.sweep(z, dvals, out, device_pointers=True)
        ctx.synchronize()
        host = np.empty((Rv, N))
        _lib.check(_lib.load().gsp_memcpy_to_host(ctx.handle, host.ctypes.data, out.ptr, Rv * N * 8))
        outs.append(host.astype(np.float32))
        out.free()
    lags = list(range(1, 21))
    paired, vs_model, exact_vs_model, masked_vs_model = [], [], [], []
    gm = model_variogram(spec, lags)
    for axis in range(len(dims)):
        gl_ = axis_variogram(outs[0], dims, axis, lags)
        ge_ = axis_variogram(outs[1], dims, axis, lags)
        gk_ = axis_variogram(outs[2], dims, axis, lags)
        paired.append(float(np.linalg.norm(gl_ - ge_) / np.linalg.norm(ge_)))
        vs_model.append(float(np.linalg.norm(gl_ - gm) / np.linalg.norm(gm)))
        exact_vs_model.append(float(np.linalg.norm(ge_ - gm) / np.linalg.norm(gm)))
        masked_vs_model.append(float(np.linalg.norm(gk_ - gm) / np.linalg.norm(gm)))
    lp.close()
    ep.close()
    mp.close()
    return {"rms_weight_error_learned_nodes": rms_learned, "rms_weight_error_all_nodes": rms_all,
            "sigma_rel_error_median": float(np.median(sig)), "sigma_rel_error_p99": float(np.quantile(sig, 0.99)),
            "learned_fraction": float(learned.mean()),
            "nodes_with_data_neighbours": data_nodes,
            "learned_with_data_neighbours": int(with_data.sum()),
            "rms_weight_error_learned_with_data": rms_data,
            "variogram_error_paired_per_axis": [round(e, 5) for e in paired],
            "variogram_error_vs_model_per_axis": [round(e, 5) for e in vs_model],
            "exact_variogram_error_vs_model_per_axis": [round(e, 5) for e in exact_vs_model],
            "masked_knn_exact_variogram_error_vs_model_per_axis": [round(e, 5) for e in masked_vs_model],
            "variogram_sample": f"{Rv} realizations, lags 1..20 cells along each axis"}


def train_batch_surrogates(ctx, _lib, spec, precision):
    """One surrogate per variogram family for a mixed batch: exact plans of
    `train_requests` other requests drawn like the batch's (another seed), the
    requests of each family in one batch plan, then device training on up to
    `samples` of that family's kriging systems. Each surrogate then serves the
    batch requests of its family inside the ranges and nuggets it saw."""
    L = spec["learned"]
    treqs = batch_specs(dict(spec, batch=L["train_requests"] * 3, seed=L["train_seed"]))
    surs, reports = [], []
    for fam in (0, 1, 2):
        fr = [r for r in treqs if r["family"] == fam]
        sizes = [int(np.prod(r["dims"])) for r in fr]
        dpath = _lib.DeviceBuffer.from_array(ctx, np.concatenate([r["path"] for r in fr]))
        ddata = _lib.DeviceBuffer.from_array(ctx, np.concatenate([r["data"].ravel() for r in fr]
                                                                 + [np.zeros(1)]))
        probs, noff, coff = [], 0, 0
        for r, n in zip(fr, sizes):
            probs.append(problem(r, dpath.ptr + 8 * noff, ddata.ptr + 8 * coff if r["n_data"] else None,
                                 precision, n))
            noff += n
            coff += r["data"].size
        plan = _lib.SeqsimPlan(ctx, probs)
        S = _lib.Surrogate(ctx, [plan], max_samples=L["samples"], epochs=L["epochs"],
                           batch=L["batch"], lr=L["lr"], seed=L.get("seed", 0) + fam,
                           data_slots=L.get("data_slots", 0))
        plan.close()
        dpath.free()
        ddata.free()
        r = S.report
        info = S.info
        surs.append(S)
        reports.append({"family": ["spherical", "exponential", "gaussian"][fam], "requests": len(fr),
                        "systems": int(r.n_samples), "candidates": int(r.n_candidates),
                        "steps": int(r.steps), "ms": round(r.ms, 1),
                        "data_slots": info["n_data_slots"],
                        "rms_weight_last_epoch": r.rms_weight_last_epoch,
                        "nugget_ratio_span": [round(x, 4) for x in info["nugget_ratio"]],
                        "range_span_cells": [round(1 / info["scale"][0][1], 2),
                                             round(1 / info["scale"][0][0], 2)]})
    return surs, {"per_family": reports, "epochs": L["epochs"], "batch": L["batch"],
                  "ms": round(sum(x["ms"] for x in reports), 1),
                  "source": f"{L['train_requests']} x 3 other requests of the same generator "
                            f"(seed {L['train_seed']}), exact plans"}


def batch_learned_accuracy(ctx, _lib, reqs, probs32, probs64, S, tol, R, n_req):
    """Learned against exact FP64 on a sample of the batch's requests (every
    (512 / n_req)-th): RMS weight error of learned nodes, sigma error, and the
    paired experimental variograms (same normals) along both axes at lags
    1..10, per request, aggregated as the RMS of the relative errors."""
    pick = list(range(0, len(reqs), max(1, len(reqs) // n_req)))[:n_req]
    p32 = [probs32[q] for q in pick]
    p64 = [probs64[q] for q in pick]
    lp = _lib.SeqsimPlan(ctx, p32, surrogate=S, tolerance=tol)
    ep = _lib.SeqsimPlan(ctx, p64)
    gl = lp.export(fields=("nbr", "lam", "sigma", "flags"))
    ge = ep.export(fields=("lam", "sigma"))
    learned = (gl["flags"] & _lib.FLAG_LEARNED) != 0
    valid = gl["nbr"] >= 0
    d = gl["lam"] - ge["lam"]
    rms = float(np.sqrt(np.mean(d[learned][valid[learned]] ** 2))) if learned.any() else None
    sig = np.abs(gl["sigma"] - ge["sigma"])[learned] / np.maximum(ge["sigma"][learned], 1e-300)
    no = lp.node_offset
    NT = int(no[-1])
    rng = np.random.default_rng(77)
    Rv = min(R, 4)
    z = rng.standard_normal((Rv, NT))
    vals = np.concatenate([reqs[q]["vals"] for q in pick] + [np.zeros(0)])
    xl = lp.sweep(z, vals)
    xe = ep.sweep(z, vals)
    errs = []
    for i, q in enumerate(pick):
        dims = reqs[q]["dims"]
        a, b = int(no[i]), int(no[i + 1])
        lags = list(range(1, 11))
        for axis in range(2):
            gl_ = axis_variogram(xl[:, a:b], dims, axis, lags)
            ge_ = axis_variogram(xe[:, a:b], dims, axis, lags)
            errs.append(float(np.linalg.norm(gl_ - ge_) / np.linalg.norm(ge_)))
    lp.close()
    ep.close()
    return {"requests_sampled": len(pick), "learned_fraction_sampled": float(learned.mean()),
            "rms_weight_error_learned_nodes": rms,
            "sigma_rel_error_median": float(np.median(sig)) if learned.any() else None,
            "variogram_error_paired_rms": float(np.sqrt(np.mean(np.square(errs)))),
            "variogram_error_paired_max": float(np.max(errs)),
            "variogram_sample": f"{Rv} realizations per request, lags 1..10 cells along each axis"}


def train_surrogate(ctx, _lib, spec, precision):
    """Exact plans of the same variogram on smaller grids (random paths; with
    `train_data` uniform conditioning points each when given), then device
    training on up to `samples` of their kriging systems."""
    L = spec["learned"]
    plans, bufs = [], []
    nd_train = L.get("train_data", 0)
    for t in range(L["train_plans"]):
        tdims = L["train_dims"]
        n = int(np.prod(tdims))
        path = np.random.default_rng(2000 + t).permutation(n).astype(np.int64)
        dp = _lib.DeviceBuffer.from_array(ctx, path)
        bufs.append(dp)
        dd = None
        if nd_train:
            xy = np.random.default_rng(2100 + t).uniform(0, 1, (nd_train, len(tdims))) \
                * np.asarray(tdims, dtype=np.float64)
            dd = _lib.DeviceBuffer.from_array(ctx, np.ascontiguousarray(xy))
            bufs.append(dd)
        plans.append(_lib.SeqsimPlan(ctx, problem(dict(spec, dims=tdims, n_data=nd_train), dp.ptr,
                                                  dd.ptr if dd is not None else None, precision, n)))
    S = _lib.Surrogate(ctx, plans, max_samples=L["samples"], epochs=L["epochs"], batch=L["batch"],
                       lr=L["lr"], seed=L.get("seed", 0), data_slots=L.get("data_slots", 0))
    r = S.report
    for p in plans:
        p.close()
    return S, {"systems": int(r.n_samples), "candidates": int(r.n_candidates), "steps": int(r.steps),
               "epochs": L["epochs"], "batch": L["batch"], "ms": round(r.ms, 1),
               "rms_weight_first_epoch": r.rms_weight_first_epoch,
               "rms_weight_last_epoch": r.rms_weight_last_epoch,
               "data_slots": S.info["n_data_slots"],
               "source": f"{L['train_plans']} exact plans of a {'x'.join(map(str, L['train_dims']))} grid, "
                         f"same variogram, random paths, {nd_train} data each"}


def fftsim_problem(_lib, spec):
    dims = spec["dims"]
    nd = len(dims)
    p = _lib.GspFftsimProblem()
    p.grid.ndim = nd
    for d in range(3):
        p.grid.dims[d] = dims[d] if d < nd else 1
        p.grid.spacing[d] = 1.0
        p.model.ranges[d] = spec["ranges"][d] if d < nd else 1.0
    p.model.family = spec["family"]
    p.model.kind = _lib.VARIOGRAM
    p.model.sill = spec["sill"]
    p.model.nugget = spec["nugget"]
    p.minneighbors, p.maxneighbors = 1, 26
    return p


def bench_fftsim(args, spec):
    """FFTSIM: plan + R realizations per step; the transforms are rocFFT's."""
    from geostatsprocesses import _lib
    dims, R = spec["dims"], spec["R"]
    N = int(np.prod(dims))
    ctx = _lib.default_context(0)
    w = _lib.DeviceBuffer(ctx, (R, N))
    for r in range(R):
        row = np.random.default_rng(spec["seed"] + r).random(N)
        _lib.check(_lib.load().gsp_memcpy_to_device(ctx.handle, w.ptr + r * N * 8, row.ctypes.data, N * 8))
    out = _lib.DeviceBuffer(ctx, (R, N))
    prob = fftsim_problem(_lib, spec)

    class Step:
        def __init__(self):
            self.plan = _lib.FftsimPlan(ctx, prob)
            self.plan.sample(w, out, device_pointers=True)
            ctx.synchronize()

        def close(self):
            self.plan.close()

    ms_per_step, tim = timed(ctx, Step, args.steps, args.warmup)
    Nh = (dims[0] // 2 + 1) * int(np.prod(dims[1:]))
    kern = {}
    for slot, name in ((_lib.K_FFT, "fft"), (_lib.K_SPECTRAL, "spectral")):
        ms, launches = tim[slot]
        if launches:
            kern[name] = {"avg_ms": round(ms / launches, 4), "launches_per_step": launches // args.steps,
                          "ms_per_step": round(ms / args.steps, 3)}
    # transforms: a real-to-complex pass reads N doubles and writes Nh complex
    # values, the inverse the reverse (one plan transform + 2 per realization)
    fft_bytes = (N * 8 + Nh * 16) * (2 * R + 1)
    fft_launches = kern["fft"]["launches_per_step"]
    per_launch = fft_bytes / fft_launches
    achieved = per_launch / (kern["fft"]["avg_ms"] * 1e-3) / 1e9
    kern["fft"] |= {"algorithmic_bytes_per_launch": int(per_launch), "achieved_GBs": round(achieved, 1)}
    roofline = {"bound": "hbm", "kernel": "fft (rocFFT)", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        F = O.fftsim_amplitude(dims, spec["family"], spec["sill"], spec["nugget"], spec["ranges"])
        nr = spec["cpu_reals"]
        O.fftsim_unconditional(F, np.random.default_rng(0).random((nr, N)), spec["sill"], 0.0)
        dt = time.perf_counter() - t0
        cpu = {"value": round(N * nr / dt, 1), "unit": "cells*realizations/s", "cores": 1,
               "kind": "port",
               "sample": f"one plan and {nr} realizations of the same grid with numpy's FFT "
                         f"({dt:.1f} s, one thread)"}
    line = {
        "metric": "FFTSIM cells*realizations/s",
        "value": round(N * R / (ms_per_step / 1e3), 1),
        "unit": "cells*realizations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "grid": list(dims), "realizations": R,
                   "variogram": "gaussian", "range": spec["ranges"][0]},
        "roofline": roofline, "kernels": kern, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_stats(args, spec):
    """Ensemble statistics: one gsp_ensemble_stats call per step."""
    from geostatsprocesses import _lib
    dims, R = spec["dims"], spec["R"]
    N = int(np.prod(dims))
    ctx = _lib.default_context(0)
    reals = _lib.DeviceBuffer(ctx, (R, N))
    for r in range(R):
        row = np.random.default_rng(spec["seed"] + r).standard_normal(N)
        _lib.check(_lib.load().gsp_memcpy_to_device(ctx.handle, reals.ptr + r * N * 8, row.ctypes.data, N * 8))
    full = [(_lib.STAT_MEAN, 0.0), (_lib.STAT_VAR, 0.0), (_lib.STAT_CDF, 0.0), (_lib.STAT_CCDF, 1.0),
            (_lib.STAT_QUANTILE, 0.05), (_lib.STAT_QUANTILE, 0.5), (_lib.STAT_QUANTILE, 0.95)]
    moments = full[:4]
    out = _lib.DeviceBuffer(ctx, (len(full), N))

    def run(stats):
        class Step:
            def __init__(self):
                _lib.ensemble_stats(ctx, reals, stats, out=out, device_pointers=True, shape=(R, N))
                ctx.synchronize()

            def close(self):
                pass
        return timed(ctx, Step, args.steps, args.warmup)

    ms_full, tim_full = run(full)
    ms_mom, tim_mom = run(moments)

    def kern(tim, n_stats):
        ms, launches = tim[_lib.K_STATS]
        avg = ms / launches
        nbytes = R * N * 8 + n_stats * N * 8
        return {"avg_ms": round(avg, 4), "launches_per_step": launches // args.steps,
                "algorithmic_bytes_per_launch": nbytes,
                "achieved_GBs": round(nbytes / (avg * 1e-3) / 1e9, 1)}
    kf, km = kern(tim_full, len(full)), kern(tim_mom, len(moments))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        n = spec["cpu_cells"]
        sample = np.random.default_rng(0).standard_normal((R, n))
        t0 = time.perf_counter()
        for kind, p in (("mean", 0), ("var", 0), ("cdf", 0.0), ("ccdf", 1.0), ("quantile", 0.05),
                        ("quantile", 0.5), ("quantile", 0.95)):
            O.ensemble_stat(sample, kind, p)
        dt = time.perf_counter() - t0
        cpu = {"value": round(n * R / dt, 1), "unit": "cells*realizations/s", "cores": 1,
               "kind": "port", "sample": f"the same 7 statistics over {n} cells x {R} realizations "
                                         f"with numpy ({dt:.1f} s, one thread)"}
    line = {
        "metric": "ensemble statistics cells*realizations/s",
        "value": round(N * R / (ms_full / 1e3), 1),
        "unit": "cells*realizations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_full, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "grid": list(dims), "realizations": R,
                   "statistics": "mean, var, cdf(0), ccdf(1), quantile(0.05, 0.5, 0.95)"},
        "roofline": {"bound": "hbm", "kernel": "k_ens_stats (moments only)",
                     "achieved": km["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(km["achieved_GBs"] / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(args.workload, "stats")},
        "kernels": {"stats_with_quantiles": kf, "stats_moments_only": km},
        "moments_only_ms_per_step": round(ms_mom, 3),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_indicator(args, spec):
    """IndicatorProcess SEQSIM: plan + categorical sweep per step."""
    from geostatsprocesses import _lib
    precision = _lib.FP32_REFINE if (args.precision or "fp32r") == "fp32r" else _lib.FP64
    N, path, data, _ = make_inputs(spec)
    R = spec["R"]
    prob = np.asarray(spec["prob"])
    cats = np.random.default_rng(spec["seed"] + 2).integers(0, len(prob), spec["n_data"]).astype(np.int32)
    ctx = _lib.default_context(0)
    dpath = _lib.DeviceBuffer.from_array(ctx, path)
    ddata = _lib.DeviceBuffer.from_array(ctx, data)
    dcats = _lib.DeviceBuffer.from_array(ctx, cats)
    u = _lib.DeviceBuffer(ctx, (R, N))
    for r in range(R):
        row = np.random.default_rng(spec["seed"] + 3 + r).random(N)
        _lib.check(_lib.load().gsp_memcpy_to_device(ctx.handle, u.ptr + r * N * 8, row.ctypes.data, N * 8))
    out = _lib.DeviceBuffer(ctx, (R, N), np.int32)
    prob_arg = problem(spec, dpath.ptr, ddata.ptr, precision, N)

    class Step:
        def __init__(self):
            self.plan = _lib.SeqsimPlan(ctx, prob_arg)
            self.plan.sweep_indicator(u, prob, dcats, out, device_pointers=True)
            ctx.synchronize()

        def close(self):
            self.plan.close()

    step = Step()
    ex = step.plan.export(fields=("nfound", "flags"))
    kbar = float(np.where(ex["flags"] & 3, 0, ex["nfound"]).mean())
    kfl = krige_flops(ex["nfound"], ex["flags"])
    K = step.plan.info.maxneighbors
    depth = step.plan.info.depth
    step.close()
    ms, tim = timed(ctx, Step, args.steps, args.warmup)
    kern = kernel_table(_lib, tim, args.steps, N, R, K, kbar, kflops=kfl)
    line = {
        "metric": "indicator SEQSIM nodes*realizations/s",
        "value": round(N * R / (ms / 1e3), 1),
        "unit": "nodes*realizations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 categories (solves fp32+fp64 refinement)",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid": list(spec["dims"]), "realizations": R,
                   "categories": len(prob), "maxneighbors": spec["kmax"], "radius": spec["radius"],
                   "n_data": spec["n_data"], "depth": depth, "mean_neighbours": round(kbar, 2)},
        "roofline": roofline_of(kern, args.workload, precision == _lib.FP64), "kernels": kern,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def cpu_info():
    """The host CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def bench_lusim(args, spec):
    """LUSIM factorization + sampling (FP64 and mixed) on one GPU; the FP64
    numpy/scipy restatement is the CPU baseline and the per-realization oracle."""
    import geostatsprocesses as gsp
    from geostatsprocesses import _lib
    dims, R = spec["dims"], spec["R"]
    N = int(np.prod(dims))
    pts = np.random.default_rng(spec["seed"]).uniform(0, 1, (spec["n_data"], 2)) * np.asarray(dims)
    vals = np.random.default_rng(spec["seed"] + 2).standard_normal(spec["n_data"])
    grid = gsp.CartesianGrid(*dims)
    proc = gsp.GaussianProcess(gsp.ExponentialCovariance(sill=spec["sill"], range=spec["ranges"][0],
                                                         nugget=spec["nugget"]))
    real, mask = gsp.initialize(proc, grid, gsp.georef({"Z": vals}, [tuple(p) for p in pts]),
                                gsp.NearestInit())
    dinds = np.flatnonzero(mask["Z"]).astype(np.int64)
    z1 = np.ascontiguousarray(real["Z"][dinds])
    nD = len(dinds)
    nS = N - nD
    ctx = _lib.default_context(0)
    ddinds = _lib.DeviceBuffer.from_array(ctx, dinds)
    dz1 = _lib.DeviceBuffer.from_array(ctx, z1)
    w = np.random.default_rng(spec["seed"] + 3).standard_normal((R, nS))
    dw = _lib.DeviceBuffer.from_array(ctx, w)
    dout = _lib.DeviceBuffer(ctx, (R, N))

    def problem(prec):
        p = _lib.GspLusimProblem()
        p.grid.ndim = 2
        for d in range(3):
            p.grid.dims[d] = dims[d] if d < 2 else 1
            p.grid.spacing[d] = 1.0
            p.model.ranges[d] = spec["ranges"][d] if d < 2 else 1.0
        p.model.family = spec["family"]
        p.model.kind = _lib.COVARIANCE
        p.model.sill = spec["sill"]
        p.model.nugget = spec["nugget"]
        p.n_data = nD
        p.dinds = ddinds.ptr
        p.z1 = dz1.ptr
        p.precision = prec
        p.device_pointers = 1
        return p

    def stepper(prec):
        prob = problem(prec)

        def step():
            plan = _lib.LusimPlan(ctx, prob)
            plan.sample(dw, dout, device_pointers=True)
            ctx.synchronize()
            return plan
        return step

    def run(prec):
        step = stepper(prec)
        for _ in range(args.warmup):
            step().close()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step().close()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        # per-kernel times from one more pass with the timing events on
        ctx.enable_timing(True)
        ctx.timings(reset=True)
        step().close()
        tim = ctx.timings(reset=True)
        ctx.enable_timing(False)
        return ms, tim, dout.download()

    ms64, tim64, out64 = run(_lib.LU_FP64)
    msmx, timmx, outmx = run(_lib.LU_MIXED)
    # algorithmic flops (factorization and sampling of lusim.jl:66-107,145-175)
    flops = nS ** 3 / 3 + 2.0 * nS * nS * nD + nD * nD * nS + R * nS * nS
    trail_flops = sum(64.0 * r * (r + 1) for r in range(nS - 64, 0, -64))
    t_ms, t_n = tim64[_lib.K_LU_TRAIL]
    trail_tf = trail_flops / (t_ms * 1e-3) / 1e12 if t_ms > 0 else None

    def kern(tim):
        out = {}
        for slot, name in ((_lib.K_LU_FACTOR, "factor (whole call)"), (_lib.K_LU_TRAIL, "k_trail"),
                           (_lib.K_LU_SAMPLE, "sample (whole call)")):
            ms, n = tim[slot]
            if n:
                out[name] = {"ms": round(ms, 4), "launches": n, "avg_ms": round(ms / n, 4)}
        return out

    cpu = accuracy = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        from threadpoolctl import threadpool_limits
        model, ncores, share = cpu_share()
        cents = O.grid_centroids(dims)
        res = {}
        for threads in (1, share):
            with threadpool_limits(limits=threads):
                t0 = time.perf_counter()
                d2, L22, sinds = O.lusim_preprocess(cents, dinds, z1, spec["family"], spec["sill"],
                                                    spec["nugget"], spec["ranges"])
                ref = O.lusim_sample(d2, L22, dinds, sinds, z1, 0.0, w)
                res[threads] = time.perf_counter() - t0
        cpu = {"value": N * R / res[share], "unit": "nodes*realizations/s", "cores": share,
               "kind": "port",
               "sample": (f"the whole workload: numpy/scipy (LAPACK potrf/trsm, BLAS gemm) "
                          f"restatement of lusim.jl, FP64, {res[share]:.2f} s at {share} threads"),
               "one_thread": {"value": N * R / res[1], "cores": 1,
                              "sample": f"the same at 1 thread, {res[1]:.2f} s"},
               "cpu_model": model, "cores_visible": ncores}

        def per_real(out):
            e = np.max(np.abs(out - ref), axis=1) / np.max(np.abs(ref), axis=1)
            return {"max": float(e.max()), "median": float(np.median(e))}
        accuracy = {"fp64_vs_oracle_per_realization": per_real(out64),
                    "mixed_vs_oracle_per_realization": per_real(outmx),
                    "data_exact": bool(np.array_equal(out64[:, dinds], np.broadcast_to(z1, (R, nD))))}
    value = N * R / (ms64 / 1e3)
    line = {
        "metric": "LUSIM nodes*realizations/s (FP64 Cholesky)",
        "value": round(value, 1), "unit": "nodes*realizations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms64, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid": list(dims), "realizations": R,
                   "n_data": nD, "n_sim": nS, "covariance": "exponential", "range": spec["ranges"][0],
                   "nugget": spec["nugget"]},
        "roofline": {"bound": "mfma", "kernel": "k_trail", "achieved": round(trail_tf, 2) if trail_tf else None,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(trail_tf / FP64_PEAK_TFLOPS, 4) if trail_tf else None,
                     "traffic": None},
        "step_tflops": round(flops / (ms64 * 1e-3) / 1e12, 2),
        "step_frac_of_fp64_peak": round(flops / (ms64 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
        "algorithmic_gflop_per_step": round(flops / 1e9, 2),
        "kernels": kern(tim64),
        "mixed": {"value": round(N * R / (msmx / 1e3), 1), "ms_per_step": round(msmx, 3),
                  "kernels": kern(timmx)},
        "accuracy": accuracy,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="seqsim-learned-256x256x128", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default=None, choices=["fp32r", "fp64"],
                    help="kriging solves (default: the workload's: fp64 for the 100x100 SEQSIM and the "
                         "posterior mean, fp32r otherwise)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the learned-vs-exact accuracy pass")
    ap.add_argument("--epochs", type=int, default=None,
                    help="surrogate training epochs (default: the workload's; fewer for profiler runs)")
    ap.add_argument("--data-slots", type=int, default=None, choices=[0, 4],
                    help="surrogate data slots (learned batch workload; default 0: nodes with "
                         "conditioning data among their neighbours are solved exactly)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (traffic-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU: this build has no multi-GPU path "
                 "(realizations would shard across GPUs with no exchange); run with --gpus 1")

    from geostatsprocesses import _lib

    if args.copy_calibration:
        ctx = _lib.default_context(0)
        nbytes = args.copy_calibration << 20
        src = _lib.DeviceBuffer(ctx, (nbytes // 8,))
        dst = _lib.DeviceBuffer(ctx, (nbytes // 8,))
        for _ in range(3):
            _lib.check(_lib.load().gsp_selftest_copy(ctx.handle, dst.ptr, src.ptr, nbytes))
        print(json.dumps({"copy_bytes": nbytes, "launches": 3}), flush=True)
        return

    spec = WORKLOADS[args.workload]
    if spec.get("method") == "fftsim":
        return bench_fftsim(args, spec)
    if spec.get("method") == "stats":
        return bench_stats(args, spec)
    if spec.get("method") == "indicator":
        return bench_indicator(args, spec)
    if spec.get("method") == "lusim":
        return bench_lusim(args, spec)
    args.precision = args.precision or spec.get("precision", "fp32r")
    precision = _lib.FP32_REFINE if args.precision == "fp32r" else _lib.FP64
    R = spec["R"]
    mean_only = spec.get("mode") == "data_only"
    if "batch" in spec:
        reqs = batch_specs(spec)
    else:
        _, path, data, vals = make_inputs(spec)
        reqs = [dict(spec, data=data, vals=vals, path=path)]
    sizes = [int(np.prod(r["dims"])) for r in reqs]
    N = int(sum(sizes))
    ctx = _lib.default_context(0)
    # inputs resident in HBM before timing: paths, data, R x N standard normals
    # (one buffer each; request q's arrays at its offsets)
    path_all = np.concatenate([r["path"] for r in reqs])
    coords_all = np.concatenate([r["data"].ravel() for r in reqs] + [np.zeros(1)])
    vals_all = np.concatenate([r["vals"] for r in reqs] + [np.zeros(1)])
    dpath = _lib.DeviceBuffer.from_array(ctx, path_all)
    ddata = _lib.DeviceBuffer.from_array(ctx, coords_all)
    dvals = _lib.DeviceBuffer.from_array(ctx, vals_all)
    del path_all
    if mean_only:  # the posterior mean: one realization of zero normals
        z = _lib.DeviceBuffer.from_array(ctx, np.zeros((R, N)))
    else:
        z = upload_normals(ctx, _lib, R, N, spec["seed"])
    out = _lib.DeviceBuffer(ctx, (R, N))

    def problems(prec):
        probs, noff, coff = [], 0, 0
        for r, n in zip(reqs, sizes):
            probs.append(problem(r, dpath.ptr + 8 * noff, ddata.ptr + 8 * coff if r["n_data"] else None,
                                 prec, n))
            noff += n
            coff += r["data"].size
        return probs[0] if len(probs) == 1 else probs

    plan_arg = problems(precision)
    S, training, tol = None, None, None
    if "learned" in spec:
        if args.epochs is not None:
            spec = dict(spec, learned=dict(spec["learned"], epochs=args.epochs))
        if args.data_slots is not None:
            spec = dict(spec, learned=dict(spec["learned"], data_slots=args.data_slots))
        if "batch" in spec:
            S, training = train_batch_surrogates(ctx, _lib, spec, precision)
        else:
            S, training = train_surrogate(ctx, _lib, spec, precision)
        tol = spec["learned"]["tolerance"]

    def step():
        plan = _lib.SeqsimPlan(ctx, plan_arg, surrogate=S, tolerance=tol if tol is not None else 0.05)
        plan.sweep(z, dvals, out, device_pointers=True)
        ctx.synchronize()
        return plan

    # realization-invariant statistics of the plan (for the byte counts)
    plan = step()
    ex = plan.export(fields=("nfound", "flags"))
    info = plan.info
    plan.close()
    keff = np.where(ex["flags"] & 3, 0, ex["nfound"]).astype(np.float64)
    kbar = float(keff.mean())
    kfl = krige_flops(ex["nfound"], ex["flags"])
    kfl_exact = krige_flops(ex["nfound"], ex["flags"] & 7)   # every system solved exactly
    del ex

    ms_per_step, tim = timed(ctx, step, args.steps, args.warmup)
    value = N * R / (ms_per_step / 1e3)
    K = info.maxneighbors                                        # row stride of nbr / lambda
    slots = (S[0] if isinstance(S, list) else S).info["n_slots"] if S is not None else 0
    kern = kernel_table(_lib, tim, args.steps, N, R, K, kbar, n_mlp=info.n_learned, kflops=kfl,
                        mlp_fl=mlp_flops(slots) if slots else 0)
    roofline = roofline_of(kern, args.workload, precision == _lib.FP64)

    # repeated sweeps of one plan (a serving loop): captured HIP graph vs one
    # launch per level
    sweep_only = None
    if S is None and not mean_only:
        plan = _lib.SeqsimPlan(ctx, plan_arg)
        res = {}
        for mode in ("0", "1"):
            os.environ["GSP_SWEEP_GRAPH"] = mode
            for _ in range(3):
                plan.sweep(z, dvals, out, device_pointers=True)
            ctx.synchronize()
            reps = 20
            t0 = time.perf_counter()
            for _ in range(reps):
                plan.sweep(z, dvals, out, device_pointers=True)
            ctx.synchronize()
            res[mode] = (time.perf_counter() - t0) * 1e3 / reps
        os.environ.pop("GSP_SWEEP_GRAPH", None)
        plan.close()
        sweep_only = {"eager_ms": round(res["0"], 4), "graph_ms": round(res["1"], 4),
                      "graph_value": round(N * R / (res["1"] / 1e3), 1), "levels": info.depth}

    exact = exact64 = accuracy = None
    if S is not None:
        def exact_step(arg):
            def step_exact():
                plan = _lib.SeqsimPlan(ctx, arg)
                plan.sweep(z, dvals, out, device_pointers=True)
                ctx.synchronize()
                return plan
            return step_exact
        ems, etim = timed(ctx, exact_step(plan_arg), args.steps, args.warmup)
        exact = {"value": round(N * R / (ems / 1e3), 1), "ms_per_step": round(ems, 3),
                 "precision": "fp32r",
                 "kernels": kernel_table(_lib, etim, args.steps, N, R, K, kbar, kflops=kfl_exact)}
        ems64, etim64 = timed(ctx, exact_step(problems(_lib.FP64)), args.steps, args.warmup)
        exact64 = {"value": round(N * R / (ems64 / 1e3), 1), "ms_per_step": round(ems64, 3),
                   "precision": "fp64",
                   "kernels": kernel_table(_lib, etim64, args.steps, N, R, K, kbar,
                                           kflops=kfl_exact)}
        if not args.no_accuracy and "batch" in spec:
            accuracy = batch_learned_accuracy(ctx, _lib, reqs, problems(precision),
                                              problems(_lib.FP64), S, tol, R,
                                              spec["learned"]["accuracy_requests"])
            accuracy["tolerance"] = tol
        elif not args.no_accuracy:
            masked = problems(_lib.FP32_REFINE)
            masked.search_mode = _lib.MASKED_KNN
            accuracy = learned_accuracy(ctx, _lib, spec, plan_arg, problems(_lib.FP64), masked, S,
                                        tol, z, dvals, N, R)
            accuracy["tolerance"] = tol

    cpu = None
    if not args.no_cpu:
        if "batch" in spec:
            cpu = cpu_baseline(reqs, R_batch=R, node_budget=1_500_000)
        elif mean_only:
            cpu = cpu_baseline(reqs, node_limit=spec["cpu_nodes"])
            cpu["unit"] = "nodes/s"
            cpu["one_thread"]["unit"] = "nodes/s"
        else:
            cpu = cpu_baseline(reqs)

    if "batch" in spec:
        config = {"workload": args.workload, "requests": len(reqs), "nodes": N,
                  "realizations_per_request": R, "side_range": list(spec["side"]),
                  "n_data": int(sum(r["n_data"] for r in reqs)), "maxneighbors": [8, 16, 32]}
    else:
        config = {"workload": args.workload, "grid": list(spec["dims"]), "realizations": R,
                  "maxneighbors": spec["kmax"], "radius": spec["radius"], "n_data": spec["n_data"]}
    config |= {"search": "k nearest data" if mean_only else "k nearest then mask",
               "precision": args.precision, "depth": info.depth, "mean_neighbours": round(kbar, 2)}
    if S is not None:
        metric = "SEQSIM nodes*realizations/s (learned kriging weights, exact fallback)"
    elif mean_only:
        metric = "kriging posterior mean nodes/s"
    else:
        metric = "SEQSIM nodes*realizations/s (exact kriging)"
    line = {
        "metric": metric,
        "value": round(value, 1),
        "unit": "nodes/s" if mean_only else "nodes*realizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (solves fp32+fp64 refinement)" if precision == _lib.FP32_REFINE else "f64",
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    if sweep_only is not None:
        line["sweep_only"] = sweep_only
    if S is not None:
        line["dtype"] = "bf16 mlp (fp32 accumulate), f64 sweep; fallback solves fp32+fp64 refinement"
        line["exact"] = exact
        line["exact_fp64"] = exact64
        line["speedup_vs_exact"] = round(value / exact["value"], 3)
        line["speedup_vs_exact_fp64"] = round(value / exact64["value"], 3)
        line["learned_nodes"] = int(info.n_learned)
        line["learned_fraction"] = round(info.n_learned / N, 4)
        line["training"] = training
        line["accuracy"] = accuracy
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
