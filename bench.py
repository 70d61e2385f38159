#!/usr/bin/env python3
"""SEQSIM throughput on one MI355X: nodes x realizations per second.

Workload (default, "seqsim-3d-128x128x64"): a 128 x 128 x 64 unit grid
(1,048,576 nodes), anisotropic exponential variogram sill 1, ranges
(40, 40, 8), nugget 0.05, 2000 conditioning points uniform in the domain with
N(0,1) values, maxneighbors 24 with the :range neighbourhood (radius 40), a
random path, 64 realizations, local solves in FP32 with one FP64-residual
refinement step, realizations in FP64. One step is one full
`rand(process, grid, 64; method=SEQSIM(...))`: plan (neighbour search +
kriging solves) and the 64-realization sweep, with the path, the data and the
64 x N standard normals already resident in HBM.

Workload "seqsim-batch-512": 512 independent 2D requests in one plan and one
sweep: square grids with side log-uniform in [32, 512] (about 23M nodes in
total), spherical/exponential/gaussian variograms (sill 1) with range uniform
in [5, 60] cells and nugget uniform in [0, 0.1], 0-500 uniform data each,
maxneighbors from {8, 16, 32}, :range neighbourhood, random paths, 8
realizations each, FP32 solves with refinement. One step is one batched plan
plus the batched sweep.

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
                              n_data=50, kmax=12, radius=30.0, R=10, seed=1001),
    "seqsim-batch-512": dict(batch=512, side=(32, 512), R=8, seed=1005),
}


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
    path = rng.permutation(N).astype(np.int64)
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
    p.search_mode = _lib.KNN_THEN_MASK
    p.precision = precision
    p.device_pointers = 1
    return p


# kernels of each bench name in the committed traffic profiles
PROFILE_KERNELS = {"search": ["k_search"], "krige": ["k_krige16<1>", "k_krige<24, 1>"],
                   "sweep": ["k_sweep_level"]}


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` (read + written, counter-calibrated) from
    profiles/<workload>_traffic.md, which rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    runs of this script produced; None when no profile covers it."""
    path = os.path.join(ROOT, "profiles", workload.replace("seqsim-", "seqsim_").replace("-", "_")
                        + "_traffic.md")
    if not os.path.exists(path):
        return None
    total, found = 0.0, False
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] in PROFILE_KERNELS.get(kernel, []):
            try:
                dispatches = int(cells[1])
                per = (float(cells[2]) + float(cells[3])) * 1e6
            except ValueError:
                continue
            # the profiled run made two plans and two sweeps (export + timed step)
            total += per * dispatches / 2
            found = True
    if not found:
        return None
    launches = {"search": 1, "krige": 1, "sweep": 1}[kernel]
    return int(total / launches)


def cpu_baseline(reqs, precision, R_batch=None, node_budget=None):
    """The oracle's reference-cost restatement (KD-free cell-hash search +
    Cholesky per node per realization), OpenMP over realizations.

    Single workloads: one full realization per thread. Batches: the first
    requests up to `node_budget` nodes, with the workload's R realizations
    each (every path position costs differently, so samples are whole
    realizations)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    if R_batch is not None:
        threads = min(threads, R_batch)
    n_total, dt_total, used, nodes = 0, 0.0, 0, 0
    for spec in reqs:
        dims = spec["dims"]
        case = O.SeqsimCase(dims=dims, family=spec["family"], sill=spec["sill"],
                            ranges=spec["ranges"][:len(dims)], nugget=spec["nugget"],
                            kmax=spec["kmax"], radius=spec["radius"], data_coords=spec["data"],
                            data_vals=spec["vals"], path=spec["path"], precision=precision)
        N = case.n_nodes
        R = R_batch if R_batch is not None else threads
        z = np.random.default_rng(7).standard_normal((R, N))
        t0 = time.perf_counter()
        _, n = O.seqsim_reference_cost(case, z, threads=threads, n_nodes_limit=N)
        dt_total += time.perf_counter() - t0
        n_total += n
        used += 1
        nodes += N
        if node_budget is None or nodes >= node_budget:
            break
    if R_batch is None:
        sample = (f"{threads} full realizations ({nodes} path nodes each) of the same workload "
                  f"({dt_total:.1f} s; search + kriging solve per node per realization, "
                  f"OpenMP over realizations, {threads} threads)")
    else:
        sample = (f"the first {used} requests of the batch ({nodes} nodes), {R_batch} full "
                  f"realizations each ({dt_total:.1f} s; search + kriging solve per node per "
                  f"realization, OpenMP over realizations, {threads} threads)")
    return {"value": n_total / dt_total, "unit": "nodes*realizations/s", "cores": threads,
            "kind": "port", "sample": sample}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="seqsim-3d-128x128x64", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32r", choices=["fp32r", "fp64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
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
    precision = _lib.FP32_REFINE if args.precision == "fp32r" else _lib.FP64
    R = spec["R"]
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
    z = _lib.DeviceBuffer(ctx, (R, N))
    for r0 in range(R):  # row by row keeps host memory small
        z_row = np.random.default_rng(spec["seed"] + 3 + r0).standard_normal(N)
        _lib.check(_lib.load().gsp_memcpy_to_device(ctx.handle, z.ptr + r0 * N * 8,
                                                    z_row.ctypes.data, N * 8))
    out = _lib.DeviceBuffer(ctx, (R, N))
    probs, noff, coff = [], 0, 0
    for r, n in zip(reqs, sizes):
        probs.append(problem(r, dpath.ptr + 8 * noff, ddata.ptr + 8 * coff if r["n_data"] else None,
                             precision, n))
        noff += n
        coff += r["data"].size
    plan_arg = probs[0] if len(probs) == 1 else probs

    def step():
        plan = _lib.SeqsimPlan(ctx, plan_arg)
        plan.sweep(z, dvals, out, device_pointers=True)
        ctx.synchronize()
        return plan

    for _ in range(args.warmup):
        step().close()
    # realization-invariant statistics of the plan (for the byte counts)
    plan = step()
    ex = plan.export(fields=("nfound", "flags"))
    info = plan.info
    plan.close()
    keff = np.where(ex["flags"] & 3, 0, ex["nfound"]).astype(np.float64)
    kbar = float(keff.mean())

    ctx.synchronize()
    ctx.enable_timing(True)
    ctx.timings(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step().close()
    ctx.synchronize()
    t1 = time.perf_counter()
    tim = ctx.timings(reset=True)
    ctx.enable_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = N * R / (ms_per_step / 1e3)

    # algorithmic bytes (realizations resident; neighbour gathers are re-reads)
    sweep_bytes = N * R * (8 + 8) + N * (kbar * (4 + 8) + 8)     # z read, x write, plan read
    K = info.maxneighbors                                        # row stride of nbr / lambda
    search_bytes = N * (8 + K * 4 + 4)                           # path, nbr, nfound
    plan_bytes = N * (K * (4 + 8) + 4 + 8 + 1 + 4)               # nbr + nfound in, lambda, sigma, flags, degrees
    kern = {}
    for slot, name, nbytes in ((_lib.K_SEARCH, "search", search_bytes), (_lib.K_PLAN, "krige", plan_bytes),
                               (_lib.K_SWEEP, "sweep", sweep_bytes),
                               (_lib.K_LEVELS, "levels", None), (_lib.K_TRANSPOSE, "transpose", 2 * 8 * N * R)):
        ms, launches = tim[slot]
        if launches == 0:
            continue
        avg = ms / launches
        kern[name] = {"avg_ms": round(avg, 4), "launches_per_step": launches // args.steps}
        if nbytes is not None:
            per_launch = nbytes / (launches / args.steps)
            kern[name]["algorithmic_bytes_per_launch"] = int(per_launch)
            kern[name]["achieved_GBs"] = round(per_launch / (avg * 1e-3) / 1e9, 1)
    dominant = max(("search", "krige", "sweep"), key=lambda k: kern.get(k, {}).get("avg_ms", 0) *
                   kern.get(k, {}).get("launches_per_step", 0))
    d = kern[dominant]
    achieved = d["achieved_GBs"]
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(args.workload, dominant)}

    cpu = None
    if not args.no_cpu:
        if "batch" in spec:
            cpu = cpu_baseline(reqs, precision, R_batch=R, node_budget=1_500_000)
        else:
            cpu = cpu_baseline(reqs, precision)

    line = {
        "metric": "SEQSIM nodes*realizations/s (exact kriging)",
        "value": round(value, 1),
        "unit": "nodes*realizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (solves fp32+fp64 refinement)" if precision == _lib.FP32_REFINE else "f64",
        "data": "synthetic",
        "config": ({"workload": args.workload, "grid": list(spec["dims"]), "realizations": R,
                    "maxneighbors": spec["kmax"], "radius": spec["radius"], "n_data": spec["n_data"]}
                   if "batch" not in spec else
                   {"workload": args.workload, "requests": len(reqs), "nodes": N,
                    "realizations_per_request": R, "side_range": list(spec["side"]),
                    "n_data": int(sum(r["n_data"] for r in reqs)), "maxneighbors": [8, 16, 32]})
        | {"search": "k nearest then mask", "precision": args.precision,
           "depth": info.depth, "mean_neighbours": round(kbar, 2)},
        "roofline": roofline,
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
