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
}


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


def cpu_baseline(spec, path, data, vals, precision):
    """The oracle's reference-cost restatement (KD-free cell-hash search +
    Cholesky per node per realization), OpenMP over realizations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    dims = spec["dims"]
    case = O.SeqsimCase(dims=dims, family=spec["family"], sill=spec["sill"],
                        ranges=spec["ranges"][:len(dims)], nugget=spec["nugget"], kmax=spec["kmax"],
                        radius=spec["radius"], data_coords=data, data_vals=vals, path=path,
                        precision=precision)
    N = case.n_nodes
    R = threads
    z = np.random.default_rng(7).standard_normal((R, N))
    # every path position costs differently (early nodes have few simulated
    # neighbours), so the sample is whole realizations: one per thread
    limit = N
    t0 = time.perf_counter()
    _, n = O.seqsim_reference_cost(case, z, threads=threads, n_nodes_limit=limit)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "nodes*realizations/s", "cores": threads, "kind": "port",
            "sample": f"{R} full realizations ({limit} path nodes each) of the same workload "
                      f"({dt:.1f} s; search + kriging solve per node per realization, "
                      f"OpenMP over realizations, {threads} threads)"}


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
    N, path, data, vals = make_inputs(spec)
    R = spec["R"]
    ctx = _lib.default_context(0)
    # inputs resident in HBM before timing: path, data, R x N standard normals
    z_host = np.random.default_rng(spec["seed"] + 3).standard_normal((R, N))
    dpath = _lib.DeviceBuffer.from_array(ctx, path)
    ddata = _lib.DeviceBuffer.from_array(ctx, data)
    dvals = _lib.DeviceBuffer.from_array(ctx, vals)
    z = _lib.DeviceBuffer.from_array(ctx, z_host)
    del z_host
    out = _lib.DeviceBuffer(ctx, (R, N))

    prob = problem(spec, dpath.ptr, ddata.ptr, precision, N)

    def step():
        plan = _lib.SeqsimPlan(ctx, prob)
        plan.sweep(z, dvals, out, device_pointers=True)
        ctx.synchronize()
        return plan

    for _ in range(args.warmup):
        step().close()
    # realization-invariant statistics of the plan (for the byte counts)
    plan = step()
    ex = plan.export()
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
    search_bytes = N * (8 + spec["kmax"] * 4 + 4)                # path, nbr, nfound
    plan_bytes = N * (spec["kmax"] * (4 + 8) + 4 + 8 + 1 + 4)    # nbr + nfound in, lambda, sigma, flags, degrees
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
        cpu = cpu_baseline(spec, path, data, vals, precision)

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
        "config": {"workload": args.workload, "grid": list(spec["dims"]), "realizations": R,
                   "maxneighbors": spec["kmax"], "radius": spec["radius"], "n_data": spec["n_data"],
                   "search": "k nearest then mask", "precision": args.precision,
                   "depth": info.depth, "mean_neighbours": round(kbar, 2)},
        "roofline": roofline,
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
