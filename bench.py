"""Benchmark: converged N-1 power flows per second on the 2000-bus synthetic grid.

One step = one full N-1 sweep: every non-bridge single-branch outage of
case2000_synth (2128 contingencies), each a Ybus variant on the base pattern,
warm-started from the base-case V, solved to ||F||inf < 1e-8 pu with at most
10 NR iterations, in the FP32-factor / FP64-mismatch form (the FP64 form is
timed beside it). Inputs are resident in HBM before the timed region; the
timed region is K back-to-back sweeps on the engine's stream.

Prints one JSON line (see the driver contract): value = converged solves / s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(info, iters, w):
    """Bytes per launch: sum over solves of it * [w (nnzJ + 3 nnzLU) + 16*4n + 8*2N]."""
    n, N = info["n_bus"], info["n_jac"]
    per_it = w * (info["nnz_jac"] + 3 * info["nnz_lu"]) + 64 * n + 16 * N
    return float(np.sum(iters)) * per_it


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", default="case2000_synth")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads for the CPU baseline (0: min(16, nproc))")
    ap.add_argument("--instances-per-wg", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the N-1 sweep would shard by "
              "contingency across GPUs, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    from lightsim2grid_amd import abi, gridgen

    g = gridgen.make(args.grid)
    d = g.solver_inputs()
    nb = np.nonzero(~gridgen.bridges(g))[0]
    conts = [[int(k)] for k in nb]
    dptr, dpos, dval = g.ybus_deltas(conts)
    ninst = len(conts)

    ctx = abi.Context(0)
    gr = abi.Grid(ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    info = gr.info
    Yx = gr.ybus_values(d["Ybus"])
    # base case (the reference's "n powerflow", BaseBatchSolverSynch.hpp:296-313), FP64
    base = gr.solve(Yx, d["Sbus"], d["V0"], 10, 1e-8, precision=abi.FP64)
    assert base["status"][0] == 0, "base case did not converge"
    Vb = base["V"][0]

    # device-resident inputs
    dY = ctx.to_device(Yx)
    dS = ctx.to_device(np.ascontiguousarray(d["Sbus"], np.complex128))
    dV0 = ctx.to_device(np.ascontiguousarray(Vb, np.complex128))
    dV = ctx.empty((ninst, g.n), np.complex128)
    ddp, ddq, ddv = ctx.to_device(dptr), ctx.to_device(dpos), ctx.to_device(dval)
    dIt = ctx.empty(ninst, np.int32)
    dSt = ctx.empty(ninst, np.int32)

    def sweep(prec):
        return gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, ninst, dIt, dSt, 10, 1e-8, precision=prec,
                               deltas=(ddp, ddq, ddv), instances_per_wg=args.instances_per_wg)

    results = {}
    tuned = {}
    for prec, name in ((abi.FP64, "fp64"), (abi.FP32, "fp32")):
        if not args.instances_per_wg:
            # the project's launch autotuner (instances per workgroup x block size), untimed
            tuned[name] = gr.autotune_device(dY, 0, dS, 0, dV0, 0, dV, ninst, dIt, dSt, 10, 1e-8, precision=prec,
                                             deltas=(ddp, ddq, ddv))
        for _ in range(args.warmup):
            sweep(prec)
        ksum = 0.0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = sweep(prec)
            ksum += st["kernel_s"]
        t1 = time.perf_counter()
        iters = dIt.to_host()
        status = dSt.to_host()
        conv = int(np.sum(status == 0))
        wall = t1 - t0
        w = 4 if prec == abi.FP32 else 8
        bytes_launch = algorithmic_bytes(info, iters, w)
        kavg = ksum / args.steps
        results[name] = dict(
            converged=conv, ninst=ninst, wall_s=wall, ms_per_step=1e3 * wall / args.steps,
            solves_per_s=conv * args.steps / wall, kernel_ms=1e3 * kavg, iters_mean=float(np.mean(iters)),
            achieved_gbs=bytes_launch / kavg / 1e9, bytes_per_launch=bytes_launch, lds_slots=st["lds_slots"],
            instances_per_wg=st["instances_per_wg"], status_hist=np.bincount(status, minlength=8).tolist(),
            autotune=tuned.get(name))

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        # bounded sample: the full sweep on nthreads threads (each thread solves the base case
        # first, like an independent ContingencyAnalysis object), plus a 1-thread sample
        deltas = (dptr, dpos, dval)
        _, _, _, cconv, _, wall_mt = oracle.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"],
                                                        d["pv"], d["pq"], Vb, deltas, 10, 1e-8, False, False, nthreads)
        k1 = min(256, ninst)
        sub = (dptr[:k1 + 1], dpos[:dptr[k1]], dval[:dptr[k1]])
        _, _, _, cconv1, _, wall_1 = oracle.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"],
                                                        d["pv"], d["pq"], Vb, sub, 10, 1e-8, False, False, 1)
        cpu = dict(value=float(np.sum(cconv)) / wall_mt, unit="converged power flows/s", cores=nthreads,
                   kind="port",
                   sample=f"oracle (C++ FP64 restatement of BaseNRAlgo<KLU> + ContingencyAnalysis loop), full "
                          f"{ninst}-outage sweep on {nthreads} threads, {wall_mt:.2f} s",
                   one_thread=dict(value=float(np.sum(cconv1)) / wall_1, sample=f"{k1} outages, {wall_1:.2f} s"))

    r32, r64 = results["fp32"], results["fp64"]
    out = {
        "metric": "power flows/sec at 1e-8 pu mismatch (N-1, 2000 bus)",
        "value": r32["solves_per_s"],
        "unit": "converged power flows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor / f64 mismatch",
        "data": "synthetic",
        "config": {"workload": f"{args.grid} N-1 sweep: {ninst} non-bridge single-branch outages, warm start "
                               f"from the base-case V, tol 1e-8 pu, max 10 iterations",
                   "n_bus": g.n, "n_branch": g.nbr, "n_jac": info["n_jac"], "nnz_jac": info["nnz_jac"],
                   "nnz_lu": info["nnz_lu"], "instances_per_step": ninst},
        "roofline": {"bound": "hbm", "achieved": r32["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r32["achieved_gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_nr<float,G>", "kernel_ms": r32["kernel_ms"]},
        "cpu_baseline": cpu,
        "fp64_form": {"value": r64["solves_per_s"], "ms_per_step": r64["ms_per_step"], "kernel_ms": r64["kernel_ms"],
                      "achieved_gbs": r64["achieved_gbs"], "iters_mean": r64["iters_mean"],
                      "converged": r64["converged"]},
        "fp32_over_fp64": r32["solves_per_s"] / r64["solves_per_s"],
        "details": {"fp32": r32, "fp64": r64},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
