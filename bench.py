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
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X vector FLOP peaks by factor precision (bytes per value): FP32 157.3, FP64 78.6 TFLOP/s
VEC_PEAK_TFLOPS = {4: 157.3, 8: 78.6}
# LDS: 128 bytes per clock per compute unit at 2.4 GHz
LDS_PEAK_GBS_PER_CU = 128 * 2.4
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_nr_case2000_n1_hbm_traffic.md")
PUBLISHED_14BUS_MS = 0.0112  # NR (KLU) time in algo, l2rpn_case14_sandbox, 1 core (docs/benchmarks.rst:212)


def measured_traffic(form: str):
    """HBM bytes per sweep launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, each scaled by
    the same counter's reading for a copy kernel of known size), from the committed profile of
    this workload; None when the profile is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {form}:"):
                    return float(line.split(":")[1])
    except OSError:
        return None
    return None


def cpu_info() -> dict:
    """The host CPU as this process sees it: logical CPUs, the CPUs it may run on (affinity and
    the cgroup CPU quota), physical cores of the machine and the model name."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = os.cpu_count()
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, p = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(p)
    except (OSError, ValueError):
        pass
    info["cgroup_cpus"] = quota
    usable = info["affinity"] or 1
    if quota:
        usable = min(usable, max(1, int(quota)))
    info["usable"] = usable
    cores, model = set(), None
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    continue
                k, v = [x.strip() for x in line.split(":", 1)]
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    info["physical_cores"] = len(cores) or None
    info["model"] = model
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def algorithmic_bytes(info, iters, w):
    """Bytes of solves with `iters` iterations: sum of it * [w (nnzJ + 3 nnzLU) + 16*4n + 8*2N]."""
    n, N = info["n_bus"], info["n_jac"]
    per_it = w * (info["nnz_jac"] + 3 * info["nnz_lu"]) + 64 * n + 16 * N
    return float(np.sum(iters)) * per_it


def algorithmic_flops(info, iters):
    """FLOPs of solves with `iters` iterations: sum of it * [refactor flops (the plan's count:
    multiply-adds as 2, divisions 1) + 2 nnzLU (the two triangular solves) + 16 nnzY (two complex
    SpMVs)]."""
    per_it = info["flops"] + 2 * info["nnz_lu"] + 16 * info["nnz_ybus"]
    return float(np.sum(iters)) * per_it


def flop_rate(info, iters, seconds, w):
    """Achieved FLOP rate of solves against MI355X's vector peak of their factor precision."""
    fl = algorithmic_flops(info, iters)
    peak = VEC_PEAK_TFLOPS[w]
    return {"flops": fl, "achieved_tflops": fl / seconds / 1e12, "peak_tflops": peak,
            "frac": fl / seconds / 1e12 / peak, "peak": "fp%d vector" % (8 * w)}


def parity(V, iters, status, Vr, itr, convr):
    """Converged flags, iteration deltas and max |V - V_ref| of GPU rows against the oracle."""
    conv = status == 0
    both = conv & convr
    dit = (iters[both].astype(int) - itr[both].astype(int)) if both.any() else np.zeros(0, int)
    hist = {str(k): int(v) for k, v in zip(*np.unique(dit, return_counts=True))}
    return {"flags_equal": bool(np.array_equal(conv, convr)), "n_flag_diff": int(np.sum(conv != convr)),
            "max_dV": float(np.max(np.abs(V[both] - Vr[both]))) if both.any() else None,
            "iter_delta_hist": hist}


def single_solve_latency(reps: int = 300) -> dict:
    """Config 0: one 14-bus FP64 flat-start solve (tol 1e-8, max 10 iterations) through the solver
    class (host arrays in, V out: the drop-in call) and through the device entry point."""
    from lightsim2grid_amd import abi, gridgen, solver
    g = gridgen.make("case14_synth")
    d = g.solver_inputs()
    s = solver.AMDGPUSolver()
    args = (d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    assert s.compute_pf(*args)
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        s.compute_pf(*args)
        t.append(time.perf_counter() - t0)
    ctx = abi.Context(0)
    gr = abi.Grid(ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    dY, dS, dV0 = (ctx.to_device(gr.ybus_values(d["Ybus"])), ctx.to_device(np.ascontiguousarray(d["Sbus"])),
                   ctx.to_device(d["V0"]))
    dV, dIt, dSt = ctx.empty((1, g.n), np.complex128), ctx.empty(1, np.int32), ctx.empty(1, np.int32)
    td, tk = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        st = gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, 1, dIt, dSt, 10, 1e-8)
        td.append(time.perf_counter() - t0)
        tk.append(st["kernel_s"])
    out = {"workload": "case14_synth, FP64, flat start, tol 1e-8, max 10 iterations, one solve per call",
           "iters": int(s.get_nb_iter()),
           "solver_class_ms_p50": 1e3 * float(np.median(t)), "solver_class_ms_p99": 1e3 * float(np.percentile(t, 99)),
           "device_call_ms_p50": 1e3 * float(np.median(td)), "kernel_ms_p50": 1e3 * float(np.median(tk)),
           "published_cpu_ms": PUBLISHED_14BUS_MS,
           "published_source": "NR (KLU) time in algo, l2rpn_case14_sandbox, i7-13700H 1 core (docs/benchmarks.rst:212)"}
    gr.close()
    ctx.close()
    return out


def serving_run(seconds: float, rate: float = 120.0, seed: int = 7) -> dict:
    """Serving mix: Poisson arrivals over resident 14/118/2000/10000-bus grids (weights
    0.4/0.3/0.2/0.1); single power flow 50 %, contingency batch of 1-256 outages 35 %,
    time-series chunk of 1-96 steps 15 %; both precision forms; p50 / p99 per grid size."""
    import threading

    from lightsim2grid_amd import gridgen
    from lightsim2grid_amd.lightsim2grid_amd_cpp import ServeLoop

    names = ["case14_synth", "case118_synth", "case2000_synth", "case10000_synth"]
    weights = np.array([0.4, 0.3, 0.2, 0.1])
    grids = [gridgen.make(nm) for nm in names]
    inputs = [g.solver_inputs() for g in grids]
    nonbridge = [np.nonzero(~gridgen.bridges(g))[0] for g in grids]
    out = {"workload": "Poisson arrivals %.0f req/s for %.1f s; grids 14/118/2000/10000 buses (0.4/0.3/0.2/0.1); "
                       "single PF 50%%, N-1 batch 1-256 35%%, time-series chunk 1-96 15%%" % (rate, seconds)}
    base_v = []
    tmp = ServeLoop(0, 16, 0)
    for d in inputs:
        t0 = tmp.add_grid(d["Ybus"], d["Sbus"], d["V0"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
        base_v.append(tmp.result(tmp.submit_pf(t0, None), 120.0)["V"][0])
    tmp.close()
    for prec, pname in ((1, "fp32"), (0, "fp64")):
        sl = ServeLoop(0, 4096, prec)
        gids = [sl.add_grid(d["Ybus"], d["Sbus"], vb, d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
                for d, vb in zip(inputs, base_v)]
        rng = np.random.default_rng(seed)
        reqs = []
        t_start = time.perf_counter()
        t_next = t_start
        while True:
            t_next += rng.exponential(1.0 / rate)
            if t_next - t_start > seconds:
                break
            k = int(rng.choice(4, p=weights))
            g = grids[k]
            u = rng.random()
            if u < 0.5:
                kind, cnt = "pf", 1
                payload = None
            elif u < 0.85:
                cnt = int(rng.integers(1, 257))
                sel = rng.choice(nonbridge[k], min(cnt, len(nonbridge[k])), replace=False)
                payload = g.ybus_deltas([[int(b)] for b in sel])
                kind, cnt = "ca", len(sel)
            else:
                cnt = int(rng.integers(1, 97))
                payload = gridgen.time_series_sbus(g, cnt, seed=int(rng.integers(1 << 30)))
                kind = "ts"
            reqs.append((t_next, k, kind, cnt, payload))
        ids = []

        def submitter():
            for (ta, k, kind, cnt, payload) in reqs:
                dt = ta - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                if kind == "pf":
                    ids.append((k, kind, sl.submit_pf(gids[k], None)))
                elif kind == "ca":
                    ids.append((k, kind, sl.submit_contingencies(gids[k], *payload)))
                else:
                    ids.append((k, kind, sl.submit_time_series(gids[k], payload)))

        th = threading.Thread(target=submitter)
        t0 = time.perf_counter()
        th.start()
        th.join()
        lat, solves, conv = [], 0, 0
        per_grid = {nm: [] for nm in names}
        per_kind = {"pf": [], "ca": [], "ts": []}
        for (k, kind, i) in ids:
            r = sl.result(i, 300.0)
            lat.append(r["latency_s"])
            per_grid[names[k]].append(r["latency_s"])
            per_kind[kind].append(r["latency_s"])
            solves += len(r["status"])
            conv += int(np.sum(r["status"] == 0))
        wall = time.perf_counter() - t0
        launches = sl.launches()
        sl.close()
        lat = np.array(lat)

        def pct(v):
            return {"n": len(v), "p50_ms": 1e3 * float(np.percentile(v, 50)), "p99_ms": 1e3 * float(np.percentile(v, 99))} \
                if v else {"n": 0}
        out[pname] = {"requests": len(ids), "solves": solves, "converged": conv, "wall_s": wall,
                      "solves_per_s": conv / wall, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
                      "p99_ms": 1e3 * float(np.percentile(lat, 99)), "launches": launches,
                      "per_grid": {nm: pct(v) for nm, v in per_grid.items()},
                      "per_kind": {nm: pct(v) for nm, v in per_kind.items()}}
    return out


def time_series_run(with_cpu: bool, nthreads: int, T: int = 8760) -> dict:
    """118-bus, 8760 hourly steps (TimeSeriesCPP), both forms: the reference's single warm-started
    chain, and chains of 8 steps side by side; the CPU oracle on the same layouts."""
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    g = gridgen.make("case118_synth")
    Sb = gridgen.time_series_sbus(g, T, seed=1)
    gen_bus = np.r_[g.slack, g.pv]
    gen_p = Sb[:, gen_bus].real * (gen_bus != g.slack)
    load_p, load_q = -Sb[:, g.pq].real, -Sb[:, g.pq].imag
    out = {"workload": "case118_synth (1 slack, 53 PV), %d hourly steps, warm start from the previous step "
                       "(the single chain in two batched passes; chains of 8 in one launch)" % T}
    from lightsim2grid_amd import abi
    d = g.solver_inputs()
    ctx = abi.Context(0)
    ncu = ctx.device_info()["n_cu"]
    gr = abi.Grid(ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    info = gr.info
    gr.close()
    ctx.close()
    for chain, lname in ((0, "single_chain"), (8, "chains_of_8")):
        res, gpu_V = {}, {}
        for st, nm in ((L.SolverType.AMDGPUMixed, "fp32"), (L.SolverType.AMDGPU, "fp64")):
            ts = L.TimeSeriesCPP(gridgen.to_gridmodel(g))
            ts.change_solver(st)
            ts.set_chain_length(chain)
            ts.compute_Vs(gen_p, np.zeros((T, 0)), load_p, load_q, g.v0(), 10, 1e-8)  # warm-up
            t0 = time.perf_counter()
            status = ts.compute_Vs(gen_p, np.zeros((T, 0)), load_p, load_q, g.v0(), 10, 1e-8)
            wall = time.perf_counter() - t0
            res[nm] = {"status": int(status), "steps": T, "wall_s": wall, "solves_per_s": T / wall,
                       "solver_time_s": ts.solver_time()}
            gpu_V[nm] = np.asarray(ts.get_voltages())
        if with_cpu:
            from oracle import oracle
            d = g.solver_inputs()
            nch = 1 if chain == 0 else (T + chain - 1) // chain
            Vr, itr, _, conv, status, wall = oracle.time_series(d["Ybus"], Sb, d["slack_ids"], d["slack_weights"],
                                                                d["pv"], d["pq"], g.v0(), 10, 1e-8, False, nch,
                                                                1 if nch == 1 else nthreads)
            res["cpu_baseline"] = {"value": float(np.sum(conv)) / wall, "cores": 1 if nch == 1 else nthreads,
                                   "kind": "port", "wall_s": wall}
            # every step against the oracle on the same chain layout (TimeSeriesCPP keeps no
            # per-step iteration counts: flags and voltages only)
            # LDS roof of the solver time: the factors live in LDS, so the algorithmic bytes (at the
            # oracle's iteration counts: one solve per step) are priced against the LDS bandwidth of
            # every compute unit (the single chain runs in two batched passes over all the steps,
            # the chains side by side in one launch)
            cus = ncu
            for nm, w in (("fp32", 4), ("fp64", 8)):
                b = algorithmic_bytes(info, itr[conv], w)
                t = res[nm]["solver_time_s"]
                res[nm]["lds_roof"] = {"achieved_gbs": b / t / 1e9, "peak_gbs": LDS_PEAK_GBS_PER_CU * cus,
                                       "frac": b / t / 1e9 / (LDS_PEAK_GBS_PER_CU * cus), "compute_units": cus,
                                       "iterations": "the oracle's on the same layout (the FP32 form may differ by 1)"}
            res["parity"] = {}
            for nm, Vg in gpu_V.items():
                ok = np.all(np.isfinite(Vg), axis=1) & (np.abs(Vg).sum(axis=1) > 0)
                both = ok & conv
                res["parity"][nm] = {"flags_equal": bool(np.array_equal(ok, conv)),
                                     "max_dV": float(np.max(np.abs(Vg[both] - Vr[both]))) if both.any() else None}
        out[lname] = res
    return out


def large_grid_run(with_cpu: bool, nthreads: int, n_out: int = 5000, n_snap: int = 4) -> dict:
    """10000-bus grid: n_out seeded non-bridge outages x n_snap load snapshots (load x U(0.8, 1.1)),
    plus the latency of one solve."""
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case10000_synth")
    d = g.solver_inputs()
    rng = np.random.default_rng(10000)
    nb = np.nonzero(~gridgen.bridges(g))[0]
    outs = np.sort(rng.choice(nb, min(n_out, len(nb)), replace=False))
    scales = rng.uniform(0.8, 1.1, n_snap)
    conts = [[int(b)] for b in outs for _ in range(n_snap)]
    dptr, dpos, dval = g.ybus_deltas(conts)
    S = np.stack([g.sbus(load_scale=sc, gen_scale=sc) for sc in scales])
    Sall = np.ascontiguousarray(np.tile(S, (len(outs), 1)))
    ninst = len(conts)
    ctx = abi.Context(0)
    t0 = time.perf_counter()
    gr = abi.Grid(ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    t_load = time.perf_counter() - t0
    Yx = gr.ybus_values(d["Ybus"])
    base = gr.solve(Yx, d["Sbus"], d["V0"], 10, 1e-8)
    Vb = base["V"][0]
    dY, dS, dV0 = ctx.to_device(Yx), ctx.to_device(Sall), ctx.to_device(Vb)
    dV = ctx.empty((ninst, g.n), np.complex128)
    dIt, dSt = ctx.empty(ninst, np.int32), ctx.empty(ninst, np.int32)
    dd = (ctx.to_device(dptr), ctx.to_device(dpos), ctx.to_device(dval))
    out = {"workload": "case10000_synth (distributed slack over the 2500 PV generators, loading 0.25): "
                       "%d non-bridge outages x %d load snapshots = %d solves" % (len(outs), n_snap, ninst),
           "grid_load_s": t_load, "n_jac": gr.info["n_jac"], "nnz_lu": gr.info["nnz_lu"]}
    # the launch autotuner on the first 2048 solves of the batch (untimed), as for the 2000-bus sweep
    kt = min(2048, ninst)
    dd_t = (ctx.to_device(dptr[:kt + 1]), ctx.to_device(dpos[:dptr[kt]]), ctx.to_device(dval[:dptr[kt]]))
    for prec, nm in ((abi.FP32, "fp32"), (abi.FP64, "fp64")):
        tuned = gr.autotune_device(dY, 0, dS, g.n, dV0, 0, dV, kt, dIt, dSt, 10, 1e-8, precision=prec, deltas=dd_t)
        gr.solve_device(dY, 0, dS, g.n, dV0, 0, dV, ninst, dIt, dSt, 10, 1e-8, precision=prec, deltas=dd)
        t0 = time.perf_counter()
        st = gr.solve_device(dY, 0, dS, g.n, dV0, 0, dV, ninst, dIt, dSt, 10, 1e-8, precision=prec, deltas=dd)
        wall = time.perf_counter() - t0
        status = dSt.to_host()
        iters = dIt.to_host()
        conv = int(np.sum(status == 0))
        w = 4 if prec == abi.FP32 else 8
        info = gr.info
        out[nm] = {"wall_s": wall, "kernel_ms": 1e3 * st["kernel_s"], "converged": conv, "solves_per_s": conv / wall,
                   "instances_per_wg": st["instances_per_wg"], "lds_slots": st["lds_slots"],
                   "fallback_fp64": st["fallback_fp64"], "resident_workgroups": st["resident_workgroups"],
                   "iters_mean": float(np.mean(iters[status == 0])), "autotune": list(tuned),
                   "achieved_gbs": algorithmic_bytes(info, iters[status == 0], w) / st["kernel_s"] / 1e9,
                   "hbm_frac": algorithmic_bytes(info, iters[status == 0], w) / st["kernel_s"] / 1e9 / HBM_PEAK_GBS,
                   "flop_rate": flop_rate(info, iters[status == 0], st["kernel_s"], w)}
        # one solve: the default launch (a wide launch: every workgroup on the one instance, level
        # by level) and the batched kernel's one workgroup per instance
        one = (dptr[:2], dpos[:dptr[1]], dval[:dptr[1]])
        for G, key in ((0, "single_solve_kernel_ms"), (1, "single_solve_one_workgroup_kernel_ms")):
            lat = []
            for _ in range(5):
                r = gr.solve(Yx, d["Sbus"], Vb, 10, 1e-8, precision=prec, deltas=one, instances_per_wg=G)
                lat.append(r["stats"]["kernel_s"])
            out[nm][key] = 1e3 * float(np.median(lat))
            if G == 0:
                out[nm]["single_solve_launch"] = "wide" if r["stats"]["instances_per_wg"] == -1 else "batched"
    if with_cpu:
        from oracle import oracle
        k = 16 * nthreads
        sub = (dptr[:k + 1], dpos[:dptr[k]], dval[:dptr[k]])
        Vr, itr, _, conv, _, wall = oracle.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"],
                                                       d["pv"], d["pq"], Vb, sub, 10, 1e-8, False, False, nthreads)
        out["cpu_baseline"] = {"value": float(np.sum(conv)) / wall, "cores": nthreads, "kind": "port",
                               "sample": "%d outage solves (base injections), %.2f s" % (k, wall)}
        # the same sample on the GPU, both forms, against the oracle
        out["parity"] = {}
        for prec, nm in ((abi.FP32, "fp32"), (abi.FP64, "fp64")):
            r = gr.solve(Yx, d["Sbus"], Vb, 10, 1e-8, precision=prec, deltas=sub)
            out["parity"][nm] = parity(r["V"], r["iters"], r["status"], Vr, itr, conv)
            out["parity"][nm]["sample"] = "the %d outages of the CPU sample" % k
    gr.close()
    ctx.close()
    return out


def dc_fdpf_run(with_cpu: bool, reps: int = 20) -> dict:
    """The other two solvers on the same factor / solve kernels: the DC N-1 sweep of the 2000-bus
    grid (one linear solve per outage, BaseDCAlgo.tpp:15-179) and the fast-decoupled (XB) 8760-step
    time series of the 118-bus grid as one warm-started chain (BaseFDPFAlgo.tpp:12-214); inputs
    resident in HBM, FP64 (the reference's only form for both). The CPU oracle times a sample of
    the DC outages and the whole FDPF chain on one core, and checks the GPU's results."""
    import scipy.sparse as sp
    from lightsim2grid_amd import abi, gridgen
    from lightsim2grid_amd import lightsim2grid_amd_cpp as C
    out = {}
    ctx = abi.Context(0)
    # ---- DC N-1: B = Im(1/(jx)) stamps of the lines, each outage an edit of its four entries
    g = gridgen.make("case2000_synth")
    d = g.solver_inputs()
    y = 1.0 / g.x
    Y = sp.csc_matrix((np.r_[y, y, -y, -y].astype(np.complex128),
                       (np.r_[g.f, g.t, g.f, g.t], np.r_[g.f, g.t, g.t, g.f])), shape=(g.n, g.n))
    Y.sum_duplicates()
    Y.sort_indices()
    nb = np.nonzero(~gridgen.bridges(g))[0]

    def pos(r, c):
        a, b = Y.indptr[c], Y.indptr[c + 1]
        return a + int(np.searchsorted(Y.indices[a:b], r))

    dptr, dpos, dval = [0], [], []
    for k in nb:
        f, t = int(g.f[k]), int(g.t[k])
        for rr, cc, v in ((f, f, y[k]), (t, t, y[k]), (f, t, -y[k]), (t, f, -y[k])):
            dpos.append(pos(rr, cc))
            dval.append(v)
        dptr.append(len(dpos))
    dptr, dpos, dval = np.array(dptr, np.int32), np.array(dpos, np.int32), np.array(dval, np.complex128)
    ninst = len(nb)
    gr = abi.Grid(ctx, Y, d["slack_ids"], None, d["pv"], d["pq"], dc=True)
    dY = ctx.to_device(gr.ybus_values(Y))
    dS = ctx.to_device(np.ascontiguousarray(d["Sbus"], np.complex128))
    dV0 = ctx.to_device(np.ascontiguousarray(d["V0"], np.complex128))
    dV = ctx.empty((ninst, g.n), np.complex128)
    dIt, dSt = ctx.empty(ninst, np.int32), ctx.empty(ninst, np.int32)
    dd = tuple(ctx.to_device(x) for x in (dptr, dpos, dval))
    for _ in range(3):
        gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, ninst, dIt, dSt, 10, 1e-8, precision=abi.FP64, deltas=dd)
    ks = 0.0
    t0 = time.perf_counter()
    for _ in range(reps):
        ks += gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, ninst, dIt, dSt, 10, 1e-8, precision=abi.FP64,
                              deltas=dd)["kernel_s"]
    wall = (time.perf_counter() - t0) / reps
    Vg, stg = dV.to_host(), dSt.to_host()
    dc = {"workload": "case2000_synth DC N-1: %d non-bridge single-branch outages, B from the line reactances, "
                      "one factor + solve each, in one launch" % ninst,
          "solves_per_s": int(np.sum(stg == 0)) / wall, "ms_per_sweep": 1e3 * wall, "kernel_ms": 1e3 * ks / reps,
          "converged": int(np.sum(stg == 0))}
    gr.close()
    if with_cpu:
        from oracle import oracle
        sample = np.arange(0, ninst, 8)
        dmax, flags_eq, tcpu = 0.0, True, 0.0
        for i in sample:
            Yc = Y.copy()
            Yc.data[dpos[dptr[i]:dptr[i + 1]]] -= dval[dptr[i]:dptr[i + 1]]
            t1 = time.perf_counter()
            Vr, err = oracle.dc_pf(Yc, d["V0"], d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
            tcpu += time.perf_counter() - t1
            flags_eq &= (err == 0) == (stg[i] == 0)
            if err == 0 and stg[i] == 0:
                dmax = max(dmax, float(np.max(np.abs(Vg[i] - Vr))))
        dc["cpu_baseline"] = {"value": len(sample) / tcpu, "unit": "DC power flows/s", "cores": 1, "kind": "port",
                              "sample": "every 8th outage (%d), one thread" % len(sample)}
        dc["parity"] = {"flags_equal": bool(flags_eq), "max_dV": dmax, "outages_checked": int(len(sample))}
    out["dc_n1_2000"] = dc
    # ---- FDPF (XB) time series: B' and B'' from the host mirror's GridModel (fillBp_Bpp)
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    T = 8760
    Sb = np.ascontiguousarray(gridgen.time_series_sbus(g, T, seed=1), np.complex128)
    gm = gridgen.to_gridmodel(g)
    Bp, Bpp = gm.debug_get_Bp_python(C.FDPFMethod.XB), gm.debug_get_Bpp_python(C.FDPFMethod.XB)
    gr = abi.Grid(ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], fdpf=(Bp, Bpp))
    dY = ctx.to_device(gr.ybus_values(d["Ybus"]))
    dS = ctx.to_device(Sb)
    dV0 = ctx.to_device(np.ascontiguousarray(g.v0(), np.complex128))
    dV = ctx.empty((T, g.n), np.complex128)
    dIt, dSt = ctx.empty(T, np.int32), ctx.empty(T, np.int32)
    gr.solve_chains_device(dY, 0, dS, dV0, 0, dV, T, T, dIt, dSt, 30, 1e-8, precision=abi.FP64)
    walls, ks = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        st = gr.solve_chains_device(dY, 0, dS, dV0, 0, dV, T, T, dIt, dSt, 30, 1e-8, precision=abi.FP64)
        walls.append(time.perf_counter() - t0)
        ks.append(st["kernel_s"])
    wall = float(np.median(walls))
    Vg, itg, stg = dV.to_host(), dIt.to_host(), dSt.to_host()
    fd = {"workload": "case118_synth fast-decoupled XB, %d hourly steps, one chain warm-started from the previous "
                      "step (two batched passes), max 30 iterations" % T,
          "steps_per_s": int(np.sum(stg == 0)) / wall, "ms_per_chain": 1e3 * wall,
          "kernel_ms": 1e3 * float(np.median(ks)), "converged": int(np.sum(stg == 0))}
    gr.close()
    if with_cpu:
        from oracle import oracle
        V = np.asarray(g.v0(), np.complex128)
        itr, err = np.zeros(T, np.int32), np.zeros(T, np.int32)
        Vr = np.zeros((T, g.n), np.complex128)
        t1 = time.perf_counter()
        for t in range(T):
            V, itr[t], err[t] = oracle.fdpf(d["Ybus"], V, Sb[t], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"],
                                            Bp, Bpp, 30, 1e-8)
            Vr[t] = V
            if err[t] != 0:  # the chain stops at its first failure, as TimeSeries does
                break
        tcpu = time.perf_counter() - t1
        ok = err == 0
        fd["cpu_baseline"] = {"value": float(np.sum(ok)) / tcpu, "unit": "FDPF steps/s", "cores": 1, "kind": "port",
                              "sample": "the whole chain, step by step, one thread"}
        both = ok & (stg == 0)
        fd["parity"] = {"flags_equal": bool(np.array_equal(ok, stg == 0)),
                        "iterations_equal": int(np.sum(itg[both] == itr[both])), "steps": T,
                        "max_dV": float(np.max(np.abs(Vg[both] - Vr[both]))) if both.any() else None}
    out["fdpf_xb_time_series_118"] = fd
    ctx.close()
    return out


def single_solves(name: str, reps: int = 5) -> dict:
    """One solve of a synthetic grid from its V0, reps times per form: the default launch (a wide
    launch, every workgroup on the one instance, for a large plan) and one workgroup per instance."""
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make(name)
    d = g.solver_inputs()
    ctx = abi.Context(0)
    gr = abi.Grid(ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    Yx = gr.ybus_values(d["Ybus"])
    out = {"workload": f"{name}, one solve from V0, tol 1e-8, max 10 iterations"}
    for prec, nm in ((abi.FP64, "fp64"), (abi.FP32_STRICT, "fp32_strict")):
        for G, key in ((0, "default"), (1, "one_workgroup")):
            ms = []
            for _ in range(reps):
                r = gr.solve(Yx, d["Sbus"], d["V0"], 10, 1e-8, precision=prec, instances_per_wg=G)
                ms.append(1e3 * r["stats"]["kernel_s"])
            out[f"{nm}_{key}"] = {"launch": "wide" if r["stats"]["instances_per_wg"] == -1 else "batched",
                                  "iters": int(r["iters"][0]), "status": int(r["status"][0]),
                                  "kernel_ms": ms}
    gr.close()
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", default="case2000_synth")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads for the CPU baseline (0: every CPU this process may use)")
    ap.add_argument("--instances-per-wg", type=int, default=0)
    ap.add_argument("--no-extra", action="store_true", help="skip the other configurations")
    ap.add_argument("--serving-seconds", type=float, default=4.0)
    ap.add_argument("--single-solve", metavar="GRID", default=None,
                    help="only time single solves of GRID from its V0 (default launch, then one workgroup) "
                         "and print one JSON line")
    ap.add_argument("--time-series", action="store_true",
                    help="only run the 118-bus time-series leg (GPU only) and print one JSON line")
    ap.add_argument("--dc-fdpf", action="store_true",
                    help="only run the DC N-1 and fast-decoupled time-series legs (with their CPU samples "
                         "unless --no-cpu) and print one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the N-1 sweep would shard by "
              "contingency across GPUs, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    from lightsim2grid_amd import abi, gridgen

    if args.time_series:
        print(json.dumps(time_series_run(False, 0)))
        return
    if args.dc_fdpf:
        print(json.dumps(dc_fdpf_run(not args.no_cpu)))
        return
    if args.single_solve:
        print(json.dumps(single_solves(args.single_solve)))
        return

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

    results, tuned, gpu_rows = {}, {}, {}
    for prec, name in ((abi.FP64, "fp64"), (abi.FP32, "fp32")):
        if not args.instances_per_wg:
            # the project's launch autotuner (instances per workgroup x block size), untimed
            tuned[name] = gr.autotune_device(dY, 0, dS, 0, dV0, 0, dV, ninst, dIt, dSt, 10, 1e-8, precision=prec,
                                             deltas=(ddp, ddq, ddv))
        for _ in range(args.warmup):
            sweep(prec)
        ksum = 0.0
        fb = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = sweep(prec)
            ksum += st["kernel_s"]
            fb += st["fallback_fp64"]
        t1 = time.perf_counter()
        iters = dIt.to_host()
        status = dSt.to_host()
        gpu_rows[name] = (dV.to_host(), iters, status)
        # the same sweep with the work queue in index order (no iteration counts of an earlier
        # sweep to take the longest solves first), untimed by the rate above
        gr.reset()
        first = sweep(prec)
        assert np.array_equal(dIt.to_host(), iters) and np.array_equal(dSt.to_host(), status)
        conv = status == 0
        wall = t1 - t0
        w = 4 if prec == abi.FP32 else 8
        bytes_conv = algorithmic_bytes(info, iters[conv], w)  # converged solves only
        kavg = ksum / args.steps
        results[name] = dict(
            converged=int(conv.sum()), ninst=ninst, wall_s=wall, ms_per_step=1e3 * wall / args.steps,
            solves_per_s=int(conv.sum()) * args.steps / wall, kernel_ms=1e3 * kavg,
            iters_mean=float(np.mean(iters[conv])), achieved_gbs=bytes_conv / kavg / 1e9, bytes_per_launch=bytes_conv,
            flop_rate=flop_rate(info, iters[conv], kavg, w),
            lds_slots=st["lds_slots"], instances_per_wg=st["instances_per_wg"], block_size=st["block_size"],
            resident_workgroups=st["resident_workgroups"], fallback_fp64_per_step=fb / args.steps,
            status_hist=np.bincount(status, minlength=9).tolist(), autotune=tuned.get(name),
            kernel_ms_index_order=1e3 * first["kernel_s"],
            phase_ms_wg0={k: 1e3 * v for k, v in st["phase_s"].items()})

    cinfo = cpu_info()
    nthreads = args.cpu_threads or cinfo["usable"]
    cpu, par = None, None
    if not args.no_cpu:
        from oracle import oracle
        # the full sweep on every CPU this process may use (each thread solves the base case first,
        # like an independent ContingencyAnalysis object), plus a 1-thread sample
        deltas = (dptr, dpos, dval)
        Vr, itr, _, cconv, _, wall_mt = oracle.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"],
                                                           d["pv"], d["pq"], Vb, deltas, 10, 1e-8, False, False,
                                                           nthreads)
        k1 = min(256, ninst)
        sub = (dptr[:k1 + 1], dpos[:dptr[k1]], dval[:dptr[k1]])
        _, _, _, cconv1, _, wall_1 = oracle.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"],
                                                        d["pv"], d["pq"], Vb, sub, 10, 1e-8, False, False, 1)
        one = float(np.sum(cconv1)) / wall_1
        value = float(np.sum(cconv)) / wall_mt
        phys = cinfo["physical_cores"]
        cpu = dict(value=value, unit="converged power flows/s", cores=nthreads, kind="port",
                   sample=f"oracle (C++ FP64 restatement of BaseNRAlgo<KLU> + ContingencyAnalysis loop), full "
                          f"{ninst}-outage sweep on {nthreads} threads, {wall_mt:.2f} s",
                   host=cinfo, one_thread=dict(value=one, sample=f"{k1} outages, {wall_1:.2f} s",
                                               ms_per_contingency=1e3 / one),
                   all_physical_cores_extrapolated=(one * phys) if phys else None,
                   speedup={"fp32": results["fp32"]["solves_per_s"] / value,
                            "fp64": results["fp64"]["solves_per_s"] / value,
                            "fp32_vs_all_cores_extrapolated": (results["fp32"]["solves_per_s"] / (one * phys))
                            if phys else None},
                   note="threads = the CPUs this process may use on the GPU box (affinity and cgroup quota); "
                        "the all-core figure scales the 1-thread rate by the machine's physical cores (an upper "
                        "bound for the CPU)")
        par = {name: parity(*gpu_rows[name], Vr, itr, cconv) for name in ("fp32", "fp64")}

    extras = {}
    if not args.no_extra:
        for name, fn in (("single_pf_14", single_solve_latency),
                         ("serving", lambda: serving_run(args.serving_seconds)),
                         ("time_series_118", lambda: time_series_run(not args.no_cpu, nthreads)),
                         ("contingency_10000", lambda: large_grid_run(not args.no_cpu, nthreads)),
                         ("dc_and_fdpf", lambda: dc_fdpf_run(not args.no_cpu))):
            try:
                extras[name] = fn()
            except Exception as exc:  # an extra configuration never hides the headline line
                extras[name] = {"error": repr(exc)}

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
                               f"from the base-case V, tol 1e-8 pu, max 10 iterations; distributed slack over the "
                               f"reference slack and the 400 PV generators, loads and dispatch at 0.35 of the "
                               f"generated values (full loading does not converge on this synthetic grid); FP32 "
                               f"factors with FP64 re-solve of stalled rows",
                   "n_bus": g.n, "n_branch": g.nbr, "n_jac": info["n_jac"], "nnz_jac": info["nnz_jac"],
                   "nnz_lu": info["nnz_lu"], "instances_per_step": ninst},
        "roofline": {"bound": "hbm", "achieved": r32["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r32["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic("fp32") if args.grid == "case2000_synth" else None,
                     "traffic_source": "profiles/k_nr_case2000_n1_hbm_traffic.md",
                     "bytes_per_launch": r32["bytes_per_launch"], "priced_on": "converged solves",
                     "kernel": "k_nr<float,G,block,lean>", "kernel_ms": r32["kernel_ms"]},
        "cpu_baseline": cpu,
        "parity": par,
        "fp64_form": {"value": r64["solves_per_s"], "ms_per_step": r64["ms_per_step"], "kernel_ms": r64["kernel_ms"],
                      "achieved_gbs": r64["achieved_gbs"], "iters_mean": r64["iters_mean"],
                      "converged": r64["converged"], "flop_rate": r64["flop_rate"]},
        "fp32_over_fp64": r32["solves_per_s"] / r64["solves_per_s"],
        "details": {"fp32": r32, "fp64": r64},
        "other_configs": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
