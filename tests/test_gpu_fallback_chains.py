"""GPU tests of the solve paths around the NR kernel, each against the oracle:

* FP32 -> FP64 re-solve of rows that stall with FP32 factors (PFGPU_FP32), on the device path;
* zero-pivot statuses: SolverFactor (5) on a fresh solver, SolverReFactor (6) afterwards, as
  klu_factor / klu_refactor report them (src/linear_solvers/KLUSolver.cpp:38-57);
* one solver reused across changed pv / pq sets (BaseNRAlgo.hpp:155-166);
* time-series chains inside one launch (TimeSeries.cpp:80-103), the reference's single chain
  and several chains side by side, including a chain that fails;
* the single chain's two batched passes when steps solved from the head diverge where the
  warm-started chain converges (restarts, the sequential rest, a failure after restarts).
"""
import sys

import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu

FP64, FP32, FP32_STRICT = 0, 1, 2


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _n1(grid_name, max_cont=None, seed=0):
    from lightsim2grid_amd import gridgen
    g = gridgen.make(grid_name)
    nb = np.nonzero(~gridgen.bridges(g))[0]
    if max_cont is not None and len(nb) > max_cont:
        nb = np.sort(np.random.default_rng(seed).choice(nb, max_cont, replace=False))
    conts = [[int(k)] for k in nb]
    return g, conts


def test_fp32_stall_falls_back_to_fp64(gpu_ctx, oracle_mod):
    """Rows whose FP32-factor solve needs one iteration more than FP64, run with max_iter = the
    FP64 count: the FP32 launch flags them as stalled (last mismatch below sqrt(tol)), the
    FP64 pass re-solves them, and the statuses, iterations and voltages equal the oracle's."""
    from lightsim2grid_amd import abi
    found = None
    # at tol 1e-8 the FP32-factor iterations equal the FP64 ones on these grids; closer to FP64
    # rounding, FP64 converges quadratically while the inexact FP32 steps converge linearly
    for tol in (1e-8, 1e-10, 1e-11, 1e-12):
        for grid_name, maxc in (("case118_synth", None), ("case2000_synth", 768)):
            g, conts = _n1(grid_name, maxc)
            d = g.solver_inputs()
            _say("fallback search", grid_name, len(conts), tol)
            o = oracle_mod.OracleNR(False)
            assert o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10,
                                1e-8)
            Vb = o.get_V()
            deltas = g.ybus_deltas(conts)
            _, itr, _, convr, _, _ = oracle_mod.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"],
                                                            d["pv"], d["pq"], Vb, deltas, 10, tol, False, False, 4)
            gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
            r32 = gr.solve(gr.ybus_values(d["Ybus"]), d["Sbus"], Vb, 10, tol, precision=FP32_STRICT, deltas=deltas)
            cand = convr & (r32["status"] == 0) & (r32["iters"] == itr + 1)
            if cand.any():
                found = (g, d, conts, Vb, itr, cand, gr, tol)
                break
        if found:
            break
    assert found is not None, "no outage needs one more FP32 iteration than FP64 on these grids"
    g, d, conts, Vb, itr, cand, gr, tol = found
    k = int(np.bincount(itr[cand]).argmax())
    sel = np.nonzero(cand & (itr == k))[0]
    sub = g.ybus_deltas([conts[i] for i in sel])
    _say("fallback", len(sel), "rows, max_iter", k)
    Yx = gr.ybus_values(d["Ybus"])
    strict = gr.solve(Yx, d["Sbus"], Vb, k, tol, precision=FP32_STRICT, deltas=sub)
    assert np.all(strict["status"] == 2)
    flags = gpu_ctx.last_flags(len(sel))
    assert np.all(flags & abi.FLAG_STALL)
    r = gr.solve(Yx, d["Sbus"], Vb, k, tol, precision=FP32, deltas=sub)
    assert r["stats"]["fallback_fp64"] == len(sel)
    Vr, itr2, _, convr2, _, _ = oracle_mod.contingency(d["Ybus"], d["Sbus"], d["slack_ids"], d["slack_weights"],
                                                       d["pv"], d["pq"], Vb, sub, k, tol, False, False, 2)
    assert convr2.all()
    assert np.all(r["status"] == 0)
    assert np.array_equal(r["iters"], itr2)
    assert np.max(np.abs(r["V"] - Vr)) <= 1e-8
    # the same on device buffers (pfgpu_solve_batch_device runs the re-solve pass on the device)
    n, m = g.n, len(sel)
    dY, dS, dV0 = gpu_ctx.to_device(Yx), gpu_ctx.to_device(np.ascontiguousarray(d["Sbus"])), gpu_ctx.to_device(Vb)
    dV = gpu_ctx.empty((m, n), np.complex128)
    dIt, dSt = gpu_ctx.empty(m, np.int32), gpu_ctx.empty(m, np.int32)
    dd = tuple(gpu_ctx.to_device(x) for x in sub)
    st = gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, m, dIt, dSt, k, tol, precision=FP32, deltas=dd)
    assert st["fallback_fp64"] == m
    assert np.all(dSt.to_host() == 0) and np.array_equal(dIt.to_host(), itr2)
    assert np.max(np.abs(dV.to_host() - Vr)) <= 1e-8
    # in place (V0 == V, one row per instance): the re-solve starts from the kept initial rows
    dVi = gpu_ctx.to_device(np.repeat(Vb[None, :], m, axis=0))
    st = gr.solve_device(dY, 0, dS, 0, dVi, n, dVi, m, dIt, dSt, k, tol, precision=FP32, deltas=dd)
    assert st["fallback_fp64"] == m
    assert np.max(np.abs(dVi.to_host() - Vr)) <= 1e-8


def _singular_case14():
    """case14 with every Ybus value of one PQ bus zeroed (pattern kept): J has a zero column."""
    d = load_fixture("case14")
    Y = d["Ybus"].tocsc().copy()
    Y.sort_indices()
    b = int(d["pq"][3])
    for c in range(Y.shape[1]):
        for q in range(Y.indptr[c], Y.indptr[c + 1]):
            if c == b or Y.indices[q] == b:
                Y.data[q] = 0.0
    d["Ybus"] = Y
    return d


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_zero_pivot_status_factor_then_refactor(gpu_ctx, oracle_mod, prec):
    from lightsim2grid_amd import abi
    d = _singular_case14()
    _say("zero pivot", prec)
    o = oracle_mod.OracleNR(False)
    a = (d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    assert not o.compute_pf(*a)
    e1 = o.get_error()
    assert not o.compute_pf(*a)
    e2 = o.get_error()
    assert (e1, e2) == (5, 6)
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    Yx = gr.ybus_values(d["Ybus"])
    s = [int(gr.solve(Yx, d["Sbus"], d["V0"], 10, 1e-8, precision=prec)["status"][0]) for _ in range(2)]
    assert s == [e1, e2]
    gr.reset()
    assert int(gr.solve(Yx, d["Sbus"], d["V0"], 10, 1e-8, precision=prec)["status"][0]) == 5
    # the host mirror: a fresh solver object reports SolverFactor
    from lightsim2grid_amd import solver
    sv = solver.AMDGPUSolver() if prec == FP64 else solver.AMDGPUMixedSolver()
    assert not sv.compute_pf(*a)
    assert int(sv.get_error()) == 5
    assert not sv.compute_pf(*a)
    assert int(sv.get_error()) == 6
    sv.reset()
    assert not sv.compute_pf(*a)
    assert int(sv.get_error()) == 5


def test_solver_reused_with_changed_pv_pq(oracle_mod):
    """One solver object, three calls: the fixture's pv / pq, then a PV bus turned PQ (the grid is
    reloaded), then the original sets again; each equals a fresh oracle on the same inputs."""
    from lightsim2grid_amd import solver
    d = load_fixture("case118")
    pv, pq = np.asarray(d["pv"]), np.asarray(d["pq"])
    pv2 = pv[1:]
    pq2 = np.sort(np.r_[pq, pv[0]]).astype(np.int32)
    sv = solver.AMDGPUSolver()
    for k, (pvs, pqs) in enumerate(((pv, pq), (pv2, pq2), (pv, pq))):
        _say("reuse", k)
        args = (d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], pvs, pqs, 10, 1e-8)
        o = oracle_mod.OracleNR(False)
        conv = o.compute_pf(*args)
        assert conv
        assert sv.compute_pf(*args) == conv
        assert sv.get_nb_iter() == o.get_nb_iter()
        assert np.max(np.abs(sv.get_V() - o.get_V())) <= 1e-8
        Jg, Jo = sv.get_J().toarray(), o.get_J().toarray()
        assert Jg.shape == Jo.shape
        assert np.max(np.abs(Jg - Jo)) <= 1e-9 * max(1.0, np.abs(Jo).max())


@pytest.mark.parametrize("prec", [FP64, FP32_STRICT, FP32])
@pytest.mark.parametrize("chain", [None, 8])
def test_time_series_chains_match_oracle(gpu_ctx, oracle_mod, prec, chain):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    T = 48
    Sb = gridgen.time_series_sbus(g, T, seed=3)
    _say("chains", chain, prec)
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    r = gr.solve_chains(gr.ybus_values(d["Ybus"]), Sb, g.v0(), chain_len=chain, precision=prec)
    nch = 1 if chain is None else T // chain
    Vr, itr, _, convr, status, _ = oracle_mod.time_series(d["Ybus"], Sb, d["slack_ids"], d["slack_weights"], d["pv"],
                                                          d["pq"], g.v0(), 10, 1e-8, False, nch)
    assert status == 1 and convr.all()
    assert np.all(r["status"] == 0)
    tol_it = 0 if prec == FP64 else 1
    assert np.max(np.abs(r["iters"] - itr)) <= tol_it
    assert np.max(np.abs(r["V"] - Vr)) <= (1e-8 if prec == FP64 else 5e-8)
    if prec == FP32_STRICT:
        assert r["stats"]["fallback_fp64"] == 0


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_time_series_chain_stops_at_first_failure(gpu_ctx, oracle_mod, prec):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    T = 24
    Sb = gridgen.time_series_sbus(g, T, seed=4)
    Sb[9] *= 40.0  # an infeasible step
    _say("chain failure", prec)
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    r = gr.solve_chains(gr.ybus_values(d["Ybus"]), Sb, g.v0(), chain_len=None, precision=prec)
    Vr, itr, _, convr, status, _ = oracle_mod.time_series(d["Ybus"], Sb, d["slack_ids"], d["slack_weights"], d["pv"],
                                                          d["pq"], g.v0(), 10, 1e-8, False, 1)
    assert status == 0 and not convr[9] and convr[:9].all()
    assert np.all(r["status"][:9] == 0) and r["status"][9] != 0
    assert np.all(r["status"][10:] == abi.STATUS_NOT_RUN) and np.all(r["iters"][10:] == 0)
    assert np.all(r["V"][10:] == 0)
    assert np.max(np.abs(r["V"][:9] - Vr[:9])) <= (1e-8 if prec == FP64 else 5e-8)
    # two chains of 12: the second one runs on
    r2 = gr.solve_chains(gr.ybus_values(d["Ybus"]), Sb, g.v0(), chain_len=12, precision=prec)
    assert r2["status"][9] != 0 and np.all(r2["status"][10:12] == abi.STATUS_NOT_RUN)
    assert np.all(r2["status"][12:] == 0)


@pytest.mark.parametrize("window", [False, True])
@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("T,amp,max_iter", [(40, 0.6, 3), (40, 0.3, 2), (80, 1.0, 2)])
def test_single_chain_restarts_match_sequential_chain(gpu_ctx, oracle_mod, prec, T, amp, max_iter, window):
    """A load ramp (Sbus x (1 + amp t / T)) from the converged base case with a small max_iter: the
    steps solved from the head diverge after the first few while the warm-started chain converges
    (40, 0.6, 3: one restart; 40, 0.3, 2: restarts, then the sequential rest; 80, 1.0, 2: the
    chain itself fails at step 58 after restarts). window: a context whose HBM budget holds the
    pass buffers of 19 steps (a quarter of it), so the chain is taken in windows. Every step equals the oracle's
    sequential chain: flags, iterations (FP64: equal; FP32: within 1), voltages, and the not-run
    rows."""
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    o = oracle_mod.OracleNR(False)
    assert o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    Vb = o.get_V()
    S0 = np.asarray(d["Sbus"])
    Sb = np.stack([S0 * (1 + amp * t / T) for t in range(T)])
    _say("two-pass restarts", T, amp, max_iter, prec)
    Vr, itr, _, convr, status, _ = oracle_mod.time_series(d["Ybus"], Sb, d["slack_ids"], d["slack_weights"], d["pv"],
                                                          d["pq"], Vb, max_iter, 1e-8, False, 1)
    # from the head alone, later steps diverge (the case the restarts are for)
    _, _, _, conv_head, _, _ = oracle_mod.time_series(d["Ybus"], Sb, d["slack_ids"], d["slack_weights"], d["pv"],
                                                      d["pq"], Vb, max_iter, 1e-8, False, T)
    assert not conv_head[3 * T // 4:].any()
    ctx = abi.Context(0, 4 * 19 * (32 * g.n + 16)) if window else gpu_ctx
    gr = abi.Grid(ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    r = gr.solve_chains(gr.ybus_values(d["Ybus"]), Sb, Vb, chain_len=None, max_iter=max_iter, precision=prec)
    conv = r["status"] == 0
    assert np.array_equal(conv[convr], np.ones(int(convr.sum()), bool))
    if convr.all():
        assert conv.all()
    else:
        f = int(np.argmin(convr))  # the oracle chain's first failure
        assert conv[:f].all() and not conv[f]
        assert np.all(r["status"][f + 1:] == abi.STATUS_NOT_RUN) and np.all(r["iters"][f + 1:] == 0)
        assert np.all(r["V"][f + 1:] == 0)
    tol_it = 0 if prec == FP64 else 1
    assert np.max(np.abs(r["iters"][convr] - itr[convr])) <= tol_it
    assert np.max(np.abs(r["V"][convr] - Vr[convr])) <= (1e-8 if prec == FP64 else 5e-8)
    gr.close()
    if window:
        ctx.close()


def test_phase_timers_cover_the_kernel(gpu_ctx):
    """The per-phase wall clocks of workgroup 0 add up to at most the launch time, and every
    phase of an NR iteration is seen."""
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    _say("phase timers")
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    r = gr.solve(gr.ybus_values(d["Ybus"]), d["Sbus"], d["V0"], 10, 1e-8)
    ph = r["stats"]["phase_s"]
    assert r["iters"][0] >= 2
    assert all(v > 0 for k, v in ph.items() if k != "tail"), ph  # the tail phase is empty without a dense tile
    assert sum(ph.values()) <= 1.05 * r["stats"]["kernel_s"] + 1e-5


def test_timers_jacobian_of_host_solver(oracle_mod):
    from lightsim2grid_amd import solver
    d = load_fixture("case118")
    _say("timers jacobian")
    s = solver.AMDGPUSolver()
    assert s.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    t = s.get_timers_jacobian()
    assert len(t) == 10
    fx, solve, refactor, init, check, dsbus, fillj, vavm, pre, total = t
    assert fx > 0 and solve > 0 and refactor > 0 and fillj > 0 and vavm > 0 and total > 0
    assert fx + solve + refactor + fillj + vavm <= total


def test_graph_replays_equal_eager_launches(gpu_ctx):
    """The launch sequence of a repeated call is captured as a HIP graph on its second call and
    replayed afterwards: every replay, with other (eager) solves and synchronisations between the
    replays, gives the eager result bit for bit."""
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    _say("graph replays")
    nb = np.nonzero(~gridgen.bridges(g))[0][:64]
    deltas = g.ybus_deltas([[int(k)] for k in nb])
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    Yx = gr.ybus_values(d["Ybus"])
    n, m = g.n, len(nb)
    dY, dS, dV0 = gpu_ctx.to_device(Yx), gpu_ctx.to_device(np.ascontiguousarray(d["Sbus"])), gpu_ctx.to_device(d["V0"])
    dV = gpu_ctx.empty((m, n), np.complex128)
    dIt, dSt = gpu_ctx.empty(m, np.int32), gpu_ctx.empty(m, np.int32)
    dd = tuple(gpu_ctx.to_device(x) for x in deltas)
    Sb = gridgen.time_series_sbus(g, 12, seed=9)
    for prec in (FP64, FP32):
        ref = None
        replays = 0
        for rep in range(6):
            dV.copy_from(np.zeros((m, n), np.complex128))
            st = gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, m, dIt, dSt, 10, 1e-8, precision=prec, deltas=dd)
            out = (dV.to_host(), dIt.to_host(), dSt.to_host())
            replays += st["graph_launches"]
            if ref is None:
                assert st["graph_launches"] == 0  # the first call runs eagerly
                ref = out
            else:
                assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]) and np.array_equal(out[2], ref[2])
            # eager work of another shape between the replays
            gr.solve_chains(Yx, Sb[: 3 + rep], d["V0"], chain_len=None, precision=prec)
        assert replays >= 4


def test_graph_capture_beside_allocations_of_another_thread(gpu_ctx):
    """One thread's solves capture HIP graphs (each launch sequence on its second call) while a
    second thread, on a context of its own, allocates and frees device memory, as the serving
    loop's workers do. HIP refuses a free that would wait on a capturing stream; the library keeps
    its captures and such calls apart, so neither thread sees an error and every captured solve
    gives the eager result bit for bit."""
    import threading
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    d = g.solver_inputs()
    _say("graph capture beside another thread's allocations")
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    Yx = gr.ybus_values(d["Ybus"])
    n, M = g.n, 48
    dY, dS, dV0 = gpu_ctx.to_device(Yx), gpu_ctx.to_device(np.ascontiguousarray(d["Sbus"])), gpu_ctx.to_device(d["V0"])
    dV = gpu_ctx.empty((M, n), np.complex128)
    dIt, dSt = gpu_ctx.empty(M, np.int32), gpu_ctx.empty(M, np.int32)
    other = abi.Context(0)
    stop, errors, churns = threading.Event(), [], [0]

    def churn():
        try:
            while not stop.is_set():
                bufs = [other.empty(1 << 16, np.uint8) for _ in range(4)]
                for b in bufs:
                    abi._check(other._L.pfgpu_dev_free(other.h, b.ptr))
                    b.ptr = None
                churns[0] += 1
        except Exception as e:  # reported by the main thread
            errors.append(e)

    # a first solve of the fresh grid (its key differs from later calls of the same shape)
    gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, M, dIt, dSt, 10, 1e-8, precision=FP64)
    t = threading.Thread(target=churn)
    t.start()
    captured = 0
    try:
        for m in range(1, M):
            outs = []
            for _ in range(2):  # eager, then captured and replayed
                st = gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, m, dIt, dSt, 10, 1e-8, precision=FP64)
                outs.append((dV.to_host()[:m], dIt.to_host()[:m], dSt.to_host()[:m]))
            captured += st["graph_launches"]
            for a, b in zip(outs[0], outs[1]):
                assert np.array_equal(a, b)
            assert (outs[0][2] == 0).all()
    finally:
        stop.set()
        t.join()
        other.close()
    assert not errors, errors
    assert captured == M - 1 and churns[0] > 0
