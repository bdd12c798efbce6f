"""GPU tests of the device kernels around the solve (csrc/batchprep.hip), each against an
independent evaluation on the host:

* connectivity after each contingency vs Tarjan bridges (N-1) and scipy connected components
  of the edited Ybus (N-2), the |y| > 1e-8 rule of ContingencyAnalysis::check_invertible;
* contingency Ybus edits vs the branch stamps gridgen computes (ContingencyAnalysis.cpp:47-110);
* time-series Sbus rows vs the same sums in numpy, bit for bit (TimeSeries.cpp:45-53);
* Ybus values from branch parameters vs the host GridModel's pre_process_solver, bit for bit;
* batched flows with clean-up vs the branch equations on the same voltages, and through
  TimeSeriesCPP.compute_flows.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _edits_inputs(g, conts):
    """Per-branch CSC positions (ff, tt, ft, tf) and stamps, the contingency CSR and dptr."""
    Y = g.ybus().tocsc()
    Y.sort_indices()
    yff, ytt, yft, ytf = g.branch_admittances()

    def pos(r, c):
        s, e = Y.indptr[c], Y.indptr[c + 1]
        return s + int(np.searchsorted(Y.indices[s:e], r))

    bpos = np.zeros((g.nbr, 4), np.int32)
    by = np.zeros((g.nbr, 4), np.complex128)
    for k in range(g.nbr):
        f, t = int(g.f[k]), int(g.t[k])
        bpos[k] = [pos(f, f), pos(t, t), pos(f, t), pos(t, f)]
        by[k] = [yff[k], ytt[k], yft[k], ytf[k]]
    cptr = np.cumsum([0] + [len(c) for c in conts]).astype(np.int32)
    cbr = np.array([k for c in conts for k in c], np.int32)
    dptr = (4 * cptr).astype(np.int32)
    return Y, bpos, by, cptr, cbr, dptr


@pytest.mark.parametrize("grid_name", ["case118_synth", "case2000_synth"])
def test_device_edits_and_connectivity(gpu_ctx, grid_name):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make(grid_name)
    _say("edits+connectivity", grid_name)
    rng = np.random.default_rng(5)
    n1 = [[k] for k in range(g.nbr)]
    n2 = [sorted(rng.choice(g.nbr, 2, replace=False).tolist()) for _ in range(200)]
    conts = n1 + n2
    Y, bpos, by, cptr, cbr, dptr = _edits_inputs(g, conts)
    L = gpu_ctx._L
    d = {k: gpu_ctx.to_device(v) for k, v in dict(cptr=cptr, cbr=cbr, bpos=bpos, by=by, dptr=dptr,
                                                    Yp=Y.indptr.astype(np.int32), Yi=Y.indices.astype(np.int32),
                                                    Yx=Y.data.astype(np.complex128)).items()}
    ne = int(dptr[-1])
    ddq, ddv = gpu_ctx.empty(ne, np.int32), gpu_ctx.empty(ne, np.complex128)
    abi._check(L.pfgpu_contingency_edits_device(gpu_ctx.h, len(conts), d["cptr"].ptr, d["cbr"].ptr, d["bpos"].ptr,
                                                d["by"].ptr, d["dptr"].ptr, ddq.ptr, ddv.ptr))
    dq, dv = ddq.to_host(), ddv.to_host()
    rp, rq, rv = g.ybus_deltas(conts)
    assert np.array_equal(rp, dptr) and np.array_equal(dq, rq) and np.array_equal(dv, rv)
    dconn = gpu_ctx.empty(len(conts), np.int32)
    abi._check(L.pfgpu_connectivity_device(gpu_ctx.h, len(conts), g.n, d["Yp"].ptr, d["Yi"].ptr, d["Yx"].ptr,
                                           d["dptr"].ptr, ddq.ptr, ddv.ptr, dconn.ptr))
    conn = dconn.to_host().astype(bool)
    assert np.array_equal(conn[:g.nbr], ~gridgen.bridges(g))
    for c in range(g.nbr, len(conts)):
        Yc = Y.copy()
        np.subtract.at(Yc.data, dq[dptr[c]:dptr[c + 1]], dv[dptr[c]:dptr[c + 1]])  # repeated positions add up
        A = sp.csc_matrix(((np.abs(Yc.data) > 1e-8).astype(np.int8), Yc.indices, Yc.indptr), shape=Yc.shape)
        A.eliminate_zeros()  # csgraph counts stored zeros as edges
        ncomp, _ = connected_components(A, directed=False)
        assert conn[c] == (ncomp == 1), c


def test_device_sbus_rows_bitwise(gpu_ctx):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    _say("sbus rows")
    rng = np.random.default_rng(2)
    T, ns = 37, g.n
    gen_bus = np.r_[g.slack, g.pv, g.pv[:5]].astype(np.int32)       # buses with two generators
    sgen_bus = rng.choice(ns, 11).astype(np.int32)
    load_bus = np.r_[g.pq, g.pq[:7]].astype(np.int32)
    gp, sg = rng.normal(size=(T, len(gen_bus))), rng.normal(size=(T, len(sgen_bus)))
    lp, lq = rng.normal(size=(T, len(load_bus))), rng.normal(size=(T, len(load_bus)))

    def csr(bus):
        per = [[] for _ in range(ns)]
        for e, b in enumerate(bus):
            per[b].append(e)
        return np.cumsum([0] + [len(p) for p in per]).astype(np.int32), np.array([e for p in per for e in p], np.int32)

    (gptr, gel), (sptr, sel), (lptr, lel) = csr(gen_bus), csr(sgen_bus), csr(load_bus)
    for sn in (1.0, 100.0):
        D = [gpu_ctx.to_device(x) for x in (gptr, gel, gp, sptr, sel, sg, lptr, lel, lp, lq)]
        dS = gpu_ctx.empty((T, ns), np.complex128)
        abi._check(gpu_ctx._L.pfgpu_sbus_rows_device(gpu_ctx.h, T, ns, D[0].ptr, D[1].ptr, len(gen_bus), D[2].ptr,
                                                      D[3].ptr, D[4].ptr, len(sgen_bus), D[5].ptr, D[6].ptr, D[7].ptr,
                                                      len(load_bus), D[8].ptr, D[9].ptr, sn, dS.ptr))
        S = dS.to_host()
        re = np.zeros((T, ns))
        im = np.zeros((T, ns))
        for b in range(ns):  # element order per bus: generators, static generators, loads
            for e in gel[gptr[b]:gptr[b + 1]]:
                re[:, b] += gp[:, e]
            for e in sel[sptr[b]:sptr[b + 1]]:
                re[:, b] += sg[:, e]
            for e in lel[lptr[b]:lptr[b + 1]]:
                re[:, b] -= lp[:, e]
            for e in lel[lptr[b]:lptr[b + 1]]:
                im[:, b] -= lq[:, e]
        if sn != 1.0:
            re, im = re / sn, im / sn
        assert np.array_equal(S.real, re) and np.array_equal(S.imag, im)


def test_device_ybus_values_bitwise(gpu_ctx):
    """Contributions listed per CSC entry in the host assembly's order (branch order; ff, tt, ft, tf),
    with one parallel branch and one branch switched off."""
    import lightsim2grid_amd as L
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    _say("ybus values")
    r, x, h = g.r.copy(), g.x.copy(), 1j * g.b
    f, t = g.f.astype(np.int32).copy(), g.t.astype(np.int32).copy()
    # a parallel branch to branch 3
    r, x, h = np.r_[r, 0.7 * r[3]], np.r_[x, 0.9 * x[3]], np.r_[h, 0.5 * h[3]]
    f, t = np.r_[f, f[3]].astype(np.int32), np.r_[t, t[3]].astype(np.int32)
    on = np.ones(len(r), np.int32)
    gm = L.GridModel()
    gm.init_bus(g.n, 1, np.ones(g.n), len(r), 0)
    gm.init_powerlines(r, x, h, f, t)
    gm.init_generators(np.zeros(1), np.ones(1), -np.ones(1), np.ones(1), np.array([g.slack], np.int32))
    gm.init_loads(np.zeros(0), np.zeros(0), np.zeros(0, np.int32))
    gm.add_gen_slackbus(0, 1.0)
    gm.ac_pf(np.ones(g.n, complex), 0, 1e-8)
    Yh = gm.get_Ybus_solver().tocsc()
    # contributions in the host assembly's insertion order, stably sorted by row per column
    cols = [[] for _ in range(g.n)]
    for k in range(len(r)):
        for u, (rr, cc) in enumerate(((f[k], f[k]), (t[k], t[k]), (f[k], t[k]), (t[k], f[k]))):
            cols[cc].append((rr, 4 * k + u))
    cptr, cidx = [0], []
    for c in range(g.n):
        lst = sorted(cols[c], key=lambda e: e[0])
        rows = sorted(set(e[0] for e in lst) | {c})
        assert np.array_equal(rows, Yh.indices[Yh.indptr[c]:Yh.indptr[c + 1]])
        for rr in rows:
            cidx += [e[1] for e in lst if e[0] == rr]
            cptr.append(len(cidx))
    cptr, cidx = np.array(cptr, np.int32), np.array(cidx, np.int32)
    D = [gpu_ctx.to_device(v) for v in (cptr, cidx, r, x, h.astype(np.complex128), on)]
    dY = gpu_ctx.empty(Yh.nnz, np.complex128)
    abi._check(gpu_ctx._L.pfgpu_ybus_values_device(gpu_ctx.h, Yh.nnz, D[0].ptr, D[1].ptr, len(r), D[2].ptr, D[3].ptr,
                                                    D[4].ptr, D[5].ptr, dY.ptr))
    Yd = dY.to_host()
    assert np.array_equal(Yd.view(np.float64), np.asarray(Yh.data, np.complex128).view(np.float64))
    # a branch switched off: its entries keep explicit zeros where nothing else contributes
    on[5] = 0
    D[5].copy_from(on)
    abi._check(gpu_ctx._L.pfgpu_ybus_values_device(gpu_ctx.h, Yh.nnz, D[0].ptr, D[1].ptr, len(r), D[2].ptr, D[3].ptr,
                                                    D[4].ptr, D[5].ptr, dY.ptr))
    Yoff = dY.to_host()
    gm.deactivate_powerline(5)
    gm.ac_pf(np.ones(g.n, complex), 0, 1e-8)
    Yh2 = gm.get_Ybus_solver().tocsc()
    Yd2 = sp.csc_matrix((Yoff, Yh.indices, Yh.indptr), shape=Yh.shape)
    Yd2.eliminate_zeros()
    Yh2.eliminate_zeros()
    assert (abs(Yd2 - Yh2)).max() == 0.0


def test_batch_flows_rows_and_clean(gpu_ctx):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case118_synth")
    _say("batch flows")
    rng = np.random.default_rng(3)
    nrows = 9
    V = (1.0 + 0.05 * rng.normal(size=(5, g.n))) * np.exp(1j * 0.1 * rng.normal(size=(5, g.n)))
    row_src = np.array([0, 1, -1, 2, 3, 4, -1, 0, 2], np.int32)
    yff, ytt, yft, ytf = g.branch_admittances()
    f, t = g.f.astype(np.int32).copy(), g.t.astype(np.int32).copy()
    f[7] = -1  # out of service
    vn = rng.uniform(20, 400, g.nbr)
    cptr = np.arange(nrows + 1, dtype=np.int32)
    cbr = rng.choice(g.nbr, nrows).astype(np.int32)
    D = [gpu_ctx.to_device(v) for v in (V, row_src, f, t, yff, yft, vn, cptr, cbr)]
    dA, dP = gpu_ctx.empty((nrows, g.nbr), np.float64), gpu_ctx.empty((nrows, g.nbr), np.float64)
    abi._check(gpu_ctx._L.pfgpu_batch_flows_device(gpu_ctx.h, nrows, g.n, D[0].ptr, D[1].ptr, g.nbr, D[2].ptr,
                                                    D[3].ptr, D[4].ptr, D[5].ptr, D[6].ptr, 100.0, D[7].ptr, D[8].ptr,
                                                    dA.ptr, dP.ptr))
    A, P = dA.to_host(), dP.to_host()
    Vr = np.where(row_src[:, None] >= 0, V[np.maximum(row_src, 0)], 0.0)
    Ef, Et = Vr[:, g.f], Vr[:, g.t]
    S = Ef * np.conj(yff * Ef + yft * Et)
    with np.errstate(invalid="ignore", divide="ignore"):
        Aref = (np.abs(S) * 100.0) / (np.sqrt(3) * (np.abs(Ef) * vn))
    Pref = S.real * 100.0
    Aref[:, 7] = 0.0
    Pref[:, 7] = 0.0
    for r in range(nrows):
        k = cbr[r]
        if np.isfinite(Aref[r, k]):
            Aref[r, k] = 0.0
        Pref[r, k] = 0.0
    ok = np.isfinite(Aref)
    assert np.array_equal(ok, np.isfinite(A))
    assert np.allclose(A[ok], Aref[ok], rtol=1e-13, atol=0)
    assert np.allclose(P, Pref, rtol=1e-13, atol=1e-12)
    assert np.all(np.isnan(A[2, np.arange(g.nbr) != 7]))


def test_time_series_flows_on_device(oracle_mod):
    import lightsim2grid_amd as L
    from lightsim2grid_amd import gridgen
    g = gridgen.make("case118_synth")
    T = 16
    _say("TimeSeriesCPP flows")
    Sb = gridgen.time_series_sbus(g, T, seed=8)
    gen_bus = np.r_[g.slack, g.pv]
    gen_p = Sb[:, gen_bus].real * (gen_bus != g.slack)
    ts = L.TimeSeriesCPP(gridgen.to_gridmodel(g))
    assert ts.compute_Vs(gen_p, np.zeros((T, 0)), -Sb[:, g.pq].real, -Sb[:, g.pq].imag, g.v0(), 10, 1e-8) == 1
    ts.compute_flows()
    ts.compute_power_flows()
    A, P = ts.get_flows(), ts.get_power_flows()
    V = ts.get_voltages()
    yff, ytt, yft, ytf = g.branch_admittances()
    Ef, Et = V[:, g.f], V[:, g.t]
    S = Ef * np.conj(yff * Ef + yft * Et)
    assert np.allclose(A, np.abs(S) / (np.sqrt(3) * np.abs(Ef)), rtol=1e-9, atol=1e-12)
    assert np.allclose(P, S.real, rtol=1e-9, atol=1e-10)
