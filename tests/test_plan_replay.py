"""CPU replay of the device plan (the exact slots, tasks and pairs the HIP
kernel executes) against the oracle. Pins the host analysis (J pattern,
scatter map, ordering, symbolic LU, level schedule) without a GPU."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import FIXTURES, build_replay, load_fixture, solver_args

_i32 = ctypes.POINTER(ctypes.c_int)
_f64 = ctypes.POINTER(ctypes.c_double)


def _arrays(d):
    Y = sp.csc_matrix(d["Ybus"])
    Y.sort_indices()
    a = lambda x, t: np.ascontiguousarray(x, dtype=t)
    return dict(n=Y.shape[0], Yp=a(Y.indptr, np.int32), Yi=a(Y.indices, np.int32), Yx=a(Y.data, np.complex128),
                S=a(d["Sbus"], np.complex128), sl=a(np.atleast_1d(d["slack_ids"]), np.int32),
                w=a(d["slack_weights"], np.float64), pv=a(d["pv"], np.int32), pq=a(d["pq"], np.int32))


def _pattern_args(A, single):
    return (A["n"], A["Yp"].ctypes.data_as(_i32), A["Yi"].ctypes.data_as(_i32), A["sl"].ctypes.data_as(_i32),
            len(A["sl"]), A["w"].ctypes.data_as(_f64), A["pv"].ctypes.data_as(_i32), len(A["pv"]),
            A["pq"].ctypes.data_as(_i32), len(A["pq"]), int(single))


def replay(d, single, fp32, max_iter=10, tol=1e-8, want_J=False, G=1, nslot=1024, lean=0):
    L = ctypes.CDLL(build_replay())
    A = _arrays(d)
    V = np.ascontiguousarray(d["V0"], dtype=np.complex128).copy()
    Jx = np.zeros(4 * len(A["Yi"]) + 2 * A["n"] + 8) if want_J else None
    st, cf = ctypes.c_int(), ctypes.c_int()
    L.replay_solve.restype = ctypes.c_int
    it = L.replay_solve(A["n"], A["Yp"].ctypes.data_as(_i32), A["Yi"].ctypes.data_as(_i32), A["Yx"].ctypes.data_as(_f64),
                        A["S"].ctypes.data_as(_f64), A["sl"].ctypes.data_as(_i32), len(A["sl"]),
                        A["w"].ctypes.data_as(_f64), A["pv"].ctypes.data_as(_i32), len(A["pv"]),
                        A["pq"].ctypes.data_as(_i32), len(A["pq"]), int(single),
                        V.ctypes.data_as(_f64), int(max_iter), ctypes.c_double(tol), int(fp32), int(G), int(nslot), int(lean),
                        ctypes.byref(st), ctypes.byref(cf), Jx.ctypes.data_as(_f64) if want_J else None)
    if want_J:
        return it, st.value, cf.value, V, Jx
    return it, st.value, cf.value, V


def repivot(d, single, Jx, tau):
    """The product's threshold_pivot_rows on J values Jx: rowpos, or None if singular."""
    L = ctypes.CDLL(build_replay())
    A = _arrays(d)
    Jx = np.ascontiguousarray(Jx, dtype=np.float64)
    out = np.zeros(2 * A["n"] + 2, dtype=np.int32)
    N = L.replay_repivot(*_pattern_args(A, single), Jx.ctypes.data_as(_f64), ctypes.c_double(tau),
                         out.ctypes.data_as(_i32))
    return out[:N] if N > 0 else None


def column_order(d, single):
    L = ctypes.CDLL(build_replay())
    A = _arrays(d)
    out = np.zeros(2 * A["n"] + 2, dtype=np.int32)
    N = L.replay_column_order(*_pattern_args(A, single), out.ctypes.data_as(_i32))
    return out[:N]


def set_rowpos(rowpos):
    L = ctypes.CDLL(build_replay())
    if rowpos is None:
        L.replay_set_rowpos(None, 0)
    else:
        r = np.ascontiguousarray(rowpos, dtype=np.int32)
        L.replay_set_rowpos(r.ctypes.data_as(_i32), len(r))


CASES = list(FIXTURES) + ["case14_synth", "case118_synth", "case2000_synth"]


def _data(name):
    if name.endswith("_synth"):
        from lightsim2grid_amd import gridgen
        return gridgen.make(name).solver_inputs()
    return load_fixture(name)


# launch shapes (instances per workgroup G, lanes per instance): 1024 and 256 lanes per workgroup
SHAPES = [(1, 1024), (8, 128), (32, 8)]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"G{s[0]}x{s[1]}")
@pytest.mark.parametrize("lean", [0, 1], ids=["general", "lean"])
def test_replay_matches_oracle(oracle_mod, name, fp32, shape, lean):
    d = _data(name)
    singles = [False, True] if len(np.atleast_1d(d["slack_ids"])) == 1 else [False]
    for single in singles:
        o = oracle_mod.OracleNR(single)
        conv = o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"],
                            10, 1e-8)
        it, st, conflicts, V = replay(d, single, fp32, G=shape[0], nslot=shape[1], lean=lean)
        assert conflicts == 0
        assert (st == 0) == conv
        assert abs(it - o.get_nb_iter()) <= (0 if not fp32 else 1)
        assert np.max(np.abs(V - o.get_V())) <= 1e-8


@pytest.mark.parametrize("single", [False, True])
def test_repivot_on_zero_pivot_matches_oracle(oracle_mod, single):
    """A flat start on a grid with an exactly zero diagonal J block: static diagonal pivots divide
    by zero, the oracle's first factorisation pivots off the diagonal (as klu_factor does) and
    converges. Pivots chosen on the flat-start J by threshold_pivot_rows make the plan replay
    converge in the oracle's iterations to its V."""
    from conftest import zero_diagonal_grid
    d = zero_diagonal_grid()
    o = oracle_mod.OracleNR(single)
    assert o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    set_rowpos(None)
    it, st, _, _ = replay(d, single, 0)
    assert st != 0  # the diagonal pivot is zero
    _, _, _, _, Jx = replay(d, single, 0, max_iter=1, want_J=True)
    rowpos = repivot(d, single, Jx, 1e-3)
    N = len(rowpos)
    # tau = 0 (the oracle's pivot tolerance): off the diagonal only where the diagonal is zero
    r0 = repivot(d, single, Jx, 0.0)
    assert sorted(r0) == list(range(N)) and np.any(r0 != np.arange(N))
    assert sorted(rowpos) == list(range(N)) and np.any(rowpos != np.arange(N))
    set_rowpos(rowpos)
    try:
        for fp32 in (0, 1):
            for shape, lean in (((1, 1024), 0), ((8, 128), 1)):
                it, st, cf, V = replay(d, single, fp32, G=shape[0], nslot=shape[1], lean=lean)
                assert cf == 0 and st == 0
                assert abs(it - o.get_nb_iter()) <= (0 if not fp32 else 1)
                assert np.max(np.abs(V - o.get_V())) <= 1e-8
    finally:
        set_rowpos(None)


@pytest.mark.parametrize("name", ["case14_synth", "case118_synth"])
def test_threshold_pivoting_equals_lapack_partial_pivoting(oracle_mod, name):
    """With tau = 1 the threshold rule is plain partial pivoting: on a J with distinct magnitudes
    its pivot rows equal LAPACK getrf's on the same columns in the same order. With tau = 0 every
    nonzero diagonal pivot is kept. A zero column is reported singular."""
    import scipy.linalg as sla
    d = _data(name)
    o = oracle_mod.OracleNR(False)
    assert o.compute_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"], 10, 1e-8)
    J = sp.csc_matrix(o.get_J())
    J.sort_indices()
    rng = np.random.default_rng(3)
    Jx = J.data * (1.0 + 0.25 * rng.random(J.nnz))
    order = column_order(d, False)
    N = J.shape[0]
    A = sp.csc_matrix((Jx, J.indices, J.indptr), shape=J.shape).toarray()[:, order]
    _, piv = sla.lu_factor(A)
    rows = np.arange(N)
    for k, p in enumerate(piv):
        rows[[k, p]] = rows[[p, k]]
    expect = np.empty(N, dtype=np.int64)
    expect[rows] = order
    assert np.array_equal(repivot(d, False, Jx, 1.0), expect)
    assert np.array_equal(repivot(d, False, Jx, 0.0), np.arange(N))
    Jz = Jx.copy()
    Jz[J.indptr[5]:J.indptr[6]] = 0.0
    assert repivot(d, False, Jz, 1e-3) is None


def _dc_inputs(d):
    """The DC layout of a grid (as the DC grid load builds it): one slack bus, every other bus an
    angle unknown."""
    n = d["Ybus"].shape[0]
    sl = int(np.atleast_1d(d["slack_ids"])[0])
    ang = np.array([b for b in range(n) if b != sl], np.int32)
    return dict(Ybus=d["Ybus"], V0=d["V0"], Sbus=d["Sbus"], pv=ang, pq=np.zeros(0, np.int32),
                slack_ids=np.array([sl], np.int32), slack_weights=np.eye(n)[sl]), sl, ang


def _nopivot_solve(M, b):
    """Dense LU without pivoting, then the two triangular solves (what a plan with static pivots
    computes, in the plan's elimination order)."""
    U = M.copy()
    y = b.copy()
    N = len(b)
    for k in range(N):
        U[k + 1:, k] /= U[k, k]
        U[k + 1:, k + 1:] -= np.outer(U[k + 1:, k], U[k, k + 1:])
        y[k + 1:] -= U[k + 1:, k] * y[k]
    x = np.zeros(N)
    for k in range(N - 1, -1, -1):
        x[k] = (y[k] - U[k, k + 1:] @ x[k + 1:]) / U[k, k]
    return x


@pytest.mark.parametrize("name", ["case14_synth", "case118_synth"])
def test_row_permuted_plans_compute_their_static_lu(name):
    """Plans built on row permutations (pivots chosen on perturbed J values, so that many rows
    move and pivots can be created by fill alone) compute exactly the unpivoted LU of the permuted
    matrix in the plan's order: one DC solve of the plan replay equals the dense one."""
    import ctypes as C
    from lightsim2grid_amd import gridgen
    L = C.CDLL(build_replay())
    g = gridgen.make(name)
    d0 = g.solver_inputs()
    y = 1.0 / g.x
    Y = sp.csc_matrix((np.r_[y, y, -y, -y].astype(np.complex128),
                       (np.r_[g.f, g.t, g.f, g.t], np.r_[g.f, g.t, g.t, g.f])), shape=(g.n, g.n))
    Y.sum_duplicates()
    Y.sort_indices()
    d0 = dict(d0, Ybus=Y, V0=d0["V0"] * np.exp(1j * 0.05))
    dd, sl, ang = _dc_inputs(d0)
    A = Y.real.toarray()[np.ix_(ang, ang)]
    order = column_order(dd, True)
    rng = np.random.default_rng(0)
    L.replay_set_dc(1)
    set_rowpos(None)
    try:
        _, _, _, _, Jx = replay(dd, True, 0, max_iter=1, want_J=True)
        moved = 0
        for trial in range(6):
            rp = repivot(dd, True, Jx * (1 + 3 * rng.random(len(Jx))), 1.0)
            moved = max(moved, int(np.sum(rp != np.arange(len(rp)))))
            set_rowpos(rp)
            for lean in (0, 1):
                _, st, cf, V = replay(dd, True, 0, max_iter=1, lean=lean)
                assert cf == 0
                M = np.zeros_like(A)
                M[rp] = A
                b = np.zeros(len(ang))
                b[rp] = d0["Sbus"].real[ang]  # theta0 - theta_slack = 0: F = -P, theta = theta_slack + M^-1 P
                x = np.zeros(len(ang))
                x[order] = _nopivot_solve(M[np.ix_(order, order)], b[order])
                th = np.full(g.n, 0.05)
                th[ang] += x
                assert np.max(np.abs(V - np.abs(d0["V0"]) * np.exp(1j * th))) <= 1e-9
            set_rowpos(None)
        assert moved >= 5
    finally:
        set_rowpos(None)
        L.replay_set_dc(0)


def test_repivot_dc_zero_diagonal(oracle_mod):
    """The DC form of the zero-diagonal grid: the static plan's solve is not finite, the
    re-pivoted plan's equals the oracle's restatement of BaseDCAlgo<KLU>."""
    import ctypes as C
    from conftest import zero_diagonal_grid
    L = C.CDLL(build_replay())
    d = zero_diagonal_grid(dc=True)
    d["V0"] = d["V0"] * np.exp(1j * 0.05)
    Vr, err = oracle_mod.dc_pf(d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["pv"], d["pq"])
    assert err == 0
    dd, _, _ = _dc_inputs(d)
    L.replay_set_dc(1)
    set_rowpos(None)
    try:
        _, st, _, _, Jx = replay(dd, True, 0, max_iter=1, want_J=True)
        assert st == 7
        rp = repivot(dd, True, Jx, 1e-3)
        set_rowpos(rp)
        for lean in (0, 1):
            _, st, cf, V = replay(dd, True, 0, max_iter=1, lean=lean)
            assert cf == 0 and st == 0
            assert np.max(np.abs(V - Vr)) <= 1e-12
    finally:
        set_rowpos(None)
        L.replay_set_dc(0)
