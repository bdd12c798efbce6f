"""ctypes front-end of the C++ FP64 oracle (test infrastructure only).

The oracle restates lightsim2grid's KLU Newton-Raphson path on the CPU (see
the header of ``nr_oracle.cpp`` for the reference file:line it follows). Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use
this module, always as the checker / baseline, never as the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile the oracle with the committed Makefile (g++ only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "_build/liboracle.so"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_solver_create.restype = ctypes.c_void_p
        L.oracle_solver_create.argtypes = [ctypes.c_int]
        L.oracle_solver_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_solver_reset.argtypes = [ctypes.c_void_p]
        L.oracle_solver_compute_pf.restype = ctypes.c_int
        L.oracle_solver_compute_pf.argtypes = [
            ctypes.c_void_p, ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _i32p, ctypes.c_int,
            _f64p, _i32p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        for nm in ("oracle_solver_nb_iter", "oracle_solver_error", "oracle_solver_J_dim", "oracle_solver_J_nnz"):
            getattr(L, nm).restype = ctypes.c_int
            getattr(L, nm).argtypes = [ctypes.c_void_p]
        L.oracle_solver_nnz_lu.restype = ctypes.c_long
        L.oracle_solver_nnz_lu.argtypes = [ctypes.c_void_p]
        L.oracle_solver_get_J.argtypes = [ctypes.c_void_p, _i32p, _i32p, _f64p]
        L.oracle_solver_get_timers.argtypes = [ctypes.c_void_p, _f64p]
        L.oracle_solver_get_lu_dense.restype = ctypes.c_int
        L.oracle_solver_get_lu_dense.argtypes = [ctypes.c_void_p, _i32p, _i32p, _f64p, _f64p]
        L.oracle_lu_solve.restype = ctypes.c_int
        L.oracle_lu_solve.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, ctypes.c_int]
        L.oracle_contingency.restype = ctypes.c_double
        L.oracle_contingency.argtypes = [
            ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _i32p, ctypes.c_int, _f64p, _i32p, ctypes.c_int,
            _i32p, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p,
            ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p, _i32p, _i32p, _i32p, _i32p]
        L.oracle_time_series.restype = ctypes.c_double
        L.oracle_time_series.argtypes = [
            ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_int, _f64p, _i32p, ctypes.c_int, _f64p,
            _i32p, ctypes.c_int, _i32p, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
            ctypes.c_int, ctypes.c_int, _f64p, _i32p, _i32p, _i32p, _i32p]
        L.oracle_dc.restype = ctypes.c_int
        L.oracle_dc.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _i32p, ctypes.c_int, _i32p,
                                ctypes.c_int, _i32p, ctypes.c_int]
        _lib = L
    return _lib


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_f64p)


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(_f64p)


def _csc(Y):
    Y = sp.csc_matrix(Y)
    Y.sort_indices()
    return Y


class OracleNR:
    """Oracle NR solver with the reference solver-class surface
    (compute_pf / get_V / get_Va / get_Vm / get_J / get_nb_iter / get_error /
    converged / reset / get_timers), cf. src/main.cpp:228-240."""

    def __init__(self, single_slack: bool = False):
        self._L = lib()
        self._h = self._L.oracle_solver_create(1 if single_slack else 0)
        self._V = None
        self._conv = False
        self.single_slack = single_slack

    def __del__(self):
        try:
            self._L.oracle_solver_destroy(self._h)
        except Exception:
            pass

    def reset(self):
        self._L.oracle_solver_reset(self._h)

    def compute_pf(self, Ybus, V0, Sbus, slack_ids, slack_weights, pv, pq, max_iter, tol):
        Y = _csc(Ybus)
        n = Y.shape[0]
        Yp, pYp = _i(Y.indptr)
        Yi, pYi = _i(Y.indices)
        Yx, pYx = _c(Y.data)
        V = np.array(V0, dtype=np.complex128, copy=True)
        pV = V.ctypes.data_as(_f64p)
        S, pS = _c(Sbus)
        sl, psl = _i(np.atleast_1d(slack_ids))
        w, pw = _d(slack_weights)
        pv_, ppv = _i(pv)
        pq_, ppq = _i(pq)
        ok = self._L.oracle_solver_compute_pf(self._h, n, pYp, pYi, pYx, pV, pS, psl, len(sl), pw,
                                              ppv, len(pv_), ppq, len(pq_), int(max_iter), float(tol))
        self._V = V
        self._conv = bool(ok)
        return self._conv

    solve = compute_pf

    def get_V(self):
        return self._V.copy()

    def get_Va(self):
        return np.angle(self._V)

    def get_Vm(self):
        return np.abs(self._V)

    def get_nb_iter(self):
        return self._L.oracle_solver_nb_iter(self._h)

    def get_error(self):
        return self._L.oracle_solver_error(self._h)

    def converged(self):
        return self._conv

    def get_timers(self):
        out = np.zeros(10)
        self._L.oracle_solver_get_timers(self._h, out.ctypes.data_as(_f64p))
        return tuple(out)

    def nnz_lu(self):
        return int(self._L.oracle_solver_nnz_lu(self._h))

    def get_J(self):
        n = self._L.oracle_solver_J_dim(self._h)
        if n <= 0:
            return sp.csc_matrix((0, 0))
        nnz = self._L.oracle_solver_J_nnz(self._h)
        Jp = np.zeros(n + 1, np.int32)
        Ji = np.zeros(nnz, np.int32)
        Jx = np.zeros(nnz)
        self._L.oracle_solver_get_J(self._h, Jp.ctypes.data_as(_i32p), Ji.ctypes.data_as(_i32p),
                                    Jx.ctypes.data_as(_f64p))
        return sp.csc_matrix((Jx, Ji, Jp), shape=(n, n))

    def get_lu_dense(self):
        n = self._L.oracle_solver_J_dim(self._h)
        P = np.zeros(n, np.int32)
        Q = np.zeros(n, np.int32)
        Lm = np.zeros((n, n))
        Um = np.zeros((n, n))
        r = self._L.oracle_solver_get_lu_dense(self._h, P.ctypes.data_as(_i32p), Q.ctypes.data_as(_i32p),
                                               Lm.ctypes.data_as(_f64p), Um.ctypes.data_as(_f64p))
        if r < 0:
            raise RuntimeError("oracle LU not factored")
        return P, Q, Lm, Um


def lu_solve(A, b, reps: int = 1):
    """Factor the real CSC matrix A with the oracle's KLU-contract LU and solve A x = b."""
    A = _csc(A)
    n = A.shape[0]
    Ap, pAp = _i(A.indptr)
    Ai, pAi = _i(A.indices)
    Ax, pAx = _d(A.data)
    x = np.array(b, dtype=np.float64, copy=True)
    err = lib().oracle_lu_solve(n, pAp, pAi, pAx, x.ctypes.data_as(_f64p), int(reps))
    return x, err


def contingency(Ybus, Sbus, slack_ids, slack_weights, pv, pq, Vinit, deltas, max_iter=10, tol=1e-8,
                init_from_n=True, single_slack=False, nthreads=1):
    """Contingency sweep on the oracle. ``deltas`` is (dptr, dpos, dval): for
    contingency c, Ybus.data[dpos[dptr[c]:dptr[c+1]]] -= dval[...] (CSC value
    positions of the sorted-index CSC Ybus). Returns (V, iters, errs, conv,
    connected, wall_seconds)."""
    Y = _csc(Ybus)
    n = Y.shape[0]
    dptr, dpos, dval = deltas
    ncont = len(dptr) - 1
    Yp, pYp = _i(Y.indptr)
    Yi, pYi = _i(Y.indices)
    Yx, pYx = _c(Y.data)
    S, pS = _c(Sbus)
    sl, psl = _i(np.atleast_1d(slack_ids))
    w, pw = _d(slack_weights)
    pv_, ppv = _i(pv)
    pq_, ppq = _i(pq)
    V0, pV0 = _c(Vinit)
    dp, pdp = _i(dptr)
    dq, pdq = _i(dpos)
    dv, pdv = _c(dval)
    Vout = np.zeros((ncont, n), np.complex128)
    iters = np.zeros(ncont, np.int32)
    errs = np.zeros(ncont, np.int32)
    conv = np.zeros(ncont, np.int32)
    conn = np.zeros(ncont, np.int32)
    wall = lib().oracle_contingency(n, pYp, pYi, pYx, pS, psl, len(sl), pw, ppv, len(pv_), ppq, len(pq_),
                                    pV0, int(init_from_n), int(single_slack), ncont, pdp, pdq, pdv,
                                    int(max_iter), float(tol), int(nthreads), Vout.ctypes.data_as(_f64p),
                                    iters.ctypes.data_as(_i32p), errs.ctypes.data_as(_i32p),
                                    conv.ctypes.data_as(_i32p), conn.ctypes.data_as(_i32p))
    return Vout, iters, errs, conv.astype(bool), conn.astype(bool), wall


def time_series(Ybus, Sbuses, slack_ids, slack_weights, pv, pq, Vinit, max_iter=10, tol=1e-8,
                single_slack=False, nchains=1, nthreads=0):
    """Time series on the oracle (warm start from the previous step inside each
    chain; nchains chains of ceil(T / nchains) steps on nthreads threads, 0: one per
    chain). Returns (V, iters, errs, conv, status, wall_seconds)."""
    Y = _csc(Ybus)
    n = Y.shape[0]
    Sb, pSb = _c(Sbuses)
    T = Sb.shape[0]
    Yp, pYp = _i(Y.indptr)
    Yi, pYi = _i(Y.indices)
    Yx, pYx = _c(Y.data)
    sl, psl = _i(np.atleast_1d(slack_ids))
    w, pw = _d(slack_weights)
    pv_, ppv = _i(pv)
    pq_, ppq = _i(pq)
    V0, pV0 = _c(Vinit)
    Vout = np.zeros((T, n), np.complex128)
    iters = np.zeros(T, np.int32)
    errs = np.zeros(T, np.int32)
    conv = np.zeros(T, np.int32)
    status = np.zeros(1, np.int32)
    wall = lib().oracle_time_series(n, pYp, pYi, pYx, T, pSb, psl, len(sl), pw, ppv, len(pv_), ppq, len(pq_),
                                    pV0, int(single_slack), int(max_iter), float(tol), int(nchains), int(nthreads),
                                    Vout.ctypes.data_as(_f64p), iters.ctypes.data_as(_i32p),
                                    errs.ctypes.data_as(_i32p), conv.ctypes.data_as(_i32p),
                                    status.ctypes.data_as(_i32p))
    return Vout, iters, errs, conv.astype(bool), int(status[0]), wall


def dc_pf(Ybus, V0, Sbus, slack_ids, pv, pq):
    """DC power flow on the oracle (BaseDCAlgo<KLU>::compute_pf). Returns (V, error)."""
    Y = _csc(Ybus)
    n = Y.shape[0]
    Yp, pYp = _i(Y.indptr)
    Yi, pYi = _i(Y.indices)
    Yx, pYx = _c(Y.data)
    V = np.array(V0, dtype=np.complex128, copy=True)
    S, pS = _c(Sbus)
    sl, psl = _i(np.atleast_1d(slack_ids))
    pv_, ppv = _i(pv)
    pq_, ppq = _i(pq)
    err = lib().oracle_dc(n, pYp, pYi, pYx, V.ctypes.data_as(_f64p), pS, psl, len(sl), ppv, len(pv_), ppq, len(pq_))
    return V, int(err)


def fdpf_matrices(n, f, t, r, x, h, method, shunt_bus=(), shunt_q_mvar=(), sn_mva=1.0):
    """grid_Bp and grid_Bpp (n x n CSC) of GridModel::fillBp_Bpp (GridModel.cpp:1037-1060) for a
    grid of lines (total charging h, split half and half between the two ends, LineContainer.cpp:
    13-34) and shunts: per line, TwoSidesContainer_rxh_A::get_fdpf_coeffs (TwoSidesContainer_rxh_A
    .hpp:818-843): ys_bp = 1/(jx) (XB) or 1/(r+jx) (BX), ys_bpp the other one; Bp gets -Im(ys_bp)
    on both diagonals and +Im(ys_bp) off them, Bpp -(Im(ys_bpp) + Im(h)/2) on the diagonals and
    +Im(ys_bpp) off them (fillBp_Bpp, :503-598); a shunt adds q_mvar / sn_mva to its Bpp diagonal
    (ShuntContainer.cpp:64-95). method: "XB" or "BX"."""
    f, t = np.asarray(f, int), np.asarray(t, int)
    r, x, h = np.asarray(r, float), np.asarray(x, float), np.asarray(h, complex)
    y_jx = (1.0 / (1j * x)).imag
    y_rx = (1.0 / (r + 1j * x)).imag
    ybp, ybpp = (y_jx, y_rx) if method == "XB" else (y_rx, y_jx)
    half = np.imag(h) * 0.5
    rows = np.r_[f, t, f, t]
    cols = np.r_[f, t, t, f]
    Bp = sp.csc_matrix((np.r_[-ybp, -ybp, ybp, ybp], (rows, cols)), shape=(n, n))
    Bpp = sp.csc_matrix((np.r_[-(ybpp + half), -(ybpp + half), ybpp, ybpp], (rows, cols)), shape=(n, n))
    if len(shunt_bus):
        sb = np.asarray(shunt_bus, int)
        Bpp = Bpp + sp.csc_matrix((np.asarray(shunt_q_mvar, float) / sn_mva, (sb, sb)), shape=(n, n))
    Bp.sum_duplicates()
    Bpp.sum_duplicates()
    Bp.sort_indices()
    Bpp.sort_indices()
    return Bp, Bpp


def fdpf(Ybus, V0, Sbus, slack_ids, slack_weights, pv, pq, Bp, Bpp, max_iter=10, tol=1e-8):
    """Fast-decoupled power flow on the oracle (BaseFDPFAlgo<KLU>::compute_pf).
    Returns (V, iterations, error)."""
    L = lib()
    if not getattr(L, "_fdpf_typed", False):
        L.oracle_fdpf.restype = ctypes.c_int
        L.oracle_fdpf.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _i32p, ctypes.c_int, _f64p,
                                  _i32p, ctypes.c_int, _i32p, ctypes.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p,
                                  ctypes.c_int, ctypes.c_double, _i32p]
        L._fdpf_typed = True
    Y = _csc(Ybus)
    n = Y.shape[0]
    Yp, pYp = _i(Y.indptr)
    Yi, pYi = _i(Y.indices)
    Yx, pYx = _c(Y.data)
    V = np.array(V0, dtype=np.complex128, copy=True)
    S, pS = _c(Sbus)
    sl, psl = _i(np.atleast_1d(slack_ids))
    w, pw = _d(slack_weights)
    pv_, ppv = _i(pv)
    pq_, ppq = _i(pq)
    Bp, Bpp = sp.csc_matrix(Bp), sp.csc_matrix(Bpp)
    Bp.sort_indices()
    Bpp.sort_indices()
    a1, pa1 = _i(Bp.indptr)
    a2, pa2 = _i(Bp.indices)
    a3, pa3 = _d(Bp.data)
    b1, pb1 = _i(Bpp.indptr)
    b2, pb2 = _i(Bpp.indices)
    b3, pb3 = _d(Bpp.data)
    it = np.zeros(1, np.int32)
    err = L.oracle_fdpf(n, pYp, pYi, pYx, V.ctypes.data_as(_f64p), pS, psl, len(sl), pw, ppv, len(pv_), ppq, len(pq_),
                        pa1, pa2, pa3, pb1, pb2, pb3, int(max_iter), float(tol), it.ctypes.data_as(_i32p))
    return V, int(it[0]), int(err)
