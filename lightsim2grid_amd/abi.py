"""ctypes front-end of the pfgpu C ABI (include/pfgpu.h).

This is a thin, typed wrapper: every call goes to libpfgpu.so (HIP kernels).
There is no CPU path behind it; if the library is missing or no GPU is
present the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfgpu.so")

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

FP64 = 0
FP32 = 1          # FP32 factors; rows that stall or break down are re-solved with FP64 factors
FP32_STRICT = 2   # FP32 factors, no re-solve

FLAG_STALL = 1
FLAG_GROWTH = 2
STATUS_NOT_RUN = 8

ERROR_NAMES = {0: "NoError", 1: "SingularMatrix", 2: "TooManyIterations", 3: "InifiniteValue", 4: "SolverAnalyze",
               5: "SolverFactor", 6: "SolverReFactor", 7: "SolverSolve", 8: "NotInitError", 9: "LicenseError"}


class GridInfo(ctypes.Structure):
    _fields_ = [("n_bus", ctypes.c_int), ("n_jac", ctypes.c_int), ("nnz_ybus", ctypes.c_int),
                ("nnz_jac", ctypes.c_int), ("nnz_lu", ctypes.c_int64), ("n_slots", ctypes.c_int),
                ("n_levels", ctypes.c_int), ("n_back_levels", ctypes.c_int), ("n_pairs", ctypes.c_int64),
                ("flops", ctypes.c_double), ("single_slack", ctypes.c_int), ("dc", ctypes.c_int),
                ("fdpf", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("total_s", ctypes.c_double), ("h2d_s", ctypes.c_double), ("kernel_s", ctypes.c_double),
                ("d2h_s", ctypes.c_double), ("instances_per_wg", ctypes.c_int), ("lds_slots", ctypes.c_int),
                ("workgroups", ctypes.c_int), ("fallback_fp64", ctypes.c_int), ("phase_s", ctypes.c_double * 6),
                ("block_size", ctypes.c_int), ("resident_workgroups", ctypes.c_int),
                ("graph_launches", ctypes.c_int), ("repivot_rows", ctypes.c_int)]

    PHASES = ("jfill", "factor", "tail", "backward", "update", "mismatch")

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "phase_s"}
        d["phase_s"] = dict(zip(self.PHASES, list(self.phase_s)))
        return d


class RequestInfo(ctypes.Structure):
    _fields_ = [("ninst", ctypes.c_int), ("launch_instances", ctypes.c_int), ("latency_s", ctypes.c_double),
                ("queue_s", ctypes.c_double)]


REQ_PF, REQ_CONTINGENCIES, REQ_TIME_SERIES = 0, 1, 2

# exported symbols and their signatures (restype, argtypes)
SIGNATURES = {
    "pfgpu_last_error": (ctypes.c_char_p, []),
    "pfgpu_abi_version": (ctypes.c_int, []),
    "pfgpu_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "pfgpu_ctx_destroy": (ctypes.c_int, [_vp]),
    "pfgpu_ctx_device_info": (ctypes.c_int, [_vp, _i32p, _i32p, ctypes.POINTER(ctypes.c_size_t)]),
    "pfgpu_grid_load": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int, _f64p, _i32p,
                                       ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "pfgpu_grid_load_dc": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int, _i32p, ctypes.c_int,
                                          _i32p, ctypes.c_int, ctypes.POINTER(_vp)]),
    "pfgpu_grid_load_fdpf": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int, _f64p, _i32p,
                                            ctypes.c_int, _i32p, ctypes.c_int, _i32p, _i32p, _f64p, _i32p, _i32p,
                                            _f64p, ctypes.POINTER(_vp)]),
    "pfgpu_grid_destroy": (ctypes.c_int, [_vp]),
    "pfgpu_grid_reset": (ctypes.c_int, [_vp]),
    "pfgpu_grid_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(GridInfo)]),
    "pfgpu_grid_get_J_pattern": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "pfgpu_solve_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _f64p, ctypes.c_int, _i32p, _i32p, _f64p, _f64p,
                                         ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, ctypes.POINTER(Stats)]),
    "pfgpu_solve_batch_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                                ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_double,
                                                ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.POINTER(Stats)]),
    "pfgpu_solve_batch_device_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp,
                                                        _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int,
                                                        ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                                        ctypes.POINTER(Stats)]),
    "pfgpu_solve_chains": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p, ctypes.c_int,
                                          _f64p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _i32p,
                                          _i32p, ctypes.POINTER(Stats)]),
    "pfgpu_solve_chains_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp,
                                                 ctypes.c_int, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                 ctypes.c_int, _vp, _vp, ctypes.POINTER(Stats)]),
    "pfgpu_ctx_last_flags": (ctypes.c_int, [_vp, _i32p, ctypes.c_longlong]),
    "pfgpu_autotune_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                             ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, _vp, _vp, _i32p, _i32p, _f64p]),
    "pfgpu_branch_flows_device": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp,
                                                 _vp, _vp, ctypes.c_double, _vp, _vp]),
    "pfgpu_batch_flows_device": (ctypes.c_int, [_vp, ctypes.c_longlong, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp,
                                                _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp]),
    "pfgpu_batch_flows_dc_device": (ctypes.c_int, [_vp, ctypes.c_longlong, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                                   _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp]),
    "pfgpu_sbus_rows_device": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp,
                                              ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_double,
                                              _vp]),
    "pfgpu_ybus_values_device": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "pfgpu_contingency_edits_device": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pfgpu_connectivity_device": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pfgpu_server_create": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "pfgpu_server_destroy": (ctypes.c_int, [_vp]),
    "pfgpu_server_add_grid": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _i32p,
                                             ctypes.c_int, _f64p, _i32p, ctypes.c_int, _i32p, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    "pfgpu_submit": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _i32p, _i32p, _f64p,
                                    ctypes.POINTER(ctypes.c_uint64)]),
    "pfgpu_poll": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_int), _f64p, _i32p,
                                  _i32p, ctypes.POINTER(RequestInfo)]),
    "pfgpu_dev_alloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "pfgpu_dev_free": (ctypes.c_int, [_vp, _vp]),
    "pfgpu_memcpy_h2d": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "pfgpu_memcpy_d2h": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "pfgpu_memcpy_d2d": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
}

_lib = None


class PfgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pfgpu error {code}: {msg}")
        self.code = code


def lib():
    """Load libpfgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m lightsim2grid_amd.build` "
                              "(the engine has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise PfgpuError(rc, lib().pfgpu_last_error().decode(errors="replace"))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_f64p)


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(_f64p)


class Context:
    """One GPU + one HIP stream + an HBM budget (pfgpu_ctx)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._L = lib()
        h = _vp()
        _check(self._L.pfgpu_ctx_create(int(device), int(hbm_budget), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._L.pfgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_flags(self, nrows: int) -> np.ndarray:
        """PFGPU_FLAG_* per row of the last solve call on this context."""
        f = np.zeros(int(nrows), np.int32)
        _check(self._L.pfgpu_ctx_last_flags(self.h, f.ctypes.data_as(_i32p), int(nrows)))
        return f

    def device_info(self):
        cu = ctypes.c_int()
        lds = ctypes.c_int()
        fr = ctypes.c_size_t()
        _check(self._L.pfgpu_ctx_device_info(self.h, ctypes.byref(cu), ctypes.byref(lds), ctypes.byref(fr)))
        return {"n_cu": cu.value, "lds_bytes": lds.value, "free_bytes": fr.value}

    # device buffers -----------------------------------------------------
    def to_device(self, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = DeviceArray(self, arr.nbytes, arr.dtype, arr.shape)
        _check(self._L.pfgpu_memcpy_h2d(self.h, d.ptr, arr.ctypes.data_as(_vp), arr.nbytes))
        return d

    def empty(self, shape, dtype) -> "DeviceArray":
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        return DeviceArray(self, nbytes, dt, shape)


class DeviceArray:
    def __init__(self, ctx: Context, nbytes: int, dtype, shape):
        self.ctx = ctx
        self.nbytes = nbytes
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        p = _vp()
        _check(ctx._L.pfgpu_dev_alloc(ctx.h, max(int(nbytes), 1), ctypes.byref(p)))
        self.ptr = p

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(self.ctx._L.pfgpu_memcpy_d2h(self.ctx.h, out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def copy_from(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        _check(self.ctx._L.pfgpu_memcpy_h2d(self.ctx.h, self.ptr, arr.ctypes.data_as(_vp), arr.nbytes))

    def free(self):
        if self.ptr:
            self.ctx._L.pfgpu_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ctx.h is not None:
                self.free()
        except Exception:
            pass


def _csc(Ybus):
    Y = sp.csc_matrix(Ybus)
    Y.sum_duplicates()
    Y.sort_indices()
    return Y


class Grid:
    """A loaded grid (pfgpu_grid): J pattern, ordering, symbolic LU and schedule on the device.
    dc=True loads the DC power flow of the same buses (pfgpu_grid_load_dc; slack_weights unused);
    fdpf=(Bp, Bpp) the fast-decoupled power flow with these n x n matrices (pfgpu_grid_load_fdpf)."""

    def __init__(self, ctx: Context, Ybus, slack_ids, slack_weights, pv, pq, single_slack: bool = False,
                 dc: bool = False, fdpf=None):
        self.ctx = ctx
        Y = _csc(Ybus)
        self.n = Y.shape[0]
        self.Yp, pYp = _i(Y.indptr)
        self.Yi, pYi = _i(Y.indices)
        sl, psl = _i(np.atleast_1d(slack_ids))
        pv_, ppv = _i(pv)
        pq_, ppq = _i(pq)
        h = _vp()
        if dc:
            _check(ctx._L.pfgpu_grid_load_dc(ctx.h, self.n, pYp, pYi, psl, len(sl), ppv, len(pv_), ppq, len(pq_),
                                             ctypes.byref(h)))
        elif fdpf is not None:
            w, pw = _d(slack_weights)
            Bp, Bpp = _csc(fdpf[0]), _csc(fdpf[1])
            keep = [_i(Bp.indptr), _i(Bp.indices), _d(Bp.data), _i(Bpp.indptr), _i(Bpp.indices), _d(Bpp.data)]
            _check(ctx._L.pfgpu_grid_load_fdpf(ctx.h, self.n, pYp, pYi, psl, len(sl), pw, ppv, len(pv_), ppq,
                                               len(pq_), *[k[1] for k in keep], ctypes.byref(h)))
        else:
            w, pw = _d(slack_weights)
            _check(ctx._L.pfgpu_grid_load(ctx.h, self.n, pYp, pYi, psl, len(sl), pw, ppv, len(pv_), ppq, len(pq_),
                                          1 if single_slack else 0, ctypes.byref(h)))
        self.h = h
        self.single_slack = single_slack
        info = GridInfo()
        _check(ctx._L.pfgpu_grid_get_info(self.h, ctypes.byref(info)))
        self.info = info.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.ctx._L.pfgpu_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        """The next solve factors afresh (pfgpu_grid_reset)."""
        _check(self.ctx._L.pfgpu_grid_reset(self.h))

    def solve_chains(self, Yx, Sbus_rows, V0, chain_len=None, max_iter=10, tol=1e-8, precision=FP64,
                     instances_per_wg=0):
        """Time series: Sbus_rows (T, n); chains of chain_len warm-started steps (None: one chain).
        V0: (n,) shared head, or (nchains, n). Returns dict(V, iters, status, stats)."""
        Yx = np.ascontiguousarray(Yx, np.complex128)
        S = np.ascontiguousarray(Sbus_rows, np.complex128)
        T = S.shape[0]
        L = T if chain_len is None else int(chain_len)
        V0 = np.ascontiguousarray(V0, np.complex128)
        V = np.zeros((T, self.n), np.complex128)
        iters = np.zeros(T, np.int32)
        status = np.zeros(T, np.int32)
        st = Stats()
        _check(self.ctx._L.pfgpu_solve_chains(
            self.ctx.h, self.h, int(T), int(max(L, 1)), Yx.ctypes.data_as(_f64p), S.ctypes.data_as(_f64p),
            V0.ctypes.data_as(_f64p), self.n if V0.ndim == 2 else 0, V.ctypes.data_as(_f64p), int(max_iter),
            float(tol), int(precision), int(instances_per_wg), iters.ctypes.data_as(_i32p),
            status.ctypes.data_as(_i32p), ctypes.byref(st)))
        return {"V": V, "iters": iters, "status": status, "stats": st.as_dict()}

    def solve_chains_device(self, d_Yx, yx_stride, d_S, d_V0, v0_stride, d_V, nrows, chain_len, d_iters, d_status,
                            max_iter=10, tol=1e-8, precision=FP64, instances_per_wg=0):
        st = Stats()
        _check(self.ctx._L.pfgpu_solve_chains_device(
            self.ctx.h, self.h, int(nrows), int(chain_len), d_Yx.ptr, int(yx_stride), d_S.ptr, d_V0.ptr,
            int(v0_stride), d_V.ptr, int(max_iter), float(tol), int(precision), int(instances_per_wg), d_iters.ptr,
            d_status.ptr, ctypes.byref(st)))
        return st.as_dict()

    def J_pattern(self):
        N, nnz = self.info["n_jac"], self.info["nnz_jac"]
        Jp = np.zeros(N + 1, np.int32)
        Ji = np.zeros(nnz, np.int32)
        _check(self.ctx._L.pfgpu_grid_get_J_pattern(self.h, Jp.ctypes.data_as(_i32p), Ji.ctypes.data_as(_i32p)))
        return Jp, Ji

    def ybus_values(self, Ybus) -> np.ndarray:
        """Values of Ybus in this grid's CSC order (pattern must match)."""
        Y = _csc(Ybus)
        if not (np.array_equal(Y.indptr, self.Yp) and np.array_equal(Y.indices, self.Yi)):
            raise ValueError("Ybus pattern differs from the loaded grid")
        return np.ascontiguousarray(Y.data, np.complex128)

    def solve(self, Yx, Sbus, V0, max_iter=10, tol=1e-8, precision=FP64, deltas=None, instances_per_wg=0,
              want_J=False):
        """Batched solve with host arrays.

        Yx: (nnz,) shared values or (ninst, nnz) per instance; Sbus: (n,) or (ninst, n);
        V0: (n,) or (ninst, n); deltas: optional (dptr, dpos, dval) Ybus edits per instance.
        Returns dict(V, iters, status, stats[, Jx])."""
        Yx = np.ascontiguousarray(Yx, np.complex128)
        Sbus = np.ascontiguousarray(Sbus, np.complex128)
        V0 = np.ascontiguousarray(V0, np.complex128)
        ninst = 1
        for arr in (Yx if Yx.ndim == 2 else None, Sbus if Sbus.ndim == 2 else None, V0 if V0.ndim == 2 else None):
            if arr is not None:
                ninst = max(ninst, arr.shape[0])
        if deltas is not None:
            ninst = max(ninst, len(deltas[0]) - 1)
        n = self.n
        V = np.empty((ninst, n), np.complex128)
        if V0.ndim == 2:
            V[:] = V0
            v0s = n
        else:
            V[0] = V0
            v0s = 0
        iters = np.zeros(ninst, np.int32)
        status = np.zeros(ninst, np.int32)
        Jx = np.zeros(self.info["nnz_jac"]) if want_J else None
        st = Stats()
        dp = dq = dv = None
        if deltas is not None:
            dp, _ = _i(deltas[0])
            dq, _ = _i(deltas[1])
            dv, _ = _c(deltas[2])
        _check(self.ctx._L.pfgpu_solve_batch(
            self.ctx.h, self.h, ninst, Yx.ctypes.data_as(_f64p), self.info["nnz_ybus"] if Yx.ndim == 2 else 0,
            dp.ctypes.data_as(_i32p) if dp is not None else None, dq.ctypes.data_as(_i32p) if dq is not None else None,
            dv.ctypes.data_as(_f64p) if dv is not None else None, Sbus.ctypes.data_as(_f64p),
            n if Sbus.ndim == 2 else 0, V.ctypes.data_as(_f64p), v0s, int(max_iter), float(tol), int(precision),
            int(instances_per_wg), iters.ctypes.data_as(_i32p), status.ctypes.data_as(_i32p),
            Jx.ctypes.data_as(_f64p) if Jx is not None else None, ctypes.byref(st)))
        out = {"V": V, "iters": iters, "status": status, "stats": st.as_dict()}
        if want_J:
            Jp, Ji = self.J_pattern()
            N = self.info["n_jac"]
            out["J"] = sp.csc_matrix((Jx, Ji, Jp), shape=(N, N))
        return out

    def autotune_device(self, d_Yx, yx_stride, d_S, s_stride, d_V0, v0_stride, d_V, ninst, d_iters, d_status,
                        max_iter=10, tol=1e-8, precision=FP64, deltas=None):
        """Time the launch shapes on this batch; the fastest becomes the default. Returns (G, block, seconds)."""
        bG, bNT, bt = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        dp, dq, dv = deltas if deltas is not None else (None, None, None)
        _check(self.ctx._L.pfgpu_autotune_device(
            self.ctx.h, self.h, int(ninst), d_Yx.ptr, int(yx_stride), dp.ptr if dp else None, dq.ptr if dq else None,
            dv.ptr if dv else None, d_S.ptr, int(s_stride), d_V0.ptr, int(v0_stride), d_V.ptr, int(max_iter),
            float(tol), int(precision), d_iters.ptr, d_status.ptr, ctypes.byref(bG), ctypes.byref(bNT), ctypes.byref(bt)))
        return bG.value, bNT.value, bt.value

    def solve_device(self, d_Yx, yx_stride, d_S, s_stride, d_V0, v0_stride, d_V, ninst, d_iters, d_status,
                     max_iter=10, tol=1e-8, precision=FP64, deltas=None, instances_per_wg=0, noalloc=False):
        """Batched solve on device-resident arrays (DeviceArray); returns the stats dict.
        noalloc: the entry point that allocates nothing (pfgpu_solve_batch_device_noalloc)."""
        st = Stats()
        dp, dq, dv = deltas if deltas is not None else (None, None, None)
        fn = self.ctx._L.pfgpu_solve_batch_device_noalloc if noalloc else self.ctx._L.pfgpu_solve_batch_device
        _check(fn(
            self.ctx.h, self.h, int(ninst), d_Yx.ptr, int(yx_stride), dp.ptr if dp else None, dq.ptr if dq else None,
            dv.ptr if dv else None, d_S.ptr, int(s_stride), d_V0.ptr, int(v0_stride), d_V.ptr, int(max_iter),
            float(tol), int(precision), int(instances_per_wg), d_iters.ptr, d_status.ptr, ctypes.byref(st)))
        return st.as_dict()


def branch_flows_device(ctx: Context, d_V: DeviceArray, ninst: int, n: int, d_f, d_t, d_yff, d_yft, d_vn, sn_mva: float,
                        d_amps=None, d_p=None):
    """Origin-side amps / active power of every branch for every instance (device arrays)."""
    nbr = d_f.shape[0]
    _check(ctx._L.pfgpu_branch_flows_device(ctx.h, int(ninst), int(n), d_V.ptr, int(nbr), d_f.ptr, d_t.ptr, d_yff.ptr,
                                            d_yft.ptr, d_vn.ptr, float(sn_mva), d_amps.ptr if d_amps else None,
                                            d_p.ptr if d_p else None))


class Server:
    """Serving loop through its C ABI (pfgpu_server_* / pfgpu_submit / pfgpu_poll)."""

    def __init__(self, hbm_budget: int = 0, max_instances_per_launch: int = 4096, precision: int = FP32):
        self._L = lib()
        h = _vp()
        _check(self._L.pfgpu_server_create(int(hbm_budget), int(max_instances_per_launch), int(precision),
                                           ctypes.byref(h)))
        self.h = h
        self._n = {}
        self._rows = {}

    def close(self):
        if getattr(self, "h", None):
            self._L.pfgpu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_grid(self, Ybus, Sbus, V0, slack_ids, slack_weights, pv, pq, max_iter=10, tol=1e-8) -> int:
        Y = _csc(Ybus)
        n = Y.shape[0]
        Yp, pYp = _i(Y.indptr)
        Yi, pYi = _i(Y.indices)
        Yx, pYx = _c(Y.data)
        S, pS = _c(Sbus)
        V, pV = _c(V0)
        sl, psl = _i(np.atleast_1d(slack_ids))
        w, pw = _d(slack_weights)
        pv_, ppv = _i(pv)
        pq_, ppq = _i(pq)
        gid = ctypes.c_int()
        _check(self._L.pfgpu_server_add_grid(self.h, n, pYp, pYi, pYx, pS, pV, psl, len(sl), pw, ppv, len(pv_), ppq,
                                             len(pq_), int(max_iter), float(tol), ctypes.byref(gid)))
        self._n[gid.value] = n
        return gid.value

    def _submit(self, gid, kind, ninst, S=None, deltas=None):
        rid = ctypes.c_uint64()
        Sp = dp = dq = dv = None
        if S is not None:
            S = np.ascontiguousarray(S, np.complex128)
            Sp = S.ctypes.data_as(_f64p)
        if deltas is not None:
            dp_, dp = _i(deltas[0])
            dq_, dq = _i(deltas[1])
            dv_, dv = _c(deltas[2])
        _check(self._L.pfgpu_submit(self.h, int(gid), int(kind), int(ninst), Sp, dp, dq, dv, ctypes.byref(rid)))
        self._rows[rid.value] = (ninst, self._n[gid])
        return rid.value

    def submit_pf(self, gid, Sbus=None):
        return self._submit(gid, REQ_PF, 1, Sbus)

    def submit_contingencies(self, gid, dptr, dpos, dval):
        return self._submit(gid, REQ_CONTINGENCIES, len(dptr) - 1, None, (dptr, dpos, dval))

    def submit_time_series(self, gid, Sbus_rows):
        S = np.atleast_2d(Sbus_rows)
        return self._submit(gid, REQ_TIME_SERIES, S.shape[0], S)

    def poll(self, rid, timeout=-1.0):
        """dict(V, iters, status, latency_s, queue_s, launch_instances), or None if not ready."""
        ninst, n = self._rows[rid]
        V = np.zeros((ninst, n), np.complex128)
        it = np.zeros(ninst, np.int32)
        st = np.zeros(ninst, np.int32)
        ready = ctypes.c_int()
        info = RequestInfo()
        rc = self._L.pfgpu_poll(self.h, ctypes.c_uint64(rid), float(timeout), ctypes.byref(ready),
                                V.ctypes.data_as(_f64p), it.ctypes.data_as(_i32p), st.ctypes.data_as(_i32p),
                                ctypes.byref(info))
        if ready.value:
            self._rows.pop(rid, None)
        _check(rc)
        if not ready.value:
            return None
        return {"V": V, "iters": it, "status": st, "latency_s": info.latency_s, "queue_s": info.queue_s,
                "launch_instances": info.launch_instances}
