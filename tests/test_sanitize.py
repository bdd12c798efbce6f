"""Host sanitizer build: the oracle, the product's host analysis and the CPU plan replay compiled
with AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only) into
tests/cpp/_build/sanitize_driver, run on reference fixtures and synthetic grids (with contingency
deltas for the threaded batch loops). A memory error, a leak or undefined behaviour aborts the
driver; its own checks compare the replay with the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT, load_fixture, zero_diagonal_grid

BUILD = os.path.join(ROOT, "tests", "cpp", "_build")
EXE = os.path.join(BUILD, "sanitize_driver")
SOURCES = [os.path.join(ROOT, "tests", "cpp", "sanitize_driver.cpp"), os.path.join(ROOT, "tests", "cpp", "plan_replay.cpp"),
           os.path.join(ROOT, "oracle", "nr_oracle.cpp"), os.path.join(ROOT, "lightsim2grid_amd", "csrc", "analysis.cpp")]


def _build():
    os.makedirs(BUILD, exist_ok=True)
    newest = max(os.path.getmtime(s) for s in SOURCES + [os.path.join(ROOT, "lightsim2grid_amd", "csrc", "analysis.hpp")])
    if os.path.exists(EXE) and os.path.getmtime(EXE) >= newest:
        return EXE
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-pthread", "-fno-omit-frame-pointer",
                           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-o", EXE + ".tmp"] + SOURCES)
    os.replace(EXE + ".tmp", EXE)
    return EXE


def _write(path, d, deltas=None):
    Y = sp.csc_matrix(d["Ybus"])
    Y.sort_indices()
    n = Y.shape[0]
    i32 = lambda a: np.ascontiguousarray(a, dtype="<i4").tobytes()
    f64 = lambda a: np.ascontiguousarray(a, dtype="<f8").tobytes()
    c128 = lambda a: np.ascontiguousarray(a, dtype=np.complex128).view("<f8").tobytes()
    sl = np.atleast_1d(d["slack_ids"])
    dptr, dpos, dval = deltas if deltas is not None else (np.zeros(1, np.int32), np.zeros(0, np.int32),
                                                          np.zeros(0, np.complex128))
    with open(path, "wb") as fh:
        fh.write(struct.pack("<ii", n, Y.nnz))
        fh.write(i32(Y.indptr) + i32(Y.indices) + c128(Y.data) + c128(d["Sbus"]) + c128(d["V0"]))
        fh.write(struct.pack("<i", len(sl)) + i32(sl) + f64(d["slack_weights"]))
        fh.write(struct.pack("<i", len(d["pv"])) + i32(d["pv"]))
        fh.write(struct.pack("<i", len(d["pq"])) + i32(d["pq"]))
        fh.write(struct.pack("<i", len(dptr) - 1) + i32(dptr) + i32(dpos))
        dv = np.asarray(dval)
        fh.write(f64(dv) if dv.dtype != np.complex128 else c128(dv))


def _cases():
    from lightsim2grid_amd import gridgen
    out = [("case14", load_fixture("case14"), None), ("case118", load_fixture("case118"), None),
           ("zero_diagonal", zero_diagonal_grid(), None)]
    g = gridgen.make("case118_synth")
    nb = np.nonzero(~gridgen.bridges(g))[0][:12]
    out.append(("case118_synth_n1", g.solver_inputs(), g.ybus_deltas([[int(k)] for k in nb])))
    return out


@pytest.mark.parametrize("case", range(4))
def test_host_code_under_asan_ubsan(tmp_path, case):
    exe = _build()
    name, d, deltas = _cases()[case]
    path = tmp_path / f"{name}.grid"
    _write(str(path), d, deltas)
    # verify_asan_link_order=0: the environment may preload libraries of its own ahead of ASan's runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    args = [exe, str(path)] + (["--repivot"] if name == "zero_diagonal" else [])
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "OK" in r.stdout, (name, r.stdout[-2000:], r.stderr[-4000:])
