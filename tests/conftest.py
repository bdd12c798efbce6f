import io
import os
import subprocess
import sys
import zipfile

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURES = {"case14": 3, "case30": 3, "case118": 4, "case300": 5, "case1888": 8}
REPLAY_SO = os.path.join(ROOT, "tests", "cpp", "libreplay.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine through the C ABI)")


def load_fixture(name):
    """Reference golden vectors (lightsim2grid/tests/<name>.zip), numpy arrays only."""
    with zipfile.ZipFile(os.path.join(GOLDEN, name + ".zip")) as z:
        def g(nm):
            return np.load(io.BytesIO(z.read(nm)), allow_pickle=False)
        d = {k: g(k + ".npy") for k in ("Ybus", "Sbus", "V0", "pv", "pq")}
        k = [nm for nm in z.namelist() if nm.startswith("J_")][0]
        it = int(k[2:-4])
        d["iters"] = it
        d["J"] = g("J_%d.npy" % it)
        d["V"] = g("V_%d.npy" % it)
    n = d["Sbus"].shape[0]
    ref = np.array(sorted(set(range(n)) - set(d["pv"].tolist()) - set(d["pq"].tolist())), dtype=np.int32)
    w = np.zeros(n)
    w[ref] = 1.0 / len(ref)
    d["Ybus"] = sp.csc_matrix(d["Ybus"])
    d["slack_ids"] = ref
    d["slack_weights"] = w
    return d


def solver_args(d):
    return (d["Ybus"], d["V0"], d["Sbus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])


def zero_diagonal_grid(a=4, c=8, x=0.05, name="case14_synth", dc=False):
    """A grid whose Jacobian has an exactly zero diagonal block at a flat start: a new PQ bus with
    no injection joined to buses a and c by lines of reactance x and -x, so that its Ybus diagonal
    is 1/(jx) + 1/(-jx) = 0 (kept as a stored entry), and its current, and with it every entry of
    its diagonal J block, is 0 while V is flat. Static diagonal pivots fail on it where a
    partial-pivoting LU picks an off-diagonal row. Returns solver inputs with a flat V0.
    dc: the DC form, Ybus = B with real entries 1/x per branch (its diagonal at the new bus
    1/x - 1/x = 0)."""
    import scipy.sparse as sp
    from lightsim2grid_amd import gridgen
    g = gridgen.make(name)
    d = g.solver_inputs()
    if dc:
        y = 1.0 / g.x
        Y = sp.coo_matrix((np.r_[y, y, -y, -y].astype(np.complex128),
                           (np.r_[g.f, g.t, g.f, g.t], np.r_[g.f, g.t, g.t, g.f])), shape=(g.n, g.n))
        Y = sp.coo_matrix(sp.csc_matrix(Y))
    else:
        Y = sp.coo_matrix(d["Ybus"])
    n = Y.shape[0]
    y1, y2 = (1 / x, -1 / x) if dc else (1 / (1j * x), 1 / (-1j * x))
    r = np.concatenate([Y.row, [n, a, n, a, c, n, c]])
    cc = np.concatenate([Y.col, [n, a, a, n, c, c, n]])
    v = np.concatenate([Y.data, [y1 + y2, y1, -y1, -y1, y2, -y2, -y2]])
    Y2 = sp.csc_matrix((v, (r, cc)), shape=(n + 1, n + 1))
    Y2.sum_duplicates()
    Y2.sort_indices()
    assert Y2[n, n] == 0 and Y2.nnz == Y.nnz + 5
    V0 = np.ones(n + 1, complex)
    V0[d["pv"]] = np.abs(d["V0"][d["pv"]])
    V0[d["slack_ids"]] = np.abs(d["V0"][d["slack_ids"]])
    return {"Ybus": Y2, "V0": V0, "Sbus": np.append(d["Sbus"], 0), "slack_ids": d["slack_ids"],
            "slack_weights": np.append(d["slack_weights"], 0), "pv": d["pv"],
            "pq": np.append(d["pq"], n).astype(d["pq"].dtype)}


def build_replay():
    src = os.path.join(ROOT, "tests", "cpp", "plan_replay.cpp")
    ana = os.path.join(ROOT, "lightsim2grid_amd", "csrc", "analysis.cpp")
    if (not os.path.exists(REPLAY_SO) or os.path.getmtime(REPLAY_SO) < max(os.path.getmtime(src), os.path.getmtime(ana))):
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", REPLAY_SO, src, ana])
    return REPLAY_SO


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    from lightsim2grid_amd import abi
    ctx = abi.Context(0)
    yield ctx
    ctx.close()
