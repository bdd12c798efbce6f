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
