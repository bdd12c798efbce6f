"""The entry point that allocates nothing (pfgpu_solve_batch_device_noalloc), through the C ABI,
against pfgpu_solve_batch_device on the same inputs: two no-allocation calls in a row, the second
with a smaller batch, each equal to the allocating call (statuses, iterations, voltages bit for
bit: the same kernels at the same launch shape)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("grid_name,sizes", [("case2000_synth", (600, 150)), ("case118_synth", (160, 40))])
@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_allocating_entry_point(gpu_ctx, grid_name, sizes, prec):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make(grid_name)
    d = g.solver_inputs()
    nb = np.nonzero(~gridgen.bridges(g))[0]
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    Yx = gr.ybus_values(d["Ybus"])
    base = gr.solve(Yx, d["Sbus"], d["V0"], 10, 1e-8)
    Vb = base["V"][0]
    dY, dS, dV0 = gpu_ctx.to_device(Yx), gpu_ctx.to_device(np.ascontiguousarray(d["Sbus"])), gpu_ctx.to_device(Vb)
    runs = []
    for m in sizes:
        sel = nb[:m] if len(nb) >= m else np.resize(nb, m)
        deltas = g.ybus_deltas([[int(k)] for k in sel])
        dd = tuple(gpu_ctx.to_device(x) for x in deltas)
        dV = gpu_ctx.empty((m, g.n), np.complex128)
        dIt, dSt = gpu_ctx.empty(m, np.int32), gpu_ctx.empty(m, np.int32)
        _say("noalloc reference", grid_name, m, prec)
        st = gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, m, dIt, dSt, 10, 1e-8, precision=prec, deltas=dd)
        runs.append((m, dd, dV, dIt, dSt, (dV.to_host(), dIt.to_host(), dSt.to_host()), st))
    for (m, dd, dV, dIt, dSt, ref, st_ref) in runs:  # two no-allocation calls in a row, the second smaller
        _say("noalloc", grid_name, m, prec)
        dV.copy_from(np.zeros((m, g.n), np.complex128))
        st = gr.solve_device(dY, 0, dS, 0, dV0, 0, dV, m, dIt, dSt, 10, 1e-8, precision=prec, deltas=dd, noalloc=True)
        assert st["instances_per_wg"] == st_ref["instances_per_wg"] and st["block_size"] == st_ref["block_size"]
        assert np.array_equal(dSt.to_host(), ref[2])
        assert np.array_equal(dIt.to_host(), ref[1])
        assert np.array_equal(dV.to_host(), ref[0])
    # in place (V0 == V, one row per instance) with the FP32 re-solve's kept initial rows in the scratch
    m, dd, dV, dIt, dSt, ref, _ = runs[-1]
    dVi = gpu_ctx.to_device(np.repeat(Vb[None, :], m, axis=0))
    gr.solve_device(dY, 0, dS, 0, dVi, g.n, dVi, m, dIt, dSt, 10, 1e-8, precision=prec, deltas=dd, noalloc=True)
    assert np.array_equal(dSt.to_host(), ref[2]) and np.array_equal(dVi.to_host(), ref[0])


def test_noalloc_rejects_what_does_not_fit(gpu_ctx):
    from lightsim2grid_amd import abi, gridgen
    g = gridgen.make("case14_synth")
    d = g.solver_inputs()
    gr = abi.Grid(gpu_ctx, d["Ybus"], d["slack_ids"], d["slack_weights"], d["pv"], d["pq"])
    Yx = gr.ybus_values(d["Ybus"])
    dY, dS, dV0 = gpu_ctx.to_device(Yx), gpu_ctx.to_device(np.ascontiguousarray(d["Sbus"])), gpu_ctx.to_device(d["V0"])
    m = 4_500_000  # 8 bytes of rows per instance: 36 MB of rows plus 16 n per instance in place > 64 MiB
    dV = gpu_ctx.empty((1, g.n), np.complex128)
    dIt, dSt = gpu_ctx.empty(1, np.int32), gpu_ctx.empty(1, np.int32)
    with pytest.raises(abi.PfgpuError) as e:
        gr.solve_device(dY, 0, dS, 0, dV, g.n, dV, m, dIt, dSt, 10, 1e-8, precision=1, noalloc=True)
    assert e.value.code == -4  # PFGPU_E_BUDGET, raised before any launch
