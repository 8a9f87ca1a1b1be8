"""Batched execution (qlx_batch_*): every mode gives the states of the circuits run alone.

complex64 batches: infidelity <= 1e-6 against the complex128 plain-path state of the same
circuit (itself pinned to the CPU oracle by test_gpu_parity.py); complex128 batches: max abs
error <= 1e-12 against the CPU oracle.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def ref_state(qlx, prog):
    st = qlx.State(prog.num_qubits, "c128")
    st.execute(prog)
    st.sync()
    return st


def workload(qlx, count, qubits, depths, seed):
    return [qlx.layered(n, d, s) for n, d, s in qlx.batch_workload(count, qubits, depths, seed, 410000)]


@pytest.mark.parametrize("mode", ["persistent", "levels", "per_circuit", "per_circuit_graph"])
def test_batch_c64_matches_c128(gpu, mode):
    progs = workload(gpu, 24, (10, 16), (3, 12), 4321)
    b = gpu.Batch(progs, "c64")
    assert b.info["circuits"] == 24
    b.run(mode)
    ms = b.sync()
    assert ms > 0
    for i, p in enumerate(progs):
        ref = ref_state(gpu, p)
        inf, mx = b.compare(i, ref)
        ref.close()
        assert inf <= 1e-6, (i, p.num_qubits, inf)
        assert mx <= 1e-5
    b.close()


def test_batch_c128_vs_oracle(gpu):
    progs = workload(gpu, 12, (10, 14), (3, 8), 99)
    b = gpu.Batch(progs, "c128")
    b.run("persistent")
    b.sync()
    for i, p in enumerate(progs):
        want, _ = oracle.execute(p.num_qubits, p.instrs)
        assert np.abs(b.state(i) - want).max() <= 1e-12
    b.close()


def test_batch_result_independent_of_the_mix(gpu):
    """A circuit gives the same state alone and inside a mixed batch (same geometry)."""
    progs = workload(gpu, 40, (10, 17), (5, 15), 777)
    probe = progs[7]
    alone = gpu.Batch([probe], "c64")
    alone.run("persistent")
    alone.sync()
    a = alone.state(0)
    mixed = gpu.Batch(progs, "c64")
    mixed.run("persistent")
    mixed.sync()
    assert np.array_equal(mixed.state(7), a)
    alone.close()
    mixed.close()


def test_batch_repeated_runs_and_graph_replays(gpu):
    progs = workload(gpu, 16, (10, 15), (3, 10), 5)
    b = gpu.Batch(progs, "c64")
    b.run("per_circuit_graph")
    b.sync()
    first = [b.state(i) for i in range(len(progs))]
    for _ in range(3):
        b.run("persistent")
        b.sync()
        b.run("per_circuit_graph")
        b.sync()
        for i in range(len(progs)):
            np.testing.assert_allclose(b.state(i), first[i], atol=1e-6)
    b.close()


def test_batch_rejects_unsupported(gpu):
    with pytest.raises(gpu.QlxUnsupported):
        gpu.Batch([gpu.layered(8, 3, 1)], "c64")
    c = gpu.Circuit(12).h(0).ccx(0, 1, 2).release()
    with pytest.raises(gpu.QlxUnsupported):
        gpu.Batch([c], "c64")


def test_batch_autotune_and_fused_batch(gpu):
    """The autotuner returns a configuration it measured; a batch of width-2-fused circuits
    gives the states of the unfused circuits."""
    from qlx import batch as qb
    progs = workload(gpu, 20, (10, 15), (4, 10), 31)
    tune = qb.autotune(progs, "c64", sample=8, reps=1)
    assert (tune["fusion_width"], tune["window_qubits"], tune["out_of_place"]) in [
        (fw, w, oop) for fw, w, oop, _ in tune["trials"]]
    fused = qb.prepare_programs(progs, 2)
    assert any((p.meta["num_qubits"] == 2).any() for p in fused if len(p.meta))
    b = gpu.Batch(fused, "c64", window_qubits=12)
    b.run("levels")
    b.sync()
    for i, p in enumerate(progs):
        ref = ref_state(gpu, p)
        inf, mx = b.compare(i, ref)
        ref.close()
        assert inf <= 1e-6 and mx <= 1e-5, (i, inf, mx)
    b.close()


@pytest.mark.parametrize("mode", ["persistent", "levels"])
def test_batch_out_of_place_matches_in_place_and_c128(gpu, mode):
    """Circuits wider than one window run their sweeps between two buffers (relabelled layouts,
    windows free of the low qubits): every state against the complex128 plain path and against
    the in-place batch; after a per-circuit run the results are read from the first buffer."""
    work = gpu.batch_workload(48, (10, 20), (5, 30), 4010, 401000)
    progs = [gpu.layered(n, d, s) for n, d, s in work]
    assert any(p.num_qubits > 14 for p in progs)
    a = gpu.Batch(progs, "c64", out_of_place=True)
    b = gpu.Batch(progs, "c64")
    assert a.info["sweeps"] < b.info["sweeps"]
    a.run(mode)
    a.sync()
    b.run(mode)
    b.sync()
    for i, p in enumerate(progs):
        ref = gpu.State(p.num_qubits, "c128")
        ref.execute(p)
        ref.sync()
        inf, mx = a.compare(i, ref)
        ref.close()
        assert inf <= 1e-6 and mx <= 1e-5, (i, inf)
        x, y = a.state(i), b.state(i)
        fid = abs(np.vdot(x, y)) ** 2 / (np.vdot(x, x).real * np.vdot(y, y).real)
        assert 1.0 - fid <= 1e-6, i
    a.run("per_circuit")
    a.sync()
    for i in (0, len(progs) - 1):
        ref = gpu.State(progs[i].num_qubits, "c128")
        ref.execute(progs[i])
        ref.sync()
        assert a.compare(i, ref)[0] <= 1e-6
        ref.close()
    a.close()
    b.close()
