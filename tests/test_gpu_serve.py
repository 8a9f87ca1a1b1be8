"""The continuous-batching server (qlx_server_*): results alone and in a mix are the same,
every result matches the complex128 path, admission respects the HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def progs(qlx, count, qubits=(12, 16), depths=(5, 12), shape_seed=55):
    return [qlx.layered(n, d, s) for n, d, s in qlx.serve_workload(count, qubits, depths, shape_seed, 510000)]


def drain(srv, expect, limit=200000):
    done = {}
    for _ in range(limit):
        srv.pump()
        for rid, lat in srv.take_completed():
            done[rid] = lat
        if len(done) == expect:
            return done
    raise AssertionError(f"server did not drain: {len(done)} of {expect}")


def test_alone_and_in_the_mix_are_identical(gpu):
    ps = progs(gpu, 30)
    probe = ps[11]
    srv = gpu.Server(1 << 30, "c64")
    rid = srv.submit(probe)
    drain(srv, 1)
    alone = srv.state(rid, probe.num_qubits)
    srv.release(rid)
    ids = [srv.submit(p) for p in ps]          # all arrive while nothing is running
    drain(srv, len(ps))
    mixed = srv.state(ids[11], probe.num_qubits)
    assert np.array_equal(alone, mixed)
    for p, i in zip(ps, ids):
        ref = gpu.State(p.num_qubits, "c128")
        ref.execute(p)
        ref.sync()
        inf, mx = srv.compare(i, ref)
        ref.close()
        assert inf <= 1e-6 and mx <= 1e-5
        srv.release(i)
    srv.close()


def test_requests_join_while_others_run(gpu):
    """Continuous batching: submissions interleaved with pumps (work in flight) all complete."""
    ps = progs(gpu, 40, (12, 18), (5, 20), 66)
    srv = gpu.Server(1 << 30, "c64")
    ids, done = [], {}
    for p in ps:
        ids.append(srv.submit(p))
        q, a, f = srv.pump()
        for rid, lat in srv.take_completed():
            done[rid] = lat
    for _ in range(200000):
        if len(done) == len(ps):
            break
        srv.pump()
        for rid, lat in srv.take_completed():
            done[rid] = lat
    assert len(done) == len(ps) and all(v >= 0 for v in done.values())
    for rid in ids:
        srv.release(rid)
    srv.close()


def test_hbm_budget_limits_admission(gpu):
    # 16 requests of 16 qubits (512 KiB each, c64) through a 1.25 MiB budget: two fit, three do not
    ps = [gpu.layered(16, 6, 520000 + j) for j in range(16)]
    srv = gpu.Server(5 << 18, "c64")
    ids = [srv.submit(p) for p in ps]
    max_active = 0
    done = {}
    for _ in range(200000):
        q, a, f = srv.pump()
        max_active = max(max_active, a)
        for rid, lat in srv.take_completed():
            done[rid] = lat
            srv.release(rid)
        if len(done) == len(ps):
            break
    assert len(done) == len(ps)
    assert max_active == 2
    srv.close()
    with pytest.raises(gpu.QlxOutOfMemory):
        gpu.Server(1 << 20, "c64").submit(gpu.layered(20, 5, 1))


def test_poisson_driver_small(gpu):
    ps = progs(gpu, 50, (12, 16), (5, 10), 77)
    arr = gpu.poisson_arrivals(len(ps), rate=2000.0, seed=5006)
    out = gpu.run_server(ps, arr, 1 << 30, "c64", check_every=10)
    assert out["requests"] == 50 and out["p99_ms"] >= out["p50_ms"] > 0
    assert out["max_infidelity"] <= 1e-6
