"""The continuous-batching server (qlx_server_*): results alone and in a mix are the same bit
for bit, also for requests admitted while others run; results match the complex128 path and
the oracle; admission respects the HBM budget."""
import numpy as np
import pytest

import oracle

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


def oracle_infidelity(p, got):
    want, _ = oracle.execute(p.num_qubits, p.instrs)
    ip = np.vdot(want, got)
    return 1.0 - abs(ip) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)


def alone_states(srv, ps):
    out = []
    for p in ps:
        rid = srv.submit(p)
        drain(srv, 1)
        out.append(srv.state(rid, p.num_qubits))
        srv.release(rid)
    return out


def test_alone_and_in_the_mix_are_identical(gpu):
    ps = progs(gpu, 30)
    srv = gpu.Server(1 << 30, "c64")
    alone = alone_states(srv, ps)
    ids = [srv.submit(p) for p in ps]          # all arrive while nothing is running
    drain(srv, len(ps))
    for j, (p, i) in enumerate(zip(ps, ids)):
        got = srv.state(i, p.num_qubits)
        assert np.array_equal(alone[j], got), j
        ref = gpu.State(p.num_qubits, "c128")
        ref.execute(p)
        ref.sync()
        inf, mx = srv.compare(i, ref)
        ref.close()
        assert inf <= 1e-6 and mx <= 1e-5
        if p.num_qubits <= 16:
            assert oracle_infidelity(p, got) <= 1e-6, j
        srv.release(i)
    srv.close()


def test_requests_admitted_while_others_run_match_alone_and_oracle(gpu):
    """Continuous batching: two long 26-qubit requests are admitted first; small requests are then
    submitted one per pump while launches are in flight, and at least one is admitted and
    finished while the long ones are still on the GPU. Every result equals its result alone, bit for bit,
    and the oracle's to infidelity <= 1e-6 (complex64)."""
    big = [gpu.layered(26, 40, 530000 + k) for k in range(2)]
    small = progs(gpu, 16, (12, 15), (5, 12), 88)
    srv = gpu.Server(3 << 30, "c64")
    alone_small = alone_states(srv, small)
    alone_big = alone_states(srv, big)
    big_ids = [srv.submit(p) for p in big]
    for _ in range(200000):                      # both long requests admitted and running
        q, a, f = srv.pump()
        if all(srv.info(r)[0] == "active" for r in big_ids) and f > 0:
            break
    assert all(srv.info(r)[0] == "active" for r in big_ids)
    small_ids, submitted_in_flight, done = [], 0, set()
    j = 0
    for _ in range(400000):
        q, a, f = srv.pump()
        for rid, _lat in srv.take_completed():
            done.add(rid)
        if j < len(small):
            small_ids.append(srv.submit(small[j]))
            submitted_in_flight += 1 if f > 0 else 0
            j += 1
        if j == len(small) and len(done) == len(small) + len(big):
            break
    assert submitted_in_flight >= 1
    assert len(done) == len(small) + len(big)
    big_done = min(srv.info(r)[3] for r in big_ids)
    big_admit = max(srv.info(r)[2] for r in big_ids)
    overlapped = [r for r in small_ids if big_admit < srv.info(r)[2] < big_done and srv.info(r)[3] < big_done]
    print(f"serve: {len(overlapped)} of {len(small)} small requests admitted and finished while the long "
          f"ones ran (long requests done at {big_done * 1e3:.1f} ms)", flush=True)
    assert len(overlapped) >= 1
    for p, rid, want in zip(small, small_ids, alone_small):
        got = srv.state(rid, p.num_qubits)
        assert np.array_equal(got, want)
        assert oracle_infidelity(p, got) <= 1e-6
        srv.release(rid)
    for p, rid, want in zip(big, big_ids, alone_big):
        assert np.array_equal(srv.state(rid, p.num_qubits), want)
        ref = gpu.State(p.num_qubits, "c128")
        ref.execute(p, gpu.exec_opts("windowed"))
        ref.sync()
        inf, _ = srv.compare(rid, ref)
        ref.close()
        assert inf <= 1e-6
        srv.release(rid)
    srv.close()


@pytest.mark.parametrize("in_place,fit", [(True, 2), (False, 1)])
def test_hbm_budget_limits_admission(gpu, in_place, fit):
    # 16 requests of 16 qubits (512 KiB each, c64) through a 1.25 MiB budget: in place two fit,
    # three do not; out of place (two buffers per request) one fits
    ps = [gpu.layered(16, 6, 520000 + j) for j in range(16)]
    srv = gpu.Server(5 << 18, "c64", in_place=in_place)
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
    assert max_active == fit
    srv.close()
    with pytest.raises(gpu.QlxOutOfMemory):
        gpu.Server(1 << 20, "c64").submit(gpu.layered(20, 5, 1))


def test_poisson_driver_small(gpu):
    ps = progs(gpu, 50, (12, 16), (5, 10), 77)
    arr = gpu.poisson_arrivals(len(ps), rate=2000.0, seed=5006)
    out = gpu.run_server(ps, arr, 1 << 30, "c64", check_every=10)
    assert out["requests"] == 50 and out["p99_ms"] >= out["p50_ms"] > 0
    assert out["max_infidelity"] <= 1e-6


def test_out_of_place_requests_match_in_place_and_oracle(gpu):
    """Requests larger than one window run out of place by default (relabelled sweeps between
    two buffers, windows free of the low qubits): each result against the complex128 path, the
    in-place server's result and, up to 18 qubits, the oracle."""
    ps = progs(gpu, 24, (14, 22), (5, 30), 91)
    results = {}
    for in_place in (False, True):
        srv = gpu.Server(1 << 31, "c64", in_place=in_place)
        ids = [srv.submit(p) for p in ps]
        drain(srv, len(ps))
        for j, (p, i) in enumerate(zip(ps, ids)):
            results[(in_place, j)] = srv.state(i, p.num_qubits)
            if not in_place:
                ref = gpu.State(p.num_qubits, "c128")
                ref.execute(p)
                ref.sync()
                inf, mx = srv.compare(i, ref)
                ref.close()
                assert inf <= 1e-6 and mx <= 1e-5, (j, inf)
            srv.release(i)
        srv.close()
    for j, p in enumerate(ps):
        a, b = results[(False, j)], results[(True, j)]
        fid = abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
        assert 1.0 - fid <= 1e-6, j
        if p.num_qubits <= 18:
            assert oracle_infidelity(p, a) <= 1e-6, j
