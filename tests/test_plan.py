"""Host-only checks of the windowed planner (qlx_plan_windowed): the execution order
it produces only moves gates past gates on disjoint qubits, so running the reordered
stream on the CPU oracle reproduces the original program."""
import numpy as np
import pytest

import bruteforce as bf
import oracle
import qlx
from test_oracle import random_program


def reordered(prog, order):
    return prog.instrs[order]


@pytest.mark.parametrize("n,d,seed", [(12, 6, 1), (14, 10, 2), (17, 8, 3)])
def test_layered_plan_is_valid_reordering(n, d, seed):
    p = qlx.layered(n, d, seed)
    order, sweep, info = qlx.plan_windowed(p)
    assert sorted(order.tolist()) == list(range(len(p)))
    assert info["window_qubits"] == 12 and info["sweeps"] >= 1
    init = bf.random_state(n, np.random.default_rng(seed))
    a, _ = oracle.execute(n, p.instrs, amps=init)
    b, _ = oracle.execute(n, reordered(p, order), amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)
    # within a sweep, gates only touch that sweep's W qubits
    for s in set(sweep.tolist()) - {-1}:
        qs = set()
        for g in p.instrs[order[sweep == s]]:
            k = 1 if g["type"] < 0x20 else 2
            qs |= set(int(x) for x in g["qubits"][:k])
        assert len(qs) <= info["window_qubits"]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_random_program_plan_is_valid_reordering(seed, dtype):
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(13, 16))
    ins, pool, meta = random_program(n, 120, rng)
    prog = qlx.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                       meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    order, sweep, info = qlx.plan_windowed(prog, dtype)
    assert sorted(order.tolist()) == list(range(len(ins)))
    init = bf.random_state(n, rng)
    a, _ = oracle.execute(n, ins, pool, meta, amps=init)
    b, _ = oracle.execute(n, ins[order], pool, meta, amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)
    # 3-qubit and dense k>=3 gates run as their own kernels
    for k, g in zip(sweep, ins[order]):
        if g["type"] in (0x30, 0x31) or (g["type"] == 0x80 and (g["flags_arity"] & 15) >= 3):
            assert k == -1


def test_config2_plan_shape():
    p = qlx.layered(30, 20, 1002)
    order, sweep, info = qlx.plan_windowed(p)
    print(info)
    assert info["gates_in_sweeps"] == 920 and info["plain_steps"] == 0
    # far fewer HBM sweeps than the 775 pass-equivalents of one-kernel-per-gate execution:
    # the beam search over windows needs 14 (the first-gate greedy window alone needs 21)
    assert info["sweeps"] <= 14


def test_cheap_sweeps_run_as_plain_kernels():
    """A sweep whose gates would cost less than one pass as plain kernels (here the QFT's
    trailing SWAPs, half a pass each) runs them as their own kernels; the order stays valid."""
    n = 17
    p = qlx.qft(n)
    order, sweep, info = qlx.plan_windowed(p)
    assert info["plain_steps"] >= 1
    plain = p.instrs[order[sweep == -1]]
    assert all(g["type"] in (0x20, 0x21) for g in plain)
    init = bf.random_state(n, np.random.default_rng(5))
    a, _ = oracle.execute(n, p.instrs, amps=init)
    b, _ = oracle.execute(n, reordered(p, order), amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)


def test_small_states_plan_nothing():
    order, sweep, info = qlx.plan_windowed(qlx.layered(8, 3, 1))
    assert len(order) == 0 and info["sweeps"] == 0


@pytest.mark.parametrize("seed", range(4))
def test_dense_sweep_plan_is_valid_reordering(seed):
    """Dense sweeps (every gate, up to 5 qubits, inside a 13-qubit window holding qubits
    0-3): the execution order is a valid reordering and every window holds its gates."""
    rng = np.random.default_rng(70 + seed)
    n = int(rng.integers(14, 18))
    ins, pool, meta = random_program(n, 120, rng)
    p = qlx.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                    meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    fused = qlx.transpile_gpu(qlx.layered(n, 6, seed), max_width=5, rotations=True)
    for prog in (p, fused):
        order, sweep, info = qlx.plan_windowed(prog, "c64", qlx.exec_opts("dense"))
        assert sorted(order.tolist()) == list(range(len(prog))) and info["window_qubits"] == 13
        assert (sweep >= 0).all() and info["sweeps"] == sweep.max() + 1
        init = bf.random_state(n, np.random.default_rng(seed))
        pl = prog.pool if prog.pool.size else None
        me = prog.meta if len(prog.meta) else None
        a, _ = oracle.execute(n, prog.instrs, pl, me, amps=init)
        b, _ = oracle.execute(n, prog.instrs[order], pl, me, amps=init)
        np.testing.assert_allclose(b, a, atol=1e-12)
        for s_ in range(info["sweeps"]):
            qs = set()
            for g in prog.instrs[order[sweep == s_]]:
                a_ = int(g["flags_arity"]) & 15
                t = int(g["type"])
                k = 2 if t in (0x20, 0x21, 0xE1) else 3 if t in (0x30, 0x31) else max(a_, 1)
                qs |= set(int(x) for x in g["qubits"][:min(k, 4)]) | ({int(g["pad"]) & 255} if k == 5 else set())
            assert len(qs | {0, 1, 2, 3}) <= 13


def test_dense_sweeps_of_the_fused_32q_circuit():
    """The 32-qubit depth-30 layered circuit fused to width 5 needs far fewer dense sweeps
    than its dense gates (one HBM pass each when run gate by gate)."""
    f = qlx.transpile_gpu(qlx.layered(32, 30, 1003), max_width=5, rotations=True)
    n_dense = int(((f.instrs["type"] == 0xE0) & ((f.instrs["flags_arity"] & 15) >= 4)).sum())
    _, _, info = qlx.plan_windowed(f, "c64", qlx.exec_opts("dense"))
    print(len(f), n_dense, info)
    assert info["sweeps"] * 3 < n_dense


def test_c64_generic_plans_fuse_heavy_register_phases():
    """complex64 generic sweeps (5 register qubits, 256 work-items) run every phase with at least
    five 1-qubit ops' worth of work as one 32 x 32 product on the matrix cores; the specialised
    sweeps and complex128 keep their ops."""
    p = qlx.layered(28, 40, 5)
    _, _, info = qlx.plan_windowed(p, "c64", qlx.exec_opts("windowed"))
    assert (info["register_qubits"], info["threads_per_group"]) == (5, 256)
    assert 0 < info["mfma_phases"] < info["phases"]
    print(info["mfma_phases"], "of", info["phases"])
    _, _, spec = qlx.plan_windowed(p, "c64", qlx.exec_opts("windowed", specialize=True))
    _, _, c128 = qlx.plan_windowed(p, "c128", qlx.exec_opts("windowed"))
    assert spec["mfma_phases"] == 0 and c128["mfma_phases"] == 0
    # a phase of CX gates alone stays on the vector units
    c = qlx.Circuit(14)
    for q in range(13):
        c.cx(q, q + 1)
    _, _, cx = qlx.plan_windowed(c.release(), "c64", qlx.exec_opts("windowed"))
    assert cx["mfma_phases"] == 0


def rotation_rich(n, count, rng):
    """Random program on n qubits, mostly rotations, with every gate kind that commutes with
    some axis on some operand (X, Y, Z, S, CX, CCX, CSWAP) and kinds that commute with none
    (H, SWAP), so that merges both happen and are blocked."""
    c = qlx.Circuit(n)
    for _ in range(count):
        k = int(rng.integers(12))
        q = [int(x) for x in rng.permutation(n)[:3]]
        th = float(rng.uniform(-7.0, 7.0))
        if k < 4:
            [c.rx, c.ry, c.rz, c.rx][k](q[0], th)
        elif k == 4:
            [c.x, c.y, c.z, c.s][int(rng.integers(4))](q[0])
        elif k < 7:
            c.cx(q[0], q[1])
        elif k == 7:
            c.ccx(q[0], q[1], q[2])
        elif k == 8:
            c.cswap(q[0], q[1], q[2])
        elif k == 9:
            c.h(q[0])
        elif k == 10:
            c.swap(q[0], q[1])
        else:
            c.rz(q[0], th)
    return c.release()


@pytest.mark.parametrize("seed", range(8))
def test_merged_rotations_are_a_valid_reordering(seed):
    """Rotations merged through the gates they commute with: the order lists each merged
    rotation right after the one it joined, and that reordered stream gives the original
    state on the oracle (a wrong commutation rule would not)."""
    rng = np.random.default_rng(4200 + seed)
    n = 13
    prog = rotation_rich(n, 160, rng)
    order, sweep, info = qlx.plan_windowed(prog, "c128", qlx.exec_opts("windowed", specialize=True))
    kept = qlx.plan_windowed(prog, "c128", qlx.exec_opts("windowed", specialize=True, keep_rotations=True))[2]
    assert info["rotations_merged"] > 0 and kept["rotations_merged"] == 0
    assert sorted(order.tolist()) == list(range(len(prog)))
    init = bf.random_state(n, rng)
    a, _ = oracle.execute(n, prog.instrs, amps=init)
    b, _ = oracle.execute(n, prog.instrs[order], amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)


@pytest.mark.parametrize("build,merges", [
    (lambda c: c.rx(3, 0.3).x(3).cx(1, 3).ccx(0, 2, 3).rx(3, 0.4), 1),   # X, CX/CCX targets
    (lambda c: c.rx(3, 0.3).cx(3, 1).rx(3, 0.4), 0),                     # CX control
    (lambda c: c.rx(3, 0.3).h(3).rx(3, 0.4), 0),
    (lambda c: c.rx(3, 0.3).rz(3, 0.2).rx(3, 0.4), 0),
    (lambda c: c.rz(3, 0.3).z(3).s(3).cx(3, 1).ccx(3, 0, 2).cswap(3, 0, 1).rz(3, 0.4), 1),
    (lambda c: c.rz(3, 0.3).cx(1, 3).rz(3, 0.4), 0),                     # CX target
    (lambda c: c.rz(3, 0.3).cswap(0, 3, 1).rz(3, 0.4), 0),               # CSWAP swapped operand
    (lambda c: c.ry(3, 0.3).y(3).ry(3, 0.4).ry(3, 0.5), 2),
    (lambda c: c.ry(3, 0.3).x(3).ry(3, 0.4), 0),
    (lambda c: c.rx(3, 0.3).swap(3, 4).rx(4, 0.4).rx(3, 0.5), 0),
])
def test_rotation_merge_rules(build, merges):
    """Which gates a rotation merges through, and which stop it."""
    n = 12
    prog = build(qlx.Circuit(n)).release()
    order, _, info = qlx.plan_windowed(prog, "c128", qlx.exec_opts("windowed", specialize=True))
    assert info["rotations_merged"] == merges
    init = bf.random_state(n, np.random.default_rng(7))
    a, _ = oracle.execute(n, prog.instrs, amps=init)
    b, _ = oracle.execute(n, prog.instrs[order], amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)


def test_config2_merges_half_of_its_rotations():
    """The 30-qubit depth-20 layered circuit: an RX on the target of a CX joins the RX before
    that CX (294 of its 600 RX)."""
    p = qlx.layered(30, 20, 1002)
    o = qlx.exec_opts("windowed", specialize=True, relabel=True)
    order, sweep, info = qlx.plan_windowed(p, "c128", o)
    assert info["rotations_merged"] == 294 and info["gates_in_sweeps"] == 920
    assert sorted(order.tolist()) == list(range(len(p)))


@pytest.mark.parametrize("case", ["layered20", "layered24", "random0", "random1", "random2"])
def test_free_window_plans_are_valid_reorderings(case):
    """Out-of-place specialised sweeps take windows that leave out qubits 0-2 when that needs
    fewer passes: the execution order is still a valid reordering (oracle), every sweep's gates
    touch only W qubits, and the pass count is no higher than with fixed windows."""
    from test_oracle import random_program
    rng = np.random.default_rng(hash(case) % 1000)
    if case.startswith("layered"):
        n = int(case[7:])
        prog = qlx.layered(n, 20, 8)
    else:
        n = 16
        ins, pool, meta = random_program(n, 200, np.random.default_rng(int(case[6:]) + 50))
        prog = qlx.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                           meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    o = qlx.exec_opts("windowed", specialize=True, relabel=True)
    order, sweep, info = qlx.plan_windowed(prog, "c128", o)
    fixed = qlx.plan_windowed(prog, "c128", qlx.exec_opts("windowed", specialize=True, relabel=True,
                                                          fixed_windows=True))[2]
    assert info["sweeps"] + info["plain_steps"] <= fixed["sweeps"] + fixed["plain_steps"]
    if case.startswith("layered"):
        assert info["sweeps"] < fixed["sweeps"]
    assert sorted(order.tolist()) == list(range(len(prog)))
    init = bf.random_state(n, np.random.default_rng(n))
    pl = prog.pool if prog.pool.size else None
    me = prog.meta if len(prog.meta) else None
    a, _ = oracle.execute(n, prog.instrs, pl, me, amps=init)
    b, _ = oracle.execute(n, prog.instrs[order], pl, me, amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)
    for s in set(sweep.tolist()) - {-1}:
        qs = set()
        for g in prog.instrs[order[sweep == s]]:
            k = g["flags_arity"] & 15
            qs |= set(int(x) for x in g["qubits"][:k])
        assert len(qs) <= info["window_qubits"]


@pytest.mark.parametrize("n,d,seed", [(22, 16, 5), (18, 12, 3)])
def test_relabelled_dense_sweep_plans_with_free_windows(n, d, seed):
    """Relabelled dense sweeps take windows without qubits 0-3 when that needs fewer sweeps:
    still a valid reordering, every window holds its gates (13 qubits)."""
    prog = qlx.transpile_gpu(qlx.layered(n, d, seed), max_width=5, rotations=True)
    order, sweep, info = qlx.plan_windowed(prog, "c64", qlx.exec_opts("dense", relabel=True))
    fixed = qlx.plan_windowed(prog, "c64", qlx.exec_opts("dense", relabel=True, fixed_windows=True))[2]
    assert info["sweeps"] < fixed["sweeps"]
    assert sorted(order.tolist()) == list(range(len(prog))) and (sweep >= 0).all()
    init = bf.random_state(n, np.random.default_rng(seed))
    a, _ = oracle.execute(n, prog.instrs, prog.pool, prog.meta, amps=init)
    b, _ = oracle.execute(n, prog.instrs[order], prog.pool, prog.meta, amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)
    for s_ in range(info["sweeps"]):
        qs = set()
        for g in prog.instrs[order[sweep == s_]]:
            a_ = int(g["flags_arity"]) & 15
            qs |= set(int(x) for x in g["qubits"][:min(a_, 4)]) | ({int(g["pad"]) & 255} if a_ == 5 else set())
        assert len(qs) <= 13
