"""GPU parity: the gfx950 kernels against the CPU oracle, through the C ABI.

Bars: permutation gates (X, Y, Z, S, CX, SWAP, CCX, CSWAP, GLOBAL_SWAP) and H are
bit-exact; every floating-point result is within 1e-12 max abs amplitude error of
the FP64 oracle (the stated tolerance for complex128).
"""
import os
import zlib

import numpy as np
import pytest

import bruteforce as bf
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def gpu_run(qlx, prog, init=None, n=None, dtype="c128", opts=None):
    n = prog.num_qubits if n is None else n
    st = qlx.State(n, dtype) if init is None else qlx.State(dtype=dtype, amps=init)
    stop = st.execute(prog, opts)
    assert stop == len(prog)
    st.sync()
    out = st.to_host()
    stats = st.last_stats()
    st.close()
    return out, stats


def oracle_run(prog, init=None):
    s, _ = oracle.execute(prog.num_qubits, prog.instrs, prog.pool if prog.pool.size else None,
                          prog.meta if len(prog.meta) else None, amps=init)
    return s


# ---- state vector -----------------------------------------------------------------
def test_state_init_and_reset(gpu):
    st = gpu.State(5)
    a = st.to_host()
    assert a[0] == 1 and np.all(a[1:] == 0)
    st.execute(gpu.Circuit(5).h(0).h(3).release())
    st.reset()
    a = st.to_host()
    assert a[0] == 1 and np.all(a[1:] == 0)
    assert st.norm_sq() == 1.0
    st.close()


def test_state_from_host_validation(gpu):
    with pytest.raises(gpu.QlxInvalidArgument):
        gpu.State(amps=np.ones(3) / np.sqrt(3))          # not a power of two
    with pytest.raises(gpu.QlxInvalidArgument):
        gpu.State(amps=np.ones(4))                       # not normalised
    v = bf.random_state(7, np.random.default_rng(1))
    st = gpu.State(amps=v)
    assert np.array_equal(st.to_host(), v)
    assert abs(st.norm_sq() - 1) < 1e-13
    w = bf.random_state(9, np.random.default_rng(2))
    st.load(w)                                            # reset(span) may change the size
    assert st.num_qubits == 9 and np.array_equal(st.to_host(), w)
    st.close()


def test_inner_and_precision_copy(gpu):
    v = bf.random_state(12, np.random.default_rng(3))
    a = gpu.State(amps=v)
    b = gpu.State(12, "c64")
    b.copy_from(a)
    ov = a.inner(b)
    assert abs(abs(ov) - 1) < 1e-6
    assert a.max_abs_diff(b) < 1e-7
    np.testing.assert_allclose(b.to_host(), v, atol=1e-7)
    a.close()
    b.close()


# ---- per-gate kernels ----------------------------------------------------------------
PERM = {"x": 1, "y": 1, "z": 1, "s": 1, "cx": 2, "swap": 2, "ccx": 3, "cswap": 3}
ROT = {"h": 1, "rx": 1, "ry": 1, "rz": 1}


@pytest.mark.parametrize("n", [3, 11, 16])
@pytest.mark.parametrize("gate", sorted(PERM) + sorted(ROT))
def test_single_gate_vs_oracle(gpu, gate, n):
    rng = np.random.default_rng(zlib.crc32(f"{gate}-{n}".encode()))
    k = {**PERM, **ROT}[gate]
    init = bf.random_state(n, rng)
    for trial in range(4):
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        c = gpu.Circuit(n)
        args = qs + ([float(rng.uniform(-7, 7))] if gate in ("rx", "ry", "rz") else [])
        getattr(c, gate)(*args)
        p = c.release()
        got, _ = gpu_run(gpu, p, init)
        want = oracle_run(p, init)
        if gate in PERM or gate == "h":
            assert np.array_equal(got, want), (gate, qs)
        else:
            assert np.abs(got - want).max() <= TOL, (gate, qs)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_dense_unitary_vs_oracle(gpu, k):
    rng = np.random.default_rng(40 + k)
    for n in (k, k + 3, 14):
        for _ in range(3):
            qs = [int(x) for x in rng.choice(n, k, replace=False)]
            c = gpu.Circuit(n).unitary(qs, bf.random_unitary(k, rng), max_arity=5)
            p = c.release()
            init = bf.random_state(n, rng)
            got, _ = gpu_run(gpu, p, init)
            want = oracle_run(p, init)
            assert np.abs(got - want).max() <= TOL, (n, qs)


@pytest.mark.parametrize("k", [4, 5])
def test_dense_unitary_fp64_matrix_cores_vs_oracle(gpu, k):
    """complex128 4/5-qubit gates on >= k + 4 qubits run on the FP64 matrix cores
    (k_dense_f64mfma): the smallest state (one 16-group tile), low, high, spread and unsorted
    target sets, several gates in a row; against the oracle and the vector-unit kernel."""
    rng = np.random.default_rng(4400 + k)
    cases = [(k + 4, list(range(k))), (k + 4, list(range(4, 4 + k))[::-1]), (k + 5, None),
             (17, [0, 3, 7, 12, 16][:k]), (17, None), (20, [19, 1, 10, 5, 14][:k]), (20, None)]
    for n, qs in cases:
        c = gpu.Circuit(n)
        for _ in range(3):
            qq = qs if qs is not None else [int(x) for x in rng.choice(n, k, replace=False)]
            c.unitary(qq, bf.random_unitary(k, rng), max_arity=5)
            qs = None
        p = c.release()
        init = bf.random_state(n, rng)
        got, _ = gpu_run(gpu, p, init)
        valu, _ = gpu_run(gpu, p, init, opts=gpu.exec_opts("plain", dense="valu"))
        want = oracle_run(p, init)
        assert np.abs(got - want).max() <= TOL, (n, k)
        assert np.abs(got - valu).max() <= TOL, (n, k)


def test_custom_swap_unitary_gather_scatter(gpu):
    sw = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], complex)
    p = gpu.Circuit(3).x(0).unitary([0, 2], sw).release()        # test_engine_kernels.cpp:173-202
    got, _ = gpu_run(gpu, p)
    assert abs(got[1]) < 1e-12 and abs(got[4] - 1) < 1e-12


def test_random_programs_vs_oracle(gpu):
    from test_oracle import random_program
    for seed in range(5):
        rng = np.random.default_rng(500 + seed)
        n = int(rng.integers(5, 15))
        ins, pool, meta = random_program(n, 80, rng)
        p = gpu.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                        meta if meta is not None else np.zeros(0, gpu.META_DTYPE))
        init = bf.random_state(n, rng)
        got, _ = gpu_run(gpu, p, init)
        want = oracle_run(p, init)
        assert np.abs(got - want).max() <= TOL


# ---- the layered h/rx/cx generator -------------------------------------------------------
def fixtures():
    return sorted(f for f in os.listdir(GOLD) if f.endswith(".npz"))


@pytest.mark.parametrize("fname", fixtures())
def test_golden_fixtures(gpu, fname):
    z = np.load(os.path.join(GOLD, fname))
    ins = z["instrs"].reshape(-1).view(gpu.GATE_DTYPE)
    n = int(fname.split("_n")[1].split("_")[0])
    p = gpu.Program(n, ins)
    got, _ = gpu_run(gpu, p)
    assert np.abs(got - z["state"]).max() <= TOL
    got, _ = gpu_run(gpu, p, z["init_random"])
    assert np.abs(got - z["state_random"]).max() <= TOL


@pytest.mark.parametrize("start", ["zero", "random"])
def test_config1_20q_depth10_vs_oracle(gpu, start):
    """20-qubit c128 layered circuit, depth 10, seed 1001, no fusion: max abs error <= 1e-12."""
    p = gpu.layered(20, 10, 1001)
    assert len(p) == 315
    init = None if start == "zero" else bf.random_state(20, np.random.default_rng(1001))
    got, stats = gpu_run(gpu, p, init)
    want = oracle_run(p, init)
    err = np.abs(got - want).max()
    print(f"20q depth10 start={start}: max abs err {err:.3e}")
    assert err <= TOL
    assert stats["sweeps_full"] == 220 and stats["sweeps_half"] == 95


def test_layered_closed_form_26q(gpu):
    """From |0...0> the layered circuit leaves |+>^N times exp(-i sum(theta)/2)."""
    p = gpu.layered(26, 8, 77)
    got, _ = gpu_run(gpu, p)
    th = p.instrs["payload"][p.instrs["type"] == 0x10].view(np.float32).astype(np.float64)
    want = np.exp(-0.5j * th.sum()) * 2.0 ** -13
    assert np.abs(got - want).max() <= TOL


def test_inverse_identity_24q(gpu):
    n = 24
    p = gpu.layered(n, 12, 99)
    init = bf.random_state(n, np.random.default_rng(5))
    st = gpu.State(amps=init)
    st.execute(p)
    st.execute(gpu.inverse_layered(p))
    got = st.to_host()
    st.close()
    assert np.abs(got - init).max() <= TOL


# ---- errors ---------------------------------------------------------------------------------
def test_out_of_range_qubit_rejected(gpu):
    st = gpu.State(3)
    bad = gpu.Circuit(8).h(5).release()
    with pytest.raises(gpu.QlxOutOfRange):
        st.execute(bad)
    st.close()


def test_unknown_opcode_runs_prefix_then_raises(gpu):
    p = gpu.Circuit(2).x(0).h(1).release()
    p.instrs = p.instrs.copy()
    p.instrs["type"][1] = 0x7F
    st = gpu.State(2)
    with pytest.raises(gpu.QlxUnsupported):
        st.execute(p)
    got = st.to_host()
    assert got[1] == 1                                    # X(0) ran before the bad opcode
    st.close()


def test_global_swap_modes(gpu):
    p = gpu.Circuit(10).x(0).x(1).release()
    gs = np.zeros(2, gpu.GATE_DTYPE)
    gs["type"] = 0xE1
    gs["flags_arity"] = 2
    gs["qubits"] = [[8, 0, 255, 255], [9, 1, 255, 255]]
    cx = gpu.Circuit(10).cx(0, 1).release().instrs
    p.instrs = np.concatenate([p.instrs, gs, cx])
    got, _ = gpu_run(gpu, p, opts=gpu.exec_opts(gswap="noop"))
    assert got[1] == 1                                    # the reference engine's answer
    got, _ = gpu_run(gpu, p, opts=gpu.exec_opts(gswap="permute"))
    assert got[768] == 1                                  # real permutation


# ---- measurement (test_engine_measure.cpp, algos_no_optimization.cpp) -------------------------
def test_measure_born_rule_and_seed(gpu):
    p = gpu.Circuit(1).h(0).measure(0).release()
    for seed in range(8):
        e = gpu.Engine(1, seed=seed)
        e.execute(p)
        m = e.measurements()
        s = e.state()
        assert len(m) == 1 and abs(abs(s[m[0]]) ** 2 - 1) < 1e-6
        _, om = oracle.execute(1, p.instrs, seed=seed)
        assert m == om
        e.reset()
        assert e.measurements() == [] and e.state()[0] == 1


def test_bell_100_shots(gpu):
    p = gpu.Circuit(2).h(0).cx(0, 1).measure(0).measure(1).release()
    e = gpu.Engine(2, seed=123)
    for _ in range(100):
        e.reset()
        e.execute(p)
        m = e.measurements()
        assert len(m) == 2 and m[0] == m[1]


def test_teleportation_matches_oracle(gpu):
    alpha, beta = 1 / np.sqrt(3), 1j * np.sqrt(2 / 3)
    init = np.zeros(8, complex)
    init[0], init[1] = alpha, beta
    alice = gpu.Circuit(3).h(1).cx(1, 2).cx(0, 1).h(0).measure(0).measure(1).release()
    e = gpu.Engine(initial_state=init, seed=37, reserve_meas=2)
    e.execute(alice)
    m = e.measurements()
    _, om = oracle.execute(3, alice.instrs, amps=init, seed=37)
    assert m == om
    bob = gpu.Circuit(3)
    if m[1]:
        bob.x(2)
    if m[0]:
        bob.z(2)
    e.execute(bob.release())
    s = e.state()
    i0, i1 = (m[1] << 1) | m[0], 4 | (m[1] << 1) | m[0]
    assert abs(s[i1] / s[i0] - beta / alpha) < 1e-6
    for i in range(8):
        if i not in (i0, i1):
            assert abs(s[i]) < 1e-6


def test_measure_collapse_vs_oracle_large(gpu):
    n = 14
    rng = np.random.default_rng(8)
    init = bf.random_state(n, rng)
    c = gpu.Circuit(n).h(3).cx(3, 9).measure(9).rx(2, 0.3).measure(0)
    p = c.release()
    e = gpu.Engine(initial_state=init, seed=2024)
    e.execute(p)
    want, om = oracle.execute(n, p.instrs, amps=init, seed=2024)
    assert e.measurements() == om
    assert np.abs(e.state() - want).max() <= TOL


@pytest.mark.parametrize("mode", ["auto", "generic_relabel", "specialised_relabel"])
def test_measurement_between_out_of_place_segments(gpu, mode):
    """MEASURE inside a 21-qubit layered circuit with the segments between measurements run out
    of place (the state moves between the two buffers): outcomes equal the oracle's for the same
    seed and the final state matches it to 1e-12."""
    n = 21
    base = gpu.layered(n, 8, 2101)
    instrs = list(base.instrs)
    cut = {len(instrs) // 3: 5, 2 * len(instrs) // 3: 17}
    prog_instrs = []
    for k, g in enumerate(instrs):
        prog_instrs.append(g)
        if k in cut:
            m = gpu.Circuit(n).measure(cut[k]).release().instrs[0]
            prog_instrs.append(m)
    prog = gpu.Program(n, np.array(prog_instrs, dtype=gpu.GATE_DTYPE), np.zeros(0, complex),
                       np.zeros(0, gpu.META_DTYPE))
    init = bf.random_state(n, np.random.default_rng(2102))
    opts = {"auto": gpu.exec_opts("auto"),
            "generic_relabel": gpu.exec_opts("windowed", relabel=True),
            "specialised_relabel": gpu.exec_opts("windowed", specialize=True, relabel=True)}[mode]
    e = gpu.Engine(initial_state=init, seed=91, opts=opts)
    e.execute(prog)
    m = e.measurements()
    got = e.state()
    want, om = oracle.execute(n, prog.instrs, amps=init, seed=91)
    assert m == om and len(m) == 2
    assert np.abs(got - want).max() <= 1e-12
