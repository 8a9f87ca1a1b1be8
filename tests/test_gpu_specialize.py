"""Circuit-specialised window sweeps (qlx_exec_opts::specialize; straight-line kernels
compiled at run time from the sweep descriptors) against the CPU oracle and the generic
sweep kernel. complex128: max abs error <= 1e-12 vs the oracle; complex64: state infidelity
<= 1e-6. Random programs mix every gate kind the sweeps run in registers (1-qubit gates,
CX, SWAP, dense 1/2-qubit matrices) with gates that run as their own kernels between sweeps."""
import numpy as np
import pytest

import bruteforce as bf
import oracle

pytestmark = pytest.mark.gpu


def run(qlx, prog, init, dtype, specialize):
    st = qlx.State(amps=init, dtype=dtype)
    st.timing(True)
    st.execute(prog, qlx.exec_opts("windowed", specialize=specialize))
    t = st.timing_read()
    out = st.to_host()
    st.close()
    return out, t


def programs(qlx):
    from test_oracle import random_program
    rng = np.random.default_rng(2024)
    ins, pool, meta = random_program(14, 120, rng)
    rnd = qlx.Program(14, ins, pool if pool is not None else np.zeros(0, complex),
                      meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    return [("layered16", qlx.layered(16, 8, 1234)), ("hea14", qlx.hea(14, 4)), ("qft13", qlx.qft(13)),
            ("random14", rnd), ("fused2", qlx.transpile_gpu(qlx.layered(15, 6, 77), max_width=2, rotations=True))]


def test_specialized_sweeps_vs_oracle_and_generic(gpu):
    for name, prog in programs(gpu):
        n = prog.num_qubits
        init = bf.random_state(n, np.random.default_rng(n))
        want, _ = oracle.execute(n, prog.instrs, prog.pool if prog.pool.size else None,
                                 prog.meta if len(prog.meta) else None, amps=init)
        got, t = run(gpu, prog, init, "c128", True)
        assert "k_sweep_spec" in t and "k_sweep" not in t, (name, t)
        assert np.abs(got - want).max() <= 1e-12, name
        gen, _ = run(gpu, prog, init, "c128", False)
        assert np.abs(got - gen).max() <= 1e-13, name
        g64, _ = run(gpu, prog, init, "c64", True)
        inf = 1 - abs(np.vdot(want, g64)) ** 2 / np.vdot(g64, g64).real
        assert inf <= 1e-6, (name, inf)


def test_specialized_kernels_are_reused(gpu):
    prog = gpu.layered(16, 6, 99)
    st = gpu.State(16, "c128")
    o = gpu.exec_opts("windowed", specialize=True)
    st.execute(prog, o)          # compiles
    st.sync()
    st.timing(True)
    for _ in range(3):
        st.execute(prog, o)      # reuses the compiled sweeps
    t = st.timing_read()
    assert t["k_sweep_spec"][0] == 3 * st.last_stats()["sweeps_full"]
    assert abs(st.norm_sq() - 1) < 1e-12
    st.close()


@pytest.mark.parametrize("dtype,w,rb", [("c128", 13, 4), ("c128", 13, 5), ("c128", 11, 3), ("c64", 14, 5),
                                        ("c64", 12, 4)])
def test_specialized_window_geometries(gpu, dtype, w, rb):
    """Windows only the specialised sweeps take (c128: 13 qubits, 128 KiB of LDS; c64: 14)
    and smaller ones, against the oracle."""
    o = gpu.exec_opts("windowed", window_qubits=w, register_qubits=rb, specialize=True)
    for prog in (gpu.layered(18, 10, 4321), gpu.clustered(17, 6, 15), gpu.hea(16, 5)):
        n = prog.num_qubits
        init = bf.random_state(n, np.random.default_rng(n + w))
        want, _ = oracle.execute(n, prog.instrs, amps=init)
        st = gpu.State(amps=init, dtype=dtype)
        st.timing(True)
        st.execute(prog, o)
        assert "k_sweep_spec" in st.timing_read()
        got = st.to_host()
        st.close()
        if dtype == "c128":
            assert np.abs(got - want).max() <= 1e-12, (w, rb, n)
        else:
            inf = 1 - abs(np.vdot(want, got)) ** 2 / np.vdot(got, got).real
            assert inf <= 1e-6, (w, rb, n, inf)


@pytest.mark.parametrize("dtype,relabel", [("c128", False), ("c128", True), ("c64", True)])
def test_pauli_frames_vs_oracle(gpu, dtype, relabel):
    """CX gates whose control is outside the registers leave pending X frames; H turns them into
    Z frames, CX and SWAP move them, RX carries X and runs with a per-work-item sign on its angle
    under Z, other gates, re-layouts and stores apply them. Random H/RX/CX/SWAP/S/Z/RY circuits
    with long-range CX, and the clustered H/CX circuit, against the oracle."""
    rng = np.random.default_rng(7100 + relabel)
    progs = []
    for n in (13, 16):
        c = gpu.Circuit(n)
        for _ in range(300):
            r = rng.random()
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            if r < 0.35:
                c.cx(a, b)
            elif r < 0.6:
                c.h(a)
            elif r < 0.82:
                c.rx(a, float(rng.uniform(0, 2 * np.pi)))
            elif r < 0.87:
                c.swap(a, b)
            elif r < 0.92:
                c.s(a)
            elif r < 0.96:
                c.z(a)
            else:
                c.ry(a, float(rng.uniform(0, 2 * np.pi)))
        progs.append(c.release())
    progs.append(gpu.clustered(18))
    o = gpu.exec_opts("windowed", specialize=True, relabel=relabel)
    for prog in progs:
        n = prog.num_qubits
        init = bf.random_state(n, rng)
        want, _ = oracle.execute(n, prog.instrs, amps=init)
        st = gpu.State(amps=init, dtype=dtype)
        st.timing(True)
        st.execute(prog, o)
        assert "k_sweep_spec" in st.timing_read()
        got = st.to_host()
        st.close()
        if dtype == "c128":
            assert np.abs(got - want).max() <= 1e-12, n
        else:
            assert 1 - abs(np.vdot(want, got)) ** 2 / np.vdot(got, got).real <= 1e-6, n


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_long_runs_of_factored_gates(gpu, dtype):
    """Hundreds of H and RX gates in one sweep: their factors are applied in several steps
    (the registers stay within 2^30 of the amplitudes' magnitude, also in FP32)."""
    n = 14
    c = gpu.Circuit(n)
    rng = np.random.default_rng(17)
    for k in range(400):
        q = int(rng.integers(0, 6))
        if k % 3:
            c.h(q)
        else:
            c.rx(q, float(rng.choice([np.pi, np.pi / 2, 3 * np.pi / 2 + 1e-3, 1e-4, rng.uniform(0, 2 * np.pi)])))
    prog = c.release()
    init = bf.random_state(n, np.random.default_rng(3))
    want, _ = oracle.execute(n, prog.instrs, amps=init)
    st = gpu.State(amps=init, dtype=dtype)
    st.execute(prog, gpu.exec_opts("windowed", specialize=True))
    got = st.to_host()
    st.close()
    if dtype == "c128":
        assert np.abs(got - want).max() <= 1e-12
    else:
        assert 1 - abs(np.vdot(want, got)) ** 2 / np.vdot(got, got).real <= 1e-6
        assert abs(np.vdot(got, got).real - 1) <= 1e-5


def relabel_programs(qlx):
    from test_oracle import random_program
    rng = np.random.default_rng(77)
    ins, pool, meta = random_program(15, 160, rng)      # CCX / 3-qubit dense gates run between sweeps
    rnd = qlx.Program(15, ins, pool if pool is not None else np.zeros(0, complex),
                      meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    return [("layered18", qlx.layered(18, 10, 4321)), ("hea16", qlx.hea(16, 5)), ("qft15", qlx.qft(15)),
            ("clustered17", qlx.clustered(17)), ("random15", rnd),
            ("fused5", qlx.transpile_gpu(qlx.layered(16, 6, 9), max_width=5, rotations=True))]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_relabelled_sweeps_vs_oracle(gpu, dtype):
    """Out-of-place sweeps that re-lay out the qubits between two buffers (relabel), with
    plain kernels between sweeps running in the current layout: the state ends in logical
    order and matches the oracle, also on a second execution of the cached plan."""
    o = gpu.exec_opts("windowed", specialize=True, relabel=True)
    for name, prog in relabel_programs(gpu):
        n = prog.num_qubits
        init = bf.random_state(n, np.random.default_rng(n + 5))
        pool = prog.pool if prog.pool.size else None
        meta = prog.meta if len(prog.meta) else None
        want, _ = oracle.execute(n, prog.instrs, pool, meta, amps=init)
        want2, _ = oracle.execute(n, prog.instrs, pool, meta, amps=want)
        st = gpu.State(amps=init, dtype=dtype)
        st.timing(True)
        st.execute(prog, o)
        assert "k_sweep_spec" in st.timing_read(), name
        got = st.to_host()
        st.execute(prog, o)
        got2 = st.to_host()
        st.close()
        for w, g in ((want, got), (want2, got2)):
            if dtype == "c128":
                assert np.abs(g - w).max() <= 1e-12, name
            else:
                inf = 1 - abs(np.vdot(w, g)) ** 2 / np.vdot(g, g).real
                assert inf <= 1e-6, (name, inf)


def test_relabel_then_in_place_on_one_state(gpu):
    """A state that ran out-of-place sweeps (and now owns two buffers) runs in-place programs
    and reads back correctly; its device pointer names the buffer holding the state."""
    import ctypes as C
    n = 16
    p = gpu.layered(n, 8, 31)
    init = bf.random_state(n, np.random.default_rng(2))
    want, _ = oracle.execute(n, p.instrs, amps=init)
    want, _ = oracle.execute(n, p.instrs, amps=want)
    want, _ = oracle.execute(n, p.instrs, amps=want)
    st = gpu.State(amps=init)
    st.execute(p, gpu.exec_opts("windowed", specialize=True, relabel=True))
    st.execute(p, gpu.exec_opts("windowed", specialize=True))
    st.execute(p, gpu.exec_opts("plain"))
    ptr = C.c_void_p()
    gpu._abi.check(gpu._abi.lib().qlx_state_device_ptr(st.handle, C.byref(ptr)))
    assert ptr.value
    got = st.to_host()
    st.close()
    assert np.abs(got - want).max() <= 1e-12
