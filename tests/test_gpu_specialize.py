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
