"""GPU parity of windowed execution (k_sweep): many gates per HBM sweep, same results.

complex128: max abs error <= 1e-12 against the FP64 oracle running the original
(unreordered, unwindowed) program. complex64: state infidelity <= 1e-6 against the
complex128 result.
"""
import numpy as np
import pytest

import bruteforce as bf
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def run(gpu, prog, init=None, dtype="c128", mode="windowed"):
    n = prog.num_qubits
    st = gpu.State(n, dtype) if init is None else gpu.State(dtype=dtype, amps=init)
    assert st.execute(prog, gpu.exec_opts(mode)) == len(prog)
    st.sync()
    out = st.to_host()
    stats = st.last_stats()
    st.close()
    return out, stats


@pytest.mark.parametrize("n,d,seed", [(12, 10, 1001), (13, 7, 5), (16, 12, 6), (20, 10, 1001)])
@pytest.mark.parametrize("start", ["zero", "random"])
def test_layered_windowed_vs_oracle(gpu, n, d, seed, start):
    p = gpu.layered(n, d, seed)
    init = None if start == "zero" else bf.random_state(n, np.random.default_rng(seed))
    got, stats = run(gpu, p, init)
    want, _ = oracle.execute(n, p.instrs, amps=init)
    err = np.abs(got - want).max()
    print(f"windowed {n}q d{d} {start}: sweeps {stats['sweeps_full']} err {err:.2e}")
    assert err <= TOL
    _, _, info = gpu.plan_windowed(p)
    assert stats["sweeps_full"] == info["sweeps"]


@pytest.mark.parametrize("seed", range(6))
def test_random_programs_windowed_vs_oracle(gpu, seed):
    from test_oracle import random_program
    rng = np.random.default_rng(700 + seed)
    n = int(rng.integers(12, 17))
    ins, pool, meta = random_program(n, 150, rng)
    p = gpu.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                    meta if meta is not None else np.zeros(0, gpu.META_DTYPE))
    init = bf.random_state(n, rng)
    got, _ = run(gpu, p, init)
    want, _ = oracle.execute(n, ins, pool, meta, amps=init)
    assert np.abs(got - want).max() <= TOL


def test_global_swap_in_window_is_real_permutation(gpu):
    n = 14
    rng = np.random.default_rng(11)
    c = gpu.Circuit(n)
    for q in range(n):
        c.ry(q, float(rng.uniform(0, 3)))
    p = c.release()
    gs = np.zeros(3, gpu.GATE_DTYPE)
    gs["type"] = 0xE1
    gs["flags_arity"] = 2
    gs["qubits"] = [[13, 0, 255, 255], [12, 5, 255, 255], [0, 13, 255, 255]]
    tail = gpu.Circuit(n).cx(0, 5).h(12).release().instrs
    p.instrs = np.concatenate([p.instrs, gs, tail])
    got, _ = run(gpu, p)
    want, _ = oracle.execute(n, p.instrs, gswap="permute")
    assert np.abs(got - want).max() <= TOL


def test_inverse_identity_26q_windowed(gpu):
    n = 26
    p = gpu.layered(n, 16, 314)
    init = bf.random_state(n, np.random.default_rng(3))
    st = gpu.State(amps=init)
    o = gpu.exec_opts("windowed")
    st.execute(p, o)
    st.execute(gpu.inverse_layered(p), o)
    got = st.to_host()
    st.close()
    assert np.abs(got - init).max() <= TOL


def test_closed_form_28q_windowed(gpu):
    p = gpu.layered(28, 12, 2718)
    got, stats = run(gpu, p)
    th = p.instrs["payload"][p.instrs["type"] == 0x10].view(np.float32).astype(np.float64)
    assert np.abs(got - np.exp(-0.5j * th.sum()) * 2.0 ** -14).max() <= TOL


def test_windowed_matches_plain_bitwise_for_permutations(gpu):
    n = 15
    rng = np.random.default_rng(4)
    c = gpu.Circuit(n)
    for _ in range(200):
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        [lambda: c.cx(a, b), lambda: c.swap(a, b), lambda: c.x(a), lambda: c.y(b), lambda: c.z(a),
         lambda: c.s(b)][int(rng.integers(6))]()
    p = c.release()
    init = bf.random_state(n, rng)
    got, _ = run(gpu, p, init)
    want, _ = oracle.execute(n, p.instrs, amps=init)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n,d,seed", [(14, 10, 400001), (20, 20, 400002)])
def test_c64_windowed_fidelity(gpu, n, d, seed):
    p = gpu.layered(n, d, seed)
    init = bf.random_state(n, np.random.default_rng(seed))
    ref, _ = run(gpu, p, init, "c128", "plain")
    got, _ = run(gpu, p, init, "c64", "windowed")
    inf = 1.0 - abs(np.vdot(ref, got)) ** 2
    print(f"c64 windowed {n}q: infidelity {inf:.2e}, max abs {np.abs(ref - got).max():.2e}")
    assert inf <= 1e-6


def _infidelity(a, b):
    return 1.0 - abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


@pytest.mark.parametrize("seed", range(4))
def test_c64_phases_on_matrix_cores_vs_oracle(gpu, seed):
    """complex64 generic sweeps (13-qubit windows, 5 register qubits): each register phase with
    enough work runs as one fused 32 x 32 product on the matrix cores (FP64 product of its ops on
    the host, FP16 split operands, FP32 accumulation). Random programs with every gate kind and a
    layered circuit, from random states, against the FP64 oracle."""
    from test_oracle import random_program
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(13, 17))
    ins, pool, meta = random_program(n, 200, rng)
    rand = gpu.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                       meta if meta is not None else np.zeros(0, gpu.META_DTYPE))
    for p in (rand, gpu.layered(n, 12, 40 + seed)):
        _, _, info = gpu.plan_windowed(p, "c64", gpu.exec_opts("windowed"))
        init = bf.random_state(n, rng)
        got, _ = run(gpu, p, init, "c64")
        want, _ = oracle.execute(n, p.instrs, p.pool if p.pool.size else None,
                                 p.meta if len(p.meta) else None, amps=init)
        inf = _infidelity(got, want)
        print(f"c64 {n}q {len(p)} gates: {info['mfma_phases']} of {info['phases']} phases fused, infidelity {inf:.2e}")
        # both programs run some phases on the matrix cores, so the comparison covers them
        assert info["mfma_phases"] > 0, (len(p), info)
        assert inf <= 1e-10
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_generic_kernel_out_of_place_vs_oracle(gpu, dtype):
    """The generic sweep kernel out of place (relabel without specialize: windows free of the
    low qubits, the qubits re-laid out between two buffers): layered circuits and random
    programs from random states against the oracle (complex128: max |d| <= 1e-12; complex64:
    infidelity <= 1e-6), and against the same kernel in place."""
    from test_oracle import random_program
    rng = np.random.default_rng(808)
    ins, pool, meta = random_program(16, 150, rng)
    rnd = gpu.Program(16, ins, pool if pool is not None else np.zeros(0, complex),
                      meta if meta is not None else np.zeros(0, gpu.META_DTYPE))
    for prog in (gpu.layered(22, 16, 8), gpu.layered(24, 20, 8), rnd):
        n = prog.num_qubits
        init = bf.random_state(n, np.random.default_rng(n))
        pl = prog.pool if prog.pool.size else None
        me = prog.meta if len(prog.meta) else None
        want, _ = oracle.execute(n, prog.instrs, pl, me, amps=init)
        outs = {}
        for relabel in (True, False):
            st = gpu.State(amps=init, dtype=dtype)
            st.timing(True)
            st.execute(prog, gpu.exec_opts("windowed", relabel=relabel))
            t = st.timing_read()
            stats = st.last_stats()
            outs[relabel] = st.to_host()
            st.close()
            assert "k_sweep" in t and "k_sweep_spec" not in t, t
            assert stats["relabelled"] == int(relabel), stats
        if dtype == "c128":
            assert np.abs(outs[True] - want).max() <= 1e-12
            assert np.abs(outs[True] - outs[False]).max() <= 1e-13
        else:
            inf = 1 - abs(np.vdot(want, outs[True])) ** 2 / np.vdot(outs[True], outs[True]).real
            assert inf <= 1e-6, inf
