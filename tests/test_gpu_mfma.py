"""Dense 4/5-qubit gates on complex64 states through the matrix cores (k_dense_mfma),
and FP64 fused-block construction on the device.

Reference: apply_unitary (include/backend/cpu_openmp_engine.hpp:539-606) and
pass_gate_fusion / FusionMath (include/middleware/middleware.hpp:526-753, 241-344).
The complex64 results are compared with the complex128 vector-unit result of the same
program on the same state (pinned to the CPU oracle by test_gpu_parity.py), and whole
circuits with the CPU oracle of the unfused circuit.

Tolerances. BF16 splits the matrix into three bf16 terms (24 significand bits, like
FP32; the split is made once per gate in FP64 on the device) and each state component
into two (hi + lo: 16 bits, relative error <= 2^-16), so one gate moves the state by a
relative error of order 1e-5 that differs from amplitude to amplitude (stated bound:
max |diff| <= 2e-4 * max |amp|, infidelity <= 1e-8 per gate). FP32 MFMA is plain FP32
arithmetic: max |diff| <= 2e-6 * max |amp|, infidelity <= 1e-11 per gate. FP16 splits both
operands into two fp16 terms (22 significand bits each; the state scaled by 2^14 first) and
drops the lo x lo product (<= 2^-22 relative): max |diff| <= 4e-6 * max |amp|, infidelity
<= 1e-10 per gate. Infidelity is the state infidelity 1 - |<a|b>|^2 / (<a|a><b|b>); the norm
is checked on its own.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = {"bf16": (2e-4, 1e-8), "f32": (2e-6, 1e-11), "f16": (4e-6, 1e-10), "valu": (2e-6, 1e-11)}
# infidelity bounds for whole circuits
CIRCUIT_BOUND = {"bf16": 1e-6, "f16": 1e-8, "f32": 1e-9, "valu": 1e-9}


def rand_unitary(rng, k):
    d = 1 << k
    z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def rand_state(rng, n):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def infidelity(a, b):
    return 1.0 - abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def run_c64(qlx, prog, psi, dense, mode="plain", relabel=False):
    st = qlx.State(amps=psi, dtype="c64")
    st.timing(True)
    st.execute(prog, qlx.exec_opts(mode, dense=dense, relabel=relabel))
    t = st.timing_read()
    out = st.to_host()
    st.close()
    return out, t


def run_c128(qlx, prog, psi):
    st = qlx.State(amps=psi, dtype="c128")
    st.execute(prog)
    out = st.to_host()
    st.close()
    return out


CASES = [
    (5, (0, 1, 2, 3, 4)),        # one group: vector-unit kernel
    (10, (0, 1, 2, 3, 4)),       # exactly one 32-group tile
    (6, (5, 0, 3, 1, 4)),
    (9, (8, 7, 6, 5, 4)),        # descending, 16 groups: vector-unit kernel
    (11, (10, 9, 8, 7, 6)),      # descending, two tiles
    (12, (2, 9, 0, 11, 5)),
    (14, (13, 0, 7, 1, 12)),
    (16, (3, 4, 5, 6, 7)),
    (4, (3, 1, 0, 2)),
    (10, (9, 0, 5, 2)),
    (15, (14, 13, 1, 0)),
]


@pytest.mark.parametrize("dense", ["bf16", "f32", "f16"])
@pytest.mark.parametrize("n,qubits", CASES)
def test_dense_gate_matches_c128(gpu, dense, n, qubits):
    rng = np.random.default_rng(1000 * n + len(qubits))
    k = len(qubits)
    u = rand_unitary(rng, k)
    prog = gpu.Circuit(n).unitary(list(qubits), u, max_arity=5).release()
    psi = rand_state(rng, n)
    got, t = run_c64(gpu, prog, psi, dense)
    if n >= k + 5:
        assert "k_dense_mfma" in t and "k_dense" not in t, t
    else:  # fewer than 32 groups: the vector-unit kernel
        assert "k_dense" in t and "k_dense_mfma" not in t, t
    want = run_c128(gpu, prog, psi)
    # independent check of the c128 path: numpy brute force at small n
    if n <= 10:
        ref, _ = oracle.execute(n, prog.instrs, prog.pool, prog.meta, amps=psi)
        assert np.abs(want - ref).max() <= 1e-12
    tol_abs, tol_inf = TOL[dense]
    assert np.abs(got - want).max() <= tol_abs * np.abs(want).max(), (dense, n, qubits)
    assert infidelity(got, want) <= tol_inf


def test_dense_mfma_repeated_gates_keep_norm(gpu):
    """200 random 5-qubit gates: error grows like a random walk, the norm stays 1."""
    rng = np.random.default_rng(77)
    n = 14
    c = gpu.Circuit(n)
    for _ in range(200):
        qs = [int(x) for x in rng.choice(n, 5, replace=False)]
        c.unitary(qs, rand_unitary(rng, 5), max_arity=5)
    prog = c.release()
    psi = rand_state(rng, n)
    want = run_c128(gpu, prog, psi)
    for dense, bound in [("bf16", 1e-6), ("f32", 1e-9), ("f16", 1e-8)]:
        got, _ = run_c64(gpu, prog, psi, dense)
        assert infidelity(got, want) <= bound, dense
        assert abs(np.vdot(got, got).real - 1.0) <= 1e-4


@pytest.mark.parametrize("dense", ["bf16", "f32", "f16", "valu"])
@pytest.mark.parametrize("mode", ["plain", "windowed"])
def test_fused_layered_circuit_c64_vs_oracle(gpu, dense, mode):
    n, d = 16, 12
    prog = gpu.layered(n, d, 1003)
    fused = gpu.transpile_gpu(prog, max_width=5, rotations=True)
    assert (fused.meta["num_qubits"] == 5).any()
    rng = np.random.default_rng(5)
    psi = rand_state(rng, n)
    want, _ = oracle.execute(n, prog.instrs, amps=psi)
    got, t = run_c64(gpu, fused, psi, dense, mode)
    assert ("k_dense_mfma" in t) == (dense != "valu"), t
    assert infidelity(got, want) <= CIRCUIT_BOUND[dense]
    assert np.abs(got - want).max() <= 50 * TOL[dense][0] * np.abs(want).max()


def test_device_built_fusion_matches_host(gpu):
    for n, d, seed, w in [(12, 8, 1, 5), (10, 6, 2, 4), (20, 4, 3, 5)]:
        prog = gpu.layered(n, d, seed)
        host = gpu.transpile_gpu(prog, max_width=w, rotations=True)
        dev = gpu.transpile_gpu(prog, max_width=w, rotations=True, on_device=True)
        assert host.instrs.tobytes() == dev.instrs.tobytes()
        assert host.meta.tobytes() == dev.meta.tobytes()
        assert np.abs(host.pool - dev.pool).max() <= 1e-13
    # dense members and the reference's own gate set (no rotations)
    rng = np.random.default_rng(9)
    c = gpu.Circuit(6).h(0).cx(0, 1).unitary([2, 4], rand_unitary(rng, 2)).cx(1, 2).s(4).swap(3, 5).z(2)
    prog = c.release()
    for rot in (False, True):
        host = gpu.transpile_gpu(prog, max_width=5, rotations=rot)
        dev = gpu.transpile_gpu(prog, max_width=5, rotations=rot, on_device=True)
        assert host.instrs.tobytes() == dev.instrs.tobytes()
        assert np.abs(host.pool - dev.pool).max() <= 1e-13


def test_dense_kernel_choice_rejected_when_unknown(gpu):
    st = gpu.State(6, "c64")
    o = gpu.exec_opts("plain")
    o.dense_kernel = 9
    prog = gpu.Circuit(6).unitary([0, 1, 2, 3], np.eye(16)).release()
    with pytest.raises(gpu.QlxError):
        st.execute(prog, o)
    st.close()


# ---- dense window sweeps (QLX_EXEC_DENSE_SWEEPS) ---------------------------------------------
@pytest.mark.parametrize("relabel", [False, True])
@pytest.mark.parametrize("dense", ["bf16", "f32", "f16"])
def test_dense_sweeps_fused_circuits_vs_oracle(gpu, dense, relabel):
    """Fused programs (widths 5 and 4) in 13-qubit LDS windows, 4/5-qubit gates on the matrix
    cores, the rest as FP32 matrices: against the oracle's unfused circuit."""
    rng = np.random.default_rng(91)
    for n, d, seed, w in [(16, 12, 1003, 5), (18, 8, 7, 4), (14, 10, 3, 5)]:
        prog = gpu.layered(n, d, seed)
        fused = gpu.transpile_gpu(prog, max_width=w, rotations=True)
        psi = rand_state(rng, n)
        want, _ = oracle.execute(n, prog.instrs, amps=psi)
        got, t = run_c64(gpu, fused, psi, dense, "dense", relabel=relabel)
        assert "k_dsweep" in t and "k_dense_mfma" not in t, t
        assert infidelity(got, want) <= CIRCUIT_BOUND[dense], (n, w)
        assert np.abs(got - want).max() <= 50 * TOL[dense][0] * np.abs(want).max()
        assert abs(np.vdot(got, got).real - 1.0) <= 1e-5


def test_dense_sweeps_every_gate_kind_vs_oracle(gpu):
    """Random programs with every gate kind (1-3 qubit standard gates, CCX, CSWAP, dense
    1-5 qubit unitaries) in dense sweeps, from random states; states below one window run
    gate by gate."""
    from test_oracle import random_program
    for seed in range(4):
        rng = np.random.default_rng(300 + seed)
        n = 13 if seed == 0 else int(rng.integers(11, 17))   # 13 qubits: one chunk per sweep
        ins, pool, meta = random_program(n, 150, rng)
        p = gpu.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                        meta if meta is not None else np.zeros(0, gpu.META_DTYPE))
        c = gpu.Circuit(n)
        for _ in range(6):
            qs = [int(x) for x in rng.choice(n, 5, replace=False)]
            c.unitary(qs, rand_unitary(rng, 5), max_arity=5)
        u5 = c.release()
        for prog in (p, u5):
            psi = rand_state(rng, n)
            want, _ = oracle.execute(n, prog.instrs, prog.pool if prog.pool.size else None,
                                     prog.meta if len(prog.meta) else None, amps=psi)
            for dense, relabel in (("bf16", False), ("f16", False), ("f16", True)):
                got, t = run_c64(gpu, prog, psi, dense, "dense", relabel=relabel)
                assert ("k_dsweep" in t) == (n >= 13), (n, t)
                assert infidelity(got, want) <= CIRCUIT_BOUND[dense], (seed, n, dense, relabel)


def test_dense_sweeps_repeat_and_switch_programs(gpu):
    """The cached dense plan follows the program: alternating two programs on one state."""
    n = 15
    rng = np.random.default_rng(5)
    a = gpu.transpile_gpu(gpu.layered(n, 6, 1), max_width=5, rotations=True)
    b = gpu.transpile_gpu(gpu.layered(n, 5, 2), max_width=4, rotations=True)
    psi = rand_state(rng, n)
    want = psi
    st = gpu.State(amps=psi, dtype="c64")
    for prog in (a, b, a, a, b):
        want, _ = oracle.execute(n, prog.instrs, prog.pool, prog.meta, amps=want)
        st.execute(prog, gpu.exec_opts("dense", dense="f32"))
    got = st.to_host()
    st.close()
    assert infidelity(got, want) <= 1e-9


@pytest.mark.parametrize("mode", ["plain", "dense"])
def test_f16_form_on_concentrated_and_basis_states(gpu, mode):
    """The FP16 form scales amplitudes by 2^14 before the split: a state with one amplitude of
    modulus ~1 (scaled: ~2^14, below the FP16 maximum 65504) and a basis state (zeros
    everywhere else) stay exact to the FP16 bound; a state whose amplitudes span many orders
    of magnitude keeps the small ones to the absolute bound."""
    rng = np.random.default_rng(17)
    n = 15
    c = gpu.Circuit(n)
    for _ in range(8):
        qs = [int(x) for x in rng.choice(n, 5, replace=False)]
        c.unitary(qs, rand_unitary(rng, 5), max_arity=5)
    prog = c.release()
    basis = np.zeros(1 << n, complex)
    basis[12345] = 1.0
    peaked = rand_state(rng, n) * 1e-3
    peaked[77] = 1.0
    peaked /= np.linalg.norm(peaked)
    graded = rand_state(rng, n) * np.exp(-np.arange(1 << n) / 2000.0)
    graded /= np.linalg.norm(graded)
    for psi in (basis, peaked, graded):
        want, _ = oracle.execute(n, prog.instrs, prog.pool, prog.meta, amps=psi)
        got, t = run_c64(gpu, prog, psi, "f16", mode)
        assert ("k_dsweep" in t) == (mode == "dense"), t
        assert infidelity(got, want) <= CIRCUIT_BOUND["f16"]
        assert np.abs(got - want).max() <= 8 * TOL["f16"][0] * np.abs(want).max()


def test_dense_sweeps_relabel_mixed_with_in_place(gpu):
    """Relabelled (out-of-place) dense sweeps leave the state in logical order in whichever of
    the two buffers the last sweep wrote: in-place runs, relabelled runs with an odd and an even
    number of sweeps and host reads interleave on one state."""
    n = 16
    rng = np.random.default_rng(23)
    progs = [gpu.transpile_gpu(gpu.layered(n, d, 40 + d), max_width=5, rotations=True) for d in (3, 7, 9)]
    psi = rand_state(rng, n)
    want = psi
    st = gpu.State(amps=psi, dtype="c64")
    sweeps = set()
    for prog, relabel in [(progs[0], True), (progs[1], False), (progs[2], True), (progs[1], True),
                          (progs[0], False)]:
        want, _ = oracle.execute(n, prog.instrs, prog.pool, prog.meta, amps=want)
        st.execute(prog, gpu.exec_opts("dense", dense="f16", relabel=relabel))
        sweeps.add(st.last_stats()["sweeps_full"] % 2)
        got = st.to_host()
        assert infidelity(got, want) <= CIRCUIT_BOUND["f16"], relabel
    st.close()
    assert sweeps == {0, 1}          # (depths 3, 7, 9: 2, 3 and 4 sweeps)


@pytest.mark.parametrize("gswap", ["permute", "noop"])
def test_dense_sweeps_global_swap_modes(gpu, gswap):
    """GLOBAL_SWAP inside dense sweeps: a qubit permutation (default) or the reference's no-op,
    against the complex128 plain path in the same mode."""
    n = 15
    rng = np.random.default_rng(61)
    base = gpu.transpile_gpu(gpu.layered(n, 6, 11), max_width=5, rotations=True)
    gs = np.zeros(3, gpu.GATE_DTYPE)
    gs["type"] = 0xE1
    gs["flags_arity"] = 2
    gs["qubits"] = [[14, 0, 255, 255], [13, 2, 255, 255], [7, 12, 255, 255]]
    half = len(base) // 2
    p = gpu.Program(n, np.concatenate([base.instrs[:half], gs, base.instrs[half:]]), base.pool, base.meta)
    psi = rand_state(rng, n)
    st = gpu.State(amps=psi, dtype="c128")
    st.execute(p, gpu.exec_opts("plain", gswap=gswap))
    want = st.to_host()
    st.close()
    for relabel in (False, True):
        st = gpu.State(amps=psi, dtype="c64")
        st.execute(p, gpu.exec_opts("dense", dense="f16", gswap=gswap, relabel=relabel))
        got = st.to_host()
        st.close()
        assert infidelity(got, want) <= CIRCUIT_BOUND["f16"], (gswap, relabel)


@pytest.mark.parametrize("n,d,seed", [(22, 16, 5), (20, 12, 4)])
def test_dense_sweeps_free_windows_vs_oracle(gpu, n, d, seed):
    """Relabelled dense sweeps whose windows leave out qubits 0-3 (fewer sweeps than windows
    that hold them; each store is coalesced on qubits the next window holds, walked in its own
    order through LDS): against the oracle's unfused circuit, and against the same program in
    windows that hold qubits 0-3."""
    prog = gpu.layered(n, d, seed)
    fused = gpu.transpile_gpu(prog, max_width=5, rotations=True)
    free = gpu.plan_windowed(fused, "c64", gpu.exec_opts("dense", relabel=True))[2]["sweeps"]
    fixed = gpu.plan_windowed(fused, "c64", gpu.exec_opts("dense", relabel=True, fixed_windows=True))[2]["sweeps"]
    assert free < fixed
    psi = rand_state(np.random.default_rng(seed), n)
    want, _ = oracle.execute(n, prog.instrs, amps=psi)
    got, t = run_c64(gpu, fused, psi, "f16", "dense", relabel=True)
    st = gpu.State(amps=psi, dtype="c64")
    st.execute(fused, gpu.exec_opts("dense", dense="f16", relabel=True, fixed_windows=True))
    ref = st.to_host()
    st.close()
    print(f"{n}q depth {d}: {free} free-window sweeps ({fixed} fixed): infidelity "
          f"{infidelity(got, want):.2e} (fixed windows {infidelity(ref, want):.2e})", flush=True)
    assert "k_dsweep" in t
    assert infidelity(got, want) <= CIRCUIT_BOUND["f16"] and infidelity(ref, want) <= CIRCUIT_BOUND["f16"]
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(want).max()
