"""The reference's benchmark generators (HEA: RY/RZ layers + CX chain; QFT: H + CX fan-in +
SWAP reversal; include/frontend/circuit_generators.hpp:22-68; the clustered H/CX circuit of
benchmarks/bench_hw_awareness.cpp:30-53) on the GPU paths, against the CPU oracle from seeded
random initial states: complex128 max abs error <= 1e-12 (plain and windowed), complex64
state infidelity 1 - |<want|got>|^2/<got|got> <= 1e-6 (windowed, and fused width 5 on the
matrix cores) and norm drift <= 1e-5. At the benchmark's own sizes (24 qubits) the clustered
circuit is checked against the product of its brute-forced cluster states."""
import numpy as np
import pytest

import bruteforce as bf
import oracle

pytestmark = pytest.mark.gpu

CIRCUITS = [("hea", 14, 6), ("hea", 16, 3), ("qft", 12, 0), ("qft", 15, 0)]


def build(qlx, kind, n, d):
    return qlx.hea(n, d) if kind == "hea" else qlx.qft(n)


def infidelity(want, got):
    return 1 - abs(np.vdot(want, got)) ** 2 / np.vdot(got, got).real


def run(qlx, prog, init, dtype, opts):
    st = qlx.State(amps=init, dtype=dtype)
    st.execute(prog, opts)
    out = st.to_host()
    st.close()
    return out


@pytest.mark.parametrize("kind,n,d", CIRCUITS)
def test_generator_circuits_vs_oracle(gpu, kind, n, d):
    prog = build(gpu, kind, n, d)
    init = bf.random_state(n, np.random.default_rng(n * 31 + d))
    want, _ = oracle.execute(n, prog.instrs, amps=init)
    for mode in ("plain", "windowed"):
        got = run(gpu, prog, init, "c128", gpu.exec_opts(mode))
        assert np.abs(got - want).max() <= 1e-12, mode
    got = run(gpu, prog, init, "c64", gpu.exec_opts("windowed"))
    assert infidelity(want, got) <= 1e-6
    fused = gpu.transpile_gpu(prog, max_width=5, rotations=True)
    for dense in ("bf16", "f32", "f16"):
        got = run(gpu, fused, init, "c64", gpu.exec_opts("plain", dense=dense))
        assert infidelity(want, got) <= 1e-6, dense
        # norm drift of the matrix-core path (BF16: the state's 16-bit split; FP32: rounding)
        assert abs(np.vdot(got, got).real - 1) <= 1e-5, dense
    got = run(gpu, fused, init, "c128", gpu.exec_opts("windowed"))
    assert np.abs(got - want).max() <= 1e-12


# ---- the clustered benchmark circuit ---------------------------------------------------------
def run_zero(qlx, prog, dtype, opts):
    st = qlx.State(prog.num_qubits, dtype)
    st.execute(prog, opts)
    out = st.to_host()
    st.close()
    return out


@pytest.mark.parametrize("n,cluster,depth", [(14, 6, 15), (17, 6, 15), (19, 5, 9)])
def test_clustered_vs_oracle(gpu, n, cluster, depth):
    prog = gpu.clustered(n, cluster, depth)
    want, _ = oracle.execute(n, prog.instrs)
    got = run_zero(gpu, prog, "c128", gpu.exec_opts("plain"))
    assert np.array_equal(got, want)                       # H and CX: bit-exact
    for spec in (False, True):
        got = run_zero(gpu, prog, "c128", gpu.exec_opts("windowed", specialize=spec))
        assert np.abs(got - want).max() <= 1e-12, spec
    # the reference's fusion (width 4, H/CX blocks) and its SWAP windowing, restored at the end
    ref_fused = gpu.transpile_gpu(prog, max_width=4, rotations=False)
    swapped = gpu.transpile_gpu(prog, fusion=True, max_width=4, rotations=False, windowing=True, k_safe=n - 4)
    assert (swapped.instrs["type"] == 0xE1).any()
    fused5 = gpu.transpile_gpu(prog, max_width=5, rotations=True)
    for p in (ref_fused, swapped, fused5):
        for mode in ("plain", "windowed"):
            got = run_zero(gpu, p, "c128", gpu.exec_opts(mode, specialize=(mode == "windowed")))
            assert np.abs(got - want).max() <= 1e-12, (len(p), mode)
    got = run_zero(gpu, fused5, "c64", gpu.exec_opts("plain", dense="bf16"))
    assert infidelity(want, got) <= 1e-6


def test_clustered_24q_product_form(gpu):
    """The benchmark's smallest size (24 qubits, clusters of 6, depth 15) in each strategy
    the reference times (no pass, fusion width 4) on each execution mode."""
    n = 24
    prog = gpu.clustered(n)
    want = bf.clustered_state(n, 6, 15, lambda s: gpu.clustered(s, s, 15).instrs)
    ref_fused = gpu.transpile_gpu(prog, max_width=4, rotations=False)
    assert len(ref_fused) == 72
    for p in (prog, ref_fused):
        for mode, spec in (("plain", False), ("windowed", False), ("windowed", True)):
            got = run_zero(gpu, p, "c128", gpu.exec_opts(mode, specialize=spec))
            err = np.abs(got - want).max()
            print(f"clustered 24q, {len(p)} instructions, {mode} specialise={spec}: max abs err {err:.2e}")
            assert err <= 1e-12
