"""The reference's benchmark generators (HEA: RY/RZ layers + CX chain; QFT: H + CX fan-in +
SWAP reversal; include/frontend/circuit_generators.hpp:22-68) on the GPU paths, against the
CPU oracle from seeded random initial states: complex128 max abs error <= 1e-12 (plain and
windowed), complex64 state infidelity 1 - |<want|got>|^2/<got|got> <= 1e-6 (windowed, and fused
width 5 on the matrix cores) and norm drift <= 1e-5."""
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
    for dense in ("bf16", "f32"):
        got = run(gpu, fused, init, "c64", gpu.exec_opts("plain", dense=dense))
        assert infidelity(want, got) <= 1e-6, dense
        # norm drift of the matrix-core path (BF16: the state's 16-bit split; FP32: rounding)
        assert abs(np.vdot(got, got).real - 1) <= 1e-5, dense
    got = run(gpu, fused, init, "c128", gpu.exec_opts("windowed"))
    assert np.abs(got - want).max() <= 1e-12
