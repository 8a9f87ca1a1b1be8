"""Merged rotations on the GPU (the window-sweep planner's default): programs rich in
rotations with gates that commute with them in between, in every window-sweep form
(generic kernel, circuit-specialised in place and relabelled, complex128 and complex64),
against the CPU oracle on the program as written (complex128: max abs error <= 1e-12;
complex64: state infidelity <= 1e-6), and against the same form with every rotation kept."""
import numpy as np
import pytest

import bruteforce as bf
import oracle
from test_plan import rotation_rich

pytestmark = pytest.mark.gpu

FORMS = [("generic", dict()), ("specialised", dict(specialize=True)),
         ("relabelled", dict(specialize=True, relabel=True))]


def run(qlx, prog, init, dtype, **kw):
    st = qlx.State(amps=init, dtype=dtype)
    st.execute(prog, qlx.exec_opts("windowed", **kw))
    out = st.to_host()
    st.close()
    return out


@pytest.mark.parametrize("seed", range(3))
def test_merged_rotations_vs_oracle(gpu, seed):
    rng = np.random.default_rng(7100 + seed)
    n = 14
    prog = rotation_rich(n, 200, rng)
    merged = gpu.plan_windowed(prog, "c128", gpu.exec_opts("windowed", specialize=True))[2]["rotations_merged"]
    assert merged > 10
    init = bf.random_state(n, rng)
    want, _ = oracle.execute(n, prog.instrs, amps=init)
    for name, kw in FORMS:
        got = run(gpu, prog, init, "c128", **kw)
        err = float(np.abs(got - want).max())
        kept = run(gpu, prog, init, "c128", keep_rotations=True, **kw)
        err_kept = float(np.abs(kept - want).max())
        print(f"seed {seed} {name}: {merged} merged, max|d| {err:.2e} (kept {err_kept:.2e})", flush=True)
        assert err <= 1e-12 and err_kept <= 1e-12, name
        g64 = run(gpu, prog, init, "c64", **kw)
        inf = 1 - abs(np.vdot(want, g64)) ** 2 / np.vdot(g64, g64).real
        assert inf <= 1e-6, (name, inf)


def test_layered_circuit_merged_vs_oracle(gpu):
    """The layered h/rx/cx circuit (half of its RX merge) from a random start."""
    prog = gpu.layered(20, 10, 1001)
    init = bf.random_state(20, np.random.default_rng(1001))
    want, _ = oracle.execute(20, prog.instrs, amps=init)
    for name, kw in FORMS:
        got = run(gpu, prog, init, "c128", **kw)
        assert np.abs(got - want).max() <= 1e-12, name
