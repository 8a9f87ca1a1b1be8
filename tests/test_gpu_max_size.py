"""The largest states one MI355X holds (256 GiB): 34 qubits complex128 and 35 qubits complex64.

No second buffer fits beside them, so the out-of-place sweeps the options ask for run in place
(the fallback of the relabelled form). The checks use device reductions only (a host copy would
need 256 GiB): an RX layer on |0...0> gives a product state whose P(1) on qubit q is
sin^2(theta_q / 2) (theta stored as float); the layered circuit and its inverse must return to it.
From this product state the layered circuit does not just multiply the state by a phase, as it does
from |0...0>."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
@pytest.mark.parametrize("n,dtype,tol", [(34, "c128", 1e-12), (35, "c64", 1e-4)])
def test_largest_state_fits_and_runs(gpu, n, dtype, tol):
    try:
        st = gpu.State(n, dtype)
    except gpu.QlxOutOfMemory:
        pytest.skip(f"{n} qubits {dtype} (256 GiB) do not fit this device")
    try:
        rng = np.random.default_rng(n)
        theta = [float(np.float32(x)) for x in rng.uniform(0.2, 3.0, n)]
        c = gpu.Circuit(n)
        for q in range(n):
            c.rx(q, theta[q])
        layer = c.release()
        o = gpu.exec_opts("windowed", specialize=True, relabel=True)
        st.execute(layer, o)
        st.sync()
        want = np.sin(np.asarray(theta) / 2) ** 2
        qs = (0, 1, 12, n // 2, n - 2, n - 1)
        for q in qs:
            assert abs(st.prob1(q) - want[q]) <= tol, (q, st.prob1(q), want[q])
        prog = gpu.layered(n, 3, 3400 + n)
        st.execute(prog, o)
        st.execute(gpu.inverse_layered(prog), o)
        st.sync()
        s = st.last_stats()
        norm = st.norm_sq()
        got = [st.prob1(q) for q in qs]
        print(f"{n} qubits {dtype}: sweeps {s['sweeps_full']}, specialised {s['specialised']}, "
              f"relabelled {s['relabelled']}, norm^2 - 1 = {norm - 1:.2e}, "
              f"max |P1 - want| = {max(abs(g - want[q]) for g, q in zip(got, qs)):.2e}", flush=True)
        assert s["specialised"] == 1 and s["relabelled"] == 0, s
        assert abs(norm - 1) <= tol
        for g, q in zip(got, qs):
            assert abs(g - want[q]) <= tol, (q, g, want[q])
    finally:
        st.close()
