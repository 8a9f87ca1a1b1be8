"""qlx_execute_noalloc: one kernel per gate with the unitary pool staged in the library's fixed device array instead of a
device allocation. Its results equal qlx_execute's in plain mode (the
same kernels) bit for bit, on two calls in a row, the second with a smaller state and pool."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def execute_noalloc(qlx, st, prog, opts=None):
    f = qlx.lib().qlx_execute_noalloc
    ins = np.ascontiguousarray(prog.instrs)
    pool = np.ascontiguousarray(prog.pool, np.complex128)
    meta = np.ascontiguousarray(prog.meta)
    stop = C.c_size_t(0)
    qlx._abi.check(f(st.handle, ins.ctypes.data if len(ins) else None, len(ins),
                     pool.ctypes.data if pool.size else None, pool.size,
                     meta.ctypes.data if len(meta) else None, len(meta),
                     C.byref(opts) if opts is not None else None, C.byref(stop)))
    return stop.value


def random_unitary(k, rng):
    d = 1 << k
    z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def mixed_program(qlx, n, seed, n_dense):
    """H/RX/CX/CCX/SWAP/RY gates with `n_dense` random 1..5-qubit matrices (pool grows with it)."""
    rng = np.random.default_rng(seed)
    c = qlx.Circuit(n)
    for q in range(n):
        c.h(q)
    for j in range(n_dense):
        k = 1 + j % 5
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        c.unitary(qs, random_unitary(k, rng), max_arity=5)
        a, b, t = (int(x) for x in rng.choice(n, 3, replace=False))
        c.rx(a, float(rng.uniform(0, 6.28))).cx(a, b).ccx(a, b, t).swap(b, t).ry(t, float(rng.uniform(0, 3)))
    return c.release()


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_equals_execute_two_calls(gpu, dtype):
    for n, n_dense, seed in ((15, 12, 9001), (11, 5, 9002)):     # second call: smaller state and pool
        prog = mixed_program(gpu, n, seed, n_dense)
        init = oracle.random_state(n, seed + 1)
        want = gpu.State(dtype=dtype, amps=init)
        want.execute(prog, gpu.exec_opts("plain", dense="valu"))
        want.sync()
        got = gpu.State(dtype=dtype, amps=init)
        assert execute_noalloc(gpu, got, prog) == len(prog)
        got.sync()
        assert got.last_stats()["gates_applied"] == want.last_stats()["gates_applied"]
        a, b = got.to_host(), want.to_host()
        print(f"noalloc {dtype} {n}q, pool {prog.pool.size} elements: max |d| = {np.abs(a - b).max():.3e}",
              flush=True)
        assert np.array_equal(a, b)
        if dtype == "c128":
            ref, _ = oracle.execute(n, prog.instrs, prog.pool, prog.meta, amps=init)
            assert np.abs(a - ref).max() <= 1e-12
        got.close()
        want.close()


def test_noalloc_limits(gpu):
    st = gpu.State(12, "c128")
    prog = mixed_program(gpu, 12, 9003, 2)
    with pytest.raises(gpu.QlxUnsupported):
        execute_noalloc(gpu, st, prog, gpu.exec_opts("windowed"))
    # a pool larger than the staging area (65536 complex elements): 69 five-qubit matrices
    big = gpu.Circuit(12)
    rng = np.random.default_rng(5)
    u = random_unitary(5, rng)
    for _ in range(69):
        big.unitary([0, 1, 2, 3, 4], u, max_arity=5)
    p = big.release()
    assert p.pool.size > 65536
    with pytest.raises(gpu.QlxInvalidArgument):
        execute_noalloc(gpu, st, p)
    # timing on: refused (events would be created)
    st.timing(True)
    with pytest.raises(gpu.QlxInvalidArgument):
        execute_noalloc(gpu, st, prog)
    st.timing(False)
    # MEASURE ends the run at its index, like qlx_execute
    m = gpu.Circuit(12).h(0).measure(0).x(1).release()
    assert execute_noalloc(gpu, st, m) == 1
    st.close()
