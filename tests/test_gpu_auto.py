"""The drop-in default execution (QLX_EXEC_AUTO, what HIPEngineOptions{} selects) and background
specialisation: results equal the oracle's both while the generic sweep kernel runs and after the
compiled sweeps take over; small states stay on the generic kernel; out-of-place sweeps fall back
to in place when the second buffer does not fit."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def test_auto_mode_matches_oracle_before_and_after_background_compile(gpu):
    n = 22
    prog = gpu.layered(n, 8, 2201)
    init = oracle.random_state(n, 2202)
    want, _ = oracle.execute(n, prog.instrs, amps=init)
    o = gpu.exec_opts("auto")
    st = gpu.State(dtype="c128", amps=init)
    st.execute(prog, o)
    st.sync()
    first = st.last_stats()
    assert np.abs(st.to_host() - want).max() <= TOL
    assert gpu.specialize_wait(300.0) == 0
    st.load(init)
    st.execute(prog, o)
    st.sync()
    after = st.last_stats()
    print(f"auto 22q: first execution {'compiled' if first['specialised'] else 'generic'} sweeps, "
          f"then specialised={after['specialised']} relabelled={after['relabelled']}", flush=True)
    assert after["specialised"] == 1 and after["relabelled"] == 1, after
    assert after["gates_applied"] == len(prog)
    assert np.abs(st.to_host() - want).max() <= TOL
    # another state (another engine) picks the compiled sweeps up at once
    st2 = gpu.State(dtype="c128", amps=init)
    st2.execute(prog, o)
    st2.sync()
    assert st2.last_stats()["specialised"] == 1
    assert np.abs(st2.to_host() - want).max() <= TOL
    st2.close()
    st.close()


def test_auto_mode_complex64_matches_oracle(gpu):
    n = 21
    prog = gpu.layered(n, 6, 2301)
    init = oracle.random_state(n, 2302)
    want, _ = oracle.execute(n, prog.instrs, amps=init)
    o = gpu.exec_opts("auto")
    st = gpu.State(dtype="c64", amps=init)
    for _ in range(2):
        st.load(init)
        st.execute(prog, o)
        st.sync()
        got = st.to_host()
        fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
        assert 1.0 - fid <= 1e-6
        gpu.specialize_wait(300.0)
    assert st.last_stats()["specialised"] == 1
    st.close()


def test_auto_mode_small_state_stays_generic(gpu):
    n = 14
    prog = gpu.layered(n, 5, 2401)
    init = oracle.random_state(n, 2402)
    want, _ = oracle.execute(n, prog.instrs, amps=init)
    st = gpu.State(dtype="c128", amps=init)
    st.execute(prog, gpu.exec_opts("auto"))
    st.sync()
    gpu.specialize_wait(300.0)
    st.load(init)
    st.execute(prog, gpu.exec_opts("auto"))
    st.sync()
    s = st.last_stats()
    assert s["specialised"] == 0 and s["sweeps_full"] >= 1
    assert np.abs(st.to_host() - want).max() <= TOL
    st.close()


def test_background_flag_in_place(gpu):
    """windowed + specialize + background without relabel: compiled in-place sweeps afterwards."""
    n = 17
    prog = gpu.layered(n, 7, 2501)
    init = oracle.random_state(n, 2502)
    want, _ = oracle.execute(n, prog.instrs, amps=init)
    o = gpu.exec_opts("windowed", specialize=True, background=True)
    st = gpu.State(dtype="c128", amps=init)
    st.execute(prog, o)
    st.sync()
    assert np.abs(st.to_host() - want).max() <= TOL
    gpu.specialize_wait(300.0)
    st.load(init)
    st.execute(prog, o)
    st.sync()
    s = st.last_stats()
    assert s["specialised"] == 1 and s["relabelled"] == 0, s
    assert np.abs(st.to_host() - want).max() <= TOL
    st.close()


def test_relabel_falls_back_in_place_when_the_second_buffer_does_not_fit(gpu):
    n = 26
    prog = gpu.layered(n, 6, 2601)
    init = oracle.random_state(n, 2602)
    ref = gpu.State(dtype="c128", amps=init)
    ref.execute(prog, gpu.exec_opts("windowed"))     # generic sweeps, pinned to the oracle elsewhere
    want = ref.to_host()
    ref.close()
    target = gpu.State(dtype="c128", amps=init)      # 1 GiB
    fillers = []
    try:
        # take HBM until well under the target's size is free: its second buffer cannot be
        # allocated (the reported free bytes can understate what an allocation finds)
        for _ in range(64):
            free = gpu.device_info(0)["hbm_free_bytes"]
            if free < (3 << 27):                        # < 0.375 GiB free
                break
            k = int(np.floor(np.log2(max(free - (1 << 27), 1 << 24) / 8)))
            try:
                fillers.append(gpu.State(k, "c64"))
            except gpu.QlxOutOfMemory:
                k -= 1
                fillers.append(gpu.State(k, "c64"))
        free = gpu.device_info(0)["hbm_free_bytes"]
        print(f"fallback test: {len(fillers)} filler states, {free / 2**20:.0f} MiB free", flush=True)
        assert free < (1 << 29)
        target.execute(prog, gpu.exec_opts("windowed", specialize=True, relabel=True))
        target.sync()
        s = target.last_stats()
        assert s["specialised"] == 1 and s["relabelled"] == 0, s
    finally:
        for f in fillers:
            f.close()
    assert np.abs(target.to_host() - want).max() <= TOL
    # with memory back, the same options run out of place
    target.load(init)
    target.execute(prog, gpu.exec_opts("windowed", specialize=True, relabel=True))
    target.sync()
    assert target.last_stats()["relabelled"] == 1
    assert np.abs(target.to_host() - want).max() <= TOL
    target.close()
