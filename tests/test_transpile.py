"""Middleware::transpile_gpu (qlx_transpile): the reference's middleware known answers
(tests/middleware/test_middleware.cpp) and numerical equivalence on the CPU oracle."""
import numpy as np
import pytest

import bruteforce as bf
import oracle
import qlx
from qlx import ir
from test_oracle import random_program

CPU_STYLE = dict(rotations=False, max_width=4, restore=False)


def names(p):
    return [n for n, _ in p.describe()]


# ---- test_middleware.cpp:60-92 (fusion heuristic) ---------------------------------------
def test_fusion_skips_two_standard_gates():
    p = qlx.transpile_gpu(qlx.Circuit(2).h(0).x(1).release(), **CPU_STYLE)
    assert names(p) == ["H", "X"]


def test_fusion_three_connected_gates_into_one_block():
    p = qlx.transpile_gpu(qlx.Circuit(2).h(0).cx(0, 1).x(1).release(), **CPU_STYLE)
    assert names(p) == ["FUSED_BLOCK"] and ir.arity(p.instrs[0]) == 2
    want = bf.expand(np.array([[0, 1], [1, 0]]), [1], 2) @ bf.expand(
        bf.gate_matrix(qlx.Circuit(2).cx(0, 1).release().instrs[0])[0], [0, 1], 2) @ bf.expand(
        np.array([[1, 1], [1, -1]]) / np.sqrt(2), [0], 2)
    np.testing.assert_allclose(p.pool.reshape(4, 4), want, atol=1e-15)


def test_reference_cannot_fuse_rotations_but_gpu_pass_can():
    c = qlx.Circuit(2).h(0).rx(0, 0.3).cx(0, 1).rx(1, 1.1).release()
    assert names(qlx.transpile_gpu(c, **CPU_STYLE)) == ["H", "RX", "CX", "RX"]
    p = qlx.transpile_gpu(c, rotations=True, max_width=2)
    assert names(p) == ["FUSED_BLOCK"]


# ---- test_middleware.cpp:94-140, 164-235, 296-323 (windowing stream shape) ---------------
def windowed(k_safe, numa=1, swap=0, restore=False):
    p = qlx.Circuit(10).x(0).x(1).cx(8, 9).release()
    return qlx.transpile_gpu(p, fusion=False, windowing=True, k_safe=k_safe, numa_penalty=numa,
                             swap_penalty=swap, restore=restore)


def test_windowing_tlb_hazard_stream():
    p = windowed(2, 1, 0)
    d = p.describe()
    assert [n for n, _ in d] == ["X", "X", "GLOBAL_SWAP", "GLOBAL_SWAP", "CX"]
    assert d[0][1] == [0] and d[1][1] == [1]
    for k in (2, 3):
        assert d[k][1][0] in (8, 9) and d[k][1][1] in (0, 1)
    assert d[2][1][0] != d[3][1][0] and d[2][1][1] != d[3][1][1]
    assert set(d[4][1]) == {0, 1}


def test_windowing_penalties():
    assert names(windowed(2, 1, 100)) == ["X", "X", "CX"]                     # :168-189
    assert names(windowed(2, 10, 1))[2:] == ["GLOBAL_SWAP", "GLOBAL_SWAP", "CX"]  # :191-212
    assert len(windowed(16, 10, 1)) == 3                                        # :304-312
    assert len(windowed(2, 10, 1)) == 5                                         # :314-322


def test_windowing_restore_appends_inverse_layout():
    p = windowed(2, 1, 0, restore=True)
    assert names(p)[:5] == ["X", "X", "GLOBAL_SWAP", "GLOBAL_SWAP", "CX"]
    assert all(n == "GLOBAL_SWAP" for n in names(p)[5:]) and len(p) > 5
    # with real swaps + restore the program computes the logical answer |..0011> = index 3
    s, _ = oracle.execute(10, p.instrs, gswap="permute")
    assert abs(s[3] - 1) < 1e-12
    # the reference's own semantics (no-op swap, no restore) give index 1 (a wrong state)
    s, _ = oracle.execute(10, windowed(2, 1, 0).instrs, gswap="noop")
    assert abs(s[1] - 1) < 1e-12


# ---- test_middleware.cpp:237-294 (custom cost model) and the greedy model ----------------
def threshold_model(k_safe, numa, swap):
    """ThresholdCostModel::evaluate_swap (middleware.hpp:76-112), restated for the checks."""
    def evaluate(freq, l2p, p2l):
        cand, best = None, 0
        for q, f in enumerate(freq):
            if f > 0 and l2p[q] > k_safe:
                r = f * numa
                if r >= swap and r > best:
                    best, cand = r, q
        if cand is None:
            return None
        for v, f in enumerate(freq):
            if l2p[v] <= k_safe and f == 0:
                return cand, v
        return None
    return evaluate


def test_custom_cost_model_is_invoked():
    calls = []

    def spy(freq, l2p, p2l):          # CustomSpyCostModel: evaluates, never swaps
        calls.append((list(freq), list(l2p), list(p2l)))
        return None
    p = qlx.Circuit(6).h(0).cx(4, 5).release()
    out = qlx.transpile_gpu(p, fusion=False, windowing=True, k_safe=3, cost_model=spy)
    assert calls and len(out) == 2
    freq, l2p, p2l = calls[0]
    assert len(freq) == 6 and l2p == list(range(6)) and p2l == list(range(6))
    assert freq[0] == 1 and freq[4] == 1 and freq[5] == 1


def test_custom_cost_model_type_without_a_model_is_rejected():
    with pytest.raises(ValueError):
        qlx.transpile_gpu(qlx.Circuit(4).h(0).release(), fusion=False, windowing=True, cost_model="custom")


def test_custom_threshold_model_equals_the_built_in_one():
    rng = np.random.default_rng(21)
    for trial in range(4):
        ins, pool, meta = random_program(12, 60, rng)
        prog = qlx.Program(12, ins, pool if pool is not None else np.zeros(0, complex),
                           meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
        for k_safe, numa, swap in ((4, 10, 1), (6, 1, 3), (3, 1, 0)):
            built = qlx.transpile_gpu(prog, fusion=False, windowing=True, k_safe=k_safe, numa_penalty=numa,
                                      swap_penalty=swap)
            custom = qlx.transpile_gpu(prog, fusion=False, windowing=True, k_safe=k_safe,
                                       cost_model=threshold_model(k_safe, numa, swap))
            assert built.instrs.tobytes() == custom.instrs.tobytes(), (trial, k_safe)


def test_greedy_cost_model_is_threshold_with_penalties_one_and_zero():
    p = qlx.Circuit(10).x(0).x(1).cx(8, 9).release()
    greedy = qlx.transpile_gpu(p, fusion=False, windowing=True, k_safe=2, numa_penalty=100, swap_penalty=100,
                               restore=False, cost_model="greedy")
    assert names(greedy) == ["X", "X", "GLOBAL_SWAP", "GLOBAL_SWAP", "CX"]
    assert greedy.instrs.tobytes() == windowed(2, 1, 0).instrs.tobytes()


def test_custom_cost_model_errors_abort_the_pass():
    p = qlx.Circuit(10).x(0).x(1).cx(8, 9).release()

    def broken(freq, l2p, p2l):
        raise RuntimeError("model failure")
    with pytest.raises(ValueError):
        qlx.transpile_gpu(p, fusion=False, windowing=True, k_safe=2, cost_model=broken)
    with pytest.raises(ValueError):         # never converges: always proposes the same swap
        qlx.transpile_gpu(p, fusion=False, windowing=True, k_safe=2, cost_model=lambda f, a, b: (8, 0))
    with pytest.raises(ValueError):         # out-of-range qubit
        qlx.transpile_gpu(p, fusion=False, windowing=True, k_safe=2, cost_model=lambda f, a, b: (8, 40))


def test_custom_windowed_program_keeps_the_state():
    """A custom model's swaps are real permutations with the layout restored: same state."""
    rng = np.random.default_rng(5)
    ins, pool, meta = random_program(10, 50, rng)
    prog = qlx.Program(10, ins, pool if pool is not None else np.zeros(0, complex),
                       meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    out = qlx.transpile_gpu(prog, fusion=False, windowing=True, k_safe=4, restore=True,
                            cost_model=threshold_model(4, 1, 0))
    assert any(n == "GLOBAL_SWAP" for n in names(out))
    init = bf.random_state(10, rng)
    want, _ = oracle.execute(10, prog.instrs, prog.pool, prog.meta, amps=init)
    got, _ = oracle.execute(10, out.instrs, out.pool, out.meta, amps=init, gswap="permute")
    assert np.abs(got - want).max() <= 1e-12


def test_gpu_safe_qubit_limit():
    assert qlx.gpu_safe_qubit_limit(0) == 12
    assert qlx.gpu_safe_qubit_limit(1) == 13


# ---- numerical equivalence -------------------------------------------------------------------
@pytest.mark.parametrize("width", [2, 3, 4, 5])
@pytest.mark.parametrize("seed", range(4))
def test_fusion_preserves_the_state(width, seed):
    rng = np.random.default_rng(seed * 10 + width)
    n = int(rng.integers(5, 10))
    ins, pool, meta = random_program(n, 60, rng)
    prog = qlx.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                       meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    fused = qlx.transpile_gpu(prog, max_width=width, rotations=True)
    assert len(fused) <= len(prog)
    for g in fused.instrs:
        assert ir.arity(g) <= max(width, 5 if g["type"] == 0x80 else 0) or g["type"] != 0xE0
    init = bf.random_state(n, rng)
    a, _ = oracle.execute(n, ins, pool, meta, amps=init)
    b, _ = oracle.execute(n, fused.instrs, fused.pool, fused.meta, amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)


@pytest.mark.parametrize("n,d,seed,width", [(8, 6, 1, 5), (10, 10, 1001, 4), (12, 5, 3, 3)])
def test_layered_fusion_width(n, d, seed, width):
    p = qlx.layered(n, d, seed)
    f = qlx.transpile_gpu(p, max_width=width, rotations=True)
    blocks = [g for g in f.instrs if g["type"] == 0xE0]
    assert blocks and max(ir.arity(g) for g in blocks) == width
    assert len(f) < len(p) // 2
    init = bf.random_state(n, np.random.default_rng(seed))
    a, _ = oracle.execute(n, p.instrs, amps=init)
    b, _ = oracle.execute(n, f.instrs, f.pool, f.meta, amps=init)
    np.testing.assert_allclose(b, a, atol=1e-12)
    np.testing.assert_allclose(bf.run(n, f.instrs, f.pool, f.meta, psi=init), a, atol=1e-12)


@pytest.mark.parametrize("k_safe", [3, 5, 8])
def test_windowing_with_restore_preserves_the_state(k_safe):
    n = 10
    p = qlx.layered(n, 6, 77)
    w = qlx.transpile_gpu(p, fusion=False, windowing=True, k_safe=k_safe, numa_penalty=10, swap_penalty=1,
                          restore=True)
    assert (w.instrs["type"] == 0xE1).sum() > 0
    init = bf.random_state(n, np.random.default_rng(k_safe))
    a, _ = oracle.execute(n, p.instrs, amps=init)
    b, _ = oracle.execute(n, w.instrs, amps=init, gswap="permute")
    np.testing.assert_allclose(b, a, atol=1e-12)


def test_fusion_then_windowing_preserves_the_state():
    n = 11
    p = qlx.layered(n, 8, 5)
    w = qlx.transpile_gpu(p, max_width=3, rotations=True, windowing=True, k_safe=6, restore=True)
    init = bf.random_state(n, np.random.default_rng(1))
    a, _ = oracle.execute(n, p.instrs, amps=init)
    b, _ = oracle.execute(n, w.instrs, w.pool, w.meta, amps=init, gswap="permute")
    np.testing.assert_allclose(b, a, atol=1e-12)


# ---- the clustered benchmark circuit (benchmarks/bench_hw_awareness.cpp:30-53) -----------------
def test_clustered_circuit_shape():
    p = qlx.clustered(28)
    assert len(p) == 765                                   # bench.txt:9 "Gates: 765"
    t = p.instrs["type"]
    assert (t == 0x00).sum() == 28 * 15 and (t == 0x20).sum() == 23 * 15
    q = p.instrs["qubits"]
    assert list(q[6][:2]) == [0, 1] and list(q[764][:2]) == [26, 27]
    assert len(qlx.clustered(5, 2, 1)) == 5 + 2            # clusters {0,1}, {2,3}, {4}


def test_clustered_strategy_matrix_gate_counts():
    """The gate counts the reference prints at 28 qubits (bench.txt:9, 15, 21, 27): SWAP
    windowing at k_safe 18 -> 774, fusion -> 76, both -> 85 (no layout restore, as there)."""
    p = qlx.clustered(28)
    style = dict(max_width=4, rotations=False, restore=False, k_safe=18)
    assert len(qlx.transpile_gpu(p, fusion=False, windowing=True, **style)) == 774
    assert len(qlx.transpile_gpu(p, fusion=True, windowing=False, **style)) == 76
    assert len(qlx.transpile_gpu(p, fusion=True, windowing=True, **style)) == 85


@pytest.mark.parametrize("n,cluster,depth", [(8, 3, 4), (13, 6, 15)])
def test_clustered_product_form_on_oracle(n, cluster, depth):
    """Oracle (plain, fused to width 4 and 5, windowed with restore) against the product of
    brute-forced cluster states."""
    p = qlx.clustered(n, cluster, depth)
    want = bf.clustered_state(n, cluster, depth, lambda s: qlx.clustered(s, s, depth).instrs)
    got, _ = oracle.execute(n, p.instrs)
    assert np.abs(got - want).max() <= 1e-12
    for width in (4, 5):
        f = qlx.transpile_gpu(p, max_width=width, rotations=True, windowing=True, k_safe=n - 3)
        got, _ = oracle.execute(n, f.instrs, f.pool, f.meta, gswap="permute")
        assert np.abs(got - want).max() <= 1e-12, width
