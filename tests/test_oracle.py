"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Sources of truth, none of which is the oracle's own formula:
  * the reference's known-answer tests (tests/backend/test_engine_kernels.cpp,
    test_engine_measure.cpp, tests/integration/algos_no_optimization.cpp,
    tests/frontend/test_circuit_runner.cpp, tests/middleware/test_fusion_math.cpp);
  * the reference's own headers built in oracle/_ref (builder bytes, layout, state init);
  * an independent numpy dense brute force (tests/bruteforce.py);
  * closed forms (H^N|0>, the layered circuit's global phase, inverse circuits).
"""
import os

import numpy as np
import pytest

import bruteforce as bf
import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
INV_SQRT2 = 1.0 / np.sqrt(2.0)


def G(t, q0=255, q1=255, q2=255, param=0.0, mat=0):
    """make_gate helper of the reference tests (test_engine_kernels.cpp:22-46)."""
    g = np.zeros(1, oracle.GATE_DTYPE)
    g["type"] = t
    qs = [q0, q1, q2, 255]
    g["qubits"] = qs
    ar = 0
    for i, q in enumerate((q0, q1, q2)):
        if q != 255:
            ar = i + 1
    g["flags_arity"] = ar
    if t in (0x10, 0x11, 0x12):
        g["payload"] = np.array([param], np.float32).view(np.uint32)[0]
    elif t in (0x80, 0xE0):
        g["payload"] = mat
    return g


def prog(*gates):
    return np.concatenate(gates)


def run(n, *gates, **kw):
    return oracle.execute(n, prog(*gates), **kw)


def meta_for(mats):
    meta = np.zeros(len(mats), np.dtype([("offset", "<u4"), ("num_elements", "<u4"), ("num_qubits", "u1"),
                                          ("pad", "u1", (3,))]))
    pool, off = [], 0
    for i, (M, k) in enumerate(mats):
        meta[i] = (off, M.size, k, (0, 0, 0))
        pool.append(M.reshape(-1))
        off += M.size
    return np.concatenate(pool).astype(np.complex128), meta


# ---------------- known answers: test_engine_kernels.cpp -------------------------------
def test_single_qubit_known_answers():
    s, _ = run(1, G(0x01, 0))
    assert abs(s[0]) < 1e-6 and abs(s[1].real - 1) < 1e-6                          # :52-58
    s, _ = run(1, G(0x00, 0))
    assert abs(s[0].real - INV_SQRT2) < 1e-6 and abs(s[1].real - INV_SQRT2) < 1e-6   # :60-67
    s, _ = run(1, G(0x02, 0))
    assert abs(s[0]) < 1e-6 and abs(s[1].imag - 1) < 1e-6                           # :69-76
    s, _ = run(1, G(0x01, 0), G(0x03, 0))
    assert abs(s[1].real + 1) < 1e-6                                                # :78-84
    s, _ = run(1, G(0x01, 0), G(0x04, 0))
    assert abs(s[1].imag - 1) < 1e-6                                                # :86-92


def test_parametric_known_answers():
    s, _ = run(1, G(0x10, 0, param=np.pi))
    assert abs(s[1].imag + 1) < 1e-6                                                # :99-105
    s, _ = run(1, G(0x11, 0, param=np.pi / 2))
    assert abs(s[0].real - INV_SQRT2) < 1e-6 and abs(s[1].real - INV_SQRT2) < 1e-6   # :107-115


def test_routing_known_answers():
    assert abs(run(3, G(0x01, 2), G(0x20, 2, 0))[0][5].real - 1) < 1e-6              # :122-127
    assert abs(run(3, G(0x01, 0), G(0x20, 0, 2))[0][5].real - 1) < 1e-6              # :129-134
    assert abs(run(3, G(0x01, 0), G(0x21, 0, 2))[0][4].real - 1) < 1e-6              # :136-140
    assert abs(run(3, G(0x01, 2), G(0x21, 2, 0))[0][1].real - 1) < 1e-6              # :144-147
    assert abs(run(3, G(0x01, 0), G(0x01, 1), G(0x30, 1, 0, 2))[0][7].real - 1) < 1e-6  # :150-159
    assert abs(run(3, G(0x01, 0), G(0x01, 1), G(0x31, 1, 0, 2))[0][6].real - 1) < 1e-6  # :161-170


def test_custom_unitary_gather_scatter():
    # 4x4 SWAP on non-adjacent (q0, q2) (test_engine_kernels.cpp:173-202)
    sw = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], np.complex128)
    pool, meta = meta_for([(sw, 2)])
    s, _ = oracle.execute(3, prog(G(0x01, 0), G(0x80, 0, 2, mat=0)), pool, meta)
    assert abs(s[1]) < 1e-6 and abs(s[4].real - 1) < 1e-6


def test_phase_kickback():
    s, _ = run(2, G(0x00, 0), G(0x01, 1), G(0x00, 1), G(0x20, 0, 1))               # :204-229
    np.testing.assert_allclose(s.real, [0.5, -0.5, -0.5, 0.5], atol=1e-6)


# ---------------- known answers: measurement and algorithms -----------------------------
def test_born_rule_and_seed_determinism():
    outs = []
    for seed in range(20):
        s, m = run(1, G(0x00, 0), G(0xF1, 0), seed=seed)
        assert len(m) == 1
        assert abs(abs(s[m[0]]) ** 2 - 1) < 1e-6 and abs(s[1 - m[0]]) < 1e-6       # test_engine_measure.cpp:46-71
        s2, m2 = run(1, G(0x00, 0), G(0xF1, 0), seed=seed)
        assert m == m2                                                             # :73-87
        outs.append(m[0])
    assert 0 < sum(outs) < 20


def test_bell_correlation():
    for seed in range(100):                                                        # :103-123
        _, m = run(2, G(0x00, 0), G(0x20, 0, 1), G(0xF1, 0), G(0xF1, 1), seed=seed)
        assert m[0] == m[1]


def test_deutsch_all_oracles():
    init = np.zeros(4, complex)
    init[2] = 1.0                                                                  # algos_no_optimization.cpp:18-95
    cases = [([], True), ([G(0x01, 1)], True), ([G(0x20, 0, 1)], False), ([G(0x20, 0, 1), G(0x01, 0)], False)]
    for orc, const in cases:
        gates = [G(0x00, 0), G(0x00, 1)] + orc + [G(0x00, 0)]
        s, _ = oracle.execute(2, prog(*gates), amps=init, seed=37)
        p = np.abs(s) ** 2
        exp = [0.5, 0, 0.5, 0] if const else [0, 0.5, 0, 0.5]
        np.testing.assert_allclose(p, exp, atol=1e-6)


def test_teleportation_seed37():
    alpha, beta = 1 / np.sqrt(3), 1j * np.sqrt(2 / 3)                             # :98-160
    init = np.zeros(8, complex)
    init[0], init[1] = alpha, beta
    alice = prog(G(0x00, 1), G(0x20, 1, 2), G(0x20, 0, 1), G(0x00, 0), G(0xF1, 0), G(0xF1, 1))
    s, m = oracle.execute(3, alice, amps=init, seed=37)
    bob = [G(0x01, 2)] if m[1] else []
    bob += [G(0x03, 2)] if m[0] else []
    if bob:
        s, _ = oracle.execute(3, prog(*bob), amps=s)
    i0, i1 = (m[1] << 1) | m[0], 4 | (m[1] << 1) | m[0]
    assert abs(abs(s[i0]) ** 2 - abs(alpha) ** 2) < 1e-6
    assert abs(abs(s[i1]) ** 2 - abs(beta) ** 2) < 1e-6
    r = s[i1] / s[i0]
    assert abs(r - beta / alpha) < 1e-6


# ---------------- closed forms (test_circuit_runner.cpp:355-395 and the generator) -------
@pytest.mark.parametrize("n", [4, 6, 11])
def test_uniform_superposition(n):
    s, _ = oracle.execute(n, prog(*[G(0x00, q) for q in range(n)]))
    np.testing.assert_allclose(s, np.full(1 << n, 2 ** (-n / 2)), atol=1e-12)
    assert abs(np.sum(np.abs(s) ** 2) - 1) < 1e-12


def layered_global_phase(ins):
    th = ins["payload"][ins["type"] == 0x10].view(np.float32).astype(np.float64)
    return np.exp(-0.5j * np.sum(th))


@pytest.mark.parametrize("n,d,seed", [(5, 7, 3), (10, 10, 1001), (14, 6, 9)])
def test_layered_from_zero_is_global_phase(n, d, seed):
    ins = oracle.gen_layered(n, d, seed)
    s, _ = oracle.execute(n, ins)
    np.testing.assert_allclose(s, np.full(1 << n, layered_global_phase(ins) * 2 ** (-n / 2)), atol=1e-12)


@pytest.mark.parametrize("n,d,seed", [(9, 8, 4), (13, 10, 5)])
def test_inverse_circuit_identity(n, d, seed):
    ins = oracle.gen_layered(n, d, seed)
    inv = ins[::-1].copy()
    rx = inv["type"] == 0x10
    inv["payload"][rx] = (-inv["payload"][rx].view(np.float32)).view(np.uint32)
    init = bf.random_state(n, np.random.default_rng(seed))
    s, _ = oracle.execute(n, ins, amps=init)
    assert np.abs(s - init).max() > 1e-3
    s, _ = oracle.execute(n, inv, amps=s)
    np.testing.assert_allclose(s, init, atol=1e-12)


# ---------------- numpy dense brute force ---------------------------------------------------
def random_program(n, ngates, rng, with_unitaries=True):
    gates, mats = [], []
    types = [0x00, 0x01, 0x02, 0x03, 0x04, 0x10, 0x11, 0x12, 0x20, 0x21, 0x30, 0x31]
    for _ in range(ngates):
        t = int(rng.choice(types + ([0x80] if with_unitaries else [])))
        if t == 0x80:
            k = int(rng.integers(1, min(5, n) + 1))
            qs = [int(x) for x in rng.choice(n, size=k, replace=False)]
            mats.append((bf.random_unitary(k, rng), k))
            g = G(0x80, mat=len(mats) - 1)
            q4 = qs[:4] + [255] * (4 - min(4, k))
            g["qubits"] = q4
            g["flags_arity"] = k
            if k == 5:
                g["pad"] = qs[4]
            gates.append(g)
            continue
        k = 1 if t < 0x20 else (2 if t < 0x30 else 3)
        qs = [int(x) for x in rng.choice(n, size=k, replace=False)] + [255] * (3 - k)
        gates.append(G(t, *qs, param=float(rng.uniform(-7, 7))))
    pool, meta = meta_for(mats) if mats else (None, None)
    return prog(*gates), pool, meta


@pytest.mark.parametrize("seed", range(6))
def test_random_programs_vs_bruteforce(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(3, 9))
    ins, pool, meta = random_program(n, 40, rng)
    init = bf.random_state(n, rng)
    s, _ = oracle.execute(n, ins, pool, meta, amps=init)
    ref = bf.run(n, ins, pool, meta, psi=init)
    np.testing.assert_allclose(s, ref, atol=1e-12)


def test_global_swap_reference_semantics():
    # Worked example: x(0).x(1).cx(8,9) windowed with k_safe=2 -> X0 X1 GS(8,0) GS(9,1) CX(0,1)
    # (test_middleware.cpp:94-140). The reference engine ignores GLOBAL_SWAP.
    p = prog(G(0x01, 0), G(0x01, 1), G(0xE1, 8, 0), G(0xE1, 9, 1), G(0x20, 0, 1))
    p["flags_arity"][2:4] = 2
    s, _ = oracle.execute(10, p, gswap="noop")
    assert abs(s[1] - 1) < 1e-12                   # the reference's (wrong) answer
    s, _ = oracle.execute(10, p, gswap="permute")
    assert abs(s[768] - 1) < 1e-12                 # a real permutation without restore


# ---------------- fixtures -------------------------------------------------------------------
def fixtures():
    return sorted(f for f in os.listdir(GOLD) if f.endswith(".npz"))


@pytest.mark.parametrize("fname", fixtures())
def test_fixture_reproduced_and_bruteforced(fname):
    z = np.load(os.path.join(GOLD, fname))
    ins = z["instrs"].reshape(-1).view(oracle.GATE_DTYPE)
    n = int(fname.split("_n")[1].split("_")[0])
    d = int(fname.split("_d")[1].split("_")[0])
    seed = int(fname.split("_s")[1].split(".")[0])
    assert oracle.gen_layered(n, d, seed).tobytes() == ins.tobytes()
    s, _ = oracle.execute(n, ins)
    assert np.array_equal(s, z["state"])
    s, _ = oracle.execute(n, ins, amps=z["init_random"])
    assert np.array_equal(s, z["state_random"])
    if n <= 10:
        np.testing.assert_allclose(bf.run(n, ins, psi=z["init_random"]), z["state_random"], atol=1e-12)
        np.testing.assert_allclose(bf.run(n, ins), z["state"], atol=1e-12)


# ---------------- FusionMath (test_fusion_math.cpp) --------------------------------------------
def test_fuse_identical_footprint():
    X = np.array([[0, 1], [1, 0]], complex)
    Z = np.array([[1, 0], [0, -1]], complex)
    F, q = oracle.fuse_matrices(X, [0], Z, [0])                                       # :18-47
    assert q == [0]
    np.testing.assert_allclose(F.real, [[0, 1], [-1, 0]])


def test_fuse_disjoint_kronecker_order():
    X = np.array([[0, 1], [1, 0]], complex)
    Z = np.array([[1, 0], [0, -1]], complex)
    F, q = oracle.fuse_matrices(X, [0], Z, [1])                                       # :49-89
    assert q == [0, 1]
    np.testing.assert_allclose(F, np.kron(Z, X))


def test_fuse_partial_overlap_cnot_cz():
    CN = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]], complex)
    CZ = np.diag([1, 1, 1, -1]).astype(complex)
    F, q = oracle.fuse_matrices(CN, [0, 1], CZ, [1, 2])                               # :91-149
    assert q == [0, 1, 2] and F.shape == (8, 8)
    assert F[1, 3] == 1 and F[3, 3] == 0 and F[5, 7] == 1 and F[7, 5] == -1
    # independent check: product of the brute-force expansions
    np.testing.assert_allclose(F, bf.expand(CZ, [1, 2], 3) @ bf.expand(CN, [0, 1], 3))


@pytest.mark.parametrize("seed", range(4))
def test_fuse_random_vs_bruteforce(seed):
    rng = np.random.default_rng(seed)
    k1, k2 = int(rng.integers(1, 4)), int(rng.integers(1, 3))
    q1 = [int(x) for x in rng.choice(5, k1, replace=False)]
    q2 = [int(x) for x in rng.choice(5, k2, replace=False)]
    U1, U2 = bf.random_unitary(k1, rng), bf.random_unitary(k2, rng)
    F, q = oracle.fuse_matrices(U1, q1, U2, q2)
    assert q == sorted(set(q1) | set(q2))
    full = bf.expand(U2, [q.index(x) for x in q2], len(q)) @ bf.expand(U1, [q.index(x) for x in q1], len(q))
    np.testing.assert_allclose(F, full, atol=1e-13)


# ---------------- state init (cpu_state_vector.hpp) ----------------------------------------------
def test_state_checks():
    assert oracle.check_state(np.ones(3) / np.sqrt(3)) == -1
    assert oracle.check_state(np.ones(4)) == -2
    assert oracle.check_state(np.ones(4) / 2) == 0


# ---------------- the reference's own headers (oracle/_ref) --------------------------------------
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference tree not built here")


@needs_ref
def test_reference_layout():
    # sizeof, alignof, offsets of GateInstr; sizeof MatrixEntry (frontend.hpp:84-123)
    assert oracle.ref_layout() == [16, 16, 1, 2, 6, 8, 12, 12]


@needs_ref
@pytest.mark.parametrize("n,d,seed", [(4, 3, 11), (12, 10, 1001), (20, 10, 1001), (30, 20, 1002)])
def test_generator_matches_reference_builder(n, d, seed):
    assert oracle.ref_gen_layered(n, d, seed).tobytes() == oracle.gen_layered(n, d, seed).tobytes()


@needs_ref
def test_insert_zero_bit_matches_reference():
    rng = np.random.default_rng(7)
    for _ in range(200):
        idx, t = int(rng.integers(0, 2 ** 40)), int(rng.integers(0, 40))
        lo = idx & ((1 << t) - 1)
        assert oracle.ref().ref_insert_zero_bit(idx, t) == lo | ((idx ^ lo) << 1)


@needs_ref
def test_state_init_matches_reference():
    import ctypes as C
    out = np.zeros(8, complex)
    oracle.ref().ref_state_zero(3, out.ctypes.data)
    np.testing.assert_array_equal(out, oracle.zero_state(3))
    for v in (np.ones(3) / np.sqrt(3), np.ones(4), np.ones(4) / 2, np.ones(8) * (1 / np.sqrt(8) + 1e-8)):
        a = np.ascontiguousarray(v, complex)
        ref_rc = oracle.ref().ref_state_check(a.ctypes.data, C.c_size_t(a.size))
        assert (ref_rc == 0) == (oracle.check_state(a) == 0)
