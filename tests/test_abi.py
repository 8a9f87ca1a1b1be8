"""CPU-side checks of the C ABI: the library loads, exports every declared symbol,
and its host-only helpers (the layered generator) match the oracle and the
reference builder. No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
import qlx
from qlx import ir

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "qlx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.findall(r"\b(qlx_[a-z0-9_]+)\s*\(", src)


def test_library_exports_every_declared_symbol():
    L = qlx.lib()
    names = declared_functions()
    assert len(names) >= 20
    import ctypes
    raw = ctypes.CDLL(qlx.LIB_PATH)
    for n in names:
        assert hasattr(raw, n), n
    # the ctypes signature table covers the header exactly
    assert sorted(n for n, _ in qlx._abi.SIGNATURES) == sorted(set(names))


def test_version_and_error_string():
    L = qlx.lib()
    assert L.qlx_version().startswith(b"qlx-gfx950")
    assert isinstance(L.qlx_last_error(), bytes)


def test_struct_sizes_match_header():
    assert C.sizeof(qlx._abi.ExecOpts) == 36
    assert C.sizeof(qlx._abi.ExecStats) == 48
    assert ir.GATE_DTYPE.itemsize == 16 and ir.META_DTYPE.itemsize == 12


@pytest.mark.parametrize("n,d,seed", [(1, 0, 1), (2, 3, 5), (20, 10, 1001), (30, 20, 1002), (32, 30, 1003)])
def test_library_generator_matches_oracle(n, d, seed):
    assert qlx.layered(n, d, seed).instrs.tobytes() == oracle.gen_layered(n, d, seed).tobytes()


def test_config_gate_counts():
    # 20q depth 10: 315 gates (20 H, 200 RX, 95 CX); 30q depth 20: 920; 32q depth 30: 1457
    for n, d, total in [(20, 10, 315), (30, 20, 920), (32, 30, 1457)]:
        assert len(qlx.layered(n, d, 1)) == total


@pytest.mark.skipif(not oracle.ref_available(), reason="reference tree not built here")
def test_python_builder_matches_reference_builder():
    rng = np.random.default_rng(3)
    n = 6
    ops, params = [], []
    c = qlx.Circuit(n)
    for _ in range(60):
        t = int(rng.choice([0x00, 0x01, 0x02, 0x03, 0x04, 0x10, 0x11, 0x12, 0x20, 0x21, 0x30, 0x31, 0xF1]))
        k = 1 if t < 0x20 or t == 0xF1 else (2 if t < 0x30 else 3)
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        th = float(rng.uniform(-10, 10))
        ops.append([t] + qs + [0] * (3 - k))
        params.append(th)
        name = {0x00: "h", 0x01: "x", 0x02: "y", 0x03: "z", 0x04: "s", 0x10: "rx", 0x11: "ry", 0x12: "rz",
                0x20: "cx", 0x21: "swap", 0x30: "ccx", 0x31: "cswap", 0xF1: "measure"}[t]
        getattr(c, name)(*qs, *([th] if t in (0x10, 0x11, 0x12) else []))
    assert c.release().instrs.tobytes() == oracle.ref_build(n, ops, params).tobytes()


@pytest.mark.skipif(not oracle.ref_available(), reason="reference tree not built here")
def test_python_builder_errors_match_reference():
    # std::out_of_range -> IndexError, std::invalid_argument -> ValueError (frontend.hpp:410-477)
    for ops, exc in [([[0x00, 5, 0, 0]], IndexError), ([[0x20, 1, 1, 0]], ValueError),
                     ([[0x30, 0, 1, 1]], ValueError), ([[0x00, -1, 0, 0]], IndexError)]:
        with pytest.raises(exc):
            oracle.ref_build(3, ops)
        c = qlx.Circuit(3)
        t, q0, q1, q2 = ops[0]
        with pytest.raises(exc):
            {0x00: lambda: c.h(q0), 0x20: lambda: c.cx(q0, q1), 0x30: lambda: c.ccx(q0, q1, q2)}[t]()


def test_mt19937_64_matches_standard():
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937_64
    g = qlx.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_measure_draw_matches_oracle_engine():
    # The host draw (uniform_real_distribution over mt19937_64) decides H|0> then MEASURE
    # exactly as the oracle's std:: types do, for many seeds.
    c = qlx.Circuit(1).h(0).measure(0).release()
    for seed in range(64):
        _, m = oracle.execute(1, c.instrs, seed=seed)
        r = qlx.MT19937_64(seed).uniform01()
        assert m[0] == (1 if r < 0.5 else 0)


def test_batch_workload_shapes():
    w = qlx.batch_workload(4096)
    ns = [n for n, _, _ in w]
    ds = [d for _, d, _ in w]
    assert min(ns) == 10 and max(ns) == 20 and min(ds) == 5 and max(ds) == 40
    assert w[0][2] == 400000 and w[-1][2] == 404095
    # the draw is the mt19937_64 stream of the stated seed
    g = qlx.MT19937_64(4004)
    assert w[0][0] == 10 + g() % 11 and w[0][1] == 5 + g() % 36


@pytest.mark.skipif(not oracle.ref_available(), reason="reference tree not built here")
@pytest.mark.parametrize("kind,n,d", [("hea", 2, 1), ("hea", 5, 3), ("hea", 12, 4), ("hea", 20, 7),
                                      ("qft", 1, 0), ("qft", 2, 0), ("qft", 5, 0), ("qft", 16, 0)])
def test_benchmark_generators_match_reference(kind, n, d):
    """make_hea_circuit / make_qft_circuit (circuit_generators.hpp:22-68): identical bytes."""
    mine = qlx.hea(n, d) if kind == "hea" else qlx.qft(n)
    assert mine.instrs.tobytes() == oracle.ref_gen_bench(kind, n, d).tobytes()
