"""The Qlassical IR as numpy records, and a builder mirroring ``QuantumCircuit``.

Layouts are byte-identical to the reference:
  GateInstr    16 bytes   include/frontend/frontend.hpp:84-105
  MatrixEntry  12 bytes   include/frontend/frontend.hpp:118-123
  UnitaryPool  complex128 row-major matrices, 4^k entries each (frontend.hpp:117-164)

``Circuit`` follows ``QuantumCircuit`` (frontend.hpp:218-497): same gate methods,
argument order (qubit first, then angle), float narrowing of angles, validation
errors (IndexError for std::out_of_range, ValueError for std::invalid_argument)
and a running uid.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi

GATE_DTYPE = np.dtype([
    ("type", "u1"), ("flags_arity", "u1"), ("qubits", "u1", (4,)), ("pad", "<u2"),
    ("payload", "<u4"), ("uid", "<u4"),
])
META_DTYPE = np.dtype([
    ("offset", "<u4"), ("num_elements", "<u4"), ("num_qubits", "u1"), ("pad", "u1", (3,)),
])
assert GATE_DTYPE.itemsize == 16 and META_DTYPE.itemsize == 12

# GateType (frontend.hpp:26-51)
H, X, Y, Z, S = 0x00, 0x01, 0x02, 0x03, 0x04
RX, RY, RZ = 0x10, 0x11, 0x12
CX, SWAP = 0x20, 0x21
CCX, CSWAP = 0x30, 0x31
UNITARY, FUSED_BLOCK, GLOBAL_SWAP = 0x80, 0xE0, 0xE1
BARRIER, MEASURE = 0xF0, 0xF1
NO_MATRIX = 0xFFFFFFFF

GATE_NAMES = {H: "H", X: "X", Y: "Y", Z: "Z", S: "S", RX: "RX", RY: "RY", RZ: "RZ", CX: "CX",
              SWAP: "SWAP", CCX: "CCX", CSWAP: "CSWAP", UNITARY: "UNITARY", FUSED_BLOCK: "FUSED_BLOCK",
              GLOBAL_SWAP: "GLOBAL_SWAP", BARRIER: "BARRIER", MEASURE: "MEASURE"}


def f32_bits(x: float) -> int:
    return int(np.array([x], dtype=np.float32).view(np.uint32)[0])


def bits_f32(u: int) -> float:
    return float(np.array([u], dtype=np.uint32).view(np.float32)[0])


def arity(g) -> int:
    return int(g["flags_arity"]) & 0x0F


@dataclass
class Program:
    """An IRModule: gate stream + unitary pool."""
    num_qubits: int
    instrs: np.ndarray = field(default_factory=lambda: np.zeros(0, GATE_DTYPE))
    pool: np.ndarray = field(default_factory=lambda: np.zeros(0, np.complex128))
    meta: np.ndarray = field(default_factory=lambda: np.zeros(0, META_DTYPE))

    def __len__(self) -> int:
        return len(self.instrs)

    def store(self, matrix: np.ndarray, k: int) -> int:
        """UnitaryPool::store (frontend.hpp:132-150): returns the matrix index."""
        m = np.ascontiguousarray(matrix, dtype=np.complex128).reshape(-1)
        expected = 1 << (2 * k)
        if m.size != expected:
            raise ValueError(f"UnitaryPool::store: matrix size mismatch. Expected {expected}, got {m.size}")
        idx = len(self.meta)
        e = np.zeros(1, META_DTYPE)
        e["offset"] = len(self.pool)
        e["num_elements"] = expected
        e["num_qubits"] = k
        self.pool = np.concatenate([self.pool, m])
        self.meta = np.concatenate([self.meta, e])
        return idx

    def slice(self, lo: int, hi: int) -> "Program":
        return Program(self.num_qubits, self.instrs[lo:hi].copy(), self.pool, self.meta)

    def describe(self) -> list:
        out = []
        for g in self.instrs:
            a = arity(g)
            qs = [int(q) for q in g["qubits"][:min(a, 4)]]
            if a == 5:
                qs.append(int(g["pad"]) & 0xFF)
            out.append((GATE_NAMES.get(int(g["type"]), hex(int(g["type"]))), qs))
        return out


class Circuit:
    """Builder mirroring qlassical::QuantumCircuit (frontend.hpp:218-497)."""

    def __init__(self, num_qubits: int):
        self.num_qubits = int(num_qubits)
        self._rows: list = []
        self._prog = Program(self.num_qubits)
        self._uid = 0

    # -- validation (frontend.hpp:410-416, 446-477)
    def _validate(self, q: int) -> None:
        if q < 0 or q >= 256 or q >= self.num_qubits:
            raise IndexError(f"Qubit index {q} out of range [0, min(256, {self.num_qubits}))")

    def _push(self, type_: int, qubits, payload: int = 0) -> "Circuit":
        g = np.zeros(1, GATE_DTYPE)
        g["type"] = type_
        g["flags_arity"] = len(qubits) & 0x0F
        q4 = [255, 255, 255, 255]
        for i, q in enumerate(qubits[:4]):
            q4[i] = q
        g["qubits"] = q4
        if len(qubits) == 5:
            g["pad"] = qubits[4]
        g["payload"] = payload
        g["uid"] = self._uid
        self._uid += 1
        self._rows.append(g)
        return self

    def _1q(self, t, q):
        self._validate(q)
        return self._push(t, [q])

    def _1qp(self, t, q, theta):
        self._validate(q)
        return self._push(t, [q], f32_bits(float(np.float32(theta))))

    def _2q(self, t, q0, q1):
        self._validate(q0)
        self._validate(q1)
        if q0 == q1:
            raise ValueError(f"Two-qubit gate requires distinct qubits, got {q0} == {q1}")
        return self._push(t, [q0, q1])

    def _3q(self, t, q0, q1, q2):
        for q in (q0, q1, q2):
            self._validate(q)
        if q0 == q1 or q0 == q2 or q1 == q2:
            raise ValueError("Three-qubit gate requires distinct qubits")
        return self._push(t, [q0, q1, q2])

    def h(self, q): return self._1q(H, q)
    def x(self, q): return self._1q(X, q)
    def y(self, q): return self._1q(Y, q)
    def z(self, q): return self._1q(Z, q)
    def s(self, q): return self._1q(S, q)
    def rx(self, q, theta): return self._1qp(RX, q, theta)
    def ry(self, q, theta): return self._1qp(RY, q, theta)
    def rz(self, q, theta): return self._1qp(RZ, q, theta)
    def cx(self, c, t): return self._2q(CX, c, t)
    cnot = cx
    def swap(self, a, b): return self._2q(SWAP, a, b)
    def ccx(self, c0, c1, t): return self._3q(CCX, c0, c1, t)
    toffoli = ccx
    def cswap(self, c, a, b): return self._3q(CSWAP, c, a, b)
    fredkin = cswap

    def unitary(self, qubits, matrix, max_arity: int = 4) -> "Circuit":
        """QuantumCircuit::unitary (frontend.hpp:340-363): arity 1..4 as in the reference.
        max_arity=5 builds the width-5 form the GPU fusion pass emits (fifth qubit in the
        pad byte)."""
        k = len(qubits)
        if k == 0 or k > max_arity:
            raise ValueError(f"unitary: arity must be 1-{max_arity}, got {k}")
        midx = self._prog.store(matrix, k)
        for q in qubits:
            self._validate(q)
        if len(set(qubits)) != k:
            raise ValueError("unitary requires distinct qubits")
        return self._push(UNITARY, list(qubits), midx)

    def barrier(self, qubits=()) -> "Circuit":
        qs = list(qubits)[:4]
        for q in qs:
            self._validate(q)
        return self._push(BARRIER, qs)

    def measure(self, q) -> "Circuit":
        if isinstance(q, (list, tuple)):
            for x in q:
                self.measure(x)
            return self
        return self._1q(MEASURE, q)

    def release(self) -> Program:
        p = self._prog
        p.instrs = np.concatenate(self._rows) if self._rows else np.zeros(0, GATE_DTYPE)
        self._rows = []
        self._prog = Program(self.num_qubits)
        return p


def layered(num_qubits: int, depth: int, seed: int) -> Program:
    """Seeded layered h/rx/cx circuit, generated by the library (qlx_gen_layered)."""
    L = _abi.lib()
    n = C.c_size_t(0)
    _abi.check(L.qlx_gen_layered(num_qubits, depth, seed, None, 0, C.byref(n)))
    arr = np.zeros(n.value, GATE_DTYPE)
    _abi.check(L.qlx_gen_layered(num_qubits, depth, seed, arr.ctypes.data, n.value, C.byref(n)))
    return Program(num_qubits, arr)


def hea(num_qubits: int, depth: int) -> Program:
    """make_hea_circuit (reference include/frontend/circuit_generators.hpp:22-48): per layer
    RY(q, theta) and RZ(q, 1.5 theta) on every qubit, then CX(q, q+1) along the chain;
    theta = 0.1234 + 0.05 * layer, accumulated in double as the reference does."""
    c = Circuit(num_qubits)
    theta = 0.1234
    for _ in range(depth):
        for q in range(num_qubits):
            c.ry(q, theta)
            c.rz(q, theta * 1.5)
        for q in range(num_qubits - 1):
            c.cx(q, q + 1)
        theta += 0.05
    return c.release()


def qft(num_qubits: int) -> Program:
    """make_qft_circuit (reference circuit_generators.hpp:50-68): H(i) then CX(j, i) for
    every j > i, for i = 0..n-1; then SWAP(i, n-1-i)."""
    c = Circuit(num_qubits)
    for i in range(num_qubits):
        c.h(i)
        for j in range(i + 1, num_qubits):
            c.cx(j, i)
    for i in range(num_qubits // 2):
        c.swap(i, num_qubits - 1 - i)
    return c.release()


def clustered(num_qubits: int, cluster_size: int = 6, depth: int = 15) -> Program:
    """The strategy-matrix benchmark circuit (reference benchmarks/bench_hw_awareness.cpp:30-53,
    run there with cluster_size 6, depth 15 at 24-28 qubits): the qubits split into clusters of
    `cluster_size` consecutive qubits (the last one may be shorter); each cluster in turn gets
    `depth` layers of H on every member followed by the CX chain (q, q+1) inside it."""
    if cluster_size < 1:
        raise ValueError("cluster_size must be positive")
    c = Circuit(num_qubits)
    for lo in range(0, num_qubits, cluster_size):
        hi = min(lo + cluster_size, num_qubits)
        for _ in range(depth):
            for q in range(lo, hi):
                c.h(q)
            for q in range(lo, hi - 1):
                c.cx(q, q + 1)
    return c.release()


def transpile_gpu(prog: Program, fusion: bool = True, max_width: int = 4, rotations: bool = True,
                  windowing: bool = False, k_safe: int = 0, numa_penalty: int = 10, swap_penalty: int = 1,
                  restore: bool = True, dtype: int = 0, on_device: bool = False, cost_model="threshold") -> Program:
    """Middleware::transpile_gpu through the library (qlx_transpile). on_device: form the
    fused-block matrices in FP64 on the current GPU (the blocks are chosen on the host).
    cost_model: "threshold" (ThresholdCostModel), "greedy" (GreedyCostModel), "custom" (no model:
    rejected, as the reference rejects CUSTOM without an instance), or a callable
    evaluate_swap(window_frequencies, log_to_phys, phys_to_log) -> (active, victim) | None,
    WindowingCostModel::evaluate_swap (middleware.hpp:44-61); its k_safe is `k_safe`."""
    L = _abi.lib()
    o = _abi.TranspileOpts()
    o.build_on_device = int(on_device)
    cb = None
    if callable(cost_model):
        model = cost_model

        def _eval(_ctx, freq, l2p, p2l, nq, active, victim):
            try:
                got = model([freq[i] for i in range(nq)], [l2p[i] for i in range(nq)],
                             [p2l[i] for i in range(nq)])
            except Exception:  # an exception in the model aborts the pass
                return -1
            if got is None:
                return 0
            active[0], victim[0] = int(got[0]), int(got[1])
            return 1
        cb = _abi.SWAP_EVAL_FN(_eval)
        o.cost_model = _abi.QLX_COST_CUSTOM
        o.custom_evaluate = C.cast(cb, C.c_void_p)
    else:
        o.cost_model = {"threshold": _abi.QLX_COST_THRESHOLD, "greedy": _abi.QLX_COST_GREEDY,
                        "custom": _abi.QLX_COST_CUSTOM}[cost_model]
    o.enable_gate_fusion, o.max_fusion_width, o.fuse_rotations = int(fusion), max_width, int(rotations)
    o.enable_windowing, o.k_safe, o.numa_penalty, o.swap_penalty = int(windowing), k_safe, numa_penalty, swap_penalty
    o.restore_layout, o.dtype = int(restore), dtype
    ins = np.ascontiguousarray(prog.instrs)
    pool = np.ascontiguousarray(prog.pool, np.complex128)
    meta = np.ascontiguousarray(prog.meta)
    h = C.c_void_p()
    _abi.check(L.qlx_transpile(prog.num_qubits, ins.ctypes.data if len(ins) else None, len(ins),
                               pool.ctypes.data if pool.size else None, pool.size,
                               meta.ctypes.data if len(meta) else None, len(meta), C.byref(o), C.byref(h)))
    try:
        ni, npool, nm = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _abi.check(L.qlx_program_sizes(h, C.byref(ni), C.byref(npool), C.byref(nm)))
        out = Program(prog.num_qubits, np.zeros(ni.value, GATE_DTYPE), np.zeros(npool.value, np.complex128),
                      np.zeros(nm.value, META_DTYPE))
        _abi.check(L.qlx_program_copy(h, out.instrs.ctypes.data if ni.value else None,
                                      out.pool.ctypes.data if npool.value else None,
                                      out.meta.ctypes.data if nm.value else None))
    finally:
        L.qlx_program_destroy(h)
    return out


def gpu_safe_qubit_limit(dtype: int = 0, device: int = -1) -> int:
    k = C.c_uint32()
    _abi.check(_abi.lib().qlx_gpu_safe_qubit_limit(device, dtype, C.byref(k)))
    return k.value


def inverse_layered(prog: Program) -> Program:
    """The inverse of an H/RX/CX program: reversed order, RX(theta) -> RX(-theta)."""
    inv = prog.instrs[::-1].copy()
    for g in inv:
        if g["type"] == RX:
            g["payload"] = f32_bits(-bits_f32(int(g["payload"])))
        elif g["type"] not in (H, CX, X, Z, SWAP):
            raise ValueError("inverse_layered only handles H/RX/CX/X/Z/SWAP")
    return Program(prog.num_qubits, inv, prog.pool, prog.meta)


class MT19937_64:
    """std::mt19937_64 (64-bit Mersenne Twister, the C++ standard's parameters)."""

    _M = 0xFFFFFFFFFFFFFFFF

    def __init__(self, seed: int = 5489):
        self.seed(seed)

    def seed(self, seed: int) -> None:
        mt = [0] * 312
        mt[0] = seed & self._M
        for i in range(1, 312):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & self._M
        self._mt, self._i = mt, 312

    def _twist(self) -> None:
        mt = self._mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self._i = 0

    def __call__(self) -> int:
        if self._i >= 312:
            self._twist()
        y = self._mt[self._i]
        self._i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self._M

    def uniform01(self) -> float:
        """std::uniform_real_distribution<double>(0, 1) over this engine, as libstdc++
        evaluates it: one 64-bit draw, double(x) / 2^64, clamped below 1."""
        r = float(self()) / 18446744073709551616.0
        if r >= 1.0:
            r = float(np.nextafter(1.0, 0.0))
        return r
