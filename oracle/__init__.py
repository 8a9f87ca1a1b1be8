"""ctypes loader of the CPU oracle (oracle.cpp) and the reference-header probe.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STRICT = os.path.join(HERE, "_build", "liboracle.so")
FAST = os.path.join(HERE, "_build", "liboracle_fast.so")
REF = os.path.join(HERE, "_ref", "libqlx_ref.so")

_libs: dict = {}


def build() -> None:
    subprocess.run(["make", "-C", HERE, "-j4"], check=True, stdout=subprocess.DEVNULL)


def _load(path: str) -> C.CDLL:
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise ImportError(f"oracle library missing: {path} (run make -C oracle)")
    L = C.CDLL(path)
    if path == REF:
        L.ref_gen_layered.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_void_p, C.c_size_t,
                                      C.POINTER(C.c_size_t)]
        L.ref_gen_bench.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.ref_build.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                C.POINTER(C.c_size_t)]
        L.ref_insert_zero_bit.argtypes = [C.c_uint64, C.c_uint8]
        L.ref_insert_zero_bit.restype = C.c_uint64
        L.ref_state_check.argtypes = [C.c_void_p, C.c_size_t]
        L.ref_state_zero.argtypes = [C.c_uint32, C.c_void_p]
        L.ref_layout.argtypes = [C.c_void_p]
        L.ref_layout.restype = None
    else:
        L.orc_last_error.restype = C.c_char_p
        L.orc_init_zero.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_check_state.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_double)]
        L.orc_execute.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                  C.c_void_p, C.c_size_t, C.c_int, C.c_uint64, C.c_void_p, C.c_size_t,
                                  C.POINTER(C.c_size_t)]
        L.orc_time_execute.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                       C.c_void_p, C.c_size_t, C.c_int]
        L.orc_time_execute.restype = C.c_double
        L.orc_gen_layered.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_void_p, C.c_size_t,
                                      C.POINTER(C.c_size_t)]
        L.orc_fuse_matrices.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                        C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        L.orc_num_threads.restype = C.c_int
        L.orc_random_state.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64]
    _libs[path] = L
    return L


def strict() -> C.CDLL:
    return _load(STRICT)


def fast() -> C.CDLL:
    return _load(FAST)


def ref_available() -> bool:
    return os.path.exists(REF)


def ref() -> C.CDLL:
    return _load(REF)


GATE_DTYPE = np.dtype([
    ("type", "u1"), ("flags_arity", "u1"), ("qubits", "u1", (4,)), ("pad", "<u2"),
    ("payload", "<u4"), ("uid", "<u4"),
])


class OracleError(RuntimeError):
    pass


def zero_state(nq: int) -> np.ndarray:
    a = np.empty(1 << nq, np.complex128)
    strict().orc_init_zero(a.ctypes.data, nq)
    return a


def execute(num_qubits: int, instrs: np.ndarray, pool=None, meta=None, amps=None, gswap: str = "noop",
            seed: int = 5489, lib: str = "strict"):
    """Runs the program on the CPU oracle; returns (state, measurements).
    gswap='noop' is the reference's GLOBAL_SWAP; 'permute' applies it."""
    L = strict() if lib == "strict" else fast()
    a = zero_state(num_qubits) if amps is None else np.array(amps, dtype=np.complex128, copy=True)
    ins = np.ascontiguousarray(instrs)
    pool = np.zeros(0, np.complex128) if pool is None else np.ascontiguousarray(pool, np.complex128)
    meta_b = np.zeros(0, np.uint8) if meta is None else np.ascontiguousarray(meta)
    meas = np.zeros(max(1, len(ins)), np.uint8)
    nm = C.c_size_t(0)
    rc = L.orc_execute(a.ctypes.data, num_qubits, ins.ctypes.data if len(ins) else None, len(ins),
                       pool.ctypes.data if pool.size else None, pool.size,
                       meta_b.ctypes.data if len(meta_b) else None, len(meta_b) if meta is not None else 0,
                       1 if gswap == "permute" else 0, seed, meas.ctypes.data, meas.size, C.byref(nm))
    if rc != 0:
        msg = L.orc_last_error().decode()
        raise {-1: ValueError, -2: IndexError}.get(rc, RuntimeError)(msg)
    return a, [int(x) for x in meas[:nm.value]]


def random_state(num_qubits: int, seed: int, out: np.ndarray = None) -> np.ndarray:
    """Seeded random normalised complex128 state (orc_random_state: SplitMix64 of (seed, index),
    OpenMP fill, thread-count independent). Fills `out` when given (a C-contiguous complex128
    array of 2^num_qubits elements)."""
    a = np.empty(1 << num_qubits, np.complex128) if out is None else out
    if a.dtype != np.complex128 or a.size != (1 << num_qubits) or not a.flags.c_contiguous:
        raise ValueError("random_state: out must be C-contiguous complex128 of 2^num_qubits elements")
    strict().orc_random_state(a.ctypes.data, num_qubits, seed)
    return a


def execute_inplace(num_qubits: int, instrs: np.ndarray, amps: np.ndarray, pool=None, meta=None,
                    gswap: str = "noop", seed: int = 5489, lib: str = "strict") -> list:
    """As execute(), but evolves `amps` (C-contiguous complex128, 2^num_qubits) in place, so a
    30-qubit run needs one 16 GiB host buffer. Returns the measurements."""
    if amps.dtype != np.complex128 or amps.size != (1 << num_qubits) or not amps.flags.c_contiguous:
        raise ValueError("execute_inplace: amps must be C-contiguous complex128 of 2^num_qubits elements")
    L = strict() if lib == "strict" else fast()
    ins = np.ascontiguousarray(instrs)
    pool = np.zeros(0, np.complex128) if pool is None else np.ascontiguousarray(pool, np.complex128)
    meta_b = np.zeros(0, np.uint8) if meta is None else np.ascontiguousarray(meta)
    meas = np.zeros(max(1, len(ins)), np.uint8)
    nm = C.c_size_t(0)
    rc = L.orc_execute(amps.ctypes.data, num_qubits, ins.ctypes.data if len(ins) else None, len(ins),
                       pool.ctypes.data if pool.size else None, pool.size,
                       meta_b.ctypes.data if len(meta_b) else None, len(meta_b) if meta is not None else 0,
                       1 if gswap == "permute" else 0, seed, meas.ctypes.data, meas.size, C.byref(nm))
    if rc != 0:
        msg = L.orc_last_error().decode()
        raise {-1: ValueError, -2: IndexError}.get(rc, RuntimeError)(msg)
    return [int(x) for x in meas[:nm.value]]


def time_execute(num_qubits: int, instrs: np.ndarray, reps: int = 1, lib: str = "fast", pool=None,
                 meta=None) -> float:
    L = strict() if lib == "strict" else fast()
    a = np.empty(1 << num_qubits, np.complex128)
    ins = np.ascontiguousarray(instrs)
    pool = None if pool is None or len(pool) == 0 else np.ascontiguousarray(pool, np.complex128)
    meta = None if meta is None or len(meta) == 0 else np.ascontiguousarray(meta)
    return L.orc_time_execute(a.ctypes.data, num_qubits, ins.ctypes.data, len(ins),
                              pool.ctypes.data if pool is not None else None, 0 if pool is None else pool.size,
                              meta.ctypes.data if meta is not None else None, 0 if meta is None else len(meta), reps)


def num_threads(lib: str = "fast") -> int:
    return (strict() if lib == "strict" else fast()).orc_num_threads()


def gen_layered(n: int, depth: int, seed: int) -> np.ndarray:
    L = strict()
    cnt = C.c_size_t(0)
    L.orc_gen_layered(n, depth, seed, None, 0, C.byref(cnt))
    out = np.zeros(cnt.value, GATE_DTYPE)
    if L.orc_gen_layered(n, depth, seed, out.ctypes.data, cnt.value, C.byref(cnt)) != 0:
        raise OracleError("generator failed")
    return out


def check_state(amps: np.ndarray) -> int:
    a = np.ascontiguousarray(amps, np.complex128)
    ns = C.c_double()
    return strict().orc_check_state(a.ctypes.data, a.size, C.byref(ns))


def fuse_matrices(U1, q1, U2, q2):
    U1 = np.ascontiguousarray(U1, np.complex128)
    U2 = np.ascontiguousarray(U2, np.complex128)
    q1a = np.array(q1, np.uint16)
    q2a = np.array(q2, np.uint16)
    nq = len(set(q1) | set(q2))
    out = np.zeros((1 << nq) ** 2, np.complex128)
    oq = np.zeros(8, np.uint16)
    on = C.c_int(0)
    if strict().orc_fuse_matrices(U1.ctypes.data, q1a.ctypes.data, len(q1), U2.ctypes.data, q2a.ctypes.data,
                                  len(q2), out.ctypes.data, oq.ctypes.data, C.byref(on)) != 0:
        raise OracleError(strict().orc_last_error().decode())
    d = 1 << on.value
    return out.reshape(d, d), [int(x) for x in oq[:on.value]]


# ---- the reference's own headers (oracle/_ref) ------------------------------------------
def ref_gen_layered(n: int, depth: int, seed: int) -> np.ndarray:
    L = ref()
    cnt = C.c_size_t(0)
    L.ref_gen_layered(n, depth, seed, None, 0, C.byref(cnt))
    out = np.zeros(cnt.value, GATE_DTYPE)
    L.ref_gen_layered(n, depth, seed, out.ctypes.data, cnt.value, C.byref(cnt))
    return out


def ref_gen_bench(kind: str, n: int, depth: int = 0) -> np.ndarray:
    """The reference's make_hea_circuit(n, depth) / make_qft_circuit(n) instruction bytes."""
    L = ref()
    k = {"hea": 0, "qft": 1}[kind]
    cnt = C.c_size_t(0)
    L.ref_gen_bench(k, n, depth, None, 0, C.byref(cnt))
    out = np.zeros(cnt.value, GATE_DTYPE)
    L.ref_gen_bench(k, n, depth, out.ctypes.data, cnt.value, C.byref(cnt))
    return out


def ref_build(n: int, ops, params=None):
    """ops: list of (type, q0, q1, q2). Returns GateInstr records or raises like the builder."""
    L = ref()
    o = np.array(ops, np.int32).reshape(-1, 4)
    p = np.zeros(len(o), np.float64) if params is None else np.array(params, np.float64)
    out = np.zeros(max(1, len(o)), GATE_DTYPE)
    cnt = C.c_size_t(0)
    rc = L.ref_build(n, o.ctypes.data, p.ctypes.data, len(o), out.ctypes.data, C.byref(cnt))
    if rc == -1:
        raise ValueError("reference builder: invalid_argument")
    if rc == -2:
        raise IndexError("reference builder: out_of_range")
    if rc != 0:
        raise OracleError("unsupported op")
    return out[:cnt.value]


def ref_layout():
    out = np.zeros(8, np.uint64)
    ref().ref_layout(out.ctypes.data)
    return [int(x) for x in out]
