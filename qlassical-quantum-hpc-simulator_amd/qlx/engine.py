"""Device state handle and a Python mirror of the C++ HIPEngine over the C ABI.

``State`` wraps ``qlx_state`` (GPUStateVector). ``Engine`` follows the
``ExecutionEngine`` contract of the reference (include/backend/execution_engine.hpp:20-79):
reset(), reset(amps), set_seed(), reserve_measurements(), measurements(),
execute(program), state(), num_qubits(). MEASURE is split out of the device
stream and decided on the host with mt19937_64 + uniform_real_distribution, as
CPUOpenMPEngine::apply_measure does (include/backend/cpu_openmp_engine.hpp:617-671).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional

import numpy as np

from . import _abi
from .ir import MEASURE, MT19937_64, Program

DTYPES = {"c128": _abi.QLX_C128, "c64": _abi.QLX_C64}


def _dt(dtype) -> int:
    if isinstance(dtype, str):
        return DTYPES[dtype]
    return int(dtype)


DENSE = {"auto": _abi.QLX_DENSE_AUTO, "valu": _abi.QLX_DENSE_VALU, "bf16": _abi.QLX_DENSE_MFMA_BF16,
         "f32": _abi.QLX_DENSE_MFMA_F32, "f16": _abi.QLX_DENSE_MFMA_F16}


def exec_opts(mode: str = "plain", window_qubits: int = 0, coalesce_qubits: int = 0,
              gswap: str = "permute", register_qubits: int = 0, dense: str = "auto",
              sweep_workgroups: int = 0, specialize: bool = False, relabel: bool = False,
              background: bool = False, keep_rotations: bool = False,
              fixed_windows: bool = False) -> _abi.ExecOpts:
    """qlx_exec_opts. mode: 'plain', 'windowed', 'dense' or 'auto' (the C++ HIPEngine's default:
    window sweeps, with specialised relabelled sweeps compiled in the background for states of
    >= 20 qubits). dense: arithmetic of 4/5-qubit dense gates on complex64 states
    ('auto' = 'f16' matrix cores (FP16 hi/lo split), 'bf16' matrix cores (BF16 split), 'f32'
    matrix cores, 'valu' vector units). relabel (windowed, with or without specialize; dense):
    out-of-place sweeps that re-lay out the qubits between two buffers. background (windowed,
    with specialize): compile on a host thread and run the generic kernel until ready.
    keep_rotations (window sweeps): apply every rotation as written instead of merging
    same-axis rotations through the gates they commute with (QLX_PLAN_KEEP_ROTATIONS).
    fixed_windows (relabelled sweeps): every window holds the low qubits
    (QLX_PLAN_FIXED_WINDOWS)."""
    if background and not specialize:
        raise ValueError("background needs specialize")
    o = _abi.ExecOpts()
    o.dense_kernel = DENSE[dense]
    o.sweep_workgroups = sweep_workgroups
    o.specialize = ((_abi.QLX_SPECIALIZE if specialize else 0) | (_abi.QLX_SPECIALIZE_RELABEL if relabel else 0) |
                    (_abi.QLX_SPECIALIZE_ASYNC if background else 0))
    o.mode = {"plain": _abi.QLX_EXEC_PLAIN, "windowed": _abi.QLX_EXEC_WINDOWED,
              "dense": _abi.QLX_EXEC_DENSE_SWEEPS, "auto": _abi.QLX_EXEC_AUTO}[mode]
    o.window_qubits = window_qubits
    o.register_qubits = register_qubits
    o.coalesce_qubits = coalesce_qubits
    o.gswap_mode = {"permute": _abi.QLX_GSWAP_PERMUTE, "noop": _abi.QLX_GSWAP_NOOP}[gswap]
    o.plan_flags = ((_abi.QLX_PLAN_KEEP_ROTATIONS if keep_rotations else 0) |
                    (_abi.QLX_PLAN_FIXED_WINDOWS if fixed_windows else 0))
    return o


def plan_windowed(prog: Program, dtype="c128", opts: Optional[_abi.ExecOpts] = None):
    """Host-only windowed plan: (order, sweep_of, info) — see qlx_plan_windowed."""
    L = _abi.lib()
    o = opts if opts is not None else exec_opts("windowed")
    ins = np.ascontiguousarray(prog.instrs)
    pool = np.ascontiguousarray(prog.pool, np.complex128)
    meta = np.ascontiguousarray(prog.meta)
    n = C.c_size_t(0)
    info = _abi.PlanInfo()
    args = (prog.num_qubits, _dt(dtype), ins.ctypes.data if len(ins) else None, len(ins),
            pool.ctypes.data if pool.size else None, pool.size, meta.ctypes.data if len(meta) else None, len(meta),
            C.byref(o))
    _abi.check(L.qlx_plan_windowed(*args, None, None, 0, C.byref(n), C.byref(info)))
    order = np.zeros(n.value, np.int64)
    sweep = np.zeros(n.value, np.int32)
    _abi.check(L.qlx_plan_windowed(*args, order.ctypes.data, sweep.ctypes.data, n.value, C.byref(n),
                                   C.byref(info)))
    return order, sweep, {k: getattr(info, k) for k, _ in info._fields_}


def sweep_source(prog: Program, sweep_index: int, dtype="c128", opts: Optional[_abi.ExecOpts] = None) -> str:
    """HIP source of one circuit-specialised sweep (host-only; see qlx_plan_sweep_source)."""
    L = _abi.lib()
    o = opts if opts is not None else exec_opts("windowed", specialize=True)
    ins = np.ascontiguousarray(prog.instrs)
    pool = np.ascontiguousarray(prog.pool, np.complex128)
    meta = np.ascontiguousarray(prog.meta)
    args = (prog.num_qubits, _dt(dtype), ins.ctypes.data if len(ins) else None, len(ins),
            pool.ctypes.data if pool.size else None, pool.size, meta.ctypes.data if len(meta) else None, len(meta),
            C.byref(o), sweep_index)
    n = C.c_size_t(0)
    _abi.check(L.qlx_plan_sweep_source(*args, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    _abi.check(L.qlx_plan_sweep_source(*args, buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def specialize_wait(timeout_s: float = -1.0) -> int:
    """Waits for background sweep compilations (qlx_specialize_wait); returns how many
    programs are still pending. Raises if a finished compilation failed."""
    n = C.c_uint32(0)
    _abi.check(_abi.lib().qlx_specialize_wait(float(timeout_s), C.byref(n)))
    return n.value


def device_info(device: int = 0) -> dict:
    info = _abi.DeviceInfo()
    _abi.check(_abi.lib().qlx_device_query(device, C.byref(info)))
    return {k: (getattr(info, k).decode() if k == "arch" else getattr(info, k)) for k, _ in info._fields_}


class State:
    """An HBM-resident state vector (qlx_state)."""

    def __init__(self, num_qubits: Optional[int] = None, dtype="c128", amps: Optional[np.ndarray] = None):
        L = _abi.lib()
        h = C.c_void_p()
        if amps is not None:
            a = np.ascontiguousarray(amps, dtype=np.complex128)
            _abi.check(L.qlx_state_create_from_host(a.ctypes.data, a.size, _dt(dtype), C.byref(h)))
        else:
            _abi.check(L.qlx_state_create(int(num_qubits), _dt(dtype), C.byref(h)))
        self._h = h
        self.dtype = _dt(dtype)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            _abi.check(_abi.lib().qlx_state_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def num_qubits(self) -> int:
        n = C.c_uint32()
        _abi.check(_abi.lib().qlx_state_num_qubits(self._h, C.byref(n)))
        return n.value

    def reset(self) -> None:
        _abi.check(_abi.lib().qlx_state_reset(self._h))

    def load(self, amps: np.ndarray) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        _abi.check(_abi.lib().qlx_state_load(self._h, a.ctypes.data, a.size))

    def to_host(self) -> np.ndarray:
        out = np.empty(1 << self.num_qubits, np.complex128)
        _abi.check(_abi.lib().qlx_state_to_host(self._h, out.ctypes.data, out.size))
        return out

    def norm_sq(self) -> float:
        v = C.c_double()
        _abi.check(_abi.lib().qlx_state_norm_sq(self._h, C.byref(v)))
        return v.value

    def inner(self, other: "State") -> complex:
        re, im = C.c_double(), C.c_double()
        _abi.check(_abi.lib().qlx_state_inner(self._h, other._h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def max_abs_diff(self, other: "State") -> float:
        v = C.c_double()
        _abi.check(_abi.lib().qlx_state_max_abs_diff(self._h, other._h, C.byref(v)))
        return v.value

    def copy_from(self, other: "State") -> None:
        _abi.check(_abi.lib().qlx_state_copy(self._h, other._h))

    def sync(self) -> None:
        _abi.check(_abi.lib().qlx_sync(self._h))

    def execute(self, prog: Program, opts: Optional[_abi.ExecOpts] = None, start: int = 0) -> int:
        """Enqueues prog.instrs[start:] up to the first MEASURE; returns its absolute index."""
        ins = prog.instrs[start:]
        ins = np.ascontiguousarray(ins)
        pool = np.ascontiguousarray(prog.pool, dtype=np.complex128)
        meta = np.ascontiguousarray(prog.meta)
        stop = C.c_size_t(0)
        _abi.check(_abi.lib().qlx_execute(
            self._h, ins.ctypes.data if len(ins) else None, len(ins),
            pool.ctypes.data if pool.size else None, pool.size,
            meta.ctypes.data if len(meta) else None, len(meta),
            C.byref(opts) if opts is not None else None, C.byref(stop)))
        return start + stop.value

    def last_stats(self) -> dict:
        s = _abi.ExecStats()
        _abi.check(_abi.lib().qlx_last_exec_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def timing(self, on: bool) -> None:
        _abi.check(_abi.lib().qlx_timing_enable(self._h, 1 if on else 0))

    def timing_reset(self) -> None:
        _abi.check(_abi.lib().qlx_timing_reset(self._h))

    def timing_read(self) -> dict:
        """{kernel: (launches, total_ms, algorithmic_bytes)} since the last reset (synchronises)."""
        buf = (_abi.KernelTiming * 32)()
        n = C.c_size_t(0)
        _abi.check(_abi.lib().qlx_timing_read(self._h, buf, 32, C.byref(n)))
        return {buf[i].kernel.decode(): (buf[i].launches, buf[i].total_ms, buf[i].algorithmic_bytes)
                for i in range(min(n.value, 32))}

    def prob1(self, q: int) -> float:
        v = C.c_double()
        _abi.check(_abi.lib().qlx_measure_prob1(self._h, q, C.byref(v)))
        return v.value

    def collapse(self, q: int, outcome: int, inv_norm: float) -> None:
        _abi.check(_abi.lib().qlx_collapse(self._h, q, int(outcome), float(inv_norm)))


class Engine:
    """Python mirror of HIPEngine (the C++ class lives in include/backend/hip_engine.hpp)."""

    def __init__(self, num_qubits: Optional[int] = None, seed: Optional[int] = None, reserve_meas: int = 0,
                 initial_state: Optional[np.ndarray] = None, dtype="c128",
                 opts: Optional[_abi.ExecOpts] = None):
        if initial_state is not None:
            self.sv = State(dtype=dtype, amps=initial_state)
        else:
            self.sv = State(num_qubits, dtype=dtype)
        self.rng = MT19937_64(seed if seed is not None else int.from_bytes(os.urandom(8), "little"))
        self._meas: list = []
        self.opts = opts

    def reset(self, initial_state: Optional[np.ndarray] = None) -> None:
        if initial_state is None:
            self.sv.reset()
        else:
            self.sv.load(initial_state)
        self._meas = []

    def set_seed(self, seed: int) -> None:
        self.rng.seed(seed)

    def reserve_measurements(self, n: int) -> None:
        pass

    def measurements(self) -> list:
        return list(self._meas)

    def num_qubits(self) -> int:
        return self.sv.num_qubits

    def execute(self, prog: Program) -> None:
        i, n = 0, len(prog.instrs)
        while i < n:
            stop = self.sv.execute(prog, self.opts, start=i)
            if stop >= n:
                break
            self._measure(int(prog.instrs[stop]["qubits"][0]))
            i = stop + 1
        self.sv.sync()

    def _measure(self, q: int) -> None:
        p1 = self.sv.prob1(q)
        r = self.rng.uniform01()
        outcome = 1 if r < p1 else 0
        self._meas.append(outcome)
        p = p1 if outcome else 1.0 - p1
        if p < 1e-12:
            p = 1e-12
        self.sv.collapse(q, outcome, 1.0 / math.sqrt(p))

    def state(self) -> np.ndarray:
        return self.sv.to_host()
