"""Batched execution of many independent circuits (qlx_batch_*), and the seeded
workload generators of the batching and serving configurations."""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Tuple

import numpy as np

from . import _abi
from .engine import State, _dt, exec_opts
from .ir import MT19937_64, Program, layered

MODES = {"persistent": _abi.QLX_BATCH_PERSISTENT, "per_circuit": _abi.QLX_BATCH_PER_CIRCUIT,
         "per_circuit_graph": _abi.QLX_BATCH_PER_CIRCUIT_GRAPH, "levels": _abi.QLX_BATCH_LEVELS}


def batch_workload(count: int = 4096, qubits=(10, 20), depths=(5, 40), shape_seed: int = 4004,
                   circuit_seed_base: int = 400000) -> List[Tuple[int, int, int]]:
    """(num_qubits, depth, seed) of each circuit: N and D drawn uniformly from an
    mt19937_64(shape_seed) stream as lo + rng() % (hi - lo + 1); circuit i uses seed base + i."""
    rng = MT19937_64(shape_seed)
    out = []
    for i in range(count):
        n = qubits[0] + rng() % (qubits[1] - qubits[0] + 1)
        d = depths[0] + rng() % (depths[1] - depths[0] + 1)
        out.append((int(n), int(d), circuit_seed_base + i))
    return out


class Batch:
    """A set of circuits whose states live in one HBM arena."""

    def __init__(self, programs: Iterable[Program], dtype="c64", window_qubits: int = 0, register_qubits: int = 0,
                 out_of_place: bool = False):
        """out_of_place: circuits larger than one window run their sweeps between two state
        buffers (relabelled layouts, windows free of the low qubits: fewer sweeps)."""
        L = _abi.lib()
        h = C.c_void_p()
        _abi.check(L.qlx_batch_create(_dt(dtype), C.byref(h)))
        self._h = h
        self.dtype = _dt(dtype)
        self.programs = []
        try:
            for p in programs:
                ins = np.ascontiguousarray(p.instrs)
                pool = np.ascontiguousarray(p.pool, np.complex128)
                meta = np.ascontiguousarray(p.meta)
                cid = C.c_uint32()
                _abi.check(L.qlx_batch_add(h, p.num_qubits, ins.ctypes.data if len(ins) else None, len(ins),
                                           pool.ctypes.data if pool.size else None, pool.size,
                                           meta.ctypes.data if len(meta) else None, len(meta), C.byref(cid)))
                self.programs.append(p)
            o = exec_opts("windowed", window_qubits=window_qubits, register_qubits=register_qubits,
                          relabel=out_of_place)
            info = _abi.BatchInfo()
            _abi.check(L.qlx_batch_prepare(h, C.byref(o), C.byref(info)))
            self.info = {k: getattr(info, k) for k, _ in info._fields_}
        except Exception:
            self.close()
            raise

    def run(self, mode: str = "persistent") -> None:
        _abi.check(_abi.lib().qlx_batch_run(self._h, MODES[mode]))

    def sync(self) -> float:
        ms = C.c_double()
        _abi.check(_abi.lib().qlx_batch_sync(self._h, C.byref(ms)))
        return ms.value

    def state(self, i: int) -> np.ndarray:
        out = np.empty(1 << self.programs[i].num_qubits, np.complex128)
        _abi.check(_abi.lib().qlx_batch_state_to_host(self._h, i, out.ctypes.data, out.size))
        return out

    def compare(self, i: int, ref: State) -> Tuple[float, float]:
        inf, m = C.c_double(), C.c_double()
        _abi.check(_abi.lib().qlx_batch_compare(self._h, i, ref.handle, C.byref(inf), C.byref(m)))
        return inf.value, m.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            _abi.check(_abi.lib().qlx_batch_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- autotuning -------------------------------------------------------------------------
# Candidate (fusion width, window qubits, out of place) configurations for complex64 batches.
# Fusion width 0 keeps the circuits as they are; width 2 merges runs of 1- and 2-qubit gates
# into 4x4 matrices (fewer, heavier in-register ops: the sweep kernel's per-op dispatch cost
# falls, its arithmetic per op doubles). The window sets the chunk of circuits wider than it.
# Out of place, circuits wider than the window take two buffers and fewer sweeps.
TUNE_SPACE = [(fw, w, oop) for oop in (False, True) for w in (13, 12, 11) for fw in (0, 2)]


def prepare_programs(programs: Iterable[Program], fusion_width: int) -> List[Program]:
    """The circuits as the batch runs them: fused to the given width (0: unchanged)."""
    from .ir import transpile_gpu
    if not fusion_width:
        return list(programs)
    return [transpile_gpu(p, fusion=True, max_width=fusion_width, rotations=True) for p in programs]


def autotune(programs: List[Program], dtype="c64", sample: int = 256, reps: int = 3, seed: int = 7,
             space=TUNE_SPACE, mode: str = "persistent") -> dict:
    """Picks the fusion width, window and buffer form for a batch by timing a random sample of
    its circuits in every candidate configuration (median of `reps` runs of one persistent
    launch). Returns {"fusion_width", "window_qubits", "out_of_place",
    "trials": [(fw, w, oop, ms), ...]}."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(programs), size=min(sample, len(programs)), replace=False)
    base = [programs[i] for i in sorted(idx)]
    trials = []
    for fw, w, oop in space:
        progs = prepare_programs(base, fw)
        try:
            b = Batch(progs, dtype, window_qubits=w, out_of_place=oop)
        except _abi.QlxUnsupported:
            continue
        try:
            b.run(mode)
            b.sync()
            times = []
            for _ in range(reps):
                b.run(mode)
                times.append(b.sync())
            trials.append((fw, w, oop, float(np.median(times))))
        finally:
            b.close()
    if not trials:
        raise _abi.QlxUnsupported("no batch configuration runs these circuits")
    fw, w, oop, _ = min(trials, key=lambda t: t[3])
    return {"fusion_width": fw, "window_qubits": w, "out_of_place": oop, "trials": trials}
