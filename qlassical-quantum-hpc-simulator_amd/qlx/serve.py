"""Continuous-batching circuit server (qlx_server_*) and a Poisson-arrival driver."""
from __future__ import annotations

import ctypes as C
import math
import time
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _abi
from .engine import State, _dt
from .ir import MT19937_64, Program, layered


class Server:
    def __init__(self, hbm_budget_bytes: int, dtype="c64", in_place: bool = False):
        """in_place: one state buffer per request (QLX_SERVE_IN_PLACE); by default requests
        larger than one window run out of place between two buffers (fewer sweeps)."""
        h = C.c_void_p()
        _abi.check(_abi.lib().qlx_server_create(_dt(dtype), int(hbm_budget_bytes), C.byref(h)))
        self._h = h
        if in_place:
            _abi.check(_abi.lib().qlx_server_set_flags(self._h, _abi.QLX_SERVE_IN_PLACE))

    def submit(self, prog: Program) -> int:
        ins = np.ascontiguousarray(prog.instrs)
        rid = C.c_uint64()
        _abi.check(_abi.lib().qlx_server_submit(self._h, prog.num_qubits, ins.ctypes.data, len(ins), C.byref(rid)))
        return rid.value

    def pump(self) -> Tuple[int, int, int]:
        q, a, f = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _abi.check(_abi.lib().qlx_server_pump(self._h, C.byref(q), C.byref(a), C.byref(f)))
        return q.value, a.value, f.value

    def info(self, rid: int) -> Tuple[str, float, float, float]:
        """(status, submit_s, admit_s, done_s): status 'queued' | 'active' | 'done'; times in
        seconds since the server was created, -1 for a step not reached."""
        st = C.c_int(0)
        t = (C.c_double * 3)()
        _abi.check(_abi.lib().qlx_server_request_info(self._h, rid, C.byref(st), t))
        return ({0: "queued", 1: "active", 2: "done"}[st.value], t[0], t[1], t[2])

    def take_completed(self, cap: int = 4096) -> List[Tuple[int, float]]:
        ids = (C.c_uint64 * cap)()
        lat = (C.c_double * cap)()
        n = C.c_size_t()
        _abi.check(_abi.lib().qlx_server_take_completed(self._h, ids, lat, cap, C.byref(n)))
        return [(ids[i], lat[i]) for i in range(n.value)]

    def compare(self, rid: int, ref: State) -> Tuple[float, float]:
        inf, m = C.c_double(), C.c_double()
        _abi.check(_abi.lib().qlx_server_compare(self._h, rid, ref.handle, C.byref(inf), C.byref(m)))
        return inf.value, m.value

    def state(self, rid: int, num_qubits: int) -> np.ndarray:
        out = np.empty(1 << num_qubits, np.complex128)
        _abi.check(_abi.lib().qlx_server_state_to_host(self._h, rid, out.ctypes.data, out.size))
        return out

    def release(self, rid: int) -> None:
        _abi.check(_abi.lib().qlx_server_release(self._h, rid))

    def close(self) -> None:
        if self._h:
            _abi.check(_abi.lib().qlx_server_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_workload(count: int, qubits=(12, 28), depths=(5, 40), shape_seed: int = 5005,
                   seed_base: int = 500000) -> List[Tuple[int, int, int]]:
    """(num_qubits, depth, seed) per request; N, D from mt19937_64(shape_seed) as lo + rng() % span."""
    rng = MT19937_64(shape_seed)
    out = []
    for j in range(count):
        n = qubits[0] + rng() % (qubits[1] - qubits[0] + 1)
        d = depths[0] + rng() % (depths[1] - depths[0] + 1)
        out.append((int(n), int(d), seed_base + j))
    return out


def poisson_arrivals(count: int, rate: float, seed: int = 5006) -> np.ndarray:
    """Arrival times (s) of a Poisson process: exponential gaps from mt19937_64(seed)."""
    rng = MT19937_64(seed)
    t, out = 0.0, []
    for _ in range(count):
        u = rng.uniform01()
        t += -math.log(1.0 - u) / rate
        out.append(t)
    return np.array(out)


def run_server(programs: List[Program], arrivals: np.ndarray, budget_bytes: int, dtype="c64",
               check_every: int = 0, timeout_s: float = 300.0, in_place: bool = False) -> Dict:
    """Replays the arrival schedule against a Server in real time; returns latency stats.
    Every check_every-th request is compared (FP64, on the GPU) with a complex128 state of the
    same circuit; those reference states are computed before the clock starts, so a check
    inside the loop costs one FP64 reduction."""
    from .engine import exec_opts
    refs: Dict[int, State] = {}
    if check_every:
        for j in range(0, len(programs), check_every):
            ref = State(programs[j].num_qubits, "c128")
            ref.execute(programs[j], exec_opts("windowed") if programs[j].num_qubits >= 12 else None)
            ref.sync()
            refs[j] = ref
    srv = Server(budget_bytes, dtype, in_place)
    rid_to_j: Dict[int, int] = {}
    lat: Dict[int, float] = {}
    checks = []
    nxt = 0
    t0 = time.perf_counter()
    try:
        while len(lat) < len(programs):
            now = time.perf_counter() - t0
            if now > timeout_s:
                raise TimeoutError(f"server run exceeded {timeout_s} s ({len(lat)} of {len(programs)} done)")
            while nxt < len(programs) and arrivals[nxt] <= now:
                rid_to_j[srv.submit(programs[nxt])] = nxt
                nxt += 1
            srv.pump()
            for rid, _ in srv.take_completed():
                j = rid_to_j[rid]
                lat[j] = (time.perf_counter() - t0) - arrivals[j]
                if j in refs:
                    checks.append((j,) + srv.compare(rid, refs[j]))
                srv.release(rid)
        wall = time.perf_counter() - t0
    finally:
        srv.close()
        for ref in refs.values():
            ref.close()
    lats = np.array([lat[j] for j in range(len(programs))])
    return {"requests": len(programs), "wall_s": wall, "circuits_per_s": len(programs) / wall,
            "p50_ms": float(np.percentile(lats, 50) * 1e3), "p99_ms": float(np.percentile(lats, 99) * 1e3),
            "mean_ms": float(lats.mean() * 1e3), "checks": checks,
            "max_infidelity": max([c[1] for c in checks], default=None),
            "max_abs": max([c[2] for c in checks], default=None)}
