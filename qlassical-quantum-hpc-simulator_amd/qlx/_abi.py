"""ctypes binding of libqlx.so (the C ABI declared in include/qlx.h).

The library is loaded from ``qlx/_lib/libqlx.so`` next to this file. There is no
fallback: if the HIP library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libqlx.so")

QLX_OK = 0
QLX_EINVAL = -1
QLX_ENOMEM = -2
QLX_ERANGE = -3
QLX_EUNSUPPORTED = -4
QLX_EHIP = -5

QLX_C128 = 0
QLX_C64 = 1

QLX_EXEC_PLAIN = 0
QLX_EXEC_WINDOWED = 1
QLX_GSWAP_PERMUTE = 0
QLX_GSWAP_NOOP = 1
QLX_EXEC_DENSE_SWEEPS = 2
QLX_EXEC_AUTO = 3
QLX_AUTO_SPECIALIZE_QUBITS = 20
QLX_SPECIALIZE = 1
QLX_SPECIALIZE_RELABEL = 2
QLX_SPECIALIZE_ASYNC = 4
QLX_PLAN_KEEP_ROTATIONS = 1
QLX_PLAN_FIXED_WINDOWS = 2
QLX_SERVE_IN_PLACE = 1
QLX_DENSE_AUTO = 0
QLX_DENSE_VALU = 1
QLX_DENSE_MFMA_BF16 = 2
QLX_DENSE_MFMA_F32 = 3
QLX_DENSE_MFMA_F16 = 4


class QlxError(RuntimeError):
    """Base class of errors reported by the C ABI."""

    code = QLX_EHIP


class QlxInvalidArgument(QlxError, ValueError):
    code = QLX_EINVAL


class QlxOutOfMemory(QlxError, MemoryError):
    code = QLX_ENOMEM


class QlxOutOfRange(QlxError, IndexError):
    code = QLX_ERANGE


class QlxUnsupported(QlxError):
    code = QLX_EUNSUPPORTED


_ERR = {
    QLX_EINVAL: QlxInvalidArgument,
    QLX_ENOMEM: QlxOutOfMemory,
    QLX_ERANGE: QlxOutOfRange,
    QLX_EUNSUPPORTED: QlxUnsupported,
    QLX_EHIP: QlxError,
}


class ExecOpts(C.Structure):
    _fields_ = [
        ("mode", C.c_uint32),
        ("window_qubits", C.c_uint32),
        ("coalesce_qubits", C.c_uint32),
        ("gswap_mode", C.c_uint32),
        ("register_qubits", C.c_uint32),
        ("dense_kernel", C.c_uint32),
        ("sweep_workgroups", C.c_uint32),
        ("specialize", C.c_uint32),
        ("plan_flags", C.c_uint32),
    ]


class ExecStats(C.Structure):
    _fields_ = [
        ("sweeps_full", C.c_uint64),
        ("sweeps_half", C.c_uint64),
        ("sweeps_quarter", C.c_uint64),
        ("gates_applied", C.c_uint64),
        ("algorithmic_bytes", C.c_double),
        ("specialised", C.c_uint32),
        ("relabelled", C.c_uint32),
    ]


class KernelTiming(C.Structure):
    _fields_ = [
        ("kernel", C.c_char * 48),
        ("launches", C.c_uint64),
        ("total_ms", C.c_double),
        ("algorithmic_bytes", C.c_double),
    ]


class PlanInfo(C.Structure):
    _fields_ = [
        ("sweeps", C.c_uint64),
        ("phases", C.c_uint64),
        ("plain_steps", C.c_uint64),
        ("gates_in_sweeps", C.c_uint64),
        ("window_qubits", C.c_uint32),
        ("register_qubits", C.c_uint32),
        ("threads_per_group", C.c_uint32),
        ("coalesce_qubits", C.c_uint32),
        ("mfma_phases", C.c_uint64),
        ("rotations_merged", C.c_uint64),
    ]


class TranspileOpts(C.Structure):
    _fields_ = [
        ("enable_gate_fusion", C.c_uint32),
        ("max_fusion_width", C.c_uint32),
        ("fuse_rotations", C.c_uint32),
        ("enable_windowing", C.c_uint32),
        ("k_safe", C.c_uint32),
        ("numa_penalty", C.c_uint32),
        ("swap_penalty", C.c_uint32),
        ("restore_layout", C.c_uint32),
        ("dtype", C.c_int32),
        ("build_on_device", C.c_uint32),
        ("cost_model", C.c_uint32),
        ("reserved", C.c_uint32),
        ("custom_evaluate", C.c_void_p),
        ("custom_ctx", C.c_void_p),
    ]


QLX_COST_THRESHOLD = 0
QLX_COST_GREEDY = 1
QLX_COST_CUSTOM = 2
# WindowingCostModel::evaluate_swap as a C callback (qlx_swap_eval_fn)
SWAP_EVAL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint16),
                           C.POINTER(C.c_uint16), C.c_uint32, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16))


class BatchInfo(C.Structure):
    _fields_ = [
        ("circuits", C.c_uint64),
        ("tasks", C.c_uint64),
        ("sweeps", C.c_uint64),
        ("state_bytes", C.c_uint64),
        ("gates", C.c_uint64),
        ("workgroups", C.c_uint32),
        ("max_sweeps", C.c_uint32),
    ]


QLX_BATCH_PERSISTENT = 0
QLX_BATCH_PER_CIRCUIT = 1
QLX_BATCH_PER_CIRCUIT_GRAPH = 2
QLX_BATCH_LEVELS = 3


class DeviceInfo(C.Structure):
    _fields_ = [
        ("hbm_bytes", C.c_uint64),
        ("hbm_free_bytes", C.c_uint64),
        ("lds_bytes_per_cu", C.c_uint32),
        ("num_cus", C.c_uint32),
        ("wavefront_size", C.c_uint32),
        ("clock_khz", C.c_uint32),
        ("arch", C.c_char * 64),
    ]


# (name, argtypes) of every function include/qlx.h declares, in header order.
SIGNATURES = [
    ("qlx_last_error", []),
    ("qlx_version", []),
    ("qlx_device_query", [C.c_int, C.POINTER(DeviceInfo)]),
    ("qlx_state_create", [C.c_uint32, C.c_int, C.POINTER(C.c_void_p)]),
    ("qlx_state_create_from_host", [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    ("qlx_state_destroy", [C.c_void_p]),
    ("qlx_state_reset", [C.c_void_p]),
    ("qlx_state_load", [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("qlx_state_to_host", [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("qlx_state_num_qubits", [C.c_void_p, C.POINTER(C.c_uint32)]),
    ("qlx_state_dtype", [C.c_void_p, C.POINTER(C.c_int)]),
    ("qlx_state_device_ptr", [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("qlx_state_norm_sq", [C.c_void_p, C.POINTER(C.c_double)]),
    ("qlx_state_inner", [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("qlx_state_max_abs_diff", [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    ("qlx_state_copy", [C.c_void_p, C.c_void_p]),
    ("qlx_execute", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                     C.POINTER(ExecOpts), C.POINTER(C.c_size_t)]),
    ("qlx_execute_noalloc", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                             C.POINTER(ExecOpts), C.POINTER(C.c_size_t)]),
    ("qlx_plan_windowed", [C.c_uint32, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                           C.c_size_t, C.POINTER(ExecOpts), C.c_void_p, C.c_void_p, C.c_size_t,
                           C.POINTER(C.c_size_t), C.POINTER(PlanInfo)]),
    ("qlx_plan_sweep_source", [C.c_uint32, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                               C.c_size_t, C.POINTER(ExecOpts), C.c_uint32, C.c_char_p, C.c_size_t,
                               C.POINTER(C.c_size_t)]),
    ("qlx_specialize_wait", [C.c_double, C.POINTER(C.c_uint32)]),
    ("qlx_last_exec_stats", [C.c_void_p, C.POINTER(ExecStats)]),
    ("qlx_measure_prob1", [C.c_void_p, C.c_uint32, C.POINTER(C.c_double)]),
    ("qlx_collapse", [C.c_void_p, C.c_uint32, C.c_int, C.c_double]),
    ("qlx_sync", [C.c_void_p]),
    ("qlx_timing_enable", [C.c_void_p, C.c_int]),
    ("qlx_timing_reset", [C.c_void_p]),
    ("qlx_timing_read", [C.c_void_p, C.POINTER(KernelTiming), C.c_size_t, C.POINTER(C.c_size_t)]),
    ("qlx_transpile", [C.c_uint32, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                       C.POINTER(TranspileOpts), C.POINTER(C.c_void_p)]),
    ("qlx_program_sizes", [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("qlx_program_copy", [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("qlx_program_destroy", [C.c_void_p]),
    ("qlx_gpu_safe_qubit_limit", [C.c_int, C.c_int, C.POINTER(C.c_uint32)]),
    ("qlx_batch_create", [C.c_int, C.POINTER(C.c_void_p)]),
    ("qlx_batch_add", [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                       C.c_size_t, C.POINTER(C.c_uint32)]),
    ("qlx_batch_prepare", [C.c_void_p, C.POINTER(ExecOpts), C.POINTER(BatchInfo)]),
    ("qlx_batch_run", [C.c_void_p, C.c_int]),
    ("qlx_batch_sync", [C.c_void_p, C.POINTER(C.c_double)]),
    ("qlx_batch_state_to_host", [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t]),
    ("qlx_batch_compare", [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("qlx_batch_destroy", [C.c_void_p]),
    ("qlx_server_create", [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("qlx_server_set_flags", [C.c_void_p, C.c_uint32]),
    ("qlx_server_submit", [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint64)]),
    ("qlx_server_pump", [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("qlx_server_request_info", [C.c_void_p, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("qlx_server_take_completed", [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("qlx_server_compare", [C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("qlx_server_state_to_host", [C.c_void_p, C.c_uint64, C.c_void_p, C.c_size_t]),
    ("qlx_server_release", [C.c_void_p, C.c_uint64]),
    ("qlx_server_destroy", [C.c_void_p]),
    ("qlx_gen_layered", [C.c_uint32, C.c_uint32, C.c_uint64, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
]

_lib = None


def lib() -> C.CDLL:
    """Loads libqlx.so once (RTLD_GLOBAL so the C++ test binaries share it)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libqlx.so is not built at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    h = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    missing = []
    for name, args in SIGNATURES:
        try:
            fn = getattr(h, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.argtypes = args
        fn.restype = C.c_char_p if name in ("qlx_last_error", "qlx_version") else C.c_int
    if missing:
        # an older build: everything it has works; calling a missing entry point raises
        for name in missing:
            setattr(h, name, _missing(name))
    _lib = h
    return h


def _missing(name):
    def call(*_a, **_k):
        raise ImportError(f"{name} is not in {LIB_PATH}: rebuild the library")
    return call


def check(rc: int) -> None:
    if rc == QLX_OK:
        return
    msg = lib().qlx_last_error().decode(errors="replace")
    raise _ERR.get(rc, QlxError)(f"[{rc}] {msg}")
