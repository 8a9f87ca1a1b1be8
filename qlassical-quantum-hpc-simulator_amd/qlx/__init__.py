"""qlx — Python binding of the gfx950 Qlassical execution backend (libqlx.so).

The product is the HIP library and its C ABI (include/qlx.h) plus the C++20
HIPEngine/GPUStateVector headers; this package is the ctypes plumbing used by
the tests, the benchmark and the serving loop.
"""
from . import _abi
from ._abi import (LIB_PATH, QlxError, QlxInvalidArgument, QlxOutOfMemory, QlxOutOfRange, QlxUnsupported,
                   lib)
from .batch import Batch, batch_workload
from .serve import Server, poisson_arrivals, run_server, serve_workload
from .engine import Engine, State, device_info, exec_opts, plan_windowed, specialize_wait, sweep_source
from .ir import (GATE_DTYPE, META_DTYPE, MT19937_64, Circuit, Program, bits_f32, f32_bits, gpu_safe_qubit_limit,
                 clustered, hea, inverse_layered, layered, qft, transpile_gpu)

__all__ = [
    "_abi", "LIB_PATH", "QlxError", "QlxInvalidArgument", "QlxOutOfMemory", "QlxOutOfRange", "QlxUnsupported",
    "lib", "Batch", "batch_workload", "Server", "poisson_arrivals", "run_server", "serve_workload", "Engine", "State", "device_info", "exec_opts", "plan_windowed", "specialize_wait", "sweep_source", "GATE_DTYPE", "META_DTYPE", "MT19937_64", "Circuit",
    "Program", "bits_f32", "clustered", "f32_bits", "gpu_safe_qubit_limit", "hea", "inverse_layered", "layered", "qft",
    "transpile_gpu",
]
