/*
 * qlx.h — C ABI of the MI355X (gfx950) state-vector execution backend.
 *
 * This is the drop-in boundary under Qlassical's C++ `ExecutionEngine` /
 * `StateVector` interfaces. The C++20 classes `HIPEngine` / `GPUStateVector`
 * (header-only, qlassical-quantum-hpc-simulator_amd/include/backend/) call these
 * functions; nothing here uses C++ or torch types.
 *
 * Conventions
 *  - Every function returns an int status: QLX_OK (0) or a negative QLX_E* code.
 *    qlx_last_error() returns a thread-local message for the last failure.
 *  - Amplitudes cross the boundary as interleaved complex128 (re, im doubles),
 *    little-endian qubit order: bit q of the basis index is qubit q
 *    (reference: include/backend/state_vector_utils.hpp:23-34).
 *  - Gate streams cross as the reference's 16-byte GateInstr records
 *    (reference: include/frontend/frontend.hpp:84-105) and the UnitaryPool as
 *    (data, metadata) arrays (reference: include/frontend/frontend.hpp:117-164).
 *  - Work is ordered on one HIP stream per state. Functions that return data
 *    to the host synchronise that stream; qlx_execute* only enqueue
 *    (qlx_sync waits and reports kernel-side errors).
 */
#ifndef QLX_ABI_H_
#define QLX_ABI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped back to C++ exceptions by HIPEngine) ---------- */
#define QLX_OK            0
#define QLX_EINVAL       -1  /* std::invalid_argument */
#define QLX_ENOMEM       -2  /* std::bad_alloc */
#define QLX_ERANGE       -3  /* std::out_of_range */
#define QLX_EUNSUPPORTED -4  /* std::runtime_error (unknown opcode etc.) */
#define QLX_EHIP         -5  /* std::runtime_error (HIP runtime failure) */

/* ---- amplitude precision of a device state ----------------------------- */
#define QLX_C128 0   /* complex128: the reference's std::complex<double> */
#define QLX_C64  1   /* complex64 working state (precision/batch/serve paths) */

/* ---- IR records, byte-identical to the reference ------------------------ */
/* GateInstr: reference include/frontend/frontend.hpp:84-105 (16 bytes).
 * Extension: when arity()==5 (FUSED_BLOCK only) the fifth qubit is stored in
 * the low byte of `pad` (the reference's `_pad`, bytes 6-7). */
typedef struct qlx_gate_instr {
    uint8_t  type;
    uint8_t  flags_arity;   /* low nibble = arity, high nibble = flags */
    uint8_t  qubits[4];
    uint16_t pad;
    union { float param; uint32_t matrix_idx; } payload;
    uint32_t uid;
} qlx_gate_instr;

/* UnitaryPool::MatrixEntry: reference include/frontend/frontend.hpp:118-123 (12 bytes). */
typedef struct qlx_matrix_entry {
    uint32_t offset;        /* start index (in complex elements) into pool data */
    uint32_t num_elements;  /* 4^k */
    uint8_t  num_qubits;    /* k */
    uint8_t  pad[3];
} qlx_matrix_entry;

/* GateType values: reference include/frontend/frontend.hpp:26-51. */
enum {
    QLX_H = 0x00, QLX_X = 0x01, QLX_Y = 0x02, QLX_Z = 0x03, QLX_S = 0x04,
    QLX_RX = 0x10, QLX_RY = 0x11, QLX_RZ = 0x12,
    QLX_CX = 0x20, QLX_SWAP = 0x21,
    QLX_CCX = 0x30, QLX_CSWAP = 0x31,
    QLX_UNITARY = 0x80, QLX_FUSED_BLOCK = 0xE0, QLX_GLOBAL_SWAP = 0xE1,
    QLX_BARRIER = 0xF0, QLX_MEASURE = 0xF1
};

/* ---- execution options --------------------------------------------------- */
#define QLX_EXEC_PLAIN    0  /* one HBM sweep per instruction (pair/quad/dense kernels) */
#define QLX_EXEC_WINDOWED 1  /* gates grouped into LDS/register window sweeps */
#define QLX_EXEC_DENSE_SWEEPS 2  /* complex64: 13-qubit windows held in LDS; every gate of a window
                                    applied there as a dense matrix, 4/5-qubit UNITARY / FUSED_BLOCK
                                    gates on the matrix cores (dense_kernel) — for fused programs */
#define QLX_EXEC_AUTO     3  /* the default of the C++ HIPEngine: window sweeps; for states of at
                                least QLX_AUTO_SPECIALIZE_QUBITS qubits (window and register qubits
                                left at their defaults) also QLX_SPECIALIZE | QLX_SPECIALIZE_RELABEL |
                                QLX_SPECIALIZE_ASYNC: a program's first executions run the generic
                                sweep kernel while its specialised sweeps compile on a host thread,
                                later executions run the compiled sweeps */
#define QLX_AUTO_SPECIALIZE_QUBITS 20

/* GLOBAL_SWAP semantics. The reference engine treats GLOBAL_SWAP as a no-op
 * (include/backend/cpu_openmp_engine.hpp:217-219); the GPU backend applies it
 * as a real qubit permutation by default. */
#define QLX_GSWAP_PERMUTE 0
#define QLX_GSWAP_NOOP    1

typedef struct qlx_exec_opts {
    uint32_t mode;            /* QLX_EXEC_PLAIN | QLX_EXEC_WINDOWED | QLX_EXEC_DENSE_SWEEPS | QLX_EXEC_AUTO */
    uint32_t window_qubits;   /* 0 = default for the precision (c128: 12, c64: 13) */
    uint32_t coalesce_qubits; /* low qubits always kept in the window; 0 = default (c128: 3, c64: 4:
                                 one 128-byte cache line) */
    uint32_t gswap_mode;      /* QLX_GSWAP_PERMUTE | QLX_GSWAP_NOOP */
    uint32_t register_qubits; /* qubits per work-item's registers in a window; 0 = default
                                 (c128: 4, c64: 5). Generic-kernel (window, register) pairs:
                                 c128 (12,4) (12,5) (11,4) (11,3); c64 (13,5) (12,5) (12,4);
                                 specialised sweeps take any window of up to 128 KiB with 3-5
                                 register qubits and 64-1024 work-items */
    uint32_t dense_kernel;    /* QLX_DENSE_*: arithmetic of UNITARY / FUSED_BLOCK gates */
    uint32_t sweep_workgroups; /* window sweeps with <= 16 registers per work-item: work-groups
                                  per launch (each loops over chunks); 0 = every resident slot */
    uint32_t specialize;      /* QLX_SPECIALIZE_* flags. windowed: without QLX_SPECIALIZE the
                                 generic sweep kernel; QLX_SPECIALIZE_RELABEL (windowed or dense
                                 sweeps) selects out-of-place relabelled sweeps */
    uint32_t plan_flags;      /* QLX_PLAN_* flags of the window-sweep planner; 0 = defaults */
} qlx_exec_opts;

/* qlx_exec_opts::plan_flags.
 * By default the window-sweep planner (windowed and AUTO modes) merges a rotation RX/RY/RZ(b)
 * into an earlier rotation of the same axis on the same qubit, RX(a) -> RX(a + b), when every
 * gate between the two on that qubit commutes with it: for RX an X or the target of a CX/CCX,
 * for RY a Y, for RZ a Z, S or a control of a CX/CCX/CSWAP. The state is the same up to
 * rounding (the angles are summed in double precision, as the cosines and sines are formed).
 * QLX_PLAN_KEEP_ROTATIONS applies every rotation as written. Plain execution never merges.
 * Out-of-place sweeps (QLX_SPECIALIZE_RELABEL: window sweeps, generic or specialised, and dense
 * sweeps) may take windows that leave out the low qubits: a sweep's loads and stores stay coalesced on qubits it shares
 * with the next sweep, which the store moves to the low address bits. The planner keeps such a
 * plan when it needs fewer passes. QLX_PLAN_FIXED_WINDOWS keeps the low qubits in every window. */
#define QLX_PLAN_KEEP_ROTATIONS 1u
#define QLX_PLAN_FIXED_WINDOWS  2u

/* qlx_exec_opts::specialize flags.
 * QLX_SPECIALIZE: compile the program's sweeps into straight-line kernels (hipRTC, once per
 *   program and state; later executions of the same program reuse them). The ROCm code-object
 *   compiler keeps its own on-disk cache, so a program compiled once is loaded in well under a
 *   second by later processes.
 * QLX_SPECIALIZE_RELABEL (windowed, with or without QLX_SPECIALIZE; dense sweeps): every
 *   sweep reads the state from one buffer
 *   and writes it to a second one with its qubits re-laid out so that the next sweep's window
 *   sits on the low address bits (short strides, fewer DRAM pages per chunk); the last sweep
 *   writes the logical order back. Needs a second device buffer of the state's size (allocated
 *   on first use, kept with the state); when that allocation fails the sweeps run in place.
 *   qlx_state_device_ptr may return the other buffer after such an execution.
 * QLX_SPECIALIZE_ASYNC (windowed, with QLX_SPECIALIZE): the program's sweeps are planned and
 *   compiled on a background host thread, shared by every state of the process; until they are
 *   ready each execution runs the generic sweep kernel (out of place when the second buffer is
 *   there, else in place; the same gates, to rounding).
 *   qlx_specialize_wait waits for the background work. */
#define QLX_SPECIALIZE 1u
#define QLX_SPECIALIZE_RELABEL 2u
#define QLX_SPECIALIZE_ASYNC 4u

/* Dense-gate arithmetic (qlx_exec_opts::dense_kernel). The split and FP32 MFMA forms apply
 * to 4- and 5-qubit gates on complex64 states. On complex128 states every value except
 * QLX_DENSE_VALU runs 4- and 5-qubit gates on the FP64 matrix cores (v_mfma_f64_16x16x4f64,
 * FP64 products and sums) on states of >= k+4 qubits. Every other dense gate runs on the
 * vector units (FP64 for complex128, FP32 for complex64). The matrices stay FP64 in the pool
 * and are converted on the device. */
#define QLX_DENSE_AUTO        0  /* complex64 4/5-qubit gates on states of >= k+5 qubits: MFMA FP16;
                                    complex128: FP64 MFMA */
#define QLX_DENSE_VALU        1  /* vector units only */
#define QLX_DENSE_MFMA_BF16   2  /* BF16 split operands (matrix hi+mid+lo, state hi+lo), four
                                    products accumulated in FP32 */
#define QLX_DENSE_MFMA_F32    3  /* FP32 MFMA (v_mfma_f32_32x32x2_f32) */
#define QLX_DENSE_MFMA_F16    4  /* FP16 split operands (matrix hi+lo, state scaled by 2^14 then
                                    hi+lo: 22 bits each), three products accumulated in FP32 */

/* Statistics of the last qlx_execute on a state (sweep accounting). */
typedef struct qlx_exec_stats {
    uint64_t sweeps_full;     /* kernels touching the whole state (f = 1) */
    uint64_t sweeps_half;     /* kernels touching half (CX/SWAP/Z/S, f = 1/2) */
    uint64_t sweeps_quarter;  /* kernels touching a quarter (CCX/CSWAP, f = 1/4) */
    uint64_t gates_applied;   /* source instructions executed (excl. BARRIER) */
    double   algorithmic_bytes; /* sum of f * 2 * sizeof(amp) * 2^N */
    uint32_t specialised;     /* 1: the window sweeps ran as circuit-specialised kernels */
    uint32_t relabelled;      /* 1: out of place through the second buffer */
} qlx_exec_stats;

/* Device description backing the GPU HardwareTopology
 * (replaces the hwloc NUMA query of reference include/middleware/hw_topology.hpp:60-102). */
typedef struct qlx_device_info {
    uint64_t hbm_bytes;        /* total device memory */
    uint64_t hbm_free_bytes;   /* free device memory now */
    uint32_t lds_bytes_per_cu; /* LDS per compute unit (workgroup max) */
    uint32_t num_cus;
    uint32_t wavefront_size;
    uint32_t clock_khz;
    char     arch[64];         /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
} qlx_device_info;

typedef struct qlx_state qlx_state;

/* ---- library / device ---------------------------------------------------- */
const char* qlx_last_error(void);
/* Version string of this library. */
const char* qlx_version(void);
/* Replaces CPUHwlocTopology (reference include/middleware/hw_topology.hpp:60-112). */
int qlx_device_query(int device, qlx_device_info* out);

/* ---- StateVector (reference include/backend/state_vector.hpp:17-47) ------ */
/* initialize(num_qubits): HBM allocation + |0...0>.
 * Reference: include/backend/cpu_state_vector.hpp:54-78. */
int qlx_state_create(uint32_t num_qubits, int dtype, qlx_state** out);
/* initialize(span): power-of-two dimension, sum |c|^2 within 1e-6 of 1, copy.
 * Reference: include/backend/cpu_state_vector.hpp:82-115. */
int qlx_state_create_from_host(const double* amps_c128, size_t dim, int dtype, qlx_state** out);
int qlx_state_destroy(qlx_state* st);
/* ExecutionEngine::reset() -> |0...0> (reference include/backend/cpu_openmp_engine.hpp:119-122). */
int qlx_state_reset(qlx_state* st);
/* ExecutionEngine::reset(span) (reference include/backend/cpu_openmp_engine.hpp:125-128);
 * the dimension may differ from the current one (the state is re-allocated). */
int qlx_state_load(qlx_state* st, const double* amps_c128, size_t dim);
/* Backs ExecutionEngine::state() (reference include/backend/execution_engine.hpp:66-67):
 * synchronises and copies the whole state to host as complex128 (c64 is widened). */
int qlx_state_to_host(qlx_state* st, double* out_c128, size_t dim);
int qlx_state_num_qubits(const qlx_state* st, uint32_t* out);
int qlx_state_dtype(const qlx_state* st, int* out);
/* Device pointer of the amplitude array (for zero-copy hand-off between engines). Valid until
 * the next qlx_execute with QLX_SPECIALIZE_RELABEL, which may move the state to its second
 * buffer. */
int qlx_state_device_ptr(const qlx_state* st, void** out);
/* FP64 sum |c|^2 on the GPU (reference reduction: cpu_state_vector.hpp:97-100). */
int qlx_state_norm_sq(qlx_state* st, double* out);
/* FP64 <a|b> on the GPU; a and b may differ in precision (c64 vs c128). */
int qlx_state_inner(qlx_state* a, qlx_state* b, double* re, double* im);
/* FP64 max |a_i - b_i| on the GPU. */
int qlx_state_max_abs_diff(qlx_state* a, qlx_state* b, double* out);
/* Copy b into a (same qubit count; precisions may differ). */
int qlx_state_copy(qlx_state* dst, qlx_state* src);

/* ---- ExecutionEngine::execute (reference include/backend/execution_engine.hpp:55-56,
 *      implemented by include/backend/cpu_openmp_engine.hpp:143-225) ---------
 * Enqueues instructions [0, n_instr) up to (not including) the first MEASURE;
 * *stop_at receives its index (or n_instr). The host then draws the outcome with
 * its own mt19937_64 (reference :632-633) via qlx_measure_prob1/qlx_collapse and
 * resumes after it. An unknown opcode ends the run like the reference's runtime_error
 * (:221-222): the instructions before it are executed (as the reference's loop has applied
 * them when it throws), then QLX_EUNSUPPORTED is returned. opts may be NULL (plain mode,
 * real GLOBAL_SWAP). */
int qlx_execute(qlx_state* st,
                const qlx_gate_instr* instrs, size_t n_instr,
                const double* pool_c128, size_t pool_elems,
                const qlx_matrix_entry* meta, size_t n_meta,
                const qlx_exec_opts* opts, size_t* stop_at);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops). qlx_execute in QLX_EXEC_PLAIN mode (opts NULL or
 * mode QLX_EXEC_PLAIN; any other mode -> QLX_EUNSUPPORTED) that allocates nothing: the matrices
 * of UNITARY / FUSED_BLOCK gates are staged in a fixed 1 MiB device array of the library (one
 * per GPU, held until the process ends) instead of the state's program buffer, so pool_elems
 * may be at most 65536 (larger -> QLX_EINVAL); dense gates run on the vector units; a state
 * with kernel timing enabled is refused (QLX_EINVAL). Same results as qlx_execute in plain
 * mode with QLX_DENSE_VALU. Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (states on different streams) race. */
int qlx_execute_noalloc(qlx_state* st,
                        const qlx_gate_instr* instrs, size_t n_instr,
                        const double* pool_c128, size_t pool_elems,
                        const qlx_matrix_entry* meta, size_t n_meta,
                        const qlx_exec_opts* opts, size_t* stop_at);
/* Host-only planning of QLX_EXEC_WINDOWED (or, with opts->mode = QLX_EXEC_DENSE_SWEEPS, of
 * dense-sweep) execution (no GPU work): the order in which the instructions of a
 * MEASURE-free stream would run. order[k] is the source index of the k-th executed
 * instruction and sweep_of[k] the window sweep that runs it (-1: its own kernel). Gates
 * are only reordered past gates on disjoint qubits, except merged rotations (see
 * qlx_exec_opts::plan_flags): a rotation merged into an earlier one is listed right after it,
 * having moved back past gates on its qubit that commute with it. */
typedef struct qlx_plan_info {
    uint64_t sweeps;           /* window-sweep launches (one full HBM pass each) */
    uint64_t phases;           /* register phases over all sweeps (LDS re-layouts = phases - sweeps) */
    uint64_t plain_steps;      /* instructions run by their own kernel */
    uint64_t gates_in_sweeps;
    uint32_t window_qubits;    /* W: qubits per chunk held by one work-group */
    uint32_t register_qubits;  /* qubits held in each work-item's registers */
    uint32_t threads_per_group;
    uint32_t coalesce_qubits;  /* low qubits always in the window */
    uint64_t mfma_phases;      /* complex64 generic sweeps: phases fused into one 32 x 32 matrix and
                                  applied on the matrix cores */
    uint64_t rotations_merged; /* rotations merged into an earlier one (applied as one op) */
} qlx_plan_info;
int qlx_plan_windowed(uint32_t num_qubits, int dtype, const qlx_gate_instr* instrs, size_t n,
                      const double* pool_c128, size_t pool_elems,
                      const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts* opts,
                      int64_t* order, int32_t* sweep_of, size_t cap, size_t* n_out, qlx_plan_info* info);

/* Host-only: the HIP source that QLX_EXEC_WINDOWED with opts->specialize compiles for sweep
 * `sweep_index` of this stream (for inspection and tests). *len = source length; `out`
 * (cap bytes) receives it NUL-terminated, truncated to cap - 1. */
int qlx_plan_sweep_source(uint32_t num_qubits, int dtype, const qlx_gate_instr* instrs, size_t n,
                          const double* pool_c128, size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta,
                          const qlx_exec_opts* opts, uint32_t sweep_index, char* out, size_t cap, size_t* len);

/* Waits up to timeout_s seconds (< 0: no limit) for the background planning and compilation
 * that QLX_SPECIALIZE_ASYNC executions started; *pending (may be NULL) receives the number of
 * programs still queued or compiling. Returns the first failure of a finished compilation
 * (QLX_EHIP, with its log in qlx_last_error) once, else QLX_OK. */
int qlx_specialize_wait(double timeout_s, uint32_t* pending);

/* Sweep accounting of the last qlx_execute on this state. */
int qlx_last_exec_stats(const qlx_state* st, qlx_exec_stats* out);

/* ---- MEASURE (reference include/backend/cpu_openmp_engine.hpp:617-671) ---- */
/* P(qubit q = 1), FP64 reduction on the GPU (reference :621-630). Synchronous. */
int qlx_measure_prob1(qlx_state* st, uint32_t q, double* p1);
/* Zero the other half and scale the kept half by inv_norm (reference :649-670). */
int qlx_collapse(qlx_state* st, uint32_t q, int outcome, double inv_norm);

/* Wait for all work on the state's stream; reports asynchronous kernel errors. */
int qlx_sync(qlx_state* st);

/* ---- kernel timing (no reference counterpart: the reference times execute() with
 *      Google Benchmark wall clock, benchmarks/bench_hw_awareness.cpp:102-107) ---
 * When enabled, every kernel the state launches is bracketed by a HIP event pair on
 * the state's stream; qlx_timing_read synchronises and returns per-kernel totals
 * accumulated since the last qlx_timing_reset. */
typedef struct qlx_kernel_timing {
    char     kernel[48];
    uint64_t launches;
    double   total_ms;
    double   algorithmic_bytes;  /* sum over launches of f * 2 * sizeof(amp) * 2^N */
} qlx_kernel_timing;
int qlx_timing_enable(qlx_state* st, int on);
int qlx_timing_reset(qlx_state* st);
int qlx_timing_read(qlx_state* st, qlx_kernel_timing* out, size_t cap, size_t* n_out);

/* ---- Middleware::transpile_gpu (host C++; reference middleware.hpp:466-481, whose
 *      GPU branch :474-476 calls the empty stub :866) ----------------------------
 * Same pass pipeline as transpile_cpu_openmp (middleware.hpp:488-524): DAG build,
 * pass_gate_fusion (:526-753), pass_qubit_windowing (:759-860), DAG release. */

/* Windowing cost model (TranspilerConfig::cost_model_type, reference middleware.hpp:352-356). */
#define QLX_COST_THRESHOLD 0  /* ThresholdCostModel(k_safe, numa_penalty, swap_penalty) (:63-119) */
#define QLX_COST_GREEDY    1  /* GreedyCostModel(k_safe): penalties 1 and 0 (:121-131) */
#define QLX_COST_CUSTOM    2  /* the caller's model, through custom_evaluate */
/* WindowingCostModel::evaluate_swap (reference middleware.hpp:44-61). The three arrays hold
 * num_qubits entries each. Returns 1 with the (active, victim) logical qubits of the swap to
 * inject, 0 for none, a negative value to abort the transpilation (QLX_EINVAL). */
typedef int (*qlx_swap_eval_fn)(void* ctx, const uint32_t* window_frequencies, const uint16_t* log_to_phys,
                                const uint16_t* phys_to_log, uint32_t num_qubits, uint16_t* active,
                                uint16_t* victim);

typedef struct qlx_transpile_opts {
    uint32_t enable_gate_fusion;  /* TranspilerConfig::enable_gate_fusion */
    uint32_t max_fusion_width;    /* 1..5; 0 = 4 (the reference's MAX_GATE_ARITY, middleware.hpp:137) */
    uint32_t fuse_rotations;      /* also fuse RX/RY/RZ/SWAP/CCX/CSWAP (not fusable in the reference, :442-448) */
    uint32_t enable_windowing;    /* TranspilerConfig::enable_hw_awareness */
    uint32_t k_safe;              /* TranspilerConfig::k_safe; 0 = qlx_gpu_safe_qubit_limit */
    uint32_t numa_penalty;        /* ThresholdCostModel (middleware.hpp:63-119); 0 = 10 */
    uint32_t swap_penalty;
    uint32_t restore_layout;      /* append GLOBAL_SWAPs restoring logical qubit order */
    int32_t  dtype;               /* precision the window is sized for (QLX_C128 / QLX_C64) */
    uint32_t build_on_device;     /* form the fused-block matrices in FP64 on the current GPU
                                     (the host decides the blocks); 0 = on the host */
    uint32_t cost_model;          /* QLX_COST_*; with QLX_COST_CUSTOM, k_safe is the model's
                                     get_k_safe() (0 = qlx_gpu_safe_qubit_limit) */
    uint32_t reserved;
    qlx_swap_eval_fn custom_evaluate;  /* QLX_COST_CUSTOM: the model (NULL -> QLX_EINVAL, as the
                                          reference's transpile throws invalid_argument, :510-511) */
    void*    custom_ctx;          /* passed to custom_evaluate */
} qlx_transpile_opts;
typedef struct qlx_program qlx_program;
int qlx_transpile(uint32_t num_qubits, const qlx_gate_instr* instrs, size_t n,
                  const double* pool_c128, size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta,
                  const qlx_transpile_opts* opts, qlx_program** out);
int qlx_program_sizes(const qlx_program* p, size_t* n_instr, size_t* pool_elems, size_t* n_meta);
int qlx_program_copy(const qlx_program* p, qlx_gate_instr* instrs, double* pool_c128, qlx_matrix_entry* meta);
int qlx_program_destroy(qlx_program* p);
/* GPU HardwareTopology::get_safe_qubit_limit (reference hw_topology.hpp:47-53): the number
 * of qubits one work-group's window holds for this precision (device < 0: no query). */
int qlx_gpu_safe_qubit_limit(int device, int dtype, uint32_t* k_safe);

/* ---- batched execution (no reference counterpart: the reference runs one circuit
 *      per engine) ---------------------------------------------------------------
 * Many independent circuits, each with its own state in one HBM arena. States start
 * from |0...0>. Circuits must have >= 10 qubits and contain only gates that run inside
 * a window sweep (1-qubit gates, CX, SWAP, dense gates on <= 2 qubits). */
#define QLX_BATCH_PERSISTENT         0  /* one persistent launch over a queue of (circuit, sweep, chunk) tasks */
#define QLX_BATCH_PER_CIRCUIT        1  /* every circuit as its own stream of one-kernel-per-gate launches */
#define QLX_BATCH_PER_CIRCUIT_GRAPH  2  /* the same launches captured once into a hipGraph, then replayed */
#define QLX_BATCH_LEVELS             3  /* one launch per sweep level over all circuits */
typedef struct qlx_batch qlx_batch;
typedef struct qlx_batch_info {
    uint64_t circuits;
    uint64_t tasks;            /* (circuit, sweep, chunk) tasks of one run */
    uint64_t sweeps;           /* window sweeps over all circuits */
    uint64_t state_bytes;      /* arena size */
    uint64_t gates;            /* instructions over all circuits */
    uint32_t workgroups;       /* persistent grid */
    uint32_t max_sweeps;       /* longest circuit, in sweeps */
} qlx_batch_info;
int qlx_batch_create(int dtype, qlx_batch** out);
int qlx_batch_add(qlx_batch* b, uint32_t num_qubits, const qlx_gate_instr* instrs, size_t n,
                  const double* pool_c128, size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta,
                  uint32_t* circuit_id);
/* Plans every circuit, uploads descriptors, allocates the arena. opts: window/register
 * qubits for circuits at least one window wide (smaller circuits use their own size);
 * specialize = QLX_SPECIALIZE_RELABEL: circuits larger than one window get two state buffers
 * and run their sweeps out of place (windows free of the low qubits unless plan_flags has
 * QLX_PLAN_FIXED_WINDOWS); other specialize flags are ignored (batched sweeps always run the
 * generic kernel). */
int qlx_batch_prepare(qlx_batch* b, const qlx_exec_opts* opts, qlx_batch_info* info);
/* Enqueues one run of every circuit from |0...0> (asynchronous). */
int qlx_batch_run(qlx_batch* b, int mode);
/* Waits for the last run; *elapsed_ms = its GPU time (HIP events around the run). */
int qlx_batch_sync(qlx_batch* b, double* elapsed_ms);
int qlx_batch_state_to_host(qlx_batch* b, uint32_t circuit_id, double* out_c128, size_t dim);
/* FP64 comparison (on the GPU) of one batch state with a reference state of the same
 * size: the state infidelity 1 - |<ref|psi>|^2 / (<ref|ref> <psi|psi>) and
 * max |ref_i - psi_i|. */
int qlx_batch_compare(qlx_batch* b, uint32_t circuit_id, qlx_state* ref, double* infidelity, double* max_abs);
int qlx_batch_destroy(qlx_batch* b);

/* ---- serving (no reference counterpart) -----------------------------------------
 * A continuous-batching request loop: submitted circuits (>= 10 qubits; 1-qubit gates,
 * CX, SWAP, GLOBAL_SWAP) are admitted FIFO while their states fit the HBM budget, grouped
 * by qubit count, and every launch of a group advances all of its active requests by one
 * window sweep; requests join and leave between launches. Results stay in HBM until
 * released. qlx_server_pump never blocks on device work. */
typedef struct qlx_server qlx_server;
int qlx_server_create(int dtype, uint64_t hbm_budget_bytes, qlx_server** out);
/* A request larger than one window runs out of place by default: its arena block holds two
 * state buffers, and each sweep re-lays out the qubits for the next (relabelled sweeps,
 * windows free of the low qubits: fewer sweeps). QLX_SERVE_IN_PLACE keeps one buffer per
 * request (half the HBM per request, more sweeps). Requests whose two buffers exceed the budget
 * run in place. Set before the first submit. */
#define QLX_SERVE_IN_PLACE 1u
int qlx_server_set_flags(qlx_server* s, uint32_t flags);
int qlx_server_submit(qlx_server* s, uint32_t num_qubits, const qlx_gate_instr* instrs, size_t n,
                      uint64_t* request_id);
/* Harvests finished launches, admits queued requests, launches the next sweep of every
 * group with room (two launches in flight per group). */
int qlx_server_pump(qlx_server* s, uint32_t* queued, uint32_t* active, uint32_t* inflight);
/* Requests completed since the last call, with submit-to-completion latency in seconds. */
int qlx_server_take_completed(qlx_server* s, uint64_t* ids, double* latency_s, size_t cap, size_t* n_out);
/* Where a request is: QLX_REQ_QUEUED (submitted, not yet admitted), QLX_REQ_ACTIVE (admitted:
 * its state is in the arena and its sweeps are being launched), QLX_REQ_DONE (its last sweep
 * has completed). times_s (may be NULL) receives {submit, admit, done} in seconds since the
 * server was created, -1 for a step not reached; admission and completion are the times the
 * server's pump observed them. */
#define QLX_REQ_QUEUED 0
#define QLX_REQ_ACTIVE 1
#define QLX_REQ_DONE   2
int qlx_server_request_info(qlx_server* s, uint64_t id, int* status, double* times_s);
/* As qlx_batch_compare, for a completed request. */
int qlx_server_compare(qlx_server* s, uint64_t id, qlx_state* ref, double* infidelity, double* max_abs);
int qlx_server_state_to_host(qlx_server* s, uint64_t id, double* out_c128, size_t dim);
int qlx_server_release(qlx_server* s, uint64_t id);
int qlx_server_destroy(qlx_server* s);

/* ---- host-side helpers (CPU; no GPU needed) ------------------------------ */
/* Seeded layered h/rx/cx circuit: H on every qubit, then `depth` layers of
 * RX(theta) on every qubit (theta = (mt19937_64() >> 11) * 2^-53 * 2*pi, stored
 * as float, layer-major) followed by CX(q, q+1) for q = l%2, l%2+2, ... < n-1.
 * Writes at most cap instructions; *n_out = required count. */
int qlx_gen_layered(uint32_t num_qubits, uint32_t depth, uint64_t seed,
                    qlx_gate_instr* out, size_t cap, size_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* QLX_ABI_H_ */
