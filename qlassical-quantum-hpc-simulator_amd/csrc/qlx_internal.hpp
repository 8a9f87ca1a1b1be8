// qlx_internal.hpp — shared definitions of the gfx950 backend (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "qlx.h"

namespace qlx {

// ---- error reporting ---------------------------------------------------------
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

// Throw inside the library; converted to a status code at the ABI edge.
[[noreturn]] void fail(int code, const std::string& msg);

#define QLX_HIP_CHECK(expr)                                                        \
    do {                                                                           \
        hipError_t qlx_e_ = (expr);                                                \
        if (qlx_e_ != hipSuccess)                                                  \
            ::qlx::fail(QLX_EHIP, std::string(#expr ": ") + hipGetErrorString(qlx_e_)); \
    } while (0)

// Wraps an ABI entry point body: maps exceptions to status codes.
template <class F>
int guard(F&& f) noexcept {
    try {
        f();
        return QLX_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return QLX_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return QLX_EHIP;
    } catch (...) {
        set_error("unknown error");
        return QLX_EHIP;
    }
}

// ---- device buffer that grows on demand ---------------------------------------
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    // Grows the buffer; synchronises `stream` first so no queued kernel still reads it.
    void reserve(size_t bytes, hipStream_t stream);
    void release();
};

}  // namespace qlx

// ---- kernel timing records --------------------------------------------------------
namespace qlx {
struct TimingRec {
    const char* kernel;
    hipEvent_t e0, e1;
    double bytes;
};
struct TimingAgg {
    uint64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
};
}  // namespace qlx

// ---- the opaque state handle -----------------------------------------------------
struct qlx_state {
    bool timing = false;
    std::vector<qlx::TimingRec> pending;      // event pairs not yet read back
    std::vector<hipEvent_t> free_events;      // recycled events
    std::vector<std::pair<std::string, qlx::TimingAgg>> timing_agg;
    uint32_t nq = 0;
    int dtype = QLX_C128;
    int device = 0;
    void* amp = nullptr;          // 2^nq amplitudes (double2 or float2), HBM resident
    void* alt = nullptr;          // second buffer of `bytes` for out-of-place sweeps (or null)
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    qlx::DevBuf scratch;          // reduction partials
    qlx::DevBuf prog;             // uploaded gate descriptors / matrices
    std::shared_ptr<void> sweep_cache;  // last windowed plan (host plan + device image)
    std::shared_ptr<void> dense_cache;  // last dense-sweep plan (host plan + device image)
    qlx::DevBuf frags;            // MFMA operand fragments of the current program's dense gates
    qlx::DevBuf frag_desc;        // their build descriptors
    std::vector<int64_t> frag_index;  // per matrix_idx: byte offset into frags, or -1
    int frag_variant = 0;         // QLX_DENSE_MFMA_* the fragments were built for
    qlx_exec_stats stats{};
};

namespace qlx {

inline size_t amp_size(int dtype) { return dtype == QLX_C128 ? 16 : 8; }

// Brackets one kernel launch with a HIP event pair when timing is enabled.
class TimedLaunch {
public:
    TimedLaunch(qlx_state* st, const char* kernel, double bytes);
    ~TimedLaunch() noexcept(false);
    TimedLaunch(const TimedLaunch&) = delete;
    TimedLaunch& operator=(const TimedLaunch&) = delete;

private:
    qlx_state* st_;
    TimingRec rec_{};
};

// Algorithmic bytes of one full-state HBM pass: read + write of every amplitude.
inline double pass_bytes(const qlx_state* st) {
    return 2.0 * (double)amp_size(st->dtype) * (double)(1ull << st->nq);
}

// Kernel launch helpers (defined in qlx_state.hip).
void launch_fill_zero_state(qlx_state* st);
double reduce_norm_sq(qlx_state* st);
// FP64 <a|b>, max |a-b| and <b|b> over n amplitudes of (possibly different) precision, on `stream`.
void compare_amplitudes(const void* a, int dta, const void* b, int dtb, uint64_t n, hipStream_t stream,
                        DevBuf& scratch, double* re, double* im, double* maxd, double* norm_b = nullptr);
// State infidelity 1 - |<ref|psi>|^2 / (<ref|ref> <psi|psi>) from the FP64 sums.
inline double infidelity_from(double re, double im, double norm_ref, double norm_psi) {
    return 1.0 - (re * re + im * im) / (norm_ref * norm_psi);
}

// Plain (one sweep per instruction) executor, qlx_gates.hip.
void execute_plain(qlx_state* st, const qlx_gate_instr* instrs, size_t n,
                   const double* pool, size_t pool_elems,
                   const qlx_matrix_entry* meta, size_t n_meta,
                   const qlx_exec_opts& opts, qlx_exec_stats& stats);

// Windowed executor, qlx_sweep.hip.
void execute_windowed(qlx_state* st, const qlx_gate_instr* instrs, size_t n,
                      const double* pool, size_t pool_elems,
                      const qlx_matrix_entry* meta, size_t n_meta,
                      const qlx_exec_opts& opts, qlx_exec_stats& stats);

// Waits for background sweep compilations (QLX_SPECIALIZE_ASYNC), qlx_sweep.hip; returns the
// programs still pending and the first unreported failure in *failure.
uint32_t specialize_wait(double timeout_s, std::string* failure);

// Dense-sweep executor (QLX_EXEC_DENSE_SWEEPS), qlx_dsweep.hip.
void execute_dense_sweeps(qlx_state* st, const qlx_gate_instr* instrs, size_t n, const double* pool,
                          size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts& opts,
                          qlx_exec_stats& stats);

// Uploads the unitary pool (complex128) to the state's program buffer; returns it.
const double* upload_pool(qlx_state* st, const double* pool, size_t pool_elems);
// Launches the one-sweep kernel of a single instruction (BARRIER: nothing).
void launch_plain_instr(qlx_state* st, const qlx_gate_instr& g, const double* dpool, const qlx_matrix_entry* meta,
                        const qlx_exec_opts& opts, qlx_exec_stats& stats);

// Dense k-qubit unitary on the device state (k <= 5), matrix already on device
// as row-major complex128. Used by both executors.
// matrix_cores: complex128 4/5-qubit gates on states of >= k + 4 qubits use the FP64
// matrix-core kernel; otherwise (and with QLX_DENSE_VALU) the vector-unit kernel
void launch_dense(qlx_state* st, const double* d_matrix, const uint8_t* qubits, int k, bool matrix_cores);

// Matrix-core path for 4/5-qubit dense gates on complex64 states (qlx_mfma.hip).
struct MfmaSplitJob {
    uint64_t pool_off;  // complex elements into the device pool
    uint64_t frag_off;  // bytes into the fragment buffer
    int k;              // 4 or 5
    int variant;        // QLX_DENSE_MFMA_BF16 | QLX_DENSE_MFMA_F32 | QLX_DENSE_MFMA_F16
};
size_t mfma_fragment_bytes(int k, int variant);
// The FP16-form A fragments of k_mfma_split (hi block, then lo block) of a row-major
// interleaved complex 2^k x 2^k matrix, formed on the host; `out` holds
// mfma_fragment_bytes(k, QLX_DENSE_MFMA_F16) bytes.
void mfma_f16_fragments_host(const double* U, int k, unsigned char* out);
// Builds the A-operand fragments of each job from the FP64 pool (synchronous).
void mfma_split(const double* d_pool, const MfmaSplitJob* jobs, size_t n, void* d_frags, DevBuf& desc_buf,
                hipStream_t stream);
void launch_dense_mfma(qlx_state* st, const void* d_frag, const uint8_t* qubits, int k, int variant);
// Resolves opts.dense_kernel for the program's dense gates and builds the MFMA
// fragments they need (st->frag_index); call after upload_pool.
void prepare_dense(qlx_state* st, const qlx_gate_instr* instrs, size_t n, const double* dpool,
                   const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts& opts);

// FP64 fused-block products on the GPU (qlx_fusion.hip). A recipe lists a block's
// members in program order; member i's matrix (2^a x 2^a complex128, row-major, bit j
// <-> operand j) starts at complex element mat_off of member_mats and operand j sits on
// block-local bit pos[j]. `out` receives the blocks' 4^k-element matrices back to back
// (interleaved re/im), in recipe order.
struct FusedMember {
    uint32_t mat_off;
    int a;
    int pos[5];
};
struct FusedRecipe {
    int k;
    std::vector<FusedMember> members;
};
void build_fused_on_device(const std::vector<FusedRecipe>& recipes, const std::vector<double>& member_mats,
                           std::vector<double>& out);

// Instruction helpers shared by the executors.
inline int instr_arity(const qlx_gate_instr& g) {
    int a = g.flags_arity & 0x0F;
    return a;
}
inline int instr_qubit(const qlx_gate_instr& g, int i) {
    return i < 4 ? g.qubits[i] : (g.pad & 0xFF);
}

// Validates qubit operands of a standard/matrix instruction; throws QLX_ERANGE/EINVAL.
void validate_instr(const qlx_gate_instr& g, uint32_t nq, const qlx_matrix_entry* meta,
                    size_t n_meta, size_t pool_elems);

}  // namespace qlx
