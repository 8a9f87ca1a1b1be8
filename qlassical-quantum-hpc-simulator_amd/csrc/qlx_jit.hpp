// qlx_jit.hpp — circuit-specialised window-sweep kernels compiled at run time (hipRTC).
#pragma once

#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "qlx_sweep_core.hpp"

namespace qlx {

// Compiled sweep kernels of one plan (one module per sweep; unloaded with the object).
struct JitSweeps {
    std::vector<hipModule_t> modules;
    std::vector<hipFunction_t> fns;
    std::vector<uint64_t> resident;  // work-groups of each kernel the device holds at once
    unsigned threads = 0;
    bool persistent = false;         // kernels loop over chunks (grid = resident work-groups)
    bool relabel = false;            // out-of-place sweeps: read st->amp, write st->alt, then swap
    double compile_s = 0.0;
    JitSweeps() = default;
    JitSweeps(const JitSweeps&) = delete;
    JitSweeps& operator=(const JitSweeps&) = delete;
    ~JitSweeps();
};

// Kernel name of a generated sweep: qlx_sweep_spec_<program tag, 6 hex digits>_<sweep>, so
// that traces tell the sweeps of different programs apart.
std::string jit_kernel_name(const SweepPlan& p, int sweep_index);
// Whether the kernels of this geometry are persistent (see the generator).
bool jit_persistent(const Geometry& g, int dtype);
// HIP source of sweep `sweep_index` of the plan, specialised to its descriptors.
std::string jit_sweep_source(const SweepPlan& p, int sweep_index, int dtype);
// Generates and compiles every sweep of the plan for gfx950 (host threads; no device work):
// one code object per sweep. `cancel` (may be null) stops before the next sweep's compilation.
std::vector<std::vector<char>> compile_jit_sweeps(const SweepPlan& p, int dtype, const std::atomic<bool>* cancel);
// Loads compiled sweeps on the current device.
std::shared_ptr<JitSweeps> load_jit_sweeps(const SweepPlan& p, int dtype, const std::vector<std::vector<char>>& code,
                                           double compile_s);
// compile_jit_sweeps + load_jit_sweeps.
std::shared_ptr<JitSweeps> build_jit_sweeps(const SweepPlan& p, int dtype);
// One launch of a compiled sweep over the state: one work-group per chunk, or, for the
// persistent kernels, as many work-groups as the device holds at once (wg_override: that
// many instead, when nonzero). Out-of-place sweeps read st->amp, write st->alt and swap the
// two pointers.
void launch_jit_sweep(const JitSweeps& j, int sweep_index, qlx_state* st, int window_qubits, uint32_t wg_override);

}  // namespace qlx
