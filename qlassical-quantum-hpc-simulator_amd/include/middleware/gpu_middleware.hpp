// gpu_middleware.hpp — the GPU branch of Middleware::transpile and the GPU HardwareTopology.
//
// Middleware::transpile(module, Backend::GPU, cfg) calls transpile_gpu (reference
// include/middleware/middleware.hpp:474-476), which the reference leaves empty (:866).
// qlassical::gpu::transpile_gpu below runs the same pipeline as transpile_cpu_openmp
// (:488-524) through the library (qlx_transpile): DAG build, pass_gate_fusion
// (RX-aware, width <= 5), pass_qubit_windowing with a GPU cost model, DAG release.
// GPUTopology replaces CPUHwlocTopology (hw_topology.hpp:60-112): its safe qubit limit
// is the window one work-group holds on chip (LDS), not a NUMA domain.
#ifndef QLASSICAL_MIDDLEWARE_GPU_MIDDLEWARE_HPP
#define QLASSICAL_MIDDLEWARE_GPU_MIDDLEWARE_HPP

#include <cmath>
#include <complex>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include <backend/gpu_state_vector.hpp>
#include <frontend/frontend.hpp>

#include "qlx.h"

namespace qlassical::gpu {

struct GpuTranspilerConfig {
    bool enable_gate_fusion = true;     // TranspilerConfig::enable_gate_fusion
    uint8_t max_fusion_width = 4;       // 1..5 (the reference caps at 4)
    bool fuse_rotations = false;        // also fuse RX/RY/RZ/SWAP/CCX/CSWAP (GPU extension)
    bool enable_windowing = false;      // GLOBAL_SWAP windowing (the engine already plans LDS windows)
    std::optional<uint16_t> k_safe;     // TranspilerConfig::k_safe; default: GPUTopology
    uint32_t numa_penalty = 10;         // TranspilerConfig::numa_penalty
    uint32_t swap_penalty = 1;          // TranspilerConfig::swap_penalty
    bool restore_layout = true;         // append GLOBAL_SWAPs that restore logical order
    backend::Precision precision = backend::Precision::C128;
};

class GPUTopology {
public:
    explicit GPUTopology(int device = 0) : device_(device) {
        backend::qlx_throw_on_error(qlx_device_query(device, &info_));
    }
    [[nodiscard]] std::size_t get_hbm_bytes() const noexcept { return info_.hbm_bytes; }
    [[nodiscard]] std::size_t get_lds_bytes_per_cu() const noexcept { return info_.lds_bytes_per_cu; }
    [[nodiscard]] uint32_t get_num_cus() const noexcept { return info_.num_cus; }
    // Window qubits one work-group holds on chip (the GPU analogue of k_safe).
    [[nodiscard]] uint16_t get_safe_qubit_limit(backend::Precision p = backend::Precision::C128) const {
        uint32_t k = 0;
        backend::qlx_throw_on_error(qlx_gpu_safe_qubit_limit(device_, static_cast<int>(p), &k));
        return static_cast<uint16_t>(k);
    }
    // Largest state that fits 85% of free HBM (the reference's 15% reserve, hw_topology.hpp:47-53).
    [[nodiscard]] uint16_t get_max_state_qubits(backend::Precision p = backend::Precision::C128) const {
        const double amp = p == backend::Precision::C128 ? 16.0 : 8.0;
        const double usable = static_cast<double>(info_.hbm_free_bytes) * 0.85;
        return usable < amp ? 0 : static_cast<uint16_t>(std::floor(std::log2(usable / amp)));
    }

private:
    int device_;
    qlx_device_info info_{};
};

// Runs the GPU pass pipeline on `module` in place (same contract as Middleware::transpile).
inline void transpile_gpu(IRModule& module, const GpuTranspilerConfig& cfg = {}) {
    qlx_transpile_opts o{};
    o.enable_gate_fusion = cfg.enable_gate_fusion;
    o.max_fusion_width = cfg.max_fusion_width;
    o.fuse_rotations = cfg.fuse_rotations;
    o.enable_windowing = cfg.enable_windowing;
    o.k_safe = cfg.k_safe.value_or(0);
    o.numa_penalty = cfg.numa_penalty;
    o.swap_penalty = cfg.swap_penalty;
    o.restore_layout = cfg.restore_layout;
    o.dtype = static_cast<int>(cfg.precision);
    qlx_program* p = nullptr;
    const auto& pool = module.unitary_pool;
    backend::qlx_throw_on_error(qlx_transpile(
        module.num_qubits, reinterpret_cast<const qlx_gate_instr*>(module.gate_stream.data()),
        module.gate_stream.size(), reinterpret_cast<const double*>(pool.data.data()), pool.data.size(),
        reinterpret_cast<const qlx_matrix_entry*>(pool.metadata.data()), pool.metadata.size(), &o, &p));
    std::size_t ni = 0, np = 0, nm = 0;
    int rc = qlx_program_sizes(p, &ni, &np, &nm);
    if (rc == QLX_OK) {
        module.gate_stream.resize(ni);
        module.unitary_pool.data.resize(np);
        module.unitary_pool.metadata.resize(nm);
        rc = qlx_program_copy(p, reinterpret_cast<qlx_gate_instr*>(module.gate_stream.data()),
                              reinterpret_cast<double*>(module.unitary_pool.data.data()),
                              reinterpret_cast<qlx_matrix_entry*>(module.unitary_pool.metadata.data()));
    }
    qlx_program_destroy(p);
    backend::qlx_throw_on_error(rc);
}

}  // namespace qlassical::gpu

#endif  // QLASSICAL_MIDDLEWARE_GPU_MIDDLEWARE_HPP
