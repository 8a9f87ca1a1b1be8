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
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <backend/gpu_state_vector.hpp>
#include <frontend/frontend.hpp>

#include "qlx.h"

namespace qlassical::gpu {

// The windowing pass's cost model: the interface of the reference's WindowingCostModel
// (include/middleware/middleware.hpp:44-61), evaluate_swap returning the (active, victim)
// logical qubits of the next GLOBAL_SWAP for a window, or nullopt; get_k_safe the window's
// qubit limit. In the reference tree a reference model is passed through
// cost_model_adapter() below.
class WindowingCostModel {
public:
    virtual ~WindowingCostModel() = default;
    virtual std::optional<std::pair<uint16_t, uint16_t>> evaluate_swap(
        std::span<const uint32_t> window_frequencies, std::span<const uint16_t> log_to_phys,
        std::span<const uint16_t> phys_to_log) const = 0;
    virtual uint16_t get_k_safe() const = 0;
};

// Wraps any object with the reference's evaluate_swap / get_k_safe members (for example the
// reference's own ThresholdCostModel, GreedyCostModel or a user subclass of its
// WindowingCostModel) as a gpu::WindowingCostModel.
template <class Model>
class CostModelAdapter final : public WindowingCostModel {
public:
    explicit CostModelAdapter(std::shared_ptr<const Model> m) : m_(std::move(m)) {}
    std::optional<std::pair<uint16_t, uint16_t>> evaluate_swap(std::span<const uint32_t> f,
                                                               std::span<const uint16_t> l2p,
                                                               std::span<const uint16_t> p2l) const override {
        return m_->evaluate_swap(f, l2p, p2l);
    }
    uint16_t get_k_safe() const override { return m_->get_k_safe(); }

private:
    std::shared_ptr<const Model> m_;
};
template <class Model>
std::shared_ptr<WindowingCostModel> cost_model_adapter(std::shared_ptr<const Model> m) {
    return std::make_shared<CostModelAdapter<Model>>(std::move(m));
}

// TranspilerConfig::cost_model_type (reference middleware.hpp:352-356).
enum class CostModelType : uint8_t { THRESHOLD = 0, GREEDY = 1, CUSTOM = 2 };

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
    // windowing cost model (TranspilerConfig::cost_model_type / custom_cost_model,
    // middleware.hpp:366-382): a custom model is used whenever it is set
    CostModelType cost_model_type = CostModelType::THRESHOLD;
    std::shared_ptr<WindowingCostModel> custom_cost_model = nullptr;
    void set_cost_model(std::shared_ptr<WindowingCostModel> m) {
        custom_cost_model = std::move(m);
        cost_model_type = CostModelType::CUSTOM;
    }
};

namespace detail {
// qlx_swap_eval_fn over a gpu::WindowingCostModel; an exception aborts the pass (-1)
inline int evaluate_swap_cb(void* ctx, const uint32_t* f, const uint16_t* l2p, const uint16_t* p2l, uint32_t n,
                            uint16_t* active, uint16_t* victim) {
    try {
        const auto* m = static_cast<const WindowingCostModel*>(ctx);
        const auto r = m->evaluate_swap(std::span<const uint32_t>(f, n), std::span<const uint16_t>(l2p, n),
                                        std::span<const uint16_t>(p2l, n));
        if (!r) return 0;
        *active = r->first;
        *victim = r->second;
        return 1;
    } catch (...) {
        return -1;
    }
}
}  // namespace detail

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
    if (cfg.custom_cost_model) {
        o.cost_model = QLX_COST_CUSTOM;
        o.custom_evaluate = &detail::evaluate_swap_cb;
        o.custom_ctx = const_cast<WindowingCostModel*>(cfg.custom_cost_model.get());
        o.k_safe = cfg.custom_cost_model->get_k_safe();  // the pass's window limit is the model's (:771)
    } else {
        o.cost_model = static_cast<uint32_t>(cfg.cost_model_type);
    }
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
