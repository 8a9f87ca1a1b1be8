// circuit_runner.hpp — CircuitRunner for the GPU backend of this package.
//
// Mirrors the reference's orchestrator (include/frontend/circuit_runner.hpp:50-261):
// same constructors and validation, same run() loop (release each block's IRModule,
// transpile it, build the engine for block 0 or move the previous block's state into a
// new engine, execute, keep the last engine), same accessors and error types. The
// Backend::GPU case is the one the reference leaves throwing (:186-187): here it runs
// HIPEngine, and the state moves between blocks without leaving HBM (the
// GPUStateVector's device allocation is moved, never copied).
// This package carries no CPU engine: Backend::CPU_OPENMP and Backend::MPI throw
// std::runtime_error here; in the reference tree the same GPU case is added to the
// reference's own runner instead (see the integration notes).
#ifndef QLASSICAL_FRONTEND_CIRCUIT_RUNNER_HPP
#define QLASSICAL_FRONTEND_CIRCUIT_RUNNER_HPP

#include <complex>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <backend/execution_engine.hpp>
#include <backend/gpu_state_vector.hpp>
#include <backend/hip_engine.hpp>
#include <frontend/frontend.hpp>
#include <middleware/gpu_middleware.hpp>

namespace qlassical {

enum class Backend : uint8_t { CPU_OPENMP = 0x00, MPI = 0x01, GPU = 0x02 };

class CircuitRunner {
public:
    explicit CircuitRunner(std::shared_ptr<QuantumCircuit> circuit, Backend backend = Backend::GPU,
                           gpu::GpuTranspilerConfig config = {}, backend::HIPEngineOptions engine_opts = {})
        : backend_(backend), config_(config), engine_opts_(engine_opts) {
        if (!circuit) throw std::invalid_argument("CircuitRunner: circuit shared pointer cannot be null");
        circuits_.push_back(std::move(circuit));
    }

    explicit CircuitRunner(std::vector<std::shared_ptr<QuantumCircuit>> circuits, Backend backend = Backend::GPU,
                           gpu::GpuTranspilerConfig config = {}, backend::HIPEngineOptions engine_opts = {})
        : circuits_(std::move(circuits)), backend_(backend), config_(config), engine_opts_(engine_opts) {
        if (circuits_.empty()) throw std::invalid_argument("CircuitRunner: circuit block sequence cannot be empty");
        for (std::size_t i = 0; i < circuits_.size(); ++i)
            if (!circuits_[i])
                throw std::invalid_argument("CircuitRunner: null circuit pointer found at block index " +
                                            std::to_string(i));
    }

    CircuitRunner(CircuitRunner&&) noexcept = default;
    CircuitRunner& operator=(CircuitRunner&&) noexcept = default;
    CircuitRunner(const CircuitRunner&) = delete;
    CircuitRunner& operator=(const CircuitRunner&) = delete;
    ~CircuitRunner() = default;

    void run() {
        if (circuits_.empty()) throw std::runtime_error("CircuitRunner::run: no circuits provided to execute");
        if (backend_ == Backend::MPI) throw std::runtime_error("CircuitRunner::run: MPI backend is not yet implemented");
        if (backend_ == Backend::CPU_OPENMP)
            throw std::runtime_error("CircuitRunner::run: this package provides the GPU backend only");
        if (backend_ != Backend::GPU) throw std::invalid_argument("CircuitRunner::run: unknown backend type specified");

        backend::GPUStateVector carried(engine_opts_.precision);
        bool first = true;
        for (std::size_t b = 0; b < circuits_.size(); ++b) {
            auto& circuit = circuits_[b];
            if (!circuit)
                throw std::invalid_argument("CircuitRunner::run: circuit pointer at block index " + std::to_string(b) +
                                            " is null");
            IRModule module = std::move(*circuit).release();
            gpu::transpile_gpu(module, config_);
            const uint32_t nq = module.num_qubits;
            std::unique_ptr<backend::HIPEngine> engine;
            if (first) {
                engine = std::make_unique<backend::HIPEngine>(nq, std::nullopt, 0, engine_opts_);
                first = false;
            } else {
                if (carried.num_qubits() != nq) {
                    throw std::invalid_argument("CircuitRunner::run: qubit count mismatch between block " +
                                                std::to_string(b - 1) + " (" + std::to_string(carried.num_qubits()) +
                                                " qubits) and block " + std::to_string(b) + " (" +
                                                std::to_string(nq) + " qubits)");
                }
                engine = std::make_unique<backend::HIPEngine>(std::move(carried), engine_opts_);
            }
            engine->execute(module.program(), module.unitaries());
            if (b + 1 == circuits_.size()) engine_ = std::move(engine);
            else carried = std::move(*engine).extract_state_vector();
        }
    }

    [[nodiscard]] std::span<const std::complex<double>> state() const {
        if (!engine_) throw std::runtime_error("CircuitRunner::state: cannot inspect state before calling run()");
        return engine_->state();
    }
    [[nodiscard]] uint32_t num_qubits() const {
        if (!engine_)
            throw std::runtime_error("CircuitRunner::num_qubits: cannot inspect qubit count before calling run()");
        return engine_->num_qubits();
    }
    [[nodiscard]] const backend::ExecutionEngine* engine() const noexcept { return engine_.get(); }
    [[nodiscard]] const gpu::GpuTranspilerConfig& transpiler_config() const noexcept { return config_; }
    [[nodiscard]] gpu::GpuTranspilerConfig& transpiler_config() noexcept { return config_; }
    [[nodiscard]] Backend backend() const noexcept { return backend_; }
    [[nodiscard]] std::size_t num_blocks() const noexcept { return circuits_.size(); }
    [[nodiscard]] const std::vector<std::shared_ptr<QuantumCircuit>>& circuits() const noexcept { return circuits_; }

private:
    std::vector<std::shared_ptr<QuantumCircuit>> circuits_;
    Backend backend_ = Backend::GPU;
    gpu::GpuTranspilerConfig config_;
    backend::HIPEngineOptions engine_opts_;
    std::unique_ptr<backend::HIPEngine> engine_;
};

}  // namespace qlassical

#endif  // QLASSICAL_FRONTEND_CIRCUIT_RUNNER_HPP
