// execution_engine.hpp — the abstract ExecutionEngine interface, restated
// (reference include/backend/execution_engine.hpp:20-79; identical API).
//
// The reference defines the non-virtual execute(const IRModule&) in
// cpu_openmp_engine.hpp:678-680; this package has no CPU engine, so the same
// one-line definition lives here.
#ifndef QLASSICAL_BACKEND_EXECUTION_ENGINE_HPP
#define QLASSICAL_BACKEND_EXECUTION_ENGINE_HPP

#include <complex>
#include <cstdint>
#include <span>
#include <vector>

#include <frontend/frontend.hpp>

namespace qlassical::backend {

class ExecutionEngine {
public:
    virtual ~ExecutionEngine() = default;
    virtual void reset() = 0;
    virtual void reset(std::span<const std::complex<double>> initial_state) = 0;
    virtual void set_seed(uint64_t seed) = 0;
    virtual void reserve_measurements(std::size_t count) = 0;
    [[nodiscard]] virtual const std::vector<uint8_t>& measurements() const = 0;
    virtual void execute(std::span<const GateInstr> program, const UnitaryPool& pool) = 0;
    void execute(const IRModule& module);
    [[nodiscard]] virtual std::span<const std::complex<double>> state() const = 0;
    [[nodiscard]] virtual uint32_t num_qubits() const = 0;

protected:
    ExecutionEngine() = default;
    ExecutionEngine(ExecutionEngine&&) = default;
    ExecutionEngine& operator=(ExecutionEngine&&) = default;
    ExecutionEngine(const ExecutionEngine&) = delete;
    ExecutionEngine& operator=(const ExecutionEngine&) = delete;
};

inline void ExecutionEngine::execute(const IRModule& module) { execute(module.program(), module.unitaries()); }

}  // namespace qlassical::backend

#endif  // QLASSICAL_BACKEND_EXECUTION_ENGINE_HPP
