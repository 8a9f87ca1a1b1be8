// hip_engine.hpp — HIPEngine: the ExecutionEngine for one MI355X (gfx950).
//
// Drop-in for CPUOpenMPEngine (reference include/backend/cpu_openmp_engine.hpp:51-672)
// behind the ExecutionEngine interface (include/backend/execution_engine.hpp:20-79):
//   * same constructors: (num_qubits, seed, reserve), (initial span, seed, reserve),
//     (GPUStateVector&&) — which, like the reference's (CPUStateVector&&), leaves the
//     measurement RNG default-seeded (:105-107);
//   * execute(span<GateInstr>, UnitaryPool) runs the stream on the device through
//     qlx_execute and is synchronous on return, as the reference's is (:143-225);
//   * MEASURE: P(1) is reduced on the GPU in FP64, the outcome is drawn on the host
//     with the same std::mt19937_64 + uniform_real_distribution<double>(0,1) pair
//     (:252-253, :632-633), and the collapse runs on the GPU (:617-671);
//   * unknown opcodes execute the prefix and then throw std::runtime_error (:221-222);
//   * state() returns a host span (synchronised copy of the HBM state).
// Differences, by design: GLOBAL_SWAP is a real qubit permutation (the reference's is
// a no-op, :217-219) unless HIPEngineOptions::global_swap_noop is set; qubit operands
// are validated before launch (out-of-range -> std::out_of_range) instead of being
// undefined behaviour; set_seed reseeds the measurement generator (the reference's
// body is empty, :130-131, while its seeded-determinism tests construct engines with a
// seed: reseeding is what a caller of set_seed asks for).
#ifndef QLASSICAL_BACKEND_HIP_ENGINE_HPP
#define QLASSICAL_BACKEND_HIP_ENGINE_HPP

#include <cmath>
#include <complex>
#include <cstdint>
#include <optional>
#include <random>
#include <span>
#include <vector>

#include <backend/execution_engine.hpp>
#include <backend/gpu_state_vector.hpp>
#include <frontend/frontend.hpp>

#include "qlx.h"

namespace qlassical::backend {

static_assert(sizeof(GateInstr) == sizeof(qlx_gate_instr), "GateInstr layout must match qlx_gate_instr");
static_assert(sizeof(UnitaryPool::MatrixEntry) == sizeof(qlx_matrix_entry), "MatrixEntry layout must match");
static_assert(sizeof(std::complex<double>) == 2 * sizeof(double), "complex128 is two doubles");

struct HIPEngineOptions {
    bool windowed = true;            // LDS/register window sweeps (else one kernel per gate)
    bool global_swap_noop = false;   // reproduce the reference's no-op GLOBAL_SWAP
    Precision precision = Precision::C128;
    // windowed: compile each program's sweeps into straight-line kernels (seconds of hipRTC
    // time on a program's first execution, reused afterwards: for programs run many times)
    bool specialize = false;
    // window sweeps (generic or specialised) or dense sweeps: out of place through a second
    // buffer of the state's size, every sweep's window on the low address bits
    bool relabel = false;
    // complex64 fused programs: every gate inside 13-qubit LDS windows, 4/5-qubit gates on the
    // matrix cores (QLX_EXEC_DENSE_SWEEPS; takes precedence over windowed)
    bool dense_sweeps = false;
    // windowed without specialize (the default): states of >= QLX_AUTO_SPECIALIZE_QUBITS qubits
    // get specialised, relabelled sweeps compiled on a background host thread (QLX_EXEC_AUTO).
    // A program's first executions run the generic sweep kernel; once its kernels are compiled
    // (once per process; the ROCm compiler's disk cache also serves later processes) every
    // execution of it, from any engine, runs them. Off: the generic sweep kernel only.
    bool auto_specialize = true;
    // window sweeps: merge a rotation into an earlier one of the same axis on the same qubit
    // when the gates between them on that qubit commute with it (QLX_PLAN_KEEP_ROTATIONS off);
    // the state is the same up to rounding
    bool merge_rotations = true;
};

class HIPEngine final : public ExecutionEngine {
public:
    using ExecutionEngine::execute;

    explicit HIPEngine(uint32_t num_qubits, std::optional<uint64_t> seed = std::nullopt, std::size_t reserve_meas = 0,
                       HIPEngineOptions opts = {})
        : sv_(opts.precision), opts_(opts) {
        sv_.initialize(num_qubits);
        seed_rng(seed);
        if (reserve_meas) measurements_.reserve(reserve_meas);
    }

    explicit HIPEngine(std::span<const std::complex<double>> initial_state, std::optional<uint64_t> seed = std::nullopt,
                       std::size_t reserve_meas = 0, HIPEngineOptions opts = {})
        : sv_(opts.precision), opts_(opts) {
        sv_.initialize(initial_state);
        seed_rng(seed);
        if (reserve_meas) measurements_.reserve(reserve_meas);
    }

    explicit HIPEngine(GPUStateVector&& sv, HIPEngineOptions opts = {}) noexcept : sv_(std::move(sv)), opts_(opts) {
        opts_.precision = sv_.precision();
    }

    void set_state_vector(GPUStateVector&& sv) noexcept { sv_ = std::move(sv); }

    void reset() override {
        sv_.initialize(sv_.num_qubits());
        measurements_.clear();
    }
    void reset(std::span<const std::complex<double>> initial_state) override {
        sv_.initialize(initial_state);
        measurements_.clear();
    }
    void set_seed(uint64_t seed) override { rng_.seed(seed); }
    void reserve_measurements(std::size_t count) override { measurements_.reserve(count); }
    [[nodiscard]] const std::vector<uint8_t>& measurements() const override { return measurements_; }

    void execute(std::span<const GateInstr> program, const UnitaryPool& pool) override {
        qlx_exec_opts o{};
        o.mode = opts_.dense_sweeps ? QLX_EXEC_DENSE_SWEEPS
                 : !opts_.windowed  ? QLX_EXEC_PLAIN
                 : (opts_.auto_specialize && !opts_.specialize) ? QLX_EXEC_AUTO
                                                                 : QLX_EXEC_WINDOWED;
        o.gswap_mode = opts_.global_swap_noop ? QLX_GSWAP_NOOP : QLX_GSWAP_PERMUTE;
        o.plan_flags = opts_.merge_rotations ? 0u : QLX_PLAN_KEEP_ROTATIONS;
        if (opts_.dense_sweeps)
            o.specialize = opts_.relabel ? QLX_SPECIALIZE_RELABEL : 0u;
        else
            o.specialize = (opts_.specialize ? QLX_SPECIALIZE : 0u) | (opts_.relabel ? QLX_SPECIALIZE_RELABEL : 0u);
        const auto* instrs = reinterpret_cast<const qlx_gate_instr*>(program.data());
        const auto* pdata = reinterpret_cast<const double*>(pool.data.data());
        const auto* meta = reinterpret_cast<const qlx_matrix_entry*>(pool.metadata.data());
        std::size_t i = 0;
        const std::size_t n = program.size();
        sv_.invalidate();
        while (i < n) {
            std::size_t stop = 0;
            qlx_throw_on_error(qlx_execute(sv_.handle(), instrs + i, n - i, pdata, pool.data.size(), meta,
                                           pool.metadata.size(), &o, &stop));
            i += stop;
            if (i >= n) break;
            apply_measure(program[i].qubits[0]);
            ++i;
        }
        qlx_throw_on_error(qlx_sync(sv_.handle()));
    }

    [[nodiscard]] std::span<const std::complex<double>> state() const override { return sv_.amplitudes(); }
    [[nodiscard]] uint32_t num_qubits() const override { return sv_.num_qubits(); }

    [[nodiscard]] GPUStateVector& state_vector() noexcept { return sv_; }
    [[nodiscard]] GPUStateVector extract_state_vector() && noexcept { return std::move(sv_); }
    [[nodiscard]] const HIPEngineOptions& options() const noexcept { return opts_; }
    void set_options(const HIPEngineOptions& o) noexcept { opts_ = o; }

private:
    GPUStateVector sv_;
    HIPEngineOptions opts_;
    std::vector<uint8_t> measurements_;
    std::mt19937_64 rng_;
    std::uniform_real_distribution<double> dist_{0.0, 1.0};

    void seed_rng(std::optional<uint64_t> seed) {
        if (seed.has_value()) {
            rng_.seed(*seed);
        } else {
            std::random_device rd;
            rng_.seed(rd());
        }
    }

    void apply_measure(uint8_t target) {
        double p1 = 0.0;
        qlx_throw_on_error(qlx_measure_prob1(sv_.handle(), target, &p1));
        const double r = dist_(rng_);
        const uint8_t outcome = (r < p1) ? 1 : 0;
        measurements_.push_back(outcome);
        double p = outcome ? p1 : (1.0 - p1);
        if (p < 1e-12) p = 1e-12;
        qlx_throw_on_error(qlx_collapse(sv_.handle(), target, outcome, 1.0 / std::sqrt(p)));
    }
};

}  // namespace qlassical::backend

#endif  // QLASSICAL_BACKEND_HIP_ENGINE_HPP
