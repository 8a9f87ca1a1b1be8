// check_reference_headers.cpp — compile-only drop-in check: the GPU backend's headers
// compiled against the reference's OWN frontend/backend interface headers
// (include path: the reference's include/ first, then this package's include/).
// The reference defines ExecutionEngine::execute(const IRModule&) in
// cpu_openmp_engine.hpp (Eigen-dependent, not built here); this TU does not call it.
#include <backend/execution_engine.hpp>
#include <backend/hip_engine.hpp>
#include <frontend/frontend.hpp>
#include <middleware/gpu_middleware.hpp>

#include <memory>
#include <type_traits>

static_assert(std::is_base_of_v<qlassical::backend::ExecutionEngine, qlassical::backend::HIPEngine>);
static_assert(std::is_base_of_v<qlassical::backend::StateVector, qlassical::backend::GPUStateVector>);
static_assert(!std::is_copy_constructible_v<qlassical::backend::HIPEngine>);
static_assert(std::is_move_constructible_v<qlassical::backend::GPUStateVector>);

// A CircuitRunner-style block loop, as the reference's Backend::GPU case would run it.
[[maybe_unused]] static void gpu_block(qlassical::IRModule& module, qlassical::backend::GPUStateVector* carried,
                                       std::unique_ptr<qlassical::backend::ExecutionEngine>& keep) {
    qlassical::gpu::transpile_gpu(module);
    std::unique_ptr<qlassical::backend::HIPEngine> engine =
        carried ? std::make_unique<qlassical::backend::HIPEngine>(std::move(*carried))
                : std::make_unique<qlassical::backend::HIPEngine>(module.num_qubits);
    engine->execute(module.program(), module.unitaries());
    keep = std::move(engine);
}
