// test_hip_engine.cpp — the reference's engine, measurement, integration, state-vector and
// runner tests, restated against HIPEngine / GPUStateVector / CircuitRunner on the GPU:
//   tests/backend/test_engine_kernels.cpp, test_engine_measure.cpp, test_state_vector.cpp,
//   tests/integration/algos_no_optimization.cpp, tests/frontend/test_circuit_runner.cpp
// Every case runs twice: one kernel per gate, and windowed (LDS/register sweeps).
#include <cmath>
#include <complex>
#include <memory>
#include <type_traits>
#include <vector>

#include <backend/hip_engine.hpp>
#include <frontend/circuit_runner.hpp>
#include <frontend/frontend.hpp>
#include <middleware/gpu_middleware.hpp>

#include "mini_test.hpp"

using namespace qlassical;
using namespace qlassical::backend;

static bool g_windowed = false;
static bool g_relabel = false;

static HIPEngineOptions opts() {
    HIPEngineOptions o;
    o.windowed = g_windowed;
    o.specialize = g_relabel;
    o.relabel = g_relabel;
    return o;
}

static GateInstr G(GateType t, uint8_t q0 = 255, uint8_t q1 = 255, uint8_t q2 = 255, float param = 0.0f,
                   uint32_t mat = 0) {
    GateInstr g{};
    g.type = t;
    uint8_t ar = 0;
    g.qubits[0] = g.qubits[1] = g.qubits[2] = g.qubits[3] = 255;
    if (q0 != 255) { g.qubits[0] = q0; ar = 1; }
    if (q1 != 255) { g.qubits[1] = q1; ar = 2; }
    if (q2 != 255) { g.qubits[2] = q2; ar = 3; }
    g.flags_arity = ar;
    if (t == GateType::RX || t == GateType::RY || t == GateType::RZ) g.payload.param = param;
    else if (t == GateType::UNITARY || t == GateType::FUSED_BLOCK) g.payload.matrix_idx = mat;
    return g;
}

const double S2 = 1.0 / std::sqrt(2.0);

TEST_CASE("GPUStateVector (test_state_vector.cpp:14-47)") {
    GPUStateVector sv;
    CHECK(sv.empty());
    sv.initialize(3);
    CHECK(sv.dimension() == 8 && sv.num_qubits() == 3 && !sv.empty());
    auto a = sv.amplitudes();
    CHECK(a[0] == std::complex<double>(1.0, 0.0));
    for (std::size_t i = 1; i < a.size(); ++i) CHECK(a[i] == std::complex<double>(0.0, 0.0));
    sv.initialize(5);
    CHECK(sv.dimension() == 32 && sv.amplitudes()[0] == std::complex<double>(1.0, 0.0));
    std::vector<std::complex<double>> bad = {1.0, 1.0, 0.0};
    CHECK_THROWS_AS(sv.initialize(bad), std::invalid_argument);
    std::vector<std::complex<double>> unnorm = {1.0, 1.0};
    CHECK_THROWS_AS(sv.initialize(unnorm), std::invalid_argument);
    CHECK(std::is_move_constructible_v<GPUStateVector> && !std::is_copy_constructible_v<GPUStateVector>);
}

TEST_CASE("single qubit kernels (test_engine_kernels.cpp:48-93)") {
    UnitaryPool pool;
    auto run1 = [&](std::vector<GateInstr> prog) {
        HIPEngine e(1, std::nullopt, 0, opts());
        e.execute(prog, pool);
        return std::vector<std::complex<double>>(e.state().begin(), e.state().end());
    };
    auto s = run1({G(GateType::X, 0)});
    CHECK_NEAR(s[0].real(), 0.0, 1e-6);
    CHECK_NEAR(s[1].real(), 1.0, 1e-6);
    s = run1({G(GateType::H, 0)});
    CHECK_NEAR(s[0].real(), S2, 1e-6);
    CHECK_NEAR(s[1].real(), S2, 1e-6);
    s = run1({G(GateType::Y, 0)});
    CHECK_NEAR(s[0].real(), 0.0, 1e-6);
    CHECK_NEAR(s[1].imag(), 1.0, 1e-6);
    s = run1({G(GateType::X, 0), G(GateType::Z, 0)});
    CHECK_NEAR(s[1].real(), -1.0, 1e-6);
    s = run1({G(GateType::X, 0), G(GateType::S, 0)});
    CHECK_NEAR(s[1].imag(), 1.0, 1e-6);
    s = run1({G(GateType::RX, 0, 255, 255, (float)M_PI)});
    CHECK_NEAR(s[1].imag(), -1.0, 1e-6);
    s = run1({G(GateType::RY, 0, 255, 255, (float)(M_PI / 2.0))});
    CHECK_NEAR(s[0].real(), S2, 1e-6);
    CHECK_NEAR(s[1].real(), S2, 1e-6);
}

TEST_CASE("multi qubit routing (test_engine_kernels.cpp:118-171)") {
    UnitaryPool pool;
    HIPEngine e(3, std::nullopt, 0, opts());
    std::vector<GateInstr> p1 = {G(GateType::X, 2), G(GateType::CX, 2, 0)};
    e.execute(p1, pool);
    CHECK_NEAR(e.state()[5].real(), 1.0, 1e-6);
    e.reset();
    std::vector<GateInstr> p2 = {G(GateType::X, 0), G(GateType::CX, 0, 2)};
    e.execute(p2, pool);
    CHECK_NEAR(e.state()[5].real(), 1.0, 1e-6);
    e.reset();
    std::vector<GateInstr> p3 = {G(GateType::X, 0), G(GateType::SWAP, 0, 2)};
    e.execute(p3, pool);
    CHECK_NEAR(e.state()[4].real(), 1.0, 1e-6);
    e.reset();
    std::vector<GateInstr> p4 = {G(GateType::X, 2), G(GateType::SWAP, 2, 0)};
    e.execute(p4, pool);
    CHECK_NEAR(e.state()[1].real(), 1.0, 1e-6);
    e.reset();
    std::vector<GateInstr> p5 = {G(GateType::X, 0), G(GateType::X, 1), G(GateType::CCX, 1, 0, 2)};
    e.execute(p5, pool);
    CHECK_NEAR(e.state()[7].real(), 1.0, 1e-6);
    e.reset();
    std::vector<GateInstr> p6 = {G(GateType::X, 0), G(GateType::X, 1), G(GateType::CSWAP, 1, 0, 2)};
    e.execute(p6, pool);
    CHECK_NEAR(e.state()[6].real(), 1.0, 1e-6);
}

TEST_CASE("custom unitary gather/scatter (test_engine_kernels.cpp:173-202)") {
    HIPEngine e(3, std::nullopt, 0, opts());
    UnitaryPool pool;
    std::vector<std::complex<double>> sw = {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1};
    const uint32_t idx = pool.store(sw, 2);
    std::vector<GateInstr> prog = {G(GateType::X, 0), G(GateType::UNITARY, 0, 2, 255, 0.0f, idx)};
    e.execute(prog, pool);
    CHECK_NEAR(e.state()[1].real(), 0.0, 1e-6);
    CHECK_NEAR(e.state()[4].real(), 1.0, 1e-6);
}

TEST_CASE("phase kickback (test_engine_kernels.cpp:204-229)") {
    HIPEngine e(2, std::nullopt, 0, opts());
    UnitaryPool pool;
    std::vector<GateInstr> prog = {G(GateType::H, 0), G(GateType::X, 1), G(GateType::H, 1), G(GateType::CX, 0, 1)};
    e.execute(prog, pool);
    auto s = e.state();
    CHECK_NEAR(s[0].real(), 0.5, 1e-6);
    CHECK_NEAR(s[1].real(), -0.5, 1e-6);
    CHECK_NEAR(s[2].real(), -0.5, 1e-6);
    CHECK_NEAR(s[3].real(), 0.5, 1e-6);
}

TEST_CASE("measurement (test_engine_measure.cpp:46-123)") {
    UnitaryPool pool;
    std::vector<GateInstr> prog = {G(GateType::H, 0), G(GateType::MEASURE, 0)};
    {
        HIPEngine e(1, std::nullopt, 0, opts());
        e.execute(prog, pool);
        CHECK(e.measurements().size() == 1);
        const uint8_t o = e.measurements()[0];
        CHECK_NEAR(std::norm(e.state()[o]), 1.0, 1e-6);
        CHECK_NEAR(std::norm(e.state()[1 - o]), 0.0, 1e-6);
    }
    {
        HIPEngine e(1, std::nullopt, 0, opts());
        e.set_seed(42);
        e.execute(prog, pool);
        const uint8_t o1 = e.measurements()[0];
        e.reset();
        e.set_seed(42);
        e.execute(prog, pool);
        CHECK(o1 == e.measurements()[0]);
        e.reset();
        CHECK(e.measurements().empty());
        CHECK_NEAR(std::norm(e.state()[0]), 1.0, 1e-6);
    }
    {
        HIPEngine e(2, std::nullopt, 0, opts());
        std::vector<GateInstr> bell = {G(GateType::H, 0), G(GateType::CX, 0, 1), G(GateType::MEASURE, 0),
                                       G(GateType::MEASURE, 1)};
        int ones = 0;
        for (int i = 0; i < 100; ++i) {
            e.reset();
            e.execute(bell, pool);
            CHECK(e.measurements().size() == 2);
            CHECK(e.measurements()[0] == e.measurements()[1]);
            ones += e.measurements()[0];
        }
        CHECK(ones > 0 && ones < 100);
    }
}

TEST_CASE("Deutsch algorithm (algos_no_optimization.cpp:18-95)") {
    std::vector<std::complex<double>> init = {0.0, 0.0, 1.0, 0.0};
    for (int variant = 0; variant < 4; ++variant) {
        HIPEngine e(init, std::optional<uint64_t>(37), 1, opts());
        QuantumCircuit c(2);
        c.h(0).h(1);
        if (variant == 1) c.x(1);
        if (variant == 2) c.cx(0, 1);
        if (variant == 3) c.cx(0, 1).x(0);
        c.h(0);
        IRModule m = std::move(c).release();
        e.execute(m);
        auto s = e.state();
        const bool constant = variant < 2;
        CHECK_NEAR(std::norm(s[0]), constant ? 0.5 : 0.0, 1e-6);
        CHECK_NEAR(std::norm(s[1]), constant ? 0.0 : 0.5, 1e-6);
        CHECK_NEAR(std::norm(s[2]), constant ? 0.5 : 0.0, 1e-6);
        CHECK_NEAR(std::norm(s[3]), constant ? 0.0 : 0.5, 1e-6);
    }
}

TEST_CASE("quantum teleportation (algos_no_optimization.cpp:98-160)") {
    const std::complex<double> alpha(1.0 / std::sqrt(3.0), 0.0), beta(0.0, std::sqrt(2.0 / 3.0));
    std::vector<std::complex<double>> init(8, 0.0);
    init[0] = alpha;
    init[1] = beta;
    HIPEngine e(init, std::optional<uint64_t>(37), 2, opts());
    QuantumCircuit alice(3);
    alice.h(1).cx(1, 2).cx(0, 1).h(0).measure(0).measure(1);
    e.execute(alice.release());
    CHECK(e.measurements().size() == 2);
    const uint8_t m0 = e.measurements()[0], m1 = e.measurements()[1];
    QuantumCircuit bob(3);
    if (m1) bob.x(2);
    if (m0) bob.z(2);
    e.execute(bob.release());
    const std::size_t i0 = (m1 << 1) | m0, i1 = 4 | (m1 << 1) | m0;
    auto s = e.state();
    CHECK_NEAR(std::norm(s[i0]), std::norm(alpha), 1e-6);
    CHECK_NEAR(std::norm(s[i1]), std::norm(beta), 1e-6);
    const auto ratio = s[i1] / s[i0], want = beta / alpha;
    CHECK_NEAR(ratio.real(), want.real(), 1e-6);
    CHECK_NEAR(ratio.imag(), want.imag(), 1e-6);
    for (std::size_t i = 0; i < 8; ++i)
        if (i != i0 && i != i1) CHECK_NEAR(std::norm(s[i]), 0.0, 1e-6);
}

TEST_CASE("unknown opcode throws runtime_error (cpu_openmp_engine.hpp:221-222)") {
    HIPEngine e(2, std::nullopt, 0, opts());
    UnitaryPool pool;
    GateInstr bad = G(GateType::X, 0);
    bad.type = static_cast<GateType>(0x7F);
    std::vector<GateInstr> prog = {G(GateType::X, 1), bad};
    CHECK_THROWS_AS(e.execute(prog, pool), std::runtime_error);
    CHECK_NEAR(e.state()[2].real(), 1.0, 1e-12);  // the prefix ran
    std::vector<GateInstr> oor = {G(GateType::H, 7)};
    CHECK_THROWS_AS(e.execute(oor, pool), std::out_of_range);
}

TEST_CASE("CircuitRunner construction and guards (test_circuit_runner.cpp:28-126)") {
    auto qc = std::make_shared<QuantumCircuit>(2);
    CircuitRunner r(qc);
    CHECK(r.num_blocks() == 1 && r.backend() == Backend::GPU && r.engine() == nullptr);
    CHECK_THROWS_AS(r.state(), std::runtime_error);
    CHECK_THROWS_AS(r.num_qubits(), std::runtime_error);
    std::shared_ptr<QuantumCircuit> null_qc;
    CHECK_THROWS_AS(CircuitRunner(null_qc), std::invalid_argument);
    std::vector<std::shared_ptr<QuantumCircuit>> empty;
    CHECK_THROWS_AS(CircuitRunner(empty), std::invalid_argument);
    std::vector<std::shared_ptr<QuantumCircuit>> with_null = {qc, nullptr};
    CHECK_THROWS_AS(CircuitRunner(with_null), std::invalid_argument);
    CHECK(std::is_move_constructible_v<CircuitRunner> && !std::is_copy_constructible_v<CircuitRunner>);
    CircuitRunner mpi(std::make_shared<QuantumCircuit>(2), Backend::MPI);
    CHECK_THROWS_AS(mpi.run(), std::runtime_error);
}

static CircuitRunner runner(std::vector<std::shared_ptr<QuantumCircuit>> blocks) {
    return CircuitRunner(std::move(blocks), Backend::GPU, gpu::GpuTranspilerConfig{}, opts());
}

TEST_CASE("CircuitRunner single block (test_circuit_runner.cpp:132-223)") {
    {
        auto r = runner({std::make_shared<QuantumCircuit>(1)});
        r.run();
        CHECK(r.num_qubits() == 1 && r.state().size() == 2);
        CHECK_NEAR(r.state()[0].real(), 1.0, 1e-12);
        CHECK_NEAR(r.state()[1].real(), 0.0, 1e-12);
    }
    {
        auto qc = std::make_shared<QuantumCircuit>(2);
        qc->h(0).x(1);
        auto r = runner({qc});
        r.run();
        auto s = r.state();
        CHECK_NEAR(s[2].real(), S2, 1e-6);
        CHECK_NEAR(s[3].real(), S2, 1e-6);
        double sum = 0;
        for (auto a : s) sum += std::norm(a);
        CHECK_NEAR(sum, 1.0, 1e-12);
    }
    {
        auto qc = std::make_shared<QuantumCircuit>(2);
        qc->h(0).cx(0, 1);  // Bell through the runner with fusion on (test_circuit_runner.cpp:465-491)
        auto r = runner({qc});
        r.run();
        CHECK_NEAR(r.state()[0].real(), S2, 1e-6);
        CHECK_NEAR(r.state()[3].real(), S2, 1e-6);
    }
}

TEST_CASE("CircuitRunner multi block device hand-off (test_circuit_runner.cpp:229-329)") {
    {
        auto b1 = std::make_shared<QuantumCircuit>(2);
        b1->h(0);
        auto b2 = std::make_shared<QuantumCircuit>(2);
        b2->x(1);
        auto b3 = std::make_shared<QuantumCircuit>(2);
        b3->h(0);
        auto r = runner({b1, b2, b3});
        r.run();
        CHECK_NEAR(r.state()[2].real(), 1.0, 1e-6);
    }
    {
        auto b1 = std::make_shared<QuantumCircuit>(1);
        b1->x(0);
        auto b2 = std::make_shared<QuantumCircuit>(1);
        auto b3 = std::make_shared<QuantumCircuit>(1);
        b3->h(0);
        auto r = runner({b1, b2, b3});
        r.run();
        CHECK_NEAR(r.state()[0].real(), S2, 1e-6);
        CHECK_NEAR(r.state()[1].real(), -S2, 1e-6);
    }
    {
        std::vector<std::shared_ptr<QuantumCircuit>> blocks;
        for (int i = 0; i < 6; ++i) {
            auto b = std::make_shared<QuantumCircuit>(1);
            b->x(0);
            blocks.push_back(b);
        }
        auto r = runner(blocks);
        r.run();
        CHECK_NEAR(r.state()[0].real(), 1.0, 1e-6);
    }
    {
        auto r = runner({std::make_shared<QuantumCircuit>(2), std::make_shared<QuantumCircuit>(3)});
        CHECK_THROWS_AS(r.run(), std::invalid_argument);
    }
}

TEST_CASE("uniform superposition 4q and 14q (test_circuit_runner.cpp:355-395)") {
    for (int n : {4, 14}) {
        auto qc = std::make_shared<QuantumCircuit>(n);
        for (int q = 0; q < n; ++q) qc->h(q);
        auto r = runner({qc});
        r.run();
        auto s = r.state();
        const double want = std::pow(2.0, -n / 2.0);
        double total = 0.0, err = 0.0;
        for (auto a : s) {
            err = std::max(err, std::abs(a - std::complex<double>(want, 0.0)));
            total += std::norm(a);
        }
        CHECK(err < 1e-12);
        CHECK_NEAR(total, 1.0, 1e-12);
    }
}

TEST_CASE("complex64 dense sweeps on a width-5 fused module (HIPEngineOptions::dense_sweeps)") {
    // H on every qubit, then 8 layers of RX and brick CX; fused to width 5 with rotations
    const int n = 15;
    QuantumCircuit qc(n);
    for (int q = 0; q < n; ++q) qc.h(q);
    for (int l = 0; l < 8; ++l) {
        for (int q = 0; q < n; ++q) qc.rx(q, 0.1 + 0.37 * q + 0.91 * l);
        for (int q = l % 2; q + 1 < n; q += 2) qc.cx(q, q + 1);
    }
    IRModule m = std::move(qc).release();
    gpu::GpuTranspilerConfig cfg;
    cfg.max_fusion_width = 5;
    cfg.fuse_rotations = true;
    cfg.precision = Precision::C64;
    gpu::transpile_gpu(m, cfg);
    HIPEngine ref(n);
    ref.execute(m);
    const auto want = ref.state();
    for (bool relabel : {false, true}) {
        HIPEngineOptions o;
        o.precision = Precision::C64;
        o.dense_sweeps = true;
        o.relabel = relabel;
        HIPEngine e(n, std::nullopt, 0, o);
        e.execute(m);
        const auto got = e.state();
        std::complex<double> ip = 0.0;
        double na = 0.0, nb = 0.0;
        for (std::size_t i = 0; i < got.size(); ++i) {
            ip += std::conj(want[i]) * got[i];
            na += std::norm(want[i]);
            nb += std::norm(got[i]);
        }
        CHECK(1.0 - std::norm(ip) / (na * nb) <= 1e-9);
        CHECK_NEAR(nb, 1.0, 1e-5);
    }
}

TEST_CASE("block hand-off moves the device allocation (circuit_runner.hpp:154-163: move, never copy)") {
    // the sequence CircuitRunner::run performs between blocks: execute, extract the state
    // vector, build the next engine from it
    const int n = 12;
    QuantumCircuit a(n), b(n);
    for (int q = 0; q < n; ++q) a.h(q);
    for (int q = 0; q + 1 < n; ++q) b.cx(q, q + 1);
    IRModule ma = std::move(a).release(), mb = std::move(b).release();
    HIPEngine e1(n, std::nullopt, 0, opts());
    e1.execute(ma);
    void* p1 = nullptr;
    CHECK(qlx_state_device_ptr(e1.state_vector().handle(), &p1) == QLX_OK && p1 != nullptr);
    qlx_state* h1 = e1.state_vector().handle();
    GPUStateVector carried = std::move(e1).extract_state_vector();
    HIPEngine e2(std::move(carried), opts());
    void* p2 = nullptr;
    CHECK(e2.state_vector().handle() == h1);
    CHECK(qlx_state_device_ptr(e2.state_vector().handle(), &p2) == QLX_OK && p2 == p1);
    e2.execute(mb);
    const auto s = e2.state();
    double total = 0.0;
    for (auto v : s) total += std::norm(v);
    CHECK_NEAR(total, 1.0, 1e-12);
}

TEST_CASE("HIPEngineOptions{} at 22 qubits: generic sweeps, then the background-compiled ones") {
    if (g_windowed && !g_relabel) {
        const int n = 22;
        QuantumCircuit qc(n);
        for (int q = 0; q < n; ++q) qc.h(q);
        for (int l = 0; l < 6; ++l) {
            for (int q = 0; q < n; ++q) qc.rx(q, 0.3 + 0.11 * q + 0.7 * l);
            for (int q = l % 2; q + 1 < n; q += 2) qc.cx(q, q + 1);
        }
        IRModule m = std::move(qc).release();
        // a non-uniform start (the circuit maps |+...+> to itself up to a phase)
        std::vector<std::complex<double>> init(std::size_t(1) << n);
        double ns = 0.0;
        for (std::size_t i = 0; i < init.size(); ++i) {
            init[i] = {std::cos(0.37 * (double)i), std::sin(0.11 * (double)(i % 977))};
            ns += std::norm(init[i]);
        }
        for (auto& v : init) v /= std::sqrt(ns);
        HIPEngineOptions plain;
        plain.windowed = false;
        HIPEngine ref(std::span<const std::complex<double>>(init), std::nullopt, 0, plain);
        ref.execute(m);
        const std::vector<std::complex<double>> want(ref.state().begin(), ref.state().end());
        const std::span<const std::complex<double>> init_span(init);
        HIPEngine e{init_span};  // HIPEngineOptions{}
        CHECK(e.options().auto_specialize && e.options().windowed && !e.options().specialize);
        e.execute(m);
        double err = 0.0;
        for (std::size_t i = 0; i < want.size(); ++i) err = std::max(err, std::abs(e.state()[i] - want[i]));
        CHECK(err <= 1e-12);
        uint32_t pending = 1;
        CHECK(qlx_specialize_wait(300.0, &pending) == QLX_OK && pending == 0);
        e.reset(init_span);
        e.execute(m);
        qlx_exec_stats st{};
        CHECK(qlx_last_exec_stats(e.state_vector().handle(), &st) == QLX_OK);
        CHECK(st.specialised == 1 && st.relabelled == 1);
        err = 0.0;
        for (std::size_t i = 0; i < want.size(); ++i) err = std::max(err, std::abs(e.state()[i] - want[i]));
        CHECK(err <= 1e-12);
        // another engine (a later CircuitRunner block) runs the compiled sweeps at once
        HIPEngine e2{init_span};
        e2.execute(m);
        CHECK(qlx_last_exec_stats(e2.state_vector().handle(), &st) == QLX_OK && st.specialised == 1);
    }
}

int main(int argc, char** argv) {
    const char* filter = argc > 1 ? argv[1] : nullptr;
    std::printf("== one kernel per gate ==\n");
    g_windowed = false;
    int f = mt::run_all(filter);
    std::printf("== windowed sweeps ==\n");
    g_windowed = true;
    f += mt::run_all(filter);
    std::printf("== specialised out-of-place sweeps ==\n");
    g_relabel = true;
    f += mt::run_all(filter);
    return f ? 1 : 0;
}
