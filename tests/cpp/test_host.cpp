// test_host.cpp — CPU-only C++ tests of the package's host side: the restated IR and
// builder (reference tests/frontend/test_frontend.cpp) and transpile_gpu (reference
// tests/middleware/test_middleware.cpp). Links libqlx.so but never touches a GPU.
#include <complex>
#include <cstddef>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <utility>
#include <type_traits>
#include <vector>

#include <frontend/circuit_generators.hpp>
#include <frontend/frontend.hpp>
#include <middleware/gpu_middleware.hpp>

#include "mini_test.hpp"

using namespace qlassical;

TEST_CASE("GateInstr layout (test_frontend.cpp:26-42)") {
    CHECK(sizeof(GateInstr) == 16);
    CHECK(alignof(GateInstr) == 16);
    CHECK(offsetof(GateInstr, type) == 0);
    CHECK(offsetof(GateInstr, flags_arity) == 1);
    CHECK(offsetof(GateInstr, qubits) == 2);
    CHECK(offsetof(GateInstr, payload) == 8);
    CHECK(offsetof(GateInstr, uid) == 12);
    CHECK(std::is_trivially_copyable_v<GateInstr>);
    GateInstr g;
    CHECK(g.type == GateType::H && g.arity() == 0 && g.qubits[0] == 255 && g.uid == 0);
    g.set_arity(3);
    g.set_flags(gate_flags::COMPILER_GEN);
    CHECK(g.arity() == 3 && g.flags() == gate_flags::COMPILER_GEN);
}

TEST_CASE("UnitaryPool store/get (test_frontend.cpp:61-129)") {
    UnitaryPool pool;
    std::vector<std::complex<double>> m2(4, 1.0), m4(16, 2.0);
    CHECK(pool.store(m2, 1) == 0);
    CHECK(pool.store(m4, 2) == 1);
    CHECK(pool.num_matrices() == 2 && pool.total_elements() == 20);
    CHECK(pool.metadata[1].offset == 4 && pool.metadata[1].num_elements == 16 && pool.metadata[1].num_qubits == 2);
    CHECK(pool.get(1)[0] == std::complex<double>(2.0));
    CHECK_THROWS_AS(pool.store(m2, 2), std::invalid_argument);
}

TEST_CASE("QuantumCircuit builder (test_frontend.cpp:230-262, 581-633)") {
    QuantumCircuit qc(3);
    qc.h(0).cx(0, 1).rx(2, 3.14159265358979).measure({0, 1});
    CHECK(qc.size() == 5);
    auto prog = qc.program();
    CHECK(prog[2].type == GateType::RX);
    CHECK(prog[2].payload.param == static_cast<float>(3.14159265358979));
    for (uint32_t i = 0; i < prog.size(); ++i) CHECK(prog[i].uid == i);
    CHECK(prog[1].arity() == 2 && prog[1].qubits[0] == 0 && prog[1].qubits[1] == 1);
    CHECK_THROWS_AS(qc.h(3), std::out_of_range);
    CHECK_THROWS_AS(qc.h(-1), std::out_of_range);
    CHECK_THROWS_AS(qc.cx(1, 1), std::invalid_argument);
    CHECK_THROWS_AS(qc.ccx(0, 1, 1), std::invalid_argument);
    std::vector<int16_t> five = {0, 1, 2, 0, 1};
    std::vector<std::complex<double>> big(1024, 0.0);
    CHECK_THROWS_AS(qc.unitary(five, big), std::invalid_argument);
    IRModule m = std::move(qc).release();
    CHECK(m.num_qubits == 3 && m.gate_count() == 5);
    CHECK(gate_arity(GateType::CCX) == 3 && gate_is_parametric(GateType::RY) && gate_uses_matrix(GateType::FUSED_BLOCK));
    CHECK(gate_is_directive(GateType::GLOBAL_SWAP) && gate_is_directive(GateType::MEASURE) && !gate_is_directive(GateType::CX));
}

static std::vector<GateType> types(const IRModule& m) {
    std::vector<GateType> t;
    for (const auto& g : m.program()) t.push_back(g.type);
    return t;
}

TEST_CASE("transpile_gpu fusion heuristic (test_middleware.cpp:60-92)") {
    gpu::GpuTranspilerConfig cfg;  // reference-compatible fusion (no rotations), no windowing
    {
        QuantumCircuit qc(2);
        qc.h(0).x(1);
        IRModule m = std::move(qc).release();
        gpu::transpile_gpu(m, cfg);
        CHECK(m.gate_count() == 2);
        CHECK(m.program()[0].type == GateType::H && m.program()[1].type == GateType::X);
    }
    {
        QuantumCircuit qc(2);
        qc.h(0).cx(0, 1).x(1);
        IRModule m = std::move(qc).release();
        gpu::transpile_gpu(m, cfg);
        CHECK(m.gate_count() == 1);
        CHECK(m.program()[0].type == GateType::FUSED_BLOCK && m.program()[0].arity() == 2);
        CHECK(m.unitary_pool.num_matrices() == 1 && m.unitary_pool.metadata[0].num_elements == 16);
    }
    {  // the GPU extension fuses rotations, up to width 5
        QuantumCircuit qc(5);  // a chain the root-anchored BFS reaches from h(0)
        qc.h(0).rx(0, 0.1);
        for (int q = 0; q < 4; ++q) qc.cx(q, q + 1).rx(q + 1, 0.2 * q);
        IRModule m = std::move(qc).release();
        gpu::GpuTranspilerConfig c5;
        c5.fuse_rotations = true;
        c5.max_fusion_width = 5;
        gpu::transpile_gpu(m, c5);
        CHECK(m.gate_count() == 1);
        bool has5 = false;
        for (const auto& g : m.program()) has5 = has5 || (g.type == GateType::FUSED_BLOCK && g.arity() == 5);
        CHECK(has5);
    }
}

TEST_CASE("transpile_gpu windowing (test_middleware.cpp:94-140, 164-235, 296-323)") {
    auto make = [] {
        QuantumCircuit qc(10);
        qc.x(0).x(1).cx(8, 9);
        return std::move(qc).release();
    };
    gpu::GpuTranspilerConfig cfg;
    cfg.enable_gate_fusion = false;
    cfg.enable_windowing = true;
    cfg.restore_layout = false;
    cfg.k_safe = 2;
    cfg.numa_penalty = 1;
    cfg.swap_penalty = 0;
    IRModule m = make();
    gpu::transpile_gpu(m, cfg);
    CHECK(m.gate_count() == 5);
    auto p = m.program();
    CHECK(p[0].type == GateType::X && p[0].qubits[0] == 0);
    CHECK(p[1].type == GateType::X && p[1].qubits[0] == 1);
    CHECK(p[2].type == GateType::GLOBAL_SWAP && p[3].type == GateType::GLOBAL_SWAP);
    CHECK((p[2].qubits[0] == 8 || p[2].qubits[0] == 9) && (p[2].qubits[1] == 0 || p[2].qubits[1] == 1));
    CHECK(p[2].qubits[0] != p[3].qubits[0] && p[2].qubits[1] != p[3].qubits[1]);
    CHECK(p[4].type == GateType::CX && p[4].qubits[0] != p[4].qubits[1]);
    CHECK((p[4].qubits[0] == 0 || p[4].qubits[0] == 1) && (p[4].qubits[1] == 0 || p[4].qubits[1] == 1));

    cfg.swap_penalty = 100;  // retention 1 < 100: no swaps
    IRModule m2 = make();
    gpu::transpile_gpu(m2, cfg);
    CHECK(m2.gate_count() == 3 && m2.program()[2].qubits[0] == 8 && m2.program()[2].qubits[1] == 9);

    cfg.k_safe = 16;  // every qubit already safe
    cfg.numa_penalty = 10;
    cfg.swap_penalty = 1;
    IRModule m3 = make();
    gpu::transpile_gpu(m3, cfg);
    CHECK(m3.gate_count() == 3);

    cfg.k_safe = 2;
    cfg.restore_layout = true;  // GPU extension: the logical order is restored at the end
    IRModule m4 = make();
    gpu::transpile_gpu(m4, cfg);
    CHECK(m4.gate_count() > 5);
    CHECK(types(m4).back() == GateType::GLOBAL_SWAP);
}

namespace {
// CustomSpyCostModel of test_middleware.cpp:237-256: evaluates, never swaps
class SpyCostModel final : public gpu::WindowingCostModel {
public:
    mutable bool evaluated = false;
    uint16_t limit;
    explicit SpyCostModel(uint16_t l) : limit(l) {}
    uint16_t get_k_safe() const override { return limit; }
    std::optional<std::pair<uint16_t, uint16_t>> evaluate_swap(std::span<const uint32_t>, std::span<const uint16_t>,
                                                               std::span<const uint16_t>) const override {
        evaluated = true;
        return std::nullopt;
    }
};
// a model with the reference's member functions but no base class (as the reference's own
// ThresholdCostModel is to this package): always swaps the first active unsafe qubit with the
// first idle safe one (GreedyCostModel's choice)
struct PlainGreedy {
    uint16_t k;
    uint16_t get_k_safe() const { return k; }
    std::optional<std::pair<uint16_t, uint16_t>> evaluate_swap(std::span<const uint32_t> f, std::span<const uint16_t> l2p,
                                                               std::span<const uint16_t>) const {
        int cand = -1;
        uint32_t best = 0;
        for (std::size_t q = 0; q < f.size(); ++q)
            if (f[q] > 0 && l2p[q] > k && f[q] > best) {
                best = f[q];
                cand = (int)q;
            }
        if (cand < 0) return std::nullopt;
        for (std::size_t v = 0; v < f.size(); ++v)
            if (l2p[v] <= k && f[v] == 0) return std::make_pair((uint16_t)cand, (uint16_t)v);
        return std::nullopt;
    }
};
}  // namespace

TEST_CASE("custom windowing cost model (test_middleware.cpp:258-294)") {
    QuantumCircuit qc(6);
    qc.h(0).cx(4, 5);
    IRModule m = std::move(qc).release();
    auto spy = std::make_shared<SpyCostModel>(3);
    gpu::GpuTranspilerConfig cfg;
    cfg.enable_gate_fusion = false;
    cfg.enable_windowing = true;
    cfg.set_cost_model(spy);
    CHECK(cfg.cost_model_type == gpu::CostModelType::CUSTOM && cfg.custom_cost_model == spy);
    gpu::transpile_gpu(m, cfg);
    CHECK(spy->evaluated);
    CHECK(m.gate_count() == 2);

    // CUSTOM without a model: invalid_argument, as the reference's transpile throws
    QuantumCircuit qc2(4);
    qc2.h(0);
    IRModule m2 = std::move(qc2).release();
    gpu::GpuTranspilerConfig bad;
    bad.enable_windowing = true;
    bad.cost_model_type = gpu::CostModelType::CUSTOM;
    CHECK_THROWS_AS(gpu::transpile_gpu(m2, bad), std::invalid_argument);

    // an adapted model object and CostModelType::GREEDY give the greedy stream of the
    // windowing test above
    auto make = [] {
        QuantumCircuit q(10);
        q.x(0).x(1).cx(8, 9);
        return std::move(q).release();
    };
    gpu::GpuTranspilerConfig g1;
    g1.enable_gate_fusion = false;
    g1.enable_windowing = true;
    g1.restore_layout = false;
    g1.set_cost_model(gpu::cost_model_adapter(std::make_shared<const PlainGreedy>(PlainGreedy{2})));
    IRModule a = make();
    gpu::transpile_gpu(a, g1);
    gpu::GpuTranspilerConfig g2;
    g2.enable_gate_fusion = false;
    g2.enable_windowing = true;
    g2.restore_layout = false;
    g2.k_safe = 2;
    g2.cost_model_type = gpu::CostModelType::GREEDY;
    IRModule b = make();
    gpu::transpile_gpu(b, g2);
    CHECK(a.gate_count() == 5 && b.gate_count() == 5);
    for (std::size_t i = 0; i < 5; ++i) {
        CHECK(a.program()[i].type == b.program()[i].type);
        CHECK(a.program()[i].qubits[0] == b.program()[i].qubits[0] && a.program()[i].qubits[1] == b.program()[i].qubits[1]);
    }
}

TEST_CASE("clustered benchmark circuit: gate counts of the strategy matrix (bench.txt:9, 15, 21, 27)") {
    // 28 qubits, clusters of 6, depth 15: the reference prints 765 -> 765 / 774 / 76 / 85 gates
    // for no pass / SWAP windowing at k_safe 18 / fusion / both.
    auto make = [] { return frontend::make_clustered_circuit(28, 6, 15).release(); };
    IRModule base = make();
    CHECK(base.gate_count() == 765);
    auto p = base.program();
    CHECK(p[0].type == GateType::H && p[0].qubits[0] == 0);
    CHECK(p[6].type == GateType::CX && p[6].qubits[0] == 0 && p[6].qubits[1] == 1);
    CHECK(p[764].type == GateType::CX && p[764].qubits[0] == 26 && p[764].qubits[1] == 27);
    const struct { bool fuse, window; uint32_t gates; } rows[] = {
        {false, true, 774}, {true, false, 76}, {true, true, 85}};
    for (const auto& r : rows) {
        gpu::GpuTranspilerConfig cfg;
        cfg.enable_gate_fusion = r.fuse;
        cfg.fuse_rotations = false;
        cfg.max_fusion_width = 4;
        cfg.enable_windowing = r.window;
        cfg.restore_layout = false;
        cfg.k_safe = 18;
        IRModule m = make();
        gpu::transpile_gpu(m, cfg);
        CHECK(m.gate_count() == r.gates);
    }
}

int main(int argc, char** argv) { return mt::run_all(argc > 1 ? argv[1] : nullptr) ? 1 : 0; }
