// frontend.hpp — Qlassical IR (GateType, GateInstr, UnitaryPool, IRModule) and the
// QuantumCircuit builder, restated for the gfx950 backend package.
//
// API- and layout-identical to the reference (include/frontend/frontend.hpp:26-529):
// same enum values, the 16-byte GateInstr with the same field offsets, the float
// angle payload, the builder's validation exceptions and the running uid. When this
// package is dropped into the reference tree, the reference's own header is used in
// its place (see the include order in the integration notes); everything the GPU
// backend needs from the IR is pinned by the static_asserts below.
#ifndef QLASSICAL_FRONTEND_HPP
#define QLASSICAL_FRONTEND_HPP

#include <algorithm>
#include <cassert>
#include <complex>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace qlassical {

enum class GateType : uint8_t {
    H = 0x00, X = 0x01, Y = 0x02, Z = 0x03, S = 0x04,
    RX = 0x10, RY = 0x11, RZ = 0x12,
    CX = 0x20, SWAP = 0x21,
    CCX = 0x30, CSWAP = 0x31,
    UNITARY = 0x80,
    FUSED_BLOCK = 0xE0, GLOBAL_SWAP = 0xE1,
    BARRIER = 0xF0, MEASURE = 0xF1,
};

namespace gate_flags {
inline constexpr uint8_t NONE = 0x00;
inline constexpr uint8_t ADJOINT = 0x01;
inline constexpr uint8_t FUSED_HEAD = 0x02;
inline constexpr uint8_t FUSED_TAIL = 0x03;
inline constexpr uint8_t COMPILER_GEN = 0x04;
}  // namespace gate_flags

inline constexpr uint32_t NO_MATRIX = UINT32_MAX;

// 16 bytes: type | flags_arity | qubits[4] | _pad | payload | uid.
struct alignas(16) GateInstr {
    GateType type = GateType::H;
    uint8_t flags_arity = 0;
    uint8_t qubits[4] = {255, 255, 255, 255};
    uint16_t _pad = 0;  // low byte: fifth qubit of a width-5 FUSED_BLOCK (GPU transpiler)
    union Payload {
        float param;
        uint32_t matrix_idx;
    } payload = {0.0f};
    uint32_t uid = 0;

    [[nodiscard]] uint8_t arity() const noexcept { return flags_arity & 0x0F; }
    void set_arity(uint8_t a) noexcept { flags_arity = static_cast<uint8_t>((flags_arity & 0xF0) | (a & 0x0F)); }
    [[nodiscard]] uint8_t flags() const noexcept { return (flags_arity >> 4) & 0x0F; }
    void set_flags(uint8_t f) noexcept { flags_arity = static_cast<uint8_t>((flags_arity & 0x0F) | ((f & 0x0F) << 4)); }
};
static_assert(sizeof(GateInstr) == 16, "GateInstr must stay 16 bytes");

struct UnitaryPool {
    struct MatrixEntry {
        uint32_t offset;
        uint32_t num_elements;
        uint8_t num_qubits;
        uint8_t _pad[3] = {};
    };

    std::vector<std::complex<double>> data;
    std::vector<MatrixEntry> metadata;

    [[nodiscard]] uint32_t store(std::span<const std::complex<double>> matrix, uint8_t num_qubits) {
        const uint32_t want = uint32_t{1} << (2u * num_qubits);
        if (matrix.size() != want) {
            throw std::invalid_argument("UnitaryPool::store: matrix size mismatch. Expected " + std::to_string(want) +
                                        ", got " + std::to_string(matrix.size()));
        }
        const auto idx = static_cast<uint32_t>(metadata.size());
        const auto off = static_cast<uint32_t>(data.size());
        data.insert(data.end(), matrix.begin(), matrix.end());
        metadata.push_back(MatrixEntry{off, want, num_qubits});
        return idx;
    }

    [[nodiscard]] std::span<const std::complex<double>> get(uint32_t matrix_idx) const {
        assert(matrix_idx < metadata.size() && "matrix_idx out of bounds");
        const MatrixEntry& e = metadata[matrix_idx];
        return {data.data() + e.offset, e.num_elements};
    }

    [[nodiscard]] std::size_t num_matrices() const noexcept { return metadata.size(); }
    [[nodiscard]] std::size_t total_elements() const noexcept { return data.size(); }
    [[nodiscard]] bool empty() const noexcept { return metadata.empty(); }
};

struct IRModule {
    uint32_t num_qubits = 0;
    std::vector<GateInstr> gate_stream;
    UnitaryPool unitary_pool;

    IRModule() = default;
    explicit IRModule(uint32_t nq, std::size_t gate_hint = 256) : num_qubits(nq) { gate_stream.reserve(gate_hint); }
    IRModule(IRModule&&) noexcept = default;
    IRModule& operator=(IRModule&&) noexcept = default;
    IRModule(const IRModule&) = delete;
    IRModule& operator=(const IRModule&) = delete;

    [[nodiscard]] std::span<const GateInstr> program() const noexcept { return gate_stream; }
    [[nodiscard]] std::span<GateInstr> program() noexcept { return gate_stream; }
    [[nodiscard]] const UnitaryPool& unitaries() const noexcept { return unitary_pool; }
    [[nodiscard]] UnitaryPool& unitaries() noexcept { return unitary_pool; }
    [[nodiscard]] std::size_t gate_count() const noexcept { return gate_stream.size(); }
    [[nodiscard]] bool empty() const noexcept { return gate_stream.empty(); }
};

class QuantumCircuit {
public:
    explicit QuantumCircuit(uint32_t num_qubits, std::size_t gate_count_hint = 256)
        : module_(num_qubits, gate_count_hint), next_uid_(0) {}
    QuantumCircuit(QuantumCircuit&&) noexcept = default;
    QuantumCircuit& operator=(QuantumCircuit&&) noexcept = default;
    QuantumCircuit(const QuantumCircuit&) = delete;
    QuantumCircuit& operator=(const QuantumCircuit&) = delete;

    [[nodiscard]] IRModule release() noexcept { return std::move(module_); }
    [[nodiscard]] uint32_t num_qubits() const noexcept { return module_.num_qubits; }
    [[nodiscard]] std::size_t size() const noexcept { return module_.gate_count(); }
    [[nodiscard]] std::span<const GateInstr> program() const noexcept { return module_.program(); }

    QuantumCircuit& h(int16_t q) { return one(GateType::H, q); }
    QuantumCircuit& x(int16_t q) { return one(GateType::X, q); }
    QuantumCircuit& y(int16_t q) { return one(GateType::Y, q); }
    QuantumCircuit& z(int16_t q) { return one(GateType::Z, q); }
    QuantumCircuit& s(int16_t q) { return one(GateType::S, q); }
    // the angle is narrowed to float, as the IR stores it
    QuantumCircuit& rx(int16_t q, double theta) { return one(GateType::RX, q, static_cast<float>(theta)); }
    QuantumCircuit& ry(int16_t q, double theta) { return one(GateType::RY, q, static_cast<float>(theta)); }
    QuantumCircuit& rz(int16_t q, double theta) { return one(GateType::RZ, q, static_cast<float>(theta)); }
    QuantumCircuit& cx(int16_t c, int16_t t) { return two(GateType::CX, c, t); }
    QuantumCircuit& cnot(int16_t c, int16_t t) { return cx(c, t); }
    QuantumCircuit& swap(int16_t a, int16_t b) { return two(GateType::SWAP, a, b); }
    QuantumCircuit& ccx(int16_t c0, int16_t c1, int16_t t) { return three(GateType::CCX, c0, c1, t); }
    QuantumCircuit& toffoli(int16_t c0, int16_t c1, int16_t t) { return ccx(c0, c1, t); }
    QuantumCircuit& cswap(int16_t c, int16_t a, int16_t b) { return three(GateType::CSWAP, c, a, b); }
    QuantumCircuit& fredkin(int16_t c, int16_t a, int16_t b) { return cswap(c, a, b); }

    QuantumCircuit& unitary(std::span<const int16_t> target_qubits, std::span<const std::complex<double>> matrix) {
        const auto k = static_cast<uint8_t>(target_qubits.size());
        if (k == 0 || k > 4) throw std::invalid_argument("unitary: arity must be 1–4, got " + std::to_string(k));
        const uint32_t midx = module_.unitary_pool.store(matrix, k);
        GateInstr g{};
        g.type = GateType::UNITARY;
        g.set_arity(k);
        g.payload.matrix_idx = midx;
        g.uid = next_uid_++;
        for (uint8_t i = 0; i < k; ++i) {
            check(target_qubits[i]);
            g.qubits[i] = static_cast<uint8_t>(target_qubits[i]);
        }
        module_.gate_stream.push_back(g);
        return *this;
    }

    QuantumCircuit& barrier(std::initializer_list<int16_t> qubits = {}) {
        GateInstr g{};
        g.type = GateType::BARRIER;
        g.set_arity(static_cast<uint8_t>(std::min<std::size_t>(qubits.size(), 4)));
        g.uid = next_uid_++;
        uint8_t i = 0;
        for (int16_t q : qubits) {
            if (i >= 4) break;
            check(q);
            g.qubits[i++] = static_cast<uint8_t>(q);
        }
        module_.gate_stream.push_back(g);
        return *this;
    }

    QuantumCircuit& measure(int16_t q) { return one(GateType::MEASURE, q); }
    QuantumCircuit& measure(std::initializer_list<int16_t> qubits) {
        for (int16_t q : qubits) measure(q);
        return *this;
    }

private:
    IRModule module_;
    uint32_t next_uid_;

    void check(int16_t q) const {
        if (q < 0 || q >= 256 || static_cast<uint32_t>(q) >= module_.num_qubits) {
            throw std::out_of_range("Qubit index " + std::to_string(q) + " out of range [0, min(256, " +
                                    std::to_string(module_.num_qubits) + "))");
        }
    }
    void append(GateInstr g) {
        g.uid = next_uid_++;
        module_.gate_stream.push_back(g);
    }
    QuantumCircuit& one(GateType t, int16_t q, float param = 0.0f) {
        check(q);
        GateInstr g;
        g.type = t;
        g.set_arity(1);
        g.qubits[0] = static_cast<uint8_t>(q);
        g.payload.param = param;
        append(g);
        return *this;
    }
    QuantumCircuit& two(GateType t, int16_t a, int16_t b) {
        check(a);
        check(b);
        if (a == b) {
            throw std::invalid_argument("Two-qubit gate requires distinct qubits, got " + std::to_string(a) +
                                        " == " + std::to_string(b));
        }
        GateInstr g;
        g.type = t;
        g.set_arity(2);
        g.qubits[0] = static_cast<uint8_t>(a);
        g.qubits[1] = static_cast<uint8_t>(b);
        append(g);
        return *this;
    }
    QuantumCircuit& three(GateType t, int16_t a, int16_t b, int16_t c) {
        check(a);
        check(b);
        check(c);
        if (a == b || a == c || b == c) throw std::invalid_argument("Three-qubit gate requires distinct qubits");
        GateInstr g;
        g.type = t;
        g.set_arity(3);
        g.qubits[0] = static_cast<uint8_t>(a);
        g.qubits[1] = static_cast<uint8_t>(b);
        g.qubits[2] = static_cast<uint8_t>(c);
        append(g);
        return *this;
    }
};

[[nodiscard]] constexpr uint8_t gate_arity(GateType type) noexcept {
    const auto v = static_cast<uint8_t>(type);
    return v < 0x20 ? 1 : v < 0x30 ? 2 : v < 0x40 ? 3 : 0;
}
[[nodiscard]] constexpr bool gate_is_parametric(GateType type) noexcept {
    const auto v = static_cast<uint8_t>(type);
    return v >= 0x10 && v < 0x20;
}
[[nodiscard]] constexpr bool gate_uses_matrix(GateType type) noexcept {
    return type == GateType::UNITARY || type == GateType::FUSED_BLOCK;
}
[[nodiscard]] constexpr bool gate_is_directive(GateType type) noexcept { return static_cast<uint8_t>(type) >= 0xE0; }

}  // namespace qlassical

#endif  // QLASSICAL_FRONTEND_HPP
