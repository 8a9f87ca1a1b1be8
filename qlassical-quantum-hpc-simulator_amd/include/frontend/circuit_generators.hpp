// circuit_generators.hpp — the reference's synthetic benchmark circuits
// (include/frontend/circuit_generators.hpp:22-70), restated over this package's builder:
//   make_hea_circuit(n, depth): per layer RY(q, theta), RZ(q, 1.5 theta) on every qubit, then
//       the linear CX chain (q, q+1); theta starts at 0.1234 and grows by 0.05 per layer.
//   make_qft_circuit(n): H(i) followed by CX(j, i) for j > i, for every i; then SWAP(i, n-1-i).
//   make_clustered_circuit(n, cluster, depth): the circuit of the reference's strategy-matrix
//       benchmark (benchmarks/bench_hw_awareness.cpp:30-53). The qubits split into clusters of
//       `cluster` consecutive qubits (the last may be shorter); each cluster in turn gets
//       `depth` layers of H on every member followed by the CX chain (q, q+1) inside it.
// The angle arithmetic (double, narrowed to float by the builder) matches the reference
// so the instruction bytes are identical.
#ifndef QLASSICAL_FRONTEND_CIRCUIT_GENERATORS_HPP
#define QLASSICAL_FRONTEND_CIRCUIT_GENERATORS_HPP

#include <cstdint>

#include <frontend/frontend.hpp>

namespace qlassical::frontend {

inline QuantumCircuit make_hea_circuit(uint32_t num_qubits, uint32_t depth) {
    QuantumCircuit qc(num_qubits, std::size_t{num_qubits} * depth * 3);
    double theta = 0.1234;
    for (uint32_t layer = 0; layer < depth; ++layer) {
        for (uint32_t q = 0; q < num_qubits; ++q) {
            qc.ry(static_cast<int16_t>(q), theta);
            qc.rz(static_cast<int16_t>(q), theta * 1.5);
        }
        for (uint32_t q = 0; q + 1 < num_qubits; ++q) qc.cx(static_cast<int16_t>(q), static_cast<int16_t>(q + 1));
        theta += 0.05;
    }
    return qc;
}

inline QuantumCircuit make_qft_circuit(uint32_t num_qubits) {
    QuantumCircuit qc(num_qubits, std::size_t{num_qubits} * num_qubits);
    for (uint32_t i = 0; i < num_qubits; ++i) {
        qc.h(static_cast<int16_t>(i));
        for (uint32_t j = i + 1; j < num_qubits; ++j) qc.cx(static_cast<int16_t>(j), static_cast<int16_t>(i));
    }
    for (uint32_t i = 0; i < num_qubits / 2; ++i)
        qc.swap(static_cast<int16_t>(i), static_cast<int16_t>(num_qubits - 1 - i));
    return qc;
}

inline QuantumCircuit make_clustered_circuit(uint32_t num_qubits, uint32_t cluster_size, uint32_t depth) {
    QuantumCircuit qc(num_qubits, std::size_t{num_qubits} * depth * 2);
    for (uint32_t lo = 0; lo < num_qubits; lo += cluster_size) {
        const uint32_t hi = lo + cluster_size < num_qubits ? lo + cluster_size : num_qubits;
        for (uint32_t d = 0; d < depth; ++d) {
            for (uint32_t q = lo; q < hi; ++q) qc.h(static_cast<int16_t>(q));
            for (uint32_t q = lo; q + 1 < hi; ++q) qc.cx(static_cast<int16_t>(q), static_cast<int16_t>(q + 1));
        }
    }
    return qc;
}

}  // namespace qlassical::frontend

#endif  // QLASSICAL_FRONTEND_CIRCUIT_GENERATORS_HPP
