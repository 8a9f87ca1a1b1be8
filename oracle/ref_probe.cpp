// ref_probe.cpp — thin C shim over the reference's own Eigen-free headers.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile against
// /root/reference/include (never copied) into oracle/_ref/libqlx_ref.so, so the
// tests can pin the oracle and the product against the reference's real code for
// the parts of the hot path that build without Eigen:
//   QuantumCircuit builder  include/frontend/frontend.hpp:218-497
//   UnitaryPool::store      include/frontend/frontend.hpp:132-150
//   insert_zero_bit         include/backend/state_vector_utils.hpp:23-28
//   CPUStateVector          include/backend/cpu_state_vector.hpp:54-115
//   make_hea/qft_circuit    include/frontend/circuit_generators.hpp:22-68
// The engine (cpu_openmp_engine.hpp) and middleware need Eigen 3.4 and are not built.

#include <backend/cpu_state_vector.hpp>
#include <backend/state_vector_utils.hpp>
#include <frontend/circuit_generators.hpp>
#include <frontend/frontend.hpp>

#include <cmath>
#include <cstring>
#include <random>

extern "C" {

// The layered h/rx/cx circuit built through the reference's QuantumCircuit.
int ref_gen_layered(uint32_t n, uint32_t depth, uint64_t seed, void* out, size_t cap, size_t* n_out) {
    qlassical::QuantumCircuit qc(n);
    std::mt19937_64 rng(seed);
    for (uint32_t q = 0; q < n; ++q) qc.h((int16_t)q);
    for (uint32_t l = 0; l < depth; ++l) {
        for (uint32_t q = 0; q < n; ++q) {
            const double theta = std::ldexp((double)(rng() >> 11), -53) * (2.0 * M_PI);
            qc.rx((int16_t)q, theta);
        }
        for (uint32_t q = l % 2; q + 1 < n; q += 2) qc.cx((int16_t)q, (int16_t)(q + 1));
    }
    qlassical::IRModule m = std::move(qc).release();
    *n_out = m.gate_stream.size();
    if (out) {
        if (cap < m.gate_stream.size()) return -1;
        std::memcpy(out, m.gate_stream.data(), m.gate_stream.size() * sizeof(qlassical::GateInstr));
    }
    return 0;
}

// The reference's own benchmark generators (circuit_generators.hpp): kind 0 = HEA, 1 = QFT.
int ref_gen_bench(int kind, uint32_t n, uint32_t depth, void* out, size_t cap, size_t* n_out) {
    qlassical::QuantumCircuit qc = kind == 0 ? qlassical::frontend::make_hea_circuit(n, depth)
                                             : qlassical::frontend::make_qft_circuit(n);
    qlassical::IRModule m = std::move(qc).release();
    *n_out = m.gate_stream.size();
    if (out) {
        if (cap < m.gate_stream.size()) return -1;
        std::memcpy(out, m.gate_stream.data(), m.gate_stream.size() * sizeof(qlassical::GateInstr));
    }
    return 0;
}

// Generic builder entry: ops = (type, q0, q1, q2) + params, one record per call.
// Returns the GateInstr bytes the reference builder would emit, or a negative code
// for the exception it throws: -1 invalid_argument, -2 out_of_range.
int ref_build(uint32_t n, const int32_t* ops, const double* params, size_t n_ops, void* out, size_t* n_out) {
    try {
        qlassical::QuantumCircuit qc(n);
        for (size_t i = 0; i < n_ops; ++i) {
            const int32_t* o = ops + 4 * i;
            const auto q0 = (int16_t)o[1], q1 = (int16_t)o[2], q2 = (int16_t)o[3];
            switch (o[0]) {
                case 0x00: qc.h(q0); break;
                case 0x01: qc.x(q0); break;
                case 0x02: qc.y(q0); break;
                case 0x03: qc.z(q0); break;
                case 0x04: qc.s(q0); break;
                case 0x10: qc.rx(q0, params[i]); break;
                case 0x11: qc.ry(q0, params[i]); break;
                case 0x12: qc.rz(q0, params[i]); break;
                case 0x20: qc.cx(q0, q1); break;
                case 0x21: qc.swap(q0, q1); break;
                case 0x30: qc.ccx(q0, q1, q2); break;
                case 0x31: qc.cswap(q0, q1, q2); break;
                case 0xF1: qc.measure(q0); break;
                default: return -3;
            }
        }
        qlassical::IRModule m = std::move(qc).release();
        *n_out = m.gate_stream.size();
        if (out) std::memcpy(out, m.gate_stream.data(), m.gate_stream.size() * sizeof(qlassical::GateInstr));
        return 0;
    } catch (const std::out_of_range&) {
        return -2;
    } catch (const std::invalid_argument&) {
        return -1;
    }
}

uint64_t ref_insert_zero_bit(uint64_t idx, uint8_t t) {
    return qlassical::backend::utils::insert_zero_bit((std::size_t)idx, t);
}

// CPUStateVector::initialize(span): 0 ok, -1 invalid_argument.
int ref_state_check(const double* amps, size_t dim) {
    try {
        qlassical::backend::CPUStateVector sv;
        sv.initialize(std::span<const std::complex<double>>(reinterpret_cast<const std::complex<double>*>(amps), dim));
        return 0;
    } catch (const std::invalid_argument&) {
        return -1;
    }
}

// CPUStateVector::initialize(n) writes |0...0>.
int ref_state_zero(uint32_t n, double* out) {
    qlassical::backend::CPUStateVector sv;
    sv.initialize(n);
    std::memcpy(out, sv.data(), sv.dimension() * sizeof(std::complex<double>));
    return 0;
}

// sizeof/alignof/offsets of GateInstr and MatrixEntry as the reference compiles them.
void ref_layout(uint64_t* out) {
    out[0] = sizeof(qlassical::GateInstr);
    out[1] = alignof(qlassical::GateInstr);
    out[2] = offsetof(qlassical::GateInstr, flags_arity);
    out[3] = offsetof(qlassical::GateInstr, qubits);
    out[4] = offsetof(qlassical::GateInstr, _pad);
    out[5] = offsetof(qlassical::GateInstr, payload);
    out[6] = offsetof(qlassical::GateInstr, uid);
    out[7] = sizeof(qlassical::UnitaryPool::MatrixEntry);
}

}  // extern "C"
