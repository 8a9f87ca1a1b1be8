// qlx_host.cpp — host-only parts of the C ABI: version, device query (the GPU
// HardwareTopology source) and the seeded layered h/rx/cx circuit generator.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>

#include "qlx_internal.hpp"

using namespace qlx;

namespace {

// GateInstr construction exactly as QuantumCircuit::push_1q / push_1q_param / push_2q
// (reference include/frontend/frontend.hpp:419-460): default qubits 255, _pad 0,
// payload 0.0f, arity in the low nibble, uid = running counter.
qlx_gate_instr make_instr(uint8_t type, uint8_t arity, uint32_t uid) {
    qlx_gate_instr g;
    std::memset(&g, 0, sizeof(g));
    g.type = type;
    g.flags_arity = arity & 0x0F;
    g.qubits[0] = g.qubits[1] = g.qubits[2] = g.qubits[3] = 255;
    g.pad = 0;
    g.payload.param = 0.0f;
    g.uid = uid;
    return g;
}

}  // namespace

extern "C" {

const char* qlx_version(void) { return "qlx-gfx950 0.1"; }

int qlx_device_query(int device, qlx_device_info* out) {
    return guard([&] {
        if (!out) fail(QLX_EINVAL, "null output");
        std::memset(out, 0, sizeof(*out));
        hipDeviceProp_t p;
        QLX_HIP_CHECK(hipGetDeviceProperties(&p, device));
        int cur = 0;
        QLX_HIP_CHECK(hipGetDevice(&cur));
        if (cur != device) QLX_HIP_CHECK(hipSetDevice(device));
        size_t free_b = 0, total_b = 0;
        QLX_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
        if (cur != device) QLX_HIP_CHECK(hipSetDevice(cur));
        out->hbm_bytes = p.totalGlobalMem;
        out->hbm_free_bytes = free_b;
        out->lds_bytes_per_cu = (uint32_t)p.maxSharedMemoryPerMultiProcessor;
        out->num_cus = (uint32_t)p.multiProcessorCount;
        out->wavefront_size = (uint32_t)p.warpSize;
        out->clock_khz = (uint32_t)p.clockRate;
        std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
    });
}

int qlx_gen_layered(uint32_t n, uint32_t depth, uint64_t seed, qlx_gate_instr* out, size_t cap,
                    size_t* n_out) {
    return guard([&] {
        if (!n_out) fail(QLX_EINVAL, "null count output");
        if (n == 0 || n > 255) fail(QLX_EINVAL, "qubit count must be in 1..255");
        size_t cx_total = 0;
        for (uint32_t l = 0; l < depth; ++l)
            for (uint32_t q = l % 2; q + 1 < n; q += 2) ++cx_total;
        const size_t total = (size_t)n + (size_t)depth * n + cx_total;
        *n_out = total;
        if (!out) return;
        if (cap < total) fail(QLX_EINVAL, "output capacity too small");
        std::mt19937_64 rng(seed);
        const double two_pi = 6.283185307179586;  // nearest double to 2*pi
        uint32_t uid = 0;
        size_t k = 0;
        for (uint32_t q = 0; q < n; ++q) {
            qlx_gate_instr g = make_instr(QLX_H, 1, uid++);
            g.qubits[0] = (uint8_t)q;
            out[k++] = g;
        }
        for (uint32_t l = 0; l < depth; ++l) {
            for (uint32_t q = 0; q < n; ++q) {
                const double theta = std::ldexp((double)(rng() >> 11), -53) * two_pi;
                qlx_gate_instr g = make_instr(QLX_RX, 1, uid++);
                g.qubits[0] = (uint8_t)q;
                g.payload.param = static_cast<float>(theta);
                out[k++] = g;
            }
            for (uint32_t q = l % 2; q + 1 < n; q += 2) {
                qlx_gate_instr g = make_instr(QLX_CX, 2, uid++);
                g.qubits[0] = (uint8_t)q;
                g.qubits[1] = (uint8_t)(q + 1);
                out[k++] = g;
            }
        }
    });
}

}  // extern "C"
