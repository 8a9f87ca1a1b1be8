// qlx_dsweep.hip — dense window sweeps (QLX_EXEC_DENSE_SWEEPS) for complex64 states.
//
// A fused program (pass_gate_fusion with width 5, reference include/middleware/
// middleware.hpp:526-753) is mostly dense 4- and 5-qubit unitaries; applied one HBM pass
// per gate (UNITARY / FUSED_BLOCK of the reference's apply_unitary,
// include/backend/cpu_openmp_engine.hpp:539-606) the pass count equals the gate count.
// Here each work-group loads one chunk of 2^13 amplitudes (a 13-qubit window that holds
// the low 4 qubits, so every HBM access is a 128-byte run) into LDS, applies every gate
// of its sweep to the chunk in LDS, and stores it: one HBM pass per window instead of one
// per gate.
//   * 4/5-qubit gates run on the matrix cores with the same arithmetic as k_dense_mfma
//     (real-embedded GEMM per wave over 32 groups; BF16 split operands accumulated in
//     FP32, or FP32 operands), the amplitudes read from and written back to LDS;
//   * every other gate (1-3 qubits) is a dense complex matrix in FP32 applied by the
//     work-items, one group of 2^k amplitudes at a time.
// With dense_kernel = QLX_DENSE_VALU the program runs as QLX_EXEC_WINDOWED instead.
// Gates of one sweep run in program order with a barrier between them; within a gate the
// groups are disjoint, and a wave (MFMA) or work-item (vector) reads its groups
// completely before writing them.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <complex>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "qlx_internal.hpp"
#include "qlx_mfma_core.hpp"
#include "qlx_sweep_core.hpp"

namespace qlx {
namespace {

constexpr int kDW = 13;              // window qubits
constexpr int kDC = 4;               // low qubits in every window (128-byte runs of float2)
constexpr int kDThreads = 256;
constexpr int kDChunk = 1 << kDW;


constexpr int kMaxChunkBits = 32;    // chunk-index bits of one launch (2^32 work-groups at most)

// A sweep's addressing: local bit l of the chunk (local bit l <-> window qubit S[l]; the
// first kDC are the qubits on address bits 0..kDC-1 at the load) and chunk-index bit j (the
// qubits outside the window, in the order of their load address bits) sit at address bit
// ld_l[l] / ld_c[j] when the chunk is read. The store walks the chunk in its own order: bit j
// of the store index is local bit st_loc[j] (the first kDC: the qubits on address bits
// 0..kDC-1 at the store) and sits at address bit st_l[j]; chunk-index bit j at st_c[j]. In place
// both layouts are the identity; with relabelling (out of place, between two buffers) the store
// layout puts the next window on the low address bits.
struct DenseSweepDesc {
    uint32_t op_begin, n_ops;
    uint32_t nc;                    // chunk-index bits (state qubits - kDW)
    uint32_t pad;
    uint32_t ld_l[kDW], st_l[kDW];
    uint32_t ld_c[kMaxChunkBits], st_c[kMaxChunkBits];
    uint32_t st_loc[kDW];
    uint32_t pad2[1];
};
static_assert(sizeof(DenseSweepDesc) % 16 == 0, "descriptor layout");
static_assert(offsetof(DenseSweepDesc, st_c) == offsetof(DenseSweepDesc, ld_c) + 4 * kMaxChunkBits,
              "st_c follows ld_c (read as one 64-entry array)");

enum : uint32_t { DOP_VALU = 0, DOP_MFMA = 1 };

struct DenseOpDesc {
    uint32_t kind;          // DOP_VALU | DOP_MFMA
    uint32_t k;             // gate qubits (1..5)
    uint32_t mat;           // DOP_VALU: first complex element of the FP32 matrix
    uint32_t pad0;
    uint64_t frag_off;      // DOP_MFMA: byte offset of the A fragments
    uint32_t bits[5];       // local bit of gate qubit i (matrix index bit i)
    uint32_t sorted[5];     // the same, ascending (group base by zero-bit insertion)
};

// LDS position of local index x: bits 5-9 folded by XOR into bits 0-4 (a bijection on the
// chunk). Groups of a gate on low bits sit 2^5 or more apart, and unswizzled all 32 columns of
// a wave half would hit one bank; swz is linear over GF(2), so swz(base | off) = swz(base) ^
// swz(off) for disjoint bit sets and every offset is swizzled once per gate.
__device__ __forceinline__ uint32_t swz(uint32_t x) { return x ^ ((x >> 5) & 31u); }

// The amplitude at byte offset `byte` of the chunk (offsets kept in bytes, so a swizzled
// address costs one XOR and no shift).
__device__ __forceinline__ float2& at(float2* lds, uint32_t byte) {
    return *reinterpret_cast<float2*>(reinterpret_cast<char*>(lds) + byte);
}

__device__ __forceinline__ uint32_t ins0_32(uint32_t idx, uint32_t t) {
    const uint32_t lo = idx & ((1u << t) - 1);
    return lo | ((idx ^ lo) << 1);
}

// One 4/5-qubit gate on the chunk in LDS, on the matrix cores. Wave w takes tiles w, w+4, ...
// of 32 groups; lane (h, j) holds group j's amplitudes [h 2^(K-1), (h+1) 2^(K-1)).
template <int K, int FORM>
__device__ void mfma_gate(float2* __restrict__ lds, const DenseOpDesc& op, const unsigned char* __restrict__ frags) {
    constexpr int D = 1 << K, HALF = D / 2, MT = (2 * D) / 32;
    constexpr int TB = (2 * D) / 16, TF = D;
    constexpr int NB = MT * TB * 64;
    constexpr int NTILES = (kDChunk >> K) / 32;
    constexpr bool F32 = FORM == FORM_F32, BF16 = FORM == FORM_BF16, F16 = FORM == FORM_F16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 5, j = lane & 31;
    // swizzled byte offset of group amplitude l: the XOR of 8 swz(2^bits[i]) over the set bits i
    // of l, formed where it is used from five values (a table of 32 offsets kept the scalar
    // registers full)
    uint32_t pb[K];
#pragma unroll
    for (int i = 0; i < K; ++i) pb[i] = swz(1u << op.bits[i]) << 3;
    auto off = [&](int l) {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < K; ++i)
            if ((l >> i) & 1) o ^= pb[i];
        return o;
    };
    const uint32_t hin = h ? pb[K - 1] : 0u, hout = h ? pb[1] : 0u;
    const unsigned char* fr = frags + op.frag_off;
    // the gate's A fragments, once per wave (the wave then takes 2 (K = 5) or 4 (K = 4) tiles)
    constexpr int RB = BF16 ? MT : 1, TBB = BF16 ? TB : 1;
    constexpr int RH = F16 ? MT : 1, TBH = F16 ? TB : 1;
    bf16x8 ah[RB][TBB], am[RB][TBB], al[RB][TBB];
    f16x8 hh[RH][TBH], hl[RH][TBH];
    float af[F32 ? MT : 1][F32 ? TF : 1];
    if constexpr (F32) {
        const float* f = (const float*)fr;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < TF; ++t) af[m][t] = f[(m * TF + t) * 64 + lane];
    } else if constexpr (BF16) {
        const bf16x8* fh = (const bf16x8*)fr;
        const bf16x8* fm = fh + NB;
        const bf16x8* fl = fm + NB;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                ah[m][t] = fh[(m * TB + t) * 64 + lane];
                am[m][t] = fm[(m * TB + t) * 64 + lane];
                al[m][t] = fl[(m * TB + t) * 64 + lane];
            }
    } else {
        const f16x8* fh = (const f16x8*)fr;
        const f16x8* fl = fh + NB;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                hh[m][t] = fh[(m * TB + t) * 64 + lane];
                hl[m][t] = fl[(m * TB + t) * 64 + lane];
            }
    }
    // The wave reads all of its tiles before it writes any (its groups are its own), so a tile's
    // reads and operand split overlap the previous tile's products.
    constexpr int TPW = NTILES / (kDThreads / 64);
    auto tile_base = [&](int tile) {
        uint32_t b = (uint32_t)tile * 32 + j;
#pragma unroll
        for (int q = 0; q < K; ++q) b = ins0_32(b, op.sorted[q]);
        return swz(b) << 3;
    };
    float2 xa[TPW][HALF];
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        const uint32_t b = tile_base(wave + i * (kDThreads / 64));
#pragma unroll
        for (int r = 0; r < HALF; ++r) xa[i][r] = at(lds, b ^ hin ^ off(r));
    }
#pragma unroll
    for (int ti = 0; ti < TPW; ++ti) {
        const uint32_t base = tile_base(wave + ti * (kDThreads / 64));
        const float2* x = xa[ti];
        f32x16 acc[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[m][v] = 0.f;
        if constexpr (F32) {
#pragma unroll
            for (int t = 0; t < TF; ++t) {
                const float xv = (t & 1) ? x[t >> 1].y : x[t >> 1].x;
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[m][t], xv, acc[m], 0, 0, 0);
            }
        } else if constexpr (BF16) {
            bf16x8 xh[TB], xl[TB];
#pragma unroll
            for (int t = 0; t < TB; ++t) split_bf16(x + 4 * t, xh[t], xl[t]);
            // small terms first (as k_dense_mfma)
#pragma unroll
            for (int m = 0; m < MT; ++m) {
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[m][t], xh[t], acc[m], 0, 0, 0);
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[m][t], xh[t], acc[m], 0, 0, 0);
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m][t], xl[t], acc[m], 0, 0, 0);
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m][t], xh[t], acc[m], 0, 0, 0);
            }
        } else {
            // the chunk is held scaled by 2^14 in LDS (applied at the load, removed at the store)
            f16x8 xh[TB], xl[TB];
#pragma unroll
            for (int t = 0; t < TB; ++t) split_f16(x + 4 * t, 1.0f, xh[t], xl[t]);
#pragma unroll
            for (int m = 0; m < MT; ++m) {
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hl[m][t], xh[t], acc[m], 0, 0, 0);
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hh[m][t], xl[t], acc[m], 0, 0, 0);
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hh[m][t], xh[t], acc[m], 0, 0, 0);
            }
        }
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    at(lds, base ^ hout ^ off(16 * m + 4 * a + b)) =
                        float2{acc[m][4 * a + 2 * b], acc[m][4 * a + 2 * b + 1]};
    }
}

// A k-qubit gate on the chunk in LDS on the vector units: work-item t takes groups t, t+256, ...
template <int K>
__device__ void valu_gate(float2* __restrict__ lds, const DenseOpDesc& op, const float2* __restrict__ mats) {
    constexpr int D = 1 << K;
    constexpr int NG = kDChunk >> K;
    const float2* M = mats + op.mat;
    uint32_t off[D];
#pragma unroll
    for (int l = 0; l < D; ++l) {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < K; ++i)
            if ((l >> i) & 1) o |= 1u << op.bits[i];
        off[l] = swz(o);
    }
    for (int g = threadIdx.x; g < NG; g += kDThreads) {
        uint32_t base = (uint32_t)g;
#pragma unroll
        for (int q = 0; q < K; ++q) base = ins0_32(base, op.sorted[q]);
        base = swz(base);
        float2 v[D];
#pragma unroll
        for (int l = 0; l < D; ++l) v[l] = lds[base ^ off[l]];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const float2 m = M[r * D + c];
                re = fmaf(m.x, v[c].x, fmaf(-m.y, v[c].y, re));
                im = fmaf(m.x, v[c].y, fmaf(m.y, v[c].x, im));
            }
            lds[base ^ off[r]] = float2{re, im};
        }
    }
}

// 3-qubit gates (CCX, CSWAP, 3-qubit matrices; rare in fused programs): the 64 matrix
// entries are staged in LDS and read row by row, which keeps them out of the registers of
// the main loop (uniform entries held in scalar registers spilled the others).
__device__ void valu_gate3(float2* __restrict__ lds, const DenseOpDesc& op, const float2* __restrict__ mats,
                           float2* __restrict__ mbuf) {
    constexpr int K = 3, D = 8, NG = kDChunk >> K;
    if (threadIdx.x < D * D) mbuf[threadIdx.x] = mats[op.mat + threadIdx.x];
    __syncthreads();
    uint32_t off[D];
#pragma unroll
    for (int l = 0; l < D; ++l) {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < K; ++i)
            if ((l >> i) & 1) o |= 1u << op.bits[i];
        off[l] = swz(o);
    }
#pragma unroll 1
    for (int g = threadIdx.x; g < NG; g += kDThreads) {
        uint32_t base = (uint32_t)g;
#pragma unroll
        for (int q = 0; q < K; ++q) base = ins0_32(base, op.sorted[q]);
        base = swz(base);
        float2 v[D];
#pragma unroll
        for (int l = 0; l < D; ++l) v[l] = lds[base ^ off[l]];
        // the group's amplitudes are all in registers: each row is written as it is formed
#pragma unroll 1
        for (int r = 0; r < D; ++r) {
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const float2 m = mbuf[r * D + c];
                re = fmaf(m.x, v[c].x, fmaf(-m.y, v[c].y, re));
                im = fmaf(m.x, v[c].y, fmaf(m.y, v[c].x, im));
            }
            const uint32_t o = ((r & 1) ? off[1] : 0u) ^ ((r & 2) ? off[2] : 0u) ^ ((r & 4) ? off[4] : 0u);
            lds[base ^ o] = float2{re, im};
        }
    }
}

// One chunk per work-group: load into LDS, every gate of the sweep, store. In the FP16 form
// the chunk is held in LDS scaled by 2^14 (exact; every gate is linear). `src` and `dst` are
// the same buffer in place; a work-group reads all of its chunk before writing any of it.
template <int FORM>
__global__ void __launch_bounds__(kDThreads) __attribute__((amdgpu_waves_per_eu(2))) k_dsweep(const float2* src, float2* dst,
                                                     const DenseSweepDesc* __restrict__ sw,
                                                     const DenseOpDesc* __restrict__ ops,
                                                     const float2* __restrict__ mats,
                                                     const unsigned char* __restrict__ frags) {
    __shared__ float2 lds[kDChunk];
    __shared__ float2 mbuf[64];  // the current 3-qubit gate's matrix
    const unsigned t = threadIdx.x;
    // the chunk's base address at the load and at the store: lane j < 32 of every wave places
    // bit j of the work-group index at its load address bit, lane 32 + j at its store address bit
    // (bits at or above nc are zero), and each half of the wave ORs its values together
    uint64_t base, sbase;
    {
        const uint32_t lane = t & 63u;
        const uint64_t bit = (blockIdx.x >> (lane & 31u)) & 1u;
        const uint32_t* cpos = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(sw) +
                                                                 offsetof(DenseSweepDesc, ld_c));
        uint64_t v = bit << cpos[lane];   // cpos[32 + j] is st_c[j]
#pragma unroll
        for (int m = 1; m < 32; m <<= 1) v |= __shfl_xor(v, m);
        base = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), 0) << 32) |
               (uint32_t)__builtin_amdgcn_readlane((uint32_t)v, 0);
        sbase = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), 32) << 32) |
                (uint32_t)__builtin_amdgcn_readlane((uint32_t)v, 32);
    }
    // local index i = t + 256 r: bits 0-7 from the work-item, bits 8-12 from r
    uint64_t tl = 0, stl = 0;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        tl |= (uint64_t)((t >> l) & 1u) << sw->ld_l[l];
        stl |= (uint64_t)((t >> l) & 1u) << sw->st_l[l];
    }
    constexpr int R = kDChunk / kDThreads;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint64_t ro = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l)
            if ((r >> l) & 1) ro |= 1ull << sw->ld_l[8 + l];
        float2 v = src[base | tl | ro];
        if constexpr (FORM == FORM_F16) v = float2{v.x * kF16Scale, v.y * kF16Scale};
        lds[swz(t + kDThreads * r)] = v;
    }
    __syncthreads();
    const uint32_t n_ops = sw->n_ops;
    for (uint32_t o = 0; o < n_ops; ++o) {
        const DenseOpDesc& op = ops[sw->op_begin + o];
        if (op.kind == DOP_MFMA) {
            if (op.k == 5) mfma_gate<5, FORM>(lds, op, frags);
            else mfma_gate<4, FORM>(lds, op, frags);
        } else {
            switch (op.k) {
                case 1: valu_gate<1>(lds, op, mats); break;
                case 2: valu_gate<2>(lds, op, mats); break;
                default: valu_gate3(lds, op, mats, mbuf); break;
            }
        }
        __syncthreads();
    }
    // the store's register-bit positions are read here (volatile): read with the others, the 32
    // register offsets would hold scalar registers through every gate. Store index t + 256 r:
    // bit j is local bit st_loc[j] (the identity unless the store's coalesced qubits differ
    // from the load's); swz is linear over GF(2), so its work-item and register parts are
    // swizzled apart
    const volatile uint32_t* vd = reinterpret_cast<const volatile uint32_t*>(sw);
    constexpr int kStL = offsetof(DenseSweepDesc, st_l) / 4;
    constexpr int kStLoc = offsetof(DenseSweepDesc, st_loc) / 4;
    uint32_t sp[5], sl[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
        sp[l] = vd[kStL + 8 + l];
        sl[l] = vd[kStLoc + 8 + l];
    }
    uint32_t it = 0;
#pragma unroll
    for (int l = 0; l < 8; ++l) it |= ((t >> l) & 1u) << vd[kStLoc + l];
    const uint32_t st_t = swz(it);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint64_t ro = 0;
        uint32_t ir = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l)
            if ((r >> l) & 1) {
                ro |= 1ull << sp[l];
                ir |= 1u << sl[l];
            }
        float2 v = lds[st_t ^ swz(ir)];
        if constexpr (FORM == FORM_F16) v = float2{v.x * kF16Unscale, v.y * kF16Unscale};
        dst[sbase | stl | ro] = v;
    }
}

// Row-major complex128 matrix of a gate, index bit i <-> qubits[i] (the convention of the
// reference's get_matrix and apply_unitary, middleware.hpp:399-439, cpu_openmp_engine.hpp:
// 539-606). Standard gates from their closed forms with the same FP64 expressions as the
// other kernels; UNITARY / FUSED_BLOCK from the pool.
std::vector<double> gate_matrix(const qlx_gate_instr& g, const double* pool, const qlx_matrix_entry* meta, int& k) {
    auto perm = [&](int kk, const std::vector<int>& to) {
        k = kk;
        const int D = 1 << kk;
        std::vector<double> m(2 * D * D, 0.0);
        for (int c = 0; c < D; ++c) m[2 * (to[c] * D + c)] = 1.0;
        return m;
    };
    const double half = static_cast<double>(g.payload.param) * 0.5;
    switch (g.type) {
        case QLX_H: {
            k = 1;
            const double r = 1.0 / std::sqrt(2.0);
            return {r, 0, r, 0, r, 0, -r, 0};
        }
        case QLX_X: k = 1; return {0, 0, 1, 0, 1, 0, 0, 0};
        case QLX_Y: k = 1; return {0, 0, 0, -1, 0, 1, 0, 0};
        case QLX_Z: k = 1; return {1, 0, 0, 0, 0, 0, -1, 0};
        case QLX_S: k = 1; return {1, 0, 0, 0, 0, 0, 0, 1};
        case QLX_RX: {
            k = 1;
            const double c = std::cos(half), s = std::sin(half);
            return {c, 0, 0, -s, 0, -s, c, 0};
        }
        case QLX_RY: {
            k = 1;
            const double c = std::cos(half), s = std::sin(half);
            return {c, 0, -s, 0, s, 0, c, 0};
        }
        case QLX_RZ:
            k = 1;
            return {std::cos(-half), std::sin(-half), 0, 0, 0, 0, std::cos(half), std::sin(half)};
        case QLX_CX: return perm(2, {0, 3, 2, 1});                    // control bit 0, target bit 1
        case QLX_SWAP: case QLX_GLOBAL_SWAP: return perm(2, {0, 2, 1, 3});
        case QLX_CCX: return perm(3, {0, 1, 2, 7, 4, 5, 6, 3});       // controls bits 0, 1; target bit 2
        case QLX_CSWAP: return perm(3, {0, 1, 2, 5, 4, 3, 6, 7});     // control bit 0; swap bits 1, 2
        case QLX_UNITARY:
        case QLX_FUSED_BLOCK: {
            k = instr_arity(g);
            const qlx_matrix_entry& e = meta[g.payload.matrix_idx];
            const double* m = pool + 2 * (size_t)e.offset;
            return std::vector<double>(m, m + 2 * (size_t)e.num_elements);
        }
        default:
            fail(QLX_EUNSUPPORTED, "gate cannot run in a dense sweep");
    }
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ---- merging the gates of one sweep ------------------------------------------------------
// Every application costs a pass over the chunk in LDS whatever its width, so the gates of a
// sweep are merged (FP64 products, on the host) into as few blocks of at most five qubits as a
// greedy rule finds: a gate joins the latest block whose qubits together with its own number at
// most five, passing over later blocks that share none of its qubits (with which it commutes).
constexpr int kMaxBlock = 5;

struct Block {
    std::vector<int> q;                  // window bits; matrix index bit i <-> q[i]
    std::vector<std::complex<double>> m;  // row-major 2^|q| x 2^|q|
};

// The matrix m (on bits gq) as an operator on the ordered bit list u, a superset of gq.
std::vector<std::complex<double>> embed_on(const std::vector<int>& gq, const std::vector<std::complex<double>>& m,
                                           const std::vector<int>& u) {
    const int kg = (int)gq.size(), ku = (int)u.size();
    const int DG = 1 << kg, DU = 1 << ku;
    int pos[kMaxBlock];
    uint32_t gmask = 0;
    for (int i = 0; i < kg; ++i) {
        pos[i] = (int)(std::find(u.begin(), u.end(), gq[i]) - u.begin());
        if (pos[i] >= ku) fail(QLX_EHIP, "dense sweep block merge: qubit missing from the union");
        gmask |= 1u << pos[i];
    }
    auto gather = [&](uint32_t x) {
        uint32_t r = 0;
        for (int i = 0; i < kg; ++i) r |= ((x >> pos[i]) & 1u) << i;
        return r;
    };
    std::vector<std::complex<double>> e((size_t)DU * DU);
    for (uint32_t r = 0; r < (uint32_t)DU; ++r)
        for (uint32_t c = 0; c < (uint32_t)DU; ++c)
            if ((r & ~gmask) == (c & ~gmask)) e[(size_t)r * DU + c] = m[(size_t)gather(r) * DG + gather(c)];
    return e;
}

// b <- g b on the union of their bits (g applied after b).
void absorb(Block& b, const Block& g) {
    std::vector<int> u = b.q;
    for (int x : g.q)
        if (std::find(u.begin(), u.end(), x) == u.end()) u.push_back(x);
    const std::vector<std::complex<double>> B = embed_on(b.q, b.m, u), G = embed_on(g.q, g.m, u);
    const size_t D = size_t{1} << u.size();
    std::vector<std::complex<double>> C(D * D);
    for (size_t r = 0; r < D; ++r)
        for (size_t k = 0; k < D; ++k) {
            const std::complex<double> a = G[r * D + k];
            if (a == 0.0) continue;
            for (size_t c = 0; c < D; ++c) C[r * D + c] += a * B[k * D + c];
        }
    b.q = std::move(u);
    b.m = std::move(C);
}

std::vector<Block> merge_blocks(std::vector<Block> gates, int max_q) {
    std::vector<Block> out;
    for (Block& g : gates) {
        bool merged = false;
        for (size_t j = out.size(); j-- > 0;) {
            Block& b = out[j];
            int shared = 0;
            for (int x : g.q) shared += std::find(b.q.begin(), b.q.end(), x) != b.q.end();
            if ((int)(b.q.size() + g.q.size()) - shared <= max_q) {
                absorb(b, g);
                merged = true;
                break;
            }
            if (shared) break;
        }
        if (!merged) out.push_back(std::move(g));
    }
    return out;
}

struct DensePlan {
    uint64_t key = 0;
    std::vector<DenseSweepDesc> sweeps;
    std::vector<DenseOpDesc> ops;
    std::vector<float> mats;     // interleaved complex FP32
    size_t gates = 0;            // program gates in the sweeps
    size_t blocks = 0;           // applications after merging
    DevBuf dev;                  // sweeps | ops | mats
    DevBuf fpool, frags, fdesc;  // FP64 matrices of the matrix-core blocks, their A fragments
    size_t off_ops = 0, off_mats = 0;
};

}  // namespace

void execute_dense_sweeps(qlx_state* st, const qlx_gate_instr* instrs, size_t n, const double* pool,
                          size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts& opts,
                          qlx_exec_stats& stats) {
    if (st->dtype != QLX_C64) fail(QLX_EUNSUPPORTED, "dense sweeps run on complex64 states");
    if (st->nq < (uint32_t)kDW) {
        execute_plain(st, instrs, n, pool, pool_elems, meta, n_meta, opts, stats);
        return;
    }
    if (opts.dense_kernel == QLX_DENSE_VALU) {
        qlx_exec_opts o = opts;
        o.mode = QLX_EXEC_WINDOWED;
        execute_windowed(st, instrs, n, pool, pool_elems, meta, n_meta, o, stats);
        return;
    }
    const uint32_t dk = opts.dense_kernel;
    if (dk > QLX_DENSE_MFMA_F16) fail(QLX_EINVAL, "unknown dense_kernel");
    const int variant = dk == QLX_DENSE_AUTO ? QLX_DENSE_MFMA_F16 : (int)dk;
    const bool f32 = variant == QLX_DENSE_MFMA_F32;
    bool relabel = (opts.specialize & QLX_SPECIALIZE_RELABEL) != 0;
    // relabelled sweeps need a second buffer of the state's size; without it they run in place
    if (relabel && !st->alt && hipMalloc(&st->alt, st->bytes) != hipSuccess) {
        (void)hipGetLastError();
        st->alt = nullptr;
        relabel = false;
    }
    const uint32_t eff_specialize = relabel ? QLX_SPECIALIZE_RELABEL : 0u;

    uint64_t key = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t b) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < b; ++i) {
            key ^= c[i];
            key *= 1099511628211ull;
        }
    };
    mix(instrs, n * sizeof(qlx_gate_instr));
    if (pool) mix(pool, pool_elems * 16);
    if (meta) mix(meta, n_meta * sizeof(qlx_matrix_entry));
    mix(&opts, sizeof(opts));
    mix(&eff_specialize, sizeof(eff_specialize));
    mix(&st->nq, sizeof(st->nq));
    if (!st->dense_cache) st->dense_cache = std::shared_ptr<void>(new DensePlan(), [](void* p) {
        auto* d = static_cast<DensePlan*>(p);
        d->dev.release();
        d->fpool.release();
        d->frags.release();
        d->fdesc.release();
        delete d;
    });
    DensePlan& P = *static_cast<DensePlan*>(st->dense_cache.get());
    if (P.key != key || P.sweeps.empty()) {
        P.sweeps.clear();
        P.ops.clear();
        P.mats.clear();
        P.gates = 0;
        P.blocks = 0;
        std::vector<double> fpool;             // FP64 matrices of the matrix-core blocks
        std::vector<MfmaSplitJob> jobs;
        size_t frag_total = 0;
        const int nq = (int)st->nq;
        const std::vector<DenseWindow> wins = plan_dense_windows(st->nq, instrs, n, opts, kDW, kDC);
        // pos[q]: address bit of logical qubit q before the current sweep (identity at the start
        // and, after the last sweep, at the end)
        std::vector<int> pos(nq);
        for (int q = 0; q < nq; ++q) pos[q] = q;
        for (size_t si = 0; si < wins.size(); ++si) {
            const DenseWindow& w = wins[si];
            if (w.S.size() != (size_t)kDW || w.st_low.size() != (size_t)kDC)
                fail(QLX_EHIP, "dense sweep window of the wrong size");
            if (!relabel)
                for (int l = 0; l < kDC; ++l)
                    if (w.S[l] != l || w.st_low[l] != l) fail(QLX_EHIP, "dense sweep window without the low qubits");
            std::vector<int> next(nq);
            for (int q = 0; q < nq; ++q) next[q] = q;
            if (relabel && si + 1 < wins.size()) {
                // the store's coalesced qubits on address bits 0..kDC-1; then the next window,
                // this window, the rest in load order
                std::vector<char> placed(nq, 0);
                int at = 0;
                auto place = [&](int q) {
                    if (!placed[q]) {
                        placed[q] = 1;
                        next[q] = at++;
                    }
                };
                for (int q : w.st_low) place(q);
                for (int q : wins[si + 1].S) place(q);
                for (int q : w.S) place(q);
                std::vector<int> rest;
                for (int q = 0; q < nq; ++q)
                    if (!placed[q]) rest.push_back(q);
                std::sort(rest.begin(), rest.end(), [&](int a, int b) { return pos[a] < pos[b]; });
                for (int q : rest) place(q);
            }
            DenseSweepDesc d{};
            std::vector<char> inw(nq, 0);
            for (int q : w.S) inw[q] = 1;
            std::vector<int> outside;
            for (int q = 0; q < nq; ++q)
                if (!inw[q]) outside.push_back(q);
            std::sort(outside.begin(), outside.end(), [&](int a, int b) { return pos[a] < pos[b]; });
            if (outside.size() > (size_t)kMaxChunkBits) fail(QLX_EUNSUPPORTED, "state too large for dense sweeps");
            d.nc = (uint32_t)outside.size();
            for (size_t j = 0; j < outside.size(); ++j) {
                d.ld_c[j] = (uint32_t)pos[outside[j]];
                d.st_c[j] = (uint32_t)next[outside[j]];
            }
            // store order: the store's coalesced local bits first, then the others ascending
            std::vector<int> sorder;
            for (int q : w.st_low)
                sorder.push_back((int)(std::find(w.S.begin(), w.S.end(), q) - w.S.begin()));
            for (int l = 0; l < kDW; ++l)
                if (std::find(w.st_low.begin(), w.st_low.end(), w.S[l]) == w.st_low.end()) sorder.push_back(l);
            if (sorder.size() != (size_t)kDW) fail(QLX_EHIP, "dense sweep store qubits outside the window");
            for (int l = 0; l < kDW; ++l) {
                d.ld_l[l] = (uint32_t)pos[w.S[l]];
                d.st_loc[l] = (uint32_t)sorder[l];
                d.st_l[l] = (uint32_t)next[w.S[sorder[l]]];
            }
            // coalesced: local bits 0..kDC-1 at the load and store-index bits 0..kDC-1 at the
            // store sit on address bits 0..kDC-1
            for (int l = 0; l < kDC; ++l)
                if (d.ld_l[l] >= (uint32_t)kDC || d.st_l[l] >= (uint32_t)kDC)
                    fail(QLX_EHIP, "dense sweep layout leaves a load or store uncoalesced");
            pos = next;
            d.op_begin = (uint32_t)P.ops.size();
            std::vector<int> loc(st->nq, -1);
            for (int l = 0; l < kDW; ++l) loc[w.S[l]] = l;
            std::vector<Block> gates;
            for (int src : w.instrs) {
                const qlx_gate_instr& g = instrs[src];
                int k = 0;
                const std::vector<double> m = gate_matrix(g, pool, meta, k);
                Block b;
                for (int i = 0; i < k; ++i) {
                    const int q = (int)instr_qubit(g, i);
                    if (loc[q] < 0) fail(QLX_EHIP, "dense sweep gate outside its window");
                    b.q.push_back(loc[q]);
                }
                b.m.resize(m.size() / 2);
                for (size_t e = 0; e < b.m.size(); ++e) b.m[e] = {m[2 * e], m[2 * e + 1]};
                gates.push_back(std::move(b));
                ++P.gates;
            }
            for (Block& b : merge_blocks(std::move(gates), kMaxBlock)) {
                // blocks of 1-3 qubits run on the matrix cores as 4-qubit blocks (identity on
                // the highest free window bits) except in FP32, where the vector form costs
                // no more than an FP32 4-qubit product
                if (b.q.size() < 4 && !f32) {
                    std::vector<int> u = b.q;
                    for (int l = kDW - 1; l >= 0 && u.size() < 4; --l)
                        if (std::find(u.begin(), u.end(), l) == u.end()) u.push_back(l);
                    b.m = embed_on(b.q, b.m, u);
                    b.q = u;
                }
                const int k = (int)b.q.size();
                DenseOpDesc o{};
                o.k = (uint32_t)k;
                for (int i = 0; i < k; ++i) o.bits[i] = o.sorted[i] = (uint32_t)b.q[i];
                std::sort(o.sorted, o.sorted + k);
                if (k >= 4) {
                    o.kind = DOP_MFMA;
                    o.frag_off = frag_total;
                    jobs.push_back(MfmaSplitJob{fpool.size() / 2, frag_total, k, variant});
                    frag_total += (mfma_fragment_bytes(k, variant) + 255) & ~size_t{255};
                    for (const std::complex<double>& v : b.m) {
                        fpool.push_back(v.real());
                        fpool.push_back(v.imag());
                    }
                } else {
                    o.kind = DOP_VALU;
                    o.mat = (uint32_t)(P.mats.size() / 2);
                    for (const std::complex<double>& v : b.m) {
                        P.mats.push_back((float)v.real());
                        P.mats.push_back((float)v.imag());
                    }
                }
                P.ops.push_back(o);
                ++P.blocks;
            }
            d.n_ops = (uint32_t)(P.ops.size() - d.op_begin);
            P.sweeps.push_back(d);
        }
        const size_t b_sw = P.sweeps.size() * sizeof(DenseSweepDesc);
        P.off_ops = align_up(b_sw, 256);
        P.off_mats = align_up(P.off_ops + P.ops.size() * sizeof(DenseOpDesc), 256);
        const size_t total = align_up(P.off_mats + P.mats.size() * sizeof(float), 256) + 256;
        std::vector<char> img(total, 0);
        std::memcpy(img.data(), P.sweeps.data(), b_sw);
        if (!P.ops.empty()) std::memcpy(img.data() + P.off_ops, P.ops.data(), P.ops.size() * sizeof(DenseOpDesc));
        if (!P.mats.empty()) std::memcpy(img.data() + P.off_mats, P.mats.data(), P.mats.size() * sizeof(float));
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));  // the old image may still be in use
        P.dev.reserve(total, st->stream);
        QLX_HIP_CHECK(hipMemcpyAsync(P.dev.ptr, img.data(), total, hipMemcpyHostToDevice, st->stream));
        if (!jobs.empty()) {
            P.fpool.reserve(fpool.size() * sizeof(double), st->stream);
            QLX_HIP_CHECK(hipMemcpyAsync(P.fpool.ptr, fpool.data(), fpool.size() * sizeof(double),
                                         hipMemcpyHostToDevice, st->stream));
            P.frags.reserve(frag_total, st->stream);
            mfma_split((const double*)P.fpool.ptr, jobs.data(), jobs.size(), P.frags.ptr, P.fdesc, st->stream);
        }
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
        P.key = key;
    }
    const char* dev = (const char*)P.dev.ptr;
    const uint64_t chunks = 1ull << (st->nq - kDW);
    if (chunks > 0xffffffffull) fail(QLX_EUNSUPPORTED, "state too large for one dense-sweep launch");
    const double pass = pass_bytes(st);
    const unsigned char* frags = (const unsigned char*)P.frags.ptr;
    if (relabel && !st->alt) fail(QLX_EHIP, "relabelled dense sweeps without a second buffer");
    for (size_t s = 0; s < P.sweeps.size(); ++s) {
        TimedLaunch tl(st, "k_dsweep", pass);
        const DenseSweepDesc* sw = (const DenseSweepDesc*)dev + s;
        const DenseOpDesc* ops = (const DenseOpDesc*)(dev + P.off_ops);
        const float2* mats = (const float2*)(dev + P.off_mats);
        const float2* src = (const float2*)st->amp;
        float2* dst = (float2*)(relabel ? st->alt : st->amp);
        if (variant == QLX_DENSE_MFMA_F32)
            k_dsweep<FORM_F32><<<(unsigned)chunks, kDThreads, 0, st->stream>>>(src, dst, sw, ops, mats, frags);
        else if (variant == QLX_DENSE_MFMA_F16)
            k_dsweep<FORM_F16><<<(unsigned)chunks, kDThreads, 0, st->stream>>>(src, dst, sw, ops, mats, frags);
        else
            k_dsweep<FORM_BF16><<<(unsigned)chunks, kDThreads, 0, st->stream>>>(src, dst, sw, ops, mats, frags);
        QLX_HIP_CHECK(hipGetLastError());
        if (relabel) std::swap(st->amp, st->alt);
        stats.sweeps_full++;
        stats.algorithmic_bytes += pass;
    }
    stats.gates_applied += P.gates;
    stats.relabelled = relabel && !P.sweeps.empty() ? 1u : 0u;
}

}  // namespace qlx
