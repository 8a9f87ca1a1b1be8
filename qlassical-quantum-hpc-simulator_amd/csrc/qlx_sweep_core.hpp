// qlx_sweep_core.hpp — device side of the window sweep (shared by the windowed executor
// and the batched persistent kernel) and the host planner's interface.
#pragma once

#include <memory>
#include <type_traits>
#include <vector>

#include "qlx_internal.hpp"
#include "qlx_mfma_core.hpp"

namespace qlx {

constexpr int kMaxW = 14;
constexpr int kMaxLogT = 10;
constexpr int kMaxR = 32;

// For the circuit-specialised sweeps the planner rewrites OP_RX and OP_H into factored forms
// (the generic kernel keeps OP_RX / OP_H: its cost is the runtime op dispatch, and the extra
// cases measured slower there). Every gate is linear and H and RX scale every amplitude of the
// chunk by the same real factor, so the factor is taken out and applied once:
//   H  = h [[1, 1], [1, -1]]                  -> OP_HB  butterfly (adds only), factor h
//   RX = c [[1, -i t], [-i t, 1]], t = s / c  -> OP_RXT two FMAs per amplitude, factor c
//      = s [[u, -i], [-i, u]],     u = c / s  -> OP_RXU (when |s| > |c|), factor s
// (c = cos(theta/2), s = sin(theta/2)); OP_SCALE multiplies every register by the product of
// the pending factors, at the end of each sweep or as soon as the product falls below 2^-30
// (each factor is at least 1/sqrt 2 in magnitude, so registers stay within 2^30 of the
// amplitudes' magnitude, also in FP32).
//
// Circuit-specialised sweeps also take OP_CXT: a CX whose control is work-item bit b0 of the
// phase's layout (a window qubit outside the registers) and whose target is register bit b1.
// The generator applies no instruction for it: it records a pending X on the target register
// bit, conditioned per work-item on the parity of selected work-item bits (the control), and
// carries it through later ops (a CX passes it on from its control to its target, a SWAP moves
// it, RX commutes with it). The pending X is a per-work-item permutation of the registers, so
// the next LDS re-layout or the final store takes it up in its addresses; only an op that does
// not commute with it (H, dense matrices) applies it first, as conditional register swaps.
enum OpKind : uint8_t {
    OP_U2 = 1, OP_RX = 2, OP_H = 3, OP_CX = 4, OP_SWAP = 5, OP_U4 = 6,
    OP_RXT = 7, OP_RXU = 8, OP_HB = 9, OP_SCALE = 10, OP_CXT = 11
};

// 32-bit fields only: the descriptor is read with scalar loads (a 16-bit field would be
// fetched with a vector load, and the wait for it would also wait for every HBM load
// in flight).
struct DevOp {
    uint32_t key;   // (kind << 8) | (register bit 0 << 4) | register bit 1
    uint32_t coef;  // index into the coefficient array (doubles): U2/U4 matrices
    double c0, c1;  // inline coefficients: RX (cos, sin), H (1/sqrt 2); OP_RXT t, OP_RXU u,
                    // OP_SCALE the factor
};
static_assert(sizeof(DevOp) == 24, "DevOp layout");

// Descriptors use 32/64-bit fields throughout so that kernels read them with scalar loads.
struct DevPhase {
    uint32_t n_ops, op_begin, remap;
    // complex64 phases of 5 register bits (256 work-items): 0, or 1 + the index into the
    // coefficient array of the phase's fused 32 x 32 matrix as FP16 matrix-core fragments (the
    // phase then runs as one dense product instead of its ops)
    uint32_t mfma;
    uint32_t roff[kMaxR];        // local-index contribution of register index i
    uint32_t tlocal[kMaxLogT];   // local bit held by work-item bit b
    uint32_t pad2[2];
    // LDS byte offset of register index i's part of a local index: swz(roff[i]) x the element
    // size (the swizzle is linear over XOR, so a re-layout address is the work-item part's
    // offset XOR this; precomputed on the host instead of per register and phase in the kernel)
    uint32_t rslot[kMaxR];
};

constexpr int kMaxOopChunkBits = 24;  // chunk-index bits of an out-of-place sweep (states of <= 24 + W qubits)

struct DevSweep {
    uint32_t n_phases, phase_begin;
    uint32_t S[kMaxW];           // global qubit of local bit j (ascending, except in free-window
                                 // plans, which run out of place only)
    uint32_t ld_tglob[kMaxLogT]; // global qubit at work-item bit b in the first layout
    uint32_t st_tglob[kMaxLogT]; // ... in the last layout
    uint64_t ld_roff[kMaxR];     // global offset of register i in the first layout
    uint64_t st_roff[kMaxR];     // ... in the last layout
    // out of place (relabelled plans): address bit of chunk-index bit j at the load and at the store
    uint32_t nc, pad[3];
    uint32_t ld_cpos[kMaxOopChunkBits], st_cpos[kMaxOopChunkBits];
};
static_assert(sizeof(DevPhase) % 16 == 0 && sizeof(DevSweep) % 16 == 0, "descriptor alignment");

// ---------------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------------
template <class A>
struct Real;
template <>
struct Real<double2> { using T = double; };
template <>
struct Real<float2> { using T = float; };

__device__ __forceinline__ uint64_t dins0(uint64_t idx, uint32_t t) {
    const uint64_t lo = idx & ((1ull << t) - 1);
    return lo | ((idx ^ lo) << 1);
}

// LDS bank swizzle: XOR the low bits of the element index with a fold of its high
// bits (a bijection on [0, 2^W)), so work-items whose local indices differ only in
// high bits still land in different banks.
template <int ELEM>
__device__ __forceinline__ uint32_t swz(uint32_t l) {
    constexpr int SH = ELEM == 16 ? 4 : 5;
    constexpr uint32_t M = (1u << SH) - 1;
    return l ^ (((l >> SH) ^ (l >> (2 * SH)) ^ (l >> (3 * SH))) & M);
}

// swz for a runtime element size (host side of the descriptors)
inline uint32_t swz_bytes(uint32_t l, int elem) {
    const int sh = elem == 16 ? 4 : 5;
    const uint32_t m = (1u << sh) - 1;
    return (l ^ (((l >> sh) ^ (l >> (2 * sh)) ^ (l >> (3 * sh))) & m)) * (uint32_t)elem;
}

template <int B, class A, int R>
__device__ __forceinline__ void g_rx(A (&v)[R], typename Real<A>::T c, typename Real<A>::T s) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (i & (1 << B)) continue;
        const A a0 = v[i], a1 = v[i | (1 << B)];
        v[i] = A{c * a0.x + s * a1.y, c * a0.y + (-s * a1.x)};
        v[i | (1 << B)] = A{s * a0.y + c * a1.x, -s * a0.x + c * a1.y};
    }
}

template <int B, class A, int R>
__device__ __forceinline__ void g_h(A (&v)[R], typename Real<A>::T r) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (i & (1 << B)) continue;
        const A a0 = v[i], a1 = v[i | (1 << B)];
        v[i] = A{r * (a0.x + a1.x), r * (a0.y + a1.y)};
        v[i | (1 << B)] = A{r * (a0.x - a1.x), r * (a0.y - a1.y)};
    }
}

template <int B, class A, int R>
__device__ __forceinline__ void g_u2(A (&v)[R], const double* __restrict__ m) {
    using T = typename Real<A>::T;
    const T m0r = (T)m[0], m0i = (T)m[1], m1r = (T)m[2], m1i = (T)m[3];
    const T m2r = (T)m[4], m2i = (T)m[5], m3r = (T)m[6], m3i = (T)m[7];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (i & (1 << B)) continue;
        const A a0 = v[i], a1 = v[i | (1 << B)];
        v[i] = A{m0r * a0.x - m0i * a0.y + m1r * a1.x - m1i * a1.y, m0r * a0.y + m0i * a0.x + m1r * a1.y + m1i * a1.x};
        v[i | (1 << B)] =
            A{m2r * a0.x - m2i * a0.y + m3r * a1.x - m3i * a1.y, m2r * a0.y + m2i * a0.x + m3r * a1.y + m3i * a1.x};
    }
}

template <int BC, int BT, class A, int R>
__device__ __forceinline__ void g_cx(A (&v)[R]) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (!(i & (1 << BC)) || (i & (1 << BT))) continue;
        const A t = v[i];
        v[i] = v[i | (1 << BT)];
        v[i | (1 << BT)] = t;
    }
}

template <int B0, int B1, class A, int R>
__device__ __forceinline__ void g_swap(A (&v)[R]) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (!(i & (1 << B0)) || (i & (1 << B1))) continue;
        const int j = i ^ (1 << B0) ^ (1 << B1);
        const A t = v[i];
        v[i] = v[j];
        v[j] = t;
    }
}

// Dense 4x4 on register bits (B0 = matrix bit 0, B1 = matrix bit 1), row-major complex128.
template <int B0, int B1, class A, int R>
__device__ __forceinline__ void g_u4(A (&v)[R], const double* __restrict__ m) {
    using T = typename Real<A>::T;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if ((i & (1 << B0)) || (i & (1 << B1))) continue;
        const int idx[4] = {i, i | (1 << B0), i | (1 << B1), i | (1 << B0) | (1 << B1)};
        A x[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) x[c] = v[idx[c]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            T re = T(0), im = T(0);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T ur = (T)m[2 * (4 * r + c)], ui = (T)m[2 * (4 * r + c) + 1];
                re += ur * x[c].x - ui * x[c].y;
                im += ur * x[c].y + ui * x[c].x;
            }
            v[idx[r]] = A{re, im};
        }
    }
}

// One flat switch over (kind, bit0, bit1): every case indexes the register array with
// compile-time constants only, so the array stays in VGPRs. (Nested or lambda-based
// dispatch made the compiler keep the array in scratch memory.)
#define QLX_K1(K, B, CALL)                     \
    case ((K) << 8) | ((B) << 4):              \
        if constexpr ((B) < RB) { CALL; }       \
        break;
#define QLX_K2(K, B0, B1, CALL)                \
    case ((K) << 8) | ((B0) << 4) | (B1):      \
        if constexpr ((B0) < RB && (B1) < RB) { CALL; } \
        break;
#define QLX_ALL1(K, F) QLX_K1(K, 0, F(0)) QLX_K1(K, 1, F(1)) QLX_K1(K, 2, F(2)) QLX_K1(K, 3, F(3)) QLX_K1(K, 4, F(4))
#define QLX_ROW2(K, B0, F)                                                                      \
    QLX_K2(K, B0, (B0 + 1) % 5, F(B0, (B0 + 1) % 5)) QLX_K2(K, B0, (B0 + 2) % 5, F(B0, (B0 + 2) % 5)) \
    QLX_K2(K, B0, (B0 + 3) % 5, F(B0, (B0 + 3) % 5)) QLX_K2(K, B0, (B0 + 4) % 5, F(B0, (B0 + 4) % 5))
#define QLX_ALL2(K, F) QLX_ROW2(K, 0, F) QLX_ROW2(K, 1, F) QLX_ROW2(K, 2, F) QLX_ROW2(K, 3, F) QLX_ROW2(K, 4, F)

template <class A, int R, int RB>
__device__ __forceinline__ void apply_op(A (&v)[R], const DevOp op, const double* __restrict__ coef) {
    using T = typename Real<A>::T;
    const double* m = coef + op.coef;
    const T c0 = (T)op.c0, c1 = (T)op.c1;
#define F_RX(B) g_rx<B>(v, c0, c1)
#define F_H(B) g_h<B>(v, c0)
#define F_U2(B) g_u2<B>(v, m)
#define F_CX(B0, B1) g_cx<B0, B1>(v)
#define F_SW(B0, B1) g_swap<B0, B1>(v)
#define F_U4(B0, B1) g_u4<B0, B1>(v, m)
    switch (op.key) {
        QLX_ALL1(OP_RX, F_RX)
        QLX_ALL1(OP_H, F_H)
        QLX_ALL1(OP_U2, F_U2)
        QLX_ALL2(OP_CX, F_CX)
        QLX_ALL2(OP_SWAP, F_SW)
        QLX_ALL2(OP_U4, F_U4)
        default: break;
    }
#undef F_RX
#undef F_H
#undef F_U2
#undef F_CX
#undef F_SW
#undef F_U4
}

// One sweep on one 2^W-amplitude chunk, in three steps: load the chunk into registers
// (or synthesise |0...0>), run the register phases with their LDS re-layouts, store.
template <int LOGT, int RB>
__device__ __forceinline__ uint64_t chunk_base(const DevSweep* __restrict__ sw, uint64_t chunk) {
    uint64_t base = chunk;
#pragma unroll
    for (int j = 0; j < LOGT + RB; ++j) base = dins0(base, sw->S[j]);
    return base;
}

// Out of place: the chunk index's bits scattered to the layout's chunk-index address bits.
__device__ __forceinline__ uint64_t chunk_base_oop(const uint32_t* __restrict__ cpos, uint32_t nc, uint64_t chunk) {
    uint64_t base = 0;
    for (uint32_t j = 0; j < nc; ++j) base |= ((chunk >> j) & 1ull) << cpos[j];
    return base;
}

template <class A, int LOGT, int RB>
__device__ __forceinline__ void load_chunk(A (&v)[1 << RB], const A* __restrict__ amp,
                                           const DevSweep* __restrict__ sw, uint64_t base) {
    const uint32_t t = threadIdx.x;
    uint64_t toff = 0;
#pragma unroll
    for (int b = 0; b < LOGT; ++b) toff |= (uint64_t)((t >> b) & 1u) << sw->ld_tglob[b];
#pragma unroll
    for (int i = 0; i < (1 << RB); ++i) v[i] = amp[base | toff | sw->ld_roff[i]];
}

template <class A, int LOGT, int RB>
__device__ __forceinline__ void zero_chunk(A (&v)[1 << RB], const DevSweep* __restrict__ sw, uint64_t base) {
    using T = typename Real<A>::T;
    const uint32_t t = threadIdx.x;
    uint64_t toff = 0;
#pragma unroll
    for (int b = 0; b < LOGT; ++b) toff |= (uint64_t)((t >> b) & 1u) << sw->ld_tglob[b];
#pragma unroll
    for (int i = 0; i < (1 << RB); ++i) v[i] = A{(base | toff | sw->ld_roff[i]) == 0 ? T(1) : T(0), T(0)};
}

template <class A, int LOGT, int RB>
__device__ __forceinline__ void store_chunk(const A (&v)[1 << RB], A* __restrict__ amp,
                                            const DevSweep* __restrict__ sw, uint64_t base) {
    const uint32_t t = threadIdx.x;
    uint64_t toff = 0;
#pragma unroll
    for (int b = 0; b < LOGT; ++b) toff |= (uint64_t)((t >> b) & 1u) << sw->st_tglob[b];
#pragma unroll
    for (int i = 0; i < (1 << RB); ++i) amp[base | toff | sw->st_roff[i]] = v[i];
}

// A phase's ops fused into one 32 x 32 matrix U over the register bits, applied on the matrix
// cores (complex64, 5 register bits, waves of 64 work-items): work-item j of a wave holds the
// 32 amplitudes of its group. v_permlane32_swap of register r with register r + 16 gives the
// two MFMA B operands of the wave's 64 groups (tile A: groups 0-31, lane half h holding their
// amplitudes 16h + r; tile B: groups 32-63), the FP16 split (state scaled by 2^14) and three
// products per K-step accumulate U X in FP32, and the same swap on the results returns every
// group's 32 outputs to its own work-item. The fragments are those of k_mfma_split's FP16 form
// (matrix hi + lo terms; M-tile m, K-step t, lane: 16-byte entry (m 4 + t) 64 + lane).
__device__ __forceinline__ void mfma_phase(float2 (&v)[32], const unsigned char* __restrict__ frags) {
    const int lane = threadIdx.x & 63;
    const f16x8* fh = reinterpret_cast<const f16x8*>(frags);
    const f16x8* fl = fh + 2 * 4 * 64;
    f16x8 xh[2][4], xl[2][4];   // [tile][K-step]
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float2 a[4], b[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int r = 4 * t + p;
            const auto sx = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, v[r].x),
                                                             __builtin_bit_cast(uint32_t, v[r + 16].x), false, false);
            const auto sy = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, v[r].y),
                                                             __builtin_bit_cast(uint32_t, v[r + 16].y), false, false);
            a[p] = float2{__builtin_bit_cast(float, (uint32_t)sx[0]), __builtin_bit_cast(float, (uint32_t)sy[0])};
            b[p] = float2{__builtin_bit_cast(float, (uint32_t)sx[1]), __builtin_bit_cast(float, (uint32_t)sy[1])};
        }
        split_f16(a, kF16Scale, xh[0][t], xl[0][t]);
        split_f16(b, kF16Scale, xh[1][t], xl[1][t]);
    }
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        f32x16 acc[2];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[k][e] = 0.f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const f16x8 ah = fh[(m * 4 + t) * 64 + lane], al = fl[(m * 4 + t) * 64 + lane];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, xh[k][t], acc[k], 0, 0, 0);
                acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, xl[k][t], acc[k], 0, 0, 0);
                acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, xh[k][t], acc[k], 0, 0, 0);
            }
        }
        // acc[k][4a + 2b + c], lane half h: output 16m + 4a + 2h + b (component c) of tile k's
        // group; after the swap tile A's register holds bit-1-clear outputs of the work-item's
        // own group, tile B's the bit-1-set ones
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                float o[2][2];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    // scaled first: the swap then takes ordinary VGPR values (taken straight from the
                    // accumulator vectors, the swaps of different elements were merged into one)
                    const float y0 = acc[0][4 * a + 2 * b + c] * kF16Unscale;
                    const float y1 = acc[1][4 * a + 2 * b + c] * kF16Unscale;
                    const auto sw = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, y0),
                                                                     __builtin_bit_cast(uint32_t, y1), false, false);
                    o[0][c] = __builtin_bit_cast(float, (uint32_t)sw[0]);
                    o[1][c] = __builtin_bit_cast(float, (uint32_t)sw[1]);
                }
                v[16 * m + 4 * a + b] = float2{o[0][0], o[0][1]};
                v[16 * m + 4 * a + 2 + b] = float2{o[1][0], o[1][1]};
            }
    }
}

// The register phases of one sweep. No global-memory access: HBM loads issued before
// the call (the next chunk's) stay in flight across it.
template <class A, int LOGT, int RB>
__device__ __forceinline__ void compute_chunk(A (&v)[1 << RB], const DevSweep* __restrict__ sw,
                                              const DevPhase* __restrict__ ph, const DevOp* __restrict__ ops,
                                              const double* __restrict__ coef, A* __restrict__ lds) {
    constexpr int R = 1 << RB;
    constexpr int ELEM = (int)sizeof(A);
    const uint32_t t = threadIdx.x;
    const DevPhase* cur = ph + sw->phase_begin;
    uint32_t lt = 0;
#pragma unroll
    for (int b = 0; b < LOGT; ++b) lt |= ((t >> b) & 1u) << cur->tlocal[b];

    const uint32_t nph = sw->n_phases;
    for (uint32_t p = 0; p < nph; ++p) {
        const DevPhase* P = ph + sw->phase_begin + p;
        if (P->remap) {
            // the work-item and register parts of a local index are disjoint bit sets and the
            // swizzle is linear over XOR on them: swz(lt | roff) = swz(lt) ^ swz(roff), the
            // register part uniform (scalar arithmetic on the descriptor)
            // (byte addresses; the register parts come precomputed with the descriptor)
            char* const ldsb = reinterpret_cast<char*>(lds);
            const uint32_t s0 = swz<ELEM>(lt) * ELEM;
#pragma unroll
            for (int i = 0; i < R; ++i) *reinterpret_cast<A*>(ldsb + (s0 ^ cur->rslot[i])) = v[i];
            __syncthreads();
            lt = 0;
#pragma unroll
            for (int b = 0; b < LOGT; ++b) lt |= ((t >> b) & 1u) << P->tlocal[b];
            const uint32_t s1 = swz<ELEM>(lt) * ELEM;
#pragma unroll
            for (int i = 0; i < R; ++i) v[i] = *reinterpret_cast<const A*>(ldsb + (s1 ^ P->rslot[i]));
            __syncthreads();
        }
        cur = P;
        if constexpr (std::is_same_v<A, float2> && LOGT == 8 && RB == 5) {
            if (P->mfma) {
                mfma_phase(v, reinterpret_cast<const unsigned char*>(coef + (P->mfma - 1)));
                continue;
            }
        }
        const uint32_t nops = P->n_ops, ob = P->op_begin;
        // the next descriptor's scalar loads are issued before the current op runs
        DevOp op = nops ? ops[ob] : DevOp{};
        for (uint32_t o = 0; o < nops; ++o) {
            const DevOp nop = ops[ob + (o + 1 < nops ? o + 1 : o)];
            apply_op<A, R, RB>(v, op, coef);
            op = nop;
        }
    }
}

// One chunk start to end (the batched and served kernels). In place (dst == src) the chunk's
// base comes from the window's qubits; out of place from the layouts' chunk-index bits.
template <class A, int LOGT, int RB>
__device__ __forceinline__ void run_chunk(const A* src, A* dst, bool oop,
                                          const DevSweep* __restrict__ sw, uint64_t chunk,
                                          const DevPhase* __restrict__ ph, const DevOp* __restrict__ ops,
                                          const double* __restrict__ coef, A* __restrict__ lds, bool init_zero) {
    A v[1 << RB];
    const uint64_t base = oop ? chunk_base_oop(sw->ld_cpos, sw->nc, chunk) : chunk_base<LOGT, RB>(sw, chunk);
    if (init_zero) zero_chunk<A, LOGT, RB>(v, sw, base);
    else load_chunk<A, LOGT, RB>(v, src, sw, base);
    compute_chunk<A, LOGT, RB>(v, sw, ph, ops, coef, lds);
    const uint64_t sbase = oop ? chunk_base_oop(sw->st_cpos, sw->nc, chunk) : base;
    store_chunk<A, LOGT, RB>(v, dst, sw, sbase);
}

// ---------------------------------------------------------------------------------
// host planner interface (qlx_sweep.hip)
// ---------------------------------------------------------------------------------
struct Geometry {
    int logt, rb, w, c;
    int elem = 16;  // bytes per amplitude
    // out-of-place sweeps: a window need not hold qubits 0..c-1. Each sweep's loads and stores
    // are coalesced on c qubits it holds that sit on address bits 0..c-1 (the first sweep loads
    // and the last stores the logical layout, so their windows hold qubits 0..c-1)
    bool free = false;
};

struct PlanOp {
    int src;
    uint8_t kind;
    int l0, l1;     // register bits (l1 = -1 for 1q)
    int tctl = -1;  // a CX whose control is this work-item bit (OP_CXT; l0 unused)
    int merged = -1;  // a merged rotation: its index in the planner's table of merged rotations
};

struct PlanPhase {
    std::vector<int> st2loc;  // storage bit -> local bit (size W)
    std::vector<PlanOp> ops;
};

struct PlanSweep {
    // global qubits, size W: local bit j is S[j]. The first c are the qubits on address bits
    // 0..c-1 at the load (ascending), then the rest ascending
    std::vector<int> S;
    // the c qubits on address bits 0..c-1 at the store (the next sweep's coalesced qubits),
    // ascending; they are the low work-item bits of the last phase
    std::vector<int> st_low;
    std::vector<PlanPhase> phases;
    // out-of-place sweeps (SweepPlan::relabel): address bit of chunk-index bit j at the load
    // and at the store, and of local bit l at the load and at the store
    std::vector<int> ld_cpos, st_cpos, ld_lpos, st_lpos;
};

struct PlanStep {
    bool sweep;
    int src;  // plain instruction
    int sweep_index;
};

struct SweepPlan {
    Geometry geo{};
    std::vector<PlanStep> steps;
    std::vector<PlanSweep> sweeps;
    std::vector<int> order;
    std::vector<int> order_sweep;
    std::vector<DevSweep> dsw;
    std::vector<DevPhase> dph;
    std::vector<DevOp> dop;
    std::vector<double> coef;
    uint64_t phases = 0, gates_in_sweeps = 0, rotations_merged = 0;
    double passes = 0.0;  // full-state passes: one per sweep, plus the plain steps' fractions
    uint64_t tag = 0;     // hash of the planned program (names its generated kernels)
    // out-of-place sweeps with per-sweep qubit layouts (assign_layouts): step_pos[s][q] is the
    // address bit of logical qubit q before step s; the layout after the last step is the identity
    bool relabel = false;
    std::vector<std::vector<int>> step_pos;
};

// Windows of dense sweeps (QLX_EXEC_DENSE_SWEEPS): every instruction of arity <= 5 runs
// inside a window of `w` qubits that holds the low `c` qubits (out-of-place sweeps: the `c`
// qubits its load and its store are coalesced on); the windows come from the beam search
// over window sequences. Each window lists its instructions in program order.
struct DenseWindow {
    // window qubits: the c qubits on address bits 0..c-1 at the load first (ascending), then
    // the rest ascending
    std::vector<int> S;
    std::vector<int> st_low;  // the c qubits on address bits 0..c-1 at the store, ascending
    std::vector<int> instrs;  // indices into the instruction stream
};
std::vector<DenseWindow> plan_dense_windows(uint32_t nq, const qlx_gate_instr* instrs, size_t n,
                                            const qlx_exec_opts& opts, int w, int c);

// Gives every sweep of the plan a load layout and a store layout (the relabel option): each
// sweep stores through a layout that puts the next sweep's window on the low address bits;
// the low C qubits never move and the last sweep stores through the identity. Also rewrites
// the sweeps' device descriptors (work-item and register address bits) to the new layouts.
void assign_layouts(SweepPlan& p, uint32_t nq);

// Complex64 plans of 5 register bits and 256 work-items (generic kernels): every phase whose ops
// would cost at least kDensePhaseCost single-qubit ops on the vector units (CX/SWAP a quarter,
// a 2-qubit matrix four) is fused into one 32 x 32 matrix over its register bits (FP64 product
// on the host) and applied on the matrix cores (DevPhase::mfma; FP16 fragments appended to the
// coefficient array). Not for the circuit-specialised sweeps.
constexpr double kDensePhaseCost = 5.0;
void densify_phases(SweepPlan& p);

// Plans windowed execution of a MEASURE-free stream with an explicit geometry (W <= nq).
// thorough: also try the classic greedy window chooser and a beam search, let sweeps that
// cost less than a pass run as plain kernels, and keep the cheapest plan (several times the
// planning time; for programs planned once and executed many times). Batched and served
// circuits plan without it: their executors run sweeps only.
std::unique_ptr<SweepPlan> build_plan_geo(uint32_t nq, const Geometry& geo, const qlx_gate_instr* instrs, size_t n,
                                          const double* pool, const qlx_matrix_entry* meta,
                                          const qlx_exec_opts& opts, bool thorough);

}  // namespace qlx
