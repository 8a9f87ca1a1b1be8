// qlx_gates.hip — one-sweep-per-instruction gate kernels for gfx950 and the plain executor.
//
// Each kernel restates one loop of CPUOpenMPEngine (include/backend/cpu_openmp_engine.hpp):
//   single_target_loop  :260-266  -> k_1q      (one work-item per amplitude pair)
//   double_hole_loop    :270-297  -> k_2hole   (one work-item per quad; CX/SWAP touch 1/2 of the state)
//   triple_hole_loop    :301-334  -> k_3hole   (one work-item per octet; CCX/CSWAP touch 1/4)
//   apply_unitary       :539-606  -> k_dense   (LDS-staged 2^k-amplitude groups, k <= 5);
//                                        k_dense_f64mfma (complex128, k = 4, 5: FP64 matrix cores)
//   apply_measure       :617-671  -> k_prob1 / k_collapse
// The per-gate arithmetic follows the reference expressions term by term so that
// permutation gates (X, Y, Z, S, CX, SWAP, CCX, CSWAP) and H are bit-exact with it.
// Index arithmetic is 64-bit: states up to 2^34 amplitudes (256 GiB c128) fit one MI355X.

#include <algorithm>
#include <mutex>
#include <cmath>

#include "qlx_internal.hpp"

namespace qlx {

template <class A>
struct RealOf;
template <>
struct RealOf<double2> { using T = double; };
template <>
struct RealOf<float2> { using T = float; };

enum Kind1 : int { K_H, K_X, K_Y, K_Z, K_S, K_RX, K_RY, K_RZ, K_U2 };

// Host-computed coefficients (FP64, narrowed for c64), passed by value.
struct Coef1 {
    double c, s;            // RX/RY: cos(theta/2), sin(theta/2); H: 1/sqrt(2)
    double p0r, p0i, p1r, p1i;  // RZ phases
    double m[8];            // U2: row-major 2x2 complex (re, im)
};

constexpr int kThreads = 256;

__device__ inline uint64_t insert_zero(uint64_t idx, uint32_t t) {
    const uint64_t lo = idx & ((1ull << t) - 1);
    return lo | ((idx ^ lo) << 1);
}

template <class A, int KIND>
__global__ void __launch_bounds__(kThreads) k_1q(A* __restrict__ amp, uint64_t npairs, uint32_t t, Coef1 cf) {
    using R = typename RealOf<A>::T;
    // grid-stride: a launch holds fewer than 2^32 work-items, a 32-qubit state has 2^31 pairs
    for (uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x; k < npairs;
         k += (uint64_t)gridDim.x * kThreads) {
    const uint64_t i0 = insert_zero(k, t);
    const uint64_t i1 = i0 | (1ull << t);
    if constexpr (KIND == K_Z) {
        A a1 = amp[i1];
        amp[i1] = A{-a1.x, -a1.y};
    } else if constexpr (KIND == K_S) {
        A a1 = amp[i1];
        amp[i1] = A{-a1.y, a1.x};
    } else {
        const A a0 = amp[i0];
        const A a1 = amp[i1];
        A o0, o1;
        if constexpr (KIND == K_H) {
            const R r = (R)cf.c;
            o0 = A{r * (a0.x + a1.x), r * (a0.y + a1.y)};
            o1 = A{r * (a0.x - a1.x), r * (a0.y - a1.y)};
        } else if constexpr (KIND == K_X) {
            o0 = a1;
            o1 = a0;
        } else if constexpr (KIND == K_Y) {
            o0 = A{a1.y, -a1.x};
            o1 = A{-a0.y, a0.x};
        } else if constexpr (KIND == K_RX) {
            const R c = (R)cf.c, s = (R)cf.s;
            o0 = A{c * a0.x + s * a1.y, c * a0.y + (-s * a1.x)};
            o1 = A{s * a0.y + c * a1.x, -s * a0.x + c * a1.y};
        } else if constexpr (KIND == K_RY) {
            const R c = (R)cf.c, s = (R)cf.s;
            o0 = A{c * a0.x - s * a1.x, c * a0.y - s * a1.y};
            o1 = A{s * a0.x + c * a1.x, s * a0.y + c * a1.y};
        } else if constexpr (KIND == K_RZ) {
            const R p0r = (R)cf.p0r, p0i = (R)cf.p0i, p1r = (R)cf.p1r, p1i = (R)cf.p1i;
            o0 = A{a0.x * p0r - a0.y * p0i, a0.x * p0i + a0.y * p0r};
            o1 = A{a1.x * p1r - a1.y * p1i, a1.x * p1i + a1.y * p1r};
        } else {  // K_U2
            const R m0r = (R)cf.m[0], m0i = (R)cf.m[1], m1r = (R)cf.m[2], m1i = (R)cf.m[3];
            const R m2r = (R)cf.m[4], m2i = (R)cf.m[5], m3r = (R)cf.m[6], m3i = (R)cf.m[7];
            o0 = A{m0r * a0.x - m0i * a0.y + m1r * a1.x - m1i * a1.y,
                   m0r * a0.y + m0i * a0.x + m1r * a1.y + m1i * a1.x};
            o1 = A{m2r * a0.x - m2i * a0.y + m3r * a1.x - m3i * a1.y,
                   m2r * a0.y + m2i * a0.x + m3r * a1.y + m3i * a1.x};
        }
        amp[i0] = o0;
        amp[i1] = o1;
    }
    }
}

// Two zero holes at qlo < qhi; swaps amp[base|ma] <-> amp[base|mb].
// CX: ma = ctrl, mb = ctrl|tgt (cpu_openmp_engine.hpp:468-480).
// SWAP: ma = m1, mb = m0 (cpu_openmp_engine.hpp:483-494).
template <class A>
__global__ void __launch_bounds__(kThreads) k_2hole(A* __restrict__ amp, uint64_t nquads, uint32_t qlo,
                                                   uint32_t qhi, uint64_t ma, uint64_t mb) {
    for (uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x; k < nquads;
         k += (uint64_t)gridDim.x * kThreads) {
        const uint64_t base = insert_zero(insert_zero(k, qlo), qhi);
        const uint64_t ia = base | ma, ib = base | mb;
        const A x = amp[ia];
        const A y = amp[ib];
        amp[ia] = y;
        amp[ib] = x;
    }
}

// Three zero holes (sorted q0 < q1 < q2); swaps amp[base|ma] <-> amp[base|mb].
// CCX: ma = c0|c1, mb = c0|c1|t (:497-511). CSWAP: ma = c|m1, mb = c|m0 (:514-528).
template <class A>
__global__ void __launch_bounds__(kThreads) k_3hole(A* __restrict__ amp, uint64_t nocts, uint32_t q0,
                                                   uint32_t q1, uint32_t q2, uint64_t ma, uint64_t mb) {
    for (uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x; k < nocts;
         k += (uint64_t)gridDim.x * kThreads) {
        const uint64_t base = insert_zero(insert_zero(insert_zero(k, q0), q1), q2);
        const uint64_t ia = base | ma, ib = base | mb;
        const A x = amp[ia];
        const A y = amp[ib];
        amp[ia] = y;
        amp[ib] = x;
    }
}

// ---- dense k-qubit unitary (UNITARY / FUSED_BLOCK), k <= 5 ------------------------
// Work-group = 4 waves, tile = 64 groups of D = 2^k amplitudes staged in LDS as
// in[j][g] (g contiguous). Wave w computes rows j = w, w+4, ... for the 64 groups
// (lane = group); matrix rows are wave-uniform, so U is read with scalar loads.
struct DenseArgs {
    uint64_t off[32];   // offset of matrix index j: OR of (1 << qubits[i]) for set bits i of j
    uint32_t sorted[5]; // target qubits ascending (zero-hole insertion order)
    uint32_t k;
};

template <class A, int K>
__global__ void __launch_bounds__(kThreads) k_dense(A* __restrict__ amp, uint64_t ngroups,
                                                   const double* __restrict__ U, DenseArgs da) {
    using R = typename RealOf<A>::T;
    constexpr int D = 1 << K;
    constexpr int RPW = D >= 4 ? D / 4 : 1;   // rows per wave
    __shared__ A tile[D][64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const uint64_t ntiles = (ngroups + 63) / 64;
    for (uint64_t tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const uint64_t g0 = tix * 64;
        for (int e = threadIdx.x; e < D * 64; e += kThreads) {
            const int j = e >> 6, g = e & 63;
            const uint64_t gg = g0 + g;
            if (gg < ngroups) {
                uint64_t base = gg;
#pragma unroll
                for (int q = 0; q < K; ++q) base = insert_zero(base, da.sorted[q]);
                tile[j][g] = amp[base | da.off[j]];
            }
        }
        __syncthreads();
        if (wave < D) {  // D == 2: two waves idle
            R accr[RPW], acci[RPW];
#pragma unroll
            for (int r = 0; r < RPW; ++r) accr[r] = acci[r] = R(0);
#pragma unroll 4
            for (int m = 0; m < D; ++m) {
                const A x = tile[m][lane];
#pragma unroll
                for (int r = 0; r < RPW; ++r) {
                    const int j = wave + 4 * r;
                    const R ur = (R)U[2 * (j * D + m)], ui = (R)U[2 * (j * D + m) + 1];
                    accr[r] += ur * x.x - ui * x.y;
                    acci[r] += ur * x.y + ui * x.x;
                }
            }
            const uint64_t gg = g0 + lane;
            if (gg < ngroups) {
                uint64_t base = gg;
#pragma unroll
                for (int q = 0; q < K; ++q) base = insert_zero(base, da.sorted[q]);
#pragma unroll
                for (int r = 0; r < RPW; ++r) {
                    const int j = wave + 4 * r;
                    amp[base | da.off[j]] = A{accr[r], acci[r]};
                }
            }
        }
        __syncthreads();  // the tile is read completely before the next one is staged
    }
}

// ---- dense 4/5-qubit unitary on complex128, FP64 matrix cores -----------------------
// v_mfma_f64_16x16x4f64 computes D[16x16] += A[16x4] B[4x16] in FP64; work-item l holds
// A[l%16][l/16], B[l/16][l%16] and D[l/16 + 4v][l%16] in element v (rows interleaved; a first
// version assumed D[4(l/16) + v], and a CPU emulation with the interleaved rows reproduced its
// results on the GPU exactly). The gate runs as the real embedding E (2D x 2D, D = 2^K) of U,
// E = [[Re U, -Im U], [Im U, Re U]], with rows and columns ordered so that work-items load
// and store whole amplitudes:
//   column c = 4s + h (k-step s, lane quarter h) <-> input amplitude 4(s/2) + h, part s%2;
//   row 16r + i (block r, i = h + 4v) <-> output amplitude 8r + 2h + v/2, part v%2, so lane
//   quarter h holds outputs 8r + 2h (elements 0, 1) and 8r + 2h + 1 (elements 2, 3).
// The 16 columns of B are 16 amplitude groups (a tile). E is built once per work-group in LDS
// (K = 5: 32 KiB), fragment (r, s) as 64 consecutive doubles in lane order, so every MFMA
// reads its A operand with one conflict-free ds_read_b64. Wave w walks tiles w, w + W, ...
// (W = waves in the grid, all resident). Every group of a tile is read and written by one
// wave, so the update is in place. Measured on a 30-qubit state (16 GiB, copy 6.47 ms), K = 5:
// E in LDS without prefetch (4 waves per SIMD) 6.91 / 6.91 / 7.49 ms for targets 0-4, 25-29
// and 3,9,15,21,27; E in VGPRs (2 waves per SIMD) 6.85 / 6.89 / 8.05 ms. K = 4 (E: 8 KiB)
// also loads the next tile before computing the current one (spread targets: 7.19 ms; E in
// VGPRs 7.37 ms).
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int K>
__global__ void __launch_bounds__(kThreads) k_dense_f64mfma(double2* __restrict__ amp, uint64_t ntiles,
                                                           const double* __restrict__ U, DenseArgs da) {
    constexpr int D = 1 << K;
    constexpr int RB = 2 * D / 16;  // row blocks of E
    constexpr int KS = 2 * D / 4;   // k-steps
    constexpr int NP = D / 4;       // amplitudes per work-item per tile
    constexpr bool kPrefetch = K == 4;
    __shared__ double e[RB * KS * 64];
    for (int idx = threadIdx.x; idx < RB * KS * 64; idx += kThreads) {
        const int l = idx & 63, f = idx >> 6, r = f / KS, s = f % KS;
        const int jj = l & 15, hh = l >> 4;
        const int mo = 8 * r + 2 * (jj & 3) + (jj >> 3), po = (jj >> 2) & 1;
        const int mi = 4 * (s >> 1) + hh, pi = s & 1;
        const double ur = U[2 * (mo * D + mi)], ui = U[2 * (mo * D + mi) + 1];
        e[idx] = po == pi ? ur : (po == 0 ? -ui : ui);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, h = lane >> 4;
    // matrix index bits 0-1 of an input amplitude are h; bits 1-2 of an output amplitude are h
    const uint64_t in_h = h == 0 ? da.off[0] : h == 1 ? da.off[1] : h == 2 ? da.off[2] : da.off[3];
    const uint64_t out_h = h == 0 ? da.off[0] : h == 1 ? da.off[2] : h == 2 ? da.off[4] : da.off[6];
    const uint64_t nwaves = (uint64_t)gridDim.x * (kThreads / 64);
    uint64_t tile = (uint64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    auto group_base = [&](uint64_t t) {
        uint64_t b = t * 16 + (uint64_t)j;
#pragma unroll
        for (int q = 0; q < K; ++q) b = insert_zero(b, da.sorted[q]);
        return b;
    };
    // one tile: 2D/4 k-steps x RB row blocks of MFMA, then the outputs in place
    auto apply = [&](uint64_t base, const double2 (&x)[NP]) {
        f64x4 acc[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const double b = (s & 1) ? x[s >> 1].y : x[s >> 1].x;
#pragma unroll
            for (int r = 0; r < RB; ++r)
                acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(e[(r * KS + s) * 64 + lane], b, acc[r], 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const uint64_t o = base | out_h | da.off[8 * r];
            amp[o] = double2{acc[r][0], acc[r][1]};
            amp[o | da.off[1]] = double2{acc[r][2], acc[r][3]};
        }
    };
    if constexpr (kPrefetch) {
        uint64_t base = group_base(tile);
        double2 x[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) x[p] = amp[base | in_h | da.off[4 * p]];
        for (;;) {
            const uint64_t next = tile + nwaves;
            const bool more = next < ntiles;
            uint64_t nbase = base;
            double2 xn[NP];
            if (more) {
                nbase = group_base(next);
#pragma unroll
                for (int p = 0; p < NP; ++p) xn[p] = amp[nbase | in_h | da.off[4 * p]];
            }
            apply(base, x);
            if (!more) break;
            tile = next;
            base = nbase;
#pragma unroll
            for (int p = 0; p < NP; ++p) x[p] = xn[p];
        }
    } else {
        for (; tile < ntiles; tile += nwaves) {
            // E's reads stay inside the loop (hoisted, they would take 128 VGPRs at K = 5)
            asm volatile("" ::: "memory");
            const uint64_t base = group_base(tile);
            double2 x[NP];
#pragma unroll
            for (int p = 0; p < NP; ++p) x[p] = amp[base | in_h | da.off[4 * p]];
            apply(base, x);
        }
    }
}

// ---- measurement ----------------------------------------------------------------
constexpr int kRedBlocks = 2048;

template <class A>
__global__ void __launch_bounds__(kThreads) k_prob1(const A* __restrict__ amp, uint64_t half, uint32_t t,
                                                   double* __restrict__ part) {
    __shared__ double sh[kThreads / 64];
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < half; i += stride) {
        const A v = amp[insert_zero(i, t) | (1ull << t)];
        const double re = (double)v.x, im = (double)v.y;
        acc += re * re + im * im;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) s += sh[w];
        part[blockIdx.x] = s;
    }
}

template <class A>
__global__ void __launch_bounds__(kThreads) k_collapse(A* __restrict__ amp, uint64_t half, uint32_t t,
                                                      int outcome, double inv_norm) {
    using R = typename RealOf<A>::T;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < half;
         i += (uint64_t)gridDim.x * kThreads) {
        const uint64_t i0 = insert_zero(i, t), i1 = i0 | (1ull << t);
        const uint64_t keep = outcome ? i1 : i0, drop = outcome ? i0 : i1;
        A v = amp[keep];
        if constexpr (sizeof(R) == 8) {
            v = A{v.x * inv_norm, v.y * inv_norm};
        } else {
            v = A{(R)((double)v.x * inv_norm), (R)((double)v.y * inv_norm)};
        }
        amp[keep] = v;
        amp[drop] = A{R(0), R(0)};
    }
}

// ---- host side -----------------------------------------------------------------------

// Blocks for n work-items, capped so a launch stays below 2^32 work-items (the kernels
// are grid-stride loops; above 2^30 items each work-item takes several).
static unsigned grid_for(uint64_t n) {
    uint64_t b = (n + kThreads - 1) / kThreads;
    if (b == 0) b = 1;
    return (unsigned)std::min<uint64_t>(b, uint64_t{1} << 22);
}

template <class A>
static void launch_1q_t(qlx_state* st, int kind, uint32_t t, const Coef1& cf) {
    A* amp = (A*)st->amp;
    const uint64_t np = 1ull << (st->nq - 1);
    const unsigned g = grid_for(np);
    switch (kind) {
        case K_H: k_1q<A, K_H><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
        case K_X: k_1q<A, K_X><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
        case K_Y: k_1q<A, K_Y><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
        case K_Z: k_1q<A, K_Z><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
        case K_S: k_1q<A, K_S><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
        case K_RX: k_1q<A, K_RX><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
        case K_RY: k_1q<A, K_RY><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
        case K_RZ: k_1q<A, K_RZ><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
        default: k_1q<A, K_U2><<<g, kThreads, 0, st->stream>>>(amp, np, t, cf); break;
    }
    QLX_HIP_CHECK(hipGetLastError());
}

static void launch_1q(qlx_state* st, int kind, uint32_t t, const Coef1& cf) {
    if (st->dtype == QLX_C128) launch_1q_t<double2>(st, kind, t, cf);
    else launch_1q_t<float2>(st, kind, t, cf);
}

static void launch_2hole(qlx_state* st, uint32_t a, uint32_t b, uint64_t ma, uint64_t mb) {
    const uint32_t lo = std::min(a, b), hi = std::max(a, b);
    const uint64_t n = 1ull << (st->nq - 2);
    if (st->dtype == QLX_C128)
        k_2hole<double2><<<grid_for(n), kThreads, 0, st->stream>>>((double2*)st->amp, n, lo, hi, ma, mb);
    else
        k_2hole<float2><<<grid_for(n), kThreads, 0, st->stream>>>((float2*)st->amp, n, lo, hi, ma, mb);
    QLX_HIP_CHECK(hipGetLastError());
}

static void launch_3hole(qlx_state* st, uint32_t a, uint32_t b, uint32_t c, uint64_t ma, uint64_t mb) {
    uint32_t q[3] = {a, b, c};
    std::sort(q, q + 3);
    const uint64_t n = 1ull << (st->nq - 3);
    if (st->dtype == QLX_C128)
        k_3hole<double2><<<grid_for(n), kThreads, 0, st->stream>>>((double2*)st->amp, n, q[0], q[1], q[2], ma, mb);
    else
        k_3hole<float2><<<grid_for(n), kThreads, 0, st->stream>>>((float2*)st->amp, n, q[0], q[1], q[2], ma, mb);
    QLX_HIP_CHECK(hipGetLastError());
}

// complex128, 4/5-qubit gates on at least 16 groups: the FP64 matrix-core kernel, one grid of
// resident work-groups (its waves split the tiles statically)
template <int K>
static void launch_dense_f64mfma(qlx_state* st, const double* dU, const DenseArgs& da) {
    const uint64_t ntiles = (1ull << (st->nq - K)) / 16;
    static thread_local int per_cu = 0;
    if (!per_cu) {
        QLX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dense_f64mfma<K>, kThreads, 0));
        per_cu = std::max(per_cu, 1);
    }
    int dev = 0, cus = 0;
    QLX_HIP_CHECK(hipGetDevice(&dev));
    QLX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint64_t grid = std::min<uint64_t>((ntiles + 3) / 4, (uint64_t)per_cu * (uint64_t)cus);
    k_dense_f64mfma<K><<<(unsigned)grid, kThreads, 0, st->stream>>>((double2*)st->amp, ntiles, dU, da);
    QLX_HIP_CHECK(hipGetLastError());
}

template <class A>
static void launch_dense_t(qlx_state* st, const double* dU, const DenseArgs& da) {
    const uint64_t ngroups = 1ull << (st->nq - da.k);
    const unsigned g = (unsigned)std::min<uint64_t>((ngroups + 63) / 64, uint64_t{1} << 22);
    A* amp = (A*)st->amp;
    switch (da.k) {
        case 1: k_dense<A, 1><<<g, kThreads, 0, st->stream>>>(amp, ngroups, dU, da); break;
        case 2: k_dense<A, 2><<<g, kThreads, 0, st->stream>>>(amp, ngroups, dU, da); break;
        case 3: k_dense<A, 3><<<g, kThreads, 0, st->stream>>>(amp, ngroups, dU, da); break;
        case 4: k_dense<A, 4><<<g, kThreads, 0, st->stream>>>(amp, ngroups, dU, da); break;
        case 5: k_dense<A, 5><<<g, kThreads, 0, st->stream>>>(amp, ngroups, dU, da); break;
        default: fail(QLX_EUNSUPPORTED, "dense gate arity must be 1..5");
    }
    QLX_HIP_CHECK(hipGetLastError());
}

void launch_dense(qlx_state* st, const double* dU, const uint8_t* qubits, int k, bool matrix_cores) {
    if (k < 1 || k > 5) fail(QLX_EUNSUPPORTED, "dense gate arity must be 1..5");
    if ((uint32_t)k > st->nq) fail(QLX_ERANGE, "gate arity exceeds the qubit count");
    DenseArgs da{};
    da.k = (uint32_t)k;
    for (int j = 0; j < (1 << k); ++j) {
        uint64_t off = 0;
        for (int i = 0; i < k; ++i)
            if ((j >> i) & 1) off |= 1ull << qubits[i];
        da.off[j] = off;
    }
    uint32_t s[5];
    for (int i = 0; i < k; ++i) s[i] = qubits[i];
    std::sort(s, s + k);
    for (int i = 0; i < k; ++i) da.sorted[i] = s[i];
    if (st->dtype == QLX_C128 && matrix_cores && k >= 4 && st->nq >= (uint32_t)k + 4) {
        if (k == 5) launch_dense_f64mfma<5>(st, dU, da);
        else launch_dense_f64mfma<4>(st, dU, da);
    } else if (st->dtype == QLX_C128) {
        launch_dense_t<double2>(st, dU, da);
    } else {
        launch_dense_t<float2>(st, dU, da);
    }
}

// ---- validation ---------------------------------------------------------------------
static int fixed_arity(uint8_t type) {
    switch (type) {
        case QLX_H: case QLX_X: case QLX_Y: case QLX_Z: case QLX_S:
        case QLX_RX: case QLX_RY: case QLX_RZ: case QLX_MEASURE: return 1;
        case QLX_CX: case QLX_SWAP: case QLX_GLOBAL_SWAP: return 2;
        case QLX_CCX: case QLX_CSWAP: return 3;
        case QLX_BARRIER: return 0;
        default: return -1;  // matrix gates: arity from the instruction
    }
}

static bool known_type(uint8_t t) {
    switch (t) {
        case QLX_H: case QLX_X: case QLX_Y: case QLX_Z: case QLX_S: case QLX_RX: case QLX_RY:
        case QLX_RZ: case QLX_CX: case QLX_SWAP: case QLX_CCX: case QLX_CSWAP: case QLX_UNITARY:
        case QLX_FUSED_BLOCK: case QLX_GLOBAL_SWAP: case QLX_BARRIER: case QLX_MEASURE: return true;
        default: return false;
    }
}

void validate_instr(const qlx_gate_instr& g, uint32_t nq, const qlx_matrix_entry* meta, size_t n_meta,
                    size_t pool_elems) {
    int ar = fixed_arity(g.type);
    if (ar < 0) {
        ar = instr_arity(g);
        if (ar < 1 || ar > 5)
            fail(QLX_EUNSUPPORTED, "matrix gate arity " + std::to_string(ar) + " not in 1..5");
        const uint32_t idx = g.payload.matrix_idx;
        if (!meta || idx >= n_meta) fail(QLX_ERANGE, "matrix_idx " + std::to_string(idx) + " out of range");
        const qlx_matrix_entry& e = meta[idx];
        if (e.num_qubits != ar || e.num_elements != (1u << (2 * ar)))
            fail(QLX_EINVAL, "matrix size does not match the gate arity");
        if ((size_t)e.offset + e.num_elements > pool_elems) fail(QLX_ERANGE, "matrix exceeds the unitary pool");
    }
    for (int i = 0; i < ar; ++i) {
        const int q = instr_qubit(g, i);
        if ((uint32_t)q >= nq)
            fail(QLX_ERANGE, "Qubit index " + std::to_string(q) + " out of range [0, " + std::to_string(nq) + ")");
        for (int j = 0; j < i; ++j)
            if (instr_qubit(g, j) == q) fail(QLX_EINVAL, "gate requires distinct qubits");
    }
}

// Index of the first instruction with an unknown opcode (or n).
static size_t first_unknown(const qlx_gate_instr* instrs, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!known_type(instrs[i].type)) return i;
    return n;
}

static Coef1 coef_for(const qlx_gate_instr& g) {
    Coef1 cf{};
    // Coefficients are evaluated on the host exactly as the reference does
    // (cpu_openmp_engine.hpp:339, :409-411, :430-432, :451-453).
    switch (g.type) {
        case QLX_H: cf.c = 1.0 / std::sqrt(2.0); break;
        case QLX_RX:
        case QLX_RY: {
            const double half = static_cast<double>(g.payload.param) * 0.5;
            cf.c = std::cos(half);
            cf.s = std::sin(half);
            break;
        }
        case QLX_RZ: {
            const double half = static_cast<double>(g.payload.param) * 0.5;
            cf.p0r = std::cos(-half);
            cf.p0i = std::sin(-half);
            cf.p1r = std::cos(half);
            cf.p1i = std::sin(half);
            break;
        }
        default: break;
    }
    return cf;
}

static int kind_for(uint8_t type) {
    switch (type) {
        case QLX_H: return K_H;
        case QLX_X: return K_X;
        case QLX_Y: return K_Y;
        case QLX_Z: return K_Z;
        case QLX_S: return K_S;
        case QLX_RX: return K_RX;
        case QLX_RY: return K_RY;
        case QLX_RZ: return K_RZ;
        default: return K_U2;
    }
}

const double* upload_pool(qlx_state* st, const double* pool, size_t pool_elems) {
    if (!pool_elems) return nullptr;
    st->prog.reserve(pool_elems * 16, st->stream);
    QLX_HIP_CHECK(hipMemcpyAsync(st->prog.ptr, pool, pool_elems * 16, hipMemcpyHostToDevice, st->stream));
    // the source is pageable caller memory: complete the copy before returning
    QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
    return (const double*)st->prog.ptr;
}

void launch_plain_instr(qlx_state* st, const qlx_gate_instr& g, const double* dpool, const qlx_matrix_entry* meta,
                        const qlx_exec_opts& opts, qlx_exec_stats& stats) {
    const double pass = pass_bytes(st);
    auto account = [&](double f) {
        if (f >= 1.0) stats.sweeps_full++;
        else if (f >= 0.5) stats.sweeps_half++;
        else stats.sweeps_quarter++;
        stats.algorithmic_bytes += f * pass;
        stats.gates_applied++;
    };
    const uint8_t* q = g.qubits;
    switch (g.type) {
        case QLX_H: case QLX_X: case QLX_Y: case QLX_RX: case QLX_RY: case QLX_RZ: {
            TimedLaunch tl(st, "k_1q", pass);
            launch_1q(st, kind_for(g.type), q[0], coef_for(g));
            account(1.0);
            break;
        }
        case QLX_Z: case QLX_S: {
            TimedLaunch tl(st, "k_1q_phase", 0.5 * pass);
            launch_1q(st, kind_for(g.type), q[0], coef_for(g));
            account(0.5);
            break;
        }
        case QLX_CX: {
            TimedLaunch tl(st, "k_2hole", 0.5 * pass);
            launch_2hole(st, q[0], q[1], 1ull << q[0], (1ull << q[0]) | (1ull << q[1]));
            account(0.5);
            break;
        }
        case QLX_SWAP: {
            TimedLaunch tl(st, "k_2hole", 0.5 * pass);
            launch_2hole(st, q[0], q[1], 1ull << q[1], 1ull << q[0]);
            account(0.5);
            break;
        }
        case QLX_GLOBAL_SWAP: {
            if (opts.gswap_mode == QLX_GSWAP_NOOP) break;
            TimedLaunch tl(st, "k_2hole", 0.5 * pass);
            launch_2hole(st, q[0], q[1], 1ull << q[1], 1ull << q[0]);
            account(0.5);
            break;
        }
        case QLX_CCX: {
            TimedLaunch tl(st, "k_3hole", 0.25 * pass);
            const uint64_t cm = (1ull << q[0]) | (1ull << q[1]);
            launch_3hole(st, q[0], q[1], q[2], cm, cm | (1ull << q[2]));
            account(0.25);
            break;
        }
        case QLX_CSWAP: {
            TimedLaunch tl(st, "k_3hole", 0.25 * pass);
            const uint64_t c = 1ull << q[0];
            launch_3hole(st, q[0], q[1], q[2], c | (1ull << q[2]), c | (1ull << q[1]));
            account(0.25);
            break;
        }
        case QLX_UNITARY:
        case QLX_FUSED_BLOCK: {
            const int k = instr_arity(g);
            uint8_t qs[5];
            for (int j = 0; j < k; ++j) qs[j] = (uint8_t)instr_qubit(g, j);
            const uint32_t mi = g.payload.matrix_idx;
            const qlx_matrix_entry& e = meta[mi];
            if (mi < st->frag_index.size() && st->frag_index[mi] >= 0) {
                TimedLaunch tl(st, "k_dense_mfma", pass);
                launch_dense_mfma(st, (const unsigned char*)st->frags.ptr + st->frag_index[mi], qs, k,
                                  st->frag_variant);
            } else {
                const bool mc = st->dtype == QLX_C128 && opts.dense_kernel != QLX_DENSE_VALU && k >= 4 &&
                                st->nq >= (uint32_t)k + 4;
                TimedLaunch tl(st, mc ? "k_dense_f64mfma" : "k_dense", pass);
                launch_dense(st, dpool + 2 * (size_t)e.offset, qs, k, mc);
            }
            account(1.0);
            break;
        }
        case QLX_BARRIER:
            break;
        default:
            fail(QLX_EUNSUPPORTED, "CPUOpenMPEngine: Unrecognized instruction type");
    }
}

void execute_plain(qlx_state* st, const qlx_gate_instr* instrs, size_t n, const double* pool,
                   size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta,
                   const qlx_exec_opts& opts, qlx_exec_stats& stats) {
    bool need_pool = false;
    for (size_t i = 0; i < n; ++i)
        if (instrs[i].type == QLX_UNITARY || instrs[i].type == QLX_FUSED_BLOCK) need_pool = true;
    const double* dpool = need_pool ? upload_pool(st, pool, pool_elems) : nullptr;
    prepare_dense(st, instrs, n, dpool, meta, n_meta, opts);
    for (size_t i = 0; i < n; ++i) launch_plain_instr(st, instrs[i], dpool, meta, opts, stats);
}

}  // namespace qlx

using namespace qlx;

extern "C" {

int qlx_execute(qlx_state* st, const qlx_gate_instr* instrs, size_t n_instr, const double* pool,
                size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts* opts,
                size_t* stop_at) {
    return guard([&] {
        if (!st || !st->amp) fail(QLX_EINVAL, "null state handle");
        if (n_instr && !instrs) fail(QLX_EINVAL, "null instruction stream");
        qlx_exec_opts o{};
        if (opts) o = *opts;
        if (o.mode == QLX_EXEC_AUTO) {
            // window sweeps; large states also get specialised, relabelled sweeps compiled in the
            // background (the default geometry only: the generic kernel runs until they are ready)
            o.mode = QLX_EXEC_WINDOWED;
            if (st->nq >= QLX_AUTO_SPECIALIZE_QUBITS && !o.window_qubits && !o.register_qubits && !o.coalesce_qubits)
                o.specialize |= QLX_SPECIALIZE | QLX_SPECIALIZE_RELABEL | QLX_SPECIALIZE_ASYNC;
        }
        if (o.mode != QLX_EXEC_PLAIN && o.mode != QLX_EXEC_WINDOWED && o.mode != QLX_EXEC_DENSE_SWEEPS)
            fail(QLX_EINVAL, "unknown execution mode");
        if ((o.specialize & QLX_SPECIALIZE_ASYNC) && o.mode != QLX_EXEC_WINDOWED)
            fail(QLX_EINVAL, "QLX_SPECIALIZE_ASYNC applies to window sweeps");
        // The run ends at the first MEASURE (handled by the host) or the first
        // unknown opcode (executed prefix, then the reference's runtime_error).
        size_t stop = n_instr;
        for (size_t i = 0; i < n_instr; ++i)
            if (instrs[i].type == QLX_MEASURE) {
                stop = i;
                break;
            }
        const size_t bad = first_unknown(instrs, stop);
        const size_t run = std::min(stop, bad);
        for (size_t i = 0; i < run; ++i) validate_instr(instrs[i], st->nq, meta, n_meta, pool_elems);
        if (stop < n_instr && stop == run) validate_instr(instrs[stop], st->nq, meta, n_meta, pool_elems);
        qlx_exec_stats stats{};
        if (run > 0) {
            if (o.mode == QLX_EXEC_PLAIN)
                execute_plain(st, instrs, run, pool, pool_elems, meta, n_meta, o, stats);
            else if (o.mode == QLX_EXEC_DENSE_SWEEPS)
                execute_dense_sweeps(st, instrs, run, pool, pool_elems, meta, n_meta, o, stats);
            else
                execute_windowed(st, instrs, run, pool, pool_elems, meta, n_meta, o, stats);
        }
        st->stats = stats;
        if (bad < stop) {
            QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
            fail(QLX_EUNSUPPORTED, "CPUOpenMPEngine: Unrecognized instruction type");
        }
        if (stop_at) *stop_at = stop;
    });
}

}  // extern "C"

// ---- execution without device allocation ------------------------------------------------
// The matrices of UNITARY / FUSED_BLOCK gates are staged in `scratch`, a file-scope device
// array of kNoAllocPoolBytes per GPU (held until the process ends), instead of the state's
// growable program buffer. In host code `scratch` names a host-side stand-in, not a device
// address: host code uses only its device address, taken once per device with
// hipGetSymbolAddress. Every call on one GPU shares that GPU's array; calls that overlap on
// one GPU (states on different streams) race on it.
namespace qlx {
constexpr size_t kNoAllocPoolBytes = size_t(1) << 20;  // 65536 complex128 matrix elements
static __device__ __align__(256) unsigned char scratch[kNoAllocPoolBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU

void* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) fail(QLX_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        QLX_HIP_CHECK(hipGetDevice(&cur));
        if (cur != device) QLX_HIP_CHECK(hipSetDevice(device));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) QLX_HIP_CHECK(hipSetDevice(cur));
        if (e != hipSuccess || !ptr)
            fail(QLX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}
}  // namespace
}  // namespace qlx

extern "C" {

int qlx_execute_noalloc(qlx_state* st, const qlx_gate_instr* instrs, size_t n_instr, const double* pool,
                        size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts* opts,
                        size_t* stop_at) {
    return guard([&] {
        if (!st || !st->amp) fail(QLX_EINVAL, "null state handle");
        if (n_instr && !instrs) fail(QLX_EINVAL, "null instruction stream");
        qlx_exec_opts o{};
        if (opts) o = *opts;
        if (o.mode != QLX_EXEC_PLAIN)
            fail(QLX_EUNSUPPORTED, "qlx_execute_noalloc runs one kernel per gate (QLX_EXEC_PLAIN) only");
        if (st->timing) fail(QLX_EINVAL, "qlx_execute_noalloc does not time kernels: disable timing on the state");
        if (pool_elems > kNoAllocPoolBytes / 16)
            fail(QLX_EINVAL, "qlx_execute_noalloc: the unitary pool holds " + std::to_string(pool_elems) +
                                 " complex elements; at most " + std::to_string(kNoAllocPoolBytes / 16) +
                                 " fit the staging area");
        if (pool_elems && !pool) fail(QLX_EINVAL, "null unitary pool");
        size_t stop = n_instr;
        for (size_t i = 0; i < n_instr; ++i)
            if (instrs[i].type == QLX_MEASURE) {
                stop = i;
                break;
            }
        const size_t bad = first_unknown(instrs, stop);
        const size_t run = std::min(stop, bad);
        for (size_t i = 0; i < run; ++i) validate_instr(instrs[i], st->nq, meta, n_meta, pool_elems);
        if (stop < n_instr && stop == run) validate_instr(instrs[stop], st->nq, meta, n_meta, pool_elems);
        bool need_pool = false;
        for (size_t i = 0; i < run; ++i)
            if (instrs[i].type == QLX_UNITARY || instrs[i].type == QLX_FUSED_BLOCK) need_pool = true;
        const double* dpool = nullptr;
        if (need_pool && pool_elems) {
            void* dst = scratch_on(st->device);
            // same stream as the kernels below; the source is pageable caller memory, so the copy
            // completes before this returns
            QLX_HIP_CHECK(hipMemcpyAsync(dst, pool, pool_elems * 16, hipMemcpyHostToDevice, st->stream));
            QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
            dpool = (const double*)dst;
        }
        // dense gates on the vector units (the matrix-core forms need fragment buffers)
        st->frag_index.clear();
        o.dense_kernel = QLX_DENSE_VALU;
        qlx_exec_stats stats{};
        for (size_t i = 0; i < run; ++i) launch_plain_instr(st, instrs[i], dpool, meta, o, stats);
        st->stats = stats;
        if (bad < stop) {
            QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
            fail(QLX_EUNSUPPORTED, "CPUOpenMPEngine: Unrecognized instruction type");
        }
        if (stop_at) *stop_at = stop;
    });
}

int qlx_measure_prob1(qlx_state* st, uint32_t q, double* p1) {
    return guard([&] {
        if (!st || !st->amp) fail(QLX_EINVAL, "null state handle");
        if (q >= st->nq) fail(QLX_ERANGE, "Qubit index " + std::to_string(q) + " out of range");
        if (!p1) fail(QLX_EINVAL, "null output");
        const uint64_t half = 1ull << (st->nq - 1);
        uint64_t nb = (half + kThreads - 1) / kThreads;
        nb = std::min<uint64_t>(std::max<uint64_t>(nb, 1), kRedBlocks);
        st->scratch.reserve(nb * sizeof(double), st->stream);
        if (st->dtype == QLX_C128)
            k_prob1<double2><<<(unsigned)nb, kThreads, 0, st->stream>>>((const double2*)st->amp, half, q,
                                                                        (double*)st->scratch.ptr);
        else
            k_prob1<float2><<<(unsigned)nb, kThreads, 0, st->stream>>>((const float2*)st->amp, half, q,
                                                                       (double*)st->scratch.ptr);
        QLX_HIP_CHECK(hipGetLastError());
        std::vector<double> h(nb);
        QLX_HIP_CHECK(hipMemcpyAsync(h.data(), st->scratch.ptr, nb * sizeof(double), hipMemcpyDeviceToHost, st->stream));
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
        double s = 0.0;
        for (double v : h) s += v;
        *p1 = s;
    });
}

int qlx_collapse(qlx_state* st, uint32_t q, int outcome, double inv_norm) {
    return guard([&] {
        if (!st || !st->amp) fail(QLX_EINVAL, "null state handle");
        if (q >= st->nq) fail(QLX_ERANGE, "Qubit index " + std::to_string(q) + " out of range");
        if (outcome != 0 && outcome != 1) fail(QLX_EINVAL, "outcome must be 0 or 1");
        const uint64_t half = 1ull << (st->nq - 1);
        if (st->dtype == QLX_C128)
            k_collapse<double2><<<grid_for(half), kThreads, 0, st->stream>>>((double2*)st->amp, half, q, outcome,
                                                                            inv_norm);
        else
            k_collapse<float2><<<grid_for(half), kThreads, 0, st->stream>>>((float2*)st->amp, half, q, outcome,
                                                                           inv_norm);
        QLX_HIP_CHECK(hipGetLastError());
    });
}

}  // extern "C"
