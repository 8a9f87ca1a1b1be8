// qlx_mfma.hip — dense 4/5-qubit unitaries on complex64 states as matrix-core (MFMA)
// contractions on gfx950.
//
// The reference applies a UNITARY / FUSED_BLOCK gate group by group as a dense
// matrix-vector product (include/backend/cpu_openmp_engine.hpp:539-606: local index
// bit i <-> qubits[i], v_out = U * v_in). Here 32 such groups form the columns of one
// real GEMM per wave:
//
//     D[2^(K+1) x 32] = A[2^(K+1) x 2^(K+1)] * X[2^(K+1) x 32]
//
// with the real embedding A[2o+0][2l+0] = Re U[o][l], A[2o+0][2l+1] = -Im U[o][l],
// A[2o+1][2l+0] = Im U[o][l], A[2o+1][2l+1] = Re U[o][l], and X[2l+c][j] the real (c=0)
// and imaginary (c=1) part of amplitude l of group j. Two arithmetic forms:
//   * BF16: v_mfma_f32_32x32x16_bf16 on bf16 splits of both operands. The matrix,
//     split once per gate, keeps three terms (a = a_hi + a_mid + a_lo: 24 significand
//     bits, as FP32): its error is the same for every group, so a coarser matrix
//     would add up coherently over the gates and drift the norm (two terms: ~4e-7
//     per gate). The state keeps two (x = x_hi + x_lo: 16 bits), an error that differs
//     from amplitude to amplitude. Four products, small terms first, accumulated in
//     FP32: a_lo x_hi + a_mid x_hi + a_hi x_lo + a_hi x_hi (dropped: a_mid x_lo and
//     below, <= 2^-24 relative);
//   * F32:    v_mfma_f32_32x32x2_f32 on the FP32 operands directly.
// A is built once per gate from the FP64 matrix on the device (k_mfma_split) and held
// in VGPRs for the whole launch; X is read from HBM straight into MFMA operand
// registers and the result is written back in place (a wave reads its 32 groups
// completely before writing them; waves own disjoint groups).
//
// MFMA operand layout used (wave64; CDNA3/CDNA4 32x32 forms): lane l holds row/column
// l % 32 of A/B and the K-slice selected by l / 32; the 32x32 FP32 result sits in 16
// VGPRs with element (row 8*(v/4) + 4*(l/32) + v%4, column l%32) in VGPR v of lane l.
// Only the pairing of A and B slices matters for K, so the K order below is chosen for
// the loads: lane half h of column j holds amplitudes [h*2^(K-1), (h+1)*2^(K-1)) of
// group j, real and imaginary parts interleaved as in memory.

#include <algorithm>

#include "qlx_internal.hpp"
#include "qlx_mfma_core.hpp"

namespace qlx {
namespace {


constexpr int kWaves = 4;
constexpr int kBlock = 64 * kWaves;

__device__ inline uint64_t ins0(uint64_t idx, uint32_t t) {
    const uint64_t lo = idx & ((1ull << t) - 1);
    return lo | ((idx ^ lo) << 1);
}

// Real-embedded element (row, col) of the complex D x D row-major matrix U.
__device__ inline double embed(const double* __restrict__ U, int D, int row, int col) {
    const int o = row >> 1, l = col >> 1;
    const double ur = U[2 * (o * D + l)], ui = U[2 * (o * D + l) + 1];
    if ((row & 1) == 0) return (col & 1) ? -ui : ur;
    return (col & 1) ? ur : ui;
}

// BF16: the column of X (and of A) that slot s of K-step t of lane half h carries.
__host__ __device__ constexpr int kcol_bf16(int K, int t, int h, int s) {
    return 2 * ((h << (K - 1)) + 4 * t + (s >> 1)) + (s & 1);
}
// F32: K-step t of lane half h (one value per lane).
__host__ __device__ constexpr int kcol_f32(int K, int t, int h) {
    return 2 * ((h << (K - 1)) + (t >> 1)) + (t & 1);
}

struct SplitDesc {
    uint64_t pool_off;   // complex elements into the device pool
    uint64_t frag_off;   // bytes into the fragment buffer
    uint32_t k;          // 4 or 5
    uint32_t form;       // FORM_BF16 (hi/mid/lo), FORM_F32, FORM_F16 (hi/lo)
};

// One work-group per matrix: writes the A-operand fragments of every (M-tile, K-step, lane).
__global__ void __launch_bounds__(kBlock) k_mfma_split(const double* __restrict__ pool,
                                                      const SplitDesc* __restrict__ descs,
                                                      unsigned char* __restrict__ frags) {
    const SplitDesc d = descs[blockIdx.x];
    const int K = (int)d.k, D = 1 << K;
    const double* U = pool + 2 * d.pool_off;
    const int MT = (2 * D) / 32;
    if (d.form == FORM_F32) {
        const int T = D;  // 2D / 2 K-steps
        float* out = (float*)(frags + d.frag_off);
        for (int e = threadIdx.x; e < MT * T * 64; e += kBlock) {
            const int lane = e & 63, t = (e >> 6) % T, m = (e >> 6) / T;
            out[e] = (float)embed(U, D, 32 * m + (lane & 31), kcol_f32(K, t, lane >> 5));
        }
    } else if (d.form == FORM_F16) {
        const int T = (2 * D) / 16;
        f16x8* hi = (f16x8*)(frags + d.frag_off);
        f16x8* lo = hi + MT * T * 64;
        for (int e = threadIdx.x; e < MT * T * 64; e += kBlock) {
            const int lane = e & 63, t = (e >> 6) % T, m = (e >> 6) / T;
            f16x8 vh, vl;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const double v = embed(U, D, 32 * m + (lane & 31), kcol_bf16(K, t, lane >> 5, s));
                const _Float16 h = (_Float16)(float)v;
                vh[s] = h;
                vl[s] = (_Float16)(float)(v - (double)(float)h);   // the residual is exact in FP64
            }
            hi[e] = vh;
            lo[e] = vl;
        }
    } else {
        const int T = (2 * D) / 16;
        bf16x8* hi = (bf16x8*)(frags + d.frag_off);
        bf16x8* mid = hi + MT * T * 64;
        bf16x8* lo = mid + MT * T * 64;
        for (int e = threadIdx.x; e < MT * T * 64; e += kBlock) {
            const int lane = e & 63, t = (e >> 6) % T, m = (e >> 6) / T;
            bf16x8 vh, vm, vl;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const double v = embed(U, D, 32 * m + (lane & 31), kcol_bf16(K, t, lane >> 5, s));
                const __bf16 h = (__bf16)(float)v;
                const double r1 = v - (double)(float)h;       // exact
                const __bf16 md = (__bf16)(float)r1;
                const double r2 = r1 - (double)(float)md;     // exact
                vh[s] = h;
                vm[s] = md;
                vl[s] = (__bf16)(float)r2;
            }
            hi[e] = vh;
            mid[e] = vm;
            lo[e] = vl;
        }
    }
}

struct MfmaArgs {
    uint32_t sorted[5];   // target qubits ascending (zero-hole insertion order)
    uint32_t in_off[16];  // offset of local amplitude r < 2^(K-1) (bit K-1 of r clear)
    uint32_t out_off[16]; // offset of output slot (m, a, b): local 16m + 4a + b (bit 1 clear)
    uint32_t hbit_in;     // 1 << qubits[K-1]: lane half 1 reads the upper half of a group
    uint32_t hbit_out;    // 1 << qubits[1]:   lane half 1 writes locals with bit 1 set
};

// ALDS (FP32 form only): the A fragments stay in LDS (read per tile) instead of VGPRs: fewer
// registers per wave, more waves per SIMD and more HBM loads in flight.
template <int K, int FORM, bool ALDS>
__global__ void __launch_bounds__(kBlock) k_dense_mfma(float2* __restrict__ amp, uint64_t ngroups,
                                                      const unsigned char* __restrict__ frags, MfmaArgs ma) {
    constexpr int D = 1 << K;          // amplitudes per group
    constexpr int HALF = D / 2;        // amplitudes per lane
    constexpr int MT = (2 * D) / 32;   // 32-row tiles of the real embedding
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5, j = lane & 31;
    const uint64_t hin = h ? ma.hbit_in : 0u, hout = h ? ma.hbit_out : 0u;

    constexpr bool F32 = FORM == FORM_F32, BF16 = FORM == FORM_BF16, F16 = FORM == FORM_F16;
    static_assert(!ALDS || F32, "LDS-resident fragments are used in the FP32 form only");
    constexpr int TB = (2 * D) / 16;   // 16-bit K-steps
    constexpr int TF = D;              // F32 K-steps
    constexpr int NB = MT * TB * 64;   // 8 x 16-bit fragments per split term
    constexpr int NF = MT * TF * 64;   // float fragments
    constexpr int RA = ALDS ? 1 : MT;  // register copies
    bf16x8 ah[BF16 ? MT : 1][BF16 ? TB : 1], am[BF16 ? MT : 1][BF16 ? TB : 1], al[BF16 ? MT : 1][BF16 ? TB : 1];
    f16x8 hh[F16 ? MT : 1][F16 ? TB : 1], hl[F16 ? MT : 1][F16 ? TB : 1];
    float af[F32 ? RA : 1][F32 ? TF : 1];
    __shared__ __align__(16) unsigned char sA[ALDS ? NF * 4 : 16];
    if constexpr (ALDS) {
        const float* f = (const float*)frags;
        float* d = (float*)sA;
        for (int e = threadIdx.x; e < NF; e += kBlock) d[e] = f[e];
        __syncthreads();
    } else if constexpr (F32) {
        const float* f = (const float*)frags;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < TF; ++t) af[m][t] = f[(m * TF + t) * 64 + lane];
    } else if constexpr (BF16) {
        const bf16x8* fh = (const bf16x8*)frags;
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
        const f16x8* fh = (const f16x8*)frags;
        const f16x8* fl = fh + NB;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                hh[m][t] = fh[(m * TB + t) * 64 + lane];
                hl[m][t] = fl[(m * TB + t) * 64 + lane];
            }
    }
    auto fragf = [&](int m, int t) -> float {
        if constexpr (ALDS) return ((const float*)sA)[(m * TF + t) * 64 + lane];
        else return af[m][t];
    };

    const uint64_t ntiles = ngroups / 32;
    const uint64_t wave0 = (uint64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * kWaves;
    if (wave0 >= ntiles) return;  // whole wave
    // Two register sets alternate: the next tile's amplitudes are loaded into one while the
    // tile in the other is multiplied and stored (a wave owns its tiles, so the reads never
    // see its own writes). Loads and stores are unconditional (whole tiles; a wave with no
    // next tile re-reads its current one and ignores it), so the wait before a tile covers
    // that tile's loads only.
    auto group_base = [&](uint64_t tile) {
        uint64_t base = tile * 32 + j;
#pragma unroll
        for (int q = 0; q < K; ++q) base = ins0(base, ma.sorted[q]);
        return base;
    };
    auto load = [&](float2 (&x)[HALF], uint64_t base) {
#pragma unroll
        for (int r = 0; r < HALF; ++r) x[r] = amp[base + hin + ma.in_off[r]];
        // keep the loads here: the scheduler would otherwise sink them below the MFMAs
        __builtin_amdgcn_sched_barrier(0);
    };
    auto multiply_store = [&](const float2 (&x)[HALF], uint64_t base) {
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
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x2f32(fragf(m, t), xv, acc[m], 0, 0, 0);
            }
        } else if constexpr (BF16) {
            bf16x8 xh[TB], xl[TB];
#pragma unroll
            for (int t = 0; t < TB; ++t) split_bf16(x + 4 * t, xh[t], xl[t]);
            // small terms first
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
            f16x8 xh[TB], xl[TB];
#pragma unroll
            for (int t = 0; t < TB; ++t) split_f16(x + 4 * t, kF16Scale, xh[t], xl[t]);
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
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m] *= kF16Unscale;
        }
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    amp[base + hout + ma.out_off[m * 8 + a * 2 + b]] =
                        float2{acc[m][4 * a + 2 * b], acc[m][4 * a + 2 * b + 1]};
    };
    float2 xa[HALF], xb[HALF];
    uint64_t tile = wave0;
    uint64_t ba = group_base(tile);
    load(xa, ba);
    for (;;) {
        const uint64_t t1 = tile + nwaves;
        const uint64_t bb = group_base(t1 < ntiles ? t1 : tile);
        load(xb, bb);
        multiply_store(xa, ba);
        if (t1 >= ntiles) break;
        const uint64_t t2 = t1 + nwaves;
        ba = group_base(t2 < ntiles ? t2 : t1);
        load(xa, ba);
        multiply_store(xb, bb);
        if (t2 >= ntiles) break;
        tile = t2;
    }
}

int form_of(int variant) {
    return variant == QLX_DENSE_MFMA_F32 ? FORM_F32 : variant == QLX_DENSE_MFMA_F16 ? FORM_F16 : FORM_BF16;
}

size_t frag_bytes(int k, int form) {
    const size_t D = size_t{1} << k, MT = 2 * D / 32;
    if (form == FORM_F32) return MT * D * 64 * sizeof(float);
    return (form == FORM_F16 ? 2 : 3) * MT * (2 * D / 16) * 64 * 16;
}

}  // namespace

void mfma_f16_fragments_host(const double* U, int k, unsigned char* out) {
    if (k != 4 && k != 5) fail(QLX_EUNSUPPORTED, "MFMA dense gates take 4 or 5 qubits");
    const int D = 1 << k, MT = (2 * D) / 32, T = (2 * D) / 16, NB = MT * T * 64;
    auto embed_h = [&](int row, int col) {
        const int o = row >> 1, l = col >> 1;
        const double ur = U[2 * (o * D + l)], ui = U[2 * (o * D + l) + 1];
        if ((row & 1) == 0) return (col & 1) ? -ui : ur;
        return (col & 1) ? ur : ui;
    };
    _Float16* hi = reinterpret_cast<_Float16*>(out);
    _Float16* lo = hi + 8 * NB;
    for (int e = 0; e < NB; ++e) {
        const int lane = e & 63, t = (e >> 6) % T, m = (e >> 6) / T;
        for (int s = 0; s < 8; ++s) {
            const double v = embed_h(32 * m + (lane & 31), kcol_bf16(k, t, lane >> 5, s));
            const _Float16 h = (_Float16)(float)v;
            hi[8 * e + s] = h;
            lo[8 * e + s] = (_Float16)(float)(v - (double)(float)h);
        }
    }
}

size_t mfma_fragment_bytes(int k, int variant) {
    if (k != 4 && k != 5) fail(QLX_EUNSUPPORTED, "MFMA dense gates take 4 or 5 qubits");
    return frag_bytes(k, form_of(variant));
}

void mfma_split(const double* d_pool, const MfmaSplitJob* jobs, size_t n, void* d_frags, DevBuf& desc_buf,
                hipStream_t stream) {
    if (!n) return;
    std::vector<SplitDesc> h(n);
    for (size_t i = 0; i < n; ++i) {
        if (jobs[i].k != 4 && jobs[i].k != 5) fail(QLX_EUNSUPPORTED, "MFMA dense gates take 4 or 5 qubits");
        h[i] = SplitDesc{jobs[i].pool_off, jobs[i].frag_off, (uint32_t)jobs[i].k, (uint32_t)form_of(jobs[i].variant)};
    }
    desc_buf.reserve(n * sizeof(SplitDesc), stream);
    QLX_HIP_CHECK(hipMemcpyAsync(desc_buf.ptr, h.data(), n * sizeof(SplitDesc), hipMemcpyHostToDevice, stream));
    k_mfma_split<<<(unsigned)n, kBlock, 0, stream>>>(d_pool, (const SplitDesc*)desc_buf.ptr,
                                                     (unsigned char*)d_frags);
    QLX_HIP_CHECK(hipGetLastError());
    // the descriptors live in pageable host memory
    QLX_HIP_CHECK(hipStreamSynchronize(stream));
}

void launch_dense_mfma(qlx_state* st, const void* d_frag, const uint8_t* qubits, int k, int variant) {
    if (st->dtype != QLX_C64) fail(QLX_EUNSUPPORTED, "MFMA dense gates run on complex64 states");
    if (k != 4 && k != 5) fail(QLX_EUNSUPPORTED, "MFMA dense gates take 4 or 5 qubits");
    if ((uint32_t)k + 5 > st->nq) fail(QLX_EUNSUPPORTED, "MFMA dense gates need at least 32 groups (k + 5 qubits)");
    MfmaArgs ma{};
    auto off = [&](int loc) {
        uint32_t o = 0;
        for (int i = 0; i < k; ++i)
            if ((loc >> i) & 1) o |= 1u << qubits[i];
        return o;
    };
    const int half = 1 << (k - 1);
    for (int r = 0; r < half; ++r) ma.in_off[r] = off(r);
    const int mt = (1 << (k + 1)) / 32;
    for (int m = 0; m < mt; ++m)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 2; ++b) ma.out_off[m * 8 + a * 2 + b] = off(16 * m + 4 * a + b);
    ma.hbit_in = 1u << qubits[k - 1];
    ma.hbit_out = 1u << qubits[1];
    uint32_t s[5];
    for (int i = 0; i < k; ++i) s[i] = qubits[i];
    std::sort(s, s + k);
    for (int i = 0; i < k; ++i) ma.sorted[i] = s[i];

    const uint64_t ngroups = 1ull << (st->nq - k);
    const uint64_t tiles = (ngroups + 31) / 32;
    // enough resident waves to keep HBM busy; every wave loops over its tiles
    const uint64_t want = (tiles + kWaves - 1) / kWaves;
    const unsigned grid = (unsigned)std::min<uint64_t>(want, 256ull * 8);
    float2* amp = (float2*)st->amp;
    const auto* f = (const unsigned char*)d_frag;
    // measured on MI355X (32-qubit complex64, 5-qubit gates): the FP32 form gains from the
    // extra occupancy of LDS-resident fragments (18.2 -> 15.5 ms per gate), the BF16 form does
    // not (14.6 ms either way) and keeps them in registers
    switch (form_of(variant)) {
        case FORM_F32:
            if (k == 5) k_dense_mfma<5, FORM_F32, true><<<grid, kBlock, 0, st->stream>>>(amp, ngroups, f, ma);
            else k_dense_mfma<4, FORM_F32, true><<<grid, kBlock, 0, st->stream>>>(amp, ngroups, f, ma);
            break;
        case FORM_F16:
            if (k == 5) k_dense_mfma<5, FORM_F16, false><<<grid, kBlock, 0, st->stream>>>(amp, ngroups, f, ma);
            else k_dense_mfma<4, FORM_F16, false><<<grid, kBlock, 0, st->stream>>>(amp, ngroups, f, ma);
            break;
        default:
            if (k == 5) k_dense_mfma<5, FORM_BF16, false><<<grid, kBlock, 0, st->stream>>>(amp, ngroups, f, ma);
            else k_dense_mfma<4, FORM_BF16, false><<<grid, kBlock, 0, st->stream>>>(amp, ngroups, f, ma);
    }
    QLX_HIP_CHECK(hipGetLastError());
}

void prepare_dense(qlx_state* st, const qlx_gate_instr* instrs, size_t n, const double* dpool,
                   const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts& opts) {
    st->frag_index.clear();
    const uint32_t dk = opts.dense_kernel;
    if (dk > QLX_DENSE_MFMA_F16) fail(QLX_EINVAL, "unknown dense_kernel");
    if (st->dtype != QLX_C64 || dk == QLX_DENSE_VALU || !dpool || !meta) return;
    const int variant = dk == QLX_DENSE_AUTO ? QLX_DENSE_MFMA_F16 : (int)dk;
    std::vector<MfmaSplitJob> jobs;
    size_t total = 0;
    for (size_t i = 0; i < n; ++i) {
        const qlx_gate_instr& g = instrs[i];
        if (g.type != QLX_UNITARY && g.type != QLX_FUSED_BLOCK) continue;
        const int k = instr_arity(g);
        const uint32_t mi = g.payload.matrix_idx;
        // states with fewer than 32 groups of 2^k amplitudes keep the vector-unit kernel
        if ((k != 4 && k != 5) || (uint32_t)k + 5 > st->nq || mi >= n_meta || meta[mi].num_qubits != k) continue;
        if (st->frag_index.empty()) st->frag_index.assign(n_meta, -1);
        if (st->frag_index[mi] >= 0) continue;
        st->frag_index[mi] = (int64_t)total;
        jobs.push_back(MfmaSplitJob{meta[mi].offset, total, k, variant});
        total += (mfma_fragment_bytes(k, variant) + 255) & ~size_t{255};
    }
    if (jobs.empty()) return;
    st->frags.reserve(total, st->stream);
    mfma_split(dpool, jobs.data(), jobs.size(), st->frags.ptr, st->frag_desc, st->stream);
    st->frag_variant = variant;
}

}  // namespace qlx
