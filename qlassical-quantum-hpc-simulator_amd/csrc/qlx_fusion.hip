// qlx_fusion.hip — FP64 construction of fused-block matrices on the GPU.
//
// pass_gate_fusion multiplies the member gates of a block into one matrix over the
// sorted union of their qubits (include/middleware/middleware.hpp:636-662, with
// FusionMath::expand_matrix / fuse_matrices :241-344: little-endian scatter expansion,
// later gate on the left, M = G_n ... G_2 G_1). The host transpiler decides the blocks;
// this file forms their products on the device: one work-group per block keeps the
// 2^k x 2^k complex128 matrix in LDS and applies the members in program order, each
// element of the new matrix being the sum over the member's own 2^a rows
//     M'[r][c] = sum_m G[r|a][m] * M[(r with member bits replaced by m)][c].

#include <algorithm>

#include "qlx_internal.hpp"

namespace qlx {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxD = 32;

struct FuseBlockDev {
    uint32_t member_begin;
    uint32_t n_members;
    uint32_t k;
    uint32_t pad;
    uint64_t out_off;   // complex elements into the output array
};

struct FuseMemberDev {
    uint32_t mat_off;   // complex elements into the member-matrix array
    uint32_t a;         // member arity (matrix is 2^a x 2^a, row-major, bit i <-> operand i)
    uint32_t pos[5];    // block-local bit of operand i
    uint32_t pad;
};

__global__ void __launch_bounds__(kThreads) k_fuse_blocks(const FuseBlockDev* __restrict__ blocks,
                                                         const FuseMemberDev* __restrict__ members,
                                                         const double2* __restrict__ mats,
                                                         double2* __restrict__ out) {
    __shared__ double2 M[2][kMaxD * kMaxD];
    __shared__ double2 G[kMaxD * kMaxD];
    const FuseBlockDev b = blocks[blockIdx.x];
    const int d = 1 << b.k, dd = d * d;
    for (int e = threadIdx.x; e < dd; e += kThreads)
        M[0][e] = make_double2((e / d) == (e % d) ? 1.0 : 0.0, 0.0);
    int cur = 0;
    for (uint32_t mi = 0; mi < b.n_members; ++mi) {
        const FuseMemberDev mem = members[b.member_begin + mi];
        const int da = 1 << mem.a;
        __syncthreads();  // previous product complete, G free
        for (int e = threadIdx.x; e < da * da; e += kThreads) G[e] = mats[mem.mat_off + e];
        // operands beyond the arity get an empty bit (bit 31 of an index < 2^10)
        uint32_t p[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) p[i] = (uint32_t)i < mem.a ? mem.pos[i] : 31u;
        uint32_t mask = 0;
#pragma unroll
        for (int i = 0; i < 5; ++i) mask |= (uint32_t)i < mem.a ? 1u << p[i] : 0u;
        __syncthreads();
        const double2* src = M[cur];
        double2* dst = M[cur ^ 1];
        for (int e = threadIdx.x; e < dd; e += kThreads) {
            const uint32_t r = (uint32_t)(e / d), c = (uint32_t)(e % d);
            uint32_t rl = 0;
#pragma unroll
            for (int i = 0; i < 5; ++i) rl |= ((r >> p[i]) & 1u) << i;
            const uint32_t rbase = r & ~mask;
            double sr = 0.0, si = 0.0;
            for (int ml = 0; ml < da; ++ml) {
                uint32_t m = rbase;
#pragma unroll
                for (int i = 0; i < 5; ++i) m |= (uint32_t)i < mem.a ? (((uint32_t)ml >> i) & 1u) << p[i] : 0u;
                const double2 g = G[rl * da + ml];
                const double2 v = src[m * d + c];
                sr += g.x * v.x - g.y * v.y;
                si += g.x * v.y + g.y * v.x;
            }
            dst[e] = make_double2(sr, si);
        }
        cur ^= 1;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < dd; e += kThreads) out[b.out_off + e] = M[cur][e];
}

}  // namespace

void build_fused_on_device(const std::vector<FusedRecipe>& recipes, const std::vector<double>& member_mats,
                           std::vector<double>& out) {
    if (recipes.empty()) return;
    std::vector<FuseBlockDev> hb;
    std::vector<FuseMemberDev> hm;
    uint64_t out_elems = 0;
    for (const FusedRecipe& r : recipes) {
        if (r.k < 1 || r.k > 5) fail(QLX_EINVAL, "fused block width must be 1..5");
        FuseBlockDev b{};
        b.member_begin = (uint32_t)hm.size();
        b.n_members = (uint32_t)r.members.size();
        b.k = (uint32_t)r.k;
        b.out_off = out_elems;
        out_elems += uint64_t{1} << (2 * r.k);
        for (const FusedMember& m : r.members) {
            if (m.a < 1 || m.a > r.k) fail(QLX_EINVAL, "member arity exceeds its block");
            FuseMemberDev d{};
            d.mat_off = m.mat_off;
            d.a = (uint32_t)m.a;
            for (int i = 0; i < m.a; ++i) {
                if (m.pos[i] < 0 || m.pos[i] >= r.k) fail(QLX_EINVAL, "member operand outside its block");
                d.pos[i] = (uint32_t)m.pos[i];
            }
            hm.push_back(d);
        }
        hb.push_back(b);
    }
    hipStream_t s = nullptr;
    QLX_HIP_CHECK(hipStreamCreate(&s));
    void *db = nullptr, *dm = nullptr, *dx = nullptr, *dout = nullptr;
    auto cleanup = [&] {
        if (s) (void)hipStreamSynchronize(s);
        for (void* p : {db, dm, dx, dout})
            if (p) (void)hipFree(p);
        if (s) (void)hipStreamDestroy(s);
    };
    try {
        const size_t mats_bytes = member_mats.size() * sizeof(double);
        QLX_HIP_CHECK(hipMalloc(&db, hb.size() * sizeof(FuseBlockDev)));
        QLX_HIP_CHECK(hipMalloc(&dm, std::max<size_t>(hm.size(), 1) * sizeof(FuseMemberDev)));
        QLX_HIP_CHECK(hipMalloc(&dx, std::max<size_t>(mats_bytes, 16)));
        QLX_HIP_CHECK(hipMalloc(&dout, out_elems * 16));
        QLX_HIP_CHECK(hipMemcpyAsync(db, hb.data(), hb.size() * sizeof(FuseBlockDev), hipMemcpyHostToDevice, s));
        if (!hm.empty())
            QLX_HIP_CHECK(hipMemcpyAsync(dm, hm.data(), hm.size() * sizeof(FuseMemberDev), hipMemcpyHostToDevice, s));
        if (mats_bytes) QLX_HIP_CHECK(hipMemcpyAsync(dx, member_mats.data(), mats_bytes, hipMemcpyHostToDevice, s));
        k_fuse_blocks<<<(unsigned)hb.size(), kThreads, 0, s>>>((const FuseBlockDev*)db, (const FuseMemberDev*)dm,
                                                               (const double2*)dx, (double2*)dout);
        QLX_HIP_CHECK(hipGetLastError());
        out.resize(2 * out_elems);
        QLX_HIP_CHECK(hipMemcpyAsync(out.data(), dout, out_elems * 16, hipMemcpyDeviceToHost, s));
        QLX_HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

}  // namespace qlx
