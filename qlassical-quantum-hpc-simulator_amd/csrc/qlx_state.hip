// qlx_state.hip — GPUStateVector storage, initialisation and FP64 reductions.
//
// Reference behaviour restated here:
//   initialize(n)    include/backend/cpu_state_vector.hpp:54-78   (|0...0>)
//   initialize(span) include/backend/cpu_state_vector.hpp:82-115  (pow2, norm 1e-6, copy)
// The state lives in HBM as 2^N interleaved amplitudes (double2 for c128,
// float2 for c64), one contiguous allocation, one HIP stream per state.

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "qlx_internal.hpp"

namespace qlx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void DevBuf::reserve(size_t bytes, hipStream_t stream) {
    if (bytes <= cap) return;
    if (ptr) {
        QLX_HIP_CHECK(hipStreamSynchronize(stream));
        QLX_HIP_CHECK(hipFree(ptr));
        ptr = nullptr;
        cap = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    if (hipMalloc(&ptr, want) != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        fail(QLX_ENOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
    }
    cap = want;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

// ---- kernel timing -----------------------------------------------------------------
static hipEvent_t take_event(qlx_state* st) {
    if (!st->free_events.empty()) {
        hipEvent_t e = st->free_events.back();
        st->free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    QLX_HIP_CHECK(hipEventCreate(&e));
    return e;
}

TimedLaunch::TimedLaunch(qlx_state* st, const char* kernel, double bytes) : st_(st) {
    if (!st_->timing) return;
    rec_.kernel = kernel;
    rec_.bytes = bytes;
    rec_.e0 = take_event(st_);
    rec_.e1 = take_event(st_);
    QLX_HIP_CHECK(hipEventRecord(rec_.e0, st_->stream));
}

TimedLaunch::~TimedLaunch() noexcept(false) {
    if (!st_->timing || !rec_.e0) return;
    QLX_HIP_CHECK(hipEventRecord(rec_.e1, st_->stream));
    st_->pending.push_back(rec_);
}

static void drain_timing(qlx_state* st) {
    QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
    for (const TimingRec& r : st->pending) {
        float ms = 0.f;
        QLX_HIP_CHECK(hipEventElapsedTime(&ms, r.e0, r.e1));
        TimingAgg* agg = nullptr;
        for (auto& kv : st->timing_agg)
            if (kv.first == r.kernel) agg = &kv.second;
        if (!agg) {
            st->timing_agg.emplace_back(r.kernel, TimingAgg{});
            agg = &st->timing_agg.back().second;
        }
        agg->launches++;
        agg->ms += ms;
        agg->bytes += r.bytes;
        st->free_events.push_back(r.e0);
        st->free_events.push_back(r.e1);
    }
    st->pending.clear();
}

// ---- kernels -------------------------------------------------------------------

// |0...0>: amp[0] = 1, everything else 0 (vectorised 16-byte stores).
__global__ void k_fill_basis0(uint4* __restrict__ p, uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}
// Plain streaming copy (16-byte vectors); also the reference point for HBM counters.
// Whole 64 KiB chunks: one work-group per chunk, each work-item issues its 16 loads before
// its 16 stores (the access shape of the window sweeps). Other sizes: grid-stride.
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}
constexpr uint64_t kCopyChunk16 = 4096;  // 16-byte vectors per chunk
__global__ void __launch_bounds__(256) k_copy_chunks(const uint4* __restrict__ src, uint4* __restrict__ dst) {
    const uint64_t base = (uint64_t)blockIdx.x * kCopyChunk16 + threadIdx.x;
    uint4 v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = src[base + 256u * r];
#pragma unroll
    for (int r = 0; r < 16; ++r) dst[base + 256u * r] = v[r];
}

__global__ void k_set_first_c128(double2* p) { p[0] = make_double2(1.0, 0.0); }
__global__ void k_set_first_c64(float2* p) { p[0] = make_float2(1.0f, 0.0f); }

__global__ void k_narrow_c128_to_c64(const double2* __restrict__ src, float2* __restrict__ dst,
                                     uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double2 v = src[i];
        dst[i] = make_float2((float)v.x, (float)v.y);
    }
}
__global__ void k_widen_c64_to_c128(const float2* __restrict__ src, double2* __restrict__ dst,
                                    uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float2 v = src[i];
        dst[i] = make_double2((double)v.x, (double)v.y);
    }
}

template <class A>
__device__ inline double2 ld_amp(const A* p, uint64_t i);
template <>
__device__ inline double2 ld_amp<double2>(const double2* p, uint64_t i) { return p[i]; }
template <>
__device__ inline double2 ld_amp<float2>(const float2* p, uint64_t i) {
    float2 v = p[i];
    return make_double2((double)v.x, (double)v.y);
}

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 2048;

__device__ inline double block_sum(double v, double* sh) {
    // wave64 reduction, then across the 4 waves of the block
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kRedThreads / 64; ++w) t += sh[w];
    }
    return t;
}

// Partial sums of |c|^2 (FP64) per block; the host adds kRedBlocks partials.
template <class A>
__global__ void __launch_bounds__(kRedThreads) k_norm_partial(const A* __restrict__ p, uint64_t n,
                                                             double* __restrict__ part) {
    __shared__ double sh[kRedThreads / 64];
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double2 v = ld_amp(p, i);
        acc = fma(v.x, v.x, fma(v.y, v.y, acc));
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Partial sums of conj(a_i) * b_i (FP64).
template <class A, class B>
__global__ void __launch_bounds__(kRedThreads) k_inner_partial(const A* __restrict__ a,
                                                              const B* __restrict__ b, uint64_t n,
                                                              double* __restrict__ part) {
    __shared__ double sh[kRedThreads / 64];
    double re = 0.0, im = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double2 x = ld_amp(a, i), y = ld_amp(b, i);
        re = fma(x.x, y.x, fma(x.y, y.y, re));
        im = fma(x.x, y.y, fma(-x.y, y.x, im));
    }
    double t = block_sum(re, sh);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = t;
    __syncthreads();
    t = block_sum(im, sh);
    if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = t;
}

template <class A, class B>
__global__ void __launch_bounds__(kRedThreads) k_maxdiff_partial(const A* __restrict__ a,
                                                                const B* __restrict__ b, uint64_t n,
                                                                double* __restrict__ part) {
    __shared__ double sh[kRedThreads / 64];
    double m = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double2 x = ld_amp(a, i), y = ld_amp(b, i);
        double dr = x.x - y.x, di = x.y - y.y;
        m = fmax(m, sqrt(dr * dr + di * di));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kRedThreads / 64; ++w) t = fmax(t, sh[w]);
        part[blockIdx.x] = t;
    }
}

static unsigned red_blocks(uint64_t n) {
    uint64_t b = (n + kRedThreads - 1) / kRedThreads;
    if (b > (uint64_t)kRedBlocks) b = kRedBlocks;
    if (b == 0) b = 1;
    return (unsigned)b;
}

static unsigned fill_blocks(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return (unsigned)b;
}

void launch_fill_zero_state(qlx_state* st) {
    const uint64_t n16 = st->bytes / 16;
    if (n16 > 0) {
        k_fill_basis0<<<fill_blocks(n16), 256, 0, st->stream>>>((uint4*)st->amp, n16);
    }
    if (st->dtype == QLX_C128) {
        k_set_first_c128<<<1, 1, 0, st->stream>>>((double2*)st->amp);
    } else {
        // (a single c64 amplitude, nq = 0, is 8 bytes: fully written here)
        k_set_first_c64<<<1, 1, 0, st->stream>>>((float2*)st->amp);
    }
    QLX_HIP_CHECK(hipGetLastError());
}

static double sum_partials(qlx_state* st, unsigned nb, int per_block) {
    const size_t bytes = (size_t)nb * per_block * sizeof(double);
    std::vector<double> h(nb * per_block);
    QLX_HIP_CHECK(hipMemcpyAsync(h.data(), st->scratch.ptr, bytes, hipMemcpyDeviceToHost, st->stream));
    QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
    double s = 0.0;
    for (unsigned i = 0; i < nb; ++i) s += h[(size_t)i * per_block];
    return s;
}

double reduce_norm_sq(qlx_state* st) {
    const uint64_t n = 1ull << st->nq;
    const unsigned nb = red_blocks(n);
    st->scratch.reserve((size_t)nb * 2 * sizeof(double), st->stream);
    if (st->dtype == QLX_C128)
        k_norm_partial<<<nb, kRedThreads, 0, st->stream>>>((const double2*)st->amp, n, (double*)st->scratch.ptr);
    else
        k_norm_partial<<<nb, kRedThreads, 0, st->stream>>>((const float2*)st->amp, n, (double*)st->scratch.ptr);
    QLX_HIP_CHECK(hipGetLastError());
    return sum_partials(st, nb, 1);
}

static void check_state(const qlx_state* st) {
    if (!st) fail(QLX_EINVAL, "null state handle");
    if (!st->amp) fail(QLX_EINVAL, "state has no storage");
}

static void alloc_storage(qlx_state* st, uint32_t nq, int dtype) {
    if (dtype != QLX_C128 && dtype != QLX_C64) fail(QLX_EINVAL, "dtype must be QLX_C128 or QLX_C64");
    if (nq > 40) fail(QLX_ENOMEM, "qubit count " + std::to_string(nq) + " exceeds device memory");
    const size_t bytes = (size_t(1) << nq) * amp_size(dtype);
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        fail(QLX_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes for a " +
                             std::to_string(nq) + "-qubit state failed");
    }
    if (st->amp) QLX_HIP_CHECK(hipFree(st->amp));
    if (st->alt) QLX_HIP_CHECK(hipFree(st->alt));
    st->alt = nullptr;
    st->amp = p;
    st->bytes = bytes;
    st->nq = nq;
    st->dtype = dtype;
}

static qlx_state* new_state() {
    auto* st = new qlx_state();
    QLX_HIP_CHECK(hipGetDevice(&st->device));
    QLX_HIP_CHECK(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
    return st;
}

static void destroy_state(qlx_state* st) {
    if (!st) return;
    if (st->stream) (void)hipStreamSynchronize(st->stream);
    if (st->amp) (void)hipFree(st->amp);
    if (st->alt) (void)hipFree(st->alt);
    st->scratch.release();
    st->prog.release();
    st->frags.release();
    st->frag_desc.release();
    for (auto& r : st->pending) {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    for (hipEvent_t e : st->free_events) (void)hipEventDestroy(e);
    if (st->stream) (void)hipStreamDestroy(st->stream);
    delete st;
}

// Validates a host complex128 array as in CPUStateVector::initialize(span)
// (include/backend/cpu_state_vector.hpp:86-106) and returns N.
static uint32_t validate_host_state(const double* amps, size_t dim) {
    if (dim == 0 || (dim & (dim - 1)) != 0)
        fail(QLX_EINVAL, "GPUStateVector: Input state dimension must be a power of 2.");
    if (!amps) fail(QLX_EINVAL, "null amplitude pointer");
    double norm_sq = 0.0;
    for (size_t i = 0; i < dim; ++i) norm_sq += amps[2 * i] * amps[2 * i] + amps[2 * i + 1] * amps[2 * i + 1];
    if (std::abs(norm_sq - 1.0) > 1e-6)
        fail(QLX_EINVAL, "GPUStateVector: Provided state vector is not normalized (norm squared != 1.0).");
    uint32_t nq = 0;
    while ((size_t(1) << nq) < dim) ++nq;
    return nq;
}

static void upload_host_state(qlx_state* st, const double* amps, size_t dim) {
    if (st->dtype == QLX_C128) {
        QLX_HIP_CHECK(hipMemcpyAsync(st->amp, amps, dim * 16, hipMemcpyHostToDevice, st->stream));
    } else {
        // stage complex128 on the device in slices, then narrow to complex64
        const size_t slice = size_t(1) << 24;
        DevBuf tmp;
        tmp.reserve(std::min(dim, slice) * 16, st->stream);
        for (size_t off = 0; off < dim; off += slice) {
            size_t cnt = std::min(slice, dim - off);
            QLX_HIP_CHECK(hipMemcpyAsync(tmp.ptr, amps + 2 * off, cnt * 16, hipMemcpyHostToDevice, st->stream));
            k_narrow_c128_to_c64<<<fill_blocks(cnt), 256, 0, st->stream>>>((const double2*)tmp.ptr,
                                                                           (float2*)st->amp + off, cnt);
            QLX_HIP_CHECK(hipGetLastError());
        }
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
        tmp.release();
    }
    QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
}

void compare_amplitudes(const void* a, int dta, const void* b, int dtb, uint64_t n, hipStream_t stream,
                        DevBuf& scratch, double* re, double* im, double* maxd, double* norm_b) {
    const unsigned nb = red_blocks(n);
    scratch.reserve((size_t)nb * 4 * sizeof(double), stream);
    double* part = (double*)scratch.ptr;
    auto go = [&](auto pa, auto pb) {
        k_inner_partial<<<nb, kRedThreads, 0, stream>>>(pa, pb, n, part);
        k_maxdiff_partial<<<nb, kRedThreads, 0, stream>>>(pa, pb, n, part + 2 * nb);
        k_norm_partial<<<nb, kRedThreads, 0, stream>>>(pb, n, part + 3 * nb);
    };
    if (dta == QLX_C128 && dtb == QLX_C128) go((const double2*)a, (const double2*)b);
    else if (dta == QLX_C128) go((const double2*)a, (const float2*)b);
    else if (dtb == QLX_C128) go((const float2*)a, (const double2*)b);
    else go((const float2*)a, (const float2*)b);
    QLX_HIP_CHECK(hipGetLastError());
    std::vector<double> h(4 * nb);
    QLX_HIP_CHECK(hipMemcpyAsync(h.data(), part, 4 * nb * sizeof(double), hipMemcpyDeviceToHost, stream));
    QLX_HIP_CHECK(hipStreamSynchronize(stream));
    double r = 0, i = 0, m = 0, nrm = 0;
    for (unsigned k = 0; k < nb; ++k) {
        r += h[2 * k];
        i += h[2 * k + 1];
        m = std::fmax(m, h[2 * nb + k]);
        nrm += h[3 * nb + k];
    }
    *re = r;
    *im = i;
    *maxd = m;
    if (norm_b) *norm_b = nrm;
}

template <class F>
static void dispatch_pair(qlx_state* a, qlx_state* b, F&& f) {
    if (a->dtype == QLX_C128 && b->dtype == QLX_C128) f((const double2*)a->amp, (const double2*)b->amp);
    else if (a->dtype == QLX_C128) f((const double2*)a->amp, (const float2*)b->amp);
    else if (b->dtype == QLX_C128) f((const float2*)a->amp, (const double2*)b->amp);
    else f((const float2*)a->amp, (const float2*)b->amp);
}

}  // namespace qlx

using namespace qlx;

extern "C" {

const char* qlx_last_error(void) { return g_last_error.c_str(); }

int qlx_state_create(uint32_t num_qubits, int dtype, qlx_state** out) {
    return guard([&] {
        if (!out) fail(QLX_EINVAL, "null output handle");
        *out = nullptr;
        qlx_state* st = new_state();
        try {
            alloc_storage(st, num_qubits, dtype);
            launch_fill_zero_state(st);
        } catch (...) {
            destroy_state(st);
            throw;
        }
        *out = st;
    });
}

int qlx_state_create_from_host(const double* amps, size_t dim, int dtype, qlx_state** out) {
    return guard([&] {
        if (!out) fail(QLX_EINVAL, "null output handle");
        *out = nullptr;
        const uint32_t nq = validate_host_state(amps, dim);
        qlx_state* st = new_state();
        try {
            alloc_storage(st, nq, dtype);
            upload_host_state(st, amps, dim);
        } catch (...) {
            destroy_state(st);
            throw;
        }
        *out = st;
    });
}

int qlx_state_destroy(qlx_state* st) {
    return guard([&] { destroy_state(st); });
}

int qlx_state_reset(qlx_state* st) {
    return guard([&] {
        check_state(st);
        launch_fill_zero_state(st);
    });
}

int qlx_state_load(qlx_state* st, const double* amps, size_t dim) {
    return guard([&] {
        if (!st) fail(QLX_EINVAL, "null state handle");
        const uint32_t nq = validate_host_state(amps, dim);
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
        if (nq != st->nq || !st->amp) alloc_storage(st, nq, st->dtype);
        upload_host_state(st, amps, dim);
    });
}

int qlx_state_to_host(qlx_state* st, double* out, size_t dim) {
    return guard([&] {
        check_state(st);
        const size_t n = size_t(1) << st->nq;
        if (!out || dim < n) fail(QLX_EINVAL, "output buffer smaller than the state dimension");
        if (st->dtype == QLX_C128) {
            QLX_HIP_CHECK(hipMemcpyAsync(out, st->amp, n * 16, hipMemcpyDeviceToHost, st->stream));
        } else {
            const size_t slice = size_t(1) << 24;
            DevBuf tmp;
            tmp.reserve(std::min(n, slice) * 16, st->stream);
            for (size_t off = 0; off < n; off += slice) {
                size_t cnt = std::min(slice, n - off);
                k_widen_c64_to_c128<<<fill_blocks(cnt), 256, 0, st->stream>>>((const float2*)st->amp + off,
                                                                               (double2*)tmp.ptr, cnt);
                QLX_HIP_CHECK(hipGetLastError());
                QLX_HIP_CHECK(hipMemcpyAsync(out + 2 * off, tmp.ptr, cnt * 16, hipMemcpyDeviceToHost, st->stream));
                QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
            }
            tmp.release();
        }
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
    });
}

int qlx_state_num_qubits(const qlx_state* st, uint32_t* out) {
    return guard([&] {
        if (!st || !out) fail(QLX_EINVAL, "null argument");
        *out = st->nq;
    });
}

int qlx_state_dtype(const qlx_state* st, int* out) {
    return guard([&] {
        if (!st || !out) fail(QLX_EINVAL, "null argument");
        *out = st->dtype;
    });
}

int qlx_state_device_ptr(const qlx_state* st, void** out) {
    return guard([&] {
        if (!st || !out) fail(QLX_EINVAL, "null argument");
        *out = st->amp;
    });
}

int qlx_state_norm_sq(qlx_state* st, double* out) {
    return guard([&] {
        check_state(st);
        if (!out) fail(QLX_EINVAL, "null output");
        *out = reduce_norm_sq(st);
    });
}

int qlx_state_inner(qlx_state* a, qlx_state* b, double* re, double* im) {
    return guard([&] {
        check_state(a);
        check_state(b);
        if (a->nq != b->nq) fail(QLX_EINVAL, "states differ in qubit count");
        if (!re || !im) fail(QLX_EINVAL, "null output");
        QLX_HIP_CHECK(hipStreamSynchronize(b->stream));
        const uint64_t n = 1ull << a->nq;
        const unsigned nb = red_blocks(n);
        a->scratch.reserve((size_t)nb * 2 * sizeof(double), a->stream);
        dispatch_pair(a, b, [&](auto pa, auto pb) {
            k_inner_partial<<<nb, kRedThreads, 0, a->stream>>>(pa, pb, n, (double*)a->scratch.ptr);
        });
        QLX_HIP_CHECK(hipGetLastError());
        std::vector<double> h(2 * nb);
        QLX_HIP_CHECK(hipMemcpyAsync(h.data(), a->scratch.ptr, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, a->stream));
        QLX_HIP_CHECK(hipStreamSynchronize(a->stream));
        double r = 0.0, i = 0.0;
        for (unsigned k = 0; k < nb; ++k) {
            r += h[2 * k];
            i += h[2 * k + 1];
        }
        *re = r;
        *im = i;
    });
}

int qlx_state_max_abs_diff(qlx_state* a, qlx_state* b, double* out) {
    return guard([&] {
        check_state(a);
        check_state(b);
        if (a->nq != b->nq) fail(QLX_EINVAL, "states differ in qubit count");
        if (!out) fail(QLX_EINVAL, "null output");
        QLX_HIP_CHECK(hipStreamSynchronize(b->stream));
        const uint64_t n = 1ull << a->nq;
        const unsigned nb = red_blocks(n);
        a->scratch.reserve((size_t)nb * 2 * sizeof(double), a->stream);
        dispatch_pair(a, b, [&](auto pa, auto pb) {
            k_maxdiff_partial<<<nb, kRedThreads, 0, a->stream>>>(pa, pb, n, (double*)a->scratch.ptr);
        });
        QLX_HIP_CHECK(hipGetLastError());
        std::vector<double> h(nb);
        QLX_HIP_CHECK(hipMemcpyAsync(h.data(), a->scratch.ptr, nb * sizeof(double), hipMemcpyDeviceToHost, a->stream));
        QLX_HIP_CHECK(hipStreamSynchronize(a->stream));
        double m = 0.0;
        for (double v : h) m = std::fmax(m, v);
        *out = m;
    });
}

int qlx_state_copy(qlx_state* dst, qlx_state* src) {
    return guard([&] {
        check_state(dst);
        check_state(src);
        if (dst->nq != src->nq) fail(QLX_EINVAL, "states differ in qubit count");
        QLX_HIP_CHECK(hipStreamSynchronize(src->stream));
        const uint64_t n = 1ull << src->nq;
        if (dst->dtype == src->dtype && dst->bytes % 16 == 0) {
            TimedLaunch tl(dst, "k_copy", 2.0 * (double)dst->bytes);
            const uint64_t n16 = dst->bytes / 16;
            int dev = 0;
            QLX_HIP_CHECK(hipGetDevice(&dev));
            int cus = 0;
            QLX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            if (n16 % kCopyChunk16 == 0 && n16 / kCopyChunk16 <= 0x7fffffffull) {
                k_copy_chunks<<<(unsigned)(n16 / kCopyChunk16), 256, 0, dst->stream>>>((const uint4*)src->amp,
                                                                                      (uint4*)dst->amp);
            } else {
                const uint64_t blocks = std::min<uint64_t>((n16 + 255) / 256, (uint64_t)cus * 16);
                k_copy<<<(unsigned)std::max<uint64_t>(blocks, 1), 256, 0, dst->stream>>>((const uint4*)src->amp,
                                                                                      (uint4*)dst->amp, n16);
            }
        } else if (dst->dtype == src->dtype) {
            QLX_HIP_CHECK(hipMemcpyAsync(dst->amp, src->amp, dst->bytes, hipMemcpyDeviceToDevice, dst->stream));
        } else if (src->dtype == QLX_C128) {
            k_narrow_c128_to_c64<<<fill_blocks(n), 256, 0, dst->stream>>>((const double2*)src->amp,
                                                                          (float2*)dst->amp, n);
        } else {
            k_widen_c64_to_c128<<<fill_blocks(n), 256, 0, dst->stream>>>((const float2*)src->amp,
                                                                         (double2*)dst->amp, n);
        }
        QLX_HIP_CHECK(hipGetLastError());
        QLX_HIP_CHECK(hipStreamSynchronize(dst->stream));
    });
}

int qlx_sync(qlx_state* st) {
    return guard([&] {
        if (!st) fail(QLX_EINVAL, "null state handle");
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
        QLX_HIP_CHECK(hipGetLastError());
    });
}

int qlx_timing_enable(qlx_state* st, int on) {
    return guard([&] {
        if (!st) fail(QLX_EINVAL, "null state handle");
        st->timing = on != 0;
    });
}

int qlx_timing_reset(qlx_state* st) {
    return guard([&] {
        if (!st) fail(QLX_EINVAL, "null state handle");
        drain_timing(st);
        st->timing_agg.clear();
    });
}

int qlx_timing_read(qlx_state* st, qlx_kernel_timing* out, size_t cap, size_t* n_out) {
    return guard([&] {
        if (!st) fail(QLX_EINVAL, "null state handle");
        drain_timing(st);
        if (n_out) *n_out = st->timing_agg.size();
        for (size_t i = 0; i < st->timing_agg.size() && out && i < cap; ++i) {
            std::memset(&out[i], 0, sizeof(out[i]));
            std::strncpy(out[i].kernel, st->timing_agg[i].first.c_str(), sizeof(out[i].kernel) - 1);
            out[i].launches = st->timing_agg[i].second.launches;
            out[i].total_ms = st->timing_agg[i].second.ms;
            out[i].algorithmic_bytes = st->timing_agg[i].second.bytes;
        }
    });
}

int qlx_last_exec_stats(const qlx_state* st, qlx_exec_stats* out) {
    return guard([&] {
        if (!st || !out) fail(QLX_EINVAL, "null argument");
        *out = st->stats;
    });
}

}  // extern "C"
