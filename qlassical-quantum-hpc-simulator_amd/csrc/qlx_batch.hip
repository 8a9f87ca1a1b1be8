// qlx_batch.hip — many independent circuits on one GPU.
//
// Every circuit gets its own state in one HBM arena and its own window-sweep plan
// (qlx_sweep.hip). Circuits at least one window wide (c64: 13 qubits, c128: 12) use
// 2^W-amplitude chunks; smaller circuits are one chunk of their own size (W = N), so a
// whole small circuit is one on-chip task: its state is synthesised as |0...0> in
// registers, every gate runs there, and it is written to HBM once.
//
// QLX_BATCH_PERSISTENT: one launch of `k_batch` with one work-group per LDS slot of the
// device. A work-group takes whole circuits from a queue (largest amount of work first,
// so the long circuits start early) and runs every sweep of its circuit itself: chunk
// after chunk, with a work-group barrier between sweeps. No work-group ever waits for
// another; the launch ends when the queue is empty.
// QLX_BATCH_LEVELS: the (circuit, sweep, chunk) tasks, one launch per sweep level
// (kernel boundaries order the levels; chunks of a level run in parallel).
// QLX_BATCH_PER_CIRCUIT(_GRAPH): the baseline — each circuit's gates as
// one-kernel-per-gate launches (optionally captured once into a hipGraph).
// With QLX_SPECIALIZE_RELABEL in the prepare options, a circuit larger than one window gets two
// state buffers and its sweeps run out of place (relabelled layouts, windows free of the low
// qubits: fewer sweeps); after a sweep mode its result is in the buffer its last sweep wrote.

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "qlx_sweep_core.hpp"

namespace qlx {

struct BatchCircuitDev {
    uint64_t amp_off;      // element offset in the arena
    uint64_t alt_off;      // out of place: the second buffer (== amp_off in place)
    uint32_t sweep_begin;  // first DevSweep of this circuit
    uint32_t n_sweeps;
    uint32_t rb;           // register qubits of this circuit's geometry
    uint32_t log_chunks;   // chunks per sweep = 2^log_chunks
    uint32_t oop, pad;     // out of place: sweep k reads buffer k % 2 and writes the other
};

struct BatchTaskDev {
    uint32_t circuit, sweep, chunk, pad;
};

template <class A>
__device__ __forceinline__ void run_task_chunk(A* arena, const BatchCircuitDev& c, uint32_t sweep, uint32_t chunk,
                                               const DevSweep* sw, const DevPhase* ph, const DevOp* ops,
                                               const double* coef, A* lds) {
    const DevSweep* s = sw + c.sweep_begin + sweep;
    const bool odd = (sweep & 1u) != 0, oop = c.oop != 0;
    A* src = arena + (odd ? c.alt_off : c.amp_off);
    A* dst = oop ? arena + (odd ? c.amp_off : c.alt_off) : src;
    const bool init = sweep == 0;
    if constexpr (sizeof(A) == 8) {
        switch (c.rb) {
            case 2: run_chunk<A, 8, 2>(src, dst, oop, s, chunk, ph, ops, coef, lds, init); break;
            case 3: run_chunk<A, 8, 3>(src, dst, oop, s, chunk, ph, ops, coef, lds, init); break;
            case 4: run_chunk<A, 8, 4>(src, dst, oop, s, chunk, ph, ops, coef, lds, init); break;
            default: run_chunk<A, 8, 5>(src, dst, oop, s, chunk, ph, ops, coef, lds, init); break;
        }
    } else {
        switch (c.rb) {
            case 2: run_chunk<A, 8, 2>(src, dst, oop, s, chunk, ph, ops, coef, lds, init); break;
            case 3: run_chunk<A, 8, 3>(src, dst, oop, s, chunk, ph, ops, coef, lds, init); break;
            default: run_chunk<A, 8, 4>(src, dst, oop, s, chunk, ph, ops, coef, lds, init); break;
        }
    }
}

// ctl[0]: next queue position (zeroed before the launch).
template <class A>
__global__ void __launch_bounds__(256) k_batch(A* __restrict__ arena, const BatchCircuitDev* __restrict__ circ,
                                              const uint32_t* __restrict__ queue, uint32_t n_circ,
                                              const DevSweep* __restrict__ sw, const DevPhase* __restrict__ ph,
                                              const DevOp* __restrict__ ops, const double* __restrict__ coef,
                                              uint32_t* __restrict__ ctl) {
    extern __shared__ __align__(16) unsigned char smem[];
    A* lds = reinterpret_cast<A*>(smem);
    __shared__ uint32_t s_job;
    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(ctl, 1u);
        __syncthreads();
        const uint32_t j = s_job;
        __syncthreads();  // every work-item holds j before s_job is written again
        if (j >= n_circ) break;
        const BatchCircuitDev c = circ[queue[j]];
        const uint32_t nch = 1u << c.log_chunks;
        for (uint32_t s = 0; s < c.n_sweeps; ++s) {
            for (uint32_t k = 0; k < nch; ++k) run_task_chunk<A>(arena, c, s, k, sw, ph, ops, coef, lds);
            // the sweep's stores are visible to the whole work-group before the next sweep reads
            __syncthreads();
        }
    }
}

// One launch per sweep level: block b runs task tasks[first + b].
template <class A>
__global__ void __launch_bounds__(256) k_batch_level(A* __restrict__ arena, const BatchCircuitDev* __restrict__ circ,
                                                    const BatchTaskDev* __restrict__ tasks, uint32_t first,
                                                    const DevSweep* __restrict__ sw, const DevPhase* __restrict__ ph,
                                                    const DevOp* __restrict__ ops, const double* __restrict__ coef) {
    extern __shared__ __align__(16) unsigned char smem[];
    const BatchTaskDev tk = tasks[first + blockIdx.x];
    run_task_chunk<A>(arena, circ[tk.circuit], tk.sweep, tk.chunk, sw, ph, ops, coef, reinterpret_cast<A*>(smem));
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace qlx

using namespace qlx;

struct qlx_batch {
    int dtype = QLX_C64;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    struct Circ {
        uint32_t nq;
        std::vector<qlx_gate_instr> instrs;
        std::vector<double> pool;
        std::vector<qlx_matrix_entry> meta;
        uint64_t amp_off = 0;
        uint64_t alt_off = 0;     // out of place: the second buffer (== amp_off in place)
        uint64_t result_off = 0;  // after a sweep mode: the buffer the last sweep wrote
        uint64_t pool_off = 0;  // complex elements into the device pool
        std::unique_ptr<SweepPlan> plan;
    };
    std::vector<Circ> circs;
    bool prepared = false;
    // device
    DevBuf arena, desc, pool, ctl, scratch;
    size_t off_circ = 0, off_task = 0, off_queue = 0, off_sw = 0, off_ph = 0, off_op = 0, off_cf = 0;
    std::vector<BatchTaskDev> tasks;
    std::vector<uint32_t> queue;  // circuits in persistent-queue order
    std::vector<uint32_t> level_first, level_count;
    uint32_t grid = 0;
    size_t lds_bytes = 0;
    qlx_batch_info info{};
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<qlx_state> views;  // non-owning per-circuit states over the arena (plain modes)
    int last_mode = -1;            // the last qlx_batch_run mode (where the results are)
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

void destroy_batch(qlx_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->arena.release();
    b->desc.release();
    b->pool.release();
    b->ctl.release();
    b->scratch.release();
    for (auto& v : b->views) {
        v.amp = nullptr;  // non-owning
        v.prog.release();
        v.scratch.release();
    }
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

Geometry batch_geometry(int dtype, uint32_t nq, const qlx_exec_opts& o) {
    const int wmax = o.window_qubits ? (int)o.window_qubits : (dtype == QLX_C128 ? 12 : 13);
    Geometry g{};
    g.elem = dtype == QLX_C128 ? 16 : 8;
    g.logt = 8;
    g.w = std::min<int>((int)nq, wmax);
    g.rb = g.w - g.logt;
    g.c = dtype == QLX_C64 ? 4 : 3;  // one 128-byte cache line of amplitudes
    const int rbmax = dtype == QLX_C128 ? 4 : 5;
    if (g.rb < 2) fail(QLX_EUNSUPPORTED, "batched circuits need at least 10 qubits");
    if (g.rb > rbmax) fail(QLX_EUNSUPPORTED, "window too large for the batched kernel");
    return g;
}

void launch_persistent(qlx_batch* b) {
    k_zero_u32<<<1, 64, 0, b->stream>>>((uint32_t*)b->ctl.ptr, 1u);
    const char* d = (const char*)b->desc.ptr;
    const uint32_t nc = (uint32_t)b->queue.size();
    auto go = [&](auto* arena) {
        using A = std::remove_pointer_t<decltype(arena)>;
        k_batch<A><<<b->grid, 256, b->lds_bytes, b->stream>>>(
            arena, (const BatchCircuitDev*)(d + b->off_circ), (const uint32_t*)(d + b->off_queue), nc,
            (const DevSweep*)(d + b->off_sw), (const DevPhase*)(d + b->off_ph), (const DevOp*)(d + b->off_op),
            (const double*)(d + b->off_cf), (uint32_t*)b->ctl.ptr);
    };
    if (b->dtype == QLX_C64) go((float2*)b->arena.ptr);
    else go((double2*)b->arena.ptr);
    QLX_HIP_CHECK(hipGetLastError());
}

void launch_levels(qlx_batch* b) {
    const char* d = (const char*)b->desc.ptr;
    for (size_t L = 0; L < b->level_first.size(); ++L) {
        auto go = [&](auto* arena) {
            using A = std::remove_pointer_t<decltype(arena)>;
            k_batch_level<A><<<b->level_count[L], 256, b->lds_bytes, b->stream>>>(
                arena, (const BatchCircuitDev*)(d + b->off_circ), (const BatchTaskDev*)(d + b->off_task),
                b->level_first[L], (const DevSweep*)(d + b->off_sw), (const DevPhase*)(d + b->off_ph),
                (const DevOp*)(d + b->off_op), (const double*)(d + b->off_cf));
        };
        if (b->dtype == QLX_C64) go((float2*)b->arena.ptr);
        else go((double2*)b->arena.ptr);
        QLX_HIP_CHECK(hipGetLastError());
    }
}

// A circuit's result: the buffer its last sweep wrote after a sweep mode, its first buffer
// after a per-circuit mode (in place).
const void* result_ptr(const qlx_batch* b, uint32_t id) {
    const bool sweeps = b->last_mode == QLX_BATCH_PERSISTENT || b->last_mode == QLX_BATCH_LEVELS;
    const uint64_t off = sweeps ? b->circs[id].result_off : b->circs[id].amp_off;
    return (const char*)b->arena.ptr + off * amp_size(b->dtype);
}

void launch_per_circuit(qlx_batch* b) {
    qlx_exec_opts o{};
    qlx_exec_stats stats{};
    for (size_t i = 0; i < b->circs.size(); ++i) {
        qlx_state& v = b->views[i];
        launch_fill_zero_state(&v);
        const auto& c = b->circs[i];
        const double* dpool = c.pool.empty() ? nullptr : (const double*)b->pool.ptr + 2 * c.pool_off;
        for (const qlx_gate_instr& g : c.instrs) launch_plain_instr(&v, g, dpool, c.meta.data(), o, stats);
    }
}

}  // namespace

extern "C" {

int qlx_batch_create(int dtype, qlx_batch** out) {
    return guard([&] {
        if (!out) fail(QLX_EINVAL, "null output");
        if (dtype != QLX_C64 && dtype != QLX_C128) fail(QLX_EINVAL, "bad dtype");
        auto* b = new qlx_batch();
        b->dtype = dtype;
        try {
            QLX_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            QLX_HIP_CHECK(hipEventCreate(&b->ev0));
            QLX_HIP_CHECK(hipEventCreate(&b->ev1));
        } catch (...) {
            destroy_batch(b);
            throw;
        }
        *out = b;
    });
}

int qlx_batch_add(qlx_batch* b, uint32_t nq, const qlx_gate_instr* instrs, size_t n, const double* pool,
                  size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta, uint32_t* id) {
    return guard([&] {
        if (!b) fail(QLX_EINVAL, "null batch");
        if (b->prepared) fail(QLX_EINVAL, "batch already prepared");
        if (nq < 10 || nq > 34) fail(QLX_EUNSUPPORTED, "batched circuits need 10..34 qubits");
        if (n == 0 || !instrs) fail(QLX_EINVAL, "a batched circuit needs at least one instruction");
        qlx_batch::Circ c;
        c.nq = nq;
        for (size_t i = 0; i < n; ++i) {
            const qlx_gate_instr& g = instrs[i];
            if (g.type == QLX_MEASURE || g.type == QLX_CCX || g.type == QLX_CSWAP ||
                ((g.type == QLX_UNITARY || g.type == QLX_FUSED_BLOCK) && instr_arity(g) > 2))
                fail(QLX_EUNSUPPORTED, "batched circuits take only gates that run inside a window sweep");
            validate_instr(g, nq, meta, n_meta, pool_elems);
            c.instrs.push_back(g);
        }
        if (pool && pool_elems) c.pool.assign(pool, pool + 2 * pool_elems);
        if (meta && n_meta) c.meta.assign(meta, meta + n_meta);
        if (id) *id = (uint32_t)b->circs.size();
        b->circs.push_back(std::move(c));
    });
}

int qlx_batch_prepare(qlx_batch* b, const qlx_exec_opts* opts, qlx_batch_info* info) {
    return guard([&] {
        if (!b) fail(QLX_EINVAL, "null batch");
        if (b->prepared) fail(QLX_EINVAL, "batch already prepared");
        if (b->circs.empty()) fail(QLX_EINVAL, "empty batch");
        qlx_exec_opts o{};
        if (opts) o = *opts;
        // batched sweeps run the generic kernel's descriptors, out of place on request
        const bool want_oop = (o.specialize & QLX_SPECIALIZE_RELABEL) != 0;
        o.specialize = 0;
        o.gswap_mode = QLX_GSWAP_PERMUTE;
        const size_t amp = amp_size(b->dtype);
        // plan every circuit, lay out the arena and the device pool
        uint64_t amp_total = 0, pool_total = 0;
        std::vector<BatchCircuitDev> cdev;
        std::vector<DevSweep> dsw;
        std::vector<DevPhase> dph;
        std::vector<DevOp> dop;
        std::vector<double> coef;
        uint32_t max_sweeps = 0;
        uint64_t gates = 0;
        for (auto& c : b->circs) {
            Geometry geo = batch_geometry(b->dtype, c.nq, o);
            const bool oop = want_oop && (int)c.nq > geo.w;
            qlx_exec_opts co = o;
            if (oop) {
                geo.free = !(o.plan_flags & QLX_PLAN_FIXED_WINDOWS);
                co.specialize = QLX_SPECIALIZE_RELABEL;
            }
            c.plan = build_plan_geo(c.nq, geo, c.instrs.data(), c.instrs.size(),
                                    c.pool.empty() ? nullptr : c.pool.data(), c.meta.empty() ? nullptr : c.meta.data(), co,
                                    /*thorough=*/false);
            if (b->dtype == QLX_C64) densify_phases(*c.plan);
            if (oop) assign_layouts(*c.plan, c.nq);
            for (const PlanStep& s : c.plan->steps)
                if (!s.sweep) fail(QLX_EUNSUPPORTED, "a batched circuit needs a kernel outside the window sweep");
            c.amp_off = amp_total;
            amp_total += 1ull << c.nq;
            c.alt_off = c.amp_off;
            if (oop) {
                c.alt_off = amp_total;
                amp_total += 1ull << c.nq;
            }
            c.result_off = (oop && (c.plan->dsw.size() & 1)) ? c.alt_off : c.amp_off;
            c.pool_off = pool_total;
            pool_total += c.pool.size() / 2;
            gates += c.instrs.size();
            BatchCircuitDev d{};
            d.amp_off = c.amp_off;
            d.alt_off = c.alt_off;
            d.oop = oop ? 1u : 0u;
            d.sweep_begin = (uint32_t)dsw.size();
            d.n_sweeps = (uint32_t)c.plan->dsw.size();
            d.rb = (uint32_t)geo.rb;
            d.log_chunks = c.nq - (uint32_t)geo.w;
            max_sweeps = std::max(max_sweeps, d.n_sweeps);
            if (coef.size() & 1) coef.push_back(0.0);  // matrix-core fragments stay 16-byte aligned
            const uint32_t ph0 = (uint32_t)dph.size(), op0 = (uint32_t)dop.size(), cf0 = (uint32_t)coef.size();
            for (DevSweep s : c.plan->dsw) {
                s.phase_begin += ph0;
                dsw.push_back(s);
            }
            for (DevPhase p : c.plan->dph) {
                p.op_begin += op0;
                if (p.mfma) p.mfma += cf0;
                dph.push_back(p);
            }
            for (DevOp op : c.plan->dop) {
                op.coef += cf0;
                dop.push_back(op);
            }
            coef.insert(coef.end(), c.plan->coef.begin(), c.plan->coef.end());
            cdev.push_back(d);
        }
        // tasks by sweep level
        b->tasks.clear();
        b->level_first.clear();
        b->level_count.clear();
        for (uint32_t L = 0; L < max_sweeps; ++L) {
            b->level_first.push_back((uint32_t)b->tasks.size());
            for (size_t ci = 0; ci < cdev.size(); ++ci) {
                const BatchCircuitDev& d = cdev[ci];
                if (L >= d.n_sweeps) continue;
                const uint32_t nch = 1u << d.log_chunks;
                for (uint32_t k = 0; k < nch; ++k) b->tasks.push_back(BatchTaskDev{(uint32_t)ci, L, k, 0});
            }
            b->level_count.push_back((uint32_t)b->tasks.size() - b->level_first.back());
        }
        // persistent queue: most chunk-sweeps first (longest-processing-time order)
        b->queue.resize(cdev.size());
        for (size_t ci = 0; ci < cdev.size(); ++ci) b->queue[ci] = (uint32_t)ci;
        auto work = [&](uint32_t ci) { return (uint64_t)cdev[ci].n_sweeps << cdev[ci].log_chunks; };
        std::stable_sort(b->queue.begin(), b->queue.end(), [&](uint32_t x, uint32_t y) { return work(x) > work(y); });
        // descriptor image
        b->off_circ = 0;
        b->off_task = align_up(cdev.size() * sizeof(BatchCircuitDev), 256);
        b->off_queue = align_up(b->off_task + b->tasks.size() * sizeof(BatchTaskDev), 256);
        b->off_sw = align_up(b->off_queue + b->queue.size() * sizeof(uint32_t), 256);
        b->off_ph = align_up(b->off_sw + dsw.size() * sizeof(DevSweep), 256);
        b->off_op = align_up(b->off_ph + dph.size() * sizeof(DevPhase), 256);
        b->off_cf = align_up(b->off_op + dop.size() * sizeof(DevOp), 256);
        const size_t total = align_up(b->off_cf + coef.size() * sizeof(double), 256) + 256;
        std::vector<char> img(total, 0);
        std::memcpy(img.data() + b->off_circ, cdev.data(), cdev.size() * sizeof(BatchCircuitDev));
        std::memcpy(img.data() + b->off_task, b->tasks.data(), b->tasks.size() * sizeof(BatchTaskDev));
        std::memcpy(img.data() + b->off_queue, b->queue.data(), b->queue.size() * sizeof(uint32_t));
        std::memcpy(img.data() + b->off_sw, dsw.data(), dsw.size() * sizeof(DevSweep));
        std::memcpy(img.data() + b->off_ph, dph.data(), dph.size() * sizeof(DevPhase));
        if (!dop.empty()) std::memcpy(img.data() + b->off_op, dop.data(), dop.size() * sizeof(DevOp));
        if (!coef.empty()) std::memcpy(img.data() + b->off_cf, coef.data(), coef.size() * sizeof(double));
        b->desc.reserve(total, b->stream);
        QLX_HIP_CHECK(hipMemcpyAsync(b->desc.ptr, img.data(), total, hipMemcpyHostToDevice, b->stream));
        b->arena.reserve(amp_total * amp, b->stream);
        b->ctl.reserve(sizeof(uint32_t), b->stream);
        if (pool_total) {
            std::vector<double> pool_img(2 * pool_total);
            for (const auto& c : b->circs)
                if (!c.pool.empty()) std::memcpy(pool_img.data() + 2 * c.pool_off, c.pool.data(), c.pool.size() * 8);
            b->pool.reserve(pool_img.size() * sizeof(double), b->stream);
            QLX_HIP_CHECK(hipMemcpyAsync(b->pool.ptr, pool_img.data(), pool_img.size() * 8, hipMemcpyHostToDevice,
                                         b->stream));
        }
        QLX_HIP_CHECK(hipStreamSynchronize(b->stream));
        // non-owning per-circuit views for the plain baselines
        b->views.resize(b->circs.size());
        for (size_t i = 0; i < b->circs.size(); ++i) {
            qlx_state& v = b->views[i];
            v.nq = b->circs[i].nq;
            v.dtype = b->dtype;
            v.amp = (char*)b->arena.ptr + b->circs[i].amp_off * amp;
            v.bytes = (size_t(1) << v.nq) * amp;
            v.stream = b->stream;
        }
        // persistent grid: every LDS slot of the device
        b->lds_bytes = b->dtype == QLX_C64 ? (size_t(1) << 13) * 8 : (size_t(1) << 12) * 16;
        int per_cu = 0;
        if (b->dtype == QLX_C64)
            QLX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch<float2>, 256, b->lds_bytes));
        else
            QLX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch<double2>, 256, b->lds_bytes));
        int dev = 0, cus = 0;
        QLX_HIP_CHECK(hipGetDevice(&dev));
        QLX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        b->grid = (uint32_t)std::max(1, per_cu) * (uint32_t)cus;
        b->grid = std::min<uint32_t>(b->grid, (uint32_t)b->queue.size());
        b->prepared = true;
        b->info = qlx_batch_info{};
        b->info.circuits = b->circs.size();
        b->info.tasks = b->tasks.size();
        b->info.sweeps = dsw.size();
        b->info.state_bytes = amp_total * amp;
        b->info.gates = gates;
        b->info.workgroups = b->grid;
        b->info.max_sweeps = max_sweeps;
        if (info) *info = b->info;
    });
}

int qlx_batch_run(qlx_batch* b, int mode) {
    return guard([&] {
        if (!b || !b->prepared) fail(QLX_EINVAL, "batch not prepared");
        QLX_HIP_CHECK(hipEventRecord(b->ev0, b->stream));
        switch (mode) {
            case QLX_BATCH_PERSISTENT: launch_persistent(b); break;
            case QLX_BATCH_LEVELS: launch_levels(b); break;
            case QLX_BATCH_PER_CIRCUIT: launch_per_circuit(b); break;
            case QLX_BATCH_PER_CIRCUIT_GRAPH: {
                if (!b->graph_exec) {
                    // capture the whole per-circuit launch sequence once (no memset nodes: the
                    // |0...0> fill is a kernel)
                    QLX_HIP_CHECK(hipStreamSynchronize(b->stream));
                    hipStream_t cap = nullptr;
                    QLX_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
                    for (auto& v : b->views) v.stream = cap;
                    QLX_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                    try {
                        launch_per_circuit(b);
                    } catch (...) {
                        hipGraph_t g = nullptr;
                        (void)hipStreamEndCapture(cap, &g);
                        if (g) (void)hipGraphDestroy(g);
                        for (auto& v : b->views) v.stream = b->stream;
                        (void)hipStreamDestroy(cap);
                        throw;
                    }
                    QLX_HIP_CHECK(hipStreamEndCapture(cap, &b->graph));
                    for (auto& v : b->views) v.stream = b->stream;
                    QLX_HIP_CHECK(hipStreamDestroy(cap));
                    QLX_HIP_CHECK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
                    QLX_HIP_CHECK(hipEventRecord(b->ev0, b->stream));
                }
                QLX_HIP_CHECK(hipGraphLaunch(b->graph_exec, b->stream));
                break;
            }
            default: fail(QLX_EINVAL, "unknown batch mode");
        }
        QLX_HIP_CHECK(hipEventRecord(b->ev1, b->stream));
        b->last_mode = mode;
    });
}

int qlx_batch_sync(qlx_batch* b, double* elapsed_ms) {
    return guard([&] {
        if (!b) fail(QLX_EINVAL, "null batch");
        QLX_HIP_CHECK(hipStreamSynchronize(b->stream));
        QLX_HIP_CHECK(hipGetLastError());
        if (elapsed_ms) {
            float ms = 0.f;
            QLX_HIP_CHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            *elapsed_ms = ms;
        }
    });
}

int qlx_batch_state_to_host(qlx_batch* b, uint32_t id, double* out, size_t dim) {
    return guard([&] {
        if (!b || !b->prepared) fail(QLX_EINVAL, "batch not prepared");
        if (id >= b->circs.size()) fail(QLX_ERANGE, "circuit id out of range");
        QLX_HIP_CHECK(hipStreamSynchronize(b->stream));
        qlx_state& v = b->views[id];
        const size_t n = size_t(1) << v.nq;
        if (!out || dim < n) fail(QLX_EINVAL, "output buffer too small");
        const void* res = result_ptr(b, id);
        if (b->dtype == QLX_C128) {
            QLX_HIP_CHECK(hipMemcpy(out, res, n * 16, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> tmp(2 * n);
            QLX_HIP_CHECK(hipMemcpy(tmp.data(), res, n * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < 2 * n; ++i) out[i] = tmp[i];
        }
    });
}

int qlx_batch_compare(qlx_batch* b, uint32_t id, qlx_state* ref, double* infidelity, double* max_abs) {
    return guard([&] {
        if (!b || !b->prepared) fail(QLX_EINVAL, "batch not prepared");
        if (id >= b->circs.size()) fail(QLX_ERANGE, "circuit id out of range");
        if (!ref || !ref->amp) fail(QLX_EINVAL, "null reference state");
        qlx_state& v = b->views[id];
        if (ref->nq != v.nq) fail(QLX_EINVAL, "reference state has a different qubit count");
        QLX_HIP_CHECK(hipStreamSynchronize(b->stream));
        QLX_HIP_CHECK(hipStreamSynchronize(ref->stream));
        double re = 0, im = 0, m = 0, nb = 0;
        compare_amplitudes(ref->amp, ref->dtype, result_ptr(b, id), v.dtype, 1ull << v.nq, b->stream, b->scratch, &re,
                           &im, &m, &nb);
        if (infidelity) *infidelity = infidelity_from(re, im, reduce_norm_sq(ref), nb);
        if (max_abs) *max_abs = m;
    });
}

int qlx_batch_destroy(qlx_batch* b) {
    return guard([&] { destroy_batch(b); });
}

}  // extern "C"
