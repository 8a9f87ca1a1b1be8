// qlx_serve.hip — a continuous-batching circuit server on one GPU.
//
// Requests (MEASURE-free circuits of >= 10 qubits made of window-sweep gates) are queued
// on submit and admitted FIFO while their state fits an HBM byte budget (first-fit
// allocator over one arena). An admitted request's window-sweep plan is uploaded once and
// its state stays in HBM until the client releases it. Requests are grouped by qubit
// count; each group owns a HIP stream, and one launch (`k_serve_step`) advances every
// active request of the group by one sweep (all its chunks as independent work-groups).
// New requests join their group's next launch, finished ones leave: iteration-level
// (continuous) batching. Groups run on different streams, so their work overlaps on the
// device. Up to two launches per group are in flight; a request's sweeps are ordered by
// its group's stream. Plans are formed on host planner threads (a few ms per circuit of
// 20-28 qubits): submit returns at once, and admission takes the queue's head as soon as
// its plan is ready, so planning overlaps the device's work. A request larger than one window
// runs out of place between two halves of its arena block (each sweep re-lays out the qubits so
// that the next window sits on the low address bits, and windows need not hold the low qubits),
// unless the server runs in place (QLX_SERVE_IN_PLACE) or the two halves exceed the budget.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "qlx_sweep_core.hpp"

namespace qlx {

struct ServeTask {
    uint64_t src_off;    // elements into the arena: the buffer this sweep reads
    uint64_t dst_off;    // ... and the one it writes (the same in place)
    uint64_t block_off;  // bytes into the descriptor pool (this request's image)
    uint32_t sweep;      // sweep index within the request
    uint32_t chunk;
    uint32_t rb;
    uint32_t init;       // first sweep: synthesise |0...0>
    uint32_t off_ph, off_op, off_cf;  // byte offsets of phases/ops/coefs inside the image
    uint32_t oop;        // out of place: the sweep's load and store layouts differ
};

template <class A>
__global__ void __launch_bounds__(256) k_serve_step(A* __restrict__ arena, const ServeTask* __restrict__ tasks,
                                                   const char* __restrict__ pool) {
    extern __shared__ __align__(16) unsigned char smem[];
    A* lds = reinterpret_cast<A*>(smem);
    const ServeTask tk = tasks[blockIdx.x];
    const char* img = pool + tk.block_off;
    const DevSweep* sw = reinterpret_cast<const DevSweep*>(img) + tk.sweep;
    const DevPhase* ph = reinterpret_cast<const DevPhase*>(img + tk.off_ph);
    const DevOp* ops = reinterpret_cast<const DevOp*>(img + tk.off_op);
    const double* coef = reinterpret_cast<const double*>(img + tk.off_cf);
    const A* src = arena + tk.src_off;
    A* dst = arena + tk.dst_off;
    const bool init = tk.init != 0, oop = tk.oop != 0;
    if constexpr (sizeof(A) == 8) {
        switch (tk.rb) {
            case 2: run_chunk<A, 8, 2>(src, dst, oop, sw, tk.chunk, ph, ops, coef, lds, init); break;
            case 3: run_chunk<A, 8, 3>(src, dst, oop, sw, tk.chunk, ph, ops, coef, lds, init); break;
            case 4: run_chunk<A, 8, 4>(src, dst, oop, sw, tk.chunk, ph, ops, coef, lds, init); break;
            default: run_chunk<A, 8, 5>(src, dst, oop, sw, tk.chunk, ph, ops, coef, lds, init); break;
        }
    } else {
        switch (tk.rb) {
            case 2: run_chunk<A, 8, 2>(src, dst, oop, sw, tk.chunk, ph, ops, coef, lds, init); break;
            case 3: run_chunk<A, 8, 3>(src, dst, oop, sw, tk.chunk, ph, ops, coef, lds, init); break;
            default: run_chunk<A, 8, 4>(src, dst, oop, sw, tk.chunk, ph, ops, coef, lds, init); break;
        }
    }
}

// First-fit allocator over [0, size) with 256-byte granularity.
class RangeAlloc {
public:
    void reset(uint64_t size) {
        free_.clear();
        free_[0] = size;
        size_ = size;
    }
    bool alloc(uint64_t n, uint64_t& off) {
        n = (n + 255) & ~uint64_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const uint64_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = (n + 255) & ~uint64_t(255);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0;
};

// A request's window-sweep plan and descriptor image, formed on a planner thread. `state`
// is written last (release) and read by the server's thread (acquire) before the results.
struct PlanJob {
    uint32_t nq = 0;
    int dtype = QLX_C64;
    uint64_t pool_bytes = 0;
    bool allow_oop = false;  // the request may run out of place (two state buffers fit the budget)
    bool oop = false;        // planned out of place
    std::vector<qlx_gate_instr> instrs;
    std::unique_ptr<SweepPlan> plan;
    std::vector<char> image;
    uint32_t off_ph = 0, off_op = 0, off_cf = 0, n_sweeps = 0, rb = 0, log_chunks = 0;
    uint64_t block_bytes = 0;
    int err_code = QLX_OK;
    std::string err;
    std::atomic<int> state{0};  // 0 pending, 1 planned, 2 failed
};

inline size_t align_to(size_t x, size_t a) { return (x + a - 1) / a * a; }

void plan_job(PlanJob& j) {
    qlx_exec_opts o{};
    Geometry geo{};
    geo.elem = j.dtype == QLX_C128 ? 16 : 8;
    geo.logt = 8;
    geo.w = std::min<int>((int)j.nq, j.dtype == QLX_C64 ? 13 : 12);
    geo.rb = geo.w - 8;
    geo.c = j.dtype == QLX_C64 ? 4 : 3;  // one 128-byte cache line of amplitudes
    // out of place: relabelled sweeps, whose windows need not hold the low qubits
    j.oop = j.allow_oop && (int)j.nq > geo.w;
    geo.free = j.oop;
    if (j.oop) o.specialize = QLX_SPECIALIZE_RELABEL;
    j.plan = build_plan_geo(j.nq, geo, j.instrs.data(), j.instrs.size(), nullptr, nullptr, o, /*thorough=*/false);
    if (j.dtype == QLX_C64) densify_phases(*j.plan);
    if (j.oop) assign_layouts(*j.plan, j.nq);
    const SweepPlan& p = *j.plan;
    j.n_sweeps = (uint32_t)p.dsw.size();
    j.rb = (uint32_t)geo.rb;
    j.log_chunks = j.nq - (uint32_t)geo.w;
    const size_t off_ph = align_to(p.dsw.size() * sizeof(DevSweep), 256);
    const size_t off_op = align_to(off_ph + p.dph.size() * sizeof(DevPhase), 256);
    const size_t off_cf = align_to(off_op + p.dop.size() * sizeof(DevOp), 256);
    const size_t total = align_to(off_cf + p.coef.size() * sizeof(double), 256) + 256;
    if (total > j.pool_bytes) fail(QLX_ENOMEM, "descriptor image larger than the pool");
    j.image.assign(total, 0);
    std::memcpy(j.image.data(), p.dsw.data(), p.dsw.size() * sizeof(DevSweep));
    std::memcpy(j.image.data() + off_ph, p.dph.data(), p.dph.size() * sizeof(DevPhase));
    if (!p.dop.empty()) std::memcpy(j.image.data() + off_op, p.dop.data(), p.dop.size() * sizeof(DevOp));
    if (!p.coef.empty()) std::memcpy(j.image.data() + off_cf, p.coef.data(), p.coef.size() * sizeof(double));
    j.off_ph = (uint32_t)off_ph;
    j.off_op = (uint32_t)off_op;
    j.off_cf = (uint32_t)off_cf;
    j.block_bytes = total;
    j.plan.reset();  // the image holds everything the device needs
}

// Planner threads: take jobs in submission order until stopped.
class Planner {
public:
    void start(unsigned n) {
        for (unsigned i = 0; i < n; ++i) threads_.emplace_back([this] { loop(); });
    }
    void push(std::shared_ptr<PlanJob> j) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            todo_.push_back(std::move(j));
        }
        cv_.notify_one();
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& t : threads_) t.join();
        threads_.clear();
    }

private:
    void loop() {
        for (;;) {
            std::shared_ptr<PlanJob> j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stop_ || !todo_.empty(); });
                if (stop_) return;
                j = std::move(todo_.front());
                todo_.pop_front();
            }
            try {
                plan_job(*j);
                j->state.store(1, std::memory_order_release);
            } catch (const Error& e) {
                j->err_code = e.code;
                j->err = e.msg;
                j->state.store(2, std::memory_order_release);
            } catch (const std::exception& e) {
                j->err_code = QLX_EHIP;
                j->err = e.what();
                j->state.store(2, std::memory_order_release);
            }
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::shared_ptr<PlanJob>> todo_;
    std::vector<std::thread> threads_;
    bool stop_ = false;
};

}  // namespace qlx

using namespace qlx;

struct qlx_server {
    int dtype = QLX_C64;
    size_t amp = 8;
    DevBuf arena, pool;
    RangeAlloc arena_alloc, pool_alloc;
    uint64_t budget = 0, pool_bytes = 0, next_id = 1;
    std::chrono::steady_clock::time_point t0;
    enum Status { QUEUED, ACTIVE, DONE };
    struct Req {
        uint64_t id;
        uint32_t nq;
        Status status = QUEUED;
        std::shared_ptr<PlanJob> job;  // the plan and descriptor image (until the upload is complete)
        uint64_t amp_off = 0, state_bytes = 0, block_off = 0, block_bytes = 0;
        // out of place: the second buffer follows the first (alloc_bytes = 2 state_bytes) and the
        // result sits in the buffer the last sweep wrote
        bool oop = false;
        uint64_t alloc_bytes = 0, alt_off = 0, result_off = 0;
        uint32_t off_ph = 0, off_op = 0, off_cf = 0;
        uint32_t next_sweep = 0, n_sweeps = 0, rb = 0, log_chunks = 0;
        double submit_s = 0, admit_s = 0, done_s = 0;
    };
    std::map<uint64_t, Req> reqs;
    std::deque<uint64_t> queue;
    struct Step {
        hipEvent_t ev;
        int slot;
        std::vector<uint64_t> finishing;  // requests whose last sweep is in this launch
    };
    struct Group {
        hipStream_t stream = nullptr;
        std::vector<uint64_t> active;
        std::deque<Step> inflight;
        ServeTask* host_tasks[2] = {nullptr, nullptr};
        DevBuf dev_tasks[2];
        size_t cap[2] = {0, 0};
        int next_slot = 0;
    };
    std::map<uint32_t, Group> groups;
    std::vector<uint64_t> completed;
    std::vector<hipEvent_t> free_events;
    DevBuf scratch;
    hipStream_t aux = nullptr;
    Planner planner;
    uint32_t flags = 0;  // QLX_SERVE_*
};

namespace {

using Group = qlx_server::Group;
using Req = qlx_server::Req;

double now_s(const qlx_server* s) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

hipEvent_t get_event(qlx_server* s) {
    if (!s->free_events.empty()) {
        hipEvent_t e = s->free_events.back();
        s->free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    QLX_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

void destroy_server(qlx_server* s) {
    if (!s) return;
    s->planner.stop();
    for (auto& kv : s->groups) {
        Group& g = kv.second;
        if (g.stream) (void)hipStreamSynchronize(g.stream);
        for (auto& st : g.inflight) (void)hipEventDestroy(st.ev);
        for (int k = 0; k < 2; ++k) {
            if (g.host_tasks[k]) (void)hipHostFree(g.host_tasks[k]);
            g.dev_tasks[k].release();
        }
        if (g.stream) (void)hipStreamDestroy(g.stream);
    }
    for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
    if (s->aux) (void)hipStreamDestroy(s->aux);
    s->arena.release();
    s->pool.release();
    s->scratch.release();
    delete s;
}

// Admission: FIFO while the head request is planned and its state and descriptors fit.
void admit(qlx_server* s) {
    while (!s->queue.empty()) {
        Req& r = s->reqs.at(s->queue.front());
        const int ps = r.job->state.load(std::memory_order_acquire);
        if (ps == 0) return;
        if (ps == 2) {
            // a request that cannot be planned leaves the queue; its error is reported once
            const int code = r.job->err_code;
            const std::string msg = "request " + std::to_string(r.id) + ": " + r.job->err;
            s->queue.pop_front();
            s->reqs.erase(r.id);
            fail(code, msg);
        }
        PlanJob& j = *r.job;
        if (!r.block_bytes) {
            r.oop = j.oop;
            r.alloc_bytes = r.state_bytes * (j.oop ? 2 : 1);
            r.block_bytes = j.block_bytes;
            r.n_sweeps = j.n_sweeps;
            r.rb = j.rb;
            r.log_chunks = j.log_chunks;
            r.off_ph = j.off_ph;
            r.off_op = j.off_op;
            r.off_cf = j.off_cf;
        }
        uint64_t aoff = 0, boff = 0;
        if (!s->arena_alloc.alloc(r.alloc_bytes, aoff)) return;
        if (!s->pool_alloc.alloc(r.block_bytes, boff)) {
            s->arena_alloc.release(aoff, r.alloc_bytes);
            return;
        }
        r.amp_off = aoff / s->amp;
        r.alt_off = r.oop ? r.amp_off + (r.state_bytes / s->amp) : r.amp_off;
        // sweep k reads buffer k % 2 and writes the other: the result is in buffer n_sweeps % 2
        r.result_off = (r.oop && (r.n_sweeps & 1)) ? r.alt_off : r.amp_off;
        r.block_off = boff;
        Group& g = s->groups[r.nq];
        if (!g.stream) QLX_HIP_CHECK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        // upload on the auxiliary stream (never waits for the groups' launches); it is
        // complete before the request's first launch is enqueued
        QLX_HIP_CHECK(hipMemcpyAsync((char*)s->pool.ptr + boff, j.image.data(), j.image.size(),
                                     hipMemcpyHostToDevice, s->aux));
        QLX_HIP_CHECK(hipStreamSynchronize(s->aux));
        r.job.reset();
        r.status = qlx_server::ACTIVE;
        r.admit_s = now_s(s);
        g.active.push_back(r.id);
        s->queue.pop_front();
    }
}

void harvest(qlx_server* s) {
    for (auto& kv : s->groups) {
        Group& g = kv.second;
        while (!g.inflight.empty()) {
            const hipError_t q = hipEventQuery(g.inflight.front().ev);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) QLX_HIP_CHECK(q);
            qlx_server::Step st = std::move(g.inflight.front());
            g.inflight.pop_front();
            s->free_events.push_back(st.ev);
            const double t = now_s(s);
            for (uint64_t id : st.finishing) {
                Req& r = s->reqs.at(id);
                r.status = qlx_server::DONE;
                r.done_s = t;
                s->pool_alloc.release(r.block_off, r.block_bytes);  // descriptors no longer read
                s->completed.push_back(id);
            }
        }
    }
}

void launch_steps(qlx_server* s) {
    for (auto& kv : s->groups) {
        Group& g = kv.second;
        if (g.inflight.size() >= 2) continue;
        std::vector<ServeTask> tasks;
        std::vector<uint64_t> finishing, still;
        for (uint64_t id : g.active) {
            Req& r = s->reqs.at(id);
            const uint32_t nch = 1u << r.log_chunks;
            for (uint32_t k = 0; k < nch; ++k) {
                ServeTask t{};
                const bool odd = (r.next_sweep & 1) != 0;
                t.src_off = odd ? r.alt_off : r.amp_off;
                t.dst_off = r.oop ? (odd ? r.amp_off : r.alt_off) : r.amp_off;
                t.oop = r.oop ? 1u : 0u;
                t.block_off = r.block_off;
                t.sweep = r.next_sweep;
                t.chunk = k;
                t.rb = r.rb;
                t.init = r.next_sweep == 0;
                t.off_ph = r.off_ph;
                t.off_op = r.off_op;
                t.off_cf = r.off_cf;
                tasks.push_back(t);
            }
            if (++r.next_sweep == r.n_sweeps) finishing.push_back(id);
            else still.push_back(id);
        }
        if (tasks.empty()) continue;
        g.active.swap(still);
        const int slot = g.next_slot;
        g.next_slot ^= 1;
        if (g.cap[slot] < tasks.size()) {
            if (g.host_tasks[slot]) QLX_HIP_CHECK(hipHostFree(g.host_tasks[slot]));
            g.host_tasks[slot] = nullptr;
            const size_t cap = std::max<size_t>(tasks.size() * 2, 256);
            QLX_HIP_CHECK(hipHostMalloc((void**)&g.host_tasks[slot], cap * sizeof(ServeTask), hipHostMallocDefault));
            g.dev_tasks[slot].reserve(cap * sizeof(ServeTask), g.stream);
            g.cap[slot] = cap;
        }
        std::memcpy(g.host_tasks[slot], tasks.data(), tasks.size() * sizeof(ServeTask));
        QLX_HIP_CHECK(hipMemcpyAsync(g.dev_tasks[slot].ptr, g.host_tasks[slot], tasks.size() * sizeof(ServeTask),
                                     hipMemcpyHostToDevice, g.stream));
        const size_t lds = s->dtype == QLX_C64 ? (size_t(1) << 13) * 8 : (size_t(1) << 12) * 16;
        if (s->dtype == QLX_C64)
            k_serve_step<float2><<<(unsigned)tasks.size(), 256, lds, g.stream>>>(
                (float2*)s->arena.ptr, (const ServeTask*)g.dev_tasks[slot].ptr, (const char*)s->pool.ptr);
        else
            k_serve_step<double2><<<(unsigned)tasks.size(), 256, lds, g.stream>>>(
                (double2*)s->arena.ptr, (const ServeTask*)g.dev_tasks[slot].ptr, (const char*)s->pool.ptr);
        QLX_HIP_CHECK(hipGetLastError());
        qlx_server::Step st;
        st.ev = get_event(s);
        st.slot = slot;
        st.finishing = std::move(finishing);
        QLX_HIP_CHECK(hipEventRecord(st.ev, g.stream));
        g.inflight.push_back(std::move(st));
    }
}

Req& get_req(qlx_server* s, uint64_t id) {
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) fail(QLX_ERANGE, "unknown request id");
    return it->second;
}

}  // namespace

extern "C" {

int qlx_server_create(int dtype, uint64_t hbm_budget_bytes, qlx_server** out) {
    return guard([&] {
        if (!out) fail(QLX_EINVAL, "null output");
        if (dtype != QLX_C64 && dtype != QLX_C128) fail(QLX_EINVAL, "bad dtype");
        if (hbm_budget_bytes < (1u << 20)) fail(QLX_EINVAL, "HBM budget too small");
        auto* s = new qlx_server();
        try {
            s->dtype = dtype;
            s->amp = amp_size(dtype);
            s->budget = hbm_budget_bytes;
            s->arena.reserve(hbm_budget_bytes, nullptr);
            s->arena_alloc.reset(hbm_budget_bytes);
            s->pool_bytes = 256ull << 20;
            s->pool.reserve(s->pool_bytes, nullptr);
            s->pool_alloc.reset(s->pool_bytes);
            QLX_HIP_CHECK(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
            s->t0 = std::chrono::steady_clock::now();
            s->planner.start(std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2)));
        } catch (...) {
            destroy_server(s);
            throw;
        }
        *out = s;
    });
}

int qlx_server_submit(qlx_server* s, uint32_t nq, const qlx_gate_instr* instrs, size_t n, uint64_t* request_id) {
    return guard([&] {
        if (!s) fail(QLX_EINVAL, "null server");
        if (nq < 10 || nq > 34) fail(QLX_EUNSUPPORTED, "served circuits need 10..34 qubits");
        if (!instrs || n == 0) fail(QLX_EINVAL, "empty circuit");
        Req r;
        r.id = s->next_id++;
        r.nq = nq;
        r.state_bytes = (1ull << nq) * s->amp;
        if (r.state_bytes > s->budget) fail(QLX_ENOMEM, "state larger than the HBM budget");
        for (size_t i = 0; i < n; ++i) {
            const qlx_gate_instr& g = instrs[i];
            if (g.type == QLX_MEASURE || g.type == QLX_CCX || g.type == QLX_CSWAP || g.type == QLX_UNITARY ||
                g.type == QLX_FUSED_BLOCK)
                fail(QLX_EUNSUPPORTED, "served circuits take 1-qubit gates, CX, SWAP and GLOBAL_SWAP");
            validate_instr(g, nq, nullptr, 0, 0);
        }
        auto job = std::make_shared<PlanJob>();
        job->nq = nq;
        job->dtype = s->dtype;
        job->allow_oop = !(s->flags & QLX_SERVE_IN_PLACE) && 2 * r.state_bytes <= s->budget;
        job->pool_bytes = s->pool_bytes;
        job->instrs.assign(instrs, instrs + n);
        r.job = job;
        r.submit_s = now_s(s);
        const uint64_t id = r.id;
        s->reqs.emplace(id, std::move(r));
        s->queue.push_back(id);
        s->planner.push(std::move(job));
        if (request_id) *request_id = id;
    });
}

int qlx_server_set_flags(qlx_server* s, uint32_t flags) {
    return guard([&] {
        if (!s) fail(QLX_EINVAL, "null server");
        if (!s->reqs.empty()) fail(QLX_EINVAL, "server flags are set before the first request");
        if (flags & ~QLX_SERVE_IN_PLACE) fail(QLX_EINVAL, "unknown server flag");
        s->flags = flags;
    });
}

int qlx_server_pump(qlx_server* s, uint32_t* queued, uint32_t* active, uint32_t* inflight) {
    return guard([&] {
        if (!s) fail(QLX_EINVAL, "null server");
        harvest(s);
        admit(s);
        launch_steps(s);
        uint32_t a = 0, f = 0;
        for (auto& kv : s->groups) {
            a += (uint32_t)kv.second.active.size();
            f += (uint32_t)kv.second.inflight.size();
        }
        if (queued) *queued = (uint32_t)s->queue.size();
        if (active) *active = a;
        if (inflight) *inflight = f;
    });
}

int qlx_server_request_info(qlx_server* s, uint64_t id, int* status, double* times_s) {
    return guard([&] {
        if (!s) fail(QLX_EINVAL, "null server");
        const Req& r = get_req(s, id);
        if (status)
            *status = r.status == qlx_server::QUEUED ? QLX_REQ_QUEUED
                      : r.status == qlx_server::ACTIVE ? QLX_REQ_ACTIVE
                                                       : QLX_REQ_DONE;
        if (times_s) {
            times_s[0] = r.submit_s;
            times_s[1] = r.status == qlx_server::QUEUED ? -1.0 : r.admit_s;
            times_s[2] = r.status == qlx_server::DONE ? r.done_s : -1.0;
        }
    });
}

int qlx_server_take_completed(qlx_server* s, uint64_t* ids, double* latency_s, size_t cap, size_t* n_out) {
    return guard([&] {
        if (!s) fail(QLX_EINVAL, "null server");
        const size_t n = std::min(cap, s->completed.size());
        for (size_t i = 0; i < n; ++i) {
            const Req& r = s->reqs.at(s->completed[i]);
            if (ids) ids[i] = r.id;
            if (latency_s) latency_s[i] = r.done_s - r.submit_s;
        }
        s->completed.erase(s->completed.begin(), s->completed.begin() + (long)n);
        if (n_out) *n_out = n;
    });
}

int qlx_server_compare(qlx_server* s, uint64_t id, qlx_state* ref, double* infidelity, double* max_abs) {
    return guard([&] {
        if (!s) fail(QLX_EINVAL, "null server");
        Req& r = get_req(s, id);
        if (r.status != qlx_server::DONE) fail(QLX_EINVAL, "request not completed");
        if (!ref || !ref->amp || ref->nq != r.nq) fail(QLX_EINVAL, "reference state does not match");
        QLX_HIP_CHECK(hipStreamSynchronize(ref->stream));
        double re = 0, im = 0, m = 0, nb = 0;
        compare_amplitudes(ref->amp, ref->dtype, (char*)s->arena.ptr + r.result_off * s->amp, s->dtype, 1ull << r.nq,
                           s->aux, s->scratch, &re, &im, &m, &nb);
        if (infidelity) *infidelity = infidelity_from(re, im, reduce_norm_sq(ref), nb);
        if (max_abs) *max_abs = m;
    });
}

int qlx_server_state_to_host(qlx_server* s, uint64_t id, double* out, size_t dim) {
    return guard([&] {
        if (!s) fail(QLX_EINVAL, "null server");
        Req& r = get_req(s, id);
        if (r.status != qlx_server::DONE) fail(QLX_EINVAL, "request not completed");
        const size_t n = size_t(1) << r.nq;
        if (!out || dim < n) fail(QLX_EINVAL, "output buffer too small");
        const char* src = (const char*)s->arena.ptr + r.result_off * s->amp;
        if (s->dtype == QLX_C128) {
            QLX_HIP_CHECK(hipMemcpy(out, src, n * 16, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> tmp(2 * n);
            QLX_HIP_CHECK(hipMemcpy(tmp.data(), src, n * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < 2 * n; ++i) out[i] = tmp[i];
        }
    });
}

int qlx_server_release(qlx_server* s, uint64_t id) {
    return guard([&] {
        if (!s) fail(QLX_EINVAL, "null server");
        Req& r = get_req(s, id);
        if (r.status != qlx_server::DONE) fail(QLX_EINVAL, "request not completed");
        s->arena_alloc.release(r.amp_off * s->amp, r.alloc_bytes);
        s->reqs.erase(id);
    });
}

int qlx_server_destroy(qlx_server* s) {
    return guard([&] { destroy_server(s); });
}

}  // extern "C"
