// qlx_sweep.hip — windowed execution: many gates per HBM sweep.
//
// This is the GPU re-targeting of the reference's qubit-windowing idea
// (include/middleware/middleware.hpp:759-860: windows of <= k_safe active qubits).
// Instead of moving high qubits into a fixed low window with GLOBAL_SWAP passes,
// each sweep chooses its own window S of W qubits (always containing the low C
// qubits, so every HBM access is a contiguous 2^C-amplitude segment). One
// work-group owns one 2^W-amplitude chunk of the state (all amplitudes that share
// the bits outside S), keeps it in registers (2^RB amplitudes per work-item,
// T = 2^LOGT work-items, W = LOGT + RB) and applies every gate of the sweep there:
//   * a gate whose qubits are "register bits" of the current layout is pure VALU work;
//   * between phases the chunk is re-laid out through LDS (one write + one read of
//     the chunk), which moves a new set of qubits into the register bits.
// The state is read and written once per sweep: algorithmic bytes 2 * sizeof(amp) * 2^N.
//
// The host planner (below) builds the sweeps from the gate DAG (gates commute when
// their qubits are disjoint), the register phases inside each sweep, and the device
// descriptors. GLOBAL_SWAP is a real qubit permutation here (a SWAP inside a sweep);
// CCX/CSWAP and dense gates on 3+ qubits run as their own plain kernels between sweeps.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <complex>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "qlx_jit.hpp"
#include "qlx_sweep_core.hpp"

namespace qlx {

// Persistent sweep: work-group b runs chunks b, b + G, b + 2G, ... (G = grid size, the
// resident work-groups of the device). With up to 16 amplitude registers per work-item,
// the next chunk's HBM loads are issued into a second register set before the current
// chunk's phases run, so they overlap its compute. OOP: out of place (relabelled plans),
// reading `src` and writing `dst` through the sweep's load and store layouts; in place
// src == dst.
template <class A, int LOGT, int RB, bool OOP>
__global__ void __launch_bounds__(1 << LOGT) k_sweep(const A* src, A* dst, const DevSweep* __restrict__ sw,
                                                    const DevPhase* __restrict__ ph, const DevOp* __restrict__ ops,
                                                    const double* __restrict__ coef, uint64_t nchunks) {
    extern __shared__ __align__(16) unsigned char smem[];
    A* lds = reinterpret_cast<A*>(smem);
    uint64_t c = blockIdx.x;
    if (c >= nchunks) return;
    auto ld_base = [&](uint64_t k) { return OOP ? chunk_base_oop(sw->ld_cpos, sw->nc, k) : chunk_base<LOGT, RB>(sw, k); };
    auto st_base = [&](uint64_t k, uint64_t lb) { return OOP ? chunk_base_oop(sw->st_cpos, sw->nc, k) : lb; };
    // 32 registers of amplitudes per work-item leave no room for a second set
    constexpr bool kPrefetch = RB <= 4;
    if constexpr (!kPrefetch) {
        // one chunk per work-group (grid = chunks): a loop here would keep every register's
        // offsets live across iterations and push the kernel past the register file
        A v[1 << RB];
        const uint64_t base = ld_base(c);
        load_chunk<A, LOGT, RB>(v, src, sw, base);
        compute_chunk<A, LOGT, RB>(v, sw, ph, ops, coef, lds);
        store_chunk<A, LOGT, RB>(v, dst, sw, st_base(c, base));
        return;
    }
    A v[1 << RB], w[1 << RB];
    uint64_t base = ld_base(c);
    load_chunk<A, LOGT, RB>(v, src, sw, base);
    for (;;) {
        const uint64_t nc = c + gridDim.x;
        const bool more = nc < nchunks;
        const uint64_t nbase = more ? ld_base(nc) : 0;
        if (more) load_chunk<A, LOGT, RB>(w, src, sw, nbase);
        compute_chunk<A, LOGT, RB>(v, sw, ph, ops, coef, lds);
        store_chunk<A, LOGT, RB>(v, dst, sw, st_base(c, base));
        if (!more) break;
#pragma unroll
        for (int i = 0; i < (1 << RB); ++i) v[i] = w[i];
        c = nc;
        base = nbase;
    }
}

// ---------------------------------------------------------------------------------
// host side: planning
// ---------------------------------------------------------------------------------
// (window qubits, register qubits) pairs with a compiled k_sweep instantiation.
// Geometries of the generic sweep kernel (template instances below). Circuit-specialised
// sweeps are generated per geometry, so they also take any window whose chunk fits in
// 128 KiB of LDS with 64..1024 work-items and 8..32 registers per work-item.
static bool geometry_built(int dtype, int w, int rb, bool specialised) {
    if (specialised) {
        const size_t elem = dtype == QLX_C128 ? 16 : 8;
        const int logt = w - rb;
        return w <= kMaxW && rb >= 3 && rb <= 5 && logt >= 6 && logt <= kMaxLogT && (elem << w) <= (size_t(128) << 10);
    }
    if (dtype == QLX_C128) return (w == 12 && (rb == 4 || rb == 5)) || (w == 11 && (rb == 4 || rb == 3));
    return (w == 13 && rb == 5) || (w == 12 && (rb == 5 || rb == 4));
}

static Geometry geometry_for(int dtype, uint32_t nq, const qlx_exec_opts& o) {
    Geometry g{};
    g.elem = (int)amp_size(dtype);
    g.w = o.window_qubits ? (int)o.window_qubits : (dtype == QLX_C128 ? 12 : 13);
    g.rb = o.register_qubits ? (int)o.register_qubits : (dtype == QLX_C128 ? 4 : 5);
    if (!geometry_built(dtype, g.w, g.rb, (o.specialize & QLX_SPECIALIZE) != 0))
        fail(QLX_EUNSUPPORTED, "window " + std::to_string(g.w) + " with " + std::to_string(g.rb) +
                                   " register qubits is not built for this precision" +
                                   ((o.specialize & QLX_SPECIALIZE) ? "" : " (specialised sweeps take more windows)"));
    g.logt = g.w - g.rb;
    // low qubits kept in the window: one 128-byte cache line of amplitudes (8 x 16 B, 16 x 8 B)
    g.c = o.coalesce_qubits ? (int)o.coalesce_qubits : (dtype == QLX_C64 ? 4 : 3);
    if (g.c < 0 || g.c > g.logt) fail(QLX_EINVAL, "coalesce_qubits out of range");
    g.free = (o.specialize & QLX_SPECIALIZE_RELABEL) && !(o.plan_flags & QLX_PLAN_FIXED_WINDOWS);
    (void)nq;
    return g;
}

struct PGate {
    int src;            // index in the instruction stream
    uint8_t type;
    int nq;             // operand count
    int q[5];
    bool sweepable;
    // a CX of a circuit-specialised sweep: its control may sit on a work-item bit (OP_CXT), so a
    // register phase can run it when its target alone is in the registers
    bool ctl_free = false;
    int merged = -1;    // a rotation that absorbed later ones: its index in the MergedRot table
};

// A rotation RX/RY/RZ that absorbed later rotations of the same axis on its qubit: the summed
// angle and the source instructions in program order (the first is the PGate's own src).
struct MergedRot {
    double theta;
    std::vector<int> srcs;
};

// Whether operand k of a gate of type `t` commutes with a rotation about axis a (0 X, 1 Y,
// 2 Z) on that operand's qubit: X with X and with the target of CX/CCX (X on the target,
// conditioned on the controls); Y with Y; Z with the diagonal Z and S and with the controls
// of CX/CCX/CSWAP (the gate is block-diagonal in the control's basis states).
static bool commutes_with_rotation(uint8_t t, int k, int a) {
    switch (a) {
        case 0: return t == QLX_X || (t == QLX_CX && k == 1) || (t == QLX_CCX && k == 2);
        case 1: return t == QLX_Y;
        default:
            return t == QLX_Z || t == QLX_S || (t == QLX_CX && k == 0) || (t == QLX_CCX && k < 2) ||
                   (t == QLX_CSWAP && k == 0);
    }
}

// Merges every sweepable rotation into the latest earlier rotation of the same axis on its
// qubit when each gate between the two on that qubit commutes with it: the later rotation
// moves back past those gates (and past gates on other qubits, trivially), and
// R(b)·R(a) = R(a + b) about one axis. The angles are the instructions' floats, summed in
// double precision (exact for angles within a few turns of each other). Removes the merged
// gates from `gates`; their sources stay listed in the table entry of the gate they joined.
static void merge_rotations(std::vector<PGate>& gates, const qlx_gate_instr* instrs, uint32_t nq,
                            std::vector<MergedRot>& merged) {
    std::vector<int> open[3];
    for (auto& o : open) o.assign(nq, -1);
    std::vector<char> drop(gates.size(), 0);
    for (size_t i = 0; i < gates.size(); ++i) {
        PGate& g = gates[i];
        const int axis = g.type == QLX_RX ? 0 : g.type == QLX_RY ? 1 : g.type == QLX_RZ ? 2 : -1;
        if (axis >= 0 && g.sweepable) {
            const int q = g.q[0];
            const int j = open[axis][q];
            if (j >= 0) {
                PGate& h = gates[j];
                if (h.merged < 0) {
                    h.merged = (int)merged.size();
                    merged.push_back(MergedRot{static_cast<double>(instrs[h.src].payload.param), {h.src}});
                }
                merged[h.merged].theta += static_cast<double>(instrs[g.src].payload.param);
                merged[h.merged].srcs.push_back(g.src);
                drop[i] = 1;
                continue;
            }
            for (int a = 0; a < 3; ++a) open[a][q] = a == axis ? (int)i : -1;
            continue;
        }
        for (int k = 0; k < g.nq; ++k)
            for (int a = 0; a < 3; ++a)
                if (!commutes_with_rotation(g.type, k, a)) open[a][g.q[k]] = -1;
    }
    size_t w = 0;
    for (size_t i = 0; i < gates.size(); ++i)
        if (!drop[i]) gates[w++] = gates[i];
    gates.resize(w);
}

static bool is_sweepable(const qlx_gate_instr& g, const qlx_matrix_entry* meta) {
    switch (g.type) {
        case QLX_H: case QLX_X: case QLX_Y: case QLX_Z: case QLX_S: case QLX_RX: case QLX_RY: case QLX_RZ:
        case QLX_CX: case QLX_SWAP: case QLX_GLOBAL_SWAP:
            return true;
        case QLX_UNITARY:
        case QLX_FUSED_BLOCK:
            return meta && instr_arity(g) <= 2;
        default:
            return false;
    }
}

static int operand_count(const qlx_gate_instr& g) {
    switch (g.type) {
        case QLX_CX: case QLX_SWAP: case QLX_GLOBAL_SWAP: return 2;
        case QLX_CCX: case QLX_CSWAP: return 3;
        case QLX_UNITARY: case QLX_FUSED_BLOCK: return instr_arity(g);
        default: return 1;
    }
}

// Marks (in `run`) the pending gates one sweep over the qubit set `in_set` can apply: a
// gate runs when it is sweepable, its qubits are in the set and no earlier pending gate on
// any of its qubits stays behind (a ctl_free CX: its target is in the set and neither of its
// qubits is held back). Scanning in program order, a gate that stays behind blocks its
// qubits for every later gate (each later gate on such a qubit depends on it), so the scan
// ends once every qubit of the set is blocked. Returns the count.
static int count_runnable(const std::vector<PGate>& gates, const std::vector<int>& pending,
                          const std::vector<char>& done, const std::vector<char>& in_set,
                          std::vector<char>& run, const std::vector<std::vector<int>>& preds) {
    (void)done;
    (void)preds;
    const int nbits = (int)in_set.size();
    int open = 0;
    for (int b = 0; b < nbits; ++b) open += in_set[b] != 0;
    std::vector<char> blocked(nbits, 0);
    int cnt = 0;
    size_t k = 0;
    for (; k < pending.size() && open > 0; ++k) {
        const int gi = pending[k];
        const PGate& g = gates[gi];
        bool ok = g.sweepable;
        if (g.ctl_free)
            ok = ok && in_set[g.q[1]] && !blocked[g.q[1]] && !blocked[g.q[0]];
        else
            for (int j = 0; ok && j < g.nq; ++j) ok = in_set[g.q[j]] && !blocked[g.q[j]];
        run[gi] = ok;
        if (ok) {
            ++cnt;
            continue;
        }
        for (int j = 0; j < g.nq; ++j) {
            const int q = g.q[j];
            if (!blocked[q]) {
                blocked[q] = 1;
                open -= in_set[q] != 0;
            }
        }
    }
    for (; k < pending.size(); ++k) run[pending[k]] = 0;
    return cnt;
}

// `theta`, when given, replaces a rotation's angle (a merged rotation).
static void fill_coef(std::vector<double>& coef, const qlx_gate_instr& g, const double* pool,
                      const qlx_matrix_entry* meta, uint8_t& kind, const double* theta = nullptr) {
    const double angle = theta ? *theta : static_cast<double>(g.payload.param);
    auto push_u2 = [&](double m00r, double m00i, double m01r, double m01i, double m10r, double m10i, double m11r,
                       double m11i) {
        kind = OP_U2;
        double m[8] = {m00r, m00i, m01r, m01i, m10r, m10i, m11r, m11i};
        coef.insert(coef.end(), m, m + 8);
    };
    switch (g.type) {
        case QLX_H:
            kind = OP_H;
            coef.push_back(1.0 / std::sqrt(2.0));
            break;
        case QLX_RX: {
            kind = OP_RX;
            const double half = angle * 0.5;
            coef.push_back(std::cos(half));
            coef.push_back(std::sin(half));
            break;
        }
        case QLX_RY: {
            const double half = angle * 0.5;
            const double c = std::cos(half), s = std::sin(half);
            push_u2(c, 0, -s, 0, s, 0, c, 0);
            break;
        }
        case QLX_RZ: {
            const double half = angle * 0.5;
            push_u2(std::cos(-half), std::sin(-half), 0, 0, 0, 0, std::cos(half), std::sin(half));
            break;
        }
        case QLX_X: push_u2(0, 0, 1, 0, 1, 0, 0, 0); break;
        case QLX_Y: push_u2(0, 0, 0, -1, 0, 1, 0, 0); break;
        case QLX_Z: push_u2(1, 0, 0, 0, 0, 0, -1, 0); break;
        case QLX_S: push_u2(1, 0, 0, 0, 0, 0, 0, 1); break;
        case QLX_CX: kind = OP_CX; break;
        case QLX_SWAP: case QLX_GLOBAL_SWAP: kind = OP_SWAP; break;
        case QLX_UNITARY:
        case QLX_FUSED_BLOCK: {
            const int k = instr_arity(g);
            const qlx_matrix_entry& e = meta[g.payload.matrix_idx];
            const double* m = pool + 2 * (size_t)e.offset;
            coef.insert(coef.end(), m, m + 2 * (size_t)e.num_elements);
            kind = k == 1 ? OP_U2 : OP_U4;
            break;
        }
        default:
            fail(QLX_EUNSUPPORTED, "gate cannot run inside a sweep");
    }
}

// Greedily chooses `want` qubits from `allowed` (seeded with `seed`) maximising the
// number of runnable pending gates.
static std::vector<int> choose_set(const std::vector<PGate>& gates, const std::vector<int>& pending,
                                   const std::vector<char>& done, const std::vector<std::vector<int>>& preds,
                                   int nbits, std::vector<char> in_set, int have, int want,
                                   const std::vector<char>& allowed) {
    std::vector<char> run(gates.size(), 0);
    // first-use rank of each bit among pending gates, for tie-breaking
    std::vector<int> rank(nbits, 1 << 30);
    int r = 0;
    for (int gi : pending)
        for (int k = 0; k < gates[gi].nq; ++k) rank[gates[gi].q[k]] = std::min(rank[gates[gi].q[k]], r++);
    while (have < want) {
        int best = -1, best_score = -1, best_rank = 1 << 30;
        for (int b = 0; b < nbits; ++b) {
            if (in_set[b] || !allowed[b]) continue;
            in_set[b] = 1;
            const int sc = count_runnable(gates, pending, done, in_set, run, preds);
            in_set[b] = 0;
            if (sc > best_score || (sc == best_score && rank[b] < best_rank)) {
                best = b;
                best_score = sc;
                best_rank = rank[b];
            }
        }
        if (best < 0) break;
        in_set[best] = 1;
        ++have;
    }
    std::vector<int> out;
    for (int b = 0; b < nbits; ++b)
        if (in_set[b]) out.push_back(b);
    return out;
}

// Plain-kernel cost of a gate in full-state passes (what launch_plain_instr touches).
static double plain_passes(uint8_t type) {
    switch (type) {
        case QLX_Z: case QLX_S: case QLX_CX: case QLX_SWAP: case QLX_GLOBAL_SWAP: return 0.5;
        case QLX_CCX: case QLX_CSWAP: return 0.25;
        case QLX_BARRIER: return 0.0;
        default: return 1.0;
    }
}

// Vector-unit work of a gate inside a sweep, in units of one one-qubit rotation or H over the
// chunk (CX and SWAP: a quarter, their cost being the register phases they add).
static double gate_work(const PGate& g) {
    switch (g.type) {
        case QLX_CX: case QLX_SWAP: case QLX_GLOBAL_SWAP: return 0.25;
        default: return g.nq >= 2 ? 4.0 : 1.0;
    }
}

// The planner's walk over the gate DAG: plain-only gates run as soon as they are ready;
// otherwise the next sweep takes every pending gate its window can run.
struct Walk {
    const std::vector<PGate>* gates;
    const std::vector<std::vector<int>>* preds;
    std::vector<char> done;
    std::vector<int> pending;
    size_t n_done = 0;

    Walk(const std::vector<PGate>& g, const std::vector<std::vector<int>>& p)
        : gates(&g), preds(&p), done(g.size(), 0), pending(g.size()) {
        for (size_t i = 0; i < g.size(); ++i) pending[i] = (int)i;
    }
    bool ready(int gi) const {
        for (int p : (*preds)[gi])
            if (!done[p]) return false;
        return true;
    }
    void prune() {
        std::vector<int> still;
        still.reserve(pending.size());
        for (int gi : pending)
            if (!done[gi]) still.push_back(gi);
        pending.swap(still);
    }
    // runs every plain-only gate that is (or becomes) ready; returns them in the order run
    std::vector<int> run_plain() {
        std::vector<int> ran;
        bool progress = true;
        while (progress) {
            progress = false;
            for (int gi : pending) {
                if (done[gi] || (*gates)[gi].sweepable || !ready(gi)) continue;
                done[gi] = 1;
                ++n_done;
                ran.push_back(gi);
                progress = true;
            }
            prune();
        }
        return ran;
    }
    // the first ready sweepable gate, or -1
    int seed_gate() const {
        for (int gi : pending)
            if ((*gates)[gi].sweepable && ready(gi)) return gi;
        return -1;
    }
    // the gates one sweep over window S runs, in program order; marks them done
    // with a work budget (> 0): the runnable gates in program order until their work reaches
    // it (a program-order prefix of runnable gates holds every predecessor it needs)
    double budget = 0.0;
    std::vector<int> take(const std::vector<int>& S, int nq) {
        std::vector<char> sset(nq, 0);
        for (int q : S) sset[q] = 1;
        std::vector<char> run(gates->size(), 0);
        count_runnable(*gates, pending, done, sset, run, *preds);
        std::vector<int> out;
        double work = 0.0;
        for (int gi : pending)
            if (run[gi]) {
                if (budget > 0.0 && work >= budget) break;
                out.push_back(gi);
                done[gi] = 1;
                ++n_done;
                work += gate_work((*gates)[gi]);
            }
        prune();
        return out;
    }
};

// Register sets of one sweep's phases, by beam search: each step extends every kept walk
// by each candidate set and keeps the `width` walks that have run the most gates. The
// candidates are the greedy extension of the first ready gate's qubits (the first phase:
// the greedy set) and every run of RB consecutive local qubits (every RB-subset instead
// gave no fewer phases on the layered circuits); the first phase may not
// hold the low C local qubits in registers (its layout is the coalesced load's). Among
// the walks that finish first, one whose last set leaves the low C qubits out of the
// registers is preferred (otherwise the store needs one more re-layout).
static std::vector<std::vector<int>> beam_regsets(const std::vector<PGate>& lg, const std::vector<std::vector<int>>& preds,
                                                  const Geometry& geo, int width, bool runs,
                                                  const std::vector<int>& st_low) {
    const int W = geo.w, RB = geo.rb, C = geo.c;
    struct Node {
        Walk w;
        std::vector<std::vector<int>> sets;
    };
    auto candidates = [&](const Walk& w, bool first) {
        std::vector<std::vector<int>> out;
        std::vector<char> allowed(W, 1), seed(W, 0);
        int have = 0;
        if (first) {
            for (int l = 0; l < C; ++l) allowed[l] = 0;
        } else {
            const int g0 = w.seed_gate();
            if (g0 >= 0)
                for (int k = 0; k < lg[g0].nq; ++k)
                    if (!seed[lg[g0].q[k]]) {
                        seed[lg[g0].q[k]] = 1;
                        ++have;
                    }
        }
        out.push_back(choose_set(lg, w.pending, w.done, preds, W, seed, have, RB, allowed));
        for (int a = first ? C : 0; runs && a + RB <= W; ++a) {
            std::vector<int> S;
            for (int l = a; l < a + RB; ++l) S.push_back(l);
            if (S != out.front()) out.push_back(std::move(S));
        }
        return out;
    };
    // the store's coalesced local bits (st_low) must be work-item bits of the last layout
    auto low_in = [&](const std::vector<int>& S) {
        for (int l : S)
            if (std::find(st_low.begin(), st_low.end(), l) != st_low.end()) return true;
        return false;
    };
    std::vector<Node> beam;
    beam.push_back(Node{Walk(lg, preds), {}});
    for (size_t step = 0; step <= lg.size() + 1; ++step) {
        const Node* fin = nullptr;
        for (const Node& nd : beam)
            if (nd.w.pending.empty() && (!fin || (low_in(fin->sets.back()) && !low_in(nd.sets.back())))) fin = &nd;
        if (fin) return fin->sets;
        std::vector<Node> next;
        for (const Node& nd : beam) {
            const bool first = nd.sets.empty();
            for (const std::vector<int>& S : candidates(nd.w, first)) {
                Node c = nd;
                if (c.w.take(S, W).empty() && !first) continue;
                c.sets.push_back(S);
                next.push_back(std::move(c));
            }
        }
        std::stable_sort(next.begin(), next.end(), [](const Node& x, const Node& y) { return x.w.n_done > y.w.n_done; });
        beam.clear();
        for (Node& c : next) {
            bool dup = false;
            for (const Node& k : beam) dup = dup || k.w.done == c.w.done;
            if (dup) continue;
            beam.push_back(std::move(c));
            if ((int)beam.size() >= width) break;
        }
        if (beam.empty()) fail(QLX_EHIP, "sweep phase planner made no progress");
    }
    fail(QLX_EHIP, "sweep phase planner did not finish");
}

// Register phases of one sweep (gates given in local bits). thorough: the beam search over
// register sets (width 8, with the runs of consecutive qubits); otherwise the greedy set
// alone at each step (width 1), which plans several times faster.
static void plan_phases(PlanSweep& sw, const std::vector<PGate>& lg, const Geometry& geo, bool thorough,
                        const std::vector<int>& st_low) {
    const int W = geo.w, RB = geo.rb, LT = geo.logt;
    const int n = (int)lg.size();
    std::vector<std::vector<int>> preds(n);
    {
        std::vector<int> last(W, -1);
        for (int i = 0; i < n; ++i) {
            for (int k = 0; k < lg[i].nq; ++k) {
                const int b = lg[i].q[k];
                if (last[b] >= 0) preds[i].push_back(last[b]);
                last[b] = i;
            }
        }
    }
    std::vector<char> done(n, 0);
    std::vector<int> pending(n);
    for (int i = 0; i < n; ++i) pending[i] = i;

    auto make_layout = [&](const std::vector<int>& regset) {
        std::vector<int> st2loc(W, -1);
        std::vector<char> isreg(W, 0);
        for (int b : regset) isreg[b] = 1;
        int ti = 0;
        for (int l = 0; l < W; ++l)
            if (!isreg[l]) st2loc[ti++] = l;          // work-item bits in ascending local order
        int ri = 0;
        for (int l : regset) st2loc[LT + ri++] = l;    // register bits
        return st2loc;
    };
    auto regset_of = [&](const std::vector<int>& st2loc) {
        std::vector<int> r;
        for (int s = LT; s < W; ++s) r.push_back(st2loc[s]);
        return r;
    };

    const std::vector<std::vector<int>> sets = beam_regsets(lg, preds, geo, thorough ? 8 : 1, thorough, st_low);
    bool first = true;
    for (const std::vector<int>& regset : sets) {
        PlanPhase phase;
        phase.st2loc = make_layout(regset);
        std::vector<int> loc2st(W);
        for (int s = 0; s < W; ++s) loc2st[phase.st2loc[s]] = s;
        std::vector<char> in_set(W, 0);
        for (int b : regset) in_set[b] = 1;
        std::vector<char> run(n, 0);
        count_runnable(lg, pending, done, in_set, run, preds);
        std::vector<int> still;
        for (int gi : pending) {
            if (run[gi]) {
                PlanOp op{};
                op.src = lg[gi].src;
                op.l0 = loc2st[lg[gi].q[0]] - LT;
                op.l1 = lg[gi].nq > 1 ? loc2st[lg[gi].q[1]] - LT : -1;
                if (lg[gi].ctl_free && op.l0 < 0) {
                    op.tctl = loc2st[lg[gi].q[0]];  // the control is a work-item bit
                    op.l0 = -1;
                }
                op.kind = lg[gi].type;  // instruction type; resolved to an OpKind later
                op.merged = lg[gi].merged;
                phase.ops.push_back(op);
                done[gi] = 1;
            } else {
                still.push_back(gi);
            }
        }
        pending.swap(still);
        if (!phase.ops.empty() || first) sw.phases.push_back(std::move(phase));
        first = false;
    }
    if (!pending.empty()) fail(QLX_EHIP, "sweep phase planner left gates behind");
    // the last layout holds the store's coalesced local bits (st_low) as work-item bits
    // 0..C-1: one more re-layout when its registers hold one of them, or when the sweep has a
    // single phase (the load's layout) and the store's coalesced bits differ from the load's
    auto is_st_low = [&](int l) { return std::find(st_low.begin(), st_low.end(), l) != st_low.end(); };
    std::vector<int> last = regset_of(sw.phases.back().st2loc);
    bool low_in_regs = false;
    for (int l : last) low_in_regs = low_in_regs || is_st_low(l);
    bool same_low = true;
    for (size_t j = 0; j < st_low.size(); ++j) same_low = same_low && st_low[j] == (int)j;
    if (low_in_regs || (sw.phases.size() == 1 && !same_low)) {
        std::vector<int> rs;
        if (low_in_regs) {
            for (int l = W - 1; l >= 0 && (int)rs.size() < RB; --l)
                if (!is_st_low(l)) rs.push_back(l);
            std::sort(rs.begin(), rs.end());
        } else {
            rs = last;
        }
        PlanPhase phase;
        phase.st2loc = make_layout(rs);
        sw.phases.push_back(std::move(phase));
    }
    // order the last layout's work-item bits: st_low first, then the others ascending (a
    // work-item bit named by an op of that phase is renumbered)
    PlanPhase& lp = sw.phases.back();
    std::vector<int> st2loc;
    for (int l : st_low) st2loc.push_back(l);
    for (int b = 0; b < LT; ++b)
        if (!is_st_low(lp.st2loc[b])) st2loc.push_back(lp.st2loc[b]);
    for (int b = LT; b < W; ++b) st2loc.push_back(lp.st2loc[b]);
    if ((int)st2loc.size() != W) fail(QLX_EHIP, "sweep phase planner: store layout");
    std::vector<int> loc2st(W);
    for (int b = 0; b < W; ++b) loc2st[st2loc[b]] = b;
    for (PlanOp& op : lp.ops)
        if (op.tctl >= 0) op.tctl = loc2st[lp.st2loc[op.tctl]];
    lp.st2loc = std::move(st2loc);
}

// Candidate windows for the next sweep, each the low C qubits plus W - C more: (a) the
// greedy extension of the first ready gate's operands (the classic choice), (b) every run
// of W - C consecutive qubits. The one that runs the most pending gates wins (ties: the
// earlier candidate). On circuits of nearest-neighbour layers, (b) lets a sweep advance a
// band of the qubit line by several layers at once: the 30-qubit layered h/rx/cx circuit
// of depth 20 takes 15 sweeps of a 12-qubit window instead of 21.
static std::vector<int> choose_window_wide(const std::vector<PGate>& gates, const std::vector<int>& pending,
                                          const std::vector<char>& done, const std::vector<std::vector<int>>& preds,
                                          int nq, const Geometry& geo, int seed_gate) {
    const int W = geo.w, C = std::min(geo.c, nq);
    std::vector<std::vector<char>> cands;
    auto base = [&] {
        std::vector<char> s(nq, 0);
        for (int l = 0; l < C; ++l) s[l] = 1;
        return s;
    };
    {
        std::vector<char> s = base();
        int have = C;
        for (int k = 0; k < gates[seed_gate].nq; ++k)
            if (!s[gates[seed_gate].q[k]]) {
                s[gates[seed_gate].q[k]] = 1;
                ++have;
            }
        const std::vector<char> allowed(nq, 1);
        const std::vector<int> S = choose_set(gates, pending, done, preds, nq, s, have, W, allowed);
        std::vector<char> c(nq, 0);
        for (int q : S) c[q] = 1;
        cands.push_back(std::move(c));
    }
    for (int a = C; a + (W - C) <= nq; ++a) {
        std::vector<char> c = base();
        for (int q = a; q < a + (W - C); ++q) c[q] = 1;
        cands.push_back(std::move(c));
    }
    int best = 0, best_n = -1;
    std::vector<char> run(gates.size(), 0);
    for (size_t k = 0; k < cands.size(); ++k) {
        const int cnt = count_runnable(gates, pending, done, cands[k], run, preds);
        if (cnt > best_n) {
            best = (int)k;
            best_n = cnt;
        }
    }
    std::vector<int> S;
    for (int q = 0; q < nq; ++q)
        if (cands[best][q]) S.push_back(q);
    return S;
}

static int overlap(const std::vector<int>& a, const std::vector<int>& b) {
    int n = 0;
    for (int x : a)
        n += std::find(b.begin(), b.end(), x) != b.end();
    return n;
}

// Candidate windows for the next sweep (see choose_window_wide), as qubit lists. Free
// windows (Geometry::free; `prev` the previous window, if any): also every run of W
// consecutive qubits and the greedy extension of the first ready gate's operands alone,
// each kept only when it shares at least C qubits with `prev` (the previous sweep's store
// is coalesced on shared qubits). The first window holds qubits 0..C-1 (the logical layout's
// load is coalesced on them): the fixed candidates.
static std::vector<std::vector<int>> window_candidates(const Walk& w, int nq, const Geometry& geo, int seed_gate,
                                                       const std::vector<int>* prev = nullptr) {
    const int W = geo.w, C = std::min(geo.c, nq);
    std::vector<std::vector<int>> out;
    if (geo.free && prev && !prev->empty()) {
        std::vector<std::vector<int>> cands;
        {
            std::vector<char> s(nq, 0);
            int have = 0;
            const PGate& g = (*w.gates)[seed_gate];
            for (int k = 0; k < g.nq; ++k)
                if (!s[g.q[k]]) {
                    s[g.q[k]] = 1;
                    ++have;
                }
            const std::vector<char> allowed(nq, 1);
            cands.push_back(choose_set(*w.gates, w.pending, w.done, *w.preds, nq, s, have, W, allowed));
        }
        for (int a = 0; a + W <= nq; ++a) {
            std::vector<int> S;
            for (int q = a; q < a + W; ++q) S.push_back(q);
            cands.push_back(std::move(S));
        }
        Geometry fixed = geo;
        fixed.free = false;
        for (std::vector<int>& S : window_candidates(w, nq, fixed, seed_gate)) cands.push_back(std::move(S));
        for (std::vector<int>& S : cands) {
            if (overlap(S, *prev) < C) continue;
            bool dup = false;
            for (const std::vector<int>& o : out) dup = dup || o == S;
            if (!dup) out.push_back(std::move(S));
        }
        return out;
    }
    {
        std::vector<char> s(nq, 0);
        int have = 0;
        for (int l = 0; l < C; ++l, ++have) s[l] = 1;
        const PGate& g = (*w.gates)[seed_gate];
        for (int k = 0; k < g.nq; ++k)
            if (!s[g.q[k]]) {
                s[g.q[k]] = 1;
                ++have;
            }
        const std::vector<char> allowed(nq, 1);
        out.push_back(choose_set(*w.gates, w.pending, w.done, *w.preds, nq, s, have, W, allowed));
    }
    for (int a = C; a + (W - C) <= nq; ++a) {
        std::vector<int> S;
        for (int l = 0; l < C; ++l) S.push_back(l);
        for (int q = a; q < a + (W - C); ++q) S.push_back(q);
        if (S != out.front()) out.push_back(std::move(S));
    }
    return out;
}

// The last window of a free-window walk (Geometry::free): qubits 0..C-1, the qubits its
// gates `took` use, then qubits of S that the previous window holds, of S, and of the
// previous window, up to W; sorted. Empty when 0..C-1 and the gates' qubits exceed W, or
// when the result shares fewer than C qubits with the previous window.
static std::vector<int> close_window(const std::vector<int>& S, const std::vector<int>& took,
                                     const std::vector<PGate>& gates, const std::vector<int>& prev, int C, int W) {
    std::vector<int> out;
    auto add = [&](int q) {
        if ((int)out.size() < W && std::find(out.begin(), out.end(), q) == out.end()) out.push_back(q);
    };
    for (int q = 0; q < C; ++q) add(q);
    for (int gi : took)
        for (int k = 0; k < gates[gi].nq; ++k) {
            const int q = gates[gi].q[k];
            if (std::find(out.begin(), out.end(), q) != out.end()) continue;
            if ((int)out.size() == W) return {};
            out.push_back(q);
        }
    for (int q : S)
        if (std::find(prev.begin(), prev.end(), q) != prev.end()) add(q);
    for (int q : S) add(q);
    for (int q : prev) add(q);
    if ((int)out.size() != W || overlap(out, prev) < C) return {};
    std::sort(out.begin(), out.end());
    return out;
}

// Beam search over window sequences: each step extends every kept walk by each candidate
// window and keeps the `width` walks that have run the most gates (distinct states only).
// Returns the windows of the first walk to finish.
static std::vector<std::vector<int>> beam_windows(const std::vector<PGate>& gates,
                                                  const std::vector<std::vector<int>>& preds, int nq,
                                                  const Geometry& geo, int width, double budget = 0.0) {
    struct Node {
        Walk w;
        std::vector<std::vector<int>> wins;
    };
    std::vector<Node> beam;
    beam.push_back(Node{Walk(gates, preds), {}});
    beam[0].w.budget = budget;
    beam[0].w.run_plain();
    for (size_t step = 0; step <= gates.size(); ++step) {
        for (const Node& nd : beam)
            if (nd.w.pending.empty()) return nd.wins;
        std::vector<Node> next;
        for (const Node& nd : beam) {
            const int seed = nd.w.seed_gate();
            if (seed < 0) fail(QLX_EHIP, "sweep planner: no runnable gate");
            const std::vector<int>* prev = nd.wins.empty() ? nullptr : &nd.wins.back();
            for (std::vector<int> S : window_candidates(nd.w, nq, geo, seed, prev)) {
                Node c = nd;
                const std::vector<int> took = c.w.take(S, nq);
                if (took.empty()) continue;
                c.w.run_plain();
                if (geo.free && c.w.pending.empty() && prev) {
                    // the last window stores the logical layout: it must hold qubits 0..C-1
                    S = close_window(S, took, gates, *prev, std::min(geo.c, nq), geo.w);
                    if (S.empty()) continue;
                }
                c.wins.push_back(S);
                next.push_back(std::move(c));
            }
        }
        std::stable_sort(next.begin(), next.end(), [](const Node& x, const Node& y) { return x.w.n_done > y.w.n_done; });
        beam.clear();
        for (Node& c : next) {
            bool dup = false;
            for (const Node& k : beam) dup = dup || k.w.done == c.w.done;
            if (dup) continue;
            beam.push_back(std::move(c));
            if ((int)beam.size() >= width) break;
        }
        if (beam.empty()) fail(QLX_EHIP, "sweep planner: beam search made no progress");
    }
    fail(QLX_EHIP, "sweep planner: beam search did not finish");
}

// Arithmetic a complex128 sweep may carry before it outlasts its HBM pass, in vector-unit work
// units (one unit: a one-qubit rotation or H on every amplitude of the chunk; a CX counts a
// quarter, its cost being the register phases it adds). With windows holding the low qubits
// (12-qubit window, 16 registers) sweeps of up to ~85 units ran at the streaming rate; with
// free windows each sweep also carries more LDS re-layouts, and a budget of 55 measured faster
// on the 30-qubit layered circuit (two kernel traces each: 65.1 and 66.4 ms against 66.6 and
// 67.0 ms with 85; the heaviest sweeps 0 and 3 give gates to sweeps 4 and 5).
// A complex64 pass moves half the bytes in the same instructions: half the budget.
constexpr double kBalanceWork = 55.0;
// The walk's cap on one sweep's arithmetic (same units): a sweep stops taking runnable gates,
// in program order, once it holds this much, leaving later layers to later sweeps. Caps of
// 90-130 units measured slower on the 30-qubit layered circuit (102.4-103.2 ms against
// 102.2 ms: the window planner then moves the heavy work to other sweeps), so no cap is set.
constexpr double kTakeWork = 1e9;

// Evens out the arithmetic of the circuit-specialised sweeps. The window planner gives each
// sweep every gate its window can run, so sweeps on the edge of a circuit (whose light cone is
// open on one side) take far more gates than the rest and outlast their HBM pass, while
// others are light. A gate moves from an over-budget sweep to the earliest later sweep whose
// window holds its qubits and that precedes none of its successors (each successor runs in a
// later step, or in that sweep after it, in program order), as long as the receiving sweep
// stays within the budget or ends up with less work than the giving one. Gates are tried from the end of the sweep backwards, so a gate's
// successors in the same sweep have moved before it is considered. Plain steps in between
// keep their places; the moved gate still runs after all its predecessors.
template <class IsSweep, class Window, class Taken>
static void balance_sweeps(const std::vector<PGate>& gates, const std::vector<std::vector<int>>& preds, size_t nsteps,
                           IsSweep is_sweep, Window window, Taken taken, std::vector<int>& step_of_gate, int nq,
                           double budget) {
    const int G = (int)gates.size();
    std::vector<std::vector<int>> succs(G);
    for (int i = 0; i < G; ++i)
        for (int p : preds[i]) succs[p].push_back(i);
    std::vector<double> work(nsteps, 0.0);
    std::vector<std::vector<char>> inwin(nsteps);
    for (size_t k = 0; k < nsteps; ++k) {
        if (!is_sweep(k)) continue;
        for (int gi : taken(k)) work[k] += gate_work(gates[gi]);
        inwin[k].assign(nq, 0);
        for (int q : window(k)) inwin[k][q] = 1;
    }
    for (size_t k = 0; k < nsteps; ++k) {
        if (!is_sweep(k) || work[k] <= budget) continue;
        std::vector<int>& src = taken(k);
        for (int idx = (int)src.size() - 1; idx >= 0 && work[k] > budget; --idx) {
            const int gi = src[idx];
            // the earliest step any successor runs in
            size_t limit = nsteps;
            bool held = false;
            for (int sc : succs[gi]) {
                const int ss = step_of_gate[sc];
                if (ss == (int)k) held = true;  // a successor stayed in this sweep
                limit = std::min(limit, (size_t)ss);
            }
            if (held) continue;
            for (size_t k2 = k + 1; k2 <= limit && k2 < nsteps; ++k2) {
                if (!is_sweep(k2)) continue;
                bool fits = true;
                for (int j = 0; j < gates[gi].nq && fits; ++j) fits = inwin[k2][gates[gi].q[j]] != 0;
                if (!fits) continue;
                // the receiving sweep stays within the budget, or at least below the giving one
                const double w2 = work[k2] + gate_work(gates[gi]);
                if (w2 > budget && w2 > work[k] - gate_work(gates[gi])) continue;
                // insert in program order (gate indices follow program order)
                std::vector<int>& dst = taken(k2);
                dst.insert(std::lower_bound(dst.begin(), dst.end(), gi), gi);
                work[k2] += gate_work(gates[gi]);
                work[k] -= gate_work(gates[gi]);
                step_of_gate[gi] = (int)k2;
                src.erase(src.begin() + idx);
                break;
            }
        }
    }
}

// The coalesced qubits of a sequence of out-of-place sweeps: ld[k], on address bits 0..C-1
// when sweep k loads, and st[k], when it stores (both ascending). The first load and the last
// store are the logical layout's (qubits 0..C-1); between two sweeps, C qubits both windows
// hold, the previous ones kept while the next window holds them. Fails when the first or last
// window lacks qubits 0..C-1 or two consecutive windows share fewer than C qubits.
static void coalesce_chain(const std::vector<std::vector<int>>& wins, int C, std::vector<std::vector<int>>& ld,
                    std::vector<std::vector<int>>& st) {
    auto holds = [](const std::vector<int>& S, int q) { return std::find(S.begin(), S.end(), q) != S.end(); };
    ld.assign(wins.size(), {});
    st.assign(wins.size(), {});
    std::vector<int> L;
    for (int q = 0; q < C; ++q) L.push_back(q);
    for (size_t k = 0; k < wins.size(); ++k) {
        std::vector<int> out;
        if (k + 1 == wins.size()) {
            for (int q = 0; q < C; ++q) out.push_back(q);
        } else {
            const std::vector<int>& N = wins[k + 1];
            for (int q : L)
                if (holds(N, q)) out.push_back(q);
            std::vector<int> rest;
            for (int q : wins[k])
                if (holds(N, q) && !holds(out, q)) rest.push_back(q);
            std::sort(rest.begin(), rest.end());
            for (size_t i = 0; (int)out.size() < C && i < rest.size(); ++i) out.push_back(rest[i]);
        }
        std::sort(out.begin(), out.end());
        bool ok = (int)out.size() == C;
        for (int q : out) ok = ok && holds(wins[k], q);
        for (int q : L) ok = ok && holds(wins[k], q);
        if (!ok) fail(QLX_EHIP, "sweep planner: consecutive windows share fewer than the coalesced qubits");
        ld[k] = L;
        st[k] = out;
        L = out;
    }
}

enum class Chooser { Greedy, Wide, Given };

// Builds the plan along one walk. Windows come from the greedy chooser, the wide chooser,
// or a given sequence (a beam search's). thorough: a sweep whose gates cost less than one
// pass as plain kernels (e.g. a lone CX: half a pass) runs them as plain kernels, and the
// register phases come from a beam search.
static std::unique_ptr<SweepPlan> build_plan_with(uint32_t nq, const Geometry& geo, const qlx_gate_instr* instrs,
                                                  size_t n, const double* pool, const qlx_matrix_entry* meta,
                                                  const qlx_exec_opts& opts, Chooser chooser,
                                                  const std::vector<std::vector<int>>* given, bool thorough) {
    auto plan = std::make_unique<SweepPlan>();
    if ((int)nq < geo.w) fail(QLX_EINVAL, "window larger than the state");
    plan->geo = geo;
    const int W = geo.w;
    const bool specialised = (opts.specialize & QLX_SPECIALIZE) != 0;

    // gate list (BARRIER and no-op GLOBAL_SWAP dropped)
    std::vector<PGate> gates;
    for (size_t i = 0; i < n; ++i) {
        const qlx_gate_instr& g = instrs[i];
        if (g.type == QLX_BARRIER) continue;
        if (g.type == QLX_GLOBAL_SWAP && opts.gswap_mode == QLX_GSWAP_NOOP) continue;
        PGate p{};
        p.src = (int)i;
        p.type = g.type;
        p.nq = operand_count(g);
        for (int k = 0; k < p.nq; ++k) p.q[k] = instr_qubit(g, k);
        p.sweepable = is_sweepable(g, meta) && (int)nq >= W;
        gates.push_back(p);
    }
    std::vector<MergedRot> merged;
    if (!(opts.plan_flags & QLX_PLAN_KEEP_ROTATIONS)) merge_rotations(gates, instrs, nq, merged);
    for (const MergedRot& m : merged) plan->rotations_merged += m.srcs.size() - 1;
    const int G = (int)gates.size();
    std::vector<std::vector<int>> preds(G);
    {
        std::vector<int> last(nq, -1);
        for (int i = 0; i < G; ++i)
            for (int k = 0; k < gates[i].nq; ++k) {
                const int q = gates[i].q[k];
                if (last[q] >= 0) preds[i].push_back(last[q]);
                last[q] = i;
            }
    }
    // circuit-specialised sweeps: each takes at most about kTakeWork units of arithmetic
    const double take_budget = specialised && thorough ? kTakeWork * geo.elem / 16.0 : 0.0;
    std::vector<std::vector<int>> beam;
    if (chooser == Chooser::Given && !given) {
        // (a width of 64 finds 13 sweeps instead of 14 for the 30-qubit depth-20 layered
        // circuit, but one of them is a lone CX, run as a half-pass kernel, and the others are
        // heavier: 102.4 ms against 101-102 ms, for 3x the planning time)
        beam = beam_windows(gates, preds, (int)nq, geo, 16, take_budget);
        given = &beam;
    }
    auto plain_step = [&](int gi) {
        // a merged rotation runs as the instructions it holds (consecutive: they commute to here)
        std::vector<int> srcs{gates[gi].src};
        if (gates[gi].merged >= 0) srcs = merged[gates[gi].merged].srcs;
        for (int src : srcs) {
            plan->steps.push_back(PlanStep{false, src, -1});
            plan->order.push_back(src);
            plan->order_sweep.push_back(-1);
            plan->passes += plain_passes(gates[gi].type);
        }
    };

    // 1. the walk: plain steps and sweeps (window + the gates it takes, in program order)
    struct RawStep {
        bool sweep;
        int gate;                // plain step: the gate
        std::vector<int> S;      // sweep: window qubits
        std::vector<int> taken;  // sweep: gates, program order
    };
    std::vector<RawStep> raw;
    Walk walk(gates, preds);
    walk.budget = take_budget;
    size_t next_window = 0;
    std::vector<int> prev_S;  // the last sweep's window (free windows)
    while (true) {
        // plain-only gates whose predecessors are done run as their own kernels
        for (int gi : walk.run_plain()) raw.push_back(RawStep{false, gi, {}, {}});
        if (walk.pending.empty()) break;

        // the window
        const int seed_gate = walk.seed_gate();
        if (seed_gate < 0) fail(QLX_EHIP, "sweep planner: no runnable gate");
        std::vector<int> S;
        if (chooser == Chooser::Given) {
            if (next_window >= given->size()) fail(QLX_EHIP, "sweep planner: window sequence too short");
            S = (*given)[next_window++];
        } else if (chooser == Chooser::Wide) {
            if (geo.free && !prev_S.empty()) {
                // free windows: the candidate (sharing C qubits with the previous window) that
                // runs the most pending gates
                int best_n = -1;
                std::vector<char> run(gates.size(), 0);
                for (const std::vector<int>& c : window_candidates(walk, (int)nq, geo, seed_gate, &prev_S)) {
                    std::vector<char> in_set(nq, 0);
                    for (int q : c) in_set[q] = 1;
                    const int cnt = count_runnable(gates, walk.pending, walk.done, in_set, run, preds);
                    if (cnt > best_n) {
                        best_n = cnt;
                        S = c;
                    }
                }
                if (S.empty()) fail(QLX_EHIP, "sweep planner: no window shares the coalesced qubits");
            } else {
                S = choose_window_wide(gates, walk.pending, walk.done, preds, (int)nq, geo, seed_gate);
            }
        } else {
            std::vector<char> in_set(nq, 0);
            int have = 0;
            for (int l = 0; l < geo.c && l < (int)nq; ++l, ++have) in_set[l] = 1;
            for (int k = 0; k < gates[seed_gate].nq; ++k)
                if (!in_set[gates[seed_gate].q[k]]) {
                    in_set[gates[seed_gate].q[k]] = 1;
                    ++have;
                }
            std::vector<char> allowed(nq, 1);
            S = choose_set(gates, walk.pending, walk.done, preds, (int)nq, in_set, have, W, allowed);
        }
        if ((int)S.size() != W) fail(QLX_EHIP, "sweep planner: window smaller than W");

        // the sweep's gates, in program order
        const std::vector<int> taken = walk.take(S, (int)nq);
        if (taken.empty()) fail(QLX_EHIP, "sweep planner: empty sweep");
        if (geo.free && chooser == Chooser::Wide && !prev_S.empty() && walk.pending.empty()) {
            // the last window stores the logical layout: it holds qubits 0..C-1
            S = close_window(S, taken, gates, prev_S, std::min(geo.c, (int)nq), W);
            if (S.empty()) fail(QLX_EHIP, "sweep planner: the last window cannot hold the coalesced qubits");
        }
        raw.push_back(RawStep{true, -1, S, taken});
        prev_S = S;
    }
    // thorough: a sweep whose gates cost less than a pass as plain kernels runs them as plain
    // kernels. With free windows only where the sweeps around it still chain: the first
    // and the last sweep hold qubits 0..C-1, and consecutive sweeps share C qubits.
    if (thorough) {
        const int Cq = std::min(geo.c, (int)nq);
        auto has_low = [&](const std::vector<int>& S) {
            for (int q = 0; q < Cq; ++q)
                if (std::find(S.begin(), S.end(), q) == S.end()) return false;
            return true;
        };
        std::vector<char> plain(raw.size(), 0);
        for (size_t k = 0; k < raw.size(); ++k) {
            if (!raw[k].sweep) continue;
            double cost = 0.0;
            for (int gi : raw[k].taken) cost += plain_passes(gates[gi].type);
            if (cost >= 1.0) continue;
            bool ok = true;
            if (geo.free) {
                int kp = (int)k - 1, kn = (int)k + 1;
                while (kp >= 0 && (!raw[kp].sweep || plain[kp])) --kp;
                while (kn < (int)raw.size() && !raw[kn].sweep) ++kn;
                if (kp < 0)
                    ok = kn == (int)raw.size() || has_low(raw[kn].S);
                else if (kn == (int)raw.size())
                    ok = has_low(raw[kp].S);
                else
                    ok = overlap(raw[kp].S, raw[kn].S) >= Cq;
            }
            plain[k] = ok;
        }
        std::vector<RawStep> out;
        for (size_t k = 0; k < raw.size(); ++k) {
            if (!plain[k]) {
                out.push_back(std::move(raw[k]));
                continue;
            }
            for (int gi : raw[k].taken) out.push_back(RawStep{false, gi, {}, {}});
        }
        raw.swap(out);
    }

    // 2. circuit-specialised sweeps: even out the arithmetic (see balance_sweeps)
    if (specialised && thorough) {
        std::vector<std::vector<int>> sets;
        std::vector<std::vector<int>*> lists;
        std::vector<int> step_of_gate(G, -1);
        for (size_t k = 0; k < raw.size(); ++k) {
            if (raw[k].sweep)
                for (int gi : raw[k].taken) step_of_gate[gi] = (int)k;
            else
                step_of_gate[raw[k].gate] = (int)k;
        }
        balance_sweeps(gates, preds, raw.size(), [&](size_t k) { return raw[k].sweep; },
                       [&](size_t k) -> const std::vector<int>& { return raw[k].S; },
                       [&](size_t k) -> std::vector<int>& { return raw[k].taken; }, step_of_gate, (int)nq,
                       kBalanceWork * geo.elem / 16.0);
    }

    // 3. the coalesced qubits of each sweep's load and store: qubits 0..C-1 at the first load
    // and the last store (the logical layout); between two sweeps, C qubits both windows hold,
    // kept from one sweep to the next while they can be. With windows that all hold 0..C-1 these
    // are 0..C-1 throughout.
    std::vector<std::vector<int>> ld_low(raw.size()), st_low(raw.size());
    {
        std::vector<std::vector<int>> wins, ld, st;
        for (const RawStep& r : raw)
            if (r.sweep) wins.push_back(r.S);
        coalesce_chain(wins, std::min(geo.c, (int)nq), ld, st);
        for (size_t k = 0, i = 0; k < raw.size(); ++k)
            if (raw[k].sweep) {
                ld_low[k] = ld[i];
                st_low[k] = st[i++];
            }
    }

    // 4. the steps: plain kernels, and sweeps with their register phases
    for (size_t k = 0; k < raw.size(); ++k) {
        RawStep& st = raw[k];
        if (!st.sweep) {
            plain_step(st.gate);
            continue;
        }
        // local bits: the load's coalesced qubits first, then the rest ascending
        std::vector<int> S = ld_low[k];
        {
            std::vector<int> rest;
            for (int q : st.S)
                if (std::find(S.begin(), S.end(), q) == S.end()) rest.push_back(q);
            std::sort(rest.begin(), rest.end());
            S.insert(S.end(), rest.begin(), rest.end());
        }
        const std::vector<int>& taken = st.taken;
        std::vector<int> loc(nq, -1);
        for (int j = 0; j < W; ++j) loc[S[j]] = j;
        std::vector<int> st_local;
        for (int q : st_low[k]) st_local.push_back(loc[q]);
        std::sort(st_local.begin(), st_local.end());
        std::vector<PGate> lg;
        for (int gi : taken) {
            PGate l = gates[gi];
            for (int k = 0; k < l.nq; ++k) l.q[k] = loc[l.q[k]];
            l.ctl_free = specialised && l.type == QLX_CX;
            lg.push_back(l);
        }
        PlanSweep sw;
        sw.S = S;
        sw.st_low = st_low[k];
        plan_phases(sw, lg, geo, thorough, st_local);
        const int si = (int)plan->sweeps.size();
        for (const PlanPhase& ph : sw.phases)
            for (const PlanOp& op : ph.ops) {
                if (op.merged >= 0) {
                    for (int src : merged[op.merged].srcs) {
                        plan->order.push_back(src);
                        plan->order_sweep.push_back(si);
                    }
                    plan->gates_in_sweeps += merged[op.merged].srcs.size() - 1;
                    continue;
                }
                plan->order.push_back(op.src);
                plan->order_sweep.push_back(si);
            }
        plan->gates_in_sweeps += lg.size();
        plan->phases += sw.phases.size();
        plan->passes += 1.0;
        plan->sweeps.push_back(std::move(sw));
        plan->steps.push_back(PlanStep{true, -1, si});
    }

    // device image (factored H / RX ops for the circuit-specialised sweeps only; see OpKind)
    const bool factored = (opts.specialize & QLX_SPECIALIZE) != 0;
    for (const PlanSweep& sw : plan->sweeps) {
        DevSweep d{};
        d.n_phases = (uint32_t)sw.phases.size();
        d.phase_begin = (uint32_t)plan->dph.size();
        for (int j = 0; j < W; ++j) d.S[j] = (uint32_t)sw.S[j];
        auto glob_of_local = [&](int l) { return sw.S[l]; };
        auto fill_ends = [&](const PlanPhase& ph, uint32_t* tglob, uint64_t* roff) {
            for (int b = 0; b < geo.logt; ++b) tglob[b] = (uint32_t)glob_of_local(ph.st2loc[b]);
            for (int i = 0; i < (1 << geo.rb); ++i) {
                uint64_t o = 0;
                for (int b = 0; b < geo.rb; ++b)
                    if ((i >> b) & 1) o |= 1ull << glob_of_local(ph.st2loc[geo.logt + b]);
                roff[i] = o;
            }
        };
        fill_ends(sw.phases.front(), d.ld_tglob, d.ld_roff);
        fill_ends(sw.phases.back(), d.st_tglob, d.st_roff);
        plan->dsw.push_back(d);
        // pending real factor of the factored H / RX ops (see OpKind); applied by OP_SCALE
        double scale = 1.0;
        auto scale_op = [&] {
            DevOp o{};
            o.key = (uint32_t)OP_SCALE << 8;
            o.c0 = scale;
            plan->dop.push_back(o);
            scale = 1.0;
        };
        for (size_t p = 0; p < sw.phases.size(); ++p) {
            const PlanPhase& ph = sw.phases[p];
            DevPhase dp{};
            dp.remap = p > 0 ? 1u : 0u;
            dp.op_begin = (uint32_t)plan->dop.size();
            for (int b = 0; b < geo.logt; ++b) dp.tlocal[b] = (uint32_t)ph.st2loc[b];
            for (int i = 0; i < (1 << geo.rb); ++i) {
                uint32_t o = 0;
                for (int b = 0; b < geo.rb; ++b)
                    if ((i >> b) & 1) o |= 1u << ph.st2loc[geo.logt + b];
                dp.roff[i] = o;
                dp.rslot[i] = swz_bytes(o, geo.elem);
            }
            plan->dph.push_back(dp);
            for (const PlanOp& op : ph.ops) {
                DevOp o{};
                o.coef = (uint32_t)plan->coef.size();
                uint8_t kind = 0;
                fill_coef(plan->coef, instrs[op.src], pool, meta, kind,
                          op.merged >= 0 ? &merged[op.merged].theta : nullptr);
                if (op.tctl >= 0) {
                    // CX with its control on work-item bit tctl (circuit-specialised sweeps only)
                    if (kind != OP_CX || !factored || op.tctl >= geo.logt || op.l1 < 0 || op.l1 >= geo.rb)
                        fail(QLX_EHIP, "sweep planner: bad work-item-controlled CX");
                    o.key = (uint32_t)((OP_CXT << 8) | (op.tctl << 4) | op.l1);
                    plan->dop.push_back(o);
                    continue;
                }
                const bool two = kind == OP_CX || kind == OP_SWAP || kind == OP_U4;
                if (two && (op.l1 < 0 || op.l0 == op.l1)) fail(QLX_EHIP, "sweep planner: bad two-qubit operands");
                if (op.l0 < 0 || op.l0 >= geo.rb || (two && op.l1 >= geo.rb))
                    fail(QLX_EHIP, "sweep planner: operand is not a register bit");
                double factor = 1.0;
                if (kind == OP_RX && !factored) {
                    o.c0 = plan->coef[o.coef];
                    o.c1 = plan->coef[o.coef + 1];
                } else if (kind == OP_H && !factored) {
                    o.c0 = plan->coef[o.coef];
                } else if (kind == OP_RX) {
                    const double c = plan->coef[o.coef], sn = plan->coef[o.coef + 1];
                    if (std::fabs(c) >= std::fabs(sn)) {
                        kind = OP_RXT;
                        o.c0 = sn / c;
                        factor = c;
                    } else {
                        kind = OP_RXU;
                        o.c0 = c / sn;
                        factor = sn;
                    }
                } else if (kind == OP_H) {
                    kind = OP_HB;
                    factor = plan->coef[o.coef];
                }
                o.key = (uint32_t)((kind << 8) | (op.l0 << 4) | (two ? op.l1 : 0));
                plan->dop.push_back(o);
                scale *= factor;
                if (std::fabs(scale) < 0x1p-30) scale_op();
            }
            // the last phase applies what is left of the factor before the store
            if (p + 1 == sw.phases.size() && scale != 1.0) scale_op();
            plan->dph.back().n_ops = (uint32_t)(plan->dop.size() - dp.op_begin);
        }
    }
    return plan;
}


std::vector<DenseWindow> plan_dense_windows(uint32_t nq, const qlx_gate_instr* instrs, size_t n,
                                            const qlx_exec_opts& opts, int w, int c) {
    std::vector<PGate> gates;
    for (size_t i = 0; i < n; ++i) {
        const qlx_gate_instr& g = instrs[i];
        if (g.type == QLX_BARRIER) continue;
        if (g.type == QLX_GLOBAL_SWAP && opts.gswap_mode == QLX_GSWAP_NOOP) continue;
        PGate p{};
        p.src = (int)i;
        p.type = g.type;
        p.nq = operand_count(g);
        for (int k = 0; k < p.nq; ++k) p.q[k] = instr_qubit(g, k);
        p.sweepable = true;
        gates.push_back(p);
    }
    const int G = (int)gates.size();
    std::vector<std::vector<int>> preds(G);
    {
        std::vector<int> last(nq, -1);
        for (int i = 0; i < G; ++i)
            for (int k = 0; k < gates[i].nq; ++k) {
                const int q = gates[i].q[k];
                if (last[q] >= 0) preds[i].push_back(last[q]);
                last[q] = i;
            }
    }
    Geometry geo{};
    geo.w = w;
    geo.c = c;
    std::vector<DenseWindow> out;
    if (gates.empty()) return out;
    // out-of-place sweeps may take free windows (Geometry::free), kept when fewer
    std::vector<std::vector<int>> wins = beam_windows(gates, preds, (int)nq, geo, 8);
    if ((opts.specialize & QLX_SPECIALIZE_RELABEL) && !(opts.plan_flags & QLX_PLAN_FIXED_WINDOWS)) {
        Geometry fg = geo;
        fg.free = true;
        try {
            std::vector<std::vector<int>> fw = beam_windows(gates, preds, (int)nq, fg, 8);
            std::vector<std::vector<int>> ld, st;
            coalesce_chain(fw, std::min(c, (int)nq), ld, st);
            if (fw.size() < wins.size()) wins = std::move(fw);
        } catch (const Error&) {
        }
    }
    std::vector<std::vector<int>> ld, st;
    coalesce_chain(wins, std::min(c, (int)nq), ld, st);
    Walk walk(gates, preds);
    for (size_t k = 0; k < wins.size(); ++k) {
        DenseWindow d;
        // local bits: the load's coalesced qubits first, then the rest ascending
        d.S = ld[k];
        std::vector<int> rest;
        for (int q : wins[k])
            if (std::find(d.S.begin(), d.S.end(), q) == d.S.end()) rest.push_back(q);
        std::sort(rest.begin(), rest.end());
        d.S.insert(d.S.end(), rest.begin(), rest.end());
        d.st_low = st[k];
        for (int gi : walk.take(wins[k], (int)nq)) d.instrs.push_back(gates[gi].src);
        if (d.instrs.empty()) fail(QLX_EHIP, "dense sweep planner: empty sweep");
        out.push_back(std::move(d));
    }
    if (!walk.pending.empty()) fail(QLX_EHIP, "dense sweep planner left gates behind");
    return out;
}

// Plans with the wider window search. `thorough` also plans with the classic greedy
// chooser alone and with a beam search over window sequences (width 16), runs sweeps that
// would cost less than one pass as plain kernels, and keeps the plan with the fewest
// passes (then fewest phases). Over 3,198
// layered, HEA, QFT and clustered circuits of 13-28 qubits (complex64, 13-qubit window)
// the wider search alone needs fewer sweeps than the greedy chooser for 1,382 of them
// (18,732 sweeps in all instead of 20,340) and one more sweep for 18.
std::unique_ptr<SweepPlan> build_plan_geo(uint32_t nq, const Geometry& geo, const qlx_gate_instr* instrs, size_t n,
                                          const double* pool, const qlx_matrix_entry* meta,
                                          const qlx_exec_opts& opts, bool thorough) {
    Geometry fixed = geo;
    fixed.free = false;
    if (!thorough) {
        // planned once (served and batched circuits): free windows when the walk can close on
        // qubits 0..C-1, which it nearly always can, else windows that hold them
        if (geo.free) {
            try {
                return build_plan_with(nq, geo, instrs, n, pool, meta, opts, Chooser::Wide, nullptr, false);
            } catch (const Error&) {
            }
        }
        return build_plan_with(nq, fixed, instrs, n, pool, meta, opts, Chooser::Wide, nullptr, false);
    }
    auto best = build_plan_with(nq, fixed, instrs, n, pool, meta, opts, Chooser::Greedy, nullptr, true);
    if (best->sweeps.size() <= 1) return best;
    auto consider = [&](std::unique_ptr<SweepPlan> p) {
        if (p->passes < best->passes - 1e-9 || (p->passes < best->passes + 1e-9 && p->phases < best->phases))
            best = std::move(p);
    };
    for (Chooser c : {Chooser::Wide, Chooser::Given})
        consider(build_plan_with(nq, fixed, instrs, n, pool, meta, opts, c, nullptr, true));
    if (geo.free) {
        // free windows (out-of-place sweeps): kept when they need fewer passes; a window
        // sequence whose consecutive windows share too few qubits is not a plan
        try {
            consider(build_plan_with(nq, geo, instrs, n, pool, meta, opts, Chooser::Given, nullptr, true));
        } catch (const Error&) {
        }
    }
    return best;
}

namespace {

// One op of a generic phase applied to 32 register amplitudes in FP64 (the same action as the
// kernel's g_* functions).
void apply_op_host(std::complex<double> (&v)[32], const DevOp& op, const std::vector<double>& coef) {
    using C = std::complex<double>;
    const int kind = (int)(op.key >> 8), b0 = (int)((op.key >> 4) & 15), b1 = (int)(op.key & 15);
    const int m0 = 1 << b0, m1 = 1 << b1;
    switch (kind) {
        case OP_RX: {
            const C c(op.c0, 0.0), ms(0.0, -op.c1);   // [[c, -i s], [-i s, c]]
            for (int i = 0; i < 32; ++i)
                if (!(i & m0)) {
                    const C a0 = v[i], a1 = v[i | m0];
                    v[i] = c * a0 + ms * a1;
                    v[i | m0] = ms * a0 + c * a1;
                }
            break;
        }
        case OP_H:
            for (int i = 0; i < 32; ++i)
                if (!(i & m0)) {
                    const C a0 = v[i], a1 = v[i | m0];
                    v[i] = op.c0 * (a0 + a1);
                    v[i | m0] = op.c0 * (a0 - a1);
                }
            break;
        case OP_U2: {
            const double* m = coef.data() + op.coef;
            const C u00(m[0], m[1]), u01(m[2], m[3]), u10(m[4], m[5]), u11(m[6], m[7]);
            for (int i = 0; i < 32; ++i)
                if (!(i & m0)) {
                    const C a0 = v[i], a1 = v[i | m0];
                    v[i] = u00 * a0 + u01 * a1;
                    v[i | m0] = u10 * a0 + u11 * a1;
                }
            break;
        }
        case OP_CX:
            for (int i = 0; i < 32; ++i)
                if ((i & m0) && !(i & m1)) std::swap(v[i], v[i | m1]);
            break;
        case OP_SWAP:
            for (int i = 0; i < 32; ++i)
                if ((i & m0) && !(i & m1)) std::swap(v[i], v[i ^ m0 ^ m1]);
            break;
        case OP_U4: {
            const double* m = coef.data() + op.coef;
            for (int i = 0; i < 32; ++i) {
                if ((i & m0) || (i & m1)) continue;
                const int idx[4] = {i, i | m0, i | m1, i | m0 | m1};
                C x[4];
                for (int c = 0; c < 4; ++c) x[c] = v[idx[c]];
                for (int r = 0; r < 4; ++r) {
                    C acc = 0.0;
                    for (int c = 0; c < 4; ++c) acc += C(m[2 * (4 * r + c)], m[2 * (4 * r + c) + 1]) * x[c];
                    v[idx[r]] = acc;
                }
            }
            break;
        }
        default:
            fail(QLX_EHIP, "phase fusion: op kind without a host form");
    }
}

}  // namespace

void densify_phases(SweepPlan& p) {
    if (p.geo.rb != 5 || p.geo.logt != 8) return;
    const size_t fdoubles = mfma_fragment_bytes(5, QLX_DENSE_MFMA_F16) / sizeof(double);
    for (DevPhase& ph : p.dph) {
        double cost = 0.0;
        bool fusable = true;
        for (uint32_t o = 0; o < ph.n_ops && fusable; ++o) {
            switch (p.dop[ph.op_begin + o].key >> 8) {
                case OP_RX: case OP_H: case OP_U2: cost += 1.0; break;
                case OP_CX: case OP_SWAP: cost += 0.25; break;
                case OP_U4: cost += 4.0; break;
                default: fusable = false;
            }
        }
        if (!fusable || cost < kDensePhaseCost) continue;
        std::vector<double> U(2 * 32 * 32);
        for (int c = 0; c < 32; ++c) {
            std::complex<double> v[32] = {};
            v[c] = 1.0;
            for (uint32_t o = 0; o < ph.n_ops; ++o) apply_op_host(v, p.dop[ph.op_begin + o], p.coef);
            for (int r = 0; r < 32; ++r) {
                U[2 * (32 * r + c)] = v[r].real();
                U[2 * (32 * r + c) + 1] = v[r].imag();
            }
        }
        if (p.coef.size() & 1) p.coef.push_back(0.0);  // 16-byte fragments
        const size_t at = p.coef.size();
        if (at + 1 > 0xffffffffull) fail(QLX_EUNSUPPORTED, "phase fusion: coefficient array too large");
        p.coef.resize(at + fdoubles);
        mfma_f16_fragments_host(U.data(), 5, reinterpret_cast<unsigned char*>(p.coef.data() + at));
        ph.mfma = (uint32_t)(at + 1);
    }
}

void assign_layouts(SweepPlan& p, uint32_t nq) {
    const Geometry& g = p.geo;
    const int n = (int)nq;
    std::vector<int> pos(n);
    for (int q = 0; q < n; ++q) pos[q] = q;
    p.relabel = true;
    p.step_pos.clear();
    for (size_t s = 0; s < p.steps.size(); ++s) {
        p.step_pos.push_back(pos);
        if (!p.steps[s].sweep) continue;
        const int si = p.steps[s].sweep_index;
        PlanSweep& sw = p.sweeps[si];
        std::vector<int> next(n);
        if (si + 1 == (int)p.sweeps.size()) {
            for (int q = 0; q < n; ++q) next[q] = q;
        } else {
            // the store's coalesced qubits on address bits 0..C-1; then the next window, this
            // window, the rest in current order
            std::vector<char> placed(n, 0);
            int at = 0;
            auto place = [&](int q) {
                if (placed[q]) return;
                placed[q] = 1;
                next[q] = at++;
            };
            for (int q : sw.st_low) place(q);
            for (int q : p.sweeps[si + 1].S) place(q);
            for (int q : sw.S) place(q);
            std::vector<int> rest;
            for (int q = 0; q < n; ++q)
                if (!placed[q]) rest.push_back(q);
            std::sort(rest.begin(), rest.end(), [&](int a, int b) { return pos[a] < pos[b]; });
            for (int q : rest) place(q);
        }
        // loads and stores coalesced: the load's coalesced qubits (the window's first local
        // bits) and the store's sit on address bits 0..C-1
        const int C = std::min(g.c, n);
        if ((int)sw.st_low.size() != C) fail(QLX_EHIP, "qubit layout: no coalesced store qubits");
        for (int j = 0; j < C; ++j)
            if (pos[sw.S[j]] >= C || next[sw.st_low[j]] >= C) fail(QLX_EHIP, "qubit layout: uncoalesced load or store");
        // chunk-index bits: the qubits outside the window, in the order of their load position
        std::vector<char> inw(n, 0);
        for (int q : sw.S) inw[q] = 1;
        std::vector<int> outside;
        for (int q = 0; q < n; ++q)
            if (!inw[q]) outside.push_back(q);
        std::sort(outside.begin(), outside.end(), [&](int a, int b) { return pos[a] < pos[b]; });
        sw.ld_cpos.clear();
        sw.st_cpos.clear();
        for (int q : outside) {
            sw.ld_cpos.push_back(pos[q]);
            sw.st_cpos.push_back(next[q]);
        }
        if (outside.size() > (size_t)kMaxOopChunkBits) fail(QLX_EUNSUPPORTED, "state too large for out-of-place sweeps");
        {
            DevSweep& d = p.dsw[si];
            d.nc = (uint32_t)outside.size();
            for (size_t j = 0; j < outside.size(); ++j) {
                d.ld_cpos[j] = (uint32_t)sw.ld_cpos[j];
                d.st_cpos[j] = (uint32_t)sw.st_cpos[j];
            }
        }
        sw.ld_lpos.clear();
        sw.st_lpos.clear();
        for (int q : sw.S) {
            sw.ld_lpos.push_back(pos[q]);
            sw.st_lpos.push_back(next[q]);
        }
        // device descriptors: address bits of the first layout's work-item and register bits
        // through the load layout, of the last layout's through the store layout
        DevSweep& d = p.dsw[si];
        auto fill = [&](const PlanPhase& ph, const std::vector<int>& lpos, uint32_t* tglob, uint64_t* roff) {
            for (int b = 0; b < g.logt; ++b) tglob[b] = (uint32_t)lpos[ph.st2loc[b]];
            for (int i = 0; i < (1 << g.rb); ++i) {
                uint64_t o = 0;
                for (int b = 0; b < g.rb; ++b)
                    if ((i >> b) & 1) o |= 1ull << lpos[ph.st2loc[g.logt + b]];
                roff[i] = o;
            }
        };
        fill(sw.phases.front(), sw.ld_lpos, d.ld_tglob, d.ld_roff);
        fill(sw.phases.back(), sw.st_lpos, d.st_tglob, d.st_roff);
        pos = next;
    }
    for (int q = 0; q < n; ++q)
        if (pos[q] != q) fail(QLX_EHIP, "qubit layout does not end in logical order");
}

static std::unique_ptr<SweepPlan> build_plan(uint32_t nq, int dtype, const qlx_gate_instr* instrs, size_t n,
                                             const double* pool, const qlx_matrix_entry* meta,
                                             const qlx_exec_opts& opts) {
    return build_plan_geo(nq, geometry_for(dtype, nq, opts), instrs, n, pool, meta, opts, true);
}

// ---------------------------------------------------------------------------------
// host side: execution
// ---------------------------------------------------------------------------------
static uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Identifies a program on a state (the plan cache key; also names its generated kernels).
static uint64_t program_key(uint32_t nq, int dtype, const qlx_gate_instr* instrs, size_t n, const double* pool,
                            size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts& opts) {
    uint64_t key = 1469598103934665603ull;
    key = fnv1a(key, instrs, n * sizeof(qlx_gate_instr));
    key = fnv1a(key, pool, pool ? pool_elems * 16 : 0);
    key = fnv1a(key, meta, meta ? n_meta * sizeof(qlx_matrix_entry) : 0);
    key = fnv1a(key, &opts, sizeof(opts));
    key = fnv1a(key, &nq, sizeof(nq));
    key = fnv1a(key, &dtype, sizeof(dtype));
    return key;
}

// One generic sweep; out of place (oop) it reads st->amp, writes st->alt and swaps the two.
template <class A, int LOGT, int RB>
static void launch_sweep(qlx_state* st, const char* dev, size_t off_sw, size_t off_ph, size_t off_op, size_t off_cf,
                         int si, uint32_t wg_override, bool oop) {
    const uint64_t chunks = 1ull << (st->nq - (LOGT + RB));
    const size_t lds = (size_t(1) << (LOGT + RB)) * sizeof(A);
    static thread_local int resident[2][2][8][8] = {};  // work-groups per CU, by (dtype, oop, LOGT-5, RB-1)
    int& per_cu = resident[sizeof(A) == 16][oop][LOGT - 5][RB - 1];
    if (!per_cu) {
        if (oop)
            QLX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep<A, LOGT, RB, true>, 1 << LOGT, lds));
        else
            QLX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep<A, LOGT, RB, false>, 1 << LOGT, lds));
        per_cu = std::max(per_cu, 1);
    }
    int dev_id = 0, cus = 0;
    QLX_HIP_CHECK(hipGetDevice(&dev_id));
    QLX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id));
    uint64_t grid = RB <= 4 ? std::min<uint64_t>(chunks, (uint64_t)per_cu * (uint64_t)cus) : chunks;
    if (RB <= 4 && wg_override) grid = std::min<uint64_t>(chunks, wg_override);
    const DevSweep* sw = (const DevSweep*)(dev + off_sw) + si;
    const DevPhase* ph = (const DevPhase*)(dev + off_ph);
    const DevOp* ops = (const DevOp*)(dev + off_op);
    const double* cf = (const double*)(dev + off_cf);
    if (oop) {
        if (!st->alt || st->alt == st->amp) fail(QLX_EHIP, "out-of-place sweep without a second buffer");
        k_sweep<A, LOGT, RB, true><<<(unsigned)grid, 1 << LOGT, lds, st->stream>>>(
            (const A*)st->amp, (A*)st->alt, sw, ph, ops, cf, chunks);
        QLX_HIP_CHECK(hipGetLastError());
        std::swap(st->amp, st->alt);
    } else {
        k_sweep<A, LOGT, RB, false><<<(unsigned)grid, 1 << LOGT, lds, st->stream>>>(
            (const A*)st->amp, (A*)st->amp, sw, ph, ops, cf, chunks);
        QLX_HIP_CHECK(hipGetLastError());
    }
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct CachedPlan {
    uint64_t key = 0;
    std::unique_ptr<SweepPlan> plan;
    DevBuf dev;
    size_t off_sw = 0, off_ph = 0, off_op = 0, off_cf = 0;
    std::shared_ptr<JitSweeps> jit;  // specialised kernels of `plan` (built on first use)
};

static CachedPlan& plan_cache(qlx_state* st) {
    if (!st->sweep_cache) {
        auto* c = new CachedPlan();
        st->sweep_cache = std::shared_ptr<void>(c, [](void* p) {
            auto* cp = static_cast<CachedPlan*>(p);
            cp->dev.release();
            delete cp;
        });
    }
    return *static_cast<CachedPlan*>(st->sweep_cache.get());
}

// Host plan of a windowed program: sweeps, phases, descriptors; complex64 generic plans fuse
// heavy register phases for the matrix cores; out-of-place plans get their qubit layouts.
static std::unique_ptr<SweepPlan> make_plan(uint32_t nq, int dtype, const qlx_gate_instr* instrs, size_t n,
                                            const double* pool, const qlx_matrix_entry* meta,
                                            const qlx_exec_opts& opts, uint64_t key) {
    std::unique_ptr<SweepPlan> p = build_plan(nq, dtype, instrs, n, pool, meta, opts);
    p->tag = key;
    if (dtype == QLX_C64 && !(opts.specialize & QLX_SPECIALIZE)) densify_phases(*p);
    if (opts.specialize & QLX_SPECIALIZE_RELABEL) assign_layouts(*p, nq);
    return p;
}

// The second buffer of out-of-place sweeps; false (and no error) when it does not fit.
static bool ensure_alt(qlx_state* st) {
    if (st->alt) return true;
    if (hipMalloc(&st->alt, st->bytes) != hipSuccess) {
        (void)hipGetLastError();
        st->alt = nullptr;
        return false;
    }
    return true;
}

// ---- background specialisation (QLX_SPECIALIZE_ASYNC) ----------------------------------
// One entry per program (key: stream, pool, options, size, precision), shared by every state
// of the process. A single host thread plans each queued program and compiles its sweeps;
// executions pick the compiled sweeps up once `status` is ready and load them on their device.
struct AsyncSpec {
    uint64_t key = 0;
    std::atomic<int> status{0};  // 0 queued or running, 1 ready, 2 failed
    bool reported = false;       // failure returned by qlx_specialize_wait already
    std::string error;
    double compile_s = 0.0;
    // inputs (copies: the caller's arrays are borrowed only for its call)
    uint32_t nq = 0;
    int dtype = QLX_C128;
    qlx_exec_opts opts{};
    std::vector<qlx_gate_instr> instrs;
    std::vector<double> pool;
    std::vector<qlx_matrix_entry> meta;
    // results (written by the worker before status = 1)
    std::unique_ptr<SweepPlan> plan;
    std::vector<std::vector<char>> code;
    // loaded modules per device (executing threads, under mu)
    std::mutex mu;
    std::vector<std::pair<int, std::shared_ptr<JitSweeps>>> loaded;
};

class SpecService {
public:
    static SpecService& get() {
        static SpecService s;
        return s;
    }
    ~SpecService() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            cancel_.store(true);
        }
        cv_.notify_all();
        if (worker_.joinable()) worker_.join();
    }

    // The program's entry, queued for planning and compilation on first sight.
    std::shared_ptr<AsyncSpec> find_or_queue(uint64_t key, uint32_t nq, int dtype, const qlx_gate_instr* instrs,
                                             size_t n, const double* pool, size_t pool_elems,
                                             const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts& opts) {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = entries_.begin(); it != entries_.end(); ++it)
            if ((*it)->key == key) {
                std::shared_ptr<AsyncSpec> e = *it;
                entries_.erase(it);
                entries_.push_back(e);  // most recently used last
                return e;
            }
        auto e = std::make_shared<AsyncSpec>();
        e->key = key;
        e->nq = nq;
        e->dtype = dtype;
        e->opts = opts;
        e->instrs.assign(instrs, instrs + n);
        if (pool && pool_elems) e->pool.assign(pool, pool + 2 * pool_elems);
        if (meta && n_meta) e->meta.assign(meta, meta + n_meta);
        entries_.push_back(e);
        // keep the most recent programs; entries still queued or compiling stay
        while (entries_.size() > kMaxEntries) {
            auto victim = std::find_if(entries_.begin(), entries_.end(),
                                       [](const std::shared_ptr<AsyncSpec>& x) { return x->status.load() != 0; });
            if (victim == entries_.end()) break;
            entries_.erase(victim);
        }
        queue_.push_back(e);
        if (!worker_.joinable()) worker_ = std::thread([this] { run(); });
        cv_.notify_all();
        return e;
    }

    // Waits for the queue to drain (timeout_s < 0: no limit); returns the programs still pending
    // and the first unreported failure.
    uint32_t wait(double timeout_s, std::string* failure) {
        std::unique_lock<std::mutex> lk(mu_);
        auto idle = [&] { return queue_.empty() && !busy_; };
        if (timeout_s < 0) cv_.wait(lk, idle);
        else cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), idle);
        for (auto& e : entries_)
            if (e->status.load() == 2 && !e->reported) {
                e->reported = true;
                if (failure && failure->empty()) *failure = e->error;
            }
        return (uint32_t)queue_.size() + (busy_ ? 1u : 0u);
    }

private:
    static constexpr size_t kMaxEntries = 32;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::shared_ptr<AsyncSpec>> queue_;
    std::vector<std::shared_ptr<AsyncSpec>> entries_;
    std::thread worker_;
    bool stop_ = false;
    bool busy_ = false;
    std::atomic<bool> cancel_{false};

    void run() {
        for (;;) {
            std::shared_ptr<AsyncSpec> e;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
                if (stop_) return;
                e = queue_.front();
                queue_.pop_front();
                busy_ = true;
            }
            try {
                const auto t0 = std::chrono::steady_clock::now();
                std::unique_ptr<SweepPlan> p =
                    make_plan(e->nq, e->dtype, e->instrs.data(), e->instrs.size(),
                              e->pool.empty() ? nullptr : e->pool.data(), e->meta.empty() ? nullptr : e->meta.data(),
                              e->opts, e->key);
                std::vector<std::vector<char>> code = compile_jit_sweeps(*p, e->dtype, &cancel_);
                e->compile_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                e->plan = std::move(p);
                e->code = std::move(code);
                e->status.store(1, std::memory_order_release);
            } catch (const Error& ex) {
                e->error = ex.msg;
                e->status.store(2, std::memory_order_release);
            } catch (const std::exception& ex) {
                e->error = ex.what();
                e->status.store(2, std::memory_order_release);
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                busy_ = false;
            }
            cv_.notify_all();
        }
    }
};

// Compiled sweeps of a ready entry, loaded on the calling thread's current device.
static std::shared_ptr<JitSweeps> async_kernels(AsyncSpec& e, int device) {
    std::lock_guard<std::mutex> lk(e.mu);
    for (auto& d : e.loaded)
        if (d.first == device) return d.second;
    std::shared_ptr<JitSweeps> j = load_jit_sweeps(*e.plan, e.dtype, e.code, e.compile_s);
    e.loaded.emplace_back(device, j);
    return j;
}

// Runs the steps of a plan: window sweeps through the compiled kernels (`jit`) or the generic
// kernel (descriptor image at `dev`), plain kernels between them.
static void run_plan(qlx_state* st, const SweepPlan& p, const JitSweeps* jit, const char* dev, const CachedPlan* cp,
                     const qlx_gate_instr* instrs, size_t n, const double* pool, size_t pool_elems,
                     const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts& opts, qlx_exec_stats& stats) {
    bool need_pool = false;
    for (const PlanStep& s : p.steps)
        if (!s.sweep && (instrs[s.src].type == QLX_UNITARY || instrs[s.src].type == QLX_FUSED_BLOCK)) need_pool = true;
    const double* dpool = need_pool ? upload_pool(st, pool, pool_elems) : nullptr;
    prepare_dense(st, instrs, n, dpool, meta, n_meta, opts);
    const double pass = pass_bytes(st);
    const Geometry& geo = p.geo;
    if (p.relabel && !p.sweeps.empty() && !st->alt) fail(QLX_EHIP, "out-of-place plan without a second buffer");
    for (size_t k = 0; k < p.steps.size(); ++k) {
        const PlanStep& s = p.steps[k];
        if (!s.sweep) {
            if (p.relabel) {
                // the gate in the current qubit layout
                qlx_gate_instr g = instrs[s.src];
                const std::vector<int>& pos = p.step_pos[k];
                const int na = operand_count(g);
                for (int i = 0; i < na && i < 4; ++i) g.qubits[i] = (uint8_t)pos[g.qubits[i]];
                if (na == 5) g.pad = (uint16_t)((g.pad & 0xFF00) | (uint16_t)pos[g.pad & 0xFF]);
                launch_plain_instr(st, g, dpool, meta, opts, stats);
            } else {
                launch_plain_instr(st, instrs[s.src], dpool, meta, opts, stats);
            }
            continue;
        }
        if (jit) {
            TimedLaunch tl(st, "k_sweep_spec", pass);
            launch_jit_sweep(*jit, s.sweep_index, st, geo.w, opts.sweep_workgroups);
        } else {
            TimedLaunch tl(st, "k_sweep", pass);
            const int key = (st->dtype == QLX_C128 ? 1000 : 0) + geo.w * 10 + geo.rb;
            const size_t a = cp->off_sw, b = cp->off_ph, c = cp->off_op, d = cp->off_cf;
            const int si = s.sweep_index;
            switch (key) {
                case 1124: launch_sweep<double2, 8, 4>(st, dev, a, b, c, d, si, opts.sweep_workgroups, p.relabel); break;
                case 1125: launch_sweep<double2, 7, 5>(st, dev, a, b, c, d, si, opts.sweep_workgroups, p.relabel); break;
                case 1114: launch_sweep<double2, 7, 4>(st, dev, a, b, c, d, si, opts.sweep_workgroups, p.relabel); break;
                case 1113: launch_sweep<double2, 8, 3>(st, dev, a, b, c, d, si, opts.sweep_workgroups, p.relabel); break;
                case 135: launch_sweep<float2, 8, 5>(st, dev, a, b, c, d, si, opts.sweep_workgroups, p.relabel); break;
                case 125: launch_sweep<float2, 7, 5>(st, dev, a, b, c, d, si, opts.sweep_workgroups, p.relabel); break;
                case 124: launch_sweep<float2, 8, 4>(st, dev, a, b, c, d, si, opts.sweep_workgroups, p.relabel); break;
                default: fail(QLX_EUNSUPPORTED, "sweep geometry not built");
            }
        }
        stats.sweeps_full++;
        stats.algorithmic_bytes += pass;
    }
    stats.gates_applied += p.gates_in_sweeps;
    stats.specialised = jit ? 1u : 0u;
    stats.relabelled = p.relabel && !p.sweeps.empty() ? 1u : 0u;
}

void execute_windowed(qlx_state* st, const qlx_gate_instr* instrs, size_t n, const double* pool, size_t pool_elems,
                      const qlx_matrix_entry* meta, size_t n_meta, const qlx_exec_opts& opts, qlx_exec_stats& stats) {
    qlx_exec_opts o = opts;
    if (o.specialize & QLX_SPECIALIZE_ASYNC) o.specialize |= QLX_SPECIALIZE;
    const Geometry geo = geometry_for(st->dtype, st->nq, o);
    if ((int)st->nq < geo.w) {
        // the state is smaller than one window: every gate runs as its own kernel
        execute_plain(st, instrs, n, pool, pool_elems, meta, n_meta, o, stats);
        return;
    }
    // out-of-place sweeps need the second buffer; without it they run in place
    if ((o.specialize & QLX_SPECIALIZE_RELABEL) && !ensure_alt(st)) o.specialize &= ~QLX_SPECIALIZE_RELABEL;
    if (o.specialize & QLX_SPECIALIZE_ASYNC) {
        // keyed (and its kernels named) as the same program compiled in the foreground, so the
        // two share the compiler's code-object cache
        qlx_exec_opts ko = o;
        ko.specialize &= ~QLX_SPECIALIZE_ASYNC;
        ko.mode = QLX_EXEC_WINDOWED;
        const uint64_t key = program_key(st->nq, st->dtype, instrs, n, pool, pool_elems, meta, n_meta, ko);
        std::shared_ptr<AsyncSpec> e = SpecService::get().find_or_queue(key, st->nq, st->dtype, instrs, n, pool,
                                                                        pool_elems, meta, n_meta, o);
        if (e->status.load(std::memory_order_acquire) == 1) {
            std::shared_ptr<JitSweeps> j = async_kernels(*e, st->device);
            run_plan(st, *e->plan, j.get(), nullptr, nullptr, instrs, n, pool, pool_elems, meta, n_meta, o, stats);
            return;
        }
        // until the compiled sweeps are ready (or if their compilation failed): the generic
        // kernel (out of place when the second buffer is there), on the default geometry when
        // the requested one is not built for it
        o.specialize &= QLX_SPECIALIZE_RELABEL;
        if (!geometry_built(st->dtype, geo.w, geo.rb, false)) {
            o.window_qubits = 0;
            o.register_qubits = 0;
        }
    }
    const uint64_t key = program_key(st->nq, st->dtype, instrs, n, pool, pool_elems, meta, n_meta, o);
    CachedPlan& cp = plan_cache(st);
    if (!cp.plan || cp.key != key) {
        cp.plan = make_plan(st->nq, st->dtype, instrs, n, pool, meta, o, key);
        cp.key = key;
        cp.jit.reset();
        SweepPlan& p = *cp.plan;
        cp.off_sw = 0;
        cp.off_ph = align_up(cp.off_sw + p.dsw.size() * sizeof(DevSweep), 256);
        cp.off_op = align_up(cp.off_ph + p.dph.size() * sizeof(DevPhase), 256);
        cp.off_cf = align_up(cp.off_op + p.dop.size() * sizeof(DevOp), 256);
        const size_t total = align_up(cp.off_cf + p.coef.size() * sizeof(double), 256) + 256;
        std::vector<char> img(total, 0);
        if (!p.dsw.empty()) std::memcpy(img.data() + cp.off_sw, p.dsw.data(), p.dsw.size() * sizeof(DevSweep));
        if (!p.dph.empty()) std::memcpy(img.data() + cp.off_ph, p.dph.data(), p.dph.size() * sizeof(DevPhase));
        if (!p.dop.empty()) std::memcpy(img.data() + cp.off_op, p.dop.data(), p.dop.size() * sizeof(DevOp));
        if (!p.coef.empty()) std::memcpy(img.data() + cp.off_cf, p.coef.data(), p.coef.size() * sizeof(double));
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));  // the old image may still be in use
        cp.dev.reserve(total, st->stream);
        QLX_HIP_CHECK(hipMemcpyAsync(cp.dev.ptr, img.data(), total, hipMemcpyHostToDevice, st->stream));
        QLX_HIP_CHECK(hipStreamSynchronize(st->stream));
    }
    const SweepPlan& p = *cp.plan;
    if ((o.specialize & QLX_SPECIALIZE) && !cp.jit) cp.jit = build_jit_sweeps(p, st->dtype);
    run_plan(st, p, (o.specialize & QLX_SPECIALIZE) ? cp.jit.get() : nullptr, (const char*)cp.dev.ptr, &cp, instrs, n,
             pool, pool_elems, meta, n_meta, o, stats);
}

uint32_t specialize_wait(double timeout_s, std::string* failure) { return SpecService::get().wait(timeout_s, failure); }

}  // namespace qlx

using namespace qlx;

extern "C" int qlx_plan_windowed(uint32_t num_qubits, int dtype, const qlx_gate_instr* instrs, size_t n,
                                 const double* pool, size_t pool_elems, const qlx_matrix_entry* meta, size_t n_meta,
                                 const qlx_exec_opts* opts, int64_t* order, int32_t* sweep_of, size_t cap,
                                 size_t* n_out, qlx_plan_info* info) {
    return guard([&] {
        if (n && !instrs) fail(QLX_EINVAL, "null instruction stream");
        qlx_exec_opts o{};
        if (opts) o = *opts;
        for (size_t i = 0; i < n; ++i) {
            if (instrs[i].type == QLX_MEASURE) fail(QLX_EINVAL, "plans cover MEASURE-free streams");
            validate_instr(instrs[i], num_qubits, meta, n_meta, pool_elems);
        }
        if (o.mode == QLX_EXEC_DENSE_SWEEPS) {
            // dense sweeps: 13-qubit windows holding the low 4 qubits, every gate inside one
            std::vector<DenseWindow> wins;
            if (num_qubits >= 13) wins = plan_dense_windows(num_qubits, instrs, n, o, 13, 4);
            size_t cnt = 0;
            for (const DenseWindow& w : wins) cnt += w.instrs.size();
            if (n_out) *n_out = cnt;
            size_t k = 0;
            for (size_t si = 0; si < wins.size(); ++si)
                for (int src : wins[si].instrs) {
                    if (k < cap) {
                        if (order) order[k] = src;
                        if (sweep_of) sweep_of[k] = (int32_t)si;
                    }
                    ++k;
                }
            if (info) {
                std::memset(info, 0, sizeof(*info));
                info->window_qubits = 13;
                info->coalesce_qubits = 4;
                info->threads_per_group = 256;
                info->sweeps = wins.size();
                info->gates_in_sweeps = cnt;
            }
            return;
        }
        const Geometry geo = geometry_for(dtype, num_qubits, o);
        std::unique_ptr<SweepPlan> p;
        if ((int)num_qubits >= geo.w) p = build_plan(num_qubits, dtype, instrs, n, pool, meta, o);
        if (p && dtype == QLX_C64 && !(o.specialize & QLX_SPECIALIZE)) densify_phases(*p);
        const size_t cnt = p ? p->order.size() : 0;
        if (n_out) *n_out = cnt;
        for (size_t k = 0; p && k < cnt && k < cap; ++k) {
            if (order) order[k] = p->order[k];
            if (sweep_of) sweep_of[k] = p->order_sweep[k];
        }
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->window_qubits = (uint32_t)geo.w;
            info->register_qubits = (uint32_t)geo.rb;
            info->threads_per_group = 1u << geo.logt;
            info->coalesce_qubits = (uint32_t)geo.c;
            if (p) {
                info->sweeps = p->sweeps.size();
                info->phases = p->phases;
                info->plain_steps = p->steps.size() - p->sweeps.size();
                info->gates_in_sweeps = p->gates_in_sweeps;
                for (const DevPhase& ph : p->dph) info->mfma_phases += ph.mfma ? 1 : 0;
                info->rotations_merged = p->rotations_merged;
            }
        }
    });
}

extern "C" int qlx_specialize_wait(double timeout_s, uint32_t* pending) {
    return guard([&] {
        std::string failure;
        const uint32_t left = specialize_wait(timeout_s, &failure);
        if (pending) *pending = left;
        if (!failure.empty()) fail(QLX_EHIP, "background sweep compilation failed: " + failure);
    });
}

extern "C" int qlx_plan_sweep_source(uint32_t num_qubits, int dtype, const qlx_gate_instr* instrs, size_t n,
                                     const double* pool, size_t pool_elems, const qlx_matrix_entry* meta,
                                     size_t n_meta, const qlx_exec_opts* opts, uint32_t sweep_index, char* out,
                                     size_t cap, size_t* len) {
    return guard([&] {
        if (n && !instrs) fail(QLX_EINVAL, "null instruction stream");
        qlx_exec_opts o{};
        if (opts) o = *opts;
        o.specialize |= QLX_SPECIALIZE;  // the source is a specialised sweep's
        for (size_t i = 0; i < n; ++i) {
            if (instrs[i].type == QLX_MEASURE) fail(QLX_EINVAL, "plans cover MEASURE-free streams");
            validate_instr(instrs[i], num_qubits, meta, n_meta, pool_elems);
        }
        const Geometry geo = geometry_for(dtype, num_qubits, o);
        if ((int)num_qubits < geo.w) fail(QLX_EINVAL, "the state is smaller than one window: no sweeps");
        const auto p = build_plan(num_qubits, dtype, instrs, n, pool, meta, o);
        p->tag = program_key(num_qubits, dtype, instrs, n, pool, pool_elems, meta, n_meta, o);
        if (o.specialize & QLX_SPECIALIZE_RELABEL) assign_layouts(*p, num_qubits);
        const std::string src = jit_sweep_source(*p, (int)sweep_index, dtype);
        if (len) *len = src.size();
        if (out && cap) {
            const size_t k = std::min(cap - 1, src.size());
            std::memcpy(out, src.data(), k);
            out[k] = 0;
        }
    });
}
