// qlx_transpile.cpp — Middleware::transpile_gpu (host C++, no Eigen).
//
// The reference leaves transpile_gpu empty (include/middleware/middleware.hpp:866).
// This fills it with the same pass pipeline as transpile_cpu_openmp (:488-524),
// re-targeted to the GPU:
//   DAG::build / DAG::release          middleware.hpp:165-230   (restated)
//   pass_gate_fusion                    middleware.hpp:526-753   (restated: BFS growth, causality
//                                       check, union width, the "skip 1-gate and 2-standard-gate
//                                       blocks" heuristic, root morphs into FUSED_BLOCK)
//       + extensions: RX/RY/RZ/SWAP/CCX/CSWAP matrices (the reference cannot fuse them,
//         :442-448), width up to 5 (the fifth qubit in GateInstr::_pad's low byte)
//   FusionMath::expand/fuse_matrices    middleware.hpp:241-344   (restated, M2 * M1)
//   pass_qubit_windowing                middleware.hpp:759-860   (restated with
//                                       ThresholdCostModel :63-119)
//       + extension: optional trailing GLOBAL_SWAPs that restore logical qubit order
//         (the reference discards its logical->physical map, so its swapped programs
//         compute a permuted state)
//   HardwareTopology::get_safe_qubit_limit  hw_topology.hpp:47-53: for the GPU the "safe
//       zone" is the LDS/register window of one work-group (qlx_gpu_safe_qubit_limit).

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <vector>

#include "qlx_internal.hpp"

namespace qlx {
namespace {

using cd = std::complex<double>;
constexpr uint32_t kNull = 0xFFFFFFFFu;
constexpr int kMaxArity = 5;

struct Node {
    qlx_gate_instr instr;
    uint32_t in[kMaxArity] = {kNull, kNull, kNull, kNull, kNull};
    uint32_t out[kMaxArity] = {kNull, kNull, kNull, kNull, kNull};
    bool fused = false;
};

int node_arity(const qlx_gate_instr& g) { return instr_arity(g); }
int node_qubit(const qlx_gate_instr& g, int i) { return instr_qubit(g, i); }
void set_qubit(qlx_gate_instr& g, int i, int q) {
    if (i < 4) g.qubits[i] = (uint8_t)q;
    else g.pad = (uint16_t)((g.pad & 0xFF00) | (q & 0xFF));
}

struct Pool {
    std::vector<cd> data;
    std::vector<qlx_matrix_entry> meta;
    uint32_t store(const std::vector<cd>& m, int k) {
        const uint32_t idx = (uint32_t)meta.size();
        qlx_matrix_entry e{};
        e.offset = (uint32_t)data.size();
        e.num_elements = (uint32_t)m.size();
        e.num_qubits = (uint8_t)k;
        data.insert(data.end(), m.begin(), m.end());
        meta.push_back(e);
        return idx;
    }
};

// ---- matrices (row-major, matrix bit i <-> qubit operand i) ------------------------
using Mat = std::vector<cd>;

bool fusable(const qlx_gate_instr& g, bool rotations, int max_width) {
    switch (g.type) {
        case QLX_H: case QLX_X: case QLX_Y: case QLX_Z: case QLX_S: case QLX_CX:
            return true;  // the reference's set (middleware.hpp:442-448)
        case QLX_UNITARY: case QLX_FUSED_BLOCK:
            return node_arity(g) <= max_width;
        case QLX_RX: case QLX_RY: case QLX_RZ: case QLX_SWAP:
            return rotations;
        case QLX_CCX: case QLX_CSWAP:
            return rotations && max_width >= 3;
        default:
            return false;
    }
}

Mat gate_matrix(const qlx_gate_instr& g, const Pool& pool) {
    const double r2 = 1.0 / std::sqrt(2.0);
    switch (g.type) {
        case QLX_UNITARY:
        case QLX_FUSED_BLOCK: {
            const qlx_matrix_entry& e = pool.meta[g.payload.matrix_idx];
            return Mat(pool.data.begin() + e.offset, pool.data.begin() + e.offset + e.num_elements);
        }
        case QLX_H: return {r2, r2, r2, -r2};
        case QLX_X: return {0.0, 1.0, 1.0, 0.0};
        case QLX_Y: return {0.0, cd(0, -1), cd(0, 1), 0.0};
        case QLX_Z: return {1.0, 0.0, 0.0, -1.0};
        case QLX_S: return {1.0, 0.0, 0.0, cd(0, 1)};
        case QLX_RX: case QLX_RY: case QLX_RZ: {
            const double h = static_cast<double>(g.payload.param) * 0.5;
            const double c = std::cos(h), s = std::sin(h);
            if (g.type == QLX_RX) return {c, cd(0, -s), cd(0, -s), c};
            if (g.type == QLX_RY) return {c, -s, s, c};
            return {cd(std::cos(-h), std::sin(-h)), 0.0, 0.0, cd(std::cos(h), std::sin(h))};
        }
        default: break;
    }
    // permutation gates: column c maps to row f(c)
    int k = 0;
    switch (g.type) {
        case QLX_CX: case QLX_SWAP: case QLX_GLOBAL_SWAP: k = 2; break;
        case QLX_CCX: case QLX_CSWAP: k = 3; break;
        default: fail(QLX_EUNSUPPORTED, "Middleware::get_matrix unsupported gate type.");
    }
    const int d = 1 << k;
    Mat m(d * d, 0.0);
    for (int c = 0; c < d; ++c) {
        int b[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
        if (g.type == QLX_CX && b[0]) b[1] ^= 1;
        else if (g.type == QLX_SWAP || g.type == QLX_GLOBAL_SWAP) std::swap(b[0], b[1]);
        else if (g.type == QLX_CCX && b[0] && b[1]) b[2] ^= 1;
        else if (g.type == QLX_CSWAP && b[0]) std::swap(b[1], b[2]);
        const int r = b[0] | (b[1] << 1) | (b[2] << 2);
        m[r * d + c] = 1.0;
    }
    return m;
}

// FusionMath::expand_matrix: U acts on q_sub; returns its action on q_full (little-endian).
Mat expand(const Mat& U, const std::vector<int>& q_sub, const std::vector<int>& q_full) {
    const size_t df = size_t{1} << q_full.size(), ds = size_t{1} << q_sub.size();
    std::vector<int> rest;
    for (int q : q_full)
        if (std::find(q_sub.begin(), q_sub.end(), q) == q_sub.end()) rest.push_back(q);
    auto pos = [&](int q) { return (int)(std::find(q_full.begin(), q_full.end(), q) - q_full.begin()); };
    std::vector<size_t> sub(ds), rst(size_t{1} << rest.size());
    for (size_t r = 0; r < ds; ++r) {
        size_t m = 0;
        for (size_t i = 0; i < q_sub.size(); ++i)
            if ((r >> i) & 1) m |= size_t{1} << pos(q_sub[i]);
        sub[r] = m;
    }
    for (size_t r = 0; r < rst.size(); ++r) {
        size_t m = 0;
        for (size_t i = 0; i < rest.size(); ++i)
            if ((r >> i) & 1) m |= size_t{1} << pos(rest[i]);
        rst[r] = m;
    }
    Mat M(df * df, 0.0);
    for (size_t b : rst)
        for (size_t i = 0; i < ds; ++i)
            for (size_t j = 0; j < ds; ++j) M[(b | sub[i]) * df + (b | sub[j])] = U[i * ds + j];
    return M;
}

// FusionMath::fuse_matrices: U2 applied after U1 over the sorted union: M2 * M1.
Mat fuse(const Mat& U1, const std::vector<int>& q1, const Mat& U2, const std::vector<int>& q2,
         std::vector<int>& out_q) {
    out_q = q1;
    for (int q : q2)
        if (std::find(out_q.begin(), out_q.end(), q) == out_q.end()) out_q.push_back(q);
    std::sort(out_q.begin(), out_q.end());
    const Mat A = expand(U1, q1, out_q), B = expand(U2, q2, out_q);
    const size_t d = size_t{1} << out_q.size();
    Mat R(d * d, 0.0);
    for (size_t i = 0; i < d; ++i)
        for (size_t m = 0; m < d; ++m) {
            const cd b = B[i * d + m];
            if (b == cd(0.0, 0.0)) continue;
            for (size_t j = 0; j < d; ++j) R[i * d + j] += b * A[m * d + j];
        }
    return R;
}

// ---- DAG (middleware.hpp:153-231) ---------------------------------------------------------
struct Dag {
    std::vector<Node> nodes;
    std::vector<qlx_gate_instr> rewritten;
    bool use_rewritten = false;
};

Dag build_dag(const qlx_gate_instr* instrs, size_t n, uint32_t nq) {
    Dag dag;
    dag.nodes.reserve(n);
    std::vector<uint32_t> frontier(nq, kNull);
    for (size_t k = 0; k < n; ++k) {
        Node node;
        node.instr = instrs[k];
        const uint32_t self = (uint32_t)dag.nodes.size();
        const int a = node_arity(instrs[k]);
        for (int i = 0; i < a && i < kMaxArity; ++i) {
            const int q = node_qubit(instrs[k], i);
            if (q < 0 || (uint32_t)q >= nq) continue;
            const uint32_t prev = frontier[q];
            node.in[i] = prev;
            if (prev != kNull) {
                Node& p = dag.nodes[prev];
                for (int j = 0; j < node_arity(p.instr) && j < kMaxArity; ++j)
                    if (node_qubit(p.instr, j) == q) {
                        p.out[j] = self;
                        break;
                    }
            }
            frontier[q] = self;
        }
        dag.nodes.push_back(node);
    }
    return dag;
}

std::vector<qlx_gate_instr> release_dag(Dag&& dag) {
    if (dag.use_rewritten) return std::move(dag.rewritten);
    std::vector<qlx_gate_instr> out;
    out.reserve(dag.nodes.size());
    for (const Node& nd : dag.nodes)
        if (!nd.fused) out.push_back(nd.instr);
    return out;
}

bool in_list(const std::vector<uint32_t>& v, uint32_t x) { return std::find(v.begin(), v.end(), x) != v.end(); }

// Blocks whose matrices are formed on the device after the pass (build_on_device).
struct DeviceFusion {
    std::vector<FusedRecipe> recipes;
    std::vector<uint32_t> pool_offset;  // where each block's matrix goes in the pool
    std::vector<double> member_mats;
};

// ---- pass_gate_fusion (middleware.hpp:526-753) ------------------------------------------------
void gate_fusion(Dag& dag, Pool& pool, int max_width, bool rotations, DeviceFusion* dev) {
    auto& N = dag.nodes;
    for (size_t i = 0; i < N.size(); ++i) {
        if (N[i].fused || !fusable(N[i].instr, rotations, max_width)) continue;
        std::vector<uint32_t> block = {(uint32_t)i};
        std::vector<int> bq;
        for (int k = 0; k < node_arity(N[i].instr); ++k) bq.push_back(node_qubit(N[i].instr, k));
        if ((int)bq.size() > max_width) continue;
        bool dense = N[i].instr.type == QLX_UNITARY || N[i].instr.type == QLX_FUSED_BLOCK;
        for (size_t head = 0; head < block.size(); ++head) {
            const uint32_t cand = block[head];
            for (int o = 0; o < node_arity(N[cand].instr); ++o) {
                const uint32_t nx = N[cand].out[o];
                if (nx == kNull || N[nx].fused || !fusable(N[nx].instr, rotations, max_width)) continue;
                if (in_list(block, nx)) continue;
                bool causal = true;
                for (int k = 0; k < node_arity(N[nx].instr) && causal; ++k) {
                    const uint32_t p = N[nx].in[k];
                    if (p != kNull && p > i && !in_list(block, p)) causal = false;
                }
                if (!causal) continue;
                std::vector<int> uq = bq;
                bool too_wide = false;
                for (int k = 0; k < node_arity(N[nx].instr); ++k) {
                    const int q = node_qubit(N[nx].instr, k);
                    if (std::find(uq.begin(), uq.end(), q) == uq.end()) {
                        if ((int)uq.size() >= max_width) {
                            too_wide = true;
                            break;
                        }
                        uq.push_back(q);
                    }
                }
                if (too_wide) continue;
                block.push_back(nx);
                bq = uq;
                if (N[nx].instr.type == QLX_UNITARY || N[nx].instr.type == QLX_FUSED_BLOCK) dense = true;
            }
        }
        // skip 1-gate blocks and 2-gate blocks of standard gates (middleware.hpp:631-634)
        if (block.size() == 1) continue;
        if (block.size() == 2 && !dense) continue;

        auto operands = [&](uint32_t b) {
            std::vector<int> q;
            for (int k = 0; k < node_arity(N[b].instr); ++k) q.push_back(node_qubit(N[b].instr, k));
            return q;
        };
        Mat M;
        std::vector<int> cq;
        if (dev) {
            // the union of the members' qubits, sorted (fuse() builds the same order)
            for (uint32_t b : block)
                for (int q : operands(b))
                    if (std::find(cq.begin(), cq.end(), q) == cq.end()) cq.push_back(q);
            std::sort(cq.begin(), cq.end());
            FusedRecipe rc;
            rc.k = (int)cq.size();
            for (uint32_t b : block) {
                const Mat g = gate_matrix(N[b].instr, pool);
                FusedMember fm{};
                fm.mat_off = (uint32_t)(dev->member_mats.size() / 2);
                const std::vector<int> oq = operands(b);
                fm.a = (int)oq.size();
                for (int j = 0; j < fm.a; ++j) fm.pos[j] = (int)(std::find(cq.begin(), cq.end(), oq[j]) - cq.begin());
                for (const cd& z : g) {
                    dev->member_mats.push_back(z.real());
                    dev->member_mats.push_back(z.imag());
                }
                rc.members.push_back(fm);
            }
            dev->recipes.push_back(std::move(rc));
            M.assign(size_t{1} << (2 * cq.size()), cd(0.0, 0.0));  // filled in after the pass
        } else {
            M = gate_matrix(N[block[0]].instr, pool);
            cq = operands(block[0]);
            for (size_t k = 1; k < block.size(); ++k) {
                std::vector<int> oq;
                M = fuse(M, cq, gate_matrix(N[block[k]].instr, pool), operands(block[k]), oq);
                cq = oq;
            }
        }
        // stitch the merged node's edges (middleware.hpp:665-703)
        const int sz = (int)cq.size();
        uint32_t bin[kMaxArity], bout[kMaxArity];
        for (int u = 0; u < sz; ++u) {
            const int q = cq[u];
            uint32_t first = kNull, last = kNull;
            for (size_t k = 0; k < block.size(); ++k) {
                const Node& nd = N[block[k]];
                bool found = false;
                for (int j = 0; j < node_arity(nd.instr); ++j)
                    if (node_qubit(nd.instr, j) == q) {
                        first = nd.in[j];
                        found = true;
                        break;
                    }
                if (found) break;
            }
            for (size_t k = block.size(); k-- > 0;) {
                const Node& nd = N[block[k]];
                bool found = false;
                for (int j = 0; j < node_arity(nd.instr); ++j)
                    if (node_qubit(nd.instr, j) == q) {
                        last = nd.out[j];
                        found = true;
                        break;
                    }
                if (found) break;
            }
            bin[u] = first;
            bout[u] = last;
        }
        const uint32_t midx = pool.store(M, sz);
        if (dev) dev->pool_offset.push_back(pool.meta[midx].offset);
        Node& A = N[block[0]];
        A.instr.type = QLX_FUSED_BLOCK;
        A.instr.flags_arity = (uint8_t)((A.instr.flags_arity & 0xF0) | (sz & 0x0F));
        A.instr.payload.matrix_idx = midx;
        A.instr.pad = 0;
        for (int u = 0; u < kMaxArity; ++u) {
            if (u < sz) {
                set_qubit(A.instr, u, cq[u]);
                A.in[u] = bin[u];
                A.out[u] = bout[u];
                if (bin[u] != kNull) {
                    Node& P = N[bin[u]];
                    for (int j = 0; j < node_arity(P.instr); ++j)
                        if (in_list(block, P.out[j])) P.out[j] = block[0];
                }
                if (bout[u] != kNull) {
                    Node& S = N[bout[u]];
                    for (int j = 0; j < node_arity(S.instr); ++j)
                        if (in_list(block, S.in[j])) S.in[j] = block[0];
                }
            } else {
                if (u < 4) A.instr.qubits[u] = 255;
                A.in[u] = kNull;
                A.out[u] = kNull;
            }
        }
        for (size_t k = 1; k < block.size(); ++k) N[block[k]].fused = true;
    }
}

// ---- pass_qubit_windowing (middleware.hpp:759-860) + ThresholdCostModel (:63-119) ---------
// (GreedyCostModel, :121-131, is the threshold model with penalties 1 and 0; a custom model
// is the caller's evaluate_swap through a C callback.)
struct Threshold {
    uint32_t k_safe, numa_penalty, swap_penalty;
    qlx_swap_eval_fn custom = nullptr;
    void* ctx = nullptr;
    bool evaluate(const std::vector<uint32_t>& freq, const std::vector<uint16_t>& l2p,
                  const std::vector<uint16_t>& p2l, int& cand, int& victim) const {
        if (custom) {
            uint16_t a = 0, v = 0;
            const int rc = custom(ctx, freq.data(), l2p.data(), p2l.data(), (uint32_t)freq.size(), &a, &v);
            if (rc < 0) fail(QLX_EINVAL, "custom windowing cost model failed");
            if (rc == 0) return false;
            if (a >= freq.size() || v >= freq.size() || a == v)
                fail(QLX_EINVAL, "custom windowing cost model proposed an invalid swap (" + std::to_string(a) + ", " +
                                     std::to_string(v) + ")");
            cand = a;
            victim = v;
            return true;
        }
        return threshold(freq, l2p, cand, victim);
    }
    bool threshold(const std::vector<uint32_t>& freq, const std::vector<uint16_t>& l2p, int& cand, int& victim) const {
        cand = -1;
        uint32_t best = 0;
        for (size_t q = 0; q < freq.size(); ++q) {
            if (freq[q] > 0 && l2p[q] > k_safe) {
                const uint32_t retention = freq[q] * numa_penalty;
                if (retention >= swap_penalty && retention > best) {
                    best = retention;
                    cand = (int)q;
                }
            }
        }
        if (cand < 0) return false;
        victim = -1;
        for (size_t q = 0; q < freq.size(); ++q)
            if (l2p[q] <= k_safe && freq[q] == 0) {
                victim = (int)q;
                break;
            }
        return victim >= 0;
    }
};

void qubit_windowing(Dag& dag, const Threshold& cm, uint32_t nq, bool restore) {
    std::vector<uint16_t> l2p(nq), p2l(nq);
    for (uint32_t i = 0; i < nq; ++i) l2p[i] = p2l[i] = (uint16_t)i;
    auto gswap = [&](int pa, int pb) {
        qlx_gate_instr s{};
        s.type = QLX_GLOBAL_SWAP;
        s.flags_arity = 2;
        s.qubits[0] = (uint8_t)pa;
        s.qubits[1] = (uint8_t)pb;
        s.qubits[2] = s.qubits[3] = 255;
        dag.rewritten.push_back(s);
    };
    std::vector<uint32_t> freq(nq, 0);
    const auto& N = dag.nodes;
    size_t i = 0;
    while (i < N.size()) {
        if (N[i].fused) {
            ++i;
            continue;
        }
        std::fill(freq.begin(), freq.end(), 0u);
        uint32_t unique = 0;
        size_t j = i;
        while (j < N.size()) {
            if (N[j].fused) {
                ++j;
                continue;
            }
            uint32_t add = 0;
            for (int k = 0; k < node_arity(N[j].instr); ++k) {
                const int q = node_qubit(N[j].instr, k);
                if ((uint32_t)q < nq && freq[q] == 0) ++add;
            }
            if (unique + add > cm.k_safe && j > i) break;
            for (int k = 0; k < node_arity(N[j].instr); ++k) {
                const int q = node_qubit(N[j].instr, k);
                if ((uint32_t)q < nq) {
                    if (freq[q] == 0) ++unique;
                    ++freq[q];
                }
            }
            ++j;
        }
        int cand, victim;
        uint32_t swaps = 0;
        while (cm.evaluate(freq, l2p, p2l, cand, victim)) {
            // a model that keeps proposing swaps for one window does not converge (every useful
            // swap moves one more active qubit into the safe zone)
            if (cm.custom && ++swaps > nq)
                fail(QLX_EINVAL, "custom windowing cost model proposed more than num_qubits swaps for one window");
            const uint16_t pc = l2p[cand], pv = l2p[victim];
            gswap(pc, pv);
            l2p[cand] = pv;
            l2p[victim] = pc;
            p2l[pc] = (uint16_t)victim;
            p2l[pv] = (uint16_t)cand;
        }
        for (size_t w = i; w < j; ++w) {
            if (N[w].fused) continue;
            qlx_gate_instr g = N[w].instr;
            for (int k = 0; k < node_arity(g); ++k) {
                const int q = node_qubit(g, k);
                if ((uint32_t)q < nq) set_qubit(g, k, l2p[q]);
            }
            dag.rewritten.push_back(g);
        }
        i = j;
    }
    if (restore) {
        for (uint32_t p = 0; p < nq; ++p) {
            if (p2l[p] == p) continue;
            const uint16_t x = l2p[p];  // where logical p currently lives
            gswap((int)p, (int)x);
            const uint16_t lp = p2l[p];
            l2p[lp] = x;
            p2l[x] = lp;
            l2p[p] = (uint16_t)p;
            p2l[p] = (uint16_t)p;
        }
    }
    dag.use_rewritten = true;
}

}  // namespace
}  // namespace qlx

struct qlx_program {
    std::vector<qlx_gate_instr> instrs;
    qlx::Pool pool;
};

using namespace qlx;

extern "C" {

int qlx_gpu_safe_qubit_limit(int device, int dtype, uint32_t* k_safe) {
    return guard([&] {
        if (!k_safe) fail(QLX_EINVAL, "null output");
        if (dtype != QLX_C128 && dtype != QLX_C64) fail(QLX_EINVAL, "bad dtype");
        // One work-group holds its window in LDS for re-layouts: the window is the
        // largest power of two of amplitudes that fits half of the CU's LDS (two
        // work-groups per CU overlap HBM streaming with compute). With 160 KiB LDS:
        // 64 KiB -> 12 qubits (complex128) / 13 (complex64).
        qlx_device_info info{};
        uint32_t lds = 160u * 1024u;
        if (device >= 0 && qlx_device_query(device, &info) == QLX_OK && info.lds_bytes_per_cu) lds = info.lds_bytes_per_cu;
        const double usable = lds / 2.0;
        const double amp = dtype == QLX_C128 ? 16.0 : 8.0;
        uint32_t w = (uint32_t)std::floor(std::log2(usable / amp));
        const uint32_t built = dtype == QLX_C128 ? 12u : 13u;  // the sweep kernels compiled in
        *k_safe = std::min(w, built);
    });
}

int qlx_transpile(uint32_t nq, const qlx_gate_instr* instrs, size_t n, const double* pool, size_t pool_elems,
                  const qlx_matrix_entry* meta, size_t n_meta, const qlx_transpile_opts* opts, qlx_program** out) {
    return guard([&] {
        if (!out) fail(QLX_EINVAL, "null output");
        *out = nullptr;
        if (n && !instrs) fail(QLX_EINVAL, "null instruction stream");
        qlx_transpile_opts o{};
        if (opts) o = *opts;
        const int width = o.max_fusion_width ? (int)o.max_fusion_width : 4;
        if (width < 1 || width > 5) fail(QLX_EINVAL, "max_fusion_width must be 1..5");
        for (size_t i = 0; i < n; ++i) {
            if (instrs[i].type == QLX_UNITARY || instrs[i].type == QLX_FUSED_BLOCK) {
                if (!meta || instrs[i].payload.matrix_idx >= n_meta) fail(QLX_ERANGE, "matrix_idx out of range");
                const qlx_matrix_entry& e = meta[instrs[i].payload.matrix_idx];
                if ((size_t)e.offset + e.num_elements > pool_elems) fail(QLX_ERANGE, "matrix exceeds the pool");
            }
        }
        auto prog = std::make_unique<qlx_program>();
        prog->pool.data.assign(reinterpret_cast<const cd*>(pool), reinterpret_cast<const cd*>(pool) + pool_elems);
        if (meta) prog->pool.meta.assign(meta, meta + n_meta);
        Dag dag = build_dag(instrs, n, nq);
        if (o.enable_gate_fusion) {
            DeviceFusion dev;
            gate_fusion(dag, prog->pool, width, o.fuse_rotations != 0, o.build_on_device ? &dev : nullptr);
            if (!dev.recipes.empty()) {
                std::vector<double> mats;
                build_fused_on_device(dev.recipes, dev.member_mats, mats);
                size_t at = 0;
                for (size_t b = 0; b < dev.recipes.size(); ++b) {
                    const size_t ne = size_t{1} << (2 * dev.recipes[b].k);
                    for (size_t e = 0; e < ne; ++e)
                        prog->pool.data[dev.pool_offset[b] + e] = cd(mats[2 * (at + e)], mats[2 * (at + e) + 1]);
                    at += ne;
                }
            }
        }
        if (o.enable_windowing) {
            uint32_t ks = o.k_safe;
            if (!ks) {
                if (qlx_gpu_safe_qubit_limit(-1, o.dtype == QLX_C64 ? QLX_C64 : QLX_C128, &ks) != QLX_OK) ks = 12;
            }
            Threshold cm{ks, o.numa_penalty ? o.numa_penalty : 10u, o.swap_penalty};
            if (o.cost_model == QLX_COST_GREEDY) {
                cm.numa_penalty = 1;
                cm.swap_penalty = 0;
            } else if (o.cost_model == QLX_COST_CUSTOM) {
                if (!o.custom_evaluate)
                    fail(QLX_EINVAL, "TranspilerConfig: CostModelType::CUSTOM specified but custom_cost_model is null.");
                cm.custom = o.custom_evaluate;
                cm.ctx = o.custom_ctx;
            } else if (o.cost_model != QLX_COST_THRESHOLD) {
                fail(QLX_EINVAL, "unknown windowing cost model");
            }
            qubit_windowing(dag, cm, nq, o.restore_layout != 0);
        }
        prog->instrs = release_dag(std::move(dag));
        *out = prog.release();
    });
}

int qlx_program_sizes(const qlx_program* p, size_t* n_instr, size_t* pool_elems, size_t* n_meta) {
    return guard([&] {
        if (!p) fail(QLX_EINVAL, "null program");
        if (n_instr) *n_instr = p->instrs.size();
        if (pool_elems) *pool_elems = p->pool.data.size();
        if (n_meta) *n_meta = p->pool.meta.size();
    });
}

int qlx_program_copy(const qlx_program* p, qlx_gate_instr* instrs, double* pool, qlx_matrix_entry* meta) {
    return guard([&] {
        if (!p) fail(QLX_EINVAL, "null program");
        if (instrs && !p->instrs.empty()) std::memcpy(instrs, p->instrs.data(), p->instrs.size() * sizeof(qlx_gate_instr));
        if (pool && !p->pool.data.empty()) std::memcpy(pool, p->pool.data.data(), p->pool.data.size() * sizeof(cd));
        if (meta && !p->pool.meta.empty())
            std::memcpy(meta, p->pool.meta.data(), p->pool.meta.size() * sizeof(qlx_matrix_entry));
    });
}

int qlx_program_destroy(qlx_program* p) {
    return guard([&] { delete p; });
}

}  // extern "C"
