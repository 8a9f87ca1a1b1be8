// qlx_jit.cpp — circuit-specialised window-sweep kernels, compiled at run time (hipRTC).
//
// The generic sweep kernel (qlx_sweep_core.hpp) interprets a sweep's descriptors: a
// runtime switch per op, register moves at the switch merge, LDS addresses built per
// register. For a program that is executed many times on a large state, the same
// sweeps can instead be emitted as straight-line HIP source — every register index,
// qubit position, LDS offset and gate coefficient a compile-time constant, CX/SWAP
// reduced to renaming registers in the generator — and compiled once for gfx950.
// The generated code follows the generic kernel operation by operation (the same
// descriptors: DevSweep / DevPhase / DevOp and the coefficient array, including the
// factored H / RX forms and the OP_SCALE steps the planner writes), so both compute the
// same gates in the same order with the same expressions.

#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "qlx_jit.hpp"

namespace qlx {
namespace {

std::string lit(double x, bool f32) {
    char b[64];
    if (f32) std::snprintf(b, sizeof b, "%af", (double)(float)x);
    else std::snprintf(b, sizeof b, "%a", x);
    return b;
}

std::string hex64(uint64_t x) {
    char b[32];
    std::snprintf(b, sizeof b, "0x%llxull", (unsigned long long)x);
    return b;
}

// swz of the generic kernel (linear over XOR on disjoint bit sets).
uint32_t swz_host(uint32_t l, int elem) {
    const int sh = elem == 16 ? 4 : 5;
    const uint32_t m = (1u << sh) - 1;
    return l ^ (((l >> sh) ^ (l >> (2 * sh)) ^ (l >> (3 * sh))) & m);
}

// The LDS slot of a local index during one re-layout: an invertible linear map over GF(2),
// given by the slot bits of each local bit (col[j]); slot(l) = XOR of col[j] over the bits j of l.
struct LdsMap {
    std::vector<uint32_t> col;
    // the reading layout's register bits map to slot bits that its work-item bits never set, so
    // a read's slot is the work-item part plus a constant (an immediate offset of the LDS read)
    bool read_disjoint = false;
    uint32_t apply(uint32_t l) const {
        uint32_t a = 0;
        for (size_t j = 0; j < col.size(); ++j)
            if ((l >> j) & 1u) a ^= col[j];
        return a;
    }
};

// rank over GF(2) of a set of bit vectors
int gf2_rank(std::vector<uint32_t> v) {
    int r = 0;
    for (int bit = 31; bit >= 0; --bit) {
        size_t piv = r;
        while (piv < v.size() && !((v[piv] >> bit) & 1u)) ++piv;
        if (piv == v.size()) continue;
        std::swap(v[r], v[piv]);
        for (size_t k = 0; k < v.size(); ++k)
            if (k != (size_t)r && ((v[k] >> bit) & 1u)) v[k] ^= v[r];
        ++r;
    }
    return r;
}

// Bank rules of 16-byte LDS accesses on gfx950, fitted exactly to the SQ_LDS_BANK_CONFLICT
// counter over 30 lane -> slot maps that permute address bits and the 52 re-layouts of the 30-qubit
// layered circuit: a ds_write_b128 serves the 8 lanes of each lane-bit-0..2 group in one cycle
// when their slot bits 0-2 differ; a ds_read_b128 serves each 32-lane half in two cycles, the
// lanes whose lane bits 2, 3 and 4 have even parity in one and the others in the other, each
// without conflicts when its 16 lanes differ in slot bits 0-3. `w` holds the local-index
// vectors of lane bits 0-2 of the writing layout, `r` those of lane bits 0-4 of the reading one.
bool lds_conflict_free(const LdsMap& m, const std::vector<uint32_t>& w, const std::vector<uint32_t>& r) {
    std::vector<uint32_t> a;
    for (uint32_t x : w) a.push_back(m.apply(x) & 7u);
    if (gf2_rank(a) < 3) return false;
    // the even-parity lanes of a half: spanned by lane bits 0, 1, 2^3 and 3^4
    const uint32_t e[4] = {r[0], r[1], r[2] ^ r[3], r[3] ^ r[4]};
    std::vector<uint32_t> b;
    for (uint32_t x : e) b.push_back(m.apply(x) & 15u);
    return gf2_rank(b) == 4;
}

// A slot map for one re-layout of 16-byte elements (w local bits; `rregs`: the local bits the
// reading layout keeps in registers). First choice (a seeded search): the reading layout's
// register bits on slot bits 4.. of their own, its work-item bits mixed over the other slot
// bits, conflict-free for this write and read, so the reads take immediate offsets. Otherwise
// the generic kernel's swizzle when it is conflict-free, or one whose slot bits 0-3 are parities
// found by a seeded search and whose other slot bits are local bits completing a basis.
LdsMap choose_lds_map(int w, const std::vector<uint32_t>& wv, const std::vector<uint32_t>& rv, uint32_t rregs,
                      uint64_t seed) {
    uint64_t s = seed * 0x9e3779b97f4a7c15ull + 1;
    auto next = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    };
    std::vector<int> rb, tb;  // register / work-item local bits of the reading layout
    for (int j = 0; j < w; ++j) ((rregs >> j) & 1u ? rb : tb).push_back(j);
    if (!rb.empty() && (int)rb.size() <= w - 4) {
        for (int attempt = 0; attempt < 20000; ++attempt) {
            // register bits -> distinct slot bits >= 4; work-item bits -> independent vectors over the rest
            std::vector<int> hi;
            for (int k = 4; k < w; ++k) hi.push_back(k);
            for (size_t k = hi.size(); k > 1; --k) std::swap(hi[k - 1], hi[next() % k]);
            uint32_t bmask = 0;
            LdsMap c;
            c.col.assign(w, 0u);
            for (size_t k = 0; k < rb.size(); ++k) {
                c.col[rb[k]] = 1u << hi[k];
                bmask |= 1u << hi[k];
            }
            const uint32_t free = ((1u << w) - 1) & ~bmask;
            std::vector<uint32_t> tv;
            for (int j : tb) {
                uint32_t v = 0;
                while (!v) v = (uint32_t)next() & free;
                c.col[j] = v;
                tv.push_back(v);
            }
            if (gf2_rank(tv) < (int)tv.size()) continue;
            if (!lds_conflict_free(c, wv, rv)) continue;
            c.read_disjoint = true;
            return c;
        }
    }
    LdsMap m;
    m.col.resize(w);
    for (int j = 0; j < w; ++j) m.col[j] = swz_host(1u << j, 16);
    if (lds_conflict_free(m, wv, rv)) return m;
    const uint32_t full = (1u << w) - 1;
    for (int attempt = 0; attempt < 100000; ++attempt) {
        uint32_t rho[4];
        for (uint32_t& x : rho) x = (uint32_t)next() & full;
        if (gf2_rank({rho[0], rho[1], rho[2], rho[3]}) < 4) continue;
        // slot bits 4.. : the local bits that complete rho to a basis, ascending
        std::vector<uint32_t> basis(rho, rho + 4);
        std::vector<int> extra;
        for (int j = 0; j < w && (int)basis.size() < w; ++j) {
            basis.push_back(1u << j);
            if (gf2_rank(basis) < (int)basis.size()) basis.pop_back();
            else extra.push_back(j);
        }
        LdsMap c;
        c.col.assign(w, 0u);
        for (int j = 0; j < w; ++j) {
            for (int k = 0; k < 4; ++k) c.col[j] |= ((rho[k] >> j) & 1u) << k;
            for (size_t e = 0; e < extra.size(); ++e)
                if (extra[e] == j) c.col[j] |= 1u << (4 + e);
        }
        if (lds_conflict_free(c, wv, rv)) return c;
    }
    return m;
}

struct Gen {
    const SweepPlan& p;
    int si;
    bool f32;
    int logt, rb, w, R, elem;
    std::string out;
    std::vector<int> map;  // register index -> variable
    // Pauli frame of the registers: per register bit r, xf[r] and zf[r] are masks of work-item
    // bits whose parities, per work-item, give the exponents of X and Z on that bit. The chunk
    // is (prod_r X_r^x Z_r^z) applied to the registers: register i holds the amplitude of
    // register position i ^ f (f = the bits with x = 1), negated for each bit r with z = 1 that
    // is set in i ^ f. A CX whose control is a work-item bit (OP_CXT) adds an X; Clifford ops
    // on register bits move the frame (H swaps a bit's X and Z, CX spreads X from control to
    // target and Z from target to control), RX carries X and runs with a per-work-item sign on
    // its angle under Z; other ops, LDS re-layouts and the stores apply it.
    std::vector<uint32_t> xf, zf;
    int tmp = 0;
    void line(const std::string& s) { out += s; out += '\n'; }
    std::string var(int i) const { return "r" + std::to_string(map[i]); }
    bool frame_pending() const {
        for (uint32_t m : xf)
            if (m) return true;
        return false;
    }
    // applies the pending Z on register bit r: negates, per work-item, the registers with bit r set
    void materialize_z(int r) {
        if (!zf[r]) return;
        const std::string U = f32 ? "unsigned" : "unsigned long long";
        const std::string F = f32 ? "float" : "double";
        const std::string sign = f32 ? "0x80000000u" : "0x8000000000000000ull";
        line("  { const " + U + " zs = " + parity(zf[r]) + " ? " + sign + " : 0;");
        for (int i = 0; i < R; ++i) {
            if (!(i & (1 << r))) continue;
            const std::string x = var(i);
            line("    " + x + ".x = __builtin_bit_cast(" + F + ", __builtin_bit_cast(" + U + ", " + x + ".x) ^ zs); " +
                 x + ".y = __builtin_bit_cast(" + F + ", __builtin_bit_cast(" + U + ", " + x + ".y) ^ zs);");
        }
        line("  }");
        zf[r] = 0;
    }
    void materialize_all_z() {
        for (int r = 0; r < rb; ++r) materialize_z(r);
    }
    std::string parity(uint32_t mask) const {
        return "((__builtin_popcount(" + tv + " & " + std::to_string(mask) + "u) & 1u) != 0u)";
    }
    // applies the pending X on register bit r: conditional swaps of the register pairs
    void materialize(int r) {
        if (!xf[r]) return;
        const std::string T2 = f32 ? "float2" : "double2";
        line("  { const bool fc = " + parity(xf[r]) + ";");
        for (int i = 0; i < R; ++i) {
            if (i & (1 << r)) continue;
            const std::string x = var(i), y = var(i | (1 << r));
            line("    { const " + T2 + " a0 = " + x + ", a1 = " + y + "; " + x + " = fc ? a1 : a0; " + y +
                 " = fc ? a0 : a1; }");
        }
        line("  }");
        xf[r] = 0;
    }
    // per-work-item XOR of `offset(bit r)` over the flipped register bits, as `name`
    template <class F>
    void frame_offset(const std::string& type, const std::string& name, F offset) {
        line("  " + type + " " + name + " = 0;");
        for (int r = 0; r < rb; ++r)
            if (xf[r]) line("  if " + parity(xf[r]) + " " + name + " ^= " + offset(r) + ";");
    }

    std::string tv = "t";  // the variable holding the work-item index

    std::string thread_bits(const uint32_t* pos) const {
        std::string e = "0u";
        for (int b = 0; b < logt; ++b)
            e += " | (((" + tv + " >> " + std::to_string(b) + ") & 1u) << " + std::to_string(pos[b]) + ")";
        return e;
    }
    // LDS byte address of the work-item part of a layout: the XOR of the slots (times the element
    // size) of the local bits its work-item bits hold (each work-item bit selects a constant by a
    // sign-extended bit mask)
    std::string thread_slot(const uint32_t* pos, const LdsMap& M) const {
        std::string e = "0u";
        for (int b = 0; b < logt; ++b) {
            const uint32_t c = M.apply(1u << pos[b]) * (uint32_t)elem;
            if (c) e += " ^ ((0u - ((" + tv + " >> " + std::to_string(b) + ") & 1u)) & " + std::to_string(c) + "u)";
        }
        return e;
    }
    // Complex128 out-of-place sweeps re-lay out in two rounds through half a chunk of LDS (32 KiB):
    // slot bit k (>= 4, so the banks are those of the one-round map) picks the round, and the
    // buffer holds the slots without bit k. Registers read in round 0 go to new variables, so the
    // values that round 1 still writes stay intact. With 32 KiB per work-group the compute unit
    // holds 3-4 work-groups instead of 2 (the sweeps use 117-168 VGPRs): on the 30-qubit layered
    // circuit 65.1 -> 63.4 ms of sweep time (kernel traces on one box), every sweep equal or faster.
    bool two_round_ok() const { return elem == 16 && logt == 8 && p.relabel; }
    void emit_two_round(const LdsMap& M, const DevPhase& cur, const DevPhase& P) {
        const std::string T2 = f32 ? "float2" : "double2";
        // work-item columns (write side with the frame) and the round bit with the fewest of them
        std::vector<uint32_t> wcol(logt), rcol(logt);
        for (int b = 0; b < logt; ++b) {
            uint32_t u = 1u << cur.tlocal[b];
            for (int r = 0; r < rb; ++r)
                if ((xf[r] >> b) & 1u) u ^= cur.roff[1 << r];
            wcol[b] = M.apply(u);
            rcol[b] = M.apply(1u << P.tlocal[b]);
        }
        int k = -1, best = 1 << 30;
        for (int kk = 4; kk < w; ++kk) {
            int cost = 0;
            for (int b = 0; b < logt; ++b) {
                const int wt = b >= 6 ? 1 : 100;  // lane-dependent rounds diverge
                cost += (((wcol[b] >> kk) & 1u) + ((rcol[b] >> kk) & 1u)) * wt;
            }
            if (cost < best) { best = cost; k = kk; }
        }
        const uint32_t kb = (uint32_t)k + 4;  // in byte addresses (16-byte elements)
        auto comp = [&](uint32_t x) { return ((x >> (kb + 1)) << kb) | (x & ((1u << kb) - 1)); };
        line("  {");
        line("    const unsigned s0 = " + thread_slot(cur.tlocal, M) + ";");
        std::string s0v = "s0";
        if (frame_pending()) {
            frame_offset("unsigned", "fx", [&](int r) {
                return std::to_string(M.apply(cur.roff[1 << r]) * elem) + "u";
            });
            line("    const unsigned s0f = s0 ^ fx;");
            s0v = "s0f";
            std::fill(xf.begin(), xf.end(), 0u);
        }
        const std::string K = std::to_string(kb), K1 = std::to_string(kb + 1);
        line("    const unsigned rw = (" + s0v + " >> " + K + ") & 1u;");
        line("    const unsigned a0 = ((" + s0v + " >> " + K1 + ") << " + K + ") | (" + s0v + " & ((1u << " + K + ") - 1u));");
        line("    const unsigned s1 = " + thread_slot(P.tlocal, M) + ";");
        line("    const unsigned rr = (s1 >> " + K + ") & 1u;");
        line("    const unsigned a1 = ((s1 >> " + K1 + ") << " + K + ") | (s1 & ((1u << " + K + ") - 1u));");
        std::string decl = "    " + T2;
        for (int i = 0; i < R; ++i) decl += std::string(i ? "," : "") + " n" + std::to_string(i);
        line(decl + ";");
        for (int round = 0; round < 2; ++round) {
            // writes: register i is in this round when rw ^ (bit k of its constant) == round
            for (int side = 0; side < 2; ++side) {
                line(std::string("    ") + (side == 0 ? "if" : "else") + (side == 0 ? " (rw == " + std::to_string(round) + "u) {" : " {"));
                for (int i = 0; i < R; ++i) {
                    const uint32_t c = M.apply(cur.roff[i]) * elem;
                    if ((int)((c >> kb) & 1u) != side) continue;
                    line("      *reinterpret_cast<" + T2 + "*>(ldsb + (a0 ^ " + std::to_string(comp(c)) + "u)) = " + var(i) + ";");
                }
                line("    }");
            }
            line("    __syncthreads();");
            for (int side = 0; side < 2; ++side) {
                line(std::string("    ") + (side == 0 ? "if" : "else") + (side == 0 ? " (rr == " + std::to_string(round) + "u) {" : " {"));
                for (int i = 0; i < R; ++i) {
                    const uint32_t c = M.apply(P.roff[i]) * elem;
                    if ((int)((c >> kb) & 1u) != side) continue;
                    line("      n" + std::to_string(i) + " = *reinterpret_cast<const " + T2 + "*>(ldsb + (a1 " +
                         (M.read_disjoint ? "+ " : "^ ") + std::to_string(comp(c)) + "u));");
                }
                line("    }");
            }
            line("    __syncthreads();");
        }
        for (int i = 0; i < R; ++i) {
            map[i] = i;
            line("    r" + std::to_string(i) + " = n" + std::to_string(i) + ";");
        }
        line("  }");
    }
    std::string thread_bits64(const uint32_t* pos) const {
        std::string e = "0ull";
        for (int b = 0; b < logt; ++b)
            e += " | ((unsigned long long)((t >> " + std::to_string(b) + ") & 1u) << " + std::to_string(pos[b]) + ")";
        return e;
    }

    void op(const DevOp& o) {
        const int kind = (int)(o.key >> 8), b0 = (int)((o.key >> 4) & 15), b1 = (int)(o.key & 15);
        const std::string T = f32 ? "float" : "double", T2 = f32 ? "float2" : "double2";
        const double* m = p.coef.data() + o.coef;
        const std::string id = std::to_string(tmp++);
        // the pending frame through the op (Z before X where both are applied: the registers
        // hold X^x Z^z of the chunk)
        const bool rx_sign = kind == OP_RXT && zf[b0] != 0u;  // RX(theta) Z = Z RX(-theta)
        switch (kind) {
            case OP_RXT: case OP_SCALE: break;
            case OP_RXU: materialize_z(b0); break;
            case OP_CXT: materialize_z(b1); break;               // Z on the target would reach the control
            case OP_CX:                                          // control b0, target b1
                xf[b1] ^= xf[b0];
                zf[b0] ^= zf[b1];
                break;
            case OP_SWAP:
                std::swap(xf[b0], xf[b1]);
                std::swap(zf[b0], zf[b1]);
                break;
            case OP_HB:                                          // H X = Z H, H Z = X H
                if (xf[b0] && zf[b0]) materialize_z(b0);
                std::swap(xf[b0], zf[b0]);
                break;
            case OP_U4:
                materialize_z(b0);
                materialize(b0);
                materialize_z(b1);
                materialize(b1);
                break;
            default:                                             // OP_U2
                materialize_z(b0);
                materialize(b0);
                break;
        }
        switch (kind) {
            case OP_CXT:
                // control on work-item bit b0, target register bit b1: recorded, applied later
                xf[b1] ^= 1u << b0;
                break;
            case OP_RXT: {
                // RX / c: a0' = a0 - i t a1, a1' = a1 - i t a0
                const std::string F = f32 ? "__builtin_fmaf" : "__builtin_fma";
                // (the coefficient is `tc`: `t` is the work-item index the frame parities read)
                if (rx_sign)
                    line("  { const " + T + " tc = " + parity(zf[b0]) + " ? -(" + lit(o.c0, f32) + ") : (" +
                         lit(o.c0, f32) + ");");
                else
                    line("  { const " + T + " tc = " + lit(o.c0, f32) + ";");
                for (int i = 0; i < R; ++i) {
                    if (i & (1 << b0)) continue;
                    const std::string x = var(i), y = var(i | (1 << b0));
                    line("    { const " + T2 + " a0 = " + x + ", a1 = " + y + "; " + x + " = " + T2 + "{" + F +
                         "(tc, a1.y, a0.x), " + F + "(-tc, a1.x, a0.y)}; " + y + " = " + T2 + "{" + F +
                         "(tc, a0.y, a1.x), " + F + "(-tc, a0.x, a1.y)}; }");
                }
                line("  }");
                break;
            }
            case OP_RXU: {
                // RX / s: a0' = u a0 - i a1, a1' = u a1 - i a0
                const std::string F = f32 ? "__builtin_fmaf" : "__builtin_fma";
                line("  { const " + T + " u = " + lit(o.c0, f32) + ";");
                for (int i = 0; i < R; ++i) {
                    if (i & (1 << b0)) continue;
                    const std::string x = var(i), y = var(i | (1 << b0));
                    line("    { const " + T2 + " a0 = " + x + ", a1 = " + y + "; " + x + " = " + T2 + "{" + F +
                         "(u, a0.x, a1.y), " + F + "(u, a0.y, -a1.x)}; " + y + " = " + T2 + "{" + F +
                         "(u, a1.x, a0.y), " + F + "(u, a1.y, -a0.x)}; }");
                }
                line("  }");
                break;
            }
            case OP_HB: {
                // H without its factor: the butterfly
                line("  {");
                for (int i = 0; i < R; ++i) {
                    if (i & (1 << b0)) continue;
                    const std::string x = var(i), y = var(i | (1 << b0));
                    line("    { const " + T2 + " a0 = " + x + ", a1 = " + y + "; " + x + " = " + T2 +
                         "{a0.x + a1.x, a0.y + a1.y}; " + y + " = " + T2 + "{a0.x - a1.x, a0.y - a1.y}; }");
                }
                line("  }");
                break;
            }
            case OP_SCALE: {
                line("  { const " + T + " k = " + lit(o.c0, f32) + ";");
                std::string l = "   ";
                for (int i = 0; i < R; ++i) l += " " + var(i) + ".x *= k; " + var(i) + ".y *= k;";
                line(l);
                line("  }");
                break;
            }
            case OP_U2: {
                std::string d = "  { const " + T + " ";
                const char* nm[8] = {"m0r", "m0i", "m1r", "m1i", "m2r", "m2i", "m3r", "m3i"};
                for (int k = 0; k < 8; ++k) d += std::string(k ? ", " : "") + nm[k] + " = " + lit(m[k], f32);
                line(d + ";");
                for (int i = 0; i < R; ++i) {
                    if (i & (1 << b0)) continue;
                    const std::string x = var(i), y = var(i | (1 << b0));
                    line("    { const " + T2 + " a0 = " + x + ", a1 = " + y + "; " + x + " = " + T2 +
                         "{m0r * a0.x - m0i * a0.y + m1r * a1.x - m1i * a1.y, m0r * a0.y + m0i * a0.x + m1r * a1.y + m1i * a1.x}; " +
                         y + " = " + T2 +
                         "{m2r * a0.x - m2i * a0.y + m3r * a1.x - m3i * a1.y, m2r * a0.y + m2i * a0.x + m3r * a1.y + m3i * a1.x}; }");
                }
                line("  }");
                break;
            }
            case OP_CX:  // control b0, target b1: rename
                for (int i = 0; i < R; ++i)
                    if ((i & (1 << b0)) && !(i & (1 << b1))) std::swap(map[i], map[i | (1 << b1)]);
                break;
            case OP_SWAP:
                for (int i = 0; i < R; ++i)
                    if ((i & (1 << b0)) && !(i & (1 << b1))) std::swap(map[i], map[i ^ (1 << b0) ^ (1 << b1)]);
                break;
            case OP_U4: {
                std::string d = "  { const " + T + " ";
                for (int k = 0; k < 32; ++k) d += std::string(k ? ", " : "") + "u" + std::to_string(k) + " = " + lit(m[k], f32);
                line(d + ";");
                for (int i = 0; i < R; ++i) {
                    if ((i & (1 << b0)) || (i & (1 << b1))) continue;
                    const int idx[4] = {i, i | (1 << b0), i | (1 << b1), i | (1 << b0) | (1 << b1)};
                    std::string s = "    { const " + T2 + " x0 = " + var(idx[0]) + ", x1 = " + var(idx[1]) + ", x2 = " +
                                    var(idx[2]) + ", x3 = " + var(idx[3]) + ";";
                    for (int r = 0; r < 4; ++r) {
                        s += " { " + T + " re = " + T + "(0), im = " + T + "(0);";
                        for (int c = 0; c < 4; ++c) {
                            const std::string ur = "u" + std::to_string(2 * (4 * r + c));
                            const std::string ui = "u" + std::to_string(2 * (4 * r + c) + 1);
                            const std::string xc = "x" + std::to_string(c);
                            s += " re += " + ur + " * " + xc + ".x - " + ui + " * " + xc + ".y; im += " + ur + " * " + xc +
                                 ".y + " + ui + " * " + xc + ".x;";
                        }
                        s += " " + var(idx[r]) + " = " + T2 + "{re, im}; }";
                    }
                    line(s + " }");
                }
                line("  }");
                break;
            }
            default:
                fail(QLX_EUNSUPPORTED, "specialised sweep: unknown op kind");
        }
    }

    // address bits of chunk `c` outside the window: in place, zero bits inserted at the
    // window's qubit positions (ascending); out of place (relabel), chunk-index bit j moved to
    // address bit cpos[j] of the load (store = false) or store layout
    void emit_base(const std::string& c, const std::string& out, const std::string& indent, bool store = false) {
        const DevSweep& sw = p.dsw[si];
        if (p.relabel) {
            const std::vector<int>& cpos = store ? p.sweeps[si].st_cpos : p.sweeps[si].ld_cpos;
            std::string e = "0ull";
            for (size_t j = 0; j < cpos.size(); ++j)
                e += " | (((" + c + ") >> " + std::to_string(j) + " & 1ull) << " + std::to_string(cpos[j]) + ")";
            line(indent + "const unsigned long long " + out + " = " + e + ";");
            return;
        }
        line(indent + "unsigned long long " + out + " = " + c + ";");
        for (int j = 0; j < w; ++j) {
            const unsigned q = sw.S[j];
            line(indent + out + " = ((" + out + " >> " + std::to_string(q) + ") << " + std::to_string(q + 1) + ") | (" +
                 out + " & " + hex64((1ull << q) - 1) + ");");
        }
    }

    // Persistent kernels (a chunk of more than 64 KiB, so one work-group per CU, and at
    // most 16 registers): the work-group runs chunks c, c + G, ... and loads the next chunk
    // into a second register set before computing the current one, which gives the CU the
    // overlap of load and compute that two resident work-groups give otherwise. With two
    // work-groups per CU the hardware's own overlap measured faster (30-qubit layered
    // circuit, 12-qubit window: 117 ms against 128 ms persistent; 13-qubit window: 119 ms
    // persistent against 136 ms with one chunk per work-group).
    void emit() {
        const DevSweep& sw = p.dsw[si];
        const std::string T2 = f32 ? "float2" : "double2";
        // the persistent form is not generated for out-of-place sweeps
        const bool persistent = jit_persistent(p.geo, f32 ? QLX_C64 : QLX_C128) && !p.relabel;
        // persistent kernels hold two register sets: ask for at least two waves per SIMD
        // in place: one pointer; out of place (relabel): read src, write dst
        const std::string ld = p.relabel ? "src" : "amp", st = p.relabel ? "dst" : "amp";
        const std::string params = p.relabel ? "const " + T2 + "* __restrict__ src, " + T2 + "* __restrict__ dst"
                                             : T2 + "* __restrict__ amp";
        line("extern \"C\" __global__ void __launch_bounds__(" + std::to_string(1 << logt) + ")" +
             (persistent ? " __attribute__((amdgpu_waves_per_eu(2)))" : "") + " " + jit_kernel_name(p, si) + "(" +
             params + ", unsigned long long nchunks) {");
        const bool two = two_round_ok();
        line("  __shared__ " + T2 + " lds[" + std::to_string(1 << (two ? w - 1 : w)) + "];");
        line("  char* const ldsb = reinterpret_cast<char*>(lds);");
        line("  const unsigned t = threadIdx.x;");
        line("  unsigned long long c = blockIdx.x;");
        line("  if (c >= nchunks) return;");
        line("  const unsigned long long tl = " + thread_bits64(sw.ld_tglob) + ";");
        line("  const unsigned long long ts = " + thread_bits64(sw.st_tglob) + ";");
        emit_base("c", "base", "  ");
        line("  const unsigned long long lo = base | tl;");
        map.resize(R);
        xf.assign(rb, 0u);
        zf.assign(rb, 0u);
        for (int i = 0; i < R; ++i) {
            map[i] = i;
            line("  " + T2 + " r" + std::to_string(i) + " = " + ld + "[lo | " + hex64(sw.ld_roff[i]) + "];");
        }
        if (persistent) {
            line("  for (;;) {");
            // an opaque copy of the work-item index: the LDS addresses of every phase derive
            // from it, and hoisting them all out of the loop would not fit the register file
            line("  unsigned tt = t;");
            line("  asm volatile(\"\" : \"+v\"(tt));");
            tv = "tt";
            line("  const unsigned long long nc = c + gridDim.x;");
            line("  const bool more = nc < nchunks;");
            // the last iteration reloads its own chunk (loaded before the stores, never used)
            emit_base("(more ? nc : c)", "nbase", "  ");
            line("  const unsigned long long nlo = nbase | tl;");
            for (int i = 0; i < R; ++i)
                line("  const " + T2 + " p" + std::to_string(i) + " = " + ld + "[nlo | " + hex64(sw.ld_roff[i]) + "];");
        }
        const DevPhase* cur = &p.dph[sw.phase_begin];
        for (uint32_t ph = 0; ph < sw.n_phases; ++ph) {
            const DevPhase& P = p.dph[sw.phase_begin + ph];
            if (P.remap) {
                materialize_all_z();  // Z: sign flips in the registers; X: folded into the addresses
                // the re-layout's LDS slot map: linear, so a local index's slot is the XOR of the
                // slots of its work-item part, its register part and (frame) the flipped register bits
                LdsMap M;
                if (elem == 16 && logt >= 5) {
                    std::vector<uint32_t> wv, rv;
                    for (int b = 0; b < 3; ++b) {
                        uint32_t u = 1u << cur->tlocal[b];
                        for (int r = 0; r < rb; ++r)
                            if ((xf[r] >> b) & 1u) u ^= cur->roff[1 << r];
                        wv.push_back(u);
                    }
                    for (int b = 0; b < 5; ++b) rv.push_back(1u << P.tlocal[b]);
                    uint32_t rregs = 0;
                    for (int r = 0; r < rb; ++r) rregs |= P.roff[1 << r];
                    M = choose_lds_map(w, wv, rv, rregs, (uint64_t)si * 1000003ull + ph);
                } else {
                    M.col.resize(w);
                    for (int j = 0; j < w; ++j) M.col[j] = swz_host(1u << j, elem);
                }
                if (two) {
                    emit_two_round(M, *cur, P);
                    cur = &P;
                    for (uint32_t k = 0; k < P.n_ops; ++k) op(p.dop[P.op_begin + k]);
                    continue;
                }
                line("  {");
                // byte addresses: slot x element size
                line("    const unsigned s0 = " + thread_slot(cur->tlocal, M) + ";");
                std::string s0v = "s0";
                if (frame_pending()) {
                    // register i holds the amplitude of position i ^ f: write it there (the slot
                    // map is linear, so the frame adds a per-work-item XOR to every address)
                    frame_offset("unsigned", "fx", [&](int r) {
                        return std::to_string(M.apply(cur->roff[1 << r]) * elem) + "u";
                    });
                    line("    const unsigned s0f = s0 ^ fx;");
                    s0v = "s0f";
                    std::fill(xf.begin(), xf.end(), 0u);
                }
                for (int i = 0; i < R; ++i)
                    line("    *reinterpret_cast<" + T2 + "*>(ldsb + (" + s0v + " ^ " +
                         std::to_string(M.apply(cur->roff[i]) * elem) + "u)) = " + var(i) + ";");
                line("    __syncthreads();");
                line("    const unsigned s1 = " + thread_slot(P.tlocal, M) + ";");
                for (int i = 0; i < R; ++i) {
                    map[i] = i;
                    const std::string off = std::to_string(M.apply(P.roff[i]) * elem) + "u";
                    line("    r" + std::to_string(i) + " = *reinterpret_cast<const " + T2 + "*>(ldsb + (s1 " +
                         (M.read_disjoint ? "+ " : "^ ") + off + "));");
                }
                line("    __syncthreads();");
                line("  }");
            }
            cur = &P;
            for (uint32_t k = 0; k < P.n_ops; ++k) op(p.dop[P.op_begin + k]);
        }
        materialize_all_z();
        if (p.relabel) {
            emit_base("c", "sbase", "  ", true);
            line("  const unsigned long long so = sbase | ts;");
        } else {
            line("  const unsigned long long so = base | ts;");
        }
        if (frame_pending()) {
            // register i holds the amplitude of register position i ^ f: store it there
            frame_offset("unsigned long long", "fo", [&](int r) { return hex64(sw.st_roff[1 << r]); });
            line("  const unsigned long long sf = so ^ fo;");
            for (int i = 0; i < R; ++i) line("  " + st + "[sf ^ " + hex64(sw.st_roff[i]) + "] = " + var(i) + ";");
            std::fill(xf.begin(), xf.end(), 0u);
        } else {
            for (int i = 0; i < R; ++i) line("  " + st + "[so | " + hex64(sw.st_roff[i]) + "] = " + var(i) + ";");
        }
        if (persistent) {
            line("  if (!more) break;");
            std::string l = " ";
            for (int i = 0; i < R; ++i) l += " r" + std::to_string(i) + " = p" + std::to_string(i) + ";";
            line(l);
            line("  c = nc;");
            line("  base = nbase;");
            line("  }");
        }
        line("}");
    }
};

std::vector<char> compile(const std::string& src, const std::string& name) {
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), name.c_str(), 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        fail(QLX_EHIP, "hiprtcCreateProgram failed");
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, 6, opts);
    std::vector<char> code;
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, log.data());
        hiprtcDestroyProgram(&prog);
        fail(QLX_EHIP, "specialised sweep compilation failed: " + log.substr(0, 2000));
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.resize(n);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return code;
}

}  // namespace

std::string jit_kernel_name(const SweepPlan& p, int sweep_index) {
    char b[64];
    std::snprintf(b, sizeof b, "qlx_sweep_spec_%06llx_%d", (unsigned long long)(p.tag & 0xffffff), sweep_index);
    return b;
}

bool jit_persistent(const Geometry& g, int dtype) {
    const size_t chunk_bytes = (size_t(dtype == QLX_C64 ? 8 : 16)) << g.w;
    return g.rb <= 4 && chunk_bytes > (size_t(64) << 10);
}

std::string jit_sweep_source(const SweepPlan& p, int sweep_index, int dtype) {
    if (sweep_index < 0 || (size_t)sweep_index >= p.dsw.size()) fail(QLX_ERANGE, "sweep index out of range");
    Gen g{p, sweep_index, dtype == QLX_C64, p.geo.logt, p.geo.rb, p.geo.w, 1 << p.geo.rb, dtype == QLX_C64 ? 8 : 16, {}, {}};
    g.emit();
    return g.out;
}

JitSweeps::~JitSweeps() {
    for (hipModule_t m : modules)
        if (m) (void)hipModuleUnload(m);
}

std::vector<std::vector<char>> compile_jit_sweeps(const SweepPlan& p, int dtype, const std::atomic<bool>* cancel) {
    const size_t n = p.dsw.size();
    std::vector<std::string> src(n);
    for (size_t s = 0; s < n; ++s) src[s] = jit_sweep_source(p, (int)s, dtype);
    std::vector<std::vector<char>> code(n);
    std::vector<std::string> err(n);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    std::mutex mu;
    size_t next = 0;
    for (unsigned k = 0; k < std::min<size_t>(hw, n); ++k)
        pool.emplace_back([&] {
            for (;;) {
                size_t s;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    if (next >= n) return;
                    s = next++;
                }
                if (cancel && cancel->load(std::memory_order_relaxed)) {
                    err[s] = "specialised sweep compilation cancelled";
                    continue;
                }
                try {
                    code[s] = compile(src[s], jit_kernel_name(p, (int)s) + ".hip");
                } catch (const Error& e) {
                    err[s] = e.msg;
                }
            }
        });
    for (auto& th : pool) th.join();
    for (size_t s = 0; s < n; ++s)
        if (!err[s].empty()) fail(QLX_EHIP, err[s]);
    return code;
}

std::shared_ptr<JitSweeps> load_jit_sweeps(const SweepPlan& p, int dtype, const std::vector<std::vector<char>>& code,
                                           double compile_s) {
    const size_t n = p.dsw.size();
    if (code.size() != n) fail(QLX_EINVAL, "specialised sweeps: code objects do not match the plan");
    auto j = std::make_shared<JitSweeps>();
    j->fns.resize(n);
    j->modules.resize(n, nullptr);
    for (size_t s = 0; s < n; ++s) {
        QLX_HIP_CHECK(hipModuleLoadData(&j->modules[s], code[s].data()));
        const std::string name = jit_kernel_name(p, (int)s);
        QLX_HIP_CHECK(hipModuleGetFunction(&j->fns[s], j->modules[s], name.c_str()));
    }
    j->threads = 1u << p.geo.logt;
    j->relabel = p.relabel;
    j->persistent = jit_persistent(p.geo, dtype) && !p.relabel;
    j->resident.resize(n, 0);
    if (j->persistent) {
        int dev = 0, cus = 0;
        QLX_HIP_CHECK(hipGetDevice(&dev));
        QLX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        for (size_t s = 0; s < n; ++s) {
            int per_cu = 0;
            QLX_HIP_CHECK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, j->fns[s], (int)j->threads, 0));
            j->resident[s] = (uint64_t)std::max(per_cu, 1) * (uint64_t)std::max(cus, 1);
        }
    }
    j->compile_s = compile_s;
    return j;
}

std::shared_ptr<JitSweeps> build_jit_sweeps(const SweepPlan& p, int dtype) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::vector<std::vector<char>> code = compile_jit_sweeps(p, dtype, nullptr);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return load_jit_sweeps(p, dtype, code, s);
}

void launch_jit_sweep(const JitSweeps& j, int sweep_index, qlx_state* st, int w, uint32_t wg_override) {
    unsigned long long chunks = 1ull << (st->nq - (uint32_t)w);
    uint64_t grid = chunks;
    if (j.persistent) grid = std::min<uint64_t>(chunks, wg_override ? wg_override : j.resident[sweep_index]);
    void* amp = st->amp;
    void* alt = st->alt;
    void* in_place[] = {&amp, &chunks};
    void* out_of_place[] = {&amp, &alt, &chunks};
    if (j.relabel && (!alt || alt == amp)) fail(QLX_EHIP, "out-of-place sweep without a second buffer");
    QLX_HIP_CHECK(hipModuleLaunchKernel(j.fns[sweep_index], (unsigned)grid, 1, 1, j.threads, 1, 1, 0, st->stream,
                                        j.relabel ? out_of_place : in_place, nullptr));
    if (j.relabel) std::swap(st->amp, st->alt);
}

}  // namespace qlx
