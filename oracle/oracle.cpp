// oracle.cpp — CPU restatement of Qlassical's CPU execution path.
//
// TEST INFRASTRUCTURE ONLY. This file is the oracle the GPU backend is checked
// against; it is never linked into, loaded by, or called from the product
// library. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// use it, and only as the checker / the timed CPU baseline.
//
// It restates, statement by statement, the reference's algorithm:
//   state init            include/backend/cpu_state_vector.hpp:54-115
//   bit helpers           include/backend/state_vector_utils.hpp:23-34
//   dispatch switch       include/backend/cpu_openmp_engine.hpp:143-225
//   loop templates        include/backend/cpu_openmp_engine.hpp:260-334
//   gate kernels          include/backend/cpu_openmp_engine.hpp:337-528
//   apply_unitary         include/backend/cpu_openmp_engine.hpp:539-606
//                         (Eigen's row-major matvec replaced by an explicit
//                          row-by-row complex dot product; k <= 5 buffers)
//   apply_measure         include/backend/cpu_openmp_engine.hpp:617-671
//   FusionMath            include/middleware/middleware.hpp:241-344
//   QuantumCircuit push   include/frontend/frontend.hpp:419-460 (for the generator)
//
// GLOBAL_SWAP: mode 0 follows the reference exactly (a no-op, :217-219); mode 1
// applies the permutation the instruction names (the "correct" oracle mode used to
// check windowed GPU programs).
//
// Parity notes (bit level):
//   * permutation gates and H are exact restatements;
//   * RX/RY/RZ use the same expression trees; FMA contraction is compiler-chosen;
//   * the Eigen matvec summation order and the OpenMP reduction order are not
//     reproducible without Eigen/the same thread count: tolerance-level only.

#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

using cd = std::complex<double>;

struct Instr {
    uint8_t type;
    uint8_t flags_arity;
    uint8_t qubits[4];
    uint16_t pad;
    union {
        float param;
        uint32_t matrix_idx;
    } payload;
    uint32_t uid;
    int arity() const { return flags_arity & 0x0F; }
    int qubit(int i) const { return i < 4 ? qubits[i] : (pad & 0xFF); }
};
static_assert(sizeof(Instr) == 16, "GateInstr is 16 bytes");

struct MatEntry {
    uint32_t offset;
    uint32_t num_elements;
    uint8_t num_qubits;
    uint8_t pad[3];
};
static_assert(sizeof(MatEntry) == 12, "MatrixEntry is 12 bytes");

enum : uint8_t {
    H = 0x00, X = 0x01, Y = 0x02, Z = 0x03, S = 0x04, RX = 0x10, RY = 0x11, RZ = 0x12,
    CX = 0x20, SWAP = 0x21, CCX = 0x30, CSWAP = 0x31, UNITARY = 0x80, FUSED_BLOCK = 0xE0,
    GLOBAL_SWAP = 0xE1, BARRIER = 0xF0, MEASURE = 0xF1
};

// state_vector_utils.hpp:23-34
inline size_t insert_zero_bit(size_t idx, uint8_t target) {
    const size_t low_mask = (size_t{1} << target) - 1;
    return (idx & low_mask) | ((idx & ~low_mask) << 1);
}
inline size_t flip_target_bit(size_t idx, uint8_t target) { return idx ^ (size_t{1} << target); }

class Engine {
public:
    Engine(cd* amp, size_t dim, uint64_t seed, int gswap_mode)
        : amp_(amp), dim_(dim), gswap_mode_(gswap_mode) {
        rng_.seed(seed);
    }

    std::vector<uint8_t>& measurements() { return meas_; }

    void execute(const Instr* prog, size_t n, const cd* pool, const MatEntry* meta, size_t n_meta) {
        for (size_t k = 0; k < n; ++k) {
            const Instr& g = prog[k];
            switch (g.type) {
                case H: apply_h(g.qubits[0]); break;
                case X: apply_x(g.qubits[0]); break;
                case Y: apply_y(g.qubits[0]); break;
                case Z: apply_z(g.qubits[0]); break;
                case S: apply_s(g.qubits[0]); break;
                case RX: apply_rx(g.qubits[0], g.payload.param); break;
                case RY: apply_ry(g.qubits[0], g.payload.param); break;
                case RZ: apply_rz(g.qubits[0], g.payload.param); break;
                case CX: apply_cx(g.qubits[0], g.qubits[1]); break;
                case SWAP: apply_swap(g.qubits[0], g.qubits[1]); break;
                case CCX: apply_ccx(g.qubits[0], g.qubits[1], g.qubits[2]); break;
                case CSWAP: apply_cswap(g.qubits[0], g.qubits[1], g.qubits[2]); break;
                case UNITARY:
                case FUSED_BLOCK: {
                    if (g.payload.matrix_idx >= n_meta) throw std::out_of_range("matrix_idx out of range");
                    apply_unitary(g, pool + meta[g.payload.matrix_idx].offset);
                    break;
                }
                case BARRIER: break;
                case MEASURE: apply_measure(g.qubits[0]); break;
                case GLOBAL_SWAP:
                    if (gswap_mode_ == 1) apply_swap(g.qubits[0], g.qubits[1]);
                    break;
                default: throw std::runtime_error("CPUOpenMPEngine: Unrecognized instruction type");
            }
        }
    }

private:
    cd* amp_;
    size_t dim_;
    int gswap_mode_;
    std::vector<uint8_t> meas_;
    std::mt19937_64 rng_;
    std::uniform_real_distribution<double> dist_{0.0, 1.0};

    template <class F>
    void single_target_loop(size_t iters, F&& kernel) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)iters; ++i) kernel((size_t)i);
    }

    template <class F>
    void double_hole_loop(uint8_t q0, uint8_t q1, F&& kernel) {
        const size_t iters = dim_ / 4;
        const uint8_t qa = std::min(q0, q1), qb = std::max(q0, q1);
        const size_t m_low = (size_t{1} << qa) - 1;
        const size_t m_mid = (size_t{1} << (qb - qa - 1)) - 1;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)iters; ++i) {
            const size_t idx = (size_t)i;
            const size_t lo = idx & m_low;
            const size_t mid = (idx >> qa) & m_mid;
            const size_t hi = idx >> (qb - 1);
            kernel(lo | (mid << (qa + 1)) | (hi << (qb + 1)));
        }
    }

    template <class F>
    void triple_hole_loop(uint8_t q0, uint8_t q1, uint8_t q2, F&& kernel) {
        const size_t iters = dim_ / 8;
        const uint8_t qa = std::min({q0, q1, q2});
        const uint8_t qc = std::max({q0, q1, q2});
        const uint8_t qb = (uint8_t)(q0 + q1 + q2 - qa - qc);
        const size_t m_low = (size_t{1} << qa) - 1;
        const size_t m_mid1 = (size_t{1} << (qb - qa - 1)) - 1;
        const size_t m_mid2 = (size_t{1} << (qc - qb - 1)) - 1;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)iters; ++i) {
            const size_t idx = (size_t)i;
            const size_t lo = idx & m_low;
            const size_t mid1 = (idx >> qa) & m_mid1;
            const size_t mid2 = (idx >> (qb - 1)) & m_mid2;
            const size_t hi = idx >> (qc - 2);
            kernel(lo | (mid1 << (qa + 1)) | (mid2 << (qb + 1)) | (hi << (qc + 1)));
        }
    }

    void apply_h(uint8_t t) {
        const double r = 1.0 / std::sqrt(2.0);
        cd* a = amp_;
        single_target_loop(dim_ / 2, [a, t, r](size_t k) {
            const size_t i0 = insert_zero_bit(k, t), i1 = flip_target_bit(i0, t);
            const cd a0 = a[i0], a1 = a[i1];
            a[i0] = r * (a0 + a1);
            a[i1] = r * (a0 - a1);
        });
    }
    void apply_x(uint8_t t) {
        cd* a = amp_;
        single_target_loop(dim_ / 2, [a, t](size_t k) {
            const size_t i0 = insert_zero_bit(k, t), i1 = flip_target_bit(i0, t);
            std::swap(a[i0], a[i1]);
        });
    }
    void apply_y(uint8_t t) {
        cd* a = amp_;
        single_target_loop(dim_ / 2, [a, t](size_t k) {
            const size_t i0 = insert_zero_bit(k, t), i1 = flip_target_bit(i0, t);
            const cd a0 = a[i0], a1 = a[i1];
            a[i0] = cd(a1.imag(), -a1.real());
            a[i1] = cd(-a0.imag(), a0.real());
        });
    }
    void apply_z(uint8_t t) {
        cd* a = amp_;
        single_target_loop(dim_ / 2, [a, t](size_t k) {
            const size_t i1 = flip_target_bit(insert_zero_bit(k, t), t);
            a[i1] = -a[i1];
        });
    }
    void apply_s(uint8_t t) {
        cd* a = amp_;
        single_target_loop(dim_ / 2, [a, t](size_t k) {
            const size_t i1 = flip_target_bit(insert_zero_bit(k, t), t);
            const double re = a[i1].real(), im = a[i1].imag();
            a[i1] = cd(-im, re);
        });
    }
    void apply_rx(uint8_t t, float theta) {
        const double half = static_cast<double>(theta) * 0.5;
        const double c = std::cos(half), s = std::sin(half);
        cd* a = amp_;
        single_target_loop(dim_ / 2, [a, t, c, s](size_t k) {
            const size_t i0 = insert_zero_bit(k, t), i1 = flip_target_bit(i0, t);
            const cd a0 = a[i0], a1 = a[i1];
            a[i0] = c * a0 + cd(s * a1.imag(), -s * a1.real());
            a[i1] = cd(s * a0.imag(), -s * a0.real()) + c * a1;
        });
    }
    void apply_ry(uint8_t t, float theta) {
        const double half = static_cast<double>(theta) * 0.5;
        const double c = std::cos(half), s = std::sin(half);
        cd* a = amp_;
        single_target_loop(dim_ / 2, [a, t, c, s](size_t k) {
            const size_t i0 = insert_zero_bit(k, t), i1 = flip_target_bit(i0, t);
            const cd a0 = a[i0], a1 = a[i1];
            a[i0] = c * a0 - s * a1;
            a[i1] = s * a0 + c * a1;
        });
    }
    void apply_rz(uint8_t t, float theta) {
        const double half = static_cast<double>(theta) * 0.5;
        const cd p0(std::cos(-half), std::sin(-half));
        const cd p1(std::cos(half), std::sin(half));
        cd* a = amp_;
        single_target_loop(dim_ / 2, [a, t, p0, p1](size_t k) {
            const size_t i0 = insert_zero_bit(k, t), i1 = flip_target_bit(i0, t);
            a[i0] *= p0;
            a[i1] *= p1;
        });
    }
    void apply_cx(uint8_t c, uint8_t t) {
        cd* a = amp_;
        const size_t cm = size_t{1} << c, tm = size_t{1} << t;
        double_hole_loop(c, t, [a, cm, tm](size_t base) { std::swap(a[base | cm], a[base | cm | tm]); });
    }
    void apply_swap(uint8_t q0, uint8_t q1) {
        cd* a = amp_;
        const size_t m0 = size_t{1} << q0, m1 = size_t{1} << q1;
        double_hole_loop(q0, q1, [a, m0, m1](size_t base) { std::swap(a[base | m1], a[base | m0]); });
    }
    void apply_ccx(uint8_t c0, uint8_t c1, uint8_t t) {
        cd* a = amp_;
        const size_t cm = (size_t{1} << c0) | (size_t{1} << c1), tm = size_t{1} << t;
        triple_hole_loop(c0, c1, t, [a, cm, tm](size_t base) {
            const size_t i0 = base | cm;
            std::swap(a[i0], a[i0 | tm]);
        });
    }
    void apply_cswap(uint8_t c, uint8_t q0, uint8_t q1) {
        cd* a = amp_;
        const size_t cm = size_t{1} << c, m0 = size_t{1} << q0, m1 = size_t{1} << q1;
        triple_hole_loop(c, q0, q1, [a, cm, m0, m1](size_t base) { std::swap(a[base | cm | m1], a[base | cm | m0]); });
    }

    void apply_unitary(const Instr& g, const cd* U) {
        const int k = g.arity();
        if (k < 1 || k > 5) throw std::invalid_argument("unitary arity must be 1..5");
        const size_t D = size_t{1} << k;
        std::array<uint8_t, 5> sorted{};
        for (int i = 0; i < k; ++i) sorted[i] = (uint8_t)g.qubit(i);
        std::sort(sorted.begin(), sorted.begin() + k);
        std::array<size_t, 32> offs{};
        for (size_t j = 0; j < D; ++j) {
            size_t o = 0;
            for (int q = 0; q < k; ++q)
                if ((j >> q) & 1) o |= size_t{1} << g.qubit(q);
            offs[j] = o;
        }
        const size_t iters = dim_ >> k;
        cd* a = amp_;
#pragma omp parallel
        {
            std::array<cd, 32> vin, vout;
#pragma omp for schedule(static)
            for (int64_t i = 0; i < (int64_t)iters; ++i) {
                size_t base = (size_t)i;
                for (int q = 0; q < k; ++q) {
                    const size_t lm = (size_t{1} << sorted[q]) - 1;
                    base = (base & lm) | ((base & ~lm) << 1);
                }
                for (size_t j = 0; j < D; ++j) vin[j] = a[base | offs[j]];
                for (size_t r = 0; r < D; ++r) {
                    cd acc(0.0, 0.0);
                    for (size_t c = 0; c < D; ++c) acc += U[r * D + c] * vin[c];
                    vout[r] = acc;
                }
                for (size_t j = 0; j < D; ++j) a[base | offs[j]] = vout[j];
            }
        }
    }

    void apply_measure(uint8_t t) {
        const size_t half = dim_ / 2;
        cd* a = amp_;
        double p1 = 0.0;
#pragma omp parallel for reduction(+ : p1) schedule(static)
        for (int64_t i = 0; i < (int64_t)half; ++i) {
            const size_t i1 = insert_zero_bit((size_t)i, t) | (size_t{1} << t);
            p1 += std::norm(a[i1]);
        }
        const double r = dist_(rng_);
        const uint8_t outcome = (r < p1) ? 1 : 0;
        meas_.push_back(outcome);
        double p = outcome ? p1 : (1.0 - p1);
        if (p < 1e-12) p = 1e-12;
        const double inv = 1.0 / std::sqrt(p);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)half; ++i) {
            const size_t i0 = insert_zero_bit((size_t)i, t), i1 = i0 | (size_t{1} << t);
            if (outcome == 1) {
                a[i0] = {0.0, 0.0};
                a[i1] *= inv;
            } else {
                a[i0] *= inv;
                a[i1] = {0.0, 0.0};
            }
        }
    }
};

// ---- FusionMath (middleware.hpp:241-344), k <= 5 ----------------------------------
using Mat = std::vector<cd>;  // row-major, dim x dim

Mat expand_matrix(const Mat& U, const std::vector<uint16_t>& q_sub, const std::vector<uint16_t>& q_full) {
    const size_t dim_full = size_t{1} << q_full.size();
    const size_t dim_sub = size_t{1} << q_sub.size();
    std::vector<uint16_t> q_rest;
    for (uint16_t qf : q_full)
        if (std::find(q_sub.begin(), q_sub.end(), qf) == q_sub.end()) q_rest.push_back(qf);
    auto position = [&](uint16_t q) {
        for (size_t p = 0; p < q_full.size(); ++p)
            if (q_full[p] == q) return p;
        throw std::invalid_argument("qubit not in the full footprint");
    };
    std::vector<size_t> sub_to_full(dim_sub), rest_to_full(size_t{1} << q_rest.size());
    for (size_t r = 0; r < dim_sub; ++r) {
        size_t m = 0;
        for (size_t i = 0; i < q_sub.size(); ++i)
            if ((r >> i) & 1) m |= size_t{1} << position(q_sub[i]);
        sub_to_full[r] = m;
    }
    for (size_t r = 0; r < rest_to_full.size(); ++r) {
        size_t m = 0;
        for (size_t i = 0; i < q_rest.size(); ++i)
            if ((r >> i) & 1) m |= size_t{1} << position(q_rest[i]);
        rest_to_full[r] = m;
    }
    Mat M(dim_full * dim_full, cd(0.0, 0.0));
    for (size_t base : rest_to_full)
        for (size_t rs = 0; rs < dim_sub; ++rs)
            for (size_t cs = 0; cs < dim_sub; ++cs)
                M[(base | sub_to_full[rs]) * dim_full + (base | sub_to_full[cs])] = U[rs * dim_sub + cs];
    return M;
}

Mat fuse_matrices(const Mat& U1, const std::vector<uint16_t>& q1, const Mat& U2, const std::vector<uint16_t>& q2,
                  std::vector<uint16_t>& out_q) {
    out_q = q1;
    for (uint16_t q : q2)
        if (std::find(out_q.begin(), out_q.end(), q) == out_q.end()) out_q.push_back(q);
    std::sort(out_q.begin(), out_q.end());
    const Mat M1 = expand_matrix(U1, q1, out_q);
    const Mat M2 = expand_matrix(U2, q2, out_q);
    const size_t d = size_t{1} << out_q.size();
    Mat R(d * d, cd(0.0, 0.0));
    for (size_t i = 0; i < d; ++i)
        for (size_t j = 0; j < d; ++j) {
            cd acc(0.0, 0.0);
            for (size_t m = 0; m < d; ++m) acc += M2[i * d + m] * M1[m * d + j];
            R[i * d + j] = acc;
        }
    return R;  // U2 applied after U1: M2 * M1
}

static thread_local std::string g_err;

}  // namespace oracle

using namespace oracle;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

int orc_num_threads() { return omp_get_max_threads(); }

// |0...0> (cpu_state_vector.hpp:54-78)
int orc_init_zero(double* amps, uint32_t nq) {
    const size_t dim = size_t{1} << nq;
    cd* a = reinterpret_cast<cd*>(amps);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)dim; ++i) a[i] = cd(0.0, 0.0);
    a[0] = cd(1.0, 0.0);
    return 0;
}

// initialize(span) validation (cpu_state_vector.hpp:86-106): 0 ok, -1 not pow2, -2 not normalised.
int orc_check_state(const double* amps, size_t dim, double* norm_sq_out) {
    if (dim == 0 || (dim & (dim - 1)) != 0) {
        g_err = "CPUStateVector: Input state dimension must be a power of 2.";
        return -1;
    }
    const cd* a = reinterpret_cast<const cd*>(amps);
    double ns = 0.0;
#pragma omp parallel for reduction(+ : ns) schedule(static)
    for (int64_t i = 0; i < (int64_t)dim; ++i) ns += std::norm(a[i]);
    if (norm_sq_out) *norm_sq_out = ns;
    if (std::abs(ns - 1.0) > 1e-6) {
        g_err = "CPUStateVector: Provided state vector is not normalized (norm squared != 1.0).";
        return -2;
    }
    return 0;
}

// Runs a full program (including MEASURE) on amps in place.
// Returns 0, -1 invalid_argument, -2 out_of_range, -3 runtime_error.
int orc_execute(double* amps, uint32_t nq, const Instr* prog, size_t n, const double* pool, size_t pool_elems,
                const MatEntry* meta, size_t n_meta, int gswap_mode, uint64_t seed, uint8_t* meas_out,
                size_t meas_cap, size_t* n_meas) {
    (void)pool_elems;
    try {
        Engine e(reinterpret_cast<cd*>(amps), size_t{1} << nq, seed, gswap_mode);
        e.execute(prog, n, reinterpret_cast<const cd*>(pool), meta, n_meta);
        auto& m = e.measurements();
        if (n_meas) *n_meas = m.size();
        if (meas_out)
            for (size_t i = 0; i < m.size() && i < meas_cap; ++i) meas_out[i] = m[i];
        return 0;
    } catch (const std::invalid_argument& ex) {
        g_err = ex.what();
        return -1;
    } catch (const std::out_of_range& ex) {
        g_err = ex.what();
        return -2;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -3;
    }
}

// Seeded random input state for full-size parity runs (test input; the reference has no such
// routine). Amplitude i is (u(2i), u(2i+1)) with u(k) a SplitMix64 hash of (seed, k) mapped to
// [-1, 1), scaled by 1/sqrt(sum |a_i|^2). The norm is summed over a fixed partition of 4096
// blocks, so the bytes do not depend on the thread count.
int orc_random_state(double* amps, uint32_t nq, uint64_t seed) {
    const int64_t dim = int64_t{1} << nq;
    auto mix = [](uint64_t z) {
        z += 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const uint64_t s = mix(seed ^ 0x5DEECE66Dull);
    const int64_t nblk = dim < 4096 ? dim : 4096, per = dim / nblk;
    std::vector<double> part(nblk, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nblk; ++b) {
        double acc = 0.0;
        for (int64_t i = b * per; i < (b + 1) * per; ++i) {
            const double re = (double)(mix(s + 2 * (uint64_t)i) >> 11) * 0x1.0p-52 - 1.0;
            const double im = (double)(mix(s + 2 * (uint64_t)i + 1) >> 11) * 0x1.0p-52 - 1.0;
            amps[2 * i] = re;
            amps[2 * i + 1] = im;
            acc += re * re + im * im;
        }
        part[b] = acc;
    }
    double ns = 0.0;
    for (double p : part) ns += p;
    const double k = 1.0 / std::sqrt(ns);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 2 * dim; ++i) amps[i] *= k;
    return 0;
}

// Wall-clock seconds of orc_execute over `reps` runs from |0...0> (init outside the timer),
// as bench_hw_awareness.cpp:102-107 times execute() only. Returns the median.
double orc_time_execute(double* amps, uint32_t nq, const Instr* prog, size_t n, const double* pool,
                        size_t pool_elems, const MatEntry* meta, size_t n_meta, int reps) {
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        orc_init_zero(amps, nq);
        const auto t0 = std::chrono::steady_clock::now();
        if (orc_execute(amps, nq, prog, n, pool, pool_elems, meta, n_meta, 0, 0, nullptr, 0, nullptr) != 0) return -1.0;
        const auto t1 = std::chrono::steady_clock::now();
        t.push_back(std::chrono::duration<double>(t1 - t0).count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

// The seeded layered h/rx/cx generator, written independently of the product's:
// the instruction bytes must match QuantumCircuit::h/rx/cx pushes (frontend.hpp:251-302, 419-460).
int orc_gen_layered(uint32_t nqb, uint32_t depth, uint64_t seed, Instr* out, size_t cap, size_t* n_out) {
    std::vector<Instr> v;
    auto push = [&](uint8_t type, int arity, int q0, int q1, float param) {
        Instr g;
        std::memset(&g, 0, sizeof g);
        g.type = type;
        g.flags_arity = (uint8_t)arity;
        g.qubits[0] = g.qubits[1] = g.qubits[2] = g.qubits[3] = 255;
        g.qubits[0] = (uint8_t)q0;
        if (q1 >= 0) g.qubits[1] = (uint8_t)q1;
        g.payload.param = param;
        g.uid = (uint32_t)v.size();
        v.push_back(g);
    };
    std::mt19937_64 rng(seed);
    const double TWO_PI = 2.0 * 3.14159265358979323846;
    for (uint32_t q = 0; q < nqb; ++q) push(H, 1, (int)q, -1, 0.0f);
    for (uint32_t l = 0; l < depth; ++l) {
        for (uint32_t q = 0; q < nqb; ++q) {
            const uint64_t r = rng();
            const double theta = std::ldexp(static_cast<double>(r >> 11), -53) * TWO_PI;
            push(RX, 1, (int)q, -1, static_cast<float>(theta));
        }
        for (uint32_t q = l & 1u; q + 1 < nqb; q += 2) push(CX, 2, (int)q, (int)q + 1, 0.0f);
    }
    if (n_out) *n_out = v.size();
    if (out) {
        if (cap < v.size()) return -1;
        std::memcpy(out, v.data(), v.size() * sizeof(Instr));
    }
    return 0;
}

// FusionMath::fuse_matrices (middleware.hpp:315-344). Matrices interleaved complex128.
int orc_fuse_matrices(const double* U1, const uint16_t* q1, int n1, const double* U2, const uint16_t* q2, int n2,
                      double* out, uint16_t* out_q, int* out_n) {
    try {
        const size_t d1 = size_t{1} << n1, d2 = size_t{1} << n2;
        Mat A(reinterpret_cast<const cd*>(U1), reinterpret_cast<const cd*>(U1) + d1 * d1);
        Mat B(reinterpret_cast<const cd*>(U2), reinterpret_cast<const cd*>(U2) + d2 * d2);
        std::vector<uint16_t> qa(q1, q1 + n1), qb(q2, q2 + n2), qo;
        Mat R = fuse_matrices(A, qa, B, qb, qo);
        for (size_t i = 0; i < qo.size(); ++i) out_q[i] = qo[i];
        *out_n = (int)qo.size();
        std::memcpy(out, R.data(), R.size() * sizeof(cd));
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

}  // extern "C"
