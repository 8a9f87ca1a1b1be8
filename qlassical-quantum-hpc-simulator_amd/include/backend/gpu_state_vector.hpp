// gpu_state_vector.hpp — GPUStateVector: an HBM-resident StateVector on one MI355X.
//
// The GPU twin of CPUStateVector (reference include/backend/cpu_state_vector.hpp:54-166):
//   initialize(n)     -> qlx_state_create / qlx_state_reset   (|0...0>, on the device)
//   initialize(span)  -> qlx_state_create_from_host / qlx_state_load
//                        (power-of-two and 1e-6 normalisation checks, same messages)
// The amplitudes never leave HBM unless amplitudes() is called: that copies them into a
// host mirror (complex128; a complex64 state is widened) which stays valid until the
// state is modified again. Move-only; moving transfers the device allocation.
#ifndef QLASSICAL_BACKEND_GPU_STATE_VECTOR_HPP
#define QLASSICAL_BACKEND_GPU_STATE_VECTOR_HPP

#include <complex>
#include <cstdint>
#include <new>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <backend/state_vector.hpp>

#include "qlx.h"

namespace qlassical::backend {

// Maps a qlx status code to the exception type the reference uses for the same failure.
inline void qlx_throw_on_error(int rc) {
    if (rc == QLX_OK) return;
    const std::string msg = qlx_last_error();
    switch (rc) {
        case QLX_EINVAL: throw std::invalid_argument(msg);
        case QLX_ERANGE: throw std::out_of_range(msg);
        case QLX_ENOMEM: throw std::bad_alloc();
        default: throw std::runtime_error(msg);
    }
}

enum class Precision : int { C128 = QLX_C128, C64 = QLX_C64 };

class GPUStateVector final : public StateVector {
public:
    using Amplitude = std::complex<double>;

    explicit GPUStateVector(Precision p = Precision::C128) : precision_(p) {}
    ~GPUStateVector() override { release(); }

    GPUStateVector(GPUStateVector&& o) noexcept
        : st_(std::exchange(o.st_, nullptr)), precision_(o.precision_), mirror_(std::move(o.mirror_)),
          mirror_valid_(std::exchange(o.mirror_valid_, false)) {}
    GPUStateVector& operator=(GPUStateVector&& o) noexcept {
        if (this != &o) {
            release();
            st_ = std::exchange(o.st_, nullptr);
            precision_ = o.precision_;
            mirror_ = std::move(o.mirror_);
            mirror_valid_ = std::exchange(o.mirror_valid_, false);
        }
        return *this;
    }
    GPUStateVector(const GPUStateVector&) = delete;
    GPUStateVector& operator=(const GPUStateVector&) = delete;

    void initialize(uint32_t num_qubits) override {
        if (st_ && num_qubits_cached() == num_qubits) {
            qlx_throw_on_error(qlx_state_reset(st_));
        } else {
            release();
            qlx_throw_on_error(qlx_state_create(num_qubits, static_cast<int>(precision_), &st_));
        }
        invalidate();
    }

    void initialize(std::span<const Amplitude> initial) override {
        const auto* p = reinterpret_cast<const double*>(initial.data());
        if (st_) {
            qlx_throw_on_error(qlx_state_load(st_, p, initial.size()));
        } else {
            qlx_throw_on_error(qlx_state_create_from_host(p, initial.size(), static_cast<int>(precision_), &st_));
        }
        invalidate();
    }

    [[nodiscard]] uint32_t num_qubits() const noexcept override { return st_ ? num_qubits_cached() : 0; }
    [[nodiscard]] std::size_t dimension() const noexcept override {
        return st_ ? (std::size_t{1} << num_qubits_cached()) : 0;
    }
    [[nodiscard]] bool empty() const noexcept override { return st_ == nullptr; }

    // Host view of the amplitudes (synchronises and copies on first use after a change).
    [[nodiscard]] std::span<const Amplitude> amplitudes() const {
        if (!st_) return {};
        if (!mirror_valid_) {
            mirror_.resize(dimension());
            qlx_throw_on_error(qlx_state_to_host(st_, reinterpret_cast<double*>(mirror_.data()), mirror_.size()));
            mirror_valid_ = true;
        }
        return mirror_;
    }

    [[nodiscard]] double norm_sq() const {
        double v = 0.0;
        qlx_throw_on_error(qlx_state_norm_sq(st_, &v));
        return v;
    }

    [[nodiscard]] qlx_state* handle() const noexcept { return st_; }
    [[nodiscard]] Precision precision() const noexcept { return precision_; }
    // Called by the engine after it modified the device state.
    void invalidate() noexcept { mirror_valid_ = false; }

private:
    qlx_state* st_ = nullptr;
    Precision precision_ = Precision::C128;
    mutable std::vector<Amplitude> mirror_;
    mutable bool mirror_valid_ = false;

    uint32_t num_qubits_cached() const noexcept {
        uint32_t n = 0;
        qlx_state_num_qubits(st_, &n);
        return n;
    }
    void release() noexcept {
        if (st_) qlx_state_destroy(st_);
        st_ = nullptr;
        mirror_.clear();
        mirror_.shrink_to_fit();
        mirror_valid_ = false;
    }
};

}  // namespace qlassical::backend

#endif  // QLASSICAL_BACKEND_GPU_STATE_VECTOR_HPP
