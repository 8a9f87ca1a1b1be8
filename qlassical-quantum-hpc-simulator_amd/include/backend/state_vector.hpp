// state_vector.hpp — the abstract StateVector interface, restated
// (reference include/backend/state_vector.hpp:17-47; identical API).
#ifndef QLASSICAL_BACKEND_STATE_VECTOR_BASE_HPP
#define QLASSICAL_BACKEND_STATE_VECTOR_BASE_HPP

#include <complex>
#include <cstddef>
#include <cstdint>
#include <span>

namespace qlassical::backend {

class StateVector {
public:
    virtual ~StateVector() = default;
    virtual void initialize(uint32_t num_qubits) = 0;
    virtual void initialize(std::span<const std::complex<double>> initial_amplitudes) = 0;
    [[nodiscard]] virtual uint32_t num_qubits() const noexcept = 0;
    [[nodiscard]] virtual std::size_t dimension() const noexcept = 0;
    [[nodiscard]] virtual bool empty() const noexcept = 0;

protected:
    StateVector() = default;
    StateVector(const StateVector&) = delete;
    StateVector& operator=(const StateVector&) = delete;
    StateVector(StateVector&&) = default;
    StateVector& operator=(StateVector&&) = default;
};

}  // namespace qlassical::backend

#endif  // QLASSICAL_BACKEND_STATE_VECTOR_BASE_HPP
