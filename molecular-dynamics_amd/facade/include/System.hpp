#pragma once

#include "Particle.hpp"

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <span>
#include <vector>

struct md_state;

// A periodic cubic box of particles whose state lives in GPU memory (reference: include/System.hpp:12-136).
// The public interface is the reference's. Host copies of each array are kept coherent lazily: a non-const
// accessor (packedPositions(), velocity(k), force(k), setPosition, setVelocity) makes the host copy current and
// marks the device copy stale; the next GPU operation uploads it. A span returned by an accessor is valid until
// the next GPU operation on this System (integrator step, force computation, list build, thermostat); take a
// fresh one afterwards, as the reference's own callers do.
class System {
public:
    // Ids are assigned in input order (0, 1, 2, ...); ids in the input records are ignored (System.cpp:9-29).
    System(const std::vector<Particle>& particles, double boxLength);
    System(const System& other);
    System& operator=(const System& other);
    System(System&& other) noexcept;
    System& operator=(System&& other) noexcept;
    ~System();

    std::size_t size() const {
        return mass_.size();
    }

    static constexpr std::size_t positionStride = 4;
    double* packedPositions();
    const double* packedPositions() const;

    std::span<double> velocity(int k);
    std::span<const double> velocity(int k) const;
    std::span<double> force(int k);
    std::span<const double> force(int k) const;
    std::span<const double> mass() const;
    std::span<const std::size_t> ids() const;

    std::array<double, 3> positionOf(std::size_t i) const;
    std::array<double, 3> velocityOf(std::size_t i) const;
    std::array<double, 3> forceOf(std::size_t i) const;
    void setPosition(std::size_t i, const std::array<double, 3>& r);
    void setVelocity(std::size_t i, const std::array<double, 3>& v);
    Particle particle(std::size_t i) const;

    double getBoxLength() const {
        return boxLength_;
    }
    double getInverseBoxLength() const {
        return inverseBoxLength_;
    }

    // d = ri - rj; d -= L * round(d / L) per component (System.hpp:86-97)
    std::array<double, 3> minimumImageDistance(const std::array<double, 3>& ri, const std::array<double, 3>& rj) const {
        std::array<double, 3> dr {};
        for (int k = 0; k < 3; ++k) {
            const double d = ri[k] - rj[k];
            const double images = std::round(d * inverseBoxLength_);
            dr[k] = d - boxLength_ * images;
        }
        return dr;
    }
    std::array<double, 3> minimumImageDistance(std::size_t i, std::size_t j) const {
        return minimumImageDistance(positionOf(i), positionOf(j));
    }
    void wrapPosition(std::array<double, 3>& r) const {
        for (int k = 0; k < 3; ++k) r[k] = wrapCoordinate(r[k]);
    }
    double wrapCoordinate(double x) const {
        const double images = std::floor(x * inverseBoxLength_);
        return x - boxLength_ * images;
    }

    // new particle k is old particle order[k] (System.cpp:42-64), done on the GPU
    void reorder(std::span<const std::size_t> order);
    // reference summation order (1024-term blocks), computed on the GPU (System.cpp:67-76)
    double kineticEnergy() const;
    // FNV-1a over x, vx, y, vy, z, vz of each particle (System.cpp:85-105)
    std::uint64_t stateHash() const;
    double temperature() const;

    // ---- mdgpu extensions -----------------------------------------------------------------------------
    // The device state with every stale field uploaded; a GPU operation that writes fields must then call
    // deviceWrote(fields).
    md_state* deviceState() const;
    enum Field : unsigned { Positions = 1u, Velocities = 2u, Forces = 4u, Masses = 8u, Ids = 16u, All = 31u };
    void deviceWrote(unsigned fields) const;

private:
    void hostRead(unsigned fields) const;
    void hostWrite(unsigned fields);
    void release();

    double boxLength_;
    double inverseBoxLength_;
    mutable std::vector<double> positions_;
    mutable std::vector<double> velocities_;  // 3n, component-major
    mutable std::vector<double> forces_;      // 3n, component-major
    mutable std::vector<double> mass_;
    mutable std::vector<std::size_t> id_;
    mutable md_state* device_ = nullptr;
    mutable unsigned hostValid_ = All;
    mutable unsigned deviceValid_ = 0;
};
