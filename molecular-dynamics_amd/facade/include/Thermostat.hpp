#pragma once

#include "System.hpp"

#include <cstdint>
#include <random>

// Rescale the velocities to the target temperature (reference: src/Thermostat.cpp:17-30)
void rescaleVelocities(System& system, double targetTemperature);

// Bussi stochastic velocity rescaling (Bussi, Donadio & Parrinello, J. Chem. Phys. 126, 014101 (2007), eq. A7)
// (reference: include/Thermostat.hpp, src/Thermostat.cpp:32-58). The kinetic energy is summed on the GPU in the
// reference's order, the two draws come from the same std::mt19937_64 stream and distributions, and the scale is
// applied on the GPU.
class BussiThermostat {
public:
    BussiThermostat(double temperature, double tau, std::uint64_t seed);

    // Call once per timestep, after the integrator step
    void apply(System& system, double dt);

    void setTemperature(double temperature) {
        temperature_ = temperature;
    }

    // ---- mdgpu extension: the next `count` (r1, sum_sq) draw pairs for a System of n particles, advancing the
    // stream exactly as `count` apply() calls with a positive kinetic energy would (for md_run) ----
    void drawPairs(std::size_t n, int count, double* out);

private:
    double temperature_;
    double tau_;
    std::mt19937_64 rng_;
};
