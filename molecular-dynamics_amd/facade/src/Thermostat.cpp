#include "Thermostat.hpp"

#include <cmath>

#include "MdgpuRuntime.hpp"

using mdgpu_facade::check;

namespace {

void scaleOnDevice(System& system, double scale) {
    check(md_scale_velocities(system.deviceState(), scale));
    system.deviceWrote(System::Velocities);
}

}  // namespace

void rescaleVelocities(System& system, double targetTemperature) {
    const double current = system.temperature();
    if (current <= 0.0) return;
    scaleOnDevice(system, std::sqrt(targetTemperature / current));
}

BussiThermostat::BussiThermostat(double temperature, double tau, std::uint64_t seed)
    : temperature_(temperature), tau_(tau), rng_(seed) {
}

void BussiThermostat::apply(System& system, double dt) {
    const double kinetic = system.kineticEnergy();
    if (kinetic <= 0.0) return;
    const double dof = 3.0 * static_cast<double>(system.size()) - 3.0;
    const double target = 0.5 * dof * temperature_;
    const double c = std::exp(-dt / tau_);
    // one Gaussian plus the chi-squared sum of the other dof - 1, fresh distributions per call
    std::normal_distribution<double> gaussian(0.0, 1.0);
    std::gamma_distribution<double> gamma(0.5 * (dof - 1.0), 2.0);
    const double r1 = gaussian(rng_);
    const double sumSquares = gamma(rng_);
    const double relax = (1.0 - c) * (target * (sumSquares + r1 * r1) / dof - kinetic);
    const double noise = 2.0 * r1 * std::sqrt(target / dof * kinetic * (1.0 - c) * c);
    const double newKinetic = kinetic + relax + noise;
    scaleOnDevice(system, std::sqrt(newKinetic / kinetic));
}

void BussiThermostat::drawPairs(std::size_t n, int count, double* out) {
    const double dof = 3.0 * static_cast<double>(n) - 3.0;
    for (int k = 0; k < count; ++k) {
        std::normal_distribution<double> gaussian(0.0, 1.0);
        std::gamma_distribution<double> gamma(0.5 * (dof - 1.0), 2.0);
        out[2 * k] = gaussian(rng_);
        out[2 * k + 1] = gamma(rng_);
    }
}
