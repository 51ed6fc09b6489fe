// Host-side input generation of the reference (Initialisation.cpp:7-76): the FCC lattice and Maxwell-Boltzmann
// velocities, with the same std:: calls in the same order, so identical seeds give identical inputs.
#include <cmath>
#include <random>
#include <vector>

#include "internal.hpp"

using namespace mdgpu;

extern "C" {

int md_fcc_lattice(int cells, double a, double* pos) {
    MD_API_BEGIN
    if (cells < 0 || (!pos && cells > 0)) fail(MD_ERR_ARG, "bad lattice arguments");
    // basis (0,0,0), (.5,.5,0), (.5,0,.5), (0,.5,.5) shifted by a/4; cells in ix, iy, iz order (Initialisation.cpp:9-40)
    static const double basis[4][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
    size_t p = 0;
    for (int ix = 0; ix < cells; ++ix)
        for (int iy = 0; iy < cells; ++iy)
            for (int iz = 0; iz < cells; ++iz)
                for (const auto& b : basis) {
                    pos[3 * p + 0] = (ix + b[0] + 0.25) * a;
                    pos[3 * p + 1] = (iy + b[1] + 0.25) * a;
                    pos[3 * p + 2] = (iz + b[2] + 0.25) * a;
                    ++p;
                }
    MD_API_END
}

int md_maxwell_boltzmann(size_t n, const double* mass, double temperature, uint32_t seed, double* vel) {
    MD_API_BEGIN
    if (!vel && n) fail(MD_ERR_ARG, "vel is NULL");
    auto m = [&](size_t i) { return mass ? mass[i] : 1.0; };
    std::mt19937 rng(seed);
    double momentum[3] = {0.0, 0.0, 0.0};
    double total_mass = 0.0;
    for (size_t i = 0; i < n; ++i) {
        std::normal_distribution<double> gauss(0.0, std::sqrt(temperature / m(i)));
        for (int k = 0; k < 3; ++k) {
            const double v = gauss(rng);
            vel[k * n + i] = v;
            momentum[k] += m(i) * v;
        }
        total_mass += m(i);
    }
    for (int k = 0; k < 3; ++k)
        for (size_t i = 0; i < n; ++i) vel[k * n + i] -= momentum[k] / total_mass;
    // rescaleVelocities: T from the kinetic energy in 1024-term blocks (Parallel.hpp:32-57, Thermostat.cpp:17-30)
    double ke = 0.0;
    for (size_t b = 0; b < n; b += 1024) {
        double part = 0.0;
        const size_t e = b + 1024 < n ? b + 1024 : n;
        for (size_t i = b; i < e; ++i) {
            const double vx = vel[i], vy = vel[n + i], vz = vel[2 * n + i];
            const double half_m = 0.5 * m(i);
            part += half_m * (vx * vx + vy * vy + vz * vz);
        }
        ke += part;
    }
    const double dof = 3.0 * static_cast<double>(n) - 3.0;
    const double current = 2.0 * ke / dof;
    if (current > 0.0) {
        const double s = std::sqrt(temperature / current);
        for (size_t i = 0; i < 3 * n; ++i) vel[i] *= s;
    }
    MD_API_END
}

}  // extern "C"
