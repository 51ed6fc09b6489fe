#include "Initialisation.hpp"

#include "MdgpuRuntime.hpp"

std::vector<Particle> createFCCLattice(int cellsPerSide, double latticeConstant) {
    const std::size_t n = 4 * static_cast<std::size_t>(cellsPerSide) * cellsPerSide * cellsPerSide;
    std::vector<double> xyz(3 * n);
    mdgpu_facade::check(md_fcc_lattice(cellsPerSide, latticeConstant, xyz.data()));
    std::vector<Particle> particles(n);
    for (std::size_t i = 0; i < n; ++i) particles[i].position = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    return particles;
}

void initializeVelocities(System& system, double temperature, std::uint32_t seed) {
    const std::size_t n = system.size();
    const auto mass = system.mass();
    std::vector<double> vel(3 * n);
    mdgpu_facade::check(md_maxwell_boltzmann(n, mass.data(), temperature, seed, vel.data()));
    for (int k = 0; k < 3; ++k) {
        auto v = system.velocity(k);
        for (std::size_t i = 0; i < n; ++i) v[i] = vel[k * n + i];
    }
}
