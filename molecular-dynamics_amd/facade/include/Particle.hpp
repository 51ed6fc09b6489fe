#pragma once

#include <array>
#include <cstddef>

// One particle as a plain record, used to construct a System and to read single particles back
// (reference: include/Particle.hpp:8-14). `id` is the stable identity that survives re-sorting.
struct Particle {
    std::array<double, 3> position {0.0, 0.0, 0.0};
    std::array<double, 3> velocity {0.0, 0.0, 0.0};
    std::array<double, 3> force {0.0, 0.0, 0.0};
    double mass {1.0};
    std::size_t id {0};
};
