#include "Trajectory.hpp"

#include <vector>

void writeXyzFrame(std::FILE* file, const System& system, double time) {
    const std::size_t n = system.size();
    const double l = system.getBoxLength();
    // storage index of each id: the inverse of the cell re-sort permutation
    const auto ids = system.ids();
    std::vector<std::size_t> slot(n);
    for (std::size_t i = 0; i < n; ++i) slot[ids[i]] = i;
    std::fprintf(file, "%zu\n", n);
    std::fprintf(file,
                 "Lattice=\"%.10g 0 0 0 %.10g 0 0 0 %.10g\" Properties=species:S:1:pos:R:3:velo:R:3:id:I:1 "
                 "Time=%.6f pbc=\"T T T\"\n",
                 l, l, l, time);
    for (std::size_t id = 0; id < n; ++id) {
        const auto r = system.positionOf(slot[id]);
        const auto v = system.velocityOf(slot[id]);
        std::fprintf(file, "Ar %.8f %.8f %.8f %.8f %.8f %.8f %zu\n", r[0], r[1], r[2], v[0], v[1], v[2], id);
    }
}
