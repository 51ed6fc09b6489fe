#include "CellList.hpp"

#include "MdgpuRuntime.hpp"

CellList::CellList(double minCellSize) : minCellSize_(minCellSize) {
}

void CellList::build(const System& system) {
    md_state* s = system.deviceState();
    int side = 0;
    mdgpu_facade::check(md_cells_build(s, minCellSize_, &side, nullptr, nullptr, nullptr));
    const std::size_t n = system.size();
    const std::size_t cells = static_cast<std::size_t>(side) * side * side;
    particleCell_.resize(n);
    cellStart_.resize(cells + 1);
    cellParticles_.resize(n);
    static_assert(sizeof(std::size_t) == sizeof(std::uint64_t));
    mdgpu_facade::check(md_cells_build(s, minCellSize_, &side, particleCell_.data(),
                                       reinterpret_cast<std::uint64_t*>(cellStart_.data()),
                                       reinterpret_cast<std::uint64_t*>(cellParticles_.data())));
    cellsPerSide_ = side;
}

std::array<int, 27> CellList::neighbourCells(int cell) const {
    const int n = cellsPerSide_;
    const int cx = cell / (n * n), cy = (cell / n) % n, cz = cell % n;
    std::array<int, 27> out {};
    int k = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
                out[k++] = (((cx + dx + n) % n) * n + (cy + dy + n) % n) * n + (cz + dz + n) % n;
    return out;
}
