#pragma once

#include "System.hpp"

#include <array>
#include <cstddef>
#include <span>
#include <vector>

// A grid of cubic cells with a stable counting sort of the particles into them (reference: include/CellList.hpp
// and src/CellList.cpp). build() runs on the GPU; its results are kept on the host for the accessors.
class CellList {
public:
    explicit CellList(double minCellSize);

    void build(const System& system);

    int getCellsPerSide() const {
        return cellsPerSide_;
    }
    int cellOf(std::size_t particle) const {
        return particleCell_[particle];
    }
    // The 3x3x3 block around `cell`: dx outermost, then dy, then dz, each -1..1, wrapped (CellList.cpp:55-83)
    std::array<int, 27> neighbourCells(int cell) const;
    std::span<const std::size_t> particlesIn(int cell) const {
        return {cellParticles_.data() + cellStart_[cell], cellStart_[cell + 1] - cellStart_[cell]};
    }
    std::span<const std::size_t> sortedOrder() const {
        return cellParticles_;
    }

private:
    double minCellSize_;
    int cellsPerSide_ {0};
    std::vector<std::size_t> cellStart_;
    std::vector<std::size_t> cellParticles_;
    std::vector<int> particleCell_;
};
