#pragma once

#include "System.hpp"

#include <cstddef>
#include <span>
#include <vector>

struct md_list;

// Verlet neighbour list (reference: include/NeighbourList.hpp and src/NeighbourList.cpp). On the GPU it is a
// padded FULL list (both directions of every pair) built by one kernel after re-sorting the System by cell; the
// reference's half-list views (neighboursOf, firstPairOf, pairsAsSecond) are materialised on the host on first
// use after each build. Each full-list row starts with exactly the reference's half-list row.
class NeighbourList {
public:
    NeighbourList(double cutOff, double skin, bool reverseIndex = false);
    // A copy is deep: parameters, counters and the device rows (md_list_copy), usable for forces at once.
    NeighbourList(const NeighbourList& other);
    NeighbourList& operator=(const NeighbourList& other);
    NeighbourList(NeighbourList&& other) noexcept;
    NeighbourList& operator=(NeighbourList&& other) noexcept;
    ~NeighbourList();

    // Rebuild only if a particle moved more than skin / 2 since the last build; true if rebuilt.
    bool update(System& system);
    // Unconditional rebuild; re-sorts the particles by cell, so indices change.
    void build(System& system);

    std::span<const std::size_t> neighboursOf(std::size_t i) const {
        materialise();
        return {neighbours_.data() + offsets_[i], offsets_[i + 1] - offsets_[i]};
    }
    std::size_t firstPairOf(std::size_t i) const {
        materialise();
        return offsets_[i];
    }
    bool hasReverseIndex() const {
        return reverseIndex_;
    }
    std::span<const std::size_t> pairsAsSecond(std::size_t j) const;
    int getRebuildCount() const;
    std::size_t getPairCount() const;

    // ---- mdgpu extension: the device list (nullptr before the first build) ------------------------------
    md_list* deviceList() const {
        return device_;
    }
    // after a device-driven run (md_run) rebuilt the device list: the host views are materialised again
    void deviceUpdated() {
        hostCurrent_ = false;
    }

private:
    void materialise() const;
    void ensureDevice();

    double cutOff_;
    double skin_;
    bool reverseIndex_;
    md_list* device_ = nullptr;
    int rebuildsBefore_ = 0;  // builds counted by earlier device lists (copies)
    mutable bool hostCurrent_ = false;
    mutable std::vector<std::size_t> offsets_;
    mutable std::vector<std::size_t> neighbours_;
    mutable std::vector<std::size_t> reverseOffsets_;
    mutable std::vector<std::size_t> reversePairs_;
};
