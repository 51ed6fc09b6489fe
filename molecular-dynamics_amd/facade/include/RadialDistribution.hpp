#pragma once

#include "NeighbourList.hpp"
#include "System.hpp"

#include <cstdint>
#include <vector>

// g(r) from the neighbour-list pairs (reference: include/RadialDistribution.hpp, src/RadialDistribution.cpp). The
// pair histogram of each sample is counted exactly on the GPU; normalisation and output are the reference's.
class RadialDistribution {
public:
    RadialDistribution(double rMax, int bins);

    void sample(const System& system, const NeighbourList& neighbourList);
    // mdgpu extension: adds `samples` samples' pair counts taken on the device (md_run's in-run histogram)
    void addCounts(const std::uint64_t* counts, int samples, const System& system);
    std::vector<double> radii() const;
    std::vector<double> values() const;
    bool write(const char* path) const;

private:
    double rMax_;
    double binWidth_;
    std::vector<double> histogram_;
    int samples_ {0};
    double particles_ {0.0};
    double density_ {0.0};
};
