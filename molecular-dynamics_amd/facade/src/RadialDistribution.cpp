#include "RadialDistribution.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <numbers>

#include "MdgpuRuntime.hpp"

RadialDistribution::RadialDistribution(double rMax, int bins)
    : rMax_(rMax), binWidth_(rMax / bins), histogram_(static_cast<std::size_t>(bins), 0.0) {
}

void RadialDistribution::sample(const System& system, const NeighbourList& neighbourList) {
    std::vector<std::uint64_t> counts(histogram_.size(), 0);
    mdgpu_facade::check(md_rdf_sample(system.deviceState(), neighbourList.deviceList(), rMax_,
                                      static_cast<int>(histogram_.size()), counts.data()));
    for (std::size_t b = 0; b < counts.size(); ++b) histogram_[b] += static_cast<double>(counts[b]);
    particles_ = static_cast<double>(system.size());
    density_ = particles_ / std::pow(system.getBoxLength(), 3);
    ++samples_;
}

void RadialDistribution::addCounts(const std::uint64_t* counts, int samples, const System& system) {
    for (std::size_t b = 0; b < histogram_.size(); ++b) histogram_[b] += static_cast<double>(counts[b]);
    particles_ = static_cast<double>(system.size());
    density_ = particles_ / std::pow(system.getBoxLength(), 3);
    samples_ += samples;
}

std::vector<double> RadialDistribution::radii() const {
    std::vector<double> r(histogram_.size());
    for (std::size_t b = 0; b < r.size(); ++b) r[b] = (static_cast<double>(b) + 0.5) * binWidth_;
    return r;
}

// g = pairs in the shell / (samples * N/2 * density * shell volume) (RadialDistribution.cpp:42-54)
std::vector<double> RadialDistribution::values() const {
    std::vector<double> g(histogram_.size(), 0.0);
    if (samples_ == 0) return g;
    for (std::size_t b = 0; b < g.size(); ++b) {
        const double lo = static_cast<double>(b) * binWidth_;
        const double hi = lo + binWidth_;
        const double shell = 4.0 / 3.0 * std::numbers::pi * (hi * hi * hi - lo * lo * lo);
        g[b] = histogram_[b] / (samples_ * 0.5 * particles_ * density_ * shell);
    }
    return g;
}

bool RadialDistribution::write(const char* path) const {
    std::FILE* file = std::fopen(path, "w");
    if (!file) return false;
    const auto r = radii();
    const auto g = values();
    std::fprintf(file, "r,g\n");
    for (std::size_t b = 0; b < g.size(); ++b) std::fprintf(file, "%.6f,%.6f\n", r[b], g[b]);
    std::fclose(file);
    return true;
}
