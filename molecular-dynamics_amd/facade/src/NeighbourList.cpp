#include "NeighbourList.hpp"

#include <utility>

#include "MdgpuRuntime.hpp"

using mdgpu_facade::check;

NeighbourList::NeighbourList(double cutOff, double skin, bool reverseIndex)
    : cutOff_(cutOff), skin_(skin), reverseIndex_(reverseIndex) {
}

NeighbourList::NeighbourList(const NeighbourList& other)
    : cutOff_(other.cutOff_), skin_(other.skin_), reverseIndex_(other.reverseIndex_),
      rebuildsBefore_(other.rebuildsBefore_) {
    if (other.device_) {
        ensureDevice();
        try {
            check(md_list_copy(device_, other.device_));
        } catch (...) {
            md_list_destroy(device_);
            throw;
        }
    }
}

NeighbourList& NeighbourList::operator=(const NeighbourList& other) {
    if (this != &other) {
        NeighbourList copy(other);
        *this = std::move(copy);
    }
    return *this;
}

NeighbourList::NeighbourList(NeighbourList&& other) noexcept
    : cutOff_(other.cutOff_), skin_(other.skin_), reverseIndex_(other.reverseIndex_), device_(other.device_),
      rebuildsBefore_(other.rebuildsBefore_), hostCurrent_(other.hostCurrent_), offsets_(std::move(other.offsets_)),
      neighbours_(std::move(other.neighbours_)), reverseOffsets_(std::move(other.reverseOffsets_)),
      reversePairs_(std::move(other.reversePairs_)) {
    other.device_ = nullptr;
    other.hostCurrent_ = false;
}

NeighbourList& NeighbourList::operator=(NeighbourList&& other) noexcept {
    if (this != &other) {
        if (device_) md_list_destroy(device_);
        cutOff_ = other.cutOff_;
        skin_ = other.skin_;
        reverseIndex_ = other.reverseIndex_;
        device_ = other.device_;
        rebuildsBefore_ = other.rebuildsBefore_;
        hostCurrent_ = other.hostCurrent_;
        offsets_ = std::move(other.offsets_);
        neighbours_ = std::move(other.neighbours_);
        reverseOffsets_ = std::move(other.reverseOffsets_);
        reversePairs_ = std::move(other.reversePairs_);
        other.device_ = nullptr;
        other.hostCurrent_ = false;
    }
    return *this;
}

NeighbourList::~NeighbourList() {
    if (device_) md_list_destroy(device_);
}

void NeighbourList::ensureDevice() {
    if (!device_) check(md_list_create(mdgpu_facade::context(), cutOff_, skin_, 0, &device_));
}

void NeighbourList::build(System& system) {
    ensureDevice();
    md_state* s = system.deviceState();
    check(md_list_build(device_, s));
    system.deviceWrote(System::All);  // re-sorted
    hostCurrent_ = false;
}

bool NeighbourList::update(System& system) {
    ensureDevice();
    md_state* s = system.deviceState();
    int rebuilt = 0;
    check(md_list_update(device_, s, &rebuilt));
    if (rebuilt) {
        system.deviceWrote(System::All);
        hostCurrent_ = false;
    }
    return rebuilt != 0;
}

// Half list in CSR from the device rows, and (with reverseIndex) the reverse CSR of NeighbourList.cpp:114-127:
// for every particle j, the indices of the pairs (i, j) in increasing order.
void NeighbourList::materialise() const {
    if (hostCurrent_ || !device_) return;
    md_list_info info {};
    check(md_list_get_info(device_, &info));
    const std::size_t n = info.particles;
    static_assert(sizeof(std::size_t) == sizeof(std::uint64_t));
    offsets_.assign(n + 1, 0);
    neighbours_.resize(info.half_pairs);
    check(md_list_half(device_, reinterpret_cast<std::uint64_t*>(offsets_.data()),
                       reinterpret_cast<std::uint64_t*>(neighbours_.data())));
    if (reverseIndex_) {
        reverseOffsets_.assign(n + 1, 0);
        for (std::size_t j : neighbours_) ++reverseOffsets_[j + 1];
        for (std::size_t i = 0; i < n; ++i) reverseOffsets_[i + 1] += reverseOffsets_[i];
        reversePairs_.resize(neighbours_.size());
        std::vector<std::size_t> next(reverseOffsets_.begin(), reverseOffsets_.end() - 1);
        for (std::size_t k = 0; k < neighbours_.size(); ++k) reversePairs_[next[neighbours_[k]]++] = k;
    }
    hostCurrent_ = true;
}

std::span<const std::size_t> NeighbourList::pairsAsSecond(std::size_t j) const {
    materialise();
    return {reversePairs_.data() + reverseOffsets_[j], reverseOffsets_[j + 1] - reverseOffsets_[j]};
}

int NeighbourList::getRebuildCount() const {
    if (!device_) return rebuildsBefore_;
    md_list_info info {};
    check(md_list_get_info(device_, &info));
    return rebuildsBefore_ + info.rebuilds;
}

std::size_t NeighbourList::getPairCount() const {
    if (!device_) return 0;
    md_list_info info {};
    check(md_list_get_info(device_, &info));
    return info.half_pairs;
}
