#include "System.hpp"

#include <cstring>
#include <utility>

#include "MdgpuRuntime.hpp"

using mdgpu_facade::check;

System::System(const std::vector<Particle>& particles, double boxLength)
    : boxLength_(boxLength), inverseBoxLength_(1.0 / boxLength) {
    const std::size_t n = particles.size();
    positions_.assign(positionStride * n, 0.0);
    velocities_.assign(3 * n, 0.0);
    forces_.assign(3 * n, 0.0);
    mass_.resize(n);
    id_.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            positions_[positionStride * i + k] = particles[i].position[k];
            velocities_[k * n + i] = particles[i].velocity[k];
            forces_[k * n + i] = particles[i].force[k];
        }
        mass_[i] = particles[i].mass;
        id_[i] = i;
    }
}

System::System(const System& other)
    : boxLength_(other.boxLength_), inverseBoxLength_(other.inverseBoxLength_) {
    other.hostRead(All);
    positions_ = other.positions_;
    velocities_ = other.velocities_;
    forces_ = other.forces_;
    mass_ = other.mass_;
    id_ = other.id_;
}

System& System::operator=(const System& other) {
    if (this != &other) {
        System copy(other);
        *this = std::move(copy);
    }
    return *this;
}

System::System(System&& other) noexcept
    : boxLength_(other.boxLength_), inverseBoxLength_(other.inverseBoxLength_),
      positions_(std::move(other.positions_)), velocities_(std::move(other.velocities_)),
      forces_(std::move(other.forces_)), mass_(std::move(other.mass_)), id_(std::move(other.id_)),
      device_(other.device_), hostValid_(other.hostValid_), deviceValid_(other.deviceValid_) {
    other.device_ = nullptr;
    other.hostValid_ = All;
    other.deviceValid_ = 0;
}

System& System::operator=(System&& other) noexcept {
    if (this != &other) {
        release();
        boxLength_ = other.boxLength_;
        inverseBoxLength_ = other.inverseBoxLength_;
        positions_ = std::move(other.positions_);
        velocities_ = std::move(other.velocities_);
        forces_ = std::move(other.forces_);
        mass_ = std::move(other.mass_);
        id_ = std::move(other.id_);
        device_ = other.device_;
        hostValid_ = other.hostValid_;
        deviceValid_ = other.deviceValid_;
        other.device_ = nullptr;
        other.hostValid_ = All;
        other.deviceValid_ = 0;
    }
    return *this;
}

System::~System() {
    release();
}

void System::release() {
    if (device_) md_state_destroy(device_);
    device_ = nullptr;
}

namespace {

struct FieldMap {
    unsigned flag;
    int field;
};
constexpr FieldMap kFields[] = {{System::Positions, MD_FIELD_POS},
                                {System::Velocities, MD_FIELD_VEL},
                                {System::Forces, MD_FIELD_FORCE},
                                {System::Masses, MD_FIELD_MASS},
                                {System::Ids, MD_FIELD_ID}};

}  // namespace

void System::hostRead(unsigned fields) const {
    const unsigned missing = fields & ~hostValid_;
    if (!missing) return;
    for (const auto& f : kFields) {
        if (!(missing & f.flag)) continue;
        void* dst = nullptr;
        switch (f.field) {
            case MD_FIELD_POS: dst = positions_.data(); break;
            case MD_FIELD_VEL: dst = velocities_.data(); break;
            case MD_FIELD_FORCE: dst = forces_.data(); break;
            case MD_FIELD_MASS: dst = mass_.data(); break;
            default: dst = id_.data(); break;
        }
        if (size()) check(md_state_download(device_, f.field, dst));
        hostValid_ |= f.flag;
    }
}

void System::hostWrite(unsigned fields) {
    hostRead(fields);
    deviceValid_ &= ~fields;
}

md_state* System::deviceState() const {
    if (!device_) {
        check(md_state_create(mdgpu_facade::context(), size(), boxLength_, &device_));
        deviceValid_ = 0;
    }
    const unsigned missing = All & ~deviceValid_;
    for (const auto& f : kFields) {
        if (!(missing & f.flag)) continue;
        const void* src = nullptr;
        switch (f.field) {
            case MD_FIELD_POS: src = positions_.data(); break;
            case MD_FIELD_VEL: src = velocities_.data(); break;
            case MD_FIELD_FORCE: src = forces_.data(); break;
            case MD_FIELD_MASS: src = mass_.data(); break;
            default: src = id_.data(); break;
        }
        if (size()) check(md_state_upload(device_, f.field, src));
        deviceValid_ |= f.flag;
    }
    return device_;
}

void System::deviceWrote(unsigned fields) const {
    hostValid_ &= ~fields;
    deviceValid_ |= fields;
}

double* System::packedPositions() {
    hostWrite(Positions);
    return positions_.data();
}

const double* System::packedPositions() const {
    hostRead(Positions);
    return positions_.data();
}

std::span<double> System::velocity(int k) {
    hostWrite(Velocities);
    return {velocities_.data() + k * size(), size()};
}

std::span<const double> System::velocity(int k) const {
    hostRead(Velocities);
    return {velocities_.data() + k * size(), size()};
}

std::span<double> System::force(int k) {
    hostWrite(Forces);
    return {forces_.data() + k * size(), size()};
}

std::span<const double> System::force(int k) const {
    hostRead(Forces);
    return {forces_.data() + k * size(), size()};
}

std::span<const double> System::mass() const {
    hostRead(Masses);
    return mass_;
}

std::span<const std::size_t> System::ids() const {
    hostRead(Ids);
    return id_;
}

std::array<double, 3> System::positionOf(std::size_t i) const {
    hostRead(Positions);
    const double* r = positions_.data() + positionStride * i;
    return {r[0], r[1], r[2]};
}

std::array<double, 3> System::velocityOf(std::size_t i) const {
    hostRead(Velocities);
    const std::size_t n = size();
    return {velocities_[i], velocities_[n + i], velocities_[2 * n + i]};
}

std::array<double, 3> System::forceOf(std::size_t i) const {
    hostRead(Forces);
    const std::size_t n = size();
    return {forces_[i], forces_[n + i], forces_[2 * n + i]};
}

void System::setPosition(std::size_t i, const std::array<double, 3>& r) {
    hostWrite(Positions);
    for (int k = 0; k < 3; ++k) positions_[positionStride * i + k] = r[k];
}

void System::setVelocity(std::size_t i, const std::array<double, 3>& v) {
    hostWrite(Velocities);
    const std::size_t n = size();
    for (int k = 0; k < 3; ++k) velocities_[k * n + i] = v[k];
}

Particle System::particle(std::size_t i) const {
    Particle p;
    p.position = positionOf(i);
    p.velocity = velocityOf(i);
    p.force = forceOf(i);
    p.mass = mass()[i];
    p.id = ids()[i];
    return p;
}

void System::reorder(std::span<const std::size_t> order) {
    md_state* s = deviceState();
    static_assert(sizeof(std::size_t) == sizeof(std::uint64_t));
    check(md_state_reorder(s, reinterpret_cast<const std::uint64_t*>(order.data())));
    deviceWrote(All);
}

double System::kineticEnergy() const {
    double ke = 0.0;
    check(md_kinetic_energy(deviceState(), &ke));
    return ke;
}

std::uint64_t System::stateHash() const {
    std::uint64_t h = 0;
    check(md_state_hash(deviceState(), &h));
    return h;
}

double System::temperature() const {
    const double dof = 3.0 * static_cast<double>(size()) - 3.0;
    return 2.0 * kineticEnergy() / dof;
}
