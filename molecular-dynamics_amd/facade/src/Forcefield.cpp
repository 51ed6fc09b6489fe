#include "Forcefield.hpp"

#include <cmath>
#include <stdexcept>

#include "MdgpuRuntime.hpp"

using mdgpu_facade::check;

LennardJonesForceField::LennardJonesForceField(double epsilon, double sigma, double cut_off)
    : epsilon_(epsilon), sigma_(sigma), cut_off_(cut_off), cut_off_squared_(cut_off * cut_off) {
    // V(rc), so the shifted potential is zero at the cutoff (Forcefield.cpp:12-14)
    const double s6 = std::pow(sigma_ / cut_off_, 6);
    energy_shift_ = 4.0 * epsilon_ * (s6 * s6 - s6);
}

void LennardJonesForceField::checkCutOff(const System& system) const {
    if (2.0 * cut_off_ > system.getBoxLength()) {
        throw std::runtime_error("Cutoff must be less than half the box length");
    }
}

int LennardJonesForceField::precision() const {
    return (mixedPrecision_ && !deterministic_) ? MD_MIXED : MD_FP64;
}

double LennardJonesForceField::computeForces(System& system) const {
    checkCutOff(system);
    const md_lj lj {epsilon_, sigma_, cut_off_};
    md_force_result r {};
    check(md_forces_allpairs(system.deviceState(), &lj, &r));
    system.deviceWrote(System::Forces);
    virial_ = r.virial;
    insidePairs_ = r.inside_pairs;
    return r.potential;
}

double LennardJonesForceField::computeForces(System& system, const NeighbourList& neighbourList) const {
    checkCutOff(system);
    if (deterministic_ && !neighbourList.hasReverseIndex()) {
        throw std::runtime_error("Deterministic forces need a NeighbourList built with reverseIndex = true");
    }
    const md_lj lj {epsilon_, sigma_, cut_off_};
    md_force_result r {};
    check(md_forces(system.deviceState(), neighbourList.deviceList(), &lj, precision(), &r));
    system.deviceWrote(System::Forces);
    virial_ = r.virial;
    insidePairs_ = r.inside_pairs;
    return r.potential;
}
