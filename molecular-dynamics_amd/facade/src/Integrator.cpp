#include "Integrator.hpp"

#include "MdgpuRuntime.hpp"

using mdgpu_facade::check;

VelocityVerletIntegrator::VelocityVerletIntegrator(double dt) : dt_(dt) {
}

void VelocityVerletIntegrator::kickAndDrift(System& system) const {
    check(md_kick_drift(system.deviceState(), dt_));
    system.deviceWrote(System::Positions | System::Velocities);
}

void VelocityVerletIntegrator::kick(System& system) const {
    check(md_kick(system.deviceState(), dt_));
    system.deviceWrote(System::Velocities);
}

// Integrator.cpp:7-16
double VelocityVerletIntegrator::step(System& system, const LennardJonesForceField& forceField) const {
    kickAndDrift(system);
    const double potential = forceField.computeForces(system);
    kick(system);
    return potential;
}

// Integrator.cpp:18-30: the same calls in the same order, each running on the GPU
double VelocityVerletIntegrator::step(System& system, const LennardJonesForceField& forceField,
                                      NeighbourList& neighbourList) const {
    kickAndDrift(system);
    neighbourList.update(system);
    const double potential = forceField.computeForces(system, neighbourList);
    kick(system);
    return potential;
}
