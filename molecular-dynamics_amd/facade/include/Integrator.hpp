#pragma once

#include "Forcefield.hpp"
#include "NeighbourList.hpp"
#include "System.hpp"

// Velocity Verlet on the GPU (reference: include/Integrator.hpp, src/Integrator.cpp).
class VelocityVerletIntegrator {
public:
    explicit VelocityVerletIntegrator(double dt);

    // half kick + drift + wrap, all-pairs forces, half kick; returns the new potential energy
    double step(System& system, const LennardJonesForceField& forceField) const;
    // half kick + drift + wrap, list update, list forces, half kick; returns the new potential energy
    double step(System& system, const LennardJonesForceField& forceField, NeighbourList& neighbourList) const;

    double getTimestep() const {
        return dt_;
    }

private:
    void kickAndDrift(System& system) const;
    void kick(System& system) const;
    double dt_;
};
