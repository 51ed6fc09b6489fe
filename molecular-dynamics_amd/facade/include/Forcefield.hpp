#pragma once

#include "NeighbourList.hpp"
#include "System.hpp"

// Truncated and shifted Lennard-Jones (reference: include/Forcefield.hpp, src/Forcefield.cpp), computed on the GPU.
// The list path is one gather kernel over the full list: in FP64 its forces are bit-equal to the reference's
// deterministic mode (and to its default mode up to the default mode's own thread-count-dependent rounding); with
// setMixedPrecision(true) the pair maths runs in float. Both settings keep the reference's meaning and errors.
class LennardJonesForceField {
public:
    LennardJonesForceField(double epsilon, double sigma, double cut_off);

    // O(N^2) over all pairs (Forcefield.cpp:39-78); returns the potential energy
    double computeForces(System& system) const;
    // over the neighbour list (Forcefield.cpp:81-450); returns the potential energy
    double computeForces(System& system, const NeighbourList& neighbourList) const;

    double getCutOff() const {
        return cut_off_;
    }
    double getEnergyShift() const {
        return energy_shift_;
    }
    double getVirial() const {
        return virial_;
    }
    double getInsidePairs() const {
        return insidePairs_;
    }
    void setMixedPrecision(bool mixed) {
        mixedPrecision_ = mixed;
    }
    void setDeterministic(bool deterministic) {
        deterministic_ = deterministic;
    }

    // ---- mdgpu extensions ----
    double getEpsilon() const {
        return epsilon_;
    }
    double getSigma() const {
        return sigma_;
    }
    bool isMixedPrecision() const {
        return mixedPrecision_;
    }
    bool isDeterministic() const {
        return deterministic_;
    }
    // The md_precision the list path runs in: deterministic mode wins over mixed, as in the reference's dispatch.
    int precision() const;

private:
    void checkCutOff(const System& system) const;

    mutable double virial_ {0.0};
    mutable double insidePairs_ {0.0};
    bool deterministic_ {false};
    bool mixedPrecision_ {false};
    double epsilon_;
    double sigma_;
    double cut_off_;
    double cut_off_squared_;
    double energy_shift_;
};
