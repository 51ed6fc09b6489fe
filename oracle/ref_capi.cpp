// Test infrastructure: the orc_* C ABI (orc_capi.h) implemented over the reference engine's own classes, so the
// reference sources, compiled in place by oracle/Makefile into oracle/_ref/, can be driven by the same Python
// loader as the clean-room oracle. This file is the only code of ours in that library; it includes the reference
// headers from the reference checkout and holds no reference code itself.

#include "orc_capi.h"

#include "Forcefield.hpp"
#include "Initialisation.hpp"
#include "Integrator.hpp"
#include "NeighbourList.hpp"
#include "RadialDistribution.hpp"
#include "System.hpp"
#include "Thermostat.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {
thread_local std::string g_error;
}

struct orc_sim {
    std::unique_ptr<System> sys;
    std::unique_ptr<LennardJonesForceField> ff = std::make_unique<LennardJonesForceField>(1.0, 1.0, 2.5);
    bool mixed = false, det = false;
    std::unique_ptr<NeighbourList> nl = std::make_unique<NeighbourList>(2.5, 0.2, false);
    std::unique_ptr<BussiThermostat> bussi;
    std::unique_ptr<RadialDistribution> rdf;
    double rc = 2.5, skin = 0.2;
    bool reverse = false;
    double rdf_rmax = 0.0;
    int rdf_bins = 0;
};

#define ORC_TRY(body)                                                                                          \
    try {                                                                                                      \
        body;                                                                                                  \
        return 0;                                                                                              \
    } catch (const std::exception& e) {                                                                        \
        g_error = e.what();                                                                                    \
        return -1;                                                                                             \
    }

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

void orc_set_threads(int threads) {
#ifdef _OPENMP
    omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

orc_sim* orc_create(size_t n, double box_length, const double* pos, const double* vel, const double* mass) {
    std::vector<Particle> ps(n);
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            ps[i].position[k] = pos ? pos[3 * i + k] : 0.0;
            ps[i].velocity[k] = vel ? vel[3 * i + k] : 0.0;
        }
        ps[i].mass = mass ? mass[i] : 1.0;
    }
    auto* s = new orc_sim;
    s->sys = std::make_unique<System>(ps, box_length);
    return s;
}

orc_sim* orc_create_lattice(int cells, double a, double box_length) {
    auto* s = new orc_sim;
    s->sys = std::make_unique<System>(createFCCLattice(cells, a), box_length);
    return s;
}

orc_sim* orc_create_fcc(int cells, double density, double temp, uint32_t seed) {
    const double a = std::cbrt(4.0 / density);
    orc_sim* s = orc_create_lattice(cells, a, cells * a);
    initializeVelocities(*s->sys, temp, seed);
    return s;
}

orc_sim* orc_clone(const orc_sim* s) {
    auto* c = new orc_sim;
    c->sys = std::make_unique<System>(*s->sys);
    c->ff = std::make_unique<LennardJonesForceField>(*s->ff);
    c->mixed = s->mixed;
    c->det = s->det;
    c->nl = std::make_unique<NeighbourList>(*s->nl);
    if (s->bussi) c->bussi = std::make_unique<BussiThermostat>(*s->bussi);
    if (s->rdf) c->rdf = std::make_unique<RadialDistribution>(*s->rdf);
    c->rc = s->rc;
    c->skin = s->skin;
    c->reverse = s->reverse;
    return c;
}

void orc_destroy(orc_sim* s) { delete s; }
size_t orc_size(const orc_sim* s) { return s->sys->size(); }
double orc_box_length(const orc_sim* s) { return s->sys->getBoxLength(); }

void orc_get_state(const orc_sim* s, double* pos4, double* vel, double* force, double* mass, uint64_t* ids) {
    const System& sys = *s->sys;
    const size_t n = sys.size();
    if (pos4) std::copy(sys.packedPositions(), sys.packedPositions() + 4 * n, pos4);
    for (int k = 0; k < 3; ++k) {
        if (vel) std::copy(sys.velocity(k).begin(), sys.velocity(k).end(), vel + k * n);
        if (force) std::copy(sys.force(k).begin(), sys.force(k).end(), force + k * n);
    }
    if (mass) std::copy(sys.mass().begin(), sys.mass().end(), mass);
    if (ids)
        for (size_t i = 0; i < n; ++i) ids[i] = sys.ids()[i];
}

void orc_set_positions(orc_sim* s, const double* pos4) {
    std::copy(pos4, pos4 + 4 * s->sys->size(), s->sys->packedPositions());
}

void orc_set_velocities(orc_sim* s, const double* vel) {
    const size_t n = s->sys->size();
    for (int k = 0; k < 3; ++k) std::copy(vel + k * n, vel + (k + 1) * n, s->sys->velocity(k).begin());
}

int orc_reorder(orc_sim* s, const uint64_t* order) {
    std::vector<std::size_t> o(order, order + s->sys->size());
    s->sys->reorder(o);
    return 0;
}

void orc_min_image(const orc_sim* s, const double* ri, const double* rj, double* dr) {
    const auto d = s->sys->minimumImageDistance({ri[0], ri[1], ri[2]}, {rj[0], rj[1], rj[2]});
    for (int k = 0; k < 3; ++k) dr[k] = d[k];
}

double orc_wrap(const orc_sim* s, double x) { return s->sys->wrapCoordinate(x); }
double orc_kinetic_energy(const orc_sim* s) { return s->sys->kineticEnergy(); }
double orc_temperature(const orc_sim* s) { return s->sys->temperature(); }
uint64_t orc_state_hash(const orc_sim* s) { return s->sys->stateHash(); }

int orc_cells_build(orc_sim* s, double min_cell, int* side, int* cell_of, uint64_t* start, uint64_t* order) {
    CellList cells(min_cell);
    cells.build(*s->sys);
    const int n = cells.getCellsPerSide();
    if (side) *side = n;
    const size_t count = s->sys->size();
    if (cell_of)
        for (size_t i = 0; i < count; ++i) cell_of[i] = cells.cellOf(i);
    if (start) {
        uint64_t acc = 0;
        for (int c = 0; c < n * n * n; ++c) {
            start[c] = acc;
            acc += cells.particlesIn(c).size();
        }
        start[n * n * n] = acc;
    }
    if (order)
        for (size_t i = 0; i < count; ++i) order[i] = cells.sortedOrder()[i];
    return 0;
}

void orc_cells_stencil(int side, int cell, int* out27) {
    // neighbourCells needs a built CellList of the right side: a box of side * 1.0 with cells of size 1.0
    System box({}, static_cast<double>(side));
    CellList cells(1.0);
    cells.build(box);
    const auto st = cells.neighbourCells(cell);
    std::copy(st.begin(), st.end(), out27);
}

void orc_forcefield(orc_sim* s, double eps, double sigma, double rc) {
    s->ff = std::make_unique<LennardJonesForceField>(eps, sigma, rc);
    s->ff->setMixedPrecision(s->mixed);
    s->ff->setDeterministic(s->det);
}

void orc_set_mixed(orc_sim* s, int mixed) {
    s->mixed = mixed != 0;
    s->ff->setMixedPrecision(s->mixed);
}

void orc_set_deterministic(orc_sim* s, int det) {
    s->det = det != 0;
    s->ff->setDeterministic(s->det);
}

double orc_energy_shift(const orc_sim* s) { return s->ff->getEnergyShift(); }
double orc_virial(const orc_sim* s) { return s->ff->getVirial(); }
double orc_inside_pairs(const orc_sim* s) { return s->ff->getInsidePairs(); }
int orc_forces_allpairs(orc_sim* s, double* u) { ORC_TRY(*u = s->ff->computeForces(*s->sys)) }
int orc_forces_list(orc_sim* s, double* u) { ORC_TRY(*u = s->ff->computeForces(*s->sys, *s->nl)) }

void orc_list(orc_sim* s, double cutoff, double skin, int reverse) {
    s->nl = std::make_unique<NeighbourList>(cutoff, skin, reverse != 0);
    s->rc = cutoff;
    s->skin = skin;
    s->reverse = reverse != 0;
}

int orc_list_build(orc_sim* s) { ORC_TRY(s->nl->build(*s->sys)) }
int orc_list_update(orc_sim* s, int* rebuilt) { ORC_TRY(*rebuilt = s->nl->update(*s->sys) ? 1 : 0) }
size_t orc_list_pairs(const orc_sim* s) { return s->nl->getPairCount(); }
int orc_list_rebuilds(const orc_sim* s) { return s->nl->getRebuildCount(); }

void orc_list_half(const orc_sim* s, uint64_t* offsets, uint64_t* neigh) {
    const size_t n = s->sys->size();
    size_t at = 0;
    for (size_t i = 0; i < n; ++i) {
        const auto row = s->nl->neighboursOf(i);
        if (offsets) offsets[i] = s->nl->firstPairOf(i);
        if (neigh)
            for (size_t j : row) neigh[at++] = j;
    }
    if (offsets) offsets[n] = s->nl->getPairCount();
}

int orc_list_reverse(const orc_sim* s, uint64_t* roffsets, uint64_t* rpairs) {
    if (!s->nl->hasReverseIndex()) return -1;
    const size_t n = s->sys->size();
    size_t at = 0;
    for (size_t j = 0; j < n; ++j) {
        if (roffsets) roffsets[j] = at;
        for (size_t k : s->nl->pairsAsSecond(j)) {
            if (rpairs) rpairs[at] = k;
            ++at;
        }
    }
    if (roffsets) roffsets[n] = at;
    return 0;
}

int orc_step_list(orc_sim* s, double dt, double* u) {
    ORC_TRY(*u = VelocityVerletIntegrator(dt).step(*s->sys, *s->ff, *s->nl))
}

int orc_step_allpairs(orc_sim* s, double dt, double* u) {
    ORC_TRY(*u = VelocityVerletIntegrator(dt).step(*s->sys, *s->ff))
}

void orc_rescale(orc_sim* s, double t) { rescaleVelocities(*s->sys, t); }

void orc_bussi(orc_sim* s, double t, double tau, uint64_t seed) {
    s->bussi = std::make_unique<BussiThermostat>(t, tau, seed);
}

void orc_bussi_apply(orc_sim* s, double dt) { s->bussi->apply(*s->sys, dt); }

void orc_rdf(orc_sim* s, double rmax, int bins) { s->rdf = std::make_unique<RadialDistribution>(rmax, bins); }
void orc_rdf_sample(orc_sim* s) { s->rdf->sample(*s->sys, *s->nl); }

void orc_rdf_values(const orc_sim* s, double* radii, double* g) {
    const auto r = s->rdf->radii();
    const auto v = s->rdf->values();
    if (radii) std::copy(r.begin(), r.end(), radii);
    if (g) std::copy(v.begin(), v.end(), g);
}

} // extern "C"
