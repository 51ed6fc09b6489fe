/* Test infrastructure: the C ABI shared by the CPU oracle (liboracle.so, a clean-room restatement of the
 * reference engine) and by oracle/_ref (the reference's own sources compiled here, wrapped by ref_capi.cpp).
 * Both libraries export exactly these symbols, so one Python loader (oracle/oracle.py) drives either and the
 * tests can compare them bit for bit. Nothing on the product path includes or links this header.
 *
 * One orc_sim bundles what a reference driver holds: System, LennardJonesForceField, NeighbourList,
 * VelocityVerletIntegrator, BussiThermostat and RadialDistribution (reference: src/main.cpp:169-243,
 * tests/TestHelpers.hpp:30-58). Arrays use the reference's storage: positions packed x,y,z,pad per particle
 * (System.hpp:22), velocity and force one array per component (System.hpp:131-132).
 * Functions returning int give 0 on success and -1 after a std::runtime_error, whose text orc_last_error()
 * returns (the reference's error messages, Forcefield.cpp:19-21, NeighbourList.cpp:57-59, 217-219).
 */
#ifndef MD_ORACLE_CAPI_H
#define MD_ORACLE_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

const char* orc_last_error(void);
/* Number of OpenMP threads used by the library's parallel loops (1 when built without OpenMP). */
void orc_set_threads(int threads);
int orc_get_threads(void);

/* System(particles, L) from raw arrays: pos[3n], vel[3n] (per particle x,y,z; may be NULL = 0), mass[n] (NULL = 1). */
orc_sim* orc_create(size_t n, double box_length, const double* pos, const double* vel, const double* mass);
/* createFCCLattice(cells, cbrt(4/density)) + initializeVelocities(T, seed) (Initialisation.cpp:7-76). */
orc_sim* orc_create_fcc(int cells, double density, double temperature, uint32_t seed);
/* createFCCLattice(cells, a) only, no velocities (test_dynamics.cpp:7-23). */
orc_sim* orc_create_lattice(int cells, double lattice_constant, double box_length);
orc_sim* orc_clone(const orc_sim* s);
void orc_destroy(orc_sim* s);

size_t orc_size(const orc_sim* s);
double orc_box_length(const orc_sim* s);
/* pos4[4n] packed positions, vel[3n] / force[3n] component-major (vx..., vy..., vz...), mass[n], ids[n]. */
void orc_get_state(const orc_sim* s, double* pos4, double* vel, double* force, double* mass, uint64_t* ids);
void orc_set_positions(orc_sim* s, const double* pos4);
void orc_set_velocities(orc_sim* s, const double* vel);
int orc_reorder(orc_sim* s, const uint64_t* order);
void orc_min_image(const orc_sim* s, const double* ri, const double* rj, double* dr);
double orc_wrap(const orc_sim* s, double x);
double orc_kinetic_energy(const orc_sim* s);
double orc_temperature(const orc_sim* s);
uint64_t orc_state_hash(const orc_sim* s);

/* Standalone CellList(min_cell).build(System) on the current order (CellList.cpp:9-52). */
int orc_cells_build(orc_sim* s, double min_cell, int* cells_per_side, int* cell_of, uint64_t* cell_start,
                    uint64_t* order);
void orc_cells_stencil(int cells_per_side, int cell, int* out27);

/* LennardJonesForceField(eps, sigma, rc) with its mode setters (Forcefield.hpp:8-45). */
void orc_forcefield(orc_sim* s, double eps, double sigma, double rc);
void orc_set_mixed(orc_sim* s, int mixed);
void orc_set_deterministic(orc_sim* s, int det);
double orc_energy_shift(const orc_sim* s);
int orc_forces_allpairs(orc_sim* s, double* potential);
int orc_forces_list(orc_sim* s, double* potential);
double orc_virial(const orc_sim* s);
double orc_inside_pairs(const orc_sim* s);

/* NeighbourList(rc, skin, reverse) (NeighbourList.cpp:5-134). */
void orc_list(orc_sim* s, double cutoff, double skin, int reverse);
int orc_list_build(orc_sim* s);
int orc_list_update(orc_sim* s, int* rebuilt);
size_t orc_list_pairs(const orc_sim* s);
int orc_list_rebuilds(const orc_sim* s);
/* offsets[n+1], neigh[pairs]: the half list in CSR (NeighbourList.hpp:25-31). */
void orc_list_half(const orc_sim* s, uint64_t* offsets, uint64_t* neigh);
/* reverse CSR (NeighbourList.hpp:38-41); roffsets[n+1], rpairs[pairs]; returns -1 without a reverse index. */
int orc_list_reverse(const orc_sim* s, uint64_t* roffsets, uint64_t* rpairs);

/* Oracle only (no reference function; oracle/_ref does not export these):
   orc_list_fixed: after orc_list, builds keep the particle order and full rows (j > i part, then j < i part,
   each in increasing j), and the list forces sum along those rows (the GPU adjoint's forward arithmetic).
   orc_pair_histogram: counts[bins] of every pair i < j of the box with r^2 < rmax^2, in bin
   size_t(sqrt(r^2) / (rmax / bins)), by brute force. */
void orc_list_fixed(orc_sim* s, int fixed);
int orc_pair_histogram(const orc_sim* s, double rmax, int bins, uint64_t* counts);

/* VelocityVerletIntegrator(dt).step(...) (Integrator.cpp:7-30). */
int orc_step_list(orc_sim* s, double dt, double* potential);
int orc_step_allpairs(orc_sim* s, double dt, double* potential);
void orc_rescale(orc_sim* s, double temperature);
void orc_bussi(orc_sim* s, double temperature, double tau, uint64_t seed);
void orc_bussi_apply(orc_sim* s, double dt);

/* RadialDistribution(rmax, bins) (RadialDistribution.cpp:6-69). */
void orc_rdf(orc_sim* s, double rmax, int bins);
void orc_rdf_sample(orc_sim* s);
void orc_rdf_values(const orc_sim* s, double* radii, double* g);

#ifdef __cplusplus
}
#endif
#endif
