/* mdgpu — C ABI of the MI355X (gfx950) Lennard-Jones step engine.
 *
 * The reference engine (nhantran1711/molecular-dynamics) has no FFI: its boundary is the C++20 class API of
 * its include/ headers. This header is the thin C layer underneath a host C++ facade that keeps those classes
 * (molecular-dynamics_amd/facade/), and it is what any other host binds (ctypes, cgo, JNI). Each entry point
 * below names the reference function it replaces.
 *
 * Conventions
 *  - Every function returns an md_status; on failure md_last_error() holds a message. Where the reference throws
 *    std::runtime_error the same text is reported with MD_ERR_RUNTIME.
 *  - One md_ctx owns one HIP stream on one device; everything created from it is stream-ordered on it. Handles
 *    are not thread-safe (the reference's classes are not either, Forcefield.hpp:61-66).
 *  - Host arrays use the reference's storage: positions packed x,y,z,pad (4 doubles) per particle
 *    (System.hpp:22), velocities and forces component-major (vx[0..n), vy[..], vz[..]) (System.hpp:131-132),
 *    ids as uint64 (System.hpp:134).
 *  - Particle indices change on every list build (the state is re-sorted by cell), as in the reference
 *    (NeighbourList.hpp:21-23).
 */
#ifndef MDGPU_H
#define MDGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum md_status {
    MD_OK = 0,
    MD_ERR_ARG = 1,      /* bad argument or handle */
    MD_ERR_HIP = 2,      /* HIP runtime failure */
    MD_ERR_RUNTIME = 3,  /* the reference's std::runtime_error cases */
    MD_ERR_OVERFLOW = 4, /* list rows kept overflowing after repeated regrowth (md_run, md_batch_run and
                            md_adjoint_run grow the rows and resume by themselves) */
    MD_ERR_STATE = 5     /* call order violated (e.g. forces before any list build) */
} md_status;

typedef enum md_precision {
    MD_FP64 = 0,  /* double pair maths; forces bit-equal to the reference's deterministic mode */
    MD_MIXED = 1  /* double minimum image, float pair maths, double accumulation (Forcefield.cpp:333-450) */
} md_precision;

typedef enum md_field {
    MD_FIELD_POS = 0,   /* double[4n] */
    MD_FIELD_VEL = 1,   /* double[3n] */
    MD_FIELD_FORCE = 2, /* double[3n] */
    MD_FIELD_MASS = 3,  /* double[n] */
    MD_FIELD_ID = 4     /* uint64[n] */
} md_field;

typedef struct md_ctx md_ctx;
typedef struct md_state md_state;
typedef struct md_list md_list;

/* LennardJonesForceField(epsilon, sigma, cut_off) parameters (Forcefield.hpp:8). */
typedef struct md_lj {
    double epsilon;
    double sigma;
    double cutoff;
} md_lj;

/* What computeForces leaves behind: return value, getVirial(), getInsidePairs() (Forcefield.hpp:11-33). */
typedef struct md_force_result {
    double potential;
    double virial;
    double inside_pairs;
} md_force_result;

typedef struct md_list_info {
    size_t half_pairs;     /* NeighbourList::getPairCount() (NeighbourList.hpp:47-49) */
    int rebuilds;          /* NeighbourList::getRebuildCount() (NeighbourList.hpp:43-45) */
    int cells_per_side;    /* CellList::getCellsPerSide() of the list's grid (CellList.hpp:20-22) */
    int max_neighbours;    /* longest full-list row at the last build */
    int padding;           /* allocated row length of the padded full list */
    size_t particles;      /* particle count of the last build (rows of the list) */
    int tiled;             /* 1 when the rows also carry the 16-bit LDS-slot view read by the tiled kernels */
    int tiles;             /* tiles of 128 particles (tiled rows only) */
    int unstaged_tiles;    /* tiles whose halo exceeds the LDS capacity (their forces gather through the int32
                              rows) */
} md_list_info;

/* ---- host-side input generation (Initialisation.hpp:9-12) --------------------------------------------- */
/* createFCCLattice(cells, a) (Initialisation.cpp:7-44): writes 4 cells^3 positions, x,y,z per particle. */
int md_fcc_lattice(int cells, double lattice_constant, double* pos);
/* initializeVelocities(system, T, seed) (Initialisation.cpp:47-76) on host arrays: mt19937(seed) Gaussian draws
   per particle, centre-of-mass removal, exact rescale to T. vel is component-major [3n]; mass may be NULL (1). */
int md_maxwell_boltzmann(size_t n, const double* mass, double temperature, uint32_t seed, double* vel);

/* ---- errors and contexts ------------------------------------------------------------------------------ */
const char* md_last_error(void);
int md_device_count(int* count);
int md_ctx_create(int device, md_ctx** out);
int md_ctx_destroy(md_ctx* ctx);
int md_ctx_synchronize(md_ctx* ctx);

/* ---- System (System.hpp:12-136) ---------------------------------------------------------------------- */
/* System(particles, boxLength) (System.cpp:9-29): n particles, zero state, unit masses, ids 0..n-1.
   Upload the arrays with md_state_upload. */
int md_state_create(md_ctx* ctx, size_t n, double box_length, md_state** out);
int md_state_destroy(md_state* s);
/* Deep copy of every per-particle array; dst must have the same n (System copy assignment). */
int md_state_copy(md_state* dst, const md_state* src);
int md_state_info(const md_state* s, size_t* n, double* box_length);
int md_state_upload(md_state* s, int field, const void* host);
int md_state_download(const md_state* s, int field, void* host);
/* System::reorder: new slot k takes old slot order[k] for every array (System.cpp:42-64). */
int md_state_reorder(md_state* s, const uint64_t* order);
/* System::kineticEnergy in the reference's fixed summation order (System.cpp:67-76, Parallel.hpp:32-57). */
int md_kinetic_energy(md_state* s, double* ke);
/* v *= scale for every component (Thermostat.cpp:6-12); used by rescaleVelocities and BussiThermostat. */
int md_scale_velocities(md_state* s, double scale);
/* System::stateHash (System.cpp:85-105). */
int md_state_hash(md_state* s, uint64_t* hash);

/* ---- CellList (CellList.hpp:11-50) --------------------------------------------------------------------- */
/* CellList(min_cell).build(system) on the current order, without re-sorting (CellList.cpp:9-52). Any output
   pointer may be NULL; cell_start has side^3 + 1 entries. */
int md_cells_build(md_state* s, double min_cell, int* cells_per_side, int* cell_of, uint64_t* cell_start,
                   uint64_t* order);

/* ---- NeighbourList (NeighbourList.hpp:12-76) ----------------------------------------------------------- */
/* NeighbourList(cutOff, skin) (NeighbourList.cpp:5-8); max_neighbours = initial row padding, 0 = automatic. */
int md_list_create(md_ctx* ctx, double cutoff, double skin, int max_neighbours, md_list** out);
int md_list_destroy(md_list* l);
/* NeighbourList::build: re-sorts the state by cell and builds the padded full list (NeighbourList.cpp:55-134). */
int md_list_build(md_list* l, md_state* s);
/* NeighbourList::update: rebuild if n changed or any particle moved > skin/2 (NeighbourList.cpp:11-36). */
int md_list_update(md_list* l, md_state* s, int* rebuilt);
int md_list_get_info(const md_list* l, md_list_info* info);
/* NeighbourList copy construction / assignment (NeighbourList.hpp:12-76, implicitly defaulted there): dst takes a
   deep copy of src's parameters, cell grid and rows, so forces through either list give the same results. */
int md_list_copy(md_list* dst, const md_list* src);
/* The half list (j > i) in CSR, as neighboursOf / firstPairOf expose it (NeighbourList.hpp:25-31):
   offsets[n+1], neighbours[half_pairs]; either may be NULL. */
int md_list_half(md_list* l, uint64_t* offsets, uint64_t* neighbours);
/* The padded full list as stored: counts[n], half_counts[n], rows[n * padding] (row-major on the host). */
int md_list_full(md_list* l, int32_t* counts, int32_t* half_counts, int32_t* rows);

/* ---- LennardJonesForceField (Forcefield.hpp:6-78) ------------------------------------------------------ */
/* computeForces(system, neighbourList) (Forcefield.cpp:81-211, 214-326, 333-450). Throws the reference's
   cutoff message as MD_ERR_RUNTIME when 2 rc > L (Forcefield.cpp:18-22). out may be NULL. */
int md_forces(md_state* s, md_list* l, const md_lj* lj, int precision, md_force_result* out);
/* computeForces(system): O(N^2) all pairs (Forcefield.cpp:39-78); forces bit-equal to the reference's. */
int md_forces_allpairs(md_state* s, const md_lj* lj, md_force_result* out);

/* ---- VelocityVerletIntegrator (Integrator.hpp:7-29) ---------------------------------------------------- */
/* kickAndDrift (Integrator.cpp:32-53) and kick (Integrator.cpp:55-71). */
int md_kick_drift(md_state* s, double dt);
int md_kick(md_state* s, double dt);
/* step(system, forceField, neighbourList) (Integrator.cpp:18-30); potential may be NULL. */
int md_step(md_state* s, md_list* l, const md_lj* lj, int precision, double dt, md_force_result* out);
/* step(system, forceField) with all-pairs forces (Integrator.cpp:7-16). */
int md_step_allpairs(md_state* s, const md_lj* lj, double dt, md_force_result* out);

/* ---- BussiThermostat / rescaleVelocities (Thermostat.hpp:8-26) ------------------------------------------ */
/* One BussiThermostat::apply with the two draws supplied by the caller (Thermostat.cpp:36-58): r1 ~ N(0,1),
   sum_sq ~ Gamma((dof-1)/2, 2). No-op when the kinetic energy is <= 0. */
int md_bussi_apply(md_state* s, double temperature, double tau, double dt, double r1, double sum_sq);

/* The thermostat's random stream: BussiThermostat's std::mt19937_64(seed) (Thermostat.cpp:32-34) and, per apply,
   a fresh normal_distribution(0, 1) then gamma_distribution(0.5 (dof - 1), 2) draw (Thermostat.cpp:47-51), with
   dof = 3 n - 3. md_bussi_rng_draws writes `count` (r1, sum_sq) pairs, advancing the stream. */
typedef struct md_rng md_rng;
int md_bussi_rng_create(uint64_t seed, md_rng** out);
int md_bussi_rng_destroy(md_rng* r);
int md_bussi_rng_draws(md_rng* r, size_t n_particles, int count, double* out);

/* ---- RadialDistribution (RadialDistribution.hpp:8-29) --------------------------------------------------- */
/* RadialDistribution::sample: adds the half-list pair counts with r^2 < rmax^2, bin size_t(sqrt(r^2) / (rmax /
   bins)), to counts[bins] (RadialDistribution.cpp:11-30). With rmax beyond the list radius, where the
   reference's list holds only part of the pairs, it counts every pair within rmax (md_rdf_sample_all). */
int md_rdf_sample(md_state* s, md_list* l, double rmax, int bins, uint64_t* counts);
/* No counterpart in the reference: the same histogram over every pair of the box within rmax (a cell pass with
   cells at least rmax wide), e.g. a g(r) to 5 sigma next to a 2.8 sigma list; rmax <= L / 2. */
int md_rdf_sample_all(md_state* s, double rmax, int bins, uint64_t* counts);

/* ---- Multi-step driver (md_sim's production loop, main.cpp:272-304, without host round trips) ---------- */
typedef struct md_run_config {
    double dt;
    int precision;        /* md_precision */
    int steps;
    md_lj lj;
    int thermostat;       /* 0 = NVE, 1 = Bussi every step (main.cpp:275-277) */
    double temperature;   /* Bussi target */
    double tau;
    const double* draws;  /* Bussi draws per step: draws[2k] = r1, draws[2k+1] = sum_sq (host, 2*steps) */
    int use_graph;        /* capture one step as a hipGraph and replay it */
    double* force_kernel_ms; /* optional out (eager runs only): summed duration of the force kernel launches,
                                measured with HIP events on the context's stream */
    /* RadialDistribution::sample every rdf_every steps inside the run (main.cpp:277-279 samples the same way),
       over every pair within rdf_rmax (md_rdf_sample_all): rdf_counts[rdf_bins] (host) and *rdf_samples (host)
       receive the run's counts and number of samples, added to what they hold. rdf_every = 0: no sampling. */
    int rdf_every;
    int rdf_bins;
    double rdf_rmax;
    uint64_t* rdf_counts;
    int* rdf_samples;
    /* 1: each record's kinetic energy is System::kineticEnergy of the step's final velocities (after the
       thermostat's scale), in its summation order (System.cpp:67-76), as md_sim prints it (main.cpp:285); the
       step's kicks then run as separate kernels (the same trajectory) */
    int exact_kinetic;
} md_run_config;

/* Advances `steps` velocity-Verlet steps entirely on the device: the list update decision, re-sort and rebuild
   are driven by a device flag. energies (host, may be NULL) receives 4 doubles per step: potential, kinetic,
   virial, inside pairs. The kinetic energy is a fixed-order tree sum after the closing half kick in NVE, and
   the thermostat's new kinetic energy K' (eq. A7) with the Bussi thermostat on. A rebuild whose rows outgrow the
   padding halts every later kernel of the run, so the state stops after that step's opening kick and drift; the
   host then grows the rows and resumes from that rebuild (NeighbourList.cpp:11-36 regrows the same way), and the
   trajectory equals the one an ample padding gives. */
int md_run(md_state* s, md_list* l, const md_run_config* cfg, double* energies);

/* No counterpart in the reference: md_run for callers that cannot allocate device memory (code under graph
   capture, serving loops, harnesses that forbid hipMalloc). The same steps and results as md_run, with the run's
   counters, per-step records and Bussi draws in a static device array of the library (3 MiB per GPU, held until the
   process ends) instead of buffers of its own, and no hipMalloc / hipFree / graph instantiation / event creation.
   Limits: steps <= 65536; the list must already be built for the state (md_list_build); eager launches only
   (use_graph and force_kernel_ms are ignored); no pair histogram (rdf_every must be 0); a mixed run folds the step
   into the force kernel only when an earlier md_run made the spare position buffer (the result is the same). A row
   overflow cannot be regrown without allocating: the call then fails with MD_ERR_OVERFLOW, the state frozen after
   the overflowing step's opening kick and drift and the steps before it recorded. Every call on one GPU shares
   that GPU's array: calls that overlap on one GPU (several contexts or threads) race. */
int md_run_static(md_state* s, md_list* l, const md_run_config* cfg, double* energies);

/* ---- Trajectory frames ------------------------------------------------------------------------------------ */
/* Asynchronous download of what writeXyzFrame prints (Trajectory.cpp:4-25): positions and velocities in particle-id
   order, the inverse of the cell re-sort permutation applied on the GPU. A frame is 6 doubles per particle, x y z
   vx vy vz, row k holding the particle whose id is k. `slots` frames (1..64) can be in flight at once; each has
   device staging and pinned host memory for `capacity` particles. */
typedef struct md_frames md_frames;
int md_frames_create(md_ctx* ctx, size_t capacity, int slots, md_frames** out);
int md_frames_destroy(md_frames* f);
/* Captures the state as it stands after the work already enqueued on the context's stream, and returns at once:
   the id-order scatter runs on the context's stream, the copy to host on a stream of the frames' own, so the
   steps enqueued afterwards overlap it. Slots are used in turn; the next one must have been released
   (MD_ERR_STATE otherwise). *slot receives the slot used. */
int md_frames_capture(md_frames* f, md_state* s, int* slot);
/* Waits for the slot's copy; *frame points at its 6n doubles (pinned memory, valid until the slot is released),
   *n and *box (may be NULL) receive the captured particle count and box length. MD_ERR_STATE if some id was not
   below n (ids must be a permutation of 0..n-1, as the reference's writer assumes). md_frames_wait and
   md_frames_release may be called from one thread other than the capturing one. */
int md_frames_wait(md_frames* f, int slot, const double** frame, size_t* n, double* box);
int md_frames_release(md_frames* f, int slot);

/* ---- Adjoint (no counterpart in the reference) ----------------------------------------------------------- */
/* Gradients of L = mean_k U(r_k) + (mean_k P_k - pressure_target)^2 over k = 1..steps, P = (2 KE + W) / (3V)
   (the reference's virial pressure, main.cpp:209-211), with respect to epsilon, sigma and the initial positions and
   velocities. FP64 velocity-Verlet forward over a fixed-order list (the state is not re-sorted; `s` itself is not
   modified), (r, v, f) checkpoints every `checkpoint_every` steps, and a reverse sweep that recomputes each
   segment and runs the Hessian-vector-product adjoint kernels through it. */
typedef struct md_adjoint_config {
    double dt;
    int steps;
    int checkpoint_every;
    md_lj lj;
    double pressure_target;
} md_adjoint_config;

typedef struct md_adjoint_result {
    double loss;
    double mean_potential;
    double mean_pressure;
    double d_epsilon;
    double d_sigma;
    double forward_ms; /* forward pass, HIP events on the context stream */
    double reverse_ms; /* recomputation + adjoint sweep */
} md_adjoint_result;

/* grad_positions / grad_velocities (may be NULL): double[3n] component-major, in the state's particle order. */
int md_adjoint_run(md_state* s, md_list* l, const md_adjoint_config* cfg, md_adjoint_result* out,
                   double* grad_positions, double* grad_velocities);

/* ---- Batched replicas (no counterpart in the reference) -------------------------------------------------- */
/* Many independent Systems, each with its own box, cell grid and neighbour list, advanced together: every
   kernel launch covers all replicas, one step is one fixed sequence of launches (optionally one hipGraph
   replay), and each replica re-sorts and rebuilds on its own half-skin trigger. Per replica the step is the
   one md_run performs on a single state (same arithmetic in FP64: the trajectories are bit-equal). It stands
   for the per-replica loop of the reference's validation driver (bench/validate.py:96-120), which runs the
   single-System API once per replica. */
typedef struct md_batch md_batch;

typedef struct md_batch_config {
    double dt;
    int precision;               /* md_precision */
    int steps;
    md_lj lj;
    int thermostat;              /* 0 = NVE, 1 = Bussi on every replica */
    double tau;
    const double* temperatures;  /* [count] Bussi targets (thermostat only) */
    const double* draws;         /* [count][steps][2] Bussi draws (r1, sum_sq), replica-major (thermostat only) */
    int use_graph;               /* capture one batched step as a hipGraph and replay it */
    int lanes;                   /* lanes per particle of the force kernel: 1, 2 or 4 (MIXED only; FP64 uses 1
                                    to keep the reference's summation order); 0 = the batch's tuned value */
    double* force_kernel_ms;     /* optional out (eager runs only): summed duration of the force kernel launches */
} md_batch_config;

typedef struct md_batch_info {
    int count;                   /* replicas */
    size_t particles;            /* sum of the replicas' particle counts */
    size_t slots;                /* particle slots allocated (each replica starts on a 256-slot boundary) */
    int padding;                 /* padded row length shared by every replica's list */
    int lanes;                   /* tuned lanes per particle of the mixed force kernel */
    long long rebuilds;          /* list rebuilds summed over replicas, including the initial ones */
    size_t half_pairs;           /* half-list pairs summed over replicas, as of each replica's last build */
    double skin;                 /* list radius = cutoff + skin (md_batch_set_skin, md_batch_autotune) */
    int tiled;                   /* 1: the LDS-tiled list build and force kernels (16 particles per wavefront) */
} md_batch_info;

/* Copies the current state of each states[r] (which stays owned by the caller) into one batch, builds every
   replica's cell grid and list with the given cutoff and skin (re-sorting each replica, as md_list_build does),
   and fails with the reference's message if some box is shorter than twice cutoff + skin. */
int md_batch_create(md_ctx* ctx, md_state* const* states, int count, double cutoff, double skin, md_batch** out);
int md_batch_destroy(md_batch* b);
/* Advances every replica cfg->steps steps. The first run of a batch first evaluates the forces of the built
   state. energies (host, may be NULL): [steps][count][4] doubles (potential, kinetic, virial, inside pairs), the
   kinetic energy as md_run reports it. */
int md_batch_run(md_batch* b, const md_batch_config* cfg, double* energies);
int md_batch_get_info(const md_batch* b, md_batch_info* info);
/* Copies replica r's current state (its particle order after the re-sorts) into dst, which must hold as many
   particles; rebuilds and half_pairs (may be NULL) receive that replica's counters. */
int md_batch_replica(md_batch* b, int r, md_state* dst, int* rebuilds, size_t* half_pairs);
/* Autotuner: times `trial_steps` steps of cfg for 1, 2 and 4 lanes per particle on a copy of the batch (the
   batch itself does not move), keeps the fastest as the batch's lanes, and writes the three timings (ms per
   step) to ms_per_step[3] (may be NULL). MIXED only. */
int md_batch_tune(md_batch* b, const md_batch_config* cfg, int trial_steps, double* ms_per_step, int* best_lanes);

/* A new list radius cutoff + skin for every replica (the rebuild cadence: a replica rebuilds once a particle has
   moved skin / 2): new cell grids and rows of `padding` entries (0 = the density estimate for that radius), then
   every replica re-sorted and rebuilt at its current state; the next md_batch_run starts with a force evaluation
   over the new rows. Fails with the reference's message if some box is shorter than twice cutoff + skin. No counterpart in the reference (its skin is fixed at construction,
   NeighbourList.hpp:15). */
int md_batch_set_skin(md_batch* b, double skin, int padding);

#define MD_TUNE_MAX_TRIALS 32
typedef struct md_batch_tune_space {
    const double* skins;     /* candidate skins (rebuild cadence); NULL = the batch's own */
    int n_skins;
    const double* headroom;  /* candidate paddings as multiples (>= 1) of the longest row the best skin built;
                                NULL = the density estimate */
    int n_headroom;
    int tune_lanes;          /* also try 1, 2 and 4 lanes per particle (the per-particle kernel of untiled
                                batches; the tiled kernels work 16 particles per wavefront, 4 lanes each) */
} md_batch_tune_space;

typedef struct md_batch_tune_result {
    double skin;             /* chosen, and now the batch's */
    int padding;             /* the batch's padding after tuning */
    int lanes;               /* lanes per particle of the force kernel the batch runs */
    double ms_per_step;      /* the chosen setting's trial */
    int trials;
    double trial_skin[MD_TUNE_MAX_TRIALS];
    int trial_padding[MD_TUNE_MAX_TRIALS];
    int trial_lanes[MD_TUNE_MAX_TRIALS];
    double trial_ms[MD_TUNE_MAX_TRIALS];  /* ms per step, wall clock over trial_steps steps of cfg */
} md_batch_tune_result;

/* Autotuner over rebuild cadence, list padding and (untiled) lanes per particle: each trial copies the batch's
   state into a scratch batch, lays it out for the candidate (md_batch_set_skin), and times `trial_steps` steps of
   cfg (any precision, NVE or Bussi); skins first, then paddings for the best skin, then lanes. The batch itself
   then takes the fastest setting (one re-sort and rebuild at its current state). FP64 forces over any list radius
   are the same sums (extra list entries beyond the cut-off add nothing), so a tuned FP64 run differs from an
   untuned one only through the order the re-sorts leave the particles in (rounding level). No counterpart in
   the reference. */
int md_batch_autotune(md_batch* b, const md_batch_config* cfg, int trial_steps, const md_batch_tune_space* space,
                      md_batch_tune_result* out);

#ifdef __cplusplus
}
#endif
#endif
