// Internal declarations of the mdgpu library (gfx950). Device layout in HBM, per particle i of n:
//   pos   double4[n]      x, y, z, pad — the reference's packed layout (System.hpp:22), one 32-byte load each
//   vel   double[3n]      vx | vy | vz component-major (System.hpp:131), coalesced per component
//   frc   double[3n]      fx | fy | fz
//   mass  double[n], id uint64[n]
// Neighbour list (one per md_list): a padded FULL list stored column-major, rows[k * n + i] = k-th neighbour of
// i, so lane i of a wavefront reads consecutive words for the same k. Each row holds first the j > i
// neighbours in the reference's 27-cell visit order (exactly the reference's half-list row), then the j < i
// neighbours in increasing index; count[i] and half[i] give both lengths.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <exception>
#include <string>

#include "mdgpu.h"

namespace mdgpu {

struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);
#define MD_HIP(expr) ::mdgpu::hip_check((expr), #expr)

template <class T>
T* dev_alloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

template <class T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Grows a device buffer to at least `count` elements (contents are not preserved).
template <class T>
void dev_reserve(T*& p, size_t& cap, size_t count) {
    if (count <= cap && p) return;
    dev_free(p);
    p = dev_alloc<T>(count);
    cap = count;
}

inline unsigned blocks_for(size_t n, unsigned threads) {
    return static_cast<unsigned>((n + threads - 1) / threads);
}

// Device-side flags of a list, shared by the eager and the captured step.
enum Flag : int {
    FLAG_STALE = 0,     // some particle moved more than skin/2 since the last build
    FLAG_OVERFLOW = 1,  // a row needed more than `padding` entries
    FLAG_MAXCOUNT = 2,  // longest row seen by the last build
    FLAG_REBUILDS = 3,  // rebuilds performed on the device
    FLAG_FORCE = 4,     // host-requested rebuild (first build, n changed)
    FLAG_STALE_ODD = 5, // FLAG_STALE of the odd steps of a fused md_run (each step's kernel sets the next one's)
    FLAG_OVERFLOW_STICKY = 6,  // FLAG_OVERFLOW and FLAG_MAXCOUNT of the device-gated builds only, cleared only by
    FLAG_MAXCOUNT_STICKY = 7,  // the caller that reads them (host-driven builds clear the other two)
    FLAG_COUNT = 8
};

// True inside a kernel of a halted md_run (see md_state::halt).
__device__ __forceinline__ bool halted(const int* halt) { return halt && *halt != 0; }
// A device-gated rebuild kernel runs only when its gate word is set and the run is not halted.
__device__ __forceinline__ bool gated_off(const int* gate, const int* halt) {
    return (gate && *gate == 0) || halted(halt);
}

}  // namespace mdgpu

// Body wrapper of every extern "C" entry point: internal errors become status codes + md_last_error().
#define MD_API_BEGIN try {
#define MD_API_END                                                                                             \
    return MD_OK;                                                                                              \
    }                                                                                                          \
    catch (const ::mdgpu::Error& e) {                                                                          \
        ::mdgpu::set_error(e.msg);                                                                             \
        return e.code;                                                                                         \
    }                                                                                                          \
    catch (const std::exception& e) {                                                                          \
        ::mdgpu::set_error(e.what());                                                                          \
        return MD_ERR_HIP;                                                                                     \
    }

struct md_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // pinned host words for small read-backs (flags, scalars)
    double* host_scalars = nullptr;  // 64 doubles
    int* host_flags = nullptr;       // 16 ints
    int live = 0;                    // states and lists created from this context and not yet destroyed
    void* run_scratch = nullptr;     // this device's address of md_run_static's scratch array
};

struct md_state {
    md_ctx* ctx = nullptr;
    size_t n = 0;
    double box = 0.0, inv_box = 0.0;
    double4* pos = nullptr;
    double* vel = nullptr;
    double* frc = nullptr;
    double* mass = nullptr;
    uint64_t* id = nullptr;
    // second copies used as the destination of a permutation
    double4* pos_alt = nullptr;
    double* vel_alt = nullptr;
    double* frc_alt = nullptr;
    double* mass_alt = nullptr;
    uint64_t* id_alt = nullptr;
    // positions after the drift of a fused md_run step (allocated by the first such run)
    double4* pos_spare = nullptr;
    // reductions: per-block partials and the device result words
    double* partials = nullptr;
    size_t partials_cap = 0;
    double* scalars = nullptr;  // [0..2] U, W, inside; [3] KE; [4] KE fast; [5..7] scratch
    double* red_blocks = nullptr;   // per-block sums of the partials reduction
    unsigned* red_ticket = nullptr; // its completion counter (zero between launches)
    // During md_run: the list's overflow word. Every kernel the run enqueues on this state exits at once while it
    // is set, so a rebuild that overflows its rows freezes the state after that step's opening kick and drift.
    const int* halt = nullptr;
};

struct md_list {
    md_ctx* ctx = nullptr;
    double cutoff = 0.0, skin = 0.0, radius = 0.0, radius2 = 0.0, max_disp2 = 0.0;
    int side = 0;
    size_t ncells = 0, ncells_cap = 0;
    size_t n = 0, n_cap = 0;
    int padding = 0;
    size_t rows_cap = 0;
    bool built = false;
    int rebuilds = 0;
    size_t half_pairs = 0;
    int max_count = 0;
    // cell grid
    uint32_t* cell_count = nullptr;  // ncells
    uint32_t* cell_start = nullptr;  // ncells + 1
    uint32_t* scan_tmp = nullptr;    // per-cell fill counters of the scatter
    uint32_t* scan_totals = nullptr; // per-block totals of the cell scan
    size_t scan_cap = 0;
    int* cell_of = nullptr;          // n (cell of particle i in the current order)
    int* cell_of_alt = nullptr;
    uint32_t* order_raw = nullptr;   // n, unstable scatter result
    uint32_t* order = nullptr;       // n, stable sorted order
    double4* ref = nullptr;          // n, positions at the last build
    // padded full list
    int32_t* rows = nullptr;         // padding * n, column-major
    int32_t* count = nullptr;        // n
    int32_t* half = nullptr;         // n
    int* flags = nullptr;            // FLAG_COUNT ints on the device
    int2* build_stats = nullptr;     // per build block: longest row, half-row sum (reduced by k_finish_build)
    size_t stats_cap = 0;
    // tiled view of the same rows (tiled.hip): LDS slot byte offsets into each tile's staged halo
    uint16_t* rows16 = nullptr;      // ceil(padding / 32) chunks of 32 entries per particle, chunk-major
    size_t rows16_cap = 0;
    int4* tile_ranges = nullptr;     // tiles * TILE_RANGES (first cell, first particle, first slot, count)
    int* tile_nr = nullptr;          // tiles; number of ranges, -1 if the tile is not staged
    int* tile_order = nullptr;       // tiles; the launch order of the tiles (cell columns in bands, see tiled.hip)
    size_t tiles_cap = 0;
    int hmax = 0;                    // LDS capacity of a tile's halo, in particles
    bool tiled = false;              // rows16 / tile_* describe the current rows
    bool rows32_valid = false;       // `rows` (int32) matches the current list
    bool slots_only = false;         // device-gated builds may write only rows16 (mixed md_run)
    mutable int unstaged = -1;       // tiles with tile_nr = -1 as of the last build (-1: not counted yet)
};

namespace mdgpu {

// An md_run step's record, written by the four-sum reduction (U, W, inside count, KE) when `energies` is set:
// energies[4 k ..] = U, KE, W, count for k = *step, then *step = k + 1.
struct StepRecord {
    double* energies = nullptr;
    int* step = nullptr;
};

// Re-sort of the state into its alt arrays: source and destination arrays (frc only when forces are carried), the
// positions at the build (ref) and the particles' cells in the new order.
enum PermuteMode : int {
    PERMUTE_NONE = 0,
    PERMUTE_STATE = 1,  // the order, the cells and the state
    PERMUTE_ORDER = 2   // the order and the cells only (the state follows in the tile setup)
};
struct PermuteArgs {
    int enabled;
    const double4* pos;
    const double* vel;
    const double* frc;
    const double* mass;
    const uint64_t* id;
    double4* pos2;
    double* vel2;
    double* frc2;
    double* mass2;
    uint64_t* id2;
    double4* ref;
    int* cell_of2;
};

// Member v of the old order goes to position k of the new one.
__device__ __forceinline__ void permute_member(const PermuteArgs& pa, size_t k, size_t v, size_t n) {
    const double4 r = pa.pos[v];
    pa.pos2[k] = r;
    pa.ref[k] = r;
    for (int c = 0; c < 3; ++c) pa.vel2[c * n + k] = pa.vel[c * n + v];
    if (pa.frc)
        for (int c = 0; c < 3; ++c) pa.frc2[c * n + k] = pa.frc[c * n + v];
    pa.mass2[k] = pa.mass[v];
    pa.id2[k] = pa.id[v];
}

// cells.hip
PermuteArgs permute_args(md_list* l, md_state* s, bool carry_forces);  // mode PERMUTE_STATE
void cells_build(md_list* l, md_state* s, double min_cell, int side);   // count, scan, scatter, stable order
void list_rebuild(md_list* l, md_state* s);                            // full host-driven rebuild + read-back
void list_enqueue_conditional_rebuild(md_list* l, md_state* s);       // device-flag-gated rebuild (md_run)
void list_enqueue_conditional_rebuild(md_list* l, md_state* s, const int* gate);  // gated by another flag word
void list_prepare(md_list* l, md_state* s);                            // grid/buffers for the state's size
void displacement_check(md_list* l, md_state* s);                      // sets FLAG_STALE
void list_rebuild_fixed(md_list* l, md_state* s);                      // list without re-sorting (host-driven)
void list_enqueue_conditional_rebuild_fixed(md_list* l, md_state* s); // same, gated by FLAG_STALE
// build totals (FLAG_MAXCOUNT, FLAG_OVERFLOW, half-pair total, FLAG_REBUILDS when gated) from `blocks` records
void finish_build_enqueue(md_list* l, int blocks, const int* gate, const int* halt);

// api.hip
void state_reserve_partials(md_state* s, size_t count);
void kinetic_exact_enqueue(md_state* s, double* out_dev);              // reference-order KE into *out_dev

// forces.hip
void forces_enqueue(md_state* s, md_list* l, const md_lj& lj, int precision, double* out_dev);
// same, and the force kernel clears *clear_flag (the list's stale flag inside md_run)
// (events, when set, are recorded on the stream right before and after the force kernel)
void forces_enqueue(md_state* s, md_list* l, const md_lj& lj, int precision, double* out_dev, int* clear_flag,
                    hipEvent_t ev_start, hipEvent_t ev_stop);
void forces_allpairs_enqueue(md_state* s, const md_lj& lj, double* out_dev);
void check_cutoff(const md_state* s, const md_lj& lj);
void validate_list(const md_state* s, const md_list* l);
void read_force_result(md_state* s, md_force_result* out);
void set_error(const std::string& m);

// tiled.hip. The tiled list and force path is the default; MDGPU_UNTILED=1 selects the per-particle global
// gather kernels instead (same results: the rows are identical).
bool tiling_enabled(size_t n);
int tile_count(size_t n);
constexpr int TILE_PARTICLES = 128;
constexpr int TILE_RANGES = 64;
// LDS capacity bound of a tile's halo, in particles: x | y | z doubles of 2184 slots keep the force kernel within
// 53248 B (13 x 4 KB) and so at three blocks per CU (160 KB of LDS); the few tiles whose halo is larger take the
// unstaged path
constexpr int TILE_HALO_MAX = 2184;
// The system a tile belongs to: the one System of the launch, or one replica of a batch (batch.hip), whose
// particles are the batch slots [first, end) and whose cell offsets are start[cell_base .. cell_base + side^3].
// Tiles never straddle systems (a replica starts on a tile boundary); tiles past a replica's end are empty.
struct TileSystem {
    uint32_t first, end;
    uint32_t cell_base;
    int side, use_cells;
    int hmax;         // LDS capacity of a tile's halo, in particles
    int interior_ok;  // side >= 7: an interior cell's 27 neighbours are nearer than L/2 along every axis
    int gate;         // batch: index of the replica's rebuild flag in TileSystems::gates
    double box, inv_box;
    double flat_cells;  // cells a tile and its halo may span with every coordinate inside half a box
    double reach;       // rc + list radius + skin, the mixed force kernel's compactness test
};

// What a launch sees: every tile in `one`, or tile t in per[rep_of_tile[t]]; `stride` is the column stride of the
// int32 rows and of the component-major velocity and force arrays (n, or the batch's slot count).
struct TileSystems {
    TileSystem one;
    const TileSystem* per;
    const uint16_t* rep_of_tile;
    const int* gates;  // batch: per replica, nonzero when its list is rebuilt (the rebuild kernels skip the others)
    // batch rebuild kernels: active[0] tiles are rebuilt this step, active[1 + k] the k-th of them (replica order)
    const int* active;
    // single system: the order in which launches take the tiles (order[k] = the k-th tile; nullptr = tile order);
    // written by the tile setup of each rebuild
    const int* order;
    size_t stride;
    int tiles;
};

// The re-sorted arrays a list build moves back into place (enabled: the build reads the re-sorted positions).
struct TiledCopyBack {
    int enabled;
    const double* vel2;
    const double* frc2;
    const double* mass2;
    const uint64_t* id2;
    const int* cell_of2;
    double4* pos;
    double* vel;
    double* frc;
    double* mass;
    uint64_t* id;
    int* cell_of;
};

// A batch's tiled list (buffers owned by batch.hip): slot rows, int32 rows of the unstaged tiles, per tile its halo
// ranges and build record.
struct BatchTiles {
    TileSystems ts;
    int4* tile_ranges;
    int* tile_nr;
    uint16_t* rows16;
    int2* stats;
    int32_t* rows;
    int32_t* count;
    int32_t* half;
    int padding;
};
int tile_halo_capacity(size_t n, size_t ncells);  // (cells.hip) LDS slots for a tile's halo of a system
int tiled_row_len(int padding);                   // slot-row entries per particle
// rebuild of the gated replicas: halo ranges from the re-sorted cells, then rows from build_pos (copied back by cb)
void tiled_batch_build_enqueue(hipStream_t st, const BatchTiles& bt, const double4* build_pos, const int* cell_of,
                               const uint32_t* start, double radius2, const TiledCopyBack& cb, const int* halt);
// forces of every replica: per tile U, W, inside count into partials[3 tile ..]; with kick_vel (mixed only) also the
// closing kick of the velocities and the kicked kinetic energy, the tile's fourth sum (partials[4 tile ..])
struct LJConst;
void tiled_batch_forces_enqueue(hipStream_t st, const BatchTiles& bt, const double4* pos, const LJConst& lj,
                                int precision, double* frc, double* partials, const int* halt,
                                double* kick_vel = nullptr, const double* mass = nullptr, double half_dt = 0.0);
size_t tiled_rows16_entries(size_t n, int padding);
void tiled_list_enqueue(md_list* l, md_state* s, bool from_alt, const int* gate, bool write32, bool carry_forces);
void tiled_rows32_enqueue(md_list* l);                          // int32 rows from the slot rows
void ensure_rows32(md_list* l);                                 // (cells.hip) materialise if stale
// mixed-precision forces over the tiled rows
struct LJConst;
void tiled_forces_enqueue(md_state* s, md_list* l, const LJConst& lj, double cutoff, double* out_dev,
                          int* clear_flag, hipEvent_t ev_start, hipEvent_t ev_stop);
// FP64 forces over the tiled rows, bit-equal to the per-particle gather kernel (k_forces_fp64)
void tiled_forces_fp64_enqueue(md_state* s, md_list* l, const LJConst& lj, double* out_dev, int* clear_flag,
                               hipEvent_t ev_start, hipEvent_t ev_stop);
// the mixed or FP64 tiled forces with the NVE step of the tile's particles folded in: U, W, inside count and KE to
// out_dev[0..3] (and the step's record), new positions to pos_next, the next displacement test to *stale_next; the
// forces are not stored
void tiled_forces_step_enqueue(md_state* s, md_list* l, const md_lj& p, int precision, double dt, double4* pos_next,
                               int* clear_flag, int* stale_next, double* out_dev, StepRecord rec, hipEvent_t ev_start,
                               hipEvent_t ev_stop);

// integrate.hip
void kick_drift_enqueue(md_state* s, double dt, md_list* check_list);
void* run_scratch_address(int device);  // md_run_static's scratch array on `device` (looked up once per device)
void kick_enqueue(md_state* s, double dt, double* ke_dev);

// rdf.hip: the cell pass over every pair within rmax (buffers for one state size, made before a run)
struct RdfPass;
RdfPass* rdf_pass_create(md_state* s, double rmax, int bins);
void rdf_pass_destroy(RdfPass* p);
// a sample of s's positions; inside md_run only when (*counter + offset) % every == 0 (counter may be NULL)
void rdf_pass_enqueue(RdfPass* p, md_state* s, const int* counter, int offset, int every);
unsigned rdf_pass_read(RdfPass* p, md_state* s, uint64_t* counts);

// Final fixed-order reduction of `blocks` rows of `width` partials (s->partials) into out_dev[0..width).
void reduce_partials_enqueue(md_state* s, int blocks, int width, double* out_dev);
void reduce_partials_enqueue(md_state* s, int blocks, int width, double* out_dev, StepRecord rec);

}  // namespace mdgpu
