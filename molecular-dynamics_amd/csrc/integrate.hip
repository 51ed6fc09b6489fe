// Velocity Verlet (reference: Integrator.cpp:7-71), the Bussi thermostat update (Thermostat.cpp:36-58) and the
// multi-step driver md_run (md_sim's production loop, main.cpp:272-304) that keeps the whole step on the GPU.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace mdgpu {

// MDGPU_FP64_GATHER=1: FP64 forces on the per-particle gather kernel, which reads every particle's int32 row
static bool fp64_on_int32_rows() {
    const char* v = std::getenv("MDGPU_FP64_GATHER");
    return v && v[0] == '1';
}

constexpr int VV_THREADS = 256;

// v += (dt/2 f) / m; r = wrap(r + dt v) per component (Integrator.cpp:32-53). With `ref` set, also the list's
// staleness test of NeighbourList::update: minimum-image |r - r_build|^2 > skin^2 / 4 (NeighbourList.cpp:25-30).
__global__ void __launch_bounds__(VV_THREADS)
    k_kick_drift(double4* __restrict__ pos, double* __restrict__ vel, const double* __restrict__ frc,
                 const double* __restrict__ mass, size_t n, double half_dt, double dt, double box, double inv_box,
                 const double4* __restrict__ ref, double max_disp2, int* __restrict__ stale, const int* halt) {
    if (halted(halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = mass[i];
    double4 r = pos[i];
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double kick = half_dt * frc[c * n + i];
        v[c] = vel[c * n + i] + kick / m;
        vel[c * n + i] = v[c];
    }
    r.x = wrap_coord(r.x + dt * v[0], box, inv_box);
    r.y = wrap_coord(r.y + dt * v[1], box, inv_box);
    r.z = wrap_coord(r.z + dt * v[2], box, inv_box);
    pos[i] = r;
    if (ref) {
        const double4 q = ref[i];
        const double dx = min_image(r.x - q.x, box, inv_box);
        const double dy = min_image(r.y - q.y, box, inv_box);
        const double dz = min_image(r.z - q.z, box, inv_box);
        if (dx * dx + dy * dy + dz * dz > max_disp2) *stale = 1;
    }
}

// v += (dt/2 f) / m (Integrator.cpp:55-71); with `partials`, each block also sums (0.5 m) v^2 in a fixed order.
__global__ void __launch_bounds__(VV_THREADS)
    k_kick(double* __restrict__ vel, const double* __restrict__ frc, const double* __restrict__ mass, size_t n,
           double half_dt, double* __restrict__ partials, const int* halt) {
    if (halted(halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double ke[1] = {0.0};
    if (i < n) {
        const double m = mass[i];
        double v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double kick = half_dt * frc[c * n + i];
            v[c] = vel[c * n + i] + kick / m;
            vel[c * n + i] = v[c];
        }
        const double half_m = 0.5 * m;
        ke[0] = half_m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    if (partials) block_sum<1, VV_THREADS>(ke, partials + blockIdx.x);
}

// The closing half kick of one step fused with the opening kick-drift of the next (NVE md_run): the same two
// updates v += (dt/2 f) / m as k_kick then k_kick_drift, in the same order, so results are bit-equal; v, f and m
// are read once instead of twice. The block partials are the closing kick's kinetic energy, as k_kick's.
__global__ void __launch_bounds__(VV_THREADS)
    k_kick_kick_drift(double4* __restrict__ pos, double* __restrict__ vel, const double* __restrict__ frc,
                      const double* __restrict__ mass, size_t n, double half_dt, double dt, double box,
                      double inv_box, const double4* __restrict__ ref, double max_disp2, int* __restrict__ stale,
                      double* __restrict__ partials, const int* halt) {
    if (halted(halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double ke[1] = {0.0};
    if (i < n) {
        const double m = mass[i];
        double v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double kick = half_dt * frc[c * n + i];
            const double closed = vel[c * n + i] + kick / m;
            v[c] = closed;
        }
        const double half_m = 0.5 * m;
        ke[0] = half_m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double kick = half_dt * frc[c * n + i];
            v[c] = v[c] + kick / m;
            vel[c * n + i] = v[c];
        }
        double4 r = pos[i];
        r.x = wrap_coord(r.x + dt * v[0], box, inv_box);
        r.y = wrap_coord(r.y + dt * v[1], box, inv_box);
        r.z = wrap_coord(r.z + dt * v[2], box, inv_box);
        pos[i] = r;
        const double4 q = ref[i];
        const double dx = min_image(r.x - q.x, box, inv_box);
        const double dy = min_image(r.y - q.y, box, inv_box);
        const double dz = min_image(r.z - q.z, box, inv_box);
        if (dx * dx + dy * dy + dz * dz > max_disp2) *stale = 1;
    }
    block_sum<1, VV_THREADS>(ke, partials + blockIdx.x);
}

// Bussi eq. A7 on the device from the reference-order kinetic energy in scalars[3] and the caller's draws
// (Thermostat.cpp:41-57). Writes the velocity scale to scalars[5] and the new kinetic energy to scalars[6].
// Draws are consumed only when KE > 0, as the reference draws only then.
__global__ void k_bussi(double* __restrict__ scalars, const double* __restrict__ draws, int* __restrict__ draw_index,
                        double dof, double kbar, double c, const int* halt) {
    if (halted(halt)) return;
    const double ke = scalars[3];
    if (!(ke > 0.0)) {
        scalars[5] = 1.0;
        scalars[6] = ke;
        return;
    }
    const int k = *draw_index;
    *draw_index = k + 1;
    const double r1 = draws[2 * k], sum_sq = draws[2 * k + 1];
    const double relax = (1.0 - c) * (kbar * (sum_sq + r1 * r1) / dof - ke);
    const double noise = 2.0 * r1 * sqrt(kbar / dof * ke * (1.0 - c) * c);
    const double ke_new = ke + relax + noise;
    scalars[5] = sqrt(ke_new / ke);
    scalars[6] = ke_new;
}

__global__ void k_scale_dev(double* __restrict__ vel, size_t count, const double* __restrict__ scale,
                            const int* halt) {
    if (halted(halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) vel[i] = vel[i] * *scale;
}

// Per-step record: potential, kinetic, virial, inside pairs -> energies[4 * step].
__global__ void k_record(const double* __restrict__ scalars, int ke_slot, double* __restrict__ energies,
                         int* __restrict__ step, const int* halt) {
    if (halted(halt)) return;
    const int k = *step;
    energies[4 * k + 0] = scalars[0];
    energies[4 * k + 1] = scalars[ke_slot];
    energies[4 * k + 2] = scalars[1];
    energies[4 * k + 3] = scalars[2];
    *step = k + 1;
}

void kick_drift_enqueue(md_state* s, double dt, md_list* check_list) {
    if (s->n == 0) return;
    k_kick_drift<<<blocks_for(s->n, VV_THREADS), VV_THREADS, 0, s->ctx->stream>>>(
        s->pos, s->vel, s->frc, s->mass, s->n, 0.5 * dt, dt, s->box, s->inv_box,
        check_list ? check_list->ref : nullptr, check_list ? check_list->max_disp2 : 0.0,
        check_list ? check_list->flags + FLAG_STALE : nullptr, s->halt);
    MD_HIP(hipGetLastError());
}

void kick_enqueue(md_state* s, double dt, double* ke_dev) {
    if (s->n == 0) return;
    const unsigned blocks = blocks_for(s->n, VV_THREADS);
    if (ke_dev) state_reserve_partials(s, blocks);
    k_kick<<<blocks, VV_THREADS, 0, s->ctx->stream>>>(s->vel, s->frc, s->mass, s->n, 0.5 * dt,
                                                      ke_dev ? s->partials : nullptr, s->halt);
    MD_HIP(hipGetLastError());
    if (ke_dev) reduce_partials_enqueue(s, static_cast<int>(blocks), 1, ke_dev);
}

// The list update of the eager step: staleness test, read-back, host-driven rebuild (NeighbourList.cpp:11-36).
static bool eager_update(md_list* l, md_state* s) {
    bool stale = !l->built || l->n != s->n;
    if (!stale) {
        hipStream_t st = l->ctx->stream;
        MD_HIP(hipMemsetAsync(l->flags + FLAG_STALE, 0, sizeof(int), st));
        displacement_check(l, s);
        MD_HIP(hipMemcpyAsync(l->ctx->host_flags, l->flags + FLAG_STALE, sizeof(int), hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
        stale = l->ctx->host_flags[0] != 0;
    }
    if (stale) list_rebuild(l, s);
    return stale;
}

struct RunBuffers {
    double* energies = nullptr;
    double* draws = nullptr;
    int* counters = nullptr;  // [0] step, [1] draw index
    ~RunBuffers() {
        dev_free(energies);
        dev_free(draws);
        dev_free(counters);
    }
};

}  // namespace mdgpu

using namespace mdgpu;


extern "C" {

int md_kick_drift(md_state* s, double dt) {
    MD_API_BEGIN
    if (!s) fail(MD_ERR_ARG, "state is NULL");
    kick_drift_enqueue(s, dt, nullptr);
    MD_API_END
}

int md_kick(md_state* s, double dt) {
    MD_API_BEGIN
    if (!s) fail(MD_ERR_ARG, "state is NULL");
    kick_enqueue(s, dt, nullptr);
    MD_API_END
}

int md_step(md_state* s, md_list* l, const md_lj* lj, int precision, double dt, md_force_result* out) {
    MD_API_BEGIN
    if (!s || !l || !lj) fail(MD_ERR_ARG, "state/list/lj is NULL");
    if (precision != MD_FP64 && precision != MD_MIXED) fail(MD_ERR_ARG, "unknown precision");
    if (!l->built || l->n != s->n || (tiling_enabled(s->n) && !l->tiled) || s->n == 0) {
        // the first build (or a new particle count): the host-driven update
        kick_drift_enqueue(s, dt, nullptr);
        eager_update(l, s);
        check_cutoff(s, *lj);
        forces_enqueue(s, l, *lj, precision, s->scalars);
        kick_enqueue(s, dt, nullptr);
        read_force_result(s, out);
        return MD_OK;
    }
    // One host synchronisation per step: the half-skin test rides on the opening kick and drift, the re-sort and
    // rebuild run gated on the device (the launches md_run makes), and the force result comes back together with
    // the list's flags. A rebuild whose rows overflow halts the step's force and kick kernels; the host then grows
    // the rows, rebuilds at the drifted state and finishes the step.
    check_cutoff(s, *lj);
    hipStream_t st = s->ctx->stream;
    if (precision == MD_FP64) ensure_rows32(l);
    struct Guard {
        md_state* s;
        md_list* l;
        ~Guard() {
            s->halt = nullptr;
            l->slots_only = false;
        }
    } guard{s, l};
    // the gated rebuild writes only the slot rows (and the int32 rows of the tiles without a staged halo, which
    // both tiled force kernels read), unless FP64 forces stay on the int32-row gather kernel
    const bool slots = (precision == MD_MIXED || !fp64_on_int32_rows()) && tiling_enabled(s->n);
    l->slots_only = slots;
    static_assert(FLAG_STALE == 0 && FLAG_OVERFLOW == 1 && FLAG_MAXCOUNT == 2 && FLAG_REBUILDS == 3,
                  "one clear covers the stale, overflow, longest-row and rebuild words");
    MD_HIP(hipMemsetAsync(l->flags, 0, 4 * sizeof(int), st));
    s->halt = l->flags + FLAG_OVERFLOW;
    kick_drift_enqueue(s, dt, l);
    list_enqueue_conditional_rebuild(l, s);
    forces_enqueue(s, l, *lj, precision, s->scalars, l->flags + FLAG_STALE, nullptr, nullptr);
    kick_enqueue(s, dt, nullptr);
    int* hf = s->ctx->host_flags;
    MD_HIP(hipMemcpyAsync(s->ctx->host_scalars, s->scalars, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    MD_HIP(hipMemcpyAsync(hf, l->flags, sizeof(int) * (FLAG_COUNT + 2), hipMemcpyDeviceToHost, st));
    MD_HIP(hipStreamSynchronize(st));
    if (hf[FLAG_OVERFLOW]) {
        // halted after the re-sort of the overflowing rebuild: longer rows, the rebuild again (the re-sort is the
        // identity now), then the step's forces and closing kick
        s->halt = nullptr;
        l->slots_only = false;
        l->padding = (hf[FLAG_MAXCOUNT] + hf[FLAG_MAXCOUNT] / 4 + 16 + 7) / 8 * 8;
        list_rebuild(l, s);
        forces_enqueue(s, l, *lj, precision, s->scalars);
        kick_enqueue(s, dt, nullptr);
        read_force_result(s, out);
        return MD_OK;
    }
    if (hf[FLAG_REBUILDS] > 0) {
        l->rebuilds += hf[FLAG_REBUILDS];
        unsigned long long total;
        std::memcpy(&total, hf + FLAG_COUNT, sizeof(total));
        l->half_pairs = static_cast<size_t>(total);
        l->max_count = hf[FLAG_MAXCOUNT];
        if (slots) l->rows32_valid = false;  // materialised again when a caller needs them
    }
    if (out) {
        out->potential = s->ctx->host_scalars[0];
        out->virial = s->ctx->host_scalars[1];
        out->inside_pairs = s->ctx->host_scalars[2];
    }
    MD_API_END
}

int md_step_allpairs(md_state* s, const md_lj* lj, double dt, md_force_result* out) {
    MD_API_BEGIN
    if (!s || !lj) fail(MD_ERR_ARG, "state/lj is NULL");
    kick_drift_enqueue(s, dt, nullptr);
    check_cutoff(s, *lj);
    forces_allpairs_enqueue(s, *lj, s->scalars);
    kick_enqueue(s, dt, nullptr);
    read_force_result(s, out);
    MD_API_END
}

}  // extern "C"

namespace mdgpu {

// Where a run keeps its per-step records, counters and Bussi draws: buffers of its own (md_run), or the static
// scratch array (md_run_static), in which case the run may not allocate at all.
struct RunMem {
    double* energies;
    int* counters;  // [0] records written, [1] Bussi draws consumed
    double* draws;
    bool no_alloc;
};

static void run_impl(md_state* s, md_list* l, const md_run_config* cfg, double* energies, const RunMem& rb) {
    if (cfg->steps < 0) fail(MD_ERR_ARG, "steps must be >= 0");
    if (cfg->precision != MD_FP64 && cfg->precision != MD_MIXED) fail(MD_ERR_ARG, "unknown precision");
    if (cfg->thermostat && !cfg->draws) fail(MD_ERR_ARG, "Bussi thermostat needs draws");
    check_cutoff(s, cfg->lj);
    // (a list last built without the tiled slot rows, e.g. by the adjoint's fixed-order build, is rebuilt first:
    // the captured step picks the tiled kernels)
    if (!l->built || l->n != s->n || (tiling_enabled(s->n) && !l->tiled)) {
        if (rb.no_alloc) fail(MD_ERR_STATE, "md_run_static needs a list built for this state (md_list_build)");
        list_rebuild(l, s);
    }
    if (cfg->steps == 0 || s->n == 0) return;
    // mixed runs on the tiled path need only the 16-bit slot rows: the device-gated rebuilds skip the int32 rows,
    // which are materialised again when a caller needs them
    if (cfg->precision == MD_FP64) ensure_rows32(l);
    struct SlotsOnly {
        md_list* l;
        bool rebuilt = false;
        ~SlotsOnly() {
            if (l->slots_only && rebuilt) l->rows32_valid = false;
            l->slots_only = false;
        }
    } slots_guard{l};
    l->slots_only = (cfg->precision == MD_MIXED || !fp64_on_int32_rows()) && tiling_enabled(s->n);
    hipStream_t st = s->ctx->stream;
    const int steps = cfg->steps;
    MD_HIP(hipMemsetAsync(rb.counters, 0, 4 * sizeof(int), st));
    if (cfg->thermostat)
        MD_HIP(hipMemcpyAsync(rb.draws, cfg->draws, 2 * sizeof(double) * steps, hipMemcpyHostToDevice, st));
    // clear the overflow word, the longest row (both accumulate over the run's builds) and the device rebuild
    // counter before the run
    MD_HIP(hipMemsetAsync(l->flags + FLAG_OVERFLOW, 0, 2 * sizeof(int), st));
    static_assert(FLAG_MAXCOUNT == FLAG_OVERFLOW + 1, "the two words are cleared together");
    MD_HIP(hipMemsetAsync(l->flags + FLAG_REBUILDS, 0, sizeof(int), st));
    MD_HIP(hipMemsetAsync(l->flags + FLAG_STALE, 0, sizeof(int), st));
    MD_HIP(hipMemsetAsync(l->flags + FLAG_STALE_ODD, 0, sizeof(int), st));
    const double dof = 3.0 * static_cast<double>(s->n) - 3.0;
    const double kbar = 0.5 * dof * cfg->temperature;
    const double c = cfg->thermostat ? std::exp(-cfg->dt / cfg->tau) : 0.0;
    const size_t nvel = 3 * s->n;
    // reserve the partials once, before any capture (no allocation may happen inside it): per-block rows of the
    // 256-thread kernels, or per-tile rows of the tiled force kernel, whichever is more (md_state_create has
    // reserved as much already)
    const size_t partials = static_cast<size_t>(std::max<size_t>(blocks_for(s->n, 256), tile_count(s->n))) * 4 + 8;
    if (rb.no_alloc && s->partials_cap < partials) fail(MD_ERR_STATE, "reduction buffer too small for md_run_static");
    state_reserve_partials(s, partials);

    // optional force-kernel timing: one event pair per force launch, eager launches only
    const bool use_graph = cfg->use_graph && !rb.no_alloc;  // instantiating a graph allocates
    const bool timed = cfg->force_kernel_ms != nullptr && !use_graph && !rb.no_alloc;
    std::vector<hipEvent_t> events;
    struct EventGuard {
        std::vector<hipEvent_t>& ev;
        ~EventGuard() {
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        }
    } guard{events};
    auto event_pair = [&](hipEvent_t& e0, hipEvent_t& e1) {
        e0 = e1 = nullptr;
        if (!timed) return;
        MD_HIP(hipEventCreate(&e0));
        events.push_back(e0);
        MD_HIP(hipEventCreate(&e1));
        events.push_back(e1);
    };

    // the pair histogram every rdf_every steps (after the step's record, on the step's final positions)
    RdfPass* rdf = nullptr;
    struct RdfGuard {
        RdfPass*& p;
        ~RdfGuard() {
            if (p) rdf_pass_destroy(p);
        }
    } rdf_guard{rdf};
    if (cfg->rdf_every > 0) {
        if (rb.no_alloc) fail(MD_ERR_ARG, "md_run_static does not sample the pair histogram");
        if (!cfg->rdf_counts) fail(MD_ERR_ARG, "rdf_every needs rdf_counts");
        rdf = rdf_pass_create(s, cfg->rdf_rmax, cfg->rdf_bins);
    }
    // a sample of s->pos when (records done + offset) is a multiple of rdf_every
    auto enqueue_rdf = [&](int offset) {
        if (rdf) rdf_pass_enqueue(rdf, s, rb.counters, offset, cfg->rdf_every);
    };

    // Every kernel of the run exits at once after a rebuild has overflowed its rows (md_state::halt): the state
    // then stands after the opening kick and drift of the step whose rebuild overflowed, and the steps before it
    // are recorded. The host grows the rows and resumes from there, repeating that rebuild.
    struct HaltGuard {
        md_state* s;
        ~HaltGuard() { s->halt = nullptr; }
    } halt_guard{s};
    s->halt = l->flags + FLAG_OVERFLOW;

    auto enqueue_forces = [&]() {
        list_enqueue_conditional_rebuild(l, s);
        hipEvent_t e0, e1;
        event_pair(e0, e1);
        forces_enqueue(s, l, cfg->lj, cfg->precision, s->scalars, l->flags + FLAG_STALE, e0, e1);
    };
    // NVE: steps k and k+1 meet in one fused kernel (closing kick of k, its kinetic energy, opening kick-drift of
    // k+1); the run is head (kick-drift, forces of step 1), steps - 1 bodies (fused kernel, record of step k,
    // forces of step k+1) and a tail (closing kick and record of the last step)
    const unsigned vv_blocks = blocks_for(s->n, VV_THREADS);
    auto enqueue_nve_body = [&]() {
        enqueue_rdf(1);  // the positions of the step whose record comes next
        k_kick_kick_drift<<<vv_blocks, VV_THREADS, 0, st>>>(s->pos, s->vel, s->frc, s->mass, s->n, 0.5 * cfg->dt,
                                                           cfg->dt, s->box, s->inv_box, l->ref, l->max_disp2,
                                                           l->flags + FLAG_STALE, s->partials, s->halt);
        reduce_partials_enqueue(s, static_cast<int>(vv_blocks), 1, s->scalars + 3);
        k_record<<<1, 1, 0, st>>>(s->scalars, 3, rb.energies, rb.counters, s->halt);
        enqueue_forces();
        MD_HIP(hipGetLastError());
    };
    // exact_kinetic: the record's kinetic energy is System::kineticEnergy of the step's final velocities, in its
    // summation order (after the thermostat's scale with Bussi on)
    const bool exact_ke = cfg->exact_kinetic != 0;
    auto enqueue_step = [&](bool kick_drift) {
        if (kick_drift) kick_drift_enqueue(s, cfg->dt, l);
        enqueue_forces();
        if (cfg->thermostat) {
            kick_enqueue(s, cfg->dt, nullptr);
            kinetic_exact_enqueue(s, s->scalars + 3);
            k_bussi<<<1, 1, 0, st>>>(s->scalars, rb.draws, rb.counters + 1, dof, kbar, c, s->halt);
            k_scale_dev<<<blocks_for(nvel, 256), 256, 0, st>>>(s->vel, nvel, s->scalars + 5, s->halt);
            if (exact_ke) kinetic_exact_enqueue(s, s->scalars + 7);
            k_record<<<1, 1, 0, st>>>(s->scalars, exact_ke ? 7 : 6, rb.energies, rb.counters, s->halt);
        } else if (exact_ke) {
            kick_enqueue(s, cfg->dt, nullptr);
            kinetic_exact_enqueue(s, s->scalars + 3);
            k_record<<<1, 1, 0, st>>>(s->scalars, 3, rb.energies, rb.counters, s->halt);
        } else {
            kick_enqueue(s, cfg->dt, s->scalars + 3);
            k_record<<<1, 1, 0, st>>>(s->scalars, 3, rb.energies, rb.counters, s->halt);
        }
        MD_HIP(hipGetLastError());
        enqueue_rdf(0);
    };
    // `body` enqueued `replays` times: captured once and replayed as a hipGraph, or launched eagerly
    auto repeat = [&](int replays, auto&& body) {
        if (replays <= 0) return;
        if (!use_graph) {
            for (int k = 0; k < replays; ++k) body();
            return;
        }
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        MD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        MD_HIP(hipStreamEndCapture(st, &graph));
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            hip_check(e, "hipGraphInstantiate");
        }
        for (int k = 0; k < replays; ++k) {
            e = hipGraphLaunch(exec, st);
            if (e != hipSuccess) break;
        }
        const hipError_t sync = hipStreamSynchronize(st);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        hip_check(e, "hipGraphLaunch");
        hip_check(sync, "hipStreamSynchronize");
    };

    const bool fused = !cfg->thermostat && !exact_ke;
    // NVE on the tiled rows (mixed and FP64) goes further: the force kernel itself carries each step (closing kick, kinetic
    // energy, opening kick and drift of the next step) for its own particles, reading the positions from one
    // buffer and writing the next step's into the other. Step k (k = 1 .. steps - 1) reads buffer (k - 1) % 2 and
    // is gated by the stale word of that parity, which its force kernel clears while setting the other one. The
    // head is the first kick-drift; the tail is the last step's forces, closing kick and record.
    const char* unfolded = std::getenv("MDGPU_UNFOLDED");  // 1: keep the separate kick-kick-drift kernel
    // (without a spare position buffer, which an allocation-free run cannot make, the separate kernels do the
    // same arithmetic)
    // (MDGPU_FP64_GATHER=1 keeps FP64 forces on the per-particle gather kernel, so FP64 runs then stay unfolded)
    const char* gather = std::getenv("MDGPU_FP64_GATHER");
    const bool fp64_gather = gather && gather[0] == '1';
    const bool can_fold = fused && tiling_enabled(s->n) && l->tiled &&
                          (cfg->precision == MD_MIXED || !fp64_gather) &&
                          !(unfolded && unfolded[0] == '1') && !(rb.no_alloc && !s->pos_spare);
    if (can_fold && steps >= 2 && !s->pos_spare) s->pos_spare = dev_alloc<double4>(s->n);
    double4* bufs[2] = {nullptr, nullptr};
    struct SpareGuard {  // on an exception the spare is whichever buffer s->pos is not
        md_state* s;
        double4** b;
        bool on = false;
        ~SpareGuard() {
            if (on) s->pos_spare = s->pos == b[0] ? b[1] : b[0];
        }
    } spare_guard{s, bufs};
    auto enqueue_fold_body = [&](int p) {
        s->pos = bufs[p];
        int* gate = l->flags + (p ? FLAG_STALE_ODD : FLAG_STALE);
        int* next = l->flags + (p ? FLAG_STALE : FLAG_STALE_ODD);
        list_enqueue_conditional_rebuild(l, s, gate);
        hipEvent_t e0, e1;
        event_pair(e0, e1);
        tiled_forces_step_enqueue(s, l, cfg->lj, cfg->precision, cfg->dt, bufs[1 - p], gate, next, s->scalars,
                                  StepRecord{rb.energies, rb.counters}, e0, e1);
        enqueue_rdf(0);  // this step's positions, still in bufs[p]
    };

    // One stretch of `n_steps` steps. With `resume` the first step's opening kick and drift are already done (the
    // state a halted stretch left). Returns true when it ran folded (positions in the two buffers).
    auto run_steps = [&](int n_steps, bool resume) -> bool {
        const bool fold = can_fold && n_steps >= 2;
        if (fold) {
            bufs[0] = s->pos;
            bufs[1] = s->pos_spare;
            spare_guard.on = true;
            if (!resume) kick_drift_enqueue(s, cfg->dt, l);
            const int bodies = n_steps - 1;
            if (use_graph && bodies >= 2) {
                // one graph holds an even and an odd step (their buffers differ)
                repeat(bodies / 2, [&]() {
                    enqueue_fold_body(0);
                    enqueue_fold_body(1);
                });
                if (bodies % 2) enqueue_fold_body(0);
            } else {
                for (int k = 1; k <= bodies; ++k) enqueue_fold_body((k - 1) % 2);
            }
            // tail: the last step's forces (stored), closing kick and record
            const int p = (n_steps - 1) % 2;
            s->pos = bufs[p];
            int* gate = l->flags + (p ? FLAG_STALE_ODD : FLAG_STALE);
            list_enqueue_conditional_rebuild(l, s, gate);
            hipEvent_t e0, e1;
            event_pair(e0, e1);
            forces_enqueue(s, l, cfg->lj, cfg->precision, s->scalars, gate, e0, e1);
            kick_enqueue(s, cfg->dt, s->scalars + 3);
            k_record<<<1, 1, 0, st>>>(s->scalars, 3, rb.energies, rb.counters, s->halt);
            MD_HIP(hipGetLastError());
            enqueue_rdf(0);
            s->pos_spare = bufs[1 - p];
            spare_guard.on = false;
            return true;
        }
        if (fused) {
            // head: the first step's kick-drift and forces; bodies; tail: the last step's closing kick and record
            if (!resume) kick_drift_enqueue(s, cfg->dt, l);
            enqueue_forces();
            repeat(n_steps - 1, enqueue_nve_body);
            kick_enqueue(s, cfg->dt, s->scalars + 3);
            k_record<<<1, 1, 0, st>>>(s->scalars, 3, rb.energies, rb.counters, s->halt);
            MD_HIP(hipGetLastError());
            enqueue_rdf(0);
            return false;
        }
        int first = 0;
        if (resume) {
            enqueue_step(false);
            first = 1;
        }
        repeat(n_steps - first, [&]() { enqueue_step(true); });
        return false;
    };

    int* hf = s->ctx->host_flags;
    int done = 0;  // steps recorded before the current stretch
    for (int attempt = 0;; ++attempt) {
        const bool folded = run_steps(steps - done, attempt > 0);
        MD_HIP(hipMemcpyAsync(hf, l->flags, sizeof(int) * (FLAG_COUNT + 2), hipMemcpyDeviceToHost, st));
        MD_HIP(hipMemcpyAsync(hf + FLAG_COUNT + 2, rb.counters, sizeof(int), hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
        if (!hf[FLAG_OVERFLOW]) break;
        if (rb.no_alloc) {
            // growing the rows allocates: leave the state frozen where the overflow halted it
            const int frozen = hf[FLAG_COUNT + 2] + 1;
            l->built = false;
            if (energies && frozen > 1) {
                MD_HIP(hipMemcpyAsync(energies, rb.energies, 4 * sizeof(double) * (frozen - 1),
                                      hipMemcpyDeviceToHost, st));
                MD_HIP(hipStreamSynchronize(st));
            }
            fail(MD_ERR_OVERFLOW, "a neighbour-list row outgrew its padding at step " + std::to_string(frozen) +
                                      " of md_run_static: the state stands after that step's opening kick and "
                                      "drift, steps before it are recorded; rebuild the list with md_list_build");
        }
        if (attempt >= 8) fail(MD_ERR_OVERFLOW, "neighbour-list rows keep overflowing during md_run");
        // the stretch froze at its step j (1-based) = the first one without a record, after that step's opening
        // kick and drift and its re-sort; folded, the positions are in the buffer step j reads
        const int recorded = hf[FLAG_COUNT + 2];
        const int j = recorded - done + 1;
        if (folded) {
            s->pos = bufs[(j - 1) % 2];
            s->pos_spare = bufs[j % 2];
        }
        done = recorded;
        // longer rows (their contents are rebuilt), then the frozen step again from its rebuild: its gate is set
        l->padding = (hf[FLAG_MAXCOUNT] + hf[FLAG_MAXCOUNT] / 4 + 16 + 7) / 8 * 8;
        list_prepare(l, s);
        MD_HIP(hipMemsetAsync(l->flags + FLAG_OVERFLOW, 0, sizeof(int), st));
        MD_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(l->flags + FLAG_STALE), 1, 1, st));
        MD_HIP(hipMemsetAsync(l->flags + FLAG_STALE_ODD, 0, sizeof(int), st));
    }
    if (energies)
        MD_HIP(hipMemcpyAsync(energies, rb.energies, 4 * sizeof(double) * steps, hipMemcpyDeviceToHost, st));
    MD_HIP(hipStreamSynchronize(st));
    if (timed) {
        double total_ms = 0.0;
        for (size_t k = 0; k + 1 < events.size(); k += 2) {
            float ms = 0.0f;
            MD_HIP(hipEventElapsedTime(&ms, events[k], events[k + 1]));
            total_ms += ms;
        }
        *cfg->force_kernel_ms = total_ms;
    }
    if (rdf) {
        const unsigned samples = rdf_pass_read(rdf, s, cfg->rdf_counts);
        if (cfg->rdf_samples) *cfg->rdf_samples += static_cast<int>(samples);
    }
    l->rebuilds += hf[FLAG_REBUILDS];
    slots_guard.rebuilt = hf[FLAG_REBUILDS] > 0;
    if (hf[FLAG_REBUILDS] > 0) {
        // the words after the flags hold the half-pair total of the last build
        unsigned long long total;
        std::memcpy(&total, hf + FLAG_COUNT, sizeof(total));
        l->half_pairs = static_cast<size_t>(total);
        l->max_count = hf[FLAG_MAXCOUNT];
    }
}

// Scratch of md_run_static: the run's counters, then its per-step records (4 doubles per step), then its Bussi
// draws (2 doubles per step). One copy per GPU, held until the process ends.
constexpr int STATIC_MAX_STEPS = 65536;
constexpr size_t STATIC_HEAD = 256;
constexpr size_t STATIC_BYTES = STATIC_HEAD + sizeof(double) * 6 * STATIC_MAX_STEPS;
static __device__ __align__(256) unsigned char scratch[STATIC_BYTES];

// The scratch array's address on `device`, looked up once per device with that device current (host code must
// use this address, not the array's name, which on the host is only a stand-in).
void* run_scratch_address(int device) {
    static std::mutex lock;
    static std::map<int, void*> known;
    std::lock_guard<std::mutex> guard(lock);
    const auto it = known.find(device);
    if (it != known.end()) return it->second;
    int current = 0;
    MD_HIP(hipGetDevice(&current));
    MD_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    const hipError_t back = hipSetDevice(current);
    hip_check(e, "hipGetSymbolAddress(scratch)");
    hip_check(back, "hipSetDevice");
    known[device] = ptr;
    return ptr;
}

}  // namespace mdgpu

extern "C" {

int md_run(md_state* s, md_list* l, const md_run_config* cfg, double* energies) {
    MD_API_BEGIN
    if (!s || !l || !cfg) fail(MD_ERR_ARG, "state/list/config is NULL");
    RunBuffers buffers;
    if (cfg->steps > 0 && s->n > 0) {
        buffers.energies = dev_alloc<double>(4 * static_cast<size_t>(cfg->steps));
        buffers.counters = dev_alloc<int>(4);
        if (cfg->thermostat) buffers.draws = dev_alloc<double>(2 * static_cast<size_t>(cfg->steps));
    }
    run_impl(s, l, cfg, energies, RunMem{buffers.energies, buffers.counters, buffers.draws, false});
    MD_API_END
}

int md_run_static(md_state* s, md_list* l, const md_run_config* cfg, double* energies) {
    MD_API_BEGIN
    if (!s || !l || !cfg) fail(MD_ERR_ARG, "state/list/config is NULL");
    if (cfg->steps > STATIC_MAX_STEPS)
        fail(MD_ERR_ARG, "md_run_static runs at most " + std::to_string(STATIC_MAX_STEPS) + " steps per call");
    if (!s->ctx->run_scratch) fail(MD_ERR_STATE, "no scratch address for this context's device");
    auto* base = static_cast<unsigned char*>(s->ctx->run_scratch);
    const size_t steps = cfg->steps > 0 ? static_cast<size_t>(cfg->steps) : 0;
    RunMem mem{reinterpret_cast<double*>(base + STATIC_HEAD), reinterpret_cast<int*>(base),
               reinterpret_cast<double*>(base + STATIC_HEAD + 4 * sizeof(double) * steps), true};
    run_impl(s, l, cfg, energies, mem);
    MD_API_END
}

}  // extern "C"
