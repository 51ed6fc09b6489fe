// Lennard-Jones force kernels (reference: Forcefield.cpp:8-450).
//
// Gather style over the padded FULL list: thread i owns particle i, walks its row and sums its own force, so
// there is one writer per particle, no atomics, and a fixed summation order. Potential energy, virial and the
// inside-cutoff count are taken over the row's first `half[i]` entries (the j > i pairs), so each pair is
// counted once, as in the reference's half list; they are reduced per wavefront by shuffles, per block through
// LDS, and across blocks by one fixed-order reduction.
//
// FP64: each pair term is F/r * d rounded once and added in row order: the j > i part reproduces the reference's
// deterministic-mode "own pairs" sum and the j < i part (increasing j) its "pairs as second" subtraction, since
// d_ij = -d_ji exactly. Forces are therefore bit-equal to Forcefield.cpp:214-326 and to the O(N^2) path.
// MIXED: double minimum image and r^2, float LJ factor (Forcefield.cpp:352-446), double accumulation.
#include <cmath>
#include <cstdlib>

#include "device.hpp"
#include "internal.hpp"
#include "pair.hpp"

namespace mdgpu {

LJConst make_lj(const md_lj& p) {
    LJConst c;
    c.sig2 = p.sigma * p.sigma;
    c.c24 = 24.0 * p.epsilon;
    c.c4 = 4.0 * p.epsilon;
    const double s6 = std::pow(p.sigma / p.cutoff, 6);
    const double s12 = s6 * s6;
    c.shift = 4.0 * p.epsilon * (s12 - s6);
    c.rc2 = p.cutoff * p.cutoff;
    c.sig2f = static_cast<float>(c.sig2);
    c.c24f = 24.0f * static_cast<float>(p.epsilon);
    c.c4f = 4.0f * static_cast<float>(p.epsilon);
    c.shiftf = static_cast<float>(c.shift);
    c.rc2f = static_cast<float>(c.rc2);
    return c;
}

constexpr int FORCE_THREADS = 256;

__global__ void __launch_bounds__(FORCE_THREADS)
    k_forces_fp64(const double4* __restrict__ pos, size_t n, const int32_t* __restrict__ rows,
                  const int32_t* __restrict__ count, const int32_t* __restrict__ half, double box, double inv_box,
                  LJConst lj, double* __restrict__ frc, double* __restrict__ partials, int* __restrict__ clear_flag,
                  const int* halt) {
    if (halted(halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    PairSums a;
    if (i < n) {
        const double4 ri = pos[i];
        const int m = count[i], h = half[i];
        const int32_t* row = rows + i;
        for (int k = 0; k < m; ++k) fp64_pair(ri, pos[row[static_cast<size_t>(k) * n]], box, inv_box, lj, k < h, a);
        frc[i] = a.fx;
        frc[n + i] = a.fy;
        frc[2 * n + i] = a.fz;
    }
    double acc[3] = {a.u, a.w, a.c};  // potential, virial, inside pairs
    if (clear_flag && i == 0) *clear_flag = 0;
    block_sum<3, FORCE_THREADS>(acc, partials + static_cast<size_t>(blockIdx.x) * 3);
}

__global__ void __launch_bounds__(FORCE_THREADS)
    k_forces_mixed(const double4* __restrict__ pos, size_t n, const int32_t* __restrict__ rows,
                   const int32_t* __restrict__ count, const int32_t* __restrict__ half, double box, double inv_box,
                   LJConst lj, double* __restrict__ frc, double* __restrict__ partials, int* __restrict__ clear_flag,
                   const int* halt) {
    if (halted(halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    PairSums a;
    if (i < n) {
        const double4 ri = pos[i];
        const int m = count[i], h = half[i];
        const int32_t* row = rows + i;
        for (int k = 0; k < m; ++k)
            mixed_pair_gather(ri, pos[row[static_cast<size_t>(k) * n]], box, inv_box, lj, k < h, a);
        frc[i] = a.fx;
        frc[n + i] = a.fy;
        frc[2 * n + i] = a.fz;
    }
    double acc[3] = {a.u, a.w, a.c};
    if (clear_flag && i == 0) *clear_flag = 0;
    block_sum<3, FORCE_THREADS>(acc, partials + static_cast<size_t>(blockIdx.x) * 3);
}

// O(N^2): particle i sums over every j != i in increasing j, skipping r^2 >= rc^2 (Forcefield.cpp:39-78); that is
// the reference's accumulation order for particle i (rows j < i subtract, then its own row adds), so the forces
// are bit-equal to it. Positions are staged through LDS one tile at a time.
__global__ void __launch_bounds__(FORCE_THREADS)
    k_forces_allpairs(const double4* __restrict__ pos, size_t n, double box, double inv_box, LJConst lj,
                      double* __restrict__ frc, double* __restrict__ partials) {
    __shared__ double4 tile[FORCE_THREADS];
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool active = i < n;
    const double4 ri = active ? pos[i] : make_double4(0, 0, 0, 0);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t base = 0; base < n; base += FORCE_THREADS) {
        __syncthreads();
        if (base + threadIdx.x < n) tile[threadIdx.x] = pos[base + threadIdx.x];
        __syncthreads();
        const size_t cnt = (n - base) < FORCE_THREADS ? (n - base) : FORCE_THREADS;
        if (!active) continue;
        for (size_t t = 0; t < cnt; ++t) {
            const size_t j = base + t;
            if (j == i) continue;
            const double4 rj = tile[t];
            const double dx = min_image(ri.x - rj.x, box, inv_box);
            const double dy = min_image(ri.y - rj.y, box, inv_box);
            const double dz = min_image(ri.z - rj.z, box, inv_box);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= lj.rc2) continue;
            const double inv_r2 = 1.0 / r2;
            const double s2 = lj.sig2 * inv_r2;
            const double s6 = s2 * s2 * s2;
            const double s12 = s6 * s6;
            const double f = lj.c24 * (2.0 * s12 - s6) * inv_r2;
            fx += f * dx;
            fy += f * dy;
            fz += f * dz;
            if (j > i) {
                acc[0] += lj.c4 * (s12 - s6) - lj.shift;
                acc[1] += f * r2;
                acc[2] += 1.0;
            }
        }
    }
    if (active) {
        frc[i] = fx;
        frc[n + i] = fy;
        frc[2 * n + i] = fz;
    }
    block_sum<3, FORCE_THREADS>(acc, partials + static_cast<size_t>(blockIdx.x) * 3);
}

void check_cutoff(const md_state* s, const md_lj& lj) {
    if (2.0 * lj.cutoff > s->box) fail(MD_ERR_RUNTIME, "Cutoff must be less than half the box length");
}

void forces_enqueue(md_state* s, md_list* l, const md_lj& p, int precision, double* out_dev, int* clear_flag,
                    hipEvent_t ev_start, hipEvent_t ev_stop) {
    hipStream_t st = s->ctx->stream;
    const size_t n = s->n;
    const unsigned blocks = blocks_for(n ? n : 1, FORCE_THREADS);
    const LJConst lj = make_lj(p);
    // both precisions run on the tiled rows (MDGPU_FP64_GATHER=1 keeps FP64 on the per-particle gather kernel,
    // whose sums the tiled FP64 kernel reproduces bit for bit)
    if (l->tiled && precision == MD_MIXED) {
        tiled_forces_enqueue(s, l, lj, p.cutoff, out_dev, clear_flag, ev_start, ev_stop);
        return;
    }
    static const bool fp64_gather = [] {
        const char* v = std::getenv("MDGPU_FP64_GATHER");
        return v && v[0] == '1';
    }();
    if (l->tiled && precision == MD_FP64 && !fp64_gather) {
        tiled_forces_fp64_enqueue(s, l, lj, out_dev, clear_flag, ev_start, ev_stop);
        return;
    }
    ensure_rows32(l);
    state_reserve_partials(s, static_cast<size_t>(blocks) * 3);
    if (ev_start) MD_HIP(hipEventRecord(ev_start, st));
    if (precision == MD_MIXED)
        k_forces_mixed<<<blocks, FORCE_THREADS, 0, st>>>(s->pos, n, l->rows, l->count, l->half, s->box, s->inv_box,
                                                          lj, s->frc, s->partials, clear_flag, s->halt);
    else
        k_forces_fp64<<<blocks, FORCE_THREADS, 0, st>>>(s->pos, n, l->rows, l->count, l->half, s->box, s->inv_box,
                                                         lj, s->frc, s->partials, clear_flag, s->halt);
    MD_HIP(hipGetLastError());
    if (ev_stop) MD_HIP(hipEventRecord(ev_stop, st));
    reduce_partials_enqueue(s, static_cast<int>(blocks), 3, out_dev);
}

void forces_enqueue(md_state* s, md_list* l, const md_lj& p, int precision, double* out_dev) {
    forces_enqueue(s, l, p, precision, out_dev, nullptr, nullptr, nullptr);
}

void forces_allpairs_enqueue(md_state* s, const md_lj& p, double* out_dev) {
    hipStream_t st = s->ctx->stream;
    const size_t n = s->n;
    const unsigned blocks = blocks_for(n ? n : 1, FORCE_THREADS);
    state_reserve_partials(s, static_cast<size_t>(blocks) * 3);
    k_forces_allpairs<<<blocks, FORCE_THREADS, 0, st>>>(s->pos, n, s->box, s->inv_box, make_lj(p), s->frc,
                                                         s->partials);
    MD_HIP(hipGetLastError());
    reduce_partials_enqueue(s, static_cast<int>(blocks), 3, out_dev);
}

void set_error(const std::string& m);

void read_force_result(md_state* s, md_force_result* out) {
    if (!out) return;
    hipStream_t st = s->ctx->stream;
    MD_HIP(hipMemcpyAsync(s->ctx->host_scalars, s->scalars, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    MD_HIP(hipStreamSynchronize(st));
    out->potential = s->ctx->host_scalars[0];
    out->virial = s->ctx->host_scalars[1];
    out->inside_pairs = s->ctx->host_scalars[2];
}

void validate_list(const md_state* s, const md_list* l) {
    if (!l->built) fail(MD_ERR_STATE, "neighbour list has not been built");
    if (l->n != s->n) fail(MD_ERR_STATE, "neighbour list was built for a different particle count");
}

}  // namespace mdgpu

using namespace mdgpu;


extern "C" {

int md_forces(md_state* s, md_list* l, const md_lj* lj, int precision, md_force_result* out) {
    MD_API_BEGIN
    if (!s || !l || !lj) fail(MD_ERR_ARG, "state/list/lj is NULL");
    if (precision != MD_FP64 && precision != MD_MIXED) fail(MD_ERR_ARG, "unknown precision");
    check_cutoff(s, *lj);
    validate_list(s, l);
    forces_enqueue(s, l, *lj, precision, s->scalars);
    read_force_result(s, out);
    MD_API_END
}

int md_forces_allpairs(md_state* s, const md_lj* lj, md_force_result* out) {
    MD_API_BEGIN
    if (!s || !lj) fail(MD_ERR_ARG, "state/lj is NULL");
    check_cutoff(s, *lj);
    forces_allpairs_enqueue(s, *lj, s->scalars);
    read_force_result(s, out);
    MD_API_END
}

}  // extern "C"
