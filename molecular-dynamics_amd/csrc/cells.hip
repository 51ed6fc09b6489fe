// Cell grid, stable counting sort into cells, re-sort of the state by cell, and the one-pass padded full Verlet
// list (reference: CellList.cpp:9-83, NeighbourList.cpp:11-134, System.cpp:42-64).
//
// Rebuild pipeline, every kernel optionally gated by a device flag so a captured step can carry it:
//   k_cell_count   cell id per particle (CellList.cpp:26-38) + atomic histogram
//   k_scan_*       exclusive prefix sum of the histogram -> cell start offsets (CellList.cpp:41-44)
//   k_cell_scatter particle index into its cell's range, unordered (atomics)
//   k_cell_sort    one wavefront per cell ranks its members by index, restoring the counting sort's stable order
//                  (CellList.cpp:47-51); when re-sorting, it records each particle's cell in the new order and
//                  gathers every particle array into that order (System::reorder) with the build-time positions
//                  (NeighbourList.cpp:129-132) — on the tiled path k_tile_setup does the gather instead
//   k_list_build   one thread per particle: the j > i candidates of the 27 cells in the reference's visit order
//                  (its half-list row), then the j < i candidates in increasing index; r^2 < (rc+skin)^2
//                  strictly (NeighbourList.cpp:75-98); also moves the re-sorted arrays back into place.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace mdgpu {

struct CopyBackArgs {
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

// the fixed-order build: lanes per particle and block shape
constexpr int FIXED_LANES = 16;
constexpr int FIXED_THREADS = 256;
constexpr int FIXED_PER_BLOCK = FIXED_THREADS / FIXED_LANES;

// Cell of every particle and the cell histogram, one atomic per run of equal cells.
__global__ void __launch_bounds__(256) k_cell_count(const double4* __restrict__ pos, size_t n, double per_length,
                                                     int side, int* __restrict__ cell_of,
                                                     uint32_t* __restrict__ counts, const int* gate,
                                                     const int* halt) {
    if (gated_off(gate, halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int c = -1;
    if (valid) {
        const double4 r = pos[i];
        c = cell_of_position(r.x, r.y, r.z, per_length, side);
        cell_of[i] = c;
    }
    const CellRun run = cell_run(c, valid);
    if (run.head) atomicAdd(&counts[c], static_cast<uint32_t>(run.len));
}

// Two-phase exclusive scan of the cell histogram, SCAN_ITEMS cells per block: block totals, then each block adds
// the totals before it and scans its own cells. start[ncells] receives the total.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 8;
constexpr int SCAN_ITEMS = SCAN_THREADS * SCAN_PER_THREAD;

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_totals(const uint32_t* __restrict__ counts, size_t ncells,
                                                               uint32_t* __restrict__ totals, const int* gate,
                                                     const int* halt) {
    if (gated_off(gate, halt)) return;
    __shared__ uint32_t wsum[SCAN_THREADS / WAVE];
    const size_t base = static_cast<size_t>(blockIdx.x) * SCAN_ITEMS + threadIdx.x * SCAN_PER_THREAD;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k)
        if (base + k < ncells) s += counts[base + k];
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if (threadIdx.x % WAVE == 0) wsum[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < SCAN_THREADS / WAVE; ++w) t += wsum[w];
        totals[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t* __restrict__ counts, size_t ncells,
                                                              const uint32_t* __restrict__ totals,
                                                              uint32_t* __restrict__ start, const int* gate,
                                                     const int* halt) {
    if (gated_off(gate, halt)) return;
    __shared__ uint32_t wsum[SCAN_THREADS / WAVE];
    __shared__ uint32_t prior[SCAN_THREADS / WAVE];
    {  // the particles of the blocks before this one
        uint32_t t = 0;
        for (unsigned b = threadIdx.x; b < blockIdx.x; b += SCAN_THREADS) t += totals[b];
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, WAVE);
        if (threadIdx.x % WAVE == 0) prior[threadIdx.x / WAVE] = t;
    }
    const size_t base = static_cast<size_t>(blockIdx.x) * SCAN_ITEMS + threadIdx.x * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        v[k] = base + k < ncells ? counts[base + k] : 0u;
        s += v[k];
    }
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    uint32_t incl = s;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, WAVE);
        if (lane >= off) incl += o;
    }
    if (lane == WAVE - 1) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < SCAN_THREADS / WAVE; ++w) before += prior[w];
    for (int w = 0; w < wave; ++w) before += wsum[w];
    uint32_t run = before + incl - s;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        if (base + k < ncells) start[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) start[ncells] = run;
}

__global__ void __launch_bounds__(256) k_cell_scatter(const int* __restrict__ cell_of, size_t n,
                                                       const uint32_t* __restrict__ start,
                                                       uint32_t* __restrict__ fill, uint32_t* __restrict__ order_raw,
                                                       const int* gate, const int* halt) {
    if (gated_off(gate, halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int c = valid ? cell_of[i] : -1;
    // one atomic per run of equal cells reserves the run's slots; its lanes take them in lane order
    const CellRun run = cell_run(c, valid);
    uint32_t base = 0;
    if (run.head) base = atomicAdd(&fill[c], static_cast<uint32_t>(run.len));
    base = static_cast<uint32_t>(__shfl(static_cast<int>(base), run.first, WAVE));
    if (valid) order_raw[start[c] + base + static_cast<uint32_t>((threadIdx.x % WAVE) - run.first)] = static_cast<uint32_t>(i);
}

// One wavefront per cell: rank = number of members with a smaller index, so the members land in increasing
// index order — exactly the order of the reference's serial counting sort. Also clears the cell's histogram
// and fill counters for the next build.
__device__ __forceinline__ void sort_cell(size_t cell, int lane, const uint32_t* __restrict__ start, size_t n,
                                          const uint32_t* __restrict__ order_raw, uint32_t* __restrict__ order,
                                          uint32_t* __restrict__ fill, uint32_t* __restrict__ counts,
                                          const PermuteArgs& pa) {
    const uint32_t b = start[cell];
    const uint32_t m = start[cell + 1] - b;
    if (lane == 0) {
        fill[cell] = 0;
        counts[cell] = 0;
    }
    for (uint32_t base = 0; base < m; base += WAVE) {
        const bool valid = base + lane < m;
        const uint32_t v = valid ? order_raw[b + base + lane] : 0xffffffffu;
        uint32_t rank = 0;
        if (m <= WAVE) {
            for (uint32_t s = 0; s < m; ++s) rank += (static_cast<uint32_t>(__shfl(static_cast<int>(v), s, WAVE)) < v);
        } else {
            for (uint32_t s = 0; s < m; ++s) rank += (order_raw[b + s] < v);
        }
        if (valid) {
            const uint32_t k = b + rank;
            order[k] = v;
            if (pa.enabled) pa.cell_of2[k] = static_cast<int>(cell);
            if (pa.enabled == PERMUTE_STATE) permute_member(pa, k, v, n);
        }
    }
}

// Wavefronts stride over the cells (a small grid, so a launch gated off costs little).
constexpr unsigned SORT_BLOCKS = 2048;
__global__ void __launch_bounds__(256) k_cell_sort(const uint32_t* __restrict__ start, size_t ncells, size_t n,
                                                    const uint32_t* __restrict__ order_raw,
                                                    uint32_t* __restrict__ order, uint32_t* __restrict__ fill,
                                                    uint32_t* __restrict__ counts, PermuteArgs pa, const int* gate,
                                                    const int* halt) {
    if (gated_off(gate, halt)) return;
    const int lane = threadIdx.x % WAVE;
    const size_t waves = static_cast<size_t>(gridDim.x) * (blockDim.x / WAVE);
    for (size_t cell = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / WAVE; cell < ncells;
         cell += waves)
        sort_cell(cell, lane, start, n, order_raw, order, fill, counts, pa);
}

__device__ __forceinline__ void sort3(int& a, int& b, int& c) {
    int t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
}

__global__ void __launch_bounds__(256) k_list_build(const double4* __restrict__ pos, size_t n,
                                                     const int* __restrict__ cell_of,
                                                     const uint32_t* __restrict__ start, int side, double box,
                                                     double inv_box, double radius2, int use_cells, int padding,
                                                     int32_t* __restrict__ rows, int32_t* __restrict__ count,
                                                     int32_t* __restrict__ half, int2* __restrict__ stats,
                                                     CopyBackArgs cb, const int* gate,
                                                     const int* halt) {
    if (gated_off(gate, halt)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int cnt = 0, hcnt = 0;
    if (i < n) {
        const double4 ri = pos[i];
        if (cb.enabled) {
            cb.pos[i] = ri;
            for (int c = 0; c < 3; ++c) cb.vel[c * n + i] = cb.vel2[c * n + i];
            if (cb.frc)
                for (int c = 0; c < 3; ++c) cb.frc[c * n + i] = cb.frc2[c * n + i];
            cb.mass[i] = cb.mass2[i];
            cb.id[i] = cb.id2[i];
            cb.cell_of[i] = cb.cell_of2[i];
        }
        auto consider = [&](uint32_t j) {
            const double4 rj = pos[j];
            const double dx = min_image(ri.x - rj.x, box, inv_box);
            const double dy = min_image(ri.y - rj.y, box, inv_box);
            const double dz = min_image(ri.z - rj.z, box, inv_box);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < radius2) {
                if (cnt < padding) rows[static_cast<size_t>(cnt) * n + i] = static_cast<int32_t>(j);
                ++cnt;
            }
        };
        const uint32_t self = static_cast<uint32_t>(i);
        if (use_cells) {
            const int c = cell_of[i];
            const int cx = c / (side * side), cy = (c / side) % side, cz = c % side;
            // j > i in the reference's stencil order: dx, then dy, then dz, each -1..1 (CellList.cpp:67-80)
            for (int ox = -1; ox <= 1; ++ox) {
                const int x = wrap_index(cx + ox, side);
                for (int oy = -1; oy <= 1; ++oy) {
                    const int y = wrap_index(cy + oy, side);
                    for (int oz = -1; oz <= 1; ++oz) {
                        const int nc = (x * side + y) * side + wrap_index(cz + oz, side);
                        if (nc < c) continue;  // lower cells hold only lower indices
                        const uint32_t e = start[nc + 1];
                        for (uint32_t j = (nc == c ? self + 1 : start[nc]); j < e; ++j) consider(j);
                    }
                }
            }
            hcnt = cnt;
            // j < i in increasing index: the 27 cells in increasing id are the sorted x, y, z offsets in
            // lexicographic order
            int xs[3] = {wrap_index(cx - 1, side), cx, wrap_index(cx + 1, side)};
            int ys[3] = {wrap_index(cy - 1, side), cy, wrap_index(cy + 1, side)};
            int zs[3] = {wrap_index(cz - 1, side), cz, wrap_index(cz + 1, side)};
            sort3(xs[0], xs[1], xs[2]);
            sort3(ys[0], ys[1], ys[2]);
            sort3(zs[0], zs[1], zs[2]);
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2)
                    for (int d = 0; d < 3; ++d) {
                        const int nc = (xs[a] * side + ys[b2]) * side + zs[d];
                        if (nc > c) continue;
                        const uint32_t e = (nc == c) ? self : start[nc + 1];
                        for (uint32_t j = start[nc]; j < e; ++j) consider(j);
                    }
        } else {
            // grid too small for a distinct 27-cell stencil: every pair (NeighbourList.cpp:46-50)
            for (uint32_t j = self + 1; j < n; ++j) consider(j);
            hcnt = cnt;
            for (uint32_t j = 0; j < self; ++j) consider(j);
        }
        // stored lengths never exceed the padding, so readers stay inside the rows even after an overflow
        // (the true maximum goes to FLAG_MAXCOUNT and the overflow flag below)
        count[i] = cnt < padding ? cnt : padding;
        half[i] = hcnt < padding ? hcnt : padding;
    }
    block_build_stat<256>(cnt, static_cast<unsigned>(hcnt), stats + blockIdx.x);
}

__global__ void __launch_bounds__(256) k_displacement(const double4* __restrict__ pos,
                                                       const double4* __restrict__ ref, size_t n, double box,
                                                       double inv_box, double max_disp2, int* __restrict__ flags) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 r = pos[i], q = ref[i];
    const double dx = min_image(r.x - q.x, box, inv_box);
    const double dy = min_image(r.y - q.y, box, inv_box);
    const double dz = min_image(r.z - q.z, box, inv_box);
    if (dx * dx + dy * dy + dz * dz > max_disp2) flags[FLAG_STALE] = 1;
}

// Totals of a build from its per-block records: longest row, overflow, half-pair total, and in device-gated
// builds the rebuild counter. One block; each thread folds a strided share of the records.
constexpr int FINISH_THREADS = 1024;
__global__ void __launch_bounds__(FINISH_THREADS) k_finish_build(const int2* __restrict__ stats, int blocks,
                                                                 int padding, int* __restrict__ flags,
                                                                 unsigned long long* __restrict__ pair_total,
                                                                 const int* gate, const int* halt) {
    if (gated_off(gate, halt)) return;
    __shared__ int s_mx[FINISH_THREADS / WAVE];
    __shared__ unsigned long long s_hs[FINISH_THREADS / WAVE];
    int mx = 0;
    unsigned long long hs = 0;
    for (int b = threadIdx.x; b < blocks; b += FINISH_THREADS) {
        const int2 v = stats[b];
        mx = v.x > mx ? v.x : mx;
        hs += static_cast<unsigned>(v.y);
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(mx, off, WAVE);
        mx = o > mx ? o : mx;
        hs += __shfl_xor(hs, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0) {
        s_mx[threadIdx.x / WAVE] = mx;
        s_hs[threadIdx.x / WAVE] = hs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < FINISH_THREADS / WAVE; ++w) {
            mx = s_mx[w] > mx ? s_mx[w] : mx;
            hs += s_hs[w];
        }
        // longest row and overflow accumulate over the builds since the host last cleared them (md_run reads
        // them after its steps, and its kernels halt on the overflow word); the sticky pair records device-gated
        // builds only and survives host-driven ones (the adjoint reads it once per call); the pair total is this
        // build's. A device-gated build that overflowed is not counted: md_run repeats it with longer rows.
        if (mx > flags[FLAG_MAXCOUNT]) flags[FLAG_MAXCOUNT] = mx;
        if (gate && mx > flags[FLAG_MAXCOUNT_STICKY]) flags[FLAG_MAXCOUNT_STICKY] = mx;
        if (mx > padding) {
            flags[FLAG_OVERFLOW] = 1;
            if (gate) flags[FLAG_OVERFLOW_STICKY] = 1;
        } else if (gate) {
            flags[FLAG_REBUILDS] += 1;
        }
        *pair_total = hs;
    }
}

namespace {

int grid_side(double box, double min_cell) {
    const int s = static_cast<int>(box / min_cell);
    return s > 1 ? s : 1;
}

// Padded row length for a fresh list: 1.35x the ideal-gas full-list count plus slack, multiple of 8.
int auto_padding(size_t n, double box, double radius) {
    const double density = static_cast<double>(n) / (box * box * box);
    const double expected = 4.0 / 3.0 * M_PI * radius * radius * radius * density;
    long p = static_cast<long>(std::ceil(expected * 1.35 + 24.0));
    p = (p + 7) / 8 * 8;
    const long cap = n > 1 ? static_cast<long>(n - 1) : 1;
    return static_cast<int>(std::min<long>(p, std::max<long>(cap, 1)));
}

}  // namespace

// Scratch words kept per list next to the flags: the running half-pair total.
static unsigned long long* pair_total_of(md_list* l) {
    return reinterpret_cast<unsigned long long*>(l->flags + FLAG_COUNT);
}

void finish_build_enqueue(md_list* l, int blocks, const int* gate, const int* halt) {
    k_finish_build<<<1, FINISH_THREADS, 0, l->ctx->stream>>>(l->build_stats, blocks, l->padding, l->flags,
                                                             pair_total_of(l), gate, halt);
    MD_HIP(hipGetLastError());
}

void list_prepare(md_list* l, md_state* s) {
    hipStream_t st = l->ctx->stream;
    const int side = grid_side(s->box, l->radius);
    const size_t ncells = static_cast<size_t>(side) * side * side;
    if (!l->flags) {
        l->flags = dev_alloc<int>(FLAG_COUNT + 4);
        MD_HIP(hipMemsetAsync(l->flags, 0, sizeof(int) * (FLAG_COUNT + 4), st));
    }
    if (ncells + 1 > l->ncells_cap || !l->cell_count) {
        dev_free(l->cell_count);
        dev_free(l->cell_start);
        dev_free(l->scan_tmp);
        dev_free(l->scan_totals);
        l->cell_count = dev_alloc<uint32_t>(ncells + 1);
        l->cell_start = dev_alloc<uint32_t>(ncells + 1);
        l->scan_tmp = dev_alloc<uint32_t>(ncells + 1);  // per-cell fill counters of the scatter
        l->scan_totals = dev_alloc<uint32_t>(ncells / 2048 + 2);
        l->ncells_cap = ncells + 1;
        MD_HIP(hipMemsetAsync(l->cell_count, 0, sizeof(uint32_t) * (ncells + 1), st));
        MD_HIP(hipMemsetAsync(l->scan_tmp, 0, sizeof(uint32_t) * (ncells + 1), st));
    }
    l->side = side;
    l->ncells = ncells;
    if (s->n > l->n_cap || !l->cell_of) {
        dev_free(l->cell_of);
        dev_free(l->cell_of_alt);
        dev_free(l->order_raw);
        dev_free(l->order);
        dev_free(l->ref);
        dev_free(l->count);
        dev_free(l->half);
        l->cell_of = dev_alloc<int>(s->n);
        l->cell_of_alt = dev_alloc<int>(s->n);
        l->order_raw = dev_alloc<uint32_t>(s->n);
        l->order = dev_alloc<uint32_t>(s->n);
        l->ref = dev_alloc<double4>(s->n);
        l->count = dev_alloc<int32_t>(s->n);
        l->half = dev_alloc<int32_t>(s->n);
        l->n_cap = s->n;
    }
    if (l->padding <= 0) l->padding = auto_padding(s->n, s->box, l->radius);
    const size_t need = static_cast<size_t>(l->padding) * (s->n ? s->n : 1);
    if (need > l->rows_cap || !l->rows) {
        dev_free(l->rows);
        l->rows = dev_alloc<int32_t>(need);
        l->rows_cap = need;
    }
    const size_t need16 = tiled_rows16_entries(s->n, l->padding);
    if (need16 > l->rows16_cap || !l->rows16) {
        dev_free(l->rows16);
        l->rows16 = dev_alloc<uint16_t>(need16);
        l->rows16_cap = need16;
    }
    const size_t tiles = static_cast<size_t>(tile_count(s->n ? s->n : 1));
    const size_t stat_blocks = std::max<size_t>(tiles, blocks_for(s->n ? s->n : 1, FIXED_PER_BLOCK));
    if (stat_blocks > l->stats_cap || !l->build_stats) {
        dev_free(l->build_stats);
        l->build_stats = dev_alloc<int2>(stat_blocks);
        l->stats_cap = stat_blocks;
    }
    if (tiles > l->tiles_cap || !l->tile_nr) {
        dev_free(l->tile_ranges);
        dev_free(l->tile_nr);
        dev_free(l->tile_order);
        l->tile_ranges = dev_alloc<int4>(tiles * TILE_RANGES);
        l->tile_nr = dev_alloc<int>(tiles);
        l->tile_order = dev_alloc<int>(tiles);
        l->tiles_cap = tiles;
    }
    l->hmax = tile_halo_capacity(s->n, ncells);
    l->n = s->n;
}

// LDS room for a tile's halo: 1.5x the expected union of the 27-cell stencils of the cells a tile spans
// (x | y | z doubles, at most TILE_HALO_MAX particles), or the whole system when that is smaller
int tile_halo_capacity(size_t n, size_t ncells) {
    const double per_cell = static_cast<double>(n) / static_cast<double>(ncells);
    const double est = 1.5 * 9.0 * (TILE_PARTICLES / (per_cell > 0.5 ? per_cell : 0.5) + 3.0) * per_cell;
    long h = static_cast<long>(std::ceil(est / 8.0)) * 8;
    h = std::max<long>(256, std::min<long>(TILE_HALO_MAX - 8, h));  // the last slot is the build's ghost
    const long all = static_cast<long>((n + 7) / 8 * 8);
    if (static_cast<long>(n) <= TILE_HALO_MAX - 8) return static_cast<int>(all > 0 ? all : 8);
    return static_cast<int>(std::min<long>(h, all > 0 ? all : 8));
}

PermuteArgs permute_args(md_list* l, md_state* s, bool carry_forces) {
    PermuteArgs pa{};
    pa.enabled = PERMUTE_STATE;
    pa.pos = s->pos;
    pa.vel = s->vel;
    pa.frc = carry_forces ? s->frc : nullptr;
    pa.mass = s->mass;
    pa.id = s->id;
    pa.pos2 = s->pos_alt;
    pa.vel2 = s->vel_alt;
    pa.frc2 = carry_forces ? s->frc_alt : nullptr;
    pa.mass2 = s->mass_alt;
    pa.id2 = s->id_alt;
    pa.ref = l->ref;
    pa.cell_of2 = l->cell_of_alt;
    return pa;
}

// Histogram, scan, scatter and stable per-cell order of the current positions; with `permute` the state is
// also gathered into its alt arrays in that order (forces only with `carry_forces`: a rebuild that is followed by
// a force evaluation leaves them behind).
static void cells_enqueue(md_list* l, md_state* s, double per_length, bool permute, const int* gate,
                          bool carry_forces) {
    hipStream_t st = l->ctx->stream;
    const size_t n = s->n;
    if (n == 0) return;
    k_cell_count<<<blocks_for(n, 256), 256, 0, st>>>(s->pos, n, per_length, l->side, l->cell_of, l->cell_count,
                                                     gate, s->halt);
    {
        const int nb = static_cast<int>((l->ncells + SCAN_ITEMS - 1) / SCAN_ITEMS);
        k_scan_totals<<<nb, SCAN_THREADS, 0, st>>>(l->cell_count, l->ncells, l->scan_totals, gate, s->halt);
        k_scan_apply<<<nb, SCAN_THREADS, 0, st>>>(l->cell_count, l->ncells, l->scan_totals, l->cell_start, gate,
                                                  s->halt);
    }
    k_cell_scatter<<<blocks_for(n, 256), 256, 0, st>>>(l->cell_of, n, l->cell_start, l->scan_tmp, l->order_raw,
                                                       gate, s->halt);
    // the tiled build's tile setup gathers the state itself (one lane per destination particle), after this sort
    PermuteArgs pa{};
    if (permute) {
        pa = permute_args(l, s, carry_forces);
        if (tiling_enabled(n)) pa.enabled = PERMUTE_ORDER;
    }
    const unsigned sort_blocks = std::min<unsigned>(blocks_for(l->ncells * WAVE, 256), SORT_BLOCKS);
    k_cell_sort<<<sort_blocks, 256, 0, st>>>(l->cell_start, l->ncells, n, l->order_raw, l->order, l->scan_tmp,
                                             l->cell_count, pa, gate, s->halt);
    MD_HIP(hipGetLastError());
}

static void list_kernel_enqueue(md_list* l, md_state* s, bool from_alt, const int* gate, bool carry_forces) {
    hipStream_t st = l->ctx->stream;
    const size_t n = s->n;
    l->unstaged = -1;
    CopyBackArgs cb{};
    cb.enabled = from_alt ? 1 : 0;
    if (from_alt) {
        cb.vel2 = s->vel_alt;
        cb.frc2 = carry_forces ? s->frc_alt : nullptr;
        cb.mass2 = s->mass_alt;
        cb.id2 = s->id_alt;
        cb.cell_of2 = l->cell_of_alt;
        cb.pos = s->pos;
        cb.vel = s->vel;
        cb.frc = carry_forces ? s->frc : nullptr;
        cb.mass = s->mass;
        cb.id = s->id;
        cb.cell_of = l->cell_of;
    }
    if (tiling_enabled(n)) {
        tiled_list_enqueue(l, s, from_alt, gate, !l->slots_only, carry_forces);
        l->tiled = true;
        if (!l->slots_only) l->rows32_valid = true;
        return;
    }
    k_list_build<<<blocks_for(n ? n : 1, 256), 256, 0, st>>>(
        from_alt ? s->pos_alt : s->pos, n, from_alt ? l->cell_of_alt : l->cell_of, l->cell_start, l->side, s->box,
        s->inv_box, l->radius2, l->side >= 3 ? 1 : 0, l->padding, l->rows, l->count, l->half, l->build_stats, cb,
        gate, s->halt);
    MD_HIP(hipGetLastError());
    finish_build_enqueue(l, static_cast<int>(blocks_for(n ? n : 1, 256)), gate, s->halt);
    l->tiled = false;
    l->rows32_valid = true;
}

void ensure_rows32(md_list* l) {
    if (l->rows32_valid || !l->built) return;
    if (!l->tiled) fail(MD_ERR_STATE, "internal: int32 rows missing without slot rows");
    tiled_rows32_enqueue(l);
    l->rows32_valid = true;
}

void cells_build(md_list* l, md_state* s, double min_cell, int side) {
    (void)min_cell;
    cells_enqueue(l, s, side / s->box, false, nullptr, true);
}

// Clears the accumulated longest row and overflow flag before a host-driven build.
static void clear_build_flags(md_list* l) {
    MD_HIP(hipMemsetAsync(l->flags + FLAG_OVERFLOW, 0, 2 * sizeof(int), l->ctx->stream));
    static_assert(FLAG_MAXCOUNT == FLAG_OVERFLOW + 1, "the two words are cleared together");
}

void list_rebuild(md_list* l, md_state* s) {
    if (2.0 * l->radius > s->box) fail(MD_ERR_RUNTIME, "Cutoff + skin must be less than half the box length");
    list_prepare(l, s);
    hipStream_t st = l->ctx->stream;
    clear_build_flags(l);
    const double per_length = l->side / s->box;
    cells_enqueue(l, s, per_length, true, nullptr, true);
    list_kernel_enqueue(l, s, true, nullptr, true);
    int* hf = l->ctx->host_flags;
    for (int attempt = 0;; ++attempt) {
        MD_HIP(hipMemcpyAsync(hf, l->flags, sizeof(int) * FLAG_COUNT, hipMemcpyDeviceToHost, st));
        MD_HIP(hipMemcpyAsync(hf + FLAG_COUNT, pair_total_of(l), sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              st));
        MD_HIP(hipStreamSynchronize(st));
        l->max_count = hf[FLAG_MAXCOUNT];
        unsigned long long total;
        std::memcpy(&total, hf + FLAG_COUNT, sizeof(total));
        l->half_pairs = static_cast<size_t>(total);
        if (!hf[FLAG_OVERFLOW] || s->n == 0) break;
        if (attempt > 2) fail(MD_ERR_HIP, "neighbour list padding did not converge");
        // grow the rows and rebuild the list from the (already re-sorted) state
        l->padding = (l->max_count + l->max_count / 4 + 16 + 7) / 8 * 8;
        list_prepare(l, s);
        clear_build_flags(l);
        list_kernel_enqueue(l, s, false, nullptr, true);
    }
    MD_HIP(hipMemsetAsync(l->flags + FLAG_STALE, 0, sizeof(int), st));
    l->built = true;
    ++l->rebuilds;
}

void displacement_check(md_list* l, md_state* s) {
    hipStream_t st = l->ctx->stream;
    if (s->n == 0) return;
    k_displacement<<<blocks_for(s->n, 256), 256, 0, st>>>(s->pos, l->ref, s->n, s->box, s->inv_box, l->max_disp2,
                                                          l->flags);
    MD_HIP(hipGetLastError());
}

// Device-flag-gated rebuild used inside md_run: every kernel reads flags[FLAG_STALE] and returns at once when it
// is clear, so the same launch sequence serves steps with and without a rebuild.
void list_enqueue_conditional_rebuild(md_list* l, md_state* s, const int* gate) {
    const double per_length = l->side / s->box;
    cells_enqueue(l, s, per_length, true, gate, false);
    list_kernel_enqueue(l, s, true, gate, false);
}

void list_enqueue_conditional_rebuild(md_list* l, md_state* s) {
    list_enqueue_conditional_rebuild(l, s, l->flags + FLAG_STALE);
}

// Fixed-order list (no re-sort): candidates are the members of the 27 cells taken through the stable cell order,
// so particle indices never change. Each row holds the j > i neighbours in increasing j, then the j < i ones in
// increasing j, so a particle's force, summed along its row, depends only on the pairs within the cut-off (the
// other entries add exact zeros) and not on when the list was built: the adjoint's forward run and its segment
// recomputations are the same trajectory whatever the checkpoint spacing. Also records the build-time
// positions. FIXED_LANES lanes per particle make one pass over the 27 cells in stencil order; the accepted
// neighbours of each part are gathered in LDS over the ballot, then each is written at its rank within its part
// (the number of smaller indices there).
__global__ void __launch_bounds__(FIXED_THREADS) k_list_build_fixed(const double4* __restrict__ pos, size_t n,
                                                                    const int* __restrict__ cell_of,
                                                                    const uint32_t* __restrict__ start,
                                                                    const uint32_t* __restrict__ order, int side,
                                                                    double box, double inv_box, double radius2,
                                                                    int use_cells, int padding,
                                                                    int32_t* __restrict__ rows,
                                                                    int32_t* __restrict__ count,
                                                                    int32_t* __restrict__ half,
                                                                    int2* __restrict__ stats,
                                                                    double4* __restrict__ ref, const int* gate,
                                                                    const int* halt) {
    if (gated_off(gate, halt)) return;
    extern __shared__ int32_t part_rows[];  // per particle of the block: `padding` j > i, then `padding` j < i
    const int g = threadIdx.x / FIXED_LANES, lane = threadIdx.x % FIXED_LANES;
    const size_t i = static_cast<size_t>(blockIdx.x) * FIXED_PER_BLOCK + g;
    const unsigned shift = (threadIdx.x % WAVE) & ~(FIXED_LANES - 1);  // the group's first lane in the wavefront
    const unsigned below = (1u << lane) - 1u;
    int32_t* const upper = part_rows + static_cast<size_t>(g) * 2 * padding;
    int32_t* const lower = upper + padding;
    int cnt = 0, lcnt = 0;  // uniform over the group
    if (i < n) {
        const double4 ri = pos[i];
        if (lane == 0) ref[i] = ri;
        const uint32_t self = static_cast<uint32_t>(i);
        // one candidate per lane (valid lanes only), classified and gathered
        auto chunk = [&](uint32_t j, bool valid) {
            bool near = false;
            if (valid && j != self) {
                const double4 rj = pos[j];
                const double dx = min_image(ri.x - rj.x, box, inv_box);
                const double dy = min_image(ri.y - rj.y, box, inv_box);
                const double dz = min_image(ri.z - rj.z, box, inv_box);
                near = dx * dx + dy * dy + dz * dz < radius2;
            }
            const bool up = near && j > self, lo = near && j < self;
            const unsigned mu = static_cast<unsigned>(__ballot(up) >> shift) & 0xFFFFu;
            const unsigned ml = static_cast<unsigned>(__ballot(lo) >> shift) & 0xFFFFu;
            if (up) {
                const int at = cnt + __popc(mu & below);
                if (at < padding) upper[at] = static_cast<int32_t>(j);
            }
            if (lo) {
                const int at = lcnt + __popc(ml & below);
                if (at < padding) lower[at] = static_cast<int32_t>(j);
            }
            cnt += __popc(mu);
            lcnt += __popc(ml);
        };
        if (use_cells) {
            const int c = cell_of[i];
            const int cx = c / (side * side), cy = (c / side) % side, cz = c % side;
            for (int ox = -1; ox <= 1; ++ox)
                for (int oy = -1; oy <= 1; ++oy)
                    for (int oz = -1; oz <= 1; ++oz) {
                        const int nc = (wrap_index(cx + ox, side) * side + wrap_index(cy + oy, side)) * side +
                                       wrap_index(cz + oz, side);
                        const uint32_t e = start[nc + 1];
                        for (uint32_t k0 = start[nc]; k0 < e; k0 += FIXED_LANES) {
                            const uint32_t k = k0 + static_cast<uint32_t>(lane);
                            chunk(k < e ? order[k] : 0u, k < e);
                        }
                    }
        } else {
            for (uint32_t j0 = 0; j0 < n; j0 += FIXED_LANES) {
                const uint32_t j = j0 + static_cast<uint32_t>(lane);
                chunk(j, j < n);
            }
        }
    }
    __syncthreads();  // the group's LDS parts are complete (written by its own lanes)
    const int total = cnt + lcnt;
    if (i < n) {
        // each stored entry goes to its rank among the stored entries of its part (indices are distinct)
        const int su = cnt < padding ? cnt : padding, sl = lcnt < padding ? lcnt : padding;
        for (int e = lane; e < su + sl; e += FIXED_LANES) {
            const bool is_up = e < su;
            const int32_t* part = is_up ? upper : lower;
            const int m = is_up ? su : sl;
            const int32_t j = part[is_up ? e : e - su];
            int rank = 0;
            for (int q = 0; q < m; ++q) rank += part[q] < j ? 1 : 0;
            const int at = (is_up ? 0 : cnt) + rank;
            if (at < padding) rows[static_cast<size_t>(at) * n + i] = j;
        }
        if (lane == 0) {
            count[i] = total < padding ? total : padding;
            half[i] = cnt < padding ? cnt : padding;
        }
    }
    block_build_stat<FIXED_THREADS>(lane == 0 ? total : 0, lane == 0 ? static_cast<unsigned>(cnt) : 0u,
                                    stats + blockIdx.x);
}

static void list_fixed_enqueue(md_list* l, md_state* s, const int* gate) {
    hipStream_t st = l->ctx->stream;
    const size_t n = s->n;
    cells_enqueue(l, s, l->side / s->box, false, gate, true);
    l->tiled = false;
    l->rows32_valid = true;
    // each block keeps both parts of FIXED_PER_BLOCK rows in LDS (64 KB at most)
    const size_t lds = static_cast<size_t>(FIXED_PER_BLOCK) * 2 * l->padding * sizeof(int32_t);
    if (lds > 65536) fail(MD_ERR_STATE, "fixed-order neighbour list: rows longer than 512 entries are not supported");
    const unsigned blocks = blocks_for(n ? n : 1, FIXED_PER_BLOCK);
    k_list_build_fixed<<<blocks, FIXED_THREADS, lds, st>>>(
        s->pos, n, l->cell_of, l->cell_start, l->order, l->side, s->box, s->inv_box, l->radius2,
        l->side >= 3 ? 1 : 0, l->padding, l->rows, l->count, l->half, l->build_stats, l->ref, gate, s->halt);
    MD_HIP(hipGetLastError());
    finish_build_enqueue(l, static_cast<int>(blocks), gate, s->halt);
}

void list_rebuild_fixed(md_list* l, md_state* s) {
    if (2.0 * l->radius > s->box) fail(MD_ERR_RUNTIME, "Cutoff + skin must be less than half the box length");
    list_prepare(l, s);
    hipStream_t st = l->ctx->stream;
    int* hf = l->ctx->host_flags;
    for (int attempt = 0;; ++attempt) {
        clear_build_flags(l);
        list_fixed_enqueue(l, s, nullptr);
        MD_HIP(hipMemcpyAsync(hf, l->flags, sizeof(int) * (FLAG_COUNT + 2), hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
        l->max_count = hf[FLAG_MAXCOUNT];
        unsigned long long total;
        std::memcpy(&total, hf + FLAG_COUNT, sizeof(total));
        l->half_pairs = static_cast<size_t>(total);
        if (!hf[FLAG_OVERFLOW] || s->n == 0) break;
        if (attempt > 2) fail(MD_ERR_HIP, "neighbour list padding did not converge");
        l->padding = (l->max_count + l->max_count / 4 + 16 + 7) / 8 * 8;
        list_prepare(l, s);
    }
    MD_HIP(hipMemsetAsync(l->flags + FLAG_STALE, 0, sizeof(int), st));
    l->built = true;
    ++l->rebuilds;
}

void list_enqueue_conditional_rebuild_fixed(md_list* l, md_state* s) {
    list_fixed_enqueue(l, s, l->flags + FLAG_STALE);
}

// The device buffers of a list, each applied to f.
template <class F>
static void for_list_buffers(md_list* l, F f) {
    f(l->cell_count);
    f(l->cell_start);
    f(l->scan_tmp);
    f(l->cell_of);
    f(l->cell_of_alt);
    f(l->order_raw);
    f(l->order);
    f(l->ref);
    f(l->count);
    f(l->half);
    f(l->rows);
    f(l->flags);
    f(l->scan_totals);
    f(l->rows16);
    f(l->tile_ranges);
    f(l->tile_nr);
    f(l->tile_order);
    f(l->build_stats);
}

static void release_list_buffers(md_list* l) {
    for_list_buffers(l, [](auto*& p) { dev_free(p); });
    l->rows16_cap = l->tiles_cap = l->stats_cap = 0;
}


void set_error(const std::string& m);

}  // namespace mdgpu

using namespace mdgpu;


extern "C" {

int md_cells_build(md_state* s, double min_cell, int* cells_per_side, int* cell_of, uint64_t* cell_start,
                   uint64_t* order) {
    MD_API_BEGIN
    if (!s) fail(MD_ERR_ARG, "state is NULL");
    if (!(min_cell > 0.0)) fail(MD_ERR_ARG, "min_cell must be positive");
    // a private grid object: the standalone CellList of the reference owns its own arrays
    md_list tmp;
    tmp.ctx = s->ctx;
    tmp.radius = min_cell;
    tmp.padding = 1;
    const size_t n = s->n;
    try {
        list_prepare(&tmp, s);
        const int side = tmp.side;
        cells_build(&tmp, s, min_cell, side);
        hipStream_t st = s->ctx->stream;
        std::vector<int> co(n);
        std::vector<uint32_t> cs(tmp.ncells + 1), od(n);
        if (n) {
            MD_HIP(hipMemcpyAsync(co.data(), tmp.cell_of, sizeof(int) * n, hipMemcpyDeviceToHost, st));
            MD_HIP(hipMemcpyAsync(od.data(), tmp.order, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
            MD_HIP(hipMemcpyAsync(cs.data(), tmp.cell_start, sizeof(uint32_t) * (tmp.ncells + 1),
                                  hipMemcpyDeviceToHost, st));
        }
        MD_HIP(hipStreamSynchronize(st));
        if (cells_per_side) *cells_per_side = side;
        if (cell_of) std::copy(co.begin(), co.end(), cell_of);
        if (order)
            for (size_t i = 0; i < n; ++i) order[i] = od[i];
        if (cell_start) {
            if (n)
                for (size_t c = 0; c <= tmp.ncells; ++c) cell_start[c] = cs[c];
            else
                for (size_t c = 0; c <= tmp.ncells; ++c) cell_start[c] = 0;
        }
    } catch (...) {
        release_list_buffers(&tmp);
        throw;
    }
    release_list_buffers(&tmp);
    MD_API_END
}

int md_list_create(md_ctx* ctx, double cutoff, double skin, int max_neighbours, md_list** out) {
    MD_API_BEGIN
    if (!ctx || !out) fail(MD_ERR_ARG, "ctx/out is NULL");
    auto* l = new md_list;
    l->ctx = ctx;
    ++ctx->live;
    l->cutoff = cutoff;
    l->skin = skin;
    l->radius = cutoff + skin;
    l->radius2 = (cutoff + skin) * (cutoff + skin);
    l->max_disp2 = 0.25 * skin * skin;
    l->padding = max_neighbours > 0 ? max_neighbours : 0;
    *out = l;
    MD_API_END
}

int md_list_destroy(md_list* l) {
    if (!l) return MD_OK;
    if (l->ctx) {
        (void)hipStreamSynchronize(l->ctx->stream);
        --l->ctx->live;
    }
    release_list_buffers(l);
    delete l;
    return MD_OK;
}

int md_list_copy(md_list* dst, const md_list* src) {
    MD_API_BEGIN
    if (!dst || !src) fail(MD_ERR_ARG, "list is NULL");
    if (dst == src) return MD_OK;
    if (dst->ctx != src->ctx) fail(MD_ERR_ARG, "both lists must belong to one context");
    hipStream_t st = src->ctx->stream;
    MD_HIP(hipStreamSynchronize(st));
    release_list_buffers(dst);
    // every scalar member, then each device buffer at the source's capacity
    md_ctx* const ctx = dst->ctx;
    md_list copy = *src;
    copy.ctx = ctx;
    // no buffer of src may be reached through copy: a failed clone releases copy's buffers
    for_list_buffers(&copy, [](auto*& p) { p = nullptr; });
    auto clone = [&](auto*& out, auto* in, size_t count) {
        using T = std::remove_const_t<std::remove_pointer_t<std::remove_reference_t<decltype(out)>>>;
        out = nullptr;
        if (!in) return;
        out = dev_alloc<T>(count);
        MD_HIP(hipMemcpyAsync(out, in, sizeof(T) * count, hipMemcpyDeviceToDevice, st));
    };
    const size_t cells = src->ncells_cap, nc = src->n_cap ? src->n_cap : 1;
    try {
        clone(copy.cell_count, src->cell_count, cells);
        clone(copy.cell_start, src->cell_start, cells);
        clone(copy.scan_tmp, src->scan_tmp, cells);
        clone(copy.scan_totals, src->scan_totals, (cells ? cells - 1 : 0) / 2048 + 2);  // as list_prepare sizes it
        clone(copy.cell_of, src->cell_of, nc);
        clone(copy.cell_of_alt, src->cell_of_alt, nc);
        clone(copy.order_raw, src->order_raw, nc);
        clone(copy.order, src->order, nc);
        clone(copy.ref, src->ref, nc);
        clone(copy.count, src->count, nc);
        clone(copy.half, src->half, nc);
        clone(copy.rows, src->rows, src->rows_cap);
        clone(copy.flags, src->flags, FLAG_COUNT + 4);
        clone(copy.build_stats, src->build_stats, src->stats_cap);
        clone(copy.rows16, src->rows16, src->rows16_cap);
        clone(copy.tile_ranges, src->tile_ranges, src->tiles_cap * TILE_RANGES);
        clone(copy.tile_nr, src->tile_nr, src->tiles_cap);
        clone(copy.tile_order, src->tile_order, src->tiles_cap);
        MD_HIP(hipStreamSynchronize(st));
    } catch (...) {
        release_list_buffers(&copy);
        throw;
    }
    *dst = copy;
    MD_API_END
}

int md_list_build(md_list* l, md_state* s) {
    MD_API_BEGIN
    if (!l || !s) fail(MD_ERR_ARG, "list/state is NULL");
    list_rebuild(l, s);
    MD_API_END
}

int md_list_update(md_list* l, md_state* s, int* rebuilt) {
    MD_API_BEGIN
    if (!l || !s) fail(MD_ERR_ARG, "list/state is NULL");
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
    if (rebuilt) *rebuilt = stale ? 1 : 0;
    MD_API_END
}

int md_list_get_info(const md_list* l, md_list_info* info) {
    MD_API_BEGIN
    if (!l || !info) fail(MD_ERR_ARG, "list/info is NULL");
    info->half_pairs = l->half_pairs;
    info->rebuilds = l->rebuilds;
    info->cells_per_side = l->side;
    info->max_neighbours = l->max_count;
    info->padding = l->padding;
    info->particles = l->built ? l->n : 0;
    info->tiled = l->built && l->tiled ? 1 : 0;
    info->tiles = 0;
    info->unstaged_tiles = 0;
    if (info->tiled && l->n > 0) {
        const int tiles = tile_count(l->n);
        info->tiles = tiles;
        if (l->unstaged < 0) {
            std::vector<int> nr(static_cast<size_t>(tiles));
            hipStream_t st = l->ctx->stream;
            MD_HIP(hipMemcpyAsync(nr.data(), l->tile_nr, sizeof(int) * nr.size(), hipMemcpyDeviceToHost, st));
            MD_HIP(hipStreamSynchronize(st));
            l->unstaged = static_cast<int>(std::count(nr.begin(), nr.end(), -1));
        }
        info->unstaged_tiles = l->unstaged;
    }
    MD_API_END
}

int md_list_half(md_list* l, uint64_t* offsets, uint64_t* neighbours) {
    MD_API_BEGIN
    if (!l) fail(MD_ERR_ARG, "list is NULL");
    if (!l->built) fail(MD_ERR_STATE, "list not built");
    const size_t n = l->n;
    hipStream_t st = l->ctx->stream;
    std::vector<int32_t> half(n), rows(neighbours ? static_cast<size_t>(l->padding) * n : 0);
    if (n) MD_HIP(hipMemcpyAsync(half.data(), l->half, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    if (neighbours && n) {
        ensure_rows32(l);
        MD_HIP(hipMemcpyAsync(rows.data(), l->rows, sizeof(int32_t) * rows.size(), hipMemcpyDeviceToHost, st));
    }
    MD_HIP(hipStreamSynchronize(st));
    uint64_t at = 0;
    for (size_t i = 0; i < n; ++i) {
        if (offsets) offsets[i] = at;
        if (neighbours)
            for (int k = 0; k < half[i]; ++k) neighbours[at + k] = static_cast<uint64_t>(rows[static_cast<size_t>(k) * n + i]);
        at += static_cast<uint64_t>(half[i]);
    }
    if (offsets) offsets[n] = at;
    MD_API_END
}

int md_list_full(md_list* l, int32_t* counts, int32_t* half_counts, int32_t* rows_out) {
    MD_API_BEGIN
    if (!l) fail(MD_ERR_ARG, "list is NULL");
    if (!l->built) fail(MD_ERR_STATE, "list not built");
    const size_t n = l->n;
    hipStream_t st = l->ctx->stream;
    if (counts && n) MD_HIP(hipMemcpyAsync(counts, l->count, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    if (half_counts && n) MD_HIP(hipMemcpyAsync(half_counts, l->half, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    std::vector<int32_t> rows(rows_out ? static_cast<size_t>(l->padding) * n : 0);
    if (rows_out && n) {
        ensure_rows32(l);
        MD_HIP(hipMemcpyAsync(rows.data(), l->rows, sizeof(int32_t) * rows.size(), hipMemcpyDeviceToHost, st));
    }
    MD_HIP(hipStreamSynchronize(st));
    if (rows_out)
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < l->padding; ++k)
                rows_out[i * l->padding + k] = rows[static_cast<size_t>(k) * n + i];
    MD_API_END
}

}  // extern "C"
