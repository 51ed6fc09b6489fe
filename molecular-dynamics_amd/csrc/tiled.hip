// Tiled neighbour list and mixed-precision force kernel: the list build and the force evaluation stage each
// tile's halo positions in LDS once, instead of gathering every neighbour from L2/MALL.
//
// A tile is TILE = 128 consecutive particles of the cell-sorted state, worked by 512 threads. Its halo is the union
// of the 27-cell stencils of the cells it spans, kept as at most RMAX ranges of whole cells (each a contiguous run
// of particle indices, since particles are sorted by cell). At rho* = 0.8442 and rc + skin = 2.8 a halo holds ~1700
// particles (~2300 when the tile crosses a cell column); the capacity `hmax` (<= TILE_HALO_MAX - 8, x | y | z
// doubles) is set from the density, and a tile that does not fit is marked tile_nr = -1 and handled through global
// gathers.
//
// A rebuild runs two kernels: k_tile_setup (one wavefront per tile) finds the halo ranges and their LDS slots (and,
// after a re-sort, gathers the tile's particles into the new order);
// k_list_build_tiled stages the halo, builds per own cell two flat candidate lists (the stencil cells above and
// below it) and walks them one wavefront per particle, 64 candidates per step, ranking the accepted ones over the
// ballot so that rows keep the reference's order (j > i in 27-cell stencil order, then j < i increasing). It writes
// each row as LDS slot byte offsets of the tile's halo (uint16, tile-major: each tile owns row_len / 32 chunks of
// 128 particles x 32 entries; entry k of the tile's particle p sits at chunk k / 32, offset 32 p + 8 (k % 4) +
// (k / 4) % 8, so each of the force kernel's 4 lanes fetches its 8 entries k = lane, lane + 4, ... of a chunk with
// one 16-byte load, a wavefront's 16 particles read 1 KB contiguously, and chunks are a constant 4096 entries
// apart) and, when a caller needs them, as global indices (int32, column-major: what the reference's half-list
// view, the RDF, the adjoint and the FP64 force kernel read). k_forces_tiled_mixed walks the slot rows; its STEP
// form also carries the velocity-Verlet step of the tile's particles.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "device.hpp"
#include "internal.hpp"
#include "pair.hpp"

namespace mdgpu {

constexpr int TILE = TILE_PARTICLES;
constexpr int LANES = 4;
constexpr int TILE_THREADS = TILE * LANES;
constexpr int RMAX = TILE_RANGES;
constexpr int CR_CAP = 192;  // cell ranges before merging (9 neighbour columns x 2 wrap pieces x segments)
constexpr int CMAX = 24;     // own cells of a tile covered by the stencil table
constexpr int CHUNK = 32;    // slot-row entries per particle and chunk
constexpr int CHUNK_STRIDE = TILE * CHUNK;  // entries between a tile's chunks
constexpr int GHOST_SLOT = TILE_HALO_MAX - 1;  // LDS slot past every halo: never a neighbour
// The mixed force kernel's ghost slot coordinate: a frame coordinate is within half a box of 0, and boxes are far
// shorter than 1e6 (a 1e6 box at liquid density would hold ~1e18 particles), so every separation from the ghost is
// finite and beyond any cut-off. (The list build's ghost holds NaN or infinity: its tests reject them.)
constexpr double GHOST_COORD = 1.0e6;
constexpr int PRE_CHUNKS = 3;  // slot chunks the force kernel loads per lane ahead of its pair loop
// The slot rows and the list build's candidate lists hold byte offsets of LDS slots (slot * 8, below 2^16), so that
// the force kernel adds an entry to a fixed halo address folded into its LDS loads.
constexpr int SLOT_BYTES = 8;
static_assert(TILE_HALO_MAX * SLOT_BYTES <= 0x10000, "slot byte offsets fit 16 bits");


// Position of a particle's first slot-row entry, and of entry k relative to it.
__device__ __forceinline__ size_t slot_row_base(size_t i, int row_len) {
    return (i / TILE) * static_cast<size_t>(TILE) * row_len + (i % TILE) * CHUNK;
}
__device__ __forceinline__ uint32_t slot_entry(uint32_t k) {
    return ((k & ~31u) << 7) | ((k & 3u) << 3) | ((k >> 2) & 7u);  // (k / 32) * 4096 + 8 (k % 4) + (k / 4) % 8
}
constexpr int SETUP_TILES = 4;
// tile_nr[t] holds the tile's number of halo ranges (-1: not staged), plus TILE_FLAT when no periodic image can
// enter the tile's frame: its cells lie in one cell column at least two cells from every face of the grid, so its
// halo cells are at least one cell from every face (a particle moves less than a cell between builds, so none
// has wrapped), and the tile and its halo span less than half the box. Then t - L rint(t / L) = t for every
// staged coordinate t = x - ref, and staging skips the image term.
constexpr int TILE_FLAT = 1 << 16;
// The same per axis (TILE_FLAT_X << a): the tile's cells lie at least two cells from both faces normal to axis a
// and span, with their stencil, less than half the box along it, so no periodic image enters that component of
// any separation (the FP64 kernel skips the minimum image along it). TILE_FLAT is set when all three are.
constexpr int TILE_FLAT_X = 1 << 17;
__device__ __forceinline__ int tile_range_count(int raw) { return raw < 0 ? -1 : (raw & (TILE_FLAT - 1)); }
__device__ __forceinline__ bool tile_is_flat(int raw) { return raw > 0 && (raw & TILE_FLAT) != 0; }
// bit a set: the minimum image is needed along axis a
__device__ __forceinline__ int tile_wrap_axes(int raw) { return raw > 0 ? (~raw >> 17) & 7 : 7; }
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform_f64(double v) {
    return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                            __builtin_amdgcn_readfirstlane(__double2loint(v)));
}

// BATCH: the tile's replica from the tables (the same for every lane, copied into scalar registers); otherwise the
// launch's one system, straight from the kernel arguments.
template <bool BATCH>
__device__ __forceinline__ TileSystem tile_system(const TileSystems& ts, int tile) {
    if constexpr (!BATCH) {
        return ts.one;
    } else {
        const TileSystem& p = ts.per[uniform_u32(ts.rep_of_tile[tile])];
        TileSystem r;
        r.first = uniform_u32(p.first);
        r.end = uniform_u32(p.end);
        r.cell_base = uniform_u32(p.cell_base);
        r.side = uniform_i32(p.side);
        r.use_cells = uniform_i32(p.use_cells);
        r.hmax = uniform_i32(p.hmax);
        r.interior_ok = uniform_i32(p.interior_ok);
        r.gate = uniform_i32(p.gate);
        r.box = uniform_f64(p.box);
        r.inv_box = uniform_f64(p.inv_box);
        r.flat_cells = uniform_f64(p.flat_cells);
        r.reach = uniform_f64(p.reach);
        return r;
    }
}
// A rebuild kernel's tile whose replica is not being rebuilt (never for the single system).
template <bool BATCH>
__device__ __forceinline__ bool tile_gated_off(const TileSystems& ts, const TileSystem& sys) {
    if constexpr (!BATCH) return false;
    else return ts.gates[sys.gate] == 0;
}

// Launch order of a single system's tiles: cell columns (x, y) in bands of ORDER_BAND x-slabs; within a band the y
// rows, each row's ORDER_BAND columns in x, each column's tiles in z order. A tile's halo spans its own column and
// the eight around it. In cell order the columns at x +- 1 lie a whole slab (side^2 cells, ~500 tiles at 4 M atoms)
// away, so an XCD's L2 has evicted their positions before the next slab stages them again; in band order they
// return within a few columns. The blocks of a launch take tiles order[xcd_item(block)].
constexpr int ORDER_BAND = 4;
__device__ __forceinline__ int launch_tile(const TileSystems& ts, int block) {
    const int k = xcd_item(block, ts.tiles);
    return ts.order ? uniform_i32(ts.order[k]) : k;
}

static_assert(RMAX == WAVE, "one lane per halo range in k_tile_setup");
static_assert(CHUNK_STRIDE == 4096, "slot_entry shifts the chunk index by 7 + 5 bits");
static_assert(TILE == 2 * WAVE, "the force kernel's energy sum: one wavefront adds two particles per lane");

// Three cell coordinates in increasing order, indexed without a private array (which would live in scratch).
struct Sorted3 {
    int v0, v1, v2;
    __device__ Sorted3(int a, int b, int c) {
        const int lo = min(a, min(b, c)), hi = max(a, max(b, c));
        v0 = lo;
        v1 = a + b + c - lo - hi;
        v2 = hi;
    }
    __device__ int operator[](int k) const { return k == 0 ? v0 : (k == 1 ? v1 : v2); }
};

__device__ __forceinline__ void push_range(int2* cr, int& ncr, int a, int b) {
    if (ncr < CR_CAP) cr[ncr] = make_int2(a, b);
    ++ncr;
}

// Cell ranges [a, b] (inclusive) covering the stencils of cells c_first..c_last, sorted and merged. Returns the
// number of merged ranges, or -1 if they do not fit.
__device__ int tile_cell_ranges(int c_first, int c_last, int s, int2* cr) {
    int ncr = 0;
    int c = c_first;
    while (c <= c_last) {
        const int x = c / (s * s), y = (c / s) % s, z0 = c % s;
        const int col_end = (x * s + y) * s + (s - 1);
        const int seg_end = c_last < col_end ? c_last : col_end;
        const int z1 = z0 + (seg_end - c);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy) {
                const int base = (wrap_index(x + dx, s) * s + wrap_index(y + dy, s)) * s;
                const int za = z0 - 1, zb = z1 + 1;
                if (zb - za + 1 >= s) {
                    push_range(cr, ncr, base, base + s - 1);
                } else if (za < 0) {
                    push_range(cr, ncr, base + za + s, base + s - 1);
                    push_range(cr, ncr, base, base + zb);
                } else if (zb >= s) {
                    push_range(cr, ncr, base + za, base + s - 1);
                    push_range(cr, ncr, base, base + zb - s);
                } else {
                    push_range(cr, ncr, base + za, base + zb);
                }
            }
        c = seg_end + 1;
    }
    if (ncr > CR_CAP) return -1;
    for (int k = 1; k < ncr; ++k) {  // insertion sort by first cell
        const int2 v = cr[k];
        int m = k - 1;
        while (m >= 0 && cr[m].x > v.x) {
            cr[m + 1] = cr[m];
            --m;
        }
        cr[m + 1] = v;
    }
    int out = 0;  // merge overlapping or adjacent ranges
    for (int k = 0; k < ncr; ++k) {
        if (out > 0 && cr[k].x <= cr[out - 1].y + 1) {
            if (cr[k].y > cr[out - 1].y) cr[out - 1].y = cr[k].y;
        } else {
            cr[out++] = cr[k];
        }
    }
    return out <= RMAX ? out : -1;
}

// The same ranges computed by the whole wavefront, for tiles whose cells span at most three cell columns (every
// lane calls it, with the same arguments): lane l < 9 nseg takes segment l / 9 and neighbour column l % 9, the one
// or two cell runs of each lane are compacted into cr, ranked by first cell, and runs that overlap or touch are
// merged (a run starts a merged range when its first cell lies past the prefix maximum of the earlier ends + 1).
__device__ __forceinline__ int tile_cell_ranges_wave(int c_first, int c_last, int s, int2* cr, int lane) {
    const int col0 = c_first / s, col1 = c_last / s;
    const int nseg = col1 - col0 + 1;
    int a0 = 0, b0 = -1, a1 = 0, b1 = -1;  // b < a: no run
    if (lane < 9 * nseg) {
        const int k = lane / 9, nb = lane % 9;
        const int col = col0 + k;
        const int x = col / s, y = col % s;
        const int z0 = k == 0 ? c_first % s : 0;
        const int z1 = col == col1 ? c_last % s : s - 1;
        const int base = (wrap_index(x + nb / 3 - 1, s) * s + wrap_index(y + nb % 3 - 1, s)) * s;
        const int za = z0 - 1, zb = z1 + 1;
        if (zb - za + 1 >= s) {
            a0 = base;
            b0 = base + s - 1;
        } else if (za < 0) {
            a0 = base + za + s;
            b0 = base + s - 1;
            a1 = base;
            b1 = base + zb;
        } else if (zb >= s) {
            a0 = base + za;
            b0 = base + s - 1;
            a1 = base;
            b1 = base + zb - s;
        } else {
            a0 = base + za;
            b0 = base + zb;
        }
    }
    const int cnt = (b0 >= a0 ? 1 : 0) + (b1 >= a1 ? 1 : 0);
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int o = __shfl_up(incl, d, WAVE);
        if (lane >= d) incl += o;
    }
    const int total = __shfl(incl, WAVE - 1, WAVE);  // at most 2 x 9 x 3 = 54 runs
    if (cnt > 0) cr[incl - cnt] = make_int2(a0, b0);
    if (cnt > 1) cr[incl - cnt + 1] = make_int2(a1, b1);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // rank of run `lane` by (first cell, position)
    int2 mine = make_int2(0, -1);
    int rank = 0;
    if (lane < total) {
        mine = cr[lane];
        for (int j = 0; j < total; ++j) {
            const int aj = cr[j].x;
            rank += (aj < mine.x || (aj == mine.x && j < lane)) ? 1 : 0;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < total) cr[rank] = mine;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const bool in = lane < total;
    const int2 r = in ? cr[lane] : make_int2(0x7fffffff, -1);
    int pmax = r.y;  // inclusive prefix maximum of the run ends
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int o = __shfl_up(pmax, d, WAVE);
        if (lane >= d && o > pmax) pmax = o;
    }
    int before = __shfl_up(pmax, 1, WAVE);
    if (lane == 0) before = -2;
    const bool head = in && r.x > before + 1;
    const unsigned long long heads = __ballot(head);
    const int nr = __popcll(heads);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (in && nr <= RMAX) {
        const int g = static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(heads >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(heads), 0u))) -
                      (head ? 0 : 1);
        const bool last = lane == total - 1 || ((heads >> (lane + 1)) & 1ull) != 0;
        if (head) cr[g].x = r.x;
        if (last) cr[g].y = pmax;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return nr <= RMAX ? nr : -1;
}

// A tile's share of the re-sort, TILE / WAVE particles per lane: every load issued before any store (the state's
// arrays are not marked as unaliased, so a store would otherwise wait for nothing and order the next loads).
constexpr int GATHER_PER_LANE = TILE / WAVE;
struct GatheredTile {  // plain doubles: a private array of double4 is not kept in registers
    double r[GATHER_PER_LANE][4];
    double v[GATHER_PER_LANE][3], f[GATHER_PER_LANE][3], m[GATHER_PER_LANE];
    uint64_t id[GATHER_PER_LANE];
};

__device__ __forceinline__ void load_gathered(const PermuteArgs& pa, const uint32_t* __restrict__ order,
                                              GatheredTile& g, int t, int lane, size_t n) {
#pragma unroll
    for (int u = 0; u < GATHER_PER_LANE; ++u) {
        const size_t k = static_cast<size_t>(t) * TILE + u * WAVE + lane;
        if (k >= n) continue;  // single system only: n is both the bound and the component stride
        const size_t v = order[k];
        const double4 p = pa.pos[v];
        g.r[u][0] = p.x;
        g.r[u][1] = p.y;
        g.r[u][2] = p.z;
        g.r[u][3] = p.w;
#pragma unroll
        for (int c = 0; c < 3; ++c) g.v[u][c] = pa.vel[c * n + v];
        if (pa.frc)
#pragma unroll
            for (int c = 0; c < 3; ++c) g.f[u][c] = pa.frc[c * n + v];
        g.m[u] = pa.mass[v];
        g.id[u] = pa.id[v];
    }
}

__device__ __forceinline__ void store_gathered(const PermuteArgs& pa, const GatheredTile& g, int t, int lane,
                                               size_t n) {
#pragma unroll
    for (int u = 0; u < GATHER_PER_LANE; ++u) {
        const size_t k = static_cast<size_t>(t) * TILE + u * WAVE + lane;
        if (k >= n) continue;
        const double4 p = make_double4(g.r[u][0], g.r[u][1], g.r[u][2], g.r[u][3]);
        if (pa.pos2) pa.pos2[k] = p;
        pa.ref[k] = p;
#pragma unroll
        for (int c = 0; c < 3; ++c) pa.vel2[c * n + k] = g.v[u][c];
        if (pa.frc)
#pragma unroll
            for (int c = 0; c < 3; ++c) pa.frc2[c * n + k] = g.f[u][c];
        pa.mass2[k] = g.m[u];
        pa.id2[k] = g.id[u];
    }
}

// Halo ranges of every tile: one wavefront per tile. The wavefront merges the cell ranges (tiles spanning more than
// three cell columns, on small grids: lane 0 alone), then lane r reads the particle span of range r and a
// wavefront scan assigns the LDS slots. tile_nr = -1 marks a tile whose halo does not fit in `hmax` slots.
template <bool BATCH>
__global__ void __launch_bounds__(SETUP_TILES * WAVE)
    k_tile_setup(const int* __restrict__ cell_of, const uint32_t* __restrict__ start_all, TileSystems ts,
                 int4* __restrict__ tile_ranges, int* __restrict__ tile_nr, const uint32_t* __restrict__ order,
                 PermuteArgs pa, int* __restrict__ tile_order, const int* gate, const int* halt) {
    if (gated_off(gate, halt)) return;
    __shared__ int2 cr_all[SETUP_TILES][CR_CAP];
    __shared__ int nr_all[SETUP_TILES];
    __shared__ int flat_all[SETUP_TILES];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    int t = blockIdx.x * SETUP_TILES + w;
    if constexpr (BATCH) t = t < uniform_i32(ts.active[0]) ? uniform_i32(ts.active[1 + t]) : ts.tiles;
    const TileSystem sys = tile_system<BATCH>(ts, t < ts.tiles ? t : 0);
    const size_t n = sys.end;  // one past the system's last particle
    // empty tiles (past a replica's end) keep tile_nr = -1 from the batch's set-up
    const bool live = t < ts.tiles && static_cast<size_t>(t) * TILE < n && !tile_gated_off<BATCH>(ts, sys);
    const int side = sys.side, use_cells = sys.use_cells, hmax = sys.hmax;
    const double flat_cells = sys.flat_cells;
    const uint32_t* const start = start_all + sys.cell_base;
    int2* cr = cr_all[w];
    // after a re-sort: the state of the tile's particles, loaded in the new order while the ranges are merged and
    // stored at the end (positions into ref, the rest into the alt arrays)
    const bool gather = live && pa.enabled == PERMUTE_STATE;
    GatheredTile g;
    if (gather) load_gathered(pa, order, g, t, lane, n);
    if (live) {
        const size_t i0 = static_cast<size_t>(t) * TILE;
        const size_t i1 = (i0 + TILE) < n ? (i0 + TILE) : n;
        if (use_cells) {
            const int c0 = cell_of[i0], c1 = cell_of[i1 - 1];
            const bool wave_path = c1 / side - c0 / side < 3;
            int nr = 0;
            if (wave_path) nr = tile_cell_ranges_wave(c0, c1, side, cr, lane);
            if (lane == 0) {
                if (!wave_path) nr = tile_cell_ranges(c0, c1, side, cr);
                // cells c = (x side + y) side + z; the tile's cells c0 .. c1 span columns col0 .. col1
                const int col0 = c0 / side, col1 = c1 / side, z0 = c0 % side, z1 = c1 % side;
                const int x0 = col0 / side, y0 = col0 % side, x1 = col1 / side, y1 = col1 % side;
                nr_all[w] = nr;
                auto axis = [&](int lo, int hi) {  // cells lo .. hi along an axis, stencil included
                    return lo >= 2 && hi <= side - 3 && static_cast<double>(hi - lo + 3) < flat_cells;
                };
                // one column, or two adjacent ones (then the z range runs through both z faces)
                const bool fx = col1 - col0 <= 1 && axis(min(x0, x1), max(x0, x1));
                const bool fy = col1 - col0 <= 1 && axis(min(y0, y1), max(y0, y1));
                const bool fz = col1 == col0 && axis(z0, z1);
                flat_all[w] = (fx ? TILE_FLAT_X : 0) | (fy ? TILE_FLAT_X << 1 : 0) | (fz ? TILE_FLAT_X << 2 : 0) |
                              (fx && fy && fz ? TILE_FLAT : 0);
            }
        } else if (lane == 0) {
            cr[0] = make_int2(0, side * side * side - 1);  // every particle
            nr_all[w] = 1;
            flat_all[w] = 0;
        }
    }
    if (tile_order && live) {
        // the tile's place in band order (see ORDER_BAND): from the cell starts alone, a tile belonging to the
        // column of its first particle; tiles(x) = tiles whose first particle lies before particle x
        auto tiles_before = [](uint32_t x) { return static_cast<int>((x + TILE - 1) / TILE); };
        int at = t;
        if (use_cells && side >= 2 * ORDER_BAND) {
            const int col = cell_of[static_cast<size_t>(t) * TILE] / side, cx = col / side, cy = col % side;
            const int b0 = cx / ORDER_BAND * ORDER_BAND, bw = min(ORDER_BAND, side - b0), k = lane;
            const size_t s2 = static_cast<size_t>(side) * side;
            int v = 0;
            if (k < bw) {
                const size_t x = b0 + k, c = (x * side + cy) * side;
                // tiles of slab x in the columns y < cy, and (columns x < cx of the band) of column (x, cy)
                v = tiles_before(start[c]) - tiles_before(start[x * s2]);
                if (x < static_cast<size_t>(cx)) v += tiles_before(start[c + side]) - tiles_before(start[c]);
            }
#pragma unroll
            for (int off = 1; off < ORDER_BAND; off <<= 1) v += __shfl_xor(v, off, WAVE);
            const size_t own = (static_cast<size_t>(cx) * side + cy) * side;
            at = tiles_before(start[b0 * s2]) + v + (t - tiles_before(start[own]));
        }
        if (lane == 0) tile_order[at] = t;
    }
    __syncthreads();
    if (!live) return;
    const int nr = nr_all[w];
    int first_cell = 0, first = 0, cnt = 0;
    if (lane < nr) {
        first_cell = cr[lane].x;
        first = static_cast<int>(start[first_cell]);
        cnt = static_cast<int>(start[cr[lane].y + 1]) - first;
    }
    int incl = cnt;  // inclusive scan of the range lengths
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int o = __shfl_up(incl, d, WAVE);
        if (lane >= d) incl += o;
    }
    const int total = __shfl(incl, WAVE - 1, WAVE);
    if (lane < nr)
        tile_ranges[static_cast<size_t>(t) * RMAX + lane] = make_int4(first_cell, first, incl - cnt, cnt);
    if (lane == 0) tile_nr[t] = (nr > 0 && total <= hmax) ? nr + flat_all[w] : -1;
    if (gather) store_gathered(pa, g, t, lane, n);
}

// Tile header shared by the build and the force kernel: the tile's halo ranges in LDS.
struct HaloRanges {
    int* ra;    // first cell of each range
    int* rgs;   // first particle of each range
    int* rofs;  // first LDS slot of each range; rofs[nr] = halo size
};

__device__ __forceinline__ void load_ranges(const int4* __restrict__ tile_ranges, int tile, int nr, HaloRanges h) {
    if (static_cast<int>(threadIdx.x) < nr) {
        const int4 r = tile_ranges[static_cast<size_t>(tile) * RMAX + threadIdx.x];
        h.ra[threadIdx.x] = r.x;
        h.rgs[threadIdx.x] = r.y;
        h.rofs[threadIdx.x] = r.z;
        if (static_cast<int>(threadIdx.x) == nr - 1) h.rofs[nr] = r.z + r.w;
    }
}

// Global particle index held by LDS slot q of the halo.
__device__ __forceinline__ int halo_particle(int q, HaloRanges h, int nr) {
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (h.rofs[mid] <= q)
            lo = mid;
        else
            hi = mid - 1;
    }
    return h.rgs[lo] + (q - h.rofs[lo]);
}

// LDS slot of the first particle of cell `nc` (a cell inside the halo), given that particle's index `first`.
__device__ __forceinline__ int slot_base_at(int nc, HaloRanges h, int first, int nr) {
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (h.ra[mid] <= nc)
            lo = mid;
        else
            hi = mid - 1;
    }
    return h.rofs[lo] + first - h.rgs[lo];
}

__device__ __forceinline__ int slot_base_of_cell(int nc, HaloRanges h, const uint32_t* __restrict__ start,
                                                 int nr) {
    return slot_base_at(nc, h, static_cast<int>(start[nc]), nr);
}

// Stages the halo: thread q of THREADS loads slots q, q + THREADS, ... (all loads issued before any is used), and
// hands each loaded position with its slot to `store`.
template <int THREADS = TILE_THREADS, class Store>
__device__ __forceinline__ void stage_halo(const double4* __restrict__ pos, HaloRanges h, int nr, int total,
                                           Store store) {
    constexpr int ROUNDS = (TILE_HALO_MAX + THREADS - 1) / THREADS;
    int src[ROUNDS];
    // the thread's slots grow by THREADS per round: one binary search, then a short forward walk per round
    int r = 0;
    {
        int lo = 0, hi = nr - 1;
        const int q = static_cast<int>(threadIdx.x);
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (h.rofs[mid] <= q)
                lo = mid;
            else
                hi = mid - 1;
        }
        r = lo;
    }
#pragma unroll
    for (int u = 0; u < ROUNDS; ++u) {
        const int q = static_cast<int>(threadIdx.x) + u * THREADS;
        if (q < total) {
            while (h.rofs[r + 1] <= q) ++r;  // rofs[nr] = total > q ends the walk
            src[u] = h.rgs[r] + (q - h.rofs[r]);
        } else {
            src[u] = -1;
        }
    }
    double4 p[ROUNDS];
#pragma unroll
    for (int u = 0; u < ROUNDS; ++u)
        if (src[u] >= 0) p[u] = pos[src[u]];
#pragma unroll
    for (int u = 0; u < ROUNDS; ++u)
        if (src[u] >= 0) store(static_cast<int>(threadIdx.x) + u * THREADS, p[u]);
}

// Stages the halo with no LDS reads: every lane of each of the block's W wavefronts holds range r = lane of the
// tile (rr: first cell, first particle, first slot, count; ranges r >= nr unused). A wavefront takes the 64-slot
// windows wave, wave + W, ...; a window's first slot lies in the last range that starts at or before it (a ballot
// of the ranges' first slots), and the ranges that start inside the window (another ballot, usually one or two)
// take over the lanes at and past their first slot. All loads are issued before any is used.
template <int W, class Store>
__device__ __forceinline__ void stage_halo_lanes(const double4* __restrict__ pos, int4 rr, int nr, Store store) {
    const int lane = static_cast<int>(threadIdx.x) % WAVE, wave = static_cast<int>(threadIdx.x) / WAVE;
    const int first = lane < nr ? rr.z : INT_MAX;  // first slot of range `lane`
    const int shift = rr.y - rr.z;                 // particle index minus slot within range `lane`
    const int total = __builtin_amdgcn_readlane(rr.z + rr.w, nr - 1);
    constexpr int ROUNDS = (TILE_HALO_MAX + W * WAVE - 1) / (W * WAVE);
    int src[ROUNDS];
#pragma unroll
    for (int u = 0; u < ROUNDS; ++u) {
        const int base = (wave + W * u) * WAVE;
        const int q = base + lane;
        int g = 0;
        if (base < total) {
            const uint64_t at_or_before = __builtin_amdgcn_ballot_w64(first <= base);  // bit 0 set: first[0] = 0
            g = __builtin_amdgcn_readlane(shift, 63 - __builtin_clzll(at_or_before));
            uint64_t inside = __builtin_amdgcn_ballot_w64(first > base && first < base + WAVE);
            while (inside) {
                const int b = __builtin_ctzll(inside);
                inside &= inside - 1;
                const int fb = __builtin_amdgcn_readlane(first, b);
                const int gb = __builtin_amdgcn_readlane(shift, b);
                g = q >= fb ? gb : g;
            }
        }
        src[u] = q < total ? g + q : -1;
    }
    double4 p[ROUNDS];
#pragma unroll
    for (int u = 0; u < ROUNDS; ++u)
        if (src[u] >= 0) p[u] = pos[src[u]];
#pragma unroll
    for (int u = 0; u < ROUNDS; ++u)
        if (src[u] >= 0) store((wave + W * u) * WAVE + lane, p[u]);
}

// One stencil segment of an own cell: neighbour cell id, LDS slot of its first particle, its particle count.
struct Segment {
    int nc;
    uint16_t base, len;
};

constexpr int LIST_CAP = 6144;  // flat candidate lists of a tile's own cells, in LDS slots
constexpr int ROWBUF_CAP = 320; // row entries a wavefront of the list walk buffers in LDS (longer rows go direct)
constexpr int OWN_STENCIL = 13; // stencil index of (0, 0, 0)

// Candidates j of particle i, walked in the order of its row. `STAGED` reads the halo from LDS, else from global
// memory; `INTERIOR` drops the minimum image where it is the identity (every candidate within 3 cells = 3L/side
// < L/2 per axis, so round(d / L) = 0 and the reference's d - L * round(d / L) is d itself). Runs on the 4 lanes
// of a particle: they test 4 consecutive candidates and a wavefront ballot compacts the accepted ones in order.
struct LaneScan {
    const double* xs;
    const double* ys;
    const double* zs;
    const double4* pos;
    double4 ri;
    double box, inv_box, radius2;
    size_t stride;  // column stride of the int32 rows
    size_t i;
    int lane, group_shift;
    unsigned lower_mask;
    int padding;
    int32_t* row;
    uint16_t* row16;  // the particle's slot_row_base
    int cnt;

    // STAGED: positions from the LDS halo, else from global memory; SLOTS: the slot entries (base + offset in the
    // cell) are written too, which a staged tile needs however its positions were read
    template <bool STAGED, bool INTERIOR, bool SLOTS = STAGED>
    __device__ __forceinline__ void scan(uint32_t b, uint32_t e, int base) {
        for (uint32_t jb = b; jb < e; jb += LANES) {
            // past the cell's end a lane re-reads its first candidate and rejects it
            const uint32_t j = jb + lane;
            const bool valid = j < e;
            const uint32_t jj = valid ? j : b;
            const int slot = SLOTS ? base + static_cast<int>(jj - b) : 0;
            double xj, yj, zj;
            if constexpr (STAGED) {
                xj = xs[slot];
                yj = ys[slot];
                zj = zs[slot];
            } else {
                const double4 rj = pos[jj];
                xj = rj.x;
                yj = rj.y;
                zj = rj.z;
            }
            double dx = ri.x - xj, dy = ri.y - yj, dz = ri.z - zj;
            if constexpr (!INTERIOR) {
                dx = min_image(dx, box, inv_box);
                dy = min_image(dy, box, inv_box);
                dz = min_image(dz, box, inv_box);
            }
            const bool near = dx * dx + dy * dy + dz * dz < radius2;
            const bool accept = valid & near;  // both evaluated: no branch around the distance test
            const unsigned group = static_cast<unsigned>((__ballot(accept) >> group_shift) & 0xFull);
            if (accept) {
                const unsigned at = static_cast<unsigned>(cnt) + __popc(group & lower_mask);
                if (at < static_cast<unsigned>(padding)) {
                    row[static_cast<uint64_t>(at) * stride] = static_cast<int32_t>(j);
                    row16[slot_entry(at)] = static_cast<uint16_t>(slot * SLOT_BYTES);
                }
            }
            cnt += __popc(group);
        }
    }
};

// List build of one tile. Fast path (staged halo, own cells' candidate lists fit in LDS): one wavefront per
// particle walks a flat list of LDS slots — for each own cell c, U_c = the slots of the stencil cells with id >= c
// in the reference's stencil order, L_c = those with id <= c in increasing id — 64 candidates at a time; the
// accepted ones (j > i from U_c, j < i from L_c, slot order being index order) are ranked by mbcnt over the
// ballot, so the row comes out in the reference's order with no divergence between lanes. Other tiles use the
// 4-lanes-per-particle scan with a binary search per stencil cell (LaneScan).
//
// FH (float halo): the halo is staged as float coordinates in the frame of the tile's first particle (half the LDS
// of doubles, three blocks per CU). A candidate is accepted when its float r^2 lies below the list radius by more
// than the rounding bound `band` of the frame, rejected when above it by more, and otherwise tested again exactly in
// double from global memory with the reference's minimum image; so the rows are those of the double walk. In a
// compact frame (2 e + 4 cells + 1 < L, e the largest own frame coordinate) a float separation is already the
// minimum image; otherwise (small boxes) the walk takes the float minimum image of each separation, with a wider
// band.
template <bool FH>
__device__ __forceinline__ void list_build_tile(
    int tile, double* lds, const double4* __restrict__ pos, const int* __restrict__ cell_of,
    const uint32_t* __restrict__ start, const TileSystem& sys, size_t stride, double radius2, int padding,
    int write32, int row_len, int32_t* __restrict__ rows, uint16_t* __restrict__ rows16, int32_t* __restrict__ count,
    int32_t* __restrict__ half, int2* __restrict__ stats, const int4* __restrict__ tile_ranges,
    const int* __restrict__ tile_nr, const TiledCopyBack& cb) {
    const size_t n = sys.end;  // one past the system's last particle
    const double box = sys.box, inv_box = sys.inv_box;
    double* xs = lds;  // fixed strides: the three reads of a slot share one address plus immediate offsets
    double* ys = lds + TILE_HALO_MAX;
    double* zs = lds + 2 * TILE_HALO_MAX;
    float* xf = reinterpret_cast<float*>(lds);  // FH: the same slots as floats
    float* yf = xf + TILE_HALO_MAX;
    float* zf = xf + 2 * TILE_HALO_MAX;
    __shared__ int s_extent;  // FH: largest own frame coordinate (float bits)
    // the candidate lists, the own slots and the row buffer hold byte offsets into the staged halo: of floats with
    // FH (converted to the slot rows' double offsets when a row is written out), of doubles otherwise
    constexpr int LIST_BYTES = FH ? 4 : SLOT_BYTES;
    __shared__ int ra[RMAX], rgs[RMAX], rofs[RMAX + 1];
    // the stencil segments are dead once the candidate lists are filled; the walk then keeps each wavefront's
    // current row in the same memory
    __shared__ union {
        Segment seg[CMAX * 27];
        uint16_t rowbuf[TILE_THREADS / WAVE][2][ROWBUF_CAP];  // two rows per wavefront (see `pair`)
    } sm;
    Segment* const seg = sm.seg;
    __shared__ uint16_t lists[LIST_CAP];  // slot byte offsets
    __shared__ int lst[CMAX][4];  // U offset, U length, L offset, L length (lengths padded to 64 with the ghost)
    __shared__ int cl_first[CMAX];
    __shared__ int cl_slot[CMAX];  // LDS slot of the own cell's first particle
    __shared__ unsigned char cl_interior[CMAX];
    __shared__ unsigned char p_cl[TILE];  // own cell of each of the tile's particles, counted from the first
    __shared__ uint16_t p_slot[TILE];     // LDS slot byte offset of each of the tile's particles
    __shared__ int s_outside, s_lists_ok;
    const HaloRanges hr{ra, rgs, rofs};
    const size_t i0 = static_cast<size_t>(tile) * TILE;
    const size_t i1 = (i0 + TILE) < n ? (i0 + TILE) : n;
    const int s = sys.side;
    const int nr = tile_range_count(tile_nr[tile]);
    const bool staged = nr > 0;
    const int c_first = sys.use_cells ? cell_of[i0] : 0;
    const int ncl = sys.use_cells ? cell_of[i1 - 1] - c_first + 1 : 0;
    const bool use_lists = staged && sys.use_cells && ncl <= CMAX;
    const uint32_t n32 = static_cast<uint32_t>(n);
    // the halo ranges, loaded together with the header (every tile owns RMAX entries; those past nr are unused)
    const int4 rr = static_cast<int>(threadIdx.x) < RMAX ? tile_ranges[static_cast<size_t>(tile) * RMAX + threadIdx.x]
                                                         : make_int4(0, 0, 0, 0);
    // the cells of the tile's own particles, loaded together with the header so the walk reads them from LDS
    const int my_cell = sys.use_cells && static_cast<size_t>(threadIdx.x) < i1 - i0 ? cell_of[i0 + threadIdx.x] : 0;
    if (threadIdx.x == 0) {
        s_outside = 0;
        s_lists_ok = 0;
        s_extent = 0;
    }
    const double4 fref = FH ? pos[i0] : make_double4(0.0, 0.0, 0.0, 0.0);  // FH: the frame's origin
    auto fcoord = [&](double x, double r) {
        const double t = x - r;
        return static_cast<float>(fma(-box, rint(t * inv_box), t));
    };
    // Moving the re-sorted arrays into place: each quarter of the block carries one group of fields of the tile's
    // particles (positions; velocities; forces; mass, id and cell). The loads leave with the tile header; the stores
    // wait until the tile is done (no thread of this kernel reads the destination arrays).
    static_assert(TILE_THREADS == 4 * TILE, "one copy-back group per quarter of the block");
    const int cb_group = threadIdx.x / TILE;
    const size_t cb_i = i0 + threadIdx.x % TILE;
    const bool cb_live = cb.enabled && cb_i < i1 && (cb_group != 2 || cb.frc != nullptr);
    uint64_t cbv[4] = {0, 0, 0, 0};
    if (cb_live) {
        if (cb_group == 0) {
            const double4 q = pos[cb_i];
            cbv[0] = __double_as_longlong(q.x);
            cbv[1] = __double_as_longlong(q.y);
            cbv[2] = __double_as_longlong(q.z);
            cbv[3] = __double_as_longlong(q.w);
        } else if (cb_group == 1) {
            for (int c = 0; c < 3; ++c) cbv[c] = __double_as_longlong(cb.vel2[c * stride + cb_i]);
        } else if (cb_group == 2) {
            for (int c = 0; c < 3; ++c) cbv[c] = __double_as_longlong(cb.frc2[c * stride + cb_i]);
        } else {
            cbv[0] = __double_as_longlong(cb.mass2[cb_i]);
            cbv[1] = cb.id2[cb_i];
            cbv[2] = static_cast<uint32_t>(cb.cell_of2[cb_i]);
        }
    }
    if (staged && static_cast<int>(threadIdx.x) < nr) {
        ra[threadIdx.x] = rr.x;
        rgs[threadIdx.x] = rr.y;
        rofs[threadIdx.x] = rr.z;
        if (static_cast<int>(threadIdx.x) == nr - 1) rofs[nr] = rr.z + rr.w;
    }
    // the stencil cells of every own cell and their particle spans: loaded now, so that the loads are in flight
    // while the halo is staged
    constexpr int SEG_ROUNDS = (CMAX * 27 + TILE_THREADS - 1) / TILE_THREADS;
    int seg_nc[SEG_ROUNDS];
    uint32_t seg_b[SEG_ROUNDS], seg_e[SEG_ROUNDS];
#pragma unroll
    for (int u = 0; u < SEG_ROUNDS; ++u) {
        const int t = threadIdx.x + u * TILE_THREADS;
        seg_nc[u] = 0;
        seg_b[u] = seg_e[u] = 0;
        if (use_lists && t < ncl * 27) {
            const int cl = t / 27, st = t % 27;
            const int c = c_first + cl;
            const int cx = c / (s * s), cy = (c / s) % s, cz = c % s;
            const int nc = (wrap_index(cx + st / 9 - 1, s) * s + wrap_index(cy + (st / 3) % 3 - 1, s)) * s +
                           wrap_index(cz + st % 3 - 1, s);
            seg_nc[u] = nc;
            seg_b[u] = start[nc];
            seg_e[u] = start[nc + 1];
        }
    }
    __syncthreads();
    if (staged && FH) {
        stage_halo(pos, hr, nr, rofs[nr], [&](int q, const double4& p) {
            xf[q] = fcoord(p.x, fref.x);
            yf[q] = fcoord(p.y, fref.y);
            zf[q] = fcoord(p.z, fref.z);
        });
        if (threadIdx.x == 0) {
            const float ghost = __builtin_inff();  // its r^2 is above every bound
            xf[GHOST_SLOT] = ghost;
            yf[GHOST_SLOT] = ghost;
            zf[GHOST_SLOT] = ghost;
        }
    } else if (staged) {
        int outside = 0;
        stage_halo(pos, hr, nr, rofs[nr], [&](int q, const double4& p) {
            xs[q] = p.x;
            ys[q] = p.y;
            zs[q] = p.z;
            outside |= !(p.x >= 0.0 && p.x < box && p.y >= 0.0 && p.y < box && p.z >= 0.0 && p.z < box);
        });
        if (outside) s_outside = 1;
        if (threadIdx.x == 0) {
            const double ghost = __builtin_nan("");
            xs[GHOST_SLOT] = ghost;
            ys[GHOST_SLOT] = ghost;
            zs[GHOST_SLOT] = ghost;
        }
    }
    // stencil segments of every own cell
#pragma unroll
    for (int u = 0; u < SEG_ROUNDS; ++u) {
        const int t = threadIdx.x + u * TILE_THREADS;
        if (!use_lists || t >= ncl * 27) continue;
        const int cl = t / 27, st = t % 27;
        Segment g;
        g.nc = seg_nc[u];
        g.base = static_cast<uint16_t>(slot_base_at(seg_nc[u], hr, static_cast<int>(seg_b[u]), nr));
        g.len = static_cast<uint16_t>(seg_e[u] - seg_b[u]);
        seg[t] = g;
        if (st == OWN_STENCIL) {
            cl_first[cl] = static_cast<int>(seg_b[u]);
            cl_slot[cl] = g.base;
        }
    }
    __syncthreads();
    // Candidate lists: list L = 2 cl + lower (U_c, then L_c). Half-wavefront h takes lists h, h + 16, h + 32; its
    // lane j holds the list's j-th segment (U_c: stencil index j, kept if nc >= c; L_c: the j-th stencil cell in
    // increasing id, kept if nc <= c), and a 32-lane scan of the kept lengths gives every segment's offset and the
    // list's length. The scan results stay in registers until the fill below.
    constexpr int HALF = WAVE / 2;
    constexpr int LIST_ROUNDS = (2 * CMAX + TILE_THREADS / HALF - 1) / (TILE_THREADS / HALF);
    const int hw = threadIdx.x / HALF, hl = threadIdx.x % HALF;
    int seg_off[LIST_ROUNDS], seg_base[LIST_ROUNDS], seg_len[LIST_ROUNDS], list_used[LIST_ROUNDS];
#pragma unroll
    for (int u = 0; u < LIST_ROUNDS; ++u) {
        const int L = hw + u * (TILE_THREADS / HALF);
        seg_off[u] = seg_base[u] = seg_len[u] = list_used[u] = 0;
        if (!use_lists || L >= 2 * ncl) continue;  // uniform over the half-wavefront
        const int cl = L >> 1, lower = L & 1;
        const int c = c_first + cl;
        int len = 0, base = 0;
        if (hl < 27) {
            int st = hl;
            if (lower) {
                // stencil offset index (0, 1, 2 for -1, 0, +1) of the a-th smallest wrapped coordinate
                auto axis = [&](int v, int a) { return v == 0 ? (a + 1) % 3 : (v == s - 1 ? (a + 2) % 3 : a); };
                const int cx = c / (s * s), cy = (c / s) % s, cz = c % s;
                st = axis(cx, hl / 9) * 9 + axis(cy, (hl / 3) % 3) * 3 + axis(cz, hl % 3);
            }
            const Segment g = seg[cl * 27 + st];
            if (lower ? g.nc <= c : g.nc >= c) {
                len = g.len;
                base = g.base;
            }
        }
        int incl = len;
#pragma unroll
        for (int d = 1; d < HALF; d <<= 1) {
            const int o = __shfl_up(incl, d, HALF);
            if (hl >= d) incl += o;
        }
        seg_off[u] = incl - len;
        seg_base[u] = base;
        seg_len[u] = len;
        list_used[u] = __shfl(incl, HALF - 1, HALF);
        if (hl == 0) lst[cl][lower ? 3 : 1] = (list_used[u] + WAVE - 1) / WAVE * WAVE;
    }
    if (use_lists && static_cast<size_t>(threadIdx.x) < i1 - i0) {
        const int cl = my_cell - c_first;
        const int slot = cl_slot[cl] + (static_cast<int>(i0 + threadIdx.x) - cl_first[cl]);
        p_cl[threadIdx.x] = static_cast<unsigned char>(cl);
        p_slot[threadIdx.x] = static_cast<uint16_t>(slot * LIST_BYTES);
        if constexpr (FH) {  // non-negative floats order as their bit patterns
            const float e = fmaxf(fabsf(xf[slot]), fmaxf(fabsf(yf[slot]), fabsf(zf[slot])));
            atomicMax(&s_extent, __float_as_int(e));
        }
    }
    if (use_lists && static_cast<int>(threadIdx.x) < ncl) {
        const int cl = threadIdx.x, c = c_first + cl;
        const int cx = c / (s * s), cy = (c / s) % s, cz = c % s;
        cl_interior[cl] = sys.interior_ok && cx >= 1 && cx <= s - 2 && cy >= 1 && cy <= s - 2 && cz >= 1 &&
                          cz <= s - 2;
    }
    __syncthreads();
    if (use_lists && threadIdx.x < WAVE) {  // list offsets in order U_0, L_0, U_1, ...: one wavefront scan
        const int L = threadIdx.x;
        const int len = L < 2 * ncl ? lst[L >> 1][(L & 1) ? 3 : 1] : 0;
        int incl = len;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const int o = __shfl_up(incl, d, WAVE);
            if (L >= d) incl += o;
        }
        if (L < 2 * ncl) lst[L >> 1][(L & 1) ? 2 : 0] = incl - len;
        if (L == WAVE - 1) s_lists_ok = incl <= LIST_CAP ? 1 : 0;
    }
    __syncthreads();
    // the walk keeps rows in its LDS row buffer; longer rows take the 4-lane scan, which stores directly
    // FH: the frame holds minimum images when every own particle lies within e of the origin and 2 e + 4 cells + 1
    // < L (a candidate is within 2 cells of its own particle); then a float r^2 is within `band` of the exact one,
    // band bounding the rounding of frame coordinates up to M = e + 2 cells + 1 (ulp(M) / 2 <= M 2^-24 each)
    // Tiles whose frame is not compact (small boxes) take the minimum image of each float separation,
    // d - L_f rint(d / L_f): a component then differs from the exact minimum-image one by at most
    // 3 x 3.5 L 2^-24 (the two frame coordinates, the difference, L_f and the fma each round once; where the
    // rounded quotient picks the other image, both images lie within that of L/2), and the band below bounds the
    // resulting r^2 error near the list radius with a factor 2 to spare.
    const double cell = box / s;
    const double fe = FH ? static_cast<double>(__int_as_float(s_extent)) : 0.0;
    const bool compact = !FH || 2.0 * fe + 4.0 * cell + 1.0 < box;
    const double big_m = fe + 2.0 * cell + 1.0, r_l = sqrt(radius2);
    constexpr double EPS32 = 5.9604644775390625e-08;  // 2^-24
    const double band = compact ? 4.0 * (2.0 * 1.7320508075688772 * r_l * (2.0 * big_m + 3.0 + r_l) + 3.0 * radius2) * EPS32
                                : (72.75 * r_l * box + 8.0 * radius2) * EPS32;
    // thresholds of the float r^2, rounded outwards: below r2_in a candidate is inside, above r2_out outside, in
    // between it is decided by the exact test
    const float r2_in = __double2float_rd(radius2 - band), r2_out = __double2float_ru(radius2 + band);
    const float box_f = static_cast<float>(box), inv_box_f = static_cast<float>(inv_box);
    const bool lists_ok = use_lists && s_lists_ok && padding <= ROWBUF_CAP;
    // fill: each lane copies its segment's slots; the list's lanes then pad it to a multiple of 64 with the ghost
#pragma unroll
    for (int u = 0; u < LIST_ROUNDS; ++u) {
        const int L = hw + u * (TILE_THREADS / HALF);
        if (!lists_ok || L >= 2 * ncl) continue;
        const int cl = L >> 1, lower = L & 1;
        const int off = lst[cl][lower ? 2 : 0];
        for (int m = 0; m < seg_len[u]; ++m)
            lists[off + seg_off[u] + m] = static_cast<uint16_t>((seg_base[u] + m) * LIST_BYTES);
        const int padded = lst[cl][lower ? 3 : 1];
        for (int q = list_used[u] + hl; q < padded; q += HALF)
            lists[off + q] = static_cast<uint16_t>(GHOST_SLOT * LIST_BYTES);
    }
    __syncthreads();
    int ep_max = 0;      // longest row seen by this lane's wavefront / lane group
    unsigned ep_half = 0;  // half-row lengths summed once per particle
    if (lists_ok) {
        const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
        // an entry is the byte offset of a slot: its x at xs + offset, y and z one and two arrays on (FH: half the
        // offset into the float arrays)
        auto coords = [&](int off) { return reinterpret_cast<const double*>(reinterpret_cast<const char*>(xs) + off); };
        auto coords_f = [&](int off) {
            return reinterpret_cast<const float*>(reinterpret_cast<const char*>(xf) + off);
        };
        // A row out of the wavefront's row buffer rb (cnt entries ranked, hcnt of them j > i): its slot chunks, count
        // and half count, and the tile's row statistics.
        auto write_row = [&](size_t i, uint16_t* rb, int cnt, int hcnt) {
            uint16_t* const row16 = rows16 + slot_row_base(i, row_len);
            {
                // the row leaves LDS as 16-byte chunk words: lane 4 c + w packs entries 32 c + w + 4 t, t = 0..7
                // (the wavefront's LDS accesses complete in issue order, so every lane's entries are in place)
                __builtin_amdgcn_wave_barrier();
                const int stored = cnt < padding ? cnt : padding;
                // the last chunk ends in ghost entries (the force kernel's compact walk reads whole chunks)
                if (stored + lane < (stored + CHUNK - 1) / CHUNK * CHUNK)
                    rb[stored + lane] = static_cast<uint16_t>(GHOST_SLOT * LIST_BYTES);
                __builtin_amdgcn_wave_barrier();
                if (lane < (stored + CHUNK - 1) / CHUNK * LANES) {
                    const int c = lane >> 2, w = lane & 3;
                    const uint16_t* e = rb + c * CHUNK + w;
                    uint4 v;
                    v.x = static_cast<unsigned>(e[0]) | (static_cast<unsigned>(e[4]) << 16);
                    v.y = static_cast<unsigned>(e[8]) | (static_cast<unsigned>(e[12]) << 16);
                    v.z = static_cast<unsigned>(e[16]) | (static_cast<unsigned>(e[20]) << 16);
                    v.w = static_cast<unsigned>(e[24]) | (static_cast<unsigned>(e[28]) << 16);
                    if constexpr (FH) {  // float offsets to double offsets: both halves doubled (each < 2^15)
                        static_assert(TILE_HALO_MAX * 4 < 0x8000, "a doubled float offset stays in its half");
                        v.x <<= 1;
                        v.y <<= 1;
                        v.z <<= 1;
                        v.w <<= 1;
                    }
                    *reinterpret_cast<uint4*>(row16 + c * CHUNK_STRIDE + 8 * w) = v;
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (lane == 0) {
                count[i] = cnt < padding ? cnt : padding;
                half[i] = hcnt < padding ? hcnt : padding;
            }
            ep_max = cnt > ep_max ? cnt : ep_max;
            if (lane == 0) ep_half += static_cast<unsigned>(hcnt);
        };
        auto single = [&](size_t p) {
            const size_t i = i0 + p;
            // uniform over the wavefront: one particle per wavefront
            const int cl = __builtin_amdgcn_readfirstlane(p_cl[p]);
            const int slot_i = __builtin_amdgcn_readfirstlane(p_slot[p]);
            double xi = 0.0, yi = 0.0, zi = 0.0;
            float xif = 0.0f, yif = 0.0f, zif = 0.0f;
            double4 ri = make_double4(0.0, 0.0, 0.0, 0.0);
            if constexpr (FH) {
                const float* const ci = coords_f(slot_i);
                xif = ci[0];
                yif = ci[TILE_HALO_MAX];
                zif = ci[2 * TILE_HALO_MAX];
                ri = pos[i];
            } else {
                const double* const ci = coords(slot_i);
                xi = ci[0];
                yi = ci[TILE_HALO_MAX];
                zi = ci[2 * TILE_HALO_MAX];
            }
            const bool interior = cl_interior[cl] && !s_outside;
            int32_t* const row = rows + i;
            uint16_t* const rb = sm.rowbuf[wave][0];
            int cnt = 0;
            // A batch is 64 candidates: each lane tests one, the accepted ones are ranked over the ballot (mbcnt
            // counts the set bits below the lane, on top of the row's count so far) and stored. Two batches are
            // loaded and tested before either stores, so their LDS reads overlap.
            // FH: the float r^2 (infinite for the ghost); a candidate is inside when r^2 < r2_in, unsure when
            // r2_in <= r^2 <= r2_out (every candidate is one of inside, unsure, outside), and an unsure one is
            // decided by the exact test in double
            // `flat_tag`: a compact frame (no image term); otherwise the float minimum image per component
            auto dist_f = [&](int sl, auto flat_tag) {
                const float* const cj = coords_f(sl);
                float dx = xif - cj[0], dy = yif - cj[TILE_HALO_MAX], dz = zif - cj[2 * TILE_HALO_MAX];
                if constexpr (!decltype(flat_tag)::value) {
                    dx = fmaf(-box_f, __builtin_rintf(dx * inv_box_f), dx);
                    dy = fmaf(-box_f, __builtin_rintf(dy * inv_box_f), dy);
                    dz = fmaf(-box_f, __builtin_rintf(dz * inv_box_f), dz);
                }
                return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            };
            auto unsure_f = [&](float d) { return d <= r2_out && !(d < r2_in); };
            auto inside_f = [&](float d) { return d < r2_in; };
            auto resolve_f = [&](int sl, float d, bool unsure) {
                bool in = inside_f(d);
                if (unsure) {
                    const double4 rj = pos[halo_particle(sl / LIST_BYTES, hr, nr)];
                    const double ex = min_image(ri.x - rj.x, box, inv_box);
                    const double ey = min_image(ri.y - rj.y, box, inv_box);
                    const double ez = min_image(ri.z - rj.z, box, inv_box);
                    in = ex * ex + ey * ey + ez * ez < radius2;
                }
                return in;
            };
            auto near = [&](int sl, auto interior_tag) {  // the double halo
                const double* const cj = coords(sl);
                double dx = xi - cj[0], dy = yi - cj[TILE_HALO_MAX], dz = zi - cj[2 * TILE_HALO_MAX];
                if constexpr (!decltype(interior_tag)::value) {
                    dx = min_image(dx, box, inv_box);
                    dy = min_image(dy, box, inv_box);
                    dz = min_image(dz, box, inv_box);
                }
                return dx * dx + dy * dy + dz * dz < radius2;  // false for the ghost (NaN)
            };
            auto commit = [&](int sl, bool accept, unsigned long long mask) {
                if (accept) {
                    const unsigned at = __builtin_amdgcn_mbcnt_hi(
                        static_cast<unsigned>(mask >> 32),
                        __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), static_cast<unsigned>(cnt)));
                    if (at < static_cast<unsigned>(padding)) {
                        rb[at] = static_cast<uint16_t>(sl);
                        if (write32) row[static_cast<uint64_t>(at) * stride] = halo_particle(sl / LIST_BYTES, hr, nr);
                    }
                }
                // the count is uniform: kept in a scalar register, the popcount and addition are scalar work
                cnt = __builtin_amdgcn_readfirstlane(cnt + __popcll(mask));
            };
            // `upper_tag`: the U list (keep j > i) or the L list (keep j < i), a compile-time choice, so the order
            // test is a single comparison whose lane mask the ballots use directly
            auto walk = [&](int off, int len, auto upper_tag, auto interior_tag) {
                constexpr bool upper = decltype(upper_tag)::value;
                int q0 = 0;
                if constexpr (FH) {
                    // A pair of batches takes the plain path unless a lane is within rounding of the cut-off
                    // (a few in a thousand batches), which re-tests those lanes exactly.
                    for (; q0 + 2 * WAVE <= len; q0 += 2 * WAVE) {
                        const int sa = lists[off + q0 + lane], sb = lists[off + q0 + WAVE + lane];
                        const float ra = dist_f(sa, interior_tag), rb2 = dist_f(sb, interior_tag);
                        const bool oa = upper ? sa > slot_i : sa < slot_i, ob = upper ? sb > slot_i : sb < slot_i;
                        const bool na = inside_f(ra), nb = inside_f(rb2);
                        // every ballot next to its comparison, where it is that comparison's own lane mask
                        const unsigned long long in_a = __ballot(na), in_b = __ballot(nb);
                        const unsigned long long ma = in_a & __ballot(oa);
                        const unsigned long long mb = in_b & __ballot(ob);
                        // unsure lanes: up to r2_out and not inside
                        if (((__ballot(ra <= r2_out) & ~in_a) | (__ballot(rb2 <= r2_out) & ~in_b)) == 0) {
                            commit(sa, na && oa, ma);
                            commit(sb, nb && ob, mb);
                        } else {
                            const bool xa = resolve_f(sa, ra, unsure_f(ra)), xb = resolve_f(sb, rb2, unsure_f(rb2));
                            commit(sa, xa && oa, __ballot(xa && oa));
                            commit(sb, xb && ob, __ballot(xb && ob));
                        }
                    }
                    if (q0 < len) {
                        const int sa = lists[off + q0 + lane];
                        const float ra = dist_f(sa, interior_tag);
                        const bool ua = unsure_f(ra), oa = upper ? sa > slot_i : sa < slot_i;
                        const bool na = __ballot(ua) == 0 ? inside_f(ra) : resolve_f(sa, ra, ua);
                        commit(sa, na && oa, __ballot(na && oa));
                    }
                    return;
                }
                for (; q0 + 2 * WAVE <= len; q0 += 2 * WAVE) {
                    const int sa = lists[off + q0 + lane], sb = lists[off + q0 + WAVE + lane];
                    const bool na = near(sa, interior_tag), nb = near(sb, interior_tag);
                    const bool oa = upper ? sa > slot_i : sa < slot_i, ob = upper ? sb > slot_i : sb < slot_i;
                    // ballots of the two compares separately: each is the compare's own lane mask
                    const unsigned long long ma = __ballot(na) & __ballot(oa), mb = __ballot(nb) & __ballot(ob);
                    commit(sa, na && oa, ma);
                    commit(sb, nb && ob, mb);
                }
                if (q0 < len) {
                    const int sa = lists[off + q0 + lane];
                    const bool na = near(sa, interior_tag), oa = upper ? sa > slot_i : sa < slot_i;
                    commit(sa, na && oa, __ballot(na) & __ballot(oa));
                }
            };
            auto both = [&](auto interior_tag) {
                walk(__builtin_amdgcn_readfirstlane(lst[cl][0]), __builtin_amdgcn_readfirstlane(lst[cl][1]),
                     std::true_type{}, interior_tag);
                const int h = cnt;
                walk(__builtin_amdgcn_readfirstlane(lst[cl][2]), __builtin_amdgcn_readfirstlane(lst[cl][3]),
                     std::false_type{}, interior_tag);
                return h;
            };
            // FH: the interior tag marks a compact frame
            const int hcnt = (FH ? compact : interior) ? both(std::true_type{}) : both(std::false_type{});
            write_row(i, rb, cnt, hcnt);
        };
        // Two particles of one own cell in one wavefront (FH): they share the cell's candidate lists, so each batch's
        // slots and positions are read once and tested against both, and the two rows' rankings, independent of each
        // other, interleave. Each row is the one `single` builds: the same candidates, tests and order. `flat_tag`: a
        // compact frame (no image term), as in `single`.
        auto pair = [&](size_t p, auto flat_tag) {
            if constexpr (FH) {
                const size_t iA = i0 + p, iB = iA + 1;
                const int cl = __builtin_amdgcn_readfirstlane(p_cl[p]);
                const int slotA = __builtin_amdgcn_readfirstlane(p_slot[p]);
                const int slotB = __builtin_amdgcn_readfirstlane(p_slot[p + 1]);
                const float* const qa = coords_f(slotA);
                const float* const qb = coords_f(slotB);
                const float xA = qa[0], yA = qa[TILE_HALO_MAX], zA = qa[2 * TILE_HALO_MAX];
                const float xB = qb[0], yB = qb[TILE_HALO_MAX], zB = qb[2 * TILE_HALO_MAX];
                uint16_t* const rbA = sm.rowbuf[wave][0];
                uint16_t* const rbB = sm.rowbuf[wave][1];
                int cntA = 0, cntB = 0;
                auto inside_f = [&](float d) { return d < r2_in; };
                auto unsure_f = [&](float d) { return d <= r2_out && !(d < r2_in); };
                auto resolve_f = [&](size_t i, int sl, float d, bool unsure) {
                    bool in = inside_f(d);
                    if (unsure) {
                        const double4 ri = pos[i];
                        const double4 rj = pos[halo_particle(sl / LIST_BYTES, hr, nr)];
                        const double ex = min_image(ri.x - rj.x, box, inv_box);
                        const double ey = min_image(ri.y - rj.y, box, inv_box);
                        const double ez = min_image(ri.z - rj.z, box, inv_box);
                        in = ex * ex + ey * ey + ez * ez < radius2;
                    }
                    return in;
                };
                auto commit = [&](size_t i, uint16_t* rb, int& cnt, int sl, bool accept, unsigned long long mask) {
                    if (accept) {
                        const unsigned at = __builtin_amdgcn_mbcnt_hi(
                            static_cast<unsigned>(mask >> 32),
                            __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), static_cast<unsigned>(cnt)));
                        if (at < static_cast<unsigned>(padding)) {
                            rb[at] = static_cast<uint16_t>(sl);
                            if (write32)
                                rows[i + static_cast<uint64_t>(at) * stride] = halo_particle(sl / LIST_BYTES, hr, nr);
                        }
                    }
                    cnt = __builtin_amdgcn_readfirstlane(cnt + __popcll(mask));
                };
                // the float r^2 of a candidate at (cx, cy, cz): in a compact frame the difference itself, otherwise its
                // float minimum image per component (single's dist_f)
                auto r2 = [&](float x, float y, float z, float cx, float cy, float cz) {
                    float dx = x - cx, dy = y - cy, dz = z - cz;
                    if constexpr (!decltype(flat_tag)::value) {
                        dx = fmaf(-box_f, __builtin_rintf(dx * inv_box_f), dx);
                        dy = fmaf(-box_f, __builtin_rintf(dy * inv_box_f), dy);
                        dz = fmaf(-box_f, __builtin_rintf(dz * inv_box_f), dz);
                    }
                    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                };
                struct Cand {
                    int s;
                    float x, y, z;
                };
                auto load = [&](int sl) {
                    const float* const c = coords_f(sl);
                    return Cand{sl, c[0], c[TILE_HALO_MAX], c[2 * TILE_HALO_MAX]};
                };
                // one particle's test of a pair of batches, as in `single`
                auto test2 = [&](size_t i, uint16_t* rb, int& cnt, int slot, float x, float y, float z, const Cand& a,
                                 const Cand& b, auto upper_tag) {
                    constexpr bool upper = decltype(upper_tag)::value;
                    const float ra = r2(x, y, z, a.x, a.y, a.z), rb2 = r2(x, y, z, b.x, b.y, b.z);
                    const bool oa = upper ? a.s > slot : a.s < slot, ob = upper ? b.s > slot : b.s < slot;
                    const bool na = inside_f(ra), nb = inside_f(rb2);
                    const unsigned long long in_a = __ballot(na), in_b = __ballot(nb);
                    const unsigned long long ma = in_a & __ballot(oa);
                    const unsigned long long mb = in_b & __ballot(ob);
                    if (((__ballot(ra <= r2_out) & ~in_a) | (__ballot(rb2 <= r2_out) & ~in_b)) == 0) {
                        commit(i, rb, cnt, a.s, na && oa, ma);
                        commit(i, rb, cnt, b.s, nb && ob, mb);
                    } else {
                        const bool xa = resolve_f(i, a.s, ra, unsure_f(ra)), xb = resolve_f(i, b.s, rb2, unsure_f(rb2));
                        commit(i, rb, cnt, a.s, xa && oa, __ballot(xa && oa));
                        commit(i, rb, cnt, b.s, xb && ob, __ballot(xb && ob));
                    }
                };
                auto test1 = [&](size_t i, uint16_t* rb, int& cnt, int slot, float x, float y, float z, const Cand& a,
                                 auto upper_tag) {
                    constexpr bool upper = decltype(upper_tag)::value;
                    const float ra = r2(x, y, z, a.x, a.y, a.z);
                    const bool ua = unsure_f(ra), oa = upper ? a.s > slot : a.s < slot;
                    const bool na = __ballot(ua) == 0 ? inside_f(ra) : resolve_f(i, a.s, ra, ua);
                    commit(i, rb, cnt, a.s, na && oa, __ballot(na && oa));
                };
                auto walk2 = [&](int off, int len, auto upper_tag) {
                    int q0 = 0;
                    for (; q0 + 2 * WAVE <= len; q0 += 2 * WAVE) {
                        const Cand a = load(lists[off + q0 + lane]), b = load(lists[off + q0 + WAVE + lane]);
                        test2(iA, rbA, cntA, slotA, xA, yA, zA, a, b, upper_tag);
                        test2(iB, rbB, cntB, slotB, xB, yB, zB, a, b, upper_tag);
                    }
                    if (q0 < len) {
                        const Cand a = load(lists[off + q0 + lane]);
                        test1(iA, rbA, cntA, slotA, xA, yA, zA, a, upper_tag);
                        test1(iB, rbB, cntB, slotB, xB, yB, zB, a, upper_tag);
                    }
                };
                walk2(__builtin_amdgcn_readfirstlane(lst[cl][0]), __builtin_amdgcn_readfirstlane(lst[cl][1]),
                      std::true_type{});
                const int hA = cntA, hB = cntB;
                walk2(__builtin_amdgcn_readfirstlane(lst[cl][2]), __builtin_amdgcn_readfirstlane(lst[cl][3]),
                      std::false_type{});
                write_row(iA, rbA, cntA, hA);
                write_row(iB, rbB, cntB, hB);
            }
        };
        // wavefront w takes particles 2 w, 2 w + 1, then 16 on: a pair of one cell together (FH), otherwise one after
        // the other
        const size_t np = i1 - i0;
        for (size_t q = 2 * static_cast<size_t>(wave); q < np; q += 2 * (TILE_THREADS / WAVE)) {
            bool two = false;
            if constexpr (FH) two = q + 1 < np && p_cl[q] == p_cl[q + 1];
            if (__builtin_amdgcn_readfirstlane(two ? 1 : 0)) {
                if (compact)
                    pair(q, std::true_type{});
                else
                    pair(q, std::false_type{});
            } else {
                const size_t q1 = q + 2 < np ? q + 2 : np;
#pragma unroll 1
                for (size_t p = q; p < q1; ++p) single(p);  // one instance of the single walk's code
            }
        }
    } else {
        const int lane = threadIdx.x & (LANES - 1);
        const size_t i = i0 + (threadIdx.x >> 2);
        if (i < i1) {
            LaneScan sc;
            sc.xs = xs;
            sc.ys = ys;
            sc.zs = zs;
            sc.pos = pos;
            sc.ri = pos[i];
            sc.box = box;
            sc.inv_box = inv_box;
            sc.radius2 = radius2;
            sc.stride = stride;
            sc.i = i;
            sc.lane = lane;
            sc.group_shift = (threadIdx.x & (WAVE - 1)) & ~(LANES - 1);
            sc.lower_mask = (1u << lane) - 1u;
            sc.padding = padding;
            sc.row = rows + i;
            sc.row16 = rows16 + slot_row_base(i, row_len);
            sc.cnt = 0;
            const uint32_t self = static_cast<uint32_t>(i);
            int hcnt = 0;
            // STAGED: positions from the LDS halo; SLOTS: a staged tile, whose slot rows are written too
            auto run = [&](auto staged_tag, auto slots_tag) {
                constexpr bool STAGED = decltype(staged_tag)::value;
                constexpr bool SLOTS = decltype(slots_tag)::value;
                auto base_of = [&](int nc) {
                    if constexpr (SLOTS)
                        return slot_base_of_cell(nc, hr, start, nr);
                    else
                        return 0;
                };
                if (sys.use_cells) {
                    const int c = cell_of[i];
                    const int cx = c / (s * s), cy = (c / s) % s, cz = c % s;
#pragma unroll 1
                    for (int st = 0; st < 27; ++st) {  // j > i: the reference's stencil order
                        const int nc = (wrap_index(cx + st / 9 - 1, s) * s + wrap_index(cy + (st / 3) % 3 - 1, s)) * s +
                                       wrap_index(cz + st % 3 - 1, s);
                        if (nc < c) continue;
                        const int base = base_of(nc);
                        const uint32_t b = start[nc], e = start[nc + 1];
                        if (nc == c)
                            sc.scan<STAGED, false, SLOTS>(self + 1, e, base + static_cast<int>(self + 1 - b));
                        else
                            sc.scan<STAGED, false, SLOTS>(b, e, base);
                    }
                    hcnt = sc.cnt;
                    const Sorted3 xo(wrap_index(cx - 1, s), cx, wrap_index(cx + 1, s));
                    const Sorted3 yo(wrap_index(cy - 1, s), cy, wrap_index(cy + 1, s));
                    const Sorted3 zo(wrap_index(cz - 1, s), cz, wrap_index(cz + 1, s));
#pragma unroll 1
                    for (int st = 0; st < 27; ++st) {  // j < i: cells in increasing id
                        const int nc = (xo[st / 9] * s + yo[(st / 3) % 3]) * s + zo[st % 3];
                        if (nc > c) break;
                        sc.scan<STAGED, false, SLOTS>(start[nc], nc == c ? self : start[nc + 1], base_of(nc));
                    }
                } else {
                    // every particle of the system is in the single staged range, at slot = its index - first
                    sc.scan<STAGED, false, SLOTS>(self + 1, n32, static_cast<int>(self + 1 - sys.first));
                    hcnt = sc.cnt;
                    sc.scan<STAGED, false, SLOTS>(sys.first, self, 0);
                }
            };
            if (staged && !FH)
                run(std::true_type{}, std::true_type{});
            else if (staged)  // FH: the halo holds floats; positions from global memory, slots from the ranges
                run(std::false_type{}, std::true_type{});
            else
                run(std::false_type{}, std::false_type{});
            if (lane == 0) {
                count[i] = sc.cnt < padding ? sc.cnt : padding;
                half[i] = hcnt < padding ? hcnt : padding;
                ep_half = static_cast<unsigned>(hcnt);
            }
            if (staged) {  // the slot row's last chunk ends in ghost entries, as on the walk above
                const int stored = sc.cnt < padding ? sc.cnt : padding;
                for (int k = stored + lane; k < (stored + CHUNK - 1) / CHUNK * CHUNK; k += LANES)
                    sc.row16[slot_entry(static_cast<uint32_t>(k))] = static_cast<uint16_t>(GHOST_SLOT * SLOT_BYTES);
            }
            ep_max = sc.cnt;
        }
    }
    if (cb_live) {
        if (cb_group == 0) {
            cb.pos[cb_i] = make_double4(__longlong_as_double(cbv[0]), __longlong_as_double(cbv[1]),
                                        __longlong_as_double(cbv[2]), __longlong_as_double(cbv[3]));
        } else if (cb_group == 1) {
            for (int c = 0; c < 3; ++c) cb.vel[c * stride + cb_i] = __longlong_as_double(cbv[c]);
        } else if (cb_group == 2) {
            for (int c = 0; c < 3; ++c) cb.frc[c * stride + cb_i] = __longlong_as_double(cbv[c]);
        } else {
            cb.mass[cb_i] = __longlong_as_double(cbv[0]);
            cb.id[cb_i] = cbv[1];
            cb.cell_of[cb_i] = static_cast<int>(static_cast<uint32_t>(cbv[2]));
        }
    }
    block_build_stat<TILE_THREADS>(ep_max, ep_half, stats + tile);
}

// six wavefronts per SIMD (three blocks per CU, which the float halo's LDS allows) cap the registers at 80
template <bool FH, bool BATCH>
__global__ void __launch_bounds__(TILE_THREADS, 6)
    k_list_build_tiled(const double4* __restrict__ pos, const int* __restrict__ cell_of,
                       const uint32_t* __restrict__ start, TileSystems ts, double radius2, int padding, int write32,
                       int row_len, int32_t* __restrict__ rows, uint16_t* __restrict__ rows16,
                       int32_t* __restrict__ count, int32_t* __restrict__ half, int2* __restrict__ stats,
                       const int4* __restrict__ tile_ranges, const int* __restrict__ tile_nr, TiledCopyBack cb,
                       const int* gate, const int* halt) {
    if (gated_off(gate, halt)) return;
    int tile;
    if constexpr (BATCH) {  // the k-th rebuilt tile, k spread over the XCDs in contiguous runs
        const int rebuilt = uniform_i32(ts.active[0]);
        if (static_cast<int>(blockIdx.x) >= rebuilt) return;
        tile = uniform_i32(ts.active[1 + xcd_item(static_cast<int>(blockIdx.x), rebuilt)]);
    } else {
        tile = launch_tile(ts, static_cast<int>(blockIdx.x));
    }
    const TileSystem sys = tile_system<BATCH>(ts, tile);
    if (tile_gated_off<BATCH>(ts, sys) || static_cast<size_t>(tile) * TILE >= sys.end) return;  // not rebuilt, or empty
    extern __shared__ double lds[];
    list_build_tile<FH>(tile, lds, pos, cell_of, start + sys.cell_base, sys, ts.stride, radius2, padding, write32,
                        row_len, rows, rows16, count, half, stats, tile_ranges, tile_nr, cb);
}

// int32 rows from the 16-bit slot rows (builds that wrote only the slots): slot -> particle through the tile's
// halo ranges. Unstaged tiles already hold int32 rows.
template <bool BATCH>
__global__ void __launch_bounds__(TILE_THREADS)
    k_rows_from_slots(TileSystems ts, const uint16_t* __restrict__ rows16, int row_len,
                      const int32_t* __restrict__ count, const int4* __restrict__ tile_ranges,
                      const int* __restrict__ tile_nr, int32_t* __restrict__ rows) {
    __shared__ int ra[RMAX], rgs[RMAX], rofs[RMAX + 1];
    const HaloRanges hr{ra, rgs, rofs};
    const int tile = blockIdx.x;
    const size_t n = tile_system<BATCH>(ts, tile).end, stride = ts.stride;
    const int nr = tile_range_count(tile_nr[tile]);
    if (nr <= 0) return;
    load_ranges(tile_ranges, tile, nr, hr);
    __syncthreads();
    const size_t i = static_cast<size_t>(tile) * TILE + (threadIdx.x >> 2);
    const int lane = threadIdx.x & (LANES - 1);
    if (i >= n) return;
    const int m = count[i];
    const uint16_t* row16 = rows16 + slot_row_base(i, row_len);
    for (int k = lane; k < m; k += LANES)
        rows[static_cast<size_t>(k) * stride + i] =
            halo_particle(row16[slot_entry(static_cast<uint32_t>(k))] / SLOT_BYTES, hr, nr);
}

constexpr int UNSTAGED_BATCH = 2;  // row entries per lane in flight on the force kernel's unstaged path

// Frame of the mixed force kernel: positions are staged relative to the tile's first particle (its nearest image),
// in double, so a separation is one double subtraction.
struct TileFrame {
    double box, inv_box;
    double reach;  // rc + list radius + skin, for the compactness test
};

// Arguments of the step that the force kernel can carry out for its own particles (see k_forces_tiled_mixed).
struct ForceStep {
    double4* pos_next;     // positions after the drift (a buffer other than the one being read)
    double* vel;           // velocities, x | y | z blocks of the stride
    const double* mass;
    const double4* ref;    // positions at the last list build
    double half_dt, dt, max_disp2;
    int* stale_next;       // set when a particle has moved more than half the skin
};

__device__ __forceinline__ double frame_coord(double x, double ref, const TileFrame& fr) {
    const double t = x - ref;
    return fma(-fr.box, rint(t * fr.inv_box), t);
}

// Mixed-precision forces over the tiled list. Lane `lane` of particle i takes row entries k = lane, lane + 4, ...:
// in staged tiles it reads 8 slots per 16-byte load (the next chunk's load is issued before the current chunk is
// evaluated) and the frame positions from LDS; unstaged tiles gather positions through the int32 rows. In a
// compact tile (its own particles within e of the reference particle, with 2e + reach < L - 1) every listed
// separation is already the minimum image; otherwise it is taken per pair.
//
// STEP: the velocity-Verlet step around the force evaluation, done for the tile's own particles as soon as their
// forces are summed (NVE md_run, mixed): the closing kick of this step and its kinetic energy, the opening kick and
// drift of the next step, and the displacement test of the next list check (the arithmetic of k_kick_kick_drift).
// The new positions go to a second buffer, since other tiles' halos still read this step's; the forces themselves
// are not stored.
// Six wavefronts per SIMD (three blocks per CU, as the halo's LDS allows) cap the registers at 80.
template <bool STEP, bool BATCH, bool UNIT>
__global__ void __launch_bounds__(TILE_THREADS, 6)
    k_forces_tiled_mixed(const double4* __restrict__ pos, TileSystems ts, const uint16_t* __restrict__ rows16,
                         const int32_t* __restrict__ rows, const int32_t* __restrict__ count,
                         const int4* __restrict__ tile_ranges, const int* __restrict__ tile_nr, int row_len,
                         LJConst lj, double* __restrict__ frc, double* __restrict__ partials,
                         int* __restrict__ clear_flag, ForceStep fs, const int* halt) {
    // The halo is a static array: its address is known when the kernel is compiled, so a slot's byte offset, as the
    // rows hold it, is the whole address computation.
    __shared__ double halo[3][TILE_HALO_MAX];
    double* xs = halo[0];
    double* ys = halo[1];
    double* zs = halo[2];
    // per particle: U, W and inside count of its row (and with STEP its kinetic energy), in the halo's memory once
    // the pair loop is done
    constexpr int SUMS = STEP ? 4 : 3;
    double(*e_part)[TILE] = reinterpret_cast<double(*)[TILE]>(&halo[0][0]);
    static_assert(4 * TILE <= TILE_HALO_MAX, "the per-particle sums fit in the halo's x array");
    __shared__ int wave_extent[TILE_THREADS / WAVE];
    const int tile = launch_tile(ts, static_cast<int>(blockIdx.x));
    const TileSystem sys = tile_system<BATCH>(ts, tile);
    const TileFrame fr{sys.box, sys.inv_box, sys.reach};
    const size_t n = sys.end, pstride = ts.stride;  // pstride: column / component stride
    const int raw_nr = tile_nr[tile];
    // halo range `wave lane` of the tile, in every wavefront (the staging below reads the ranges from registers;
    // ranges past nr are unused)
    const int4 rr = tile_ranges[static_cast<size_t>(tile) * RMAX + threadIdx.x % WAVE];
    const int nr = tile_range_count(raw_nr);
    const bool staged = nr > 0;
    const bool flat = tile_is_flat(raw_nr);
    const size_t i0 = static_cast<size_t>(tile) * TILE;
    const int lane = threadIdx.x & (LANES - 1);
    const size_t i = i0 + (threadIdx.x >> 2);
    const bool active = i < n;
    // Header loads, unconditional so that no branch separates them: lanes past n read particle i0's count and
    // position, which they do not use.
    const size_t ic = active ? i : i0;
    const int m_read = count[ic];  // row length, loaded first: the slot loads below depend on it
    const double4 ref = pos[i0];
    const double4 ri = pos[ic];  // used once the halo is staged
    const int m = active ? m_read : 0;
    // The lane's slot words of the row's first PRE_CHUNKS chunks, issued before the halo's loads, so both are in
    // flight together and the pair loop runs from registers (longer rows load their remaining chunks in the loop);
    // chunks past the row are ghost words. The loads are unconditional: a chunk index past the padded row (which
    // every slot of the tile has, unstaged tiles included) is clamped to its last chunk, whose words go unused.
    const int chunks = (m + CHUNK - 1) / CHUNK;
    const uint4* rp = reinterpret_cast<const uint4*>(rows16 + slot_row_base(i, row_len) + lane * 8);
    constexpr int stride = CHUNK_STRIDE / 8;  // uint4 words between chunks
    constexpr unsigned GHOST_WORD = (GHOST_SLOT * SLOT_BYTES) * 0x10001u;  // two ghost entries
    const uint4 ghost4 = make_uint4(GHOST_WORD, GHOST_WORD, GHOST_WORD, GHOST_WORD);
    uint4 pre[PRE_CHUNKS];
#pragma unroll
    for (int c = 0; c < PRE_CHUNKS; ++c) pre[c] = rp[max(0, min(c, row_len / CHUNK - 1)) * stride];  // row_len >= CHUNK
#pragma unroll
    for (int c = 0; c < PRE_CHUNKS; ++c) {
        const bool live = staged && c < chunks;
        pre[c].x = live ? pre[c].x : GHOST_WORD;
        pre[c].y = live ? pre[c].y : GHOST_WORD;
        pre[c].z = live ? pre[c].z : GHOST_WORD;
        pre[c].w = live ? pre[c].w : GHOST_WORD;
    }
    if (threadIdx.x == 0) {
        // the ghost slot, far from every staged coordinate: its pair term is exactly zero (see the compact walk)
        xs[GHOST_SLOT] = GHOST_COORD;
        ys[GHOST_SLOT] = GHOST_COORD;
        zs[GHOST_SLOT] = GHOST_COORD;
    }
    if (staged && flat) {
        // a flat tile: no periodic image anywhere in the tile and its halo, so the raw positions are staged and each
        // separation is the difference of raw coordinates, the reference's minimum image exactly
        stage_halo_lanes<TILE_THREADS / WAVE>(pos, rr, nr, [&](int q, const double4& p) {
            xs[q] = p.x;
            ys[q] = p.y;
            zs[q] = p.z;
        });
    } else if (staged) {
        stage_halo_lanes<TILE_THREADS / WAVE>(pos, rr, nr, [&](int q, const double4& p) {
            xs[q] = frame_coord(p.x, ref.x, fr);
            ys[q] = frame_coord(p.y, ref.y, fr);
            zs[q] = frame_coord(p.z, ref.z, fr);
        });
    }
    // the particle's own coordinates: raw in a flat tile, in the tile's frame otherwise
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (active) {
        if (flat) {
            xi = ri.x;
            yi = ri.y;
            zi = ri.z;
        } else {
            xi = frame_coord(ri.x, ref.x, fr);
            yi = frame_coord(ri.y, ref.y, fr);
            zi = frame_coord(ri.z, ref.z, fr);
        }
    }
    {  // how far the tile's own particles lie from the reference particle (frames only): per wavefront, then over
       // the block
        const float e = active && !flat ? static_cast<float>(fmax(fabs(xi), fmax(fabs(yi), fabs(zi)))) : 0.0f;
        const int e_wave = wave_max_dpp(__float_as_int(e));  // non-negative floats order as their bit patterns
        if (threadIdx.x % WAVE == 0) wave_extent[threadIdx.x / WAVE] = e_wave;
    }
    __syncthreads();
    int e_bits = 0;
#pragma unroll
    for (int w = 0; w < TILE_THREADS / WAVE; ++w) e_bits = max(e_bits, wave_extent[w]);
    const float e = __int_as_float(e_bits);
    const bool compact =
        staged && (flat || 2.0f * e + static_cast<float>(fr.reach) < static_cast<float>(fr.box) - 1.0f);
    PairSumsF a;
    // inside pairs counted per wavefront on the compact walk (a uniform sum of ballots)
    unsigned wave_inside = 0;
    // STEP: the particle's mass and its lane's component of velocity, position and list-build position (lane 3, whose
    // step only repeats z's, takes the position's fourth word, which the drift copies). The compact walk issues these
    // loads before its last preloaded chunk, when the slot words of the chunks already walked no longer hold
    // registers, so they arrive while that chunk is evaluated; the other walks load them afterwards.
    const int comp = lane < 3 ? lane : 2;
    double st_m = 0.0, st_v = 0.0, st_r = 0.0, st_ref = 0.0;
    auto load_step = [&]() {
        if constexpr (STEP) {
            if (active) {
                st_m = fs.mass[i];
                st_v = fs.vel[comp * pstride + i];
                if constexpr (!BATCH) {
                    st_r = reinterpret_cast<const double*>(pos + i)[lane];
                    st_ref = reinterpret_cast<const double*>(fs.ref + i)[comp];
                }
            }
        }
    };
    if (compact) {
        // Compact tile: every listed separation is already the minimum image. Each lane walks whole chunks: the
        // entries past its row read the ghost slot (the build ends every row's last chunk with ghost entries, and
        // chunks past the row are ghost words), whose separation is finite and beyond the cut-off, so its term
        // adds exactly zero to every sum. The wavefront stops at its longest row, a uniform bound: no per-entry
        // guard. Inactive lanes (past the system's end) walk ghost words only.
        const int mw = wave_max_dpp(m);
        const int cw = (mw + CHUNK - 1) / CHUNK;
        auto chunk_pairs = [&](const uint4& cw4, int c) {
            const unsigned words[4] = {cw4.x, cw4.y, cw4.z, cw4.w};
            // entries two at a time (both halves of a word): two independent chains to interleave; past the
            // wavefront's longest row both are ghosts for every lane
#pragma unroll
            for (int t = 0; t < 8; t += 2) {
                if (c * CHUNK + t * LANES >= mw) break;  // no lane of the wavefront has entry c * CHUNK + t * LANES on
                const unsigned w = words[t >> 1];
                const double* xa = reinterpret_cast<const double*>(reinterpret_cast<const char*>(xs) + (w & 0xffffu));
                const double* xb = reinterpret_cast<const double*>(reinterpret_cast<const char*>(xs) + (w >> 16));
                const double ax = xi - xa[0], ay = yi - xa[TILE_HALO_MAX], az = zi - xa[2 * TILE_HALO_MAX];
                const double bx = xi - xb[0], by = yi - xb[TILE_HALO_MAX], bz = zi - xb[2 * TILE_HALO_MAX];
                const bool in_a = mixed_pair_core<UNIT>(ax, ay, az, lj, a);
                const bool in_b = mixed_pair_core<UNIT>(bx, by, bz, lj, a);
                wave_inside += static_cast<unsigned>(__popcll(__ballot(in_a)) + __popcll(__ballot(in_b)));
            }
        };
#pragma unroll
        for (int c = 0; c < PRE_CHUNKS; ++c) {
            if (c == PRE_CHUNKS - 1) load_step();
            if (c < cw) chunk_pairs(pre[c], c);
        }
        for (int c = PRE_CHUNKS; c < cw; ++c) chunk_pairs(c < chunks ? rp[c * stride] : ghost4, c);
    } else if (active) {
        if (staged) {
            // staged, not compact: the minimum image of each separation, entries guarded per lane
            auto chunk_pairs = [&](const uint4& cw, int c) {
                const unsigned words[4] = {cw.x, cw.y, cw.z, cw.w};
                const int rem = m - c * CHUNK - lane;  // entry c * CHUNK + t * LANES + lane exists iff t * LANES < rem
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    if (t * LANES < rem) {
                        const unsigned off = __builtin_amdgcn_ubfe(words[t >> 1], (t & 1) * 16, 16);
                        const double* xj = reinterpret_cast<const double*>(reinterpret_cast<const char*>(xs) + off);
                        double dx = xi - xj[0], dy = yi - xj[TILE_HALO_MAX], dz = zi - xj[2 * TILE_HALO_MAX];
                        dx = fma(-fr.box, rint(dx * fr.inv_box), dx);
                        dy = fma(-fr.box, rint(dy * fr.inv_box), dy);
                        dz = fma(-fr.box, rint(dz * fr.inv_box), dz);
                        mixed_pair_term_f<UNIT>(dx, dy, dz, lj, a);
                    }
                }
            };
#pragma unroll
            for (int c = 0; c < PRE_CHUNKS; ++c) chunk_pairs(pre[c], c);
            for (int c = PRE_CHUNKS; c < chunks; ++c) chunk_pairs(rp[c * stride], c);
        } else {
            // unstaged tile: neighbours gathered through the int32 rows, UNSTAGED_BATCH entries of the lane at a time
            // (their row words, then their positions, then their terms in entry order), so the two dependent loads
            // of an entry overlap with those of the others
            const double4 ri = pos[i];
            for (int k0 = lane; k0 < m; k0 += UNSTAGED_BATCH * LANES) {
                int j[UNSTAGED_BATCH];
#pragma unroll
                for (int b = 0; b < UNSTAGED_BATCH; ++b) {
                    const int k = k0 + b * LANES;
                    j[b] = k < m ? rows[static_cast<size_t>(k) * pstride + i] : -1;
                }
                double xj[UNSTAGED_BATCH], yj[UNSTAGED_BATCH], zj[UNSTAGED_BATCH];  // plain doubles stay in registers
#pragma unroll
                for (int b = 0; b < UNSTAGED_BATCH; ++b) {
                    if (j[b] < 0) continue;
                    const double4 r = pos[j[b]];
                    xj[b] = r.x;
                    yj[b] = r.y;
                    zj[b] = r.z;
                }
#pragma unroll
                for (int b = 0; b < UNSTAGED_BATCH; ++b)
                    if (j[b] >= 0)
                        mixed_pair_gather_f<UNIT>(ri, make_double4(xj[b], yj[b], zj[b], 0.0), fr.box, fr.inv_box, lj,
                                                  a);
            }
        }
    }
    if (!compact) load_step();
    // A halted run (a list build of this step overflowed) stores nothing: the state stays at the last valid step.
    // Tested here rather than first, so the halt word's load does not delay the tile's loads; the walk above
    // read the truncated rows of such a build, which stay within the tile's halo and rows.
    if (halted(halt)) return;
    // the particle's four lanes are a quad of the wavefront: its partner lanes' values (lane ^ 1, then lane ^ 2)
    // arrive by DPP quad permutes, with no LDS round trip
    auto quad_sums = [&](auto ctrl) {
        constexpr int CTRL = decltype(ctrl)::value;
        a.gx += dpp_f64<CTRL>(a.gx);
        a.gy += dpp_f64<CTRL>(a.gy);
        a.gz += dpp_f64<CTRL>(a.gz);
        a.t += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a.t), CTRL, 0xF, 0xF, false));
        a.s += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a.s), CTRL, 0xF, 0xF, false));
        a.c += __builtin_amdgcn_update_dpp(0, a.c, CTRL, 0xF, 0xF, false);
    };
    quad_sums(std::integral_constant<int, QUAD_XOR1>{});
    quad_sums(std::integral_constant<int, QUAD_XOR2>{});
    // every lane now holds the particle's sums; lane c (c < 3) takes force component c (lane 3 repeats z)
    const double c24 = static_cast<double>(lj.c24f);
    const double fc = c24 * (comp == 0 ? a.gx : (comp == 1 ? a.gy : a.gz));
    const int pl = threadIdx.x >> 2;
    double ke = 0.0;
    if constexpr (STEP) {
        // the arithmetic of k_kick_kick_drift, one component per lane; the kinetic energy and the displacement test
        // gather the three components on lane 0 and add them in x, y, z order
        const double m = st_m;
        double kick = 0.0, v = 0.0, d = 0.0, r = 0.0;
        // x / 1 = x exactly: a wavefront of unit masses skips the division
        const bool unit_mass = __ballot(active && m != 1.0) == 0;
        if (active) {
            kick = unit_mass ? fs.half_dt * fc : fs.half_dt * fc / m;
            v = st_v + kick;  // closing kick of this step
        }
        const double vy = dpp_f64<QUAD_XOR1>(v), vz = dpp_f64<QUAD_XOR2>(v);
        if (active) ke = 0.5 * m * (v * v + vy * vy + vz * vz);  // lane 0's is the particle's
        if constexpr (BATCH) {
            // batched replicas (Bussi): the closing kick only, since the thermostat scales each replica's velocities
            // once its kinetic energy is summed; the next step's opening kick reads the stored force
            if (active && lane < 3) fs.vel[comp * pstride + i] = v;
        } else {
        if (active) {
            v = v + kick;                                 // opening kick of the next step
            r = wrap_coord(st_r + fs.dt * v, fr.box, fr.inv_box);
            d = min_image(r - st_ref, fr.box, fr.inv_box);
        }
        const double dy = dpp_f64<QUAD_XOR1>(d), dz = dpp_f64<QUAD_XOR2>(d);
        if (active) {
            double* next = reinterpret_cast<double*>(fs.pos_next + i);
            if (lane < 3) {
                fs.vel[comp * pstride + i] = v;
                next[lane] = r;
            } else {
                next[3] = st_r;
            }
            if (lane == 0 && d * d + dy * dy + dz * dz > fs.max_disp2) *fs.stale_next = 1;
        }
        }
    }
    if (active && (!STEP || BATCH) && lane < 3) frc[lane * pstride + i] = fc;
    __syncthreads();  // every wavefront is done with the halo
    if (lane == 0) {
        // every pair was met from both ends: half of each full-row sum (zero for the slots past n); the inside
        // count of the particle's row (other walks) or of the wavefront's rows (the compact walk, on its first lane)
        const double u = static_cast<double>(lj.c4f) * (static_cast<double>(a.t) - static_cast<double>(a.s));
        const unsigned inside = static_cast<unsigned>(a.c) + (threadIdx.x % WAVE == 0 ? wave_inside : 0u);
        e_part[0][pl] = 0.5 * u;
        e_part[1][pl] = 0.5 * c24 * static_cast<double>(a.t);
        e_part[2][pl] = 0.5 * static_cast<double>(inside);
        if constexpr (STEP) e_part[3][pl] = ke;
    }
    if (clear_flag && blockIdx.x == 0 && threadIdx.x == 0) *clear_flag = 0;
    __syncthreads();
    // the tile's U, W and inside count (and KE): the 128 particle sums added in double by one wavefront, in a fixed
    // order (DPP moves, no LDS round trips); the potential's shift once per inside pair (the halved count is exact)
    if (threadIdx.x < WAVE) {
        double v[SUMS];
#pragma unroll
        for (int q = 0; q < SUMS; ++q) v[q] = wave_sum_dpp(e_part[q][threadIdx.x] + e_part[q][threadIdx.x + WAVE]);
        if (threadIdx.x == 0) {
            double* out = partials + static_cast<size_t>(tile) * SUMS;
            out[0] = v[0] - static_cast<double>(lj.shiftf) * v[2];
#pragma unroll
            for (int q = 1; q < SUMS; ++q) out[q] = v[q];
        }
    }
}

// ---- FP64 forces over the tiled rows ---------------------------------------------------------------------
//
// The halo holds the raw positions (x | y | z doubles, the same slots as the mixed kernel), so each separation is
// the reference's minimum image of r_i - r_j (Forcefield.cpp:214-326), exactly; along an axis where the tile is
// flat no image can enter and the separation component is the difference itself (the image term would be
// d - L * 0 = d).
//
// Two lanes per particle: a tile's 128 particles are worked by 256 lanes (four wavefronts, three blocks per CU as
// the halo's LDS admits, so three wavefronts per SIMD). Of particle i's lane pair, the even lane evaluates the even
// row entries and the odd lane the odd ones (the even lane's words 0 and 2 of each chunk, the odd lane's words 1
// and 3); for every four entries 4g .. 4g + 3 the even lane adds t(4g), its partner's t(4g + 1) (one DPP move per
// 32-bit half), t(4g + 2) and the partner's t(4g + 3) to the particle's sums, one addition at a time in entry
// order: the sums are the gather kernel's, k_forces_fp64's, bit for bit. Against four lanes per particle this
// halves the additions and transfers that only one lane of each group uses. An entry past the row's end reads
// halo slot 0 (a finite position) with its force factor forced to zero, so its term is +-0, which leaves every sum
// unchanged (a sum that starts at +0 never becomes -0 under round to nearest); the loop body is branch free.
// U, W and the inside count are summed over the row's j > i part per lane, then per tile.

constexpr int PAIR_THREADS = 2 * TILE;

struct PairLaneSums {
    double fx = 0.0, fy = 0.0, fz = 0.0;  // valid on the particle's even lane
    double s12 = 0.0, s6 = 0.0;           // this lane's entries of the j > i part inside the cutoff
    int c = 0;
};

// One entry's term t = F/r d (fp64_pair's arithmetic): zero unless `valid`; `upper` (entry in the j > i part) adds
// the energy sums. Per pair U = 4 eps (s12 - s6) - shift and W = F r = 24 eps (2 s12 - s6), so the lanes sum s12 and
// s6 and count the pairs, and the tile's U and W are formed from the three sums. 2 s12 - s6 is one FMA: the
// doubling is exact, so it rounds once, as the reference's two operations do (for s12 below 2^1022, i.e. r above
// 2^-85 sigma). With sigma = 1 (UNIT) s2 = 1 * inv_r2 = inv_r2 exactly.
template <bool UNIT>
__device__ __forceinline__ void pair_lane_term(double dx, double dy, double dz, bool valid, bool upper,
                                               const LJConst& lj, PairLaneSums& a, double& tx, double& ty,
                                               double& tz) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double inv_r2 = recip_rn(r2);
    const double s2 = UNIT ? inv_r2 : lj.sig2 * inv_r2;
    const double s6 = s2 * s2 * s2;
    const double s12 = s6 * s6;
    const double t2 = fma(2.0, s12, -s6);
    double fr = lj.c24 * t2 * inv_r2;
    asm("" : "+v"(fr));  // computed on every lane: the select stays a select, not a branch around the arithmetic
    const bool inside = r2 < lj.rc2;
    const double f = (inside && valid) ? fr : 0.0;
    tx = f * dx;
    ty = f * dy;
    tz = f * dz;
    const bool up = inside && upper;
    a.s12 += up ? s12 : 0.0;
    a.s6 += up ? s6 : 0.0;
    a.c += up ? 1 : 0;
}

// The partner lane's value (DPP quad_perm [1, 1, 3, 3]: even lanes read the odd lane above, as two 32-bit halves).
__device__ __forceinline__ double odd_partner(double v) {
    constexpr int ctrl = 1 | (1 << 2) | (3 << 4) | (3 << 6);
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Even lane: s += t(this lane), then t(the odd partner), one addition at a time.
__device__ __forceinline__ void add_pair(PairLaneSums& a, double tx, double ty, double tz) {
    const double px = odd_partner(tx), py = odd_partner(ty), pz = odd_partner(tz);
    a.fx += tx;
    a.fy += ty;
    a.fz += tz;
    a.fx += px;
    a.fy += py;
    a.fz += pz;
}

// STEP: the velocity-Verlet step of the tile's particles folded in after their forces (NVE md_run, FP64), with
// the arithmetic of k_kick_kick_drift on the particle's even lane: the closing kick and its kinetic energy, the
// opening kick and drift of the next step into a second position buffer (other tiles' halos still read this
// step's), and the next displacement test. The forces are not stored; the tile's fourth sum is the kinetic energy.
template <bool UNIT, bool BATCH, bool STEP>
__global__ void __launch_bounds__(PAIR_THREADS, 3)
    k_forces_tiled_fp64_pair(const double4* __restrict__ pos, TileSystems ts, const uint16_t* __restrict__ rows16,
                             const int32_t* __restrict__ rows, const int32_t* __restrict__ count,
                             const int32_t* __restrict__ half, const int4* __restrict__ tile_ranges,
                             const int* __restrict__ tile_nr, int row_len, LJConst lj, double* __restrict__ frc,
                             double* __restrict__ partials, int* __restrict__ clear_flag, ForceStep fs,
                             const int* halt) {
    if (halted(halt)) return;
    constexpr int SUMS = STEP ? 4 : 3;
    __shared__ double halo[3][TILE_HALO_MAX];
    __shared__ double e_part[PAIR_THREADS / WAVE][SUMS];
    const int tile = launch_tile(ts, static_cast<int>(blockIdx.x));
    const TileSystem sys = tile_system<BATCH>(ts, tile);
    const size_t n = sys.end, pstride = ts.stride;  // pstride: column / component stride
    const double box = sys.box, inv_box = sys.inv_box;
    const int raw_nr = tile_nr[tile];
    // every wavefront holds the tile's ranges, one per lane
    const int4 rr = tile_ranges[static_cast<size_t>(tile) * RMAX + threadIdx.x % WAVE];
    const int nr = tile_range_count(raw_nr);
    const bool staged = nr > 0;
    const int wrap_axes = tile_wrap_axes(raw_nr);
    const int par = threadIdx.x & 1;  // 0: even entries, 1: odd entries
    const size_t i = static_cast<size_t>(tile) * TILE + (threadIdx.x >> 1);
    const bool active = i < n;
    const int m = active ? count[i] : 0;
    const int h = active ? half[i] : 0;
    const double4 ri = active ? pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    // STEP: the even lane's mass, velocity and list-build position, loaded with the header (the registers are there
    // at three wavefronts per SIMD), so they have arrived when the pair loop ends
    double st_m = 1.0, st_v[3] = {0.0, 0.0, 0.0};
    double4 st_ref = make_double4(0.0, 0.0, 0.0, 0.0);
    if constexpr (STEP) {
        if (active && par == 0) {
            st_m = fs.mass[i];
#pragma unroll
            for (int c = 0; c < 3; ++c) st_v[c] = fs.vel[c * pstride + i];
            st_ref = fs.ref[i];
        }
    }
    if (staged) {
        stage_halo_lanes<PAIR_THREADS / WAVE>(pos, rr, nr, [&](int q, const double4& p) {
            halo[0][q] = p.x;
            halo[1][q] = p.y;
            halo[2][q] = p.z;
        });
    }
    // the lane's words of chunk c: w0 (entries 4 g + par) and w1 (entries 4 g + par + 2), g = 0..7
    const uint4* rp = reinterpret_cast<const uint4*>(rows16 + slot_row_base(i, row_len)) + par;
    constexpr int stride = CHUNK_STRIDE / 8;  // uint4 words between chunks
    const int chunks = (m + CHUNK - 1) / CHUNK;
    uint4 w0 = staged && chunks > 0 ? rp[0] : make_uint4(0u, 0u, 0u, 0u);
    uint4 w1 = staged && chunks > 0 ? rp[2] : make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    // the wavefront walks its chunks to the end of its longest row: groups of four entries past it are skipped
    const int m_wave = wave_max_dpp(m);
    PairLaneSums a;
    if (staged) {
        const char* hb = reinterpret_cast<const char*>(&halo[0][0]);
        auto run = [&](auto wrap) {
            constexpr int W = decltype(wrap)::value;  // axes that need the minimum image
            auto term = [&](unsigned raw, int k, double& tx, double& ty, double& tz) {
                const bool valid = k < m;
                const double* xj = reinterpret_cast<const double*>(hb + (valid ? raw : 0u));
                double dx = ri.x - xj[0], dy = ri.y - xj[TILE_HALO_MAX], dz = ri.z - xj[2 * TILE_HALO_MAX];
                if constexpr ((W & 1) != 0) dx = min_image(dx, box, inv_box);
                if constexpr ((W & 2) != 0) dy = min_image(dy, box, inv_box);
                if constexpr ((W & 4) != 0) dz = min_image(dz, box, inv_box);
                pair_lane_term<UNIT>(dx, dy, dz, valid, k < h, lj, a, tx, ty, tz);
            };
#pragma unroll 1
            for (int c = 0; c < chunks; ++c) {
                const uint4 c0 = w0, c1 = w1;
                if (c + 1 < chunks) {  // the next chunk's words in flight
                    w0 = rp[(c + 1) * stride];
                    w1 = rp[(c + 1) * stride + 2];
                }
                const int kb = c * CHUNK + par;
#pragma unroll
                for (int g = 0; g < 8; ++g) {
                    if (c * CHUNK + 4 * g >= m_wave) break;
                    const uint32_t u0 = g < 2 ? c0.x : (g < 4 ? c0.y : (g < 6 ? c0.z : c0.w));
                    const uint32_t u1 = g < 2 ? c1.x : (g < 4 ? c1.y : (g < 6 ? c1.z : c1.w));
                    const unsigned s0 = (g & 1) ? u0 >> 16 : u0 & 0xFFFFu;
                    const unsigned s1 = (g & 1) ? u1 >> 16 : u1 & 0xFFFFu;
                    double ax, ay, az, bx, by, bz;
                    term(s0, kb + 4 * g, ax, ay, az);      // entry 4 g + par
                    term(s1, kb + 4 * g + 2, bx, by, bz);  // entry 4 g + par + 2
                    add_pair(a, ax, ay, az);
                    add_pair(a, bx, by, bz);
                }
            }
        };
        switch (wrap_axes) {
            case 0: run(std::integral_constant<int, 0>{}); break;
            case 1: run(std::integral_constant<int, 1>{}); break;
            case 2: run(std::integral_constant<int, 2>{}); break;
            case 3: run(std::integral_constant<int, 3>{}); break;
            case 4: run(std::integral_constant<int, 4>{}); break;
            case 5: run(std::integral_constant<int, 5>{}); break;
            case 6: run(std::integral_constant<int, 6>{}); break;
            default: run(std::integral_constant<int, 7>{}); break;
        }
    } else {
        // unstaged tile: entries 2 g + par through the int32 rows, four of the lane's in flight
        const int groups = (m + 1) / 2;
#pragma unroll 1
        for (int g0 = 0; g0 < groups; g0 += 4) {
            int j[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int k = 2 * (g0 + b) + par;
                j[b] = k < m ? rows[static_cast<size_t>(k) * pstride + i] : -1;
            }
            double4 rj[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) rj[b] = j[b] >= 0 ? pos[j[b]] : ri;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if (g0 + b >= groups) break;
                const int k = 2 * (g0 + b) + par;
                double tx, ty, tz;
                pair_lane_term<UNIT>(min_image(ri.x - rj[b].x, box, inv_box), min_image(ri.y - rj[b].y, box, inv_box),
                                     min_image(ri.z - rj[b].z, box, inv_box), j[b] >= 0, k < h, lj, a, tx, ty, tz);
                add_pair(a, tx, ty, tz);
            }
        }
    }
    double ke = 0.0;
    if constexpr (STEP) {
        const double m = st_m;
        // x / 1 = x exactly: a wavefront of unit masses skips the divisions
        const bool unit_mass = __ballot(m != 1.0) == 0;
        if (active && par == 0) {
            const double f[3] = {a.fx, a.fy, a.fz};
            double v[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double kick = fs.half_dt * f[c];
                v[c] = st_v[c] + (unit_mass ? kick : kick / m);  // closing kick of this step
            }
            const double half_m = 0.5 * m;
            ke = half_m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double kick = fs.half_dt * f[c];
                v[c] = v[c] + (unit_mass ? kick : kick / m);  // opening kick of the next step
                fs.vel[c * pstride + i] = v[c];
            }
            double4 r = ri;
            r.x = wrap_coord(r.x + fs.dt * v[0], box, inv_box);
            r.y = wrap_coord(r.y + fs.dt * v[1], box, inv_box);
            r.z = wrap_coord(r.z + fs.dt * v[2], box, inv_box);
            fs.pos_next[i] = r;
            const double4 q = st_ref;
            const double dx = min_image(r.x - q.x, box, inv_box);
            const double dy = min_image(r.y - q.y, box, inv_box);
            const double dz = min_image(r.z - q.z, box, inv_box);
            if (dx * dx + dy * dy + dz * dz > fs.max_disp2) *fs.stale_next = 1;
        }
    } else if (active && par == 0) {
        frc[i] = a.fx;
        frc[pstride + i] = a.fy;
        frc[2 * pstride + i] = a.fz;
    }
    const double cnt = wave_sum_dpp(static_cast<double>(a.c));
    const double s12 = wave_sum_dpp(a.s12), s6 = wave_sum_dpp(a.s6);
    const double ke_wave = STEP ? wave_sum_dpp(ke) : 0.0;
    const int wave = threadIdx.x / WAVE;
    if (threadIdx.x % WAVE == 0) {
        e_part[wave][0] = lj.c4 * (s12 - s6) - lj.shift * cnt;
        e_part[wave][1] = lj.c24 * (2.0 * s12 - s6);
        e_part[wave][2] = cnt;
        if constexpr (STEP) e_part[wave][3] = ke_wave;
    }
    if (clear_flag && blockIdx.x == 0 && threadIdx.x == 0) *clear_flag = 0;
    __syncthreads();
    if (threadIdx.x < SUMS) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < PAIR_THREADS / WAVE; ++q) v += e_part[q][threadIdx.x];
        partials[static_cast<size_t>(tile) * SUMS + threadIdx.x] = v;
    }
}

static int slot_row_len(int padding);

// The launch view of one System: every tile in ts.one.
static TileSystems single_system(const md_state* s, const md_list* l, double reach) {
    TileSystems ts{};
    TileSystem& o = ts.one;
    const size_t n = s->n;
    o.first = 0;
    o.end = static_cast<uint32_t>(n);
    o.cell_base = 0;
    o.side = l->side;
    o.use_cells = l->side >= 3 ? 1 : 0;
    o.hmax = l->hmax;
    o.interior_ok = l->side >= 7 ? 1 : 0;
    o.gate = 0;
    o.box = s->box;
    o.inv_box = s->inv_box;
    // cells that a tile and its halo may span with every coordinate inside half a box (skin: the largest drift)
    o.flat_cells = l->side > 0 ? (0.5 * s->box - l->skin) / (s->box / l->side) : 0.0;
    o.reach = reach;
    // MDGPU_TILE_ORDER=0: launches take the tiles in cell order (the band order is still written)
    static const bool banded = [] {
        const char* v = std::getenv("MDGPU_TILE_ORDER");
        return !(v && v[0] == '0');
    }();
    ts.order = banded ? l->tile_order : nullptr;
    ts.stride = n;
    ts.tiles = tile_count(n ? n : 1);
    return ts;
}

void tiled_forces_fp64_enqueue(md_state* s, md_list* l, const LJConst& lj, double* out_dev, int* clear_flag,
                               hipEvent_t ev_start, hipEvent_t ev_stop) {
    hipStream_t st = s->ctx->stream;
    const size_t n = s->n;
    const int tiles = tile_count(n ? n : 1);
    state_reserve_partials(s, static_cast<size_t>(tiles) * 3);
    // unstaged tiles gather through the int32 rows
    ensure_rows32(l);
    if (ev_start) MD_HIP(hipEventRecord(ev_start, st));
    if (n > 0) {
        TileSystems ts = single_system(s, l, 0.0);
        ts.order = nullptr;  // the force kernel takes the tiles in cell order
        auto kernel = lj.sig2 == 1.0 ? k_forces_tiled_fp64_pair<true, false, false>
                                     : k_forces_tiled_fp64_pair<false, false, false>;
        hipLaunchKernelGGL(kernel, dim3(tiles), dim3(PAIR_THREADS), 0, st, s->pos, ts, l->rows16, l->rows, l->count,
                           l->half, l->tile_ranges, l->tile_nr, slot_row_len(l->padding), lj, s->frc, s->partials,
                           clear_flag, ForceStep{}, s->halt);
        MD_HIP(hipGetLastError());
    }
    if (ev_stop) MD_HIP(hipEventRecord(ev_stop, st));
    reduce_partials_enqueue(s, n > 0 ? tiles : 0, 3, out_dev);
}

bool tiling_enabled(size_t n) {
    // the slot rows index chunks with 32-bit strides: 32 n < 2^32
    if (n >= (size_t{1} << 27)) return false;
    static const bool enabled = [] {
        const char* v = std::getenv("MDGPU_UNTILED");
        return !(v && v[0] == '1');
    }();
    return enabled;
}

int tile_count(size_t n) { return static_cast<int>((n + TILE - 1) / TILE); }

static int slot_row_len(int padding) { return (padding + CHUNK - 1) / CHUNK * CHUNK; }

void tiled_rows32_enqueue(md_list* l) {
    const size_t n = l->n;
    if (n == 0) return;
    const int tiles = tile_count(n);
    TileSystems ts{};
    ts.one.end = static_cast<uint32_t>(n);
    ts.stride = n;
    ts.tiles = tiles;
    hipLaunchKernelGGL(k_rows_from_slots<false>, dim3(tiles), dim3(TILE_THREADS), 0, l->ctx->stream, ts, l->rows16,
                       slot_row_len(l->padding), l->count, l->tile_ranges, l->tile_nr, l->rows);
    MD_HIP(hipGetLastError());
}

size_t tiled_rows16_entries(size_t n, int padding) {
    return static_cast<size_t>(tile_count(n ? n : 1)) * TILE * slot_row_len(padding);
}

void tiled_list_enqueue(md_list* l, md_state* s, bool from_alt, const int* gate, bool write32, bool carry_forces) {
    hipStream_t st = l->ctx->stream;
    const size_t n = s->n;
    if (n == 0) {
        finish_build_enqueue(l, 0, gate, s->halt);
        return;
    }
    if (l->hmax > TILE_HALO_MAX) fail(MD_ERR_STATE, "tile halo capacity exceeds the LDS bound");
    TiledCopyBack cb{};
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
    const int tiles = tile_count(n);
    const TileSystems ts = single_system(s, l, 0.0);
    // after a re-sort the tile setup gathers the state into the alt arrays, the positions only into `ref` (the
    // build-time positions, which they are): the build reads them there and copies them back into place
    PermuteArgs gather{};
    if (from_alt) {
        gather = permute_args(l, s, carry_forces);
        gather.pos2 = nullptr;
    }
    const double4* build_pos = from_alt ? l->ref : s->pos;
    const int* cells = from_alt ? l->cell_of_alt : l->cell_of;
    hipLaunchKernelGGL(k_tile_setup<false>, dim3((tiles + SETUP_TILES - 1) / SETUP_TILES), dim3(SETUP_TILES * WAVE), 0, st,
                       cells, l->cell_start, ts, l->tile_ranges, l->tile_nr, l->order, gather, l->tile_order, gate,
                       s->halt);
    // the float halo where the grid is wide enough for compact frames and the cells dense enough for the
    // candidate lists (MDGPU_BUILD_F64=1 keeps the double halo)
    const char* f64 = std::getenv("MDGPU_BUILD_F64");
    const bool fh = l->side >= 20 && static_cast<double>(n) >= 6.0 * static_cast<double>(l->ncells) &&
                    !(f64 && f64[0] == '1');
    auto kernel = fh ? k_list_build_tiled<true, false> : k_list_build_tiled<false, false>;
    const size_t lds = static_cast<size_t>(TILE_HALO_MAX) * 3 * (fh ? sizeof(float) : sizeof(double));
    hipLaunchKernelGGL(kernel, dim3(tiles), dim3(TILE_THREADS), lds, st, build_pos, cells, l->cell_start, ts,
                       l->radius2, l->padding, write32 ? 1 : 0, slot_row_len(l->padding), l->rows, l->rows16,
                       l->count, l->half, l->build_stats, l->tile_ranges, l->tile_nr, cb, gate, s->halt);
    MD_HIP(hipGetLastError());
    finish_build_enqueue(l, tiles, gate, s->halt);
    MD_HIP(hipGetLastError());
}

static void tiled_forces_launch(md_state* s, md_list* l, const LJConst& lj, double cutoff, double* out_dev,
                                int* clear_flag, const ForceStep* step, StepRecord rec, hipEvent_t ev_start,
                                hipEvent_t ev_stop) {
    hipStream_t st = s->ctx->stream;
    const size_t n = s->n;
    const int tiles = tile_count(n ? n : 1);
    const int sums = step ? 4 : 3;
    state_reserve_partials(s, static_cast<size_t>(tiles) * sums);
    TileSystems ts = single_system(s, l, cutoff + l->radius + l->skin);
    ts.order = nullptr;  // the force kernel takes the tiles in cell order
    if (ev_start) MD_HIP(hipEventRecord(ev_start, st));
    if (n > 0) {
        // sigma = 1 drops the sigma^2 factor (a product with 1.0f: the same values)
        const bool unit = lj.sig2f == 1.0f;
        auto kernel = step ? (unit ? k_forces_tiled_mixed<true, false, true> : k_forces_tiled_mixed<true, false, false>)
                           : (unit ? k_forces_tiled_mixed<false, false, true> : k_forces_tiled_mixed<false, false, false>);
        hipLaunchKernelGGL(kernel, dim3(tiles), dim3(TILE_THREADS), 0, st, s->pos, ts, l->rows16, l->rows, l->count,
                           l->tile_ranges, l->tile_nr, slot_row_len(l->padding), lj, s->frc, s->partials, clear_flag,
                           step ? *step : ForceStep{}, s->halt);
        MD_HIP(hipGetLastError());
    }
    if (ev_stop) MD_HIP(hipEventRecord(ev_stop, st));
    reduce_partials_enqueue(s, n > 0 ? tiles : 0, sums, out_dev, rec);
}

int tiled_row_len(int padding) { return slot_row_len(padding); }

void tiled_batch_build_enqueue(hipStream_t st, const BatchTiles& bt, const double4* build_pos, const int* cell_of,
                               const uint32_t* start, double radius2, const TiledCopyBack& cb, const int* halt) {
    const int tiles = bt.ts.tiles;
    hipLaunchKernelGGL(k_tile_setup<true>, dim3((tiles + SETUP_TILES - 1) / SETUP_TILES), dim3(SETUP_TILES * WAVE), 0,
                       st, cell_of, start, bt.ts, bt.tile_ranges, bt.tile_nr, nullptr, PermuteArgs{}, nullptr, nullptr, halt);
    MD_HIP(hipGetLastError());
    // the float halo: three blocks per CU, and float distance tests (with the minimum image in small boxes)
    hipLaunchKernelGGL((k_list_build_tiled<true, true>), dim3(tiles), dim3(TILE_THREADS),
                       static_cast<size_t>(TILE_HALO_MAX) * 3 * sizeof(float), st, build_pos, cell_of, start, bt.ts,
                       radius2, bt.padding, 0, slot_row_len(bt.padding), bt.rows, bt.rows16, bt.count, bt.half,
                       bt.stats, bt.tile_ranges, bt.tile_nr, cb, nullptr, halt);
    MD_HIP(hipGetLastError());
}

void tiled_batch_forces_enqueue(hipStream_t st, const BatchTiles& bt, const double4* pos, const LJConst& lj,
                                int precision, double* frc, double* partials, const int* halt, double* kick_vel,
                                const double* mass, double half_dt) {
    const int tiles = bt.ts.tiles;
    const int row_len = slot_row_len(bt.padding);
    if (precision == MD_MIXED) {
        // kick_vel: the closing kick of every particle and its kinetic energy as each tile's fourth sum
        ForceStep fs{};
        fs.vel = kick_vel;
        fs.mass = mass;
        fs.half_dt = half_dt;
        const bool unit = lj.sig2f == 1.0f;
        auto kernel = kick_vel ? (unit ? k_forces_tiled_mixed<true, true, true> : k_forces_tiled_mixed<true, true, false>)
                               : (unit ? k_forces_tiled_mixed<false, true, true> : k_forces_tiled_mixed<false, true, false>);
        hipLaunchKernelGGL(kernel, dim3(tiles), dim3(TILE_THREADS), 0, st, pos, bt.ts,
                           bt.rows16, bt.rows, bt.count, bt.tile_ranges, bt.tile_nr, row_len, lj, frc, partials,
                           nullptr, fs, halt);
    } else {
        auto kernel = lj.sig2 == 1.0 ? k_forces_tiled_fp64_pair<true, true, false>
                                     : k_forces_tiled_fp64_pair<false, true, false>;
        hipLaunchKernelGGL(kernel, dim3(tiles), dim3(PAIR_THREADS), 0, st, pos, bt.ts, bt.rows16, bt.rows, bt.count,
                           bt.half, bt.tile_ranges, bt.tile_nr, row_len, lj, frc, partials, nullptr, ForceStep{}, halt);
    }
    MD_HIP(hipGetLastError());
}

void tiled_forces_enqueue(md_state* s, md_list* l, const LJConst& lj, double cutoff, double* out_dev,
                          int* clear_flag, hipEvent_t ev_start, hipEvent_t ev_stop) {
    tiled_forces_launch(s, l, lj, cutoff, out_dev, clear_flag, nullptr, StepRecord{}, ev_start, ev_stop);
}

void tiled_forces_step_enqueue(md_state* s, md_list* l, const md_lj& p, int precision, double dt, double4* pos_next,
                               int* clear_flag, int* stale_next, double* out_dev, StepRecord rec, hipEvent_t ev_start,
                               hipEvent_t ev_stop) {
    ForceStep fs;
    fs.pos_next = pos_next;
    fs.vel = s->vel;
    fs.mass = s->mass;
    fs.ref = l->ref;
    fs.half_dt = 0.5 * dt;
    fs.dt = dt;
    fs.max_disp2 = l->max_disp2;
    fs.stale_next = stale_next;
    const LJConst lj = make_lj(p);
    if (precision == MD_MIXED) {
        tiled_forces_launch(s, l, lj, p.cutoff, out_dev, clear_flag, &fs, rec, ev_start, ev_stop);
        return;
    }
    // FP64: the tiled FP64 kernel with the step folded in (unstaged tiles gather through the int32 rows, which the
    // run's FP64 rebuilds write)
    hipStream_t st = s->ctx->stream;
    const size_t n = s->n;
    const int tiles = tile_count(n ? n : 1);
    state_reserve_partials(s, static_cast<size_t>(tiles) * 4);
    ensure_rows32(l);
    if (ev_start) MD_HIP(hipEventRecord(ev_start, st));
    if (n > 0) {
        TileSystems ts = single_system(s, l, 0.0);
        ts.order = nullptr;  // the force kernels take the tiles in cell order
        auto kernel = lj.sig2 == 1.0 ? k_forces_tiled_fp64_pair<true, false, true>
                                     : k_forces_tiled_fp64_pair<false, false, true>;
        hipLaunchKernelGGL(kernel, dim3(tiles), dim3(PAIR_THREADS), 0, st, s->pos, ts, l->rows16, l->rows, l->count,
                           l->half, l->tile_ranges, l->tile_nr, slot_row_len(l->padding), lj, s->frc, s->partials,
                           clear_flag, fs, s->halt);
        MD_HIP(hipGetLastError());
    }
    if (ev_stop) MD_HIP(hipEventRecord(ev_stop, st));
    reduce_partials_enqueue(s, n > 0 ? tiles : 0, 4, out_dev, rec);
}

}  // namespace mdgpu
