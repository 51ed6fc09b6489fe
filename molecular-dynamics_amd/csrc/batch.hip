// Batched replicas: many independent Systems advanced together, every kernel launch covering all of them.
//
// Layout. The replicas' particles are concatenated: replica r owns slots [first_r, first_r + n_r), first_r a
// multiple of BT = 256, so every 256-slot block belongs to exactly one replica (block_rep[]). Every per-particle
// array has `slots` entries; the component-major velocity and force arrays have stride `slots`. Cells are
// concatenated the same way (replica r owns cell words [cell_first_r, cell_first_r + ncells_r], the last one
// its end offset; cell_rep[] maps a word back to its replica) and the cell start offsets hold batch slots, so
// list rows hold batch slots too. One padded full list serves all replicas: padding shared, column-major with
// stride `slots`.
//
// A step is one fixed sequence of launches (and so one hipGraph): k_b_kick_drift raises a replica's stale flag
// (NeighbourList::update's half-skin test) → the rebuild kernels, each exiting at once for a replica that is not
// stale → forces → kick and kinetic energy (+ Bussi) → k_b_record, which also clears the stale flags. Each kernel
// repeats the arithmetic of the single-system kernel for the same phase (cells.hip, forces.hip, integrate.hip,
// api.hip), so an FP64 replica follows exactly the trajectory md_run gives its System.
//
// Lists and forces use the tiled kernels of tiled.hip (one launch over every replica's tiles of 128 slots, each
// tile reading its replica's box and cell grid from a table): the LDS-staged list build and the LDS-staged mixed
// and FP64 force kernels, with the single system's rows and sums. MDGPU_UNTILED=1 keeps the per-particle kernels
// k_b_list_build and k_b_forces below (the same rows; their forces take 1, 2 or 4 lanes per particle).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "pair.hpp"

namespace mdgpu {

constexpr int BT = 256;          // particle slots per block; replicas start on block boundaries
constexpr int KE_TERMS = 1024;   // the reference's deterministicSum block (Parallel.hpp:32-57)
constexpr int SCALARS = 8;       // per replica: U, W, inside, KE, -, alpha, KE', -
constexpr int REDUCE_WAVES = 4;  // replicas per block of the per-replica wavefront kernels

struct Rep {
    uint32_t first, n;            // particle slots [first, first + n)
    uint32_t cell_first, ncells;  // cell words [cell_first, cell_first + ncells]
    int side, use_cells;
    double box, inv_box, per_length;
    uint32_t ke_first, ke_blocks; // reference-order kinetic energy blocks of 1024 particles
};

struct BatchPermute {
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

struct BatchCopyBack {
    const double* vel2;
    const double* frc2;
    const double* mass2;
    const uint64_t* id2;
    double4* pos;
    double* vel;
    double* frc;
    double* mass;
    uint64_t* id;
};

// ---- integration -----------------------------------------------------------------------------------------

// k_kick_drift per replica box; the staleness test raises stale[r].
// With `scalars` (Bussi runs), the velocities are first scaled by the replica's alpha from the previous step's
// thermostat (k_scale of that step, moved here: the same products, one pass over the velocities fewer).
__global__ void __launch_bounds__(BT)
    k_b_kick_drift(const Rep* __restrict__ reps, const uint16_t* __restrict__ block_rep, double4* __restrict__ pos,
                   double* __restrict__ vel, const double* __restrict__ frc, const double* __restrict__ mass,
                   size_t slots, double half_dt, double dt, const double4* __restrict__ ref, double max_disp2,
                   int* __restrict__ stale, const double* __restrict__ scalars, const int* halt) {
    if (halted(halt)) return;
    const int r = block_rep[blockIdx.x];
    const Rep rp = reps[r];
    const size_t i = static_cast<size_t>(blockIdx.x) * BT + threadIdx.x;
    if (i >= static_cast<size_t>(rp.first) + rp.n) return;
    const double m = mass[i];
    double4 p = pos[i];
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double kick = half_dt * frc[c * slots + i];
        double vc = vel[c * slots + i];
        if (scalars) vc = vc * scalars[r * SCALARS + 5];
        v[c] = vc + kick / m;
        vel[c * slots + i] = v[c];
    }
    p.x = wrap_coord(p.x + dt * v[0], rp.box, rp.inv_box);
    p.y = wrap_coord(p.y + dt * v[1], rp.box, rp.inv_box);
    p.z = wrap_coord(p.z + dt * v[2], rp.box, rp.inv_box);
    pos[i] = p;
    const double4 q = ref[i];
    const double dx = min_image(p.x - q.x, rp.box, rp.inv_box);
    const double dy = min_image(p.y - q.y, rp.box, rp.inv_box);
    const double dz = min_image(p.z - q.z, rp.box, rp.inv_box);
    if (dx * dx + dy * dy + dz * dz > max_disp2) stale[r] = 1;
}

// k_kick; with `partials`, per-block sums of (0.5 m) v^2 (blocks never straddle replicas).
__global__ void __launch_bounds__(BT)
    k_b_kick(const Rep* __restrict__ reps, const uint16_t* __restrict__ block_rep, double* __restrict__ vel,
             const double* __restrict__ frc, const double* __restrict__ mass, size_t slots, double half_dt,
             double* __restrict__ partials, const int* halt) {
    if (halted(halt)) return;
    const Rep rp = reps[block_rep[blockIdx.x]];
    const size_t i = static_cast<size_t>(blockIdx.x) * BT + threadIdx.x;
    double ke[1] = {0.0};
    if (i < static_cast<size_t>(rp.first) + rp.n) {
        const double m = mass[i];
        double v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double kick = half_dt * frc[c * slots + i];
            v[c] = vel[c * slots + i] + kick / m;
            vel[c * slots + i] = v[c];
        }
        const double half_m = 0.5 * m;
        ke[0] = half_m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    if (partials) block_sum<1, BT>(ke, partials + blockIdx.x);
}

// Per replica, in block order: the kick's kinetic-energy partials -> scalars[3].
__global__ void k_b_reduce_ke(const Rep* __restrict__ reps, int count, const double* __restrict__ partials,
                              double* __restrict__ scalars, const int* halt) {
    if (halted(halt)) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    const Rep rp = reps[r];
    const uint32_t b0 = rp.first / BT, b1 = (rp.first + rp.n + BT - 1) / BT;
    double s = 0.0;
    for (uint32_t b = b0; b < b1; ++b) s += partials[b];
    scalars[r * SCALARS + 3] = s;
}

// The closing kick (k_b_kick) of one reference-order kinetic-energy block of 1024 particles, then the block's
// sequential sum of (0.5 m) v^2 from the kicked velocities (k_ke_blocks per replica; k_b_ke_final adds the block
// sums in order).
__global__ void __launch_bounds__(BT)
    k_b_kick_ke(const Rep* __restrict__ reps, const uint16_t* __restrict__ block_rep,
                const uint32_t* __restrict__ ke_start, double* __restrict__ vel, const double* __restrict__ frc,
                const double* __restrict__ mass, size_t slots, double half_dt, double* __restrict__ ke_part,
                const int* halt) {
    if (halted(halt)) return;
    __shared__ double terms[KE_TERMS];
    const uint32_t base = ke_start[blockIdx.x];
    const Rep rp = reps[block_rep[base / BT]];
    const uint32_t end = rp.first + rp.n;
#pragma unroll
    for (int t = threadIdx.x; t < KE_TERMS; t += BT) {
        const size_t i = static_cast<size_t>(base) + t;
        double term = 0.0;
        if (i < end) {
            const double m = mass[i];
            double v[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double kick = half_dt * frc[c * slots + i];
                v[c] = vel[c * slots + i] + kick / m;
                vel[c * slots + i] = v[c];
            }
            const double half_m = 0.5 * m;
            term = half_m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        terms[t] = term;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t cnt = (end - base) < static_cast<uint32_t>(KE_TERMS) ? (end - base) : KE_TERMS;
        double s = 0.0;
        for (uint32_t t = 0; t < cnt; ++t) s += terms[t];
        ke_part[blockIdx.x] = s;
    }
}

// alpha = 1 for every replica: the velocity scale the first opening kick of a Bussi run applies.
__global__ void k_b_unit_alpha(int count, double* __restrict__ scalars) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < count) scalars[r * SCALARS + 5] = 1.0;
}

__global__ void k_b_ke_final(const Rep* __restrict__ reps, int count, const double* __restrict__ ke_part,
                             double* __restrict__ scalars, const int* halt) {
    if (halted(halt)) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    const Rep rp = reps[r];
    double total = 0.0;
    for (uint32_t b = 0; b < rp.ke_blocks; ++b) total += ke_part[rp.ke_first + b];
    scalars[r * SCALARS + 3] = total;
}

// k_bussi per replica: eq. A7 from the replica's KE and its own draw stream (consumed only when KE > 0).
__global__ void k_b_bussi(int count, double* __restrict__ scalars, const double* __restrict__ draws, int steps,
                          int* __restrict__ draw_index, const double2* __restrict__ dof_kbar, double c, const int* halt) {
    if (halted(halt)) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    double* sc = scalars + r * SCALARS;
    const double ke = sc[3];
    if (!(ke > 0.0)) {
        sc[5] = 1.0;
        sc[6] = ke;
        return;
    }
    const double dof = dof_kbar[r].x, kbar = dof_kbar[r].y;
    const int k = draw_index[r];
    draw_index[r] = k + 1;
    const double* d = draws + (static_cast<size_t>(r) * steps + k) * 2;
    const double r1 = d[0], sum_sq = d[1];
    const double relax = (1.0 - c) * (kbar * (sum_sq + r1 * r1) / dof - ke);
    const double noise = 2.0 * r1 * sqrt(kbar / dof * ke * (1.0 - c) * c);
    const double ke_new = ke + relax + noise;
    sc[5] = sqrt(ke_new / ke);
    sc[6] = ke_new;
}

__global__ void __launch_bounds__(BT)
    k_b_scale(const Rep* __restrict__ reps, const uint16_t* __restrict__ block_rep, double* __restrict__ vel,
              size_t slots, const double* __restrict__ scalars, const int* halt) {
    if (halted(halt)) return;
    const int r = block_rep[blockIdx.x];
    const Rep rp = reps[r];
    const size_t i = static_cast<size_t>(blockIdx.x) * BT + threadIdx.x;
    if (i >= static_cast<size_t>(rp.first) + rp.n) return;
    const double s = scalars[r * SCALARS + 5];
#pragma unroll
    for (int c = 0; c < 3; ++c) vel[c * slots + i] = vel[c * slots + i] * s;
}

// Per-step record of every replica, the rebuild counters, and the stale flags cleared for the next step.
// One block; flags[2] is the step counter.
__global__ void __launch_bounds__(1024)
    k_b_record(int count, const double* __restrict__ scalars, int ke_slot, double* __restrict__ energies,
               int* __restrict__ stale, int* __restrict__ rebuilds, int* __restrict__ flags, const int* halt) {
    if (halted(halt)) return;
    const int k = flags[2];
    for (int r = threadIdx.x; r < count; r += blockDim.x) {
        const double* sc = scalars + r * SCALARS;
        double* e = energies + (static_cast<size_t>(k) * count + r) * 4;
        e[0] = sc[0];
        e[1] = sc[ke_slot];
        e[2] = sc[1];
        e[3] = sc[2];
        rebuilds[r] += stale[r];
        stale[r] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) flags[2] = k + 1;
}

// ---- cells and list (each kernel exits at once for a replica that is not stale) --------------------------

__global__ void __launch_bounds__(BT)
    k_b_cell_count(const Rep* __restrict__ reps, const uint16_t* __restrict__ block_rep,
                   const double4* __restrict__ pos, int* __restrict__ cell_of, uint32_t* __restrict__ counts,
                   const int* __restrict__ act_blocks, const int* halt) {
    if (halted(halt)) return;
    if (static_cast<int>(blockIdx.x) >= act_blocks[0]) return;
    const int blk = act_blocks[1 + blockIdx.x];
    const int r = block_rep[blk];
    const Rep rp = reps[r];
    const size_t i = static_cast<size_t>(blk) * BT + threadIdx.x;
    const bool valid = i < static_cast<size_t>(rp.first) + rp.n;
    int c = -1;
    if (valid) {
        const double4 p = pos[i];
        c = cell_of_position(p.x, p.y, p.z, rp.per_length, rp.side);
        cell_of[i] = c;
    }
    const CellRun run = cell_run(c, valid);
    if (run.head) atomicAdd(&counts[rp.cell_first + c], static_cast<uint32_t>(run.len));
}

// One block per replica: exclusive scan of its histogram into batch slots; also clears its pair total.
__global__ void __launch_bounds__(BT)
    k_b_scan(const Rep* __restrict__ reps, const uint32_t* __restrict__ counts, uint32_t* __restrict__ start,
             const int* __restrict__ stale, unsigned long long* __restrict__ pairs, const int* halt) {
    if (halted(halt)) return;
    const int r = blockIdx.x;
    if (!stale[r]) return;
    const Rep rp = reps[r];
    __shared__ uint32_t wsum[BT / WAVE];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) {
        carry = 0;
        pairs[r] = 0ull;
    }
    __syncthreads();
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    for (uint32_t b0 = 0; b0 < rp.ncells; b0 += BT) {
        const uint32_t c = b0 + threadIdx.x;
        const uint32_t v = c < rp.ncells ? counts[rp.cell_first + c] : 0u;
        uint32_t incl = v;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, WAVE);
            if (lane >= off) incl += o;
        }
        if (lane == WAVE - 1) wsum[wave] = incl;
        __syncthreads();
        uint32_t before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (c < rp.ncells) start[rp.cell_first + c] = rp.first + before + incl - v;
        __syncthreads();
        if (threadIdx.x == BT - 1) carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[rp.cell_first + rp.ncells] = rp.first + rp.n;
}

__global__ void __launch_bounds__(BT)
    k_b_scatter(const Rep* __restrict__ reps, const uint16_t* __restrict__ block_rep,
                const int* __restrict__ cell_of, const uint32_t* __restrict__ start, uint32_t* __restrict__ fill,
                uint32_t* __restrict__ order_raw, const int* __restrict__ act_blocks, const int* halt) {
    if (halted(halt)) return;
    if (static_cast<int>(blockIdx.x) >= act_blocks[0]) return;
    const int blk = act_blocks[1 + blockIdx.x];
    const int r = block_rep[blk];
    const Rep rp = reps[r];
    const size_t i = static_cast<size_t>(blk) * BT + threadIdx.x;
    const bool valid = i < static_cast<size_t>(rp.first) + rp.n;
    const int c = valid ? cell_of[i] : -1;
    const uint32_t cw = rp.cell_first + static_cast<uint32_t>(c);
    const CellRun run = cell_run(c, valid);
    uint32_t base = 0;
    if (run.head) base = atomicAdd(&fill[cw], static_cast<uint32_t>(run.len));
    base = static_cast<uint32_t>(__shfl(static_cast<int>(base), run.first, WAVE));
    if (valid) order_raw[start[cw] + base + static_cast<uint32_t>((threadIdx.x % WAVE) - run.first)] = static_cast<uint32_t>(i);
}

// k_cell_sort over the cells of the replicas being rebuilt: one wavefront per cell word of act_cells.
__global__ void __launch_bounds__(BT)
    k_b_sort(const Rep* __restrict__ reps, const uint16_t* __restrict__ cell_rep, size_t slots,
             const uint32_t* __restrict__ start, const uint32_t* __restrict__ order_raw, uint32_t* __restrict__ fill,
             uint32_t* __restrict__ counts, BatchPermute pa, const uint32_t* __restrict__ act_cells,
             const int* halt) {
    if (halted(halt)) return;
    const size_t w = (static_cast<size_t>(blockIdx.x) * BT + threadIdx.x) / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (w >= act_cells[0]) return;
    const uint32_t g = act_cells[1 + w];
    const int r = cell_rep[g];
    const Rep rp = reps[r];
    const uint32_t cell = g - rp.cell_first;
    const uint32_t b = start[g];
    const uint32_t m = start[g + 1] - b;
    if (lane == 0) {
        fill[g] = 0;
        counts[g] = 0;
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
            const double4 p = pa.pos[v];
            pa.pos2[k] = p;
            pa.ref[k] = p;
            for (int c = 0; c < 3; ++c) {
                pa.vel2[c * slots + k] = pa.vel[c * slots + v];
                pa.frc2[c * slots + k] = pa.frc[c * slots + v];
            }
            pa.mass2[k] = pa.mass[v];
            pa.id2[k] = pa.id[v];
            pa.cell_of2[k] = static_cast<int>(cell);
        }
    }
}

// k_list_build per replica: rows in the reference's order (j > i in stencil order, then j < i increasing),
// from the re-sorted arrays, which it also moves back into place.
__global__ void __launch_bounds__(BT)
    k_b_list_build(const Rep* __restrict__ reps, const uint16_t* __restrict__ block_rep,
                   const double4* __restrict__ pos, size_t slots, const int* __restrict__ cell_of,
                   const uint32_t* __restrict__ start, double radius2, int padding, int32_t* __restrict__ rows,
                   int32_t* __restrict__ count, int32_t* __restrict__ half, int2* __restrict__ bstats,
                   BatchCopyBack cb, const int* __restrict__ stale, const int* halt) {
    if (halted(halt)) return;
    const int r = block_rep[blockIdx.x];
    if (!stale[r]) return;
    const Rep rp = reps[r];
    const size_t i = static_cast<size_t>(blockIdx.x) * BT + threadIdx.x;
    const uint32_t end = rp.first + rp.n;
    int cnt = 0, hcnt = 0;
    if (i < end) {
        const double4 ri = pos[i];
        cb.pos[i] = ri;
        for (int c = 0; c < 3; ++c) {
            cb.vel[c * slots + i] = cb.vel2[c * slots + i];
            cb.frc[c * slots + i] = cb.frc2[c * slots + i];
        }
        cb.mass[i] = cb.mass2[i];
        cb.id[i] = cb.id2[i];
        const double box = rp.box, inv_box = rp.inv_box;
        auto consider = [&](uint32_t j) {
            const double4 rj = pos[j];
            const double dx = min_image(ri.x - rj.x, box, inv_box);
            const double dy = min_image(ri.y - rj.y, box, inv_box);
            const double dz = min_image(ri.z - rj.z, box, inv_box);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < radius2) {
                if (cnt < padding) rows[static_cast<size_t>(cnt) * slots + i] = static_cast<int32_t>(j);
                ++cnt;
            }
        };
        const uint32_t self = static_cast<uint32_t>(i);
        const int side = rp.side;
        const uint32_t* st = start + rp.cell_first;
        if (rp.use_cells) {
            const int c = cell_of[i];
            const int cx = c / (side * side), cy = (c / side) % side, cz = c % side;
            for (int ox = -1; ox <= 1; ++ox) {
                const int x = wrap_index(cx + ox, side);
                for (int oy = -1; oy <= 1; ++oy) {
                    const int y = wrap_index(cy + oy, side);
                    for (int oz = -1; oz <= 1; ++oz) {
                        const int nc = (x * side + y) * side + wrap_index(cz + oz, side);
                        if (nc < c) continue;
                        const uint32_t e = st[nc + 1];
                        for (uint32_t j = (nc == c ? self + 1 : st[nc]); j < e; ++j) consider(j);
                    }
                }
            }
            hcnt = cnt;
            int xs[3] = {wrap_index(cx - 1, side), cx, wrap_index(cx + 1, side)};
            int ys[3] = {wrap_index(cy - 1, side), cy, wrap_index(cy + 1, side)};
            int zs[3] = {wrap_index(cz - 1, side), cz, wrap_index(cz + 1, side)};
            for (int a = 0; a < 2; ++a)
                for (int k = 0; k < 2 - a; ++k) {
                    if (xs[k] > xs[k + 1]) { const int t = xs[k]; xs[k] = xs[k + 1]; xs[k + 1] = t; }
                    if (ys[k] > ys[k + 1]) { const int t = ys[k]; ys[k] = ys[k + 1]; ys[k + 1] = t; }
                    if (zs[k] > zs[k + 1]) { const int t = zs[k]; zs[k] = zs[k + 1]; zs[k + 1] = t; }
                }
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2)
                    for (int d = 0; d < 3; ++d) {
                        const int nc = (xs[a] * side + ys[b2]) * side + zs[d];
                        if (nc > c) continue;
                        const uint32_t e = (nc == c) ? self : st[nc + 1];
                        for (uint32_t j = st[nc]; j < e; ++j) consider(j);
                    }
        } else {
            for (uint32_t j = self + 1; j < end; ++j) consider(j);
            hcnt = cnt;
            for (uint32_t j = rp.first; j < self; ++j) consider(j);
        }
        count[i] = cnt < padding ? cnt : padding;
        half[i] = hcnt < padding ? hcnt : padding;
    }
    block_build_stat<BT>(cnt, static_cast<unsigned>(hcnt), bstats + blockIdx.x);
}

// Per-replica totals of a build from its records (one per `unit` slots: a block of the per-particle build, or a
// tile): half-pair count, and the longest row and overflow over the stale replicas into flags[1] and flags[0].
// One wavefront per replica (integer maximum and sum: any order gives the same totals).
__global__ void __launch_bounds__(REDUCE_WAVES * WAVE)
    k_b_build_totals(const Rep* __restrict__ reps, int count, int unit, const int2* __restrict__ bstats,
                     int padding, int* __restrict__ flags, unsigned long long* __restrict__ pairs,
                     const int* __restrict__ stale, const int* halt) {
    if (halted(halt)) return;
    const int r = blockIdx.x * REDUCE_WAVES + threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    if (r >= count || !stale[r]) return;
    const Rep rp = reps[r];
    const uint32_t b0 = rp.first / unit, b1 = (rp.first + rp.n + unit - 1) / unit;
    int mx = 0;
    unsigned long long hs = 0;
    for (uint32_t b = b0 + lane; b < b1; b += WAVE) {
        const int2 v = bstats[b];
        mx = v.x > mx ? v.x : mx;
        hs += static_cast<unsigned>(v.y);
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(mx, off, WAVE);
        mx = o > mx ? o : mx;
        hs += __shfl_xor(hs, off, WAVE);
    }
    if (lane == 0) {
        pairs[r] = hs;
        atomicMax(&flags[1], mx);
        if (mx > padding) atomicOr(&flags[0], 1);
    }
}

// The work lists of a rebuild, over the stale replicas in replica order: their non-empty tiles (act_tiles), their
// 256-slot blocks (act_blocks) and their cell words (act_cells, end words excluded); word 0 of each holds its
// length. The rebuild kernels take their k-th work item from these lists, so that a launch over the whole batch
// spends its blocks on the replicas being rebuilt only. One block.
constexpr int ACT_THREADS = 1024;
__global__ void __launch_bounds__(ACT_THREADS)
    k_b_activate(const Rep* __restrict__ reps, int count, const int* __restrict__ stale, int* __restrict__ act_tiles,
                 int* __restrict__ act_blocks, uint32_t* __restrict__ act_cells, const int* halt) {
    if (halted(halt)) return;
    // per stale replica of the chunk (compacted): first slot, particles, first cell word, cells, output offsets
    __shared__ uint32_t q_first[ACT_THREADS], q_n[ACT_THREADS], q_cell[ACT_THREADS], q_ncells[ACT_THREADS];
    __shared__ uint32_t q_ot[ACT_THREADS], q_ob[ACT_THREADS], q_oc[ACT_THREADS];
    __shared__ uint32_t wsum[4][ACT_THREADS / WAVE];
    __shared__ uint32_t base[4];  // [0] stale replicas in the chunk, [1..3] tiles, blocks, cells so far
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    if (tid < 4) base[tid] = 0;
    __syncthreads();
    for (int r0 = 0; r0 < count; r0 += ACT_THREADS) {
        const int r = r0 + tid;
        uint32_t v[4] = {0u, 0u, 0u, 0u};
        Rep rp{};
        if (r < count && stale[r]) {
            rp = reps[r];
            v[0] = 1u;
            v[1] = (rp.n + TILE_PARTICLES - 1) / TILE_PARTICLES;
            v[2] = (rp.n + BT - 1) / BT;
            v[3] = rp.ncells;
        }
        uint32_t incl[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            incl[k] = v[k];
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const uint32_t o = __shfl_up(incl[k], d, WAVE);
                if (lane >= d) incl[k] += o;
            }
            if (lane == WAVE - 1) wsum[k][wave] = incl[k];
        }
        __syncthreads();
        uint32_t excl[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t before = k == 0 ? 0u : base[k];
            for (int w = 0; w < wave; ++w) before += wsum[k][w];
            excl[k] = before + incl[k] - v[k];
        }
        if (v[0]) {
            const uint32_t q = excl[0];
            q_first[q] = rp.first;
            q_n[q] = rp.n;
            q_cell[q] = rp.cell_first;
            q_ncells[q] = rp.ncells;
            q_ot[q] = excl[1];
            q_ob[q] = excl[2];
            q_oc[q] = excl[3];
        }
        __syncthreads();
        if (tid == ACT_THREADS - 1) {
#pragma unroll
            for (int k = 0; k < 4; ++k) base[k] = excl[k] + v[k];
        }
        __syncthreads();
        const uint32_t stale_here = base[0];
        for (uint32_t q = 0; q < stale_here; ++q) {  // every thread walks the same list: uniform bounds
            const uint32_t first = q_first[q], n = q_n[q];
            const uint32_t nt = (n + TILE_PARTICLES - 1) / TILE_PARTICLES, nb = (n + BT - 1) / BT;
            for (uint32_t k = tid; k < nt; k += ACT_THREADS)
                act_tiles[1 + q_ot[q] + k] = static_cast<int>(first / TILE_PARTICLES + k);
            for (uint32_t k = tid; k < nb; k += ACT_THREADS) act_blocks[1 + q_ob[q] + k] = static_cast<int>(first / BT + k);
            for (uint32_t k = tid; k < q_ncells[q]; k += ACT_THREADS) act_cells[1 + q_oc[q] + k] = q_cell[q] + k;
        }
        __syncthreads();  // the q_* arrays and base[0] are rewritten by the next chunk
    }
    if (tid == 0) {
        act_tiles[0] = static_cast<int>(base[1]);
        act_blocks[0] = static_cast<int>(base[2]);
        act_cells[0] = base[3];
    }
}

// ---- forces --------------------------------------------------------------------------------------------

// LANES lanes per particle take row entries k = lane, lane + LANES, ...; LANES = 1 is the single-system kernel.
template <bool MIXED, int LANES>
__global__ void __launch_bounds__(BT)
    k_b_forces(const Rep* __restrict__ reps, const uint16_t* __restrict__ block_rep, const double4* __restrict__ pos,
               size_t slots, const int32_t* __restrict__ rows, const int32_t* __restrict__ count,
               const int32_t* __restrict__ half, LJConst lj, double* __restrict__ frc,
               double* __restrict__ partials, const int* halt) {
    if (halted(halt)) return;
    constexpr int PB = BT / LANES;
    const size_t p0 = static_cast<size_t>(blockIdx.x) * PB;
    const Rep rp = reps[block_rep[p0 / BT]];
    const int lane = threadIdx.x % LANES;
    const size_t i = p0 + threadIdx.x / LANES;
    const bool active = i < static_cast<size_t>(rp.first) + rp.n;
    PairSums a;
    if (active) {
        const double4 ri = pos[i];
        const int m = count[i], h = half[i];
        const int32_t* row = rows + i;
        for (int k = lane; k < m; k += LANES) {
            const double4 rj = pos[row[static_cast<size_t>(k) * slots]];
            if constexpr (MIXED)
                mixed_pair_gather(ri, rj, rp.box, rp.inv_box, lj, k < h, a);
            else
                fp64_pair(ri, rj, rp.box, rp.inv_box, lj, k < h, a);
        }
    }
#pragma unroll
    for (int off = 1; off < LANES; off <<= 1) {
        a.fx += __shfl_xor(a.fx, off, WAVE);
        a.fy += __shfl_xor(a.fy, off, WAVE);
        a.fz += __shfl_xor(a.fz, off, WAVE);
    }
    if (active && lane == 0) {
        frc[i] = a.fx;
        frc[slots + i] = a.fy;
        frc[2 * slots + i] = a.fz;
    }
    double acc[3] = {a.u, a.w, a.c};
    block_sum<3, BT>(acc, partials + static_cast<size_t>(blockIdx.x) * 3);
}

// Per replica, one wavefront: potential, virial, inside pairs -> scalars[0..2] (and with SUMS = 4 the kicked
// kinetic energy -> scalars[3]), from the force kernel's partial rows, each covering `unit` particle slots (a tile,
// or a block of the per-particle kernel). Lane l adds rows l, l + 64, ... of the replica, then a butterfly
// combines the lanes: a fixed order.
template <int SUMS>
__global__ void __launch_bounds__(REDUCE_WAVES * WAVE)
    k_b_reduce_forces(const Rep* __restrict__ reps, int count, int unit, const double* __restrict__ partials,
                      double* __restrict__ scalars, const int* halt) {
    if (halted(halt)) return;
    const int r = blockIdx.x * REDUCE_WAVES + threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    if (r >= count) return;
    const Rep rp = reps[r];
    const uint32_t b0 = rp.first / unit, b1 = (rp.first + rp.n + unit - 1) / unit;
    double s[SUMS] = {};
    for (uint32_t b = b0 + lane; b < b1; b += WAVE)
        for (int k = 0; k < SUMS; ++k) s[k] += partials[static_cast<size_t>(b) * SUMS + k];
    for (int k = 0; k < SUMS; ++k) s[k] = wave_sum(s[k]);
    if (lane == 0)
        for (int k = 0; k < SUMS; ++k) scalars[r * SCALARS + k] = s[k];
}

// After a host-driven build: count it and clear the stale flags.
__global__ void k_b_finish_build(int count, int* __restrict__ stale, int* __restrict__ rebuilds) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    rebuilds[r] += stale[r];
    stale[r] = 0;
}

__global__ void k_b_fill(int* __restrict__ p, int count, int value) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) p[k] = value;
}

// Slot rows moved to a longer row length: tile t's chunks (CHUNK_STRIDE entries each) from t * TILE * old_len to
// t * TILE * new_len; entries of the added chunks are left as they are (no row reaches them yet).
constexpr int ROWS16_CHUNK = TILE_PARTICLES * 32;
__global__ void k_b_regrow_rows16(const uint16_t* __restrict__ from, uint16_t* __restrict__ to, int tiles,
                                  int old_len, int new_len) {
    const size_t per_tile = static_cast<size_t>(TILE_PARTICLES) * old_len;
    const size_t total = per_tile * tiles;
    for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t t = e / per_tile, k = e % per_tile;
        to[t * TILE_PARTICLES * new_len + k] = from[e];
    }
}
static_assert(ROWS16_CHUNK == 4096, "a tile's slot rows are chunks of 128 particles x 32 entries");

int grid_side_of(double box, double min_cell) {
    const int s = static_cast<int>(box / min_cell);
    return s > 1 ? s : 1;
}

int padding_for(size_t n, double box, double radius) {
    const double density = static_cast<double>(n) / (box * box * box);
    const double expected = 4.0 / 3.0 * M_PI * radius * radius * radius * density;
    long p = static_cast<long>(std::ceil(expected * 1.35 + 24.0));
    p = (p + 7) / 8 * 8;
    const long cap = n > 1 ? static_cast<long>(n - 1) : 1;
    return static_cast<int>(std::min<long>(p, std::max<long>(cap, 1)));
}

}  // namespace mdgpu

struct md_batch {
    md_ctx* ctx = nullptr;
    int count = 0;
    size_t slots = 0, particles = 0, cells_total = 0;
    int blocks = 0, ke_total = 0;
    double cutoff = 0.0, skin = 0.0, radius2 = 0.0, max_disp2 = 0.0;
    int padding = 0;
    int lanes = 1;
    bool forces_valid = false;
    std::vector<mdgpu::Rep> reps;
    mdgpu::Rep* d_reps = nullptr;
    uint16_t* block_rep = nullptr;
    uint16_t* cell_rep = nullptr;
    uint32_t* ke_start = nullptr;
    double4 *pos = nullptr, *pos_alt = nullptr, *ref = nullptr;
    double *vel = nullptr, *vel_alt = nullptr, *frc = nullptr, *frc_alt = nullptr;
    double *mass = nullptr, *mass_alt = nullptr;
    uint64_t *id = nullptr, *id_alt = nullptr;
    int *cell_of = nullptr, *cell_of_alt = nullptr;
    uint32_t *counts = nullptr, *start = nullptr, *fill = nullptr, *order_raw = nullptr;
    int32_t *rows = nullptr, *rcount = nullptr, *half = nullptr;
    int* stale = nullptr;
    int* rebuilds = nullptr;
    unsigned long long* pairs = nullptr;
    int2* bstats = nullptr;     // per build block: longest row, half-row sum
    int* flags = nullptr;       // [0] overflow, [1] longest row, [2] step counter, [3] spare
    double* partials = nullptr; // force partials (up to blocks * 4 lanes * 3), then the kick's (blocks)
    // tiled lists (the default): a table of every replica's system, the replica of each tile, per tile its halo
    // ranges and build record, and the slot rows
    bool tiled = false;
    int tiles = 0;
    double tsys_cutoff = 0.0;  // LJ cut-off the table's `reach` was computed for
    std::vector<mdgpu::TileSystem> tsys;
    mdgpu::TileSystem* d_tsys = nullptr;
    uint16_t* tile_rep = nullptr;
    int4* tile_ranges = nullptr;
    int* tile_nr = nullptr;
    uint16_t* rows16 = nullptr;
    int2* tstats = nullptr;
    // a rebuild's work lists (k_b_activate): the stale replicas' tiles, 256-slot blocks and cell words
    int* act_tiles = nullptr;
    int* act_blocks = nullptr;
    uint32_t* act_cells = nullptr;
    double* ke_part = nullptr;
    double* scalars = nullptr;
    // During md_batch_run: the overflow word. The run's kernels exit at once while it is set, so a replica's rebuild
    // that overflows its rows freezes every replica after that step's opening kick and drift.
    const int* halt = nullptr;
};

namespace mdgpu {
namespace {

void batch_free(md_batch* b) {
    dev_free(b->d_reps);
    dev_free(b->block_rep);
    dev_free(b->cell_rep);
    dev_free(b->ke_start);
    dev_free(b->pos);
    dev_free(b->pos_alt);
    dev_free(b->ref);
    dev_free(b->vel);
    dev_free(b->vel_alt);
    dev_free(b->frc);
    dev_free(b->frc_alt);
    dev_free(b->mass);
    dev_free(b->mass_alt);
    dev_free(b->id);
    dev_free(b->id_alt);
    dev_free(b->cell_of);
    dev_free(b->cell_of_alt);
    dev_free(b->counts);
    dev_free(b->start);
    dev_free(b->fill);
    dev_free(b->order_raw);
    dev_free(b->rows);
    dev_free(b->rcount);
    dev_free(b->half);
    dev_free(b->stale);
    dev_free(b->rebuilds);
    dev_free(b->pairs);
    dev_free(b->bstats);
    dev_free(b->flags);
    dev_free(b->partials);
    dev_free(b->ke_part);
    dev_free(b->scalars);
    dev_free(b->d_tsys);
    dev_free(b->tile_rep);
    dev_free(b->tile_ranges);
    dev_free(b->tile_nr);
    dev_free(b->rows16);
    dev_free(b->tstats);
    dev_free(b->act_tiles);
    dev_free(b->act_blocks);
    dev_free(b->act_cells);
}

size_t rows16_entries(const md_batch* b) {
    return static_cast<size_t>(b->tiles) * TILE_PARTICLES * tiled_row_len(b->padding);
}

// The replicas' table for the tiled kernels; `cutoff` is the LJ cut-off of the forces that follow.
void upload_tile_systems(md_batch* b, double cutoff) {
    const double radius = std::sqrt(b->radius2);
    b->tsys.resize(b->count);
    for (int r = 0; r < b->count; ++r) {
        const Rep& rp = b->reps[r];
        TileSystem& t = b->tsys[r];
        t.first = rp.first;
        t.end = rp.first + rp.n;
        t.cell_base = rp.cell_first;
        t.side = rp.side;
        t.use_cells = rp.use_cells;
        t.hmax = tile_halo_capacity(rp.n, rp.ncells);
        t.interior_ok = rp.side >= 7 ? 1 : 0;
        t.gate = r;
        t.box = rp.box;
        t.inv_box = rp.inv_box;
        t.flat_cells = (0.5 * rp.box - b->skin) / (rp.box / rp.side);
        t.reach = cutoff + radius + b->skin;
    }
    b->tsys_cutoff = cutoff;
    MD_HIP(hipMemcpyAsync(b->d_tsys, b->tsys.data(), sizeof(TileSystem) * b->count, hipMemcpyHostToDevice,
                          b->ctx->stream));
    MD_HIP(hipStreamSynchronize(b->ctx->stream));  // the host table may change before the next upload
}

BatchTiles batch_tiles(md_batch* b) {
    BatchTiles bt{};
    bt.ts.one = b->tsys[0];
    bt.ts.per = b->d_tsys;
    bt.ts.rep_of_tile = b->tile_rep;
    bt.ts.gates = b->stale;
    bt.ts.active = b->act_tiles;
    bt.ts.stride = b->slots;
    bt.ts.tiles = b->tiles;
    bt.tile_ranges = b->tile_ranges;
    bt.tile_nr = b->tile_nr;
    bt.rows16 = b->rows16;
    bt.stats = b->tstats;
    bt.rows = b->rows;
    bt.count = b->rcount;
    bt.half = b->half;
    bt.padding = b->padding;
    return bt;
}

// Allocates every device buffer for the layout in b (sizes, reps) and uploads the tables.
void batch_alloc(md_batch* b) {
    hipStream_t st = b->ctx->stream;
    const size_t S = b->slots;
    b->d_reps = dev_alloc<Rep>(b->count);
    b->block_rep = dev_alloc<uint16_t>(b->blocks);
    b->cell_rep = dev_alloc<uint16_t>(b->cells_total);
    b->ke_start = dev_alloc<uint32_t>(b->ke_total);
    b->pos = dev_alloc<double4>(S);
    b->pos_alt = dev_alloc<double4>(S);
    b->ref = dev_alloc<double4>(S);
    b->vel = dev_alloc<double>(3 * S);
    b->vel_alt = dev_alloc<double>(3 * S);
    b->frc = dev_alloc<double>(3 * S);
    b->frc_alt = dev_alloc<double>(3 * S);
    b->mass = dev_alloc<double>(S);
    b->mass_alt = dev_alloc<double>(S);
    b->id = dev_alloc<uint64_t>(S);
    b->id_alt = dev_alloc<uint64_t>(S);
    b->cell_of = dev_alloc<int>(S);
    b->cell_of_alt = dev_alloc<int>(S);
    b->counts = dev_alloc<uint32_t>(b->cells_total);
    b->start = dev_alloc<uint32_t>(b->cells_total);
    b->fill = dev_alloc<uint32_t>(b->cells_total);
    b->order_raw = dev_alloc<uint32_t>(S);
    b->rows = dev_alloc<int32_t>(static_cast<size_t>(b->padding) * S);
    b->rcount = dev_alloc<int32_t>(S);
    b->half = dev_alloc<int32_t>(S);
    b->stale = dev_alloc<int>(b->count);
    b->rebuilds = dev_alloc<int>(b->count);
    b->pairs = dev_alloc<unsigned long long>(b->count);
    b->bstats = dev_alloc<int2>(b->blocks);
    b->flags = dev_alloc<int>(4);
    b->partials = dev_alloc<double>(static_cast<size_t>(b->blocks) * (4 * 3 + 1));
    b->ke_part = dev_alloc<double>(b->ke_total);
    b->scalars = dev_alloc<double>(static_cast<size_t>(b->count) * SCALARS);
    b->tiled = tiling_enabled(S);
    b->tiles = static_cast<int>(S / TILE_PARTICLES);
    b->act_tiles = dev_alloc<int>(1 + static_cast<size_t>(b->tiles));
    b->act_blocks = dev_alloc<int>(1 + static_cast<size_t>(b->blocks));
    b->act_cells = dev_alloc<uint32_t>(1 + b->cells_total);
    if (b->tiled) {
        b->d_tsys = dev_alloc<TileSystem>(b->count);
        b->tile_rep = dev_alloc<uint16_t>(b->tiles);
        b->tile_ranges = dev_alloc<int4>(static_cast<size_t>(b->tiles) * TILE_RANGES);
        b->tile_nr = dev_alloc<int>(b->tiles);
        b->rows16 = dev_alloc<uint16_t>(rows16_entries(b));
        b->tstats = dev_alloc<int2>(b->tiles);
        // empty tiles (past a replica's end) stay unstaged with no particles
        k_b_fill<<<blocks_for(b->tiles, 256), 256, 0, st>>>(b->tile_nr, b->tiles, -1);
        MD_HIP(hipGetLastError());
    }
    MD_HIP(hipMemsetAsync(b->counts, 0, sizeof(uint32_t) * b->cells_total, st));
    MD_HIP(hipMemsetAsync(b->fill, 0, sizeof(uint32_t) * b->cells_total, st));
    MD_HIP(hipMemsetAsync(b->rebuilds, 0, sizeof(int) * b->count, st));
    MD_HIP(hipMemsetAsync(b->flags, 0, sizeof(int) * 4, st));
    MD_HIP(hipMemsetAsync(b->scalars, 0, sizeof(double) * b->count * SCALARS, st));
    // padding slots: zero mass/velocity so that no kernel reads garbage there
    MD_HIP(hipMemsetAsync(b->pos, 0, sizeof(double4) * S, st));
    MD_HIP(hipMemsetAsync(b->vel, 0, sizeof(double) * 3 * S, st));
    MD_HIP(hipMemsetAsync(b->frc, 0, sizeof(double) * 3 * S, st));
    MD_HIP(hipMemsetAsync(b->mass, 0, sizeof(double) * S, st));
    MD_HIP(hipMemsetAsync(b->id, 0, sizeof(uint64_t) * S, st));

    std::vector<uint16_t> brep(b->blocks), crep(b->cells_total), trep(b->tiles);
    std::vector<uint32_t> kstart(b->ke_total);
    for (int r = 0; r < b->count; ++r) {
        const Rep& rp = b->reps[r];
        const uint32_t nb = (rp.n + BT - 1) / BT;
        for (uint32_t k = 0; k < (nb ? nb : 1); ++k) brep[rp.first / BT + k] = static_cast<uint16_t>(r);
        for (uint32_t k = 0; k < (nb ? nb : 1) * (BT / TILE_PARTICLES); ++k)
            trep[rp.first / TILE_PARTICLES + k] = static_cast<uint16_t>(r);
        for (uint32_t c = 0; c <= rp.ncells; ++c) crep[rp.cell_first + c] = static_cast<uint16_t>(r);
        for (uint32_t k = 0; k < rp.ke_blocks; ++k) kstart[rp.ke_first + k] = rp.first + k * KE_TERMS;
    }
    MD_HIP(hipMemcpyAsync(b->d_reps, b->reps.data(), sizeof(Rep) * b->count, hipMemcpyHostToDevice, st));
    MD_HIP(hipMemcpyAsync(b->block_rep, brep.data(), sizeof(uint16_t) * brep.size(), hipMemcpyHostToDevice, st));
    MD_HIP(hipMemcpyAsync(b->cell_rep, crep.data(), sizeof(uint16_t) * crep.size(), hipMemcpyHostToDevice, st));
    if (b->ke_total)
        MD_HIP(hipMemcpyAsync(b->ke_start, kstart.data(), sizeof(uint32_t) * kstart.size(), hipMemcpyHostToDevice,
                              st));
    if (b->tiled)
        MD_HIP(hipMemcpyAsync(b->tile_rep, trep.data(), sizeof(uint16_t) * trep.size(), hipMemcpyHostToDevice, st));
    MD_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
    if (b->tiled) upload_tile_systems(b, b->cutoff);
}

// Longer rows for every replica, keeping the current ones (column-major int32 rows: the same entries at the same
// places; slot rows: each tile's chunks moved to the new tile stride).
void grow_rows(md_batch* b, int grown, bool keep) {
    hipStream_t st = b->ctx->stream;
    int32_t* rows = dev_alloc<int32_t>(static_cast<size_t>(grown) * b->slots);
    uint16_t* rows16 = nullptr;
    hipError_t ce = hipSuccess;
    if (b->tiled) {
        const int old_len = tiled_row_len(b->padding), new_len = tiled_row_len(grown);
        const size_t entries = static_cast<size_t>(b->tiles) * TILE_PARTICLES * new_len;
        ce = hipMalloc(reinterpret_cast<void**>(&rows16), entries * sizeof(uint16_t));
        if (ce == hipSuccess && keep && new_len != old_len) {
            k_b_regrow_rows16<<<2048, 256, 0, st>>>(b->rows16, rows16, b->tiles, old_len, new_len);
            ce = hipGetLastError();
        } else if (ce == hipSuccess && keep) {
            ce = hipMemcpyAsync(rows16, b->rows16, entries * sizeof(uint16_t), hipMemcpyDeviceToDevice, st);
        }
    }
    if (ce == hipSuccess && keep)
        ce = hipMemcpyAsync(rows, b->rows, sizeof(int32_t) * b->padding * b->slots, hipMemcpyDeviceToDevice, st);
    if (ce == hipSuccess) ce = hipStreamSynchronize(st);
    if (ce != hipSuccess) {
        dev_free(rows);
        if (rows16) (void)hipFree(rows16);
        hip_check(ce, "growing the neighbour rows");
    }
    dev_free(b->rows);
    b->rows = rows;
    if (b->tiled) {
        dev_free(b->rows16);
        b->rows16 = rows16;
    }
    b->padding = grown;
}

void enqueue_rebuild(md_batch* b) {
    hipStream_t st = b->ctx->stream;
    const size_t S = b->slots;
    k_b_activate<<<1, ACT_THREADS, 0, st>>>(b->d_reps, b->count, b->stale, b->act_tiles, b->act_blocks, b->act_cells,
                                            b->halt);
    k_b_cell_count<<<b->blocks, BT, 0, st>>>(b->d_reps, b->block_rep, b->pos, b->cell_of, b->counts, b->act_blocks,
                                             b->halt);
    k_b_scan<<<b->count, BT, 0, st>>>(b->d_reps, b->counts, b->start, b->stale, b->pairs, b->halt);
    k_b_scatter<<<b->blocks, BT, 0, st>>>(b->d_reps, b->block_rep, b->cell_of, b->start, b->fill, b->order_raw,
                                          b->act_blocks, b->halt);
    BatchPermute pa{b->pos, b->vel, b->frc, b->mass, b->id, b->pos_alt, b->vel_alt,
                    b->frc_alt, b->mass_alt, b->id_alt, b->ref, b->cell_of_alt};
    k_b_sort<<<blocks_for(b->cells_total * WAVE, BT), BT, 0, st>>>(b->d_reps, b->cell_rep, S, b->start,
                                                                   b->order_raw, b->fill, b->counts, pa,
                                                                   b->act_cells, b->halt);
    if (b->tiled) {
        // the re-sorted arrays go back into place from the build, which reads the positions from pos_alt
        TiledCopyBack cb{};
        cb.enabled = 1;
        cb.vel2 = b->vel_alt;
        cb.frc2 = b->frc_alt;
        cb.mass2 = b->mass_alt;
        cb.id2 = b->id_alt;
        cb.cell_of2 = b->cell_of_alt;
        cb.pos = b->pos;
        cb.vel = b->vel;
        cb.frc = b->frc;
        cb.mass = b->mass;
        cb.id = b->id;
        cb.cell_of = b->cell_of;
        tiled_batch_build_enqueue(st, batch_tiles(b), b->pos_alt, b->cell_of_alt, b->start, b->radius2, cb, b->halt);
        k_b_build_totals<<<blocks_for(b->count, REDUCE_WAVES), REDUCE_WAVES * WAVE, 0, st>>>(
            b->d_reps, b->count, TILE_PARTICLES, b->tstats, b->padding, b->flags, b->pairs, b->stale, b->halt);
    } else {
        BatchCopyBack cb{b->vel_alt, b->frc_alt, b->mass_alt, b->id_alt, b->pos, b->vel, b->frc, b->mass, b->id};
        k_b_list_build<<<b->blocks, BT, 0, st>>>(b->d_reps, b->block_rep, b->pos_alt, S, b->cell_of_alt, b->start,
                                                 b->radius2, b->padding, b->rows, b->rcount, b->half, b->bstats, cb,
                                                 b->stale, b->halt);
        k_b_build_totals<<<blocks_for(b->count, REDUCE_WAVES), REDUCE_WAVES * WAVE, 0, st>>>(
            b->d_reps, b->count, BT, b->bstats, b->padding, b->flags, b->pairs, b->stale, b->halt);
    }
    MD_HIP(hipGetLastError());
}

// kick_half_dt > 0 (tiled mixed batches only, see kick_folds): the force kernel also applies the closing kick
// and the kicked kinetic energy goes to scalars[3], as k_b_kick_ke and k_b_ke_final would leave it
void enqueue_forces(md_batch* b, const LJConst& lj, int precision, int lanes, hipEvent_t e0, hipEvent_t e1,
                    double kick_half_dt) {
    hipStream_t st = b->ctx->stream;
    const size_t S = b->slots;
    if (b->tiled) {
        const bool kick = kick_half_dt > 0.0;
        if (e0) MD_HIP(hipEventRecord(e0, st));
        tiled_batch_forces_enqueue(st, batch_tiles(b), b->pos, lj, precision, b->frc, b->partials, b->halt,
                                   kick ? b->vel : nullptr, b->mass, kick_half_dt);
        if (e1) MD_HIP(hipEventRecord(e1, st));
        if (kick)
            k_b_reduce_forces<4><<<blocks_for(b->count, REDUCE_WAVES), REDUCE_WAVES * WAVE, 0, st>>>(
                b->d_reps, b->count, TILE_PARTICLES, b->partials, b->scalars, b->halt);
        else
            k_b_reduce_forces<3><<<blocks_for(b->count, REDUCE_WAVES), REDUCE_WAVES * WAVE, 0, st>>>(
                b->d_reps, b->count, TILE_PARTICLES, b->partials, b->scalars, b->halt);
        MD_HIP(hipGetLastError());
        return;
    }
    if (precision != MD_MIXED) lanes = 1;
    const unsigned grid = static_cast<unsigned>(b->blocks) * lanes;
    if (e0) MD_HIP(hipEventRecord(e0, st));
#define MD_B_FORCES(M, L)                                                                                  \
    k_b_forces<M, L><<<grid, BT, 0, st>>>(b->d_reps, b->block_rep, b->pos, S, b->rows, b->rcount, b->half, lj, \
                                          b->frc, b->partials, b->halt)
    if (precision == MD_MIXED) {
        if (lanes == 4)
            MD_B_FORCES(true, 4);
        else if (lanes == 2)
            MD_B_FORCES(true, 2);
        else
            MD_B_FORCES(true, 1);
    } else {
        MD_B_FORCES(false, 1);
    }
#undef MD_B_FORCES
    MD_HIP(hipGetLastError());
    if (e1) MD_HIP(hipEventRecord(e1, st));
    k_b_reduce_forces<3><<<blocks_for(b->count, REDUCE_WAVES), REDUCE_WAVES * WAVE, 0, st>>>(
        b->d_reps, b->count, BT / lanes, b->partials, b->scalars, b->halt);
    MD_HIP(hipGetLastError());
}

// The closing kick and its kinetic energy ride on the force kernel for tiled mixed batches under the Bussi
// thermostat (the kinetic energy is then summed per tile and per replica instead of in the reference's 1024-particle
// blocks, a rounding-level difference that mixed precision allows); FP64 keeps k_b_kick_ke's order, so its replicas
// stay bit-equal to md_run.
bool kick_folds(const md_batch* b, const md_batch_config* cfg) {
    return b->tiled && cfg->precision == MD_MIXED && cfg->thermostat;
}

void check_boxes(const md_batch* b, double cutoff) {
    for (const Rep& rp : b->reps)
        if (2.0 * cutoff > rp.box) fail(MD_ERR_RUNTIME, "Cutoff must be less than half the box length");
}

// Host-driven build of the stale replicas with overflow regrowth (list_rebuild for a batch).
void batch_build(md_batch* b) {
    hipStream_t st = b->ctx->stream;
    for (int attempt = 0;; ++attempt) {
        MD_HIP(hipMemsetAsync(b->flags, 0, 2 * sizeof(int), st));
        enqueue_rebuild(b);
        int hf[2];
        MD_HIP(hipMemcpyAsync(b->ctx->host_flags, b->flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
        std::memcpy(hf, b->ctx->host_flags, sizeof(hf));
        if (!hf[0]) break;
        if (attempt > 8) fail(MD_ERR_OVERFLOW, "neighbour-list rows keep overflowing");
        // the replicas are sorted now (a re-sort is the identity); rebuild all with longer rows
        grow_rows(b, (hf[1] + hf[1] / 4 + 16 + 7) / 8 * 8, false);
        k_b_fill<<<blocks_for(b->count, 256), 256, 0, st>>>(b->stale, b->count, 1);
    }
    k_b_finish_build<<<blocks_for(b->count, 256), 256, 0, st>>>(b->count, b->stale, b->rebuilds);
    MD_HIP(hipGetLastError());
}

// A new list radius (cutoff + skin) for every replica: new cell grids, cell words and rows (padding rows per
// particle; <= 0: the density estimate for the new radius), then every replica re-sorted and rebuilt from its
// current state. The particle slots stay where they are, so positions, velocities, masses and ids carry over; only
// the order within each replica changes, as at any rebuild, and the next run starts with a force evaluation.
void relayout(md_batch* b, double skin, int padding) {
    hipStream_t st = b->ctx->stream;
    if (!(skin >= 0.0)) fail(MD_ERR_ARG, "skin must be >= 0");
    const double radius = b->cutoff + skin;
    for (const Rep& rp : b->reps)
        if (2.0 * radius > rp.box) fail(MD_ERR_RUNTIME, "Cutoff + skin must be less than half the box length");
    MD_HIP(hipStreamSynchronize(st));
    size_t cell = 0;
    int pad = 8;
    for (Rep& rp : b->reps) {
        rp.side = grid_side_of(rp.box, radius);
        rp.use_cells = rp.side >= 3 ? 1 : 0;
        rp.ncells = static_cast<uint32_t>(rp.side) * rp.side * rp.side;
        rp.cell_first = static_cast<uint32_t>(cell);
        rp.per_length = rp.side / rp.box;
        cell += rp.ncells + 1;
        pad = std::max(pad, padding_for(rp.n, rp.box, radius));
        if (cell > 0xffffffffu) fail(MD_ERR_ARG, "batch too large for 32-bit cell words");
    }
    b->skin = skin;
    b->radius2 = radius * radius;
    b->max_disp2 = 0.25 * skin * skin;
    b->cells_total = cell;
    b->padding = padding > 0 ? (padding + 7) / 8 * 8 : pad;
    dev_free(b->counts);
    dev_free(b->start);
    dev_free(b->fill);
    dev_free(b->cell_rep);
    dev_free(b->act_cells);
    dev_free(b->rows);
    dev_free(b->rows16);
    b->counts = dev_alloc<uint32_t>(b->cells_total);
    b->start = dev_alloc<uint32_t>(b->cells_total);
    b->fill = dev_alloc<uint32_t>(b->cells_total);
    b->cell_rep = dev_alloc<uint16_t>(b->cells_total);
    b->act_cells = dev_alloc<uint32_t>(1 + b->cells_total);
    b->rows = dev_alloc<int32_t>(static_cast<size_t>(b->padding) * b->slots);
    if (b->tiled) b->rows16 = dev_alloc<uint16_t>(rows16_entries(b));
    MD_HIP(hipMemsetAsync(b->counts, 0, sizeof(uint32_t) * b->cells_total, st));
    MD_HIP(hipMemsetAsync(b->fill, 0, sizeof(uint32_t) * b->cells_total, st));
    std::vector<uint16_t> crep(b->cells_total);
    for (int r = 0; r < b->count; ++r) {
        const Rep& rp = b->reps[r];
        for (uint32_t c = 0; c <= rp.ncells; ++c) crep[rp.cell_first + c] = static_cast<uint16_t>(r);
    }
    MD_HIP(hipMemcpyAsync(b->d_reps, b->reps.data(), sizeof(Rep) * b->count, hipMemcpyHostToDevice, st));
    MD_HIP(hipMemcpyAsync(b->cell_rep, crep.data(), sizeof(uint16_t) * crep.size(), hipMemcpyHostToDevice, st));
    MD_HIP(hipStreamSynchronize(st));  // crep goes out of scope
    if (b->tiled) upload_tile_systems(b, b->tsys_cutoff > 0.0 ? b->tsys_cutoff : b->cutoff);
    k_b_fill<<<blocks_for(b->count, 256), 256, 0, st>>>(b->stale, b->count, 1);
    MD_HIP(hipGetLastError());
    batch_build(b);
    MD_HIP(hipStreamSynchronize(st));
    // the next run evaluates the forces over the new rows first (their summation order follows the new sort)
    b->forces_valid = false;
}

// A scratch batch with b's layout and contents (the autotuner's trials run on it, so b does not move).
void scratch_copy(const md_batch* b, md_batch& tmp) {
    hipStream_t st = b->ctx->stream;
    tmp.ctx = b->ctx;
    tmp.count = b->count;
    tmp.slots = b->slots;
    tmp.particles = b->particles;
    tmp.cells_total = b->cells_total;
    tmp.blocks = b->blocks;
    tmp.ke_total = b->ke_total;
    tmp.cutoff = b->cutoff;
    tmp.skin = b->skin;
    tmp.radius2 = b->radius2;
    tmp.max_disp2 = b->max_disp2;
    tmp.padding = b->padding;
    tmp.lanes = b->lanes;
    tmp.forces_valid = b->forces_valid;
    tmp.reps = b->reps;
    batch_alloc(&tmp);
    const size_t S = b->slots;
    auto copy = [&](void* dst, const void* src, size_t bytes) {
        MD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
    };
    copy(tmp.pos, b->pos, sizeof(double4) * S);
    copy(tmp.ref, b->ref, sizeof(double4) * S);
    copy(tmp.vel, b->vel, sizeof(double) * 3 * S);
    copy(tmp.frc, b->frc, sizeof(double) * 3 * S);
    copy(tmp.mass, b->mass, sizeof(double) * S);
    copy(tmp.id, b->id, sizeof(uint64_t) * S);
    copy(tmp.cell_of, b->cell_of, sizeof(int) * S);
    copy(tmp.start, b->start, sizeof(uint32_t) * b->cells_total);
    copy(tmp.rows, b->rows, sizeof(int32_t) * b->padding * S);
    copy(tmp.rcount, b->rcount, sizeof(int32_t) * S);
    copy(tmp.half, b->half, sizeof(int32_t) * S);
    copy(tmp.pairs, b->pairs, sizeof(unsigned long long) * b->count);
    if (b->tiled) {
        copy(tmp.tile_ranges, b->tile_ranges, sizeof(int4) * b->tiles * TILE_RANGES);
        copy(tmp.tile_nr, b->tile_nr, sizeof(int) * b->tiles);
        copy(tmp.rows16, b->rows16, sizeof(uint16_t) * rows16_entries(b));
    }
    MD_HIP(hipStreamSynchronize(st));
}

// Puts b's particle state (and force validity) back into the scratch batch before a trial, so every trial starts
// from the same state; the trial's relayout then re-sorts and rebuilds it.
void reset_particles(const md_batch* b, md_batch& tmp) {
    hipStream_t st = b->ctx->stream;
    const size_t S = b->slots;
    MD_HIP(hipMemcpyAsync(tmp.pos, b->pos, sizeof(double4) * S, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipMemcpyAsync(tmp.vel, b->vel, sizeof(double) * 3 * S, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipMemcpyAsync(tmp.frc, b->frc, sizeof(double) * 3 * S, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipMemcpyAsync(tmp.mass, b->mass, sizeof(double) * S, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipMemcpyAsync(tmp.id, b->id, sizeof(uint64_t) * S, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipStreamSynchronize(st));
    tmp.forces_valid = b->forces_valid;
}

// A trial's configuration: cfg over its first `steps` steps (each replica's Bussi draws taken from the start of its
// own [cfg->steps][2] block, so the trial's draws are laid out [count][steps][2] as md_batch_run reads them).
md_batch_config trial_config(const md_batch_config* cfg, int count, int steps, int lanes, std::vector<double>& draws) {
    md_batch_config c = *cfg;
    c.steps = steps;
    c.lanes = lanes;
    c.force_kernel_ms = nullptr;
    if (cfg->thermostat && cfg->draws) {
        draws.resize(static_cast<size_t>(count) * steps * 2);
        for (int r = 0; r < count; ++r)
            std::memcpy(draws.data() + static_cast<size_t>(r) * steps * 2,
                        cfg->draws + static_cast<size_t>(r) * cfg->steps * 2, sizeof(double) * steps * 2);
        c.draws = draws.data();
    }
    return c;
}

}  // namespace
}  // namespace mdgpu

using namespace mdgpu;

extern "C" {

int md_batch_create(md_ctx* ctx, md_state* const* states, int count, double cutoff, double skin, md_batch** out) {
    MD_API_BEGIN
    if (!ctx || !states || !out) fail(MD_ERR_ARG, "ctx/states/out is NULL");
    if (count <= 0 || count > 65535) fail(MD_ERR_ARG, "replica count must be in 1..65535");
    if (!(cutoff > 0.0) || !(skin >= 0.0)) fail(MD_ERR_ARG, "cutoff must be > 0 and skin >= 0");
    *out = nullptr;
    const double radius = cutoff + skin;
    auto b = new md_batch();
    struct Guard {
        md_batch* b;
        ~Guard() {
            if (b) {
                batch_free(b);
                delete b;
            }
        }
    } guard{b};
    b->ctx = ctx;
    b->count = count;
    b->cutoff = cutoff;
    b->skin = skin;
    b->radius2 = radius * radius;
    b->max_disp2 = 0.25 * skin * skin;
    b->reps.resize(count);
    size_t slot = 0, cell = 0, ke = 0;
    int padding = 8;
    for (int r = 0; r < count; ++r) {
        const md_state* s = states[r];
        if (!s) fail(MD_ERR_ARG, "a replica state is NULL");
        if (s->ctx != ctx) fail(MD_ERR_ARG, "every replica state must belong to the batch's context");
        if (s->n == 0) fail(MD_ERR_ARG, "a replica has no particles");
        if (2.0 * radius > s->box) fail(MD_ERR_RUNTIME, "Cutoff + skin must be less than half the box length");
        Rep& rp = b->reps[r];
        rp.first = static_cast<uint32_t>(slot);
        rp.n = static_cast<uint32_t>(s->n);
        rp.side = grid_side_of(s->box, radius);
        rp.use_cells = rp.side >= 3 ? 1 : 0;
        rp.ncells = static_cast<uint32_t>(rp.side) * rp.side * rp.side;
        rp.cell_first = static_cast<uint32_t>(cell);
        rp.box = s->box;
        rp.inv_box = s->inv_box;
        rp.per_length = rp.side / s->box;
        rp.ke_first = static_cast<uint32_t>(ke);
        rp.ke_blocks = static_cast<uint32_t>((s->n + KE_TERMS - 1) / KE_TERMS);
        slot += (s->n + BT - 1) / BT * BT;
        cell += rp.ncells + 1;
        ke += rp.ke_blocks;
        padding = std::max(padding, padding_for(s->n, s->box, radius));
        b->particles += s->n;
        if (slot > 0x7fffffffu || cell > 0xffffffffu) fail(MD_ERR_ARG, "batch too large for 32-bit slot indices");
    }
    b->slots = slot;
    b->cells_total = cell;
    b->blocks = static_cast<int>(slot / BT);
    b->ke_total = static_cast<int>(ke);
    b->padding = padding;
    // MDGPU_BATCH_PADDING=<rows> replaces the density-based initial row length (a short one exercises the regrowth
    // inside md_batch_run)
    if (const char* v = std::getenv("MDGPU_BATCH_PADDING")) {
        const int p = std::atoi(v);
        if (p > 0) b->padding = p;
    }
    batch_alloc(b);
    hipStream_t st = ctx->stream;
    for (int r = 0; r < count; ++r) {
        const md_state* s = states[r];
        const Rep& rp = b->reps[r];
        const size_t n = s->n;
        MD_HIP(hipMemcpyAsync(b->pos + rp.first, s->pos, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        for (int c = 0; c < 3; ++c) {
            MD_HIP(hipMemcpyAsync(b->vel + c * b->slots + rp.first, s->vel + c * n, sizeof(double) * n,
                                  hipMemcpyDeviceToDevice, st));
            MD_HIP(hipMemcpyAsync(b->frc + c * b->slots + rp.first, s->frc + c * n, sizeof(double) * n,
                                  hipMemcpyDeviceToDevice, st));
        }
        MD_HIP(hipMemcpyAsync(b->mass + rp.first, s->mass, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        MD_HIP(hipMemcpyAsync(b->id + rp.first, s->id, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
    }
    k_b_fill<<<blocks_for(count, 256), 256, 0, st>>>(b->stale, count, 1);
    MD_HIP(hipGetLastError());
    batch_build(b);
    MD_HIP(hipStreamSynchronize(st));
    ++ctx->live;
    guard.b = nullptr;
    *out = b;
    MD_API_END
}

int md_batch_destroy(md_batch* b) {
    MD_API_BEGIN
    if (!b) return MD_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    --b->ctx->live;
    batch_free(b);
    delete b;
    MD_API_END
}

int md_batch_get_info(const md_batch* b, md_batch_info* info) {
    MD_API_BEGIN
    if (!b || !info) fail(MD_ERR_ARG, "batch/info is NULL");
    std::vector<int> reb(b->count);
    std::vector<unsigned long long> pairs(b->count);
    hipStream_t st = b->ctx->stream;
    MD_HIP(hipMemcpyAsync(reb.data(), b->rebuilds, sizeof(int) * b->count, hipMemcpyDeviceToHost, st));
    MD_HIP(hipMemcpyAsync(pairs.data(), b->pairs, sizeof(unsigned long long) * b->count, hipMemcpyDeviceToHost, st));
    MD_HIP(hipStreamSynchronize(st));
    info->count = b->count;
    info->particles = b->particles;
    info->slots = b->slots;
    info->padding = b->padding;
    info->lanes = b->lanes;
    info->rebuilds = 0;
    info->half_pairs = 0;
    info->skin = b->skin;
    info->tiled = b->tiled ? 1 : 0;
    for (int r = 0; r < b->count; ++r) {
        info->rebuilds += reb[r];
        info->half_pairs += pairs[r];
    }
    MD_API_END
}

int md_batch_replica(md_batch* b, int r, md_state* dst, int* rebuilds, size_t* half_pairs) {
    MD_API_BEGIN
    if (!b) fail(MD_ERR_ARG, "batch is NULL");
    if (r < 0 || r >= b->count) fail(MD_ERR_ARG, "replica index out of range");
    const Rep& rp = b->reps[r];
    hipStream_t st = b->ctx->stream;
    if (dst) {
        if (dst->n != rp.n) fail(MD_ERR_ARG, "destination state holds a different particle count");
        const size_t n = rp.n;
        MD_HIP(hipMemcpyAsync(dst->pos, b->pos + rp.first, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        for (int c = 0; c < 3; ++c) {
            MD_HIP(hipMemcpyAsync(dst->vel + c * n, b->vel + c * b->slots + rp.first, sizeof(double) * n,
                                  hipMemcpyDeviceToDevice, st));
            MD_HIP(hipMemcpyAsync(dst->frc + c * n, b->frc + c * b->slots + rp.first, sizeof(double) * n,
                                  hipMemcpyDeviceToDevice, st));
        }
        MD_HIP(hipMemcpyAsync(dst->mass, b->mass + rp.first, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        MD_HIP(hipMemcpyAsync(dst->id, b->id + rp.first, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
        dst->box = rp.box;
        dst->inv_box = rp.inv_box;
    }
    int reb = 0;
    unsigned long long pairs = 0;
    MD_HIP(hipMemcpyAsync(&reb, b->rebuilds + r, sizeof(int), hipMemcpyDeviceToHost, st));
    MD_HIP(hipMemcpyAsync(&pairs, b->pairs + r, sizeof(pairs), hipMemcpyDeviceToHost, st));
    MD_HIP(hipStreamSynchronize(st));
    if (rebuilds) *rebuilds = reb;
    if (half_pairs) *half_pairs = static_cast<size_t>(pairs);
    MD_API_END
}

int md_batch_run(md_batch* b, const md_batch_config* cfg, double* energies) {
    MD_API_BEGIN
    if (!b || !cfg) fail(MD_ERR_ARG, "batch/config is NULL");
    if (cfg->steps < 0) fail(MD_ERR_ARG, "steps must be >= 0");
    if (cfg->precision != MD_FP64 && cfg->precision != MD_MIXED) fail(MD_ERR_ARG, "unknown precision");
    if (cfg->thermostat && (!cfg->draws || !cfg->temperatures))
        fail(MD_ERR_ARG, "Bussi thermostat needs per-replica temperatures and draws");
    const int lanes = cfg->lanes ? cfg->lanes : b->lanes;
    if (lanes != 1 && lanes != 2 && lanes != 4) fail(MD_ERR_ARG, "lanes must be 1, 2 or 4");
    check_boxes(b, cfg->lj.cutoff);
    hipStream_t st = b->ctx->stream;
    const LJConst lj = make_lj(cfg->lj);
    if (b->tiled && cfg->lj.cutoff != b->tsys_cutoff) upload_tile_systems(b, cfg->lj.cutoff);
    const size_t S = b->slots;
    const int count = b->count;
    if (!b->forces_valid) {
        enqueue_forces(b, lj, cfg->precision, lanes, nullptr, nullptr, 0.0);
        b->forces_valid = true;
    }
    const int steps = cfg->steps;
    if (steps == 0) {
        MD_HIP(hipStreamSynchronize(st));
        return MD_OK;
    }
    struct RunBuffers {
        double* energies = nullptr;
        double* draws = nullptr;
        int* draw_index = nullptr;
        double2* dof_kbar = nullptr;
        ~RunBuffers() {
            dev_free(energies);
            dev_free(draws);
            dev_free(draw_index);
            dev_free(dof_kbar);
        }
    } rb;
    rb.energies = dev_alloc<double>(static_cast<size_t>(steps) * count * 4);
    const double c = cfg->thermostat ? std::exp(-cfg->dt / cfg->tau) : 0.0;
    if (cfg->thermostat) {
        rb.draws = dev_alloc<double>(static_cast<size_t>(count) * steps * 2);
        MD_HIP(hipMemcpyAsync(rb.draws, cfg->draws, sizeof(double) * count * steps * 2, hipMemcpyHostToDevice, st));
        rb.draw_index = dev_alloc<int>(count);
        MD_HIP(hipMemsetAsync(rb.draw_index, 0, sizeof(int) * count, st));
        std::vector<double2> dk(count);
        for (int r = 0; r < count; ++r) {
            const double dof = 3.0 * static_cast<double>(b->reps[r].n) - 3.0;
            dk[r] = make_double2(dof, 0.5 * dof * cfg->temperatures[r]);
        }
        rb.dof_kbar = dev_alloc<double2>(count);
        MD_HIP(hipMemcpyAsync(rb.dof_kbar, dk.data(), sizeof(double2) * count, hipMemcpyHostToDevice, st));
        MD_HIP(hipStreamSynchronize(st));  // dk goes out of scope
    }
    MD_HIP(hipMemsetAsync(b->flags, 0, 4 * sizeof(int), st));
    MD_HIP(hipMemsetAsync(b->stale, 0, sizeof(int) * count, st));

    const bool timed = cfg->force_kernel_ms != nullptr && !cfg->use_graph;
    std::vector<hipEvent_t> events;
    struct EventGuard {
        std::vector<hipEvent_t>& ev;
        ~EventGuard() {
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        }
    } eguard{events};
    const double half_dt = 0.5 * cfg->dt;
    struct HaltGuard {
        md_batch* b;
        ~HaltGuard() { b->halt = nullptr; }
    } halt_guard{b};
    b->halt = b->flags;
    // Bussi runs: each opening kick first applies the previous step's velocity scale (1 before the first step);
    // the last step's scale is applied after the run.
    const double* alpha = cfg->thermostat ? b->scalars : nullptr;
    if (cfg->thermostat) {
        k_b_unit_alpha<<<blocks_for(count, 256), 256, 0, st>>>(count, b->scalars);
        MD_HIP(hipGetLastError());
    }
    auto enqueue_step = [&](bool kick_drift) {
        if (kick_drift)
            k_b_kick_drift<<<b->blocks, BT, 0, st>>>(b->d_reps, b->block_rep, b->pos, b->vel, b->frc, b->mass, S,
                                                     half_dt, cfg->dt, b->ref, b->max_disp2, b->stale, alpha,
                                                     b->halt);
        enqueue_rebuild(b);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) {  // one pair per force launch (a resumed step launches twice)
            MD_HIP(hipEventCreate(&e0));
            events.push_back(e0);
            MD_HIP(hipEventCreate(&e1));
            events.push_back(e1);
        }
        const bool fold = kick_folds(b, cfg);
        enqueue_forces(b, lj, cfg->precision, lanes, e0, e1, fold ? half_dt : 0.0);
        if (cfg->thermostat) {
            if (!fold) {
                k_b_kick_ke<<<b->ke_total, BT, 0, st>>>(b->d_reps, b->block_rep, b->ke_start, b->vel, b->frc,
                                                        b->mass, S, half_dt, b->ke_part, b->halt);
                k_b_ke_final<<<blocks_for(count, 256), 256, 0, st>>>(b->d_reps, count, b->ke_part, b->scalars,
                                                                     b->halt);
            }
            k_b_bussi<<<blocks_for(count, 256), 256, 0, st>>>(count, b->scalars, rb.draws, steps, rb.draw_index,
                                                              rb.dof_kbar, c, b->halt);
            k_b_record<<<1, 1024, 0, st>>>(count, b->scalars, 6, rb.energies, b->stale, b->rebuilds, b->flags,
                                           b->halt);
        } else {
            double* kick_partials = b->partials + static_cast<size_t>(b->blocks) * 4 * 3;
            k_b_kick<<<b->blocks, BT, 0, st>>>(b->d_reps, b->block_rep, b->vel, b->frc, b->mass, S, half_dt,
                                               kick_partials, b->halt);
            k_b_reduce_ke<<<blocks_for(count, 256), 256, 0, st>>>(b->d_reps, count, kick_partials, b->scalars, b->halt);
            k_b_record<<<1, 1024, 0, st>>>(count, b->scalars, 3, rb.energies, b->stale, b->rebuilds, b->flags,
                                           b->halt);
        }
        MD_HIP(hipGetLastError());
    };
    // `n_steps` steps; with `resume` the first one's opening kick and drift are already done
    auto run_steps = [&](int n_steps, bool resume) {
        int first = 0;
        if (resume) {
            enqueue_step(false);
            first = 1;
        }
        const int replays = n_steps - first;
        if (replays <= 0) return;
        if (!cfg->use_graph) {
            for (int k = 0; k < replays; ++k) enqueue_step(true);
            return;
        }
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        MD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_step(true);
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
    int* hf = b->ctx->host_flags;
    for (int attempt = 0, done = 0;; ++attempt) {
        run_steps(steps - done, attempt > 0);
        MD_HIP(hipMemcpyAsync(hf, b->flags, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
        if (!hf[0]) break;
        if (attempt >= 8) fail(MD_ERR_OVERFLOW, "neighbour-list rows keep overflowing during md_batch_run");
        // Frozen at the first unrecorded step, after its opening kick and drift and the rebuilds of its stale
        // replicas (whose stale flags are still set). Longer rows keep every other replica's entries where they
        // were (column-major, stride = slots); the resumed step repeats the stale replicas' rebuilds, re-sorting
        // nothing (they are sorted already).
        done = hf[2];
        const int longest = hf[1];
        grow_rows(b, (longest + longest / 4 + 16 + 7) / 8 * 8, true);
        MD_HIP(hipMemsetAsync(b->flags, 0, sizeof(int), st));
    }
    if (cfg->thermostat) {  // the last step's velocity scale
        k_b_scale<<<b->blocks, BT, 0, st>>>(b->d_reps, b->block_rep, b->vel, S, b->scalars, b->halt);
        MD_HIP(hipGetLastError());
    }
    if (energies)
        MD_HIP(hipMemcpyAsync(energies, rb.energies, sizeof(double) * steps * count * 4, hipMemcpyDeviceToHost, st));
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
    MD_API_END
}

int md_batch_tune(md_batch* b, const md_batch_config* cfg, int trial_steps, double* ms_per_step, int* best_lanes) {
    MD_API_BEGIN
    if (!b || !cfg) fail(MD_ERR_ARG, "batch/config is NULL");
    if (cfg->precision != MD_MIXED) fail(MD_ERR_ARG, "the autotuner picks lanes for the mixed-precision kernel");
    if (trial_steps <= 0 || trial_steps > cfg->steps) fail(MD_ERR_ARG, "trial_steps must be in 1..cfg->steps");
    // a scratch copy of the batch, so that the trials leave the batch itself where it is
    md_batch tmp;
    struct Guard {
        md_batch* t;
        ~Guard() { batch_free(t); }
    } guard{&tmp};
    scratch_copy(b, tmp);
    const int options[3] = {1, 2, 4};
    double best = 0.0;
    int best_l = b->lanes;
    std::vector<double> draws;
    for (int k = 0; k < 3; ++k) {
        const md_batch_config c = trial_config(cfg, b->count, trial_steps, options[k], draws);
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = md_batch_run(&tmp, &c, nullptr);
        if (rc != MD_OK) fail(rc, std::string("autotuner trial failed: ") + md_last_error());
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / trial_steps;
        if (ms_per_step) ms_per_step[k] = ms;
        if (k == 0 || ms < best) {
            best = ms;
            best_l = options[k];
        }
    }
    b->lanes = best_l;
    if (best_lanes) *best_lanes = best_l;
    MD_API_END
}

int md_batch_autotune(md_batch* b, const md_batch_config* cfg, int trial_steps, const md_batch_tune_space* space,
                      md_batch_tune_result* out) {
    MD_API_BEGIN
    if (!b || !cfg || !out) fail(MD_ERR_ARG, "batch/config/result is NULL");
    if (trial_steps <= 0 || trial_steps > cfg->steps) fail(MD_ERR_ARG, "trial_steps must be in 1..cfg->steps");
    const int n_skins = space && space->skins ? space->n_skins : 0;
    const int n_head = space && space->headroom ? space->n_headroom : 0;
    const bool try_lanes = space && space->tune_lanes && !b->tiled && cfg->precision == MD_MIXED;
    if (n_skins < 0 || n_head < 0) fail(MD_ERR_ARG, "negative candidate count");
    if (std::max(n_skins, 1) + n_head + (try_lanes ? 3 : 0) > MD_TUNE_MAX_TRIALS)
        fail(MD_ERR_ARG, "too many autotuner trials");
    for (int k = 0; k < n_skins; ++k) {
        const double s = space->skins[k];
        if (!(s >= 0.0)) fail(MD_ERR_ARG, "candidate skins must be >= 0");
        for (const Rep& rp : b->reps)
            if (2.0 * (b->cutoff + s) > rp.box)
                fail(MD_ERR_RUNTIME, "Cutoff + skin must be less than half the box length");
    }
    for (int k = 0; k < n_head; ++k)
        if (!(space->headroom[k] >= 1.0)) fail(MD_ERR_ARG, "padding headroom must be >= 1");
    std::memset(out, 0, sizeof(*out));
    md_batch tmp;
    struct Guard {
        md_batch* t;
        ~Guard() { batch_free(t); }
    } guard{&tmp};
    scratch_copy(b, tmp);
    hipStream_t st = b->ctx->stream;
    std::vector<double> draws;
    // one trial: b's state into the scratch batch, laid out for (skin, padding), then trial_steps timed steps
    // (the relayout's re-sort and rebuild are not timed); longest: the longest row built during the trial
    auto trial = [&](double skin, int padding, int lanes, int* longest) {
        reset_particles(b, tmp);
        relayout(&tmp, skin, padding);
        const md_batch_config c = trial_config(cfg, b->count, trial_steps, lanes, draws);
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = md_batch_run(&tmp, &c, nullptr);
        if (rc != MD_OK) fail(rc, std::string("autotuner trial failed: ") + md_last_error());
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / trial_steps;
        if (longest) {
            int f[2] = {0, 0};
            MD_HIP(hipMemcpy(f, tmp.flags, sizeof(f), hipMemcpyDeviceToHost));
            *longest = std::max(f[1], 1);
        }
        if (out->trials < MD_TUNE_MAX_TRIALS) {
            out->trial_skin[out->trials] = skin;
            out->trial_padding[out->trials] = tmp.padding;
            out->trial_lanes[out->trials] = lanes;
            out->trial_ms[out->trials] = ms;
            ++out->trials;
        }
        return ms;
    };
    const int lanes0 = b->lanes;
    // rebuild cadence: the list radius cutoff + skin, each skin with the density estimate of its padding
    double best_skin = b->skin, best_ms = 0.0;
    int longest = 0;
    if (n_skins == 0) {
        best_ms = trial(b->skin, b->padding, lanes0, &longest);
    } else {
        for (int k = 0; k < n_skins; ++k) {
            int lk = 0;
            const double ms = trial(space->skins[k], 0, lanes0, &lk);
            if (k == 0 || ms < best_ms) {
                best_ms = ms;
                best_skin = space->skins[k];
                longest = lk;
            }
        }
    }
    // list padding: multiples of the longest row the chosen skin built (a row that outgrows it costs a halt,
    // regrowth and resume inside the run, which the trial's time includes)
    int best_pad = 0;  // 0: the density estimate the skin trials used
    for (int k = 0; k < n_head; ++k) {
        const int pad = static_cast<int>(std::ceil(longest * space->headroom[k]));
        const double ms = trial(best_skin, pad, lanes0, nullptr);
        if (ms < best_ms) {
            best_ms = ms;
            best_pad = pad;
        }
    }
    // particles per wavefront of the per-particle force kernel (1, 2 or 4 lanes each)
    int best_lanes = lanes0;
    if (try_lanes) {
        const int options[3] = {1, 2, 4};
        for (int k = 0; k < 3; ++k) {
            const double ms = trial(best_skin, best_pad, options[k], nullptr);
            if (ms < best_ms) {
                best_ms = ms;
                best_lanes = options[k];
            }
        }
    }
    // the batch itself takes the chosen setting (re-sorted and rebuilt at its current state)
    const int want_pad = best_pad > 0 ? (best_pad + 7) / 8 * 8 : 0;
    if (best_skin != b->skin || (want_pad > 0 && want_pad != b->padding)) relayout(b, best_skin, want_pad);
    b->lanes = best_lanes;
    MD_HIP(hipStreamSynchronize(st));
    out->skin = best_skin;
    out->padding = b->padding;
    out->lanes = b->tiled ? 4 : best_lanes;
    out->ms_per_step = best_ms;
    MD_API_END
}

int md_batch_set_skin(md_batch* b, double skin, int padding) {
    MD_API_BEGIN
    if (!b) fail(MD_ERR_ARG, "batch is NULL");
    if (padding < 0) fail(MD_ERR_ARG, "padding must be >= 0");
    relayout(b, skin, padding);
    MD_API_END
}

}  // extern "C"
