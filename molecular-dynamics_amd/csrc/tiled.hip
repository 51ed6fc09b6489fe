// Tiled neighbour list and force kernels: the list build and the force evaluation stage each tile's halo
// positions in LDS once, instead of gathering every neighbour from L2/MALL.
//
// A tile is TILE consecutive particles of the cell-sorted state. Its halo is the union of the 27-cell stencils of
// the cells it spans, kept as at most RMAX ranges of whole cells (each range is a contiguous run of particle
// indices, since particles are sorted by cell). The halo of a tile of 128 particles at rho* = 0.8442 with
// rc + skin = 2.8 holds ~1700 particles (~2000 when the tile crosses a cell column), staged as x | y | z double
// arrays (24 B per particle).
//
// The list build writes every row twice: as global particle indices (int32, the list the rest of the library and
// the reference's half-list view use) and as LDS slots of the tile's staged halo (uint16, half the bytes, what
// the force kernel streams). Row order is the reference's (j > i in 27-cell stencil order, then j < i
// increasing), so the FP64 force sums stay bit-equal to the reference's deterministic mode. A tile whose halo
// does not fit (sparse or tiny systems) is marked with tile_nr = -1 and its rows are used through global
// gathers instead.
#include <cstdlib>
#include <cstring>

#include "device.hpp"
#include "internal.hpp"

namespace mdgpu {

constexpr int TILE = 128;
constexpr int RMAX = 64;
constexpr int CR_CAP = 192;  // cell ranges before merging (9 neighbour columns x 2 wrap pieces x segments)

struct TileArgs {
    int side;
    int use_cells;
    int hmax;  // LDS capacity in particles
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
    // insertion sort by first cell, then merge overlapping or adjacent ranges
    for (int k = 1; k < ncr; ++k) {
        const int2 v = cr[k];
        int m = k - 1;
        while (m >= 0 && cr[m].x > v.x) {
            cr[m + 1] = cr[m];
            --m;
        }
        cr[m + 1] = v;
    }
    int out = 0;
    for (int k = 0; k < ncr; ++k) {
        if (out > 0 && cr[k].x <= cr[out - 1].y + 1) {
            if (cr[k].y > cr[out - 1].y) cr[out - 1].y = cr[k].y;
        } else {
            cr[out++] = cr[k];
        }
    }
    return out <= RMAX ? out : -1;
}

// LDS offset of the first particle of cell `nc`, found by binary search over the merged cell ranges.
__device__ __forceinline__ int slot_base_of_cell(int nc, const int* ra, const int* rofs, const int* rgs,
                                                 const uint32_t* __restrict__ start, int nr) {
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ra[mid] <= nc)
            lo = mid;
        else
            hi = mid - 1;
    }
    return rofs[lo] + static_cast<int>(start[nc]) - rgs[lo];
}

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

__global__ void __launch_bounds__(TILE)
    k_list_build_tiled(const double4* __restrict__ pos, size_t n, const int* __restrict__ cell_of,
                       const uint32_t* __restrict__ start, TileArgs ta, double box, double inv_box, double radius2,
                       int padding, int32_t* __restrict__ rows, uint16_t* __restrict__ rows16,
                       int32_t* __restrict__ count, int32_t* __restrict__ half, int* __restrict__ flags,
                       unsigned long long* __restrict__ pair_total, int2* __restrict__ tile_ranges,
                       int* __restrict__ tile_nr, TiledCopyBack cb, const int* gate) {
    if (gate && *gate == 0) return;
    extern __shared__ double lds[];
    double* xs = lds;
    double* ys = lds + ta.hmax;
    double* zs = lds + 2 * ta.hmax;
    __shared__ int2 cr[CR_CAP];
    __shared__ int ra[RMAX], rofs[RMAX + 1], rgs[RMAX];
    __shared__ int s_nr;
    const size_t i0 = static_cast<size_t>(blockIdx.x) * TILE;
    const size_t i1 = (i0 + TILE) < n ? (i0 + TILE) : n;
    const int s = ta.side;
    if (threadIdx.x == 0) {
        int nr;
        if (ta.use_cells) {
            nr = tile_cell_ranges(cell_of[i0], cell_of[i1 - 1], s, cr);
        } else {
            cr[0] = make_int2(0, s * s * s - 1);  // every particle
            nr = 1;
        }
        int total = 0;
        if (nr > 0) {
            for (int r = 0; r < nr; ++r) {
                ra[r] = cr[r].x;
                rgs[r] = static_cast<int>(start[cr[r].x]);
                rofs[r] = total;
                total += static_cast<int>(start[cr[r].y + 1]) - rgs[r];
            }
            rofs[nr] = total;
            if (total > ta.hmax) nr = -1;
        }
        s_nr = nr;
        tile_nr[blockIdx.x] = nr;
        for (int r = 0; r < (nr > 0 ? nr : 0); ++r)
            tile_ranges[static_cast<size_t>(blockIdx.x) * RMAX + r] = make_int2(rgs[r], rofs[r + 1] - rofs[r]);
    }
    __syncthreads();
    const int nr = s_nr;
    const bool staged = nr > 0;
    if (staged) {
        for (int r = 0; r < nr; ++r) {
            const int cnt = rofs[r + 1] - rofs[r];
            for (int q = threadIdx.x; q < cnt; q += TILE) {
                const double4 p = pos[rgs[r] + q];
                xs[rofs[r] + q] = p.x;
                ys[rofs[r] + q] = p.y;
                zs[rofs[r] + q] = p.z;
            }
        }
    }
    __syncthreads();
    const size_t i = i0 + threadIdx.x;
    int cnt = 0, hcnt = 0;
    if (i < i1) {
        const double4 ri = pos[i];
        if (cb.enabled) {
            cb.pos[i] = ri;
            for (int c = 0; c < 3; ++c) {
                cb.vel[c * n + i] = cb.vel2[c * n + i];
                cb.frc[c * n + i] = cb.frc2[c * n + i];
            }
            cb.mass[i] = cb.mass2[i];
            cb.id[i] = cb.id2[i];
            cb.cell_of[i] = cb.cell_of2[i];
        }
        const uint32_t self = static_cast<uint32_t>(i);
        // one candidate j, whose LDS slot is `slot` when the tile is staged
        auto consider = [&](uint32_t j, int slot) {
            double xj, yj, zj;
            if (staged) {
                xj = xs[slot];
                yj = ys[slot];
                zj = zs[slot];
            } else {
                const double4 rj = pos[j];
                xj = rj.x;
                yj = rj.y;
                zj = rj.z;
            }
            const double dx = min_image(ri.x - xj, box, inv_box);
            const double dy = min_image(ri.y - yj, box, inv_box);
            const double dz = min_image(ri.z - zj, box, inv_box);
            if (dx * dx + dy * dy + dz * dz < radius2) {
                if (cnt < padding) {
                    rows[static_cast<size_t>(cnt) * n + i] = static_cast<int32_t>(j);
                    rows16[static_cast<size_t>(cnt) * n + i] = static_cast<uint16_t>(staged ? slot : 0);
                }
                ++cnt;
            }
        };
        if (ta.use_cells) {
            const int c = cell_of[i];
            const int cx = c / (s * s), cy = (c / s) % s, cz = c % s;
            for (int ox = -1; ox <= 1; ++ox) {
                const int x = wrap_index(cx + ox, s);
                for (int oy = -1; oy <= 1; ++oy) {
                    const int y = wrap_index(cy + oy, s);
                    for (int oz = -1; oz <= 1; ++oz) {
                        const int nc = (x * s + y) * s + wrap_index(cz + oz, s);
                        if (nc < c) continue;
                        const int base = staged ? slot_base_of_cell(nc, ra, rofs, rgs, start, nr) : 0;
                        const uint32_t b = start[nc], e = start[nc + 1];
                        for (uint32_t j = (nc == c ? self + 1 : b); j < e; ++j)
                            consider(j, base + static_cast<int>(j - b));
                    }
                }
            }
            hcnt = cnt;
            int xo[3] = {wrap_index(cx - 1, s), cx, wrap_index(cx + 1, s)};
            int yo[3] = {wrap_index(cy - 1, s), cy, wrap_index(cy + 1, s)};
            int zo[3] = {wrap_index(cz - 1, s), cz, wrap_index(cz + 1, s)};
            for (int a = 0; a < 2; ++a)
                for (int k = 0; k < 2 - a; ++k) {
                    if (xo[k] > xo[k + 1]) { const int t = xo[k]; xo[k] = xo[k + 1]; xo[k + 1] = t; }
                    if (yo[k] > yo[k + 1]) { const int t = yo[k]; yo[k] = yo[k + 1]; yo[k + 1] = t; }
                    if (zo[k] > zo[k + 1]) { const int t = zo[k]; zo[k] = zo[k + 1]; zo[k + 1] = t; }
                }
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2)
                    for (int d = 0; d < 3; ++d) {
                        const int nc = (xo[a] * s + yo[b2]) * s + zo[d];
                        if (nc > c) continue;
                        const int base = staged ? slot_base_of_cell(nc, ra, rofs, rgs, start, nr) : 0;
                        const uint32_t b = start[nc];
                        const uint32_t e = (nc == c) ? self : start[nc + 1];
                        for (uint32_t j = b; j < e; ++j) consider(j, base + static_cast<int>(j - b));
                    }
        } else {
            // every particle is in the single staged range, at slot = its index
            for (uint32_t j = self + 1; j < n; ++j) consider(j, static_cast<int>(j));
            hcnt = cnt;
            for (uint32_t j = 0; j < self; ++j) consider(j, static_cast<int>(j));
        }
        count[i] = cnt < padding ? cnt : padding;
        half[i] = hcnt < padding ? hcnt : padding;
    }
    int mx = cnt;
    unsigned long long hs = static_cast<unsigned long long>(hcnt);
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(mx, off, WAVE);
        mx = o > mx ? o : mx;
        hs += __shfl_xor(hs, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0) {
        atomicMax(&flags[FLAG_MAXCOUNT], mx);
        if (mx > padding) atomicOr(&flags[FLAG_OVERFLOW], 1);
        atomicAdd(pair_total, hs);
    }
    if (gate && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&flags[FLAG_REBUILDS], 1);
}

struct TiledLJ {
    double sig2, c24, c4, shift, rc2;
    float sig2f, c24f, c4f, shiftf, rc2f;
};

// Forces over the tiled list. Staged tiles read neighbour positions from LDS by slot; unstaged tiles gather them
// through the int32 rows. The arithmetic per pair is that of k_forces_fp64 / k_forces_mixed.
template <int MIXED>
__global__ void __launch_bounds__(TILE)
    k_forces_tiled(const double4* __restrict__ pos, size_t n, const uint16_t* __restrict__ rows16,
                   const int32_t* __restrict__ rows, const int32_t* __restrict__ count,
                   const int32_t* __restrict__ half, const int2* __restrict__ tile_ranges,
                   const int* __restrict__ tile_nr, int hmax, double box, double inv_box, TiledLJ lj,
                   double* __restrict__ frc, double* __restrict__ partials, int* __restrict__ clear_flag) {
    extern __shared__ double lds[];
    double* xs = lds;
    double* ys = lds + hmax;
    double* zs = lds + 2 * hmax;
    const int nr = tile_nr[blockIdx.x];
    const bool staged = nr > 0;
    if (staged) {
        int off = 0;
        for (int r = 0; r < nr; ++r) {
            const int2 rg = tile_ranges[static_cast<size_t>(blockIdx.x) * RMAX + r];
            for (int q = threadIdx.x; q < rg.y; q += TILE) {
                const double4 p = pos[rg.x + q];
                xs[off + q] = p.x;
                ys[off + q] = p.y;
                zs[off + q] = p.z;
            }
            off += rg.y;
        }
    }
    __syncthreads();
    const size_t i = static_cast<size_t>(blockIdx.x) * TILE + threadIdx.x;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const double4 ri = pos[i];
        const int m = count[i], h = half[i];
        for (int k = 0; k < m; ++k) {
            double xj, yj, zj;
            if (staged) {
                const int sl = rows16[static_cast<size_t>(k) * n + i];
                xj = xs[sl];
                yj = ys[sl];
                zj = zs[sl];
            } else {
                const double4 rj = pos[rows[static_cast<size_t>(k) * n + i]];
                xj = rj.x;
                yj = rj.y;
                zj = rj.z;
            }
            if (MIXED) {
                double ex = ri.x - xj, ey = ri.y - yj, ez = ri.z - zj;
                ex = fma(-box, rint(ex * inv_box), ex);
                ey = fma(-box, rint(ey * inv_box), ey);
                ez = fma(-box, rint(ez * inv_box), ez);
                const float dx = static_cast<float>(ex), dy = static_cast<float>(ey), dz = static_cast<float>(ez);
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const float q = __builtin_amdgcn_rcpf(r2);
                const float inv_r2 = fmaf(q, fmaf(-r2, q, 1.0f), q);
                const float s2 = lj.sig2f * inv_r2;
                const float s6 = s2 * s2 * s2;
                const float s12 = s6 * s6;
                const bool inside = r2 < lj.rc2f;
                const float f = inside ? lj.c24f * fmaf(2.0f, s12, -s6) * inv_r2 : 0.0f;
                fx += static_cast<double>(f * dx);
                fy += static_cast<double>(f * dy);
                fz += static_cast<double>(f * dz);
                if (k < h) {
                    acc[0] += inside ? static_cast<double>(fmaf(lj.c4f, s12 - s6, -lj.shiftf)) : 0.0;
                    acc[1] += static_cast<double>(f * r2);
                    acc[2] += inside ? 1.0 : 0.0;
                }
            } else {
                const double dx = min_image(ri.x - xj, box, inv_box);
                const double dy = min_image(ri.y - yj, box, inv_box);
                const double dz = min_image(ri.z - zj, box, inv_box);
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double inv_r2 = 1.0 / r2;
                const double s2 = lj.sig2 * inv_r2;
                const double s6 = s2 * s2 * s2;
                const double s12 = s6 * s6;
                const bool inside = r2 < lj.rc2;
                const double f = inside ? lj.c24 * (2.0 * s12 - s6) * inv_r2 : 0.0;
                fx += f * dx;
                fy += f * dy;
                fz += f * dz;
                if (k < h) {
                    acc[0] += inside ? lj.c4 * (s12 - s6) - lj.shift : 0.0;
                    acc[1] += f * r2;
                    acc[2] += inside ? 1.0 : 0.0;
                }
            }
        }
        frc[i] = fx;
        frc[n + i] = fy;
        frc[2 * n + i] = fz;
    }
    if (clear_flag && i == 0) *clear_flag = 0;
    block_sum<3, TILE>(acc, partials + static_cast<size_t>(blockIdx.x) * 3);
}

bool tiling_enabled() {
    static const bool enabled = [] {
        const char* v = std::getenv("MDGPU_UNTILED");
        return !(v && v[0] == '1');
    }();
    return enabled;
}

size_t tiled_lds_bytes(const md_list* l) { return static_cast<size_t>(l->hmax) * 3 * sizeof(double); }

int tile_count(size_t n) { return static_cast<int>((n + TILE - 1) / TILE); }

void tiled_list_enqueue(md_list* l, md_state* s, bool from_alt, const int* gate) {
    hipStream_t st = l->ctx->stream;
    const size_t n = s->n;
    if (n == 0) return;
    TiledCopyBack cb{};
    cb.enabled = from_alt ? 1 : 0;
    if (from_alt) {
        cb.vel2 = s->vel_alt;
        cb.frc2 = s->frc_alt;
        cb.mass2 = s->mass_alt;
        cb.id2 = s->id_alt;
        cb.cell_of2 = l->cell_of_alt;
        cb.pos = s->pos;
        cb.vel = s->vel;
        cb.frc = s->frc;
        cb.mass = s->mass;
        cb.id = s->id;
        cb.cell_of = l->cell_of;
    }
    TileArgs ta{l->side, l->side >= 3 ? 1 : 0, l->hmax};
    const int tiles = tile_count(n ? n : 1);
    hipLaunchKernelGGL(k_list_build_tiled, dim3(tiles), dim3(TILE), tiled_lds_bytes(l), st,
                       from_alt ? s->pos_alt : s->pos, n, from_alt ? l->cell_of_alt : l->cell_of, l->cell_start, ta,
                       s->box, s->inv_box, l->radius2, l->padding, l->rows, l->rows16, l->count, l->half, l->flags,
                       reinterpret_cast<unsigned long long*>(l->flags + FLAG_COUNT), l->tile_ranges, l->tile_nr, cb,
                       gate);
    MD_HIP(hipGetLastError());
}

void tiled_forces_enqueue(md_state* s, md_list* l, double sig2, double c24, double c4, double shift, double rc2,
                          int precision, double* out_dev, int* clear_flag, hipEvent_t ev_start, hipEvent_t ev_stop) {
    hipStream_t st = s->ctx->stream;
    const size_t n = s->n;
    const int tiles = tile_count(n ? n : 1);
    state_reserve_partials(s, static_cast<size_t>(tiles) * 3);
    TiledLJ lj;
    lj.sig2 = sig2;
    lj.c24 = c24;
    lj.c4 = c4;
    lj.shift = shift;
    lj.rc2 = rc2;
    lj.sig2f = static_cast<float>(sig2);
    lj.c24f = static_cast<float>(c24);
    lj.c4f = static_cast<float>(c4);
    lj.shiftf = static_cast<float>(shift);
    lj.rc2f = static_cast<float>(rc2);
    if (ev_start) MD_HIP(hipEventRecord(ev_start, st));
    if (n == 0) {
        if (ev_stop) MD_HIP(hipEventRecord(ev_stop, st));
        reduce_partials_enqueue(st, s->partials, 0, 3, out_dev);
        return;
    }
    if (precision == MD_MIXED)
        hipLaunchKernelGGL(k_forces_tiled<1>, dim3(tiles), dim3(TILE), tiled_lds_bytes(l), st, s->pos, n, l->rows16,
                           l->rows, l->count, l->half, l->tile_ranges, l->tile_nr, l->hmax, s->box, s->inv_box, lj,
                           s->frc, s->partials, clear_flag);
    else
        hipLaunchKernelGGL(k_forces_tiled<0>, dim3(tiles), dim3(TILE), tiled_lds_bytes(l), st, s->pos, n, l->rows16,
                           l->rows, l->count, l->half, l->tile_ranges, l->tile_nr, l->hmax, s->box, s->inv_box, lj,
                           s->frc, s->partials, clear_flag);
    MD_HIP(hipGetLastError());
    if (ev_stop) MD_HIP(hipEventRecord(ev_stop, st));
    reduce_partials_enqueue(st, s->partials, tiles, 3, out_dev);
}

}  // namespace mdgpu
