// Tiled neighbour list and mixed-precision force kernel: the list build and the force evaluation stage each
// tile's halo positions in LDS once, instead of gathering every neighbour from L2/MALL.
//
// A tile is TILE = 128 consecutive particles of the cell-sorted state, worked by 512 threads: LANES = 4 lanes per
// particle. Its halo is the union of the 27-cell stencils of the cells it spans, kept as at most RMAX ranges of
// whole cells (each a contiguous run of particle indices, since particles are sorted by cell). At rho* = 0.8442
// and rc + skin = 2.8 a halo holds ~1700 particles (~2400 when the tile crosses a cell column); the capacity
// `hmax` (<= 2560, 24 B each = 60 KB) is set from the density, and a tile that does not fit is marked
// tile_nr = -1 and handled through global gathers.
//
// The build writes each row twice: as global indices (int32, column-major: what the reference's half-list view,
// the RDF, the adjoint and the FP64 force kernel read) and as LDS slots of the tile's halo (uint16, in chunks of
// LANES entries: entry k of particle i at ((k / 4) * n + i) * 4 + k % 4, so a wavefront of 16 particles x 4 lanes
// reads 128 contiguous bytes per chunk). Rows keep the reference's order (j > i in 27-cell stencil order, then
// j < i increasing): the 4 lanes test 4 consecutive candidates and a wavefront ballot compacts the accepted ones
// in order.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "device.hpp"
#include "internal.hpp"

namespace mdgpu {

constexpr int TILE = 128;
constexpr int LANES = 4;
constexpr int TILE_THREADS = TILE * LANES;
constexpr int RMAX = 64;
constexpr int CR_CAP = 192;  // cell ranges before merging (9 neighbour columns x 2 wrap pieces x segments)

__device__ __forceinline__ size_t slot_index(size_t n, size_t i, int k) {
    return ((static_cast<size_t>(k) >> 2) * n + i) * LANES + (k & 3);
}

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

// LDS slot of the first particle of cell `nc`, by binary search over the merged cell ranges.
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

__global__ void __launch_bounds__(TILE_THREADS)
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
            for (int q = threadIdx.x; q < cnt; q += TILE_THREADS) {
                const double4 p = pos[rgs[r] + q];
                xs[rofs[r] + q] = p.x;
                ys[rofs[r] + q] = p.y;
                zs[rofs[r] + q] = p.z;
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & (LANES - 1);
    const int group_shift = (threadIdx.x & (WAVE - 1)) & ~(LANES - 1);
    const unsigned lower_mask = (1u << lane) - 1u;
    const size_t i = i0 + (threadIdx.x >> 2);
    int cnt = 0, hcnt = 0;
    if (i < i1) {
        const double4 ri = pos[i];
        if (cb.enabled && lane == 0) {
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
        // candidates j in [b, e) of one cell, 4 at a time; `base` is the LDS slot of j = b
        auto scan = [&](uint32_t b, uint32_t e, int base) {
            for (uint32_t jb = b; jb < e; jb += LANES) {
                const uint32_t j = jb + lane;
                bool accept = false;
                int slot = 0;
                if (j < e) {
                    double xj, yj, zj;
                    slot = base + static_cast<int>(j - b);
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
                    accept = dx * dx + dy * dy + dz * dz < radius2;
                }
                const unsigned group = static_cast<unsigned>((__ballot(accept) >> group_shift) & 0xFull);
                if (accept) {
                    const int at = cnt + __popc(group & lower_mask);
                    if (at < padding) {
                        rows[static_cast<size_t>(at) * n + i] = static_cast<int32_t>(j);
                        rows16[slot_index(n, i, at)] = static_cast<uint16_t>(staged ? slot : 0);
                    }
                }
                cnt += __popc(group);
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
                        if (nc == c)
                            scan(self + 1, e, base + static_cast<int>(self + 1 - b));
                        else
                            scan(b, e, base);
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
                        scan(b, nc == c ? self : start[nc + 1], base);
                    }
        } else {
            // every particle is in the single staged range, at slot = its index
            scan(self + 1, static_cast<uint32_t>(n), static_cast<int>(self + 1));
            hcnt = cnt;
            scan(0, self, 0);
        }
        if (lane == 0) {
            count[i] = cnt < padding ? cnt : padding;
            half[i] = hcnt < padding ? hcnt : padding;
        }
    }
    int mx = cnt;
    unsigned long long hs = lane == 0 ? static_cast<unsigned long long>(hcnt) : 0ull;
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(mx, off, WAVE);
        mx = o > mx ? o : mx;
        hs += __shfl_xor(hs, off, WAVE);
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        atomicMax(&flags[FLAG_MAXCOUNT], mx);
        if (mx > padding) atomicOr(&flags[FLAG_OVERFLOW], 1);
        atomicAdd(pair_total, hs);
    }
    if (gate && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&flags[FLAG_REBUILDS], 1);
}

struct MixedLJ {
    float sig2, c24, c4, shift, rc2;
    float box_hi, box_lo, inv_box;
    double box, inv_box_d, reach;  // reach: rc + list radius + skin, for the tile compactness test
    double frame;                  // 1.5 * 2^ceil(log2 L): the centre of the tile frame's binade
};

// Position as a float pair in the tile frame: the image of x nearest to `ref`, moved so that `ref` sits at `frame`:
// x' = x - L round((x - ref) / L) - ref + frame (double), hi = float(x'), lo = float(x' - hi). Every x' lies in
// [frame - L/2, frame + L/2], inside one binade [2^m, 2^(m+1)), so all hi values share one spacing and any
// difference of two of them (also L-sized ones across the box) is exact in float: (hi_i - hi_j) + (lo_i - lo_j)
// is then the separation rounded once to float, as accurate as narrowing the double separation.
__device__ __forceinline__ void split_near(double x, double ref, const MixedLJ& lj, float& hi, float& lo) {
    const double xs = x - lj.box * rint((x - ref) * lj.inv_box_d) - ref + lj.frame;
    hi = static_cast<float>(xs);
    lo = static_cast<float>(xs - static_cast<double>(hi));
}

// One pair in float: separation, LJ terms (Forcefield.cpp:352-446 arithmetic, with FMA), accumulation.
template <bool WRAP>
__device__ __forceinline__ void mixed_pair(float dxh, float dyh, float dzh, float dxl, float dyl, float dzl,
                                           const MixedLJ& lj, bool count_energy, double& fx, double& fy, double& fz,
                                           double& u, double& w, double& c) {
    if (WRAP) {  // minimum image on the exact hi differences (tiles not compact against the box)
        const float kx = rintf(dxh * lj.inv_box), ky = rintf(dyh * lj.inv_box), kz = rintf(dzh * lj.inv_box);
        dxh = fmaf(-kx, lj.box_hi, dxh);
        dyh = fmaf(-ky, lj.box_hi, dyh);
        dzh = fmaf(-kz, lj.box_hi, dzh);
        dxl = fmaf(-kx, lj.box_lo, dxl);
        dyl = fmaf(-ky, lj.box_lo, dyl);
        dzl = fmaf(-kz, lj.box_lo, dzl);
    }
    const float dx = dxh + dxl, dy = dyh + dyl, dz = dzh + dzl;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float q = __builtin_amdgcn_rcpf(r2);
    const float inv_r2 = fmaf(q, fmaf(-r2, q, 1.0f), q);
    const float s2 = lj.sig2 * inv_r2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    const bool inside = r2 < lj.rc2;
    const float f = inside ? lj.c24 * fmaf(2.0f, s12, -s6) * inv_r2 : 0.0f;
    // pair forces reach a few hundred near contact and cancel in the sum: accumulate in double
    fx += static_cast<double>(f * dx);
    fy += static_cast<double>(f * dy);
    fz += static_cast<double>(f * dz);
    if (count_energy) {
        u += inside ? static_cast<double>(fmaf(lj.c4, s12 - s6, -lj.shift)) : 0.0;
        w += static_cast<double>(f * r2);
        c += inside ? 1.0 : 0.0;
    }
}

// Mixed-precision forces over the tiled list: lane `lane` of particle i takes row entries k = lane, lane + 4, ...
// Staged tiles read float hi/lo positions from LDS by slot; unstaged tiles gather doubles through the int32 rows.
__global__ void __launch_bounds__(TILE_THREADS)
    k_forces_tiled_mixed(const double4* __restrict__ pos, size_t n, const uint16_t* __restrict__ rows16,
                         const int32_t* __restrict__ rows, const int32_t* __restrict__ count,
                         const int32_t* __restrict__ half, const int2* __restrict__ tile_ranges,
                         const int* __restrict__ tile_nr, int hmax, MixedLJ lj, double* __restrict__ frc,
                         double* __restrict__ partials, int* __restrict__ clear_flag) {
    extern __shared__ float4 ldsf[];
    float4* ha = ldsf;                                      // xh, yh, zh, xl
    float2* hb = reinterpret_cast<float2*>(ldsf + hmax);  // yl, zl
    __shared__ float extent[3];
    const int nr = tile_nr[blockIdx.x];
    const bool staged = nr > 0;
    const size_t i0 = static_cast<size_t>(blockIdx.x) * TILE;
    const double4 ref = pos[i0];
    if (threadIdx.x < 3) extent[threadIdx.x] = 0.0f;
    __syncthreads();
    if (staged) {
        int off = 0;
        for (int r = 0; r < nr; ++r) {
            const int2 rg = tile_ranges[static_cast<size_t>(blockIdx.x) * RMAX + r];
            for (int q = threadIdx.x; q < rg.y; q += TILE_THREADS) {
                const double4 p = pos[rg.x + q];
                float xh, yh, zh, xl, yl, zl;
                split_near(p.x, ref.x, lj, xh, xl);
                split_near(p.y, ref.y, lj, yh, yl);
                split_near(p.z, ref.z, lj, zh, zl);
                ha[off + q] = make_float4(xh, yh, zh, xl);
                hb[off + q] = make_float2(yl, zl);
            }
            off += rg.y;
        }
    }
    const int lane = threadIdx.x & (LANES - 1);
    const size_t i = i0 + (threadIdx.x >> 2);
    const bool active = i < n;
    float xih = 0.f, yih = 0.f, zih = 0.f, xil = 0.f, yil = 0.f, zil = 0.f;
    if (active) {
        const double4 ri = pos[i];
        split_near(ri.x, ref.x, lj, xih, xil);
        split_near(ri.y, ref.y, lj, yih, yil);
        split_near(ri.z, ref.z, lj, zih, zil);
        if (lane == 0) {  // how far the tile's own particles lie from the reference particle, per axis
            const float c = static_cast<float>(lj.frame);
            atomicMax(reinterpret_cast<int*>(&extent[0]), __float_as_int(fabsf(xih - c)));
            atomicMax(reinterpret_cast<int*>(&extent[1]), __float_as_int(fabsf(yih - c)));
            atomicMax(reinterpret_cast<int*>(&extent[2]), __float_as_int(fabsf(zih - c)));
        }
    }
    __syncthreads();
    // compact tile: every listed pair within the cutoff is separated by less than L/2 in the shifted frame
    const float e = fmaxf(extent[0], fmaxf(extent[1], extent[2]));
    const bool compact = staged && (2.0f * e + static_cast<float>(lj.reach) < static_cast<float>(lj.box) - 1.0f);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double acc[3] = {0.0, 0.0, 0.0};
    if (active) {
        const int m = count[i], h = half[i];
        if (staged) {
            for (int k = lane; k < m; k += LANES) {
                const int sl = rows16[slot_index(n, i, k)];
                const float4 a = ha[sl];
                const float2 b = hb[sl];
                if (compact)
                    mixed_pair<false>(xih - a.x, yih - a.y, zih - a.z, xil - a.w, yil - b.x, zil - b.y, lj, k < h, fx,
                                      fy, fz, acc[0], acc[1], acc[2]);
                else
                    mixed_pair<true>(xih - a.x, yih - a.y, zih - a.z, xil - a.w, yil - b.x, zil - b.y, lj, k < h, fx,
                                     fy, fz, acc[0], acc[1], acc[2]);
            }
        } else {
            const double4 ri = pos[i];
            for (int k = lane; k < m; k += LANES) {
                const double4 rj = pos[rows[static_cast<size_t>(k) * n + i]];
                float xjh, yjh, zjh, xjl, yjl, zjl;
                split_near(rj.x, ri.x, lj, xjh, xjl);
                split_near(rj.y, ri.y, lj, yjh, yjl);
                split_near(rj.z, ri.z, lj, zjh, zjl);
                float xh, yh, zh, xl, yl, zl;
                split_near(ri.x, ri.x, lj, xh, xl);
                split_near(ri.y, ri.y, lj, yh, yl);
                split_near(ri.z, ri.z, lj, zh, zl);
                mixed_pair<false>(xh - xjh, yh - yjh, zh - zjh, xl - xjl, yl - yjl, zl - zjl, lj, k < h, fx, fy, fz,
                                  acc[0], acc[1], acc[2]);
            }
        }
    }
    // the particle's four lanes are adjacent in the wavefront
    fx += __shfl_xor(fx, 1, WAVE);
    fy += __shfl_xor(fy, 1, WAVE);
    fz += __shfl_xor(fz, 1, WAVE);
    fx += __shfl_xor(fx, 2, WAVE);
    fy += __shfl_xor(fy, 2, WAVE);
    fz += __shfl_xor(fz, 2, WAVE);
    if (active && lane == 0) {
        frc[i] = fx;
        frc[n + i] = fy;
        frc[2 * n + i] = fz;
    }
    if (clear_flag && blockIdx.x == 0 && threadIdx.x == 0) *clear_flag = 0;
    block_sum<3, TILE_THREADS>(acc, partials + static_cast<size_t>(blockIdx.x) * 3);
}

bool tiling_enabled() {
    static const bool enabled = [] {
        const char* v = std::getenv("MDGPU_UNTILED");
        return !(v && v[0] == '1');
    }();
    return enabled;
}

int tile_count(size_t n) { return static_cast<int>((n + TILE - 1) / TILE); }

size_t tiled_rows16_entries(size_t n, int padding) {
    return static_cast<size_t>((padding + LANES - 1) / LANES) * LANES * (n ? n : 1);
}

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
    hipLaunchKernelGGL(k_list_build_tiled, dim3(tile_count(n)), dim3(TILE_THREADS),
                       static_cast<size_t>(l->hmax) * 3 * sizeof(double), st, from_alt ? s->pos_alt : s->pos, n,
                       from_alt ? l->cell_of_alt : l->cell_of, l->cell_start, ta, s->box, s->inv_box, l->radius2,
                       l->padding, l->rows, l->rows16, l->count, l->half, l->flags,
                       reinterpret_cast<unsigned long long*>(l->flags + FLAG_COUNT), l->tile_ranges, l->tile_nr, cb,
                       gate);
    MD_HIP(hipGetLastError());
}

void tiled_forces_enqueue(md_state* s, md_list* l, double sig2, double c24, double c4, double shift, double rc2,
                          double cutoff, double* out_dev, int* clear_flag, hipEvent_t ev_start, hipEvent_t ev_stop) {
    hipStream_t st = s->ctx->stream;
    const size_t n = s->n;
    const int tiles = tile_count(n ? n : 1);
    state_reserve_partials(s, static_cast<size_t>(tiles) * 3);
    MixedLJ lj;
    lj.sig2 = static_cast<float>(sig2);
    lj.c24 = static_cast<float>(c24);
    lj.c4 = static_cast<float>(c4);
    lj.shift = static_cast<float>(shift);
    lj.rc2 = static_cast<float>(rc2);
    lj.box = s->box;
    lj.inv_box_d = s->inv_box;
    lj.box_hi = static_cast<float>(s->box);
    lj.box_lo = static_cast<float>(s->box - static_cast<double>(lj.box_hi));
    lj.inv_box = static_cast<float>(s->inv_box);
    lj.reach = cutoff + l->radius + l->skin;
    lj.frame = 1.5 * std::exp2(std::ceil(std::log2(s->box)));
    if (ev_start) MD_HIP(hipEventRecord(ev_start, st));
    if (n > 0) {
        const size_t lds = static_cast<size_t>(l->hmax) * (sizeof(float4) + sizeof(float2));
        hipLaunchKernelGGL(k_forces_tiled_mixed, dim3(tiles), dim3(TILE_THREADS), lds, st, s->pos, n, l->rows16,
                           l->rows, l->count, l->half, l->tile_ranges, l->tile_nr, l->hmax, lj, s->frc, s->partials,
                           clear_flag);
        MD_HIP(hipGetLastError());
    }
    if (ev_stop) MD_HIP(hipEventRecord(ev_stop, st));
    reduce_partials_enqueue(st, s->partials, n > 0 ? tiles : 0, 3, out_dev);
}

}  // namespace mdgpu
