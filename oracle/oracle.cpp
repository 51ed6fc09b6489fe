// CPU ORACLE — test infrastructure only. A clean-room restatement of the reference engine's hot path
// (nhantran1711/molecular-dynamics), written from its semantics; used by tests/, smoke() and bench.py's
// cpu_baseline leg as the checker, never by the product. Compile with -ffp-contract=off: every expression
// below is evaluated in the reference's source order with one rounding per operation, which is what makes the
// cell ids, the permutation, the neighbour sets, the deterministic-mode forces, the integrator and the
// kinetic-energy sums bit-comparable with the GPU path and with oracle/_ref.
//
// Pinned by (tests/test_oracle.py): oracle/_ref (the reference's own sources built here) bit for bit, the
// reference's closed-form tests (two-particle LJ, FCC spacing, minimum image, wrap, KE, virial), brute-force
// neighbour sets and O(N^2) forces, momentum and energy conservation, and the Bussi canonical moments.
//
// One deliberate difference: the reference's default list path sums per-thread force arrays and uses SIMD
// reductions whose order depends on the thread count and the vector width (Forcefield.cpp:105-207); this
// restatement runs it serially in list order, so it matches the reference there to rounding (~1e-15), not bits.
//
// Threads (OpenMP) only split loops whose results do not depend on the split: per-particle list rows (counted,
// then filled at prefix-sum offsets), per-pair and per-particle force terms, element-wise kicks and drifts,
// integer counts, and the 1024-term blocks of blocked_sum (whose block sums are still added in order). Every
// result is therefore the same for any thread count, and equal to the serial evaluation.
//
// Two additions without a reference function, for checking the GPU library's own extensions:
//  - a fixed-order list (no re-sort; each full row holds the j > i neighbours, then the j < i ones, each part in
//    increasing j, forces summed along the row): the arithmetic of the adjoint's forward run;
//  - the exact pair-distance histogram over every pair of the box (O(N^2)), the check of the cell-pass g(r).

#include "orc_capi.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numbers>
#include <omp.h>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

// ---------------------------------------------------------------------------------------------------------
// Particle storage (System.hpp:12-136, System.cpp:9-29): packed positions with 4 slots per particle, one array
// per velocity / force component, masses, and stable ids equal to the input order.
struct Particles {
    double box = 0.0;
    double inv_box = 0.0; // 1 / box, stored once (System.cpp:10)
    std::vector<double> pos; // 4 * n
    std::vector<double> vel[3];
    std::vector<double> frc[3];
    std::vector<double> mass;
    std::vector<std::uint64_t> id;

    std::size_t n() const { return mass.size(); }

    void resize(std::size_t count) {
        pos.assign(4 * count, 0.0);
        for (int k = 0; k < 3; ++k) {
            vel[k].assign(count, 0.0);
            frc[k].assign(count, 0.0);
        }
        mass.assign(count, 1.0);
        id.resize(count);
        for (std::size_t i = 0; i < count; ++i) id[i] = i;
    }
};

// Minimum image of ri - rj, one component at a time: d -= L * round(d / L) with round half away from zero
// (System.hpp:86-97).
inline void min_image(const Particles& p, const double* ri, const double* rj, double* d) {
    for (int k = 0; k < 3; ++k) {
        double dk = ri[k] - rj[k];
        const double images = std::round(dk * p.inv_box);
        d[k] = dk - p.box * images;
    }
}

inline double norm2(const double* d) { return d[0] * d[0] + d[1] * d[1] + d[2] * d[2]; }

// x - L * floor(x / L) (System.hpp:110-112).
inline double wrap(const Particles& p, double x) { return x - p.box * std::floor(x * p.inv_box); }

// new slot k takes old slot order[k], for all four position slots and every other per-particle array
// (System.cpp:42-64).
void permute(Particles& p, const std::uint64_t* order) {
    const std::size_t n = p.n();
    std::vector<double> pos(4 * n);
#pragma omp parallel for schedule(static)
    for (std::size_t k = 0; k < n; ++k)
        for (int c = 0; c < 4; ++c) pos[4 * k + c] = p.pos[4 * order[k] + c];
    p.pos.swap(pos);
    auto gather = [&](auto& arr) {
        auto out = arr;
#pragma omp parallel for schedule(static)
        for (std::size_t k = 0; k < n; ++k) out[k] = arr[order[k]];
        arr.swap(out);
    };
    for (int k = 0; k < 3; ++k) {
        gather(p.vel[k]);
        gather(p.frc[k]);
    }
    gather(p.mass);
    gather(p.id);
}

// Fixed-order sum: 1024-term blocks summed left to right from 0.0, then the block sums left to right from 0.0
// (Parallel.hpp:32-57). The result does not depend on the thread count.
template <typename Term>
double blocked_sum(std::size_t count, Term term) {
    constexpr std::size_t block = 1024;
    const std::size_t blocks = (count + block - 1) / block;
    std::vector<double> parts(blocks);
#pragma omp parallel for schedule(static)
    for (std::size_t b = 0; b < blocks; ++b) {
        double part = 0.0;
        const std::size_t end = std::min(count, (b + 1) * block);
        for (std::size_t i = b * block; i < end; ++i) part += term(i);
        parts[b] = part;
    }
    double total = 0.0;
    for (std::size_t b = 0; b < blocks; ++b) total += parts[b];
    return total;
}

// KE = sum of (0.5 m) (vx^2 + vy^2 + vz^2) in blocked order (System.cpp:67-76).
double kinetic(const Particles& p) {
    return blocked_sum(p.n(), [&](std::size_t i) {
        const double vx = p.vel[0][i], vy = p.vel[1][i], vz = p.vel[2][i];
        const double half_m = 0.5 * p.mass[i];
        return half_m * (vx * vx + vy * vy + vz * vz);
    });
}

// T = 2 KE / (3N - 3) (System.cpp:79-82).
double temperature(const Particles& p) {
    const double dof = 3.0 * static_cast<double>(p.n()) - 3.0;
    return 2.0 * kinetic(p) / dof;
}

// 64-bit FNV-1a over the little-endian bytes of x, vx, y, vy, z, vz of each particle (System.cpp:85-105).
std::uint64_t fnv1a(const Particles& p) {
    std::uint64_t h = 0xcbf29ce484222325ull;
    auto feed = [&h](double x) {
        std::uint64_t bits;
        std::memcpy(&bits, &x, 8);
        for (int b = 0; b < 8; ++b) {
            h = (h ^ ((bits >> (8 * b)) & 0xffull)) * 0x100000001b3ull;
        }
    };
    for (std::size_t i = 0; i < p.n(); ++i) {
        for (int k = 0; k < 3; ++k) {
            feed(p.pos[4 * i + k]);
            feed(p.vel[k][i]);
        }
    }
    return h;
}

// ---------------------------------------------------------------------------------------------------------
// Cell grid (CellList.cpp:9-83): side count max(1, int(L / minCell)), cell of a particle
// min(int(x * side / L), side - 1) per axis, id (cx * side + cy) * side + cz, and a stable counting sort.
struct Grid {
    double min_cell = 0.0;
    int side = 0;
    std::vector<int> cell_of;
    std::vector<std::uint64_t> start; // side^3 + 1
    std::vector<std::uint64_t> order; // particle indices, cell by cell, increasing within a cell

    void build(const Particles& p) {
        const std::size_t n = p.n();
        side = std::max(1, static_cast<int>(p.box / min_cell));
        const int cells = side * side * side;
        const double per_length = side / p.box;
        cell_of.resize(n);
        start.assign(static_cast<std::size_t>(cells) + 1, 0);
        order.resize(n);
        for (std::size_t i = 0; i < n; ++i) {
            int c[3];
            for (int k = 0; k < 3; ++k) {
                c[k] = std::min(static_cast<int>(p.pos[4 * i + k] * per_length), side - 1);
            }
            cell_of[i] = (c[0] * side + c[1]) * side + c[2];
            start[cell_of[i] + 1] += 1;
        }
        for (int c = 1; c <= cells; ++c) start[c] += start[c - 1];
        std::vector<std::uint64_t> next(start.begin(), start.end() - 1);
        for (std::size_t i = 0; i < n; ++i) order[next[cell_of[i]]++] = i;
    }

    // The 27 cells around `cell`, dx outermost, then dy, then dz, each from -1 to +1, wrapped periodically
    // (CellList.cpp:55-83).
    std::array<int, 27> stencil(int cell) const { return stencil_of(side, cell); }

    static std::array<int, 27> stencil_of(int s, int cell) {
        const int cx = cell / (s * s), cy = (cell / s) % s, cz = cell % s;
        std::array<int, 27> out{};
        int k = 0;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz)
                    out[k++] = (((cx + dx + s) % s) * s + (cy + dy + s) % s) * s + (cz + dz + s) % s;
        return out;
    }
};

// ---------------------------------------------------------------------------------------------------------
// Half Verlet list (NeighbourList.cpp:5-134). With `fixed` set (no reference function: the GPU adjoint's list)
// the state is not re-sorted and full rows are kept: row i holds its j > i neighbours in increasing j, then its
// j < i neighbours in increasing j; the half list (offsets, neigh) is then the j > i part of those rows.
struct HalfList {
    double radius = 0.0, radius2 = 0.0, max_disp2 = 0.0;
    bool reverse = false;
    bool fixed = false;
    Grid grid;
    std::vector<std::uint64_t> offsets, neigh, roffsets, rpairs;
    std::vector<std::uint64_t> full_offsets, full; // fixed mode: full rows in CSR
    std::vector<std::uint32_t> half_len;           // fixed mode: j > i entries at the start of each row
    std::vector<double> ref; // 3 per particle: positions at the last build
    int rebuilds = 0;

    void init(double rc, double skin, bool rev) {
        radius = rc + skin;
        radius2 = (rc + skin) * (rc + skin);
        max_disp2 = 0.25 * skin * skin;
        reverse = rev;
        grid.min_cell = rc + skin;
    }

    // Candidates of particle i in the reference's visiting order: the 27 stencil cells in order, each cell's
    // members in increasing index; every particle in increasing index on grids too small for a distinct stencil
    // (NeighbourList.cpp:46-50, 75-98).
    template <typename Visit>
    void visit(const Particles& p, std::size_t i, bool cells, Visit&& v) const {
        if (cells) {
            for (int c : grid.stencil(grid.cell_of[i]))
                for (std::uint64_t k = grid.start[c]; k < grid.start[c + 1]; ++k) v(grid.order[k]);
        } else {
            for (std::size_t j = 0; j < p.n(); ++j) v(j);
        }
    }

    bool near(const Particles& p, std::size_t i, std::size_t j) const {
        double d[3];
        min_image(p, &p.pos[4 * i], &p.pos[4 * j], d);
        return norm2(d) < radius2;
    }

    void build(Particles& p) {
        if (2.0 * radius > p.box) throw std::runtime_error("Cutoff + skin must be less than half the box length");
        const std::size_t n = p.n();
        grid.build(p);
        if (!fixed) {
            permute(p, grid.order.data());
            grid.build(p);
        }
        const bool cells = grid.side >= 3; // smaller grids would repeat cells in the stencil
        // rows are counted, then written at their prefix-sum offsets: the result does not depend on the threads
        std::vector<std::uint64_t> up(n + 1, 0), all(n + 1, 0);
#pragma omp parallel for schedule(dynamic, 256)
        for (std::size_t i = 0; i < n; ++i) {
            std::uint64_t a = 0, u = 0;
            visit(p, i, cells, [&](std::size_t j) {
                if (j == i || (!fixed && j < i) || !near(p, i, j)) return;
                ++a;
                u += j > i ? 1 : 0;
            });
            up[i + 1] = u;
            all[i + 1] = a;
        }
        for (std::size_t i = 0; i < n; ++i) {
            up[i + 1] += up[i];
            all[i + 1] += all[i];
        }
        offsets = up;
        neigh.assign(up[n], 0);
        if (fixed) {
            full_offsets = all;
            full.assign(all[n], 0);
            half_len.assign(n, 0);
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (std::size_t i = 0; i < n; ++i) {
            if (!fixed) {
                std::uint64_t at = offsets[i];
                visit(p, i, cells, [&](std::size_t j) {
                    if (j > i && near(p, i, j)) neigh[at++] = j;
                });
                continue;
            }
            const std::uint64_t b = full_offsets[i];
            const std::uint32_t h = static_cast<std::uint32_t>(offsets[i + 1] - offsets[i]);
            std::uint64_t hi = b, lo = b + h;
            visit(p, i, cells, [&](std::size_t j) {
                if (j == i || !near(p, i, j)) return;
                full[j > i ? hi++ : lo++] = j;
            });
            std::sort(full.begin() + static_cast<std::ptrdiff_t>(b), full.begin() + static_cast<std::ptrdiff_t>(b + h));
            std::sort(full.begin() + static_cast<std::ptrdiff_t>(b + h),
                      full.begin() + static_cast<std::ptrdiff_t>(full_offsets[i + 1]));
            half_len[i] = h;
            std::copy(full.begin() + static_cast<std::ptrdiff_t>(b), full.begin() + static_cast<std::ptrdiff_t>(b + h),
                      neigh.begin() + static_cast<std::ptrdiff_t>(offsets[i]));
        }
        if (reverse) {
            roffsets.assign(n + 1, 0);
            for (std::uint64_t j : neigh) roffsets[j + 1] += 1;
            for (std::size_t i = 0; i < n; ++i) roffsets[i + 1] += roffsets[i];
            rpairs.resize(neigh.size());
            std::vector<std::uint64_t> next(roffsets.begin(), roffsets.end() - 1);
            for (std::size_t k = 0; k < neigh.size(); ++k) rpairs[next[neigh[k]]++] = k;
        }
        ref.resize(3 * n);
#pragma omp parallel for schedule(static)
        for (std::size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) ref[3 * i + k] = p.pos[4 * i + k];
        ++rebuilds;
    }

    // Rebuild when the particle count changed or any particle moved more than skin / 2 (minimum image) since
    // the last build (NeighbourList.cpp:11-36).
    bool update(Particles& p) {
        const std::size_t n = p.n();
        if (ref.size() != 3 * n) {
            build(p);
            return true;
        }
        int stale = 0;
#pragma omp parallel for schedule(static) reduction(| : stale)
        for (std::size_t i = 0; i < n; ++i) {
            double d[3];
            min_image(p, &p.pos[4 * i], &ref[3 * i], d);
            stale |= norm2(d) > max_disp2 ? 1 : 0;
        }
        if (stale) build(p);
        return stale != 0;
    }
};

// ---------------------------------------------------------------------------------------------------------
// Truncated and shifted Lennard-Jones (Forcefield.cpp:8-450).
struct LJ {
    double eps = 1.0, sigma = 1.0, rc = 2.5, rc2 = 6.25, shift = 0.0;
    bool mixed = false, det = false;
    double virial = 0.0, inside = 0.0;

    void init(double e, double s, double cut) {
        eps = e;
        sigma = s;
        rc = cut;
        rc2 = cut * cut;
        const double s6 = std::pow(sigma / rc, 6);
        shift = 4.0 * eps * (s6 * s6 - s6); // V(rc) (Forcefield.cpp:12-14)
    }

    void check(const Particles& p) const {
        if (2.0 * rc > p.box) throw std::runtime_error("Cutoff must be less than half the box length");
    }

    // F/r and shifted pair energy at squared distance r2 (Forcefield.cpp:25-36).
    void pair(double r2, double& f_over_r, double& energy) const {
        const double inv_r2 = 1.0 / r2;
        const double s2 = sigma * sigma * inv_r2;
        const double s6 = s2 * s2 * s2;
        const double s12 = s6 * s6;
        f_over_r = 24.0 * eps * (2.0 * s12 - s6) * inv_r2;
        energy = 4.0 * eps * (s12 - s6) - shift;
    }

    void zero_forces(Particles& p) const {
        for (int k = 0; k < 3; ++k) std::fill(p.frc[k].begin(), p.frc[k].end(), 0.0);
    }

    // O(N^2) over i < j, pairs at r2 >= rc2 skipped, Newton's third law (Forcefield.cpp:39-78).
    double allpairs(Particles& p) {
        check(p);
        zero_forces(p);
        const std::size_t n = p.n();
        double u = 0.0, w = 0.0, cnt = 0.0;
        for (std::size_t i = 0; i < n; ++i) {
            for (std::size_t j = i + 1; j < n; ++j) {
                double d[3];
                min_image(p, &p.pos[4 * i], &p.pos[4 * j], d);
                const double r2 = norm2(d);
                if (r2 >= rc2) continue;
                double f, e;
                pair(r2, f, e);
                for (int k = 0; k < 3; ++k) {
                    p.frc[k][i] += f * d[k];
                    p.frc[k][j] -= f * d[k];
                }
                u += e;
                w += f * r2;
                cnt += 1.0;
            }
        }
        virial = w;
        inside = cnt;
        return u;
    }

    double list(Particles& p, const HalfList& nl) {
        check(p);
        if (nl.fixed) return list_fixed(p, nl);
        if (det) return list_deterministic(p, nl);
        if (mixed) return list_mixed(p, nl);
        return list_double(p, nl);
    }

    // Default FP64 list path, serially: every listed pair is evaluated, pairs at r2 >= rc2 are masked to zero,
    // +F/r d to i (summed over its row first) and -F/r d to j (Forcefield.cpp:81-211).
    double list_double(Particles& p, const HalfList& nl) {
        const std::size_t n = p.n();
        std::vector<double> acc(3 * n, 0.0);
        double u = 0.0, w = 0.0, cnt = 0.0;
        for (std::size_t i = 0; i < n; ++i) {
            double fi[3] = {0.0, 0.0, 0.0}, ui = 0.0, wi = 0.0, ci = 0.0;
            for (std::uint64_t k = nl.offsets[i]; k < nl.offsets[i + 1]; ++k) {
                const std::uint64_t j = nl.neigh[k];
                double d[3];
                min_image(p, &p.pos[4 * i], &p.pos[4 * j], d);
                const double r2 = norm2(d);
                double f, e;
                pair(r2, f, e);
                const bool in = r2 < rc2;
                f = in ? f : 0.0;
                ui += in ? e : 0.0;
                wi += f * r2;
                ci += in ? 1.0 : 0.0;
                for (int c = 0; c < 3; ++c) {
                    fi[c] += f * d[c];
                    acc[3 * j + c] -= f * d[c];
                }
            }
            for (int c = 0; c < 3; ++c) acc[3 * i + c] += fi[c];
            u += ui;
            w += wi;
            cnt += ci;
        }
        for (std::size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) p.frc[c][i] = acc[3 * i + c];
        virial = w;
        inside = cnt;
        return u;
    }

    // Deterministic path: one slot per listed pair, then particle i sums its own slots (+) in list order and
    // then the slots where it is second (-) in increasing pair index; U and W are blocked sums over the slots
    // (Forcefield.cpp:214-326).
    double list_deterministic(Particles& p, const HalfList& nl) {
        if (!nl.reverse)
            throw std::runtime_error("Deterministic forces need a NeighbourList built with reverseIndex = true");
        const std::size_t n = p.n(), pairs = nl.neigh.size();
        std::vector<double> fx(pairs), fy(pairs), fz(pairs), e(pairs), w(pairs);
        long long cnt = 0;  // an integer count: exact in any order
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : cnt)
        for (std::size_t i = 0; i < n; ++i) {
            for (std::uint64_t k = nl.offsets[i]; k < nl.offsets[i + 1]; ++k) {
                double d[3];
                min_image(p, &p.pos[4 * i], &p.pos[4 * nl.neigh[k]], d);
                const double r2 = norm2(d);
                double f, en;
                pair(r2, f, en);
                const bool in = r2 < rc2;
                f = in ? f : 0.0;
                fx[k] = f * d[0];
                fy[k] = f * d[1];
                fz[k] = f * d[2];
                e[k] = in ? en : 0.0;
                w[k] = f * r2;
                cnt += in ? 1 : 0;
            }
        }
#pragma omp parallel for schedule(static)
        for (std::size_t i = 0; i < n; ++i) {
            double s[3] = {0.0, 0.0, 0.0};
            for (std::uint64_t k = nl.offsets[i]; k < nl.offsets[i + 1]; ++k) {
                s[0] += fx[k];
                s[1] += fy[k];
                s[2] += fz[k];
            }
            for (std::uint64_t r = nl.roffsets[i]; r < nl.roffsets[i + 1]; ++r) {
                const std::uint64_t k = nl.rpairs[r];
                s[0] -= fx[k];
                s[1] -= fy[k];
                s[2] -= fz[k];
            }
            for (int c = 0; c < 3; ++c) p.frc[c][i] = s[c];
        }
        virial = blocked_sum(pairs, [&](std::size_t k) { return w[k]; });
        inside = static_cast<double>(cnt);
        return blocked_sum(pairs, [&](std::size_t k) { return e[k]; });
    }

    // Fixed-order list (no reference function; the GPU adjoint's forward run): particle i sums F/r d over its
    // full row in row order (j > i part, then j < i part, each in increasing j). U and W are taken over the j > i
    // part of each row, per particle, then in blocked order over particles.
    double list_fixed(Particles& p, const HalfList& nl) {
        const std::size_t n = p.n();
        std::vector<double> ui(n), wi(n);
        long long cnt = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : cnt)
        for (std::size_t i = 0; i < n; ++i) {
            double s[3] = {0.0, 0.0, 0.0}, u = 0.0, w = 0.0;
            const std::uint64_t b = nl.full_offsets[i];
            for (std::uint64_t k = b; k < nl.full_offsets[i + 1]; ++k) {
                double d[3];
                min_image(p, &p.pos[4 * i], &p.pos[4 * nl.full[k]], d);
                const double r2 = norm2(d);
                double f, en;
                pair(r2, f, en);
                const bool in = r2 < rc2;
                f = in ? f : 0.0;
                for (int c = 0; c < 3; ++c) s[c] += f * d[c];
                if (k - b < nl.half_len[i]) {
                    u += in ? en : 0.0;
                    w += f * r2;
                    cnt += in ? 1 : 0;
                }
            }
            for (int c = 0; c < 3; ++c) p.frc[c][i] = s[c];
            ui[i] = u;
            wi[i] = w;
        }
        virial = blocked_sum(n, [&](std::size_t i) { return wi[i]; });
        inside = static_cast<double>(cnt);
        return blocked_sum(n, [&](std::size_t i) { return ui[i]; });
    }

    // Mixed path: double minimum image narrowed to float, float pair maths and float per-row sums, double
    // accumulation across rows (Forcefield.cpp:333-450).
    double list_mixed(Particles& p, const HalfList& nl) {
        const std::size_t n = p.n();
        const float s2f = static_cast<float>(sigma * sigma);
        const float ef = static_cast<float>(eps);
        const float rc2f = static_cast<float>(rc2);
        const float shf = static_cast<float>(shift);
        std::vector<double> acc(3 * n, 0.0);
        double u = 0.0, w = 0.0, cnt = 0.0;
        for (std::size_t i = 0; i < n; ++i) {
            float fi[3] = {0.0f, 0.0f, 0.0f}, ui = 0.0f, wi = 0.0f, ci = 0.0f;
            for (std::uint64_t k = nl.offsets[i]; k < nl.offsets[i + 1]; ++k) {
                const std::uint64_t j = nl.neigh[k];
                double dd[3];
                min_image(p, &p.pos[4 * i], &p.pos[4 * j], dd);
                const float d[3] = {static_cast<float>(dd[0]), static_cast<float>(dd[1]), static_cast<float>(dd[2])};
                const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const float inv_r2 = 1.0f / r2;
                const float q2 = s2f * inv_r2;
                const float q6 = q2 * q2 * q2;
                const float q12 = q6 * q6;
                const bool in = r2 < rc2f;
                const float f = in ? 24.0f * ef * (2.0f * q12 - q6) * inv_r2 : 0.0f;
                ui += in ? 4.0f * ef * (q12 - q6) - shf : 0.0f;
                wi += f * r2;
                ci += in ? 1.0f : 0.0f;
                for (int c = 0; c < 3; ++c) {
                    fi[c] += f * d[c];
                    acc[3 * j + c] -= f * d[c];
                }
            }
            for (int c = 0; c < 3; ++c) acc[3 * i + c] += fi[c];
            u += ui;
            w += wi;
            cnt += ci;
        }
        for (std::size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) p.frc[c][i] = acc[3 * i + c];
        virial = w;
        inside = cnt;
        return u;
    }
};

// ---------------------------------------------------------------------------------------------------------
// Velocity Verlet (Integrator.cpp:7-71): v += (dt/2 f) / m; r = wrap(r + dt v); forces; v += (dt/2 f) / m.
void half_kick(Particles& p, double dt) {
    const double h = 0.5 * dt;
    const std::size_t n = p.n();
    for (int k = 0; k < 3; ++k) {
#pragma omp parallel for schedule(static)
        for (std::size_t i = 0; i < n; ++i) p.vel[k][i] += h * p.frc[k][i] / p.mass[i];
    }
}

void drift(Particles& p, double dt) {
    const std::size_t n = p.n();
    for (int k = 0; k < 3; ++k) {
#pragma omp parallel for schedule(static)
        for (std::size_t i = 0; i < n; ++i) p.pos[4 * i + k] = wrap(p, p.pos[4 * i + k] + dt * p.vel[k][i]);
    }
}

void scale_velocities(Particles& p, double s) {
    const std::size_t n = p.n();
    for (int k = 0; k < 3; ++k) {
#pragma omp parallel for schedule(static)
        for (std::size_t i = 0; i < n; ++i) p.vel[k][i] *= s;
    }
}

// s = sqrt(T0 / T), skipped when T <= 0 (Thermostat.cpp:17-30).
void rescale(Particles& p, double target) {
    const double t = temperature(p);
    if (t <= 0.0) return;
    scale_velocities(p, std::sqrt(target / t));
}

// Bussi stochastic velocity rescaling, eq. A7 of Bussi, Donadio & Parrinello (2007): one N(0,1) draw and a
// Gamma((dof-1)/2, 2) draw per call from mt19937_64, with fresh distribution objects each call
// (Thermostat.cpp:36-58).
struct Bussi {
    double target_t = 1.0, tau = 0.1;
    std::mt19937_64 rng;

    void apply(Particles& p, double dt) {
        const double ke = kinetic(p);
        if (ke <= 0.0) return;
        const double dof = 3.0 * static_cast<double>(p.n()) - 3.0;
        const double kbar = 0.5 * dof * target_t;
        const double c = std::exp(-dt / tau);
        std::normal_distribution<double> normal(0.0, 1.0);
        std::gamma_distribution<double> chi2(0.5 * (dof - 1.0), 2.0);
        const double r1 = normal(rng);
        const double sum_sq = chi2(rng);
        const double relax = (1.0 - c) * (kbar * (sum_sq + r1 * r1) / dof - ke);
        const double noise = 2.0 * r1 * std::sqrt(kbar / dof * ke * (1.0 - c) * c);
        const double ke_new = ke + relax + noise;
        scale_velocities(p, std::sqrt(ke_new / ke));
    }
};

// g(r) from half-list pairs with r2 < rmax^2 binned at size_t(sqrt(r2) / width) (RadialDistribution.cpp:6-69).
struct Rdf {
    double rmax = 0.0, width = 0.0;
    std::vector<double> hist;
    int samples = 0;
    double particles = 0.0, density = 0.0;

    void init(double r, int bins) {
        rmax = r;
        width = r / bins;
        hist.assign(static_cast<std::size_t>(bins), 0.0);
        samples = 0;
    }

    void sample(const Particles& p, const HalfList& nl) {
        const double rmax2 = rmax * rmax;
        for (std::size_t i = 0; i < p.n(); ++i) {
            for (std::uint64_t k = nl.offsets[i]; k < nl.offsets[i + 1]; ++k) {
                double d[3];
                min_image(p, &p.pos[4 * i], &p.pos[4 * nl.neigh[k]], d);
                const double r2 = norm2(d);
                if (r2 < rmax2) hist[static_cast<std::size_t>(std::sqrt(r2) / width)] += 1.0;
            }
        }
        particles = static_cast<double>(p.n());
        density = particles / std::pow(p.box, 3);
        ++samples;
    }

    void values(double* radii, double* g) const {
        for (std::size_t b = 0; b < hist.size(); ++b) {
            if (radii) radii[b] = (static_cast<double>(b) + 0.5) * width;
            if (!g) continue;
            if (samples == 0) {
                g[b] = 0.0;
                continue;
            }
            const double lo = static_cast<double>(b) * width;
            const double hi = lo + width;
            const double shell = 4.0 / 3.0 * std::numbers::pi * (hi * hi * hi - lo * lo * lo);
            g[b] = hist[b] / (samples * 0.5 * particles * density * shell);
        }
    }
};

// FCC: 4-atom basis at (0,0,0), (.5,.5,0), (.5,0,.5), (0,.5,.5), shifted by a/4, cells in ix, iy, iz order
// (Initialisation.cpp:7-44).
std::vector<std::array<double, 3>> fcc(int cells, double a) {
    static const double basis[4][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
    std::vector<std::array<double, 3>> out;
    out.reserve(4 * static_cast<std::size_t>(cells) * cells * cells);
    for (int ix = 0; ix < cells; ++ix)
        for (int iy = 0; iy < cells; ++iy)
            for (int iz = 0; iz < cells; ++iz)
                for (const auto& b : basis)
                    out.push_back({(ix + b[0] + 0.25) * a, (iy + b[1] + 0.25) * a, (iz + b[2] + 0.25) * a});
    return out;
}

// Maxwell-Boltzmann velocities: mt19937(seed), a new normal_distribution(0, sqrt(T/m)) per particle drawn for
// x, y, z; then the centre-of-mass velocity is removed and T rescaled exactly (Initialisation.cpp:47-76).
void maxwell_boltzmann(Particles& p, double t, std::uint32_t seed) {
    std::mt19937 rng(seed);
    double mom[3] = {0.0, 0.0, 0.0}, total_mass = 0.0;
    for (std::size_t i = 0; i < p.n(); ++i) {
        std::normal_distribution<double> gauss(0.0, std::sqrt(t / p.mass[i]));
        for (int k = 0; k < 3; ++k) {
            const double v = gauss(rng);
            p.vel[k][i] = v;
            mom[k] += p.mass[i] * v;
        }
        total_mass += p.mass[i];
    }
    for (int k = 0; k < 3; ++k)
        for (double& v : p.vel[k]) v -= mom[k] / total_mass;
    rescale(p, t);
}

} // namespace

struct orc_sim {
    Particles p;
    LJ lj;
    HalfList nl;
    Bussi bussi;
    Rdf rdf;
};

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }
void orc_set_threads(int threads) { omp_set_num_threads(threads > 0 ? threads : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

orc_sim* orc_create(size_t n, double box_length, const double* pos, const double* vel, const double* mass) {
    auto* s = new orc_sim;
    s->p.box = box_length;
    s->p.inv_box = 1.0 / box_length;
    s->p.resize(n);
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            s->p.pos[4 * i + k] = pos ? pos[3 * i + k] : 0.0;
            s->p.vel[k][i] = vel ? vel[3 * i + k] : 0.0;
        }
        s->p.mass[i] = mass ? mass[i] : 1.0;
    }
    s->lj.init(1.0, 1.0, 2.5);
    s->nl.init(2.5, 0.2, false);
    return s;
}

orc_sim* orc_create_lattice(int cells, double a, double box_length) {
    const auto sites = fcc(cells, a);
    std::vector<double> flat(3 * sites.size());
    for (size_t i = 0; i < sites.size(); ++i)
        for (int k = 0; k < 3; ++k) flat[3 * i + k] = sites[i][k];
    return orc_create(sites.size(), box_length, flat.data(), nullptr, nullptr);
}

orc_sim* orc_create_fcc(int cells, double density, double temp, uint32_t seed) {
    const double a = std::cbrt(4.0 / density);
    orc_sim* s = orc_create_lattice(cells, a, cells * a);
    maxwell_boltzmann(s->p, temp, seed);
    return s;
}

orc_sim* orc_clone(const orc_sim* s) { return new orc_sim(*s); }
void orc_destroy(orc_sim* s) { delete s; }
size_t orc_size(const orc_sim* s) { return s->p.n(); }
double orc_box_length(const orc_sim* s) { return s->p.box; }

void orc_get_state(const orc_sim* s, double* pos4, double* vel, double* force, double* mass, uint64_t* ids) {
    const size_t n = s->p.n();
    if (pos4) std::copy(s->p.pos.begin(), s->p.pos.end(), pos4);
    for (int k = 0; k < 3; ++k) {
        if (vel) std::copy(s->p.vel[k].begin(), s->p.vel[k].end(), vel + k * n);
        if (force) std::copy(s->p.frc[k].begin(), s->p.frc[k].end(), force + k * n);
    }
    if (mass) std::copy(s->p.mass.begin(), s->p.mass.end(), mass);
    if (ids) std::copy(s->p.id.begin(), s->p.id.end(), ids);
}

void orc_set_positions(orc_sim* s, const double* pos4) { std::copy(pos4, pos4 + s->p.pos.size(), s->p.pos.begin()); }

void orc_set_velocities(orc_sim* s, const double* vel) {
    const size_t n = s->p.n();
    for (int k = 0; k < 3; ++k) std::copy(vel + k * n, vel + (k + 1) * n, s->p.vel[k].begin());
}

int orc_reorder(orc_sim* s, const uint64_t* order) {
    permute(s->p, order);
    return 0;
}

void orc_min_image(const orc_sim* s, const double* ri, const double* rj, double* dr) { min_image(s->p, ri, rj, dr); }
double orc_wrap(const orc_sim* s, double x) { return wrap(s->p, x); }
double orc_kinetic_energy(const orc_sim* s) { return kinetic(s->p); }
double orc_temperature(const orc_sim* s) { return temperature(s->p); }
uint64_t orc_state_hash(const orc_sim* s) { return fnv1a(s->p); }

int orc_cells_build(orc_sim* s, double min_cell, int* side, int* cell_of, uint64_t* start, uint64_t* order) {
    Grid g;
    g.min_cell = min_cell;
    g.build(s->p);
    if (side) *side = g.side;
    if (cell_of) std::copy(g.cell_of.begin(), g.cell_of.end(), cell_of);
    if (start) std::copy(g.start.begin(), g.start.end(), start);
    if (order) std::copy(g.order.begin(), g.order.end(), order);
    return 0;
}

void orc_cells_stencil(int side, int cell, int* out27) {
    const auto st = Grid::stencil_of(side, cell);
    std::copy(st.begin(), st.end(), out27);
}

void orc_forcefield(orc_sim* s, double eps, double sigma, double rc) {
    const bool m = s->lj.mixed, d = s->lj.det;
    s->lj = LJ{};
    s->lj.init(eps, sigma, rc);
    s->lj.mixed = m;
    s->lj.det = d;
}
void orc_set_mixed(orc_sim* s, int mixed) { s->lj.mixed = mixed != 0; }
void orc_set_deterministic(orc_sim* s, int det) { s->lj.det = det != 0; }
double orc_energy_shift(const orc_sim* s) { return s->lj.shift; }
double orc_virial(const orc_sim* s) { return s->lj.virial; }
double orc_inside_pairs(const orc_sim* s) { return s->lj.inside; }

#define ORC_TRY(body)                                                                                          \
    try {                                                                                                      \
        body;                                                                                                  \
        return 0;                                                                                              \
    } catch (const std::exception& e) {                                                                        \
        g_error = e.what();                                                                                    \
        return -1;                                                                                             \
    }

int orc_forces_allpairs(orc_sim* s, double* u) { ORC_TRY(*u = s->lj.allpairs(s->p)) }
int orc_forces_list(orc_sim* s, double* u) { ORC_TRY(*u = s->lj.list(s->p, s->nl)) }

void orc_list(orc_sim* s, double cutoff, double skin, int reverse) {
    s->nl = HalfList{};
    s->nl.init(cutoff, skin, reverse != 0);
}
void orc_list_fixed(orc_sim* s, int fixed) { s->nl.fixed = fixed != 0; }

int orc_pair_histogram(const orc_sim* s, double rmax, int bins, uint64_t* counts) {
    if (bins <= 0 || !(rmax > 0.0)) {
        g_error = "bins and rmax must be positive";
        return -1;
    }
    const Particles& p = s->p;
    const std::size_t n = p.n();
    const double rmax2 = rmax * rmax, width = rmax / bins;
    std::vector<std::uint64_t> total(static_cast<std::size_t>(bins), 0);
#pragma omp parallel
    {
        std::vector<std::uint64_t> local(static_cast<std::size_t>(bins), 0);
        // min_image's arithmetic, with round(t) taken as the signed zero it returns for |t| < 1/2 (no library
        // call for the near images), and pairs dropped as soon as a partial sum of squares reaches rmax^2 (the
        // full sum can only be larger)
        auto comp = [&](double dk) {
            const double t = dk * p.inv_box;
            const double images = std::fabs(t) < 0.5 ? std::copysign(0.0, t) : std::round(t);
            return dk - p.box * images;
        };
#pragma omp for schedule(dynamic, 64)
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = i + 1; j < n; ++j) {
                const double* ri = &p.pos[4 * i];
                const double* rj = &p.pos[4 * j];
                const double dx = comp(ri[0] - rj[0]);
                const double xx = dx * dx;
                if (!(xx < rmax2)) continue;
                const double dy = comp(ri[1] - rj[1]);
                const double xy = xx + dy * dy;
                if (!(xy < rmax2)) continue;
                const double dz = comp(ri[2] - rj[2]);
                const double r2 = xy + dz * dz;
                if (!(r2 < rmax2)) continue;
                const std::size_t b = static_cast<std::size_t>(std::sqrt(r2) / width);
                if (b < local.size()) ++local[b];
            }
#pragma omp critical
        for (std::size_t b = 0; b < local.size(); ++b) total[b] += local[b];
    }
    std::copy(total.begin(), total.end(), counts);
    return 0;
}
int orc_list_build(orc_sim* s) { ORC_TRY(s->nl.build(s->p)) }
int orc_list_update(orc_sim* s, int* rebuilt) { ORC_TRY(*rebuilt = s->nl.update(s->p) ? 1 : 0) }
size_t orc_list_pairs(const orc_sim* s) { return s->nl.neigh.size(); }
int orc_list_rebuilds(const orc_sim* s) { return s->nl.rebuilds; }

void orc_list_half(const orc_sim* s, uint64_t* offsets, uint64_t* neigh) {
    if (offsets) std::copy(s->nl.offsets.begin(), s->nl.offsets.end(), offsets);
    if (neigh) std::copy(s->nl.neigh.begin(), s->nl.neigh.end(), neigh);
}

int orc_list_reverse(const orc_sim* s, uint64_t* roffsets, uint64_t* rpairs) {
    if (!s->nl.reverse) return -1;
    if (roffsets) std::copy(s->nl.roffsets.begin(), s->nl.roffsets.end(), roffsets);
    if (rpairs) std::copy(s->nl.rpairs.begin(), s->nl.rpairs.end(), rpairs);
    return 0;
}

int orc_step_list(orc_sim* s, double dt, double* u) {
    ORC_TRY({
        half_kick(s->p, dt);
        drift(s->p, dt);
        s->nl.update(s->p);
        *u = s->lj.list(s->p, s->nl);
        half_kick(s->p, dt);
    })
}

int orc_step_allpairs(orc_sim* s, double dt, double* u) {
    ORC_TRY({
        half_kick(s->p, dt);
        drift(s->p, dt);
        *u = s->lj.allpairs(s->p);
        half_kick(s->p, dt);
    })
}

void orc_rescale(orc_sim* s, double t) { rescale(s->p, t); }

void orc_bussi(orc_sim* s, double t, double tau, uint64_t seed) {
    s->bussi.target_t = t;
    s->bussi.tau = tau;
    s->bussi.rng.seed(seed);
}
void orc_bussi_apply(orc_sim* s, double dt) { s->bussi.apply(s->p, dt); }

void orc_rdf(orc_sim* s, double rmax, int bins) { s->rdf.init(rmax, bins); }
void orc_rdf_sample(orc_sim* s) { s->rdf.sample(s->p, s->nl); }
void orc_rdf_values(const orc_sim* s, double* radii, double* g) { s->rdf.values(radii, g); }

} // extern "C"
