// Pair-distance histograms of RadialDistribution (reference: RadialDistribution.cpp:6-69).
//
// List pass (RadialDistribution::sample, RadialDistribution.cpp:11-30): over the half-list pairs (the first
// half[i] entries of each full-list row), r^2 < rmax^2, bin size_t(sqrt(r^2) / width). It sees only the pairs the
// list holds, which is every pair within rmax when rmax does not exceed the list radius.
//
// Cell pass (no reference function): every pair of the box within rmax, for rmax beyond the list radius (a g(r)
// to 5 sigma with a 2.8 sigma list). The particles are binned into a grid of cells at least rmax wide
// (cells = floor(L / rmax) per side, one counting pass, unordered within a cell) and each particle counts its
// partners j > i in the 27 surrounding cells; with fewer than 3 cells per side every j > i is a partner. The
// arithmetic per pair is the list pass's (minimum image, r^2, sqrt, division), so both passes give the same
// counts for the same pairs.
//
// Counts are integers, accumulated in a per-block LDS histogram and flushed with one 64-bit atomic per bin, so
// the result is exact (the reference adds 1.0 per pair to a double, exact below 2^53). Inside md_run the cell pass
// is gated on the step counter (a sample after every `every`-th step) and on the run's halt word.
#include <algorithm>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace mdgpu {

constexpr int RDF_THREADS = 256;
constexpr int RDF_LDS_BINS = 4096;
constexpr int RDF_SCAN_THREADS = 1024;

// One pair at squared distance r2 into the histogram (the LDS copy when the bins fit).
__device__ __forceinline__ void rdf_count(double r2, double rmax2, double width, int bins, bool use_lds,
                                          unsigned* local, unsigned long long* hist) {
    if (!(r2 < rmax2)) return;
    const double q = sqrt(r2) / width;
    const size_t b = static_cast<size_t>(q);
    if (b >= static_cast<size_t>(bins)) return;  // the reference would index past its array here
    if (use_lds)
        atomicAdd(&local[b], 1u);
    else
        atomicAdd(&hist[b], 1ull);
}

__device__ __forceinline__ double pair_r2(const double4& ri, const double4& rj, double box, double inv_box) {
    const double dx = min_image(ri.x - rj.x, box, inv_box);
    const double dy = min_image(ri.y - rj.y, box, inv_box);
    const double dz = min_image(ri.z - rj.z, box, inv_box);
    return dx * dx + dy * dy + dz * dz;
}

__global__ void __launch_bounds__(RDF_THREADS)
    k_rdf(const double4* __restrict__ pos, size_t n, const int32_t* __restrict__ rows,
          const int32_t* __restrict__ half, double box, double inv_box, double rmax2, double width, int bins,
          unsigned long long* __restrict__ hist) {
    __shared__ unsigned int local[RDF_LDS_BINS];
    const bool use_lds = bins <= RDF_LDS_BINS;
    if (use_lds)
        for (int b = threadIdx.x; b < bins; b += RDF_THREADS) local[b] = 0u;
    __syncthreads();
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const double4 ri = pos[i];
        const int h = half[i];
        for (int k = 0; k < h; ++k)
            rdf_count(pair_r2(ri, pos[rows[static_cast<size_t>(k) * n + i]], box, inv_box), rmax2, width, bins,
                      use_lds, local, hist);
    }
    __syncthreads();
    if (use_lds)
        for (int b = threadIdx.x; b < bins; b += RDF_THREADS)
            if (local[b]) atomicAdd(&hist[b], static_cast<unsigned long long>(local[b]));
}

// When the cell pass runs: outside md_run always; inside it after every `every`-th recorded step (the record
// counter plus `offset`, the steps whose record is still to come at that point of the step) unless halted.
struct RdfGate {
    const int* counter;
    int offset, every;
    const int* halt;
};

__device__ __forceinline__ bool rdf_skip(const RdfGate& g) {
    if (halted(g.halt)) return true;
    return g.counter && ((*g.counter + g.offset) % g.every) != 0;
}

__global__ void __launch_bounds__(RDF_THREADS)
    k_rdf_cells(const double4* __restrict__ pos, size_t n, double per_length, int side, int* __restrict__ cell_of,
                uint32_t* __restrict__ count, RdfGate gate) {
    if (rdf_skip(gate)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 r = pos[i];
    const int c = cell_of_position(r.x, r.y, r.z, per_length, side);
    cell_of[i] = c;
    atomicAdd(&count[c], 1u);
}

// One block: exclusive scan of the cell counts into start[0..ncells] (each thread a contiguous chunk), the counts
// cleared for the next pass, and the sample counted.
__global__ void __launch_bounds__(RDF_SCAN_THREADS)
    k_rdf_scan(uint32_t* __restrict__ count, size_t ncells, uint32_t* __restrict__ start,
               unsigned* __restrict__ samples, RdfGate gate) {
    if (rdf_skip(gate)) return;
    __shared__ uint32_t part[RDF_SCAN_THREADS];
    const size_t per = (ncells + RDF_SCAN_THREADS - 1) / RDF_SCAN_THREADS;
    const size_t c0 = threadIdx.x * per, c1 = std::min(ncells, c0 + per);
    uint32_t s = 0;
    for (size_t c = c0; c < c1; ++c) s += count[c];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int t = 0; t < RDF_SCAN_THREADS; ++t) {
            const uint32_t v = part[t];
            part[t] = run;
            run += v;
        }
        start[ncells] = run;
        *samples += 1u;
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (size_t c = c0; c < c1; ++c) {
        const uint32_t v = count[c];
        start[c] = run;
        run += v;
        count[c] = 0u;
    }
}

__global__ void __launch_bounds__(RDF_THREADS)
    k_rdf_scatter(const int* __restrict__ cell_of, size_t n, const uint32_t* __restrict__ start,
                  uint32_t* __restrict__ fill, uint32_t* __restrict__ members, RdfGate gate) {
    if (rdf_skip(gate)) return;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell_of[i];
    members[start[c] + atomicAdd(&fill[c], 1u)] = static_cast<uint32_t>(i);
}

// Particle i counts its partners j > i in the 27 cells around its own (every j > i on grids of fewer than 3 cells
// per side); the scatter's fill counters are cleared here for the next pass.
__global__ void __launch_bounds__(RDF_THREADS)
    k_rdf_cell_pairs(const double4* __restrict__ pos, size_t n, const int* __restrict__ cell_of,
                     const uint32_t* __restrict__ start, const uint32_t* __restrict__ members, int side,
                     size_t ncells, uint32_t* __restrict__ fill, double box, double inv_box, double rmax2,
                     double width, int bins, unsigned long long* __restrict__ hist, RdfGate gate) {
    if (rdf_skip(gate)) return;
    __shared__ unsigned int local[RDF_LDS_BINS];
    const bool use_lds = bins <= RDF_LDS_BINS;
    if (use_lds)
        for (int b = threadIdx.x; b < bins; b += RDF_THREADS) local[b] = 0u;
    __syncthreads();
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t c = i; c < ncells; c += static_cast<size_t>(gridDim.x) * RDF_THREADS) fill[c] = 0u;
    if (i < n) {
        const double4 ri = pos[i];
        if (side >= 3) {
            const int c = cell_of[i];
            const int cx = c / (side * side), cy = (c / side) % side, cz = c % side;
            for (int ox = -1; ox <= 1; ++ox)
                for (int oy = -1; oy <= 1; ++oy)
                    for (int oz = -1; oz <= 1; ++oz) {
                        const int nc = (wrap_index(cx + ox, side) * side + wrap_index(cy + oy, side)) * side +
                                       wrap_index(cz + oz, side);
                        const uint32_t e = start[nc + 1];
                        for (uint32_t k = start[nc]; k < e; ++k) {
                            const uint32_t j = members[k];
                            if (j > i) rdf_count(pair_r2(ri, pos[j], box, inv_box), rmax2, width, bins, use_lds,
                                                 local, hist);
                        }
                    }
        } else {
            for (size_t j = i + 1; j < n; ++j)
                rdf_count(pair_r2(ri, pos[j], box, inv_box), rmax2, width, bins, use_lds, local, hist);
        }
    }
    __syncthreads();
    if (use_lds)
        for (int b = threadIdx.x; b < bins; b += RDF_THREADS)
            if (local[b]) atomicAdd(&hist[b], static_cast<unsigned long long>(local[b]));
}

__global__ void k_rdf_clear(unsigned long long* __restrict__ hist, int bins, uint32_t* __restrict__ count,
                            uint32_t* __restrict__ fill, size_t ncells, unsigned* __restrict__ samples) {
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < static_cast<size_t>(bins)) hist[t] = 0ull;
    if (t < ncells) {
        count[t] = 0u;
        fill[t] = 0u;
    }
    if (t == 0) *samples = 0u;
}

// Device buffers of the cell pass for one state size and rmax (allocated before a run, so that the step can be
// captured in a graph).
struct RdfPass {
    int side = 0;
    size_t ncells = 0, n = 0;
    double per_length = 0.0, rmax2 = 0.0, width = 0.0;
    int bins = 0;
    int* cell_of = nullptr;
    uint32_t *count = nullptr, *start = nullptr, *fill = nullptr, *members = nullptr;
    unsigned long long* hist = nullptr;
    unsigned* samples = nullptr;
    ~RdfPass() {
        dev_free(cell_of);
        dev_free(count);
        dev_free(start);
        dev_free(fill);
        dev_free(members);
        dev_free(hist);
        dev_free(samples);
    }
};

RdfPass* rdf_pass_create(md_state* s, double rmax, int bins) {
    if (bins <= 0) fail(MD_ERR_ARG, "bins must be positive");
    if (!(rmax > 0.0)) fail(MD_ERR_ARG, "rmax must be positive");
    if (2.0 * rmax > s->box) fail(MD_ERR_ARG, "rmax must not exceed half the box length");
    auto* p = new RdfPass;
    try {
        const int side = std::max(1, static_cast<int>(s->box / rmax));
        p->side = side;
        p->ncells = static_cast<size_t>(side) * side * side;
        p->n = s->n;
        p->per_length = side / s->box;
        p->rmax2 = rmax * rmax;
        p->width = rmax / bins;
        p->bins = bins;
        p->cell_of = dev_alloc<int>(s->n);
        p->count = dev_alloc<uint32_t>(p->ncells);
        p->start = dev_alloc<uint32_t>(p->ncells + 1);
        p->fill = dev_alloc<uint32_t>(p->ncells);
        p->members = dev_alloc<uint32_t>(s->n);
        p->hist = dev_alloc<unsigned long long>(static_cast<size_t>(bins));
        p->samples = dev_alloc<unsigned>(1);
        const size_t m = std::max(p->ncells, static_cast<size_t>(bins));
        k_rdf_clear<<<blocks_for(m, 256), 256, 0, s->ctx->stream>>>(p->hist, bins, p->count, p->fill, p->ncells,
                                                                     p->samples);
        MD_HIP(hipGetLastError());
    } catch (...) {
        delete p;
        throw;
    }
    return p;
}

void rdf_pass_destroy(RdfPass* p) { delete p; }

// One cell-pass sample of s's current positions into p's histogram (gated as described at RdfGate).
void rdf_pass_enqueue(RdfPass* p, md_state* s, const int* counter, int offset, int every) {
    if (s->n == 0) return;
    if (s->n != p->n) fail(MD_ERR_STATE, "the pair histogram was set up for another particle count");
    hipStream_t st = s->ctx->stream;
    const RdfGate gate{counter, offset, every > 0 ? every : 1, s->halt};
    const unsigned blocks = blocks_for(s->n, RDF_THREADS);
    k_rdf_cells<<<blocks, RDF_THREADS, 0, st>>>(s->pos, s->n, p->per_length, p->side, p->cell_of, p->count, gate);
    k_rdf_scan<<<1, RDF_SCAN_THREADS, 0, st>>>(p->count, p->ncells, p->start, p->samples, gate);
    k_rdf_scatter<<<blocks, RDF_THREADS, 0, st>>>(p->cell_of, s->n, p->start, p->fill, p->members, gate);
    k_rdf_cell_pairs<<<blocks, RDF_THREADS, 0, st>>>(s->pos, s->n, p->cell_of, p->start, p->members, p->side,
                                                     p->ncells, p->fill, s->box, s->inv_box, p->rmax2, p->width,
                                                     p->bins, p->hist, gate);
    MD_HIP(hipGetLastError());
}

// Adds p's counts to the host array and returns the number of samples taken.
unsigned rdf_pass_read(RdfPass* p, md_state* s, uint64_t* counts) {
    hipStream_t st = s->ctx->stream;
    std::vector<unsigned long long> host(static_cast<size_t>(p->bins));
    unsigned samples = 0;
    MD_HIP(hipMemcpyAsync(host.data(), p->hist, sizeof(unsigned long long) * host.size(), hipMemcpyDeviceToHost,
                          st));
    MD_HIP(hipMemcpyAsync(&samples, p->samples, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    MD_HIP(hipStreamSynchronize(st));
    for (size_t b = 0; b < host.size(); ++b) counts[b] += host[b];
    return samples;
}

}  // namespace mdgpu

using namespace mdgpu;

extern "C" {

int md_rdf_sample(md_state* s, md_list* l, double rmax, int bins, uint64_t* counts) {
    MD_API_BEGIN
    if (!s || !l || !counts) fail(MD_ERR_ARG, "state/list/counts is NULL");
    if (bins <= 0) fail(MD_ERR_ARG, "bins must be positive");
    validate_list(s, l);
    if (rmax > l->radius) {
        // the list does not hold every pair out to rmax: count them all with the cell pass
        RdfPass* p = rdf_pass_create(s, rmax, bins);
        struct Guard {
            RdfPass* p;
            ~Guard() { rdf_pass_destroy(p); }
        } guard{p};
        rdf_pass_enqueue(p, s, nullptr, 0, 1);
        rdf_pass_read(p, s, counts);
        return MD_OK;
    }
    ensure_rows32(l);
    hipStream_t st = s->ctx->stream;
    unsigned long long* hist = dev_alloc<unsigned long long>(static_cast<size_t>(bins));
    std::vector<unsigned long long> host(static_cast<size_t>(bins));
    try {
        MD_HIP(hipMemsetAsync(hist, 0, sizeof(unsigned long long) * bins, st));
        if (s->n)
            k_rdf<<<blocks_for(s->n, RDF_THREADS), RDF_THREADS, 0, st>>>(s->pos, s->n, l->rows, l->half, s->box,
                                                                         s->inv_box, rmax * rmax, rmax / bins, bins,
                                                                         hist);
        MD_HIP(hipGetLastError());
        MD_HIP(hipMemcpyAsync(host.data(), hist, sizeof(unsigned long long) * bins, hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
    } catch (...) {
        dev_free(hist);
        throw;
    }
    dev_free(hist);
    for (int b = 0; b < bins; ++b) counts[b] += host[b];
    MD_API_END
}

int md_rdf_sample_all(md_state* s, double rmax, int bins, uint64_t* counts) {
    MD_API_BEGIN
    if (!s || !counts) fail(MD_ERR_ARG, "state/counts is NULL");
    RdfPass* p = rdf_pass_create(s, rmax, bins);
    struct Guard {
        RdfPass* p;
        ~Guard() { rdf_pass_destroy(p); }
    } guard{p};
    rdf_pass_enqueue(p, s, nullptr, 0, 1);
    rdf_pass_read(p, s, counts);
    MD_API_END
}

}  // extern "C"
