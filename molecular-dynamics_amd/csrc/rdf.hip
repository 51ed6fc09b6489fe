// Pair-distance histogram of RadialDistribution::sample (reference: RadialDistribution.cpp:11-30): over the
// half-list pairs (the first half[i] entries of each full-list row), r^2 < rmax^2, bin size_t(sqrt(r^2) / width).
// Counts are integers, accumulated in a per-block LDS histogram and flushed with one 64-bit atomic per bin, so
// the result is exact (the reference adds 1.0 per pair to a double, exact below 2^53).
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace mdgpu {

constexpr int RDF_THREADS = 256;
constexpr int RDF_LDS_BINS = 4096;

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
        for (int k = 0; k < h; ++k) {
            const double4 rj = pos[rows[static_cast<size_t>(k) * n + i]];
            const double dx = min_image(ri.x - rj.x, box, inv_box);
            const double dy = min_image(ri.y - rj.y, box, inv_box);
            const double dz = min_image(ri.z - rj.z, box, inv_box);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < rmax2) {
                const double q = sqrt(r2) / width;
                const size_t b = static_cast<size_t>(q);
                if (b >= static_cast<size_t>(bins)) continue;  // the reference would index past its array here
                if (use_lds)
                    atomicAdd(&local[b], 1u);
                else
                    atomicAdd(&hist[b], 1ull);
            }
        }
    }
    __syncthreads();
    if (use_lds)
        for (int b = threadIdx.x; b < bins; b += RDF_THREADS)
            if (local[b]) atomicAdd(&hist[b], static_cast<unsigned long long>(local[b]));
}

}  // namespace mdgpu

using namespace mdgpu;

extern "C" int md_rdf_sample(md_state* s, md_list* l, double rmax, int bins, uint64_t* counts) {
    try {
        if (!s || !l || !counts) fail(MD_ERR_ARG, "state/list/counts is NULL");
        if (bins <= 0) fail(MD_ERR_ARG, "bins must be positive");
        validate_list(s, l);
        ensure_rows32(l);
        hipStream_t st = s->ctx->stream;
        unsigned long long* hist = dev_alloc<unsigned long long>(static_cast<size_t>(bins));
        std::vector<unsigned long long> host(static_cast<size_t>(bins));
        try {
            MD_HIP(hipMemsetAsync(hist, 0, sizeof(unsigned long long) * bins, st));
            if (s->n)
                k_rdf<<<blocks_for(s->n, RDF_THREADS), RDF_THREADS, 0, st>>>(
                    s->pos, s->n, l->rows, l->half, s->box, s->inv_box, rmax * rmax, rmax / bins, bins, hist);
            MD_HIP(hipGetLastError());
            MD_HIP(hipMemcpyAsync(host.data(), hist, sizeof(unsigned long long) * bins, hipMemcpyDeviceToHost, st));
            MD_HIP(hipStreamSynchronize(st));
        } catch (...) {
            dev_free(hist);
            throw;
        }
        dev_free(hist);
        for (int b = 0; b < bins; ++b) counts[b] += host[b];
        return MD_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    }
}
