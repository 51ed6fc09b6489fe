// Contexts, particle state, reductions and the System-level entry points of the mdgpu C ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <random>
#include <string>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace mdgpu {

namespace {
thread_local std::string g_error;
}

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(MD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void set_error(const std::string& m) { g_error = m; }
const char* get_error() { return g_error.c_str(); }

// ----------------------------------------------------------------------------------------------------------
// Reductions

// Rows of W partials summed by RED_BLOCKS blocks: block g takes the contiguous rows [g R, (g + 1) R), R = ceil(rows /
// RED_BLOCKS); thread t sums rows g R + t, g R + t + 256, ... in order, the block adds its threads' sums (wavefront
// butterflies, then the 4 wavefronts in order) into block_sums[g], and the last block to finish (it draws the
// last ticket, and resets the counter for the next launch) adds the RED_BLOCKS block sums in order. The order is
// fixed for a given row count, so results repeat bit for bit; no memory is cleared by the host, so the launch can
// sit in a captured graph.
constexpr int RED_BLOCKS = 64;
constexpr int RED_THREADS = 256;
constexpr int RED_WIDTH_MAX = 8;

template <int W>
__global__ void __launch_bounds__(RED_THREADS) k_reduce_partials(const double* __restrict__ partials, int rows,
                                                                 double* __restrict__ block_sums,
                                                                 unsigned* __restrict__ ticket,
                                                                 double* __restrict__ out, StepRecord rec,
                                                                 const int* halt) {
    if (halted(halt)) return;
    __shared__ double lds[RED_THREADS / WAVE][W];
    __shared__ double fin[RED_BLOCKS][W];
    __shared__ int s_last;
    const int per = (rows + RED_BLOCKS - 1) / RED_BLOCKS;
    const int r0 = blockIdx.x * per;
    const int r1 = min(rows, r0 + per);
    double s[W];
#pragma unroll
    for (int k = 0; k < W; ++k) s[k] = 0.0;
    for (int r = r0 + static_cast<int>(threadIdx.x); r < r1; r += RED_THREADS)
#pragma unroll
        for (int k = 0; k < W; ++k) s[k] += partials[static_cast<size_t>(r) * W + k];
#pragma unroll
    for (int k = 0; k < W; ++k) s[k] = wave_sum(s[k]);
    if (threadIdx.x % WAVE == 0)
#pragma unroll
        for (int k = 0; k < W; ++k) lds[threadIdx.x / WAVE][k] = s[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < W; ++k) {
            double t = 0.0;
            for (int w = 0; w < RED_THREADS / WAVE; ++w) t += lds[w][k];
            block_sums[blockIdx.x * RED_WIDTH_MAX + k] = t;
        }
        __threadfence();  // the block's sums are visible before its ticket is drawn
        s_last = atomicAdd(ticket, 1u) == RED_BLOCKS - 1;
    }
    __syncthreads();
    if (!s_last) return;
    if (threadIdx.x < RED_BLOCKS) {
        __threadfence();
#pragma unroll
        for (int k = 0; k < W; ++k) fin[threadIdx.x][k] = block_sums[threadIdx.x * RED_WIDTH_MAX + k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum[W];
        for (int k = 0; k < W; ++k) {
            double t = 0.0;
            for (int g = 0; g < RED_BLOCKS; ++g) t += fin[g][k];
            out[k] = t;
            sum[k] = t;
        }
        if constexpr (W == 4) {
            if (rec.energies) {  // the run's record of this step from (U, W, inside count, KE): U, KE, W, count
                const int step = *rec.step;
                rec.energies[4 * step + 0] = sum[0];
                rec.energies[4 * step + 1] = sum[3];
                rec.energies[4 * step + 2] = sum[1];
                rec.energies[4 * step + 3] = sum[2];
                *rec.step = step + 1;
            }
        }
        *ticket = 0u;
    }
}

void reduce_partials_enqueue(md_state* s, int blocks, int width, double* out_dev, StepRecord rec) {
    hipStream_t st = s->ctx->stream;
    const double* partials = s->partials;
    double* bs = s->red_blocks;
    unsigned* tk = s->red_ticket;
    const int* h = s->halt;
    if (rec.energies && width != 4) fail(MD_ERR_ARG, "a step record needs the four-sum reduction");
    switch (width) {
        case 1: k_reduce_partials<1><<<RED_BLOCKS, RED_THREADS, 0, st>>>(partials, blocks, bs, tk, out_dev, rec, h); break;
        case 3: k_reduce_partials<3><<<RED_BLOCKS, RED_THREADS, 0, st>>>(partials, blocks, bs, tk, out_dev, rec, h); break;
        case 4: k_reduce_partials<4><<<RED_BLOCKS, RED_THREADS, 0, st>>>(partials, blocks, bs, tk, out_dev, rec, h); break;
        case 6: k_reduce_partials<6><<<RED_BLOCKS, RED_THREADS, 0, st>>>(partials, blocks, bs, tk, out_dev, rec, h); break;
        default: fail(MD_ERR_ARG, "unsupported reduction width");
    }
    MD_HIP(hipGetLastError());
}

void reduce_partials_enqueue(md_state* s, int blocks, int width, double* out_dev) {
    reduce_partials_enqueue(s, blocks, width, out_dev, StepRecord{});
}

// Kinetic energy in the reference's order (Parallel.hpp:32-57): block b of 1024 particles is summed left to
// right from 0.0 by one thread after the whole block has staged its terms (0.5 m)(vx^2 + vy^2 + vz^2) in LDS.
constexpr int KE_BLOCK = 1024;

__global__ void __launch_bounds__(256) k_ke_blocks(const double* __restrict__ vel, const double* __restrict__ mass,
                                                    size_t n, double* __restrict__ partial, const int* halt) {
    if (halted(halt)) return;
    __shared__ double terms[KE_BLOCK];
    const size_t base = static_cast<size_t>(blockIdx.x) * KE_BLOCK;
    for (int t = threadIdx.x; t < KE_BLOCK; t += 256) {
        const size_t i = base + t;
        double term = 0.0;
        if (i < n) {
            const double vx = vel[i], vy = vel[n + i], vz = vel[2 * n + i];
            const double half_m = 0.5 * mass[i];
            term = half_m * (vx * vx + vy * vy + vz * vz);
        }
        terms[t] = term;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const size_t end = (n - base) < KE_BLOCK ? (n - base) : KE_BLOCK;
        double s = 0.0;
        for (size_t t = 0; t < end; ++t) s += terms[t];
        partial[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(256) k_ke_final(const double* __restrict__ partial, int blocks,
                                                   double* __restrict__ out, const int* halt) {
    if (halted(halt)) return;
    __shared__ double stage[2048];
    double total = 0.0;
    for (int base = 0; base < blocks; base += 2048) {
        const int cnt = (blocks - base) < 2048 ? (blocks - base) : 2048;
        __syncthreads();
        for (int t = threadIdx.x; t < cnt; t += 256) stage[t] = partial[base + t];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int t = 0; t < cnt; ++t) total += stage[t];
    }
    if (threadIdx.x == 0) *out = total;
}

void state_reserve_partials(md_state* s, size_t count) {
    if (count > s->partials_cap || !s->partials) {
        // growing frees and allocates: never inside a stream capture (callers reserve before they capture)
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        MD_HIP(hipStreamIsCapturing(s->ctx->stream, &cs));
        if (cs != hipStreamCaptureStatusNone) fail(MD_ERR_STATE, "reduction buffer would grow inside a graph capture");
    }
    dev_reserve(s->partials, s->partials_cap, count);
}

void kinetic_exact_enqueue(md_state* s, double* out_dev) {
    hipStream_t st = s->ctx->stream;
    const int blocks = static_cast<int>((s->n + KE_BLOCK - 1) / KE_BLOCK);
    if (blocks == 0) {
        MD_HIP(hipMemsetAsync(out_dev, 0, sizeof(double), st));
        return;
    }
    state_reserve_partials(s, blocks);
    k_ke_blocks<<<blocks, 256, 0, st>>>(s->vel, s->mass, s->n, s->partials, s->halt);
    k_ke_final<<<1, 256, 0, st>>>(s->partials, blocks, out_dev, s->halt);
    MD_HIP(hipGetLastError());
}

// v *= scale (Thermostat.cpp:6-12); the scale comes from the host or from a device word.
__global__ void k_scale(double* __restrict__ vel, size_t count, double scale, const double* __restrict__ dev_scale) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const double s = dev_scale ? *dev_scale : scale;
    vel[i] = vel[i] * s;
}

// new slot k <- old slot order[k] for every array (System.cpp:42-64)
__global__ void k_reorder(const uint32_t* __restrict__ order, size_t n, const double4* __restrict__ pos,
                          const double* __restrict__ vel, const double* __restrict__ frc,
                          const double* __restrict__ mass, const uint64_t* __restrict__ id, double4* __restrict__ pos2,
                          double* __restrict__ vel2, double* __restrict__ frc2, double* __restrict__ mass2,
                          uint64_t* __restrict__ id2) {
    const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t p = order[k];
    pos2[k] = pos[p];
    for (int c = 0; c < 3; ++c) {
        vel2[c * n + k] = vel[c * n + p];
        frc2[c * n + k] = frc[c * n + p];
    }
    mass2[k] = mass[p];
    id2[k] = id[p];
}

__global__ void k_init_state(size_t n, double* mass, uint64_t* id) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mass[i] = 1.0;
    id[i] = i;
}

void swap_alt(md_state* s) {
    std::swap(s->pos, s->pos_alt);
    std::swap(s->vel, s->vel_alt);
    std::swap(s->frc, s->frc_alt);
    std::swap(s->mass, s->mass_alt);
    std::swap(s->id, s->id_alt);
}

}  // namespace mdgpu

using namespace mdgpu;


extern "C" {

const char* md_last_error(void) { return mdgpu::get_error(); }

int md_device_count(int* count) {
    MD_API_BEGIN
    if (!count) fail(MD_ERR_ARG, "count is NULL");
    MD_HIP(hipGetDeviceCount(count));
    MD_API_END
}

int md_ctx_create(int device, md_ctx** out) {
    MD_API_BEGIN
    if (!out) fail(MD_ERR_ARG, "out is NULL");
    auto* c = new md_ctx;
    try {
        c->device = device;
        MD_HIP(hipSetDevice(device));
        MD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        MD_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->host_scalars), 64 * sizeof(double), hipHostMallocDefault));
        MD_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->host_flags), 16 * sizeof(int), hipHostMallocDefault));
        c->run_scratch = run_scratch_address(device);
    } catch (...) {
        delete c;
        throw;
    }
    *out = c;
    MD_API_END
}

int md_ctx_destroy(md_ctx* c) {
    MD_API_BEGIN
    if (!c) return MD_OK;
    if (c->live > 0) fail(MD_ERR_STATE, "context still owns states or lists; destroy them first");
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->host_scalars) (void)hipHostFree(c->host_scalars);
    if (c->host_flags) (void)hipHostFree(c->host_flags);
    delete c;
    MD_API_END
}

int md_ctx_synchronize(md_ctx* c) {
    MD_API_BEGIN
    if (!c) fail(MD_ERR_ARG, "ctx is NULL");
    MD_HIP(hipStreamSynchronize(c->stream));
    MD_API_END
}

int md_state_create(md_ctx* ctx, size_t n, double box_length, md_state** out) {
    MD_API_BEGIN
    if (!ctx || !out) fail(MD_ERR_ARG, "ctx/out is NULL");
    if (!(box_length > 0.0)) fail(MD_ERR_ARG, "box length must be positive");
    MD_HIP(hipSetDevice(ctx->device));
    auto* s = new md_state;
    try {
        s->ctx = ctx;
        ++ctx->live;
        s->n = n;
        s->box = box_length;
        s->inv_box = 1.0 / box_length;
        s->pos = dev_alloc<double4>(n);
        s->vel = dev_alloc<double>(3 * n);
        s->frc = dev_alloc<double>(3 * n);
        s->mass = dev_alloc<double>(n);
        s->id = dev_alloc<uint64_t>(n);
        s->pos_alt = dev_alloc<double4>(n);
        s->vel_alt = dev_alloc<double>(3 * n);
        s->frc_alt = dev_alloc<double>(3 * n);
        s->mass_alt = dev_alloc<double>(n);
        s->id_alt = dev_alloc<uint64_t>(n);
        s->scalars = dev_alloc<double>(16);
        s->red_blocks = dev_alloc<double>(RED_BLOCKS * RED_WIDTH_MAX);
        s->red_ticket = dev_alloc<unsigned>(1);
        hipStream_t st = ctx->stream;
        MD_HIP(hipMemsetAsync(s->red_ticket, 0, sizeof(unsigned), st));
        MD_HIP(hipMemsetAsync(s->pos, 0, sizeof(double4) * (n ? n : 1), st));
        MD_HIP(hipMemsetAsync(s->vel, 0, sizeof(double) * 3 * (n ? n : 1), st));
        MD_HIP(hipMemsetAsync(s->frc, 0, sizeof(double) * 3 * (n ? n : 1), st));
        MD_HIP(hipMemsetAsync(s->scalars, 0, sizeof(double) * 16, st));
        if (n) k_init_state<<<blocks_for(n, 256), 256, 0, st>>>(n, s->mass, s->id);
        MD_HIP(hipGetLastError());
        // the reduction rows every kernel of this library may need for n particles (six sums per 256-particle
        // block or 128-particle tile at most), so that later calls (md_run_static among them) allocate none
        state_reserve_partials(s, std::max<size_t>(blocks_for(n ? n : 1, 256), tile_count(n ? n : 1)) * 6 + 8);
        MD_HIP(hipStreamSynchronize(st));
    } catch (...) {
        md_state_destroy(s);
        throw;
    }
    *out = s;
    MD_API_END
}

int md_state_destroy(md_state* s) {
    MD_API_BEGIN
    if (!s) return MD_OK;
    if (s->ctx) {
        (void)hipSetDevice(s->ctx->device);
        (void)hipStreamSynchronize(s->ctx->stream);
        --s->ctx->live;
    }
    dev_free(s->pos);
    dev_free(s->vel);
    dev_free(s->frc);
    dev_free(s->mass);
    dev_free(s->id);
    dev_free(s->pos_alt);
    dev_free(s->pos_spare);
    dev_free(s->vel_alt);
    dev_free(s->frc_alt);
    dev_free(s->mass_alt);
    dev_free(s->id_alt);
    dev_free(s->partials);
    dev_free(s->scalars);
    dev_free(s->red_blocks);
    dev_free(s->red_ticket);
    delete s;
    MD_API_END
}

int md_state_copy(md_state* dst, const md_state* src) {
    MD_API_BEGIN
    if (!dst || !src) fail(MD_ERR_ARG, "state is NULL");
    if (dst->n != src->n) fail(MD_ERR_ARG, "md_state_copy needs equal particle counts");
    hipStream_t st = dst->ctx->stream;
    const size_t n = src->n ? src->n : 1;
    if (src->ctx->stream != st) MD_HIP(hipStreamSynchronize(src->ctx->stream));
    MD_HIP(hipMemcpyAsync(dst->pos, src->pos, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipMemcpyAsync(dst->vel, src->vel, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipMemcpyAsync(dst->frc, src->frc, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipMemcpyAsync(dst->mass, src->mass, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    MD_HIP(hipMemcpyAsync(dst->id, src->id, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
    dst->box = src->box;
    dst->inv_box = src->inv_box;
    MD_HIP(hipStreamSynchronize(st));
    MD_API_END
}

int md_state_info(const md_state* s, size_t* n, double* box) {
    MD_API_BEGIN
    if (!s) fail(MD_ERR_ARG, "state is NULL");
    if (n) *n = s->n;
    if (box) *box = s->box;
    MD_API_END
}

static size_t field_bytes(const md_state* s, int field) {
    switch (field) {
        case MD_FIELD_POS: return sizeof(double4) * s->n;
        case MD_FIELD_VEL:
        case MD_FIELD_FORCE: return sizeof(double) * 3 * s->n;
        case MD_FIELD_MASS: return sizeof(double) * s->n;
        case MD_FIELD_ID: return sizeof(uint64_t) * s->n;
        default: fail(MD_ERR_ARG, "unknown field");
    }
}

static void* field_ptr(const md_state* s, int field) {
    switch (field) {
        case MD_FIELD_POS: return s->pos;
        case MD_FIELD_VEL: return s->vel;
        case MD_FIELD_FORCE: return s->frc;
        case MD_FIELD_MASS: return s->mass;
        case MD_FIELD_ID: return s->id;
        default: fail(MD_ERR_ARG, "unknown field");
    }
}

int md_state_upload(md_state* s, int field, const void* host) {
    MD_API_BEGIN
    if (!s || !host) fail(MD_ERR_ARG, "state/host is NULL");
    const size_t bytes = field_bytes(s, field);
    if (bytes == 0) return MD_OK;
    MD_HIP(hipMemcpyAsync(field_ptr(s, field), host, bytes, hipMemcpyHostToDevice, s->ctx->stream));
    MD_HIP(hipStreamSynchronize(s->ctx->stream));
    MD_API_END
}

int md_state_download(const md_state* s, int field, void* host) {
    MD_API_BEGIN
    if (!s || !host) fail(MD_ERR_ARG, "state/host is NULL");
    const size_t bytes = field_bytes(s, field);
    if (bytes == 0) return MD_OK;
    MD_HIP(hipMemcpyAsync(host, field_ptr(s, field), bytes, hipMemcpyDeviceToHost, s->ctx->stream));
    MD_HIP(hipStreamSynchronize(s->ctx->stream));
    MD_API_END
}

int md_state_reorder(md_state* s, const uint64_t* order) {
    MD_API_BEGIN
    if (!s || (!order && s->n)) fail(MD_ERR_ARG, "state/order is NULL");
    if (s->n == 0) return MD_OK;
    std::vector<uint32_t> o(s->n);
    std::vector<char> seen(s->n, 0);
    for (size_t k = 0; k < s->n; ++k) {
        if (order[k] >= s->n || seen[order[k]]) fail(MD_ERR_ARG, "order is not a permutation");
        seen[order[k]] = 1;
        o[k] = static_cast<uint32_t>(order[k]);
    }
    hipStream_t st = s->ctx->stream;
    uint32_t* d = dev_alloc<uint32_t>(s->n);
    try {
        MD_HIP(hipMemcpyAsync(d, o.data(), sizeof(uint32_t) * s->n, hipMemcpyHostToDevice, st));
        k_reorder<<<blocks_for(s->n, 256), 256, 0, st>>>(d, s->n, s->pos, s->vel, s->frc, s->mass, s->id,
                                                         s->pos_alt, s->vel_alt, s->frc_alt, s->mass_alt, s->id_alt);
        MD_HIP(hipGetLastError());
        MD_HIP(hipStreamSynchronize(st));
    } catch (...) {
        dev_free(d);
        throw;
    }
    dev_free(d);
    swap_alt(s);
    MD_API_END
}

int md_kinetic_energy(md_state* s, double* ke) {
    MD_API_BEGIN
    if (!s || !ke) fail(MD_ERR_ARG, "state/ke is NULL");
    hipStream_t st = s->ctx->stream;
    kinetic_exact_enqueue(s, s->scalars + 3);
    MD_HIP(hipMemcpyAsync(s->ctx->host_scalars, s->scalars + 3, sizeof(double), hipMemcpyDeviceToHost, st));
    MD_HIP(hipStreamSynchronize(st));
    *ke = s->ctx->host_scalars[0];
    MD_API_END
}

int md_scale_velocities(md_state* s, double scale) {
    MD_API_BEGIN
    if (!s) fail(MD_ERR_ARG, "state is NULL");
    if (s->n == 0) return MD_OK;
    k_scale<<<blocks_for(3 * s->n, 256), 256, 0, s->ctx->stream>>>(s->vel, 3 * s->n, scale, nullptr);
    MD_HIP(hipGetLastError());
    MD_API_END
}

int md_state_hash(md_state* s, uint64_t* hash) {
    MD_API_BEGIN
    if (!s || !hash) fail(MD_ERR_ARG, "state/hash is NULL");
    std::vector<double> pos(4 * s->n), vel(3 * s->n);
    hipStream_t st = s->ctx->stream;
    if (s->n) {
        MD_HIP(hipMemcpyAsync(pos.data(), s->pos, sizeof(double4) * s->n, hipMemcpyDeviceToHost, st));
        MD_HIP(hipMemcpyAsync(vel.data(), s->vel, sizeof(double) * 3 * s->n, hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
    }
    // 64-bit FNV-1a over the bytes of x, vx, y, vy, z, vz of each particle (System.cpp:85-105)
    uint64_t h = 0xcbf29ce484222325ull;
    auto feed = [&h](double x) {
        uint64_t bits;
        std::memcpy(&bits, &x, 8);
        for (int b = 0; b < 8; ++b) h = (h ^ ((bits >> (8 * b)) & 0xffull)) * 0x100000001b3ull;
    };
    for (size_t i = 0; i < s->n; ++i)
        for (int k = 0; k < 3; ++k) {
            feed(pos[4 * i + k]);
            feed(vel[k * s->n + i]);
        }
    *hash = h;
    MD_API_END
}

int md_bussi_apply(md_state* s, double temperature, double tau, double dt, double r1, double sum_sq) {
    MD_API_BEGIN
    if (!s) fail(MD_ERR_ARG, "state is NULL");
    double ke = 0.0;
    int rc = md_kinetic_energy(s, &ke);
    if (rc != MD_OK) return rc;
    if (ke <= 0.0) return MD_OK;
    // eq. A7 of Bussi, Donadio & Parrinello (2007), in the reference's operation order (Thermostat.cpp:41-57)
    const double dof = 3.0 * static_cast<double>(s->n) - 3.0;
    const double kbar = 0.5 * dof * temperature;
    const double c = std::exp(-dt / tau);
    const double relax = (1.0 - c) * (kbar * (sum_sq + r1 * r1) / dof - ke);
    const double noise = 2.0 * r1 * std::sqrt(kbar / dof * ke * (1.0 - c) * c);
    const double ke_new = ke + relax + noise;
    return md_scale_velocities(s, std::sqrt(ke_new / ke));
    MD_API_END
}

struct md_rng {
    std::mt19937_64 engine;
};

int md_bussi_rng_create(uint64_t seed, md_rng** out) {
    MD_API_BEGIN
    if (!out) fail(MD_ERR_ARG, "out is NULL");
    *out = new md_rng{std::mt19937_64(seed)};
    MD_API_END
}

int md_bussi_rng_destroy(md_rng* r) {
    delete r;
    return MD_OK;
}

int md_bussi_rng_draws(md_rng* r, size_t n_particles, int count, double* out) {
    MD_API_BEGIN
    if (!r || (!out && count > 0)) fail(MD_ERR_ARG, "rng/out is NULL");
    const double dof = 3.0 * static_cast<double>(n_particles) - 3.0;
    for (int k = 0; k < count; ++k) {
        std::normal_distribution<double> gaussian(0.0, 1.0);
        std::gamma_distribution<double> chi2(0.5 * (dof - 1.0), 2.0);
        out[2 * k] = gaussian(r->engine);
        out[2 * k + 1] = chi2(r->engine);
    }
    MD_API_END
}

}  // extern "C"
