// Trajectory frames downloaded asynchronously in particle-id order (the data writeXyzFrame prints,
// Trajectory.cpp:4-25). A capture is two operations on two streams:
//   1. on the context's stream, after everything already enqueued there: k_frame_scatter writes each particle's
//      position and velocity to row id[i] of a device staging frame (the inverse of the cell re-sort permutation,
//      Trajectory.cpp:9-11), 48 bytes per particle, x y z vx vy vz;
//   2. on the frames' own copy stream, once the scatter is done: one device-to-host copy of the staging frame
//      into pinned host memory.
// Later steps enqueued on the context's stream overlap the copy: they change the state, not the staging frame.
// Slots form a ring so that several frames can be in flight; a slot is reused only after its holder released it.
// md_frames_wait and md_frames_release may run on one other thread than md_frames_capture (a writer thread): the
// slot states they share are atomic, and the HIP calls involved (event synchronisation) are thread-safe.
#include <atomic>
#include <cstring>
#include <memory>

#include "internal.hpp"

struct md_frames {
    md_ctx* ctx = nullptr;
    size_t capacity = 0;
    hipStream_t copy = nullptr;
    struct Slot {
        double* staging = nullptr;  // device: 6 * capacity doubles, then one int word (an id >= n was seen)
        double* host = nullptr;     // pinned: the same layout
        hipEvent_t scattered = nullptr, copied = nullptr;
        size_t n = 0;
        double box = 0.0;
        std::atomic<int> state{0};  // 0 free, 1 captured (copy in flight or landed), 2 waited on
    };
    std::unique_ptr<Slot[]> slots;
    int count = 0;
    int next = 0;
};

namespace mdgpu {
namespace {

constexpr int FRAME_THREADS = 256;

// Row id[i] of the frame gets particle i. Ids are the System's: a permutation of 0..n-1; an id outside it is
// reported through the flag word (the row is not written).
__global__ __launch_bounds__(FRAME_THREADS) void k_frame_scatter(const double4* __restrict__ pos,
                                                                 const double* __restrict__ vel,
                                                                 const uint64_t* __restrict__ id, size_t n,
                                                                 double* __restrict__ frame, int* __restrict__ bad) {
    const size_t i = static_cast<size_t>(blockIdx.x) * FRAME_THREADS + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = id[i];
    if (k >= n) {
        *bad = 1;
        return;
    }
    const double4 r = pos[i];
    double* row = frame + 6 * k;
    row[0] = r.x;
    row[1] = r.y;
    row[2] = r.z;
    row[3] = vel[i];
    row[4] = vel[n + i];
    row[5] = vel[2 * n + i];
}

__global__ void k_frame_clear(int* bad) { *bad = 0; }

size_t frame_bytes(size_t capacity) { return 6 * capacity * sizeof(double) + sizeof(double); }

void release_slots(md_frames* f) {
    for (int k = 0; k < f->count; ++k) {
        auto& s = f->slots[k];
        if (s.scattered) (void)hipEventDestroy(s.scattered);
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.staging) (void)hipFree(s.staging);
        if (s.host) (void)hipHostFree(s.host);
        s.scattered = s.copied = nullptr;
        s.staging = s.host = nullptr;
    }
}

}  // namespace
}  // namespace mdgpu

using namespace mdgpu;

extern "C" {

int md_frames_create(md_ctx* ctx, size_t capacity, int slots, md_frames** out) {
    MD_API_BEGIN
    if (!ctx || !out) fail(MD_ERR_ARG, "ctx/out is NULL");
    if (slots < 1 || slots > 64) fail(MD_ERR_ARG, "slots must be in 1..64");
    MD_HIP(hipSetDevice(ctx->device));
    auto* f = new md_frames;
    try {
        f->ctx = ctx;
        f->capacity = capacity;
        MD_HIP(hipStreamCreateWithFlags(&f->copy, hipStreamNonBlocking));
        f->slots = std::make_unique<md_frames::Slot[]>(slots);
        f->count = slots;
        for (int k = 0; k < slots; ++k) {
            auto& s = f->slots[k];
            MD_HIP(hipMalloc(reinterpret_cast<void**>(&s.staging), frame_bytes(capacity)));
            MD_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.host), frame_bytes(capacity), hipHostMallocDefault));
            MD_HIP(hipEventCreateWithFlags(&s.scattered, hipEventDisableTiming));
            MD_HIP(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
        }
        ++ctx->live;
    } catch (...) {
        if (f->copy) (void)hipStreamDestroy(f->copy);
        release_slots(f);
        delete f;
        throw;
    }
    *out = f;
    MD_API_END
}

int md_frames_destroy(md_frames* f) {
    MD_API_BEGIN
    if (!f) return MD_OK;
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    (void)hipStreamSynchronize(f->copy);
    (void)hipStreamDestroy(f->copy);
    release_slots(f);
    --f->ctx->live;
    delete f;
    MD_API_END
}

int md_frames_capture(md_frames* f, md_state* s, int* slot) {
    MD_API_BEGIN
    if (!f || !s || !slot) fail(MD_ERR_ARG, "frames/state/slot is NULL");
    if (s->ctx != f->ctx) fail(MD_ERR_ARG, "state and frames belong to different contexts");
    if (s->n > f->capacity) fail(MD_ERR_ARG, "state has more particles than the frames hold");
    auto& sl = f->slots[f->next];
    if (sl.state.load(std::memory_order_acquire) != 0)
        fail(MD_ERR_STATE, "every frame slot is in use; release one first");
    MD_HIP(hipSetDevice(f->ctx->device));
    hipStream_t st = f->ctx->stream;
    int* bad = reinterpret_cast<int*>(sl.staging + 6 * f->capacity);
    k_frame_clear<<<1, 1, 0, st>>>(bad);
    if (s->n) {
        k_frame_scatter<<<blocks_for(s->n, FRAME_THREADS), FRAME_THREADS, 0, st>>>(s->pos, s->vel, s->id, s->n,
                                                                                  sl.staging, bad);
        MD_HIP(hipGetLastError());
    }
    MD_HIP(hipEventRecord(sl.scattered, st));
    MD_HIP(hipStreamWaitEvent(f->copy, sl.scattered, 0));
    MD_HIP(hipMemcpyAsync(sl.host, sl.staging, 6 * s->n * sizeof(double), hipMemcpyDeviceToHost, f->copy));
    MD_HIP(hipMemcpyAsync(sl.host + 6 * f->capacity, bad, sizeof(int), hipMemcpyDeviceToHost, f->copy));
    MD_HIP(hipEventRecord(sl.copied, f->copy));
    sl.n = s->n;
    sl.box = s->box;
    sl.state.store(1, std::memory_order_release);
    *slot = f->next;
    f->next = (f->next + 1) % f->count;
    MD_API_END
}

int md_frames_wait(md_frames* f, int slot, const double** frame, size_t* n, double* box) {
    MD_API_BEGIN
    if (!f || !frame) fail(MD_ERR_ARG, "frames/frame is NULL");
    if (slot < 0 || slot >= f->count) fail(MD_ERR_ARG, "no such slot");
    auto& sl = f->slots[slot];
    if (sl.state.load(std::memory_order_acquire) == 0) fail(MD_ERR_STATE, "slot holds no captured frame");
    MD_HIP(hipSetDevice(f->ctx->device));
    MD_HIP(hipEventSynchronize(sl.copied));
    sl.state.store(2, std::memory_order_relaxed);
    int bad = 0;
    std::memcpy(&bad, sl.host + 6 * f->capacity, sizeof(int));
    if (bad) fail(MD_ERR_STATE, "particle ids are not a permutation of 0..n-1");
    *frame = sl.host;
    if (n) *n = sl.n;
    if (box) *box = sl.box;
    MD_API_END
}

int md_frames_release(md_frames* f, int slot) {
    MD_API_BEGIN
    if (!f) fail(MD_ERR_ARG, "frames is NULL");
    if (slot < 0 || slot >= f->count) fail(MD_ERR_ARG, "no such slot");
    auto& sl = f->slots[slot];
    if (sl.state.load(std::memory_order_acquire) == 1) {
        // released without a wait: the copy may still be writing the slot
        MD_HIP(hipSetDevice(f->ctx->device));
        MD_HIP(hipEventSynchronize(sl.copied));
    }
    sl.state.store(0, std::memory_order_release);
    MD_API_END
}

}  // extern "C"
