// Device helpers shared by the mdgpu kernels. The library is compiled with -ffp-contract=off, so every double
// expression here rounds once per operation in source order — the same arithmetic as the reference's own
// expressions (System.hpp:86-112, Forcefield.cpp:25-36), which keeps cell ids, list membership, integrator
// updates and FP64 forces bit-equal to the CPU results.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mdgpu {

constexpr int WAVE = 64;

// Minimum image of one separation component: d - L * round(d / L), round half away from zero (System.hpp:94).
// round(x) is formed as trunc(x + copysign(0.5 - 2^-54, x)), two FP64 operations instead of the library's four:
// the addition rounds to nearest even, so x + (0.5 - 2^-54) reaches the next integer exactly when x's fraction is
// at least 1/2 (at a tie the exact sum lies half an ulp below the integer and rounds up to it; below a tie it stays
// under it), which is round()'s result for every double, ties, infinities and signed zeros included.
__device__ __forceinline__ double min_image(double d, double box, double inv_box) {
    const double x = d * inv_box;
    const double images = trunc(x + copysign(0.49999999999999994, x));
    return d - box * images;
}

// 1 / r2 as the compiler's IEEE double division computes it (v_div_scale, v_rcp_f64, two Newton steps, the
// residual step through v_div_fmas, v_div_fixup), without its scaling and special-case steps, which change nothing
// for 2^-767 <= r2 < 2^1022: the same correctly rounded quotient, in seven instructions instead of eleven. The LJ
// kernels use the value only for r2 < rc^2; below 2^-767 (|r| < 1e-115) the LJ powers overflow and the force is
// NaN either way.
__device__ __forceinline__ double recip_rn(double r2) {
    double y = __builtin_amdgcn_rcp(r2);
    double e = fma(-r2, y, 1.0);
    y = fma(y, e, y);
    e = fma(-r2, y, 1.0);
    y = fma(y, e, y);
    e = fma(-r2, y, 1.0);
    return fma(e, y, y);
}

// x - L * floor(x / L) (System.hpp:110-112).
__device__ __forceinline__ double wrap_coord(double x, double box, double inv_box) {
    const double images = floor(x * inv_box);
    return x - box * images;
}

// Cell coordinate along one axis: min(int(x * side / L), side - 1) (CellList.cpp:30-33). Coordinates below the
// box (never produced by the wrapped dynamics; the reference would index before its arrays) are clamped to 0.
__device__ __forceinline__ int cell_axis(double x, double per_length, int side) {
    const int c = static_cast<int>(x * per_length);
    return c < 0 ? 0 : (c < side - 1 ? c : side - 1);
}

// Linear cell id (cx * side + cy) * side + cz (CellList.cpp:36).
__device__ __forceinline__ int cell_of_position(double x, double y, double z, double per_length, int side) {
    return (cell_axis(x, per_length, side) * side + cell_axis(y, per_length, side)) * side +
           cell_axis(z, per_length, side);
}

__device__ __forceinline__ int wrap_index(int c, int side) {
    return c < 0 ? c + side : (c >= side ? c - side : c);
}

// gfx950 deals the workgroups of a launch to its 8 XCDs round-robin (block b runs on XCD b % 8), and each XCD has
// its own L2. Mapping block b to item (b % 8) * (items / 8) + b / 8 gives every XCD one contiguous eighth of the
// items, so neighbouring tiles (which share most of their halo) meet in the same L2. Items past the largest
// multiple of 8 map to themselves.
constexpr int XCDS = 8;
__device__ __forceinline__ int xcd_item(int b, int items) {
    const int full = items & ~(XCDS - 1);
    if (b >= full) return b;
    return (b % XCDS) * (full / XCDS) + b / XCDS;
}

// Wavefront sum by butterfly shuffles; every lane returns the same value, combined in a fixed order.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// DPP quad permutes: every lane reads lane ^ 1, or lane ^ 2, of its quad
constexpr int QUAD_XOR1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int QUAD_XOR2 = 0x4E;  // quad_perm [2,3,0,1]

// Lanes' values moved by one DPP control (CTRL) within the rows of ROW_MASK; lanes not written read 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Wavefront sum through DPP moves only (no LDS round trips): quads, half rows and rows by permutes and mirrors,
// then row 15/31 broadcasts into lane 63, read back as a uniform value.
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_f64<0xB1>(v);         // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);         // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);        // row_half_mirror
    v += dpp_f64<0x140>(v);        // row_mirror: every lane holds its row's sum
    v += dpp_f64<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3: lane 63 holds the total
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), WAVE - 1);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), WAVE - 1);
    return __hiloint2double(hi, lo);
}

// Wavefront maximum of non-negative values through the same DPP sequence (unwritten lanes read 0), uniform.
__device__ __forceinline__ int wave_max_dpp(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(v, WAVE - 1);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Block-wide sums of `W` values per thread into out[0..W) for thread 0 (fixed order: waves, then wave 0..).
template <int W, int THREADS>
__device__ __forceinline__ void block_sum(double (&v)[W], double* out) {
    __shared__ double lds[THREADS / WAVE][W];
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < W; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < W; ++k) lds[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            double s = 0.0;
            for (int w = 0; w < THREADS / WAVE; ++w) s += lds[w][k];
            out[k] = s;
        }
    }
}

// Runs of equal cell ids among a wavefront's consecutive particles (the state is still nearly sorted by cell from
// the last build): lane `head` marks the first lane of a run, `len` (valid in head lanes) its length, `first` the
// run's first lane for every lane. Lanes past the end (valid = false) belong to no run.
struct CellRun {
    bool head;
    int len, first;
};

__device__ __forceinline__ CellRun cell_run(int c, bool valid) {
    const int lane = threadIdx.x % WAVE;
    const int prev = __shfl_up(c, 1, WAVE);
    const bool head = valid && (lane == 0 || prev != c);
    const unsigned long long heads = __ballot(head);
    const int live = __popcll(__ballot(valid));  // valid lanes are a prefix of the wavefront
    CellRun r;
    r.head = head;
    const unsigned long long later = lane == WAVE - 1 ? 0ull : heads & (~0ull << (lane + 1));
    r.len = (later ? __ffsll(static_cast<long long>(later)) - 1 : live) - lane;
    const unsigned long long upto = lane == WAVE - 1 ? heads : heads & ((2ull << lane) - 1ull);
    r.first = upto ? WAVE - 1 - __clzll(static_cast<long long>(upto)) : 0;
    return r;
}

// Per-block record of a list build: (longest row seen, sum of the block's half-row lengths). A block covers at
// most 512 rows of at most `padding` entries each, and padding x n int32 rows fit in device memory, so the sum
// fits in 32 bits.
template <int THREADS>
__device__ __forceinline__ void block_build_stat(int mx, unsigned hs, int2* out) {
    __shared__ int2 lds[THREADS / WAVE];
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(mx, off, WAVE);
        mx = o > mx ? o : mx;
        hs += __shfl_xor(hs, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0) lds[threadIdx.x / WAVE] = make_int2(mx, static_cast<int>(hs));
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
        unsigned h = 0;
        for (int w = 0; w < THREADS / WAVE; ++w) {
            m = lds[w].x > m ? lds[w].x : m;
            h += static_cast<unsigned>(lds[w].y);
        }
        *out = make_int2(m, static_cast<int>(h));
    }
}

}  // namespace mdgpu
