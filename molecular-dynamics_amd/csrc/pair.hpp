// Lennard-Jones pair terms of the gather kernels (reference: Forcefield.cpp:214-326 for FP64, 352-446 for the
// mixed-precision variant), shared by the single-system and the batched force kernels so both sum exactly the
// same rounded terms.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"
#include "mdgpu.h"

namespace mdgpu {

struct LJConst {
    double sig2, c24, c4, shift, rc2;
    float sig2f, c24f, c4f, shiftf, rc2f;
};

// Constructor constants (Forcefield.cpp:8-15): shift = V(rc) with std::pow, as the reference computes it.
LJConst make_lj(const md_lj& p);

// Per-particle sums of one row: force, and over the row's j > i part the potential, virial and inside count.
struct PairSums {
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double u = 0.0, w = 0.0, c = 0.0;
};

// FP64 term: minimum image, r^2, LJ force F/r * d rounded once per component, added in row order.
__device__ __forceinline__ void fp64_pair(const double4& ri, const double4& rj, double box, double inv_box,
                                          const LJConst& lj, bool count_energy, PairSums& a) {
    const double dx = min_image(ri.x - rj.x, box, inv_box);
    const double dy = min_image(ri.y - rj.y, box, inv_box);
    const double dz = min_image(ri.z - rj.z, box, inv_box);
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double inv_r2 = recip_rn(r2);
    const double s2 = lj.sig2 * inv_r2;
    const double s6 = s2 * s2 * s2;
    const double s12 = s6 * s6;
    const bool inside = r2 < lj.rc2;
    const double f = inside ? lj.c24 * (2.0 * s12 - s6) * inv_r2 : 0.0;
    a.fx += f * dx;
    a.fy += f * dy;
    a.fz += f * dz;
    if (count_energy) {
        a.u += inside ? lj.c4 * (s12 - s6) - lj.shift : 0.0;
        a.w += f * r2;
        a.c += inside ? 1.0 : 0.0;
    }
}

// Mixed term from a double separation d (already the minimum image): r^2 in double rounded once to float, the
// LJ factor F/r in float (Newton-refined hardware reciprocal), then F/r * d accumulated in double. Rounding r^2
// once keeps the relative error of the steep r^-14 force at a few float ulps; the reference's float path
// (r^2 from the narrowed d) can lose about twice that. The cut-off test is taken on the double r^2, so exactly
// the FP64 path's pairs interact: the shifted potential's force jumps by 24 eps (2 (s/rc)^12 - (s/rc)^6) / rc
// (0.039 at rc = 2.5) at the cut-off, and a test on the float r^2 (the reference's, Forcefield.cpp:395-402)
// moves the pairs within float rounding of rc^2 across it (about ten of them at 4M atoms).
__device__ __forceinline__ void mixed_pair_term(double dx, double dy, double dz, const LJConst& lj,
                                                bool count_energy, PairSums& a) {
    const double r2d = fma(dx, dx, fma(dy, dy, dz * dz));
    const float r2 = static_cast<float>(r2d);
    const float q = __builtin_amdgcn_rcpf(r2);
    const float inv_r2 = fmaf(q, fmaf(-r2, q, 1.0f), q);
    const float s2 = lj.sig2f * inv_r2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    const bool inside = r2d < lj.rc2;
    const float ff = inside ? lj.c24f * fmaf(2.0f, s12, -s6) * inv_r2 : 0.0f;
    const double f = static_cast<double>(ff);
    a.fx = fma(f, dx, a.fx);
    a.fy = fma(f, dy, a.fy);
    a.fz = fma(f, dz, a.fz);
    if (count_energy) {
        const float up = inside ? fmaf(lj.c4f, s12 - s6, -lj.shiftf) : 0.0f;
        a.u += static_cast<double>(up);
        a.w = fma(f, r2d, a.w);
        a.c += inside ? 1.0 : 0.0;
    }
}

// Per-lane sums of the tiled mixed kernel over a particle's full row, with the constant factors left for the end:
// g = sum t inv_r2 d (in double), T = sum t and S = sum s12 (in float), where t = 2 s12 - s6, and the inside count
// as an integer. Then F = 24 eps g, W = sum f r^2 = 24 eps T and U = 4 eps (T - S) - shift * count. The reference
// sums energy_i, virial_i and inside_i in float over the half row j > i (Forcefield.cpp:397-401) and adds them in
// double across particles; here every pair is met from both of its ends, so the caller halves the tile's U, W and
// count: the same sums, with the float rounding taken in another order.
struct PairSumsF {
    double gx = 0.0, gy = 0.0, gz = 0.0;
    float t = 0.0f, s = 0.0f;
    int c = 0;
};

// One pair of a full row, without its inside count (returned: the caller counts it, per lane or per wavefront).
// Outside rc inv_r2 is zero, which makes the force and the energy terms zero; the cut-off test is on the double r^2,
// as in mixed_pair_term. UNIT (sigma = 1): s2 = inv_r2, the same value as sig2f * inv_r2 with sig2f = 1.
template <bool UNIT = false>
__device__ __forceinline__ bool mixed_pair_core(double dx, double dy, double dz, const LJConst& lj, PairSumsF& a) {
    const double r2d = fma(dx, dx, fma(dy, dy, dz * dz));
    const float r2 = static_cast<float>(r2d);
    const float q = __builtin_amdgcn_rcpf(r2);
    const bool inside = r2d < lj.rc2;
    const float inv_r2 = inside ? fmaf(q, fmaf(-r2, q, 1.0f), q) : 0.0f;
    const float s2 = UNIT ? inv_r2 : lj.sig2f * inv_r2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    const float t = fmaf(2.0f, s12, -s6);
    const double g = static_cast<double>(t * inv_r2);
    a.gx = fma(g, dx, a.gx);
    a.gy = fma(g, dy, a.gy);
    a.gz = fma(g, dz, a.gz);
    a.t += t;
    a.s += s12;
    return inside;
}

template <bool UNIT = false>
__device__ __forceinline__ void mixed_pair_term_f(double dx, double dy, double dz, const LJConst& lj, PairSumsF& a) {
    a.c += mixed_pair_core<UNIT>(dx, dy, dz, lj, a) ? 1 : 0;
}

// Mixed term of the gather kernels: double minimum image, then mixed_pair_term.
__device__ __forceinline__ void mixed_pair_gather(const double4& ri, const double4& rj, double box, double inv_box,
                                                  const LJConst& lj, bool count_energy, PairSums& a) {
    double ex = ri.x - rj.x, ey = ri.y - rj.y, ez = ri.z - rj.z;
    ex = fma(-box, rint(ex * inv_box), ex);
    ey = fma(-box, rint(ey * inv_box), ey);
    ez = fma(-box, rint(ez * inv_box), ez);
    mixed_pair_term(ex, ey, ez, lj, count_energy, a);
}

template <bool UNIT = false>
__device__ __forceinline__ void mixed_pair_gather_f(const double4& ri, const double4& rj, double box, double inv_box,
                                                    const LJConst& lj, PairSumsF& a) {
    double ex = ri.x - rj.x, ey = ri.y - rj.y, ez = ri.z - rj.z;
    ex = fma(-box, rint(ex * inv_box), ex);
    ey = fma(-box, rint(ey * inv_box), ey);
    ez = fma(-box, rint(ez * inv_box), ez);
    mixed_pair_term_f<UNIT>(ex, ey, ez, lj, a);
}

}  // namespace mdgpu
