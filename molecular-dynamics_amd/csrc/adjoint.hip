// Adjoint of K velocity-Verlet steps (no counterpart in the reference): gradients of
//   L = (1/K) sum_k U(r_k) + (Pbar - 1)^2,   Pbar = (1/K) sum_k P_k,   P_k = (2 KE_k + W_k) / (3V),
// k = 1..K, with respect to epsilon, sigma, the initial positions and the initial velocities.
//
// Forward: FP64 steps over a fixed-order list (no re-sort, so indices stay put), with (r, v, f) checkpoints every M
// steps and a list rebuild at every checkpoint so a segment can be recomputed bit for bit. Reverse: for each
// segment from the last, recompute its points from the checkpoint, then sweep its points backwards. Point k gets
// one pair kernel at r_k, with w_k = lv_k / m + P_k:
//   (H w)_i = -sum_j [ f(s) w_ij + 2 f'(s) (d . w_ij) d ],   f = F/r = 24 eps (2 q12 - q6) / s,  s = |d|^2
//   grad_i U = -F_i,   grad_i W = sum_j 2 (f'(s) s + f(s)) d
// plus the epsilon / sigma partials of U, W and sum_i (dF_i/dtheta) . w_i. The element-wise update then folds the
// two half-kicks that meet at r_k (their Hessian products are merged into the single w_k):
//   lr += cU gU + cW gW - (dt/2) H w;   lvh = lv + dt lr;   P = lvh / m;   lv = lvh + cKE m v_{k-1}
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace mdgpu {

constexpr int ADJ_THREADS = 256;
constexpr int ADJ_PARTS = 6;  // dU/deps, dU/dsig, dW/deps, dW/dsig, dFw/deps, dFw/dsig

struct AdjLJ {
    double eps, sigma, sig2, rc2, shift1, dshift_sig;
};

__global__ void __launch_bounds__(ADJ_THREADS)
    k_adj_pair(const double4* __restrict__ pos, size_t n, const int32_t* __restrict__ rows,
               const int32_t* __restrict__ count, const int32_t* __restrict__ half, double box, double inv_box, AdjLJ lj,
               const double* __restrict__ w, double* __restrict__ hw, double* __restrict__ gu, double* __restrict__ gw,
               double* __restrict__ partials) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double acc[ADJ_PARTS] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double4 ri = pos[i];
        const double wix = w[i], wiy = w[n + i], wiz = w[2 * n + i];
        double h[3] = {0, 0, 0}, g[3] = {0, 0, 0}, v[3] = {0, 0, 0};
        const int m = count[i], hcount = half[i];
        const double inv_sigma = 1.0 / lj.sigma;
        for (int k = 0; k < m; ++k) {
            const int32_t j = rows[static_cast<size_t>(k) * n + i];
            const double4 rj = pos[j];
            const double dx = min_image(ri.x - rj.x, box, inv_box);
            const double dy = min_image(ri.y - rj.y, box, inv_box);
            const double dz = min_image(ri.z - rj.z, box, inv_box);
            const double s = dx * dx + dy * dy + dz * dz;
            if (!(s < lj.rc2)) continue;
            const double inv = 1.0 / s;
            const double q2 = lj.sig2 * inv;
            const double q6 = q2 * q2 * q2;
            const double q12 = q6 * q6;
            const double f1 = 24.0 * (2.0 * q12 - q6) * inv;          // f / eps
            const double f = lj.eps * f1;
            const double fp = lj.eps * 24.0 * (-14.0 * q12 + 4.0 * q6) * inv * inv;  // df/ds
            const double wx = wix - w[j], wy = wiy - w[n + j], wz = wiz - w[2 * n + j];
            const double dw = dx * wx + dy * wy + dz * wz;
            h[0] -= f * wx + 2.0 * fp * dw * dx;
            h[1] -= f * wy + 2.0 * fp * dw * dy;
            h[2] -= f * wz + 2.0 * fp * dw * dz;
            g[0] -= f * dx;
            g[1] -= f * dy;
            g[2] -= f * dz;
            const double gwf = 2.0 * (fp * s + f);
            v[0] += gwf * dx;
            v[1] += gwf * dy;
            v[2] += gwf * dz;
            const double dfs = lj.eps * 24.0 * (24.0 * q12 - 6.0 * q6) * inv * inv_sigma;  // df/dsigma
            const double dwi = dx * wix + dy * wiy + dz * wiz;
            acc[4] += f1 * dwi;
            acc[5] += dfs * dwi;
            if (k < hcount) {
                acc[0] += 4.0 * (q12 - q6) - lj.shift1;
                acc[1] += lj.eps * 4.0 * (12.0 * q12 - 6.0 * q6) * inv_sigma - lj.dshift_sig;
                acc[2] += f1 * s;
                acc[3] += dfs * s;
            }
        }
        for (int c = 0; c < 3; ++c) {
            hw[c * n + i] = h[c];
            gu[c * n + i] = g[c];
            gw[c * n + i] = v[c];
        }
    }
    block_sum<ADJ_PARTS, ADJ_THREADS>(acc, partials + static_cast<size_t>(blockIdx.x) * ADJ_PARTS);
}

// Element-wise adjoint update after the pair kernel at point k (k >= 1). `next_initial`: point k-1 is the initial
// state (no observables there, and only one half-kick meets it).
__global__ void k_adj_update(size_t n, double* __restrict__ lr, double* __restrict__ lv, double* __restrict__ pend,
                             double* __restrict__ w, const double* __restrict__ hw, const double* __restrict__ gu,
                             const double* __restrict__ gw, const double* __restrict__ mass,
                             const double* __restrict__ v_prev, double cu, double cw, double cke, double dt,
                             int next_initial) {
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    const size_t i = t % n;
    const double m = mass[i];
    const double r_tot = lr[t] + cu * gu[t] + cw * gw[t] - 0.5 * dt * hw[t];
    const double lvh = lv[t] + dt * r_tot;
    lr[t] = r_tot;
    const double p = lvh / m;
    pend[t] = p;
    if (next_initial) {
        lv[t] = lvh;
        w[t] = p;
    } else {
        const double lv_next = lvh + cke * m * v_prev[t];
        lv[t] = lv_next;
        w[t] = lv_next / m + p;
    }
}

// First point K: lv = cKE m v_K, w = cKE v_K, lr = P = 0.
__global__ void k_adj_start(size_t n, double* __restrict__ lr, double* __restrict__ lv, double* __restrict__ pend,
                            double* __restrict__ w, const double* __restrict__ mass, const double* __restrict__ v,
                            double cke) {
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    const double m = mass[t % n];
    lr[t] = 0.0;
    pend[t] = 0.0;
    lv[t] = cke * m * v[t];
    w[t] = cke * v[t];
}

// Point 0: lr -= (dt/2) H w.
__global__ void k_adj_finish(size_t n, double* __restrict__ lr, const double* __restrict__ hw, double dt) {
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < 3 * n) lr[t] = lr[t] - 0.5 * dt * hw[t];
}

// theta += cU dU + cW dW + (dt/2) dFw from the reduced point partials (one thread: a fixed order over points).
__global__ void k_adj_theta(const double* __restrict__ point, double* __restrict__ theta, double cu, double cw,
                            double dt) {
    theta[0] += cu * point[0] + cw * point[2] + 0.5 * dt * point[4];
    theta[1] += cu * point[1] + cw * point[3] + 0.5 * dt * point[5];
}

__global__ void k_copy_pos(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_copy_f64(const double* __restrict__ src, double* __restrict__ dst, size_t count) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i];
}

// obs[3k + 0..2] = U, W, KE of point k
__global__ void k_record_obs(const double* __restrict__ scalars, double* __restrict__ obs, int k) {
    obs[3 * k + 0] = scalars[0];
    obs[3 * k + 1] = scalars[1];
    obs[3 * k + 2] = scalars[3];
}

namespace {

struct Buf {
    std::vector<void*> ptrs;
    template <class T>
    T* get(size_t count) {
        T* p = dev_alloc<T>(count);
        ptrs.push_back(p);
        return p;
    }
    ~Buf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

struct WorkState {
    md_state* s = nullptr;
    ~WorkState() {
        if (s) md_state_destroy(s);
    }
};

}  // namespace

}  // namespace mdgpu

using namespace mdgpu;

extern "C" int md_adjoint_run(md_state* src, md_list* l, const md_adjoint_config* cfg, md_adjoint_result* out,
                              double* grad_positions, double* grad_velocities) {
    MD_API_BEGIN
    if (!src || !l || !cfg || !out) fail(MD_ERR_ARG, "state/list/config/result is NULL");
    if (cfg->steps < 1 || cfg->checkpoint_every < 1) fail(MD_ERR_ARG, "steps and checkpoint_every must be >= 1");
    check_cutoff(src, cfg->lj);
    md_ctx* ctx = src->ctx;
    hipStream_t st = ctx->stream;
    const size_t n = src->n;
    const int K = cfg->steps, M = cfg->checkpoint_every;
    const double dt = cfg->dt;
    const double volume = src->box * src->box * src->box;
    const unsigned b3 = blocks_for(3 * n, 256), bn = blocks_for(n, ADJ_THREADS);

    // on every exit the list describes the (destroyed) work state: force a rebuild on its next use
    struct ListGuard {
        md_list* l;
        ~ListGuard() { l->built = false; }
    } list_guard{l};
    WorkState ws;
    {
        int rc = md_state_create(ctx, n, src->box, &ws.s);
        if (rc != MD_OK) return rc;
        rc = md_state_copy(ws.s, src);
        if (rc != MD_OK) return rc;
    }
    md_state* s = ws.s;
    state_reserve_partials(s, static_cast<size_t>(bn) * ADJ_PARTS + 8);

    const int nck = (K + M - 1) / M;  // checkpoints at 0, M, ..., (nck-1) M
    Buf buf;
    double4* ck_pos = buf.get<double4>(static_cast<size_t>(nck) * n);
    double* ck_vel = buf.get<double>(static_cast<size_t>(nck) * 3 * n);
    double* ck_frc = buf.get<double>(static_cast<size_t>(nck) * 3 * n);
    double4* seg_pos = buf.get<double4>(static_cast<size_t>(M + 1) * n);
    double* seg_vel = buf.get<double>(static_cast<size_t>(M + 1) * 3 * n);
    double* lr = buf.get<double>(3 * n);
    double* lv = buf.get<double>(3 * n);
    double* pend = buf.get<double>(3 * n);
    double* w = buf.get<double>(3 * n);
    double* hw = buf.get<double>(3 * n);
    double* gu = buf.get<double>(3 * n);
    double* gw = buf.get<double>(3 * n);
    double* obs = buf.get<double>(3 * static_cast<size_t>(K + 1));
    double* point = buf.get<double>(ADJ_PARTS);
    double* theta = buf.get<double>(2);

    auto save_checkpoint = [&](int c) {
        k_copy_pos<<<bn, ADJ_THREADS, 0, st>>>(s->pos, ck_pos + static_cast<size_t>(c) * n, n);
        k_copy_f64<<<b3, 256, 0, st>>>(s->vel, ck_vel + static_cast<size_t>(c) * 3 * n, 3 * n);
        k_copy_f64<<<b3, 256, 0, st>>>(s->frc, ck_frc + static_cast<size_t>(c) * 3 * n, 3 * n);
    };
    auto load_checkpoint = [&](int c) {
        k_copy_pos<<<bn, ADJ_THREADS, 0, st>>>(ck_pos + static_cast<size_t>(c) * n, s->pos, n);
        k_copy_f64<<<b3, 256, 0, st>>>(ck_vel + static_cast<size_t>(c) * 3 * n, s->vel, 3 * n);
        k_copy_f64<<<b3, 256, 0, st>>>(ck_frc + static_cast<size_t>(c) * 3 * n, s->frc, 3 * n);
    };
    // one FP64 step over the fixed-order list, device-gated rebuild, KE by the fixed-order tree sum
    auto step = [&]() {
        kick_drift_enqueue(s, dt, l);
        list_enqueue_conditional_rebuild_fixed(l, s);
        forces_enqueue(s, l, cfg->lj, MD_FP64, s->scalars, l->flags + FLAG_STALE, nullptr, nullptr);
        kick_enqueue(s, dt, s->scalars + 3);
    };

    hipEvent_t e0, e1, e2;
    MD_HIP(hipEventCreate(&e0));
    MD_HIP(hipEventCreate(&e1));
    MD_HIP(hipEventCreate(&e2));
    struct EvGuard {
        hipEvent_t a, b, c;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            (void)hipEventDestroy(c);
        }
    } evg{e0, e1, e2};

    // A device-gated rebuild whose rows overflow leaves truncated rows behind for the rest of its segment: the
    // whole computation is then repeated from the caller's state with longer rows.
    int* hf = ctx->host_flags;
    const double target = cfg->pressure_target;
    double ubar = 0.0, pbar = 0.0;
    double h_theta[2] = {0.0, 0.0};
    for (int attempt = 0;; ++attempt) {
        if (attempt > 0) {
            const int rc = md_state_copy(s, src);
            if (rc != MD_OK) return rc;
        }
        MD_HIP(hipMemsetAsync(theta, 0, 2 * sizeof(double), st));
        // ---- forward -------------------------------------------------------------------------------------------
        MD_HIP(hipEventRecord(e0, st));
        list_rebuild_fixed(l, s);
        // from here on the device-gated rebuilds report an overflow through the sticky words
        MD_HIP(hipMemsetAsync(l->flags + FLAG_OVERFLOW_STICKY, 0, 2 * sizeof(int), st));
        static_assert(FLAG_MAXCOUNT_STICKY == FLAG_OVERFLOW_STICKY + 1, "the two words are cleared together");
        forces_enqueue(s, l, cfg->lj, MD_FP64, s->scalars, nullptr, nullptr, nullptr);
        save_checkpoint(0);
        for (int k = 1; k <= K; ++k) {
            step();
            k_record_obs<<<1, 1, 0, st>>>(s->scalars, obs, k);
            if (k % M == 0 && k < K) {
                list_rebuild_fixed(l, s);  // the recompute starts from a fresh list here too
                save_checkpoint(k / M);
            }
        }
        MD_HIP(hipEventRecord(e1, st));
        std::vector<double> h_obs(3 * static_cast<size_t>(K + 1));
        MD_HIP(hipMemcpyAsync(h_obs.data(), obs, sizeof(double) * h_obs.size(), hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
        ubar = 0.0;
        pbar = 0.0;
        for (int k = 1; k <= K; ++k) {
            ubar += h_obs[3 * k];
            pbar += (2.0 * h_obs[3 * k + 2] + h_obs[3 * k + 1]) / (3.0 * volume);
        }
        ubar /= K;
        pbar /= K;
        const double g = 2.0 * (pbar - target);
        const double cu = 1.0 / K, cw = g / (3.0 * volume * K), cke = 2.0 * g / (3.0 * volume * K);

        // ---- reverse ---------------------------------------------------------------------------------------------
        AdjLJ alj;
        alj.eps = cfg->lj.epsilon;
        alj.sigma = cfg->lj.sigma;
        alj.sig2 = cfg->lj.sigma * cfg->lj.sigma;
        alj.rc2 = cfg->lj.cutoff * cfg->lj.cutoff;
        {
            const double q6 = std::pow(cfg->lj.sigma / cfg->lj.cutoff, 6);
            alj.shift1 = 4.0 * (q6 * q6 - q6);
            alj.dshift_sig = cfg->lj.epsilon * 4.0 * (12.0 * q6 * q6 - 6.0 * q6) / cfg->lj.sigma;
        }
        auto pair_at = [&](int slot, double coef_u, double coef_w) {
            // positions of the point into the work state; keep the list valid for them
            k_copy_pos<<<bn, ADJ_THREADS, 0, st>>>(seg_pos + static_cast<size_t>(slot) * n, s->pos, n);
            displacement_check(l, s);
            list_enqueue_conditional_rebuild_fixed(l, s);
            k_adj_pair<<<bn, ADJ_THREADS, 0, st>>>(s->pos, n, l->rows, l->count, l->half, s->box, s->inv_box, alj, w, hw,
                                                   gu, gw, s->partials);
            reduce_partials_enqueue(s, static_cast<int>(bn), ADJ_PARTS, point);
            k_adj_theta<<<1, 1, 0, st>>>(point, theta, coef_u, coef_w, dt);
            // the stale flag is left set only by a check that rebuilt; clear it for the next point
            MD_HIP(hipMemsetAsync(l->flags + FLAG_STALE, 0, sizeof(int), st));
        };

        for (int seg = nck - 1; seg >= 0; --seg) {
            const int c = seg * M;
            const int e = std::min(K, c + M);
            load_checkpoint(seg);
            list_rebuild_fixed(l, s);
            k_copy_pos<<<bn, ADJ_THREADS, 0, st>>>(s->pos, seg_pos, n);
            k_copy_f64<<<b3, 256, 0, st>>>(s->vel, seg_vel, 3 * n);
            for (int k = c + 1; k <= e; ++k) {
                step();
                k_copy_pos<<<bn, ADJ_THREADS, 0, st>>>(s->pos, seg_pos + static_cast<size_t>(k - c) * n, n);
                k_copy_f64<<<b3, 256, 0, st>>>(s->vel, seg_vel + static_cast<size_t>(k - c) * 3 * n, 3 * n);
            }
            for (int k = e; k >= c + 1; --k) {
                if (k == K)
                    k_adj_start<<<b3, 256, 0, st>>>(n, lr, lv, pend, w, s->mass,
                                                    seg_vel + static_cast<size_t>(k - c) * 3 * n, cke);
                pair_at(k - c, cu, cw);
                k_adj_update<<<b3, 256, 0, st>>>(n, lr, lv, pend, w, hw, gu, gw, s->mass,
                                                 seg_vel + static_cast<size_t>(k - 1 - c) * 3 * n, cu, cw, cke, dt,
                                                 k - 1 == 0 ? 1 : 0);
            }
            if (c == 0) {
                pair_at(0, 0.0, 0.0);
                k_adj_finish<<<b3, 256, 0, st>>>(n, lr, hw, dt);
            }
            MD_HIP(hipGetLastError());
        }
        MD_HIP(hipEventRecord(e2, st));
        MD_HIP(hipMemcpyAsync(h_theta, theta, sizeof(h_theta), hipMemcpyDeviceToHost, st));
        MD_HIP(hipMemcpyAsync(hf, l->flags, sizeof(int) * FLAG_COUNT, hipMemcpyDeviceToHost, st));
        if (grad_positions) MD_HIP(hipMemcpyAsync(grad_positions, lr, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
        if (grad_velocities) MD_HIP(hipMemcpyAsync(grad_velocities, lv, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
        MD_HIP(hipStreamSynchronize(st));
        if (!hf[FLAG_OVERFLOW_STICKY]) break;
        if (attempt >= 4) fail(MD_ERR_OVERFLOW, "neighbour-list rows keep overflowing during md_adjoint_run");
        l->padding = (hf[FLAG_MAXCOUNT_STICKY] + hf[FLAG_MAXCOUNT_STICKY] / 4 + 16 + 7) / 8 * 8;
        l->built = false;
    }
    // the list now describes the work state, not the caller's: force a rebuild on its next use
    l->built = false;
    float fwd_ms = 0.f, bwd_ms = 0.f;
    MD_HIP(hipEventElapsedTime(&fwd_ms, e0, e1));
    MD_HIP(hipEventElapsedTime(&bwd_ms, e1, e2));
    out->loss = ubar + (pbar - target) * (pbar - target);
    out->mean_potential = ubar;
    out->mean_pressure = pbar;
    out->d_epsilon = h_theta[0];
    out->d_sigma = h_theta[1];
    out->forward_ms = fwd_ms;
    out->reverse_ms = bwd_ms;
    MD_API_END
}
