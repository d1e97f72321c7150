// k_ghost.hip — the reference's periodic Barnes-Hut mode for gfx950: every
// particle within max(W, H) * theta of an edge gets periodic images
// (_generateGhosts, integrator.js:201-244), reals and images together are
// the tree's points (_buildTree, integrator.js:247-258), and the walk counts a
// leaf point and each of its images (forces.js:643-649 skip only the
// observer itself and its own images).  The tree points are a separate SoA of
// up to four times the particles, sorted and built like the particles are.
#include "geon_kernels.h"

namespace geon {

struct GhostImage {
    float x, y;
    int flip_x, flip_y;
};

// The images of a particle at (x, y): _generateGhosts's cases in its order.
__device__ __forceinline__ int ghost_images(float x, float y, const GhostArgs& a, GhostImage* img) {
    const float W = a.dom_w, H = a.dom_h, m = a.margin;
    const bool L = x < m, R = x > W - m, T = y < m, B = y > H - m;
    int k = 0;
    auto add = [&](float sx, float sy, int fx, int fy) { img[k].x = sx; img[k].y = sy; img[k].flip_x = fx; img[k].flip_y = fy; k++; };
    if (a.topo == 0) {                  // torus
        if (L) add(x + W, y, 0, 0);
        if (R) add(x - W, y, 0, 0);
        if (T) add(x, y + H, 0, 0);
        if (B) add(x, y - H, 0, 0);
        if (L && T) add(x + W, y + H, 0, 0);
        if (L && B) add(x + W, y - H, 0, 0);
        if (R && T) add(x - W, y + H, 0, 0);
        if (R && B) add(x - W, y - H, 0, 0);
    } else if (a.topo == 1) {           // Klein bottle: the y wrap mirrors x and negates vx
        if (L) add(x + W, y, 0, 0);
        if (R) add(x - W, y, 0, 0);
        if (T) add(W - x, y + H, 1, 0);
        if (B) add(W - x, y - H, 1, 0);
        if (L && T) add(W - x + W, y + H, 1, 0);
        if (L && B) add(W - x + W, y - H, 1, 0);
        if (R && T) add(W - x - W, y + H, 1, 0);
        if (R && B) add(W - x - W, y - H, 1, 0);
    } else {                            // RP^2: the x wrap mirrors y, the y wrap mirrors x
        if (L) add(x + W, H - y, 0, 1);
        if (R) add(x - W, H - y, 0, 1);
        if (T) add(W - x, y + H, 1, 0);
        if (B) add(W - x, y - H, 1, 0);
        if (L && T) add(W - x + W, y + H, 1, 0);
        if (L && B) add(W - x + W, y - H, 1, 0);
        if (R && T) add(W - x - W, y + H, 1, 0);
        if (R && B) add(W - x - W, y - H, 1, 0);
    }
    return k;
}

__global__ void k_ghost_count(GhostArgs a, int32_t* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.ctrl->n_alive) return;
    GhostImage img[8];
    cnt[i] = ghost_images(a.x[i], a.y[i], a, img);
}

// Tree points: the reals at 0..n-1, the images of real i from n + off[i].
// Aggregate moments: a real contributes the moments cached by the last force
// pass (calculateMassDistribution reads p.magMoment, quadtree.js:272-309,
// before computeAllForces refreshes them); an image carries moments computed
// at its creation (_addGhost, integrator.js:198-199).
__global__ void k_ghost_write(GhostArgs a, const int32_t* __restrict__ off, Ctrl* __restrict__ gctrl) {
    const int n = a.ctrl->n_alive;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) gctrl->n_alive = min(n + gctrl->n_tree, a.cap_points);   // exclusive scan total in n_tree
    if (i >= n) return;
    uint32_t* const* o = a.out.c;
    const float x = a.x[i], y = a.y[i];
    const float m = a.m[i], q = a.q[i], r = a.radius[i];
    const float wx = a.wx[i], wy = a.wy[i], vx = a.vx[i], vy = a.vy[i], av = a.ang_vel[i];
    const int32_t qq = a.qq[i];
    auto put = [&](int j, float px, float py, int fx, int fy, float mu, float L) {
        const bool flip = fx || fy;
        const float cols[kGhostCols - 2] = {px, py, m, q, fx ? -wx : wx, fy ? -wy : wy, fx ? -vx : vx, fy ? -vy : vy,
                                           flip ? -av : av, r, mu, L};
        for (int c = 0; c < kGhostCols - 2; c++) reinterpret_cast<float*>(o[c])[j] = cols[c];
        reinterpret_cast<int32_t*>(o[G_QQ])[j] = qq;
        reinterpret_cast<int32_t*>(o[G_ORIG])[j] = i;
    };
    put(i, x, y, 0, 0, a.mu[i], a.L[i]);
    GhostImage img[8];
    const int k = ghost_images(x, y, a, img);
    const int base = n + off[i];
    if (base + k > a.cap_points) { gctrl->overflow = 1; return; }   // not reached: sized for the most images
    for (int t = 0; t < k; t++) {
        const float gav = (img[t].flip_x || img[t].flip_y) ? -av : av;
        const float r2 = r * r;
        put(base + t, img[t].x, img[t].y, img[t].flip_x, img[t].flip_y, 0.2f * q * gav * r2, 0.4f * m * gav * r2);
    }
}

// Walk sources of the sorted tree points; moments from the current spin
// (computeAllForces' cache, forces.js:71-76; an image's own at creation).
__global__ void k_ghost_sources(GhostCols col, const float* __restrict__ ax, const float* __restrict__ yuk,
                                const float* __restrict__ higgs, float4* __restrict__ src4,
                                float4* __restrict__ src8a, float4* __restrict__ src8b,
                                const Ctrl* __restrict__ gctrl) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= gctrl->n_alive) return;
    auto f = [&](int c) { return reinterpret_cast<const float*>(col.c[c])[j]; };
    const float m = f(G_M), q = f(G_Q), av = f(G_AV), r = f(G_RAD);
    const float r2 = r * r;             // as k_prepare_sources forms the particles' moments
    src4[j] = make_float4(f(G_X), f(G_Y), m, q);
    src8a[j] = make_float4(f(G_VX), f(G_VY), av, 0.2f * q * av * r2);
    const int o = reinterpret_cast<const int32_t*>(col.c[G_ORIG])[j];
    src8b[j] = make_float4(0.4f * m * av * r2, ax ? ax[o] : 1.f, yuk ? yuk[o] : 1.f, higgs ? higgs[o] : 1.f);
}

// After the walk: the ghost tree's diagnostics into the particles' controller.
__global__ void k_ghost_finish(const Ctrl* __restrict__ gctrl, Ctrl* __restrict__ ctrl) {
    if (gctrl->overflow) ctrl->overflow = gctrl->overflow;
    ctrl->n_divided = gctrl->n_divided;
    ctrl->depth_overflow += gctrl->depth_overflow;
}

__global__ void k_ghost_reset(Ctrl* gctrl) {
    gctrl->n_alive = 0; gctrl->n_tree = 0; gctrl->n_divided = 0;
    gctrl->overflow = 0; gctrl->depth_overflow = 0;
}

void ghost_tree_points(hipStream_t s, const GhostArgs& a, int n_upper, int32_t* cnt, int32_t* off, int32_t* scan_tmp,
                       Ctrl* gctrl) {
    hipLaunchKernelGGL(k_ghost_reset, dim3(1), dim3(1), 0, s, gctrl);
    if (n_upper <= 0) return;
    const dim3 grid((n_upper + 255) / 256), block(256);
    hipLaunchKernelGGL(k_ghost_count, grid, block, 0, s, a, cnt);
    exclusive_scan_i32(s, cnt, off, n_upper, &a.ctrl->n_alive, &gctrl->n_tree, scan_tmp);
    hipLaunchKernelGGL(k_ghost_write, grid, block, 0, s, a, off, gctrl);
}

void ghost_sources(hipStream_t s, const GhostCols& col, const float* ax, const float* yuk, const float* higgs,
                   float4* src4, float4* src8a, float4* src8b, int n_upper, const Ctrl* gctrl) {
    if (n_upper <= 0) return;
    hipLaunchKernelGGL(k_ghost_sources, dim3((n_upper + 255) / 256), dim3(256), 0, s, col, ax, yuk, higgs, src4, src8a,
                       src8b, gctrl);
}

void ghost_finish(hipStream_t s, const Ctrl* gctrl, Ctrl* ctrl) {
    hipLaunchKernelGGL(k_ghost_finish, dim3(1), dim3(1), 0, s, gctrl, ctrl);
}

}  // namespace geon
