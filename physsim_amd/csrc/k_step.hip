// k_step.hip — the substep controller, the fused Boris kick/rotate/kick +
// spin/torque + drift pass, boundary handling and the FP64 conservation
// reductions, for gfx950.
#include "geon_kernels.h"

namespace geon {

constexpr double kEps = 1e-12;          // EPSILON (config.js:30)

__device__ __forceinline__ void atomic_max_pos_double(uint64_t* addr, double v) {
    // non-negative doubles order like their bit patterns; NaN never wins
    // (the reference's `aSq > maxAccelSq` is false for NaN, integrator.js:822)
    if (!(v >= 0.0)) return;
    atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Max over a 256-thread block, then one atomic per block (every thread of the
// block must call it).
__device__ __forceinline__ void block_max_atomic(uint64_t* addr, double v) {
    __shared__ double s_part[4];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = s_part[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, s_part[w]);
        atomic_max_pos_double(addr, m);
    }
    __syncthreads();
}

// Max |F/m|^2 and max cyclotron frequency over the live particles
// (integrator.js:812-849).
// FP32 per particle (the forces are FP32); only the maxima go to FP64.
__device__ __forceinline__ void dt_scan_one(const StepArgs& a, int i, float& acc2f, float& cycf) {
        const float inv_m = 1.f / a.mass[i];
        const float fx = a.fx[i] * inv_m, fy = a.fy[i] * inv_m;
        acc2f = fx * fx + fy * fy;
        if (acc2f != acc2f) acc2f = -1.f;
        cycf = 0.f;
        if (a.mag || a.ext_bz_on) {
            const float absbz = fabsf(a.bz[i]);
            if (absbz > 0.f) cycf = fmaxf(cycf, fabsf(a.charge[i]) * absbz * inv_m);
        }
        if (a.gm) {
            const float absbgz = fabsf(a.bgz[i]);
            if (absbgz > 0.f) cycf = fmaxf(cycf, 4.f * absbgz);
        }
        if (cycf != cycf) cycf = -1.f;
}

__global__ void k_dt_scan(StepArgs a) {
    const int n = a.ctrl->n_alive;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float acc2f = -1.f, cycf = -1.f;
    if (i < n) dt_scan_one(a, i, acc2f, cycf);
    const double acc2 = acc2f, cyc = cycf;
    block_max_atomic(&a.ctrl->max_acc2_fresh, acc2);
    block_max_atomic(&a.ctrl->max_cyc, cyc);
}

void launch_dt_scan(hipStream_t s, const StepArgs& a) {
    if (a.n_upper <= 0) return;
    hipLaunchKernelGGL(k_dt_scan, dim3((a.n_upper + 255) / 256), dim3(256), 0, s, a);
}

// The adaptive-substep decision of integrator.js:810-874 for one system.
__device__ __forceinline__ void ctrl_step(Ctrl* c, double softening, int first) {
    if (c->total_steps >= c->max_substeps || !(c->dt_remain > kEps)) {
        c->active = 0;
        return;
    }
    const double acc2 = __longlong_as_double((long long)(first ? c->max_acc2_fresh : c->max_acc2_kick));
    const double cyc = __longlong_as_double((long long)c->max_cyc);
    double a_max = sqrt(acc2);
    if (a_max != a_max) a_max = 0;
    double dt_safe = a_max > kEps ? sqrt(softening / a_max) : c->dt_remain;
    if (cyc > 0) {
        const double dtc = 0.7853981633974483 / cyc;
        if (dtc < dt_safe) dt_safe = dtc;
    }
    const int budget = c->max_substeps - c->total_steps;
    const double need = ceil(c->dt_remain / dt_safe);
    const int steps = need < (double)budget ? (int)need : budget;
    const double dt_sub = c->dt_remain / steps;
    c->total_steps++;
    c->dt_remain -= dt_sub;
    c->dt_sub = dt_sub;
    c->sim_time += dt_sub;
    c->active = 1;
    c->max_acc2_kick = 0;
    c->max_acc2_fresh = 0;
    c->max_cyc = 0;
}

__global__ void k_ctrl(Ctrl* c, double softening, int first) { ctrl_step(c, softening, first); }

void launch_ctrl(hipStream_t s, Ctrl* ctrl, double softening, int first) {
    hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(1), 0, s, ctrl, softening, first);
}

// Fused substep body per particle: save 1PN force, half kick, Boris rotation,
// half kick, NaN guard (integrator.js:884-928); spin-orbit and torques
// (930-981); drift (1270-1278); expansion (1281-1292).
__device__ __forceinline__ double kick_drift_one(const StepArgs& a, int i, float dt) {
        double acc2;
        const float half_dt = 0.5f * dt;
        const float m = a.mass[i];
        const float inv_m = 1.f / m;
        const float h = half_dt * inv_m;
        if (a.one_pn) { a.f1ox[i] = a.f1x[i]; a.f1oy[i] = a.f1y[i]; }
        const float fx = a.fx[i], fy = a.fy[i];
        float wx = a.wx[i] + fx * h;
        float wy = a.wy[i] + fy * h;
        const float q = a.charge[i];
        float vx = a.vx[i], vy = a.vy[i];
        if (a.mag || a.gm || a.ext_bz_on) {
            float t = 0.f;
            if (a.mag || a.ext_bz_on) t += (q * 0.5f * inv_m) * a.bz[i];
            if (a.gm) t += 2.f * a.bgz[i];
            if (t != 0.f) {
                const float gamma = a.rel ? sqrtf(1.f + wx * wx + wy * wy) : 1.f;
                t *= dt / gamma;
                const float s = 2.f * t / (1.f + t * t);
                const float wpx = wx + wy * t, wpy = wy - wx * t;
                wx = wx + wpy * s;
                wy = wy - wpx * s;
            }
        }
        wx += fx * h;
        wy += fy * h;
        if (wx != wx || wy != wy) { wx = 0.f; wy = 0.f; vx = 0.f; vy = 0.f; }
        acc2 = ((double)fx * fx + (double)fy * fy) * (double)inv_m * (double)inv_m;
        float angw = a.angw[i];
        float ang_vel = a.ang_vel[i];
        const float r = a.radius[i];
        const bool need_so = a.spin_orbit && (a.mag || a.gm);
        const bool need_tq = (a.gm && a.rel) || a.grav || a.col_bounce;
        if (need_so && fabsf(ang_vel) >= 1e-12f) {
            const float ixo = a.L[i];       // angMomentum cached by the last force pass
            if (fabsf(ixo) >= 1e-12f) {
                const float dt_m = dt * inv_m;
                if (a.mag && fabsf(q) > 1e-12f) {
                    const float mu = a.mu[i];
                    angw -= mu * (vx * a.dbzdx[i] + vy * a.dbzdy[i]) * dt / ixo;
                    wx += mu * a.dbzdx[i] * dt_m;
                    wy += mu * a.dbzdy[i] * dt_m;
                }
                if (a.gm) {
                    const float L = ixo;
                    angw -= L * (vx * a.dbgzdx[i] + vy * a.dbgzdy[i]) * dt / ixo;
                    wx -= L * a.dbgzdx[i] * dt_m;
                    wy -= L * a.dbgzdy[i] * dt_m;
                }
                if (angw != angw) angw = 0.f;
                const float sr = angw * r;
                ang_vel = a.rel ? angw / sqrtf(1.f + sr * sr) : angw;
            }
        }
        if (need_tq) {
            float torque = 0.f;
            if (a.gm && a.rel) torque += a.frame_drag[i];
            if (a.grav) torque += a.tidal[i];
            if (a.contact) torque += a.contact[i];       // integrator.js:970
            if (torque != 0.f) {
                const float I = 0.4f * m * r * r;
                if (I >= 1e-12f) {
                    angw += torque * dt / I;
                    if (angw != angw) angw = 0.f;
                    const float sr = angw * r;
                    ang_vel = a.rel ? angw / sqrtf(1.f + sr * sr) : angw;
                }
            }
        }
        if (a.rad) {            // Landau-Lifshitz radiation reaction (integrator.js:983-1086), FP64
            double de = 0.0, dpx = 0.0, dpy = 0.0;
            double frx = 0.0, fry = 0.0;        // the force shown for the selected particle (1031-1032)
            const double wm2 = (double)wx * wx + (double)wy * wy;
            if (fabsf(q) >= 1e-12f && wm2 >= 1e-12) {
                const double md = m, inv_md = 1.0 / md;
                const double gamma = a.rel ? sqrt(1.0 + wm2) : 1.0;
                const double tau = 2.0 / 3.0 * ((double)q * q) * inv_md;
                frx = tau * a.jx[i];
                fry = tau * a.jy[i];
                if (a.rel && gamma > 1.0) {
                    const double ig3 = 1.0 / (gamma * gamma * gamma);
                    frx *= ig3; fry *= ig3;
                    const double ig = 1.0 / gamma;
                    const double ux = wx * ig, uy = wy * ig;
                    const double fsq = (double)fx * fx + (double)fy * fy;
                    const double vdf = ux * fx + uy * fy;
                    const double t23 = -tau * gamma * (fsq - vdf * vdf) * inv_md;
                    frx += t23 * ux; fry += t23 * uy;
                }
                const double ke_b = wm2 / (gamma + 1.0) * md;
                double nwx = wx + frx * dt * inv_md, nwy = wy + fry * dt * inv_md;
                if (nwx != nwx || nwy != nwy) { nwx = 0.0; nwy = 0.0; }
                const double wm2a = nwx * nwx + nwy * nwy;
                const double ke_a = wm2a / (sqrt(1.0 + wm2a) + 1.0) * md;
                de = ke_b - ke_a > 0.0 ? ke_b - ke_a : 0.0;
                if (de > 0.0) {
                    const double ax = fx * inv_md, ay = fy * inv_md;
                    const double am2 = ax * ax + ay * ay;
                    if (am2 > 1e-12) {
                        const double iam = 1.0 / sqrt(am2);
                        dpx = -de * ax * iam;
                        dpy = -de * ay * iam;
                    }
                }
                wx = (float)nwx;
                wy = (float)nwy;
            }
            a.rad_e[i] = de;
            a.rad_e[a.rad_stride + i] = dpx;
            a.rad_e[2 * a.rad_stride + i] = dpy;
            if (a.sel_out && i == *a.sel) { a.sel_out[SEL_RAD] = (float)frx; a.sel_out[SEL_RAD + 1] = (float)fry; }
        }
        // drift
        const float inv_g = a.rel ? 1.f / sqrtf(1.f + wx * wx + wy * wy) : 1.f;
        vx = wx * inv_g;
        vy = wy * inv_g;
        ang_vel = a.rel ? angw / sqrtf(1.f + angw * angw * r * r) : angw;
        float x = a.x[i] + vx * dt;
        float y = a.y[i] + vy * dt;
        if (a.expansion) {
            x += a.hubble * (x - a.cx) * dt;
            y += a.hubble * (y - a.cy) * dt;
            const float decay = 1.f - a.hubble * dt;
            wx *= decay;
            wy *= decay;
            if (a.rel && a.one_pn) {
                const float ig = 1.f / sqrtf(1.f + wx * wx + wy * wy);
                vx = wx * ig; vy = wy * ig;
            }
        }
        a.x[i] = x; a.y[i] = y;
        a.wx[i] = wx; a.wy[i] = wy;
        a.vx[i] = vx; a.vy[i] = vy;
        a.angw[i] = angw; a.ang_vel[i] = ang_vel;
        return acc2;
}

__global__ void k_kick_drift(StepArgs a) {
    const Ctrl* c = a.ctrl;
    if (!c->active) return;
    const int n = c->n_alive;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc2 = -1.0;
    if (i < n) acc2 = kick_drift_one(a, i, (float)c->dt_sub);
    block_max_atomic(&a.ctrl->max_acc2_kick, acc2);
}

void launch_kick_drift(hipStream_t s, const StepArgs& a) {
    if (a.n_upper <= 0) return;
    hipLaunchKernelGGL(k_kick_drift, dim3((a.n_upper + 255) / 256), dim3(256), 0, s, a);
}

// 1PN velocity-Verlet correction w += (F1PN_new - F1PN_old) dt/(2m)
// (integrator.js:1310-1316).
__global__ void k_vv_kick(StepArgs a) {
    const Ctrl* c = a.ctrl;
    if (!c->active) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c->n_alive) return;
    const float h = 0.5f * (float)c->dt_sub / a.mass[i];
    float wx = a.wx[i] + (a.f1x[i] - a.f1ox[i]) * h;
    float wy = a.wy[i] + (a.f1y[i] - a.f1oy[i]) * h;
    if (wx != wx || wy != wy) { wx = wy = 0.f; a.vx[i] = 0.f; a.vy[i] = 0.f; }
    a.wx[i] = wx; a.wy[i] = wy;
}

void launch_vv_kick(hipStream_t s, const StepArgs& a) {
    if (a.n_upper <= 0) return;
    hipLaunchKernelGGL(k_vv_kick, dim3((a.n_upper + 255) / 256), dim3(256), 0, s, a);
}

// wrapPosition (topology.js:81-135) for the Klein bottle and RP^2: a crossing
// of a glide edge mirrors the other coordinate and negates that axis's w and
// v, angw and the angular velocity.
__device__ __forceinline__ void glide_wrap(const StepArgs& a, int i, int topo, float w, float h, float& x, float& y) {
    float wx = a.wx[i], wy = a.wy[i], vx = a.vx[i], vy = a.vy[i], aw = a.angw[i], av = a.ang_vel[i];
    bool flipped = false;
    if (topo == 1) {
        if (x < 0.f) x += w; else if (x > w) x -= w;
        if (y < 0.f || y > h) {
            y += y < 0.f ? h : -h;
            x = w - x;
            wx = -wx; vx = -vx; aw = -aw; av = -av;
            flipped = true;
        }
    } else {
        for (int it = 0; it < 2; it++) {
            bool wrapped = false;
            if (x < 0.f || x > w) {
                x += x < 0.f ? w : -w;
                y = h - y;
                wy = -wy; vy = -vy; aw = -aw; av = -av;
                wrapped = flipped = true;
            }
            if (y < 0.f || y > h) {
                y += y < 0.f ? h : -h;
                x = w - x;
                wx = -wx; vx = -vx; aw = -aw; av = -av;
                wrapped = flipped = true;
            }
            if (!wrapped) break;
        }
    }
    if (flipped) {
        a.wx[i] = wx; a.wy[i] = wy; a.vx[i] = vx; a.vy[i] = vy; a.angw[i] = aw; a.ang_vel[i] = av;
    }
}

// Boundary handling once per update (integrator.js:1798-1824, topology.js:81-135).
__global__ void k_boundary(StepArgs a, int mode, int topo, float w, float h) {
    const int n = a.ctrl->n_alive;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = a.x[i], y = a.y[i];
    int alive = 1;
    if (mode == 0) {            // despawn (DESPAWN_MARGIN = 64, config.js:138)
        if (x < -64.f || x > w + 64.f || y < -64.f || y > h + 64.f) alive = 0;
    } else if (mode == 2 && topo == 0) {     // loop, torus wrap
        if (x < 0.f) x += w; else if (x > w) x -= w;
        if (y < 0.f) y += h; else if (y > h) y -= h;
    } else if (mode == 2) {
        glide_wrap(a, i, topo, w, h, x, y);
    } else {                    // bounce safety clamp
        x = fminf(fmaxf(x, 0.f), w);
        y = fminf(fmaxf(y, 0.f), h);
    }
    a.x[i] = x; a.y[i] = y;
    a.alive[i] = alive;
    if (!alive) atomicAdd(&a.ctrl->deaths, 1);
}

__global__ void k_wall_forces(WallArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.ctrl->n_alive) return;
    const float r = a.radius[i], x = a.x[i], y = a.y[i], vx = a.vx[i], vy = a.vy[i], w = a.ang_vel[i];
    const float mu = a.friction;
    float fx = a.fx[i], fy = a.fy[i], tq = 0.f, ex = 0.f, ey = 0.f;
    auto tangential = [&](float Fn, float vt) { return -mu * Fn * fminf(1.f, fmaxf(-1.f, vt * 10.f)); };
    float d = r - x;                                  // left wall
    if (d > 0.f) {
        const float Fn = d * sqrtf(d);
        fx += Fn; ex += Fn;
        if (mu > 0.f) { const float Ft = tangential(Fn, vy + w * r); fy += Ft; tq += r * Ft; }
    }
    d = r - (a.w - x);                                // right wall
    if (d > 0.f) {
        const float Fn = d * sqrtf(d);
        fx -= Fn; ex -= Fn;
        if (mu > 0.f) { const float Ft = tangential(Fn, -vy - w * r); fy -= Ft; tq -= r * Ft; }
    }
    d = r - y;                                        // top wall
    if (d > 0.f) {
        const float Fn = d * sqrtf(d);
        fy += Fn; ey += Fn;
        if (mu > 0.f) { const float Ft = tangential(Fn, -vx + w * r); fx -= Ft; tq += r * Ft; }
    }
    d = r - (a.h - y);                                // bottom wall
    if (d > 0.f) {
        const float Fn = d * sqrtf(d);
        fy -= Fn; ey -= Fn;
        if (mu > 0.f) { const float Ft = tangential(Fn, vx - w * r); fx += Ft; tq -= r * Ft; }
    }
    a.fx[i] = fx; a.fy[i] = fy;
    a.contact[i] = (a.accumulate ? a.contact[i] : 0.f) + tq;
    // forceExternal takes the normal forces only (integrator.js:619-658)
    if (a.sel_out && i == *a.sel) { a.sel_out[SEL_EXT] += ex; a.sel_out[SEL_EXT + 1] += ey; }
}

void launch_wall_forces(hipStream_t s, const WallArgs& a, int n_upper) {
    if (n_upper <= 0) return;
    hipLaunchKernelGGL(k_wall_forces, dim3((n_upper + 255) / 256), dim3(256), 0, s, a);
}

// One contact of the pair (p1, p2): Hertz force delta^1.5 along the unit
// vector n from p1 to p2 and the friction force Ft along t = (-n_y, n_x);
// false when they do not overlap.  Both members evaluate it with the same
// operands, so both see the same numbers.
__device__ __forceinline__ bool repel_pair(const RepelArgs& a, int i1, int i2, float& nFx, float& nFy, float& Ft,
                                           float& tx, float& ty) {
    float dx = a.x[i2] - a.x[i1], dy = a.y[i2] - a.y[i1];
    if (a.periodic) { dx = torus_wrap(dx, a.dom_w, 0.5f * a.dom_w); dy = torus_wrap(dy, a.dom_h, 0.5f * a.dom_h); }
    const float r1 = a.radius[i1], r2 = a.radius[i2];
    const float dist_sq = dx * dx + dy * dy;
    const float min_dist = r1 + r2;
    if (dist_sq >= min_dist * min_dist) return false;
    const float dist = sqrtf(dist_sq);
    const float delta = min_dist - dist;
    if (delta <= 0.f) return false;
    const float Fn = delta * sqrtf(delta);
    const float safe = dist > 1e-12f ? dist : 1e-12f;
    const float nx = dx / safe, ny = dy / safe;
    nFx = nx * Fn; nFy = ny * Fn;
    tx = -ny; ty = nx;
    Ft = 0.f;
    if (a.friction > 0.f) {
        const float v1t = a.vx[i1] * tx + a.vy[i1] * ty + a.ang_vel[i1] * r1;
        const float v2t = a.vx[i2] * tx + a.vy[i2] * ty - a.ang_vel[i2] * r2;
        Ft = -a.friction * Fn * fminf(1.f, fmaxf(-1.f, (v1t - v2t) * 10.f));
    }
    return true;
}

__global__ void k_repel_zero(int32_t* __restrict__ cnt2, int32_t* __restrict__ fill2, const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ctrl->n_alive) { cnt2[i] = 0; fill2[i] = 0; }
}

// grid-stride over the pairs the list holds
__global__ void k_repel_count2(RepelArgs a) {
    const int64_t np = min((int64_t)*a.n_pairs, a.cap_pairs);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&a.cnt2[a.pb[k]], 1);
}

__global__ void k_repel_fill2(RepelArgs a) {
    const int64_t np = min((int64_t)*a.n_pairs, a.cap_pairs);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x) {
        const int b = a.pb[k];
        a.list2[a.off2[b] + atomicAdd(&a.fill2[b], 1)] = (int32_t)k;
    }
}

__global__ void k_repulsion(RepelArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.ctrl->n_alive) return;
    float fx = a.fx[i], fy = a.fy[i], tq = 0.f, ex = 0.f, ey = 0.f;
    const float ri = a.radius[i];
    // as the higher-slot member: this body's entries in pair order (the
    // atomic fill left them in any order; a short insertion sort fixes it)
    int32_t* l2 = a.list2 + a.off2[i];
    const int c2 = a.cnt2[i];
    for (int u = 1; u < c2; u++) {
        const int32_t v = l2[u];
        int w = u - 1;
        while (w >= 0 && l2[w] > v) { l2[w + 1] = l2[w]; w--; }
        l2[w + 1] = v;
    }
    float nFx, nFy, Ft, tx, ty;
    for (int u = 0; u < c2; u++) {
        if (!repel_pair(a, a.pa[l2[u]], i, nFx, nFy, Ft, tx, ty)) continue;
        fx += nFx; fy += nFy; ex += nFx; ey += nFy;
        fx -= tx * Ft; fy -= ty * Ft;
        tq -= ri * Ft;
    }
    // as the lower-slot member: its own pairs in list order
    const int64_t np = min((int64_t)*a.n_pairs, a.cap_pairs);
    const int o1 = a.off1[i], c1 = a.cnt1[i];
    for (int u = 0; u < c1 && o1 + u < np; u++) {
        if (!repel_pair(a, i, a.pb[o1 + u], nFx, nFy, Ft, tx, ty)) continue;
        fx -= nFx; fy -= nFy; ex -= nFx; ey -= nFy;
        fx += tx * Ft; fy += ty * Ft;
        tq += ri * Ft;
    }
    a.fx[i] = fx; a.fy[i] = fy;
    a.contact[i] = tq;
    if (a.sel_out && i == *a.sel) { a.sel_out[SEL_EXT] += ex; a.sel_out[SEL_EXT + 1] += ey; }
}

void launch_repulsion(hipStream_t s, const RepelArgs& a, int n_upper) {
    if (n_upper <= 0) return;
    const dim3 grid((n_upper + 255) / 256), block(256);
    const int64_t pb = (a.cap_pairs + 255) / 256;
    const dim3 pgrid((unsigned)(pb < 4096 ? pb : 4096));
    hipLaunchKernelGGL(k_repel_zero, grid, block, 0, s, a.cnt2, a.fill2, a.ctrl);
    hipLaunchKernelGGL(k_repel_count2, pgrid, block, 0, s, a);
    exclusive_scan_i32(s, a.cnt2, a.off2, n_upper, &a.ctrl->n_alive, a.total2, a.scan_tmp);
    hipLaunchKernelGGL(k_repel_fill2, pgrid, block, 0, s, a);
    hipLaunchKernelGGL(k_repulsion, grid, block, 0, s, a);
}

void launch_boundary(hipStream_t s, const StepArgs& a, int mode, int topo, double w, double h) {
    if (a.n_upper <= 0) return;
    hipLaunchKernelGGL(k_boundary, dim3((a.n_upper + 255) / 256), dim3(256), 0, s, a, mode, topo, (float)w, (float)h);
}

// ---------------------------------------------------------------------------
// FP64 conservation reductions (energy.js:11-62), deterministic two-level.
// ---------------------------------------------------------------------------
constexpr int kStatThreads = 256;
constexpr int kStatCols = 8;

__global__ __launch_bounds__(kStatThreads) void k_stats_pass1(StepArgs a, const float* __restrict__ pe,
                                                              double* __restrict__ partial, int nb, int rel) {
    __shared__ double sh[kStatCols][kStatThreads];
    const int n = a.ctrl->n_alive;
    double v[kStatCols] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * kStatThreads + threadIdx.x; i < n; i += nb * kStatThreads) {
        const double m = a.mass[i];
        const double r = a.radius[i];
        const double r2 = r * r;
        const double wx = a.wx[i], wy = a.wy[i], angw = a.angw[i];
        if (rel) {
            const double wsq = wx * wx + wy * wy;
            const double g = sqrt(1 + wsq);
            v[0] += wsq / (g + 1) * m;
            const double srsq = angw * angw * r2;
            v[1] += 0.4 * m * srsq / (sqrt(1 + srsq) + 1);
        } else {
            const double vx = a.vx[i], vy = a.vy[i], av = a.ang_vel[i];
            v[0] += 0.5 * m * (vx * vx + vy * vy);
            v[1] += 0.5 * 0.4 * m * r2 * av * av;
        }
        v[2] += m * wx;
        v[3] += m * wy;
        v[4] += m;
        v[5] += m * (double)a.x[i];
        v[6] += m * (double)a.y[i];
        v[7] += (double)pe[i];
    }
#pragma unroll
    for (int c = 0; c < kStatCols; c++) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int off = kStatThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
#pragma unroll
            for (int c = 0; c < kStatCols; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < kStatCols) partial[threadIdx.x * nb + blockIdx.x] = sh[threadIdx.x][0];
}

// One wave per column: lane-strided partial sums, then a fixed-order wave reduction.
__global__ void k_stats_final(const double* __restrict__ partial, int nb, int ncols, double* __restrict__ out) {
    const int c = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (c >= ncols) return;
    double s = 0;
    for (int b = lane; b < nb; b += 64) s += partial[c * nb + b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[c] = s;
}

__global__ __launch_bounds__(kStatThreads) void k_stats_pass2(StepArgs a, const double* __restrict__ sums,
                                                              double* __restrict__ partial, int nb) {
    __shared__ double sh[2][kStatThreads];
    const int n = a.ctrl->n_alive;
    const double tm = sums[4];
    const double comx = tm > 0 ? sums[5] / tm : 0.0, comy = tm > 0 ? sums[6] / tm : 0.0;
    double orb = 0, spin = 0;
    for (int i = blockIdx.x * kStatThreads + threadIdx.x; i < n; i += nb * kStatThreads) {
        const double m = a.mass[i];
        const double r = a.radius[i];
        const double dx = (double)a.x[i] - comx, dy = (double)a.y[i] - comy;
        orb += dx * (m * (double)a.wy[i]) - dy * (m * (double)a.wx[i]);
        spin += 0.4 * m * r * r * (double)a.angw[i];
    }
    sh[0][threadIdx.x] = orb;
    sh[1][threadIdx.x] = spin;
    __syncthreads();
    for (int off = kStatThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2) partial[threadIdx.x * nb + blockIdx.x] = sh[threadIdx.x][0];
}

// FP64 sums of two per-particle FP32 columns (deterministic two-level).
__global__ __launch_bounds__(kStatThreads) void k_sum2(const float* __restrict__ a, const float* __restrict__ b,
                                                       const Ctrl* __restrict__ ctrl, double* __restrict__ partial,
                                                       int nb) {
    __shared__ double sh[2][kStatThreads];
    const int n = ctrl->n_alive;
    double sa = 0, sb = 0;
    for (int i = blockIdx.x * kStatThreads + threadIdx.x; i < n; i += nb * kStatThreads) {
        sa += (double)a[i];
        sb += (double)b[i];
    }
    sh[0][threadIdx.x] = sa;
    sh[1][threadIdx.x] = sb;
    __syncthreads();
    for (int off = kStatThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2) partial[threadIdx.x * nb + blockIdx.x] = sh[threadIdx.x][0];
}

void launch_sum2(hipStream_t s, const float* a, const float* b, int n_upper, const Ctrl* ctrl, double* partial,
                 double* out) {
    int nb = (n_upper + kStatThreads - 1) / kStatThreads;
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_sum2, dim3(nb), dim3(kStatThreads), 0, s, a, b, ctrl, partial, nb);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(2 * 64), 0, s, partial, nb, 2, out);
}

// Darwin field energy and momentum, O(v^2/c^2) (energy.js:64-142): every
// unordered pair once, in FP64.  The pair is oriented as the reference's array
// order orients it: the particle added first (lower slot) observes, and with
// signal delay on sees the other at its retarded state (energy.js:85-93).
// Pairs are visited as j > i in the sorted order, so each thread's row starts
// at its own tile; per-thread sums, then fixed-order block and final
// reductions (deterministic).
constexpr int kDarwinTile = 256;

template <bool DELAY>
__global__ __launch_bounds__(kDarwinTile) void k_darwin(DarwinArgs d, double* __restrict__ partial, int nb) {
    __shared__ float4 tpv[kDarwinTile];     // x, y, vx, vy
    __shared__ float4 tmq[kDarwinTile];     // m, q, axMod
    __shared__ int32_t tsl[kDarwinTile];    // slot
    __shared__ double sh[3][kDarwinTile];
    const int n = d.ctrl->n_alive;
    const int i = blockIdx.x * kDarwinTile + threadIdx.x;
    const bool valid = i < n;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), mi = make_float4(0.f, 0.f, 1.f, 0.f);
    int si = 0;
    if (valid) {
        pi = make_float4(d.x[i], d.y[i], d.vx[i], d.vy[i]);
        mi = make_float4(d.mass[i], d.charge[i], d.ax ? d.ax[i] : 1.f, 0.f);
        si = d.slot[i];
    }
    const double now = DELAY ? d.ctrl->sim_time : 0.0;
    const double soft2 = d.soft2;
    double e = 0.0, fpx = 0.0, fpy = 0.0;
    for (int base = blockIdx.x * kDarwinTile; base < n; base += kDarwinTile) {
        __syncthreads();
        const int jj = base + threadIdx.x;
        if (jj < n) {
            tpv[threadIdx.x] = make_float4(d.x[jj], d.y[jj], d.vx[jj], d.vy[jj]);
            tmq[threadIdx.x] = make_float4(d.mass[jj], d.charge[jj], d.ax ? d.ax[jj] : 1.f, 0.f);
            tsl[threadIdx.x] = d.slot[jj];
        }
        __syncthreads();
        if (!valid) continue;
        const int cnt = min(kDarwinTile, n - base);
        for (int k = 0; k < cnt; k++) {
            if (base + k <= i) continue;
            const float4 pj = tpv[k], mj = tmq[k];
            const int sj = tsl[k];
            const bool i_obs = si < sj;
            const float4 po = i_obs ? pi : pj, ps = i_obs ? pj : pi;
            const float4 mo = i_obs ? mi : mj, ms = i_obs ? mj : mi;
            double dx, dy, svx = ps.z, svy = ps.w;
            if (DELAY) {
                Retarded rt;
                if (!delayed_state(d.hist, d.pf, i_obs ? sj : si, po.x, po.y, now, rt)) continue;
                dx = rt.dx; dy = rt.dy; svx = rt.vx; svy = rt.vy;
            } else {
                delay_disp64(d.pf, po.x, po.y, ps.x, ps.y, dx, dy);
            }
            const double ovx = po.z, ovy = po.w;
            const double inv_r = 1.0 / sqrt(dx * dx + dy * dy + soft2);
            const double rx = dx * inv_r, ry = dy * inv_r;
            const double vel_term = (ovx * svx + ovy * svy) + (ovx * rx + ovy * ry) * (svx * rx + svy * ry);
            const double sx = ovx + svx, sy = ovy + svy;
            const double sdr = sx * rx + sy * ry;
            const double mo_ = mo.x, ms_ = ms.x, qo = mo.y, qs = ms.y;
            if (d.mag) {
                const double qq = qo * qs * inv_r * sqrt((double)mo.z * (double)ms.z);
                if (d.em_energy) e -= 0.5 * qq * vel_term;
                fpx += 0.5 * qq * (sx + rx * sdr);
                fpy += 0.5 * qq * (sy + ry * sdr);
            }
            if (d.gm) {
                const double mm = mo_ * ms_ * inv_r;
                if (d.gm_energy) e += 0.5 * mm * vel_term;
                fpx -= 0.5 * mm * (sx + rx * sdr);
                fpy -= 0.5 * mm * (sy + ry * sdr);
            }
            if (d.bazanski) {
                const double cc = qo * qs * (mo_ + ms_) - (qo * qo * ms_ + qs * qs * mo_);
                e += 0.5 * cc * inv_r * inv_r;
            }
        }
    }
    sh[0][threadIdx.x] = e;
    sh[1][threadIdx.x] = fpx;
    sh[2][threadIdx.x] = fpy;
    __syncthreads();
    for (int off = kDarwinTile / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int c = 0; c < 3; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[threadIdx.x * nb + blockIdx.x] = sh[threadIdx.x][0];
}

void launch_darwin(hipStream_t s, const DarwinArgs& d, int n_upper, double* partial, double* out) {
    if (n_upper <= 0) return;
    const int nb = (n_upper + kDarwinTile - 1) / kDarwinTile;
    if (d.hist.on) hipLaunchKernelGGL(k_darwin<true>, dim3(nb), dim3(kDarwinTile), 0, s, d, partial, nb);
    else hipLaunchKernelGGL(k_darwin<false>, dim3(nb), dim3(kDarwinTile), 0, s, d, partial, nb);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(3 * 64), 0, s, partial, nb, 3, out);
}

void launch_stats(hipStream_t s, const StepArgs& a, const float* pe, double* partial, double* out, int rel) {
    int nb = (a.n_upper + kStatThreads - 1) / kStatThreads;
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_stats_pass1, dim3(nb), dim3(kStatThreads), 0, s, a, pe, partial, nb, rel);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kStatCols * 64), 0, s, partial, nb, kStatCols, out);
    hipLaunchKernelGGL(k_stats_pass2, dim3(nb), dim3(kStatThreads), 0, s, a, out, partial, nb);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(2 * 64), 0, s, partial, nb, 2, out + kStatCols);
}


// ---------------------------------------------------------------------------
// Batches: the same per-particle bodies with each system's own controller.
// Slices are kBatchAlign-aligned, so every wavefront lies in one system.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_max_to(uint64_t* addr, double v) {
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) atomic_max_pos_double(addr, v);
}

__global__ void k_batch_begin(Ctrl* __restrict__ c, int S, double dt, int max_sub) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    c[s].dt_remain = dt; c[s].dt_total = dt; c[s].total_steps = 0; c[s].max_substeps = max_sub;
    c[s].active = 0; c[s].deaths = 0;
    c[s].max_acc2_fresh = 0; c[s].max_acc2_kick = 0; c[s].max_cyc = 0;
}

__global__ void k_batch_ctrl(Ctrl* __restrict__ c, int S, double softening, int first) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) ctrl_step(&c[s], softening, first);
}

__global__ void k_batch_kick_drift(StepArgs a, BatchStepArgs b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.total) return;
    const int s = b.sys[i];
    const Ctrl* c = &b.ctrl[s];
    if (!c->active) return;                          // the whole wave shares s
    double acc2 = -1.0;
    if (b.valid[i]) acc2 = kick_drift_one(a, i, (float)c->dt_sub);
    wave_max_to(&b.ctrl[s].max_acc2_kick, acc2);
}

__global__ void k_batch_dt_scan(StepArgs a, BatchStepArgs b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.total) return;
    const int s = b.sys[i];
    float acc2f = -1.f, cycf = -1.f;
    if (b.valid[i]) dt_scan_one(a, i, acc2f, cycf);
    wave_max_to(&b.ctrl[s].max_acc2_fresh, (double)acc2f);
    wave_max_to(&b.ctrl[s].max_cyc, (double)cycf);
}

// Despawn once per update (integrator.js:1798-1824): the particle stops being live.
__global__ void k_batch_boundary(StepArgs a, BatchStepArgs b, int mode, float w, float h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.total || !b.valid[i]) return;
    float x = a.x[i], y = a.y[i];
    if (mode == 0) {
        if (x < -64.f || x > w + 64.f || y < -64.f || y > h + 64.f) {
            b.valid[i] = 0;
            atomicAdd(&b.ctrl[b.sys[i]].deaths, 1);
        }
    } else if (mode == 2) {
        if (x < 0.f) x += w; else if (x > w) x -= w;
        if (y < 0.f) y += h; else if (y > h) y -= h;
    } else {
        x = fminf(fmaxf(x, 0.f), w);
        y = fminf(fmaxf(y, 0.f), h);
    }
    a.x[i] = x; a.y[i] = y;
}

// FP64 conservation sums of each system (energy.js:11-62): one block per
// system, fixed-order reductions; out[s * 10 + k] in the launch_stats order.
__global__ __launch_bounds__(kStatThreads) void k_batch_stats(StepArgs a, BatchStepArgs b, const float* __restrict__ pe,
                                                              int rel, double* __restrict__ out) {
    __shared__ double sh[kStatCols][kStatThreads];
    const int s = blockIdx.x;
    const int lo = b.soff[s], hi = b.soff[s + 1];
    double v[kStatCols] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = lo + threadIdx.x; i < hi; i += kStatThreads) {
        if (!b.valid[i]) continue;
        const double m = a.mass[i];
        const double r = a.radius[i];
        const double r2 = r * r;
        const double wx = a.wx[i], wy = a.wy[i], angw = a.angw[i];
        if (rel) {
            const double wsq = wx * wx + wy * wy;
            const double g = sqrt(1 + wsq);
            v[0] += wsq / (g + 1) * m;
            const double srsq = angw * angw * r2;
            v[1] += 0.4 * m * srsq / (sqrt(1 + srsq) + 1);
        } else {
            const double vx = a.vx[i], vy = a.vy[i], av = a.ang_vel[i];
            v[0] += 0.5 * m * (vx * vx + vy * vy);
            v[1] += 0.5 * 0.4 * m * r2 * av * av;
        }
        v[2] += m * wx;
        v[3] += m * wy;
        v[4] += m;
        v[5] += m * (double)a.x[i];
        v[6] += m * (double)a.y[i];
        v[7] += (double)pe[i];
    }
    for (int c = 0; c < kStatCols; c++) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int off = kStatThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int c = 0; c < kStatCols; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + off];
        __syncthreads();
    }
    const double tm = sh[4][0];
    const double comx = tm > 0 ? sh[5][0] / tm : 0.0, comy = tm > 0 ? sh[6][0] / tm : 0.0;
    if (threadIdx.x < kStatCols) out[s * 10 + threadIdx.x] = sh[threadIdx.x][0];
    __syncthreads();
    double orb = 0, spin = 0;
    for (int i = lo + threadIdx.x; i < hi; i += kStatThreads) {
        if (!b.valid[i]) continue;
        const double m = a.mass[i];
        const double r = a.radius[i];
        const double dx = (double)a.x[i] - comx, dy = (double)a.y[i] - comy;
        orb += dx * (m * (double)a.wy[i]) - dy * (m * (double)a.wx[i]);
        spin += 0.4 * m * r * r * (double)a.angw[i];
    }
    sh[0][threadIdx.x] = orb;
    sh[1][threadIdx.x] = spin;
    __syncthreads();
    for (int off = kStatThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[s * 10 + 8] = sh[0][0]; out[s * 10 + 9] = sh[1][0]; }
}

void launch_batch_begin(hipStream_t s, Ctrl* ctrls, int S, double dt, int max_sub) {
    hipLaunchKernelGGL(k_batch_begin, dim3((S + 255) / 256), dim3(256), 0, s, ctrls, S, dt, max_sub);
}
void launch_batch_ctrl(hipStream_t s, Ctrl* ctrls, int S, double softening, int first) {
    hipLaunchKernelGGL(k_batch_ctrl, dim3((S + 255) / 256), dim3(256), 0, s, ctrls, S, softening, first);
}
void launch_batch_kick_drift(hipStream_t s, const StepArgs& a, const BatchStepArgs& b) {
    hipLaunchKernelGGL(k_batch_kick_drift, dim3((b.total + 255) / 256), dim3(256), 0, s, a, b);
}
void launch_batch_dt_scan(hipStream_t s, const StepArgs& a, const BatchStepArgs& b) {
    hipLaunchKernelGGL(k_batch_dt_scan, dim3((b.total + 255) / 256), dim3(256), 0, s, a, b);
}
void launch_batch_boundary(hipStream_t s, const StepArgs& a, const BatchStepArgs& b, int mode, double w, double h) {
    hipLaunchKernelGGL(k_batch_boundary, dim3((b.total + 255) / 256), dim3(256), 0, s, a, b, mode, (float)w, (float)h);
}
void launch_batch_stats(hipStream_t s, const StepArgs& a, const BatchStepArgs& b, const float* pe, int rel,
                        double* out) {
    hipLaunchKernelGGL(k_batch_stats, dim3(b.S), dim3(kStatThreads), 0, s, a, b, pe, rel, out);
}

}  // namespace geon
