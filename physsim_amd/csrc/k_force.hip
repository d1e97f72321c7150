// k_force.hip — force kernels for gfx950: the Barnes-Hut walk with
// wavefront-shared node stacks, and the LDS-tiled O(N^2) direct sum.
//
// Walk (forces.js:622-694 calculateForce): one lane per particle, particles in
// Morton order so the 64 lanes of a wavefront are spatial neighbours.  The wave
// walks ONE shared stack held in LDS; every stack entry carries the 64-bit mask
// of the lanes whose own walk reaches that node.  At a node each lane in the
// mask makes the reference's decision (skip light node / leaf points / accept
// aggregate / open); the lanes that open form the mask pushed with the four
// children.  Each lane therefore visits exactly the node sequence of the
// reference's per-particle stack walk, in the same order (children pushed
// NW,NE,SW,SE, popped SE first), while node records are fetched once per wave
// through the scalar cache.
#include "geon_kernels.h"

namespace geon {

constexpr int kWalkThreads = 256;
constexpr int kWavesPerBlock = kWalkThreads / kWave;

__global__ void k_prepare_sources(const float* __restrict__ x, const float* __restrict__ y,
                                  const float* __restrict__ m, const float* __restrict__ q,
                                  const float* __restrict__ vx, const float* __restrict__ vy,
                                  const float* __restrict__ ang_vel, const float* __restrict__ radius,
                                  float* __restrict__ mu, float* __restrict__ L, const float* __restrict__ ax,
                                  const float* __restrict__ yuk, const float* __restrict__ higgs,
                                  float4* __restrict__ src4, float4* __restrict__ src8a,
                                  float4* __restrict__ src8b, int update_moments, const Ctrl* __restrict__ ctrl) {
    const int n = ctrl->n_alive;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = radius[i];
    const float r2 = r * r;
    const float w = ang_vel[i];
    float mui, Li;
    if (update_moments) {
        // forces.js:71-76: cached dipole moments for this substep
        mui = 0.2f * q[i] * w * r2;
        Li = 0.4f * m[i] * w * r2;
        mu[i] = mui;
        L[i] = Li;
    } else {
        mui = mu[i];
        Li = L[i];
    }
    src4[i] = make_float4(x[i], y[i], m[i], q[i]);
    src8a[i] = make_float4(vx[i], vy[i], w, mui);
    src8b[i] = make_float4(Li, ax ? ax[i] : 1.f, yuk ? yuk[i] : 1.f, higgs ? higgs[i] : 1.f);
}

void launch_prepare_sources(hipStream_t s, const float* x, const float* y, const float* m, const float* q,
                            const float* vx, const float* vy, const float* ang_vel, const float* radius, float* mu,
                            float* L, const float* ax, const float* yuk, const float* higgs, float4* src4,
                            float4* src8a, float4* src8b, int n_upper, const Ctrl* ctrl, int update_moments) {
    if (n_upper <= 0) return;
    hipLaunchKernelGGL(k_prepare_sources, dim3((n_upper + 255) / 256), dim3(256), 0, s, x, y, m, q, vx, vy, ang_vel,
                       radius, mu, L, ax, yuk, higgs, src4, src8a, src8b, update_moments, ctrl);
}

__device__ __forceinline__ void load_obs(const ForceArgs& a, int i, Obs& o) {
    const float4 s4 = (a.obs4 ? a.obs4 : a.src4)[i];
    const float4 s8a = (a.obs8a ? a.obs8a : a.src8a)[i];
    const float4 s8b = (a.obs8b ? a.obs8b : a.src8b)[i];
    o.x = s4.x; o.y = s4.y; o.m = s4.z; o.q = s4.w;
    o.vx = s8a.x; o.vy = s8a.y; o.ang_vel = s8a.z; o.mu = s8a.w;
    o.L = s8b.x; o.ax = s8b.y; o.yuk = s8b.z; o.higgs = s8b.w;
    const float r = a.radius[i];
    const float r2 = r * r;
    o.tidal_ri5 = r2 * r2 * r;
    o.inv_m = o.m > kEpsMass ? 1.f / o.m : 0.f;
}

template <uint32_t F, bool KAHAN>
__device__ __forceinline__ void store_acc(const ForceArgs& a, int i, const Obs& o, const Acc<KAHAN>& acc) {
    if (F & F_JERK) { a.jx[i] = acc.jx; a.jy[i] = acc.jy; }
    // _applyExternalFields (integrator.js:563-592) folded into the store
    a.fx[i] = acc.fx.s + o.m * a.ext_gx + o.q * a.ext_ex;
    a.fy[i] = acc.fy.s + o.m * a.ext_gy + o.q * a.ext_ey;
    a.pe[i] = acc.pe.s;
    a.bz[i] = acc.bz + a.ext_bz;
    a.dbzdx[i] = acc.dbzdx; a.dbzdy[i] = acc.dbzdy;
    a.bgz[i] = acc.bgz; a.dbgzdx[i] = acc.dbgzdx; a.dbgzdy[i] = acc.dbgzdy;
    a.frame_drag[i] = acc.frame_drag; a.tidal[i] = acc.tidal;
    a.f1x[i] = acc.f1x; a.f1y[i] = acc.f1y;
}

__device__ __forceinline__ void wave_add_counters(Ctrl* ctrl, int64_t pairs, int64_t nodes, int64_t visits) {
    for (int off = 32; off > 0; off >>= 1) {
        pairs += __shfl_xor(pairs, off, 64);
        nodes += __shfl_xor(nodes, off, 64);
        visits += __shfl_xor(visits, off, 64);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicAdd((unsigned long long*)&ctrl->cnt_pair, (unsigned long long)pairs);
        atomicAdd((unsigned long long*)&ctrl->cnt_node, (unsigned long long)nodes);
        atomicAdd((unsigned long long*)&ctrl->cnt_visit, (unsigned long long)visits);
    }
}

// FAR: 0 monopole nodes (the reference), kFarMultipole, kFarList (compile-time,
// so the reference walk carries no far-field registers).  DELAY: leaf sources
// act from their past light cone (forces.js:651-672).  SEL: only the lane of
// the selected particle (*a.sel) walks, summing its forces per type into *split
// and storing nothing else.
template <uint32_t F, bool KAHAN, int FAR, int STACK = kWalkStack, bool DELAY = false, bool SEL = false>
__device__ __forceinline__ void walk_core(const ForceArgs& a, int64_t group, Split* split, const Node* nodes,
                                          uint32_t* far_list, const float4* src4, const float4* src8a,
                                          const float4* src8b) {
    __shared__ int32_t s_node[kWavesPerBlock][STACK];
    __shared__ uint64_t s_mask[kWavesPerBlock][STACK];
    const int n = a.ctrl->n_alive;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int i = (int)(group * kWave) + lane;
    bool valid = i < n;
    if constexpr (SEL) valid = valid && i == *a.sel;
    int my_sys = -1;
    if (a.bsys != nullptr && valid) {
        // batch: live particles of systems still inside their update
        my_sys = a.bsys[i];
        valid = a.bvalid[i] && a.bctrl[my_sys].active;
    }
    const uint64_t active = __ballot(valid);
    if (active == 0) return;                       // whole wave past the end (or idle)
    const uint64_t my_bit = 1ull << lane;
    Obs o;
    if (valid) load_obs(a, i, o);
    else { o.x = o.y = o.vx = o.vy = o.m = o.q = o.ang_vel = o.mu = o.L = 0.f; o.ax = o.yuk = o.higgs = 1.f; o.tidal_ri5 = 0.f; o.inv_m = 0.f; }
    Acc<KAHAN> acc;
    acc.init();
    const PhysFlags pf = a.pf;
    const float half_w = 0.5f * pf.dom_w, half_h = 0.5f * pf.dom_h;
    const double now = DELAY ? a.ctrl->sim_time : 0.0;
    int64_t c_pair = 0, c_node = 0, c_visit = 0;
    int far_n = 0;                                 // entries of this lane's far queue (FAR == kFarList)
    const int far_lane_cap = (int)(a.far_cap / kWave);   // a multiple of 4
    // Each lane's last queue entries wait in LDS until they fill one 16-byte
    // chunk of far_slot's layout, written with one store: the lists' HBM
    // writes fall to about their size, without registers held for them.
    __shared__ uint4 s_far[(FAR == kFarList) ? kWavesPerBlock * kWave : 1];
    int sp = 0;
    if (a.bsys == nullptr) {
        if (lane == 0) { s_node[wv][0] = 0; s_mask[wv][0] = active; }
        sp = 1;
    } else {
        // one root per system among the wave's lanes, each with its lanes' mask
        uint64_t rem = active;
        while (rem != 0 && sp < STACK - 4) {
            const int sy = __shfl(my_sys, (int)__builtin_ctzll(rem), 64);
            const uint64_t m = __ballot(valid && my_sys == sy);
            if (lane == 0) { s_node[wv][sp] = a.broot[sy]; s_mask[wv][sp] = m; }
            sp++;
            rem &= ~m;
        }
    }
    __builtin_amdgcn_wave_barrier();
    bool overflow = false;
    while (sp > 0) {
        sp--;
        const int node = __builtin_amdgcn_readfirstlane(s_node[wv][sp]);
        const uint64_t mask = s_mask[wv][sp];
        const bool mine = (mask & my_bit) != 0;
        if (node < 0 || node >= a.max_nodes) { overflow = true; continue; }
        const Node nd = nodes[node];
        c_visit += mine;
        if (nd.mass < kEpsMass) continue;          // forces.js:630
        if (nd.count > 0) {
            // leaf: every point except the particle itself (forces.js:643-676)
            const int first = nd.child, cnt = nd.count;
            for (int j = first; j < first + cnt; j++) {
                const float4 s4 = src4[j];
                const float4 s8a = src8a[j];
                float4 s8b = make_float4(0.f, 1.f, 1.f, 1.f);
                if (F & F_EXTRA) s8b = src8b[j];
                if (DELAY) {
                    const int sj = a.hist.slot[j];
                    Retarded rt;
                    if (mine && j != i && delayed_state(a.hist, pf, sj, o.x, o.y, now, rt)) {
                        const float rx = rt.dx, ry = rt.dy;
                        const float rad = a.radius[j];
                        const float r2 = rad * rad;
                        const float av = retarded_ang_vel(rt.aw, r2);
                        pair_force<F, KAHAN, true, true, SEL>(o, acc, pf, rx, ry, s4.z, s4.w, rt.vx, rt.vy, av,
                                                              0.2f * s4.w * av * r2, 0.4f * s4.z * av * r2, s8b.y,
                                                              s8b.z, s8b.w, split);
                        c_pair++;
                    }
                } else if (mine && (a.src_orig ? a.src_orig[j] : j) != i) {
                    // with periodic images in the tree: not the observer nor one of its images
                    float rx, ry;
                    min_image_f<F>(pf, o.x, o.y, s4.x, s4.y, rx, ry);
                    pair_force<F, KAHAN, false, true, SEL>(o, acc, pf, rx, ry, s4.z, s4.w, s8a.x, s8a.y, s8a.z,
                                                           s8a.w, s8b.x, s8b.y, s8b.z, s8b.w, split);
                    c_pair++;
                }
            }
        } else if (nd.count < 0) {
            const float rawx = nd.comx - o.x, rawy = nd.comy - o.y;
            float dx, dy;
            min_image_f<F>(pf, o.x, o.y, nd.comx, nd.comy, dx, dy);
            const float d2 = dx * dx + dy * dy;
            bool accept = mine && (nd.size2 < pf.theta2 * d2);   // forces.js:680
            if (window_f<F>(pf)) {
                // the node's whole box must map through the COM's image shift into the window
                const float cx = nd.bx - o.x + (dx - rawx), cy = nd.by - o.y + (dy - rawy);
                accept = accept && (fabsf(cx) + nd.half <= half_w) && (fabsf(cy) + nd.half * (pf.dom_h / pf.dom_w) <= half_h);
            }
            if (accept) {
                // aggregate: sAngVel = 0, mods = 1 (forces.js:681-685)
                pair_force<F, KAHAN, false, true, SEL>(o, acc, pf, dx, dy, nd.mass, nd.charge, nd.avx, nd.avy, 0.f,
                                                       nd.mu, nd.L, 1.f, 1.f, 1.f, split);
                c_node++;
                if (FAR == kFarMultipole) far_multipole<F, KAHAN, SEL>(o, acc, pf, dx, dy, a.node_mom[node], split);
            }
            if (FAR == kFarList && accept) {
                // each accepting lane appends the node to its own queue
                uint32_t* q = reinterpret_cast<uint32_t*>(&s_far[wv * kWave + lane]);
                q[far_n & 3] = (uint32_t)node;
                if ((far_n & 3) == 3 && far_n < far_lane_cap)
                    *reinterpret_cast<uint4*>(far_list + far_slot(group, a.far_cap, lane, far_n - 3)) =
                        s_far[wv * kWave + lane];
                far_n++;
            }
            const uint64_t open = __ballot(mine && !accept);
            if (open) {
                if (sp + 4 <= STACK) {
                    if (lane == 0) {
                        // push nw, ne, sw, se: se is popped first (forces.js:688-691)
                        s_node[wv][sp] = nd.child;     s_mask[wv][sp] = open;
                        s_node[wv][sp + 1] = nd.child + 1; s_mask[wv][sp + 1] = open;
                        s_node[wv][sp + 2] = nd.child + 2; s_mask[wv][sp + 2] = open;
                        s_node[wv][sp + 3] = nd.child + 3; s_mask[wv][sp + 3] = open;
                    }
                    __builtin_amdgcn_wave_barrier();
                    sp += 4;
                } else {
                    overflow = true;
                }
            }
        }
    }
    if constexpr (SEL) return;                     // the force pass has its counters and outputs
    if (overflow && lane == 0) a.ctrl->overflow = 2;
    if (FAR == kFarList) {
        if ((far_n & 3) != 0 && far_n < far_lane_cap)   // the partial last chunk; entries past far_n are not read
            *reinterpret_cast<uint4*>(far_list + far_slot(group, a.far_cap, lane, far_n & ~3)) =
                s_far[wv * kWave + lane];
        a.far_count[group * kWave + lane] = far_n < far_lane_cap ? far_n : far_lane_cap;
        if (__ballot(far_n > far_lane_cap) != 0 && lane == 0) a.ctrl->overflow = 3;
    }
    wave_add_counters(a.ctrl, c_pair, c_node, c_visit);
    if (valid) store_acc<F, KAHAN>(a, i, o, acc);
}

// The listing walk takes the tree and source pointers as restrict parameters:
// its far-queue stores then provably miss the node and source records, so
// those loads of a wave-uniform address stay scalar loads as in the other
// walks.  With a possibly aliasing global store in the loop the compiler turns
// them into vector loads (1.3e8 more vector memory instructions per pass at
// N = 2^20, 9% of the walk's time).  The other walks store nothing in the loop
// and keep plain pointers (restrict made the plain walk 1.5% slower).
template <uint32_t F, bool KAHAN, int STACK, bool DELAY, bool SEL>
__device__ __forceinline__ void walk_list_core(const ForceArgs& a, int64_t group, Split* split,
                                               const Node* __restrict__ nodes, uint32_t* __restrict__ far_list,
                                               const float4* __restrict__ src4, const float4* __restrict__ src8a,
                                               const float4* __restrict__ src8b) {
    walk_core<F, KAHAN, kFarList, STACK, DELAY, SEL>(a, group, split, nodes, far_list, src4, src8a, src8b);
}

template <uint32_t F, bool KAHAN, int FAR, int STACK = kWalkStack, bool DELAY = false, bool SEL = false>
__device__ __forceinline__ void walk_body(const ForceArgs& a, int64_t group, Split* split = nullptr) {
    if constexpr (FAR == kFarList)
        walk_list_core<F, KAHAN, STACK, DELAY, SEL>(a, group, split, a.nodes, a.far_list, a.src4, a.src8a, a.src8b);
    else
        walk_core<F, KAHAN, FAR, STACK, DELAY, SEL>(a, group, split, a.nodes, a.far_list, a.src4, a.src8a, a.src8b);
}

__device__ __forceinline__ int64_t walk_group() { return (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; }

template <uint32_t F, bool KAHAN>
// At least five waves per SIMD: the walks are latency-bound.  At 102 VGPRs the
// heaviest instantiations spill up to 24 bytes per lane and reload a few dwords
// inside the walk loop; the multipole walk still runs 10% faster at N = 2^20.
// The list walk was slower that way and stays at four.
__global__ __launch_bounds__(kWalkThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_walk(ForceArgs a) { walk_body<F, KAHAN, 0>(a, walk_group()); }
template <uint32_t F, bool KAHAN>
__global__ __launch_bounds__(kWalkThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_walk_multipole(ForceArgs a) {
    walk_body<F, KAHAN, kFarMultipole>(a, walk_group());
}
template <uint32_t F, bool KAHAN>
__global__ __launch_bounds__(kWalkThreads) void k_walk_list(ForceArgs a) {
    walk_body<F, KAHAN, kFarList>(a, walk_group());
}
template <uint32_t F, bool KAHAN>
__global__ __launch_bounds__(kWalkThreads) void k_walk_delay(ForceArgs a) {
    walk_body<F, KAHAN, 0, kWalkStack, true>(a, walk_group());
}

// Persistent walk: every wave pulls 64-particle groups from a queue until none
// are left, so uneven work (systems of a batch, dense regions) balances itself.
// Node-stack entries per wave: 64 or 96 (LDS per block = 4 waves x 12 B x entries).
// The tree and source pointers are restrict parameters, as in the listing walk:
// the result stores and queue atomics between groups then provably miss them,
// and the leaf-source loads of a wave-uniform address stay scalar loads.
template <uint32_t F, bool KAHAN, int STACK>
__device__ __forceinline__ void persistent_walks(const ForceArgs& a, int* queue, int64_t n_groups,
                                                 const Node* __restrict__ nodes, const float4* __restrict__ src4,
                                                 const float4* __restrict__ src8a, const float4* __restrict__ src8b) {
    for (;;) {
        int g = 0;
        if ((threadIdx.x & (kWave - 1)) == 0) g = atomicAdd(queue, 1);
        g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
        if (g >= n_groups) break;
        walk_core<F, KAHAN, 0, STACK>(a, g, nullptr, nodes, a.far_list, src4, src8a, src8b);
    }
}
template <uint32_t F, bool KAHAN>
__global__ __launch_bounds__(kWalkThreads) void k_walk_persistent64(ForceArgs a, int* queue, int64_t n_groups) {
    persistent_walks<F, KAHAN, 64>(a, queue, n_groups, a.nodes, a.src4, a.src8a, a.src8b);
}
template <uint32_t F, bool KAHAN>
__global__ __launch_bounds__(kWalkThreads) void k_walk_persistent96(ForceArgs a, int* queue, int64_t n_groups) {
    persistent_walks<F, KAHAN, 96>(a, queue, n_groups, a.nodes, a.src4, a.src8a, a.src8b);
}

// 1PN recompute for the velocity-Verlet correction: the same wavefront walk
// with _accum1PN at leaves and accepted nodes (forces.js:511-571); leaf sources
// from their past light cone when DELAY (542-546), else at the current time.
// F: 0 for the runtime-checked walk, or a fixed open-box set (its 1PN toggles as constants).
template <bool DELAY, uint32_t F = 0u>
__global__ __launch_bounds__(kWalkThreads) void k_walk1pn(ForceArgs a) {
    __shared__ int32_t s_node[kWavesPerBlock][kWalkStack];
    __shared__ uint64_t s_mask[kWavesPerBlock][kWalkStack];
    if (!a.ctrl->active) return;
    const int n = a.ctrl->n_alive;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int i = blockIdx.x * kWalkThreads + threadIdx.x;
    const bool valid = i < n;
    const uint64_t active = __ballot(valid);
    if (active == 0) return;
    const uint64_t my_bit = 1ull << lane;
    Obs o;
    if (valid) load_obs(a, i, o);
    else { o.x = o.y = o.vx = o.vy = o.m = o.q = o.ang_vel = o.mu = o.L = 0.f; o.ax = o.yuk = o.higgs = 1.f; o.tidal_ri5 = 0.f; o.inv_m = 0.f; }
    float f1x = 0.f, f1y = 0.f;
    const PhysFlags pf = a.pf;
    const float half_w = 0.5f * pf.dom_w, half_h = 0.5f * pf.dom_h;
    const double now = DELAY ? a.ctrl->sim_time : 0.0;
    int sp = 0;
    if (lane == 0) { s_node[wv][0] = 0; s_mask[wv][0] = active; }
    __builtin_amdgcn_wave_barrier();
    sp = 1;
    bool overflow = false;
    while (sp > 0) {
        sp--;
        const int node = __builtin_amdgcn_readfirstlane(s_node[wv][sp]);
        const uint64_t mask = s_mask[wv][sp];
        const bool mine = (mask & my_bit) != 0;
        if (node < 0 || node >= a.max_nodes) { overflow = true; continue; }
        const Node nd = a.nodes[node];
        if (nd.mass < kEpsMass) continue;
        if (nd.count > 0) {
            if (!DELAY && !a.pn_leaves) continue;
            for (int j = nd.child; j < nd.child + nd.count; j++) {
                const float4 s4 = a.src4[j];
                const float4 s8a = a.src8a[j];
                const float4 s8b = a.src8b[j];
                if (DELAY) {
                    Retarded rt;
                    if (mine && j != i && delayed_state(a.hist, pf, a.hist.slot[j], o.x, o.y, now, rt)) {
                        const float rx = rt.dx, ry = rt.dy;
                        accum_1pn<F>(o, f1x, f1y, pf, rx, ry, s4.z, s4.w, rt.vx, rt.vy, s8b.z, s8b.w);
                    }
                } else if (mine && j != i) {
                    float rx, ry;
                    min_image_f<F>(pf, o.x, o.y, s4.x, s4.y, rx, ry);
                    accum_1pn<F>(o, f1x, f1y, pf, rx, ry, s4.z, s4.w, s8a.x, s8a.y, s8b.z, s8b.w);
                }
            }
        } else if (nd.count < 0) {
            const float rawx = nd.comx - o.x, rawy = nd.comy - o.y;
            float dx, dy;
            min_image_f<F>(pf, o.x, o.y, nd.comx, nd.comy, dx, dy);
            bool accept = mine && (nd.size2 < pf.theta2 * (dx * dx + dy * dy));
            if (window_f<F>(pf)) {
                const float cx = nd.bx - o.x + (dx - rawx), cy = nd.by - o.y + (dy - rawy);
                accept = accept && (fabsf(cx) + nd.half <= half_w) && (fabsf(cy) + nd.half * (pf.dom_h / pf.dom_w) <= half_h);
            }
            if (accept) accum_1pn<F>(o, f1x, f1y, pf, dx, dy, nd.mass, nd.charge, nd.avx, nd.avy, 1.f, 1.f);
            const uint64_t open = __ballot(mine && !accept);
            if (open) {
                if (sp + 4 <= kWalkStack) {
                    if (lane == 0) {
                        s_node[wv][sp] = nd.child;     s_mask[wv][sp] = open;
                        s_node[wv][sp + 1] = nd.child + 1; s_mask[wv][sp + 1] = open;
                        s_node[wv][sp + 2] = nd.child + 2; s_mask[wv][sp + 2] = open;
                        s_node[wv][sp + 3] = nd.child + 3; s_mask[wv][sp + 3] = open;
                    }
                    __builtin_amdgcn_wave_barrier();
                    sp += 4;
                } else {
                    overflow = true;
                }
            }
        }
    }
    if (overflow && lane == 0) a.ctrl->overflow = 2;
    if (valid) { a.f1x[i] = f1x; a.f1y[i] = f1y; }
}

// O(N^2) 1PN recompute (forces.js:593-608); sources from their past light
// cone when DELAY (600-602), else at the current time.
template <bool DELAY>
__global__ __launch_bounds__(kWalkThreads) void k_direct1pn(ForceArgs a) {
    __shared__ float4 t4[kWalkThreads];
    __shared__ float4 t8a[kWalkThreads];
    __shared__ float4 t8b[kWalkThreads];
    __shared__ int32_t tslot[kWalkThreads];
    if (!a.ctrl->active) return;
    const int n = a.ctrl->n_alive;
    const int i = blockIdx.x * kWalkThreads + threadIdx.x;
    const bool valid = i < n;
    Obs o;
    if (valid) load_obs(a, i, o);
    else { o.x = o.y = o.vx = o.vy = o.m = o.q = o.ang_vel = o.mu = o.L = 0.f; o.ax = o.yuk = o.higgs = 1.f; o.tidal_ri5 = 0.f; o.inv_m = 0.f; }
    float f1x = 0.f, f1y = 0.f;
    const PhysFlags pf = a.pf;
    const float half_w = 0.5f * pf.dom_w, half_h = 0.5f * pf.dom_h;
    const double now = DELAY ? a.ctrl->sim_time : 0.0;
    const int n_src = (DELAY || a.pn_leaves) ? n : 0;     // uniform: every thread skips the tiles alike
    for (int base = 0; base < n_src; base += kWalkThreads) {
        const int j = base + threadIdx.x;
        __syncthreads();
        if (j < n) {
            t4[threadIdx.x] = a.src4[j]; t8a[threadIdx.x] = a.src8a[j]; t8b[threadIdx.x] = a.src8b[j];
            if (DELAY) tslot[threadIdx.x] = a.hist.slot[j];
        }
        __syncthreads();
        const int cnt = min(kWalkThreads, n - base);
        if (valid) {
            for (int k = 0; k < cnt; k++) {
                if (base + k == i) continue;
                float rx, ry, svx = t8a[k].x, svy = t8a[k].y;
                if (DELAY) {
                    Retarded rt;
                    if (!delayed_state(a.hist, pf, tslot[k], o.x, o.y, now, rt)) continue;
                    rx = rt.dx; ry = rt.dy; svx = rt.vx; svy = rt.vy;
                } else {
                    min_image(pf, o.x, o.y, t4[k].x, t4[k].y, rx, ry);
                }
                accum_1pn(o, f1x, f1y, pf, rx, ry, t4[k].z, t4[k].w, svx, svy, t8b[k].z, t8b[k].w);
            }
        }
    }
    if (valid) { a.f1x[i] = f1x; a.f1y[i] = f1y; }
}

void launch_walk1pn(hipStream_t s, const ForceArgs& a, bool tree) {
    if (a.n_upper <= 0) return;
    const int grid = (a.n_upper + kWalkThreads - 1) / kWalkThreads;
    // gravitomagnetism + 1PN in an open box, no magnetic, Yukawa or axion terms (config 3's set)
    const PhysFlags& pf = a.pf;
    if (tree && !a.hist.on && !pf.periodic && pf.gm && pf.one_pn && !pf.mag && !pf.yuk && !pf.ax_mod) {
        constexpr uint32_t kFix1PN = F_GRAV | F_EXTRA | F_OPEN | F_FIXED | F_GM | F_1PN;
        hipLaunchKernelGGL((k_walk1pn<false, kFix1PN>), dim3(grid), dim3(kWalkThreads), 0, s, a);
        return;
    }
    if (a.hist.on) {
        if (tree) hipLaunchKernelGGL(k_walk1pn<true>, dim3(grid), dim3(kWalkThreads), 0, s, a);
        else hipLaunchKernelGGL(k_direct1pn<true>, dim3(grid), dim3(kWalkThreads), 0, s, a);
    } else {
        if (tree) hipLaunchKernelGGL(k_walk1pn<false>, dim3(grid), dim3(kWalkThreads), 0, s, a);
        else hipLaunchKernelGGL(k_direct1pn<false>, dim3(grid), dim3(kWalkThreads), 0, s, a);
    }
}

// Collision broadphase (collisions.js:35-60 with quadtree.js:372-404): for each
// particle p1 a box query of half-size 2 r1 over the tree; a candidate p2 is a
// pair when slot1 < slot2 and the bodies overlap, dist^2 < (r1 + r2)^2.  The
// query rides the wavefront-shared stack of the walk, each entry carrying the
// lanes whose box still intersects the node.  On a torus the box is also
// queried at its wrapped images (the reference finds those neighbours through
// ghost particles, integrator.js:201-258).  WRITE = 0 counts, 1 writes at the
// lane's offset.  The tree, source and body arrays come in as restrict
// parameters, so the pair and flag stores in the loop provably miss them and
// the loads of a wave-uniform address stay scalar loads.
template <int WRITE>
__device__ __forceinline__ void pairs_query(const PairArgs& a, const Node* __restrict__ nodes,
                                            const float4* __restrict__ src4, const int32_t* __restrict__ slot,
                                            const float* __restrict__ radius) {
    __shared__ int32_t s_node[kWavesPerBlock][kWalkStack];
    __shared__ uint64_t s_mask[kWavesPerBlock][kWalkStack];
    const int n = a.ctrl->n_alive;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int i = blockIdx.x * kWalkThreads + threadIdx.x;
    const bool valid = i < n && a.mass[i] != 0.f;
    const uint64_t active = __ballot(valid);
    if (active == 0) {
        if (!WRITE && i < n) a.count[i] = 0;
        return;
    }
    const uint64_t my_bit = 1ull << lane;
    const float px = valid ? a.x[i] : 0.f, py = valid ? a.y[i] : 0.f;
    const float r1 = valid ? radius[i] * a.slack : 0.f;
    const int slot1 = valid ? slot[i] : 0;
    const float rw = 2.f * r1;
    int found = 0;
    int out = WRITE && valid ? a.offset[i] : 0;
    const int n_images = a.periodic ? 4 : 1;
    for (int img = 0; img < n_images; img++) {
        // image shift of the query box: (0,0), (sx,0), (0,sy), (sx,sy) toward the nearer edge
        float ox = 0.f, oy = 0.f;
        const float sx = px < 0.5f * a.dom_w ? a.dom_w : -a.dom_w;
        const float sy = py < 0.5f * a.dom_h ? a.dom_h : -a.dom_h;
        if (img & 1) ox = sx;
        if (img & 2) oy = sy;
        const float qx = px + ox, qy = py + oy;
        // lanes whose shifted box can touch the domain at all
        const bool need = valid && (img == 0 || ((img & 1) == 0 || qx - rw <= a.dom_w && qx + rw >= 0.f) &&
                                                ((img & 2) == 0 || qy - rw <= a.dom_h && qy + rw >= 0.f));
        const uint64_t start = __ballot(need);
        if (start == 0) continue;
        int sp = 0;
        if (lane == 0) { s_node[wv][0] = 0; s_mask[wv][0] = start; }
        __builtin_amdgcn_wave_barrier();
        sp = 1;
        while (sp > 0) {
            sp--;
            const int node = __builtin_amdgcn_readfirstlane(s_node[wv][sp]);
            const uint64_t mask = s_mask[wv][sp];
            if (node < 0 || node >= a.max_nodes) continue;
            const Node nd = nodes[node];
            const float bw = nd.half, bh = nd.half * (a.dom_h / a.dom_w);
            const bool mine = (mask & my_bit) != 0 &&
                              !(qx - rw > nd.bx + bw || qx + rw < nd.bx - bw || qy - rw > nd.by + bh ||
                                qy + rw < nd.by - bh);
            if (nd.count >= 0) {
                for (int j = nd.child; j < nd.child + nd.count; j++) {
                    const float4 s4 = src4[j];
                    if (!mine) continue;
                    if (!(s4.x >= qx - rw && s4.x <= qx + rw && s4.y >= qy - rw && s4.y <= qy + rw)) continue;
                    const int slot2 = slot[j];
                    if (j == i || s4.z == 0.f || slot1 >= slot2) continue;
                    float dx = s4.x - px, dy = s4.y - py;
                    if (a.periodic) { dx = torus_wrap(dx, a.dom_w, 0.5f * a.dom_w); dy = torus_wrap(dy, a.dom_h, 0.5f * a.dom_h); }
                    const float md = r1 + radius[j] * a.slack;
                    if (dx * dx + dy * dy < md * md) {
                        if (WRITE) {
                            if (out < a.cap_pairs) {
                                a.pa[out] = a.index_out ? i : slot1;
                                a.pb[out] = a.index_out ? j : slot2;
                                if (a.pl) a.pl[out] = nd.child;
                            } else {
                                a.ctrl->overflow = 4;
                            }
                            out++;
                        }
                        found++;
                    }
                }
            } else {
                const uint64_t open = __ballot(mine);
                if (open && sp + 4 <= kWalkStack) {
                    if (lane == 0) {
                        s_node[wv][sp] = nd.child;     s_mask[wv][sp] = open;
                        s_node[wv][sp + 1] = nd.child + 1; s_mask[wv][sp + 1] = open;
                        s_node[wv][sp + 2] = nd.child + 2; s_mask[wv][sp + 2] = open;
                        s_node[wv][sp + 3] = nd.child + 3; s_mask[wv][sp + 3] = open;
                    }
                    __builtin_amdgcn_wave_barrier();
                    sp += 4;
                } else if (open) {
                    if (lane == 0) a.ctrl->overflow = 2;
                }
            }
        }
    }
    if (!WRITE && i < n) a.count[i] = valid ? found : 0;
}

template <int WRITE>
__global__ __launch_bounds__(kWalkThreads) void k_pairs(PairArgs a) {
    pairs_query<WRITE>(a, a.nodes, a.src4, a.slot, a.radius);
}

void launch_pairs(hipStream_t s, const PairArgs& a, bool write) {
    if (a.n_upper <= 0) return;
    const int grid = (a.n_upper + kWalkThreads - 1) / kWalkThreads;
    if (write) hipLaunchKernelGGL(k_pairs<1>, dim3(grid), dim3(kWalkThreads), 0, s, a);
    else hipLaunchKernelGGL(k_pairs<0>, dim3(grid), dim3(kWalkThreads), 0, s, a);
}

// Direct sum (forces.js:93-126): each block stages 256 sources at a time in
// LDS.  DELAY: every source from its past light cone, skipped without history
// (forces.js:99-110).
template <uint32_t F, bool KAHAN, bool DELAY>
__device__ __forceinline__ void direct_body(const ForceArgs& a) {
    __shared__ float4 t4[kWalkThreads];
    __shared__ float4 t8a[kWalkThreads];
    __shared__ float4 t8b[kWalkThreads];
    __shared__ int32_t tslot[kWalkThreads];
    __shared__ float tr2[kWalkThreads];
    const int n = a.ctrl->n_alive;
    const int i = blockIdx.x * kWalkThreads + threadIdx.x;
    const bool valid = i < n;
    Obs o;
    if (valid) load_obs(a, i, o);
    else { o.x = o.y = o.vx = o.vy = o.m = o.q = o.ang_vel = o.mu = o.L = 0.f; o.ax = o.yuk = o.higgs = 1.f; o.tidal_ri5 = 0.f; o.inv_m = 0.f; }
    Acc<KAHAN> acc;
    acc.init();
    const PhysFlags pf = a.pf;
    const float half_w = 0.5f * pf.dom_w, half_h = 0.5f * pf.dom_h;
    const double now = DELAY ? a.ctrl->sim_time : 0.0;
    int64_t c_pair = 0;
    for (int base = 0; base < n; base += kWalkThreads) {
        const int j = base + threadIdx.x;
        __syncthreads();
        if (j < n) {
            t4[threadIdx.x] = a.src4[j];
            t8a[threadIdx.x] = a.src8a[j];
            if ((F & F_EXTRA) || DELAY) t8b[threadIdx.x] = a.src8b[j];
            if (DELAY) {
                tslot[threadIdx.x] = a.hist.slot[j];
                const float rad = a.radius[j];
                tr2[threadIdx.x] = rad * rad;
            }
        }
        __syncthreads();
        const int cnt = min(kWalkThreads, n - base);
        if (valid) {
            for (int k = 0; k < cnt; k++) {
                if (base + k == i) continue;
                const float4 s4 = t4[k];
                const float4 s8a = t8a[k];
                float4 s8b = make_float4(0.f, 1.f, 1.f, 1.f);
                if ((F & F_EXTRA) || DELAY) s8b = t8b[k];
                if (DELAY) {
                    Retarded rt;
                    if (!delayed_state(a.hist, pf, tslot[k], o.x, o.y, now, rt)) continue;
                    const float rx = rt.dx, ry = rt.dy;
                    const float r2 = tr2[k];
                    const float av = retarded_ang_vel(rt.aw, r2);
                    pair_force<F, KAHAN, true>(o, acc, pf, rx, ry, s4.z, s4.w, rt.vx, rt.vy, av,
                                               0.2f * s4.w * av * r2, 0.4f * s4.z * av * r2, s8b.y, s8b.z, s8b.w);
                    c_pair++;
                } else {
                    float rx, ry;
                    min_image(pf, o.x, o.y, s4.x, s4.y, rx, ry);
                    pair_force<F, KAHAN>(o, acc, pf, rx, ry, s4.z, s4.w, s8a.x, s8a.y, s8a.z, s8a.w, s8b.x, s8b.y,
                                         s8b.z, s8b.w);
                }
            }
        }
    }
    if (DELAY) {
        wave_add_counters(a.ctrl, c_pair, 0, 0);
    } else if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t nn = n;
        atomicAdd((unsigned long long*)&a.ctrl->cnt_pair, (unsigned long long)(nn * (nn - 1)));
    }
    if (valid) store_acc<F, KAHAN>(a, i, o, acc);
}

template <uint32_t F, bool KAHAN>
__global__ __launch_bounds__(kWalkThreads) void k_direct(ForceArgs a) { direct_body<F, KAHAN, false>(a); }
template <uint32_t F, bool KAHAN>
__global__ __launch_bounds__(kWalkThreads) void k_direct_delay(ForceArgs a) { direct_body<F, KAHAN, true>(a); }

// Retired sources (forces.js:128-153): every live particle against every dead
// record, each from its past light cone with its death mass, no PE; added to
// the force outputs the walk or direct sum wrote.
template <uint32_t F, bool KAHAN>
__global__ __launch_bounds__(kWalkThreads) void k_dead_forces(ForceArgs a) {
    const int n = a.ctrl->n_alive;
    const int i = blockIdx.x * kWalkThreads + threadIdx.x;
    const bool valid = i < n;
    if (!__ballot(valid)) return;
    Obs o;
    if (valid) load_obs(a, i, o);
    else { o.x = o.y = o.vx = o.vy = o.m = o.q = o.ang_vel = o.mu = o.L = 0.f; o.ax = o.yuk = o.higgs = 1.f; o.tidal_ri5 = 0.f; o.inv_m = 0.f; }
    Acc<false> acc;
    acc.init();
    const PhysFlags pf = a.pf;
    const float half_w = 0.5f * pf.dom_w, half_h = 0.5f * pf.dom_h;
    const double now = a.ctrl->sim_time;
    int64_t c_pair = 0;
    for (int d = 0; d < a.hist.n_dead; d++) {
        const DeadRec r = a.hist.dead[d];
        Retarded rt;
        if (!valid || !delayed_state(a.hist, pf, r.slot, o.x, o.y, now, rt)) continue;
        const float rx = rt.dx, ry = rt.dy;
        const float av = retarded_ang_vel(rt.aw, r.r2);
        pair_force<F, false, true, false>(o, acc, pf, rx, ry, r.mass, r.charge, rt.vx, rt.vy, av,
                                          0.2f * r.charge * av * r.r2, 0.4f * r.mass * av * r.r2,
                                          r.ax != 0.f ? r.ax : 1.f, r.yuk != 0.f ? r.yuk : 1.f,
                                          r.higgs != 0.f ? r.higgs : 1.f);
        c_pair++;
    }
    wave_add_counters(a.ctrl, c_pair, 0, 0);
    if (!valid) return;
    a.fx[i] += acc.fx.s; a.fy[i] += acc.fy.s;
    a.bz[i] += acc.bz; a.dbzdx[i] += acc.dbzdx; a.dbzdy[i] += acc.dbzdy;
    a.bgz[i] += acc.bgz; a.dbgzdx[i] += acc.dbgzdx; a.dbgzdy[i] += acc.dbgzdy;
    a.frame_drag[i] += acc.frame_drag; a.tidal[i] += acc.tidal;
    a.f1x[i] += acc.f1x; a.f1y[i] += acc.f1y;
    if (F & F_JERK) { a.jx[i] += acc.jx; a.jy[i] += acc.jy; }
}

// ---- The selected particle ------------------------------------------------
// readStats' `selected` (gpu-physics.js:4238-4276) shows one particle's force by
// type, which the reference keeps for every particle (forces.js:36-46).  Here
// only the selected particle's are formed: after the pair forces of a pass, one
// wave repeats that particle's walk (or direct sum) and its dead sources with
// every term runtime-gated and summed by type.

__global__ void k_sel_clear(int32_t* idx) { *idx = -1; }

__global__ void k_sel_find(const int32_t* __restrict__ slot, int32_t target, int32_t* __restrict__ idx,
                           const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ctrl->n_alive && slot[i] == target) *idx = i;
}

void launch_sel_find(hipStream_t s, const int32_t* slot, int32_t target, int32_t* idx, int n_upper, const Ctrl* ctrl) {
    hipLaunchKernelGGL(k_sel_clear, dim3(1), dim3(1), 0, s, idx);
    if (n_upper > 0)
        hipLaunchKernelGGL(k_sel_find, dim3((n_upper + 255) / 256), dim3(256), 0, s, slot, target, idx, ctrl);
}

constexpr uint32_t kSelF = F_GRAV | F_COUL | F_EXTRA | F_JERK;   // every term, each gated by the flags

// One source j of the direct sum on the selected observer (direct_body's body).
template <bool DELAY>
__device__ __forceinline__ void sel_direct_source(const ForceArgs& a, const Obs& o, Acc<false>& acc, int j,
                                                  double now, Split* sp) {
    const PhysFlags& pf = a.pf;
    const float4 s4 = a.src4[j];
    const float4 s8a = a.src8a[j];
    const float4 s8b = a.src8b[j];
    if (DELAY) {
        Retarded rt;
        if (!delayed_state(a.hist, pf, a.hist.slot[j], o.x, o.y, now, rt)) return;
        const float rad = a.radius[j];
        const float r2 = rad * rad;
        const float av = retarded_ang_vel(rt.aw, r2);
        pair_force<kSelF, false, true, true, true>(o, acc, pf, rt.dx, rt.dy, s4.z, s4.w, rt.vx, rt.vy, av,
                                                   0.2f * s4.w * av * r2, 0.4f * s4.z * av * r2, s8b.y, s8b.z, s8b.w,
                                                   sp);
    } else {
        float rx, ry;
        min_image(pf, o.x, o.y, s4.x, s4.y, rx, ry);
        pair_force<kSelF, false, false, true, true>(o, acc, pf, rx, ry, s4.z, s4.w, s8a.x, s8a.y, s8a.z, s8a.w,
                                                    s8b.x, s8b.y, s8b.z, s8b.w, sp);
    }
}

template <int FAR, bool DELAY, bool DIRECT>
__global__ __launch_bounds__(kWave) void k_sel_forces(ForceArgs a, int residual) {
    const int n = a.ctrl->n_alive;
    const int si = *a.sel;
    if (si < 0 || si >= n) return;
    const int lane = threadIdx.x;
    Split sp;
    for (int t = 0; t < kSplitTypes; t++) sp.f[t][0] = sp.f[t][1] = 0.f;
    Obs o;
    load_obs(a, si, o);
    Acc<false> acc;
    acc.init();
    const double now = a.ctrl->sim_time;
    if (DIRECT) {
        for (int j = lane; j < n; j += kWave)
            if (j != si) sel_direct_source<DELAY>(a, o, acc, j, now, &sp);
    } else {
        walk_body<kSelF, false, FAR, kWalkStack, DELAY, true>(a, si / kWave, &sp);
    }
    // retired sources (k_dead_forces), shared over the lanes
    for (int d = lane; d < a.hist.n_dead; d += kWave) {
        const DeadRec r = a.hist.dead[d];
        Retarded rt;
        if (!delayed_state(a.hist, a.pf, r.slot, o.x, o.y, now, rt)) continue;
        const float av = retarded_ang_vel(rt.aw, r.r2);
        pair_force<kSelF, false, true, false, true>(o, acc, a.pf, rt.dx, rt.dy, r.mass, r.charge, rt.vx, rt.vy, av,
                                              0.2f * r.charge * av * r.r2, 0.4f * r.mass * av * r.r2,
                                              r.ax != 0.f ? r.ax : 1.f, r.yuk != 0.f ? r.yuk : 1.f,
                                              r.higgs != 0.f ? r.higgs : 1.f, &sp);
    }
    float sum[2] = {0.f, 0.f};
    for (int t = 0; t < kSplitTypes; t++)
        for (int c = 0; c < 2; c++) {
            float v = sp.f[t][c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            sp.f[t][c] = v;
            sum[c] += v;
        }
    if (lane != 0) return;
    float* out = a.sel_out;
    for (int t = 0; t < kSplitTypes; t++) { out[SEL_SPLIT + 2 * t] = sp.f[t][0]; out[SEL_SPLIT + 2 * t + 1] = sp.f[t][1]; }
    const float ex = o.m * a.ext_gx + o.q * a.ext_ex, ey = o.m * a.ext_gy + o.q * a.ext_ey;
    out[SEL_EXT] = ex;
    out[SEL_EXT + 1] = ey;
    // the learned correction: what the pass added beyond the terms above
    out[SEL_FAR] = residual ? a.fx[si] - ex - sum[0] : 0.f;
    out[SEL_FAR + 1] = residual ? a.fy[si] - ey - sum[1] : 0.f;
    out[SEL_VALID] = 1.f;
}

#define GEON_FEAT_SWITCH(KERNEL, feat, kahan, grid, stream, args)                                     \
    do {                                                                                              \
        const dim3 g_(grid), b_(kWalkThreads);                                                        \
        switch ((feat) & 7u) {                                                                        \
            case 1: if (kahan) hipLaunchKernelGGL((KERNEL<1, true>), g_, b_, 0, stream, args);        \
                    else hipLaunchKernelGGL((KERNEL<1, false>), g_, b_, 0, stream, args); break;      \
            case 2: if (kahan) hipLaunchKernelGGL((KERNEL<2, true>), g_, b_, 0, stream, args);        \
                    else hipLaunchKernelGGL((KERNEL<2, false>), g_, b_, 0, stream, args); break;      \
            case 3: if (kahan) hipLaunchKernelGGL((KERNEL<3, true>), g_, b_, 0, stream, args);        \
                    else hipLaunchKernelGGL((KERNEL<3, false>), g_, b_, 0, stream, args); break;      \
            case 4: if (kahan) hipLaunchKernelGGL((KERNEL<4, true>), g_, b_, 0, stream, args);        \
                    else hipLaunchKernelGGL((KERNEL<4, false>), g_, b_, 0, stream, args); break;      \
            case 5: if (kahan) hipLaunchKernelGGL((KERNEL<5, true>), g_, b_, 0, stream, args);        \
                    else hipLaunchKernelGGL((KERNEL<5, false>), g_, b_, 0, stream, args); break;      \
            case 6: if (kahan) hipLaunchKernelGGL((KERNEL<6, true>), g_, b_, 0, stream, args);        \
                    else hipLaunchKernelGGL((KERNEL<6, false>), g_, b_, 0, stream, args); break;      \
            case 7: if (kahan) hipLaunchKernelGGL((KERNEL<7, true>), g_, b_, 0, stream, args);        \
                    else hipLaunchKernelGGL((KERNEL<7, false>), g_, b_, 0, stream, args); break;      \
            default: if (kahan) hipLaunchKernelGGL((KERNEL<0, true>), g_, b_, 0, stream, args);       \
                    else hipLaunchKernelGGL((KERNEL<0, false>), g_, b_, 0, stream, args); break;      \
        }                                                                                             \
    } while (0)

// As GEON_FEAT_SWITCH, with the radiation variants (gravity and Coulomb
// runtime-checked plus the jerk: 11 without, 15 with the extra forces) for the
// kernels that run with radiation on.
#define GEON_FEAT_SWITCH_R(KERNEL, feat, kahan, grid, stream, args)                                   \
    do {                                                                                              \
        if (((feat) & F_JERK) && ((feat) & F_EXTRA)) {                                                \
            const dim3 g_(grid), b_(kWalkThreads);                                                    \
            if (kahan) hipLaunchKernelGGL((KERNEL<15, true>), g_, b_, 0, stream, args);                \
            else hipLaunchKernelGGL((KERNEL<15, false>), g_, b_, 0, stream, args);                     \
        } else if ((feat) & F_JERK) {                                                                 \
            const dim3 g_(grid), b_(kWalkThreads);                                                    \
            if (kahan) hipLaunchKernelGGL((KERNEL<11, true>), g_, b_, 0, stream, args);                \
            else hipLaunchKernelGGL((KERNEL<11, false>), g_, b_, 0, stream, args);                     \
        } else {                                                                                      \
            GEON_FEAT_SWITCH(KERNEL, feat, kahan, grid, stream, args);                                \
        }                                                                                             \
    } while (0)

// The fixed instantiation that runs this launch's feature set, or 0 when none
// does and the runtime-checked kernel runs.  In an open box:
//   kFixG     gravity;
//   kFixGC    gravity + Coulomb;
//   kFixGCMY  gravity + Coulomb + magnetic + Yukawa, no axion modulation or
//             Higgs-scaled Yukawa range (the N = 65536 configuration's set);
//   kFixGCM   gravity + Coulomb + magnetic, no axion modulation;
//   kFixGP    gravity + gravitomagnetism + 1PN, no Yukawa or axion-modulated
//             charges: the plain walk only.
// (The Higgs field only scales the Yukawa range, so without Yukawa it may be
// on.)  On a torus (topology 0), minimum image or windowed: the gravity +
// Coulomb and configuration-2 sets.  The signal-delay walk has the open-box
// gravity + Coulomb (+ magnetic) ones.
constexpr uint32_t kFixG = F_GRAV | F_OPEN;
constexpr uint32_t kFixGC = F_GRAV | F_COUL | F_OPEN;
constexpr uint32_t kFixGCMY = F_GRAV | F_COUL | F_EXTRA | F_OPEN | F_FIXED | F_MAG | F_YUK;
constexpr uint32_t kFixGCM = F_GRAV | F_COUL | F_EXTRA | F_OPEN | F_FIXED | F_MAG;
constexpr uint32_t kFixGP = F_GRAV | F_EXTRA | F_OPEN | F_FIXED | F_GM | F_1PN;
constexpr uint32_t kFixGCJ = F_GRAV | F_COUL | F_JERK | F_OPEN | F_FIXED;   // radiation: gravity + Coulomb + jerk
constexpr uint32_t kFixGCT = (kFixGC & ~F_OPEN) | F_TORUS;
constexpr uint32_t kFixGCW = kFixGCT | F_WINDOW;
constexpr uint32_t kFixGCMYT = (kFixGCMY & ~F_OPEN) | F_TORUS;
constexpr uint32_t kFixGCMYW = kFixGCMYT | F_WINDOW;
static uint32_t fixed_set(const PhysFlags& pf, uint32_t feat) {
    switch (feat & 7u) {
        case F_GRAV: return kFixG;
        case F_GRAV | F_COUL: return kFixGC;
        case F_GRAV | F_COUL | F_EXTRA:
            if (!pf.mag || pf.gm || pf.one_pn || pf.ax_mod) return 0;
            if (!pf.yuk) return kFixGCM;
            return pf.higgs ? 0 : kFixGCMY;
        case F_GRAV | F_EXTRA:
            return pf.gm && pf.one_pn && !pf.mag && !pf.yuk && !pf.ax_mod ? kFixGP : 0;
        default: return 0;
    }
}
static uint32_t fixed_feat(const PhysFlags& pf, uint32_t feat, bool kahan) {
    if (!kahan) return 0;
    if (feat & F_JERK)   // the radiation walks: gravity and Coulomb both on, no extra force, open box
        return feat == (F_GRAV | F_COUL | F_JERK) && pf.grav_on && pf.coul_on && !pf.periodic ? kFixGCJ : 0;
    const uint32_t set = fixed_set(pf, feat);
    if (!pf.periodic) return set;
    if (pf.topology != 0 || (set != kFixGC && set != kFixGCMY)) return 0;
    return (set & ~F_OPEN) | F_TORUS | (pf.window ? F_WINDOW : 0u);
}
#define GEON_FIXED_LAUNCH(KERNEL, fix, grid, stream, args)                                                  \
    do {                                                                                                \
        const dim3 g_(grid), b_(kWalkThreads);                                                          \
        switch (fix) {                                                                                  \
            case kFixG: hipLaunchKernelGGL((KERNEL<kFixG, true>), g_, b_, 0, stream, args); break;        \
            case kFixGC: hipLaunchKernelGGL((KERNEL<kFixGC, true>), g_, b_, 0, stream, args); break;      \
            case kFixGCM: hipLaunchKernelGGL((KERNEL<kFixGCM, true>), g_, b_, 0, stream, args); break;    \
            case kFixGCMY: hipLaunchKernelGGL((KERNEL<kFixGCMY, true>), g_, b_, 0, stream, args); break;  \
            case kFixGCT: hipLaunchKernelGGL((KERNEL<kFixGCT, true>), g_, b_, 0, stream, args); break;    \
            case kFixGCW: hipLaunchKernelGGL((KERNEL<kFixGCW, true>), g_, b_, 0, stream, args); break;    \
            case kFixGCMYT: hipLaunchKernelGGL((KERNEL<kFixGCMYT, true>), g_, b_, 0, stream, args); break; \
            default: hipLaunchKernelGGL((KERNEL<kFixGCMYW, true>), g_, b_, 0, stream, args); break;       \
        }                                                                                               \
    } while (0)

void launch_walk(hipStream_t s, const ForceArgs& a, uint32_t feat, bool kahan) {
    if (a.n_upper <= 0) return;
    const int grid = (a.n_upper + kWalkThreads - 1) / kWalkThreads;
    const uint32_t fix = fixed_feat(a.pf, feat, kahan);
    if (a.hist.on) {
        const dim3 g_(grid), b_(kWalkThreads);
        if (fix == kFixGC) hipLaunchKernelGGL((k_walk_delay<kFixGC, true>), g_, b_, 0, s, a);
        else if (fix == kFixGCM) hipLaunchKernelGGL((k_walk_delay<kFixGCM, true>), g_, b_, 0, s, a);
        else GEON_FEAT_SWITCH_R(k_walk_delay, feat, kahan, grid, s, a);
    } else if (a.far_mode == kFarMultipole) {
        if (fix && fix != kFixGP && fix != kFixGCJ) GEON_FIXED_LAUNCH(k_walk_multipole, fix, grid, s, a);
        else GEON_FEAT_SWITCH(k_walk_multipole, feat, kahan, grid, s, a);
    } else if (a.far_mode == kFarList) {
        if (fix && fix != kFixGP && fix != kFixGCJ) GEON_FIXED_LAUNCH(k_walk_list, fix, grid, s, a);
        else GEON_FEAT_SWITCH(k_walk_list, feat, kahan, grid, s, a);
    } else if (fix == kFixGP) {
        hipLaunchKernelGGL((k_walk<kFixGP, true>), dim3(grid), dim3(kWalkThreads), 0, s, a);
    } else if (fix == kFixGCJ) {
        hipLaunchKernelGGL((k_walk<kFixGCJ, true>), dim3(grid), dim3(kWalkThreads), 0, s, a);
    } else if (fix) GEON_FIXED_LAUNCH(k_walk, fix, grid, s, a);
    else GEON_FEAT_SWITCH_R(k_walk, feat, kahan, grid, s, a);
}

#define GEON_FEAT_SWITCH_P(KERNEL, feat, kahan, grid, threads, stream, ...)                              \
    do {                                                                                              \
        const dim3 g_(grid), b_(threads);                                                             \
        switch ((feat) & 7u) {                                                                        \
            case 1: if (kahan) hipLaunchKernelGGL((KERNEL<1, true>), g_, b_, 0, stream, __VA_ARGS__); \
                    else hipLaunchKernelGGL((KERNEL<1, false>), g_, b_, 0, stream, __VA_ARGS__); break; \
            case 2: if (kahan) hipLaunchKernelGGL((KERNEL<2, true>), g_, b_, 0, stream, __VA_ARGS__); \
                    else hipLaunchKernelGGL((KERNEL<2, false>), g_, b_, 0, stream, __VA_ARGS__); break; \
            case 3: if (kahan) hipLaunchKernelGGL((KERNEL<3, true>), g_, b_, 0, stream, __VA_ARGS__); \
                    else hipLaunchKernelGGL((KERNEL<3, false>), g_, b_, 0, stream, __VA_ARGS__); break; \
            default: if (kahan) hipLaunchKernelGGL((KERNEL<7, true>), g_, b_, 0, stream, __VA_ARGS__); \
                    else hipLaunchKernelGGL((KERNEL<7, false>), g_, b_, 0, stream, __VA_ARGS__); break; \
        }                                                                                             \
    } while (0)

void launch_walk_persistent(hipStream_t s, const ForceArgs& a, uint32_t feat, bool kahan, int* queue,
                            int64_t n_groups, int grid, int threads, int stack) {
    if (n_groups <= 0) return;
    const uint32_t fix = fixed_feat(a.pf, feat, kahan);
    const dim3 g_(grid), b_(threads);
    if (fix == kFixGC && stack <= 64)
        hipLaunchKernelGGL((k_walk_persistent64<kFixGC, true>), g_, b_, 0, s, a, queue, n_groups);
    else if (fix == kFixGC)
        hipLaunchKernelGGL((k_walk_persistent96<kFixGC, true>), g_, b_, 0, s, a, queue, n_groups);
    else if (stack <= 64) GEON_FEAT_SWITCH_P(k_walk_persistent64, feat, kahan, grid, threads, s, a, queue, n_groups);
    else GEON_FEAT_SWITCH_P(k_walk_persistent96, feat, kahan, grid, threads, s, a, queue, n_groups);
}

void launch_direct(hipStream_t s, const ForceArgs& a, uint32_t feat, bool kahan) {
    if (a.n_upper <= 0) return;
    const int grid = (a.n_upper + kWalkThreads - 1) / kWalkThreads;
    if (a.hist.on) GEON_FEAT_SWITCH_R(k_direct_delay, feat, kahan, grid, s, a);
    else GEON_FEAT_SWITCH_R(k_direct, feat, kahan, grid, s, a);
}

void launch_dead_forces(hipStream_t s, const ForceArgs& a, uint32_t feat) {
    if (a.n_upper <= 0 || a.hist.n_dead <= 0) return;
    const int grid = (a.n_upper + kWalkThreads - 1) / kWalkThreads;
    GEON_FEAT_SWITCH_R(k_dead_forces, feat, false, grid, s, a);
}

void launch_sel_forces(hipStream_t s, const ForceArgs& a, bool bh, bool residual) {
    if (a.n_upper <= 0 || !a.sel || !a.sel_out) return;
    const dim3 g(1), b(kWave);
    const int r = residual ? 1 : 0;
    if (!bh) {
        if (a.hist.on) hipLaunchKernelGGL((k_sel_forces<0, true, true>), g, b, 0, s, a, r);
        else hipLaunchKernelGGL((k_sel_forces<0, false, true>), g, b, 0, s, a, r);
    } else if (a.hist.on) {
        hipLaunchKernelGGL((k_sel_forces<0, true, false>), g, b, 0, s, a, r);
    } else if (a.far_mode == kFarMultipole) {
        hipLaunchKernelGGL((k_sel_forces<kFarMultipole, false, false>), g, b, 0, s, a, r);
    } else {
        hipLaunchKernelGGL((k_sel_forces<0, false, false>), g, b, 0, s, a, r);
    }
}

}  // namespace geon
