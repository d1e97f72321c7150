// k_merge.hip — collision mode "merge" on gfx950: handleCollisions and
// resolveMerge (collisions.js:24-158) over the broadphase pair list.
//
// The reference folds over its overlap pairs one at a time: particles in
// array order, each one's candidates in the order its tree query meets them
// (quadtree.js:372-398), with masses and radii changing as it goes.
// Positions, velocities, spins and charges do not change during the fold, and
// radii only shrink (annihilation), so the pair list built from the radii at
// the start is a superset of the pairs the fold acts on; each pair is tested
// again with the state at its turn.  Pairs whose two particles take part in no
// other pair cannot see any other pair's effect: they resolve in parallel.
// The rest are folded by one thread in the reference's order, which the pair
// sort key reproduces:
//   slot of the first member (array order = creation order),
//   leaf position in the query's traversal (children pushed NW..SE, so SE
//     first: descending Morton order of the leaves),
//   the candidate's place in its leaf's stored order.
// That stored order follows the reference's iterative insert
// (quadtree.js:203-231): a node that divides on its fifth arrival passes the
// five on newest first, later arrivals in arrival order.  With the five
// smallest slots of every divided node (k_s5) a walk from the root gives any
// leaf's order (k_leaf_rank); it is computed only for leaves holding two
// candidates of one particle, the only case where it matters.
#include "geon_kernels.h"

// The fold's FP64 arithmetic follows the reference's operation order with no
// fused multiply-adds, so decisions such as "kinetic energy lost > 0" come out
// as in IEEE double evaluation of the same expressions: a particle that
// already merged merges again with exactly zero lost energy there.
#pragma clang fp contract(off)

namespace geon {

// ---- pair flags and leaf ranks ---------------------------------------------

// need[0] = 1 when some particle has two candidates in one leaf (consecutive
// in its query-ordered run of pairs).
__global__ void k_pair_same_leaf(const int32_t* __restrict__ p1, const int32_t* __restrict__ leaf,
                                 const Ctrl* __restrict__ mc, int32_t* __restrict__ need) {
    const int np = mc->n_alive;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k + 1 >= np) return;
    if (p1[k] == p1[k + 1] && leaf[k] == leaf[k + 1]) need[0] = 1;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off);
        const uint64_t o = ((uint64_t)hi << 32) | lo;
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ void insert5(uint64_t* best, uint64_t v) {
#pragma unroll
    for (int q = 0; q < 5; q++) {               // insertion into a sorted five
        const uint64_t lo = v < best[q] ? v : best[q];
        v = v < best[q] ? best[q] : v;
        best[q] = lo;
    }
}

// Pops the wave's five smallest keys from the lanes' sorted fives, ascending.
__device__ __forceinline__ void wave_pop5(uint64_t* best, uint64_t* out) {
    for (int q = 0; q < 5; q++) {
        const uint64_t m = wave_min_u64(best[0]);
        if (best[0] == m) {                     // keys are unique: one lane pops
#pragma unroll
            for (int t = 0; t < 4; t++) best[t] = best[t + 1];
            best[4] = ~0ull;
        }
        out[q] = m;
    }
}

constexpr int kChunk = 256;                      // sorted points per chunk of k_chunk5

__device__ __forceinline__ uint64_t slot_key(const int32_t* slot, int j) {
    return ((uint64_t)(uint32_t)slot[j] << 32) | (uint32_t)j;
}

// chunk5[c] = the five smallest (slot, sorted index) keys of sorted points
// [c * kChunk, (c + 1) * kChunk), ascending, ~0 past the end: one wave per chunk.
__global__ __launch_bounds__(256) void k_chunk5(const int32_t* __restrict__ slot, const Ctrl* __restrict__ ctrl,
                                                int n_chunks, uint64_t* __restrict__ chunk5) {
    const int lane = threadIdx.x & (kWave - 1);
    const int c = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (c >= n_chunks) return;
    const int n = ctrl->n_alive;
    uint64_t best[5] = {~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
#pragma unroll
    for (int t = 0; t < kChunk / kWave; t++) {
        const int j = c * kChunk + t * kWave + lane;
        if (j < n) insert5(best, slot_key(slot, j));
    }
    uint64_t out[5];
    wave_pop5(best, out);
    if (lane == 0)
        for (int q = 0; q < 5; q++) chunk5[(size_t)c * 5 + q] = out[q];
}

// The five smallest slots of the marked divided nodes' points, as sorted
// indices in ascending slot order: one wave per node, over the fives of the
// whole chunks inside the node's range and the points of the partial ends.
__global__ __launch_bounds__(256) void k_s5(const Node* __restrict__ nodes, const int32_t* __restrict__ first,
                                            const int32_t* __restrict__ npts, const int32_t* __restrict__ slot,
                                            const uint64_t* __restrict__ chunk5, int n_nodes,
                                            const int32_t* __restrict__ mark, int32_t* __restrict__ s5) {
    const int lane = threadIdx.x & (kWave - 1);
    const int r = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (r >= n_nodes) return;
    if (!mark[r] || nodes[r].count >= 0) return;   // only the divided nodes on a ranked leaf's path
    const int a = first[r], e = a + npts[r];
    const int cf = (a + kChunk - 1) / kChunk, cl = e / kChunk;     // whole chunks [cf, cl)
    uint64_t best[5] = {~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
    if (cf < cl) {
        for (int j = a + lane; j < cf * kChunk; j += kWave) insert5(best, slot_key(slot, j));
        for (int j = cl * kChunk + lane; j < e; j += kWave) insert5(best, slot_key(slot, j));
        const int64_t nk = (int64_t)(cl - cf) * 5;
        for (int64_t u = lane; u < nk; u += kWave) insert5(best, chunk5[(size_t)cf * 5 + u]);
    } else {
        for (int j = a + lane; j < e; j += kWave) insert5(best, slot_key(slot, j));
    }
    uint64_t out[5];
    wave_pop5(best, out);
    if (lane == 0)
        for (int q = 0; q < 5; q++) s5[(size_t)r * 5 + q] = out[q] == ~0ull ? -1 : (int32_t)(uint32_t)out[q];
}

// Child of divided node `node` whose point range holds sorted index `target`, or -1.
__device__ __forceinline__ int child_holding(const Node* nodes, const int32_t* first, const int32_t* npts, int node,
                                             int target) {
    const int cb = nodes[node].child;
    for (int q = 0; q < 4; q++) {
        const int a = first[cb + q];
        if (target >= a && target < a + npts[cb + q]) return cb + q;
    }
    return -1;
}

// mark[node] = 1 for every divided node on the path from the root to a leaf
// that k_leaf_rank will order (the same run heads).
__global__ void k_mark_paths(const int32_t* __restrict__ p1, const int32_t* __restrict__ leaf,
                             const Ctrl* __restrict__ mc, const Node* __restrict__ nodes,
                             const int32_t* __restrict__ first, const int32_t* __restrict__ npts,
                             int32_t* __restrict__ mark) {
    const int np = mc->n_alive;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const bool with_next = k + 1 < np && p1[k] == p1[k + 1] && leaf[k] == leaf[k + 1];
    const bool with_prev = k > 0 && p1[k] == p1[k - 1] && leaf[k] == leaf[k - 1];
    if (!with_next || with_prev) return;
    int node = 0;
    for (int depth = 0; depth <= 64 && nodes[node].count < 0; depth++) {
        mark[node] = 1;
        node = child_holding(nodes, first, npts, node, leaf[k]);
        if (node < 0) return;
    }
}

// Stored order of the leaf whose first point is leaf[k], for the pairs that
// share their leaf with a neighbour in the run: rank[j] = place of sorted
// point j in it.
__global__ void k_leaf_rank(const int32_t* __restrict__ p1, const int32_t* __restrict__ leaf,
                            const Ctrl* __restrict__ mc, const Node* __restrict__ nodes,
                            const int32_t* __restrict__ first, const int32_t* __restrict__ npts,
                            const int32_t* __restrict__ slot, const int32_t* __restrict__ s5,
                            int32_t* __restrict__ rank) {
    const int np = mc->n_alive;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const bool with_next = k + 1 < np && p1[k] == p1[k + 1] && leaf[k] == leaf[k + 1];
    const bool with_prev = k > 0 && p1[k] == p1[k - 1] && leaf[k] == leaf[k - 1];
    if (!with_next || with_prev) return;        // one thread per run of such pairs
    const int target = leaf[k];
    int f[5], nf;
    int node = 0;
    if (nodes[0].count >= 0) return;            // the root is a leaf: id order, nothing to rank
    nf = 0;
    for (int q = 0; q < 5; q++) { const int v = s5[q]; if (v >= 0) f[nf++] = v; }
    for (int depth = 0; depth < 64; depth++) {
        const int c = child_holding(nodes, first, npts, node, target);
        if (c < 0) return;
        const int ca = first[c], ce = ca + npts[c];
        // H(C): the node's first five arrivals that fall in C, newest first
        int hh[5], nh = 0;
        for (int t = nf - 1; t >= 0; t--)
            if (f[t] >= ca && f[t] < ce) hh[nh++] = f[t];
        if (nodes[c].count >= 0) {              // the leaf: H first, then the rest by slot
            for (int j = ca; j < ce; j++) {
                int r = -1;
                for (int t = 0; t < nh; t++) if (hh[t] == j) r = t;
                if (r < 0) {
                    r = nh;
                    const int sj = slot[j];
                    for (int i = ca; i < ce; i++) {
                        bool in_h = false;
                        for (int t = 0; t < nh; t++) in_h |= hh[t] == i;
                        if (!in_h && slot[i] < sj) r++;
                    }
                }
                rank[j] = r;
            }
            return;
        }
        // F5(C) = H(C), then C's smallest slots outside it, five in all
        for (int t = 0; t < nh; t++) f[t] = hh[t];
        nf = nh;
        const int32_t* sc = s5 + (size_t)c * 5;
        for (int q = 0; q < 5 && nf < 5; q++) {
            const int v = sc[q];
            if (v < 0) break;
            bool in_h = false;
            for (int t = 0; t < nh; t++) in_h |= hh[t] == v;
            if (!in_h) f[nf++] = v;
        }
        node = c;
    }
}

// Sort key of every pair: (slot of the first member, traversal place of the
// candidate's leaf, the candidate's place in that leaf).
__global__ void k_pair_keys(const int32_t* __restrict__ p1, const int32_t* __restrict__ p2,
                            const int32_t* __restrict__ leaf, const int32_t* __restrict__ slot,
                            const int32_t* __restrict__ rank, const Ctrl* __restrict__ mc,
                            const Ctrl* __restrict__ ctrl, uint64_t* __restrict__ keys, uint32_t* __restrict__ perm) {
    const int np = mc->n_alive;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t lf = (uint32_t)(ctrl->n_tree - 1 - leaf[k]);
    const uint32_t rk = rank ? (uint32_t)min(max(rank[p2[k]], 0), 7) : 0u;
    keys[k] = ((uint64_t)(uint32_t)slot[p1[k]] << 32) | (uint64_t)((lf << 3) | rk);
    perm[k] = (uint32_t)k;
}

// ---- resolution ------------------------------------------------------------

__global__ void k_fill_i32(int32_t* __restrict__ a, int32_t v, const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ctrl->n_alive) a[i] = v;
}

__global__ void k_merge_init(MergeArgs m) {
    const int np = m.mc->n_alive;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const int ends[2] = {m.p1[k], m.p2[k]};
    for (int e = 0; e < 2; e++) {
        const int i = ends[e];
        m.wm[i] = m.mass[i];
        m.wbm[i] = m.base_mass[i];
        m.wr[i] = cbrt((double)m.mass[i]);     // the body radius in FP64, as the reference holds it
        m.wdm[i] = m.mass[i];
    }
}

__global__ void k_merge_degree(MergeArgs m) {
    const int np = m.mc->n_alive;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    atomicAdd(&m.deg[m.p1[k]], 1);
    atomicAdd(&m.deg[m.p2[k]], 1);
}

__device__ __forceinline__ double js_min(double a, double b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }
__device__ __forceinline__ double js_max(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }

// _particleKE (collisions.js:17-20)
__device__ __forceinline__ double particle_ke(double wx, double wy, double m) {
    const double w2 = wx * wx + wy * wy;
    return w2 / (sqrt(1.0 + w2) + 1.0) * m;
}

// One pair at its turn of the fold (collisions.js:41-97): p1's query box
// half-size rw was fixed when p1's turn began.
__device__ void resolve_pair(const MergeArgs& m, int t, int i, int j, double rw) {
    MergeRes res;
    res.type = 0;
    res.mev = 0;
    if (m.wm[j] == 0.0) { m.res[t] = res; return; }
    double dx = (double)m.x[j] - (double)m.x[i], dy = (double)m.y[j] - (double)m.y[i];
    if (m.periodic) {
        const double W = m.dom_w, H = m.dom_h;
        if (dx > 0.5 * W) dx -= W; else if (dx < -0.5 * W) dx += W;
        if (dy > 0.5 * H) dy -= H; else if (dy < -0.5 * H) dy += H;
    }
    if (!(fabs(dx) <= rw && fabs(dy) <= rw)) { m.res[t] = res; return; }
    const double min_d = m.wr[i] + m.wr[j];
    if (!(dx * dx + dy * dy < min_d * min_d)) { m.res[t] = res; return; }
    const double w1x = m.wx[i], w1y = m.wy[i], w2x = m.wx[j], w2y = m.wy[j];
    const double m1 = m.wm[i], m2 = m.wm[j];
    const double x1 = m.x[i], y1 = m.y[i];
    if ((m.flags[i] & 1u) != (m.flags[j] & 1u)) {
        // annihilation (collisions.js:62-82)
        const double ann = js_min(m1, m2);
        res.type = 2;
        res.x = x1 + dx * 0.5;
        res.y = y1 + dy * 0.5;
        res.px = (w1x + w2x) * ann;
        res.py = (w1y + w2y) * ann;
        const double ke = (ann / m1) * particle_ke(w1x, w1y, m1) + (ann / m2) * particle_ke(w2x, w2y, m2);
        res.e = 2.0 * ann + ke;
        m.wdm[i] = m1;
        m.wdm[j] = m2;
        const double n1 = m1 - ann, n2 = m2 - ann;
        m.wm[i] = n1;
        m.wm[j] = n2;
        if (m1 > 0) m.wbm[i] *= n1 / m1;
        if (m2 > 0) m.wbm[j] *= n2 / m2;
        m.wr[i] = cbrt(n1);                     // updateColor (particle.js:110-121)
        m.wr[j] = cbrt(n2);
        m.res[t] = res;
        return;
    }
    // merge (collisions.js:83-96, resolveMerge 112-158)
    const double ke_before = particle_ke(w1x, w1y, m1) + particle_ke(w2x, w2y, m2);
    m.wdm[i] = m1;
    m.wdm[j] = m2;
    const double M = m1 + m2;
    const double nwx = (m1 * w1x + m2 * w2x) / M, nwy = (m1 * w1y + m2 * w2y) / M;
    const double x2 = x1 + dx, y2 = y1 + dy;
    const double nx = (x1 * m1 + x2 * m2) / M, ny = (y1 * m1 + y2 * m2) / M;
    const double dx1 = x1 - nx, dy1 = y1 - ny, dx2 = x2 - nx, dy2 = y2 - ny;
    const double lorb = dx1 * (m1 * w1y) - dy1 * (m1 * w1x) + dx2 * (m2 * w2y) - dy2 * (m2 * w2x);
    const double r1 = m.wr[i], r2 = m.wr[j];
    const double lspin = 0.4 * m1 * r1 * r1 * (double)m.angw[i] + 0.4 * m2 * r2 * r2 * (double)m.angw[j];
    const double nr = cbrt(M);
    const double nI = 0.4 * M * nr * nr;
    res.type = 1;
    res.x = nx; res.y = ny;
    res.wx = nwx; res.wy = nwy;
    res.mass = M;
    res.base_mass = m.wbm[i] + m.wbm[j];
    res.angw = nI > 1e-12 ? (lorb + lspin) / nI : 0.0;
    res.qq = m.qq[i] + m.qq[j];
    res.anti = (int)(m.flags[i] & 1u);
    const double ke = particle_ke(nwx, nwy, M);
    const double lost = js_max(0.0, ke_before - ke);
    res.e = lost;
    res.mev = lost > 0 ? 1 : 0;
    res.px = res.py = 0.0;
    m.wm[i] = 0.0; m.wm[j] = 0.0;
    m.wbm[i] = 0.0; m.wbm[j] = 0.0;
    m.res[t] = res;
}

// Pairs whose particles take part in no other pair, all at once.
__global__ void k_resolve_isolated(MergeArgs m) {
    const int np = m.mc->n_alive;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const int k = (int)m.order[t];
    const int i = m.p1[k], j = m.p2[k];
    if (m.deg[i] != 1 || m.deg[j] != 1) return;
    resolve_pair(m, t, i, j, 2.0 * m.wr[i]);
}

// f_fold[t] = 1 for the pairs the fold takes in order (a particle in another pair too).
__global__ void k_fold_flags(MergeArgs m) {
    const int np = m.mc->n_alive;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const int k = (int)m.order[t];
    m.f_fold[t] = !(m.deg[m.p1[k]] == 1 && m.deg[m.p2[k]] == 1);
}

__global__ void k_fold_list(MergeArgs m) {
    const int np = m.mc->n_alive;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < np && m.f_fold[t]) m.fold[m.o_fold[t]] = t;
}

// The remaining pairs in the reference's order, by one thread.
__global__ void k_resolve_fold(MergeArgs m, int only_on_fail) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (only_on_fail && !*m.uf_fail) return;
    const int nf = m.cnt[6];
    int prev = -1;
    bool live = false;
    double rw = 0.0;
    for (int u = 0; u < nf; u++) {
        const int t = m.fold[u];
        const int k = (int)m.order[t];
        const int i = m.p1[k], j = m.p2[k];
        if (i != prev) {                        // p1's turn begins (collisions.js:35-39)
            prev = i;
            live = m.wm[i] != 0.0;
            rw = 2.0 * m.wr[i];
        }
        if (!live) { MergeRes r; r.type = 0; r.mev = 0; m.res[t] = r; continue; }
        resolve_pair(m, t, i, j, rw);
    }
}

// ---- the fold by connected components ---------------------------------------
// Pairs that share no particle, directly or through a chain of pairs, do not
// see each other's effects, and a particle's pairs are contiguous in the fold
// order and all in its component.  So folding every component by its own
// thread, in fold order, gives the one-thread fold's results exactly.
// Components come from a lock-free union-find over the particles (a root is
// hooked under a smaller root by compare-and-swap), then a stable radix sort
// of the fold pairs by root.

__device__ __forceinline__ int uf_load(const int32_t* p, int i) {
    return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of v; parents only ever point to smaller indices, so the walk ends.
__device__ int uf_find(int32_t* parent, int v) {
    int p = uf_load(parent, v);
    while (p != v) {
        const int g = uf_load(parent, p);
        if (g != p) __hip_atomic_store(parent + v, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);   // halving
        v = p;
        p = g;
    }
    return v;
}

__global__ void k_uf_init(int32_t* __restrict__ parent, const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ctrl->n_alive) parent[i] = i;
}

constexpr int kUnionTries = 1 << 16;

__global__ void k_uf_union(MergeArgs m) {
    const int nf = m.cnt[6];
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nf) return;
    const int k = (int)m.order[m.fold[u]];
    const int a = m.p1[k], b = m.p2[k];
    for (int tries = 0; tries < kUnionTries; tries++) {
        const int ra = uf_find(m.uf, a), rb = uf_find(m.uf, b);
        if (ra == rb) return;
        const int hi = ra > rb ? ra : rb, lo = ra ^ rb ^ hi;
        if (atomicCAS(m.uf + hi, hi, lo) == hi) return;
    }
    atomicExch(m.uf_fail, 1);                   // not reached in practice: the host folds on one thread
}

// Sort keys: the root of each fold pair's first particle, in fold order.
__global__ void k_uf_keys(MergeArgs m, uint64_t* __restrict__ keys, uint32_t* __restrict__ perm) {
    const int nf = m.cnt[6];
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nf) return;
    const int k = (int)m.order[m.fold[u]];
    keys[u] = (uint64_t)(uint32_t)uf_find(m.uf, m.p1[k]);
    perm[u] = (uint32_t)u;
}

// One thread per component: the component's pairs in fold order.
__global__ void k_fold_components(MergeArgs m, const uint64_t* __restrict__ keys,
                                  const uint32_t* __restrict__ perm) {
    const int nf = m.cnt[6];
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (v0 >= nf || *m.uf_fail) return;
    if (v0 > 0 && keys[v0] == keys[v0 - 1]) return;
    int prev = -1;
    bool live = false;
    double rw = 0.0;
    for (int v = v0; v < nf && keys[v] == keys[v0]; v++) {
        const int t = m.fold[perm[v]];
        const int k = (int)m.order[t];
        const int i = m.p1[k], j = m.p2[k];
        if (i != prev) {                        // p1's turn begins (collisions.js:35-39)
            prev = i;
            live = m.wm[i] != 0.0;
            rw = 2.0 * m.wr[i];
        }
        if (!live) { MergeRes r; r.type = 0; r.mev = 0; m.res[t] = r; continue; }
        resolve_pair(m, t, i, j, rw);
    }
}

// The sort's count (the fold pairs) and a cleared failure flag.
__global__ void k_set_count(Ctrl* __restrict__ c, const int32_t* __restrict__ n, int32_t* __restrict__ fail) {
    c->n_alive = *n;
    *fail = 0;
}

__global__ void k_merge_flags(MergeArgs m) {
    const int np = m.mc->n_alive;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const MergeRes& r = m.res[t];
    m.f_spawn[t] = r.type == 1;
    m.f_mev[t] = r.type == 1 && r.mev;
    m.f_aev[t] = r.type == 2;
}

// Spawn records and events in fold order: merge events first, then
// annihilation events (as the excitations take them, integrator.js:1404-1413).
__global__ void k_merge_emit(MergeArgs m) {
    const int np = m.mc->n_alive;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const MergeRes& r = m.res[t];
    const int n_mev = m.cnt[1];
    if (r.type == 1) {
        m.spawns[m.o_spawn[t]] = r;
        if (r.mev) {
            double* ev = m.events + (size_t)6 * m.o_mev[t];
            ev[0] = 1.0; ev[1] = r.x; ev[2] = r.y; ev[3] = r.e; ev[4] = 0.0; ev[5] = 0.0;
        }
    } else if (r.type == 2) {
        double* ev = m.events + (size_t)6 * (n_mev + m.o_aev[t]);
        ev[0] = 2.0; ev[1] = r.x; ev[2] = r.y; ev[3] = r.e; ev[4] = r.px; ev[5] = r.py;
    }
}

// Write the fold's outcome into the particle columns.  A removed particle
// keeps its death mass and radius in place for _retireParticle
// (integrator.js:152-173): the compaction that follows retires it.
__global__ void k_merge_apply(MergeArgs m, float* __restrict__ mass, float* __restrict__ base_mass,
                              float* __restrict__ radius, int32_t* __restrict__ alive, int32_t* __restrict__ deaths,
                              const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ctrl->n_alive) return;
    if (m.deg[i] == 0) { alive[i] = 1; return; }
    const double wm = m.wm[i];
    if (wm != 0.0) {
        alive[i] = 1;
        mass[i] = (float)wm;
        base_mass[i] = (float)m.wbm[i];
        radius[i] = (float)m.wr[i];
    } else {
        alive[i] = 0;
        mass[i] = (float)m.wdm[i];
        radius[i] = (float)m.wr[i];
        atomicAdd(deaths, 1);
    }
}

// Spawned particles appended after the survivors (integrator.js:1365-1384,
// main.js:504-539): slot next_slot + k, radius cbrt(m), moments and forces 0.
__global__ void k_spawn(SpawnArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.count) return;
    const MergeRes& r = a.spawns[k];
    const int i = a.base + k;
    double x = r.x, y = r.y;
    if (a.periodic) {                           // wrapPosition on the torus (topology.js:81-93)
        if (x < 0) x += a.dom_w; else if (x >= a.dom_w) x -= a.dom_w;
        if (y < 0) y += a.dom_h; else if (y >= a.dom_h) y -= a.dom_h;
    }
    const double rad = cbrt(r.mass);
    const double ig = a.rel ? 1.0 / sqrt(1.0 + r.wx * r.wx + r.wy * r.wy) : 1.0;
    a.x[i] = (float)x; a.y[i] = (float)y;
    a.wx[i] = (float)r.wx; a.wy[i] = (float)r.wy;
    a.vx[i] = (float)(r.wx * ig); a.vy[i] = (float)(r.wy * ig);
    a.mass[i] = (float)r.mass; a.base_mass[i] = (float)r.base_mass;
    a.charge[i] = (float)(r.qq * 0.1);
    a.angw[i] = (float)r.angw;
    a.angvel[i] = (float)(a.rel ? r.angw / sqrt(1.0 + r.angw * r.angw * rad * rad) : r.angw);
    a.radius[i] = (float)rad;
    a.mu[i] = 0.f; a.L[i] = 0.f; a.f1ox[i] = 0.f; a.f1oy[i] = 0.f;
    a.ax[i] = 1.f; a.yuk[i] = 1.f; a.higgs[i] = 1.f;
    a.qq[i] = r.qq;
    a.slot[i] = a.first_slot + k;
    a.flags[i] = (uint32_t)r.anti;
    for (int c = 0; c < a.n_force; c++) a.force[c][i] = 0.f;
}

// ScalarField.depositExcitation (scalar-field.js:419-446) for every event in
// order, one thread per grid cell: the wave packets of the merges, then of the
// annihilations, added to fieldDot.
__global__ void k_excite(float* __restrict__ field_dot, int G, double W, double H, const double* __restrict__ events,
                         int n_events, double frac) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G * G) return;
    const int ix = c % G, iy = c / G;
    const double cw = W / G, ch = H / G;
    if (cw < 1e-12 || ch < 1e-12) return;
    const double sigma = 2.0, s2 = sigma * sigma;   // FIELD_EXCITATION_SIGMA (config.js:86)
    const int range = 6;                            // ceil(3 sigma)
    double acc = field_dot[c];
    for (int e = 0; e < n_events; e++) {
        const double* ev = events + (size_t)6 * e;
        const double gx = ev[1] / cw, gy = ev[2] / ch;
        const double fx = floor(gx), fy = floor(gy);
        if (ix < fx - range || ix > fx + range || iy < fy - range || iy > fy + range) continue;
        const double dx = ix - gx, dy = iy - gy;
        const double r2 = dx * dx + dy * dy;
        if (r2 > 9 * s2) continue;
        // MERGE_EXCITATION_SCALE 0.5, EXCITATION_MAX_AMPLITUDE 1 (config.js:87, 94)
        const double amp = js_min(0.5 * sqrt(ev[3] * frac), 1.0);
        acc += amp * exp(-r2 / (2 * s2));
    }
    field_dot[c] = (float)acc;
}

// ---- launchers ----------------------------------------------------------------

static inline int blocks_of(int n) { return (n + 255) / 256; }

void merge_pair_same_leaf(hipStream_t s, const int32_t* p1, const int32_t* leaf, const Ctrl* mc, int32_t* need,
                          int np) {
    if (np > 1) hipLaunchKernelGGL(k_pair_same_leaf, dim3(blocks_of(np)), dim3(256), 0, s, p1, leaf, mc, need);
}

__global__ void k_zero_i32(int32_t* __restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

void merge_leaf_ranks(hipStream_t s, const TreeBufs& tb, const int32_t* slot, int n_nodes, int32_t* s5,
                      int32_t* mark, uint64_t* chunk5, int n_upper, const Ctrl* ctrl, const int32_t* p1,
                      const int32_t* leaf, const Ctrl* mc, int32_t* rank, int np) {
    if (n_nodes <= 0 || np <= 0) return;
    const int n_chunks = (n_upper + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_chunk5, dim3((unsigned)(((int64_t)n_chunks * kWave + 255) / 256)), dim3(256), 0, s, slot,
                       ctrl, n_chunks, chunk5);
    hipLaunchKernelGGL(k_zero_i32, dim3(blocks_of(n_nodes)), dim3(256), 0, s, mark, n_nodes);
    hipLaunchKernelGGL(k_mark_paths, dim3(blocks_of(np)), dim3(256), 0, s, p1, leaf, mc, tb.nodes, tb.node_first,
                       tb.node_npts, mark);
    const int64_t threads = (int64_t)n_nodes * kWave;
    hipLaunchKernelGGL(k_s5, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, tb.nodes, tb.node_first,
                       tb.node_npts, slot, chunk5, n_nodes, mark, s5);
    hipLaunchKernelGGL(k_leaf_rank, dim3(blocks_of(np)), dim3(256), 0, s, p1, leaf, mc, tb.nodes, tb.node_first,
                       tb.node_npts, slot, s5, rank);
}

void merge_pair_keys(hipStream_t s, const int32_t* p1, const int32_t* p2, const int32_t* leaf, const int32_t* slot,
                     const int32_t* rank, const Ctrl* mc, const Ctrl* ctrl, uint64_t* keys, uint32_t* perm, int np) {
    if (np > 0)
        hipLaunchKernelGGL(k_pair_keys, dim3(blocks_of(np)), dim3(256), 0, s, p1, p2, leaf, slot, rank, mc, ctrl, keys,
                           perm);
}

hipError_t merge_resolve(hipStream_t s, const MergeArgs& m, TreeBufs& csort, Ctrl* cctrl, int np, int n) {
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_of(n)), dim3(256), 0, s, m.deg, 0, m.ctrl);
    hipLaunchKernelGGL(k_merge_init, dim3(blocks_of(np)), dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_merge_degree, dim3(blocks_of(np)), dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_resolve_isolated, dim3(blocks_of(np)), dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_fold_flags, dim3(blocks_of(np)), dim3(256), 0, s, m);
    exclusive_scan_i32(s, m.f_fold, m.o_fold, np, &m.mc->n_alive, m.cnt + 6, m.scan_tmp);
    hipLaunchKernelGGL(k_fold_list, dim3(blocks_of(np)), dim3(256), 0, s, m);
    int32_t nf = 0;
    hipError_t e = hipMemcpyAsync(&nf, m.cnt + 6, sizeof(nf), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (nf <= kFoldOneThread) {
        hipLaunchKernelGGL(k_resolve_fold, dim3(1), dim3(64), 0, s, m, 0);
    } else {
        hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, s, cctrl, m.cnt + 6, m.uf_fail);
        hipLaunchKernelGGL(k_uf_init, dim3(blocks_of(n)), dim3(256), 0, s, m.uf, m.ctrl);
        hipLaunchKernelGGL(k_uf_union, dim3(blocks_of(nf)), dim3(256), 0, s, m);
        hipLaunchKernelGGL(k_uf_keys, dim3(blocks_of(nf)), dim3(256), 0, s, m, csort.keys, csort.perm);
        radix_sort(s, csort, nf, 32, cctrl);
        hipLaunchKernelGGL(k_fold_components, dim3(blocks_of(nf)), dim3(256), 0, s, m, csort.keys, csort.perm);
        hipLaunchKernelGGL(k_resolve_fold, dim3(1), dim3(64), 0, s, m, 1);
    }
    hipLaunchKernelGGL(k_merge_flags, dim3(blocks_of(np)), dim3(256), 0, s, m);
    exclusive_scan_i32(s, m.f_spawn, m.o_spawn, np, &m.mc->n_alive, m.cnt + 0, m.scan_tmp);
    exclusive_scan_i32(s, m.f_mev, m.o_mev, np, &m.mc->n_alive, m.cnt + 1, m.scan_tmp);
    exclusive_scan_i32(s, m.f_aev, m.o_aev, np, &m.mc->n_alive, m.cnt + 2, m.scan_tmp);
    hipLaunchKernelGGL(k_merge_emit, dim3(blocks_of(np)), dim3(256), 0, s, m);
    return hipGetLastError();
}

void merge_apply(hipStream_t s, const MergeArgs& m, float* mass, float* base_mass, float* radius, int32_t* alive,
                 int n) {
    if (n > 0)
        hipLaunchKernelGGL(k_merge_apply, dim3(blocks_of(n)), dim3(256), 0, s, m, mass, base_mass, radius, alive,
                           m.cnt + 3, m.ctrl);
}

void merge_spawn(hipStream_t s, const SpawnArgs& a) {
    if (a.count > 0) hipLaunchKernelGGL(k_spawn, dim3(blocks_of(a.count)), dim3(256), 0, s, a);
}

void merge_excite(hipStream_t s, float* field_dot, int G, double W, double H, const double* events, int n_events,
                  double frac) {
    if (n_events > 0 && G > 0)
        hipLaunchKernelGGL(k_excite, dim3(blocks_of(G * G)), dim3(256), 0, s, field_dot, G, W, H, events, n_events,
                           frac);
}

}  // namespace geon
