// k_tree.hip — Morton keys, LSD radix sort, permutation gather, prefix sums and
// the bottom-up capacity-4 quadtree build, for gfx950.
//
// The quadtree reproduces the structure of the reference's insertion-built tree
// (src/quadtree.js:185-237): a node is divided iff more than QUADTREE_CAPACITY
// (4) points fall in its box, children are NW,NE,SW,SE (quadtree.js:119-172) and
// a point on a split plane goes to the lower-coordinate child because the
// reference tests `py <= cy`, `px <= cx` (quadtree.js:175-179).  The root box is
// centre (W/2, H/2), half-extent (2W, 2H) (integrator.js:727-730); points
// outside it are not inserted (quadtree.js:186).
//
// Build (no locks, no atomics on the structure):
//   1. key_i = Morton interleave of the finest cell, with cells half-open on
//      the LOW side, (b_c, b_c+1], so the `<=` tie rule holds at every level.
//   2. LSD radix sort of (key, index).
//   3. For sorted i: c_i = common prefix (in levels) of key_{i-1}, key_i;
//      D_i = common prefix of key_i, key_{i+4}.  The node at level d holding
//      point i has >= 5 points and i is its first point iff c_i < d <= D_i, so
//      point i "owns" D_i - c_i divided nodes; their pre-order ranks come from
//      one exclusive scan.  Divided node of rank r keeps its 4 children at
//      node slots 1 + 4r .. 1 + 4r + 3, like the reference's pool (c..c+3).
//   4. One thread per divided node finds its children's point ranges by binary
//      search and writes the 4 child records with FP64 aggregates (direct sums
//      for leaves, prefix-sum differences for divided children).
#include "geon_kernels.h"

namespace geon {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;   // 2048 keys per block
static_assert(kSortTile == kSortTileKeys, "radix sort tile");
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

// ---------------------------------------------------------------------------
// Morton keys
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t cell_index(double v, double vmin, double scale, int levels) {
    // Finest-level cell c with b_c < v <= b_{c+1} (b_c = vmin + c/scale); the
    // lowest cell also holds v == vmin.  u is exact when vmin and 1/scale are
    // dyadic and v is an FP32 value of moderate magnitude; the two exact
    // comparisons below repair the cases where it is not.
    const double u = (v - vmin) * scale;
    const double maxc = (double)((1u << levels) - 1);
    double c = ceil(u) - 1.0;
    c = fmin(fmax(c, 0.0), maxc);
    if (c < maxc && v > vmin + (c + 1.0) / scale) c += 1.0;
    if (c > 0.0 && v <= vmin + c / scale) c -= 1.0;
    return (uint32_t)c;
}

__device__ __forceinline__ uint64_t spread2(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

__global__ void k_keys(const float* __restrict__ x, const float* __restrict__ y, KeyParams kp,
                       uint64_t* __restrict__ keys, uint32_t* __restrict__ perm, const Ctrl* __restrict__ ctrl) {
    const int n = ctrl->n_alive;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = (double)x[i], py = (double)y[i];
    const double span_x = (double)(1u << kKeyLevels) / kp.scale_x;
    const double span_y = (double)(1u << kKeyLevels) / kp.scale_y;
    // quadtree.js:105-110 `_contains`: inclusive on both ends.
    const bool inside = px >= kp.xmin && px <= kp.xmin + span_x && py >= kp.ymin && py <= kp.ymin + span_y;
    uint64_t key;
    if (inside) {
        const uint32_t cx = cell_index(px, kp.xmin, kp.scale_x, kKeyLevels);
        const uint32_t cy = cell_index(py, kp.ymin, kp.scale_y, kKeyLevels);
        key = spread2(cx) | (spread2(cy) << 1);     // digit = 2*ybit + xbit: NW0 NE1 SW2 SE3
    } else {
        key = 1ull << (2 * kKeyLevels);             // sorts after every inserted point
    }
    keys[i] = key;
    perm[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// Radix sort (stable LSD, 8-bit digits)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t* __restrict__ keys, int shift,
                                                             uint32_t* __restrict__ hist, int nblocks,
                                                             const Ctrl* __restrict__ ctrl) {
    __shared__ uint32_t h[kRadix];
    const int n = ctrl->n_alive;
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        const int i = base + k * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Per digit (one block each): exclusive scan of that digit's per-block counts,
// in place, and the digit's total.  The scatter adds the digit's base offset.
__global__ __launch_bounds__(kSortThreads) void k_radix_scan(uint32_t* __restrict__ hist, int nblocks,
                                                             uint32_t* __restrict__ digit_total) {
    __shared__ uint32_t part[kSortThreads];
    const int t = threadIdx.x;
    uint32_t* h = hist + (size_t)blockIdx.x * nblocks;
    const int chunk = (nblocks + kSortThreads - 1) / kSortThreads;
    const int b = t * chunk, e = min(nblocks, b + chunk);
    uint32_t s = 0;
    for (int i = b; i < e; i++) s += h[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < kSortThreads; off <<= 1) {
        uint32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = b; i < e; i++) { uint32_t v = h[i]; h[i] = run; run += v; }
    if (t == kSortThreads - 1) digit_total[blockIdx.x] = part[t];
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint64_t* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in,
                                                                uint64_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out, int shift,
                                                                const uint32_t* __restrict__ hist, int nblocks,
                                                                const uint32_t* __restrict__ digit_total,
                                                                const Ctrl* __restrict__ ctrl) {
    __shared__ uint64_t s_key[kSortTile];
    __shared__ uint32_t s_val[kSortTile];
    __shared__ uint32_t s_wcnt[kSortThreads / kWave][kRadix];
    __shared__ uint32_t s_run[kRadix];
    __shared__ uint32_t s_dstart[kRadix];
    __shared__ uint32_t s_goff[kRadix];
    const int n = ctrl->n_alive;
    const int t = threadIdx.x, lane = t & (kWave - 1), wv = t / kWave;
    const int base = blockIdx.x * kSortTile;
    if (base >= n) return;
    // Global offset of (digit, block): exclusive scan of the digit totals plus
    // this block's offset inside its digit.
    {
        const uint32_t tot = digit_total[t];
        s_goff[t] = tot;
        __syncthreads();
        for (int off = 1; off < kRadix; off <<= 1) {
            uint32_t a = t >= off ? s_goff[t - off] : 0;
            __syncthreads();
            s_goff[t] += a;
            __syncthreads();
        }
        s_goff[t] = s_goff[t] - tot + hist[t * nblocks + blockIdx.x];
    }
    // Tile digit histogram -> digit start offsets inside the tile.
    s_run[t] = 0;
    __syncthreads();
    uint64_t k_reg[kSortItems];
    uint32_t v_reg[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        const int i = base + k * kSortThreads + t;
        k_reg[k] = i < n ? keys_in[i] : 0;
        v_reg[k] = i < n ? vals_in[i] : 0;
        if (i < n) atomicAdd(&s_run[(k_reg[k] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    {
        // exclusive scan of s_run (256 entries) by one wave-level pass per thread
        uint32_t v = s_run[t];
        __syncthreads();
        // Hillis-Steele over 256 entries
        s_dstart[t] = v;
        __syncthreads();
        for (int off = 1; off < kRadix; off <<= 1) {
            uint32_t a = t >= off ? s_dstart[t - off] : 0;
            __syncthreads();
            s_dstart[t] += a;
            __syncthreads();
        }
        s_dstart[t] -= v;       // exclusive
        s_run[t] = 0;           // running count per digit across rounds
    }
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        const int i = base + k * kSortThreads + t;
        const bool valid = i < n;
        const uint32_t d = (uint32_t)(k_reg[k] >> shift) & (kRadix - 1);
        // Lanes holding the same digit, by 8 ballots.
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; b++) {
            const uint64_t bb = __ballot(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? bb : ~bb;
        }
#pragma unroll
        for (int j = 0; j < kSortThreads / kWave; j++) s_wcnt[j][t] = 0;
        __syncthreads();
        const uint32_t rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) s_wcnt[wv][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t run = s_run[t];
#pragma unroll
            for (int j = 0; j < kSortThreads / kWave; j++) {
                const uint32_t c = s_wcnt[j][t];
                s_wcnt[j][t] = run;
                run += c;
            }
            s_run[t] = run;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = s_dstart[d] + s_wcnt[wv][d] + rank;
            s_key[pos] = k_reg[k];
            s_val[pos] = v_reg[k];
        }
        __syncthreads();
    }
    const int cnt = min(kSortTile, n - base);
    for (int j = t; j < cnt; j += kSortThreads) {
        const uint64_t key = s_key[j];
        const uint32_t d = (uint32_t)(key >> shift) & (kRadix - 1);
        const uint32_t dst = s_goff[d] + (uint32_t)j - s_dstart[d];
        keys_out[dst] = key;
        vals_out[dst] = s_val[j];
    }
}

void radix_sort(hipStream_t s, TreeBufs& tb, int n_upper, int key_bits, const Ctrl* ctrl) {
    const int nblocks = (n_upper + kSortTile - 1) / kSortTile;
    if (nblocks == 0) return;
    const int passes = (key_bits + kRadixBits - 1) / kRadixBits;
    uint64_t* kin = tb.keys; uint64_t* kout = tb.keys_alt;
    uint32_t* vin = tb.perm; uint32_t* vout = tb.perm_alt;
    for (int p = 0; p < passes; p++) {
        const int shift = p * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortThreads), 0, s, kin, shift, tb.hist, nblocks, ctrl);
        hipLaunchKernelGGL(k_radix_scan, dim3(kRadix), dim3(kSortThreads), 0, s, tb.hist, nblocks, tb.digit_total);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSortThreads), 0, s, kin, vin, kout, vout, shift,
                           tb.hist, nblocks, tb.digit_total, ctrl);
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    if (kin != tb.keys) {   // odd pass count: result lives in the alt buffers
        uint64_t* tk = tb.keys; tb.keys = tb.keys_alt; tb.keys_alt = tk;
        uint32_t* tv = tb.perm; tb.perm = tb.perm_alt; tb.perm_alt = tv;
    }
}

void launch_keys(hipStream_t s, const float* x, const float* y, int n, KeyParams kp, uint64_t* keys,
                 uint32_t* perm, Ctrl* ctrl) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_keys, dim3((n + 255) / 256), dim3(256), 0, s, x, y, kp, keys, perm, ctrl);
}

// ---------------------------------------------------------------------------
// Permutation gather of 32-bit columns
// ---------------------------------------------------------------------------
__global__ void k_gather(Columns cols, const uint32_t* __restrict__ perm, const Ctrl* __restrict__ ctrl) {
    const int n = ctrl->n_alive;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = perm[i];
    for (int c = 0; c < cols.ncols; c++) cols.dst[c][i] = cols.src[c][p];
}

void launch_gather(hipStream_t s, const Columns& cols, const uint32_t* perm, int n_upper, const Ctrl* ctrl) {
    if (n_upper <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3((n_upper + 255) / 256), dim3(256), 0, s, cols, perm, ctrl);
}

// ---------------------------------------------------------------------------
// Int32 exclusive scan (3 phases)
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int32_t block_excl_scan_i32(int32_t v, int32_t* sh, int32_t* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        int32_t a = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    const int32_t incl = sh[t];
    *total = sh[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t* __restrict__ in, int32_t* __restrict__ bsum,
                                                              const int32_t* __restrict__ n_dev) {
    __shared__ int32_t sh[kScanThreads];
    const int n = *n_dev;
    const int base = blockIdx.x * kScanTile;
    int32_t s = 0;
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    int32_t tot;
    block_excl_scan_i32(s, sh, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_bsums(int32_t* __restrict__ bsum, int nb, int32_t* __restrict__ total_dev) {
    __shared__ int32_t part[1024];
    const int t = threadIdx.x;
    const int chunk = (nb + 1023) / 1024;
    const int b = t * chunk, e = min(nb, b + chunk);
    int32_t s = 0;
    for (int i = b; i < e; i++) s += bsum[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = part[t] - s;
    for (int i = b; i < e; i++) { int32_t v = bsum[i]; bsum[i] = run; run += v; }
    if (t == 1023 && total_dev) *total_dev = part[1023];
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                             const int32_t* __restrict__ bsum,
                                                             const int32_t* __restrict__ n_dev) {
    __shared__ int32_t sh[kScanThreads];
    const int n = *n_dev;
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    int32_t tot;
    int32_t run = block_excl_scan_i32(s, sh, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

void exclusive_scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int n_upper, const int32_t* n_dev,
                        int32_t* total_dev, int32_t* tmp) {
    const int nb = (n_upper + kScanTile - 1) / kScanTile;
    if (nb == 0) return;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, s, in, tmp, n_dev);
    hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(1024), 0, s, tmp, nb, total_dev);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanThreads), 0, s, in, out, tmp, n_dev);
}

// ---------------------------------------------------------------------------
// FP64 multi-column prefix sums over the sorted points (tree aggregates)
// ---------------------------------------------------------------------------
constexpr int kPreCols = 8;   // m, m*x, m*y, mu, L, m*wx, m*wy, charge quanta
constexpr int kMomCols = 9;   // m X^2, m XY, m Y^2, q X, q Y, q X^2, q XY, q Y^2, |q|  (X = x - x0)

struct PreIn {
    static constexpr int kCols = kPreCols;
    const float *x, *y, *m, *wx, *wy, *mu, *L;
    const int32_t* qq;
    __device__ __forceinline__ void item(int i, double* v) const {
        const double mi = m[i];
        v[0] = mi;
        v[1] = mi * (double)x[i];
        v[2] = mi * (double)y[i];
        v[3] = mu[i];
        v[4] = L[i];
        v[5] = mi * (double)wx[i];
        v[6] = mi * (double)wy[i];
        v[7] = (double)qq[i];
    }
};

// Second moments about a fixed origin (x0, y0) near the points, for the
// far-field inputs of nodes too large for a direct sum.
struct MomIn {
    static constexpr int kCols = kMomCols;
    const float *x, *y, *m;
    const int32_t* qq;
    double x0, y0;
    __device__ __forceinline__ void item(int i, double* v) const {
        const double X = (double)x[i] - x0, Y = (double)y[i] - y0;
        const double mi = m[i], q = (double)qq[i] * kChargeQuantum;
        v[0] = mi * X * X; v[1] = mi * X * Y; v[2] = mi * Y * Y;
        v[3] = q * X; v[4] = q * Y;
        v[5] = q * X * X; v[6] = q * X * Y; v[7] = q * Y * Y;
        v[8] = fabs(q);
    }
};

template <class IN>
__global__ __launch_bounds__(kScanThreads) void k_pre_reduce(IN in, double* __restrict__ bsum, int nb,
                                                             const Ctrl* __restrict__ ctrl) {
    constexpr int NC = IN::kCols;
    __shared__ double sh[NC][kScanThreads];
    const int n = ctrl->n_tree;
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        if (i < n) {
            double v[NC];
            in.item(i, v);
#pragma unroll
            for (int c = 0; c < NC; c++) acc[c] += v[c];
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int off = kScanThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
#pragma unroll
            for (int c = 0; c < NC; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < NC) bsum[threadIdx.x * nb + blockIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(1024) void k_pre_bsums(double* __restrict__ bsum, int nb, int ncols) {
    // one wave per column: serial chunks + wave scan
    const int c = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (c >= ncols) return;
    double* b = bsum + (size_t)c * nb;
    const int chunk = (nb + 63) / 64;
    const int lo = lane * chunk, hi = min(nb, lo + chunk);
    double s = 0;
    for (int i = lo; i < hi; i++) s += b[i];
    double incl = s;
    for (int off = 1; off < 64; off <<= 1) {
        const double v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    double run = incl - s;
    for (int i = lo; i < hi; i++) { const double v = b[i]; b[i] = run; run += v; }
}

// Exclusive scan of the block sums restarting at every system of a batch
// (slices aligned to kScanTile): one thread per system, fixed order.
__global__ void k_pre_bsums_seg(double* __restrict__ bsum, int nb, int ncols, const int32_t* __restrict__ soff,
                                int S, const Ctrl* __restrict__ ctrl) {
    const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (sidx >= S || ctrl->n_tree == 0) return;
    const int b0 = soff[sidx] / kScanTile, b1 = min(nb, (soff[sidx + 1] + kScanTile - 1) / kScanTile);
    for (int c = 0; c < ncols; c++) {
        double run = 0.0;
        for (int bk = b0; bk < b1; bk++) {
            const double v = bsum[(size_t)c * nb + bk];
            bsum[(size_t)c * nb + bk] = run;
            run += v;
        }
    }
}

template <class IN>
__global__ __launch_bounds__(kScanThreads) void k_pre_apply(IN in, const double* __restrict__ bsum, int nb,
                                                            double* __restrict__ pre, int64_t stride,
                                                            const Ctrl* __restrict__ ctrl) {
    constexpr int NC = IN::kCols;
    __shared__ double sh[NC][kScanThreads];
    const int n = ctrl->n_tree;
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        if (i < n) {
            double v[NC];
            in.item(i, v);
#pragma unroll
            for (int c = 0; c < NC; c++) acc[c] += v[c];
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        double a[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) a[c] = threadIdx.x >= (unsigned)off ? sh[c][threadIdx.x - off] : 0.0;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; c++) sh[c][threadIdx.x] += a[c];
        __syncthreads();
    }
    double run[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) run[c] = sh[c][threadIdx.x] - acc[c] + bsum[(size_t)c * nb + blockIdx.x];
    if (blockIdx.x == 0 && threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < NC; c++) pre[(size_t)c * stride] = 0.0;
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k;
        if (i < n) {
            double v[NC];
            in.item(i, v);
#pragma unroll
            for (int c = 0; c < NC; c++) {
                run[c] += v[c];
                pre[(size_t)c * stride + i + 1] = run[c];
            }
        }
    }
}

template <class IN>
static void prefix_columns(hipStream_t s, const IN& in, double* bsum, double* pre, int64_t stride, int nb,
                           const Ctrl* ctrl) {
    hipLaunchKernelGGL(k_pre_reduce<IN>, dim3(nb), dim3(kScanThreads), 0, s, in, bsum, nb, ctrl);
    hipLaunchKernelGGL(k_pre_bsums, dim3(1), dim3(IN::kCols * 64), 0, s, bsum, nb, IN::kCols);
    hipLaunchKernelGGL(k_pre_apply<IN>, dim3(nb), dim3(kScanThreads), 0, s, in, bsum, nb, pre, stride, ctrl);
}

// ---------------------------------------------------------------------------
// Tree structure
// ---------------------------------------------------------------------------
// Common prefix of two keys in tree levels; -1 when they differ above the
// Morton bits (the outside-the-root-box bit, or a batch's system id).
__device__ __forceinline__ int cpl_levels(uint64_t a, uint64_t b, int L) {
    const uint64_t d = a ^ b;
    if (d == 0) return L;
    if (d >> (2 * L)) return -1;
    return (__clzll((long long)d) - (64 - 2 * L)) >> 1;
}

__global__ void k_count_tree(const uint64_t* __restrict__ keys, Ctrl* __restrict__ ctrl) {
    // first index whose key carries the "outside the root box" bit
    int lo = 0, hi = ctrl->n_alive;
    const uint64_t out_bit = 1ull << (2 * kKeyLevels);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] & out_bit) hi = mid; else lo = mid + 1;
    }
    ctrl->n_tree = lo;
}

__global__ void k_tree_levels(const uint64_t* __restrict__ keys, int32_t* __restrict__ owned, Ctrl* __restrict__ ctrl) {
    const int n = ctrl->n_tree;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    const int c = i == 0 ? -1 : cpl_levels(keys[i - 1], k, kKeyLevels);
    int D = -1;
    if (i + kLeafCap < n) {
        const int cp = cpl_levels(k, keys[i + kLeafCap], kKeyLevels);
        if (cp == kKeyLevels && c < kKeyLevels) atomicAdd((unsigned long long*)&ctrl->depth_overflow, 1ull);
        D = min(cp, kKeyLevels - 1);
    }
    owned[i] = max(0, D - c);
}

__global__ void k_tree_scatter(const uint64_t* __restrict__ keys, const int32_t* __restrict__ rank_start,
                               int2* __restrict__ divnode, int64_t cap_div, Ctrl* __restrict__ ctrl) {
    const int n = ctrl->n_tree;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    const int c = i == 0 ? -1 : cpl_levels(keys[i - 1], k, kKeyLevels);
    int D = -1;
    if (i + kLeafCap < n) D = min(cpl_levels(k, keys[i + kLeafCap], kKeyLevels), kKeyLevels - 1);
    int r = rank_start[i];
    for (int d = c + 1; d <= D; d++, r++) {
        if (r < cap_div) divnode[r] = make_int2(i, d);
        else ctrl->overflow = 1;
    }
}

struct NodeBuild {
    const uint64_t* keys;
    const int32_t* rank_start;
    const int2* divnode;
    const double* pre;
    int64_t stride;
    PreIn pts;
    Node* nodes;
    double *mass64, *comx64, *comy64, *qq64;
    uint64_t* path;
    int32_t* npts;
    int32_t* first;             // first sorted point of each node
    const int32_t* seg_first;   // batch: first slot of each point's system (prefix sums restart there)
    NodeMom* mom;               // null: moments not wanted
    const double* pre_mom;      // kMomCols columns of prefix sums about (mom_in.x0, mom_in.y0)
    MomIn mom_in;
    double root_size;           // 4W: full width of the root box
    double xmin, ymin;          // lower corner of the root box
    double aspect_x, aspect_y;  // box extent per unit of `size` along x and y (1 and H/W)
    int64_t cap_div;
    bool batch;                 // roots are written per system by k_batch_roots, not at slot 0
    int levels;                 // levels the keys resolve
};

__device__ __forceinline__ uint64_t prefix_at(uint64_t key, int level, int L) {
    return key >> (2 * (L - level));      // level 0 keeps the bits above the Morton code
}

// First index j in [lo, hi) with prefix_at(keys[j], level) >= target (keys sorted).
__device__ __forceinline__ int lower_bound_prefix(const uint64_t* keys, int lo, int hi, int level, uint64_t target,
                                                  int L) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (prefix_at(keys[mid], level, L) < target) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Moments of the points [a, e) about their COM: the inputs of the far-field
// correction.  Small nodes sum their points directly in FP64; large ones take
// prefix-sum differences about a fixed origin, then shift to the COM
// (parallel-axis terms).
constexpr int kMomDirect = 64;

__device__ NodeMom node_moments(const NodeBuild& b, int a, int e, double m, double comx, double comy, double qtot) {
    double S[kMomCols];
    double cX, cY;
    if (e - a <= kMomDirect) {
        MomIn in = b.mom_in;
        in.x0 = comx; in.y0 = comy;
        for (int c = 0; c < kMomCols; c++) S[c] = 0.0;
        for (int j = a; j < e; j++) {
            double v[kMomCols];
            in.item(j, v);
            for (int c = 0; c < kMomCols; c++) S[c] += v[c];
        }
        cX = 0.0; cY = 0.0;
    } else {
        const int64_t st = b.stride;
        for (int c = 0; c < kMomCols; c++) S[c] = b.pre_mom[c * st + e] - b.pre_mom[c * st + a];
        cX = comx - b.mom_in.x0; cY = comy - b.mom_in.y0;
    }
    NodeMom nm;
    nm.qxx = (float)(S[0] - m * cX * cX);
    nm.qxy = (float)(S[1] - m * cX * cY);
    nm.qyy = (float)(S[2] - m * cY * cY);
    nm.aq = (float)S[8];
    nm.dx = (float)(S[3] - qtot * cX);
    nm.dy = (float)(S[4] - qtot * cY);
    nm.cxx = (float)(S[5] - 2.0 * cX * S[3] + qtot * cX * cX);
    nm.cxy = (float)(S[6] - cX * S[4] - cY * S[3] + qtot * cX * cY);
    nm.cyy = (float)(S[7] - 2.0 * cY * S[4] + qtot * cY * cY);
    nm.ims2 = nm.ias = nm.ias2 = 0.f;       // set by write_node, which knows the width
    return nm;
}

__device__ void write_node(const NodeBuild& b, int slot, int level, uint64_t path, int a, int e, int ctrl_n_tree,
                           bool divided, int child_base) {
    double m, mx, my, mu, L, px, py, qq;
    const int cnt = e - a;
    if (!divided && cnt <= 8) {
        m = mx = my = mu = L = px = py = qq = 0.0;
        for (int j = a; j < e; j++) {
            double v[kPreCols];
            b.pts.item(j, v);
            m += v[0]; mx += v[1]; my += v[2]; mu += v[3]; L += v[4]; px += v[5]; py += v[6]; qq += v[7];
        }
    } else {
        const double* P = b.pre;
        const int64_t s = b.stride;
        // a batch's prefix sums restart at each system: its first point subtracts 0
        const bool z = b.seg_first != nullptr && b.seg_first[a] == a;
        m = P[e] - (z ? 0.0 : P[a]);
        mx = P[s + e] - (z ? 0.0 : P[s + a]);
        my = P[2 * s + e] - (z ? 0.0 : P[2 * s + a]);
        mu = P[3 * s + e] - (z ? 0.0 : P[3 * s + a]);
        L = P[4 * s + e] - (z ? 0.0 : P[4 * s + a]);
        px = P[5 * s + e] - (z ? 0.0 : P[5 * s + a]);
        py = P[6 * s + e] - (z ? 0.0 : P[6 * s + a]);
        qq = P[7 * s + e] - (z ? 0.0 : P[7 * s + a]);
    }
    Node nd;
    const double comx = m > 0 ? mx / m : 0.0, comy = m > 0 ? my / m : 0.0;
    nd.comx = (float)comx; nd.comy = (float)comy;
    nd.mass = (float)m;
    nd.charge = (float)(qq * 0.1);
    nd.mu = (float)mu; nd.L = (float)L;
    nd.avx = m > 0 ? (float)(px / m) : 0.f;
    nd.avy = m > 0 ? (float)(py / m) : 0.f;
    const double size = b.root_size / (double)(1ull << level);
    nd.size2 = (float)(size * size);
    nd.child = divided ? child_base : a;
    nd.count = divided ? -1 : cnt;
    nd.half = (float)(0.5 * size);
    // box centre from the path: de-interleave the level's quadrant digits
    uint32_t ix = 0, iy = 0;
    for (int l = 0; l < level; l++) {
        const uint32_t dgt = (uint32_t)(path >> (2 * (level - 1 - l))) & 3u;
        ix = (ix << 1) | (dgt & 1u);
        iy = (iy << 1) | (dgt >> 1);
    }
    nd.bx = (float)(b.xmin + (ix + 0.5) * size * b.aspect_x);
    nd.by = (float)(b.ymin + (iy + 0.5) * size * b.aspect_y);
    nd.level = level;
    nd.pad = 0.f;
    b.nodes[slot] = nd;
    b.mass64[slot] = m; b.comx64[slot] = comx; b.comy64[slot] = comy; b.qq64[slot] = qq;
    b.path[slot] = path;
    b.npts[slot] = cnt;
    b.first[slot] = a;
    if (b.mom != nullptr && divided) {
        NodeMom nm = node_moments(b, a, e, m, comx, comy, qq * kChargeQuantum);
        const double w = 2.0 * (double)nd.half, M = nd.mass, A = nm.aq;
        nm.ims2 = M > 0 ? (float)(1.0 / (M * w * w)) : 0.f;
        nm.ias = A > 0 ? (float)(1.0 / (A * w)) : 0.f;
        nm.ias2 = A > 0 ? (float)(1.0 / (A * w * w)) : 0.f;
        b.mom[slot] = nm;
    }
}

__device__ __forceinline__ int divided_rank(const NodeBuild& b, int a, int level) {
    const int ca = a == 0 ? -1 : cpl_levels(b.keys[a - 1], b.keys[a], b.levels);
    return b.rank_start[a] + (level - ca - 1);
}

__global__ void k_tree_nodes(NodeBuild b, Ctrl* __restrict__ ctrl) {
    const int ndiv = ctrl->n_divided;
    const int n = ctrl->n_tree;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0 && !b.batch) {
        // root (quadtree.js:412-419): divided iff more than 4 points were inserted
        const bool div = n > kLeafCap;
        write_node(b, 0, 0, 0ull, 0, n, n, div, 1);
        if (1 + 4ll * ndiv > 4ll * b.cap_div + 1) ctrl->overflow = 1;
    }
    for (int r = gid; r < ndiv && r < b.cap_div; r += gridDim.x * blockDim.x) {
        const int2 dn = b.divnode[r];
        const int lo = dn.x, level = dn.y;
        const int L = b.levels;
        const uint64_t P = prefix_at(b.keys[lo], level, L);
        // end of this node's range: galloping search for the first key with a larger prefix
        int step = 1, hi = lo + 1;
        while (hi < n && prefix_at(b.keys[hi], level, L) == P) { hi = lo + step; step <<= 1; if (hi > n) hi = n; }
        int lo_b = lo, hi_b = hi;
        // now prefix(keys[hi_b]) != P or hi_b == n; binary search the boundary in [lo_b, hi_b)
        hi = lower_bound_prefix(b.keys, lo_b, hi_b, level, P + 1, L);
        const int cl = level + 1;
        int bound[5];
        bound[0] = lo;
        bound[4] = hi;
        for (int q = 1; q < 4; q++) bound[q] = lower_bound_prefix(b.keys, lo, hi, cl, (P << 2) | (uint64_t)q, L);
        for (int q = 0; q < 4; q++) {
            const int a = bound[q], e = bound[q + 1];
            const int cnt = e - a;
            bool div = cnt > kLeafCap && cl <= L - 1;
            int child_base = 0;
            if (div) {
                const int rc = divided_rank(b, a, cl);
                if (rc >= b.cap_div) {      // node pool exhausted: keep it a (large) leaf
                    div = false;
                    ctrl->overflow = 1;
                } else {
                    child_base = 1 + 4 * rc;
                }
            }
            write_node(b, 1 + 4 * r + q, cl, (P << 2) | (uint64_t)q, a, e, n, div, child_base);
        }
    }
}

void build_tree(hipStream_t s, TreeBufs& tb, const float* x, const float* y, const float* m, const float* wx,
                const float* wy, const float* mu, const float* L, const int32_t* qq, int n_upper, KeyParams kp,
                Ctrl* ctrl, bool moments) {
    // n_tree (points inside the root box) from the sorted keys
    hipLaunchKernelGGL(k_count_tree, dim3(1), dim3(1), 0, s, tb.keys, ctrl);
    const int nb = (n_upper + kScanTile - 1) / kScanTile;
    PreIn in{x, y, m, wx, wy, mu, L, qq};
    // origin of the moment prefix sums: the centre of the root box
    MomIn mom_in{x, y, m, qq, kp.xmin + kp.root_hw, kp.ymin + kp.root_hw * (kp.scale_x / kp.scale_y)};
    const int64_t stride = tb.cap_particles + 1;
    if (nb > 0) {
        prefix_columns(s, in, tb.scan_tmp64, tb.pre, stride, nb, ctrl);
        if (moments) prefix_columns(s, mom_in, tb.scan_tmp64, tb.pre_mom, stride, nb, ctrl);
        hipLaunchKernelGGL(k_tree_levels, dim3((n_upper + 255) / 256), dim3(256), 0, s, tb.keys, tb.owned, ctrl);
        exclusive_scan_i32(s, tb.owned, tb.rank_start, n_upper, &ctrl->n_tree, &ctrl->n_divided, tb.scan_tmp);
        hipLaunchKernelGGL(k_tree_scatter, dim3((n_upper + 255) / 256), dim3(256), 0, s, tb.keys, tb.rank_start,
                           tb.divnode, tb.cap_divided, ctrl);
    } else {
        hipMemsetAsync(&ctrl->n_divided, 0, sizeof(int32_t), s);
    }
    NodeBuild b;
    b.keys = tb.keys; b.rank_start = tb.rank_start; b.divnode = tb.divnode;
    b.pre = tb.pre; b.stride = stride; b.pts = in;
    b.nodes = tb.nodes; b.mass64 = tb.node_mass64; b.comx64 = tb.node_comx64; b.comy64 = tb.node_comy64;
    b.qq64 = tb.node_qq64; b.path = tb.node_path; b.npts = tb.node_npts;
    b.first = tb.node_first;
    b.seg_first = nullptr;
    b.batch = false;
    b.mom = moments ? tb.node_mom : nullptr;
    b.pre_mom = tb.pre_mom;
    b.mom_in = mom_in;
    b.root_size = 2.0 * kp.root_hw;
    b.xmin = kp.xmin;
    b.ymin = kp.ymin;
    b.aspect_x = 1.0;
    b.aspect_y = (kp.scale_x / kp.scale_y);
    b.cap_div = tb.cap_divided;
    b.levels = kKeyLevels;
    int blocks = (int)((tb.cap_divided + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_tree_nodes, dim3(blocks), dim3(256), 0, s, b, ctrl);
}


// ---------------------------------------------------------------------------
// Batches: many independent systems in one sorted array.  Each system owns a
// slice of slots aligned to kScanTile; its keys carry the system id above the
// outside bit, so one sort orders by (system, inside, Morton), prefix sums
// restart at every slice, and the tree of one system never depends on the
// others (a system's results are the same alone or in any batch).
// ---------------------------------------------------------------------------
constexpr int kBatchSysShift = 2 * kBatchKeyLevels + 1;    // system id above the outside bit
static_assert(kScanTile == kBatchAlign, "batch slices must align to the prefix-scan tile");

__global__ void k_batch_keys(const float* __restrict__ x, const float* __restrict__ y,
                             const int32_t* __restrict__ valid, const int32_t* __restrict__ sys, KeyParams kp,
                             uint64_t* __restrict__ keys, uint32_t* __restrict__ perm, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double px = (double)x[i], py = (double)y[i];
    const double span_x = (double)(1u << kBatchKeyLevels) / kp.scale_x;
    const double span_y = (double)(1u << kBatchKeyLevels) / kp.scale_y;
    const bool inside = valid[i] && px >= kp.xmin && px <= kp.xmin + span_x && py >= kp.ymin && py <= kp.ymin + span_y;
    uint64_t key = 1ull << (2 * kBatchKeyLevels);
    if (inside) {
        const uint32_t cx = cell_index(px, kp.xmin, kp.scale_x, kBatchKeyLevels);
        const uint32_t cy = cell_index(py, kp.ymin, kp.scale_y, kBatchKeyLevels);
        key = spread2(cx) | (spread2(cy) << 1);
    }
    keys[i] = key | ((uint64_t)sys[i] << kBatchSysShift);
    perm[i] = (uint32_t)i;
}

// Per system: points inside the root box (the first outside key of its slice).
__global__ void k_batch_count(const uint64_t* __restrict__ keys, const int32_t* __restrict__ soff,
                              int32_t* __restrict__ n_in, int S, const Ctrl* __restrict__ ctrl) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || ctrl->n_tree == 0) return;       // no system active: nothing to build
    const uint64_t out_bit = 1ull << (2 * kBatchKeyLevels);
    int lo = soff[s], hi = soff[s + 1];
    const int a = lo;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] & out_bit) hi = mid; else lo = mid + 1;
    }
    n_in[s] = lo - a;
}

__global__ void k_batch_levels(const uint64_t* __restrict__ keys, const int32_t* __restrict__ soff,
                               int32_t* __restrict__ owned, int32_t* __restrict__ seg_first, Ctrl* __restrict__ ctrl) {
    const int total = ctrl->n_tree;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t k = keys[i];
    seg_first[i] = soff[k >> kBatchSysShift];
    if (k & (1ull << (2 * kBatchKeyLevels))) { owned[i] = 0; return; }
    const int c = i == 0 ? -1 : cpl_levels(keys[i - 1], k, kBatchKeyLevels);
    int D = -1;
    if (i + kLeafCap < total) {
        const int cp = cpl_levels(k, keys[i + kLeafCap], kBatchKeyLevels);
        if (cp == kBatchKeyLevels && c < kBatchKeyLevels) atomicAdd((unsigned long long*)&ctrl->depth_overflow, 1ull);
        D = min(cp, kBatchKeyLevels - 1);
    }
    owned[i] = max(0, D - c);
}

__global__ void k_batch_scatter(const uint64_t* __restrict__ keys, const int32_t* __restrict__ rank_start,
                                int2* __restrict__ divnode, int64_t cap_div, Ctrl* __restrict__ ctrl) {
    const int total = ctrl->n_tree;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t k = keys[i];
    if (k & (1ull << (2 * kBatchKeyLevels))) return;
    const int c = i == 0 ? -1 : cpl_levels(keys[i - 1], k, kBatchKeyLevels);
    int D = -1;
    if (i + kLeafCap < total) D = min(cpl_levels(k, keys[i + kLeafCap], kBatchKeyLevels), kBatchKeyLevels - 1);
    int r = rank_start[i];
    for (int d = c + 1; d <= D; d++, r++) {
        if (r < cap_div) divnode[r] = make_int2(i, d);
        else ctrl->overflow = 1;
    }
}

// The root of each system (quadtree.js:412-419) at slot root_base + s.
__global__ void k_batch_roots(NodeBuild b, const int32_t* __restrict__ soff, const int32_t* __restrict__ n_in,
                              int32_t* __restrict__ root_slot, int64_t root_base, int S, Ctrl* __restrict__ ctrl) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || ctrl->n_tree == 0) return;
    const int a = soff[s], e = a + n_in[s];
    bool div = e - a > kLeafCap;
    int child_base = 0;
    if (div) {
        const int r = b.rank_start[a];          // the level-0 node owned by the system's first point
        if (r >= b.cap_div) { div = false; ctrl->overflow = 1; }
        else child_base = 1 + 4 * r;
    }
    write_node(b, (int)(root_base + s), 0, (uint64_t)s << 1, a, e, e, div, child_base);
    root_slot[s] = (int32_t)(root_base + s);
}

// ctrl->n_tree is the slot count, or 0 when no system is active (then every
// kernel of the build returns at once: the graph replays idle substeps cheaply).
void build_tree_batch(hipStream_t s, TreeBufs& tb, const BatchTreeArgs& ba, const float* x, const float* y,
                      const float* m, const float* wx, const float* wy, const float* mu, const float* L,
                      const int32_t* qq, KeyParams kp, Ctrl* ctrl) {
    const int total = ba.total;
    const dim3 gt((total + 255) / 256), bt(256), gs((ba.S + 255) / 256);
    hipLaunchKernelGGL(k_batch_count, gs, bt, 0, s, tb.keys, ba.soff, ba.n_in, ba.S, ctrl);
    const int nb = (total + kScanTile - 1) / kScanTile;
    PreIn in{x, y, m, wx, wy, mu, L, qq};
    const int64_t stride = tb.cap_particles + 1;
    // FP64 prefix sums with a segmented block-sum scan: each system restarts at 0
    hipLaunchKernelGGL(k_pre_reduce<PreIn>, dim3(nb), dim3(kScanThreads), 0, s, in, tb.scan_tmp64, nb, ctrl);
    hipLaunchKernelGGL(k_pre_bsums_seg, gs, bt, 0, s, tb.scan_tmp64, nb, kPreCols, ba.soff, ba.S, ctrl);
    hipLaunchKernelGGL(k_pre_apply<PreIn>, dim3(nb), dim3(kScanThreads), 0, s, in, tb.scan_tmp64, nb, tb.pre, stride,
                       ctrl);
    hipLaunchKernelGGL(k_batch_levels, gt, bt, 0, s, tb.keys, ba.soff, tb.owned, ba.seg_first, ctrl);
    exclusive_scan_i32(s, tb.owned, tb.rank_start, total, &ctrl->n_tree, &ctrl->n_divided, tb.scan_tmp);
    hipLaunchKernelGGL(k_batch_scatter, gt, bt, 0, s, tb.keys, tb.rank_start, tb.divnode, tb.cap_divided, ctrl);
    NodeBuild b;
    b.keys = tb.keys; b.rank_start = tb.rank_start; b.divnode = tb.divnode;
    b.pre = tb.pre; b.stride = stride; b.pts = in;
    b.nodes = tb.nodes; b.mass64 = tb.node_mass64; b.comx64 = tb.node_comx64; b.comy64 = tb.node_comy64;
    b.qq64 = tb.node_qq64; b.path = tb.node_path; b.npts = tb.node_npts;
    b.first = tb.node_first;
    b.seg_first = ba.seg_first;
    b.batch = true;
    b.levels = kBatchKeyLevels;
    b.mom = nullptr;
    b.pre_mom = nullptr;
    b.root_size = 2.0 * kp.root_hw;
    b.xmin = kp.xmin;
    b.ymin = kp.ymin;
    b.aspect_x = 1.0;
    b.aspect_y = (kp.scale_x / kp.scale_y);
    b.cap_div = tb.cap_divided;
    hipLaunchKernelGGL(k_batch_roots, gs, bt, 0, s, b, ba.soff, ba.n_in, ba.root_slot, 1 + 4 * tb.cap_divided, ba.S,
                       ctrl);
    int blocks = (int)((tb.cap_divided + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_tree_nodes, dim3(blocks), dim3(256), 0, s, b, ctrl);
}

void launch_batch_keys(hipStream_t s, const float* x, const float* y, const int32_t* valid, const int32_t* sys,
                       KeyParams kp, uint64_t* keys, uint32_t* perm, int total) {
    hipLaunchKernelGGL(k_batch_keys, dim3((total + 255) / 256), dim3(256), 0, s, x, y, valid, sys, kp, keys, perm,
                       total);
}

}  // namespace geon
