// k_field.hip — Higgs and axion scalar fields on a G x G grid (G <= 128) for
// gfx950: PQS deposit, LDS-resident FFT self-gravity, Stormer-Verlet KDK
// stencils and the particle-side gathers (forces, mass modulation, axMod).
//
// Behaviour follows src/scalar-field.js, src/higgs-field.js and
// src/axion-field.js (file:line at each kernel).  Two deliberate differences
// from the CPU path, both order effects only:
//   * the deposit accumulates in 64-bit fixed point (2^32 units) with atomics,
//     so the grid sums are bit-identical run to run whatever the particle order;
//   * a half-kick reads the old fieldDot everywhere (Jacobi) where the CPU loop
//     updates fieldDot in place in raster order, so its viscosity term sees
//     already-kicked left/top neighbours.
#include "geon_kernels.h"

namespace geon {

constexpr double kFix = 4294967296.0;          // 2^32 fixed-point scale of the deposit
constexpr float kFieldMax = 2.f;               // SCALAR_FIELD_MAX (config.js:85)
constexpr float kSelfGravPhiMax = 0.25f;       // SELFGRAV_PHI_MAX (config.js:92)
constexpr float kHiggsCoupling = 1.f;          // HIGGS_COUPLING (config.js:89)
constexpr float kAxionCoupling = 0.05f;        // AXION_COUPLING (config.js:96)
constexpr float kPortal = 0.01f;               // HIGGS_AXION_COUPLING (config.js:97)
constexpr float kHiggsMassFloor = 0.05f;       // HIGGS_MASS_FLOOR (config.js:90)
constexpr float kHiggsMaxDelta = 8.f;          // HIGGS_MASS_MAX_DELTA (config.js:91)

// _nb (scalar-field.js:78-108): loop (torus; Klein bottle and RP^2 mirror the
// other index across a glide edge), bounce (clamp), despawn (-1 = Dirichlet).
__device__ __forceinline__ int nb_index(int G, int nx, int ny, int bc, int topo) {
    if (bc == 2) {
        if (topo == 0) {
            nx = nx < 0 ? nx + G : (nx >= G ? nx - G : nx);
            ny = ny < 0 ? ny + G : (ny >= G ? ny - G : ny);
        } else if (topo == 1) {
            nx = nx < 0 ? nx + G : (nx >= G ? nx - G : nx);
            if (ny < 0) { ny += G; nx = G - 1 - nx; } else if (ny >= G) { ny -= G; nx = G - 1 - nx; }
        } else {
            if (nx < 0) { nx += G; ny = G - 1 - ny; } else if (nx >= G) { nx -= G; ny = G - 1 - ny; }
            if (ny < 0) { ny += G; nx = G - 1 - nx; } else if (ny >= G) { ny -= G; nx = G - 1 - nx; }
        }
        return ny * G + nx;
    }
    if (bc == 1) {
        nx = min(max(nx, 0), G - 1);
        ny = min(max(ny, 0), G - 1);
        return ny * G + nx;
    }
    if (nx < 0 || nx >= G || ny < 0 || ny >= G) return -1;
    return ny * G + nx;
}

// _pqsCoords (scalar-field.js:114-140)
__device__ __forceinline__ void pqs_weights(float x, float y, float icw, float ich, int& ix, int& iy, float* wx,
                                            float* wy) {
    const float gx = x * icw - 0.5f, gy = y * ich - 0.5f;
    const float fx = floorf(gx), fy = floorf(gy);
    ix = (int)fx;
    iy = (int)fy;
    const float dx = gx - fx, tx = 1.f - dx, dx2 = dx * dx, dx3 = dx2 * dx;
    wx[0] = tx * tx * tx * (1.f / 6.f);
    wx[1] = (4.f - 6.f * dx2 + 3.f * dx3) * (1.f / 6.f);
    wx[2] = (1.f + 3.f * dx + 3.f * dx2 - 3.f * dx3) * (1.f / 6.f);
    wx[3] = dx3 * (1.f / 6.f);
    const float dy = gy - fy, ty = 1.f - dy, dy2 = dy * dy, dy3 = dy2 * dy;
    wy[0] = ty * ty * ty * (1.f / 6.f);
    wy[1] = (4.f - 6.f * dy2 + 3.f * dy3) * (1.f / 6.f);
    wy[2] = (1.f + 3.f * dy + 3.f * dy2 - 3.f * dy3) * (1.f / 6.f);
    wy[3] = dy3 * (1.f / 6.f);
}

// Sum of arr over the 4x4 stencil; `outside` at Dirichlet nodes
// (interpolate / gradient / gravity gather, scalar-field.js:284-415, 665-722).
__device__ __forceinline__ float pqs_gather(const float* arr, int G, int bc, int topo, int ix, int iy, const float* wx,
                                            const float* wy, float outside) {
    float v = 0.f;
    for (int jy = 0; jy < 4; jy++) {
        float row = 0.f;
        for (int jx = 0; jx < 4; jx++) {
            const int idx = nb_index(G, ix + jx - 1, iy + jy - 1, bc, topo);
            row += (idx >= 0 ? arr[idx] : outside) * wx[jx];
        }
        v += row * wy[jy];
    }
    return v;
}

// One particle's deposit values: Higgs source g*baseMass and thermal KE
// (higgs-field.js:117-143); axion g*q^2 (Coulomb) + g*m (Yukawa)
// (axion-field.js:196-210).  False when the particle deposits nothing.
__device__ __forceinline__ bool deposit_values(int i, const float* __restrict__ m, const float* __restrict__ base_mass,
                                               const float* __restrict__ q, const float* __restrict__ wx,
                                               const float* __restrict__ wy, const float* __restrict__ vx,
                                               const float* __restrict__ vy, int higgs, int rel, int coul, int yuk,
                                               float& s0, float& s1) {
    s0 = 0.f;
    s1 = 0.f;
    if (higgs) {
        const float bm = base_mass[i];
        if (bm < 1e-12f) return false;
        float ke;
        if (rel) {
            const float w2 = wx[i] * wx[i] + wy[i] * wy[i];
            ke = w2 / (sqrtf(1.f + w2) + 1.f) * m[i];
        } else {
            ke = 0.5f * m[i] * (vx[i] * vx[i] + vy[i] * vy[i]);
        }
        s0 = kHiggsCoupling * bm;
        if (ke >= 1e-12f) s1 = ke;
        return true;
    }
    if (coul) { const float q2 = q[i] * q[i]; if (q2 > 1e-12f) s0 += kAxionCoupling * q2; }
    if (yuk && m[i] > 1e-12f) s0 += kAxionCoupling * m[i];
    return s0 != 0.f;
}

__device__ __forceinline__ unsigned long long fixed_point(float value, float wx, float wy) {
    const double v = (double)value * (double)wx * (double)wy;
    return (unsigned long long)(long long)llrint(v * kFix);
}

// The deposit of kDepItems consecutive particles per thread, kDepThreads per
// block.  Particles are in Morton order, so a block's particles cover a small
// patch of the grid: the block sums its stencils in an LDS tile over the
// patch (with 64-bit LDS atomics) and then adds each touched tile cell to the
// grid with one global atomic.  Every stencil value is rounded to 2^-32 on
// its own and the sums are of integers, so the grid is bit-identical to
// depositing each stencil value straight into the grid.  A block whose patch
// exceeds the tile deposits straight into the grid.
constexpr int kDepThreads = 256;
constexpr int kDepItems = 4;
constexpr int kDepTile = 1024;        // tile cells per quantity (8 KB)

__global__ __launch_bounds__(kDepThreads) void k_field_deposit(
        FieldArgs f, const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ m,
        const float* __restrict__ base_mass, const float* __restrict__ q, const float* __restrict__ wx,
        const float* __restrict__ wy, const float* __restrict__ vx, const float* __restrict__ vy, int higgs, int rel,
        int coul, int yuk, const Ctrl* __restrict__ ctrl) {
    __shared__ unsigned long long tile[2][kDepTile];
    __shared__ int s_box[4];
    const int n = ctrl->n_alive;
    const int base = blockIdx.x * kDepThreads * kDepItems;
    if (threadIdx.x == 0) { s_box[0] = INT_MAX; s_box[1] = INT_MAX; s_box[2] = INT_MIN; s_box[3] = INT_MIN; }
    __syncthreads();
    int ixs[kDepItems], iys[kDepItems];
    float vals0[kDepItems], vals1[kDepItems];
    bool use[kDepItems];
    int lx = INT_MAX, ly = INT_MAX, hx = INT_MIN, hy = INT_MIN;
    for (int k = 0; k < kDepItems; k++) {
        const int i = base + k * kDepThreads + threadIdx.x;
        use[k] = i < n && deposit_values(i, m, base_mass, q, wx, wy, vx, vy, higgs, rel, coul, yuk, vals0[k],
                                         vals1[k]);
        ixs[k] = iys[k] = 0;
        if (use[k]) {
            const float gx = x[i] * f.icw - 0.5f, gy = y[i] * f.ich - 0.5f;   // pqs_weights' cell
            ixs[k] = (int)floorf(gx);
            iys[k] = (int)floorf(gy);
            lx = min(lx, ixs[k]); hx = max(hx, ixs[k]);
            ly = min(ly, iys[k]); hy = max(hy, iys[k]);
        }
    }
    if (lx != INT_MAX) {
        atomicMin(&s_box[0], lx); atomicMin(&s_box[1], ly);
        atomicMax(&s_box[2], hx); atomicMax(&s_box[3], hy);
    }
    __syncthreads();
    if (s_box[0] == INT_MAX) return;                  // nothing to deposit in this block (uniform)
    // stencil cells run from ix - 1 to ix + 2
    const long long tw_l = (long long)s_box[2] - s_box[0] + 4, th_l = (long long)s_box[3] - s_box[1] + 4;
    const bool tiled = tw_l * th_l <= kDepTile;
    const int tw = tiled ? (int)tw_l : 0;
    const int x0 = s_box[0] - 1, y0 = s_box[1] - 1;
    const int nq = higgs ? 2 : 1;
    if (tiled) {
        for (int c = threadIdx.x; c < nq * kDepTile; c += kDepThreads) tile[c / kDepTile][c % kDepTile] = 0ull;
        __syncthreads();
    }
    for (int k = 0; k < kDepItems; k++) {
        if (!use[k]) continue;
        const int i = base + k * kDepThreads + threadIdx.x;
        int ix, iy;
        float pw[4], qw[4];
        pqs_weights(x[i], y[i], f.icw, f.ich, ix, iy, pw, qw);
        // the tile covers the cells the first pass found; a particle whose
        // cell came out otherwise here goes straight to the grid
        const bool in_tile = tiled && ix == ixs[k] && iy == iys[k];
        for (int jy = 0; jy < 4; jy++)
            for (int jx = 0; jx < 4; jx++) {
                const unsigned long long a0 = fixed_point(vals0[k], pw[jx], qw[jy]);
                const unsigned long long a1 = vals1[k] != 0.f ? fixed_point(vals1[k], pw[jx], qw[jy]) : 0ull;
                if (in_tile) {
                    const int t = (iy + jy - 1 - y0) * tw + (ix + jx - 1 - x0);
                    atomicAdd(&tile[0][t], a0);
                    if (a1) atomicAdd(&tile[1][t], a1);
                } else {
                    const int idx = nb_index(f.G, ix + jx - 1, iy + jy - 1, f.bc, f.topo);
                    if (idx < 0) continue;
                    atomicAdd(f.dep_src + idx, a0);
                    if (a1) atomicAdd(f.dep_th + idx, a1);
                }
            }
    }
    if (!tiled) return;                               // uniform over the block
    __syncthreads();
    for (int t = threadIdx.x; t < (int)(tw_l * th_l); t += kDepThreads) {
        const unsigned long long a0 = tile[0][t], a1 = nq == 2 ? tile[1][t] : 0ull;
        if (a0 == 0ull && a1 == 0ull) continue;
        const int idx = nb_index(f.G, x0 + t % tw, y0 + t / tw, f.bc, f.topo);
        if (idx < 0) continue;                        // Dirichlet border: the value leaves the grid
        if (a0) atomicAdd(f.dep_src + idx, a0);
        if (a1) atomicAdd(f.dep_th + idx, a1);
    }
}

// Fixed point -> float, and clear the accumulators for the next deposit.
__global__ void k_field_finalize(FieldArgs f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= f.G * f.G) return;
    f.source[i] = (float)((double)(long long)f.dep_src[i] / kFix);
    f.thermal[i] = (float)((double)(long long)f.dep_th[i] / kFix);
    f.dep_src[i] = 0ull;
    f.dep_th[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Self-gravity: rho -> FFT -> x G^ -> IFFT -> Phi, grad Phi, all in one
// workgroup with the complex grid resident in LDS (scalar-field.js:600-660).
// ---------------------------------------------------------------------------
constexpr int kFftMax = 128;
constexpr int kFftThreads = 1024;
constexpr int kPitch = kFftMax + 1;            // padded row pitch: column passes hit distinct banks

struct FftLds {
    float re[kFftMax * kPitch];
    float im[kFftMax * kPitch];
    float tw_re[kFftMax / 2], tw_im[kFftMax / 2];
};

// In-place radix-2 DIT over all rows (along_x) or all columns of the LDS grid.
// Twiddles W_N^k = exp(-2 pi i k / N) come from a table computed directly per
// index (not by recurrence).
__device__ void fft_pass(FftLds& s, int N, int logN, bool along_x, bool inverse) {
    const int t = threadIdx.x;
    for (int e = t; e < N * N; e += kFftThreads) {
        const int line = e >> logN, k = e & (N - 1);
        const int j = (int)(__brev((unsigned)k) >> (32 - logN));
        if (k < j) {
            const int a = along_x ? line * kPitch + k : k * kPitch + line;
            const int b = along_x ? line * kPitch + j : j * kPitch + line;
            const float tr = s.re[a], ti = s.im[a];
            s.re[a] = s.re[b]; s.im[a] = s.im[b];
            s.re[b] = tr; s.im[b] = ti;
        }
    }
    __syncthreads();
    // N is a power of two: line, group and position by shifts and masks
    for (int lg = 1; lg <= logN; lg++) {
        const int len = 1 << lg, half = len >> 1, stride = N >> lg;
        for (int e = t; e < N * (N / 2); e += kFftThreads) {
            const int line = e >> (logN - 1), k = e & (N / 2 - 1);
            const int grp = k >> (lg - 1), jj = k & (half - 1);
            const int ia = grp * len + jj, ib = ia + half;
            const int a = along_x ? line * kPitch + ia : ia * kPitch + line;
            const int b = along_x ? line * kPitch + ib : ib * kPitch + line;
            const float wr = s.tw_re[jj * stride];
            const float wi = inverse ? -s.tw_im[jj * stride] : s.tw_im[jj * stride];
            const float br = s.re[b], bi = s.im[b];
            const float tr = wr * br - wi * bi, ti = wr * bi + wi * br;
            const float ar = s.re[a], ai = s.im[a];
            s.re[b] = ar - tr; s.im[b] = ai - ti;
            s.re[a] = ar + tr; s.im[a] = ai + ti;
        }
        __syncthreads();
    }
    if (inverse) {
        const float inv = 1.f / (float)N;          // fft.js:63-69 normalizes each 1-D inverse
        for (int e = t; e < N * N; e += kFftThreads) {
            const int a = (e >> logN) * kPitch + (e & (N - 1));
            s.re[a] *= inv; s.im[a] *= inv;
        }
        __syncthreads();
    }
}

__device__ void fft2d_lds(FftLds& s, int N, int logN, bool inverse) {
    fft_pass(s, N, logN, true, inverse);
    fft_pass(s, N, logN, false, inverse);
}

__device__ void fft_twiddles(FftLds& s, int N) {
    for (int k = threadIdx.x; k < N / 2; k += kFftThreads) {
        double sn, cs;
        sincospi(-2.0 * (double)k / (double)N, &sn, &cs);
        s.tw_re[k] = (float)cs;
        s.tw_im[k] = (float)sn;
    }
}

// mode 1: build G^ = FFT(-1/sqrt(r^2+eps^2)) (scalar-field.js:524-560; the
// min-image and wrapped-index distance grids of the reference coincide, so one
// grid serves both boundary modes); modes 2, 3: a plain 2-D FFT of f.green.
// (Once per grid, or for tests: the per-substep self-gravity runs over many
// workgroups below.)
__global__ __launch_bounds__(kFftThreads) void k_field_selfgrav(FieldArgs f, int mode, float soft2) {
    __shared__ FftLds s;
    const int N = f.G, logN = 31 - __clz(N);
    const int t = threadIdx.x;
    fft_twiddles(s, N);
    __syncthreads();
    if (mode == 1) {
        const bool glide = f.bc == 2 && f.topo != 0;
        const double W = (double)f.cw * N, H = (double)f.ch * N;
        for (int e = t; e < N * N; e += kFftThreads) {
            const int iy = e >> logN, ix = e & (N - 1);
            double ddx, ddy;
            if (glide) {
                // Klein / RP^2: the reference takes minImage of the grid offset (scalar-field.js:538-548)
                min_image_t<double>(1u, (uint32_t)f.topo, W, H, 0.0, 0.0, ix * (double)f.cw, iy * (double)f.ch, ddx, ddy);
            } else {
                ddx = (float)(ix <= N / 2 ? ix : ix - N) * f.cw;
                ddy = (float)(iy <= N / 2 ? iy : iy - N) * f.ch;
            }
            s.re[iy * kPitch + ix] = (float)(-1.0 / sqrt(ddx * ddx + ddy * ddy + (double)soft2));
            s.im[iy * kPitch + ix] = 0.f;
        }
        __syncthreads();
        fft2d_lds(s, N, logN, false);
        for (int e = t; e < N * N; e += kFftThreads) {
            const int a = (e >> logN) * kPitch + (e & (N - 1));
            f.green[e] = make_float2(s.re[a], s.im[a]);
        }
        return;
    }
    if (mode == 2 || mode == 3) {
        // plain 2-D FFT of the complex grid in f.green, in place (2 forward, 3 inverse)
        for (int e = t; e < N * N; e += kFftThreads) {
            const float2 v = f.green[e];
            s.re[(e >> logN) * kPitch + (e & (N - 1))] = v.x;
            s.im[(e >> logN) * kPitch + (e & (N - 1))] = v.y;
        }
        __syncthreads();
        fft2d_lds(s, N, logN, mode == 3);
        for (int e = t; e < N * N; e += kFftThreads) {
            const int a = (e >> logN) * kPitch + (e & (N - 1));
            f.green[e] = make_float2(s.re[a], s.im[a]);
        }
    }
}

// ---------------------------------------------------------------------------
// Self-gravity of a field (scalar-field.js:567-660): rho -> FFT -> x G^ ->
// IFFT -> Phi, grad Phi, over one workgroup per grid line, so that the grid's
// global reads and writes spread over as many CUs as it has lines:
//   k_sg_rows_fwd  the energy density of a row and its forward FFT;
//   k_sg_cols      a column's forward FFT, the product with G^, its inverse FFT;
//   k_sg_rows_inv  a row's inverse FFT: Phi = Re;
//   k_sg_grad      central-difference gradient (outside = 0).
// The 2-D transform is separable, so running the inverse's columns before its
// rows changes only the FP32 rounding order.  A grid without energy
// (no cell >= 1e-12) gives Phi = 0 and has_energy = 0, as the reference.
// ---------------------------------------------------------------------------
struct LineLds {
    float re[kFftMax], im[kFftMax];
    float tw_re[kFftMax / 2], tw_im[kFftMax / 2];
};

// blockDim.x == N: thread t holds element t of the line.
__device__ void line_twiddles(LineLds& s, int N) {
    const int t = threadIdx.x;
    if (t < N / 2) {
        double sn, cs;
        sincospi(-2.0 * (double)t / (double)N, &sn, &cs);
        s.tw_re[t] = (float)cs;
        s.tw_im[t] = (float)sn;
    }
}

// In-place radix-2 DIT of the line in LDS, with fft_pass's butterflies.  The
// caller has synchronised the line's writes.
__device__ void line_fft(LineLds& s, int N, int logN, bool inverse) {
    const int t = threadIdx.x;
    {
        const int j = (int)(__brev((unsigned)t) >> (32 - logN));
        const float vr = s.re[j], vi = s.im[j];
        __syncthreads();
        s.re[t] = vr;
        s.im[t] = vi;
        __syncthreads();
    }
    for (int lg = 1; lg <= logN; lg++) {
        const int len = 1 << lg, half = len >> 1, stride = N >> lg;
        if (t < N / 2) {
            const int grp = t >> (lg - 1), jj = t & (half - 1);
            const int ia = grp * len + jj, ib = ia + half;
            const float wr = s.tw_re[jj * stride];
            const float wi = inverse ? -s.tw_im[jj * stride] : s.tw_im[jj * stride];
            const float br = s.re[ib], bi = s.im[ib];
            const float tr = wr * br - wi * bi, ti = wr * bi + wi * br;
            const float ar = s.re[ia], ai = s.im[ia];
            s.re[ib] = ar - tr; s.im[ib] = ai - ti;
            s.re[ia] = ar + tr; s.im[ia] = ai + ti;
        }
        __syncthreads();
    }
    if (inverse) {
        const float inv = 1.f / (float)N;          // fft.js:63-69 normalizes each 1-D inverse
        s.re[t] *= inv;
        s.im[t] *= inv;
        __syncthreads();
    }
}

// Whether any row of the grid holds energy: every block reduces the row flags
// (blockDim.x >= N).
__device__ __forceinline__ bool sg_has_energy(const FieldArgs& f) {
    return __syncthreads_or(threadIdx.x < (unsigned)f.G ? f.sg_rows[threadIdx.x] : 0) != 0;
}

__global__ __launch_bounds__(kFftMax) void k_sg_rows_fwd(FieldArgs f) {
    __shared__ LineLds s;
    const int N = f.G, logN = 31 - __clz(N);
    const int y = blockIdx.x, x = threadIdx.x, e = y * N + x;
    line_twiddles(s, N);
    // energy density (scalar-field.js:486-516) times the cell area
    const float icw2 = f.icw * f.icw, ich2 = f.ich * f.ich, area = f.cw * f.ch;
    const float fd = f.field_dot[e], gx = f.grad_x[e], gy = f.grad_y[e], v = f.field[e];
    float pot;
    if (f.higgs) {
        const float v2 = v * v;
        pot = f.mu2 * (-0.5f * v2 + 0.25f * v2 * v2) + 0.25f * f.mu2;
    } else {
        pot = 0.5f * f.m2 * v * v;
    }
    const float rho = 0.5f * fd * fd + 0.5f * (gx * gx * icw2 + gy * gy * ich2) + pot;
    s.re[x] = rho * area;
    s.im[x] = 0.f;
    const int any = __syncthreads_or(rho >= 1e-12f);
    if (x == 0) f.sg_rows[y] = any;
    line_fft(s, N, logN, false);
    f.sg_c[e] = make_float2(s.re[x], s.im[x]);
}

__global__ __launch_bounds__(kFftMax) void k_sg_cols(FieldArgs f) {
    __shared__ LineLds s;
    const int N = f.G, logN = 31 - __clz(N);
    const int x = blockIdx.x, y = threadIdx.x, e = y * N + x;
    if (!sg_has_energy(f)) return;                 // uniform over the block
    line_twiddles(s, N);
    const float2 v = f.sg_c[e];
    s.re[y] = v.x;
    s.im[y] = v.y;
    __syncthreads();
    line_fft(s, N, logN, false);
    const float2 g = f.green[e];
    const float ar = s.re[y], ai = s.im[y];
    s.re[y] = ar * g.x - ai * g.y;
    s.im[y] = ar * g.y + ai * g.x;
    __syncthreads();
    line_fft(s, N, logN, true);
    f.sg_c[e] = make_float2(s.re[y], s.im[y]);
}

__global__ __launch_bounds__(kFftMax) void k_sg_rows_inv(FieldArgs f) {
    __shared__ LineLds s;
    const int N = f.G, logN = 31 - __clz(N);
    const int y = blockIdx.x, x = threadIdx.x, e = y * N + x;
    if (!sg_has_energy(f)) return;                 // uniform over the block
    line_twiddles(s, N);
    const float2 v = f.sg_c[e];
    s.re[x] = v.x;
    s.im[x] = v.y;
    __syncthreads();
    line_fft(s, N, logN, true);
    f.sg_phi[e] = s.re[x];
}

__global__ __launch_bounds__(256) void k_sg_grad(FieldArgs f) {
    const int N = f.G, logN = 31 - __clz(N);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const bool has = sg_has_energy(f);             // before any thread leaves: it synchronises the block
    if (e >= N * N) return;
    if (!has) {
        f.sg_phi[e] = 0.f;
        f.sg_gx[e] = 0.f;
        f.sg_gy[e] = 0.f;
        if (e == 0) *f.has_energy = 0;
        return;
    }
    // Phi = Re, central-difference gradient (outside = 0) (scalar-field.js:567-596)
    const int iy = e >> logN, ix = e & (N - 1);
    const int iL = nb_index(N, ix - 1, iy, f.bc, f.topo), iR = nb_index(N, ix + 1, iy, f.bc, f.topo);
    const int iT = nb_index(N, ix, iy - 1, f.bc, f.topo), iB = nb_index(N, ix, iy + 1, f.bc, f.topo);
    auto P = [&](int idx) { return idx >= 0 ? f.sg_phi[idx] : 0.f; };
    f.sg_gx[e] = (P(iR) - P(iL)) * 0.5f;
    f.sg_gy[e] = (P(iB) - P(iT)) * 0.5f;
    if (e == 0) *f.has_energy = 1;
}

// Half-kick (higgs-field.js:148-218, axion-field.js:129-190): fd_out = fd_in + ddphi * halfDt.
__global__ void k_field_kick(FieldArgs f, const float* __restrict__ fd_in, float* __restrict__ fd_out,
                             const float* __restrict__ portal, int sg_on, float half_dt) {
    const int G = f.G;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G * G) return;
    if (f.ctrl) {                       // substep length from the device controller
        if (!f.ctrl->active) { fd_out[i] = fd_in[i]; return; }
        half_dt = (float)(0.5 * f.ctrl->dt_sub);
    }
    const int iy = i / G, ix = i % G, last = G - 1;
    const float icw2 = f.icw * f.icw, ich2 = f.ich * f.ich;
    const float nu = 0.5f / sqrtf(icw2 + ich2);
    const float* fld = f.field;
    const float fC = fld[i], fdC = fd_in[i];
    float lap, visc;
    {
        const int iL = nb_index(G, ix - 1, iy, f.bc, f.topo), iR = nb_index(G, ix + 1, iy, f.bc, f.topo);
        const int iT = nb_index(G, ix, iy - 1, f.bc, f.topo), iB = nb_index(G, ix, iy + 1, f.bc, f.topo);
        lap = ((iL >= 0 ? fld[iL] : f.vac) + (iR >= 0 ? fld[iR] : f.vac) - 2.f * fC) * icw2 +
              ((iT >= 0 ? fld[iT] : f.vac) + (iB >= 0 ? fld[iB] : f.vac) - 2.f * fC) * ich2;
        // viscosity: clamp-to-edge on the border whatever the boundary (scalar-field.js:254-276)
        const float fdL = ix > 0 ? fd_in[i - 1] : fdC, fdR = ix < last ? fd_in[i + 1] : fdC;
        const float fdT = iy > 0 ? fd_in[i - G] : fdC, fdB = iy < last ? fd_in[i + G] : fdC;
        visc = nu * ((fdL + fdR - 2.f * fdC) * icw2 + (fdT + fdB - 2.f * fdC) * ich2);
    }
    const float v = fC;
    const float src = f.source[i] * f.inv_area;
    float dd;
    if (f.higgs) {
        const float v3 = v * v * v;
        const float mu_eff = f.mu2 - f.thermal[i];
        dd = lap + mu_eff * v - f.mu2 * v3 - f.damp * fdC + src + visc;
        const float pt = portal ? kPortal * portal[i] * portal[i] : 0.f;
        if (sg_on) {
            const float Phi = fminf(fmaxf(f.sg_phi[i], -kSelfGravPhiMax), kSelfGravPhiMax);
            dd += 4.f * Phi * lap + 2.f * (f.sg_gx[i] * f.grad_x[i] * icw2 + f.sg_gy[i] * f.grad_y[i] * ich2) +
                  2.f * Phi * mu_eff * v - 2.f * Phi * f.mu2 * v3 - pt * v - 2.f * Phi * pt * v;
        } else {
            dd -= pt * v;
        }
    } else {
        dd = lap - f.m2 * v - f.damp * fdC + src + visc;
        const float pt = portal ? kPortal * portal[i] * portal[i] : 0.f;
        if (sg_on) {
            const float Phi = fminf(fmaxf(f.sg_phi[i], -kSelfGravPhiMax), kSelfGravPhiMax);
            dd += 4.f * Phi * lap + 2.f * (f.sg_gx[i] * f.grad_x[i] * icw2 + f.sg_gy[i] * f.grad_y[i] * ich2) -
                  2.f * Phi * f.m2 * v - pt * v - 2.f * Phi * pt * v;
        } else {
            dd -= pt * v;
        }
    }
    float nfd = fdC + dd * half_dt;
    if (!isfinite(nfd)) { nfd = 0.f; f.field[i] = f.vac; }
    fd_out[i] = nfd;
}

// Drift with clamp (higgs-field.js:96-98).
__global__ void k_field_drift(FieldArgs f, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= f.G * f.G) return;
    if (f.ctrl) {
        if (!f.ctrl->active) return;
        dt = (float)f.ctrl->dt_sub;
    }
    f.field[i] = fminf(fmaxf(f.field[i] + f.field_dot[i] * dt, -kFieldMax), kFieldMax);
}

// Central-difference grid gradients in grid units (scalar-field.js:216-245).
__global__ void k_field_grads(FieldArgs f) {
    const int G = f.G;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G * G) return;
    const int iy = i / G, ix = i % G;
    const int iL = nb_index(G, ix - 1, iy, f.bc, f.topo), iR = nb_index(G, ix + 1, iy, f.bc, f.topo);
    const int iT = nb_index(G, ix, iy - 1, f.bc, f.topo), iB = nb_index(G, ix, iy + 1, f.bc, f.topo);
    const float* fl = f.field;
    f.grad_x[i] = ((iR >= 0 ? fl[iR] : f.vac) - (iL >= 0 ? fl[iL] : f.vac)) * 0.5f;
    f.grad_y[i] = ((iB >= 0 ? fl[iB] : f.vac) - (iT >= 0 ? fl[iT] : f.vac)) * 0.5f;
}

// modulateMasses (higgs-field.js:225-270): m -> baseMass*|phi| with a rate clamp,
// w rescaled by m_old/m_new.
__global__ void k_field_modulate(FieldArgs f, const float* __restrict__ x,
                                 const float* __restrict__ y, float* __restrict__ m,
                                 const float* __restrict__ base_mass, float* __restrict__ radius,
                                 float* __restrict__ wx, float* __restrict__ wy, float* __restrict__ vx,
                                 float* __restrict__ vy, float* __restrict__ higgs_mod, float dt,
                                 const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ctrl->n_alive) return;
    if (f.ctrl) {
        if (!f.ctrl->active) return;
        dt = (float)f.ctrl->dt_sub;
    }
    const float bm = base_mass[i];
    if (bm < 1e-12f) return;
    int ix, iy;
    float pw[4], qw[4];
    pqs_weights(x[i], y[i], f.icw, f.ich, ix, iy, pw, qw);
    const float phi = pqs_gather(f.field, f.G, f.bc, f.topo, ix, iy, pw, qw, f.vac);
    higgs_mod[i] = fmaxf(fabsf(phi), kHiggsMassFloor);
    const float target = fmaxf(bm * fabsf(phi), kHiggsMassFloor * bm);
    if (target != target) return;
    const float max_delta = kHiggsMaxDelta * dt;
    const float m0 = m[i];
    const float diff = target - m0;
    const float nm = m0 + fminf(fmaxf(diff, -max_delta), max_delta);
    const float ratio = m0 / nm;
    const float nwx = wx[i] * ratio, nwy = wy[i] * ratio;
    wx[i] = nwx; wy[i] = nwy;
    m[i] = nm;
    radius[i] = cbrtf(nm);
    const float g = sqrtf(1.f + nwx * nwx + nwy * nwy);
    vx[i] = nwx / g;
    vy[i] = nwy / g;
}

// interpolateAxMod (axion-field.js:271-309): axMod = 1 + g a, yukMod = 1 + g a.
__global__ void k_field_axmod(FieldArgs f, const float* __restrict__ x, const float* __restrict__ y,
                              float* __restrict__ ax, float* __restrict__ yuk, int coul, int yukawa,
                              const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ctrl->n_alive) return;
    if (!coul && !yukawa) { ax[i] = 1.f; yuk[i] = 1.f; return; }
    int ix, iy;
    float pw[4], qw[4];
    pqs_weights(x[i], y[i], f.icw, f.ich, ix, iy, pw, qw);
    const float ga = kAxionCoupling * pqs_gather(f.field, f.G, f.bc, f.topo, ix, iy, pw, qw, f.vac);
    ax[i] = coul ? (ga > -1.f ? 1.f + ga : 0.f) : 1.f;
    yuk[i] = yukawa ? (ga > -1.f ? 1.f + ga : 0.f) : 1.f;
}

// Field forces on particles after the pair forces: Higgs gradient force
// (higgs-field.js:301-327), axion gradient force (axion-field.js:315-353) and
// field gravity F = -m grad Phi (scalar-field.js:665-722); also the
// particle-field interaction energy and m*Phi (gravPE) per particle.
__global__ void k_field_forces(FieldArgs h, FieldArgs a, int use_h, int use_a, int grav, int coul, int yukawa,
                               const float* __restrict__ x, const float* __restrict__ y,
                               const float* __restrict__ m, const float* __restrict__ base_mass,
                               const float* __restrict__ q, float* __restrict__ fx, float* __restrict__ fy,
                               float* __restrict__ pe_grav, float* __restrict__ pfi,
                               const Ctrl* __restrict__ ctrl, const int32_t* __restrict__ sel,
                               float* __restrict__ sel_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ctrl->n_alive) return;
    float Fx = 0.f, Fy = 0.f, peg = 0.f, e_pfi = 0.f;
    // by type for the selected particle: Higgs gradient, axion gradient, field
    // gravity (higgs-field.js:324, axion-field.js:350, scalar-field.js:717)
    float* so = (sel_out && i == *sel) ? sel_out : nullptr;
    if (so) for (int k = SEL_HIGGS; k < SEL_FIELD_GRAV + 2; k++) so[k] = 0.f;
    const float mi = m[i], bm = base_mass[i], qi = q[i];
    if (use_h) {
        int ix, iy;
        float pw[4], qw[4];
        pqs_weights(x[i], y[i], h.icw, h.ich, ix, iy, pw, qw);
        const float val = pqs_gather(h.field, h.G, h.bc, h.topo, ix, iy, pw, qw, h.vac);
        if (bm >= 1e-12f) {
            const float gx = pqs_gather(h.grad_x, h.G, h.bc, h.topo, ix, iy, pw, qw, 0.f) * h.icw;
            const float gy = pqs_gather(h.grad_y, h.G, h.bc, h.topo, ix, iy, pw, qw, 0.f) * h.ich;
            if (gx == gx && gy == gy) {
                const float sgn = val >= 0.f ? 1.f : -1.f;
                Fx += kHiggsCoupling * bm * sgn * gx;
                Fy += kHiggsCoupling * bm * sgn * gy;
                if (so) { so[SEL_HIGGS] = kHiggsCoupling * bm * sgn * gx; so[SEL_HIGGS + 1] = kHiggsCoupling * bm * sgn * gy; }
            }
            e_pfi -= bm * (fabsf(val) - 1.f);
        }
        if (grav && *h.has_energy && mi >= 1e-12f) {
            const float gfx = -mi * pqs_gather(h.sg_gx, h.G, h.bc, h.topo, ix, iy, pw, qw, 0.f) * h.icw;
            const float gfy = -mi * pqs_gather(h.sg_gy, h.G, h.bc, h.topo, ix, iy, pw, qw, 0.f) * h.ich;
            Fx += gfx;
            Fy += gfy;
            if (so) { so[SEL_FIELD_GRAV] += gfx; so[SEL_FIELD_GRAV + 1] += gfy; }
            peg += mi * pqs_gather(h.sg_phi, h.G, h.bc, h.topo, ix, iy, pw, qw, 0.f);
        }
    }
    if (use_a) {
        int ix, iy;
        float pw[4], qw[4];
        pqs_weights(x[i], y[i], a.icw, a.ich, ix, iy, pw, qw);
        float coupling = 0.f;
        if (coul) { const float q2 = qi * qi; if (q2 > 1e-12f) coupling += q2; }
        if (yukawa && mi > 1e-12f) coupling += mi;
        if (coupling != 0.f) {
            const float gx = pqs_gather(a.grad_x, a.G, a.bc, a.topo, ix, iy, pw, qw, 0.f) * a.icw;
            const float gy = pqs_gather(a.grad_y, a.G, a.bc, a.topo, ix, iy, pw, qw, 0.f) * a.ich;
            if (gx == gx && gy == gy) {
                Fx += kAxionCoupling * coupling * gx;
                Fy += kAxionCoupling * coupling * gy;
                if (so) { so[SEL_AXION] = kAxionCoupling * coupling * gx; so[SEL_AXION + 1] = kAxionCoupling * coupling * gy; }
            }
        }
        const float av = pqs_gather(a.field, a.G, a.bc, a.topo, ix, iy, pw, qw, a.vac);
        if (coul) e_pfi -= kAxionCoupling * qi * qi * av;
        if (yukawa && mi > 1e-12f) e_pfi -= kAxionCoupling * mi * av;
        if (grav && *a.has_energy && mi >= 1e-12f) {
            const float gfx = -mi * pqs_gather(a.sg_gx, a.G, a.bc, a.topo, ix, iy, pw, qw, 0.f) * a.icw;
            const float gfy = -mi * pqs_gather(a.sg_gy, a.G, a.bc, a.topo, ix, iy, pw, qw, 0.f) * a.ich;
            Fx += gfx;
            Fy += gfy;
            if (so) { so[SEL_FIELD_GRAV] += gfx; so[SEL_FIELD_GRAV + 1] += gfy; }
            peg += mi * pqs_gather(a.sg_phi, a.G, a.bc, a.topo, ix, iy, pw, qw, 0.f);
        }
    }
    fx[i] += Fx;
    fy[i] += Fy;
    pe_grav[i] = peg;
    pfi[i] = e_pfi;
}

// FP64 grid sums for the readout: field energy (scalar-field.js:452-478),
// momentum (773-812) and, for Higgs with an axion partner, the portal energy
// (higgs-field.js:353-364).  One block; out[0..3] = energy, px, py, portal.
__global__ __launch_bounds__(256) void k_field_energy(FieldArgs f, const float* __restrict__ partner,
                                                      double* __restrict__ out) {
    __shared__ double sh[4][256];
    const int G = f.G, last = G - 1;
    const double cw = f.cw, ch = f.ch, area = cw * ch, icw2 = 1.0 / (cw * cw), ich2 = 1.0 / (ch * ch);
    double e = 0, px = 0, py = 0, portal = 0;
    for (int i = threadIdx.x; i < G * G; i += 256) {
        const int iy = i / G, ix = i % G;
        const double v = f.field[i], fd = f.field_dot[i];
        const double fR = ix + 1 < G ? f.field[i + 1] : v, fB = iy + 1 < G ? f.field[i + G] : v;
        const double dfx = fR - v, dfy = fB - v;
        double pot;
        if (f.higgs) { const double v2 = v * v; pot = f.mu2 * (-0.5 * v2 + 0.25 * v2 * v2) + 0.25 * f.mu2; }
        else pot = 0.5 * (double)f.m2 * v * v;
        e += (0.5 * fd * fd + 0.5 * (dfx * dfx * icw2 + dfy * dfy * ich2) + pot) * area;
        const double fl = ix > 0 ? f.field[i - 1] : v, fr = ix < last ? f.field[i + 1] : v;
        const double ft = iy > 0 ? f.field[i - G] : v, fb = iy < last ? f.field[i + G] : v;
        px -= fd * (fr - fl);
        py -= fd * (fb - ft);
        if (partner) { const double a = partner[i]; portal += v * v * a * a; }
    }
    sh[0][threadIdx.x] = e; sh[1][threadIdx.x] = px; sh[2][threadIdx.x] = py; sh[3][threadIdx.x] = portal;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int c = 0; c < 4; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sh[0][0];
        out[1] = sh[1][0] * ch * 0.5;
        out[2] = sh[2][0] * cw * 0.5;
        out[3] = 0.5 * kPortal * sh[3][0] * area;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int grid_blocks(int n) { return (n + 255) / 256; }

void field_build_green(hipStream_t s, const FieldArgs& f, float soft2) {
    hipLaunchKernelGGL(k_field_selfgrav, dim3(1), dim3(kFftThreads), 0, s, f, 1, soft2);
}

void field_selfgrav(hipStream_t s, const FieldArgs& f) {
    const int N = f.G;                          // a power of two in [8, kFftMax]
    hipLaunchKernelGGL(k_sg_rows_fwd, dim3(N), dim3(N), 0, s, f);
    hipLaunchKernelGGL(k_sg_cols, dim3(N), dim3(N), 0, s, f);
    hipLaunchKernelGGL(k_sg_rows_inv, dim3(N), dim3(N), 0, s, f);
    hipLaunchKernelGGL(k_sg_grad, dim3((N * N + 255) / 256), dim3(256), 0, s, f);
}

void field_fft(hipStream_t s, const FieldArgs& f, int inverse) {
    hipLaunchKernelGGL(k_field_selfgrav, dim3(1), dim3(kFftThreads), 0, s, f, inverse ? 3 : 2, 0.f);
}

void field_deposit(hipStream_t s, const FieldArgs& f, const ParticleView& p, int n_upper, int rel, int coul, int yuk) {
    if (n_upper > 0)
        hipLaunchKernelGGL(k_field_deposit, dim3((n_upper + kDepThreads * kDepItems - 1) / (kDepThreads * kDepItems)),
                           dim3(kDepThreads), 0, s, f, p.x, p.y, p.m, p.base_mass, p.q, p.wx, p.wy, p.vx, p.vy,
                           f.higgs, rel, coul, yuk, p.ctrl);
    hipLaunchKernelGGL(k_field_finalize, dim3(grid_blocks(f.G * f.G)), dim3(256), 0, s, f);
}

void field_grads(hipStream_t s, const FieldArgs& f) {
    hipLaunchKernelGGL(k_field_grads, dim3(grid_blocks(f.G * f.G)), dim3(256), 0, s, f);
}

// One HiggsField.update / AxionField.update (higgs-field.js:39-113,
// axion-field.js:48-124) after the deposit: [self-gravity], half-kick, drift,
// [gradients, self-gravity], half-kick, gradients.
void field_kdk(hipStream_t s, FieldArgs f, const float* portal, int sg_on, float dt) {
    const int nb = grid_blocks(f.G * f.G);
    if (sg_on) field_selfgrav(s, f);
    hipLaunchKernelGGL(k_field_kick, dim3(nb), dim3(256), 0, s, f, f.field_dot, f.field_dot_alt, portal, sg_on, 0.5f * dt);
    std::swap(f.field_dot, f.field_dot_alt);
    hipLaunchKernelGGL(k_field_drift, dim3(nb), dim3(256), 0, s, f, dt);
    if (sg_on) {
        field_grads(s, f);
        field_selfgrav(s, f);
    }
    hipLaunchKernelGGL(k_field_kick, dim3(nb), dim3(256), 0, s, f, f.field_dot, f.field_dot_alt, portal, sg_on, 0.5f * dt);
    field_grads(s, f);
}

void field_modulate(hipStream_t s, const FieldArgs& f, const ParticleView& p, float* higgs_mod, float dt, int n_upper) {
    if (n_upper <= 0) return;
    hipLaunchKernelGGL(k_field_modulate, dim3(grid_blocks(n_upper)), dim3(256), 0, s, f, p.x, p.y, p.m,
                       p.base_mass, p.radius, p.wx, p.wy, p.vx, p.vy, higgs_mod, dt, p.ctrl);
}

void field_axmod(hipStream_t s, const FieldArgs& f, const ParticleView& p, float* ax, float* yuk, int coul, int yukawa,
                 int n_upper) {
    if (n_upper <= 0) return;
    hipLaunchKernelGGL(k_field_axmod, dim3(grid_blocks(n_upper)), dim3(256), 0, s, f, p.x, p.y, ax, yuk, coul, yukawa,
                       p.ctrl);
}

void field_forces(hipStream_t s, const FieldArgs& h, const FieldArgs& a, int use_h, int use_a, int grav, int coul,
                  int yukawa, const ParticleView& p, float* fx, float* fy, float* pe_grav, float* pfi, int n_upper,
                  const int32_t* sel, float* sel_out) {
    if (n_upper <= 0) return;
    hipLaunchKernelGGL(k_field_forces, dim3(grid_blocks(n_upper)), dim3(256), 0, s, h, a, use_h, use_a, grav, coul,
                       yukawa, p.x, p.y, p.m, p.base_mass, p.q, fx, fy, pe_grav, pfi, p.ctrl, sel, sel_out);
}

void field_energy(hipStream_t s, const FieldArgs& f, const float* partner, double* out) {
    hipLaunchKernelGGL(k_field_energy, dim3(1), dim3(256), 0, s, f, partner, out);
}

}  // namespace geon
