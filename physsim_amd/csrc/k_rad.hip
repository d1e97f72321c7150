// k_rad.hip — radiation reaction on gfx950: the FP64 sums of the
// Landau-Lifshitz term's radiated energy and momentum, and the quadrupole
// radiation of Physics.update (integrator.js:1567-1740) once per update.
//
// Every sum is a fixed two-level tree (256-thread blocks, then one block over
// the block sums), so results do not depend on scheduling.
#include "geon_kernels.h"

namespace geon {

constexpr int kRadThreads = 256;
constexpr int kRadCols = 9;

__device__ __forceinline__ double block_sum_f64(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = kRadThreads / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// partial[c * nb + b] = sum of column c over block b's particles.
__global__ __launch_bounds__(kRadThreads) void k_sum_cols1(const double* __restrict__ cols, int ncols, int64_t stride,
                                                           const Ctrl* __restrict__ ctrl, double* __restrict__ partial,
                                                           int nb, int gate_active) {
    __shared__ double sh[kRadThreads];
    const int n = (gate_active && !ctrl->active) ? 0 : ctrl->n_alive;
    const int i = blockIdx.x * kRadThreads + threadIdx.x;
    for (int c = 0; c < ncols; c++) {
        const double v = i < n ? cols[(size_t)c * stride + i] : 0.0;
        const double s = block_sum_f64(v, sh);
        if (threadIdx.x == 0) partial[(size_t)c * nb + blockIdx.x] = s;
    }
}

// out[c] (+)= sum over blocks, one block per column.
__global__ __launch_bounds__(kRadThreads) void k_sum_cols2(const double* __restrict__ partial, int nb,
                                                           double* __restrict__ out, int accumulate) {
    __shared__ double sh[kRadThreads];
    const int c = blockIdx.x;
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += kRadThreads) v += partial[(size_t)c * nb + b];
    const double s = block_sum_f64(v, sh);
    if (threadIdx.x == 0) out[c] = accumulate ? out[c] + s : s;
}

void launch_sum_cols(hipStream_t s, const double* cols, int ncols, int64_t stride, int n_upper, const Ctrl* ctrl,
                     double* partial, double* out, bool accumulate, bool gate_active) {
    const int nb = (n_upper + kRadThreads - 1) / kRadThreads;
    if (nb == 0 || ncols == 0) return;
    hipLaunchKernelGGL(k_sum_cols1, dim3(nb), dim3(kRadThreads), 0, s, cols, ncols, stride, ctrl, partial, nb,
                       gate_active ? 1 : 0);
    hipLaunchKernelGGL(k_sum_cols2, dim3(ncols), dim3(kRadThreads), 0, s, partial, nb, out, accumulate ? 1 : 0);
}

// ---- quadrupole radiation (integrator.js:1567-1740) ----

__global__ void k_quad_com(QuadArgs q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.ctrl->n_alive) return;
    const double m = q.mass[i];
    q.scr[i] = (double)q.x[i] * m;
    q.scr[q.stride + i] = (double)q.y[i] * m;
    q.scr[2 * q.stride + i] = m;
}

// Per particle: d3I_ij and d3Q_ij contributions about the COM (third time
// derivatives of m x_i x_j and of (q/m) m x_i x_j from the force and the
// analytic jerk), the kinetic energy, and the squared contributions.
__global__ void k_quad_terms(QuadArgs q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.ctrl->n_alive) return;
    const double* com = q.sums;                 // [0] sum m x, [1] sum m y, [2] sum m
    const double mt = com[2];
    const double cx = mt > 1e-12 ? com[0] / mt : 0.0, cy = mt > 1e-12 ? com[1] / mt : 0.0;
    const double x = (double)q.x[i] - cx, y = (double)q.y[i] - cy;
    const double vx = q.vx[i], vy = q.vy[i], Fx = q.fx[i], Fy = q.fy[i], Jx = q.jx[i], Jy = q.jy[i];
    const double m = q.mass[i];
    const double wx = q.wx[i], wy = q.wy[i];
    const double wsq = wx * wx + wy * wy;
    double v[kRadCols] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (wsq > 1e-12) v[6] = m * wsq / (sqrt(1.0 + wsq) + 1.0);
    const double ixx = 6 * vx * Fx + 2 * x * Jx;
    const double ixy = Jx * y + 3 * Fx * vy + 3 * vx * Fy + x * Jy;
    const double iyy = 6 * vy * Fy + 2 * y * Jy;
    if (q.gw) {
        v[0] = ixx; v[1] = ixy; v[2] = iyy;
        v[7] = ixx * ixx + 2 * ixy * ixy + iyy * iyy;
    }
    if (q.em) {
        const double qm = (double)q.charge[i] / m;
        v[3] = qm * ixx; v[4] = qm * ixy; v[5] = qm * iyy;
        v[8] = v[3] * v[3] + 2 * v[4] * v[4] + v[5] * v[5];
    }
    for (int c = 0; c < kRadCols; c++) q.scr[(size_t)c * q.stride + i] = v[c];
}

// Powers, the radiated energy P dt, and its extraction from each particle in
// proportion to its contribution by exact relativistic rescaling of w.
__global__ void k_quad_apply(QuadArgs q, double dt) {
    const double* S = q.sums + 3;               // the nine sums of k_quad_terms
    const double tr = S[0] + S[2];
    const double ixx = S[0] - tr / 3, iyy = S[2] - tr / 3;
    const double gw_p = 0.2 * (ixx * ixx + 2 * S[1] * S[1] + iyy * iyy);
    const double em_p = q.em ? (1.0 / 180) * (S[3] * S[3] + 2 * S[4] * S[4] + S[5] * S[5]) : 0.0;
    const double qp = gw_p + em_p;
    if (!(qp > 0)) return;
    const double dE = qp * dt;
    const double gw_de = dE * (gw_p / qp), em_de = dE - gw_de;
    if (blockIdx.x == 0 && threadIdx.x == 0) q.rad_tot[0] += dE;
    if (!(dE > 1e-12 && S[6] > 1e-12)) return;
    const double inv_i = S[7] > 1e-12 ? 1.0 / S[7] : 0.0, inv_q = S[8] > 1e-12 ? 1.0 / S[8] : 0.0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.ctrl->n_alive) return;
    const double ci = q.scr[7 * q.stride + i], cq = q.scr[8 * q.stride + i];
    const double dke = (inv_i > 0 ? gw_de * ci * inv_i : 0.0) + (inv_q > 0 ? em_de * cq * inv_q : 0.0);
    if (dke <= 0) return;
    const double wx = q.wx[i], wy = q.wy[i];
    const double wsq = wx * wx + wy * wy;
    if (wsq < 1e-12) return;
    const double m = q.mass[i];
    const double ke = wsq / (sqrt(1.0 + wsq) + 1.0) * m;
    if (dke >= ke) {
        q.wx[i] = 0.f; q.wy[i] = 0.f;
    } else {
        const double gn = 1.0 + (ke - dke) / m;
        const double sc = sqrt((gn * gn - 1.0) / wsq);
        q.wx[i] = (float)(wx * sc); q.wy[i] = (float)(wy * sc);
    }
}

void launch_quad_radiation(hipStream_t s, const QuadArgs& q, double dt, int n_upper) {
    if (n_upper <= 0) return;
    const int nb = (n_upper + kRadThreads - 1) / kRadThreads;
    hipLaunchKernelGGL(k_quad_com, dim3(nb), dim3(kRadThreads), 0, s, q);
    launch_sum_cols(s, q.scr, 3, q.stride, n_upper, q.ctrl, q.partial, q.sums, false);
    hipLaunchKernelGGL(k_quad_terms, dim3(nb), dim3(kRadThreads), 0, s, q);
    launch_sum_cols(s, q.scr, kRadCols, q.stride, n_upper, q.ctrl, q.partial, q.sums + 3, false);
    hipLaunchKernelGGL(k_quad_apply, dim3(nb), dim3(kRadThreads), 0, s, q, dt);
}

}  // namespace geon
