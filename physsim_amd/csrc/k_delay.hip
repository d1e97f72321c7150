// k_delay.hip — signal-delay bookkeeping on gfx950: history samples of the
// live particles, death snapshots of retired ones, and slot metadata resets.
// The light-cone solve itself (geon_delay.h) runs inside the force kernels.
#include "geon_kernels.h"

namespace geon {

// integrator.js:1549-1565: every live particle appends (x, y, vx, vy, angw) as
// event ev; a particle without history starts it here (_initHistory,
// particle.js:124-134).
__global__ void k_hist_record(HistWrite w, const float* __restrict__ x, const float* __restrict__ y,
                              const float* __restrict__ vx, const float* __restrict__ vy,
                              const float* __restrict__ angw, const int32_t* __restrict__ slot, int64_t ev,
                              const Ctrl* __restrict__ ctrl) {
    const int n = ctrl->n_alive;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) w.ev_time[ev] = ctrl->sim_time;
    if (i >= n) return;
    const int s = slot[i];
    if (w.fe[s].x < 0) w.fe[s].x = (int)ev;
    const size_t k = (size_t)s * kHistN + (size_t)(ev & (kHistN - 1));
    w.h4[k] = make_float4(x[i], y[i], vx[i], vy[i]);
    w.haw[k] = angw[i];
}

void launch_hist_record(hipStream_t s, const HistWrite& w, const float* x, const float* y, const float* vx,
                        const float* vy, const float* angw, const int32_t* slot, int64_t ev, int n,
                        const Ctrl* ctrl) {
    const int grid = n > 0 ? (n + 255) / 256 : 1;
    hipLaunchKernelGGL(k_hist_record, dim3(grid), dim3(256), 0, s, w, x, y, vx, vy, angw, slot, ev, ctrl);
}

// _retireParticle (integrator.js:152-173): the final snapshot at the death
// time becomes the slot's last history entry (event ev, private to it), and
// the moments it will carry are taken from the record at use.
__global__ void k_retire(HistWrite w, const float* __restrict__ x, const float* __restrict__ y,
                         const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ angw,
                         const float* __restrict__ mass, const float* __restrict__ charge,
                         const float* __restrict__ radius, const float* __restrict__ ax,
                         const float* __restrict__ yuk, const float* __restrict__ higgs,
                         const int32_t* __restrict__ slot, const int32_t* __restrict__ alive,
                         const int32_t* __restrict__ alive_scan, int64_t ev, int dead_base, int n,
                         const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || alive[i]) return;
    const int s = slot[i];
    int2 fe = w.fe[s];
    if (fe.x < 0) fe.x = (int)ev;
    fe.y = (int)(ev + 1);
    w.fe[s] = fe;
    w.cd[s].y = ctrl->sim_time;
    const size_t k = (size_t)s * kHistN + (size_t)(ev & (kHistN - 1));
    w.h4[k] = make_float4(x[i], y[i], vx[i], vy[i]);
    w.haw[k] = angw[i];
    DeadRec r;
    r.slot = s;
    r.mass = mass[i];
    r.charge = charge[i];
    r.r2 = radius[i] * radius[i];
    r.ax = ax ? ax[i] : 1.f;
    r.yuk = yuk ? yuk[i] : 1.f;
    r.higgs = higgs ? higgs[i] : 1.f;
    r.pad = 0.f;
    w.dead[dead_base + (i - alive_scan[i])] = r;
}

void launch_retire(hipStream_t s, const HistWrite& w, const float* x, const float* y, const float* vx,
                   const float* vy, const float* angw, const float* mass, const float* charge,
                   const float* radius, const float* ax, const float* yuk, const float* higgs,
                   const int32_t* slot, const int32_t* alive, const int32_t* alive_scan, int64_t ev,
                   int dead_base, int n, const Ctrl* ctrl) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_retire, dim3((n + 255) / 256), dim3(256), 0, s, w, x, y, vx, vy, angw, mass, charge, radius,
                       ax, yuk, higgs, slot, alive, alive_scan, ev, dead_base, n, ctrl);
}

__global__ void k_hist_reset(int2* __restrict__ fe, double2* __restrict__ cd, int first_slot, int count, double t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    fe[first_slot + i] = make_int2(-1, -1);
    cd[first_slot + i] = make_double2(t, INFINITY);
}

void launch_hist_reset(hipStream_t s, int2* fe, double2* cd, int first_slot, int count, double t) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_hist_reset, dim3((count + 255) / 256), dim3(256), 0, s, fe, cd, first_slot, count, t);
}

}  // namespace geon
