// geon_delay.h — signal delay on gfx950: per-slot history rings and the
// light-cone solve of getDelayedState (signal-delay.js:16-221).
//
// Layout in HBM (indexed by the caller's slot, which never moves, so the rings
// are not re-sorted with the particles):
//   hist4[slot * 256 + (e & 255)]  {x, y, vx, vy} of history event e   (FP32)
//   hist_aw[slot * 256 + (e & 255)] angw of event e
//   hist_fe[slot]   {first event, end event} (-1: none / still alive)
//   hist_cd[slot]   {creation time, death time}                        (FP64)
//   ev_time[e]      simulation time of event e, shared by every particle
//                   sampled then                                       (FP64)
// A live particle holds events [max(first, E - 256), E) of the E events taken
// so far; a retired one holds its own [.., end) with the last entry being the
// death snapshot at its death time (integrator.js:152-173).
#pragma once
#include "geon_dev.h"

namespace geon {

constexpr int kHistN = 256;             // HISTORY_SIZE (config.js:120)
constexpr double kDelayEps = 1e-12;     // EPSILON (config.js:30)

// A retired particle still acting through its history (forces.js:128-153).
struct DeadRec {
    int32_t slot;
    float mass, charge, r2;             // _deathMass, charge, bodyRadiusSq
    float ax, yuk, higgs, pad;
};

struct HistArgs {
    const float4* h4;
    const float* haw;
    const int2* fe;
    const double2* cd;
    const double* ev_time;
    const int32_t* slot;                // sorted index -> slot
    const DeadRec* dead;
    int64_t n_events;
    int n_dead;
    int on;
};

// The retarded source seen from the observer: displacement observer -> source
// (min-imaged on a torus), its velocity and angw.
struct Retarded {
    float dx, dy, vx, vy, aw;
};

__device__ __forceinline__ void delay_disp(const PhysFlags& pf, float ox, float oy, float px, float py, float& dx,
                                           float& dy) {
    min_image(pf, ox, oy, px, py, dx, dy);
}

// The same displacement in FP64: exact for FP32 inputs, so the ill-conditioned
// light-cone geometry sees no rounding of the recorded positions.
__device__ __forceinline__ void delay_disp64(const PhysFlags& pf, double ox, double oy, double px, double py,
                                             double& dx, double& dy) {
    min_image_t<double>(pf.periodic, pf.topology, (double)pf.dom_w, (double)pf.dom_h, ox, oy, px, py, dx, dy);
}

// The retarded point (xr, yr), an absolute position as getDelayedState returns
// it, seen from the observer through pairForce's minimum image: one
// conditional shift of the absolute difference (forces.js:165-167,
// topology.js:9-12), whose image can differ from the one the light-cone solve
// used when the point was extrapolated far.
__device__ __forceinline__ void set_disp(const PhysFlags& pf, double xr, double yr, float ox, float oy,
                                         Retarded& out) {
    double dx, dy;
    delay_disp64(pf, ox, oy, xr, yr, dx, dy);
    out.dx = (float)dx;
    out.dy = (float)dy;
}

// Root u of (v^2 - 1) u^2 + 2 hb u + c = 0 picked as the reference picks it:
// BACK = 0, the largest root in [0, seg] (with EPSILON slack); BACK = 1, the
// largest root <= 0.  Returns false where the reference moves on.
template <bool BACK>
__device__ __forceinline__ bool light_cone_root(double a, double hb, double c, double seg, double& u) {
    const double disc = hb * hb - a * c;
    if (!(disc >= 0.0)) return false;
    if (fabs(a) < kDelayEps) {
        if (fabs(hb) < kDelayEps) return false;
        u = -c / (2.0 * hb);
        return true;
    }
    const double sq = sqrt(disc);
    const double u1 = (-hb + sq) / a, u2 = (-hb - sq) / a;
    const bool ok1 = BACK ? (u1 <= kDelayEps) : (u1 >= -kDelayEps && u1 <= seg + kDelayEps);
    const bool ok2 = BACK ? (u2 <= kDelayEps) : (u2 >= -kDelayEps && u2 <= seg + kDelayEps);
    if (ok1 && ok2) u = fmax(u1, u2);
    else if (ok1) u = u1;
    else if (ok2) u = u2;
    else return false;
    return true;
}

// The state of source `s` where the observer's past light cone crosses its
// history at time `now`: bisection on the segment ends (g(t) = |x(t) - x_obs|
// - (now - t) increases along the history), the exact quadratic on the found
// segment, then its lower and upper neighbours, else constant-velocity
// extrapolation before the oldest sample (not for a retired source, nor before
// its creation).  The bisection in FP32 (it only picks the first segment to
// try); displacements, times and the quadratic in FP64, the result rounded
// once to FP32.
__device__ inline bool delayed_state(const HistArgs& H, const PhysFlags& pf, int s, float ox, float oy, double now,
                                     Retarded& out) {
    const int2 fe = H.fe[s];
    if (fe.x < 0) return false;
    const bool dead = fe.y >= 0;
    const int64_t end = dead ? (int64_t)fe.y : H.n_events;
    const int64_t span_ev = end - (int64_t)fe.x;
    const int count = span_ev < kHistN ? (int)span_ev : kHistN;
    if (count < 2) return false;
    const int64_t e0 = end - count;
    const double t_death = dead ? H.cd[s].y : 0.0;
    const float4* h4 = H.h4 + (size_t)s * kHistN;
    const float* haw = H.haw + (size_t)s * kHistN;
    auto ev_t = [&](int64_t e) -> double { return (dead && e == end - 1) ? t_death : H.ev_time[e]; };
    const double t_old = ev_t(e0);
    const double span = now - t_old;
    if (span < kDelayEps) return false;
    {
        const float4 hn = h4[(end - 1) & (kHistN - 1)];
        double cdx, cdy;
        delay_disp64(pf, ox, oy, hn.x, hn.y, cdx, cdy);
        const double dist2 = cdx * cdx + cdy * cdy;
        if (!(dist2 > 4.0 * span * span)) {
            int lo = 0, hi = count - 2;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                const int64_t e = e0 + mid + 1;
                const float4 hm = h4[e & (kHistN - 1)];
                float bx, by;
                delay_disp(pf, ox, oy, hm.x, hm.y, bx, by);
                const float g = sqrtf(bx * bx + by * by) - (float)(now - ev_t(e));
                if (g < 0.f) lo = mid + 1;
                else hi = mid;
            }
            for (int att = 0; att < 3; att++) {
                const int k = att == 0 ? lo : (att == 1 ? lo - 1 : lo + 1);
                if (k < 0 || k > count - 2) continue;
                const int64_t el = e0 + k;
                const double t_lo = ev_t(el);
                const double seg = ev_t(el + 1) - t_lo;
                if (seg < kDelayEps) continue;
                const int il = (int)(el & (kHistN - 1)), ih = (int)((el + 1) & (kHistN - 1));
                const float4 a4 = h4[il], b4 = h4[ih];
                double dx, dy, sx, sy;
                delay_disp64(pf, ox, oy, a4.x, a4.y, dx, dy);
                delay_disp64(pf, a4.x, a4.y, b4.x, b4.y, sx, sy);
                const double vx = sx / seg, vy = sy / seg;
                const double T = now - t_lo;
                double u;
                if (!light_cone_root<false>(vx * vx + vy * vy - 1.0, dx * vx + dy * vy + T, dx * dx + dy * dy - T * T,
                                            seg, u))
                    continue;
                u = fmin(fmax(u, 0.0), seg);
                const double fd = u / seg;
                set_disp(pf, (double)a4.x + fd * sx, (double)a4.y + fd * sy, ox, oy, out);
                const float f = (float)fd;
                out.vx = a4.z + f * (b4.z - a4.z);
                out.vy = a4.w + f * (b4.w - a4.w);
                out.aw = haw[il] + f * (haw[ih] - haw[il]);
                return true;
            }
        }
    }
    if (dead) return false;
    const int i0 = (int)(e0 & (kHistN - 1));
    const float4 o4 = h4[i0];
    double dx, dy;
    delay_disp64(pf, ox, oy, o4.x, o4.y, dx, dy);
    const double vx = o4.z, vy = o4.w;
    double u;
    if (!light_cone_root<true>(vx * vx + vy * vy - 1.0, dx * vx + dy * vy + span, dx * dx + dy * dy - span * span, 0.0,
                               u))
        return false;
    if (u > 0.0) u = 0.0;
    if (t_old + u < H.cd[s].x) return false;
    set_disp(pf, (double)o4.x + vx * u, (double)o4.y + vy * u, ox, oy, out);
    out.vx = o4.z;
    out.vy = o4.w;
    out.aw = haw[i0];
    return true;
}

// Retarded spin of a delayed source (forces.js:105-110, 657-664): angular
// velocity from angw and the moments it carries.
__device__ __forceinline__ float retarded_ang_vel(float aw, float r2) { return aw / sqrtf(1.f + aw * aw * r2); }

}  // namespace geon
