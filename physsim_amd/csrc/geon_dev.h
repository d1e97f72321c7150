// geon_dev.h — device-side data layout and the pair interaction shared by the
// gfx950 kernels of the Geon physics backend.
//
// Layout in HBM (all particle arrays in Morton order, re-sorted every substep):
//   SoA FP32 particle state  x, y, wx, wy, vx, vy, mass, base_mass, charge,
//                            angw, ang_vel, radius, mu_cache, l_cache, f1old
//   int32                    qq (charge quanta), slot (caller's index)
//   packed sources           src4[i] = {x, y, mass, charge}  (one 16 B load)
//                            src8[i] = {vx, vy, ang_vel, mu, L, ax, yuk, higgs}
//   tree nodes               Node (64 B, one s_load_dwordx16 per visit)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace geon {

constexpr int kWave = 64;
constexpr int kKeyLevels = 31;          // quadtree levels resolved by the Morton key of one system
constexpr int kBatchKeyLevels = 24;     // of a batch, whose keys also carry the system id
constexpr int kLeafCap = 4;             // QUADTREE_CAPACITY (config.js:59)
constexpr int kWalkStack = 100;         // per-wave node stack: 3 * levels + 4 <= 97 entries
constexpr float kEpsMass = 1e-12f;      // EPSILON (config.js:30), node-mass skip

// Feature bits baked into kernel instantiations.
enum : uint32_t {
    F_GRAV = 1u,        // gravity (+ tidal torque)
    F_COUL = 2u,        // Coulomb
    F_EXTRA = 4u,       // any of magnetic / GM / 1PN / Yukawa / axion / Higgs (runtime-checked)
    F_JERK = 8u,        // radiation on: the analytic jerk too; gravity and Coulomb runtime-checked
    // Fixed instantiations for the common sets: the walk and its pair terms
    // compile without the runtime branches (and the register copies they cost).
    F_OPEN = 16u,       // open box: no minimum image, no window test
    F_FIXED = 32u,      // the extra forces are the bits below (no axion, no Higgs), not runtime toggles
    F_MAG = 64u, F_GM = 128u, F_YUK = 256u, F_1PN = 512u,
    F_TORUS = 1024u,    // periodic torus (topology 0): the minimum image without the topology branch
    F_WINDOW = 2048u,   // with F_TORUS: the windowed walk (nodes accepted only inside the min-image window)
};

// Runtime toggles of the extra forces (uniform per launch).
struct PhysFlags {
    uint32_t mag, gm, one_pn, yuk, ax_mod, higgs;
    float soft2;            // softening^2
    float yuk_g, yuk_mu;    // YUKAWA_COUPLING, yukawaMu
    float yuk_cut2;         // (6/muMin)^2
    float theta2;
    float dom_w, dom_h;     // periodic box (min-image when periodic != 0)
    uint32_t periodic;
    uint32_t window;        // periodic: approximate only nodes inside the min-image window
    uint32_t topology;      // periodic: 0 torus, 1 Klein bottle, 2 RP^2 (config.js:22-26)
    uint32_t grav_on = 0, coul_on = 0;  // gravity / Coulomb, read by the F_JERK variants
};

// The extra-force toggles as a pair_force<F> instantiation sees them: compile-time
// constants under F_FIXED, the launch's uniform PhysFlags otherwise.
template <uint32_t F>
struct Feat {
    static constexpr bool fixed = (F & F_FIXED) != 0u;
    bool mag, gm, one_pn, yuk, ax_mod, higgs;
    __device__ __forceinline__ explicit Feat(const PhysFlags& pf)
        : mag(fixed ? (F & F_MAG) != 0u : pf.mag != 0u), gm(fixed ? (F & F_GM) != 0u : pf.gm != 0u),
          one_pn(fixed ? (F & F_1PN) != 0u : pf.one_pn != 0u), yuk(fixed ? (F & F_YUK) != 0u : pf.yuk != 0u),
          ax_mod(fixed ? false : pf.ax_mod != 0u), higgs(fixed ? false : pf.higgs != 0u) {}
};


// Tree node record: 16 dwords.
constexpr double kChargeQuantum = 0.1;   // charges are integer multiples (main.js:504-539)

// Moments of a divided node's points about its COM (not in the reference: the
// inputs of the far-field correction).  Mass quadrupole q**, charge dipole d*,
// charge quadrupole c**, the node's total |charge| aq, and the feature scales
// 1/(M s^2), 1/(A s), 1/(A s^2) (0 without mass / charge; s = node width).
struct __align__(16) NodeMom {
    float qxx, qxy, qyy, aq;
    float dx, dy, cxx, cxy;
    float cyy, ims2, ias, ias2;
};

struct __align__(64) Node {
    float comx, comy, mass, charge;
    float mu, L, avx, avy;          // avx/avy: momentum / mass (proper velocity, quadtree.js:285)
    float size2;                    // (2 * half_width)^2  (forces.js:641)
    int32_t child;                  // divided: index of first child; leaf: first point
    int32_t count;                  // -1 divided, else points in leaf
    float half;                     // half width of the node box
    float bx, by;                   // box centre
    int32_t level;
    float pad;
};

// Compensated (Kahan) FP32 accumulator.
struct Kahan {
    float s, c;
    __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void add(float v) {
        float y = v - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// Observer (target particle) values held in registers for a whole walk.
struct Obs {
    float x, y, vx, vy, m, q, ang_vel, mu, L, ax, yuk, higgs;
    float tidal_ri5;        // body_radius^5 (forces.js:437-438)
    float inv_m;            // 1/m (0 if m <= EPSILON) for the tidal coupling q1q2/m
};

// Accumulated outputs of one observer.
template <bool KAHAN>
struct Acc {
    Kahan fx, fy, pe;
    float bz, dbzdx, dbzdy, bgz, dbgzdx, dbgzdy, frame_drag, tidal, f1x, f1y;
    float jx, jy;           // analytic jerk dF/dt (F_JERK only)
    __device__ __forceinline__ void init() {
        fx.init(); fy.init(); pe.init();
        bz = dbzdx = dbzdy = bgz = dbgzdx = dbgzdy = frame_drag = tidal = f1x = f1y = 0.f;
        jx = jy = 0.f;
    }
    __device__ __forceinline__ void addj(float fd, float jr, float vrx, float vry, float rx, float ry) {
        jx += vrx * fd + rx * jr;
        jy += vry * fd + ry * jr;
    }
    __device__ __forceinline__ void addf(float x, float y) {
        if (KAHAN) { fx.add(x); fy.add(y); } else { fx.s += x; fy.s += y; }
    }
    __device__ __forceinline__ void addpe(float v) {
        if (KAHAN) pe.add(v); else pe.s += v;
    }
};

// Per-type force of one observer, kept only for a selected particle: the
// reference's per-particle forceGravity, forceCoulomb, forceMagnetic,
// forceGravitomag, force1PN and forceYukawa (forces.js:36-46, 210-427).
enum SplitType { S_GRAV, S_COUL, S_MAG, S_GM, S_PN, S_YUK, kSplitTypes };
struct Split {
    float f[kSplitTypes][2];
    __device__ __forceinline__ void add(int t, float x, float y) { f[t][0] += x; f[t][1] += y; }
};

__device__ __forceinline__ float torus_wrap(float d, float full, float half) {
    // topology.js:9-12
    return d > half ? d - full : (d < -half ? d + full : d);
}

// minImage (topology.js:18-76): displacement observer (ox, oy) -> source
// (sx, sy) on the periodic domain.  Torus: each axis wrapped once.  Klein
// bottle (x periodic, y glide (x, y) ~ (W - x, y + H)) and RP^2 (glides on
// both axes): the nearest of the reference's candidate images, strict `<` in
// its order.  T is float or double.
template <typename T>
__device__ __forceinline__ T wrap_t(T d, T full, T half) { return d > half ? d - full : (d < -half ? d + full : d); }

template <typename T>
__device__ __forceinline__ void min_image_t(uint32_t periodic, uint32_t topo, T W, T H, T ox, T oy, T sx, T sy, T& dx,
                                            T& dy) {
    dx = sx - ox;
    dy = sy - oy;
    if (!periodic) return;
    const T hw = T(0.5) * W, hh = T(0.5) * H;
    if (topo == 0) {
        dx = wrap_t(dx, W, hw);
        dy = wrap_t(dy, H, hh);
        return;
    }
    if (topo == 1) {
        T bx = wrap_t(dx, W, hw), by = dy;
        const T cx = wrap_t((W - sx) - ox, W, hw), cy = wrap_t((sy + H) - oy, T(2) * H, H);
        if (cx * cx + cy * cy < bx * bx + by * by) { bx = cx; by = cy; }
        dx = bx;
        dy = by;
        return;
    }
    T bx = dx, by = dy, best = dx * dx + dy * dy;
    T cx = (W - sx) - ox, cy = wrap_t((sy + H) - oy, T(2) * H, H);
    T q = cx * cx + cy * cy;
    if (q < best) { bx = cx; by = cy; best = q; }
    cx = wrap_t((sx + W) - ox, T(2) * W, W);
    cy = (H - sy) - oy;
    q = cx * cx + cy * cy;
    if (q < best) { bx = cx; by = cy; best = q; }
    cx = wrap_t((T(2) * W - sx) - ox, T(2) * W, W);
    cy = wrap_t((T(2) * H - sy) - oy, T(2) * H, H);
    q = cx * cx + cy * cy;
    if (q < best) { bx = cx; by = cy; }
    dx = bx;
    dy = by;
}

__device__ __forceinline__ void min_image(const PhysFlags& pf, float ox, float oy, float sx, float sy, float& dx,
                                          float& dy) {
    min_image_t<float>(pf.periodic, pf.topology, pf.dom_w, pf.dom_h, ox, oy, sx, sy, dx, dy);
}
// As min_image, with the open box (F_OPEN) or the torus (F_TORUS) known at
// compile time.
template <uint32_t F>
__device__ __forceinline__ void min_image_f(const PhysFlags& pf, float ox, float oy, float sx, float sy, float& dx,
                                            float& dy) {
    if constexpr ((F & F_OPEN) != 0u) {
        dx = sx - ox;
        dy = sy - oy;
    } else if constexpr ((F & F_TORUS) != 0u) {
        dx = wrap_t(sx - ox, pf.dom_w, 0.5f * pf.dom_w);
        dy = wrap_t(sy - oy, pf.dom_h, 0.5f * pf.dom_h);
    } else {
        min_image(pf, ox, oy, sx, sy, dx, dy);
    }
}
// Whether the walk tests the min-image window: compile-time for the open box
// and the torus instantiations, the launch's flag otherwise.
template <uint32_t F>
__device__ __forceinline__ bool window_f(const PhysFlags& pf) {
    if constexpr ((F & F_OPEN) != 0u) return false;
    else if constexpr ((F & F_TORUS) != 0u) return (F & F_WINDOW) != 0u;
    else return pf.window != 0u;
}

// Far-field terms of an accepted node beyond its monopole (not in the
// reference): with the softened kernel g(d) = d h(r2), h = r2^-3/2,
// r2 = |d|^2 + eps^2, the node's points at COM + delta_j contribute
//   gravity  m_i [ 2 (Q d) h' + tr(Q) d h' + 2 d (d.Q.d) h'' ]          (1/2 Q : grad grad g)
//   Coulomb -q_i [ D h + 2 d (D.d) h' + (the same quadrupole form in C) ]
// with h' = dh/dr2 = -3/2 r2^-5/2 and h'' = 15/4 r2^-7/2; Q, D, C are the
// node's mass quadrupole, charge dipole and charge quadrupole about its COM.
// The matching potential terms (f = r2^-1/2, grad f = -d h) go to the PE:
//   gravity  m_i (tr(Q) h / 2 + (d.Q.d) h'),  Coulomb  -q_i ((D.d) h + tr(C) h / 2 + (d.C.d) h').
template <uint32_t F, bool KAHAN, bool SPLIT = false>
__device__ __forceinline__ void far_multipole(const Obs& o, Acc<KAHAN>& a, const PhysFlags& pf, float dx, float dy,
                                              const NodeMom& c, Split* sp = nullptr) {
    const float r2 = dx * dx + dy * dy + pf.soft2;
    const float ir2 = 1.f / r2;
    const float h = __builtin_amdgcn_rsqf(r2) * ir2;
    const float h1 = -1.5f * h * ir2, h2 = 3.75f * h * ir2 * ir2;
    float fx = 0.f, fy = 0.f, pe = 0.f;
    if (F & F_GRAV) {
        const float qx = c.qxx * dx + c.qxy * dy, qy = c.qxy * dx + c.qyy * dy;
        const float dqd = dx * qx + dy * qy, tr = c.qxx + c.qyy;
        fx += o.m * (2.f * qx * h1 + tr * dx * h1 + 2.f * dx * dqd * h2);
        fy += o.m * (2.f * qy * h1 + tr * dy * h1 + 2.f * dy * dqd * h2);
        pe += o.m * (0.5f * tr * h + dqd * h1);
        if constexpr (SPLIT) sp->add(S_GRAV, fx, fy);
    }
    if (F & F_COUL) {
        float qi = o.q;
        if ((F & F_EXTRA) && Feat<F>(pf).ax_mod) qi *= sqrtf(o.ax);   // node sources carry axMod = 1
        const float dd = c.dx * dx + c.dy * dy;
        const float qx = c.cxx * dx + c.cxy * dy, qy = c.cxy * dx + c.cyy * dy;
        const float dqd = dx * qx + dy * qy, tr = c.cxx + c.cyy;
        fx -= qi * (c.dx * h + 2.f * dx * dd * h1 + 2.f * qx * h1 + tr * dx * h1 + 2.f * dx * dqd * h2);
        fy -= qi * (c.dy * h + 2.f * dy * dd * h1 + 2.f * qy * h1 + tr * dy * h1 + 2.f * dy * dqd * h2);
        pe -= qi * (dd * h + 0.5f * tr * h + dqd * h1);
        if constexpr (SPLIT)
            sp->add(S_COUL, -qi * (c.dx * h + 2.f * dx * dd * h1 + 2.f * qx * h1 + tr * dx * h1 + 2.f * dx * dqd * h2),
                    -qi * (c.dy * h + 2.f * dy * dd * h1 + 2.f * qy * h1 + tr * dy * h1 + 2.f * dy * dqd * h2));
    }
    a.addf(fx, fy);
    a.addpe(pe);
}

// One source acting on the observer: the body of pairForce (forces.js:163-442)
// in FP32; F_JERK adds each term's analytic jerk dF/dt (forces.js:198-402),
// the input of the radiation reaction.  (rx, ry) is the (min-imaged)
// displacement source - observer.  ABERR: the source is signal-delayed, and the
// Lienard-Wiechert factor min((max(1 - n.v_s, 0.01))^-3, 100) scales its
// position-dependent forces (forces.js:182-189); PE and fields are unscaled.
// PE = false: no potential energy (dead sources, forces.js:130).  SPLIT: also
// sum the terms by type into *sp (the selected observer only).
template <uint32_t F, bool KAHAN, bool ABERR = false, bool PE = true, bool SPLIT = false>
__device__ __forceinline__ void pair_force(const Obs& o, Acc<KAHAN>& a, const PhysFlags& pf,
                                           float rx, float ry, float sm, float sq,
                                           float svx, float svy, float s_ang_vel,
                                           float s_mu, float s_L, float s_ax, float s_yuk,
                                           float s_higgs, Split* sp = nullptr) {
    const Feat<F> ft(pf);
    const float raw2 = rx * rx + ry * ry;
    const float r2 = raw2 + pf.soft2;
    const float inv_r = __builtin_amdgcn_rsqf(r2);   // 1/sqrt(r^2 + eps^2)
    const float inv_r2 = inv_r * inv_r;
    const float inv_r3 = inv_r * inv_r2;
    float aberr = 1.f;
    if (ABERR) {
        const float n_dot_v = -(rx * svx + ry * svy) * inv_r;
        const float den = fmaxf(1.f - n_dot_v, 0.01f);
        aberr = fminf(__builtin_amdgcn_rcpf(den * den * den), 100.f);
    }
    const float inv_r3a = ABERR ? inv_r3 * aberr : inv_r3;
    // the jerk instantiations check gravity and Coulomb at run time, unless fixed
    constexpr bool kRuntimeGC = (F & F_JERK) != 0u && (F & F_FIXED) == 0u;
    const bool do_grav = kRuntimeGC ? pf.grav_on != 0u : (F & F_GRAV) != 0u;
    const bool do_coul = kRuntimeGC ? pf.coul_on != 0u : (F & F_COUL) != 0u;
    float vrx = 0.f, vry = 0.f, rdv = 0.f, inv_r5j = 0.f;
    if (F & F_JERK) {                                               // forces.js:198-203
        vrx = svx - o.vx;
        vry = svy - o.vy;
        rdv = rx * vrx + ry * vry;
        inv_r5j = ABERR ? inv_r3 * inv_r2 * aberr : inv_r3 * inv_r2;
    }
    // The pair's central terms add into one radial coefficient fr (force =
    // r fr), its velocity-dependent terms into (fnx, fny) and its potential
    // into pe; the pair then enters the compensated sums once.
    float fr = 0.f, fnx = 0.f, fny = 0.f, pe = 0.f;
    if (do_grav) {
        const float k = o.m * sm;
        const float fd = k * inv_r3a;
        fr += fd;
        if constexpr (SPLIT) sp->add(S_GRAV, rx * fd, ry * fd);
        if (F & F_JERK) a.addj(fd, -3.f * k * rdv * inv_r5j, vrx, vry, rx, ry);
        if (PE) pe -= k * inv_r;
    }
    float ax_pair = 1.f;
    if ((F & F_EXTRA) && ft.ax_mod) ax_pair = sqrtf(o.ax * s_ax);
    if (do_coul) {
        const float k = -(o.q * sq) * ax_pair;
        const float fd = k * inv_r3a;
        fr += fd;
        if constexpr (SPLIT) sp->add(S_COUL, rx * fd, ry * fd);
        if (F & F_JERK) a.addj(fd, -3.f * k * rdv * inv_r5j, vrx, vry, rx, ry);
        if (PE) pe += o.q * sq * inv_r * ax_pair;
    }
    if (F & F_EXTRA) {
        const float inv_r5 = inv_r3 * inv_r2;
        const float inv_r5a = ABERR ? inv_r5 * aberr : inv_r5;
        const float cross_sv = svx * ry - svy * rx;
        if (ft.one_pn && (ft.gm || ft.mag)) {                       // forces.js:249-301
            const float r = r2 * inv_r;             // sqrt(r2) without a division
            const float nx = rx * inv_r, ny = ry * inv_r;
            const float v1v2 = o.vx * svx + o.vy * svy;
            if (ft.gm) {
                const float v1sq = o.vx * o.vx + o.vy * o.vy;
                const float v2sq = svx * svx + svy * svy;
                const float nv1 = nx * o.vx + ny * o.vy;
                const float nv2 = nx * svx + ny * svy;
                const float radial = -v1sq - 2.f * v2sq + 1.5f * nv2 * nv2 + 5.f * o.m * inv_r + 4.f * sm * inv_r;
                const float c1 = 4.f * nv1 - 3.f * nv2, c2 = 3.f * nv2;
                const float base = o.m * sm * inv_r3;
                const float fx = base * (rx * radial + (o.vx * c1 + svx * c2) * r);
                const float fy = base * (ry * radial + (o.vy * c1 + svy * c2) * r);
                fnx += fx; fny += fy; a.f1x += fx; a.f1y += fy;
                if constexpr (SPLIT) sp->add(S_PN, fx, fy);
                if (F & F_JERK) {               // position-only EIH term m1 m2 (5 m1 + 4 m2) r / r^4
                    const float ke = o.m * sm * (5.f * o.m + 4.f * sm);
                    const float r4 = inv_r2 * inv_r2;
                    a.addj(ke * r4, -4.f * ke * rdv * r4 * inv_r2, vrx, vry, rx, ry);
                }
                if (PE) pe -= o.m * sm * inv_r * (1.5f * (v1sq + v2sq) - 3.5f * v1v2 - 0.5f * nv1 * nv2 + o.m * inv_r + sm * inv_r);
            }
            if (ft.mag) {
                const float v2n = svx * nx + svy * ny, v1n = o.vx * nx + o.vy * ny;
                const float coeff = 0.5f * o.q * sq * inv_r2;
                const float fx = coeff * (o.vx * v2n - 3.f * nx * v1n * v2n);
                const float fy = coeff * (o.vy * v2n - 3.f * ny * v1n * v2n);
                fnx += fx; fny += fy; a.f1x += fx; a.f1y += fy;
                if constexpr (SPLIT) sp->add(S_PN, fx, fy);
                if (PE) pe -= 0.5f * o.q * sq * inv_r * (v1v2 + v1n * v2n);
            }
        }
        if (ft.one_pn && ft.gm && ft.mag) {                         // forces.js:303-320
            const float cc = o.q * sq * (o.m + sm) - (o.q * o.q * sm + sq * sq * o.m);
            const float fd = cc * inv_r2 * inv_r2;
            fr += fd; a.f1x += rx * fd; a.f1y += ry * fd;
            if constexpr (SPLIT) sp->add(S_PN, rx * fd, ry * fd);
            if (F & F_JERK) a.addj(fd, -4.f * cc * rdv * inv_r2 * inv_r2 * inv_r2, vrx, vry, rx, ry);
            if (PE) pe += 0.5f * cc * inv_r2;
        }
        if (ft.mag) {                                               // forces.js:322-353
            const float fd = -3.f * (o.mu * s_mu) * inv_r5a * ax_pair;
            fr += fd;
            if constexpr (SPLIT) sp->add(S_MAG, rx * fd, ry * fd);
            if (F & F_JERK) a.addj(fd, 15.f * (o.mu * s_mu) * rdv * inv_r5a * inv_r2 * ax_pair, vrx, vry, rx, ry);
            if (PE) pe += (o.mu * s_mu) * inv_r3 * ax_pair;
            const float bzm = sq * cross_sv * inv_r3 * ax_pair;
            a.bz += bzm;
            a.dbzdx += 3.f * bzm * rx * inv_r2 + sq * svy * inv_r3 * ax_pair;
            a.dbzdy += 3.f * bzm * ry * inv_r2 - sq * svx * inv_r3 * ax_pair;
            a.bz -= s_mu * inv_r3 * ax_pair;
            a.dbzdx -= 3.f * s_mu * rx * inv_r5 * ax_pair;
            a.dbzdy -= 3.f * s_mu * ry * inv_r5 * ax_pair;
        }
        if (ft.gm) {                                                // forces.js:355-387
            const float fd = 3.f * (o.L * s_L) * inv_r5a;
            fr += fd;
            if constexpr (SPLIT) sp->add(S_GM, rx * fd, ry * fd);
            if (F & F_JERK) a.addj(fd, -15.f * (o.L * s_L) * rdv * inv_r5a * inv_r2, vrx, vry, rx, ry);
            if (PE) pe -= (o.L * s_L) * inv_r3;
            const float bgm = -sm * cross_sv * inv_r3;
            a.bgz += bgm;
            a.dbgzdx += 3.f * bgm * rx * inv_r2 - sm * svy * inv_r3;
            a.dbgzdy += 3.f * bgm * ry * inv_r2 + sm * svx * inv_r3;
            a.bgz -= 2.f * s_L * inv_r3;
            a.dbgzdx -= 6.f * s_L * rx * inv_r5;
            a.dbgzdy -= 6.f * s_L * ry * inv_r5;
            a.frame_drag += 2.f * s_L * (s_ang_vel - o.ang_vel) * inv_r3;
        }
        if (ft.yuk && raw2 < pf.yuk_cut2) {                         // forces.js:389-427
            const float mu = ft.higgs ? pf.yuk_mu * sqrtf(o.higgs * s_higgs) : pf.yuk_mu;
            const float r = r2 * inv_r;             // sqrt(r2) without a division
            const float e = __expf(-mu * r);
            const float ym = ft.ax_mod ? sqrtf(o.yuk * s_yuk) : 1.f;
            const float gmm = pf.yuk_g * ym * o.m * sm * e;
            const float fd = gmm * (inv_r2 + mu * inv_r) * (ABERR ? inv_r * aberr : inv_r);
            fr += fd;
            if constexpr (SPLIT) sp->add(S_YUK, rx * fd, ry * fd);
            if (F & F_JERK)
                a.addj(fd, -(3.f * inv_r2 + 3.f * mu * inv_r + mu * mu) * rdv * gmm * inv_r2 *
                               (ABERR ? inv_r * aberr : inv_r), vrx, vry, rx, ry);
            if (PE) pe -= gmm * inv_r;
            if (ft.one_pn) {
                const float nx = rx * inv_r, ny = ry * inv_r;
                const float nv1 = nx * o.vx + ny * o.vy, nv2 = nx * svx + ny * svy;
                const float v1v2 = o.vx * svx + o.vy * svy;
                const float alpha = 1.f + mu * r;
                const float beta = 0.5f * gmm * inv_r2;
                const float radial = -(alpha * v1v2 + (alpha * alpha + alpha + 1.f) * nv1 * nv2);
                const float fx = beta * (radial * nx + alpha * (nv2 * o.vx + nv1 * svx));
                const float fy = beta * (radial * ny + alpha * (nv2 * o.vy + nv1 * svy));
                fnx += fx; fny += fy; a.f1x += fx; a.f1y += fy;
                if constexpr (SPLIT) sp->add(S_PN, fx, fy);
                if (PE) pe += 0.5f * gmm * inv_r * (v1v2 + nv1 * nv2 * alpha);
            }
        }
    }
    a.addf(rx * fr + fnx, ry * fr + fny);
    if (PE) a.addpe(pe);
    if (do_grav) {                                                  // forces.js:429-441
        const float cross_rv = rx * (svy - o.vy) - ry * (svx - o.vx);
        const float dw = o.ang_vel - cross_rv * inv_r2;
        float coupling = sm;
        if (do_coul) coupling += o.q * sq * o.inv_m;
        const float inv_r6 = inv_r2 * inv_r2 * inv_r2;
        a.tidal -= 64.f * coupling * coupling * o.tidal_ri5 * inv_r6 * dw;
    }
}

// _accum1PN (forces.js:445-506): the 1PN force of one source for the
// velocity-Verlet correction (EIH, Darwin, Bazanski, scalar Breit).
template <uint32_t F = 0u>
__device__ __forceinline__ void accum_1pn(const Obs& o, float& f1x, float& f1y, const PhysFlags& pf, float rx,
                                          float ry, float sm, float sq, float svx, float svy, float s_yuk,
                                          float s_higgs) {
    const Feat<F> ft(pf);
    const float r2 = rx * rx + ry * ry + pf.soft2;
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float inv_r2 = inv_r * inv_r;
    const float r = r2 * inv_r;                 // sqrt(r2) without a division
    const float nx = rx * inv_r, ny = ry * inv_r;
    if (ft.gm) {
        const float v1sq = o.vx * o.vx + o.vy * o.vy, v2sq = svx * svx + svy * svy;
        const float nv1 = nx * o.vx + ny * o.vy, nv2 = nx * svx + ny * svy;
        const float radial = -v1sq - 2.f * v2sq + 1.5f * nv2 * nv2 + 5.f * o.m * inv_r + 4.f * sm * inv_r;
        const float c1 = 4.f * nv1 - 3.f * nv2, c2 = 3.f * nv2;
        const float base = o.m * sm * inv_r2 * inv_r;
        f1x += base * (rx * radial + (o.vx * c1 + svx * c2) * r);
        f1y += base * (ry * radial + (o.vy * c1 + svy * c2) * r);
    }
    if (ft.mag) {
        const float v2n = svx * nx + svy * ny, v1n = o.vx * nx + o.vy * ny;
        const float coeff = 0.5f * o.q * sq * inv_r2;
        f1x += coeff * (o.vx * v2n - 3.f * nx * v1n * v2n);
        f1y += coeff * (o.vy * v2n - 3.f * ny * v1n * v2n);
    }
    if (ft.gm && ft.mag) {
        const float cc = o.q * sq * (o.m + sm) - (o.q * o.q * sm + sq * sq * o.m);
        const float fd = cc * inv_r2 * inv_r2;
        f1x += rx * fd;
        f1y += ry * fd;
    }
    if (ft.yuk) {
        const float mu = ft.higgs ? pf.yuk_mu * sqrtf(o.higgs * s_higgs) : pf.yuk_mu;
        const float e = __expf(-mu * r);
        const float nv1 = nx * o.vx + ny * o.vy, nv2 = nx * svx + ny * svy;
        const float v1v2 = o.vx * svx + o.vy * svy;
        const float alpha = 1.f + mu * r;
        const float beta = 0.5f * pf.yuk_g * sqrtf(o.yuk * s_yuk) * o.m * sm * e * inv_r2;
        const float radial = -(alpha * v1v2 + (alpha * alpha + alpha + 1.f) * nv1 * nv2);
        f1x += beta * (radial * nx + alpha * (nv2 * o.vx + nv1 * svx));
        f1y += beta * (radial * ny + alpha * (nv2 * o.vy + nv1 * svy));
    }
}

// Flops of one gravity+Coulomb interaction by the GPU-Gems rule (FMA = 2,
// rsqrt = 4): displacement 2, r^2 + eps 4, rsqrt 4, r^-2 r^-3 2, force 8,
// PE 4.  The tidal term adds 14 more when gravity is on.
constexpr int kFlopsGravCoul = 23;
constexpr int kFlopsTidal = 14;

}  // namespace geon
