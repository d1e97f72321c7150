// geon_api.hip — the C ABI (include/geon.h) over the gfx950 kernels.
//
// One GeonContext owns one particle system in device memory and one HIP
// stream.  The substep loop of Physics.update (integrator.js:726-1825) runs as
// a chain of kernels; the adaptive-dt controller runs on the device (k_ctrl),
// and the host only reads the controller's 64-byte record, one substep ahead,
// to know whether to enqueue another substep.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include <mutex>
#include <cmath>
#include "../../include/geon.h"
#include "geon_kernels.h"

using namespace geon;

namespace {

// State columns: gathered into Morton order every time the tree is rebuilt.
enum StateCol {
    C_X, C_Y, C_WX, C_WY, C_VX, C_VY, C_MASS, C_BASEMASS, C_CHARGE, C_ANGW, C_ANGVEL, C_RADIUS,
    C_MU, C_L, C_F1OX, C_F1OY, C_AX, C_YUK, C_HIGGS, C_QQ, C_SLOT, C_FLAGS, N_STATE
};
// Force columns: results of the last force pass, kept between updates.
enum ForceCol {
    K_FX, K_FY, K_PE, K_BZ, K_DBZDX, K_DBZDY, K_BGZ, K_DBGZDX, K_DBGZDY, K_FDRAG, K_TIDAL, K_F1X, K_F1Y,
    K_PEG, K_PFI, K_JX, K_JY, K_CONTACT, N_FORCE
};

// Device state of one scalar field (scalar-field.js:13-64).
struct FieldState {
    float *field = nullptr, *field_dot = nullptr, *field_dot_alt = nullptr, *grad_x = nullptr, *grad_y = nullptr;
    float *source = nullptr, *thermal = nullptr, *sg_phi = nullptr, *sg_gx = nullptr, *sg_gy = nullptr;
    unsigned long long *dep_src = nullptr, *dep_th = nullptr;
    float2* green = nullptr;
    float2* sg_c = nullptr;
    int* sg_rows = nullptr;
    int* has_energy = nullptr;
    double green_w = -1, green_h = -1, green_soft2 = -1;
    int green_topo = -1;        // 0 torus / open, else the glide topology the grid was built for
};

constexpr int kTimers = 8;   // walk, tree, sort, integrator, network, field deposit, field KDK, field -> particle
thread_local std::string g_create_error;

}  // namespace

// The tree points of the periodic-image mode (reals and their images).
struct GhostState {
    int64_t cap = 0;            // tree points held: factor x particles
    int factor = 0;
    GhostCols col = {}, alt = {};
    int32_t *cnt = nullptr, *off = nullptr;
    TreeBufs tb = {};
    float4 *src4 = nullptr, *src8a = nullptr, *src8b = nullptr;
    Ctrl* ctrl = nullptr;
};

struct GeonContext {
    GeonParams prm;
    GhostState gh;
    bool tree_ghost = false;    // the exported tree is the image tree
    unsigned char* na_scratch = nullptr;    // this device's copy of `scratch` (allocation-free pair list)
    int cap = 0;
    int n_host = 0;             // live particles (host mirror, exact between calls)
    int next_slot = 0;
    hipStream_t stream = nullptr;
    uint32_t* st[N_STATE] = {};
    uint32_t* st_alt[N_STATE] = {};
    float* fc[N_FORCE] = {};
    float* fc_alt[N_FORCE] = {};
    int32_t* alive = nullptr;
    int32_t* alive_scan = nullptr;
    uint32_t* compact_perm = nullptr;
    float4 *src4 = nullptr, *src8a = nullptr, *src8b = nullptr;
    Ctrl* ctrl = nullptr;
    Ctrl* ctrl_host = nullptr;  // pinned
    hipEvent_t ctrl_ev = nullptr;
    TreeBufs tb = {};
    double* stat_partial = nullptr;
    double* stat_out = nullptr;
    bool forces_init = false;
    bool tree_valid = false;
    int64_t substeps_total = 0;
    double sim_time = 0;
    std::string err;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[kTimers];
    size_t ev_used[kTimers] = {};
    int64_t last_counters[4] = {};
    FieldState fld[2];          // 0 Higgs, 1 axion
    int field_g = 0;
    double* field_out = nullptr;
    // learned far-field correction
    uint32_t* far_list = nullptr;
    int32_t* far_count = nullptr;
    int64_t far_cap = 0, far_groups = 0;
    void* far_frags = nullptr;
    float *far_params = nullptr, *far_stats = nullptr;      // FP32 master parameters, input/output statistics
    float *far_m = nullptr, *far_v = nullptr, *far_grad = nullptr, *far_loss = nullptr;
    float *far_gpart = nullptr, *far_lpart = nullptr;
    int64_t far_blocks_cap = 0;
    float* far_samples = nullptr;
    int32_t* far_obs = nullptr;         // observer slot of each sample
    int64_t far_n_samples = 0, far_cap_samples = 0;
    int64_t* far_sel = nullptr;
    int64_t far_sel_cap = 0;
    int64_t* far_rows = nullptr;
    int64_t far_rows_cap = 0;
    double* far_f2 = nullptr;
    bool far_have_model = false, far_have_stats = false;
    int64_t far_adam_t = 0;
    // signal delay: per-slot history rings (allocated at the first sample),
    // slot metadata, event times, retired sources (device + host mirror)
    float4* hist4 = nullptr;
    float* hist_aw = nullptr;
    int2* hist_fe = nullptr;
    double2* hist_cd = nullptr;
    double* ev_time = nullptr;
    int64_t ev_cap = 0, n_events = 0;
    int hist_stride_ctr = 0;
    DeadRec* dead = nullptr;
    int n_dead = 0;
    std::vector<DeadRec> dead_host;
    std::vector<double> dead_time;
    // collision mode merge: pair list, fold scratch, spawn records and events
    int64_t mg_cap_pairs = 0;
    int32_t *mg_p1 = nullptr, *mg_p2 = nullptr, *mg_leaf = nullptr;
    TreeBufs mg_sort = {};              // keys, perm and radix scratch of the pair sort
    TreeBufs mg_csort = {};             // the same for the fold pairs grouped by component
    int32_t *mg_uf = nullptr, *mg_uf_fail = nullptr;
    Ctrl* mg_cctrl = nullptr;           // n_alive = fold pairs (the component sort's count)
    int32_t *mg_fl = nullptr;           // 6 x pairs: spawn / merge-event / annihilation flags and offsets
    int32_t* mg_scan_tmp = nullptr;
    MergeRes *mg_res = nullptr, *mg_spawns = nullptr;
    double* mg_events = nullptr;
    int32_t *mg_deg = nullptr, *mg_rank = nullptr, *mg_off = nullptr, *mg_s5 = nullptr, *mg_mark = nullptr;
    uint64_t* mg_chunk5 = nullptr;
    double* mg_work = nullptr;          // 4 x particles: mass, base mass, radius, death mass
    int32_t* mg_cnt = nullptr;          // 8 ints: totals, deaths, same-leaf flag
    Ctrl* mg_ctrl = nullptr;            // n_alive = pairs of the current fold
    std::vector<double> events_host;    // events of the last update, 6 doubles each
    // radiation: per-particle FP64 scratch (9 columns), block sums, quadrupole
    // sums and the radiated energy / momentum since the last upload
    double *rad_scr = nullptr, *rad_partial = nullptr, *rad_sums = nullptr, *rad_tot = nullptr;
    // the selected particle (readStats' `selected`): its slot or -1, its index
    // on the device, and its force record followed by its state and force row
    int32_t sel_slot = -1;
    int32_t* sel_idx = nullptr;
    float* sel_rec = nullptr;
    // collision mode bounce: the overlap pairs (at most kBouncePairs per body)
    // and each body's list of the pairs it is the higher-slot member of
    int32_t *bc_pa = nullptr, *bc_pb = nullptr, *bc_cnt1 = nullptr, *bc_off1 = nullptr;
    int32_t *bc_cnt2 = nullptr, *bc_off2 = nullptr, *bc_fill2 = nullptr, *bc_list2 = nullptr, *bc_tot = nullptr;
    int64_t bc_cap = 0;
};

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            h->err = std::string(#expr) + ": " + hipGetErrorString(e_);                      \
            return GEON_EDEVICE;                                                             \
        }                                                                                    \
    } while (0)

namespace {

float* F(GeonContext* h, int c) { return reinterpret_cast<float*>(h->st[c]); }
int32_t* I(GeonContext* h, int c) { return reinterpret_cast<int32_t*>(h->st[c]); }

int set_err(GeonContext* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

template <typename T>
int dmalloc(GeonContext* h, T** p, size_t count) {
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) { h->err = std::string("hipMalloc: ") + hipGetErrorString(e); return GEON_ENOMEM; }
    return GEON_OK;
}

struct Timer {
    GeonContext* h;
    int k;
    hipEvent_t stop = nullptr;
    Timer(GeonContext* h_, int k_) : h(h_), k(k_) {
        if (!h->timing) return;
        auto& v = h->ev[k];
        if (h->ev_used[k] == v.size()) {
            hipEvent_t a, b;
            hipEventCreate(&a);
            hipEventCreate(&b);
            v.push_back({a, b});
        }
        auto& pr = v[h->ev_used[k]++];
        hipEventRecord(pr.first, h->stream);
        stop = pr.second;
    }
    ~Timer() {
        if (stop) hipEventRecord(stop, h->stream);
    }
};

uint32_t validate(const GeonParams* p, std::string* why) {
    const uint32_t t = p->toggles0;
    const uint32_t bad = GEON_DISINTEGRATION | GEON_BOSON_INTER | GEON_BLACK_HOLE;
    if (t & bad) { *why = "toggle not run by this backend (disintegration, bosons, black hole)"; return 1; }
    if ((t & GEON_RADIATION) && (t & GEON_YUKAWA)) {
        *why = "radiation with Yukawa is not run (its scalar radiation emits pions of random charge)"; return 1;
    }
    if ((t & GEON_RADIATION) && p->far_field != GEON_FAR_NONE) { *why = "radiation runs with monopole nodes only (far_field NONE)"; return 1; }
    const int G = p->field_grid;
    if ((t & (GEON_HIGGS | GEON_AXION)) && (G < 8 || G > 128 || (G & (G - 1)))) {
        *why = "field_grid must be a power of two in [8, 128]"; return 1;
    }
    if (p->collision_mode < GEON_COL_PASS || p->collision_mode > GEON_COL_BOUNCE) { *why = "unknown collision mode"; return 1; }
    if (p->collision_mode != GEON_COL_PASS && p->boundary_mode == GEON_BOUND_LOOP && p->topology != GEON_TORUS) {
        *why = "merge and bounce collisions run on open and torus domains"; return 1;
    }
    if (p->topology < GEON_TORUS || p->topology > GEON_RP2) { *why = "unknown topology"; return 1; }
    if (p->periodic_mode < GEON_PERIODIC_MIN_IMAGE || p->periodic_mode > GEON_PERIODIC_WINDOWED) { *why = "unknown periodic_mode"; return 1; }
    if (p->boundary_mode == GEON_BOUND_LOOP && p->periodic_mode == GEON_PERIODIC_GHOSTS && (t & GEON_BARNES_HUT)) {
        // the reference's periodic walk over image particles (integrator.js:201-258)
        if (p->far_field != GEON_FAR_NONE) { *why = "far-field corrections run in min-image or windowed periodic mode"; return 1; }
        if (t & (GEON_ONE_PN | GEON_RADIATION)) { *why = "1PN and radiation run in min-image or windowed periodic mode"; return 1; }
        if (p->collision_mode == GEON_COL_MERGE) { *why = "merge collisions run in min-image or windowed periodic mode"; return 1; }
        const bool dl = p->signal_delay == GEON_DELAY_ON || (p->signal_delay == GEON_DELAY_FOLLOW && (t & GEON_RELATIVITY));
        if (dl) { *why = "signal delay runs in min-image or windowed periodic mode"; return 1; }
    }
    if (p->boundary_mode == GEON_BOUND_LOOP && p->topology != GEON_TORUS) {
        if (p->periodic_mode == GEON_PERIODIC_WINDOWED) { *why = "Klein and RP2 run min-image or image-particle periodic mode"; return 1; }
        if (p->far_field != GEON_FAR_NONE) { *why = "far-field corrections run on open or torus domains"; return 1; }
    }
    if (!(p->domain_w > 0) || !(p->domain_h > 0)) { *why = "domain must be positive"; return 1; }
    if (p->max_substeps < 1) { *why = "max_substeps must be >= 1"; return 1; }
    if (p->far_field < GEON_FAR_NONE || p->far_field > GEON_FAR_LEARNED) { *why = "unknown far_field mode"; return 1; }
    if (p->far_list_cap < 0) { *why = "far_list_cap must be >= 0"; return 1; }
    if (p->signal_delay < GEON_DELAY_FOLLOW || p->signal_delay > GEON_DELAY_ON) { *why = "unknown signal_delay mode"; return 1; }
    if (p->history_stride < 0) { *why = "history_stride must be >= 0"; return 1; }
    const bool delay = p->signal_delay == GEON_DELAY_ON || (p->signal_delay == GEON_DELAY_FOLLOW && (t & GEON_RELATIVITY));
    if (delay && p->far_field != GEON_FAR_NONE) { *why = "signal delay runs with monopole nodes only (far_field NONE)"; return 1; }
    return 0;
}

bool delay_on(const GeonParams& p) {
    return p.signal_delay == GEON_DELAY_ON || (p.signal_delay == GEON_DELAY_FOLLOW && (p.toggles0 & GEON_RELATIVITY));
}

PhysFlags phys_flags(const GeonParams& p) {
    PhysFlags f;
    const uint32_t t = p.toggles0;
    f.mag = !!(t & GEON_MAGNETIC);
    f.gm = !!(t & GEON_GRAVITOMAG);
    f.one_pn = !!(t & GEON_ONE_PN);
    f.yuk = !!(t & GEON_YUKAWA);
    f.higgs = !!(t & GEON_HIGGS);
    f.ax_mod = ((t & GEON_COULOMB) || (t & GEON_MAGNETIC)) && (t & GEON_AXION);
    const double soft = (t & GEON_BLACK_HOLE) ? p.bh_softening : p.softening;
    f.soft2 = (float)(soft * soft);
    f.yuk_g = (float)p.yukawa_coupling;
    f.yuk_mu = (float)p.yukawa_mu;
    const double mu_min = f.higgs ? p.yukawa_mu * 0.05 : p.yukawa_mu;
    f.yuk_cut2 = f.yuk ? (float)((6.0 / mu_min) * (6.0 / mu_min)) : INFINITY;
    f.theta2 = (float)(p.theta * p.theta);
    f.dom_w = (float)p.domain_w;
    f.dom_h = (float)p.domain_h;
    f.periodic = p.boundary_mode == GEON_BOUND_LOOP;
    f.window = f.periodic && p.periodic_mode == GEON_PERIODIC_WINDOWED;
    f.topology = f.periodic ? (uint32_t)p.topology : 0u;
    f.grav_on = !!(t & GEON_GRAVITY);
    f.coul_on = !!(t & GEON_COULOMB);
    return f;
}

uint32_t feature_mask(const GeonParams& p) {
    const uint32_t t = p.toggles0;
    uint32_t f = 0;
    if (t & GEON_GRAVITY) f |= F_GRAV;
    if (t & GEON_COULOMB) f |= F_COUL;
    if (t & (GEON_MAGNETIC | GEON_GRAVITOMAG | GEON_YUKAWA)) f |= F_EXTRA;
    // the axion field modulates Coulomb pairs even without the other extra forces (forces.js:79, 223)
    if ((t & GEON_COULOMB) && (t & GEON_AXION)) f |= F_EXTRA;
    if (t & GEON_RADIATION) f = F_GRAV | F_COUL | F_JERK | (f & F_EXTRA);   // gravity, Coulomb runtime-checked; the jerk
    return f;
}

KeyParams key_params(const GeonParams& p) {
    KeyParams k;
    const double W = p.domain_w, H = p.domain_h;
    k.xmin = W / 2 - 2 * W;
    k.ymin = H / 2 - 2 * H;
    k.scale_x = (double)(1u << kKeyLevels) / (4 * W);
    k.scale_y = (double)(1u << kKeyLevels) / (4 * H);
    k.root_hw = 2 * W;
    return k;
}

StepArgs step_args(GeonContext* h) {
    StepArgs a;
    a.x = F(h, C_X); a.y = F(h, C_Y); a.wx = F(h, C_WX); a.wy = F(h, C_WY);
    a.vx = F(h, C_VX); a.vy = F(h, C_VY); a.angw = F(h, C_ANGW); a.ang_vel = F(h, C_ANGVEL);
    a.mass = F(h, C_MASS); a.radius = F(h, C_RADIUS); a.charge = F(h, C_CHARGE);
    a.mu = F(h, C_MU); a.L = F(h, C_L);
    a.f1x = h->fc[K_F1X]; a.f1y = h->fc[K_F1Y]; a.f1ox = F(h, C_F1OX); a.f1oy = F(h, C_F1OY);
    a.fx = h->fc[K_FX]; a.fy = h->fc[K_FY]; a.bz = h->fc[K_BZ]; a.bgz = h->fc[K_BGZ];
    a.dbzdx = h->fc[K_DBZDX]; a.dbzdy = h->fc[K_DBZDY]; a.dbgzdx = h->fc[K_DBGZDX]; a.dbgzdy = h->fc[K_DBGZDY];
    a.frame_drag = h->fc[K_FDRAG]; a.tidal = h->fc[K_TIDAL];
    a.alive = h->alive;
    a.ctrl = h->ctrl;
    const uint32_t t = h->prm.toggles0;
    a.rel = !!(t & GEON_RELATIVITY);
    a.mag = !!(t & GEON_MAGNETIC);
    a.gm = !!(t & GEON_GRAVITOMAG);
    a.ext_bz_on = h->prm.ext_bz != 0;
    a.grav = !!(t & GEON_GRAVITY);
    a.spin_orbit = !!(t & GEON_SPIN_ORBIT);
    a.one_pn = !!(t & GEON_ONE_PN);
    a.expansion = !!(t & GEON_EXPANSION);
    a.hubble = (float)h->prm.hubble;
    a.cx = (float)(h->prm.domain_w * 0.5);
    a.cy = (float)(h->prm.domain_h * 0.5);
    a.n_upper = h->n_host;
    a.rad = (t & GEON_RADIATION) && (t & GEON_COULOMB) && h->rad_scr;
    a.jx = h->fc[K_JX]; a.jy = h->fc[K_JY];
    a.rad_e = h->rad_scr;
    a.rad_stride = h->cap;
    if (h->sel_slot >= 0) { a.sel = h->sel_idx; a.sel_out = h->sel_rec; }
    a.col_bounce = h->prm.collision_mode == GEON_COL_BOUNCE;
    if (h->prm.boundary_mode == GEON_BOUND_BOUNCE || a.col_bounce) a.contact = h->fc[K_CONTACT];
    return a;
}

ForceArgs force_args(GeonContext* h) {
    ForceArgs a;
    a.src4 = h->src4; a.src8a = h->src8a; a.src8b = h->src8b;
    a.radius = F(h, C_RADIUS);
    a.nodes = h->tb.nodes;
    a.fx = h->fc[K_FX]; a.fy = h->fc[K_FY]; a.pe = h->fc[K_PE];
    a.bz = h->fc[K_BZ]; a.dbzdx = h->fc[K_DBZDX]; a.dbzdy = h->fc[K_DBZDY];
    a.bgz = h->fc[K_BGZ]; a.dbgzdx = h->fc[K_DBGZDX]; a.dbgzdy = h->fc[K_DBGZDY];
    a.frame_drag = h->fc[K_FDRAG]; a.tidal = h->fc[K_TIDAL]; a.f1x = h->fc[K_F1X]; a.f1y = h->fc[K_F1Y];
    a.jx = h->fc[K_JX]; a.jy = h->fc[K_JY];
    a.mass_for_cyc = F(h, C_MASS);
    a.charge = F(h, C_CHARGE);
    a.ctrl = h->ctrl;
    a.pf = phys_flags(h->prm);
    const GeonParams& p = h->prm;
    a.ext_gx = (float)(p.ext_gravity * cos(p.ext_gravity_angle));
    a.ext_gy = (float)(p.ext_gravity * sin(p.ext_gravity_angle));
    a.ext_ex = (float)(p.ext_electric * cos(p.ext_electric_angle));
    a.ext_ey = (float)(p.ext_electric * sin(p.ext_electric_angle));
    a.ext_bz = (float)p.ext_bz;
    a.cyc_mag = a.pf.mag || p.ext_bz != 0;
    a.cyc_gm = a.pf.gm;
    a.n_upper = h->n_host;
    a.max_nodes = 1 + 4 * h->tb.cap_divided;
    a.far_mode = 0;
    a.bsys = nullptr; a.bvalid = nullptr; a.bctrl = nullptr; a.broot = nullptr;
    a.node_mom = h->tb.node_mom;
    a.far_list = h->far_list;
    a.far_count = h->far_count;
    a.far_cap = h->far_cap;
    a.hist.h4 = h->hist4; a.hist.haw = h->hist_aw;
    a.hist.fe = h->hist_fe; a.hist.cd = h->hist_cd;
    a.hist.ev_time = h->ev_time;
    a.hist.slot = I(h, C_SLOT);
    a.hist.dead = h->dead;
    a.hist.n_events = h->n_events;
    a.hist.n_dead = h->n_dead;
    // no leaf source has two samples before the rings exist (every leaf skipped)
    a.hist.on = delay_on(h->prm) ? 1 : 0;
    if (h->sel_slot >= 0) { a.sel = h->sel_idx; a.sel_out = h->sel_rec; }
    return a;
}

// The selected particle's current index (after every reordering).
void sel_locate(GeonContext* h) {
    if (h->sel_slot >= 0) launch_sel_find(h->stream, I(h, C_SLOT), h->sel_slot, h->sel_idx, h->n_host, h->ctrl);
}

HistWrite hist_write(GeonContext* h) {
    HistWrite w;
    w.h4 = h->hist4; w.haw = h->hist_aw; w.fe = h->hist_fe; w.cd = h->hist_cd; w.ev_time = h->ev_time;
    w.dead = h->dead;
    return w;
}

int hist_alloc(GeonContext* h) {
    if (h->hist4) return GEON_OK;
    int rc;
    if ((rc = dmalloc(h, &h->hist4, (size_t)h->cap * kHistN))) return rc;
    if ((rc = dmalloc(h, &h->hist_aw, (size_t)h->cap * kHistN))) return rc;
    return GEON_OK;
}

// Room for event n_events in ev_time (grown by doubling, contents kept).
int ev_reserve(GeonContext* h) {
    if (h->n_events < h->ev_cap) return GEON_OK;
    const int64_t cap = h->ev_cap ? 2 * h->ev_cap : 4096;
    double* nb = nullptr;
    int rc;
    if ((rc = dmalloc(h, &nb, (size_t)cap))) return rc;
    if (h->ev_time) {
        HIP_OK(hipMemcpyAsync(nb, h->ev_time, sizeof(double) * (size_t)h->n_events, hipMemcpyDeviceToDevice, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        hipFree(h->ev_time);
    }
    h->ev_time = nb;
    h->ev_cap = cap;
    return GEON_OK;
}

// Forget every history and retired source (upload, reset).
void hist_clear(GeonContext* h) {
    h->n_events = 0;
    h->hist_stride_ctr = 0;
    h->n_dead = 0;
    h->dead_host.clear();
    h->dead_time.clear();
}

bool want_moments(GeonContext* h) {
    return (h->prm.toggles0 & GEON_BARNES_HUT) && h->prm.far_field != GEON_FAR_NONE;
}

int far_list_alloc(GeonContext* h) {
    const int64_t groups = ((int64_t)h->cap + kWave - 1) / kWave;
    // whole 1 KB blocks of the lane queues: a multiple of 4 entries per lane
    const int64_t blk = 4 * kWave;
    const int64_t cap = ((h->prm.far_list_cap > 0 ? (int64_t)h->prm.far_list_cap : 65536) + blk - 1) / blk * blk;
    if (h->far_list && h->far_cap == cap && h->far_groups == groups) return GEON_OK;
    hipFree(h->far_list); hipFree(h->far_count);
    h->far_list = nullptr; h->far_count = nullptr;
    int rc;
    if ((rc = dmalloc(h, &h->far_list, (size_t)(groups * cap)))) return rc;
    if ((rc = dmalloc(h, &h->far_count, (size_t)(groups * kWave)))) return rc;
    h->far_cap = cap; h->far_groups = groups;
    return GEON_OK;
}

int far_model_alloc(GeonContext* h) {
    if (h->far_params) return GEON_OK;
    int rc;
    uint8_t* fr = nullptr;
    if ((rc = dmalloc(h, &fr, far_frags_bytes()))) return rc;
    h->far_frags = fr;
    if ((rc = dmalloc(h, &h->far_params, kFarParams))) return rc;
    if ((rc = dmalloc(h, &h->far_stats, kFarStats))) return rc;
    if ((rc = dmalloc(h, &h->far_m, kFarParams))) return rc;
    if ((rc = dmalloc(h, &h->far_v, kFarParams))) return rc;
    if ((rc = dmalloc(h, &h->far_grad, kFarParams))) return rc;
    if ((rc = dmalloc(h, &h->far_loss, 1))) return rc;
    if ((rc = dmalloc(h, &h->far_f2, 2))) return rc;
    return GEON_OK;
}

void far_release(GeonContext* h) {
    hipFree(h->far_list); hipFree(h->far_count); hipFree(h->far_frags); hipFree(h->far_params);
    hipFree(h->far_stats); hipFree(h->far_m); hipFree(h->far_v); hipFree(h->far_grad); hipFree(h->far_loss);
    hipFree(h->far_gpart); hipFree(h->far_lpart); hipFree(h->far_samples); hipFree(h->far_sel);
    hipFree(h->far_rows); hipFree(h->far_f2); hipFree(h->far_obs);
}

FarArgs far_args(GeonContext* h) {
    FarArgs f;
    f.frags = (const FarFrags*)h->far_frags;
    f.src4 = h->src4;
    f.nodes = h->tb.nodes;
    f.node_mom = h->tb.node_mom;
    f.far_list = h->far_list;
    f.far_count = h->far_count;
    f.far_cap = h->far_cap;
    f.fx = h->fc[K_FX]; f.fy = h->fc[K_FY]; f.pe = h->fc[K_PE];
    f.ctrl = h->ctrl;
    const PhysFlags pf = phys_flags(h->prm);
    f.soft2 = pf.soft2;
    f.dom_w = pf.dom_w; f.dom_h = pf.dom_h;
    f.periodic = pf.periodic;
    f.grav = !!(h->prm.toggles0 & GEON_GRAVITY);
    f.coul = !!(h->prm.toggles0 & GEON_COULOMB);
    return f;
}

FieldArgs field_args(GeonContext* h, int which) {
    FieldState& s = h->fld[which];
    const GeonParams& p = h->prm;
    FieldArgs f;
    f.G = h->field_g;
    f.bc = p.boundary_mode;
    f.topo = p.boundary_mode == GEON_BOUND_LOOP ? p.topology : 0;
    f.higgs = which == 0;
    f.cw = (float)(p.domain_w / f.G);
    f.ch = (float)(p.domain_h / f.G);
    f.icw = (float)(f.G / p.domain_w);
    f.ich = (float)(f.G / p.domain_h);
    const double area = (p.domain_w / f.G) * (p.domain_h / f.G);
    f.inv_area = area > 1e-12 ? (float)(1.0 / area) : 0.f;
    f.vac = which == 0 ? 1.f : 0.f;
    f.mu2 = (float)(0.5 * p.higgs_mass * p.higgs_mass);          // higgs-field.js:79-81
    f.m2 = (float)(p.axion_mass * p.axion_mass);                 // axion-field.js:86-88
    f.damp = which == 0 ? (float)(2 * p.higgs_mass) : (float)(0.05 * p.axion_mass);
    f.field = s.field; f.field_dot = s.field_dot; f.field_dot_alt = s.field_dot_alt;
    f.grad_x = s.grad_x; f.grad_y = s.grad_y; f.source = s.source; f.thermal = s.thermal;
    f.sg_phi = s.sg_phi; f.sg_gx = s.sg_gx; f.sg_gy = s.sg_gy;
    f.dep_src = s.dep_src; f.dep_th = s.dep_th; f.green = s.green; f.has_energy = s.has_energy;
    f.sg_c = s.sg_c; f.sg_rows = s.sg_rows;
    f.ctrl = h->ctrl;
    return f;
}

ParticleView particle_view(GeonContext* h) {
    ParticleView v;
    v.x = F(h, C_X); v.y = F(h, C_Y); v.q = F(h, C_CHARGE);
    v.m = F(h, C_MASS); v.base_mass = F(h, C_BASEMASS); v.wx = F(h, C_WX); v.wy = F(h, C_WY);
    v.vx = F(h, C_VX); v.vy = F(h, C_VY); v.radius = F(h, C_RADIUS);
    v.ctrl = h->ctrl;
    return v;
}

bool fields_on(GeonContext* h) { return (h->prm.toggles0 & (GEON_HIGGS | GEON_AXION)) != 0; }

void ensure_green(GeonContext* h, int which) {
    FieldState& s = h->fld[which];
    const GeonParams& p = h->prm;
    const double soft = (p.toggles0 & GEON_BLACK_HOLE) ? p.bh_softening : p.softening;
    const int topo = p.boundary_mode == GEON_BOUND_LOOP ? p.topology : 0;
    if (s.green_w == p.domain_w && s.green_h == p.domain_h && s.green_soft2 == soft * soft && s.green_topo == topo) return;
    field_build_green(h->stream, field_args(h, which), (float)(soft * soft));
    s.green_w = p.domain_w; s.green_h = p.domain_h; s.green_soft2 = soft * soft; s.green_topo = topo;
}

// HiggsField.update + modulateMasses, AxionField.update (integrator.js:1320-1334),
// on the device; dt from the controller is mirrored on the host (dt_sub).
void fields_step(GeonContext* h, float dt) {
    const uint32_t t = h->prm.toggles0;
    const int grav = !!(t & GEON_GRAVITY), rel = !!(t & GEON_RELATIVITY);
    const int coul = !!(t & GEON_COULOMB), yuk = !!(t & GEON_YUKAWA);
    const ParticleView pv = particle_view(h);
    const int n = h->n_host;
    if (t & GEON_HIGGS) {
        if (grav) ensure_green(h, 0);
        FieldArgs f = field_args(h, 0);
        { Timer tm(h, 5); field_deposit(h->stream, f, pv, n, rel, coul, yuk); }
        { Timer tm(h, 6); field_kdk(h->stream, f, (t & GEON_AXION) ? h->fld[1].field : nullptr, grav, dt); }
        { Timer tm(h, 7); field_modulate(h->stream, f, pv, F(h, C_HIGGS), dt, n); }
    }
    if (t & GEON_AXION) {
        if (grav) ensure_green(h, 1);
        FieldArgs f = field_args(h, 1);
        { Timer tm(h, 5); field_deposit(h->stream, f, pv, n, rel, coul, yuk); }
        { Timer tm(h, 6); field_kdk(h->stream, f, (t & GEON_HIGGS) ? h->fld[0].field : nullptr, grav, dt); }
    }
}

__global__ void k_fill(float* a, int n, float v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Field reset to vacuum (HiggsField.reset / AxionField.reset, scalar-field.js:66-71).
void reset_fields(GeonContext* h) {
    const int g2 = h->field_g * h->field_g;
    for (int w = 0; w < 2; w++) {
        FieldState& s = h->fld[w];
        const dim3 gb((g2 + 255) / 256), bb(256);
        hipLaunchKernelGGL(k_fill, gb, bb, 0, h->stream, s.field, g2, w == 0 ? 1.f : 0.f);
        float* zero[] = {s.field_dot, s.field_dot_alt, s.grad_x, s.grad_y, s.source, s.thermal, s.sg_phi, s.sg_gx, s.sg_gy};
        for (float* p : zero) hipLaunchKernelGGL(k_fill, gb, bb, 0, h->stream, p, g2, 0.f);
        hipMemsetAsync(s.has_energy, 0, sizeof(int), h->stream);
    }
}

__global__ void k_reset_counters(Ctrl* c) {
    c->cnt_pair = 0; c->cnt_node = 0; c->cnt_visit = 0; c->depth_overflow = 0; c->overflow = 0;
}

__global__ void k_begin_update(Ctrl* c, double dt, int max_sub, double sim_time) {
    c->dt_remain = dt; c->dt_total = dt; c->total_steps = 0; c->max_substeps = max_sub;
    c->sim_time = sim_time; c->active = 0; c->deaths = 0;
    c->max_acc2_fresh = 0; c->max_acc2_kick = 0; c->max_cyc = 0;
}

__global__ void k_set_n(Ctrl* c, int n) { c->n_alive = n; }
__global__ void k_set_time(Ctrl* c, double t) { c->sim_time = t; }

__global__ void k_survivor_perm(const int32_t* alive, const int32_t* scan, uint32_t* perm, const Ctrl* c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c->n_alive && alive[i]) perm[scan[i]] = (uint32_t)i;
}

// Sort by Morton key and gather the state columns into that order.
int sort_and_gather(GeonContext* h) {
    const int n = h->n_host;
    if (n == 0) return GEON_OK;
    const bool one_pn = !!(h->prm.toggles0 & GEON_ONE_PN);
    const KeyParams kp = key_params(h->prm);
    {
        Timer t(h, 2);
        launch_keys(h->stream, F(h, C_X), F(h, C_Y), n, kp, h->tb.keys, h->tb.perm, h->ctrl);
        radix_sort(h->stream, h->tb, n, 2 * kKeyLevels + 1, h->ctrl);
        Columns cols;
        cols.ncols = N_STATE;
        for (int c = 0; c < N_STATE; c++) { cols.src[c] = h->st[c]; cols.dst[c] = h->st_alt[c]; }
        // the 1PN force outlives the force pass (the next kick takes it as the
        // old value): it moves with its particle
        if (one_pn) {
            for (int c : {(int)K_F1X, (int)K_F1Y}) {
                cols.src[cols.ncols] = reinterpret_cast<uint32_t*>(h->fc[c]);
                cols.dst[cols.ncols++] = reinterpret_cast<uint32_t*>(h->fc_alt[c]);
            }
        }
        launch_gather(h->stream, cols, h->tb.perm, n, h->ctrl);
    }
    for (int c = 0; c < N_STATE; c++) std::swap(h->st[c], h->st_alt[c]);
    if (one_pn) { std::swap(h->fc[K_F1X], h->fc_alt[K_F1X]); std::swap(h->fc[K_F1Y], h->fc_alt[K_F1Y]); }
    sel_locate(h);
    return GEON_OK;
}

int ghost_alloc(GeonContext* h, int factor);

bool ghosts_on(const GeonParams& p) {
    return p.boundary_mode == GEON_BOUND_LOOP && p.periodic_mode == GEON_PERIODIC_GHOSTS &&
           (p.toggles0 & GEON_BARNES_HUT);
}

// The reference's periodic tree (integrator.js:201-258): the particles sorted
// as usual, their images within max(W, H) * theta of an edge generated, and
// every tree point (particle or image) sorted and built over.
int ghost_tree(GeonContext* h) {
    const int n = h->n_host;
    const GeonParams& p = h->prm;
    // a particle has images on both sides of an axis only when the margin
    // exceeds half that side: at most (1 + sides_x)(1 + sides_y) tree points
    const double margin = std::max(p.domain_w, p.domain_h) * p.theta;     // integrator.js:204
    const int sx = margin > 0.5 * p.domain_w ? 2 : 1, sy = margin > 0.5 * p.domain_h ? 2 : 1;
    const int factor = (1 + sx) * (1 + sy);
    int rc;
    if ((rc = ghost_alloc(h, factor))) return rc;
    if ((rc = sort_and_gather(h))) return rc;
    GhostState& g = h->gh;
    GhostArgs ga;
    ga.x = F(h, C_X); ga.y = F(h, C_Y); ga.m = F(h, C_MASS); ga.q = F(h, C_CHARGE);
    ga.wx = F(h, C_WX); ga.wy = F(h, C_WY); ga.vx = F(h, C_VX); ga.vy = F(h, C_VY);
    ga.ang_vel = F(h, C_ANGVEL); ga.radius = F(h, C_RADIUS); ga.mu = F(h, C_MU); ga.L = F(h, C_L);
    ga.qq = I(h, C_QQ); ga.ctrl = h->ctrl;
    ga.dom_w = (float)p.domain_w; ga.dom_h = (float)p.domain_h;
    ga.margin = (float)margin;
    ga.topo = p.topology;
    ga.cap_points = (int)g.cap;
    ga.out = g.col;
    const int n4 = factor * n;
    const KeyParams kp = key_params(p);
    Timer t(h, 1);
    ghost_tree_points(h->stream, ga, n, g.cnt, g.off, h->tb.scan_tmp, g.ctrl);
    auto fcol = [&](int c) { return reinterpret_cast<float*>(g.col.c[c]); };
    launch_keys(h->stream, fcol(G_X), fcol(G_Y), n4, kp, g.tb.keys, g.tb.perm, g.ctrl);
    radix_sort(h->stream, g.tb, n4, 2 * kKeyLevels + 1, g.ctrl);
    Columns cols;
    cols.ncols = kGhostCols;
    for (int c = 0; c < kGhostCols; c++) { cols.src[c] = g.col.c[c]; cols.dst[c] = g.alt.c[c]; }
    launch_gather(h->stream, cols, g.tb.perm, n4, g.ctrl);
    std::swap(g.col, g.alt);
    build_tree(h->stream, g.tb, fcol(G_X), fcol(G_Y), fcol(G_M), fcol(G_WX), fcol(G_WY), fcol(G_MU), fcol(G_L),
               reinterpret_cast<int32_t*>(g.col.c[G_QQ]), n4, kp, g.ctrl, false);
    return GEON_OK;
}

constexpr int kBouncePairs = 8;     // overlap pairs per body a bounce pass holds

// _applyRepulsion (integrator.js:669-724) over the pair list of the tree the
// pass built at these positions: the pairs counted, scanned and written
// (lower slot first), then every body's contacts summed in a fixed order.
int repulsion(GeonContext* h, const ForceArgs& fa) {
    const int n = h->n_host;
    if (n == 0) return GEON_OK;
    int rc;
    if (!h->bc_pa) {
        h->bc_cap = (int64_t)kBouncePairs * h->cap;
        if ((rc = dmalloc(h, &h->bc_pa, (size_t)h->bc_cap)) || (rc = dmalloc(h, &h->bc_pb, (size_t)h->bc_cap)) ||
            (rc = dmalloc(h, &h->bc_list2, (size_t)h->bc_cap)) || (rc = dmalloc(h, &h->bc_cnt1, (size_t)h->cap)) ||
            (rc = dmalloc(h, &h->bc_off1, (size_t)h->cap)) || (rc = dmalloc(h, &h->bc_cnt2, (size_t)h->cap)) ||
            (rc = dmalloc(h, &h->bc_off2, (size_t)h->cap)) || (rc = dmalloc(h, &h->bc_fill2, (size_t)h->cap)) ||
            (rc = dmalloc(h, &h->bc_tot, 2)))
            return rc;
    }
    PairArgs pa;
    pa.x = F(h, C_X); pa.y = F(h, C_Y); pa.mass = F(h, C_MASS); pa.radius = F(h, C_RADIUS);
    pa.slot = I(h, C_SLOT); pa.src4 = h->src4; pa.nodes = h->tb.nodes; pa.max_nodes = 1 + 4 * h->tb.cap_divided;
    pa.count = h->bc_cnt1; pa.offset = h->bc_off1; pa.pa = nullptr; pa.pb = nullptr; pa.pl = nullptr;
    pa.index_out = 1;
    pa.slack = 1.f;
    pa.dom_w = (float)h->prm.domain_w; pa.dom_h = (float)h->prm.domain_h;
    pa.periodic = h->prm.boundary_mode == GEON_BOUND_LOOP;
    pa.ctrl = h->ctrl; pa.n_upper = n;
    pa.cap_pairs = h->bc_cap;
    launch_pairs(h->stream, pa, false);
    exclusive_scan_i32(h->stream, h->bc_cnt1, h->bc_off1, n, &h->ctrl->n_alive, h->bc_tot, h->tb.scan_tmp);
    pa.pa = h->bc_pa; pa.pb = h->bc_pb;
    launch_pairs(h->stream, pa, true);
    RepelArgs r;
    r.x = F(h, C_X); r.y = F(h, C_Y); r.vx = F(h, C_VX); r.vy = F(h, C_VY); r.ang_vel = F(h, C_ANGVEL);
    r.radius = F(h, C_RADIUS);
    r.pa = h->bc_pa; r.pb = h->bc_pb; r.cnt1 = h->bc_cnt1; r.off1 = h->bc_off1;
    r.cnt2 = h->bc_cnt2; r.off2 = h->bc_off2; r.fill2 = h->bc_fill2; r.list2 = h->bc_list2;
    r.total2 = h->bc_tot + 1; r.n_pairs = h->bc_tot;
    r.fx = h->fc[K_FX]; r.fy = h->fc[K_FY]; r.contact = h->fc[K_CONTACT];
    r.sel = fa.sel; r.sel_out = fa.sel_out; r.ctrl = h->ctrl;
    r.dom_w = (float)h->prm.domain_w; r.dom_h = (float)h->prm.domain_h; r.friction = (float)h->prm.bounce_friction;
    r.periodic = pa.periodic;
    r.cap_pairs = h->bc_cap;
    r.scan_tmp = h->tb.scan_tmp;
    launch_repulsion(h->stream, r, n);
    return GEON_OK;
}

// computeAllForces + external fields at the current state (forces.js:61-154,
// integrator.js:563-592), preceded by the tree rebuild (integrator.js:1340-1348).
// reuse_tree: the tree built for this substep's 1PN pass stands (integrator.js:1340-1348
// reuses it; with Higgs on, its aggregates are recomputed for the new masses).
// collect: list the accepted (particle, node) pairs for training samples
// instead of applying the far-field correction.
int force_pass(GeonContext* h, bool reuse_tree = false, bool collect = false) {
    const int n = h->n_host;
    const bool bh = !!(h->prm.toggles0 & GEON_BARNES_HUT);
    const int far = collect ? GEON_FAR_LEARNED : (bh ? h->prm.far_field : GEON_FAR_NONE);
    if (far == GEON_FAR_LEARNED) {
        if (!collect && !h->far_have_model)
            return set_err(h, GEON_ESTATE, "learned far field needs a model (geon_far_set_model or geon_far_train)");
        int rc = far_list_alloc(h);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, h->stream, h->ctrl);
    if (n == 0) return GEON_OK;
    const bool ghosts = bh && ghosts_on(h->prm);
    const bool col_bounce = h->prm.collision_mode == GEON_COL_BOUNCE;
    if (ghosts) {
        int rc = ghost_tree(h);
        if (rc) return rc;
        h->tree_valid = true;
        h->tree_ghost = true;
        if (col_bounce) {           // the contact query runs over the particles' own tree
            Timer t(h, 1);
            build_tree(h->stream, h->tb, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_WX), F(h, C_WY), F(h, C_MU),
                       F(h, C_L), I(h, C_QQ), n, key_params(h->prm), h->ctrl, false);
        }
    } else if (!bh && col_bounce) {     // the reference builds its tree every substep (integrator.js:1340-1348)
        int rc = sort_and_gather(h);
        if (rc) return rc;
        Timer t(h, 1);
        build_tree(h->stream, h->tb, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_WX), F(h, C_WY), F(h, C_MU),
                   F(h, C_L), I(h, C_QQ), n, key_params(h->prm), h->ctrl, false);
        h->tree_valid = true;
        h->tree_ghost = false;
    } else if (bh) {
        h->tree_ghost = false;
        if (!reuse_tree) {
            int rc = sort_and_gather(h);
            if (rc) return rc;
        }
        if (!reuse_tree || (h->prm.toggles0 & GEON_HIGGS)) {
            Timer t(h, 1);
            build_tree(h->stream, h->tb, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_WX), F(h, C_WY), F(h, C_MU),
                       F(h, C_L), I(h, C_QQ), n, key_params(h->prm), h->ctrl, far != GEON_FAR_NONE);
        }
        h->tree_valid = true;
    }
    const uint32_t tg = h->prm.toggles0;
    const bool axion = !!(tg & GEON_AXION), higgs = !!(tg & GEON_HIGGS);
    if (axion) {    // _syncAxionField (integrator.js:595-605)
        Timer tm(h, 7);
        field_axmod(h->stream, field_args(h, 1), particle_view(h), F(h, C_AX), F(h, C_YUK), !!(tg & GEON_COULOMB),
                    !!(tg & GEON_YUKAWA), n);
    }
    launch_prepare_sources(h->stream, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_CHARGE), F(h, C_VX), F(h, C_VY),
                           F(h, C_ANGVEL), F(h, C_RADIUS), F(h, C_MU), F(h, C_L), axion ? F(h, C_AX) : nullptr,
                           axion ? F(h, C_YUK) : nullptr, higgs ? F(h, C_HIGGS) : nullptr, h->src4, h->src8a,
                           h->src8b, n, h->ctrl);
    ForceArgs fa = force_args(h);
    if (bh && far == GEON_FAR_MULTIPOLE) fa.far_mode = kFarMultipole;
    if (bh && far == GEON_FAR_LEARNED) fa.far_mode = kFarList;
    if (ghosts) {
        // sources: the tree points; observers: the particles
        GhostState& g = h->gh;
        ghost_sources(h->stream, g.col, axion ? F(h, C_AX) : nullptr, axion ? F(h, C_YUK) : nullptr,
                      higgs ? F(h, C_HIGGS) : nullptr, g.src4, g.src8a, g.src8b, g.factor * n, g.ctrl);
        fa.obs4 = h->src4; fa.obs8a = h->src8a; fa.obs8b = h->src8b;
        fa.src4 = g.src4; fa.src8a = g.src8a; fa.src8b = g.src8b;
        fa.src_orig = reinterpret_cast<const int32_t*>(g.col.c[G_ORIG]);
        fa.nodes = g.tb.nodes;
        fa.max_nodes = 1 + 4 * g.tb.cap_divided;
        fa.node_mom = nullptr;
    }
    {
        Timer t(h, 0);
        if (bh) launch_walk(h->stream, fa, feature_mask(h->prm), h->prm.kahan != 0);
        else launch_direct(h->stream, fa, feature_mask(h->prm), h->prm.kahan != 0);
    }
    if (ghosts) ghost_finish(h->stream, h->gh.ctrl, h->ctrl);
    if (bh && far == GEON_FAR_LEARNED && !collect) {
        Timer t(h, 4);
        launch_far_mlp(h->stream, far_args(h), n);
    }
    if (fa.hist.on && h->n_dead > 0) launch_dead_forces(h->stream, fa, feature_mask(h->prm));
    if (fa.sel_out) launch_sel_forces(h->stream, fa, bh, bh && far == GEON_FAR_LEARNED && !collect);
    if (higgs || axion) {   // field forces after the pair forces (integrator.js:1494-1507)
        Timer tm(h, 7);
        field_forces(h->stream, field_args(h, 0), field_args(h, 1), higgs, axion, !!(tg & GEON_GRAVITY),
                     !!(tg & GEON_COULOMB), !!(tg & GEON_YUKAWA), particle_view(h), h->fc[K_FX], h->fc[K_FY],
                     h->fc[K_PEG], h->fc[K_PFI], n, fa.sel, fa.sel_out);
    }
    if (col_bounce) {           // _applyRepulsion (integrator.js:779, 1545)
        int rc = repulsion(h, fa);
        if (rc) return rc;
    }
    if (h->prm.boundary_mode == GEON_BOUND_BOUNCE) {    // after every other force (integrator.js:780, 1546)
        WallArgs w;
        w.accumulate = col_bounce;
        w.x = F(h, C_X); w.y = F(h, C_Y); w.vx = F(h, C_VX); w.vy = F(h, C_VY); w.ang_vel = F(h, C_ANGVEL);
        w.radius = F(h, C_RADIUS); w.fx = h->fc[K_FX]; w.fy = h->fc[K_FY]; w.contact = h->fc[K_CONTACT];
        w.sel = fa.sel; w.sel_out = fa.sel_out; w.ctrl = h->ctrl;
        w.w = (float)h->prm.domain_w; w.h = (float)h->prm.domain_h; w.friction = (float)h->prm.bounce_friction;
        launch_wall_forces(h->stream, w, n);
    }
    HIP_OK(hipGetLastError());
    return GEON_OK;
}

// 1PN velocity-Verlet correction (integrator.js:1296-1318): tree at the drifted
// positions, compute1PN, then w += (F1PN_new - F1PN_old) dt / 2m.  The moment
// caches are left as the last force pass wrote them, as in the reference.
int one_pn_step(GeonContext* h) {
    const int n = h->n_host;
    if (n == 0) return GEON_OK;
    const bool bh = !!(h->prm.toggles0 & GEON_BARNES_HUT);
    if (bh) {
        int rc = sort_and_gather(h);
        if (rc) return rc;
        Timer t(h, 1);
        build_tree(h->stream, h->tb, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_WX), F(h, C_WY), F(h, C_MU),
                   F(h, C_L), I(h, C_QQ), n, key_params(h->prm), h->ctrl, want_moments(h));
        h->tree_valid = true;
    }
    const uint32_t tg = h->prm.toggles0;
    launch_prepare_sources(h->stream, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_CHARGE), F(h, C_VX), F(h, C_VY),
                           F(h, C_ANGVEL), F(h, C_RADIUS), F(h, C_MU), F(h, C_L),
                           (tg & GEON_AXION) ? F(h, C_AX) : nullptr, (tg & GEON_AXION) ? F(h, C_YUK) : nullptr,
                           (tg & GEON_HIGGS) ? F(h, C_HIGGS) : nullptr, h->src4, h->src8a, h->src8b, n, h->ctrl,
                           /*update_moments=*/0);
    ForceArgs fa = force_args(h);
    // compute1PN takes every leaf from its past light cone, with or without
    // relativity (forces.js:542-546, 600-602); in the reference's mode without
    // relativity no history is ever recorded, so leaves add nothing.
    fa.pn_leaves = !(h->prm.signal_delay == GEON_DELAY_FOLLOW && !delay_on(h->prm));
    launch_walk1pn(h->stream, fa, bh);
    launch_vv_kick(h->stream, step_args(h));
    HIP_OK(hipGetLastError());
    return GEON_OK;
}

int check_overflow(GeonContext* h) {
    switch (h->ctrl_host->overflow) {
        case 0: return GEON_OK;
        case 3: return set_err(h, GEON_EOVERFLOW, "far-field pair list full: raise far_list_cap");
        case 4: return set_err(h, GEON_EOVERFLOW, "bounce collisions: more overlap pairs than 8 per body");
        default: return set_err(h, GEON_EOVERFLOW, "tree node pool or walk stack overflow");
    }
}

// Dead-particle GC (main.js:874-885): a retired source is dropped
// 2 sqrt(W^2 + H^2) time units after its death (once per update here, once
// per frame in the reference).
int dead_gc(GeonContext* h) {
    if (h->n_dead == 0) return GEON_OK;
    const double W = h->prm.domain_w, H = h->prm.domain_h;
    const double max_dist = 2.0 * sqrt(W * W + H * H);
    int w = 0;
    for (int k = 0; k < h->n_dead; k++) {
        if (h->sim_time - h->dead_time[k] < max_dist) {
            h->dead_host[w] = h->dead_host[k];
            h->dead_time[w] = h->dead_time[k];
            w++;
        }
    }
    if (w == h->n_dead) return GEON_OK;
    h->n_dead = w;
    h->dead_host.resize((size_t)w);
    h->dead_time.resize((size_t)w);
    if (w) HIP_OK(hipMemcpyAsync(h->dead, h->dead_host.data(), sizeof(DeadRec) * w, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

int read_ctrl(GeonContext* h) {
    HIP_OK(hipMemcpyAsync(h->ctrl_host, h->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

// Remove despawned particles, keeping the survivors' relative order
// (integrator.js:1821-1823 compacts the array the same way).
int compact(GeonContext* h, double t_death) {
    const int n = h->n_host;
    int32_t* total = &h->ctrl->n_tree;    // scratch: n_tree is rebuilt before every use
    exclusive_scan_i32(h->stream, h->alive, h->alive_scan, n, &h->ctrl->n_alive, total, h->tb.scan_tmp);
    // perm[j] = old index of the j-th survivor
    hipLaunchKernelGGL(k_survivor_perm, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->alive, h->alive_scan,
                       h->compact_perm, h->ctrl);
    HIP_OK(hipMemcpyAsync(h->ctrl_host, h->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    const int survivors = h->ctrl_host->n_tree;
    const int deaths = n - survivors;
    if (deaths > 0 && delay_on(h->prm)) {
        // _retireParticle (integrator.js:152-173): death snapshots and dead records
        int rc = hist_alloc(h);
        if (rc) return rc;
        const uint32_t tg = h->prm.toggles0;
        const bool axion = !!(tg & GEON_AXION), higgs = !!(tg & GEON_HIGGS);
        launch_retire(h->stream, hist_write(h), F(h, C_X), F(h, C_Y), F(h, C_VX), F(h, C_VY), F(h, C_ANGW),
                      F(h, C_MASS), F(h, C_CHARGE), F(h, C_RADIUS), axion ? F(h, C_AX) : nullptr,
                      axion ? F(h, C_YUK) : nullptr, higgs ? F(h, C_HIGGS) : nullptr, I(h, C_SLOT), h->alive,
                      h->alive_scan, h->n_events, h->n_dead, n, h->ctrl);
        h->dead_host.resize((size_t)(h->n_dead + deaths));
        HIP_OK(hipMemcpyAsync(h->dead_host.data() + h->n_dead, h->dead + h->n_dead, sizeof(DeadRec) * deaths,
                              hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        h->dead_time.resize((size_t)(h->n_dead + deaths), t_death);
        h->n_dead += deaths;
    }
    hipLaunchKernelGGL(k_set_n, dim3(1), dim3(1), 0, h->stream, h->ctrl, survivors);
    Columns cols;
    cols.ncols = 0;
    for (int c = 0; c < N_STATE; c++) { cols.src[cols.ncols] = h->st[c]; cols.dst[cols.ncols++] = h->st_alt[c]; }
    for (int c = 0; c < N_FORCE; c++) {
        cols.src[cols.ncols] = reinterpret_cast<uint32_t*>(h->fc[c]);
        cols.dst[cols.ncols++] = reinterpret_cast<uint32_t*>(h->fc_alt[c]);
    }
    launch_gather(h->stream, cols, h->compact_perm, survivors, h->ctrl);
    for (int c = 0; c < N_STATE; c++) std::swap(h->st[c], h->st_alt[c]);
    for (int c = 0; c < N_FORCE; c++) std::swap(h->fc[c], h->fc_alt[c]);
    h->n_host = survivors;
    h->tree_valid = false;
    sel_locate(h);
    return GEON_OK;
}


// ---- Allocation-free collision pair list ----------------------------------
// Device scratch of geon_collision_pairs_noalloc: the pair staging, the packed
// sort keys with their permutation, and the radix sort's histograms, for at
// most kNoAllocPairs pairs (about 33 MB).  Every GPU the library runs on holds
// one copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (na_scratch).  Calls that
// overlap on one GPU share it and race.
constexpr int64_t kNoAllocPairs = int64_t(1) << 20;
constexpr size_t kNaCtrl = 0;
constexpr size_t kNaPa = 256;
constexpr size_t kNaPb = kNaPa + 4 * kNoAllocPairs;
constexpr size_t kNaKeys = kNaPb + 4 * kNoAllocPairs;
constexpr size_t kNaKeysAlt = kNaKeys + 8 * kNoAllocPairs;
constexpr size_t kNaPerm = kNaKeysAlt + 8 * kNoAllocPairs;
constexpr size_t kNaPermAlt = kNaPerm + 4 * kNoAllocPairs;
constexpr size_t kNaHist = kNaPermAlt + 4 * kNoAllocPairs;
constexpr size_t kNaDigit = kNaHist + 4 * 256 * ((kNoAllocPairs + kSortTileKeys - 1) / kSortTileKeys);
constexpr size_t kNaBytes = kNaDigit + 4 * 256;
static_assert(sizeof(Ctrl) <= kNaPa, "controller record fits its scratch slot");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kMaxDevices = 64;
std::mutex g_na_mutex;
unsigned char* g_na_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
int noalloc_init(GeonContext* h) {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_err(h, GEON_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(h, GEON_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_na_ptr[dev] = static_cast<unsigned char*>(p);
    }
    h->na_scratch = g_na_ptr[dev];
    return GEON_OK;
}

// (a, b) slot pairs -> keys a << bits | b, so one radix sort orders them by (a, b).
__global__ void k_pack_pairs(const int32_t* __restrict__ pa, const int32_t* __restrict__ pb, uint64_t* __restrict__ keys,
                             uint32_t* __restrict__ perm, int bits, const Ctrl* __restrict__ c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c->n_alive) return;
    keys[i] = ((uint64_t)(uint32_t)pa[i] << bits) | (uint32_t)pb[i];
    perm[i] = (uint32_t)i;
}

__global__ void k_split_pairs(const uint64_t* __restrict__ keys, int32_t* __restrict__ pa, int32_t* __restrict__ pb,
                              int bits, const Ctrl* __restrict__ c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c->n_alive) return;
    const uint64_t k = keys[i];
    pa[i] = (int32_t)(k >> bits);
    pb[i] = (int32_t)(k & ((uint64_t(1) << bits) - 1));
}

// Buffers of one sort + tree build over up to `cap` points.
int tree_alloc(GeonContext* h, TreeBufs& tb, size_t cap) {
    int rc;
    tb.cap_particles = (int64_t)cap;
    tb.cap_divided = (int64_t)cap + 1024;
    const size_t nodes = 1 + 4 * (size_t)tb.cap_divided;
    const size_t nblk = (cap + 2047) / 2048 + 1;
    if ((rc = dmalloc(h, &tb.keys, cap))) return rc;
    if ((rc = dmalloc(h, &tb.keys_alt, cap))) return rc;
    if ((rc = dmalloc(h, &tb.perm, cap))) return rc;
    if ((rc = dmalloc(h, &tb.perm_alt, cap))) return rc;
    if ((rc = dmalloc(h, &tb.hist, 256 * nblk))) return rc;
    if ((rc = dmalloc(h, &tb.digit_total, 256))) return rc;
    if ((rc = dmalloc(h, &tb.owned, cap))) return rc;
    if ((rc = dmalloc(h, &tb.rank_start, cap))) return rc;
    if ((rc = dmalloc(h, &tb.divnode, (size_t)tb.cap_divided))) return rc;
    if ((rc = dmalloc(h, &tb.pre, 8 * (cap + 1)))) return rc;
    if ((rc = dmalloc(h, &tb.nodes, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.node_mass64, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.node_comx64, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.node_comy64, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.node_qq64, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.node_path, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.node_npts, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.node_first, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.node_mom, nodes))) return rc;
    if ((rc = dmalloc(h, &tb.pre_mom, 9 * (cap + 1)))) return rc;
    if ((rc = dmalloc(h, &tb.scan_tmp, nblk + 1024))) return rc;
    if ((rc = dmalloc(h, &tb.scan_tmp64, 9 * (nblk + 1024)))) return rc;
    return GEON_OK;
}

void tree_free(TreeBufs& tb) {
    hipFree(tb.keys); hipFree(tb.keys_alt); hipFree(tb.perm); hipFree(tb.perm_alt); hipFree(tb.hist);
    hipFree(tb.digit_total);
    hipFree(tb.owned); hipFree(tb.rank_start); hipFree(tb.divnode); hipFree(tb.pre); hipFree(tb.nodes);
    hipFree(tb.node_mass64); hipFree(tb.node_comx64); hipFree(tb.node_comy64); hipFree(tb.node_qq64);
    hipFree(tb.node_path); hipFree(tb.node_npts); hipFree(tb.scan_tmp); hipFree(tb.scan_tmp64);
    hipFree(tb.node_first); hipFree(tb.node_mom); hipFree(tb.pre_mom);
    tb = TreeBufs{};
}

void ghost_release(GeonContext* h);

// Periodic images in the tree (GEON_PERIODIC_GHOSTS): the tree points (a
// particle and its images: `factor` per particle at most), their sort and
// tree, allocated at the first such pass.
int ghost_alloc(GeonContext* h, int factor) {
    GhostState& g = h->gh;
    if (g.cap && g.factor >= factor) return GEON_OK;
    ghost_release(h);
    const size_t cap4 = (size_t)factor * (size_t)h->cap;
    int rc;
    for (int c = 0; c < kGhostCols; c++)
        if ((rc = dmalloc(h, &g.col.c[c], cap4)) || (rc = dmalloc(h, &g.alt.c[c], cap4))) return rc;
    if ((rc = dmalloc(h, &g.cnt, (size_t)h->cap)) || (rc = dmalloc(h, &g.off, (size_t)h->cap)) ||
        (rc = dmalloc(h, &g.src4, cap4)) || (rc = dmalloc(h, &g.src8a, cap4)) || (rc = dmalloc(h, &g.src8b, cap4)) ||
        (rc = dmalloc(h, &g.ctrl, 1)))
        return rc;
    if ((rc = tree_alloc(h, g.tb, cap4))) return rc;
    g.cap = (int64_t)cap4;
    g.factor = factor;
    return GEON_OK;
}

void ghost_release(GeonContext* h) {
    GhostState& g = h->gh;
    if (!g.cap) return;
    for (int c = 0; c < kGhostCols; c++) { hipFree(g.col.c[c]); hipFree(g.alt.c[c]); }
    hipFree(g.cnt); hipFree(g.off); hipFree(g.src4); hipFree(g.src8a); hipFree(g.src8b); hipFree(g.ctrl);
    tree_free(g.tb);
    g = GhostState{};
}

int allocate(GeonContext* h) {
    const size_t cap = (size_t)h->cap;
    int rc;
    for (int c = 0; c < N_STATE; c++) {
        if ((rc = dmalloc(h, &h->st[c], cap))) return rc;
        if ((rc = dmalloc(h, &h->st_alt[c], cap))) return rc;
    }
    for (int c = 0; c < N_FORCE; c++) {
        if ((rc = dmalloc(h, &h->fc[c], cap))) return rc;
        if ((rc = dmalloc(h, &h->fc_alt[c], cap))) return rc;
    }
    if ((rc = dmalloc(h, &h->alive, cap))) return rc;
    if ((rc = dmalloc(h, &h->alive_scan, cap))) return rc;
    if ((rc = dmalloc(h, &h->compact_perm, cap))) return rc;
    if ((rc = dmalloc(h, &h->src4, cap))) return rc;
    if ((rc = dmalloc(h, &h->src8a, cap))) return rc;
    if ((rc = dmalloc(h, &h->src8b, cap))) return rc;
    if ((rc = dmalloc(h, &h->ctrl, 1))) return rc;
    if ((rc = dmalloc(h, &h->hist_fe, cap))) return rc;
    if ((rc = dmalloc(h, &h->hist_cd, cap))) return rc;
    if ((rc = dmalloc(h, &h->dead, cap))) return rc;
    if ((rc = tree_alloc(h, h->tb, cap))) return rc;
    if ((rc = dmalloc(h, &h->stat_partial, std::max<size_t>(16 * 1024, 3 * ((cap + 255) / 256) + 64)))) return rc;
    if ((rc = dmalloc(h, &h->stat_out, 32))) return rc;
    h->field_g = h->prm.field_grid > 0 ? h->prm.field_grid : 128;
    const size_t g2 = (size_t)h->field_g * h->field_g;
    for (int w = 0; w < 2; w++) {
        FieldState& s = h->fld[w];
        if ((rc = dmalloc(h, &s.field, g2))) return rc;
        if ((rc = dmalloc(h, &s.field_dot, g2))) return rc;
        if ((rc = dmalloc(h, &s.field_dot_alt, g2))) return rc;
        if ((rc = dmalloc(h, &s.grad_x, g2))) return rc;
        if ((rc = dmalloc(h, &s.grad_y, g2))) return rc;
        if ((rc = dmalloc(h, &s.source, g2))) return rc;
        if ((rc = dmalloc(h, &s.thermal, g2))) return rc;
        if ((rc = dmalloc(h, &s.sg_phi, g2))) return rc;
        if ((rc = dmalloc(h, &s.sg_gx, g2))) return rc;
        if ((rc = dmalloc(h, &s.sg_gy, g2))) return rc;
        if ((rc = dmalloc(h, &s.dep_src, g2))) return rc;
        if ((rc = dmalloc(h, &s.dep_th, g2))) return rc;
        if ((rc = dmalloc(h, &s.green, g2))) return rc;
        if ((rc = dmalloc(h, &s.has_energy, 1))) return rc;
        if ((rc = dmalloc(h, &s.sg_c, g2))) return rc;
        if ((rc = dmalloc(h, &s.sg_rows, (size_t)h->field_g))) return rc;
        HIP_OK(hipMemset(s.dep_src, 0, 8 * g2));
        HIP_OK(hipMemset(s.dep_th, 0, 8 * g2));
    }
    if ((rc = dmalloc(h, &h->field_out, 8))) return rc;
    HIP_OK(hipHostMalloc((void**)&h->ctrl_host, sizeof(Ctrl), hipHostMallocDefault));
    HIP_OK(hipMemset(h->ctrl, 0, sizeof(Ctrl)));
    HIP_OK(hipEventCreateWithFlags(&h->ctrl_ev, hipEventDisableTiming));
    return GEON_OK;
}

void merge_release(GeonContext* h);
void rad_release(GeonContext* h);

void release(GeonContext* h) {
    for (int c = 0; c < N_STATE; c++) { hipFree(h->st[c]); hipFree(h->st_alt[c]); }
    for (int c = 0; c < N_FORCE; c++) { hipFree(h->fc[c]); hipFree(h->fc_alt[c]); }
    hipFree(h->alive); hipFree(h->alive_scan); hipFree(h->compact_perm);
    hipFree(h->src4); hipFree(h->src8a); hipFree(h->src8b); hipFree(h->ctrl);
    hipFree(h->hist4); hipFree(h->hist_aw); hipFree(h->hist_fe); hipFree(h->hist_cd); hipFree(h->ev_time);
    hipFree(h->dead);
    hipFree(h->sel_idx); hipFree(h->sel_rec);
    for (int32_t* p : {h->bc_pa, h->bc_pb, h->bc_cnt1, h->bc_off1, h->bc_cnt2, h->bc_off2, h->bc_fill2, h->bc_list2,
                       h->bc_tot})
        hipFree(p);
    TreeBufs& tb = h->tb;
    tree_free(tb);
    ghost_release(h);
    merge_release(h);
    rad_release(h);
    far_release(h);
    hipFree(h->stat_partial); hipFree(h->stat_out); hipFree(h->field_out);
    for (int w = 0; w < 2; w++) {
        FieldState& s = h->fld[w];
        float* fa[] = {s.field, s.field_dot, s.field_dot_alt, s.grad_x, s.grad_y, s.source, s.thermal,
                       s.sg_phi, s.sg_gx, s.sg_gy};
        for (float* p : fa) hipFree(p);
        hipFree(s.dep_src); hipFree(s.dep_th); hipFree(s.green); hipFree(s.has_energy); hipFree(s.sg_c); hipFree(s.sg_rows);
    }
    if (h->ctrl_host) hipHostFree(h->ctrl_host);
    if (h->ctrl_ev) hipEventDestroy(h->ctrl_ev);
    for (int k = 0; k < kTimers; k++)
        for (auto& pr : h->ev[k]) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
    if (h->stream) hipStreamDestroy(h->stream);
}

}  // namespace

namespace {

uint64_t splitmix(uint64_t& z) {
    uint64_t r = (z += 0x9E3779B97F4A7C15ull);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
}

// Input means / inverse deviations and output scales from (a subsample of)
// the held samples; also the loss-weight normaliser 1 / mean(w).
int far_sample_stats(GeonContext* h, std::vector<float>* stats, float* wscale) {
    const int64_t ns = h->far_n_samples;
    const int64_t take = std::min<int64_t>(ns, 1 << 18);
    const int64_t stride = std::max<int64_t>(1, ns / std::max<int64_t>(take, 1));
    std::vector<float> buf((size_t)take * kFarSample);
    HIP_OK(hipMemcpy2D(buf.data(), sizeof(float) * kFarSample, h->far_samples, sizeof(float) * kFarSample * stride,
                       sizeof(float) * kFarSample, (size_t)take, hipMemcpyDeviceToHost));
    double sum[kFarIn] = {}, sq[kFarIn] = {}, tq[kFarOut] = {}, tn[kFarOut] = {}, wsum = 0, wn = 0;
    for (int64_t r = 0; r < take; r++) {
        const float* s = &buf[(size_t)r * kFarSample];
        const float* w = s + kFarIn + kFarOut;
        for (int k = 0; k < kFarIn; k++) { sum[k] += s[k]; sq[k] += (double)s[k] * s[k]; }
        for (int j = 0; j < kFarOut; j++) {
            const float wj = w[j < 4 ? (j >> 1) : j - 2];
            if (wj > 0) { tq[j] += (double)s[kFarIn + j] * s[kFarIn + j]; tn[j] += 1; }
        }
        for (int j = 0; j < kFarW; j++) if (w[j] > 0) { wsum += w[j]; wn += 1; }
    }
    stats->assign(kFarStats, 0.f);
    for (int k = 0; k < kFarIn; k++) {
        const double m = take ? sum[k] / take : 0.0;
        const double var = take ? std::max(0.0, sq[k] / take - m * m) : 0.0;
        (*stats)[k] = (float)m;
        (*stats)[kFarIn + k] = (float)(1.0 / std::max(sqrt(var), 1e-6));
    }
    for (int j = 0; j < kFarOut; j++)
        (*stats)[2 * kFarIn + j] = (float)std::max(tn[j] > 0 ? sqrt(tq[j] / tn[j]) : 1.0, 1e-6);
    *wscale = (float)(wsum > 0 ? wn / wsum : 1.0);
    return GEON_OK;
}

int far_pack(GeonContext* h) {
    launch_far_pack(h->stream, h->far_params, h->far_stats, h->far_frags);
    HIP_OK(hipGetLastError());
    return GEON_OK;
}

int far_blocks(GeonContext* h, int64_t nblocks) {
    if (nblocks <= h->far_blocks_cap) return GEON_OK;
    hipFree(h->far_gpart); hipFree(h->far_lpart);
    h->far_gpart = nullptr; h->far_lpart = nullptr;
    int rc;
    if ((rc = dmalloc(h, &h->far_gpart, (size_t)(nblocks * kFarParams)))) return rc;
    if ((rc = dmalloc(h, &h->far_lpart, (size_t)nblocks))) return rc;
    h->far_blocks_cap = nblocks;
    return GEON_OK;
}

FarTrainArgs far_train_args(GeonContext* h, float wscale) {
    FarTrainArgs a;
    a.frags = (const FarFrags*)h->far_frags;
    a.samples = h->far_samples;
    a.n_samples = h->far_n_samples;
    a.rows = nullptr;
    a.batch = 0;
    a.seed = 0;
    a.step = 0;
    a.wscale = wscale;
    a.grad_part = h->far_gpart; a.loss_part = h->far_lpart;
    a.grad = h->far_grad; a.loss_out = h->far_loss;
    a.params = h->far_params; a.adam_m = h->far_m; a.adam_v = h->far_v;
    return a;
}

}  // namespace

namespace {

// ---- radiation reaction (integrator.js:983-1086, 1567-1740) ----

void rad_release(GeonContext* h) {
    hipFree(h->rad_scr); hipFree(h->rad_partial); hipFree(h->rad_sums); hipFree(h->rad_tot);
    h->rad_scr = h->rad_partial = h->rad_sums = h->rad_tot = nullptr;
}

// Scratch of the radiation passes, allocated when radiation is first on.
int rad_alloc(GeonContext* h) {
    if (h->rad_scr) return GEON_OK;
    const size_t cap = (size_t)h->cap, nb = cap / 256 + 1;
    int rc;
    if ((rc = dmalloc(h, &h->rad_scr, 9 * cap)) || (rc = dmalloc(h, &h->rad_partial, 9 * nb)) ||
        (rc = dmalloc(h, &h->rad_sums, 16)) || (rc = dmalloc(h, &h->rad_tot, 4)))
        return rc;
    HIP_OK(hipMemset(h->rad_tot, 0, 4 * sizeof(double)));
    return GEON_OK;
}

QuadArgs quad_args(GeonContext* h) {
    QuadArgs q;
    q.x = F(h, C_X); q.y = F(h, C_Y); q.vx = F(h, C_VX); q.vy = F(h, C_VY);
    q.fx = h->fc[K_FX]; q.fy = h->fc[K_FY]; q.jx = h->fc[K_JX]; q.jy = h->fc[K_JY];
    q.mass = F(h, C_MASS); q.charge = F(h, C_CHARGE);
    q.wx = F(h, C_WX); q.wy = F(h, C_WY);
    q.scr = h->rad_scr; q.stride = h->cap;
    q.partial = h->rad_partial; q.sums = h->rad_sums; q.rad_tot = h->rad_tot;
    q.ctrl = h->ctrl;
    q.gw = !!(h->prm.toggles0 & GEON_GRAVITY);
    q.em = !!(h->prm.toggles0 & GEON_COULOMB);
    return q;
}

// ---- collision mode merge (collisions.js:24-158, integrator.js:1336-1416) ----

void merge_free_pairs(GeonContext* h) {
    hipFree(h->mg_p1); hipFree(h->mg_p2); hipFree(h->mg_leaf);
    hipFree(h->mg_sort.keys); hipFree(h->mg_sort.keys_alt); hipFree(h->mg_sort.perm); hipFree(h->mg_sort.perm_alt);
    hipFree(h->mg_sort.hist); hipFree(h->mg_sort.digit_total);
    hipFree(h->mg_csort.keys); hipFree(h->mg_csort.keys_alt); hipFree(h->mg_csort.perm);
    hipFree(h->mg_csort.perm_alt); hipFree(h->mg_csort.hist); hipFree(h->mg_csort.digit_total);
    h->mg_csort = {};
    hipFree(h->mg_fl); hipFree(h->mg_scan_tmp); hipFree(h->mg_res); hipFree(h->mg_spawns); hipFree(h->mg_events);
    h->mg_p1 = h->mg_p2 = h->mg_leaf = nullptr;
    h->mg_sort = {};
    h->mg_fl = h->mg_scan_tmp = nullptr;
    h->mg_res = h->mg_spawns = nullptr;
    h->mg_events = nullptr;
    h->mg_cap_pairs = 0;
}

void merge_release(GeonContext* h) {
    merge_free_pairs(h);
    hipFree(h->mg_deg); hipFree(h->mg_rank); hipFree(h->mg_off); hipFree(h->mg_s5); hipFree(h->mg_mark);
    hipFree(h->mg_chunk5); hipFree(h->mg_work);
    h->mg_chunk5 = nullptr;
    hipFree(h->mg_cnt); hipFree(h->mg_ctrl); hipFree(h->mg_uf); hipFree(h->mg_uf_fail); hipFree(h->mg_cctrl);
    h->mg_uf = h->mg_uf_fail = nullptr;
    h->mg_cctrl = nullptr;
    h->mg_deg = h->mg_rank = h->mg_off = h->mg_s5 = h->mg_mark = h->mg_cnt = nullptr;
    h->mg_work = nullptr;
    h->mg_ctrl = nullptr;
}

// Scratch sized by the particle capacity, allocated at the first merge pass.
int merge_alloc(GeonContext* h) {
    if (h->mg_deg) return GEON_OK;
    const size_t cap = (size_t)h->cap;
    int rc;
    if ((rc = dmalloc(h, &h->mg_deg, cap)) || (rc = dmalloc(h, &h->mg_rank, cap)) ||
        (rc = dmalloc(h, &h->mg_off, cap)) || (rc = dmalloc(h, &h->mg_work, 4 * cap)) ||
        (rc = dmalloc(h, &h->mg_cnt, 8)) || (rc = dmalloc(h, &h->mg_ctrl, 1)) || (rc = dmalloc(h, &h->mg_uf, cap)) ||
        (rc = dmalloc(h, &h->mg_uf_fail, 1)) || (rc = dmalloc(h, &h->mg_cctrl, 1)))
        return rc;
    HIP_OK(hipMemset(h->mg_ctrl, 0, sizeof(Ctrl)));
    HIP_OK(hipMemset(h->mg_cctrl, 0, sizeof(Ctrl)));
    return GEON_OK;
}

// Buffers sized by the number of pairs, grown when a pass needs more.
int merge_reserve_pairs(GeonContext* h, int64_t np) {
    if (np <= h->mg_cap_pairs) return GEON_OK;
    HIP_OK(hipStreamSynchronize(h->stream));
    merge_free_pairs(h);
    const size_t cap = (size_t)std::max<int64_t>(std::max<int64_t>(np, 2 * h->mg_cap_pairs), 1024);
    const size_t nblk = cap / 2048 + 1;
    int rc;
    TreeBufs& sb = h->mg_sort;
    if ((rc = dmalloc(h, &h->mg_p1, cap)) || (rc = dmalloc(h, &h->mg_p2, cap)) || (rc = dmalloc(h, &h->mg_leaf, cap)) ||
        (rc = dmalloc(h, &sb.keys, cap)) || (rc = dmalloc(h, &sb.keys_alt, cap)) || (rc = dmalloc(h, &sb.perm, cap)) ||
        (rc = dmalloc(h, &sb.perm_alt, cap)) || (rc = dmalloc(h, &sb.hist, 256 * nblk)) ||
        (rc = dmalloc(h, &sb.digit_total, 256)) || (rc = dmalloc(h, &h->mg_fl, 9 * cap)) ||
        (rc = dmalloc(h, &h->mg_scan_tmp, nblk + 16)) || (rc = dmalloc(h, &h->mg_res, cap)) ||
        (rc = dmalloc(h, &h->mg_spawns, cap)) || (rc = dmalloc(h, &h->mg_events, 6 * cap)))
        return rc;
    TreeBufs& cb = h->mg_csort;
    if ((rc = dmalloc(h, &cb.keys, cap)) || (rc = dmalloc(h, &cb.keys_alt, cap)) || (rc = dmalloc(h, &cb.perm, cap)) ||
        (rc = dmalloc(h, &cb.perm_alt, cap)) || (rc = dmalloc(h, &cb.hist, 256 * nblk)) ||
        (rc = dmalloc(h, &cb.digit_total, 256)))
        return rc;
    h->mg_cap_pairs = (int64_t)cap;
    return GEON_OK;
}

// Step 6 of a substep in collision mode merge: the pair list over a tree at
// the current positions, the fold, removal (retired for signal delay) and the
// spawned particles, field excitations.  *changed: the particle set changed.
int merge_step(GeonContext* h, bool* changed) {
    *changed = false;
    const int n = h->n_host;
    if (n == 0) return GEON_OK;
    int rc;
    if ((rc = merge_alloc(h))) return rc;
    if ((rc = sort_and_gather(h))) return rc;
    // the tree the force pass would build at these positions (it walks this one
    // when the pass changes nothing)
    build_tree(h->stream, h->tb, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_WX), F(h, C_WY), F(h, C_MU), F(h, C_L),
               I(h, C_QQ), n, key_params(h->prm), h->ctrl, want_moments(h));
    launch_prepare_sources(h->stream, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_CHARGE), F(h, C_VX), F(h, C_VY),
                           F(h, C_ANGVEL), F(h, C_RADIUS), F(h, C_MU), F(h, C_L), nullptr, nullptr, nullptr, h->src4,
                           h->src8a, h->src8b, n, h->ctrl, /*update_moments=*/0);
    HIP_OK(hipMemsetAsync(h->mg_cnt, 0, 8 * sizeof(int32_t), h->stream));
    PairArgs pa;
    pa.x = F(h, C_X); pa.y = F(h, C_Y); pa.mass = F(h, C_MASS); pa.radius = F(h, C_RADIUS);
    pa.slot = I(h, C_SLOT); pa.src4 = h->src4; pa.nodes = h->tb.nodes; pa.max_nodes = 1 + 4 * h->tb.cap_divided;
    pa.count = h->mg_deg; pa.offset = h->mg_off; pa.pa = nullptr; pa.pb = nullptr; pa.pl = nullptr;
    pa.index_out = 1;
    pa.slack = 1.0f + 1.0f / 4096;      // a superset: the fold tests every pair again in FP64
    pa.dom_w = (float)h->prm.domain_w; pa.dom_h = (float)h->prm.domain_h;
    pa.periodic = h->prm.boundary_mode == GEON_BOUND_LOOP;
    pa.ctrl = h->ctrl; pa.n_upper = n;
    launch_pairs(h->stream, pa, false);
    exclusive_scan_i32(h->stream, h->mg_deg, h->mg_off, n, &h->ctrl->n_alive, h->mg_cnt + 4, h->tb.scan_tmp);
    int32_t cnt[8];
    HIP_OK(hipMemcpyAsync(cnt, h->mg_cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    if ((rc = read_ctrl(h))) return rc;
    if (h->ctrl_host->overflow) return check_overflow(h);
    const int np = cnt[4];
    const int n_div = h->ctrl_host->n_divided;
    const double t_now = h->ctrl_host->sim_time;
    if (np == 0) return GEON_OK;
    if ((rc = merge_reserve_pairs(h, np))) return rc;
    hipLaunchKernelGGL(k_set_n, dim3(1), dim3(1), 0, h->stream, h->mg_ctrl, np);
    pa.pa = h->mg_p1; pa.pb = h->mg_p2; pa.pl = h->mg_leaf;
    launch_pairs(h->stream, pa, true);
    // leaf orders only where a particle meets two candidates in one leaf
    merge_pair_same_leaf(h->stream, h->mg_p1, h->mg_leaf, h->mg_ctrl, h->mg_cnt + 5, np);
    HIP_OK(hipMemcpyAsync(cnt, h->mg_cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    const bool need_rank = cnt[5] != 0;
    if (need_rank) {
        const size_t max_nodes = 1 + 4 * (size_t)h->tb.cap_divided;
        if (!h->mg_s5 && ((rc = dmalloc(h, &h->mg_s5, 5 * max_nodes)) || (rc = dmalloc(h, &h->mg_mark, max_nodes)) ||
                          (rc = dmalloc(h, &h->mg_chunk5, 5 * ((size_t)h->cap / 256 + 1)))))
            return rc;
        merge_leaf_ranks(h->stream, h->tb, I(h, C_SLOT), 1 + 4 * n_div, h->mg_s5, h->mg_mark, h->mg_chunk5, n, h->ctrl,
                         h->mg_p1, h->mg_leaf, h->mg_ctrl, h->mg_rank, np);
    }
    merge_pair_keys(h->stream, h->mg_p1, h->mg_p2, h->mg_leaf, I(h, C_SLOT), need_rank ? h->mg_rank : nullptr,
                    h->mg_ctrl, h->ctrl, h->mg_sort.keys, h->mg_sort.perm, np);
    radix_sort(h->stream, h->mg_sort, np, 64, h->mg_ctrl);
    MergeArgs m;
    m.x = F(h, C_X); m.y = F(h, C_Y); m.wx = F(h, C_WX); m.wy = F(h, C_WY);
    m.mass = F(h, C_MASS); m.base_mass = F(h, C_BASEMASS); m.radius = F(h, C_RADIUS); m.angw = F(h, C_ANGW);
    m.qq = I(h, C_QQ); m.flags = h->st[C_FLAGS];
    m.p1 = h->mg_p1; m.p2 = h->mg_p2; m.order = h->mg_sort.perm;
    m.deg = h->mg_deg;
    const size_t cap = (size_t)h->cap;
    m.wm = h->mg_work; m.wbm = h->mg_work + cap; m.wr = h->mg_work + 2 * cap; m.wdm = h->mg_work + 3 * cap;
    m.res = h->mg_res; m.spawns = h->mg_spawns;
    const size_t pc = (size_t)h->mg_cap_pairs;
    m.f_spawn = h->mg_fl; m.f_mev = h->mg_fl + pc; m.f_aev = h->mg_fl + 2 * pc;
    m.o_spawn = h->mg_fl + 3 * pc; m.o_mev = h->mg_fl + 4 * pc; m.o_aev = h->mg_fl + 5 * pc;
    m.f_fold = h->mg_fl + 6 * pc; m.o_fold = h->mg_fl + 7 * pc; m.fold = h->mg_fl + 8 * pc;
    m.cnt = h->mg_cnt; m.scan_tmp = h->mg_scan_tmp; m.events = h->mg_events;
    m.mc = h->mg_ctrl; m.ctrl = h->ctrl;
    m.periodic = h->prm.boundary_mode == GEON_BOUND_LOOP;
    m.dom_w = h->prm.domain_w; m.dom_h = h->prm.domain_h;
    m.uf = h->mg_uf; m.uf_fail = h->mg_uf_fail;
    HIP_OK(merge_resolve(h->stream, m, h->mg_csort, h->mg_cctrl, np, n));
    // The spawn count is known once the fold has run: a pass whose merged
    // products would not fit the slot space fails before it changes any
    // particle (merge_apply below is the first write to the particle columns).
    HIP_OK(hipMemcpyAsync(cnt, h->mg_cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    const int n_spawn = cnt[0], n_ev = cnt[1] + cnt[2];
    if (n_spawn > 0 && h->next_slot + n_spawn > h->cap)
        return set_err(h, GEON_EFULL, "slot space exhausted: merged particles need new slots (capacity bounds all slots ever used)");
    merge_apply(h->stream, m, F(h, C_MASS), F(h, C_BASEMASS), F(h, C_RADIUS), h->alive, n);
    HIP_OK(hipMemcpyAsync(cnt, h->mg_cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    const int deaths = cnt[3];
    if (n_ev > 0) {
        const size_t old = h->events_host.size();
        h->events_host.resize(old + 6 * (size_t)n_ev);
        HIP_OK(hipMemcpyAsync(h->events_host.data() + old, h->mg_events, sizeof(double) * 6 * n_ev,
                              hipMemcpyDeviceToHost, h->stream));
    }
    if (deaths > 0 && (rc = compact(h, t_now))) return rc;
    if (n_spawn > 0) {
        SpawnArgs sa;
        sa.spawns = h->mg_spawns;
        sa.count = n_spawn;
        sa.base = h->n_host;
        sa.first_slot = h->next_slot;
        sa.periodic = h->prm.boundary_mode == GEON_BOUND_LOOP;
        sa.rel = !!(h->prm.toggles0 & GEON_RELATIVITY);
        sa.dom_w = h->prm.domain_w; sa.dom_h = h->prm.domain_h;
        sa.x = F(h, C_X); sa.y = F(h, C_Y); sa.wx = F(h, C_WX); sa.wy = F(h, C_WY); sa.vx = F(h, C_VX);
        sa.vy = F(h, C_VY); sa.mass = F(h, C_MASS); sa.base_mass = F(h, C_BASEMASS); sa.charge = F(h, C_CHARGE);
        sa.angw = F(h, C_ANGW); sa.angvel = F(h, C_ANGVEL); sa.radius = F(h, C_RADIUS); sa.mu = F(h, C_MU);
        sa.L = F(h, C_L); sa.f1ox = F(h, C_F1OX); sa.f1oy = F(h, C_F1OY); sa.ax = F(h, C_AX); sa.yuk = F(h, C_YUK);
        sa.higgs = F(h, C_HIGGS); sa.qq = I(h, C_QQ); sa.slot = I(h, C_SLOT); sa.flags = h->st[C_FLAGS];
        sa.n_force = N_FORCE;
        for (int c = 0; c < N_FORCE; c++) sa.force[c] = h->fc[c];
        merge_spawn(h->stream, sa);
        hipLaunchKernelGGL(k_set_n, dim3(1), dim3(1), 0, h->stream, h->ctrl, h->n_host + n_spawn);
        launch_hist_reset(h->stream, h->hist_fe, h->hist_cd, h->next_slot, n_spawn, t_now);   // main.js:514
        h->n_host += n_spawn;
        h->next_slot += n_spawn;
    }
    const uint32_t t = h->prm.toggles0;
    const bool hg = !!(t & GEON_HIGGS), ax = !!(t & GEON_AXION);
    if (n_ev > 0 && (hg || ax)) {              // integrator.js:1394-1415
        const double g2h = 1.0, g2a = 0.05 * 0.05;        // HIGGS_COUPLING^2, AXION_COUPLING^2 (config.js:90, 97)
        const double fh = (hg && ax) ? g2h / (g2h + g2a) : 1.0, fa = (hg && ax) ? g2a / (g2h + g2a) : 1.0;
        for (int w = 0; w < 2; w++) {
            if (!(w == 0 ? hg : ax)) continue;
            const FieldArgs f = field_args(h, w);
            merge_excite(h->stream, f.field_dot, f.G, h->prm.domain_w, h->prm.domain_h, h->mg_events, n_ev,
                         w == 0 ? fh : fa);
        }
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    h->tree_valid = false;
    *changed = deaths > 0 || n_spawn > 0;
    return GEON_OK;
}

// The counting half of the collision pair list (collisions.js:35-60,
// quadtree.js:372-404): the tree at the current positions, each particle's
// pair count and their exclusive scan.  *pa is then ready for the writing pass
// but for its output arrays; *np is the total.
int pairs_count(GeonContext* h, PairArgs* pa_out, int64_t* np) {
    *np = 0;
    if (h->prm.boundary_mode == GEON_BOUND_LOOP && h->prm.topology != GEON_TORUS)
        return set_err(h, GEON_EUNSUPPORTED, "the pair list runs on open and torus domains");
    const int n = h->n_host;
    if (n == 0) return GEON_OK;
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, h->stream, h->ctrl);
    int rc = sort_and_gather(h);
    if (rc) return rc;
    build_tree(h->stream, h->tb, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_WX), F(h, C_WY), F(h, C_MU), F(h, C_L),
               I(h, C_QQ), n, key_params(h->prm), h->ctrl);
    h->tree_valid = true;
    launch_prepare_sources(h->stream, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_CHARGE), F(h, C_VX), F(h, C_VY),
                           F(h, C_ANGVEL), F(h, C_RADIUS), F(h, C_MU), F(h, C_L), nullptr, nullptr, nullptr, h->src4,
                           h->src8a, h->src8b, n, h->ctrl, /*update_moments=*/0);
    PairArgs& pa = *pa_out;
    pa.x = F(h, C_X); pa.y = F(h, C_Y); pa.mass = F(h, C_MASS); pa.radius = F(h, C_RADIUS);
    pa.slot = I(h, C_SLOT); pa.src4 = h->src4; pa.nodes = h->tb.nodes; pa.max_nodes = 1 + 4 * h->tb.cap_divided;
    pa.count = h->tb.owned; pa.offset = h->tb.rank_start; pa.pa = nullptr; pa.pb = nullptr;
    pa.pl = nullptr; pa.index_out = 0; pa.slack = 1.f;
    pa.dom_w = (float)h->prm.domain_w; pa.dom_h = (float)h->prm.domain_h;
    pa.periodic = h->prm.boundary_mode == GEON_BOUND_LOOP;
    pa.ctrl = h->ctrl; pa.n_upper = n;
    launch_pairs(h->stream, pa, false);
    int32_t* total = &h->ctrl->n_divided;     // scratch: rebuilt by the next tree build
    exclusive_scan_i32(h->stream, h->tb.owned, h->tb.rank_start, n, &h->ctrl->n_alive, total, h->tb.scan_tmp);
    if ((rc = read_ctrl(h))) return rc;
    *np = h->ctrl_host->n_divided;
    h->tree_valid = false;                    // n_divided no longer describes the tree
    h->tree_ghost = false;
    return GEON_OK;
}

}  // namespace

extern "C" {

int geon_abi_version(void) { return GEON_ABI_VERSION; }

void geon_default_params(GeonParams* p) {
    memset(p, 0, sizeof(*p));
    // Physics() defaults with barnesHut on (integrator.js:65-99), CPU constants (config.js)
    p->toggles0 = GEON_GRAVITY | GEON_COULOMB | GEON_BARNES_HUT;
    p->boundary_mode = GEON_BOUND_DESPAWN;
    p->topology = GEON_TORUS;
    p->collision_mode = GEON_COL_PASS;
    p->periodic_mode = GEON_PERIODIC_MIN_IMAGE;
    p->max_substeps = 32;
    p->domain_w = 64; p->domain_h = 64;
    p->softening = 8; p->bh_softening = 4;
    p->theta = 0.5;
    p->yukawa_coupling = 14; p->yukawa_mu = 0.15;
    p->higgs_mass = 0.5; p->axion_mass = 0.05; p->hubble = 0.001;
    p->ext_gravity_angle = M_PI * 0.5;
    p->field_grid = 128;
    p->kahan = 1;
    p->bounce_friction = 0.4;
}

int geon_create(const GeonParams* params, int32_t max_particles, GeonHandle* out) {
    if (!params || !out || max_particles <= 0) { g_create_error = "bad argument"; return GEON_EINVAL; }
    std::string why;
    if (validate(params, &why)) { g_create_error = why; return GEON_EUNSUPPORTED; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device"; return GEON_EDEVICE; }
    GeonContext* h = new GeonContext();
    h->prm = *params;
    h->cap = max_particles;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        delete h;
        return GEON_EDEVICE;
    }
    int rc = allocate(h);
    if (!rc) rc = noalloc_init(h);
    if (rc) { g_create_error = h->err; release(h); delete h; return rc; }
    reset_fields(h);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        g_create_error = "device error during field reset";
        release(h); delete h;
        return GEON_EDEVICE;
    }
    *out = h;
    return GEON_OK;
}

void geon_destroy(GeonHandle h) {
    if (!h) return;
    hipStreamSynchronize(h->stream);
    release(h);
    delete h;
}

int geon_reset(GeonHandle h) {
    if (!h) return GEON_EINVAL;
    h->n_host = 0; h->next_slot = 0; h->forces_init = false; h->tree_valid = false;
    h->sel_slot = -1;
    h->sim_time = 0; h->substeps_total = 0;
    hist_clear(h);
    if (h->rad_tot) HIP_OK(hipMemsetAsync(h->rad_tot, 0, 4 * sizeof(double), h->stream));
    hipLaunchKernelGGL(k_set_n, dim3(1), dim3(1), 0, h->stream, h->ctrl, 0);
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(1), 0, h->stream, h->ctrl, 0.0);
    reset_fields(h);
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

int geon_set_params(GeonHandle h, const GeonParams* p) {
    if (!h || !p) return GEON_EINVAL;
    std::string why;
    if (validate(p, &why)) return set_err(h, GEON_EUNSUPPORTED, why);
    h->prm = *p;
    h->forces_init = false;    // forces depend on the toggles: recompute at next update
    return GEON_OK;
}

int geon_get_params(GeonHandle h, GeonParams* out) {
    if (!h || !out) return GEON_EINVAL;
    *out = h->prm;
    return GEON_OK;
}

int geon_upload_soa(GeonHandle h, const GeonSoAView* v, int32_t n) {
    if (!h || !v || n < 0) return GEON_EINVAL;
    if (n > h->cap) return set_err(h, GEON_EFULL, "more particles than max_particles");
    if (n > 0 && (!v->x || !v->y || !v->wx || !v->wy || !v->mass || !v->charge))
        return set_err(h, GEON_EINVAL, "upload needs x, y, wx, wy, mass, charge");
    const bool rel = !!(h->prm.toggles0 & GEON_RELATIVITY);
    std::vector<std::vector<uint32_t>> col(N_STATE, std::vector<uint32_t>((size_t)n));
    auto putf = [&](int c, int i, double val) { float f = (float)val; memcpy(&col[c][i], &f, 4); };
    for (int i = 0; i < n; i++) {
        const double m = v->mass[i];
        if (!(m > 0)) return set_err(h, GEON_EINVAL, "mass must be > 0");
        // main.js:511, gpu-physics.js:1874: charge quantized to multiples of 0.1
        const int32_t qq = (int32_t)llround(v->charge[i] / 0.1);
        const double q = qq * 0.1;
        const double r = cbrt(m);   // particle.js:96
        const double angw = v->angw ? v->angw[i] : 0.0;
        const double wx = v->wx[i], wy = v->wy[i];
        const double ig = rel ? 1.0 / sqrt(1 + wx * wx + wy * wy) : 1.0;
        putf(C_X, i, v->x[i]); putf(C_Y, i, v->y[i]);
        putf(C_WX, i, wx); putf(C_WY, i, wy);
        putf(C_VX, i, wx * ig); putf(C_VY, i, wy * ig);
        putf(C_MASS, i, m); putf(C_BASEMASS, i, v->base_mass ? v->base_mass[i] : m);
        putf(C_CHARGE, i, q);
        putf(C_ANGW, i, angw);
        putf(C_ANGVEL, i, rel ? angw / sqrt(1 + angw * angw * r * r) : angw);
        putf(C_RADIUS, i, r);
        putf(C_MU, i, 0.0); putf(C_L, i, 0.0);          // particle.js:49-50: moments start at 0
        putf(C_F1OX, i, 0.0); putf(C_F1OY, i, 0.0);
        putf(C_AX, i, 1.0); putf(C_YUK, i, 1.0); putf(C_HIGGS, i, 1.0);
        col[C_QQ][i] = (uint32_t)qq;
        col[C_SLOT][i] = (uint32_t)i;
        col[C_FLAGS][i] = (v->antimatter && v->antimatter[i]) ? 1u : 0u;
    }
    for (int c = 0; c < N_STATE; c++)
        if (n) HIP_OK(hipMemcpyAsync(h->st[c], col[c].data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, h->stream));
    for (int c = 0; c < N_FORCE; c++) if (n) HIP_OK(hipMemsetAsync(h->fc[c], 0, sizeof(float) * n, h->stream));
    hipLaunchKernelGGL(k_set_n, dim3(1), dim3(1), 0, h->stream, h->ctrl, n);
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(1), 0, h->stream, h->ctrl, 0.0);
    // loaded particles are treated as always existing (save-load.js:231)
    hist_clear(h);
    if (h->rad_tot) HIP_OK(hipMemsetAsync(h->rad_tot, 0, 4 * sizeof(double), h->stream));
    launch_hist_reset(h->stream, h->hist_fe, h->hist_cd, 0, n, -INFINITY);
    HIP_OK(hipStreamSynchronize(h->stream));
    h->n_host = n;
    h->next_slot = n;
    h->forces_init = false;
    h->tree_valid = false;
    h->sim_time = 0;
    h->sel_slot = -1;           // a new system: slots name other particles
    return GEON_OK;
}

int32_t geon_add_particle(GeonHandle h, const GeonParticleInit* p) {
    if (!h || !p) return GEON_EINVAL;
    if (h->n_host >= h->cap || h->next_slot >= h->cap) return GEON_EFULL;   // gpu-physics.js:1864
    if (!(p->mass > 0)) return set_err(h, GEON_EINVAL, "mass must be > 0");
    const bool rel = !!(h->prm.toggles0 & GEON_RELATIVITY);
    uint32_t w[N_STATE];
    auto putf = [&](int c, double val) { float f = (float)val; memcpy(&w[c], &f, 4); };
    const int32_t qq = (int32_t)llround(p->charge / 0.1);
    const double r = cbrt(p->mass);
    const double ig = rel ? 1.0 / sqrt(1 + p->wx * p->wx + p->wy * p->wy) : 1.0;
    putf(C_X, p->x); putf(C_Y, p->y); putf(C_WX, p->wx); putf(C_WY, p->wy);
    putf(C_VX, p->wx * ig); putf(C_VY, p->wy * ig);
    putf(C_MASS, p->mass); putf(C_BASEMASS, p->mass); putf(C_CHARGE, qq * 0.1);
    putf(C_ANGW, p->angw);
    putf(C_ANGVEL, rel ? p->angw / sqrt(1 + p->angw * p->angw * r * r) : p->angw);
    putf(C_RADIUS, r); putf(C_MU, 0); putf(C_L, 0); putf(C_F1OX, 0); putf(C_F1OY, 0);
    putf(C_AX, 1); putf(C_YUK, 1); putf(C_HIGGS, 1);
    w[C_QQ] = (uint32_t)qq;
    const int32_t slot = h->next_slot;
    w[C_SLOT] = (uint32_t)slot;
    w[C_FLAGS] = p->antimatter ? 1u : 0u;
    const int i = h->n_host;
    for (int c = 0; c < N_STATE; c++) HIP_OK(hipMemcpyAsync(h->st[c] + i, &w[c], 4, hipMemcpyHostToDevice, h->stream));
    for (int c = 0; c < N_FORCE; c++) HIP_OK(hipMemsetAsync(h->fc[c] + i, 0, 4, h->stream));
    hipLaunchKernelGGL(k_set_n, dim3(1), dim3(1), 0, h->stream, h->ctrl, i + 1);
    launch_hist_reset(h->stream, h->hist_fe, h->hist_cd, slot, 1, h->sim_time);   // main.js:514 creationTime
    HIP_OK(hipStreamSynchronize(h->stream));
    h->n_host = i + 1;
    h->next_slot = slot + 1;
    h->forces_init = false;
    h->tree_valid = false;
    return slot;
}

static int find_slot(GeonHandle h, int32_t slot, int* idx) {
    std::vector<int32_t> slots((size_t)h->n_host);
    if (h->n_host)
        HIP_OK(hipMemcpyAsync(slots.data(), h->st[C_SLOT], 4 * (size_t)h->n_host, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < h->n_host; i++) if (slots[i] == slot) { *idx = i; return GEON_OK; }
    return set_err(h, GEON_EINVAL, "no live particle in that slot");
}

// The death metadata follows the callers' `p._deathMass || p.mass` and
// `p._deathAngVel || p.angVel` (input.js:183, 208-213): zero keeps the
// particle's own value.  A retired particle acts through its dead record's
// mass (forces.js:145-147) and through the spin of its final history sample
// (_retireParticle, integrator.js:160-166), whose angw gives back the death
// angular velocity under forces.js:141-143's angVel = angw / sqrt(1 + angw^2 r^2).
int geon_remove_particle(GeonHandle h, int32_t slot, double death_mass, double death_ang_vel) {
    if (!h) return GEON_EINVAL;
    if (!(death_mass >= 0) || !std::isfinite(death_mass) || !std::isfinite(death_ang_vel))
        return set_err(h, GEON_EINVAL, "death mass must be finite and >= 0, death angular velocity finite");
    int idx;
    int rc = find_slot(h, slot, &idx);
    if (rc) return rc;
    float radius = 0.f;
    HIP_OK(hipMemcpyAsync(&radius, F(h, C_RADIUS) + idx, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    const double r2 = (double)radius * radius, s = death_ang_vel * death_ang_vel * r2;
    if (death_ang_vel != 0 && !(s < 1))
        return set_err(h, GEON_EINVAL, "death angular velocity puts the surface at or above c");
    std::vector<int32_t> alive((size_t)h->n_host, 1);
    alive[idx] = 0;
    HIP_OK(hipMemcpyAsync(h->alive, alive.data(), 4 * alive.size(), hipMemcpyHostToDevice, h->stream));
    const int dead0 = h->n_dead;
    const int64_t ev = h->n_events;
    if ((rc = compact(h, h->sim_time))) return rc;
    if (h->n_dead > dead0) {             // signal delay on: the particle left a dead record
        DeadRec& r = h->dead_host[(size_t)dead0];
        if (death_mass > 0) r.mass = (float)death_mass;
        HIP_OK(hipMemcpyAsync(h->dead + dead0, &r, sizeof(DeadRec), hipMemcpyHostToDevice, h->stream));
        if (death_ang_vel != 0) {
            const float aw = (float)(death_ang_vel / sqrt(1.0 - s));
            HIP_OK(hipMemcpyAsync(h->hist_aw + (size_t)slot * kHistN + (size_t)(ev & (kHistN - 1)), &aw, 4,
                                  hipMemcpyHostToDevice, h->stream));
        }
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

int geon_patch_mass(GeonHandle h, int32_t slot, double mass, double base_mass) {
    if (!h) return GEON_EINVAL;
    if (!(mass > 0)) return set_err(h, GEON_EINVAL, "mass must be > 0");
    int idx;
    int rc = find_slot(h, slot, &idx);
    if (rc) return rc;
    const float m = (float)mass, bm = (float)base_mass, r = (float)cbrt(mass);
    HIP_OK(hipMemcpyAsync(F(h, C_MASS) + idx, &m, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(F(h, C_BASEMASS) + idx, &bm, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(F(h, C_RADIUS) + idx, &r, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

int geon_download_soa(GeonHandle h, GeonSoAView* v, int32_t cap, int32_t* n_slots) {
    if (!h || !v) return GEON_EINVAL;
    const int n = h->n_host, ns = h->next_slot;
    if (n_slots) *n_slots = ns;
    if (cap < ns) return set_err(h, GEON_EINVAL, "download arrays hold fewer entries than the slots in use");
    std::vector<std::vector<uint32_t>> col(N_STATE, std::vector<uint32_t>((size_t)n));
    for (int c = 0; c < N_STATE; c++)
        if (n) HIP_OK(hipMemcpyAsync(col[c].data(), h->st[c], 4 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    auto getf = [&](int c, int i) { float f; memcpy(&f, &col[c][i], 4); return (double)f; };
    if (v->alive) for (int s = 0; s < ns; s++) v->alive[s] = 0;
    for (int i = 0; i < n; i++) {
        const int s = (int)col[C_SLOT][i];
        if (s < 0 || s >= ns) continue;
        if (v->x) v->x[s] = getf(C_X, i);
        if (v->y) v->y[s] = getf(C_Y, i);
        if (v->wx) v->wx[s] = getf(C_WX, i);
        if (v->wy) v->wy[s] = getf(C_WY, i);
        if (v->mass) v->mass[s] = getf(C_MASS, i);
        if (v->base_mass) v->base_mass[s] = getf(C_BASEMASS, i);
        if (v->charge) v->charge[s] = (int32_t)col[C_QQ][i] * 0.1;
        if (v->angw) v->angw[s] = getf(C_ANGW, i);
        if (v->antimatter) v->antimatter[s] = (int32_t)(col[C_FLAGS][i] & 1u);
        if (v->alive) v->alive[s] = 1;
    }
    if (n_slots) *n_slots = ns;
    return GEON_OK;
}

int geon_compute_forces(GeonHandle h) {
    if (!h) return GEON_EINVAL;
    int rc = force_pass(h);
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(h->stream));
    h->forces_init = true;
    if ((rc = read_ctrl(h))) return rc;
    return check_overflow(h);
}

int geon_update(GeonHandle h, double dt, int32_t* substeps_taken) {
    if (!h) return GEON_EINVAL;
    if (!(dt >= 0)) return set_err(h, GEON_EINVAL, "dt must be >= 0");
    int rc;
    if (!h->forces_init && h->n_host > 0) {          // integrator.js:750-782
        if ((rc = force_pass(h))) return rc;
        h->forces_init = true;
    }
    const int n = h->n_host;
    h->events_host.clear();
    const bool rad = !!(h->prm.toggles0 & GEON_RADIATION);
    if (rad && (rc = rad_alloc(h))) return rc;
    hipLaunchKernelGGL(k_begin_update, dim3(1), dim3(1), 0, h->stream, h->ctrl, dt, h->prm.max_substeps, h->sim_time);
    StepArgs sa = step_args(h);
    const double soft = (h->prm.toggles0 & GEON_BLACK_HOLE) ? h->prm.bh_softening : h->prm.softening;
    int steps = 0;
    const bool active = dt > 1e-12;     // the substep loop runs while dtRemain > EPSILON (integrator.js:810)
    if (active && n == 0) {
        // No particles: dtSafe = dtRemain gives one substep of the whole dt, so
        // simTime advances and the fields still step (integrator.js:830-836,
        // 1294, 1320-1334).
        launch_ctrl(h->stream, h->ctrl, soft, 1);
        if (fields_on(h)) fields_step(h, 0.f);
        steps = 1;
    } else if (active) {
        launch_dt_scan(h->stream, sa);
        for (int k = 0; k < h->prm.max_substeps; k++) {
            launch_ctrl(h->stream, h->ctrl, soft, k == 0);
            HIP_OK(hipMemcpyAsync(h->ctrl_host, h->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
            HIP_OK(hipEventRecord(h->ctrl_ev, h->stream));
            {
                Timer t(h, 3);
                sa = step_args(h);
                launch_kick_drift(h->stream, sa);
                if (sa.rad)         // the substep's radiated energy and momentum (integrator.js:1044-1060)
                    launch_sum_cols(h->stream, h->rad_scr, 3, h->cap, h->n_host, h->ctrl, h->rad_partial, h->rad_tot,
                                    true, true);
            }
            const bool one_pn = !!(h->prm.toggles0 & GEON_ONE_PN);
            if (one_pn && (rc = one_pn_step(h))) return rc;
            if (fields_on(h)) fields_step(h, 0.f);
            bool reuse = one_pn && (h->prm.toggles0 & GEON_BARNES_HUT);
            if (h->prm.collision_mode == GEON_COL_MERGE) {
                // After a change the force pass walks a fresh tree; the reference
                // walks its step-5 tree, where removed bodies still sit (mass 0)
                // and spawned ones are missing, for this one substep.
                // Without a change the merge pass's tree, built at these
                // positions, is the one the force pass would build.
                bool changed = false;
                if ((rc = merge_step(h, &changed))) return rc;
                reuse = !changed;
            }
            if ((rc = force_pass(h, reuse))) return rc;
            sa = step_args(h);
            launch_dt_scan(h->stream, sa);
            HIP_OK(hipEventSynchronize(h->ctrl_ev));
            const Ctrl& c = *h->ctrl_host;
            if (!c.active) break;          // nothing left (dt <= EPSILON)
            steps = c.total_steps;
            if (!(c.dt_remain > 1e-12) || c.total_steps >= c.max_substeps) break;
        }
    }
    HIP_OK(hipGetLastError());
    if (delay_on(h->prm) && h->n_host > 0) {
        // signal-delay history every history_stride updates (integrator.js:1549-1565)
        const int stride = h->prm.history_stride > 0 ? h->prm.history_stride : 64;
        if (++h->hist_stride_ctr >= stride) {
            h->hist_stride_ctr -= stride;
            if ((rc = hist_alloc(h)) || (rc = ev_reserve(h))) return rc;
            launch_hist_record(h->stream, hist_write(h), F(h, C_X), F(h, C_Y), F(h, C_VX), F(h, C_VY), F(h, C_ANGW),
                               I(h, C_SLOT), h->n_events, h->n_host, h->ctrl);
            h->n_events++;
        }
    }
    const uint32_t tg = h->prm.toggles0;
    if (rad && h->n_host >= 2 && (tg & (GEON_GRAVITY | GEON_COULOMB)))     // integrator.js:1567-1740
        launch_quad_radiation(h->stream, quad_args(h), dt, h->n_host);
    // boundary once per update (integrator.js:1798-1824)
    sa = step_args(h);
    launch_boundary(h->stream, sa, h->prm.boundary_mode, h->prm.topology, h->prm.domain_w, h->prm.domain_h);
    if ((rc = read_ctrl(h))) return rc;
    h->sim_time = h->ctrl_host->sim_time;
    h->substeps_total += steps;
    for (int k = 0; k < 3; k++) h->last_counters[k] = 0;
    h->last_counters[0] = h->ctrl_host->cnt_pair;
    h->last_counters[1] = h->ctrl_host->cnt_node;
    h->last_counters[2] = h->ctrl_host->cnt_visit;
    if ((rc = check_overflow(h))) return rc;
    if (h->ctrl_host->deaths > 0) {
        if ((rc = compact(h, h->sim_time))) return rc;
        HIP_OK(hipStreamSynchronize(h->stream));
    }
    if ((rc = dead_gc(h))) return rc;
    if (substeps_taken) *substeps_taken = steps;
    return GEON_OK;
}

int geon_get_forces(GeonHandle h, GeonForceView* v, int32_t n_slots) {
    if (!h || !v) return GEON_EINVAL;
    const int n = h->n_host;
    std::vector<int32_t> slot((size_t)n);
    std::vector<std::vector<float>> fc(N_FORCE, std::vector<float>((size_t)n));
    if (n) {
        HIP_OK(hipMemcpyAsync(slot.data(), h->st[C_SLOT], 4 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
        for (int c = 0; c < N_FORCE; c++)
            HIP_OK(hipMemcpyAsync(fc[c].data(), h->fc[c], 4 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    double* outs[N_FORCE] = {v->fx, v->fy, v->pe, v->bz, v->dbzdx, v->dbzdy, v->bgz, v->dbgzdx, v->dbgzdy,
                             v->frame_drag_torque, v->tidal_torque, v->f1pn_x, v->f1pn_y, nullptr, nullptr,
                             v->jerk_x, v->jerk_y, nullptr};
    for (int i = 0; i < n; i++) {
        const int s = slot[i];
        if (s < 0 || s >= n_slots) continue;
        for (int c = 0; c < N_FORCE; c++) if (outs[c]) outs[c][s] = fc[c][i];
    }
    return GEON_OK;
}

namespace {
// One particle's state and force columns at the selected index, after the record.
struct SelRow {
    const uint32_t* c[N_STATE + N_FORCE];
};
__global__ void k_sel_row(SelRow r, const int32_t* __restrict__ idx, uint32_t* __restrict__ out) {
    const int t = threadIdx.x;
    const int i = *idx;
    if (t < N_STATE + N_FORCE) out[t] = i >= 0 ? r.c[t][i] : 0u;
    if (t == 0) out[N_STATE + N_FORCE] = (uint32_t)i;
}
constexpr int kSelRecWords = kSelFloats + N_STATE + N_FORCE + 1;
}  // namespace

// requestStats(selectedGpuIdx) (gpu-physics.js:4146-4151): the particle whose
// per-type forces the following force passes record (-1: none).
int geon_select(GeonHandle h, int32_t slot) {
    if (!h) return GEON_EINVAL;
    if (slot < -1 || slot >= h->next_slot) return set_err(h, GEON_EINVAL, "slot out of range");
    if (slot == h->sel_slot) return GEON_OK;
    h->sel_slot = -1;
    if (slot < 0) return GEON_OK;
    int rc;
    if (!h->sel_idx && (rc = dmalloc(h, &h->sel_idx, 1))) return rc;
    if (!h->sel_rec && (rc = dmalloc(h, &h->sel_rec, (size_t)kSelRecWords))) return rc;
    HIP_OK(hipMemsetAsync(h->sel_rec, 0, sizeof(float) * kSelRecWords, h->stream));
    h->sel_slot = slot;         // the kernels see the selection only once its buffers exist
    sel_locate(h);
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

// readStats().selected (gpu-physics.js:4238-4276).  The force types follow the
// reference's per-particle vectors (forces.js:36-46) and its display pass after
// the substeps (integrator.js:1755-1796): the Lorentz and gravitomagnetic
// forces of the particle's velocity in the last pass's Bz and Bgz join the
// magnetic and gravitomagnetic dipole forces, and the Stern-Gerlach /
// Mathisson-Papapetrou forces form forceSpinCurv.
int geon_get_selected(GeonHandle h, GeonSelected* out) {
    if (!h || !out) return GEON_EINVAL;
    memset(out, 0, sizeof(*out));
    out->slot = -1;
    if (h->sel_slot < 0 || h->n_host == 0) return GEON_OK;
    SelRow r;
    for (int c = 0; c < N_STATE; c++) r.c[c] = h->st[c];
    for (int c = 0; c < N_FORCE; c++) r.c[N_STATE + c] = reinterpret_cast<const uint32_t*>(h->fc[c]);
    uint32_t* row = reinterpret_cast<uint32_t*>(h->sel_rec) + kSelFloats;
    hipLaunchKernelGGL(k_sel_row, dim3(1), dim3(64), 0, h->stream, r, h->sel_idx, row);
    std::vector<uint32_t> w((size_t)kSelRecWords);
    HIP_OK(hipMemcpyAsync(w.data(), h->sel_rec, sizeof(uint32_t) * kSelRecWords, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if ((int32_t)w[kSelRecWords - 1] < 0) return GEON_OK;         // the particle is gone
    auto rec = [&](int k) { float f; memcpy(&f, &w[k], 4); return (double)f; };
    auto st = [&](int c) { return rec(kSelFloats + c); };
    auto fc = [&](int c) { return rec(kSelFloats + N_STATE + c); };
    out->slot = h->sel_slot;
    out->forces_valid = rec(SEL_VALID) != 0.0;
    out->antimatter = (int32_t)(w[kSelFloats + C_FLAGS] & 1u);
    out->x = st(C_X); out->y = st(C_Y); out->wx = st(C_WX); out->wy = st(C_WY);
    out->mass = st(C_MASS); out->charge = st(C_CHARGE); out->angw = st(C_ANGW); out->base_mass = st(C_BASEMASS);
    out->radius = st(C_RADIUS); out->vx = st(C_VX); out->vy = st(C_VY); out->ang_vel = st(C_ANGVEL);
    out->mag_moment = st(C_MU); out->ang_momentum = st(C_L);
    if (!out->forces_valid) return GEON_OK;
    const uint32_t t = h->prm.toggles0;
    const bool mag = !!(t & GEON_MAGNETIC), gm = !!(t & GEON_GRAVITOMAG);
    double (*f)[2] = out->force;
    auto split = [&](int type, int gf) { f[gf][0] += rec(SEL_SPLIT + 2 * type); f[gf][1] += rec(SEL_SPLIT + 2 * type + 1); };
    auto add = [&](int k, int gf) { f[gf][0] += rec(k); f[gf][1] += rec(k + 1); };
    split(S_GRAV, GEON_SF_GRAVITY);
    add(SEL_FIELD_GRAV, GEON_SF_GRAVITY);                      // scalar-field.js:717-718
    split(S_COUL, GEON_SF_COULOMB);
    split(S_MAG, GEON_SF_MAGNETIC);
    split(S_GM, GEON_SF_GRAVITOMAG);
    split(S_YUK, GEON_SF_YUKAWA);
    if (t & GEON_ONE_PN) { f[GEON_SF_ONE_PN][0] = fc(K_F1X); f[GEON_SF_ONE_PN][1] = fc(K_F1Y); }   // forces.js:581
    add(SEL_EXT, GEON_SF_EXTERNAL);
    add(SEL_HIGGS, GEON_SF_HIGGS);
    add(SEL_AXION, GEON_SF_AXION);
    if ((t & GEON_RADIATION) && (t & GEON_COULOMB)) add(SEL_RAD, GEON_SF_RADIATION);
    add(SEL_FAR, GEON_SF_FAR_FIELD);
    const double vx = out->vx, vy = out->vy, q = out->charge, m = out->mass;
    if (mag || h->prm.ext_bz != 0) {                            // integrator.js:1760-1763
        const double bz = fc(K_BZ);
        f[GEON_SF_MAGNETIC][0] += q * vy * bz;
        f[GEON_SF_MAGNETIC][1] -= q * vx * bz;
    }
    if (gm) {                                                   // 1764-1767
        const double bgz = fc(K_BGZ);
        f[GEON_SF_GRAVITOMAG][0] += 4.0 * m * vy * bgz;
        f[GEON_SF_GRAVITOMAG][1] -= 4.0 * m * vx * bgz;
    }
    if ((t & GEON_SPIN_ORBIT) && fabs(out->ang_vel) >= 1e-12) {  // 1768-1786 (EPSILON, config.js:30)
        if (mag && fabs(q) >= 1e-12) {
            f[GEON_SF_SPIN_CURV][0] += out->mag_moment * fc(K_DBZDX);
            f[GEON_SF_SPIN_CURV][1] += out->mag_moment * fc(K_DBZDY);
        }
        if (gm) {
            f[GEON_SF_SPIN_CURV][0] -= out->ang_momentum * fc(K_DBGZDX);
            f[GEON_SF_SPIN_CURV][1] -= out->ang_momentum * fc(K_DBGZDY);
        }
    }
    return GEON_OK;
}

int geon_get_stats(GeonHandle h, GeonStats* out) {
    if (!h || !out) return GEON_EINVAL;
    memset(out, 0, sizeof(*out));
    StepArgs sa = step_args(h);
    double r[16] = {};
    if (h->n_host > 0) {
        launch_stats(h->stream, sa, h->fc[K_PE], h->stat_partial, h->stat_out, !!(h->prm.toggles0 & GEON_RELATIVITY));
        HIP_OK(hipMemcpyAsync(r, h->stat_out, sizeof(double) * 10, hipMemcpyDeviceToHost, h->stream));
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    out->linear_ke = r[0]; out->spin_ke = r[1]; out->px = r[2]; out->py = r[3];
    out->total_mass = r[4];
    out->com_x = r[4] > 0 ? r[5] / r[4] : 0; out->com_y = r[4] > 0 ? r[6] / r[4] : 0;
    out->pe = 0.5 * r[7];                       // getPEAccum halves the double-counted sum (forces.js:26)
    out->orbital_ang_mom = r[8]; out->spin_ang_mom = r[9];
    const uint32_t t = h->prm.toggles0;
    if ((t & (GEON_HIGGS | GEON_AXION)) && h->n_host > 0) {
        double pf[2] = {0, 0};
        launch_sum2(h->stream, h->fc[K_PEG], h->fc[K_PFI], h->n_host, h->ctrl, h->stat_partial, h->stat_out + 12);
        HIP_OK(hipMemcpyAsync(pf, h->stat_out + 12, 16, hipMemcpyDeviceToHost, h->stream));
        double fe[2][4] = {};
        for (int w = 0; w < 2; w++) {
            if (!(t & (w == 0 ? GEON_HIGGS : GEON_AXION))) continue;
            const float* partner = (w == 0 && (t & GEON_AXION)) ? h->fld[1].field : nullptr;
            field_energy(h->stream, field_args(h, w), partner, h->field_out);
            HIP_OK(hipMemcpyAsync(fe[w], h->field_out, 32, hipMemcpyDeviceToHost, h->stream));
            HIP_OK(hipStreamSynchronize(h->stream));
        }
        HIP_OK(hipStreamSynchronize(h->stream));
        if (t & GEON_GRAVITY) out->pe += pf[0];  // gravPE of the fields (integrator.js:1744-1749)
        out->pfi_energy = pf[1];
        out->pe += pf[1];                        // energy.js:186
        out->higgs_field_energy = fe[0][0] + fe[0][3];
        out->axion_field_energy = fe[1][0];
        out->field_px += fe[0][1] + fe[1][1];
        out->field_py += fe[0][2] + fe[1][2];
    }
    if ((t & (GEON_MAGNETIC | GEON_GRAVITOMAG)) && h->n_host > 0) {
        // Darwin field energy and momentum, pass 3 of computeEnergies (energy.js:64-142)
        DarwinArgs d;
        d.x = F(h, C_X); d.y = F(h, C_Y); d.vx = F(h, C_VX); d.vy = F(h, C_VY);
        d.mass = F(h, C_MASS); d.charge = F(h, C_CHARGE);
        d.ax = (t & GEON_AXION) ? F(h, C_AX) : nullptr;
        d.slot = I(h, C_SLOT);
        d.hist = force_args(h).hist;
        d.pf = phys_flags(h->prm);
        const double soft = (t & GEON_BLACK_HOLE) ? h->prm.bh_softening : h->prm.softening;
        d.soft2 = soft * soft;
        const bool one_pn = !!(t & GEON_ONE_PN);
        d.mag = !!(t & GEON_MAGNETIC); d.gm = !!(t & GEON_GRAVITOMAG);
        d.em_energy = d.mag && !one_pn; d.gm_energy = d.gm && !one_pn;
        d.bazanski = d.mag && d.gm && !one_pn;
        d.ctrl = h->ctrl;
        launch_darwin(h->stream, d, h->n_host, h->stat_partial, h->stat_out + 16);
        double dw[3];
        HIP_OK(hipMemcpyAsync(dw, h->stat_out + 16, sizeof(dw), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        out->field_energy = dw[0];
        out->field_px += dw[1];
        out->field_py += dw[2];
    }
    if (h->rad_tot) {
        double rt[4];
        HIP_OK(hipMemcpyAsync(rt, h->rad_tot, sizeof(rt), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        out->radiated_energy = rt[0]; out->radiated_px = rt[1]; out->radiated_py = rt[2];
    }
    out->sim_time = h->sim_time;
    out->alive_count = h->n_host;
    out->dead_count = h->n_dead;
    return GEON_OK;
}

int geon_particle_count(GeonHandle h, int32_t* alive) {
    if (!h || !alive) return GEON_EINVAL;
    HIP_OK(hipStreamSynchronize(h->stream));
    *alive = h->n_host;
    return GEON_OK;
}

int geon_get_history(GeonHandle h, int32_t slot, GeonHistoryView* out, int32_t* count) {
    if (!h || !out || !count) return GEON_EINVAL;
    if (slot < 0 || slot >= h->next_slot) return set_err(h, GEON_EINVAL, "slot out of range");
    int2 fe;
    double2 cd;
    HIP_OK(hipMemcpyAsync(&fe, h->hist_fe + slot, sizeof(fe), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(&cd, h->hist_cd + slot, sizeof(cd), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    *count = 0;
    if (fe.x < 0 || !h->hist4) return GEON_OK;
    const bool dead = fe.y >= 0;
    const int64_t end = dead ? fe.y : h->n_events;
    const int c = (int)std::min<int64_t>(end - fe.x, kHistN);
    if (c <= 0) return GEON_OK;
    std::vector<float4> h4(kHistN);
    std::vector<float> aw(kHistN);
    std::vector<double> et((size_t)h->n_events);
    HIP_OK(hipMemcpyAsync(h4.data(), h->hist4 + (size_t)slot * kHistN, sizeof(float4) * kHistN, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(aw.data(), h->hist_aw + (size_t)slot * kHistN, sizeof(float) * kHistN, hipMemcpyDeviceToHost, h->stream));
    if (h->n_events) HIP_OK(hipMemcpyAsync(et.data(), h->ev_time, sizeof(double) * h->n_events, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < c; k++) {
        const int64_t e = end - c + k;
        const int j = (int)(e & (kHistN - 1));
        if (out->x) out->x[k] = h4[j].x;
        if (out->y) out->y[k] = h4[j].y;
        if (out->vx) out->vx[k] = h4[j].z;
        if (out->vy) out->vy[k] = h4[j].w;
        if (out->angw) out->angw[k] = aw[j];
        if (out->t) out->t[k] = (dead && e == end - 1) ? cd.y : et[(size_t)e];
    }
    *count = c;
    return GEON_OK;
}

int geon_build_tree(GeonHandle h) {
    if (!h) return GEON_EINVAL;
    if (h->n_host == 0) return GEON_OK;
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, h->stream, h->ctrl);
    int rc;
    if (ghosts_on(h->prm)) {                    // the tree over particles and their images
        if ((rc = ghost_tree(h))) return rc;
        ghost_finish(h->stream, h->gh.ctrl, h->ctrl);
        h->tree_valid = true;
        h->tree_ghost = true;
        return read_ctrl(h);
    }
    if ((rc = sort_and_gather(h))) return rc;
    build_tree(h->stream, h->tb, F(h, C_X), F(h, C_Y), F(h, C_MASS), F(h, C_WX), F(h, C_WY), F(h, C_MU), F(h, C_L),
               I(h, C_QQ), h->n_host, key_params(h->prm), h->ctrl, want_moments(h));
    h->tree_valid = true;
    h->tree_ghost = false;
    return read_ctrl(h);
}

int geon_get_tree(GeonHandle h, GeonTreeNode* out, int64_t cap, int64_t* n_nodes) {
    if (!h) return GEON_EINVAL;
    if (!h->tree_valid) return set_err(h, GEON_ESTATE, "no tree built");
    int rc = read_ctrl(h);
    if (rc) return rc;
    const int64_t nn = h->n_host == 0 ? 0 : 1 + 4 * (int64_t)h->ctrl_host->n_divided;
    if (n_nodes) *n_nodes = nn;
    if (!out) return GEON_OK;
    const int64_t m = std::min(cap, nn);
    if (m <= 0) return GEON_OK;
    const TreeBufs& tb = h->tree_ghost ? h->gh.tb : h->tb;
    std::vector<Node> nodes((size_t)m);
    std::vector<double> mass((size_t)m), cx((size_t)m), cy((size_t)m), qq((size_t)m);
    std::vector<uint64_t> path((size_t)m);
    std::vector<int32_t> npts((size_t)m);
    HIP_OK(hipMemcpyAsync(nodes.data(), tb.nodes, sizeof(Node) * m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(mass.data(), tb.node_mass64, 8 * m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(cx.data(), tb.node_comx64, 8 * m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(cy.data(), tb.node_comy64, 8 * m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(qq.data(), tb.node_qq64, 8 * m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(path.data(), tb.node_path, 8 * m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(npts.data(), tb.node_npts, 4 * m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    for (int64_t i = 0; i < m; i++) {
        out[i].path = path[i];
        out[i].level = nodes[i].level;
        out[i].divided = nodes[i].count < 0 ? 1 : 0;
        out[i].count = npts[i];
        out[i].charge_quanta = (int32_t)llround(qq[i]);
        out[i].mass = mass[i];
        out[i].com_x = cx[i];
        out[i].com_y = cy[i];
    }
    return GEON_OK;
}

int geon_set_field(GeonHandle h, int32_t which, const double* field, const double* field_dot, int32_t n) {
    if (!h || which < 0 || which > 1) return GEON_EINVAL;
    const int g2 = h->field_g * h->field_g;
    if (n != g2) return set_err(h, GEON_EINVAL, "n must be field_grid^2");
    FieldState& s = h->fld[which];
    std::vector<float> buf((size_t)g2);
    if (field) {
        for (int i = 0; i < g2; i++) buf[i] = (float)field[i];
        HIP_OK(hipMemcpyAsync(s.field, buf.data(), 4 * (size_t)g2, hipMemcpyHostToDevice, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
    }
    if (field_dot) {
        for (int i = 0; i < g2; i++) buf[i] = (float)field_dot[i];
        HIP_OK(hipMemcpyAsync(s.field_dot, buf.data(), 4 * (size_t)g2, hipMemcpyHostToDevice, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
    }
    field_grads(h->stream, field_args(h, which));
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

int geon_get_field(GeonHandle h, int32_t which, int32_t what, double* out, int32_t n) {
    if (!h || !out || which < 0 || which > 1) return GEON_EINVAL;
    const int g2 = h->field_g * h->field_g;
    if (n != g2) return set_err(h, GEON_EINVAL, "n must be field_grid^2");
    FieldState& s = h->fld[which];
    const float* src[] = {s.field, s.field_dot, s.grad_x, s.grad_y, s.sg_phi, s.sg_gx, s.sg_gy, s.source, s.thermal};
    if (what < 0 || what > 8) return set_err(h, GEON_EINVAL, "bad field quantity");
    if (!src[what]) return set_err(h, GEON_ESTATE, "that field quantity is not held (field off)");
    std::vector<float> buf((size_t)g2);
    HIP_OK(hipMemcpyAsync(buf.data(), src[what], 4 * (size_t)g2, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < g2; i++) out[i] = buf[i];
    return GEON_OK;
}

int geon_field_selfgravity(GeonHandle h, int32_t which) {
    if (!h || which < 0 || which > 1) return GEON_EINVAL;
    ensure_green(h, which);
    field_selfgrav(h->stream, field_args(h, which));
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

int geon_fft2d(GeonHandle h, double* re, double* im, int32_t n, int32_t inverse) {
    if (!h || !re || !im) return GEON_EINVAL;
    if (n < 2 || n > 128 || (n & (n - 1)) || n > h->field_g) return set_err(h, GEON_EINVAL, "n must be a power of two <= field_grid");
    // scratch: the axion field's G^ buffer is rebuilt afterwards
    FieldState& s = h->fld[1];
    std::vector<float2> buf((size_t)n * n);
    for (int i = 0; i < n * n; i++) buf[i] = make_float2((float)re[i], (float)im[i]);
    HIP_OK(hipMemcpyAsync(s.green, buf.data(), sizeof(float2) * n * n, hipMemcpyHostToDevice, h->stream));
    FieldArgs f = field_args(h, 1);
    f.G = n;
    field_fft(h->stream, f, inverse);
    HIP_OK(hipMemcpyAsync(buf.data(), s.green, sizeof(float2) * n * n, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < n * n; i++) { re[i] = buf[i].x; im[i] = buf[i].y; }
    s.green_w = -1;     // force a rebuild of the axion G^
    return GEON_OK;
}

int geon_collision_pairs(GeonHandle h, int32_t* a, int32_t* b, int64_t cap, int64_t* n_pairs) {
    if (!h || (cap > 0 && (!a || !b))) return GEON_EINVAL;
    if (n_pairs) *n_pairs = 0;
    PairArgs pa;
    int64_t np = 0;
    int rc = pairs_count(h, &pa, &np);
    if (rc) return rc;
    if (n_pairs) *n_pairs = np;
    if (np == 0 || cap <= 0) return GEON_OK;
    int32_t *da = nullptr, *db = nullptr;
    if ((rc = dmalloc(h, &da, (size_t)np))) return rc;
    if ((rc = dmalloc(h, &db, (size_t)np))) { hipFree(da); return rc; }
    pa.pa = da; pa.pb = db;
    launch_pairs(h->stream, pa, true);
    std::vector<int64_t> keys((size_t)np);
    std::vector<int32_t> ha((size_t)np), hb((size_t)np);
    hipError_t e1 = hipMemcpyAsync(ha.data(), da, 4 * (size_t)np, hipMemcpyDeviceToHost, h->stream);
    hipError_t e2 = hipMemcpyAsync(hb.data(), db, 4 * (size_t)np, hipMemcpyDeviceToHost, h->stream);
    hipError_t e3 = hipStreamSynchronize(h->stream);
    hipFree(da); hipFree(db);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return set_err(h, GEON_EDEVICE, "pair list copy failed");
    for (int64_t k = 0; k < np; k++) keys[k] = ((int64_t)ha[k] << 32) | (uint32_t)hb[k];
    std::sort(keys.begin(), keys.end());
    const int64_t m = std::min(cap, np);
    for (int64_t k = 0; k < m; k++) { a[k] = (int32_t)(keys[k] >> 32); b[k] = (int32_t)(keys[k] & 0xffffffff); }
    return GEON_OK;
}

int geon_collision_pairs_noalloc(GeonHandle h, int32_t* a, int32_t* b, int64_t cap, int64_t* n_pairs) {
    if (!h || (cap > 0 && (!a || !b))) return GEON_EINVAL;
    if (!h->na_scratch) return set_err(h, GEON_ESTATE, "device scratch not set up");
    PairArgs pa;
    int64_t np = 0;
    int rc = pairs_count(h, &pa, &np);
    if (rc) return rc;
    if (np > kNoAllocPairs)
        return set_err(h, GEON_EOVERFLOW, "allocation-free pair list: " + std::to_string(np) +
                                              " pairs exceed its fixed scratch of " + std::to_string(kNoAllocPairs));
    if (np == 0 || cap <= 0) {
        if (n_pairs) *n_pairs = np;
        return GEON_OK;
    }
    unsigned char* const sc = h->na_scratch;
    Ctrl* const nc = reinterpret_cast<Ctrl*>(sc + kNaCtrl);
    pa.pa = reinterpret_cast<int32_t*>(sc + kNaPa);
    pa.pb = reinterpret_cast<int32_t*>(sc + kNaPb);
    launch_pairs(h->stream, pa, true);
    hipLaunchKernelGGL(k_set_n, dim3(1), dim3(1), 0, h->stream, nc, (int)np);
    TreeBufs sb = {};
    sb.keys = reinterpret_cast<uint64_t*>(sc + kNaKeys);
    sb.keys_alt = reinterpret_cast<uint64_t*>(sc + kNaKeysAlt);
    sb.perm = reinterpret_cast<uint32_t*>(sc + kNaPerm);
    sb.perm_alt = reinterpret_cast<uint32_t*>(sc + kNaPermAlt);
    sb.hist = reinterpret_cast<uint32_t*>(sc + kNaHist);
    sb.digit_total = reinterpret_cast<uint32_t*>(sc + kNaDigit);
    int bits = 1;                                   // slot numbers are < next_slot
    while (bits < 31 && (int64_t(1) << bits) < (int64_t)h->next_slot) bits++;
    const dim3 grid((unsigned)((np + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_pack_pairs, grid, block, 0, h->stream, pa.pa, pa.pb, sb.keys, sb.perm, bits, nc);
    radix_sort(h->stream, sb, (int)np, 2 * bits, nc);
    hipLaunchKernelGGL(k_split_pairs, grid, block, 0, h->stream, sb.keys, pa.pa, pa.pb, bits, nc);
    const int64_t m = std::min(cap, np);
    HIP_OK(hipMemcpyAsync(a, pa.pa, 4 * (size_t)m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(b, pa.pb, 4 * (size_t)m, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (n_pairs) *n_pairs = np;
    return GEON_OK;
}

int geon_collision_events(GeonHandle h, double* out, int64_t cap, int64_t* n_events) {
    if (!h || (cap > 0 && !out)) return GEON_EINVAL;
    const int64_t n = (int64_t)(h->events_host.size() / GEON_EVENT_STRIDE);
    if (n_events) *n_events = n;
    const int64_t m = std::min(cap, n);
    if (m > 0) memcpy(out, h->events_host.data(), sizeof(double) * GEON_EVENT_STRIDE * (size_t)m);
    return GEON_OK;
}

int geon_counters(GeonHandle h, GeonCounters* out) {
    if (!h || !out) return GEON_EINVAL;
    int rc = read_ctrl(h);
    if (rc) return rc;
    memset(out, 0, sizeof(*out));
    out->pair_interactions = h->ctrl_host->cnt_pair;
    out->node_interactions = h->ctrl_host->cnt_node;
    out->nodes_visited = h->ctrl_host->cnt_visit;
    out->tree_divided = h->ctrl_host->n_divided;
    out->tree_nodes = h->tree_valid ? 1 + 4 * (int64_t)h->ctrl_host->n_divided : 0;
    out->depth_overflow = h->ctrl_host->depth_overflow;
    out->substeps_total = h->substeps_total;
    return GEON_OK;
}

int geon_timing_enable(GeonHandle h, int32_t on) {
    if (!h) return GEON_EINVAL;
    HIP_OK(hipStreamSynchronize(h->stream));
    h->timing = on != 0;
    for (int k = 0; k < kTimers; k++) h->ev_used[k] = 0;
    return GEON_OK;
}

int geon_kernel_timing(GeonHandle h, int32_t kernel, double* total_ms, int64_t* launches) {
    if (!h || kernel < 0 || kernel >= kTimers) return GEON_EINVAL;
    HIP_OK(hipStreamSynchronize(h->stream));
    double tot = 0;
    for (size_t i = 0; i < h->ev_used[kernel]; i++) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, h->ev[kernel][i].first, h->ev[kernel][i].second));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = (int64_t)h->ev_used[kernel];
    return GEON_OK;
}

// ---- learned far-field correction --------------------------------------------
int geon_far_clear(GeonHandle h) {
    if (!h) return GEON_EINVAL;
    h->far_n_samples = 0;
    return GEON_OK;
}

int geon_far_collect(GeonHandle h, int32_t group_stride, int64_t* n_samples) {
    if (!h || group_stride < 1) return GEON_EINVAL;
    if (!(h->prm.toggles0 & GEON_BARNES_HUT)) return set_err(h, GEON_ESTATE, "far-field samples need Barnes-Hut on");
    if (delay_on(h->prm)) return set_err(h, GEON_EUNSUPPORTED, "far-field samples are taken without signal delay");
    const int n = h->n_host;
    int rc;
    if ((rc = far_model_alloc(h))) return rc;
    if (n > 0) {
        if ((rc = force_pass(h, false, true))) return rc;
        h->forces_init = false;             // forces above lack the far-field correction
        if ((rc = read_ctrl(h))) return rc;
        if ((rc = check_overflow(h))) return rc;
        const int64_t groups = (n + kWave - 1) / kWave;
        std::vector<int32_t> cnt((size_t)(groups * kWave));
        HIP_OK(hipMemcpy(cnt.data(), h->far_count, 4 * (size_t)(groups * kWave), hipMemcpyDeviceToHost));
        std::vector<int64_t> sel, off;
        int64_t rows = 0;
        for (int64_t g = 0; g < groups; g += group_stride) {
            sel.push_back(g);
            off.push_back(rows);
            for (int l = 0; l < kWave; l++) rows += cnt[(size_t)(g * kWave + l)];
        }
        const int64_t ns = (int64_t)sel.size();
        if (2 * ns > h->far_sel_cap) {
            hipFree(h->far_sel);
            h->far_sel = nullptr;
            if ((rc = dmalloc(h, &h->far_sel, (size_t)(2 * ns)))) return rc;
            h->far_sel_cap = 2 * ns;
        }
        HIP_OK(hipMemcpy(h->far_sel, sel.data(), 8 * ns, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(h->far_sel + ns, off.data(), 8 * ns, hipMemcpyHostToDevice));
        const int64_t need = h->far_n_samples + rows;
        if (need > h->far_cap_samples) {
            const int64_t cap = std::max(need, 2 * h->far_cap_samples);
            float* ns_buf = nullptr;
            int32_t* no_buf = nullptr;
            if ((rc = dmalloc(h, &ns_buf, (size_t)(cap * kFarSample)))) return rc;
            if ((rc = dmalloc(h, &no_buf, (size_t)cap))) { hipFree(ns_buf); return rc; }
            if (h->far_n_samples) {
                HIP_OK(hipMemcpy(ns_buf, h->far_samples, sizeof(float) * kFarSample * h->far_n_samples,
                                 hipMemcpyDeviceToDevice));
                HIP_OK(hipMemcpy(no_buf, h->far_obs, sizeof(int32_t) * h->far_n_samples, hipMemcpyDeviceToDevice));
            }
            hipFree(h->far_samples);
            hipFree(h->far_obs);
            h->far_samples = ns_buf;
            h->far_obs = no_buf;
            h->far_cap_samples = cap;
        }
        launch_far_f2(h->stream, h->fc[K_FX], h->fc[K_FY], h->fc[K_PE], h->ctrl, h->far_f2);
        FarSampleArgs a;
        a.x = F(h, C_X); a.y = F(h, C_Y); a.m = F(h, C_MASS); a.qq = I(h, C_QQ);
        a.nodes = h->tb.nodes; a.node_mom = h->tb.node_mom;
        a.mass64 = h->tb.node_mass64; a.comx64 = h->tb.node_comx64; a.comy64 = h->tb.node_comy64;
        a.qq64 = h->tb.node_qq64; a.first = h->tb.node_first; a.npts = h->tb.node_npts;
        a.far_list = h->far_list; a.far_count = h->far_count; a.far_cap = h->far_cap;
        a.groups = h->far_sel; a.row_off = h->far_sel + ns; a.n_groups = (int)ns; a.n_rows = rows;
        a.f2ref = h->far_f2;
        a.out = h->far_samples + (size_t)h->far_n_samples * kFarSample;
        a.slot = I(h, C_SLOT);
        a.obs = h->far_obs + h->far_n_samples;
        const PhysFlags pf = phys_flags(h->prm);
        a.soft2 = pf.soft2; a.dom_w = h->prm.domain_w; a.dom_h = h->prm.domain_h; a.periodic = pf.periodic;
        a.grav = !!(h->prm.toggles0 & GEON_GRAVITY);
        a.coul = !!(h->prm.toggles0 & GEON_COULOMB);
        launch_far_samples(h->stream, a);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(h->stream));
        h->far_n_samples = need;
    }
    if (n_samples) *n_samples = h->far_n_samples;
    return GEON_OK;
}

int geon_far_get_samples(GeonHandle h, int64_t first, int64_t n, float* out, int32_t* observer) {
    if (!h || first < 0 || n < 0) return GEON_EINVAL;
    if (first + n > h->far_n_samples) return set_err(h, GEON_EINVAL, "sample range beyond the samples held");
    if (n && out)
        HIP_OK(hipMemcpy(out, h->far_samples + (size_t)first * kFarSample, sizeof(float) * kFarSample * n,
                         hipMemcpyDeviceToHost));
    if (n && observer)
        HIP_OK(hipMemcpy(observer, h->far_obs + first, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return GEON_OK;
}

int geon_far_train(GeonHandle h, int32_t steps, int32_t batch, double lr, double lr_final, uint64_t seed,
                   float* loss_out) {
    if (!h || steps < 0 || batch < 1) return GEON_EINVAL;
    if (h->far_n_samples == 0) return set_err(h, GEON_ESTATE, "no training samples (geon_far_collect)");
    int rc;
    if ((rc = far_model_alloc(h))) return rc;
    std::vector<float> stats;
    float wscale = 1.f;
    if ((rc = far_sample_stats(h, &stats, &wscale))) return rc;
    if (!h->far_have_stats) {
        HIP_OK(hipMemcpy(h->far_stats, stats.data(), sizeof(float) * kFarStats, hipMemcpyHostToDevice));
        h->far_have_stats = true;
    }
    if (!h->far_have_model) {
        // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) per layer; W1 rows 0..10 (10 inputs + bias), rows 11..15 zero
        std::vector<float> p(kFarParams, 0.f);
        uint64_t z = seed;
        auto uni = [&](double a) { return (float)(a * (2.0 * ((splitmix(z) >> 11) * 0x1.0p-53) - 1.0)); };
        for (int k = 0; k <= kFarIn; k++)
            for (int m = 0; m < kFarHid; m++) p[kFarOffW1 + k * kFarHid + m] = uni(1.0 / sqrt((double)kFarIn));
        for (int k = 0; k < kFarHid * kFarHid; k++) p[kFarOffW2 + k] = uni(0.125);
        for (int k = 0; k < kFarHid; k++) p[kFarOffB2 + k] = uni(0.125);
        for (int k = 0; k < kFarHid * kFarOut; k++) p[kFarOffW3 + k] = uni(0.125);
        for (int k = 0; k < kFarOut; k++) p[kFarOffB3 + k] = uni(0.125);
        HIP_OK(hipMemcpy(h->far_params, p.data(), sizeof(float) * kFarParams, hipMemcpyHostToDevice));
        HIP_OK(hipMemset(h->far_m, 0, sizeof(float) * kFarParams));
        HIP_OK(hipMemset(h->far_v, 0, sizeof(float) * kFarParams));
        h->far_adam_t = 0;
        h->far_have_model = true;
    }
    const int64_t nblocks = (batch + kWave - 1) / kWave;
    if ((rc = far_blocks(h, nblocks))) return rc;
    float* dloss = nullptr;
    if (steps > 0 && (rc = dmalloc(h, &dloss, (size_t)steps))) return rc;
    FarTrainArgs a = far_train_args(h, wscale);
    a.batch = batch;
    a.seed = seed;
    for (int s = 0; s < steps; s++) {
        const double c = 0.5 * (1.0 + cos(M_PI * s / std::max(1, steps)));
        const float lr_s = (float)(lr_final + (lr - lr_final) * c);
        a.step = h->far_adam_t;
        a.loss_out = dloss + s;
        if ((rc = far_pack(h))) { hipFree(dloss); return rc; }
        launch_far_train(h->stream, a, (int)nblocks, 1, lr_s, ++h->far_adam_t);
    }
    if ((rc = far_pack(h))) { hipFree(dloss); return rc; }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && loss_out && steps > 0)
        e = hipMemcpyAsync(loss_out, dloss, sizeof(float) * steps, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(dloss);
    if (e != hipSuccess) return set_err(h, GEON_EDEVICE, std::string("far training: ") + hipGetErrorString(e));
    return GEON_OK;
}

int geon_far_grad(GeonHandle h, const int64_t* rows, int32_t n_rows, float* grad_out, float* loss_out) {
    if (!h || !rows || n_rows < 1) return GEON_EINVAL;
    if (!h->far_have_model) return set_err(h, GEON_ESTATE, "no far-field model");
    for (int32_t i = 0; i < n_rows; i++)
        if (rows[i] < 0 || rows[i] >= h->far_n_samples) return set_err(h, GEON_EINVAL, "sample row out of range");
    int rc;
    const int64_t nblocks = (n_rows + kWave - 1) / kWave;
    if ((rc = far_blocks(h, nblocks))) return rc;
    if (n_rows > h->far_rows_cap) {
        hipFree(h->far_rows);
        h->far_rows = nullptr;
        if ((rc = dmalloc(h, &h->far_rows, (size_t)n_rows))) return rc;
        h->far_rows_cap = n_rows;
    }
    HIP_OK(hipMemcpy(h->far_rows, rows, 8 * (size_t)n_rows, hipMemcpyHostToDevice));
    std::vector<float> stats;
    float wscale = 1.f;
    if ((rc = far_sample_stats(h, &stats, &wscale))) return rc;
    FarTrainArgs a = far_train_args(h, wscale);
    a.rows = h->far_rows;
    a.batch = n_rows;
    if ((rc = far_pack(h))) return rc;
    launch_far_train(h->stream, a, (int)nblocks, 0, 0.f, 1);
    HIP_OK(hipGetLastError());
    if (grad_out) HIP_OK(hipMemcpyAsync(grad_out, h->far_grad, sizeof(float) * kFarParams, hipMemcpyDeviceToHost, h->stream));
    if (loss_out) HIP_OK(hipMemcpyAsync(loss_out, h->far_loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

int geon_far_get_model(GeonHandle h, float* out, int32_t n) {
    if (!h || !out || n != GEON_FAR_MODEL_FLOATS) return GEON_EINVAL;
    if (!h->far_have_model) return set_err(h, GEON_ESTATE, "no far-field model");
    HIP_OK(hipMemcpy(out, h->far_params, sizeof(float) * kFarParams, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(out + kFarParams, h->far_stats, sizeof(float) * kFarStats, hipMemcpyDeviceToHost));
    return GEON_OK;
}

int geon_far_set_model(GeonHandle h, const float* in, int32_t n) {
    if (!h || !in || n != GEON_FAR_MODEL_FLOATS) return GEON_EINVAL;
    for (int32_t i = 0; i < n; i++)
        if (!std::isfinite(in[i])) return set_err(h, GEON_EINVAL, "model holds a non-finite value");
    int rc;
    if ((rc = far_model_alloc(h))) return rc;
    HIP_OK(hipMemcpy(h->far_params, in, sizeof(float) * kFarParams, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->far_stats, in + kFarParams, sizeof(float) * kFarStats, hipMemcpyHostToDevice));
    HIP_OK(hipMemset(h->far_m, 0, sizeof(float) * kFarParams));
    HIP_OK(hipMemset(h->far_v, 0, sizeof(float) * kFarParams));
    h->far_adam_t = 0;
    h->far_have_model = true;
    h->far_have_stats = true;
    h->forces_init = false;
    if ((rc = far_pack(h))) return rc;
    HIP_OK(hipStreamSynchronize(h->stream));
    return GEON_OK;
}

const char* geon_last_error(GeonHandle h) {
    if (!h) return g_create_error.c_str();
    return h->err.c_str();
}

}  // extern "C"
