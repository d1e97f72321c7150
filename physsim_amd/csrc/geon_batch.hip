// geon_batch.hip — the C ABI of batched ensembles: many independent systems,
// each with the semantics of one Physics.update per call (integrator.js:726-1825:
// its own adaptive substeps, despawn once per update), stepped together.
//
// Every system owns a kBatchAlign-aligned slice of one set of SoA columns.
// One substep of all systems is one pass of the usual kernels over all
// slices: a per-system controller (k_batch_ctrl), the shared kick/drift body
// under each system's dt, one sort by (system, Morton key), one tree build
// whose prefix sums restart at every slice, and a persistent walk whose waves
// take 64-particle groups from a queue (a wave never crosses a slice).  The
// substep bodies are captured into hipGraphs (two bodies each by default) and
// replayed until no system can take another substep: as many graphs as the
// last update's longest system needed go out back to back, then the host
// checks between graphs.  Systems that finished their update are predicated
// off inside the kernels (sort and tree skip entirely once none is active).
// A system's results do not depend on the other systems in the batch.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <string>
#include <vector>
#include "../../include/geon.h"
#include "geon_kernels.h"

using namespace geon;

namespace {

enum BCol {
    B_X, B_Y, B_WX, B_WY, B_VX, B_VY, B_MASS, B_CHARGE, B_ANGW, B_ANGVEL, B_RADIUS, B_MU, B_L, B_F1OX, B_F1OY,
    B_QQ, B_SLOT, B_SYS, B_VALID, NB_STATE
};
enum BForce {
    BF_FX, BF_FY, BF_PE, BF_BZ, BF_DBZDX, BF_DBZDY, BF_BGZ, BF_DBGZDX, BF_DBGZDY, BF_FDRAG, BF_TIDAL, BF_F1X, BF_F1Y,
    NB_FORCE
};
constexpr int kBatchKeyBits = 2 * kBatchKeyLevels + 1 + 9 + 1;   // Morton + outside bit + system id (< 2^10)
constexpr int kMaxSystems = 1 << 10;
constexpr int kChunk = 2;           // default substep bodies per captured graph (even: the state ping-pong returns)

thread_local std::string g_batch_create_error;

}  // namespace

struct GeonBatch_ {
    GeonParams prm;
    int S = 0, total = 0;
    std::vector<int32_t> cap, soff, n_up;   // per system: slice size, slice start, uploaded count
    hipStream_t stream = nullptr;
    uint32_t* st[NB_STATE] = {};
    uint32_t* st_alt[NB_STATE] = {};
    float* fc[NB_FORCE] = {};
    float4 *src4 = nullptr, *src8a = nullptr, *src8b = nullptr;
    TreeBufs tb = {};
    Ctrl* bctrl = nullptr;          // per system
    Ctrl* bctrl_host = nullptr;     // pinned mirror
    Ctrl* gctrl = nullptr;          // whole batch (counts, counters, overflow)
    Ctrl* gctrl_host = nullptr;
    Ctrl* actrl = nullptr;          // counts gated on any system being active (sort, tree)
    int32_t *d_soff = nullptr, *n_in = nullptr, *seg_first = nullptr, *root_slot = nullptr, *queue = nullptr;
    double* stats = nullptr;        // 10 per system
    // the substep loop as captured chunks of `bodies` substeps: the first chunk
    // (controller's first-substep rule) and the one replayed after it
    hipGraphExec_t graph = nullptr, graph_next = nullptr;
    uint32_t* graph_st0 = nullptr;  // the state buffer the graphs start from (ping-pong parity)
    int32_t* pending = nullptr;     // device flag: a system still inside its update
    int32_t* pending_host = nullptr;
    GeonBatchTune tune = {};
    int last_max_steps = 0;         // the most substeps any system took in the last update
    int cus = 0;
    bool forces_init = false;
    std::string err;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
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

float* F(GeonBatch_* h, int c) { return reinterpret_cast<float*>(h->st[c]); }
int32_t* I(GeonBatch_* h, int c) { return reinterpret_cast<int32_t*>(h->st[c]); }

int set_err(GeonBatch_* h, int code, const std::string& msg) {
    h->err = msg;
    return code;
}

template <typename T>
int dmalloc(GeonBatch_* h, T** p, size_t count) {
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) { h->err = std::string("hipMalloc: ") + hipGetErrorString(e); return GEON_ENOMEM; }
    return GEON_OK;
}

uint32_t validate(const GeonParams* p, std::string* why) {
    const uint32_t ok = GEON_GRAVITY | GEON_COULOMB | GEON_MAGNETIC | GEON_GRAVITOMAG | GEON_RELATIVITY |
                        GEON_YUKAWA | GEON_BARNES_HUT | GEON_SPIN_ORBIT;
    if (p->toggles0 & ~ok) { *why = "a batch runs gravity, Coulomb, magnetic, gravitomagnetic, Yukawa, relativity and spin-orbit only"; return 1; }
    if (!(p->toggles0 & GEON_BARNES_HUT)) { *why = "a batch runs the Barnes-Hut walk"; return 1; }
    if (p->far_field != GEON_FAR_NONE) { *why = "a batch runs monopole nodes only"; return 1; }
    if (p->signal_delay == GEON_DELAY_ON || (p->signal_delay == GEON_DELAY_FOLLOW && (p->toggles0 & GEON_RELATIVITY))) {
        *why = "a batch runs without signal delay (relativity needs signal_delay OFF)"; return 1;
    }
    if (p->collision_mode != GEON_COL_PASS) { *why = "only collision mode 'pass' is run"; return 1; }
    if (p->boundary_mode == GEON_BOUND_BOUNCE) { *why = "bounce boundary (Hertz wall forces) is not run"; return 1; }
    if (p->boundary_mode == GEON_BOUND_LOOP && p->topology != GEON_TORUS) { *why = "only the torus topology is run"; return 1; }
    if (p->boundary_mode == GEON_BOUND_LOOP && p->periodic_mode != GEON_PERIODIC_MIN_IMAGE) { *why = "a batch runs min-image periodic mode"; return 1; }
    if (!(p->domain_w > 0) || !(p->domain_h > 0)) { *why = "domain must be positive"; return 1; }
    if (p->max_substeps < 1 || p->max_substeps > 256) { *why = "max_substeps must be in [1, 256]"; return 1; }
    return 0;
}

PhysFlags flags(const GeonParams& p) {
    PhysFlags f;
    const uint32_t t = p.toggles0;
    f.mag = !!(t & GEON_MAGNETIC);
    f.gm = !!(t & GEON_GRAVITOMAG);
    f.one_pn = 0;
    f.yuk = !!(t & GEON_YUKAWA);
    f.higgs = 0;
    f.ax_mod = 0;
    const double soft = p.softening;
    f.soft2 = (float)(soft * soft);
    f.yuk_g = (float)p.yukawa_coupling;
    f.yuk_mu = (float)p.yukawa_mu;
    f.yuk_cut2 = f.yuk ? (float)((6.0 / p.yukawa_mu) * (6.0 / p.yukawa_mu)) : INFINITY;
    f.theta2 = (float)(p.theta * p.theta);
    f.dom_w = (float)p.domain_w;
    f.dom_h = (float)p.domain_h;
    f.periodic = p.boundary_mode == GEON_BOUND_LOOP;
    f.topology = 0;
    f.window = 0;
    return f;
}

uint32_t features(const GeonParams& p) {
    uint32_t f = 0;
    if (p.toggles0 & GEON_GRAVITY) f |= F_GRAV;
    if (p.toggles0 & GEON_COULOMB) f |= F_COUL;
    if (p.toggles0 & (GEON_MAGNETIC | GEON_GRAVITOMAG | GEON_YUKAWA)) f |= F_EXTRA;
    return f;
}

KeyParams keyp(const GeonParams& p) {
    KeyParams k;
    const double W = p.domain_w, H = p.domain_h;
    k.xmin = W / 2 - 2 * W;
    k.ymin = H / 2 - 2 * H;
    k.scale_x = (double)(1u << kBatchKeyLevels) / (4 * W);
    k.scale_y = (double)(1u << kBatchKeyLevels) / (4 * H);
    k.root_hw = 2 * W;
    return k;
}

StepArgs step_args(GeonBatch_* h) {
    StepArgs a;
    a.x = F(h, B_X); a.y = F(h, B_Y); a.wx = F(h, B_WX); a.wy = F(h, B_WY);
    a.vx = F(h, B_VX); a.vy = F(h, B_VY); a.angw = F(h, B_ANGW); a.ang_vel = F(h, B_ANGVEL);
    a.mass = F(h, B_MASS); a.radius = F(h, B_RADIUS); a.charge = F(h, B_CHARGE);
    a.mu = F(h, B_MU); a.L = F(h, B_L);
    a.f1x = h->fc[BF_F1X]; a.f1y = h->fc[BF_F1Y]; a.f1ox = F(h, B_F1OX); a.f1oy = F(h, B_F1OY);
    a.fx = h->fc[BF_FX]; a.fy = h->fc[BF_FY]; a.bz = h->fc[BF_BZ]; a.bgz = h->fc[BF_BGZ];
    a.dbzdx = h->fc[BF_DBZDX]; a.dbzdy = h->fc[BF_DBZDY]; a.dbgzdx = h->fc[BF_DBGZDX]; a.dbgzdy = h->fc[BF_DBGZDY];
    a.frame_drag = h->fc[BF_FDRAG]; a.tidal = h->fc[BF_TIDAL];
    a.alive = I(h, B_VALID);
    a.ctrl = h->gctrl;
    const uint32_t t = h->prm.toggles0;
    a.rel = !!(t & GEON_RELATIVITY);
    a.mag = !!(t & GEON_MAGNETIC);
    a.gm = !!(t & GEON_GRAVITOMAG);
    a.ext_bz_on = h->prm.ext_bz != 0;
    a.grav = !!(t & GEON_GRAVITY);
    a.spin_orbit = !!(t & GEON_SPIN_ORBIT);
    a.one_pn = 0;
    a.expansion = 0;
    a.hubble = 0.f;
    a.cx = (float)(h->prm.domain_w * 0.5);
    a.cy = (float)(h->prm.domain_h * 0.5);
    a.n_upper = h->total;
    return a;
}

BatchStepArgs bstep(GeonBatch_* h) {
    BatchStepArgs b;
    b.S = h->S;
    b.total = h->total;
    b.sys = I(h, B_SYS);
    b.valid = I(h, B_VALID);
    b.soff = h->d_soff;
    b.ctrl = h->bctrl;
    return b;
}

ForceArgs force_args(GeonBatch_* h) {
    ForceArgs a;
    a.src4 = h->src4; a.src8a = h->src8a; a.src8b = h->src8b;
    a.radius = F(h, B_RADIUS);
    a.nodes = h->tb.nodes;
    a.fx = h->fc[BF_FX]; a.fy = h->fc[BF_FY]; a.pe = h->fc[BF_PE];
    a.bz = h->fc[BF_BZ]; a.dbzdx = h->fc[BF_DBZDX]; a.dbzdy = h->fc[BF_DBZDY];
    a.bgz = h->fc[BF_BGZ]; a.dbgzdx = h->fc[BF_DBGZDX]; a.dbgzdy = h->fc[BF_DBGZDY];
    a.frame_drag = h->fc[BF_FDRAG]; a.tidal = h->fc[BF_TIDAL]; a.f1x = h->fc[BF_F1X]; a.f1y = h->fc[BF_F1Y];
    a.mass_for_cyc = F(h, B_MASS);
    a.charge = F(h, B_CHARGE);
    a.ctrl = h->gctrl;
    a.pf = flags(h->prm);
    const GeonParams& p = h->prm;
    a.ext_gx = (float)(p.ext_gravity * cos(p.ext_gravity_angle));
    a.ext_gy = (float)(p.ext_gravity * sin(p.ext_gravity_angle));
    a.ext_ex = (float)(p.ext_electric * cos(p.ext_electric_angle));
    a.ext_ey = (float)(p.ext_electric * sin(p.ext_electric_angle));
    a.ext_bz = (float)p.ext_bz;
    a.cyc_mag = a.pf.mag || p.ext_bz != 0;
    a.cyc_gm = a.pf.gm;
    a.n_upper = h->total;
    a.max_nodes = 1 + 4 * h->tb.cap_divided + h->S;
    a.far_mode = 0;
    a.node_mom = nullptr; a.far_list = nullptr; a.far_count = nullptr; a.far_cap = 0;
    a.bsys = I(h, B_SYS);
    a.bvalid = I(h, B_VALID);
    a.bctrl = h->bctrl;
    a.broot = h->root_slot;
    a.hist = HistArgs{};
    return a;
}

__global__ void k_batch_reset(Ctrl* g, int* queue) {
    g->cnt_pair = 0; g->cnt_node = 0; g->cnt_visit = 0;
    *queue = 0;
}

__global__ void k_batch_activate(Ctrl* c, int S) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) c[s].active = 1;
}

// *flag = 1 if a system can still take a substep (integrator.js:810-812 loop test).
__global__ __launch_bounds__(1024) void k_batch_pending(const Ctrl* __restrict__ c, int S, int32_t* flag) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int s = threadIdx.x; s < S; s += blockDim.x)
        if (c[s].total_steps < c[s].max_substeps && c[s].dt_remain > 1e-12) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) *flag = any;
}

// a->n_alive = a->n_tree = total while any system is inside its update, else 0.
__global__ __launch_bounds__(1024) void k_batch_any(const Ctrl* __restrict__ c, int S, Ctrl* __restrict__ a, int total) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int s = threadIdx.x; s < S; s += blockDim.x)
        if (c[s].active) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        a->n_alive = any ? total : 0;
        a->n_tree = any ? total : 0;
        a->active = any;
    }
}

// Sort by (system, Morton key), gather the state, build every system's tree,
// then the persistent walk.  Launches only: it is captured into the graph.
void force_pass(GeonBatch_* h, bool timed = false) {
    const hipStream_t s = h->stream;
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, h->gctrl, h->queue);
    hipLaunchKernelGGL(k_batch_any, dim3(1), dim3(1024), 0, s, h->bctrl, h->S, h->actrl, h->total);
    const KeyParams kp = keyp(h->prm);
    // keys and an identity permutation always; the sort only while a system is
    // active, so idle substeps gather with the identity
    launch_batch_keys(s, F(h, B_X), F(h, B_Y), I(h, B_VALID), I(h, B_SYS), kp, h->tb.keys, h->tb.perm, h->total);
    radix_sort(s, h->tb, h->total, kBatchKeyBits, h->actrl);
    Columns cols;
    cols.ncols = NB_STATE;
    for (int c = 0; c < NB_STATE; c++) { cols.src[c] = h->st[c]; cols.dst[c] = h->st_alt[c]; }
    launch_gather(s, cols, h->tb.perm, h->total, h->gctrl);
    for (int c = 0; c < NB_STATE; c++) std::swap(h->st[c], h->st_alt[c]);
    BatchTreeArgs ba;
    ba.S = h->S; ba.total = h->total; ba.soff = h->d_soff; ba.n_in = h->n_in; ba.seg_first = h->seg_first;
    ba.root_slot = h->root_slot;
    build_tree_batch(s, h->tb, ba, F(h, B_X), F(h, B_Y), F(h, B_MASS), F(h, B_WX), F(h, B_WY), F(h, B_MU), F(h, B_L),
                     I(h, B_QQ), kp, h->actrl);
    launch_prepare_sources(s, F(h, B_X), F(h, B_Y), F(h, B_MASS), F(h, B_CHARGE), F(h, B_VX), F(h, B_VY),
                           F(h, B_ANGVEL), F(h, B_RADIUS), F(h, B_MU), F(h, B_L), nullptr, nullptr, nullptr, h->src4,
                           h->src8a, h->src8b, h->total, h->actrl);
    const int threads = h->tune.walk_threads;
    const int grid = std::max(1, h->cus * h->tune.waves_per_cu / (threads / kWave));
    if (timed) hipEventRecord(h->ev0, s);
    launch_walk_persistent(s, force_args(h), features(h->prm), h->prm.kahan != 0, h->queue, h->total / kWave, grid,
                           threads, h->tune.stack);
    if (timed) hipEventRecord(h->ev1, s);
}

void substep(GeonBatch_* h, int k) {
    const BatchStepArgs b = bstep(h);
    launch_batch_ctrl(h->stream, h->bctrl, h->S, h->prm.softening, k == 0);
    launch_batch_kick_drift(h->stream, step_args(h), b);
    force_pass(h);                                  // swaps the state columns (sort + gather)
    launch_batch_dt_scan(h->stream, step_args(h), b);
}

int read_ctrls(GeonBatch_* h) {
    HIP_OK(hipMemcpyAsync(h->bctrl_host, h->bctrl, sizeof(Ctrl) * h->S, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(h->gctrl_host, h->gctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    Ctrl a;
    HIP_OK(hipMemcpyAsync(&a, h->actrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (h->gctrl_host->overflow || a.overflow) return set_err(h, GEON_EOVERFLOW, "tree node pool or walk stack overflow");
    return GEON_OK;
}

void drop_graph(GeonBatch_* h) {
    if (h->graph) hipGraphExecDestroy(h->graph);
    if (h->graph_next) hipGraphExecDestroy(h->graph_next);
    h->graph = nullptr;
    h->graph_next = nullptr;
}

// Substep bodies per captured graph: even, at most the substep cap rounded up to even.
int graph_bodies(const GeonBatch_* h) {
    const int cap = (h->prm.max_substeps + 1) / 2 * 2;
    const int b = h->tune.graph_bodies > 0 ? h->tune.graph_bodies : kChunk;
    return std::max(2, std::min(b, cap));
}

int capture_chunk(GeonBatch_* h, int first_body, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < graph_bodies(h); k++) substep(h, first_body + k);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess) return set_err(h, GEON_EDEVICE, std::string("substep graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) { *out = nullptr; return set_err(h, GEON_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e)); }
    return GEON_OK;
}

int capture(GeonBatch_* h) {
    drop_graph(h);
    int rc;
    const int b = graph_bodies(h);
    if ((rc = capture_chunk(h, 0, &h->graph))) return rc;
    if (b < h->prm.max_substeps && (rc = capture_chunk(h, b, &h->graph_next))) { drop_graph(h); return rc; }
    h->graph_st0 = h->st[0];
    return GEON_OK;
}

// Are there systems left inside their update?  (one small kernel + a pinned read)
int pending(GeonBatch_* h, bool* any) {
    hipLaunchKernelGGL(k_batch_pending, dim3(1), dim3(1024), 0, h->stream, h->bctrl, h->S, h->pending);
    HIP_OK(hipMemcpyAsync(h->pending_host, h->pending, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    *any = *h->pending_host != 0;
    return GEON_OK;
}

}  // namespace

extern "C" {

int geon_batch_create(const GeonParams* params, int32_t n_systems, const int32_t* max_particles, GeonBatchHandle* out) {
    if (!params || !out || n_systems <= 0 || n_systems > kMaxSystems || !max_particles) {
        g_batch_create_error = "bad argument (1 <= n_systems <= 1024)";
        return GEON_EINVAL;
    }
    std::string why;
    if (validate(params, &why)) { g_batch_create_error = why; return GEON_EUNSUPPORTED; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_batch_create_error = "no HIP device"; return GEON_EDEVICE; }
    GeonBatch_* h = new GeonBatch_();
    h->prm = *params;
    h->S = n_systems;
    h->cap.resize(n_systems);
    h->soff.resize(n_systems + 1);
    h->n_up.assign(n_systems, 0);
    int64_t total = 0;
    for (int s = 0; s < n_systems; s++) {
        if (max_particles[s] < 0) { delete h; g_batch_create_error = "negative system size"; return GEON_EINVAL; }
        const int64_t c = ((int64_t)std::max(max_particles[s], 1) + kBatchAlign - 1) / kBatchAlign * kBatchAlign;
        h->cap[s] = (int32_t)c;
        h->soff[s] = (int32_t)total;
        total += c;
    }
    if (total > (1 << 30)) { delete h; g_batch_create_error = "batch too large"; return GEON_EINVAL; }
    h->soff[n_systems] = (int32_t)total;
    h->total = (int)total;
    int rc = GEON_OK;
    auto fail = [&](int code) { geon_batch_destroy(h); return code; };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(GEON_EDEVICE);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return fail(GEON_EDEVICE);
    h->cus = prop.multiProcessorCount;
    const size_t T = (size_t)total;
    for (int c = 0; c < NB_STATE; c++) {
        if ((rc = dmalloc(h, &h->st[c], T)) || (rc = dmalloc(h, &h->st_alt[c], T))) return fail(rc);
        if (hipMemset(h->st[c], 0, 4 * T) != hipSuccess) return fail(GEON_EDEVICE);
    }
    for (int c = 0; c < NB_FORCE; c++) {
        if ((rc = dmalloc(h, &h->fc[c], T))) return fail(rc);
        if (hipMemset(h->fc[c], 0, 4 * T) != hipSuccess) return fail(GEON_EDEVICE);
    }
    if ((rc = dmalloc(h, &h->src4, T)) || (rc = dmalloc(h, &h->src8a, T)) || (rc = dmalloc(h, &h->src8b, T))) return fail(rc);
    TreeBufs& tb = h->tb;
    tb.cap_particles = total;
    tb.cap_divided = total + 1024;
    const size_t nodes = 1 + 4 * (size_t)tb.cap_divided + (size_t)n_systems;
    const size_t nblk = (T + 2047) / 2048 + 1;
    if ((rc = dmalloc(h, &tb.keys, T)) || (rc = dmalloc(h, &tb.keys_alt, T)) || (rc = dmalloc(h, &tb.perm, T)) ||
        (rc = dmalloc(h, &tb.perm_alt, T)) || (rc = dmalloc(h, &tb.hist, 256 * nblk)) ||
        (rc = dmalloc(h, &tb.digit_total, 256)) || (rc = dmalloc(h, &tb.owned, T)) ||
        (rc = dmalloc(h, &tb.rank_start, T)) || (rc = dmalloc(h, &tb.divnode, (size_t)tb.cap_divided)) ||
        (rc = dmalloc(h, &tb.pre, 8 * (T + 1))) || (rc = dmalloc(h, &tb.nodes, nodes)) ||
        (rc = dmalloc(h, &tb.node_mass64, nodes)) || (rc = dmalloc(h, &tb.node_comx64, nodes)) ||
        (rc = dmalloc(h, &tb.node_comy64, nodes)) || (rc = dmalloc(h, &tb.node_qq64, nodes)) ||
        (rc = dmalloc(h, &tb.node_path, nodes)) || (rc = dmalloc(h, &tb.node_npts, nodes)) ||
        (rc = dmalloc(h, &tb.node_first, nodes)) || (rc = dmalloc(h, &tb.scan_tmp, nblk + 1024)) ||
        (rc = dmalloc(h, &tb.scan_tmp64, 9 * (nblk + 1024))))
        return fail(rc);
    tb.node_mom = nullptr;
    tb.pre_mom = nullptr;
    if ((rc = dmalloc(h, &h->bctrl, (size_t)n_systems)) || (rc = dmalloc(h, &h->gctrl, 1)) ||
        (rc = dmalloc(h, &h->actrl, 1)) ||
        (rc = dmalloc(h, &h->d_soff, (size_t)n_systems + 1)) || (rc = dmalloc(h, &h->n_in, (size_t)n_systems)) ||
        (rc = dmalloc(h, &h->seg_first, T)) || (rc = dmalloc(h, &h->root_slot, (size_t)n_systems)) ||
        (rc = dmalloc(h, &h->queue, 1)) || (rc = dmalloc(h, &h->stats, 10 * (size_t)n_systems)) ||
        (rc = dmalloc(h, &h->pending, 1)))
        return fail(rc);
    if (hipHostMalloc((void**)&h->bctrl_host, sizeof(Ctrl) * n_systems, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->gctrl_host, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->pending_host, sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
        return fail(GEON_ENOMEM);
    if (hipMemset(h->bctrl, 0, sizeof(Ctrl) * n_systems) != hipSuccess) return fail(GEON_EDEVICE);
    Ctrl g;
    memset(&g, 0, sizeof(g));
    g.n_alive = (int32_t)total;
    g.n_tree = (int32_t)total;
    g.active = 1;
    if (hipMemcpy(h->gctrl, &g, sizeof(Ctrl), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->actrl, &g, sizeof(Ctrl), hipMemcpyHostToDevice) != hipSuccess)
        return fail(GEON_EDEVICE);
    // slot -> system, all slots empty
    std::vector<int32_t> sys(T);
    for (int s = 0; s < n_systems; s++)
        for (int k = h->soff[s]; k < h->soff[s + 1]; k++) sys[k] = s;
    if (hipMemcpy(h->st[B_SYS], sys.data(), 4 * T, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_soff, h->soff.data(), 4 * (n_systems + 1), hipMemcpyHostToDevice) != hipSuccess)
        return fail(GEON_EDEVICE);
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) return fail(GEON_EDEVICE);
    h->tune.walk_threads = 256;
    h->tune.waves_per_cu = 16;
    h->tune.stack = 96;
    h->tune.use_graph = 1;
    h->tune.graph_bodies = kChunk;
    h->tune.walk_ms = 0;
    *out = h;
    return GEON_OK;
}

void geon_batch_destroy(GeonBatchHandle h) {
    if (!h) return;
    hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (int c = 0; c < NB_STATE; c++) { hipFree(h->st[c]); hipFree(h->st_alt[c]); }
    for (int c = 0; c < NB_FORCE; c++) hipFree(h->fc[c]);
    hipFree(h->src4); hipFree(h->src8a); hipFree(h->src8b);
    TreeBufs& tb = h->tb;
    hipFree(tb.keys); hipFree(tb.keys_alt); hipFree(tb.perm); hipFree(tb.perm_alt); hipFree(tb.hist);
    hipFree(tb.digit_total); hipFree(tb.owned); hipFree(tb.rank_start); hipFree(tb.divnode); hipFree(tb.pre);
    hipFree(tb.nodes); hipFree(tb.node_mass64); hipFree(tb.node_comx64); hipFree(tb.node_comy64);
    hipFree(tb.node_qq64); hipFree(tb.node_path); hipFree(tb.node_npts); hipFree(tb.node_first);
    hipFree(tb.scan_tmp); hipFree(tb.scan_tmp64);
    hipFree(h->bctrl); hipFree(h->gctrl); hipFree(h->actrl); hipFree(h->d_soff); hipFree(h->n_in); hipFree(h->seg_first);
    hipFree(h->root_slot); hipFree(h->queue); hipFree(h->stats); hipFree(h->pending);
    if (h->pending_host) hipHostFree(h->pending_host);
    if (h->bctrl_host) hipHostFree(h->bctrl_host);
    if (h->gctrl_host) hipHostFree(h->gctrl_host);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

// GPUPhysics-style bulk upload of one system (gpu-physics.js:3886 deserialize,
// main.js:504-539 normalisation): replaces the system's particles.
int geon_batch_upload(GeonBatchHandle h, int32_t system, const GeonSoAView* v, int32_t n) {
    if (!h || !v || system < 0 || system >= h->S || n < 0) return GEON_EINVAL;
    if (n > h->cap[system]) return set_err(h, GEON_EFULL, "more particles than the system's max_particles");
    if (n > 0 && (!v->x || !v->y || !v->wx || !v->wy || !v->mass || !v->charge))
        return set_err(h, GEON_EINVAL, "upload needs x, y, wx, wy, mass, charge");
    const bool rel = !!(h->prm.toggles0 & GEON_RELATIVITY);
    const int cap = h->cap[system], off = h->soff[system];
    std::vector<std::vector<uint32_t>> col(NB_STATE, std::vector<uint32_t>((size_t)cap, 0u));
    auto putf = [&](int c, int i, double val) { float f = (float)val; memcpy(&col[c][i], &f, 4); };
    for (int i = 0; i < n; i++) {
        const double m = v->mass[i];
        if (!(m > 0)) return set_err(h, GEON_EINVAL, "mass must be > 0");
        const int32_t qq = (int32_t)llround(v->charge[i] / 0.1);     // main.js:511
        const double r = cbrt(m);                                     // particle.js:96
        const double angw = v->angw ? v->angw[i] : 0.0;
        const double wx = v->wx[i], wy = v->wy[i];
        const double ig = rel ? 1.0 / sqrt(1 + wx * wx + wy * wy) : 1.0;
        putf(B_X, i, v->x[i]); putf(B_Y, i, v->y[i]);
        putf(B_WX, i, wx); putf(B_WY, i, wy);
        putf(B_VX, i, wx * ig); putf(B_VY, i, wy * ig);
        putf(B_MASS, i, m);
        putf(B_CHARGE, i, qq * 0.1);
        putf(B_ANGW, i, angw);
        putf(B_ANGVEL, i, rel ? angw / sqrt(1 + angw * angw * r * r) : angw);
        putf(B_RADIUS, i, r);
        col[B_QQ][i] = (uint32_t)qq;
        col[B_SLOT][i] = (uint32_t)i;
        col[B_VALID][i] = 1u;
    }
    for (int i = 0; i < cap; i++) col[B_SYS][i] = (uint32_t)system;
    for (int c = 0; c < NB_STATE; c++)
        HIP_OK(hipMemcpyAsync(h->st[c] + off, col[c].data(), 4 * (size_t)cap, hipMemcpyHostToDevice, h->stream));
    for (int c = 0; c < NB_FORCE; c++) HIP_OK(hipMemsetAsync(h->fc[c] + off, 0, 4 * (size_t)cap, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    h->n_up[system] = n;
    h->forces_init = false;
    return GEON_OK;
}

int geon_batch_download(GeonBatchHandle h, int32_t system, GeonSoAView* v, int32_t out_cap, int32_t* n_slots) {
    if (!h || !v || system < 0 || system >= h->S) return GEON_EINVAL;
    const int cap = h->cap[system], off = h->soff[system], ns = h->n_up[system];
    if (n_slots) *n_slots = ns;
    if (out_cap < ns) { h->err = "download arrays hold fewer entries than the system's slots"; return GEON_EINVAL; }
    std::vector<std::vector<uint32_t>> col(NB_STATE, std::vector<uint32_t>((size_t)cap));
    for (int c = 0; c < NB_STATE; c++)
        HIP_OK(hipMemcpyAsync(col[c].data(), h->st[c] + off, 4 * (size_t)cap, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    auto getf = [&](int c, int i) { float f; memcpy(&f, &col[c][i], 4); return (double)f; };
    if (v->alive) for (int s = 0; s < ns; s++) v->alive[s] = 0;
    for (int i = 0; i < cap; i++) {
        if (col[B_MASS][i] == 0u) continue;             // an empty slot of the slice (masses are > 0)
        const int s = (int)col[B_SLOT][i];
        if (s < 0 || s >= ns) continue;
        if (v->x) v->x[s] = getf(B_X, i);
        if (v->y) v->y[s] = getf(B_Y, i);
        if (v->wx) v->wx[s] = getf(B_WX, i);
        if (v->wy) v->wy[s] = getf(B_WY, i);
        if (v->mass) v->mass[s] = getf(B_MASS, i);
        if (v->base_mass) v->base_mass[s] = getf(B_MASS, i);
        if (v->charge) v->charge[s] = (int32_t)col[B_QQ][i] * 0.1;
        if (v->angw) v->angw[s] = getf(B_ANGW, i);
        if (v->alive) v->alive[s] = (int32_t)col[B_VALID][i];     // 0: despawned
    }
    if (n_slots) *n_slots = ns;
    return GEON_OK;
}

int geon_batch_compute_forces(GeonBatchHandle h) {
    if (!h) return GEON_EINVAL;
    hipLaunchKernelGGL(k_batch_activate, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->bctrl, h->S);
    force_pass(h);
    HIP_OK(hipGetLastError());
    h->forces_init = true;
    return read_ctrls(h);
}

// One Physics.update(dt) of every system (integrator.js:726-1825).
int geon_batch_update(GeonBatchHandle h, double dt, int32_t* substeps) {
    if (!h) return GEON_EINVAL;
    if (!(dt >= 0)) return set_err(h, GEON_EINVAL, "dt must be >= 0");
    int rc;
    if (!h->forces_init) {
        hipLaunchKernelGGL(k_batch_activate, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->bctrl, h->S);
        force_pass(h);
        h->forces_init = true;
    }
    launch_batch_begin(h->stream, h->bctrl, h->S, dt, h->prm.max_substeps);
    launch_batch_dt_scan(h->stream, step_args(h), bstep(h));
    // The substep loop in chunks of graph_bodies bodies (systems that finished
    // are predicated off inside the kernels).  The chunks that the last
    // update's longest system needed go out back to back; after that the host
    // checks, after each chunk, whether any system can take another substep.
    const int B = graph_bodies(h);
    const int chunks = (h->prm.max_substeps + B - 1) / B;
    const int ahead = std::max(1, (h->last_max_steps + B - 1) / B);
    for (int c = 0; c < chunks; c++) {
        if (c >= ahead) {
            bool any = false;
            if ((rc = pending(h, &any))) return rc;
            if (!any) break;
        }
        if (h->tune.use_graph) {
            // a force pass outside the graphs (bootstrap, compute_forces) flips the ping-pong
            if (!h->graph || h->graph_st0 != h->st[0]) {
                HIP_OK(hipStreamSynchronize(h->stream));
                if ((rc = capture(h))) return rc;
            }
            HIP_OK(hipGraphLaunch(c == 0 ? h->graph : h->graph_next, h->stream));
        } else {
            for (int k = 0; k < B; k++) substep(h, c * B + k);
        }
    }
    launch_batch_boundary(h->stream, step_args(h), bstep(h), h->prm.boundary_mode, h->prm.domain_w, h->prm.domain_h);
    HIP_OK(hipGetLastError());
    if ((rc = read_ctrls(h))) return rc;
    h->last_max_steps = 0;
    for (int s = 0; s < h->S; s++) h->last_max_steps = std::max(h->last_max_steps, h->bctrl_host[s].total_steps);
    if (substeps)
        for (int s = 0; s < h->S; s++) substeps[s] = h->bctrl_host[s].total_steps;
    return GEON_OK;
}

int geon_batch_get_forces(GeonBatchHandle h, int32_t system, GeonForceView* v, int32_t n_slots) {
    if (!h || !v || system < 0 || system >= h->S) return GEON_EINVAL;
    const int cap = h->cap[system], off = h->soff[system];
    std::vector<int32_t> slot((size_t)cap), valid((size_t)cap);
    std::vector<std::vector<float>> fc(NB_FORCE, std::vector<float>((size_t)cap));
    HIP_OK(hipMemcpyAsync(slot.data(), h->st[B_SLOT] + off, 4 * (size_t)cap, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(valid.data(), h->st[B_VALID] + off, 4 * (size_t)cap, hipMemcpyDeviceToHost, h->stream));
    for (int c = 0; c < NB_FORCE; c++)
        HIP_OK(hipMemcpyAsync(fc[c].data(), h->fc[c] + off, 4 * (size_t)cap, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    double* outs[NB_FORCE] = {v->fx, v->fy, v->pe, v->bz, v->dbzdx, v->dbzdy, v->bgz, v->dbgzdx, v->dbgzdy,
                              v->frame_drag_torque, v->tidal_torque, v->f1pn_x, v->f1pn_y};
    for (int i = 0; i < cap; i++) {
        const int s = slot[i];
        if (!valid[i] || s < 0 || s >= n_slots) continue;
        for (int c = 0; c < NB_FORCE; c++) if (outs[c]) outs[c][s] = fc[c][i];
    }
    return GEON_OK;
}

int geon_batch_get_stats(GeonBatchHandle h, int32_t system, GeonStats* out) {
    if (!h || !out || system < 0 || system >= h->S) return GEON_EINVAL;
    memset(out, 0, sizeof(*out));
    launch_batch_stats(h->stream, step_args(h), bstep(h), h->fc[BF_PE], !!(h->prm.toggles0 & GEON_RELATIVITY),
                       h->stats);
    double r[10];
    HIP_OK(hipMemcpyAsync(r, h->stats + 10 * system, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(h->bctrl_host, h->bctrl, sizeof(Ctrl) * h->S, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    out->linear_ke = r[0]; out->spin_ke = r[1]; out->px = r[2]; out->py = r[3];
    out->total_mass = r[4];
    out->com_x = r[4] > 0 ? r[5] / r[4] : 0; out->com_y = r[4] > 0 ? r[6] / r[4] : 0;
    out->pe = 0.5 * r[7];                       // forces.js:26
    out->orbital_ang_mom = r[8]; out->spin_ang_mom = r[9];
    out->alive_count = h->n_up[system] - h->bctrl_host[system].deaths;
    return GEON_OK;
}

int geon_batch_get_tune(GeonBatchHandle h, GeonBatchTune* out) {
    if (!h || !out) return GEON_EINVAL;
    *out = h->tune;
    return GEON_OK;
}

int geon_batch_set_tune(GeonBatchHandle h, const GeonBatchTune* t) {
    if (!h || !t) return GEON_EINVAL;
    if ((t->walk_threads != 64 && t->walk_threads != 128 && t->walk_threads != 256) || t->waves_per_cu < 1 ||
        t->waves_per_cu > 64 || (t->stack != 64 && t->stack != 96) || t->graph_bodies < 0 || t->graph_bodies > 32 ||
        (t->graph_bodies & 1))
        return set_err(h, GEON_EINVAL,
                       "tune: walk_threads 64/128/256, waves_per_cu 1..64, stack 64/96, graph_bodies even 0..32");
    h->tune = *t;
    drop_graph(h);
    return GEON_OK;
}

// Times the persistent walk of the current state for every (threads,
// waves per CU, stack) candidate and keeps the fastest that runs without a
// stack overflow.
int geon_batch_autotune(GeonBatchHandle h, GeonBatchTune* out) {
    if (!h) return GEON_EINVAL;
    const GeonBatchTune keep = h->tune;
    const int threads[] = {64, 128, 256};
    const int waves[] = {8, 16, 32};
    const int stacks[] = {64, 96};
    GeonBatchTune best = keep;
    best.walk_ms = 1e30;
    hipLaunchKernelGGL(k_batch_activate, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->bctrl, h->S);
    for (int th : threads)
        for (int wv : waves)
            for (int st : stacks) {
                h->tune.walk_threads = th; h->tune.waves_per_cu = wv; h->tune.stack = st;
                double ms_min = 1e30;
                for (int rep = 0; rep < 3; rep++) {
                    force_pass(h, true);
                    HIP_OK(hipEventSynchronize(h->ev1));
                    float ms = 0;
                    HIP_OK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
                    if (ms < ms_min) ms_min = ms;
                }
                int rc = read_ctrls(h);
                if (rc == GEON_EOVERFLOW) {                 // this stack is too small: not a candidate
                    HIP_OK(hipMemsetAsync(&h->gctrl->overflow, 0, sizeof(int32_t), h->stream));
                    HIP_OK(hipMemsetAsync(&h->actrl->overflow, 0, sizeof(int32_t), h->stream));
                    continue;
                }
                if (rc) return rc;
                if (ms_min < best.walk_ms) {
                    best = h->tune;
                    best.walk_ms = ms_min;
                }
            }
    if (best.walk_ms >= 1e30) { h->tune = keep; return set_err(h, GEON_EOVERFLOW, "no walk configuration ran"); }
    best.use_graph = keep.use_graph;
    best.graph_bodies = keep.graph_bodies;
    h->tune = best;
    drop_graph(h);
    HIP_OK(hipStreamSynchronize(h->stream));
    h->forces_init = true;
    if (out) *out = best;
    return GEON_OK;
}

const char* geon_batch_last_error(GeonBatchHandle h) {
    if (!h) return g_batch_create_error.c_str();
    return h->err.c_str();
}

}  // extern "C"
