// geon_kernels.h — host-side launchers of the gfx950 kernels (internal to
// libgeon_hip.so; not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "geon_dev.h"
#include "geon_delay.h"

namespace geon {

// Device control block: the substep controller lives on the device so the
// adaptive-dt decision never needs the host (integrator.js:810-874).
struct Ctrl {
    double dt_remain, dt_sub, sim_time, dt_total;
    uint64_t max_acc2_fresh;    // bits of a non-negative double (ordered as uint64)
    uint64_t max_acc2_kick;     // max |F/m|^2 seen by the last kick (H8, integrator.js:924-928)
    uint64_t max_cyc;           // max cyclotron frequency (integrator.js:817-849)
    int32_t total_steps, active, n_alive, n_tree;
    int32_t max_substeps, n_divided, overflow, deaths;
    int32_t first_substep, pad0, pad1, pad2;
    int64_t cnt_pair, cnt_node, cnt_visit, depth_overflow;
};

// Word-columns gathered by a permutation (all 32-bit).
constexpr int kMaxCols = 40;
struct Columns {
    uint32_t* src[kMaxCols];
    uint32_t* dst[kMaxCols];
    int ncols;
};

// ---- sort + tree (k_tree.hip) ----
struct TreeBufs {
    uint64_t *keys, *keys_alt;
    uint32_t *perm, *perm_alt;
    uint32_t* hist;             // 256 * max_blocks
    uint32_t* digit_total;      // 256
    int32_t *owned, *rank_start;
    int2* divnode;              // (lo, level) per divided node, pre-order rank
    double* pre;                // 8 columns of (cap + 1) prefix sums
    Node* nodes;
    double *node_mass64, *node_comx64, *node_comy64, *node_qq64;
    uint64_t* node_path;
    int32_t* node_npts;         // points in each node's subtree
    int32_t* node_first;        // first sorted point of each node's subtree
    NodeMom* node_mom;          // far-field moments of divided nodes (when requested)
    double* pre_mom;            // 9 columns of (cap + 1) moment prefix sums
    int32_t* scan_tmp;          // block sums for int scans
    double* scan_tmp64;
    int64_t cap_particles, cap_divided;
};

struct KeyParams {
    double xmin, ymin, scale_x, scale_y;   // u = (x - xmin) * scale (exact for dyadic boxes)
    double root_hw;                          // root half width (node size = 4W / 2^level)
};

void launch_keys(hipStream_t s, const float* x, const float* y, int n, KeyParams kp,
                 uint64_t* keys, uint32_t* perm, Ctrl* ctrl);
// Sorts keys/perm in place (ping-pong internally); n is a host upper bound,
// the exact count is ctrl->n_alive.
constexpr int kSortTileKeys = 2048;   // keys per radix-sort block: hist holds 256 words per block
void radix_sort(hipStream_t s, TreeBufs& tb, int n_upper, int key_bits, const Ctrl* ctrl);
void launch_gather(hipStream_t s, const Columns& cols, const uint32_t* perm, int n_upper, const Ctrl* ctrl);
void build_tree(hipStream_t s, TreeBufs& tb, const float* x, const float* y, const float* m,
                const float* wx, const float* wy, const float* mu, const float* L,
                const int32_t* qq, int n_upper, KeyParams kp, Ctrl* ctrl, bool moments = false);
// Batch of independent systems (slices of the sorted arrays, kScanTile-aligned).
struct BatchTreeArgs {
    int S, total;
    const int32_t* soff;        // S + 1 slice offsets
    int32_t* n_in;              // per system: tree points (written)
    int32_t* seg_first;         // per slot: first slot of its system (written)
    int32_t* root_slot;         // per system: root node slot (written)
};
void launch_batch_keys(hipStream_t s, const float* x, const float* y, const int32_t* valid, const int32_t* sys,
                       KeyParams kp, uint64_t* keys, uint32_t* perm, int total);
void build_tree_batch(hipStream_t s, TreeBufs& tb, const BatchTreeArgs& ba, const float* x, const float* y,
                      const float* m, const float* wx, const float* wy, const float* mu, const float* L,
                      const int32_t* qq, KeyParams kp, Ctrl* ctrl);
constexpr int kBatchAlign = 2048;   // slice alignment (the prefix-scan tile)
// Exclusive scan of int32 flags (n from device ctrl->n_alive) -> out, total -> *total_dev.
void exclusive_scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int n_upper,
                        const int32_t* n_dev, int32_t* total_dev, int32_t* tmp);

// ---- forces (k_force.hip) ----
struct ForceArgs {
    const float4* src4;         // {x, y, mass, charge}
    const float4* src8a;        // {vx, vy, ang_vel, mu}
    const float4* src8b;        // {L, ax, yuk, higgs}
    const float* radius;
    const Node* nodes;
    float *fx, *fy, *pe, *bz, *dbzdx, *dbzdy, *bgz, *dbgzdx, *dbgzdy, *frame_drag, *tidal, *f1x, *f1y;
    float *jx = nullptr, *jy = nullptr;     // analytic jerk (radiation on)
    const float* mass_for_cyc;  // for cyclotron scan in the epilogue
    const float* charge;
    Ctrl* ctrl;
    PhysFlags pf;
    float ext_gx, ext_gy, ext_ex, ext_ey, ext_bz;
    uint32_t cyc_mag, cyc_gm;   // which B fields enter the cyclotron estimate
    int n_upper;
    int64_t max_nodes;          // node records allocated (walk bounds check)
    // far field of accepted nodes: 0 monopole only, 1 + quadrupole/dipole
    // terms from node_mom, 2 also list (node, lane) per 64-particle group
    int far_mode;
    const NodeMom* node_mom;
    uint32_t* far_list;         // far_cap entries per group in lane queues (far_slot)
    int32_t* far_count;         // entries per lane queue (kWave per group)
    int64_t far_cap;
    // batch of systems (null for one system): per slot system id and liveness,
    // per system controller (its `active` flag) and root node slot
    const int32_t* bsys;
    const int32_t* bvalid;
    const Ctrl* bctrl;
    const int32_t* broot;
    // signal delay of leaf sources (hist.on), and the retired sources
    HistArgs hist;
    // 1PN recompute without history: 0 takes no leaf source at all (the
    // reference takes them only from a light cone that nothing has recorded)
    int pn_leaves = 1;
    // observers when they are not the sources (periodic images in the tree):
    // null means src4/src8a/src8b; src_orig[j] is the particle source j images
    const float4 *obs4 = nullptr, *obs8a = nullptr, *obs8b = nullptr;
    const int32_t* src_orig = nullptr;
    // the selected particle (readStats' `selected`): its index, or -1, and its
    // per-type force record (kSelFloats floats, layout SelField)
    const int32_t* sel = nullptr;
    float* sel_out = nullptr;
};

// Per-type force record of the selected particle: the pair terms by type, the
// external fields, the learned far-field correction, the Higgs and axion
// gradient forces, the fields' gravity and the last radiation-reaction force.
enum SelField {
    SEL_SPLIT = 0, SEL_EXT = 2 * kSplitTypes, SEL_FAR = SEL_EXT + 2, SEL_HIGGS = SEL_FAR + 2,
    SEL_AXION = SEL_HIGGS + 2, SEL_FIELD_GRAV = SEL_AXION + 2, SEL_RAD = SEL_FIELD_GRAV + 2,
    SEL_VALID = SEL_RAD + 2,    // 1 once a force pass has filled the record
    kSelFloats = SEL_VALID + 1
};
void launch_sel_find(hipStream_t s, const int32_t* slot, int32_t target, int32_t* idx, int n_upper, const Ctrl* ctrl);
// After the pair and dead-source forces of a pass: the selected particle's
// pair forces again, by type (bh: the pass walked the tree; residual: the
// remainder of its pair force is the learned far-field correction).
void launch_sel_forces(hipStream_t s, const ForceArgs& a, bool bh, bool residual);

// ---- periodic images in the tree (k_ghost.hip) ----
enum GhostCol { G_X, G_Y, G_M, G_Q, G_WX, G_WY, G_VX, G_VY, G_AV, G_RAD, G_MU, G_L, G_QQ, G_ORIG, kGhostCols };
struct GhostCols {
    uint32_t* c[kGhostCols];
};
struct GhostArgs {
    const float *x, *y, *m, *q, *wx, *wy, *vx, *vy, *ang_vel, *radius, *mu, *L;
    const int32_t* qq;
    const Ctrl* ctrl;           // the particles' (n_alive)
    float dom_w, dom_h, margin;
    int topo;
    int cap_points;             // tree points the output columns hold
    GhostCols out;
};
// Tree points of the particles (n of ctrl) and their images, unsorted; the
// count goes to gctrl->n_alive.
void ghost_tree_points(hipStream_t s, const GhostArgs& a, int n_upper, int32_t* cnt, int32_t* off, int32_t* scan_tmp,
                       Ctrl* gctrl);
// Walk sources of the sorted tree points; the modulations (axMod, yukMod,
// higgsMod) are those of the particle a point images, current at the walk.
void ghost_sources(hipStream_t s, const GhostCols& col, const float* ax, const float* yuk, const float* higgs,
                   float4* src4, float4* src8a, float4* src8b, int n_upper, const Ctrl* gctrl);
void ghost_finish(hipStream_t s, const Ctrl* gctrl, Ctrl* ctrl);
constexpr int kFarMultipole = 1, kFarList = 2;
// Far pair lists (learned far field): a 64-particle group owns far_cap entries
// (a multiple of 4 * kWave), one queue of far_cap / kWave accepted node
// indices per lane.  Entry t of lane l sits at (t / 4) * 4 kWave + 4 l + t % 4
// in the group's slice, so the t-th entries of the group's lanes lie in one
// 1 KB block and each lane's next three in the same 16 bytes.
__host__ __device__ __forceinline__ int64_t far_slot(int64_t group, int64_t far_cap, int lane, int t) {
    return group * far_cap + (int64_t)(t >> 2) * (4 * kWave) + 4 * lane + (t & 3);
}
void launch_walk(hipStream_t s, const ForceArgs& a, uint32_t feat, bool kahan);
// Persistent walk: `grid` blocks of `threads` pull 64-particle groups from *queue
// (reset to 0 before the launch) until all n_groups are done.
void launch_walk_persistent(hipStream_t s, const ForceArgs& a, uint32_t feat, bool kahan, int* queue,
                            int64_t n_groups, int grid, int threads, int stack);
void launch_direct(hipStream_t s, const ForceArgs& a, uint32_t feat, bool kahan);
void launch_prepare_sources(hipStream_t s, const float* x, const float* y, const float* m,
                            const float* q, const float* vx, const float* vy, const float* ang_vel,
                            const float* radius, float* mu, float* L, const float* ax,
                            const float* yuk, const float* higgs, float4* src4, float4* src8a,
                            float4* src8b, int n_upper, const Ctrl* ctrl, int update_moments = 1);
// 1PN recompute for the velocity-Verlet correction: tree walk or O(N^2).
void launch_walk1pn(hipStream_t s, const ForceArgs& a, bool tree);

// ---- signal delay (k_delay.hip) ----
// Forces of the retired sources in a.hist.dead on every live particle, added
// to the force outputs (no PE, forces.js:128-153).
void launch_dead_forces(hipStream_t s, const ForceArgs& a, uint32_t feat);
// History sample of every live particle as event `ev` (integrator.js:1549-1565);
// ev_time[ev] takes the device simulation time.
struct HistWrite {
    float4* h4;
    float* haw;
    int2* fe;
    double2* cd;
    double* ev_time;
    DeadRec* dead;
};
void launch_hist_record(hipStream_t s, const HistWrite& w, const float* x, const float* y, const float* vx,
                        const float* vy, const float* angw, const int32_t* slot, int64_t ev, int n,
                        const Ctrl* ctrl);
// _retireParticle (integrator.js:152-173) for every particle with alive[i] == 0:
// death snapshot as its private event `ev`, dead record at dead_base + its rank
// among the dying (alive_scan: exclusive scan of alive).
void launch_retire(hipStream_t s, const HistWrite& w, const float* x, const float* y, const float* vx,
                   const float* vy, const float* angw, const float* mass, const float* charge,
                   const float* radius, const float* ax, const float* yuk, const float* higgs,
                   const int32_t* slot, const int32_t* alive, const int32_t* alive_scan, int64_t ev,
                   int dead_base, int n, const Ctrl* ctrl);
// Slot metadata of new particles: no history, created at `t` (-inf: loaded).
void launch_hist_reset(hipStream_t s, int2* fe, double2* cd, int first_slot, int count, double t);

// Collision broadphase over the tree (count pass, then write pass).
struct PairArgs {
    const float *x, *y, *mass, *radius;
    const int32_t* slot;
    const float4* src4;
    const Node* nodes;
    int64_t max_nodes;
    int32_t* count;             // count pass: pairs found per particle
    const int32_t* offset;      // write pass: exclusive scan of count
    int32_t *pa, *pb;           // write pass: pair slots (index_out: sorted indices)
    int32_t* pl;                // write pass, optional: first sorted point of the candidate's leaf
    int index_out;
    float slack;                // box and overlap tests widened by this factor (1: exact)
    float dom_w, dom_h;
    int periodic;
    Ctrl* ctrl;
    int n_upper;
    int64_t cap_pairs = INT64_MAX;  // write pass: pairs the output holds (beyond: overflow 4, not written)
};
void launch_pairs(hipStream_t s, const PairArgs& a, bool write);

// ---- collision mode bounce (k_step.hip) ----
// _applyRepulsion / _repelPair (integrator.js:669-724): the overlap pairs of
// the pair list (pa lower slot, pb higher, grouped by pa), each body's sum of
// its contacts in a fixed order: as the higher-slot member in pair order,
// then as the lower in list order.  Hertz force and friction into the force,
// contact torque into `contact`, normal forces into the selected particle's
// external share.
struct RepelArgs {
    const float *x, *y, *vx, *vy, *ang_vel, *radius;
    const int32_t *pa, *pb, *cnt1, *off1;
    int32_t *cnt2, *off2, *fill2, *list2;
    int32_t* total2;                // device: scratch for the scan's total
    const int32_t* n_pairs;         // device: pairs in the list
    float *fx, *fy, *contact;
    const int32_t* sel;
    float* sel_out;
    const Ctrl* ctrl;
    float dom_w, dom_h, friction;
    int periodic;
    int64_t cap_pairs;
    int32_t* scan_tmp;
};
void launch_repulsion(hipStream_t s, const RepelArgs& a, int n_upper);

// ---- collision mode merge (k_merge.hip) ----
// Outcome of one pair of the fold; spawn records use the same layout.
struct MergeRes {
    double x, y, wx, wy, mass, base_mass, angw, e, px, py;
    int32_t qq, anti, type, mev;    // type 0 none, 1 merge, 2 annihilation; mev: merge with lost KE
};
struct MergeArgs {
    const float *x, *y, *wx, *wy, *mass, *base_mass, *radius, *angw;
    const int32_t* qq;
    const uint32_t* flags;          // bit 0: antimatter
    const int32_t *p1, *p2;         // pairs as sorted particle indices
    const uint32_t* order;          // fold position -> pair
    int32_t* deg;                   // pairs per particle
    double *wm, *wbm, *wr, *wdm;    // working mass, base mass, radius, death mass (FP64)
    MergeRes *res, *spawns;
    int32_t *f_spawn, *f_mev, *f_aev, *o_spawn, *o_mev, *o_aev;
    int32_t *f_fold, *o_fold, *fold;    // the pairs of the ordered fold, compacted in fold order
    int32_t* uf;                        // union-find parents over the particles (the fold's components)
    int32_t* uf_fail;                   // set if a union gave up: the one-thread fold runs instead
    int32_t* cnt;                   // totals: [0] spawns, [1] merge events, [2] annihilation events, [3] deaths, [6] fold pairs
    int32_t* scan_tmp;
    double* events;                 // 6 doubles each: type, x, y, energy, px, py
    Ctrl* mc;                       // n_alive = number of pairs
    const Ctrl* ctrl;               // n_alive = number of particles
    int periodic;
    double dom_w, dom_h;
};
struct SpawnArgs {
    const MergeRes* spawns;
    int count, base, first_slot, periodic, rel, n_force;
    double dom_w, dom_h;
    float *x, *y, *wx, *wy, *vx, *vy, *mass, *base_mass, *charge, *angw, *angvel, *radius;
    float *mu, *L, *f1ox, *f1oy, *ax, *yuk, *higgs;
    int32_t *qq, *slot;
    uint32_t* flags;
    float* force[32];
};
void merge_pair_same_leaf(hipStream_t s, const int32_t* p1, const int32_t* leaf, const Ctrl* mc, int32_t* need, int np);
void merge_leaf_ranks(hipStream_t s, const TreeBufs& tb, const int32_t* slot, int n_nodes, int32_t* s5,
                      int32_t* mark, uint64_t* chunk5, int n_upper, const Ctrl* ctrl, const int32_t* p1,
                      const int32_t* leaf, const Ctrl* mc, int32_t* rank, int np);
void merge_pair_keys(hipStream_t s, const int32_t* p1, const int32_t* p2, const int32_t* leaf, const int32_t* slot,
                     const int32_t* rank, const Ctrl* mc, const Ctrl* ctrl, uint64_t* keys, uint32_t* perm, int np);
// Fold pairs above which the fold runs by connected components.
constexpr int kFoldOneThread = 64;      // one thread folds ~0.5 us per pair; the component path costs ~70 us
hipError_t merge_resolve(hipStream_t s, const MergeArgs& m, TreeBufs& csort, Ctrl* cctrl, int np, int n);
void merge_apply(hipStream_t s, const MergeArgs& m, float* mass, float* base_mass, float* radius, int32_t* alive,
                 int n);
void merge_spawn(hipStream_t s, const SpawnArgs& a);
void merge_excite(hipStream_t s, float* field_dot, int G, double W, double H, const double* events, int n_events,
                  double frac);

// ---- learned far-field correction (k_learn.hip) ----
constexpr int kFarIn = 10;          // features of a pair
constexpr int kFarHid = 64;         // hidden width (2 layers)
constexpr int kFarOut = 6;          // gravity force (u, v), Coulomb force (u, v), gravity PE, Coulomb PE
constexpr int kFarW = 4;            // loss weights: gravity force, Coulomb force, gravity PE, Coulomb PE
constexpr int kFarSample = kFarIn + kFarOut + kFarW;   // floats per training sample
constexpr int kFarOffW1 = 0;        // W1[16][64] (input 10 = bias of layer 1)
constexpr int kFarOffW2 = 1024;     // W2[64][64]
constexpr int kFarOffB2 = 5120;     // b2[64]
constexpr int kFarOffW3 = 5184;     // W3[64][6]
constexpr int kFarOffB3 = 5568;     // b3[6]
constexpr int kFarParams = 5574;
constexpr int kFarStats = 2 * kFarIn + kFarOut;   // feature mean[10], 1/std[10], output scale[6]

struct FarFrags;
struct FarArgs {
    const FarFrags* frags;
    const float4* src4;
    const Node* nodes;
    const NodeMom* node_mom;
    const uint32_t* far_list;
    const int32_t* far_count;
    int64_t far_cap;
    float *fx, *fy, *pe;
    const Ctrl* ctrl;
    float soft2, dom_w, dom_h;
    int periodic, grav, coul;
};
struct FarSampleArgs {
    const float *x, *y, *m;
    const int32_t* qq;
    const Node* nodes;
    const NodeMom* node_mom;
    const double *mass64, *comx64, *comy64, *qq64;
    const int32_t *first, *npts;
    const uint32_t* far_list;
    const int32_t* far_count;   // entries per lane queue
    int64_t far_cap;
    const int64_t* groups;      // selected groups
    const int64_t* row_off;     // first sample row of each selected group
    int n_groups;
    int64_t n_rows;
    const double* f2ref;        // mean |F|^2 and 2 mean PE^2 of the system (loss weight normalisers)
    float* out;                 // n_rows * kFarSample
    const int32_t* slot;        // observer slots (optional record)
    int32_t* obs;               // n_rows observer slots, or null
    double soft2, dom_w, dom_h;
    int periodic, grav, coul;
};
struct FarTrainArgs {
    const FarFrags* frags;
    const float* samples;
    int64_t n_samples;
    const int64_t* rows;        // explicit rows (gradient query) or null (random batch)
    int64_t batch;
    uint64_t seed;
    int64_t step;
    float wscale;
    float* grad_part;           // blocks * kFarParams
    float* loss_part;           // blocks
    float* grad;                // kFarParams
    float* loss_out;            // 1
    float *params, *adam_m, *adam_v;
};
size_t far_frags_bytes();
void launch_far_pack(hipStream_t s, const float* params, const float* stats, void* frags);
void launch_far_mlp(hipStream_t s, const FarArgs& a, int n_upper);
void launch_far_samples(hipStream_t s, const FarSampleArgs& a);
void launch_far_train(hipStream_t s, const FarTrainArgs& a, int nblocks, int apply, float lr, int64_t t);
void launch_far_f2(hipStream_t s, const float* fx, const float* fy, const float* pe, const Ctrl* ctrl, double* out);

// ---- scalar fields (k_field.hip) ----
struct FieldArgs {
    int G, bc, higgs;
    int topo;                   // loop boundary: 0 torus, 1 Klein bottle, 2 RP^2
    float cw, ch, icw, ich, inv_area;
    float vac;                  // vacuum value: Higgs 1, axion 0
    float mu2, m2, damp;        // Higgs: mu^2 = m_H^2/2, damp 2 m_H; axion: m_a^2, damp g m_a
    float *field, *field_dot, *field_dot_alt, *grad_x, *grad_y;
    float *source, *thermal, *sg_phi, *sg_gx, *sg_gy;
    unsigned long long *dep_src, *dep_th;   // 2^32 fixed-point deposit accumulators
    float2* green;              // G^ (FFT of the softened Green's function)
    float2* sg_c;               // self-gravity: the complex grid between its row and column passes
    int* sg_rows;               // self-gravity: per row, whether any cell holds energy
    int* has_energy;
    const Ctrl* ctrl;           // non-null: time-stepping kernels take dt_sub / active from it
};

struct ParticleView {
    const float *x, *y, *q;
    float *m, *base_mass, *wx, *wy, *vx, *vy, *radius;
    const Ctrl* ctrl;
};

void field_build_green(hipStream_t s, const FieldArgs& f, float soft2);
void field_selfgrav(hipStream_t s, const FieldArgs& f);
void field_fft(hipStream_t s, const FieldArgs& f, int inverse);   // in place on f.green
void field_deposit(hipStream_t s, const FieldArgs& f, const ParticleView& p, int n_upper, int rel, int coul, int yuk);
void field_grads(hipStream_t s, const FieldArgs& f);
void field_kdk(hipStream_t s, FieldArgs f, const float* portal, int sg_on, float dt);
void field_modulate(hipStream_t s, const FieldArgs& f, const ParticleView& p, float* higgs_mod, float dt, int n_upper);
void field_axmod(hipStream_t s, const FieldArgs& f, const ParticleView& p, float* ax, float* yuk, int coul, int yukawa,
                 int n_upper);
void field_forces(hipStream_t s, const FieldArgs& h, const FieldArgs& a, int use_h, int use_a, int grav, int coul,
                  int yukawa, const ParticleView& p, float* fx, float* fy, float* pe_grav, float* pfi, int n_upper,
                  const int32_t* sel = nullptr, float* sel_out = nullptr);
void field_energy(hipStream_t s, const FieldArgs& f, const float* partner, double* out);

// ---- integrator + reductions (k_step.hip) ----
struct StepArgs {
    float *x, *y, *wx, *wy, *vx, *vy, *angw, *ang_vel, *mass, *radius, *charge;
    float *mu, *L, *f1x, *f1y, *f1ox, *f1oy;
    const float *fx, *fy, *bz, *bgz, *dbzdx, *dbzdy, *dbgzdx, *dbgzdy, *frame_drag, *tidal;
    int32_t* alive;
    Ctrl* ctrl;
    uint32_t rel, mag, gm, ext_bz_on, grav, spin_orbit, one_pn, expansion;
    uint32_t col_bounce = 0;        // collision mode bounce: the torque step always runs (integrator.js:932)
    float hubble, cx, cy;
    int n_upper;
    // Landau-Lifshitz radiation reaction (radiation and Coulomb on): the jerk
    // of the last force pass in, each particle's radiated energy and momentum
    // out (three FP64 columns of `rad_stride`)
    uint32_t rad = 0;
    const float *jx = nullptr, *jy = nullptr;
    double* rad_e = nullptr;
    int64_t rad_stride = 0;
    // the selected particle's index and force record (its radiation reaction)
    const int32_t* sel = nullptr;
    float* sel_out = nullptr;
    // boundary mode bounce: each particle's wall contact torque of the last force pass
    const float* contact = nullptr;
};
// _applyBoundaryForces (integrator.js:606-667): Hertz wall repulsion with
// friction on [0, w] x [0, h], added to the force; the contact torque written.
struct WallArgs {
    const float *x, *y, *vx, *vy, *ang_vel, *radius;
    float *fx, *fy, *contact;
    const int32_t* sel;
    float* sel_out;
    const Ctrl* ctrl;
    float w, h, friction;
    int accumulate = 0;             // 1: add to the contact torque the body contacts left
};
void launch_wall_forces(hipStream_t s, const WallArgs& a, int n_upper);
// FP64 column sums: out[c] (+)= sum of column c over the live particles;
// gate_active: nothing is added when the controller has no substep left.
void launch_sum_cols(hipStream_t s, const double* cols, int ncols, int64_t stride, int n_upper, const Ctrl* ctrl,
                     double* partial, double* out, bool accumulate, bool gate_active = false);
// Quadrupole radiation once per update (integrator.js:1567-1740).
struct QuadArgs {
    const float *x, *y, *vx, *vy, *fx, *fy, *jx, *jy, *mass, *charge;
    float *wx, *wy;
    double* scr;                // nine FP64 columns of `stride`
    int64_t stride;
    double* partial;            // block sums
    double* sums;               // [0..2] COM sums, [3..11] quadrupole sums
    double* rad_tot;            // [0] radiated energy (accumulated)
    const Ctrl* ctrl;
    int gw, em;
};
void launch_quad_radiation(hipStream_t s, const QuadArgs& q, double dt, int n_upper);
void launch_dt_scan(hipStream_t s, const StepArgs& a);
// Batch of systems: per-slot system id and liveness, per-system controllers.
struct BatchStepArgs {
    int S, total;
    const int32_t* sys;
    int32_t* valid;
    const int32_t* soff;
    Ctrl* ctrl;
};
void launch_batch_begin(hipStream_t s, Ctrl* ctrls, int S, double dt, int max_sub);
void launch_batch_ctrl(hipStream_t s, Ctrl* ctrls, int S, double softening, int first);
void launch_batch_kick_drift(hipStream_t s, const StepArgs& a, const BatchStepArgs& b);
void launch_batch_dt_scan(hipStream_t s, const StepArgs& a, const BatchStepArgs& b);
void launch_batch_boundary(hipStream_t s, const StepArgs& a, const BatchStepArgs& b, int mode, double w, double h);
void launch_batch_stats(hipStream_t s, const StepArgs& a, const BatchStepArgs& b, const float* pe, int rel,
                        double* out);
void launch_ctrl(hipStream_t s, Ctrl* ctrl, double softening, int kick_reuse);
void launch_kick_drift(hipStream_t s, const StepArgs& a);
void launch_vv_kick(hipStream_t s, const StepArgs& a);
void launch_boundary(hipStream_t s, const StepArgs& a, int mode, int topo, double w, double h);
// FP64 two-level reductions for the conservation readout; out[0..15].
void launch_stats(hipStream_t s, const StepArgs& a, const float* pe, double* partial, double* out, int rel);
// Darwin field energy and momentum (energy.js:64-142): out[0..2] = fieldEnergy,
// fieldPx, fieldPy.  partial holds 3 * ceil(n / 256) doubles.
struct DarwinArgs {
    const float *x, *y, *vx, *vy, *mass, *charge, *ax;   // ax null: 1
    const int32_t* slot;
    HistArgs hist;              // on: sources at their retarded state
    PhysFlags pf;               // periodic box
    double soft2;
    int mag, gm, em_energy, gm_energy, bazanski;
    const Ctrl* ctrl;
};
void launch_darwin(hipStream_t s, const DarwinArgs& d, int n_upper, double* partial, double* out);
void launch_sum2(hipStream_t s, const float* a, const float* b, int n_upper, const Ctrl* ctrl, double* partial,
                 double* out);

}  // namespace geon
