/*
 * geon.h — C ABI of the gfx950 physics backend for Geon (a9lim/physsim).
 *
 * One handle owns one particle system resident in one GPU's HBM: SoA particle
 * state, the Barnes-Hut quadtree, force accumulators and the scalar-field
 * grids.  Every entry point is plain C: pointers, sizes, int return codes.
 * No exception, no torch type and no HIP type crosses this boundary.
 *
 * The entry points replace the de-facto GPU backend contract the reference app
 * calls on `GPUPhysics` (src/gpu/gpu-physics.js) and the CPU `Physics.update`
 * call site (main.js:697); each declaration cites the method it stands in for.
 *
 * Return codes: 0 = ok, negative = error (GEON_E*).  An entry point that fails
 * writes nothing to its outputs.  A handle is not thread-safe; each handle owns
 * one HIP stream.
 */
#ifndef GEON_H
#define GEON_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GEON_ABI_VERSION 1

/* ---- return codes ---- */
#define GEON_OK           0
#define GEON_EINVAL      -1   /* bad argument */
#define GEON_EFULL       -2   /* no free particle slot (gpu-physics.js:1864 returns -1) */
#define GEON_EDEVICE     -3   /* HIP runtime error or no device */
#define GEON_ENOMEM      -4   /* device allocation failed */
#define GEON_ESTATE      -5   /* call not valid in the current state */
#define GEON_EUNSUPPORTED -6  /* toggle/mode combination this backend does not run */
#define GEON_EOVERFLOW   -7   /* a capacity exceeded: node pool, walk stack, far-field pair queue, pair lists */

/* ---- toggles0 bits: same values as src/gpu/gpu-constants.js:196-213 ---- */
#define GEON_GRAVITY        (1u << 0)
#define GEON_COULOMB        (1u << 1)
#define GEON_MAGNETIC       (1u << 2)
#define GEON_GRAVITOMAG     (1u << 3)
#define GEON_ONE_PN         (1u << 4)
#define GEON_RELATIVITY     (1u << 5)
#define GEON_SPIN_ORBIT     (1u << 6)
#define GEON_RADIATION      (1u << 7)   /* radiation reaction; rejected together with GEON_YUKAWA */
#define GEON_BLACK_HOLE     (1u << 8)
#define GEON_DISINTEGRATION (1u << 9)   /* not run by this backend: rejected */
#define GEON_EXPANSION      (1u << 10)
#define GEON_YUKAWA         (1u << 11)
#define GEON_HIGGS          (1u << 12)
#define GEON_AXION          (1u << 13)
#define GEON_BARNES_HUT     (1u << 14)
#define GEON_BOSON_INTER    (1u << 15)  /* not run by this backend: rejected */

/* ---- mode enums: src/config.js:10-26 ---- */
#define GEON_COL_PASS   0
#define GEON_COL_MERGE  1
#define GEON_COL_BOUNCE 2
#define GEON_BOUND_DESPAWN 0
#define GEON_BOUND_BOUNCE  1
#define GEON_BOUND_LOOP    2
#define GEON_TORUS 0
#define GEON_KLEIN 1
#define GEON_RP2   2

/* Periodic-image semantics (the reference's BH walk counts a source once per
 * inserted ghost image, src/integrator.js:247-258 with forces.js:633-635).
 * MIN_IMAGE: no ghosts, every displacement minimum-imaged, each source counted
 * once (what the reference's own direct sum does, forces.js:93-126). */
#define GEON_PERIODIC_MIN_IMAGE 0
#define GEON_PERIODIC_GHOSTS    1
/* WINDOWED: as MIN_IMAGE, but a divided node is approximated only when its
 * whole box lies inside the observer's min-image window, so every particle of
 * the node has the image its COM has and the walk converges to the min-image
 * direct sum as theta -> 0 (with MIN_IMAGE it does not: a node straddling the
 * window edge is counted at its COM's image). */
#define GEON_PERIODIC_WINDOWED  2

typedef struct GeonContext* GeonHandle;

/* Every constant the reference hard-codes (src/config.js) as a field, so that
 * configurations that need other values can set them.  geon_default_params
 * fills the reference defaults. */
typedef struct GeonParams {
    uint32_t toggles0;          /* GEON_GRAVITY | ... */
    int32_t  boundary_mode;     /* GEON_BOUND_* */
    int32_t  topology;          /* GEON_TORUS, GEON_KLEIN, GEON_RP2 (loop boundary, topology.js) */
    int32_t  collision_mode;    /* GEON_COL_PASS, or GEON_COL_MERGE on open and torus domains */
    int32_t  periodic_mode;     /* GEON_PERIODIC_* */
    int32_t  max_substeps;      /* MAX_SUBSTEPS = 32 (config.js:36) */
    double   domain_w, domain_h;/* width/height passed to Physics.update */
    double   softening;         /* SOFTENING = 8 (config.js:44) */
    double   bh_softening;      /* BH_SOFTENING = 4 (config.js:46) */
    double   theta;             /* BH_THETA = 0.5 (config.js:57) */
    double   yukawa_coupling;   /* YUKAWA_COUPLING = 14 (config.js:78) */
    double   yukawa_mu;         /* DEFAULT_PION_MASS = 0.15 (config.js:79) */
    double   higgs_mass;        /* DEFAULT_HIGGS_MASS = 0.5 (config.js:88) */
    double   axion_mass;        /* DEFAULT_AXION_MASS = 0.05 (config.js:95) */
    double   hubble;            /* DEFAULT_HUBBLE = 0.001 (config.js:131) */
    double   ext_gravity, ext_gravity_angle;   /* integrator.js:83-87 */
    double   ext_electric, ext_electric_angle;
    double   ext_bz;
    int32_t  field_grid;        /* scalar-field grid side: 64 CPU, 128 GPU (config.js:84,125) */
    int32_t  kahan;             /* 1: compensated FP32 force accumulation */
    int32_t  far_field;         /* GEON_FAR_*: correction added to accepted nodes */
    int32_t  far_list_cap;      /* LEARNED: listed (particle, node) pairs per 64 particles, one queue of
                                   far_list_cap / 64 per particle (rounded up to 4); 0 = 65536 */
    int32_t  signal_delay;      /* GEON_DELAY_*: FOLLOW couples it to GEON_RELATIVITY (forces.js:85) */
    int32_t  history_stride;    /* HISTORY_STRIDE = 64 updates per history sample (config.js:122); 0 = 64 */
    double   bounce_friction;   /* bounceFriction = 0.4 (integrator.js:72): tangential friction of the
                                   Hertz wall contacts in boundary mode bounce */
} GeonParams;

/* Signal delay (signal-delay.js:16-221): every leaf source acts from the point
 * where the observer's past light cone crosses its recorded trajectory, a
 * 256-sample history per particle taken every history_stride updates
 * (integrator.js:1549-1565).  A source with fewer than 2 samples exerts no
 * leaf force (forces.js:652-655); despawned and removed particles keep acting
 * from their history until 2 sqrt(W^2 + H^2) time units after their death
 * (forces.js:128-153; main.js:875-885).  The relativistic aberration factor
 * (1 - n.v)^-3 scales the delayed pair's position-dependent forces
 * (forces.js:182-189).  Accepted nodes use current-time aggregates, as in
 * the reference.  FOLLOW (the reference) turns it on exactly when
 * GEON_RELATIVITY is on. */
#define GEON_DELAY_FOLLOW 0
#define GEON_DELAY_OFF    1
#define GEON_DELAY_ON     2

/* Far-field correction of accepted Barnes-Hut nodes (not in the reference,
 * whose nodes carry monopoles only, quadtree.js:272-309): NONE reproduces the
 * reference; MULTIPOLE adds the mass-quadrupole and charge dipole+quadrupole
 * terms of gravity and Coulomb; LEARNED adds the gravity and Coulomb residual
 * predicted for every (particle, accepted node) pair by the 2 x 64 hard-swish
 * network set with geon_far_set_model or trained with geon_far_train. */
#define GEON_FAR_NONE      0
#define GEON_FAR_MULTIPOLE 1
#define GEON_FAR_LEARNED   2

/* addParticle argument (gpu-physics.js:1857): vx/vy are proper velocity w. */
typedef struct GeonParticleInit {
    double x, y, wx, wy, mass, charge, angw;
    int32_t antimatter;
} GeonParticleInit;

/* Caller-owned host arrays, one entry per slot.  Any pointer may be NULL on
 * download (that field is skipped); on upload x,y,wx,wy,mass,charge must be
 * given, the rest default (base_mass = mass, angw = 0). */
typedef struct GeonSoAView {
    double *x, *y, *wx, *wy, *mass, *base_mass, *charge, *angw;
    int32_t *alive;             /* download only: 1 alive, 0 despawned/removed */
    int32_t *antimatter;        /* 1 antimatter (particle.js:21); NULL on upload: all matter */
} GeonSoAView;

/* Per-particle results of the last force pass, slot order (any may be NULL).
 * Totals and per-type display vectors as on `Particle` (particle.js:24-36). */
typedef struct GeonForceView {
    double *fx, *fy;
    double *bz, *dbzdx, *dbzdy, *bgz, *dbgzdx, *dbgzdy;
    double *frame_drag_torque, *tidal_torque;
    double *pe;                 /* per-particle PE contribution (sum over sources) */
    double *f1pn_x, *f1pn_y;    /* force1PN after the last 1PN pass */
    double *jerk_x, *jerk_y;    /* analytic dF/dt (radiation on, forces.js:198-402) */
} GeonForceView;

/* FP64 conservation readout: same fields as GPUPhysics.readStats
 * (gpu-physics.js:4146-4281) / computeEnergies (energy.js:11-195). */
typedef struct GeonStats {
    double linear_ke, spin_ke, px, py, orbital_ang_mom, spin_ang_mom;
    double com_x, com_y, total_mass, pe;
    double field_energy, higgs_field_energy, axion_field_energy, pfi_energy;
    double field_px, field_py;
    double sim_time;
    int32_t alive_count;
    int32_t dead_count;         /* retired particles still acting through signal delay */
    /* sim.totalRadiated, totalRadiatedPx/Py (radiation on): energy and momentum
     * taken from the particles by radiation reaction since the last upload */
    double radiated_energy, radiated_px, radiated_py;
} GeonStats;

/* Work counters of the last force pass, for roofline accounting. */
typedef struct GeonCounters {
    int64_t pair_interactions;  /* particle-source (leaf point / direct) pairs */
    int64_t node_interactions;  /* particle-aggregate (accepted node) pairs */
    int64_t nodes_visited;
    int64_t tree_nodes;         /* nodes in the last built tree */
    int64_t tree_divided;       /* divided (internal) nodes */
    int64_t depth_overflow;     /* leaves at the key's deepest level holding > 4 points */
    int64_t substeps_total;     /* substeps run since create/reset */
} GeonCounters;

/* Canonical tree export: one record per node, any order.  `path` packs the
 * quadrant digits root->node (2 bits each, NW=0 NE=1 SW=2 SE=3, first digit
 * most significant), `level` is the depth (root = 0). */
typedef struct GeonTreeNode {
    uint64_t path;
    int32_t  level;
    int32_t  divided;           /* 1 internal, 0 leaf */
    int32_t  count;             /* points in the subtree */
    int32_t  charge_quanta;     /* sum of round(q/0.1) over the subtree (exact) */
    double   mass, com_x, com_y;/* FP64 aggregates (quadtree.js:239-312) */
} GeonTreeNode;

int  geon_abi_version(void);

/* Reference defaults of every constant (src/config.js). */
void geon_default_params(GeonParams* p);

/* GPUPhysics constructor + init() (gpu-physics.js, main.js:31-63). */
int  geon_create(const GeonParams* params, int32_t max_particles, GeonHandle* out);
/* GPUPhysics.destroy() (gpu-physics.js:4291). */
void geon_destroy(GeonHandle h);
/* GPUPhysics.reset() (gpu-physics.js:3973): drop all particles, sim time 0. */
int  geon_reset(GeonHandle h);
/* setToggles(physics) + the writable boundaryMode/topologyMode/domainW/H
 * fields (gpu-physics.js:1969-2056). */
int  geon_set_params(GeonHandle h, const GeonParams* params);
int  geon_get_params(GeonHandle h, GeonParams* out);

/* addParticle({x,y,vx,vy,mass,charge,angw,antimatter}) -> slot | GEON_EFULL
 * (gpu-physics.js:1857-1926).  Charge is quantized to multiples of 0.1 (1874). */
int32_t geon_add_particle(GeonHandle h, const GeonParticleInit* p);
/* removeParticle(idx, deathMass, deathAngVel) (gpu-physics.js:1936-1946).
 * With signal delay on, the particle stays a retarded source: death_mass is
 * its mass as one (the reference's _deathMass) and death_ang_vel its angular
 * velocity in the final history sample; 0 keeps the particle's own value, as
 * the reference's callers pass `p._deathMass || p.mass` (input.js:183).
 * Negative or non-finite mass, or |w| r >= 1, is GEON_EINVAL.  Without signal
 * delay the particle is simply removed (integrator.js:154). */
int  geon_remove_particle(GeonHandle h, int32_t slot, double death_mass, double death_ang_vel);
/* patchMass(idx, mass, baseMass) (gpu-physics.js:1955-1963). */
int  geon_patch_mass(GeonHandle h, int32_t slot, double mass, double base_mass);

/* Bulk state transfer (the SoA the reference's serialize/deserialize read and
 * write, gpu-physics.js:3749-3836, 3886).  Upload replaces the whole system:
 * slots 0..n-1, every given array n entries long.  Download fills slots
 * 0..n_slots-1 of arrays `cap` entries long and writes n_slots; when
 * cap < n_slots it writes only *n_slots and returns GEON_EINVAL, so a caller
 * can size its arrays with a first call of cap 0. */
int  geon_upload_soa(GeonHandle h, const GeonSoAView* v, int32_t n);
int  geon_download_soa(GeonHandle h, GeonSoAView* v, int32_t cap, int32_t* n_slots);

/* One Physics.update(particles, dt, ...) call (integrator.js:726-1825):
 * adaptive substeps (<= max_substeps), boundary handling once at the end. */
int  geon_update(GeonHandle h, double dt, int32_t* substeps_taken);

/* The force pass alone, at the current state (integrator.js:750-782 bootstrap:
 * computeAllForces + external fields).  For parity checks. */
int  geon_compute_forces(GeonHandle h);
int  geon_get_forces(GeonHandle h, GeonForceView* v, int32_t n_slots);

/* One slot's signal-delay history, oldest sample first (a live particle or
 * one retired within the last 2 sqrt(W^2 + H^2) time units): position,
 * coordinate velocity, angw and time, as Particle.histX/Y/Vx/Vy/AngW/Time
 * hold them (particle.js:77-85).  Up to 256 samples; the count goes to
 * *count.  No counterpart in the reference's interface: for inspection and
 * parity checks. */
typedef struct GeonHistoryView {
    double *x, *y, *vx, *vy, *angw, *t;
} GeonHistoryView;
int  geon_get_history(GeonHandle h, int32_t slot, GeonHistoryView* out, int32_t* count);

/* getParticleCount() (gpu-physics.js:4287): live particles, after the work
 * queued on the handle's stream has finished. */
int  geon_particle_count(GeonHandle h, int32_t* alive);

/* requestStats + readStats (gpu-physics.js:4146-4281), in FP64. */
int  geon_get_stats(GeonHandle h, GeonStats* out);

/* The selected particle of readStats (gpu-physics.js:4238-4276): its state
 * and its force by type.  requestStats(selectedGpuIdx) (4146-4151) names it:
 * geon_select(h, slot), -1 for none.  The force passes from then on record the
 * selected particle's forces by type, as the reference keeps them for every
 * particle (forces.js:36-46) and rebuilds them for display after the substeps
 * (integrator.js:1755-1796); forces_valid is 0 until one has run.  Only that
 * particle's are formed: the pass repeats its walk (or direct sum) with the
 * terms summed by type, and the field, external and radiation-reaction
 * kernels write its shares.  GEON_SF_FAR_FIELD has no reference counterpart:
 * the learned far-field correction (GEON_FAR_LEARNED), 0 otherwise. */
#define GEON_SF_GRAVITY     0   /* forceGravity (+ the fields' gravity, scalar-field.js:717) */
#define GEON_SF_COULOMB     1
#define GEON_SF_MAGNETIC    2   /* dipole force + q v x Bz */
#define GEON_SF_GRAVITOMAG  3   /* GM dipole force + 4 m v x Bgz */
#define GEON_SF_ONE_PN      4   /* EIH, Darwin, Bazanski, scalar Breit (the last 1PN evaluation) */
#define GEON_SF_SPIN_CURV   5   /* Stern-Gerlach and Mathisson-Papapetrou */
#define GEON_SF_RADIATION   6   /* Landau-Lifshitz force of the last substep */
#define GEON_SF_YUKAWA      7
#define GEON_SF_EXTERNAL    8   /* uniform gravity and electric field */
#define GEON_SF_HIGGS       9
#define GEON_SF_AXION      10
#define GEON_SF_FAR_FIELD  11
#define GEON_SF_COUNT      12
typedef struct GeonSelected {
    int32_t slot;               /* the selected slot; -1: none selected, or that particle is gone */
    int32_t forces_valid;
    int32_t antimatter;
    int32_t pad;
    double x, y, wx, wy, mass, charge, angw, base_mass, radius, vx, vy, ang_vel, mag_moment, ang_momentum;
    double force[GEON_SF_COUNT][2];
} GeonSelected;
int  geon_select(GeonHandle h, int32_t slot);
int  geon_get_selected(GeonHandle h, GeonSelected* out);

/* Build the quadtree at the current positions and export it (quadtree.js:412-419). */
int  geon_build_tree(GeonHandle h);
int  geon_get_tree(GeonHandle h, GeonTreeNode* out, int64_t cap, int64_t* n_nodes);

/* ---- scalar fields (Higgs, axion) on the field_grid^2 grid ---- */
#define GEON_FIELD_HIGGS 0
#define GEON_FIELD_AXION 1
/* quantity selectors for geon_get_field */
#define GEON_FQ_FIELD     0   /* phi / a */
#define GEON_FQ_FIELD_DOT 1
#define GEON_FQ_GRAD_X    2   /* central differences, grid units */
#define GEON_FQ_GRAD_Y    3
#define GEON_FQ_SG_PHI    4   /* self-gravity potential */
#define GEON_FQ_SG_GX     5
#define GEON_FQ_SG_GY     6
#define GEON_FQ_SOURCE    7   /* last PQS deposit */
#define GEON_FQ_THERMAL   8   /* Higgs only: deposited kinetic energy */

/* Set a field's value and time derivative (n = grid^2 row-major); the grid
 * gradients are recomputed.  The reference has no such call (its fields start
 * at the vacuum, higgs-field.js:25-27, axion-field.js:39-41); this is how
 * seeded initial fields enter.  Either pointer may be NULL (left unchanged). */
int  geon_set_field(GeonHandle h, int32_t which, const double* field, const double* field_dot, int32_t n);
/* getFieldBuffers / readbackFieldData (gpu-physics.js:3307, 3844). */
int  geon_get_field(GeonHandle h, int32_t which, int32_t what, double* out, int32_t n);
/* computeSelfGravity on the current field state (scalar-field.js:600-660). */
int  geon_field_selfgravity(GeonHandle h, int32_t which);
/* The self-gravity kernel's 2-D FFT applied to a caller grid in place
 * (n x n, n a power of two <= 128; inverse scaled by 1/n per axis as fft.js:63-69). */
int  geon_fft2d(GeonHandle h, double* re, double* im, int32_t n, int32_t inverse);

/* Collision broadphase pair list (collisions.js:35-60, quadtree.js:372-404):
 * all (a, b) with slot a < slot b whose bodies overlap, found by a box query of
 * half-size 2 r_a around each particle, as the reference finds them.  Builds
 * the tree at the current positions.  Writes min(cap, total) pairs sorted by
 * (a, b) and the total to *n_pairs (a, b may be NULL to count only). */
int  geon_collision_pairs(GeonHandle h, int32_t* a, int32_t* b, int64_t cap, int64_t* n_pairs);

/* The same pair list with no device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its staging and sort
 * buffers are one fixed device array per GPU (about 33 MB, held until the
 * process ends), so a state with more than 2^20 overlapping pairs fails with
 * GEON_EOVERFLOW and writes nothing; calls on one GPU that overlap in time
 * (from several threads) share that array and race.  The pairs are sorted on
 * the device; output as geon_collision_pairs. */
int  geon_collision_pairs_noalloc(GeonHandle h, int32_t* a, int32_t* b, int64_t cap, int64_t* n_pairs);

/* Collision events of the last geon_update in collision mode merge
 * (collisions.js:62-92, integrator.js:1388-1415): merges that lost kinetic
 * energy (the field excitations' source), then annihilations (the photon
 * bursts' source), each in the order the fold met them, as rows of
 * GEON_EVENT_STRIDE doubles {type, x, y, energy, px, py}; type 1 merge, 2
 * annihilation.  Photons are not simulated by this backend: a host that
 * keeps bosons emits the bursts from these rows (emitPhotonBurst).  Writes
 * min(cap, total) rows and the total to *n_events. */
#define GEON_EVENT_STRIDE 6
#define GEON_EVENT_MERGE 1
#define GEON_EVENT_ANNIHILATION 2
int  geon_collision_events(GeonHandle h, double* out, int64_t cap, int64_t* n_events);

int  geon_counters(GeonHandle h, GeonCounters* out);

/* ---- Learned far-field correction (no counterpart in the reference) -------
 * Model: GEON_FAR_MODEL_FLOATS floats = GEON_FAR_PARAMS FP32 parameters
 * (W1[16][64] whose row 10 is the first layer's bias, W2[64][64], b2[64],
 * W3[64][6], b3[6]) followed by the feature means[10], inverse standard
 * deviations[10] and output scales[6].  The six outputs are the gravity and
 * Coulomb force residuals (two components each, in the offset frame) and the
 * gravity and Coulomb potential residuals of one (particle, node) pair; the
 * inputs and outputs are defined in physsim_amd/csrc/k_learn.hip. */
#define GEON_FAR_PARAMS       5574
#define GEON_FAR_MODEL_FLOATS 5600
#define GEON_FAR_SAMPLE       20    /* floats per sample: 10 features, 6 targets, 4 weights */
/* Training samples from the current state: builds the tree, lists the pairs
 * the walk accepts, and for every group_stride-th group of 64 particles (in
 * Morton order) appends one sample per listed pair with its exact FP64
 * monopole-truncation residual.  *n_samples = samples held afterwards. */
int  geon_far_collect(GeonHandle h, int32_t group_stride, int64_t* n_samples);
int  geon_far_clear(GeonHandle h);
/* Samples [first, first + n): out (n * GEON_FAR_SAMPLE floats) and/or the
 * observer slot of each (either may be NULL). */
int  geon_far_get_samples(GeonHandle h, int64_t first, int64_t n, float* out, int32_t* observer);
/* Adam on random batches of the held samples, learning rate cosine-decayed
 * from lr to lr_final; loss_out (may be NULL) receives `steps` losses.  The
 * first call with no model initialises the weights from seed and the input /
 * output statistics from the samples. */
int  geon_far_train(GeonHandle h, int32_t steps, int32_t batch, double lr, double lr_final, uint64_t seed,
                    float* loss_out);
/* Loss gradient (GEON_FAR_PARAMS floats) on the given samples at the current
 * model, with the same BF16 kernels as training. */
int  geon_far_grad(GeonHandle h, const int64_t* rows, int32_t n_rows, float* grad_out, float* loss_out);
int  geon_far_get_model(GeonHandle h, float* out, int32_t n);
int  geon_far_set_model(GeonHandle h, const float* in, int32_t n);
/* Kernel timing: average device duration (ms) over the launches of `kernel`
 * since the last reset, measured with HIP events on the handle's stream.
 * kernel: 0 = force walk / direct sum, 1 = tree build, 2 = sort, 3 = integrator,
 * 4 = far-field network pass, 5 = field deposits, 6 = field KDK steps (with
 * their self-gravity FFTs), 7 = field -> particle kernels. */
int  geon_timing_enable(GeonHandle h, int32_t on);
int  geon_kernel_timing(GeonHandle h, int32_t kernel, double* total_ms, int64_t* launches);

/* Last error message of this handle (or of the last failed create when h is NULL). */
const char* geon_last_error(GeonHandle h);


/* ---- Batched ensembles (no counterpart in the reference) -----------------
 * Many independent systems that share one GeonParams, each with the semantics
 * of one Physics.update per call (its own adaptive substeps, despawn once per
 * update), stepped together: every substep of all systems is one pass of the
 * kernels, replayed from captured graphs of four substep bodies until no
 * system can take another substep.
 * A system's results do not depend on the other systems of the batch.  Runs
 * gravity, Coulomb, magnetic, gravitomagnetic, Yukawa, relativity and
 * spin-orbit with the Barnes-Hut walk (GEON_EUNSUPPORTED otherwise). */
typedef struct GeonBatch_* GeonBatchHandle;
typedef struct GeonBatchTune {
    int32_t walk_threads;       /* persistent walk block size: 64, 128 or 256 */
    int32_t waves_per_cu;       /* persistent walk grid: resident waves per compute unit */
    int32_t stack;              /* walk node-stack entries per wave: 64 or 96 */
    int32_t use_graph;          /* 1: replay the substep loop from captured graphs */
    int32_t graph_bodies;       /* substep bodies per captured graph (even, 2..32; 0 = 2).  Systems that
                                   finished their update sit predicated off inside a graph.  The graphs
                                   the last update's longest system needed go out back to back; then
                                   the host checks for a system left between graphs.  graph_bodies >=
                                   max_substeps runs each update as one graph launch */
    double  walk_ms;            /* geon_batch_autotune: the chosen configuration's walk time */
} GeonBatchTune;
int  geon_batch_create(const GeonParams* params, int32_t n_systems, const int32_t* max_particles,
                       GeonBatchHandle* out);
void geon_batch_destroy(GeonBatchHandle h);
/* One system's particles, as geon_upload_soa. */
int  geon_batch_upload(GeonBatchHandle h, int32_t system, const GeonSoAView* v, int32_t n);
int  geon_batch_download(GeonBatchHandle h, int32_t system, GeonSoAView* v, int32_t cap, int32_t* n_slots);
/* One Physics.update(dt) of every system; substeps (n_systems entries, may be
 * NULL) receives each system's substep count. */
int  geon_batch_update(GeonBatchHandle h, double dt, int32_t* substeps);
int  geon_batch_compute_forces(GeonBatchHandle h);
int  geon_batch_get_forces(GeonBatchHandle h, int32_t system, GeonForceView* v, int32_t n_slots);
int  geon_batch_get_stats(GeonBatchHandle h, int32_t system, GeonStats* out);
/* Tile and stack sizes of the persistent walk: autotune times every candidate
 * on the current state and keeps the fastest. */
int  geon_batch_autotune(GeonBatchHandle h, GeonBatchTune* out);
int  geon_batch_set_tune(GeonBatchHandle h, const GeonBatchTune* t);
int  geon_batch_get_tune(GeonBatchHandle h, GeonBatchTune* out);
const char* geon_batch_last_error(GeonBatchHandle h);

#ifdef __cplusplus
}
#endif
#endif /* GEON_H */
