/*
 * geon_oracle.h — FP64 CPU restatement of the reference's CPU physics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP backend is
 * compared against (and the timed CPU baseline in bench.py).  Nothing in the
 * product path (physsim_amd/) links, loads or calls it.
 *
 * It follows src/integrator.js, src/forces.js, src/quadtree.js,
 * src/topology.js, src/relativity.js, src/energy.js and src/collisions.js of
 * a9lim/physsim; each function in geon_oracle.c cites the lines it restates.
 */
#ifndef GEON_ORACLE_H
#define GEON_ORACLE_H
#include <stdint.h>
#include "../include/geon.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrcSys OrcSys;

OrcSys* orc_create(const GeonParams* prm, int32_t capacity);
void    orc_destroy(OrcSys* s);
int     orc_set_params(OrcSys* s, const GeonParams* prm);
/* Replace the particle array (main.js:504-539 normalization applied). */
int     orc_upload(OrcSys* s, int32_t n, const double* x, const double* y,
                   const double* wx, const double* wy, const double* mass,
                   const double* charge, const double* angw);
int32_t orc_count(OrcSys* s);
/* Read state in array order.  ids[i] is the upload index of particle i. */
int     orc_download(OrcSys* s, double* x, double* y, double* wx, double* wy,
                     double* mass, double* angw, int32_t* ids);
/* Bootstrap force pass (integrator.js:750-782). */
int     orc_compute_forces(OrcSys* s);
/* Forces in array order; any pointer may be NULL. */
int     orc_get_forces(OrcSys* s, double* fx, double* fy, double* bz, double* bgz,
                       double* dbzdx, double* dbzdy, double* dbgzdx, double* dbgzdy,
                       double* frame_drag, double* tidal, double* pe_total);
/* One Physics.update call (integrator.js:726-1825). */
int     orc_update(OrcSys* s, double dt, int32_t* substeps);
int     orc_stats(OrcSys* s, GeonStats* out);
/* Tree at the current positions (quadtree.js:412-419) in canonical form. */
int64_t orc_build_tree(OrcSys* s);
int64_t orc_get_tree(OrcSys* s, GeonTreeNode* out, int64_t cap);
/* Counters of the last force pass: [pair, node, visited]. */
void    orc_counters(OrcSys* s, int64_t* out3);
/* Collision pair list (collisions.js:35-60): pairs (i<j by id) whose bodies
 * overlap, found by tree query.  Returns the number of pairs; writes at most
 * cap pairs as (a[k], b[k]) upload indices, sorted. */
int64_t orc_collision_pairs(OrcSys* s, int32_t* a, int32_t* b, int64_t cap);
/* The same pairs unsorted, in the order handleCollisions meets them: array
 * order of the first member, then the tree query order (quadtree.js:372-398)
 * over the tree with its periodic ghosts (integrator.js:247-258). */
int64_t orc_collision_pairs_ordered(OrcSys* s, int32_t* a, int32_t* b, int64_t cap);
/* Collision events of the last update in merge mode (collisions.js:62-92):
 * merges with lost kinetic energy first, then annihilations, each in the
 * order they happened, as ORC_EVENT_STRIDE doubles {type, x, y, energy, px, py}.
 * Returns the count; writes at most cap. */
#define ORC_EVENT_STRIDE 6
#define ORC_EV_MERGE 1
#define ORC_EV_ANNIHILATION 2
int64_t orc_collision_events(OrcSys* s, double* out, int64_t cap);
/* Antimatter flags in array order (particle.js:21); extra per-particle state. */
int     orc_set_antimatter(OrcSys* s, const int32_t* flags);
int     orc_download_ex(OrcSys* s, double* base_mass, double* charge, int32_t* antimatter, double* radius);
/* Leaf contents in stored order: rows (level << 56 | path, id). */
int64_t orc_leaf_points(OrcSys* s, uint64_t* key, int32_t* id, int64_t cap);
/* Analytic jerk dF/dt of the last force pass (radiation on), array order. */
int     orc_get_jerk(OrcSys* s, double* jx, double* jy);

/* Stand-alone FP64 direct sum of gravity+Coulomb+... (forces.js:93-126) over
 * `n` sources for the targets listed in `idx` (all if idx == NULL).  Uses the
 * system's toggles; min-image when the boundary is loop.  Multithreaded. */
int     orc_direct_forces(OrcSys* s, const int32_t* idx, int32_t n_idx,
                          double* fx, double* fy);
void    orc_set_threads(int32_t n);

/* Far-node samples of the learned correction (not in the reference): builds
 * the tree, then for each observer in idx lists every divided node its walk
 * accepts, as ORC_FAR_STRIDE doubles: observer, dx, dy, r^2 + eps^2, node size,
 * mass, charge, the 9 moments about the COM, and the exact monopole-truncation
 * residuals: force (gravity x, y per unit observer mass; Coulomb x, y per unit
 * observer charge) and potential (gravity, Coulomb).  Returns the number of
 * samples; writes at most cap. */
#define ORC_FAR_STRIDE 22
int64_t orc_far_samples(OrcSys* s, const int32_t* idx, int32_t n_idx, double* out, int64_t cap);
/* FP64 restatement of the learned far-field network (model layout as
 * GEON_FAR_MODEL_FLOATS in include/geon.h): set the model used when
 * far_field == GEON_FAR_LEARNED, and the training loss and its gradient on
 * samples in the GEON_FAR_SAMPLE layout. */
int     orc_far_set_model(OrcSys* s, const float* model);
int64_t orc_far_training_samples(OrcSys* s, const int32_t* idx, int32_t n_idx, double f2ref, double pe2ref,
                                 float* out, int32_t* obs, int64_t cap);
double  orc_far_loss_grad(const float* model, const float* samples, int64_t n, double wscale, double* grad);

/* Signal delay (signal-delay.js).  History of live particle i (dead = 0) or
 * of the i-th dead record (dead = 1), oldest first; returns the count.
 * set_history replaces live particle i's ring with `count` samples in time
 * order; set_time sets the simulation time; set_lifetime the creation and
 * death times; delayed_state evaluates getDelayedState for source i and the
 * observer (ox, oy) at the current time: 1 and out5 = x, y, vx, vy, angw, or
 * 0 where the reference returns null. */
int32_t orc_get_history(OrcSys* s, int32_t i, int32_t dead, double* x, double* y, double* vx, double* vy,
                        double* aw, double* t);
int     orc_set_history(OrcSys* s, int32_t i, int32_t count, const double* x, const double* y,
                        const double* vx, const double* vy, const double* aw, const double* t);
void    orc_set_time(OrcSys* s, double t);
/* minImage (topology.js:18-76) under the system's boundary and topology. */
int     orc_min_image(OrcSys* s, double ox, double oy, double sx, double sy, double* out2);
/* force1PN of the last force pass, array order. */
int     orc_get_f1pn(OrcSys* s, double* f1x, double* f1y);
int32_t orc_dead_count(OrcSys* s);
int     orc_set_lifetime(OrcSys* s, int32_t i, double creation, double death);
int     orc_delayed_state(OrcSys* s, int32_t i, double ox, double oy, double* out5);

/* Scalar fields (which: 0 Higgs, 1 axion). */
int     orc_field_set(OrcSys* s, int32_t which, const double* field, const double* field_dot);
int     orc_field_get(OrcSys* s, int32_t which, int32_t what, double* out);
int     orc_field_selfgrav(OrcSys* s, int32_t which);
int     orc_pqs_deposit(int32_t G, double W, double H, int32_t bc, int32_t n, const double* x, const double* y,
                        const double* v, double* out);
void    orc_fft2d(double* re, double* im, int32_t N, int32_t inverse);

#ifdef __cplusplus
}
#endif
#endif
