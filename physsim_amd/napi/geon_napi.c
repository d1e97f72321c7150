/*
 * geon_napi.c — Node-API addon over the C ABI of libgeon_hip.so.
 *
 * It gives the reference's JavaScript host (main.js, which drives a backend
 * object, main.js:31-63 and 613-692) the entry points of include/geon.h as
 * plain JS functions; physsim_amd/js/hip-physics.mjs wraps them in a class
 * with the GPUPhysics method surface.  Handles are JS externals with a
 * finalizer; errors are thrown as JS Errors carrying geon_last_error.
 *
 * Build (no node-gyp: only the Node-API headers are needed, the napi_*
 * symbols resolve from the node binary at load time):
 *   gcc -O2 -shared -fPIC -I/usr/include/node geon_napi.c -L.. -lgeon_hip \
 *       -Wl,-rpath,'$ORIGIN/..' -o geon_napi.node
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/geon.h"

#define CHECK(call) do { if ((call) != napi_ok) { napi_throw_error(env, NULL, "napi call failed: " #call); return NULL; } } while (0)

static napi_value throw_geon(napi_env env, GeonHandle h, int rc) {
    char msg[512];
    snprintf(msg, sizeof(msg), "geon error %d: %s", rc, geon_last_error(h));
    napi_throw_error(env, NULL, msg);
    return NULL;
}

static void finalize_handle(napi_env env, void* data, void* hint) {
    (void)env; (void)hint;
    GeonHandle* box = (GeonHandle*)data;
    if (box && *box) geon_destroy(*box);
    free(box);
}

static GeonHandle get_handle(napi_env env, napi_value v) {
    void* data = NULL;
    if (napi_get_value_external(env, v, &data) != napi_ok || !data) return NULL;
    return *(GeonHandle*)data;
}

static double get_num(napi_env env, napi_value obj, const char* key, double dflt) {
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
    napi_value v;
    napi_get_named_property(env, obj, key, &v);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_boolean) { bool b; napi_get_value_bool(env, v, &b); return b ? 1.0 : 0.0; }
    if (t != napi_number) return dflt;
    double d;
    napi_get_value_double(env, v, &d);
    return d;
}

/* params object -> GeonParams (missing keys keep the reference defaults). */
static void read_params(napi_env env, napi_value obj, GeonParams* p) {
    geon_default_params(p);
    p->toggles0 = (uint32_t)get_num(env, obj, "toggles0", p->toggles0);
    p->boundary_mode = (int32_t)get_num(env, obj, "boundaryMode", p->boundary_mode);
    p->topology = (int32_t)get_num(env, obj, "topology", p->topology);
    p->collision_mode = (int32_t)get_num(env, obj, "collisionMode", p->collision_mode);
    p->periodic_mode = (int32_t)get_num(env, obj, "periodicMode", p->periodic_mode);
    p->max_substeps = (int32_t)get_num(env, obj, "maxSubsteps", p->max_substeps);
    p->domain_w = get_num(env, obj, "domainW", p->domain_w);
    p->domain_h = get_num(env, obj, "domainH", p->domain_h);
    p->softening = get_num(env, obj, "softening", p->softening);
    p->bh_softening = get_num(env, obj, "bhSoftening", p->bh_softening);
    p->theta = get_num(env, obj, "theta", p->theta);
    p->yukawa_coupling = get_num(env, obj, "yukawaCoupling", p->yukawa_coupling);
    p->yukawa_mu = get_num(env, obj, "yukawaMu", p->yukawa_mu);
    p->higgs_mass = get_num(env, obj, "higgsMass", p->higgs_mass);
    p->axion_mass = get_num(env, obj, "axionMass", p->axion_mass);
    p->hubble = get_num(env, obj, "hubbleParam", p->hubble);
    p->ext_gravity = get_num(env, obj, "extGravity", p->ext_gravity);
    p->ext_gravity_angle = get_num(env, obj, "extGravityAngle", p->ext_gravity_angle);
    p->ext_electric = get_num(env, obj, "extElectric", p->ext_electric);
    p->ext_electric_angle = get_num(env, obj, "extElectricAngle", p->ext_electric_angle);
    p->ext_bz = get_num(env, obj, "extBz", p->ext_bz);
    p->field_grid = (int32_t)get_num(env, obj, "fieldGrid", p->field_grid);
    p->kahan = (int32_t)get_num(env, obj, "kahan", p->kahan);
    p->far_field = (int32_t)get_num(env, obj, "farField", p->far_field);
    p->far_list_cap = (int32_t)get_num(env, obj, "farListCap", p->far_list_cap);
    p->signal_delay = (int32_t)get_num(env, obj, "signalDelay", p->signal_delay);
    p->history_stride = (int32_t)get_num(env, obj, "historyStride", p->history_stride);
    p->bounce_friction = get_num(env, obj, "bounceFriction", p->bounce_friction);
}

static napi_value args_n(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
    size_t argc = want;
    napi_get_cb_info(env, info, &argc, argv, NULL, NULL);
    if (argc < want) { napi_throw_type_error(env, NULL, "too few arguments"); return NULL; }
    return argv[0];
}

static napi_value num(napi_env env, double d) { napi_value v; napi_create_double(env, d, &v); return v; }
static napi_value i32(napi_env env, int32_t d) { napi_value v; napi_create_int32(env, d, &v); return v; }

/* create(params, maxParticles) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonParams p;
    read_params(env, argv[0], &p);
    int32_t cap;
    CHECK(napi_get_value_int32(env, argv[1], &cap));
    GeonHandle h = NULL;
    int rc = geon_create(&p, cap, &h);
    if (rc) return throw_geon(env, NULL, rc);
    GeonHandle* box = (GeonHandle*)malloc(sizeof(GeonHandle));
    *box = h;
    napi_value ext;
    CHECK(napi_create_external(env, box, finalize_handle, NULL, &ext));
    return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    void* data = NULL;
    CHECK(napi_get_value_external(env, argv[0], &data));
    GeonHandle* box = (GeonHandle*)data;
    if (box && *box) { geon_destroy(*box); *box = NULL; }
    return NULL;
}

static napi_value js_set_params(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    GeonParams p;
    read_params(env, argv[1], &p);
    int rc = geon_set_params(h, &p);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

static napi_value js_reset(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int rc = geon_reset(h);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

/* addParticle(h, {x, y, vx, vy, mass, charge, angw, antimatter}) -> slot | -1 */
static napi_value js_add_particle(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    GeonParticleInit pi;
    pi.x = get_num(env, argv[1], "x", 0);
    pi.y = get_num(env, argv[1], "y", 0);
    pi.wx = get_num(env, argv[1], "vx", 0);
    pi.wy = get_num(env, argv[1], "vy", 0);
    pi.mass = get_num(env, argv[1], "mass", 1);
    pi.charge = get_num(env, argv[1], "charge", 0);
    pi.angw = get_num(env, argv[1], "angw", 0);
    pi.antimatter = (int32_t)get_num(env, argv[1], "antimatter", 0);
    int32_t rc = geon_add_particle(h, &pi);
    if (rc == GEON_EFULL) return i32(env, -1);
    if (rc < 0) return throw_geon(env, h, rc);
    return i32(env, rc);
}

static napi_value js_remove_particle(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!args_n(env, info, 4, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int32_t slot;
    double death_mass, death_ang_vel;
    CHECK(napi_get_value_int32(env, argv[1], &slot));
    CHECK(napi_get_value_double(env, argv[2], &death_mass));
    CHECK(napi_get_value_double(env, argv[3], &death_ang_vel));
    int rc = geon_remove_particle(h, slot, death_mass, death_ang_vel);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

static napi_value js_patch_mass(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!args_n(env, info, 4, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int32_t slot;
    double m, bm;
    CHECK(napi_get_value_int32(env, argv[1], &slot));
    CHECK(napi_get_value_double(env, argv[2], &m));
    CHECK(napi_get_value_double(env, argv[3], &bm));
    int rc = geon_patch_mass(h, slot, m, bm);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

/* A named typed-array column of `obj`: 0 absent (or undefined), 1 present with
 * element type `want` (data and length filled), -1 present but of another kind. */
static int column_of(napi_env env, napi_value obj, const char* key, napi_typedarray_type want, void** data,
                     size_t* len) {
    bool has = false;
    *data = NULL;
    *len = 0;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
    napi_value v;
    napi_get_named_property(env, obj, key, &v);
    napi_valuetype vt;
    napi_typeof(env, v, &vt);
    if (vt == napi_undefined) return 0;
    bool is_ta = false;
    napi_is_typedarray(env, v, &is_ta);
    if (!is_ta) return -1;
    napi_typedarray_type type;
    napi_value ab;
    size_t off;
    napi_get_typedarray_info(env, v, &type, len, data, &ab, &off);
    return type == want ? 1 : -1;
}

/* uploadSoA(h, {x, y, wx, wy, mass, charge, angw?, baseMass?} Float64Arrays, antimatter? Int32Array).
 * Every column given must be a typed array of x's length: the C ABI reads n
 * entries from each. */
static napi_value js_upload(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    GeonSoAView v;
    memset(&v, 0, sizeof(v));
    const char* names[] = {"x", "y", "wx", "wy", "mass", "charge", "angw", "baseMass"};
    double** ptrs[] = {&v.x, &v.y, &v.wx, &v.wy, &v.mass, &v.charge, &v.angw, &v.base_mass};
    size_t n = 0;
    char msg[160];
    for (int i = 0; i < 9; i++) {
        const int required = i < 6;
        const char* key = i < 8 ? names[i] : "antimatter";
        void* data;
        size_t len;
        const int r = column_of(env, argv[1], key, i < 8 ? napi_float64_array : napi_int32_array, &data, &len);
        if (r < 0 || (r == 0 && required)) {
            snprintf(msg, sizeof(msg), "uploadSoA: column '%s' must be %s%s", key,
                     i < 8 ? "a Float64Array" : "an Int32Array", required ? "" : " when given");
            napi_throw_type_error(env, NULL, msg);
            return NULL;
        }
        if (i == 0) n = len;
        if (r == 1 && len != n) {
            snprintf(msg, sizeof(msg), "uploadSoA: column '%s' has %zu entries, x has %zu", key, len, n);
            napi_throw_range_error(env, NULL, msg);
            return NULL;
        }
        if (i < 8) *ptrs[i] = (double*)data;
        else v.antimatter = (int32_t*)data;
    }
    if (n > 0x7fffffff) { napi_throw_range_error(env, NULL, "uploadSoA: too many particles"); return NULL; }
    int rc = geon_upload_soa(h, &v, (int32_t)n);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

static napi_value new_f64(napi_env env, size_t n, double** data) {
    napi_value ab, ta;
    napi_create_arraybuffer(env, n * sizeof(double), (void**)data, &ab);
    napi_create_typedarray(env, napi_float64_array, n, ab, 0, &ta);
    return ta;
}

/* downloadSoA(h) -> {x, y, wx, wy, mass, baseMass, charge, angw, alive, antimatter, nSlots}: arrays of
 * one entry per slot in use. */
static napi_value js_download(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    GeonSoAView v;
    memset(&v, 0, sizeof(v));
    int32_t cap = 0;
    int rc = geon_download_soa(h, &v, 0, &cap);     /* sizes the arrays: writes only the slot count */
    if (rc && rc != GEON_EINVAL) return throw_geon(env, h, rc);
    napi_value out;
    napi_create_object(env, &out);
    const char* names[] = {"x", "y", "wx", "wy", "mass", "baseMass", "charge", "angw"};
    double** ptrs[] = {&v.x, &v.y, &v.wx, &v.wy, &v.mass, &v.base_mass, &v.charge, &v.angw};
    for (int i = 0; i < 8; i++) napi_set_named_property(env, out, names[i], new_f64(env, (size_t)cap, ptrs[i]));
    napi_value ab, ta;
    napi_create_arraybuffer(env, (size_t)cap * 4, (void**)&v.alive, &ab);
    napi_create_typedarray(env, napi_int32_array, (size_t)cap, ab, 0, &ta);
    napi_set_named_property(env, out, "alive", ta);
    napi_create_arraybuffer(env, (size_t)cap * 4, (void**)&v.antimatter, &ab);
    napi_create_typedarray(env, napi_int32_array, (size_t)cap, ab, 0, &ta);
    napi_set_named_property(env, out, "antimatter", ta);
    int32_t ns = 0;
    rc = geon_download_soa(h, &v, cap, &ns);
    if (rc) return throw_geon(env, h, rc);
    napi_set_named_property(env, out, "nSlots", i32(env, ns));
    return out;
}

/* update(h, dt) -> substeps taken */
static napi_value js_update(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    double dt;
    CHECK(napi_get_value_double(env, argv[1], &dt));
    int32_t steps = 0;
    int rc = geon_update(h, dt, &steps);
    if (rc) return throw_geon(env, h, rc);
    return i32(env, steps);
}

static napi_value js_compute_forces(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int rc = geon_compute_forces(h);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

/* getForces(h, nSlots) -> {fx, fy, bz, pe} */
static napi_value js_get_forces(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int32_t n;
    CHECK(napi_get_value_int32(env, argv[1], &n));
    GeonForceView v;
    memset(&v, 0, sizeof(v));
    napi_value out;
    napi_create_object(env, &out);
    napi_set_named_property(env, out, "fx", new_f64(env, (size_t)n, &v.fx));
    napi_set_named_property(env, out, "fy", new_f64(env, (size_t)n, &v.fy));
    napi_set_named_property(env, out, "bz", new_f64(env, (size_t)n, &v.bz));
    napi_set_named_property(env, out, "pe", new_f64(env, (size_t)n, &v.pe));
    int rc = geon_get_forces(h, &v, n);
    if (rc) return throw_geon(env, h, rc);
    return out;
}

/* particleCount(h) -> live particles (gpu-physics.js:4287) */
static napi_value js_particle_count(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int32_t n = 0;
    int rc = geon_particle_count(h, &n);
    if (rc) return throw_geon(env, h, rc);
    return i32(env, n);
}

/* collisionEvents(h) -> Float64Array of GEON_EVENT_STRIDE doubles per event
 * {type, x, y, energy, px, py} of the last update (collisions.js:62-92) */
static napi_value js_collision_events(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int64_t n = 0;
    int rc = geon_collision_events(h, NULL, 0, &n);
    if (rc) return throw_geon(env, h, rc);
    double* data;
    napi_value ta = new_f64(env, (size_t)n * GEON_EVENT_STRIDE, &data);
    if (n > 0 && (rc = geon_collision_events(h, data, n, &n))) return throw_geon(env, h, rc);
    return ta;
}

/* getStats(h) -> readStats-shaped object (gpu-physics.js:4215-4281) */
static napi_value js_get_stats(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    GeonStats s;
    int rc = geon_get_stats(h, &s);
    if (rc) return throw_geon(env, h, rc);
    napi_value o;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "linearKE", num(env, s.linear_ke));
    napi_set_named_property(env, o, "spinKE", num(env, s.spin_ke));
    napi_set_named_property(env, o, "px", num(env, s.px));
    napi_set_named_property(env, o, "py", num(env, s.py));
    napi_set_named_property(env, o, "orbitalAngMom", num(env, s.orbital_ang_mom));
    napi_set_named_property(env, o, "spinAngMom", num(env, s.spin_ang_mom));
    napi_set_named_property(env, o, "comX", num(env, s.com_x));
    napi_set_named_property(env, o, "comY", num(env, s.com_y));
    napi_set_named_property(env, o, "totalMass", num(env, s.total_mass));
    napi_set_named_property(env, o, "pe", num(env, s.pe));
    napi_set_named_property(env, o, "fieldEnergy", num(env, s.field_energy));
    napi_set_named_property(env, o, "higgsFieldEnergy", num(env, s.higgs_field_energy));
    napi_set_named_property(env, o, "axionFieldEnergy", num(env, s.axion_field_energy));
    napi_set_named_property(env, o, "pfiEnergy", num(env, s.pfi_energy));
    napi_set_named_property(env, o, "fieldPx", num(env, s.field_px));
    napi_set_named_property(env, o, "fieldPy", num(env, s.field_py));
    napi_set_named_property(env, o, "simTime", num(env, s.sim_time));
    napi_set_named_property(env, o, "aliveCount", i32(env, s.alive_count));
    napi_set_named_property(env, o, "deadCount", i32(env, s.dead_count));
    napi_set_named_property(env, o, "totalRadiated", num(env, s.radiated_energy));
    napi_set_named_property(env, o, "totalRadiatedPx", num(env, s.radiated_px));
    napi_set_named_property(env, o, "totalRadiatedPy", num(env, s.radiated_py));
    return o;
}

/* select(h, slot): requestStats(selectedGpuIdx)'s selection (gpu-physics.js:4146); -1 for none */
static napi_value js_select(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int32_t slot;
    CHECK(napi_get_value_int32(env, argv[1], &slot));
    int rc = geon_select(h, slot);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

static napi_value vec2(napi_env env, const double* f) {
    napi_value o;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "x", num(env, f[0]));
    napi_set_named_property(env, o, "y", num(env, f[1]));
    return o;
}

/* getSelected(h) -> readStats().selected (gpu-physics.js:4238-4276), or null */
static napi_value js_get_selected(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    GeonSelected s;
    int rc = geon_get_selected(h, &s);
    if (rc) return throw_geon(env, h, rc);
    napi_value o;
    if (s.slot < 0) { napi_get_null(env, &o); return o; }
    napi_create_object(env, &o);
    const char* sn[] = {"posX", "posY", "velWX", "velWY", "mass", "charge", "angW", "baseMass", "radius", "velX",
                        "velY", "angVel", "magMoment", "angMomentum"};
    const double sv[] = {s.x, s.y, s.wx, s.wy, s.mass, s.charge, s.angw, s.base_mass, s.radius, s.vx, s.vy,
                         s.ang_vel, s.mag_moment, s.ang_momentum};
    for (int i = 0; i < 14; i++) napi_set_named_property(env, o, sn[i], num(env, sv[i]));
    napi_value b;
    napi_get_boolean(env, s.antimatter != 0, &b);
    napi_set_named_property(env, o, "antimatter", b);
    napi_set_named_property(env, o, "flags", i32(env, s.antimatter ? 1 : 0));
    napi_set_named_property(env, o, "slot", i32(env, s.slot));
    napi_get_boolean(env, s.forces_valid != 0, &b);
    napi_set_named_property(env, o, "forcesValid", b);
    const char* fn[GEON_SF_COUNT] = {"forceGravity", "forceCoulomb", "forceMagnetic", "forceGravitomag", "force1PN",
                                     "forceSpinCurv", "forceRadiation", "forceYukawa", "forceExternal",
                                     "forceHiggs", "forceAxion", "forceFarField"};
    for (int k = 0; k < GEON_SF_COUNT; k++) napi_set_named_property(env, o, fn[k], vec2(env, s.force[k]));
    return o;
}

/* getField(h, which, what) -> Float64Array(fieldGrid^2), row-major (geon_get_field) */
static napi_value js_get_field(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!args_n(env, info, 3, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int32_t which, what;
    CHECK(napi_get_value_int32(env, argv[1], &which));
    CHECK(napi_get_value_int32(env, argv[2], &what));
    GeonParams p;
    int rc = geon_get_params(h, &p);
    if (rc) return throw_geon(env, h, rc);
    const size_t n = (size_t)p.field_grid * (size_t)p.field_grid;
    double* data;
    napi_value ta = new_f64(env, n, &data);
    rc = geon_get_field(h, which, what, data, (int32_t)n);
    if (rc) return throw_geon(env, h, rc);
    return ta;
}

static napi_value js_last_error(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc = 1;
    napi_get_cb_info(env, info, &argc, argv, NULL, NULL);
    GeonHandle h = argc ? get_handle(env, argv[0]) : NULL;
    napi_value s;
    napi_create_string_utf8(env, geon_last_error(h), NAPI_AUTO_LENGTH, &s);
    return s;
}

static napi_value js_abi_version(napi_env env, napi_callback_info info) {
    (void)info;
    return i32(env, geon_abi_version());
}

/* ---- learned far-field correction ---- */

/* farCollect(h, groupStride) -> samples held */
static napi_value js_far_collect(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int32_t stride;
    CHECK(napi_get_value_int32(env, argv[1], &stride));
    int64_t n = 0;
    int rc = geon_far_collect(h, stride, &n);
    if (rc) return throw_geon(env, h, rc);
    return num(env, (double)n);
}

static napi_value js_far_clear(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int rc = geon_far_clear(h);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

/* farTrain(h, steps, batch, lr, lrFinal, seed) -> Float32Array of losses */
static napi_value js_far_train(napi_env env, napi_callback_info info) {
    napi_value argv[6];
    if (!args_n(env, info, 6, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    int32_t steps, batch;
    double lr, lr_final, seed;
    CHECK(napi_get_value_int32(env, argv[1], &steps));
    CHECK(napi_get_value_int32(env, argv[2], &batch));
    CHECK(napi_get_value_double(env, argv[3], &lr));
    CHECK(napi_get_value_double(env, argv[4], &lr_final));
    CHECK(napi_get_value_double(env, argv[5], &seed));
    if (steps < 0) steps = 0;
    float* loss = NULL;
    napi_value ab, ta;
    CHECK(napi_create_arraybuffer(env, (size_t)(steps > 0 ? steps : 1) * sizeof(float), (void**)&loss, &ab));
    CHECK(napi_create_typedarray(env, napi_float32_array, (size_t)steps, ab, 0, &ta));
    int rc = geon_far_train(h, steps, batch, lr, lr_final, (uint64_t)seed, loss);
    if (rc) return throw_geon(env, h, rc);
    return ta;
}

/* farGetModel(h) -> Float32Array(GEON_FAR_MODEL_FLOATS) */
static napi_value js_far_get_model(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args_n(env, info, 1, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    float* m = NULL;
    napi_value ab, ta;
    CHECK(napi_create_arraybuffer(env, GEON_FAR_MODEL_FLOATS * sizeof(float), (void**)&m, &ab));
    CHECK(napi_create_typedarray(env, napi_float32_array, GEON_FAR_MODEL_FLOATS, ab, 0, &ta));
    int rc = geon_far_get_model(h, m, GEON_FAR_MODEL_FLOATS);
    if (rc) return throw_geon(env, h, rc);
    return ta;
}

/* farSetModel(h, Float32Array(GEON_FAR_MODEL_FLOATS)) */
static napi_value js_far_set_model(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args_n(env, info, 2, argv)) return NULL;
    GeonHandle h = get_handle(env, argv[0]);
    napi_typedarray_type t;
    size_t len = 0;
    void* data = NULL;
    CHECK(napi_get_typedarray_info(env, argv[1], &t, &len, &data, NULL, NULL));
    if (t != napi_float32_array) { napi_throw_type_error(env, NULL, "farSetModel wants a Float32Array"); return NULL; }
    int rc = geon_far_set_model(h, (const float*)data, (int32_t)len);
    if (rc) return throw_geon(env, h, rc);
    return NULL;
}

#define EXPORT(name, fn) do { napi_value f; napi_create_function(env, name, NAPI_AUTO_LENGTH, fn, NULL, &f); \
                              napi_set_named_property(env, exports, name, f); } while (0)

static napi_value init(napi_env env, napi_value exports) {
    EXPORT("abiVersion", js_abi_version);
    EXPORT("create", js_create);
    EXPORT("destroy", js_destroy);
    EXPORT("setParams", js_set_params);
    EXPORT("reset", js_reset);
    EXPORT("addParticle", js_add_particle);
    EXPORT("removeParticle", js_remove_particle);
    EXPORT("patchMass", js_patch_mass);
    EXPORT("uploadSoA", js_upload);
    EXPORT("collisionEvents", js_collision_events);
    EXPORT("downloadSoA", js_download);
    EXPORT("update", js_update);
    EXPORT("computeForces", js_compute_forces);
    EXPORT("getForces", js_get_forces);
    EXPORT("getStats", js_get_stats);
    EXPORT("select", js_select);
    EXPORT("getSelected", js_get_selected);
    EXPORT("getField", js_get_field);
    EXPORT("particleCount", js_particle_count);
    EXPORT("lastError", js_last_error);
    EXPORT("farCollect", js_far_collect);
    EXPORT("farClear", js_far_clear);
    EXPORT("farTrain", js_far_train);
    EXPORT("farGetModel", js_far_get_model);
    EXPORT("farSetModel", js_far_set_model);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
