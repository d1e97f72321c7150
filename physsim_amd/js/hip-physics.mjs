// hip-physics.mjs — the gfx950 backend as a third physics backend for the
// reference app, beside CPUPhysics (src/cpu-physics.js) and GPUPhysics
// (src/gpu/gpu-physics.js).  It keeps GPUPhysics' method names and argument
// meanings (addParticle returns -1 when full, vx/vy are proper velocity,
// charge is quantized, readStats returns the same fields); every call goes
// through the Node-API addon to the C ABI of libgeon_hip.so.
import { createRequire } from 'module';
import { fileURLToPath } from 'url';
import path from 'path';

import { TOGGLE_BITS, paramsToState, applyState, particlesFromState } from './state-v1.mjs';

const require = createRequire(import.meta.url);
const here = path.dirname(fileURLToPath(import.meta.url));
const addon = require(path.join(here, '..', 'napi', 'geon_napi.node'));

export { TOGGLE_BITS };

export const BACKEND_HIP = 'hip';

export default class HipPhysics {
    /** @param {object} params  GeonParams fields in camelCase (missing = reference defaults) */
    constructor(params = {}, maxParticles = 1 << 16) {
        this._params = { ...params };
        this._max = maxParticles;
        this._h = null;
        this._stats = null;
    }

    /** GPUPhysics.init(): allocate the device system. */
    async init() {
        this._h = addon.create(this._params, this._max);
        return this;
    }

    get aliveCount() { return this.getParticleCount(); }

    set boundaryMode(v) { this._setParam('boundaryMode', v); }
    set topologyMode(v) { this._setParam('topology', v); }
    set domainW(v) { this._setParam('domainW', v); }
    set domainH(v) { this._setParam('domainH', v); }

    _setParam(k, v) {
        this._params[k] = v;
        if (this._h) addon.setParams(this._h, this._params);
    }

    /** setToggles(physics): the *Enabled booleans of a Physics-like object (gpu-physics.js:1969). */
    setToggles(physics) {
        let t = 0;
        for (const [name, bit] of Object.entries(TOGGLE_BITS)) if (physics[name]) t |= bit;
        this._params.toggles0 = t;
        if (physics.yukawaMu !== undefined) this._params.yukawaMu = physics.yukawaMu;
        if (physics.higgsMass !== undefined) this._params.higgsMass = physics.higgsMass;
        if (physics.axionMass !== undefined) this._params.axionMass = physics.axionMass;
        if (physics.extGravity !== undefined) this._params.extGravity = physics.extGravity;
        if (physics.extGravityAngle !== undefined) this._params.extGravityAngle = physics.extGravityAngle;
        if (physics.extElectric !== undefined) this._params.extElectric = physics.extElectric;
        if (physics.extElectricAngle !== undefined) this._params.extElectricAngle = physics.extElectricAngle;
        if (physics.extBz !== undefined) this._params.extBz = physics.extBz;
        if (this._h) addon.setParams(this._h, this._params);
    }

    /** One Physics.update call: adaptive substeps, boundary once at the end. */
    update(dt) { return addon.update(this._h, dt); }

    addParticle({ x, y, vx = 0, vy = 0, mass = 1, charge = 0, angw = 0, antimatter = false }) {
        return addon.addParticle(this._h, { x, y, vx, vy, mass, charge, angw, antimatter });
    }

    // deathMass / deathAngVel: 0 keeps the particle's own values (input.js:183)
    removeParticle(idx, deathMass = 0, deathAngVel = 0) { addon.removeParticle(this._h, idx, deathMass, deathAngVel); }

    patchMass(idx, mass, baseMass) { addon.patchMass(this._h, idx, mass, baseMass); }

    uploadSoA(arrays) { addon.uploadSoA(this._h, arrays); }

    downloadSoA() { return addon.downloadSoA(this._h); }

    computeForces() { addon.computeForces(this._h); }

    getForces(nSlots) { return addon.getForces(this._h, nSlots); }

    reset() { addon.reset(this._h); }

    /** Signal delay of leaf sources: 'follow' (on exactly when relativity is,
     *  forces.js:85), 'off' or 'on'; a history sample every `stride` updates
     *  (HISTORY_STRIDE, config.js:122). */
    setSignalDelay(mode, stride = 64) {
        this._params.signalDelay = { follow: 0, off: 1, on: 2 }[mode];
        this._params.historyStride = stride;
        addon.setParams(this._h, this._params);
    }

    // ---- learned far-field correction (no GPUPhysics counterpart) ----
    /** Far-field mode: 'none' (the reference's monopole nodes), 'multipole' or 'learned'. */
    setFarField(mode, listCap = 0) {
        this._params.farField = { none: 0, multipole: 1, learned: 2 }[mode];
        this._params.farListCap = listCap;
        addon.setParams(this._h, this._params);
    }

    /** Training samples from the current state; returns the count held. */
    farCollect(groupStride = 1) { return addon.farCollect(this._h, groupStride); }

    farClear() { addon.farClear(this._h); }

    /** Adam steps on the held samples; returns the per-step losses. */
    farTrain(steps, { batch = 16384, lr = 2e-3, lrFinal = 1e-5, seed = 3 } = {}) {
        return addon.farTrain(this._h, steps, batch, lr, lrFinal, seed);
    }

    farGetModel() { return addon.farGetModel(this._h); }

    farSetModel(model) { addon.farSetModel(this._h, model); }

    getParticleCount() { return addon.particleCount(this._h); }

    /** requestStats(selectedGpuIdx) (gpu-physics.js:4146): the stats now, and
     *  the selected particle, whose forces by type the force passes record
     *  from its selection on (`selected.forcesValid`). */
    requestStats(selIdx = -1) {
        addon.select(this._h, selIdx);
        this._stats = addon.getStats(this._h);
        this._stats.selected = addon.getSelected(this._h);
    }

    readStats() { return this._stats; }

    /** getFieldBuffers(which) (gpu-physics.js:3307): the field ('higgs' or
     *  'axion') as host arrays of fieldGrid^2 values, row-major. */
    getFieldBuffers(which) {
        const w = which === 'higgs' ? 0 : 1;
        if (!(this._params.toggles0 & (w === 0 ? TOGGLE_BITS.higgsEnabled : TOGGLE_BITS.axionEnabled))) return null;
        const q = (k) => addon.getField(this._h, w, k);
        return { field: q(0), fieldDot: q(1), gradX: q(2), gradY: q(3), source: q(7) };
    }

    /** readbackFieldData() (gpu-physics.js:3844): Float32Arrays of the fields
     *  and their time derivatives at the device grid, null for a field that is off. */
    async readbackFieldData() {
        const out = { higgsField: null, higgsFieldDot: null, axionField: null, axionFieldDot: null };
        for (const name of ['higgs', 'axion']) {
            const b = this.getFieldBuffers(name);
            if (!b) continue;
            out[name + 'Field'] = Float32Array.from(b.field);
            out[name + 'FieldDot'] = Float32Array.from(b.fieldDot);
        }
        return out;
    }

    /** Events of the last update in collision mode merge (collisions.js:62-92):
     *  [{type: 'merge'|'annihilation', x, y, energy, px, py}], merges first.
     *  A host that keeps photons emits the annihilation bursts from these
     *  (integrator.js:1388-1392). */
    getCollisionEvents() {
        const a = addon.collisionEvents(this._h);
        const out = [];
        for (let k = 0; k + 6 <= a.length; k += 6)
            out.push({ type: a[k] === 1 ? 'merge' : 'annihilation', x: a[k + 1], y: a[k + 2], energy: a[k + 3],
                       px: a[k + 4], py: a[k + 5] });
        return out;
    }

    /** serialize(sim): the v1 state (gpu-physics.js:3749-3836): live particles
     *  in slot order, toggles, parameters and modes; `sim` may carry the app's
     *  speedScale and camera (settings.friction is the bounce friction). */
    async serialize(sim = {}) {
        const d = this.downloadSoA();
        const particles = [];
        for (let i = 0; i < d.nSlots; i++) {
            if (!d.alive[i]) continue;
            particles.push({ x: d.x[i], y: d.y[i], wx: d.wx[i], wy: d.wy[i], mass: d.mass[i],
                             baseMass: d.baseMass[i], charge: d.charge[i], angw: d.angw[i],
                             antimatter: !!d.antimatter[i] });
        }
        return paramsToState(this._params, particles, sim);
    }

    /** deserialize(state): GPUPhysics.deserialize (gpu-physics.js:3886-3960)
     *  with loadState's restore of toggles, parameters and modes
     *  (save-load.js:129-190).  False for anything but version 1. */
    deserialize(state) {
        if (!state || state.version !== 1) return false;
        this._params = applyState(this._params, state);
        addon.setParams(this._h, this._params);
        this.uploadSoA(particlesFromState(state, this._max));
        return true;
    }

    destroy() {
        if (this._h) { addon.destroy(this._h); this._h = null; }
    }
}
