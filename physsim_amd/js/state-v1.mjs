// state-v1.mjs — the reference's save-state format, version 1
// (save-load.js:89-127 writes it, 129-262 reads it; gpu-physics.js:3749-3836
// and 3886-3960 are the GPU backend's copy), mapped onto the addon's
// camelCase GeonParams and the SoA upload.  No device work here.

// toggles0 bits: src/gpu/gpu-constants.js:196-213
export const TOGGLE_BITS = {
    gravityEnabled: 1, coulombEnabled: 2, magneticEnabled: 4, gravitomagEnabled: 8,
    onePNEnabled: 16, relativityEnabled: 32, spinOrbitEnabled: 64, radiationEnabled: 128,
    blackHoleEnabled: 256, disintegrationEnabled: 512, expansionEnabled: 1024, yukawaEnabled: 2048,
    higgsEnabled: 4096, axionEnabled: 8192, barnesHutEnabled: 16384, bosonInterEnabled: 32768,
};

// the toggles a state carries, in the order _cpuSaveState writes them (save-load.js:107-113)
export const TOGGLE_KEYS = ['gravityEnabled', 'bosonInterEnabled', 'coulombEnabled', 'magneticEnabled',
    'gravitomagEnabled', 'relativityEnabled', 'barnesHutEnabled', 'radiationEnabled', 'blackHoleEnabled',
    'disintegrationEnabled', 'spinOrbitEnabled', 'onePNEnabled', 'yukawaEnabled', 'axionEnabled',
    'expansionEnabled', 'higgsEnabled'];

// config.js:12-25
export const COL_NAMES = ['pass', 'merge', 'bounce'];
export const BOUND_NAMES = ['despawn', 'bounce', 'loop'];
export const TOPO_NAMES = ['torus', 'klein', 'rp2'];
export const colFromString = s => (s === 'merge' ? 1 : s === 'bounce' ? 2 : 0);
export const boundFromString = s => (s === 'loop' ? 2 : s === 'bounce' ? 1 : 0);
export const topoFromString = s => (s === 'klein' ? 1 : s === 'rp2' ? 2 : 0);

// [state key, addon param key, restored when]: 'value' = `!= null`,
// 'defined' = `!== undefined` (save-load.js:116-125, 221-230)
export const PARAM_KEYS = [
    ['yukawaMu', 'yukawaMu', 'value'], ['axionMass', 'axionMass', 'value'], ['hubbleParam', 'hubbleParam', 'value'],
    ['higgsMass', 'higgsMass', 'defined'], ['extGravity', 'extGravity', 'defined'],
    ['extGravityAngle', 'extGravityAngle', 'defined'], ['extElectric', 'extElectric', 'defined'],
    ['extElectricAngle', 'extElectricAngle', 'defined'], ['extBz', 'extBz', 'defined'],
];

// Reference defaults (src/config.js) of the fields a state carries.
const DEFAULTS = { toggles0: 16387, collisionMode: 0, boundaryMode: 0, topology: 0, yukawaMu: 0.15,
                   axionMass: 0.05, hubbleParam: 0.001, higgsMass: 0.5, extGravity: 0, extGravityAngle: Math.PI / 2,
                   extElectric: 0, extElectricAngle: 0, extBz: 0, bounceFriction: 0.4 };

/** A v1 state object from addon params and the live particles; `sim` may
 *  carry the app's speedScale and camera.  settings.friction is the
 *  physics' bounceFriction (save-load.js:103). */
export function paramsToState(params, particles, sim = {}) {
    const p = { ...DEFAULTS, ...params };
    const state = {
        version: 1,
        particles,
        toggles: {},
        settings: {
            collision: COL_NAMES[p.collisionMode],
            boundary: BOUND_NAMES[p.boundaryMode],
            topology: TOPO_NAMES[p.topology],
            speed: sim.speedScale != null ? sim.speedScale : 16,
            friction: p.bounceFriction != null ? p.bounceFriction : 0.4,
        },
    };
    for (const k of TOGGLE_KEYS) state.toggles[k] = !!(p.toggles0 & TOGGLE_BITS[k]);
    if (sim.camera) state.camera = { x: sim.camera.x, y: sim.camera.y, zoom: sim.camera.zoom };
    for (const [key, param] of PARAM_KEYS) state[key] = p[param];
    return state;
}

/** Addon params with a state's toggles, parameters and modes applied over `params`. */
export function applyState(params, state) {
    const p = { ...DEFAULTS, ...params };
    let t = p.toggles0;
    for (const [key, val] of Object.entries(state.toggles || {})) {
        const bit = TOGGLE_BITS[key];
        if (bit === undefined) continue;                     // `if (key in ph)` (save-load.js:215)
        t = val ? (t | bit) : (t & ~bit);
    }
    p.toggles0 = t >>> 0;
    for (const [key, param, when] of PARAM_KEYS) {
        if (!(key in state)) continue;
        const v = state[key];
        if (v == null && when === 'value') continue;
        p[param] = v == null ? 0 : v;
    }
    const st = state.settings;
    if (st) {
        const col = st.collision === 'repel' ? 'bounce' : (st.collision || 'pass');   // save-load.js:194
        p.collisionMode = colFromString(col);
        p.boundaryMode = boundFromString(st.boundary);
        p.topology = topoFromString(st.topology);
        if (st.friction != null) p.bounceFriction = st.friction;          // save-load.js:223
    }
    return p;
}

/** SoA columns with the GPU backend's defaults (gpu-physics.js:3896-3912). */
export function particlesFromState(state, cap) {
    const ps = state.particles.slice(0, cap);
    const n = ps.length;
    const f = () => new Float64Array(n);
    const c = { x: f(), y: f(), wx: f(), wy: f(), mass: f(), baseMass: f(), charge: f(), angw: f(),
                antimatter: new Int32Array(n) };
    ps.forEach((pd, i) => {
        const m = pd.mass || 1;
        c.x[i] = pd.x; c.y[i] = pd.y;
        c.wx[i] = pd.wx || 0; c.wy[i] = pd.wy || 0;
        c.mass[i] = m;
        c.baseMass[i] = pd.baseMass != null ? pd.baseMass : m;     // `??` (node 12 lacks it)
        c.charge[i] = pd.charge || 0;
        c.angw[i] = pd.angw || 0;
        c.antimatter[i] = pd.antimatter ? 1 : 0;
    });
    return c;
}
