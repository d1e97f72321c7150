// Runs the reference's CPU physics (Physics.update of src/integrator.js, with
// HiggsField / AxionField from src/higgs-field.js, src/axion-field.js) on the
// cases of an input JSON and prints what it computed as one JSON document.
// It runs in the build container only, under Node 12, with
//   GEON_REF_SRC=<reference>/src node --experimental-loader ./ref_loader.mjs reference_goldens.mjs in.json
// (make_reference_goldens.py does this and stores the results as fixtures).
// The browser globals the modules read at load time (window._PALETTE,
// window._parseHex, document for the theme and the fields' offscreen canvas)
// are stood in by inert values: they only pick display colours.

const palette = {
    neutral: '#808080',
    light: { text: '#000000' },
    dark: { text: '#ffffff' },
    extended: { red: '#ff0000', blue: '#0000ff', purple: '#800080', lime: '#00ff00', indigo: '#4b0082',
                yellow: '#ffff00' },
};
const parseHex = h => [1, 3, 5].map(k => parseInt(h.slice(k, k + 2), 16) / 255);
const ctx2d = { createImageData: (w, h) => ({ data: new Uint8ClampedArray(4 * w * h), width: w, height: h }),
                putImageData() {}, drawImage() {} };
globalThis.window = { _PALETTE: palette, _parseHex: parseHex };
globalThis.document = {
    documentElement: { dataset: { theme: 'light' } },
    createElement: () => ({ width: 0, height: 0, getContext: () => ctx2d }),
};

const TOGGLES = ['gravityEnabled', 'coulombEnabled', 'magneticEnabled', 'gravitomagEnabled', 'onePNEnabled',
                 'relativityEnabled', 'spinOrbitEnabled', 'radiationEnabled', 'blackHoleEnabled',
                 'disintegrationEnabled', 'expansionEnabled', 'yukawaEnabled', 'higgsEnabled', 'axionEnabled',
                 'barnesHutEnabled', 'bosonInterEnabled'];

function forcesOf(ps) {
    const f = k => ps.map(p => k(p));
    return {
        fx: f(p => p.force.x), fy: f(p => p.force.y),
        bz: f(p => p.Bz), bgz: f(p => p.Bgz),
        dbzdx: f(p => p.dBzdx), dbzdy: f(p => p.dBzdy), dbgzdx: f(p => p.dBgzdx), dbgzdy: f(p => p.dBgzdy),
        frame_drag: f(p => p._frameDragTorque), tidal: f(p => p._tidalTorque),
        f1pn_x: f(p => p.force1PN.x), f1pn_y: f(p => p.force1PN.y),
        // the per-type vectors readStats' selected shows (after the display pass, integrator.js:1751-1796)
        ...Object.fromEntries(DISPLAY.flatMap(k => [[`${k}_x`, f(p => p[k].x)], [`${k}_y`, f(p => p[k].y)]])),
    };
}

const DISPLAY = ['forceGravity', 'forceCoulomb', 'forceMagnetic', 'forceGravitomag', 'force1PN', 'forceSpinCurv',
                 'forceRadiation', 'forceYukawa', 'forceExternal', 'forceHiggs', 'forceAxion'];

function stateOf(ps) {
    const f = k => ps.map(p => k(p));
    return { id: f(p => p.id), x: f(p => p.pos.x), y: f(p => p.pos.y), wx: f(p => p.w.x), wy: f(p => p.w.y),
             mass: f(p => p.mass), angw: f(p => p.angw) };
}

async function runCase(c, mods) {
    const { Physics, Particle, HiggsField, AxionField, computeEnergies } = mods;
    Particle.nextId = 0;
    const ph = new Physics();
    for (const k of TOGGLES) ph[k] = !!c.toggles[k];
    if (c.yukawaMu !== undefined) ph.yukawaMu = c.yukawaMu;
    const sim = { photons: [], pions: [], leptons: [], deadParticles: [], totalRadiated: 0, totalRadiatedPx: 0,
                  totalRadiatedPy: 0, higgsField: null, axionField: null, physics: ph,
                  boundaryMode: c.boundaryMode };     // main.js:177 (read by computeEnergies)
    if (ph.higgsEnabled) sim.higgsField = new HiggsField();
    if (ph.axionEnabled) sim.axionField = new AxionField();
    if (c.fields) {
        // a seeded initial field, its grid gradients refreshed (as geon_set_field does)
        for (const [f, v] of [[sim.higgsField, c.fields.higgs], [sim.axionField, c.fields.axion]]) {
            if (!f || !v) continue;
            f.field.set(v);
            f._computeGridGradients(c.boundaryMode, c.topology, f._vacValue);
        }
    }
    ph.sim = sim;
    const P = c.particles;
    const ps = [];
    for (let i = 0; i < P.x.length; i++) {
        // loadState (save-load.js): mass, charge, proper velocity w, angw; v from w
        const p = new Particle(P.x[i], P.y[i], P.mass[i], Math.round(P.charge[i] / 0.1) * 0.1);
        p.w.x = P.wx[i]; p.w.y = P.wy[i];
        if (ph.relativityEnabled) {
            const ig = 1 / Math.sqrt(1 + p.w.x * p.w.x + p.w.y * p.w.y);
            p.vel.x = p.w.x * ig; p.vel.y = p.w.y * ig;
        } else { p.vel.x = p.w.x; p.vel.y = p.w.y; }
        p.angw = P.angw[i];
        p.angVel = ph.relativityEnabled ? p.angw / Math.sqrt(1 + p.angw * p.angw * p.radius * p.radius) : p.angw;
        ps.push(p);
    }
    // substeps are counted through the one method every substep calls once (step 7)
    let calls = 0;
    const sync = ph._syncAxionField.bind(ph);
    ph._syncAxionField = (...a) => { calls++; return sync(...a); };
    const args = [c.collisionMode, c.boundaryMode, c.topology, c.W, c.H];
    ph.update(ps, 0, ...args);                         // the bootstrap force pass only (dtRemain = 0)
    if (c.timeOnly) {
        // wall time of the updates alone, particle-steps counted as live particles per substep
        let steps = 0;
        const t0 = process.hrtime.bigint();
        for (let k = 0; k < c.updates; k++) {
            calls = 0;
            const n0 = ps.length;
            ph.update(ps, c.dt, ...args);
            steps += calls * n0;
        }
        const sec = Number(process.hrtime.bigint() - t0) * 1e-9;
        return { seconds: sec, particle_steps: steps, n: ps.length };
    }
    const out = { bootstrap: { forces: forcesOf(ps), pe: ph.potentialEnergy } };
    const substeps = [];
    for (let k = 0; k < c.updates; k++) {
        calls = 0;
        ph.update(ps, c.dt, ...args);
        substeps.push(calls);
    }
    out.substeps = substeps;
    out.final = { state: stateOf(ps), forces: forcesOf(ps), pe: ph.potentialEnergy, sim_time: ph.simTime };
    const e = computeEnergies(ps, ph, sim);
    out.final.energies = {};
    for (const [k, v] of Object.entries(e)) if (typeof v === 'number') out.final.energies[k] = v;
    if (sim.higgsField) out.final.higgs = { field: Array.from(sim.higgsField.field),
                                            field_dot: Array.from(sim.higgsField.fieldDot) };
    if (sim.axionField) out.final.axion = { field: Array.from(sim.axionField.field),
                                            field_dot: Array.from(sim.axionField.fieldDot) };
    return out;
}

async function main() {
    const src = process.env.GEON_REF_SRC;
    const fs = await import('fs');
    const input = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
    const mods = {
        Physics: (await import(src + '/integrator.js')).default,
        Particle: (await import(src + '/particle.js')).default,
        HiggsField: (await import(src + '/higgs-field.js')).default,
        AxionField: (await import(src + '/axion-field.js')).default,
        computeEnergies: (await import(src + '/energy.js')).computeEnergies,
    };
    const results = {};
    for (const c of input.cases) results[c.name] = await runCase(c, mods);
    process.stdout.write(JSON.stringify(results));
}

main().catch(e => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1); });
