// harness.mjs — drives HipPhysics the way the reference app's frame loop drives
// a backend (main.js:603-700: fixed PHYSICS_DT ticks of Physics.update), on a
// particle set read from a JSON file, and prints the final state and stats as
// one JSON line.  Usage: node harness.mjs <input.json> <updates> <dt>
import fs from 'fs';
import HipPhysics from './hip-physics.mjs';

async function main() {
    const [inputPath, updatesArg, dtArg] = process.argv.slice(2);
    const input = JSON.parse(fs.readFileSync(inputPath, 'utf8'));
    const updates = parseInt(updatesArg || '5', 10);
    const dt = parseFloat(dtArg || String(1 / 128));
    const sim = await new HipPhysics(input.params, input.particles.length).init();
    sim.deserialize({ version: 1, particles: input.particles });
    const substeps = [];
    for (let k = 0; k < updates; k++) substeps.push(sim.update(dt));
    sim.requestStats();
    const state = await sim.serialize();
    process.stdout.write(JSON.stringify({ substeps, stats: sim.readStats(), particles: state.particles }) + '\n');
    sim.destroy();
}

main().catch(e => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1); });
