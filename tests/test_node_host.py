"""The JavaScript host side: the Node-API addon and the HipPhysics ES module
drive the same C ABI as the Python tests.  Skipped where node is absent."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from physsim_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "physsim_amd", "napi", "geon_napi.node")
HARNESS = os.path.join(ROOT, "physsim_amd", "js", "harness.mjs")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node not installed")


def _ensure_addon():
    if not os.path.exists(ADDON):
        from physsim_amd import build
        build.build()
    if not os.path.exists(ADDON):
        pytest.skip("node headers absent: addon not built")


def test_addon_loads_and_exports():
    _ensure_addon()
    out = subprocess.run([NODE, "-e", "const g=require(%r); console.log(JSON.stringify(Object.keys(g)), g.abiVersion())"
                          % ADDON], capture_output=True, text=True, check=True).stdout
    keys, ver = out.rsplit(" ", 1)
    assert int(ver) == 1
    for k in ("create", "update", "addParticle", "removeParticle", "patchMass", "uploadSoA", "downloadSoA",
              "getStats", "computeForces", "getForces", "setParams", "reset", "destroy"):
        assert k in json.loads(keys)


def test_create_without_gpu_throws():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    _ensure_addon()
    r = subprocess.run([NODE, "-e", "const g=require(%r); try { g.create({}, 8); console.log('made') } "
                                    "catch (e) { console.log('threw:', e.message) }" % ADDON],
                       capture_output=True, text=True, check=True)
    assert "threw: geon error -3" in r.stdout


@pytest.mark.gpu
def test_node_harness_matches_python_path(tmp_path):
    """The JS host (v1 JSON in, main.js-style update ticks) gives bit-identical
    state to the Python host over the same library."""
    _ensure_addon()
    from physsim_amd import HipPhysics, default_params
    st, prm, dt = W.config1(512)
    particles = [dict(x=float(st["x"][i]), y=float(st["y"][i]), wx=float(st["wx"][i]), wy=float(st["wy"][i]),
                      mass=float(st["mass"][i]), baseMass=float(st["mass"][i]), charge=float(st["charge"][i]),
                      angw=0.0) for i in range(len(st["x"]))]
    js_params = dict(toggles0=prm["toggles0"], boundaryMode=prm["boundary_mode"], domainW=prm["domain_w"],
                     domainH=prm["domain_h"], softening=prm["softening"], theta=prm["theta"])
    inp = tmp_path / "in.json"
    inp.write_text(json.dumps(dict(params=js_params, particles=particles)))
    print("case: node harness config1 n=512", flush=True)
    r = subprocess.run([NODE, HARNESS, str(inp), "4", repr(dt)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    g = HipPhysics(default_params(**prm), max_particles=512)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    steps = [g.update(dt) for _ in range(4)]
    assert steps == js["substeps"]
    d = g.download()
    alive = np.nonzero(d["alive"])[0]
    assert len(alive) == len(js["particles"])
    assert np.array_equal(d["x"][alive], np.array([p["x"] for p in js["particles"]]))
    assert np.array_equal(d["wy"][alive], np.array([p["wy"] for p in js["particles"]]))
    assert js["stats"]["aliveCount"] == len(alive)


def test_addon_upload_rejects_bad_columns():
    """uploadSoA checks its columns before it reaches the library: a column
    shorter than x, a plain Array, or a missing required column throws."""
    _ensure_addon()
    js = ("const g=require(%r); const f=n=>new Float64Array(n);"
          "const cols=(n,m)=>({x:f(n),y:f(m),wx:f(n),wy:f(n),mass:f(n),charge:f(n)});"
          "const tryit=(c)=>{try{g.uploadSoA(null,c);console.log('ok')}catch(e){console.log(e.name+': '+e.message)}};"
          "tryit(cols(4,3)); const c=cols(4,4); c.x=[1,2,3,4]; tryit(c);"
          "const d=cols(4,4); delete d.charge; tryit(d);"
          "const e=cols(4,4); e.antimatter=new Int32Array(2); tryit(e);" % ADDON)
    out = subprocess.run([NODE, "-e", js], capture_output=True, text=True, check=True).stdout.splitlines()
    assert out[0].startswith("RangeError") and "'y' has 3 entries" in out[0]
    assert out[1].startswith("TypeError") and "'x'" in out[1]
    assert out[2].startswith("TypeError") and "'charge'" in out[2]
    assert out[3].startswith("RangeError") and "'antimatter'" in out[3]


@pytest.mark.gpu
def test_node_selection_and_field_readback():
    """requestStats(selIdx)/readStats().selected, removeParticle's death
    metadata, getFieldBuffers and readbackFieldData through the addon and the
    HipPhysics module (gpu-physics.js:1936, 3307, 3844, 4146-4276)."""
    _ensure_addon()
    module = os.path.join(ROOT, "physsim_amd", "js", "hip-physics.mjs")
    js = """
import HipPhysics, { TOGGLE_BITS } from %r;
async function main() {
  const t = TOGGLE_BITS.gravityEnabled | TOGGLE_BITS.higgsEnabled;
  const g = await new HipPhysics({ toggles0: t, fieldGrid: 32, domainW: 64, domainH: 64 }, 8).init();
  for (const [x, m] of [[20, 1], [26, 2], [40, 1]]) g.addParticle({ x, y: 32, mass: m });
  g.requestStats(1);
  for (let k = 0; k < 16; k++) g.update(0.25);
  g.requestStats(1);
  const sel = g.readStats().selected;
  const d = g.downloadSoA();
  const f = await g.readbackFieldData();
  const b = g.getFieldBuffers('higgs');
  g.removeParticle(2, 3.0, 0.0);
  g.requestStats(-1);
  const st = g.readStats();
  console.log(JSON.stringify({ slot: sel.slot, valid: sel.forcesValid, posX: sel.posX, x1: d.x[1],
    grav: sel.forceGravity, higgs: sel.forceHiggs, keys: Object.keys(sel),
    hLen: f.higgsField.length, hType: f.higgsField.constructor.name, axion: f.axionField,
    hMin: Math.min(...f.higgsField), hMax: Math.max(...f.higgsField), bLen: b.field.length,
    same: b.field.every((v, i) => Math.fround(v) === f.higgsField[i]), dead: st.deadCount, selNull: st.selected === null }));
}
main().catch(e => { console.error(e); process.exit(1); });
""" % module
    print("case: node selection and field readback", flush=True)
    r = subprocess.run([NODE, "--input-type=module", "-e", js], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)
    assert o["slot"] == 1 and o["valid"] and o["posX"] == o["x1"]
    assert o["grav"]["x"] < 0 and np.isfinite(o["higgs"]["x"])   # pulled toward the nearer neighbour at -x
    for k in ("posX", "velWX", "mass", "angVel", "magMoment", "forceGravity", "forceSpinCurv", "forceAxion"):
        assert k in o["keys"]
    assert o["hLen"] == 32 * 32 and o["hType"] == "Float32Array" and o["axion"] is None
    assert 0.5 < o["hMin"] <= o["hMax"] < 1.5                    # about the vacuum value 1
    assert o["bLen"] == 32 * 32 and o["same"]
    assert o["dead"] == 0 and o["selNull"]                       # no signal delay: nothing stays behind
