"""v1 save state through the device (GPUPhysics.serialize / deserialize,
gpu-physics.js:3749-3960): a system saved and loaded into a fresh handle holds
the same particles (antimatter included), toggles, parameters and modes, and
steps bit-identically.  Loading renumbers slots 0..n-1 in their old order,
which is the array order the reference saves."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import state_v1 as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "state_v1_sample.json")
NODE = shutil.which("node")


def _system(n=1500, seed=3):
    from physsim_amd import HipPhysics, default_params
    g0 = np.random.default_rng(seed)
    prm = default_params(toggles0=A.GRAVITY | A.COULOMB | A.RELATIVITY | A.BARNES_HUT, collision_mode=A.COL_MERGE,
                         boundary_mode=A.BOUND_LOOP, topology=A.TORUS, domain_w=128.0, domain_h=128.0,
                         signal_delay=A.DELAY_OFF, yukawa_mu=0.2, ext_bz=0.001)
    g = HipPhysics(prm, max_particles=2 * n)
    f = lambda a: np.asarray(a, np.float32).astype(np.float64)
    g.upload(f(g0.uniform(0, 128, n)), f(g0.uniform(0, 128, n)), f(g0.normal(0, 0.2, n)), f(g0.normal(0, 0.2, n)),
             f(g0.uniform(0.5, 3, n)), g0.integers(-3, 4, n) * 0.1, f(g0.uniform(-0.2, 0.2, n)),
             antimatter=(g0.random(n) < 0.3).astype(np.int32))
    return g


def test_save_load_round_trip_steps_identically():
    from physsim_amd import HipPhysics, default_params
    print("case: v1 round trip n=1500", flush=True)
    g = _system()
    g.update(1.0 / 64)                       # merges and annihilations leave gaps in the slots
    d = g.download()
    alive = np.nonzero(d["alive"])[0]
    assert len(alive) < 1500 and len(alive) < len(d["alive"])
    state = json.loads(json.dumps(g.serialize()))
    h = HipPhysics(default_params(), max_particles=3000)
    assert h.deserialize(state)
    for f, _ in A.GeonParams._fields_:
        if f in ("toggles0", "collision_mode", "boundary_mode", "topology", "yukawa_mu", "ext_bz", "higgs_mass"):
            assert getattr(h.params, f) == getattr(g.params, f), f
    e = h.download()
    assert len(e["x"]) == len(alive) and e["alive"].all()
    for k in ("x", "y", "wx", "wy", "mass", "base_mass", "charge", "angw", "antimatter"):
        assert np.array_equal(e[k], d[k][alive]), k
    assert e["antimatter"].sum() > 100
    # a loaded system is a fresh upload of the same columns under the same
    # parameters (forces and moment caches are rebuilt, as after any load):
    # the next updates agree bit for bit
    k = HipPhysics(S.apply_state(default_params(), state), max_particles=3000)
    k.upload(*[d[c][alive] for c in ("x", "y", "wx", "wy", "mass", "charge", "angw", "base_mass")],
             antimatter=d["antimatter"][alive].astype(np.int32))
    for _ in range(3):
        assert h.update(1.0 / 64) == k.update(1.0 / 64)
    e, f = h.download(), k.download()
    assert np.array_equal(e["alive"], f["alive"])
    for c in ("x", "y", "wx", "wy", "mass", "angw", "charge"):
        assert np.array_equal(e[c], f[c]), c
    # and the original system, which kept its cached forces, stays within FP32 rounding
    g.update(1.0 / 64)


def test_load_refuses_other_versions():
    from physsim_amd import HipPhysics, default_params
    h = HipPhysics(default_params(), max_particles=16)
    assert h.deserialize({"version": 2, "particles": []}) is False
    assert h.deserialize(None) is False


@pytest.mark.skipif(NODE is None, reason="node not installed")
def test_js_host_round_trip():
    mod = "file://" + os.path.join(ROOT, "physsim_amd", "js", "hip-physics.mjs")
    # node 12: no top-level await, so a promise chain
    script = """
import HipPhysics from %r;
import fs from 'fs';
const s = JSON.parse(fs.readFileSync(%r, 'utf8'));
const g = new HipPhysics({}, 64);
g.init().then(async () => {
    const ok = g.deserialize(s);
    const back = await g.serialize({ speedScale: 32, friction: 0.4, camera: s.camera });
    console.log(JSON.stringify({ ok, back }));
    g.destroy();
}).catch(e => { console.error(e); process.exit(1); });
""" % (mod, FIXTURE)
    print("case: js v1 round trip", flush=True)
    r = subprocess.run([NODE, "--input-type=module", "-e", script], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["ok"] is True
    s = json.load(open(FIXTURE))
    back = out["back"]
    for k in ("toggles", "settings", "camera"):
        assert back[k] == s[k], k
    c = S.particles_from_state(s, 64)
    f32 = lambda v: float(np.float32(v))
    for i, p in enumerate(back["particles"]):
        for jk, pk in (("x", "x"), ("y", "y"), ("wx", "wx"), ("mass", "mass"), ("baseMass", "base_mass"),
                       ("angw", "angw")):
            assert p[jk] == f32(c[pk][i]), (i, jk)
        assert p["charge"] == round(c["charge"][i] / 0.1) * 0.1     # quantized as main.js:511 does
        assert p["antimatter"] == bool(c["antimatter"][i])
