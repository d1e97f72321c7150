"""The reference's v1 save state (save-load.js:89-262, gpu-physics.js:3749-3960)
in both hosts: the Python mirror (physsim_amd/state_v1.py) and the JS module
the app would load (physsim_amd/js/state-v1.mjs) read the same fixture to the
same parameters and particle columns, and write back what they read.  The
fixture follows the layout _cpuSaveState writes, with the missing and falsy
fields the GPU backend's deserialize fills in (mass || 1, baseMass ?? mass,
charge || 0, angw || 0, wx || 0)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import state_v1 as S
from physsim_amd.geon import default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "state_v1_sample.json")
MODULE = os.path.join(ROOT, "physsim_amd", "js", "state-v1.mjs")
NODE = shutil.which("node")


def _state():
    with open(FIXTURE) as f:
        return json.load(f)


def test_toggles_parameters_and_modes():
    p = S.apply_state(default_params(toggles0=0), _state())
    # bits of gpu-constants.js:196-213 for the toggles the fixture turns on
    want = 1 | 2 | 4 | 32 | 16384 | 64 | 2048
    assert p.toggles0 == want
    assert (p.collision_mode, p.boundary_mode, p.topology) == (A.COL_MERGE, A.BOUND_LOOP, A.TORUS)
    assert (p.yukawa_mu, p.higgs_mass, p.ext_gravity, p.ext_bz) == (0.2, 0.6, 0.01, 0.002)
    assert p.ext_gravity_angle == pytest.approx(np.pi / 2)


def test_missing_keys_keep_current_values():
    base = default_params(toggles0=A.HIGGS | A.GRAVITY, yukawa_mu=0.3, boundary_mode=A.BOUND_LOOP)
    p = S.apply_state(base, {"version": 1, "particles": [], "toggles": {"gravityEnabled": False},
                             "yukawaMu": None, "extBz": None})
    assert p.toggles0 == A.HIGGS                      # only the listed toggle changes
    assert p.yukawa_mu == 0.3                         # `!= null` keys skip null
    assert p.ext_bz == 0.0                            # `!== undefined` keys take null as 0
    assert p.boundary_mode == A.BOUND_LOOP            # no settings: modes stay


def test_legacy_repel_reads_as_bounce():
    p = S.apply_state(default_params(), {"settings": {"collision": "repel"}})
    assert p.collision_mode == A.COL_BOUNCE           # save-load.js:194
    p = S.apply_state(default_params(), {"settings": {}})
    assert (p.collision_mode, p.boundary_mode, p.topology) == (A.COL_PASS, A.BOUND_DESPAWN, A.TORUS)


def test_particle_defaults_of_the_gpu_backend():
    c = S.particles_from_state(_state(), 16)
    np.testing.assert_array_equal(c["mass"], [2.0, 1.5, 1.0, 4.0])          # mass 0 -> 1
    np.testing.assert_array_equal(c["base_mass"], [2.0, 1.5, 3.0, 4.5])     # missing -> mass
    np.testing.assert_array_equal(c["charge"], [0.3, -0.2, 0.0, 0.0])       # null, missing -> 0
    np.testing.assert_array_equal(c["wx"], [0.125, -0.5, 0.0, 0.0])
    np.testing.assert_array_equal(c["angw"], [0.05, 0.0, 0.0, -0.125])
    np.testing.assert_array_equal(c["antimatter"], [0, 1, 0, 0])
    assert len(S.particles_from_state(_state(), 2)["x"]) == 2               # capacity cut


def test_write_back_what_was_read():
    s = _state()
    p = S.apply_state(default_params(), s)
    out = S.params_to_state(p, s["particles"], {"speedScale": 32, "friction": 0.4, "camera": s["camera"]})
    for k in ("version", "toggles", "settings", "camera") + tuple(k for k, _, _ in S.PARAM_KEYS):
        assert out[k] == s[k], k
    assert list(out["toggles"]) == S.TOGGLE_KEYS


@pytest.mark.skipif(NODE is None, reason="node not installed")
def test_js_module_reads_the_same():
    script = """
import { applyState, particlesFromState, paramsToState } from %r;
import fs from 'fs';
const s = JSON.parse(fs.readFileSync(%r, 'utf8'));
const p = applyState({ toggles0: 0 }, s);
const c = particlesFromState(s, 16);
const back = paramsToState(p, s.particles, { speedScale: 32, friction: 0.4, camera: s.camera });
const arr = a => Array.from(a);
console.log(JSON.stringify({ p, c: { mass: arr(c.mass), baseMass: arr(c.baseMass), charge: arr(c.charge),
                                     wx: arr(c.wx), angw: arr(c.angw), antimatter: arr(c.antimatter) },
                             back }));
""" % ("file://" + MODULE, FIXTURE)
    r = subprocess.run([NODE, "--input-type=module", "-e", script], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    py = S.apply_state(default_params(toggles0=0), _state())
    assert js["p"]["toggles0"] == py.toggles0
    assert (js["p"]["collisionMode"], js["p"]["boundaryMode"], js["p"]["topology"]) == \
        (py.collision_mode, py.boundary_mode, py.topology)
    for k, f, _ in S.PARAM_KEYS:
        assert js["p"][k if k != "hubbleParam" else "hubbleParam"] == getattr(py, f), k
    c = S.particles_from_state(_state(), 16)
    for jk, pk in (("mass", "mass"), ("baseMass", "base_mass"), ("charge", "charge"), ("wx", "wx"),
                   ("angw", "angw"), ("antimatter", "antimatter")):
        assert js["c"][jk] == list(c[pk]), jk
    s = _state()
    for k in ("toggles", "settings", "camera"):
        assert js["back"][k] == s[k], k
