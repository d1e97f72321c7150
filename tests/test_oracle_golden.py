"""The FP64 oracle against what the reference's own CPU physics computes.

tests/golden/reference_*.npz hold seeded inputs and the outputs of the
reference's Physics.update (integrator.js:726-1825), with HiggsField /
AxionField.update (higgs-field.js:39-113, axion-field.js:48-124) and
computeEnergies (energy.js:11-195), run under Node by
tests/golden/make_reference_goldens.py.  They pin the parts of the oracle no
closed form fixes: the composite substep sequence (kick order, H8 a_max
reuse, PE halving, the 1PN velocity Verlet), every pairForce term in
combination (forces.js:163-442), signal-delayed leaves after the history
warm-up, the field KDK with self-gravity and portal terms, the periodic walk
over ghost images, and the bounce walls (Hertz force, friction, contact torque,
clamp).

Each case runs the oracle on the fixture's inputs with the reference's
constants (softening 8, theta 0.5, capacity 4, 64^2 field grid): the forces of
the bootstrap pass (update(0)), the substep count of every update, and the
state, forces, energies and fields after the last update must agree to FP64
rounding (the tolerances leave room for the different summation orders a C
compiler may pick, amplified by the trajectory).
"""
import glob
import json
import os

import numpy as np
import pytest

from physsim_amd import _abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "reference_*.npz")))


def _load(path):
    z = np.load(path)
    meta = json.loads(bytes(z["meta"]).decode())
    return meta, z


def _run(oracle_lib, meta, z):
    t = 0
    for name, bit in A.TOGGLE_BITS.items():
        if meta["toggles"].get(name):
            t |= bit
    # GHOSTS: the reference's periodic walk counts every ghost image (integrator.js:201-258)
    prm = oracle_lib.default_params(toggles0=t, boundary_mode=meta["boundaryMode"], topology=meta["topology"],
                                    collision_mode=meta["collisionMode"], domain_w=meta["W"], domain_h=meta["H"],
                                    field_grid=64, periodic_mode=A.PERIODIC_GHOSTS)
    n = len(z["in_x"])
    o = oracle_lib.Oracle(prm, n)
    o.upload(z["in_x"], z["in_y"], z["in_wx"], z["in_wy"], z["in_mass"], z["in_charge"], z["in_angw"])
    if "in_field_higgs" in z:
        o.field_set(0, z["in_field_higgs"].reshape(64, 64))
    if "in_field_axion" in z:
        o.field_set(1, z["in_field_axion"].reshape(64, 64))
    return o


FORCE_KEYS = ("fx", "fy", "bz", "bgz", "dbzdx", "dbzdy", "dbgzdx", "dbgzdy", "frame_drag", "tidal")


def _close(got, want, rtol, what):
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want).max() / scale
    assert err <= rtol, "%s: max abs error %.3e of scale %.3e (rel %.2e > %.1e)" % (
        what, np.abs(got - want).max(), scale, err, rtol)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[10:-4] for p in FIXTURES])
def test_oracle_matches_reference_physics(path, oracle_lib):
    meta, z = _load(path)
    o = _run(oracle_lib, meta, z)
    assert o.update(0.0) == 0
    f = o.forces()
    for k in FORCE_KEYS:
        _close(f[k], z["bootstrap_" + k], 1e-12, "bootstrap " + k)
    _close(o.f1pn()[0], z["bootstrap_f1pn_x"], 1e-12, "bootstrap force1PN.x")
    assert f["pe_total"] == pytest.approx(float(z["bootstrap_pe"]), rel=1e-12, abs=1e-300)
    steps = [o.update(meta["dt"]) for _ in range(meta["updates"])]
    assert steps == z["substeps"].tolist()
    d = o.download()
    assert np.array_equal(d["id"], z["final_id"].astype(np.int64))
    tol = 1e-9
    for k in ("x", "y", "wx", "wy", "mass", "angw"):
        _close(d[k], z["final_" + k], tol, "final " + k)
    f = o.forces()
    for k in FORCE_KEYS:
        _close(f[k], z["final_" + k], tol, "final " + k)
    fields_on = "final_higgs_field" in z or "final_axion_field" in z
    if not fields_on:      # potentialEnergy adds the fields' gravPE (integrator.js:1743-1749)
        assert f["pe_total"] == pytest.approx(float(z["final_pe"]), rel=tol)
    st = o.stats()
    assert st["sim_time"] == pytest.approx(float(z["final_sim_time"]), rel=1e-14)
    keys = json.loads(bytes(z["energy_keys"]).decode())
    en = dict(zip(keys, z["energy_values"]))
    # computeEnergies (energy.js:11-195): pe = potentialEnergy + pfiEnergy
    for mine, ref in (("linear_ke", "linearKE"), ("spin_ke", "spinKE"), ("px", "px"), ("py", "py"),
                      ("orbital_ang_mom", "orbitalAngMom"), ("spin_ang_mom", "spinAngMom"), ("pe", "pe"),
                      ("total_mass", "totalMass"), ("com_x", "comX"), ("com_y", "comY"),
                      ("field_energy", "fieldEnergy"), ("field_px", "fieldPx"), ("field_py", "fieldPy"),
                      ("higgs_field_energy", "higgsFieldEnergy"), ("axion_field_energy", "axionFieldEnergy"),
                      ("pfi_energy", "pfiEnergy")):
        assert st[mine] == pytest.approx(en[ref], rel=tol, abs=1e-12), mine
    for w, name in ((0, "higgs"), (1, "axion")):
        if "final_%s_field" % name in z:
            _close(o.field_get(w, "field").ravel(), z["final_%s_field" % name], tol, name + " field")
            _close(o.field_get(w, "field_dot").ravel(), z["final_%s_field_dot" % name], tol, name + " fieldDot")
