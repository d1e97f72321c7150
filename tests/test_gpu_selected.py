"""The selected particle of readStats (gpu-physics.js:4146-4276) and the death
metadata of removeParticle (gpu-physics.js:1936-1946), through the C ABI.

The force types are pinned two ways: each pair type (gravity, Coulomb,
Yukawa, the magnetic dipole force) equals the FP64 oracle's whole force with
only that interaction on, at the same state and opening angle, and all types
together add up to the particle's total force of the same pass.  The learned
far-field share is the difference the network makes to the pass."""
import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def _gpu(st, prm, cap=None):
    from physsim_amd import HipPhysics, default_params
    n = len(st["x"])
    print("case: n=%d toggles=0x%x" % (n, prm["toggles0"]), flush=True)
    g = HipPhysics(default_params(**prm), max_particles=cap or n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return g


def _set(g, **kw):
    p = A.GeonParams.from_buffer_copy(g.params)
    for k, v in kw.items():
        setattr(p, k, v)
    g.set_params(p)


def _v(s, name):
    return np.array([s[name]["x"], s[name]["y"]])


SLOTS = (0, 17, 1234, 4095)


@pytest.mark.parametrize("bh", [True, False])
def test_selected_pair_types_match_single_force_oracle(bh, oracle_lib):
    """forceGravity / forceCoulomb / forceYukawa and the magnetic dipole part
    of forceMagnetic against the oracle with that interaction alone."""
    st, prm, _ = W.config2(4096, open_box=True)
    prm["toggles0"] = A.GRAVITY | A.COULOMB | A.YUKAWA | A.MAGNETIC | A.RELATIVITY | (A.BARNES_HUT if bh else 0)
    sel, bz = {}, {}
    for k in SLOTS:
        # each slot from the first pass after upload, as the oracle's: the tree's
        # dipole moments are the previous pass's (quadtree.js:272-309)
        g = _gpu(st, prm)
        g.requestStats(k)
        g.compute_forces()
        s = g.selected()
        assert s is not None and s["slot"] == k and s["forcesValid"]
        sel[k], bz[k] = s, g.forces(4096)["bz"][k]
        g.destroy()
    for tog, name in ((A.GRAVITY, "forceGravity"), (A.COULOMB, "forceCoulomb"), (A.YUKAWA, "forceYukawa"),
                      (A.MAGNETIC, "forceMagnetic")):
        p1 = dict(prm, toggles0=tog | A.RELATIVITY | (A.BARNES_HUT if bh else 0))
        o = oracle_lib.Oracle(oracle_lib.default_params(**p1), 4096)
        o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        o.compute_forces()
        fo = o.forces()
        for k in SLOTS:
            got = _v(sel[k], name)
            if tog == A.MAGNETIC:       # the display pass's Lorentz force q v x Bz (integrator.js:1760-1763)
                q = sel[k]["charge"]
                got = got - q * bz[k] * np.array([sel[k]["velY"], -sel[k]["velX"]])
            want = np.array([fo["fx"][k], fo["fy"][k]])
            assert np.linalg.norm(got - want) <= 2e-5 * np.linalg.norm(want) + 1e-9, (name, k, got, want)


def test_selected_types_add_up_to_the_pass_total():
    """Gravity, Coulomb, Yukawa, magnetic and gravitomagnetic dipoles, the 1PN
    terms and a uniform external field: the types add up to the force the walk
    gave the particle (the display terms q v x Bz, 4 m v x Bgz taken out)."""
    st, prm, _ = W.config2(4096, open_box=True)
    prm["toggles0"] = (A.GRAVITY | A.COULOMB | A.YUKAWA | A.MAGNETIC | A.GRAVITOMAG | A.ONE_PN | A.RELATIVITY |
                       A.BARNES_HUT)
    prm.update(ext_gravity=0.02, ext_gravity_angle=0.7, ext_electric=0.01, ext_electric_angle=-1.1)
    g = _gpu(st, prm)
    for k in SLOTS:
        g.requestStats(k)
        g.compute_forces()
        s = g.selected()
        f = g.forces(4096)
        d = g.download()
        q, m, vx, vy = d["charge"][k], d["mass"][k], s["velX"], s["velY"]
        ext = m * 0.02 * np.array([np.cos(0.7), np.sin(0.7)]) + q * 0.01 * np.array([np.cos(-1.1), np.sin(-1.1)])
        np.testing.assert_allclose(_v(s, "forceExternal"), ext, rtol=1e-6, atol=1e-12)
        parts = [_v(s, n) for n in ("forceGravity", "forceCoulomb", "forceYukawa", "force1PN", "forceExternal")]
        parts.append(_v(s, "forceMagnetic") - q * f["bz"][k] * np.array([vy, -vx]))
        parts.append(_v(s, "forceGravitomag") - 4 * m * f["bgz"][k] * np.array([vy, -vx]))
        total = np.array([f["fx"][k], f["fy"][k]])
        scale = sum(np.linalg.norm(p) for p in parts)
        assert np.linalg.norm(sum(parts) - total) <= 1e-5 * scale, (k, sum(parts), total)
        np.testing.assert_allclose(_v(s, "force1PN"), [f["f1pn_x"][k], f["f1pn_y"][k]], rtol=0, atol=0)
        assert np.all(_v(s, "forceFarField") == 0)
        assert s["posX"] == d["x"][k] and s["velWY"] == d["wy"][k] and s["mass"] == d["mass"][k]


def test_selected_far_field_share_is_the_learned_correction():
    """With the learned far field the monopole walk's types stay as without it
    and forceFarField carries what the network adds (no reference counterpart)."""
    st, prm = W.plummer_charged(2048, 2)
    prm = dict(prm, theta=0.5)
    g = _gpu(st, prm)
    g.far_collect(1)
    g.far_train(100, batch=4096, seed=5)
    k = 1000
    g.requestStats(k)
    g.compute_forces()
    s0, f0 = g.selected(), g.forces(2048)
    _set(g, far_field=A.FAR_LEARNED)
    g.compute_forces()
    s1, f1 = g.selected(), g.forces(2048)
    for name in ("forceGravity", "forceCoulomb"):
        assert np.array_equal(_v(s0, name), _v(s1, name)), name
    assert np.all(_v(s0, "forceFarField") == 0)
    corr = np.array([f1["fx"][k] - f0["fx"][k], f1["fy"][k] - f0["fy"][k]])
    F = np.hypot(f0["fx"][k], f0["fy"][k])
    assert np.linalg.norm(corr) > 1e-6 * F                        # the network changes this particle's force
    assert np.linalg.norm(_v(s1, "forceFarField") - corr) <= 2e-6 * F


def test_selected_follows_its_slot_through_updates():
    """Morton reordering and field kernels: the record stays with the slot;
    with Higgs and axion fields the types add up to the stored total, the
    fields' gravity inside forceGravity."""
    st, prm, dt, fields = W.config3(8192, grid=64, side=256.0, total_mass=4.0)
    prm["toggles0"] = A.GRAVITY | A.YUKAWA | A.HIGGS | A.AXION | A.BARNES_HUT
    prm.update(ext_gravity=0.01, ext_gravity_angle=2.0)
    g = _gpu(st, prm)
    g.set_field(0, fields["higgs"])
    g.set_field(1, fields["axion"])
    k = 4321
    g.requestStats(k)
    assert g.selected()["forcesValid"] is False                   # no pass since the selection
    for _ in range(3):
        g.update(dt)
    s = g.selected()
    d = g.download()
    f = g.forces(8192)
    assert s["slot"] == k and s["forcesValid"]
    assert (s["posX"], s["posY"], s["mass"]) == (d["x"][k], d["y"][k], d["mass"][k])
    assert np.linalg.norm(_v(s, "forceHiggs")) > 0 and np.linalg.norm(_v(s, "forceAxion")) > 0
    parts = [_v(s, n) for n in ("forceGravity", "forceYukawa", "forceExternal", "forceHiggs", "forceAxion")]
    total = np.array([f["fx"][k], f["fy"][k]])
    assert np.linalg.norm(sum(parts) - total) <= 1e-5 * sum(np.linalg.norm(p) for p in parts)
    g.requestStats(-1)
    assert g.readStats()["selected"] is None


def test_selected_radiation_share():
    """forceRadiation is the last substep's Landau-Lifshitz force (integrator.js:1031,
    1485, 1792): zero for a neutral particle, finite and nonzero for a charged
    one (the force itself is pinned by the radiation tests' momentum budget)."""
    st, prm, _ = W.config2(1024, open_box=True)
    prm["toggles0"] = A.GRAVITY | A.COULOMB | A.RADIATION | A.RELATIVITY | A.BARNES_HUT
    g = _gpu(st, prm)
    charged = int(np.nonzero(st["charge"] != 0)[0][0])
    neutral = int(np.nonzero(st["charge"] == 0)[0][0])
    for k in (charged, neutral):
        g.requestStats(k)
        g.update(1.0 / 128)
        s = g.selected()
        fr = _v(s, "forceRadiation")
        assert np.all(np.isfinite(fr))
        if k == neutral:
            assert np.all(fr == 0)
        else:
            assert np.linalg.norm(fr) > 0


def _pair_with_history(dt=0.25, updates=24):
    """Two light bodies 4 apart, sampled every update for 6 time units: the
    survivor's past light cone meets the other's history, and neither moves
    far (accelerations ~1e-2)."""
    from physsim_amd import HipPhysics, default_params
    prm = dict(toggles0=A.GRAVITY | A.RELATIVITY, boundary_mode=A.BOUND_DESPAWN, domain_w=64.0, domain_h=64.0,
               softening=1.0, history_stride=1)
    g = HipPhysics(default_params(**prm), max_particles=4)
    g.addParticle(30.0, 32.0, 0.0, 0.0, 0.1, 0.0)
    g.addParticle(34.0, 32.0, 0.0, 0.05, 0.2, 0.0, angw=0.1)
    for _ in range(updates):
        g.update(dt)
    return g


def test_remove_particle_death_mass_scales_the_dead_force():
    """The retired particle acts with its death mass (forces.js:145-147): the
    force on the survivor scales with it; 0 keeps the particle's own mass."""
    print("case: removeParticle death metadata", flush=True)
    fx = {}
    for dm in (0.0, 0.2, 0.6):
        g = _pair_with_history()
        g.removeParticle(1, deathMass=dm)
        assert g.stats()["deadCount"] == 1
        g.compute_forces()
        fx[dm] = g.forces(2)["fx"][0]
        g.destroy()
    assert fx[0.2] > 0                                  # the dead source still pulls (+x)
    assert fx[0.0] == fx[0.2]                           # 0: the particle's own mass, 0.2
    assert fx[0.6] / fx[0.2] == pytest.approx(3.0, rel=1e-6)


def test_remove_particle_death_ang_vel_is_the_final_spin():
    g = _pair_with_history()
    r = np.float32(0.2) ** (1.0 / 3.0)                  # radius = cbrt(mass)
    w = 0.3
    g.removeParticle(1, deathMass=0.0, deathAngVel=w)
    h = g.history(1)
    assert h["angw"][-1] == pytest.approx(w / np.sqrt(1 - (w * r) ** 2), rel=1e-6)
    assert h["angw"][-2] == pytest.approx(0.1, rel=1e-2)   # before death: its own spin (tidal torque aside)
    from physsim_amd.geon import GeonError
    with pytest.raises(GeonError):
        g.removeParticle(0, deathMass=-1.0)
    with pytest.raises(GeonError):
        g.removeParticle(0, deathAngVel=1.5 / 0.1 ** (1.0 / 3.0))   # |w| r = 1.5: the surface above c


def test_remove_particle_without_signal_delay_ignores_metadata():
    from physsim_amd import HipPhysics, default_params
    g = HipPhysics(default_params(toggles0=A.GRAVITY), max_particles=4)
    for x in (10.0, 20.0, 30.0):
        g.addParticle(x, 10.0, 0.0, 0.0, 1.0, 0.0)
    g.removeParticle(1, deathMass=5.0, deathAngVel=0.2)
    assert g.stats()["deadCount"] == 0
    assert list(g.download()["alive"]) == [True, False, True]


# ---- against the reference's own per-type vectors ---------------------------
import glob as _glob
import json as _json
import os as _os

_FIX = sorted(_glob.glob(_os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "golden", "reference_*.npz")))
_TYPES = ("forceGravity", "forceCoulomb", "forceMagnetic", "forceGravitomag", "force1PN", "forceSpinCurv",
          "forceRadiation", "forceYukawa", "forceExternal", "forceHiggs", "forceAxion")


@pytest.mark.parametrize("path", _FIX, ids=[_os.path.basename(p)[10:-4] for p in _FIX])
def test_selected_matches_reference_display_forces(path):
    """The reference's per-particle force vectors after its bootstrap pass and
    display pass (Physics.update(0), integrator.js:750-782, 1751-1796), run
    under Node into tests/golden: the GPU's selected particle shows the same
    vector of every type, within FP32 rounding of the pass."""
    from physsim_amd import HipPhysics, default_params
    z = np.load(path)
    meta = _json.loads(bytes(z["meta"]).decode())
    t = 0
    for name, bit in A.TOGGLE_BITS.items():
        if meta["toggles"].get(name):
            t |= bit
    prm = dict(toggles0=t, boundary_mode=meta["boundaryMode"], topology=meta["topology"],
               collision_mode=meta["collisionMode"], domain_w=meta["W"], domain_h=meta["H"], field_grid=64,
               periodic_mode=A.PERIODIC_GHOSTS)
    n = len(z["in_x"])
    ftot = np.hypot(z["bootstrap_fx"], z["bootstrap_fy"])
    print("case: %s n=%d toggles=0x%x" % (meta["name"], n, t), flush=True)
    for k in (0, n // 2, n - 1):
        g = HipPhysics(default_params(**prm), max_particles=n)        # the bootstrap pass of a fresh system
        g.upload(z["in_x"], z["in_y"], z["in_wx"], z["in_wy"], z["in_mass"], z["in_charge"], z["in_angw"])
        for w, f in ((0, "in_field_higgs"), (1, "in_field_axion")):
            if f in z:
                g.set_field(w, z[f].reshape(64, 64))
        g.requestStats(k)
        g.compute_forces()
        s = g.selected()
        for name in _TYPES:
            got = _v(s, name)
            want = np.array([z["bootstrap_%s_x" % name][k], z["bootstrap_%s_y" % name][k]])
            err = np.linalg.norm(got - want)
            assert err <= 3e-5 * np.linalg.norm(want) + 3e-6 * ftot[k] + 1e-12, (meta["name"], k, name, got, want)
        g.destroy()
