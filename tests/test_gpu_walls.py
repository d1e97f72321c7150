"""Boundary mode bounce and collision mode bounce on the GPU (through the C
ABI): the Hertz wall forces with friction and their contact torque
(integrator.js:606-667, 964-981), the once-per-update clamp
(integrator.js:1812-1818), and the Hertz contacts between bodies
(integrator.js:669-724), against closed forms, the reference's own forces and
the FP64 oracle (pinned to the reference by the fixtures in tests/golden)."""
import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def _pair(st, prm, oracle_lib):
    from physsim_amd import HipPhysics, default_params
    n = len(st["x"])
    print("case: bounce n=%d toggles=0x%x friction=%.2f" % (n, prm["toggles0"], prm.get("bounce_friction", 0.4)),
          flush=True)
    g = HipPhysics(default_params(**prm), max_particles=n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), n)
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return g, o


def _box(n, side, seed):
    """n bodies in a side x side box, a quarter of them pressed into the walls."""
    r = np.random.default_rng(seed)
    m = W.f32(r.lognormal(0, 0.5, n))
    rad = np.cbrt(m)
    x = r.uniform(0, side, n)
    y = r.uniform(0, side, n)
    k = n // 4
    x[:k // 2] = rad[:k // 2] * r.uniform(0.3, 0.9, k // 2)                # into the left wall
    y[k // 2:k] = side - rad[k // 2:k] * r.uniform(0.3, 0.9, k - k // 2)   # into the bottom wall
    return dict(x=W.f32(x), y=W.f32(y), wx=W.f32(r.normal(0, 0.05, n)), wy=W.f32(r.normal(0, 0.05, n)),
                mass=m, charge=W.f32(np.round(r.uniform(-1, 1, n) * 10) / 10),
                angw=W.f32(r.uniform(-0.2, 0.2, n)))


@pytest.mark.parametrize("bh", [False, True])
def test_wall_forces_match_oracle(bh, oracle_lib):
    st = _box(2048, 128.0, 3)
    prm = dict(toggles0=A.GRAVITY | A.COULOMB | (A.BARNES_HUT if bh else 0), boundary_mode=A.BOUND_BOUNCE,
               domain_w=128.0, domain_h=128.0, softening=1.0)
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(2048), o.forces()
    # the pressed bodies: their wall force dominates
    k = np.arange(512)
    err = np.hypot(fg["fx"][k] - fo["fx"][k], fg["fy"][k] - fo["fy"][k]) / np.hypot(fo["fx"][k], fo["fy"][k])
    print("pressed bodies: max rel force error %.3e" % err.max())
    assert err.max() < 2e-5
    d = np.hypot(fg["fx"] - fo["fx"], fg["fy"] - fo["fy"])
    assert np.sqrt(np.sum(d ** 2) / np.sum(fo["fx"] ** 2 + fo["fy"] ** 2)) < 2e-5


def _sparse_box(side=256.0, seed=5):
    """32 bodies pressed into each wall, 8 apart along it (far enough that the
    tidal torque stays small beside the contact torque), 128 inside."""
    r = np.random.default_rng(seed)
    along = np.linspace(12.0, side - 12.0, 32)
    m = W.f32(r.lognormal(0, 0.3, 256))
    rad = np.cbrt(m)
    depth = rad[:128] * r.uniform(0.2, 0.6, 128)
    x = np.concatenate([depth[:32], side - depth[32:64], along, along, r.uniform(20, side - 20, 128)])
    y = np.concatenate([along, along, depth[64:96], side - depth[96:128], r.uniform(20, side - 20, 128)])
    return dict(x=W.f32(x), y=W.f32(y), wx=W.f32(r.normal(0, 0.05, 256)), wy=W.f32(r.normal(0, 0.05, 256)),
                mass=m, charge=np.zeros(256), angw=W.f32(r.uniform(-0.1, 0.1, 256)))


def test_wall_contact_torque_and_clamp_match_oracle(oracle_lib):
    """Updates with friction: the contact torque turns the pressed bodies'
    spins (gravity on, so the torque step runs), and the clamp keeps every
    body inside the box."""
    side = 256.0
    st = _sparse_box(side)
    prm = dict(toggles0=A.GRAVITY, boundary_mode=A.BOUND_BOUNCE, domain_w=side, domain_h=side, softening=1.0)
    g, o = _pair(st, prm, oracle_lib)
    dt = 1.0 / 64
    for step in range(8):
        assert g.update(dt) == o.update(dt), step
    dg, do = g.download(), o.download()
    ids = do["id"]
    spin0 = st["angw"][ids]
    dspin_g, dspin_o = dg["angw"][ids] - spin0, do["angw"] - spin0
    pressed = ids < 128
    assert np.median(np.abs(dspin_o[pressed])) > 10 * np.median(np.abs(dspin_o[~pressed]))
    np.testing.assert_allclose(dspin_g[pressed], dspin_o[pressed], rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(dg["x"][ids], do["x"], rtol=0, atol=2e-4)
    np.testing.assert_allclose(dg["y"][ids], do["y"], rtol=0, atol=2e-4)
    assert dg["x"].min() >= 0 and dg["x"].max() <= side and dg["y"].min() >= 0 and dg["y"].max() <= side


def test_selected_external_share_holds_the_wall_force():
    """forceExternal gets a wall's normal force, not its friction (integrator.js:619-658)."""
    from physsim_amd import HipPhysics, default_params
    prm = dict(toggles0=A.GRAVITY, boundary_mode=A.BOUND_BOUNCE, domain_w=64.0, domain_h=64.0, softening=1.0)
    g = HipPhysics(default_params(**prm), max_particles=2)
    g.addParticle(1.5, 32.0, 0.0, 0.3, 8.0, 0.0)        # radius 2: 0.5 into the left wall
    g.requestStats(0)
    g.compute_forces()
    s = g.selected()
    Fn = 0.5 ** 1.5
    assert s["forceExternal"]["x"] == pytest.approx(Fn, rel=1e-6) and s["forceExternal"]["y"] == 0
    f = g.forces(1)
    assert f["fx"][0] == pytest.approx(Fn, rel=1e-6) and f["fy"][0] == pytest.approx(-0.4 * Fn, rel=1e-6)


# ---- collision mode bounce: Hertz contacts between bodies ------------------

def test_contact_pair_closed_form():
    """Two bodies of mass 8 (radius 2) 1.5 apart overlap by 2.5: each feels
    2.5^1.5 along the line of centres, away from the other, and the friction
    of their tangential velocities (integrator.js:687-724)."""
    from physsim_amd import HipPhysics, default_params
    prm = dict(toggles0=A.GRAVITY, collision_mode=A.COL_BOUNCE, domain_w=64.0, domain_h=64.0, softening=1.0,
               bounce_friction=0.0)
    g = HipPhysics(default_params(**prm), max_particles=2)
    g.addParticle(30.0, 32.0, 0.0, 0.0, 8.0, 0.0)
    g.addParticle(31.5, 32.0, 0.0, 0.0, 8.0, 0.0)
    g.compute_forces()
    f = g.forces(2)
    grav = 64.0 * 1.5 / (1.5 ** 2 + 1.0) ** 1.5        # the gravity pair force, softening 1
    Fn = 2.5 ** 1.5
    assert f["fx"][0] == pytest.approx(grav - Fn, rel=1e-6) and f["fx"][1] == pytest.approx(Fn - grav, rel=1e-6)
    assert f["fy"][0] == 0 and f["fy"][1] == 0
    # friction: body 2 slides up along the contact (t = (0, 1) seen from body 1), vRel = -0.3
    prm["bounce_friction"] = 0.4
    g = HipPhysics(default_params(**prm), max_particles=2)
    g.addParticle(30.0, 32.0, 0.0, 0.0, 8.0, 0.0)
    g.addParticle(31.5, 32.0, 0.0, 0.3, 8.0, 0.0)
    g.compute_forces()
    f = g.forces(2)
    Ft = -0.4 * Fn * -1.0                               # clamp(10 * -0.3) = -1
    assert f["fy"][0] == pytest.approx(Ft, rel=1e-6) and f["fy"][1] == pytest.approx(-Ft, rel=1e-6)


@pytest.mark.parametrize("case", ["bounce_collisions", "bounce_collisions_torus"])
def test_contacts_match_reference_bootstrap_and_oracle_trajectory(case, oracle_lib):
    """The reference's own bootstrap forces of a crowded box (contacts dominate
    them) and, over 10 updates, the oracle, which the same fixture pins to the
    reference to FP64 rounding."""
    import json
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_%s.npz" % case))
    meta = json.loads(bytes(z["meta"]).decode())
    t = 0
    for name, bit in A.TOGGLE_BITS.items():
        if meta["toggles"].get(name):
            t |= bit
    # the reference's periodic walk counts the ghost images (A1): the image-particle mode
    prm = dict(toggles0=t, boundary_mode=meta["boundaryMode"], topology=meta["topology"],
               collision_mode=meta["collisionMode"], domain_w=meta["W"], domain_h=meta["H"],
               periodic_mode=A.PERIODIC_GHOSTS)
    st = {k: z["in_" + k] for k in ("x", "y", "wx", "wy", "mass", "charge", "angw")}
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    n = len(st["x"])
    f = g.forces(n)
    fref = np.hypot(z["bootstrap_fx"], z["bootstrap_fy"])
    err = np.hypot(f["fx"] - z["bootstrap_fx"], f["fy"] - z["bootstrap_fy"])
    print("%s: bootstrap max |dF| %.3e of max |F| %.3e" % (case, err.max(), fref.max()))
    # FP32: the overlap delta = r1 + r2 - |d| loses digits to cancellation (positions ~30, ulp 2e-6)
    assert err.max() <= 3e-5 * fref.max()
    for step in range(10):
        assert g.update(meta["dt"]) == o.update(meta["dt"]), step
    dg, do = g.download(), o.download()
    ids = do["id"]
    np.testing.assert_allclose(dg["x"][ids], do["x"], rtol=0, atol=1e-4)
    np.testing.assert_allclose(dg["y"][ids], do["y"], rtol=0, atol=1e-4)
    np.testing.assert_allclose(dg["angw"][ids], do["angw"], rtol=0, atol=1e-4)
