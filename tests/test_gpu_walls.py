"""Boundary mode bounce on the GPU (through the C ABI): the Hertz wall forces
with friction and their contact torque (integrator.js:606-667, 964-981) and
the once-per-update clamp (integrator.js:1812-1818), against the FP64 oracle
(itself pinned by closed forms in test_oracle.py)."""
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
