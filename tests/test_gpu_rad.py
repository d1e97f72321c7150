"""Radiation reaction on gfx950 against the FP64 oracle (forces.js:198-402,
integrator.js:983-1086, 1567-1740): the analytic jerk of every force pass,
the Landau-Lifshitz kick with its radiated energy and momentum, and the
quadrupole radiation once per update."""
import numpy as np
import pytest

from physsim_amd import _abi as A

pytestmark = pytest.mark.gpu

f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)


def _state(n, seed, W, speed=0.2, qscale=1.0):
    g = np.random.default_rng(seed)
    c = g.uniform(0.3 * W, 0.7 * W, (4, 2))
    k = g.integers(0, 4, n)
    p = c[k] + g.normal(0, 0.06 * W, (n, 2))
    return dict(x=f32(p[:, 0]), y=f32(p[:, 1]), wx=f32(g.normal(0, speed, n)), wy=f32(g.normal(0, speed, n)),
                mass=f32(g.uniform(0.5, 3.0, n)), charge=g.integers(-10, 11, n) * 0.1 * qscale,
                angw=f32(g.uniform(-0.3, 0.3, n)))


def _pair(st, prm, oracle_lib):
    from physsim_amd import HipPhysics, default_params
    n = len(st["x"])
    print("case: radiation n=%d toggles=0x%x" % (n, prm["toggles0"]), flush=True)
    g = HipPhysics(default_params(**prm), max_particles=n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), n)
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return g, o


def _rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("bh", [False, True])
@pytest.mark.parametrize("toggles", [A.GRAVITY | A.COULOMB, A.GRAVITY | A.COULOMB | A.MAGNETIC | A.GRAVITOMAG])
def test_jerk_matches_oracle(bh, toggles, oracle_lib):
    n = 2048 if bh else 768
    st = _state(n, 11, 256.0)
    prm = dict(toggles0=toggles | A.RADIATION | (A.BARNES_HUT if bh else 0), boundary_mode=A.BOUND_DESPAWN,
               domain_w=256.0, domain_h=256.0, signal_delay=A.DELAY_OFF, softening=4.0, theta=0.5)
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    # a second pass on both: the dipole moments are cached by the first (forces.js:71-76)
    g.compute_forces()
    o.compute_forces()
    fg = g.forces(n)
    jx, jy = o.jerk()
    ej = _rel(np.concatenate([fg["jerk_x"], fg["jerk_y"]]), np.concatenate([jx, jy]))
    fo = o.forces()
    ef = _rel(np.concatenate([fg["fx"], fg["fy"]]), np.concatenate([fo["fx"], fo["fy"]]))
    print("bh=%d toggles=0x%x: jerk rel L2 %.3e, force %.3e" % (bh, toggles, ej, ef))
    assert ej < 1e-4 and ef < 2e-5


def test_landau_lifshitz_matches_oracle(oracle_lib):
    # charges with relativity: the LL term on top of the kick, for several updates
    n = 512
    st = _state(n, 12, 128.0, speed=0.4, qscale=3.0)
    prm = dict(toggles0=A.COULOMB | A.RELATIVITY | A.RADIATION, boundary_mode=A.BOUND_DESPAWN, domain_w=128.0,
               domain_h=128.0, signal_delay=A.DELAY_OFF, softening=4.0)
    g, o = _pair(st, prm, oracle_lib)
    for k in range(4):
        assert g.update(0.25) == o.update(0.25), k
    sg, so = g.stats(), o.stats()
    print("radiated %.6e / %.6e, px %.4e / %.4e" % (sg["totalRadiated"], so["radiated_energy"],
                                                     sg["totalRadiatedPx"], so["radiated_px"]))
    assert so["radiated_energy"] > 0
    assert sg["totalRadiated"] == pytest.approx(so["radiated_energy"], rel=2e-3)
    p_o = np.hypot(so["radiated_px"], so["radiated_py"])
    assert np.hypot(sg["totalRadiatedPx"] - so["radiated_px"], sg["totalRadiatedPy"] - so["radiated_py"]) < 5e-3 * p_o + 1e-12
    dg, do = g.download(), o.download()
    ids = do["id"]
    for k in ("x", "y", "wx", "wy"):
        np.testing.assert_allclose(dg[k][ids], do[k], rtol=2e-4, atol=2e-4 * np.abs(do[k]).max(), err_msg=k)


@pytest.mark.parametrize("bh", [False, True])
def test_quadrupole_radiation_matches_oracle(bh, oracle_lib):
    # a cold gravitating cluster with charges: GW and EM quadrupole power once per update
    n = 1024
    st = _state(n, 13, 128.0, speed=0.05)
    prm = dict(toggles0=A.GRAVITY | A.COULOMB | A.RADIATION | (A.BARNES_HUT if bh else 0),
               boundary_mode=A.BOUND_DESPAWN, domain_w=128.0, domain_h=128.0, signal_delay=A.DELAY_OFF,
               softening=4.0, theta=0.5)
    g, o = _pair(st, prm, oracle_lib)
    for k in range(3):
        assert g.update(0.5) == o.update(0.5), k
    sg, so = g.stats(), o.stats()
    print("bh=%d radiated %.6e / %.6e" % (bh, sg["totalRadiated"], so["radiated_energy"]))
    assert so["radiated_energy"] > 0
    assert sg["totalRadiated"] == pytest.approx(so["radiated_energy"], rel=5e-3)
    dg, do = g.download(), o.download()
    ids = do["id"]
    for k in ("wx", "wy"):
        np.testing.assert_allclose(dg[k][ids], do[k], rtol=5e-4, atol=5e-4 * np.abs(do[k]).max(), err_msg=k)


def test_radiation_refuses_yukawa():
    from physsim_amd import HipPhysics, default_params
    from physsim_amd.geon import GeonError
    with pytest.raises(GeonError):
        HipPhysics(default_params(toggles0=A.GRAVITY | A.YUKAWA | A.RADIATION), max_particles=16)
