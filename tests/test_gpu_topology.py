"""Klein bottle and RP^2 (topology.js) on gfx950 against the FP64 oracle, in
min-image mode (no ghosts, every source once): forces through the walk and
the direct sum, trajectories whose particles cross the glide edges
(wrapPosition flips), signal-delayed sources, and the scalar fields with
their glide neighbours and Green's function."""
import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import workloads as W

pytestmark = pytest.mark.gpu

TOPOS = [A.KLEIN, A.RP2]


def _state(n, seed, Wd, Hd, wsig=0.05):
    g = np.random.default_rng(seed)
    f = lambda a: np.asarray(a, np.float32).astype(np.float64)
    return dict(x=f(g.random(n) * Wd), y=f(g.random(n) * Hd), wx=f(g.normal(0, wsig, n)), wy=f(g.normal(0, wsig, n)),
                mass=f(g.lognormal(0, 0.5, n)), charge=f(g.integers(-10, 11, n) * 0.1),
                angw=f(g.uniform(-0.5, 0.5, n)))


def _prm(topo, Wd, Hd, bh=True, toggles=A.GRAVITY | A.COULOMB | A.YUKAWA | A.MAGNETIC, **kw):
    p = dict(toggles0=toggles | (A.BARNES_HUT if bh else 0), boundary_mode=A.BOUND_LOOP, topology=topo,
             periodic_mode=A.PERIODIC_MIN_IMAGE, domain_w=Wd, domain_h=Hd, softening=4.0, theta=0.5, yukawa_mu=1.0,
             signal_delay=A.DELAY_OFF)
    p.update(kw)
    return p


def _pair(st, prm, oracle_lib):
    from physsim_amd import HipPhysics, default_params
    n = len(st["x"])
    print("case: n=%d topology=%d toggles=0x%x" % (n, prm["topology"], prm["toggles0"]), flush=True)
    g = HipPhysics(default_params(**prm), max_particles=n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), n)
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return g, o


def _rel_l2(ax, ay, bx, by):
    return float(np.sqrt(np.sum((ax - bx) ** 2 + (ay - by) ** 2) / np.sum(bx ** 2 + by ** 2)))


@pytest.mark.parametrize("topo", TOPOS)
@pytest.mark.parametrize("bh", [True, False])
def test_forces_match_oracle(topo, bh, oracle_lib):
    n = 4096 if bh else 1024
    st = _state(n, 10 + topo, 96.0, 64.0)
    prm = _prm(topo, 96.0, 64.0, bh=bh)
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(n), o.forces()
    err = _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"])
    ebz = np.sqrt(np.sum((fg["bz"] - fo["bz"]) ** 2) / np.sum(fo["bz"] ** 2))
    print("topology %d bh %d: force %.3e Bz %.3e" % (topo, bh, err, ebz))
    assert err < 2e-5 and ebz < 1e-4
    assert 0.5 * fg["pe"].sum() == pytest.approx(fo["pe_total"], rel=2e-5)
    if bh:
        assert g.counters()["pair_interactions"] == o.counters()["pair"]


@pytest.mark.parametrize("topo", TOPOS)
def test_glide_crossings_match_oracle(topo, oracle_lib):
    # light particles streaming across every edge: wrapPosition's mirror and
    # w / v / spin flips (topology.js:81-135) on both backends
    n = 1024
    Wd, Hd = 64.0, 48.0
    st = _state(n, 20 + topo, Wd, Hd, wsig=0.6)
    st["mass"] = st["mass"] * 1e-3
    st["charge"][:] = 0.0
    prm = _prm(topo, Wd, Hd, toggles=A.GRAVITY | A.RELATIVITY)
    g, o = _pair(st, prm, oracle_lib)
    for k in range(6):
        assert g.update(1.0) == o.update(1.0), k
    dg, do = g.download(), o.download()
    ids = do["id"]
    flipped = np.sign(do["angw"]) != np.sign(np.asarray(st["angw"])[ids])
    print("topology %d: %d spins flipped by glide crossings" % (topo, flipped.sum()))
    assert flipped.sum() > 20
    for k in ("x", "y", "wx", "wy", "angw"):
        np.testing.assert_allclose(dg[k][ids], do[k], rtol=0, atol=1e-4, err_msg=k)


@pytest.mark.parametrize("topo", TOPOS)
def test_delayed_forces_on_glide_surfaces(topo, oracle_lib):
    # light-cone sources seen through the glide images (signal-delay.js uses minImage throughout)
    n = 512
    Wd, Hd = 48.0, 40.0
    st = _state(n, 30 + topo, Wd, Hd, wsig=0.3)
    prm = _prm(topo, Wd, Hd, bh=False, toggles=A.GRAVITY | A.COULOMB | A.RELATIVITY, signal_delay=A.DELAY_FOLLOW,
               history_stride=1)
    from physsim_amd import HipPhysics, default_params
    g = HipPhysics(default_params(**prm), max_particles=n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    for _ in range(16):
        g.update(0.25)
    d = g.download()
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), n)
    o.upload(d["x"], d["y"], d["wx"], d["wy"], d["mass"], d["charge"], d["angw"])
    for s in range(n):
        o.set_history(s, g.history(s))
    o.set_time(g.stats()["simTime"])
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(n), o.forces()
    err = _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"])
    print("topology %d delayed: force %.3e" % (topo, err))
    assert err < 2e-5


@pytest.mark.parametrize("topo", TOPOS)
def test_fields_on_glide_surfaces(topo, oracle_lib):
    # deposit, KDK with self-gravity (Green's function from minImage of the
    # grid offsets), gradients and field forces with the glide neighbours of _nb
    from physsim_amd import HipPhysics, default_params
    n, G = 2048, 64
    st, prm, dt, fields = W.config3(n, grid=G, side=256.0, total_mass=4.0)
    prm = dict(prm, boundary_mode=A.BOUND_LOOP, topology=topo, periodic_mode=A.PERIODIC_MIN_IMAGE)
    print("case: fields n=%d topology=%d" % (n, topo), flush=True)
    g = HipPhysics(default_params(**prm), max_particles=n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    g.set_field(0, fields["higgs"])
    g.set_field(1, fields["axion"])
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), n)
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    o.field_set(0, fields["higgs"])
    o.field_set(1, fields["axion"])
    for step in range(3):
        assert g.update(dt) == o.update(dt), step
    rel = lambda a, b: float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
    for which in (0, 1):
        dev = g.get_field(which, "field") - (1.0 if which == 0 else 0.0)
        odev = o.field_get(which, "field") - (1.0 if which == 0 else 0.0)
        e = rel(dev, odev)
        print("topology %d field %d: %.3e" % (topo, which, e))
        assert e < 2e-3, which
        assert rel(g.get_field(which, "sg_phi"), o.field_get(which, "sg_phi")) < 1e-4, which
