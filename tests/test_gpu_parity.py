"""Parity of the gfx950 backend (through the C ABI) against the FP64 oracle.

Every case names its periodic mode: all periodic cases here run min-image mode
(each source counted once, forces.js:93-126 semantics), the only mode the
device path runs.
"""
import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def _pair(st, prm, oracle_lib, cap=None):
    from physsim_amd import HipPhysics, default_params
    n = len(st["x"])
    print("case: n=%d toggles=0x%x" % (n, prm["toggles0"]), flush=True)
    g = HipPhysics(default_params(**prm), max_particles=cap or n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), n)
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return g, o


def _rel_l2(ax, ay, bx, by):
    return float(np.sqrt(np.sum((ax - bx) ** 2 + (ay - by) ** 2) / np.sum(bx ** 2 + by ** 2)))


def _canon(t):
    keys = np.stack([t["level"].astype(np.int64), t["path"].astype(np.int64)], 1)
    order = np.lexsort((keys[:, 1], keys[:, 0]))
    return t[order]


@pytest.mark.parametrize("n", [5, 64, 1024, 4096])
def test_tree_structure_bit_exact(n, oracle_lib):
    st, prm, _ = W.config1(n)
    g, o = _pair(st, prm, oracle_lib)
    g.build_tree()
    tg = _canon(g.tree())
    to = _canon(o.tree())
    assert len(tg) == len(to)
    for f in ("level", "path", "divided", "count", "charge_quanta"):
        assert np.array_equal(tg[f], to[f]), f
    np.testing.assert_allclose(tg["mass"], to["mass"], rtol=1e-12, atol=1e-12)
    m = to["mass"] > 0
    np.testing.assert_allclose(tg["com_x"][m], to["com_x"][m], rtol=0, atol=1e-9)
    np.testing.assert_allclose(tg["com_y"][m], to["com_y"][m], rtol=0, atol=1e-9)
    assert g.counters()["depth_overflow"] == 0


def test_tree_ties_on_split_planes(oracle_lib):
    """Points exactly on split planes go to the lower child (quadtree.js:175-179)."""
    W_ = 64.0
    xs = np.array([32.0, 32.0, 16.0, 48.0, 32.0, 0.0, 64.0, 8.0, 24.0, 40.0, 32.0, 16.0])
    ys = np.array([32.0, 16.0, 32.0, 32.0, 48.0, 0.0, 64.0, 32.0, 32.0, 32.0, 8.0, 16.0])
    n = len(xs)
    st = dict(x=xs, y=ys, wx=np.zeros(n), wy=np.zeros(n), mass=np.ones(n), charge=np.zeros(n), angw=np.zeros(n))
    prm = dict(toggles0=A.GRAVITY | A.BARNES_HUT, domain_w=W_, domain_h=W_)
    g, o = _pair(st, prm, oracle_lib)
    g.build_tree()
    tg, to = _canon(g.tree()), _canon(o.tree())
    assert len(tg) == len(to)
    for f in ("level", "path", "divided", "count"):
        assert np.array_equal(tg[f], to[f]), f


def test_points_outside_root_box_not_inserted(oracle_lib):
    st, prm, _ = W.config1(256)
    st["x"][:3] = [-200.0, 300.0, 10.0]
    st["y"][:3] = [10.0, 10.0, 500.0]
    g, o = _pair(st, prm, oracle_lib)
    g.build_tree()
    tg, to = _canon(g.tree()), _canon(o.tree())
    assert tg[0]["count"] == to[0]["count"] == 253
    for f in ("level", "path", "divided", "count"):
        assert np.array_equal(tg[f], to[f]), f


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.8])
def test_bh_forces_match_oracle(theta, oracle_lib):
    st, prm, _ = W.config1(4096)
    prm["theta"] = theta
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(len(st["x"])), o.forces()
    err = _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"])
    assert err < 2e-5, err      # FP32 + Kahan on the same interaction lists vs FP64
    cg, co = g.counters(), o.counters()
    # same per-particle decisions up to FP32 ties on the opening test
    assert abs(cg["pair_interactions"] - co["pair"]) <= 1e-3 * co["pair"]
    assert abs(cg["node_interactions"] - co["node"]) <= 1e-3 * max(co["node"], 1)
    pe_g = 0.5 * fg["pe"].sum()
    assert abs(pe_g - fo["pe_total"]) <= 1e-5 * abs(fo["pe_total"])
    np.testing.assert_allclose(fg["tidal_torque"], fo["tidal"], rtol=1e-3, atol=1e-6 * np.abs(fo["tidal"]).max())


def test_direct_sum_matches_oracle(oracle_lib):
    st, prm, _ = W.config1(2048)
    prm["toggles0"] = A.GRAVITY | A.COULOMB       # BH off: O(N^2) (forces.js:93-126)
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(len(st["x"])), o.forces()
    assert _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"]) < 1e-5
    assert g.counters()["pair_interactions"] == 2048 * 2047


@pytest.mark.parametrize("toggles", [
    A.GRAVITY | A.BARNES_HUT,                                               # compiled: gravity
    A.GRAVITY | A.COULOMB | A.BARNES_HUT,                                   # compiled: gravity + Coulomb
    A.GRAVITY | A.COULOMB | A.YUKAWA | A.MAGNETIC | A.RELATIVITY | A.BARNES_HUT,   # compiled: the config-2 set
    A.GRAVITY | A.GRAVITOMAG | A.ONE_PN | A.BARNES_HUT,                     # compiled: gravitomagnetism + 1PN
    A.GRAVITY | A.COULOMB | A.MAGNETIC | A.AXION | A.BARNES_HUT,            # runtime-checked (axion-modulated charges)
])
def test_feature_set_walks_match_oracle(toggles, oracle_lib):
    """Open-box walks for each feature set compiled with its toggles as
    constants, and one that falls back to the runtime-checked walk: forces,
    fields, PE and interaction counts against the oracle."""
    st, prm, _ = W.config2(4096, open_box=True)
    prm.update(toggles0=toggles, signal_delay=A.DELAY_OFF, theta=0.5)
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(len(st["x"])), o.forces()
    assert _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"]) < 2e-5
    fields = (("bz", "bz"), ("dbzdx", "dbzdx"), ("dbzdy", "dbzdy")) if toggles & A.MAGNETIC else ()
    fields += (("bgz", "bgz"),) if toggles & A.GRAVITOMAG else ()
    for k_g, k_o in fields:
        a, b = fg[k_g], fo[k_o]
        assert np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)) < 1e-4, k_g
    assert 0.5 * fg["pe"].sum() == pytest.approx(fo["pe_total"], rel=1e-5)
    cg, co = g.counters(), o.counters()
    assert abs(cg["pair_interactions"] - co["pair"]) <= 1e-3 * co["pair"]
    assert abs(cg["node_interactions"] - co["node"]) <= 1e-3 * max(co["node"], 1)


@pytest.mark.parametrize("mode", [A.PERIODIC_MIN_IMAGE, A.PERIODIC_WINDOWED])
@pytest.mark.parametrize("toggles", [A.GRAVITY | A.COULOMB | A.BARNES_HUT, A.GRAVITY | A.BARNES_HUT])
def test_torus_walks_match_oracle(toggles, mode, oracle_lib):
    """Torus walks: gravity + Coulomb compiled with the torus minimum image
    (and the window) as constants, gravity alone on the runtime-checked walk;
    forces, PE and interaction counts against the oracle (configuration-2
    state in its 256^2 torus, theta 0.5)."""
    st, prm, _ = W.config2(4096)
    prm.update(toggles0=toggles, periodic_mode=mode, signal_delay=A.DELAY_OFF)
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(len(st["x"])), o.forces()
    assert _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"]) < 1e-4
    assert 0.5 * fg["pe"].sum() == pytest.approx(fo["pe_total"], rel=1e-5)
    cg, co = g.counters(), o.counters()
    assert abs(cg["pair_interactions"] - co["pair"]) <= 1e-3 * co["pair"]
    assert abs(cg["node_interactions"] - co["node"]) <= 1e-3 * max(co["node"], 1)


@pytest.mark.parametrize("mode", [A.PERIODIC_MIN_IMAGE, A.PERIODIC_WINDOWED])
def test_config2_forces_fields_periodic(mode, oracle_lib):
    """Gravity + Coulomb + Yukawa + magnetic, relativistic, torus (min-image / windowed)."""
    st, prm, _ = W.config2(4096)
    prm["periodic_mode"] = mode
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(len(st["x"])), o.forces()
    assert _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"]) < 1e-4
    # Bz is a sum of nearly cancelling v x r terms; in windowed mode a few
    # window-edge decisions flip between FP32 and FP64 and move it visibly.
    tol = 1e-4 if mode == A.PERIODIC_MIN_IMAGE else 3e-3
    for k_g, k_o in (("bz", "bz"), ("dbzdx", "dbzdx"), ("dbzdy", "dbzdy")):
        a, b = fg[k_g], fo[k_o]
        assert np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)) < tol, k_g


def test_update_trajectory_config1(oracle_lib):
    # softening 0.01: close pairs make the dynamics chaotic, so FP32 and FP64
    # trajectories separate; compare after 5 updates pointwise and after 20 in
    # the bulk (median) only.
    st, prm, dt = W.config1(1024)
    g, o = _pair(st, prm, oracle_lib)
    for step in range(20):
        sg, so = g.update(dt), o.update(dt)
        assert sg == so, step
        if step == 4:
            dg, do = g.download(), o.download()
            ids = do["id"]
            for k in ("x", "y"):
                np.testing.assert_allclose(dg[k][ids], do[k], rtol=0, atol=1e-4)
            # FP32 positions near x = 32 are quantized to 4e-6, against a 0.01
            # softening length: the tight pairs that carry the kinetic energy
            # drift apart at the 1e-3 level within a few updates
            sg_, so_ = g.stats(), o.stats()
            for k_g, k_o in (("linearKE", "linear_ke"), ("pe", "pe"), ("px", "px"), ("py", "py")):
                scale = max(abs(so_[k_o]), 1e-3 * abs(so_["pe"]))
                assert abs(sg_[k_g] - so_[k_o]) <= 2e-3 * scale, k_g
    dg, do = g.download(), o.download()
    ids = do["id"]
    assert np.array_equal(np.nonzero(dg["alive"])[0], np.sort(ids))
    for k in ("x", "y"):
        assert np.median(np.abs(dg[k][ids] - do[k])) < 1e-5, k
        assert np.percentile(np.abs(dg[k][ids] - do[k]), 95) < 1e-4, k
    # tight pairs dominate the kinetic energy and separate fastest: bulk check only
    sg_, so_ = g.stats(), o.stats()
    assert abs(sg_["linearKE"] - so_["linear_ke"]) <= 1e-2 * abs(so_["linear_ke"])


def test_update_config2_substeps(oracle_lib):
    st, prm, dt = W.config2(2048)
    g, o = _pair(st, prm, oracle_lib)
    for step in range(3):
        sg, so = g.update(dt), o.update(dt)
        assert sg == so, step
    dg, do = g.download(), o.download()
    ids = do["id"]
    np.testing.assert_allclose(dg["x"][ids], do["x"], rtol=0, atol=1e-3)
    np.testing.assert_allclose(dg["angw"][ids], do["angw"], rtol=1e-3, atol=1e-4)


def test_despawn_compacts_and_keeps_slots(oracle_lib):
    st, prm, dt = W.config1(512)
    st["x"][10] = 64 + 64 - 1e-3
    st["wx"][10] = 5.0                     # leaves past the despawn margin in one update
    g, o = _pair(st, prm, oracle_lib)
    g.update(0.01)
    o.update(0.01)
    dg, do = g.download(), o.download()
    assert not dg["alive"][10]
    assert np.array_equal(np.nonzero(dg["alive"])[0], np.sort(do["id"]))
    assert g.getParticleCount() == o.count()


def test_deterministic_bits(oracle_lib):
    st, prm, dt = W.config1(2048)
    outs = []
    for _ in range(2):
        g, _o = _pair(st, prm, oracle_lib)
        for _s in range(3):
            g.update(dt)
        d = g.download()
        outs.append(np.concatenate([d["x"], d["y"], d["wx"], d["wy"]]))
        g.destroy()
    assert np.array_equal(outs[0], outs[1])


def test_add_remove_patch(oracle_lib):
    from physsim_amd import HipPhysics, default_params
    g = HipPhysics(default_params(), max_particles=4)
    assert g.addParticle(10, 10, 0, 0, 2.0, 0.26) == 0
    assert g.addParticle(20, 10, 0, 0, 1.0, -0.1) == 1
    assert g.addParticle(30, 10, 0, 0, 1.0, 0.0) == 2
    assert g.addParticle(40, 10, 0, 0, 1.0, 0.0) == 3
    assert g.addParticle(50, 10, 0, 0, 1.0, 0.0) == -1          # full (gpu-physics.js:1864)
    d = g.download()
    assert abs(d["charge"][0] - 0.3) < 1e-12                      # quantized to 0.1
    g.removeParticle(1)
    g.patchMass(2, 3.0, 3.0)
    d = g.download()
    assert list(d["alive"]) == [True, False, True, True]
    assert d["mass"][2] == 3.0
    g.compute_forces()
    f = g.forces(4)
    assert f["fx"][0] > 0 and f["fx"][3] < 0


def test_windowed_walk_converges_to_direct_sum(oracle_lib):
    """Windowed periodic mode: BH at small theta approaches the min-image direct sum."""
    st, prm, _ = W.config2(4096)
    prm["periodic_mode"] = A.PERIODIC_WINDOWED
    errs = []
    o = None
    for theta in (0.5, 0.2):
        prm["theta"] = theta
        g, o = _pair(st, prm, oracle_lib)
        g.compute_forces()
        fg = g.forces(len(st["x"]))
        if theta == 0.5:
            fx, fy = o.direct_forces()
        errs.append(_rel_l2(fg["fx"], fg["fy"], fx, fy))
        g.destroy()
    assert errs[1] < errs[0] and errs[1] < 0.02, errs


@pytest.mark.parametrize("bh", [True, False])
def test_one_pn_velocity_verlet_matches_oracle(bh, oracle_lib):
    """Gravity + gravitomagnetism + 1PN (EIH) with the velocity-Verlet 1PN correction
    (integrator.js:1296-1318), relativistic; BH walk or O(N^2)."""
    st, prm, dt, _ = W.config3(2048, grid=64, side=256.0, total_mass=4.0)
    prm["toggles0"] = A.GRAVITY | A.GRAVITOMAG | A.ONE_PN | A.RELATIVITY | (A.BARNES_HUT if bh else 0)
    prm["signal_delay"] = A.DELAY_OFF               # current-time leaves (light-cone leaves: test_gpu_delay.py)
    g, o = _pair(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(len(st["x"])), o.forces()
    assert _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"]) < 1e-4
    for step in range(3):
        sg, so = g.update(dt), o.update(dt)
        assert sg == so, step
    dg, do = g.download(), o.download()
    ids = do["id"]
    wg = np.hypot(dg["wx"][ids], dg["wy"][ids])
    assert np.sqrt(np.sum((dg["wx"][ids] - do["wx"]) ** 2 + (dg["wy"][ids] - do["wy"]) ** 2) / np.sum(wg ** 2)) < 1e-4


@pytest.mark.parametrize("periodic", [False, True])
def test_collision_pairs_bit_exact(periodic, oracle_lib):
    """Overlap pair list (collisions.js:35-60) equals the oracle's, pair for pair.
    Periodic: the oracle runs the reference's ghost images (integrator.js:201-258)."""
    g_ = np.random.default_rng(21)
    n = 3000
    W_ = 64.0
    x = W.f32(W_ * g_.random(n))
    y = W.f32(W_ * g_.random(n))
    m = W.f32(0.05 + 0.5 * g_.random(n))
    z = np.zeros(n)
    st = dict(x=x, y=y, wx=z, wy=z, mass=m, charge=z, angw=z)
    prm = dict(toggles0=A.GRAVITY | A.BARNES_HUT, domain_w=W_, domain_h=W_)
    if periodic:
        prm.update(boundary_mode=A.BOUND_LOOP)
    g, o = _pair(st, prm, oracle_lib)
    if periodic:
        o.set_params(oracle_lib.default_params(**dict(prm, periodic_mode=A.PERIODIC_GHOSTS)))
    ga, gb = g.collision_pairs()
    oa, ob = o.collision_pairs()
    assert len(oa) > 100
    assert np.array_equal(ga, oa) and np.array_equal(gb, ob)


def test_tree_resolves_close_points_like_the_oracle(oracle_lib):
    """Points a few FP32 ulps apart share a cell of 24 levels; the reference
    keeps subdividing until each leaf holds at most 4 (quadtree.js:203-231:
    the depth counter restarts for every re-inserted point, so the split goes
    as deep as the points need).  The keys resolve 31 levels, which separates
    FP32 points this close at these coordinates: the tree is the oracle's, bit
    for bit, with nothing left in an over-full finest cell."""
    st, prm, _ = W.config1(300)
    g0 = np.random.default_rng(17)
    xs, ys = [st["x"]], [st["y"]]
    for c in range(5):
        x0 = np.float32(g0.uniform(20, 44))
        y0 = np.float32(g0.uniform(20, 44))
        xs.append(np.array([np.nextafter(x0, np.float32(100), dtype=np.float32) if k % 2 else x0 for k in range(6)],
                           np.float32).astype(np.float64) + np.arange(6) // 2 * np.spacing(np.float32(x0)))
        ys.append(np.full(6, y0, np.float64) + np.arange(6) % 3 * np.spacing(np.float32(y0)))
    n = 300 + 30
    ext = dict(x=np.concatenate(xs), y=np.concatenate(ys))
    for k in ("wx", "wy", "mass", "charge", "angw"):
        ext[k] = np.concatenate([st[k], np.resize(st[k], 30)])
    assert len(np.unique(np.stack([ext["x"], ext["y"]], 1), axis=0)) == n      # distinct, some by one ulp
    g, o = _pair(ext, prm, oracle_lib)
    g.build_tree()
    tg, to = _canon(g.tree()), _canon(o.tree())
    print("levels: gpu %d oracle %d" % (tg["level"].max(), to["level"].max()))
    assert to["level"].max() > 24
    assert g.counters()["depth_overflow"] == 0
    assert len(tg) == len(to)
    for f in ("level", "path", "divided", "count", "charge_quanta"):
        assert np.array_equal(tg[f], to[f]), f
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(n), o.forces()
    assert _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"]) < 2e-5
