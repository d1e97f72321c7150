"""The five benchmark configurations as GPU parity workloads, each against the
FP64 oracle on the same seeded inputs, plus the precision policy (Kahan FP32
accumulation against naive FP32, both judged by FP64) and the spin-orbit
branch of the substep (integrator.js:930-981).

Periodic cases name their mode: min-image (each source once) or windowed."""
import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def _gpu(st, prm, cap=None):
    from physsim_amd import HipPhysics, default_params
    n = len(st["x"])
    print("case: n=%d toggles=0x%x theta=%.2f far=%d boundary=%d" % (
        n, prm["toggles0"], prm.get("theta", 0.5), prm.get("far_field", 0), prm.get("boundary_mode", 0)), flush=True)
    g = HipPhysics(default_params(**prm), max_particles=cap or n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return g


def _orc(st, prm, oracle_lib):
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), len(st["x"]))
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return o


def _rel_l2(ax, ay, bx, by):
    return float(np.sqrt(np.sum((ax - bx) ** 2 + (ay - by) ** 2) / np.sum(bx ** 2 + by ** 2)))


def _rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b ** 2), 1e-300)))


def _set(g, **kw):
    p = A.GeonParams.from_buffer_copy(g.params)
    for k, v in kw.items():
        setattr(p, k, v)
    g.set_params(p)


# ---- configuration 1: N=1024 disk, gravity + Coulomb, softening 0.01, 100 updates ----

def test_config1_all_100_updates(oracle_lib):
    """All 100 updates of configuration 1, in two ways.

    Shadowed: at every update k the GPU and a fresh oracle both start from
    the oracle trajectory's state after k updates (rounded to FP32, so both
    see identical inputs) and take one update: the substep count must be the
    oracle's and every position must agree to FP32 rounding, so the whole
    trajectory is checked update by update.

    Free-running: the GPU runs the 100 updates on its own.  Softening 0.01
    makes tight pairs chaotic (FP32 and FP64 pair orbits separate within a
    few updates), so only what the pairs cannot change is compared: the
    particle set, the total mass and the momentum."""
    from physsim_amd import HipPhysics, default_params
    st, prm, dt = W.config1(1024)
    o = _orc(st, prm, oracle_lib)
    g = _gpu(st, prm)
    worst, same = 0.0, 0
    for k in range(100):
        d = o.download()
        s32 = {c: W.f32(d[c]) for c in ("x", "y", "wx", "wy", "mass", "angw")}
        q = np.zeros(len(d["id"]))
        q[:] = st["charge"][d["id"]]
        g.upload(s32["x"], s32["y"], s32["wx"], s32["wy"], s32["mass"], q, s32["angw"])
        o1 = oracle_lib.Oracle(oracle_lib.default_params(**prm), len(q))
        o1.upload(s32["x"], s32["y"], s32["wx"], s32["wy"], s32["mass"], q, s32["angw"])
        sg, so = g.update(dt), o1.update(dt)
        same += sg == so
        dg, do1 = g.download(), o1.download()
        ids = do1["id"]
        worst = max(worst, float(np.abs(np.hypot(dg["x"][ids] - do1["x"], dg["y"][ids] - do1["y"])).max()))
        o.update(dt)
    print("config1 shadowed: substeps equal in %d/100 updates, worst one-update position difference %.2e"
          % (same, worst))
    assert same == 100
    assert worst < 1e-4
    g2 = _gpu(st, prm)
    o2 = _orc(st, prm, oracle_lib)
    steps_g = [g2.update(dt) for _ in range(100)]
    steps_o = [o2.update(dt) for _ in range(100)]
    dg, do = g2.download(), o2.download()
    ids = do["id"]
    assert np.array_equal(np.nonzero(dg["alive"])[0], np.sort(ids))
    sg, so = g2.stats(), o2.stats()
    print("config1 free-running: %d vs %d substeps; px %.4e vs %.4e, py %.4e vs %.4e"
          % (sum(steps_g), sum(steps_o), sg["px"], so["px"], sg["py"], so["py"]))
    assert sg["totalMass"] == pytest.approx(so["total_mass"], rel=1e-6)
    p_scale = np.sum(do["mass"] * np.hypot(do["wx"], do["wy"]))
    assert abs(sg["px"] - so["px"]) < 1e-4 * p_scale and abs(sg["py"] - so["py"]) < 1e-4 * p_scale


# ---- configuration 2 at its full size: N=65536 torus, relativistic config-2 force set ----

def test_config2_full_size_forces_and_update(oracle_lib):
    """N = 65536 in the 256^2 torus (min-image), gravity + Coulomb + Yukawa +
    magnetic dipole, relativistic, spins: forces, fields and PE of the
    bootstrap pass, then one adaptive update (substeps, positions, spins)."""
    st, prm, dt = W.config2(65536)
    g, o = _gpu(st, prm), _orc(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(65536), o.forces()
    # In min-image mode a node is taken at its COM's nearest image; for a COM
    # within rounding of half the box from the observer, FP32 and FP64 pick
    # opposite images and that node's pull flips.  Such ties touch a handful of
    # particles, so the bulk is compared per particle and the ties are counted.
    fmag = np.maximum(np.hypot(fo["fx"], fo["fy"]), 1e-300)
    per = np.hypot(fg["fx"] - fo["fx"], fg["fy"] - fo["fy"]) / fmag
    ties = int(np.sum(per > 1e-3))
    ef = _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"])
    eb = _rel(fg["bz"], fo["bz"])
    print("config2 N=65536: force rel L2 %.3e, per-particle median %.2e p99 %.2e, %d particles over 1e-3; Bz %.3e"
          % (ef, np.median(per), np.percentile(per, 99), ties, eb))
    assert np.median(per) < 1e-5 and np.percentile(per, 99) < 2e-4
    assert ties <= 1e-3 * 65536
    ok = per <= 1e-3
    # each tie particle's GPU force is the oracle's force with that particle
    # moved by 2e-5 (FP32 rounding at 128 is 8e-6) in one of four directions
    bad = np.flatnonzero(~ok)
    best = np.full(len(bad), np.inf)
    for dx, dy in ((2e-5, 0.0), (-2e-5, 0.0), (0.0, 2e-5), (0.0, -2e-5)):
        x2, y2 = st["x"].copy(), st["y"].copy()
        x2[bad] += dx
        y2[bad] += dy
        on = oracle_lib.Oracle(oracle_lib.default_params(**prm), 65536)
        on.upload(x2, y2, st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        on.compute_forces()
        fn = on.forces()
        e = np.hypot(fg["fx"][bad] - fn["fx"][bad], fg["fy"][bad] - fn["fy"][bad]) / fmag[bad]
        best = np.minimum(best, e)
    print("tie particles", bad.tolist(), "error vs the nearest nudged oracle", np.round(best, 8).tolist())
    assert np.all(best < 1e-4)
    assert _rel_l2(fg["fx"][ok], fg["fy"][ok], fo["fx"][ok], fo["fy"][ok]) < 1e-4
    assert 0.5 * fg["pe"].sum() == pytest.approx(fo["pe_total"], rel=1e-5)
    cg, co = g.counters(), o.counters()
    assert abs(cg["pair_interactions"] - co["pair"]) <= 1e-3 * co["pair"]
    assert g.update(dt) == o.update(dt)
    dg, do = g.download(), o.download()
    ids = do["id"]
    np.testing.assert_allclose(dg["x"][ids], do["x"], rtol=0, atol=1e-3)
    keep = ok[ids]                      # the tie particles were kicked by a different node image
    np.testing.assert_allclose(dg["wx"][ids][keep], do["wx"][keep], rtol=0, atol=1e-4)
    np.testing.assert_allclose(dg["angw"][ids][keep], do["angw"][keep], rtol=1e-3, atol=1e-4)


# ---- configuration 3: 1PN EIH + Higgs + axion on the 128^2 grid ----

def test_config3_one_pn_fields_grid128(oracle_lib):
    """Gravity + gravitomagnetism + 1PN (EIH, velocity-Verlet pass), Higgs and
    axion with FFT self-gravity on a 128^2 grid, Plummer sphere of 8192, 3
    updates (integrator.js:1296-1334): substeps, fields, field derivatives,
    masses (Higgs modulation), positions, velocities and field energies."""
    st, prm, dt, fields = W.config3(8192, grid=128, side=256.0, total_mass=4.0)
    assert prm["toggles0"] & A.ONE_PN and prm["toggles0"] & A.GRAVITOMAG
    g, o = _gpu(st, prm), _orc(st, prm, oracle_lib)
    for w, k in ((0, "higgs"), (1, "axion")):
        g.set_field(w, fields[k])
        o.field_set(w, fields[k])
    for step in range(3):
        assert g.update(dt) == o.update(dt), step
    for w in (0, 1):
        vac = 1.0 if w == 0 else 0.0
        e_f = _rel(g.get_field(w, "field") - vac, o.field_get(w, "field") - vac)
        e_d = _rel(g.get_field(w, "field_dot"), o.field_get(w, "field_dot"))
        print("field %d: value %.3e, rate %.3e" % (w, e_f, e_d))
        assert e_f < 2e-3 and e_d < 5e-3, w
    dg, do = g.download(), o.download()
    ids = do["id"]
    np.testing.assert_allclose(dg["mass"][ids], do["mass"], rtol=2e-5)
    np.testing.assert_allclose(dg["x"][ids], do["x"], rtol=0, atol=1e-3)
    ew = _rel_l2(dg["wx"][ids], dg["wy"][ids], do["wx"], do["wy"])
    print("velocities: %.3e" % ew)
    assert ew < 1e-4
    f1g = g.forces(8192)
    f1o = o.f1pn()
    e1 = _rel_l2(f1g["f1pn_x"][ids], f1g["f1pn_y"][ids], f1o[0], f1o[1])
    print("1PN force: %.3e" % e1)
    assert e1 < 1e-3
    sg, so = g.stats(), o.stats()
    assert sg["higgsFieldEnergy"] == pytest.approx(so["higgs_field_energy"], rel=1e-4)
    assert sg["axionFieldEnergy"] == pytest.approx(so["axion_field_energy"], rel=1e-3)


# ---- configuration 5: the learned correction on the config-2 force set ----

@pytest.fixture(scope="module")
def far_model():
    """A network trained on the GPU on exact residuals of 4 Plummer spheres of 16384."""
    from physsim_amd import HipPhysics, default_params
    _, prm = W.plummer_charged(16384, 0)
    T = HipPhysics(default_params(**dict(prm, theta=0.5)), max_particles=16384)
    for seed in range(4):
        st, _ = W.plummer_charged(16384, seed)
        T.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        T.far_collect(4)
    loss = T.far_train(2000, batch=8192, lr=2e-3, lr_final=2e-5, seed=3)
    print("far model: loss %.4e -> %.4e" % (loss[:20].mean(), loss[-20:].mean()), flush=True)
    model = T.far_get_model()
    T.destroy()
    return model


@pytest.mark.parametrize("case", ["open", "torus_windowed"])
@pytest.mark.parametrize("theta", [0.5, 0.8])
def test_config5_learned_forces_config2_set(case, theta, far_model, oracle_lib):
    """FAR_LEARNED with gravity + Coulomb + Yukawa + magnetic, relativistic:
    the BF16/FP16 network on the GPU against the same network in FP64 in the
    oracle (forces, Bz, PE), open box and windowed torus."""
    st, prm, _ = W.config2(4096, open_box=(case == "open"))
    if case != "open":
        prm["periodic_mode"] = A.PERIODIC_WINDOWED
    prm = dict(prm, theta=theta, far_field=A.FAR_LEARNED)
    g = _gpu(st, prm)
    g.far_set_model(far_model)
    g.compute_forces()
    o = _orc(st, prm, oracle_lib)
    o.far_set_model(far_model)
    o.compute_forces()
    fg, fo = g.forces(4096), o.forces()
    e = _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"])
    eb = _rel(fg["bz"], fo["bz"])
    print("%s theta %.1f: learned forces vs oracle %.3e, Bz %.3e" % (case, theta, e, eb))
    # windowed: a node whose box touches the edge of the observer's min-image
    # window is opened or accepted by an FP32 comparison the FP64 oracle can
    # decide the other way (as in the windowed config-2 test)
    assert e < (3e-4 if case == "open" else 1e-3)
    assert eb < (1e-4 if case == "open" else 3e-3)
    assert 0.5 * fg["pe"].sum() == pytest.approx(fo["pe_total"], rel=1e-4)
    assert g.counters()["node_interactions"] > 0


def test_config5_learned_gain_on_config2_set(far_model, oracle_lib):
    """What the trained network gains over the monopole walk at the same
    opening angle on the config-2 force set (Yukawa and magnetic nodes stay
    monopole), judged against the FP64 direct sum of the same forces."""
    st, prm, _ = W.config2(16384, open_box=True)
    theta = 0.5
    g = _gpu(st, dict(prm, theta=theta))
    g.compute_forces()
    fp = g.forces(16384)
    g.far_set_model(far_model)
    _set(g, far_field=A.FAR_LEARNED)
    g.compute_forces()
    fl = g.forces(16384)
    o = _orc(st, prm, oracle_lib)
    idx = np.arange(0, 16384, 32, dtype=np.int32)
    fx, fy = o.direct_forces(idx)
    e_plain = _rel_l2(fp["fx"][idx], fp["fy"][idx], fx, fy)
    e_learn = _rel_l2(fl["fx"][idx], fl["fy"][idx], fx, fy)
    print("config-2 set theta %.2f: monopole %.3e, learned %.3e (x%.1f)" % (theta, e_plain, e_learn,
                                                                           e_plain / e_learn))
    assert e_learn < e_plain / 4


# ---- precision policy: compensated FP32 accumulation ----

def test_kahan_beats_naive_fp32_against_fp64(oracle_lib):
    """N = 65536 direct sum (gravity + Coulomb, config-2 positions and
    charges, open box): the FP32 forces with Kahan-compensated accumulation
    are closer to the FP64 direct sum than the naive FP32 sums."""
    st, prm, _ = W.config2(65536, open_box=True)
    prm = dict(prm, toggles0=A.GRAVITY | A.COULOMB, signal_delay=A.DELAY_OFF)
    g = _gpu(st, dict(prm, kahan=1))
    g.compute_forces()
    fk = g.forces(65536)
    _set(g, kahan=0)
    g.compute_forces()
    fn = g.forces(65536)
    o = _orc(st, prm, oracle_lib)
    idx = np.arange(0, 65536, 16, dtype=np.int32)
    fx, fy = o.direct_forces(idx)
    ek = _rel_l2(fk["fx"][idx], fk["fy"][idx], fx, fy)
    en = _rel_l2(fn["fx"][idx], fn["fy"][idx], fx, fy)
    print("N=65536 direct sum vs FP64: Kahan %.3e, naive %.3e (x%.1f)" % (ek, en, en / ek))
    assert ek < 0.5 * en


# ---- spin-orbit coupling and torques in the substep ----

@pytest.mark.parametrize("bh", [False, True])
def test_spin_orbit_matches_oracle(bh, oracle_lib):
    """Stern-Gerlach and Mathisson-Papapetrou kicks with their spin energy
    transfer, frame-drag and tidal torques (integrator.js:930-981):
    magnetic + gravitomagnetic + spin-orbit, relativistic, current-time
    sources, 4 updates; spins (angw) and proper velocities against the oracle."""
    st, prm, dt = W.config2(2048, open_box=True)
    prm = dict(prm, toggles0=A.GRAVITY | A.COULOMB | A.MAGNETIC | A.GRAVITOMAG | A.SPIN_ORBIT | A.RELATIVITY |
               (A.BARNES_HUT if bh else 0), signal_delay=A.DELAY_OFF)
    g, o = _gpu(st, prm), _orc(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(2048), o.forces()
    for k_g, k_o in (("dbzdx", "dbzdx"), ("dbgzdx", "dbgzdx"), ("frame_drag_torque", "frame_drag"),
                     ("tidal_torque", "tidal")):
        e = _rel(fg[k_g], fo[k_o])
        print("%s: %.3e" % (k_g, e))
        assert e < 1e-4, k_g
    for step in range(4):
        assert g.update(dt) == o.update(dt), step
    dg, do = g.download(), o.download()
    ids = do["id"]
    ea = _rel(dg["angw"][ids], do["angw"])
    ew = _rel_l2(dg["wx"][ids], dg["wy"][ids], do["wx"], do["wy"])
    moved = np.abs(do["angw"] - st["angw"][ids]).max()
    print("spin-orbit bh=%d: angw %.3e (largest change %.3e), w %.3e" % (bh, ea, moved, ew))
    assert moved > 1e-6                                              # the spins moved
    assert ea < 1e-4 and ew < 1e-4
