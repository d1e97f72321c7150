"""Batched ensembles on gfx950: many config-1 systems stepped together.

A system's results must not depend on the batch it runs in (bit-for-bit
against a batch of one), must agree with the single-system backend and the
FP64 oracle, and must be the same replayed from the captured graph as
launched eagerly, and after the autotuner changes the walk's tile and stack.
"""
import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def _batch(states, prm, caps=None):
    from physsim_amd import HipBatch, default_params
    caps = caps if caps is not None else [len(s["x"]) for s in states]
    print("batch: %d systems, %d particles" % (len(states), sum(len(s["x"]) for s in states)), flush=True)
    b = HipBatch(default_params(**prm), caps)
    for k, st in enumerate(states):
        b.upload(k, st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return b


def _single(st, prm):
    from physsim_amd import HipPhysics, default_params
    g = HipPhysics(default_params(**prm), max_particles=len(st["x"]))
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return g


def _rel_l2(ax, ay, bx, by):
    return float(np.sqrt(np.sum((ax - bx) ** 2 + (ay - by) ** 2) / np.sum(bx ** 2 + by ** 2)))


def test_batch_forces_match_single_and_oracle(oracle_lib):
    states, prm, _ = W.config4(6, replicate=3, n_max=4096)
    b = _batch(states, prm)
    b.compute_forces()
    for k, st in enumerate(states):
        n = len(st["x"])
        fb = b.forces(k, n)
        g = _single(st, prm)
        g.compute_forces()
        fg = g.forces(n)
        assert _rel_l2(fb["fx"], fb["fy"], fg["fx"], fg["fy"]) < 1e-6
        o = oracle_lib.Oracle(oracle_lib.default_params(**prm), n)
        o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        o.compute_forces()
        fo = o.forces()
        assert _rel_l2(fb["fx"], fb["fy"], fo["fx"], fo["fy"]) < 1e-4


def test_batch_invariance_bit_exact():
    states, prm, dt = W.config4(8, replicate=5, n_max=4096)
    caps = [len(s["x"]) for s in states]
    B = _batch(states, prm, caps)
    subs = [B.update(dt) for _ in range(3)]
    for k in (0, 3, 7):
        solo = _batch([states[k]], prm, [caps[k]])
        ssub = [solo.update(dt) for _ in range(3)]
        assert [int(s[k]) for s in subs] == [int(s[0]) for s in ssub]
        db, ds = B.download(k), solo.download(0)
        for f in ("x", "y", "wx", "wy", "angw"):
            assert np.array_equal(db[f], ds[f]), (k, f)
        assert B.stats(k)["linear_ke"] == solo.stats(0)["linear_ke"]


def test_batch_update_matches_single_backend():
    states, prm, dt = W.config4(4, replicate=7, n_max=2048)
    B = _batch(states, prm)
    sb = [B.update(dt) for _ in range(2)]
    for k, st in enumerate(states):
        g = _single(st, prm)
        sg = [g.update(dt) for _ in range(2)]
        assert [int(s[k]) for s in sb] == sg
        db, dg = B.download(k), g.download()
        np.testing.assert_allclose(db["x"], dg["x"], rtol=0, atol=2e-4)
        np.testing.assert_allclose(db["y"], dg["y"], rtol=0, atol=2e-4)
        eb, eg = B.stats(k), g.stats()
        assert eb["linear_ke"] == pytest.approx(eg["linearKE"], rel=1e-4)
        assert eb["total_mass"] == pytest.approx(eg["totalMass"], rel=1e-12)


@pytest.mark.parametrize("bodies", [4, 32, 2])
def test_batch_graph_equals_eager(bodies):
    """Captured substep graphs of 2, 4 or all 32 bodies (one graph per update,
    the finished systems predicated off inside it) step exactly as the eager loop."""
    states, prm, dt = W.config4(5, replicate=9, n_max=2048)
    g1 = _batch(states, prm)
    g1.tune(graph_bodies=bodies)
    g2 = _batch(states, prm)
    g2.tune(use_graph=0)
    for _ in range(3):
        assert np.array_equal(g1.update(dt), g2.update(dt))
    for k in range(len(states)):
        a, b = g1.download(k), g2.download(k)
        for f in ("x", "y", "wx", "wy"):
            assert np.array_equal(a[f], b[f]), (k, f)
    with pytest.raises(Exception):
        g1.tune(graph_bodies=3)


def test_batch_autotune_keeps_results():
    states, prm, dt = W.config4(6, replicate=11, n_max=4096)
    b = _batch(states, prm)
    b.compute_forces()
    before = [b.forces(k, len(s["x"])) for k, s in enumerate(states)]
    t = b.autotune()
    print("autotune:", t)
    assert t["walk_threads"] in (64, 128, 256) and t["stack"] in (64, 96) and t["walk_ms"] > 0
    b.compute_forces()
    for k, s in enumerate(states):
        f = b.forces(k, len(s["x"]))
        assert np.array_equal(f["fx"], before[k]["fx"]) and np.array_equal(f["fy"], before[k]["fy"])


def test_batch_despawn_and_substep_cap():
    states, prm, dt = W.config4(3, replicate=13, n_max=1024)
    states[1]["wx"][0] = 1e5                     # leaves the despawn margin in one update
    b = _batch(states, prm)
    sub = b.update(dt)
    assert np.all(sub >= 1) and np.all(sub <= prm.get("max_substeps", 32))
    d = b.download(1)
    assert not d["alive"][0] and d["alive"][1:].all()
    assert b.stats(1)["alive_count"] == len(states[1]["x"]) - 1
    assert b.download(0)["alive"].all()
