"""Collision mode "merge" on gfx950 against the FP64 oracle
(collisions.js:24-158, integrator.js:1351-1416): which particles merge or
annihilate, the slots of the spawned particles and the order of the events
must be identical; masses, positions, velocities and spins within FP32
rounding.  Inputs are FP32-representable so both sides start from the same
state."""
import numpy as np
import pytest

from physsim_amd import _abi as A

pytestmark = pytest.mark.gpu

f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)


def _params(**kw):
    p = dict(toggles0=0, boundary_mode=A.BOUND_DESPAWN, domain_w=128.0, domain_h=128.0, collision_mode=A.COL_MERGE,
             signal_delay=A.DELAY_OFF, softening=4.0)
    p.update(kw)
    return p


# Without forces and relativity a drift of DT moves positions (multiples of
# 2^-10) by velocities (multiples of 2^-6) exactly in FP32 and FP64: both sides
# then test the same geometry, and only the fold decides.
DT = 2.0 ** -8
q = lambda a, s: np.round(np.asarray(a) * s) / s


def _state(n, seed, W, clustered=False, mass=(0.5, 4.0), speed=0.3, anti_frac=0.0, spread=0.04):
    g = np.random.default_rng(seed)
    if clustered:
        c = g.uniform(0.25 * W, 0.75 * W, (6, 2))
        k = g.integers(0, 6, n)
        p = c[k] + g.normal(0, spread * W, (n, 2))
        x, y = np.clip(p[:, 0], 1, W - 1), np.clip(p[:, 1], 1, W - 1)
    else:
        x, y = g.uniform(1, W - 1, n), g.uniform(1, W - 1, n)
    return dict(x=q(x, 1024), y=q(y, 1024), wx=q(g.normal(0, speed, n), 64), wy=q(g.normal(0, speed, n), 64),
                mass=f32(g.uniform(*mass, n)), charge=g.integers(-5, 6, n) * 0.1, angw=f32(g.uniform(-0.3, 0.3, n)),
                anti=(g.random(n) < anti_frac).astype(np.int32))


def _pair(st, prm, oracle_lib, slots=2):
    # every spawned particle takes a new slot: capacity = slots x initial count
    from physsim_amd import HipPhysics, default_params
    n = len(st["x"])
    print("case: n=%d toggles=0x%x boundary=%d" % (n, prm["toggles0"], prm["boundary_mode"]), flush=True)
    g = HipPhysics(default_params(**prm), max_particles=slots * n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"], antimatter=st["anti"])
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), n)
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    o.set_antimatter(st["anti"])
    return g, o


def _compare(g, o, tol=2e-5, spin_tol=None):
    # spin_tol: runs with forces integrate spins through the tidal torque of
    # close encounters, whose FP32 sensitivity exceeds that of the positions
    dg, do, eo = g.download(), o.download(), o.download_ex()
    ids = do["id"]
    assert g.getParticleCount() == o.count()
    assert np.array_equal(np.nonzero(dg["alive"])[0], np.sort(ids)), "different survivors / spawned slots"
    for k in ("x", "y", "wx", "wy", "mass", "angw"):
        t = spin_tol if (k == "angw" and spin_tol) else tol
        scale = max(1.0, np.abs(do[k]).max())
        np.testing.assert_allclose(dg[k][ids], do[k], rtol=t, atol=t * scale, err_msg=k)
    np.testing.assert_allclose(dg["base_mass"][ids], eo["base_mass"], rtol=tol, err_msg="base_mass")
    np.testing.assert_allclose(dg["charge"][ids], eo["charge"], atol=1e-6, err_msg="charge")
    assert np.array_equal(dg["antimatter"][ids].astype(np.int32), eo["antimatter"])


def _compare_events(g, o):
    # A particle that already merged in its own turn merges again with its next
    # candidate; the kinetic energy such a merge loses is zero in exact
    # arithmetic, so whether "keLost > 0" emits its event is decided by
    # rounding (collisions.js:86-90) and differs between FP32 and FP64
    # velocities.  Those events are left out of the comparison.
    noise = lambda e: (e[:, 0] == 1) & (e[:, 3] < 1e-12)
    eg, eo = g.collision_events(), o.collision_events()
    eg, eo = eg[~noise(eg)], eo[~noise(eo)]
    assert eg.shape == eo.shape
    if len(eo):
        assert np.array_equal(eg[:, 0], eo[:, 0])
        np.testing.assert_allclose(eg[:, 1:3], eo[:, 1:3], rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(eg[:, 3:], eo[:, 3:], rtol=1e-4, atol=1e-5)
    return len(eo)


def test_isolated_merges_match_oracle(oracle_lib):
    st = _state(4000, 1, 256.0)
    g, o = _pair(st, _params(domain_w=256.0, domain_h=256.0), oracle_lib)
    ng, no = g.update(DT), o.update(DT)
    assert ng == no
    _compare(g, o)
    ne = _compare_events(g, o)
    print("merges: %d particles left of 4000, %d events" % (o.count(), ne))
    assert o.count() < 3900 and ne > 50


@pytest.mark.parametrize("seed", [2, 3])
def test_clustered_fold_matches_oracle(seed, oracle_lib):
    # particles overlapping several others: the fold order (slot, query
    # traversal, leaf stored order) decides the outcome
    st = _state(3000, seed, 128.0, clustered=True, mass=(0.3, 2.5), spread=0.05)
    prm = _params()
    g, o = _pair(st, prm, oracle_lib)
    a, b = o.collision_pairs_ordered()
    key, lid = o.leaf_points()
    leaf_of = dict(zip(lid.tolist(), key.tolist()))
    same_leaf = sum(1 for k in range(len(a) - 1) if a[k] == a[k + 1] and leaf_of[b[k]] == leaf_of[b[k + 1]])
    multi = sum(1 for i in set(a.tolist()) if (a == i).sum() > 1)
    deg = np.bincount(np.concatenate([a, b]))
    folded = int(np.sum((deg[a] > 1) | (deg[b] > 1)))    # pairs sharing a particle: the ordered fold
    print("%d pairs, %d particles with several candidates, %d same-leaf neighbours, %d folded pairs"
          % (len(a), multi, same_leaf, folded))
    assert multi > 50 and same_leaf > 10
    assert folded > 64             # above this the fold runs by connected components (kFoldOneThread)
    g.update(DT)
    o.update(DT)
    _compare(g, o)
    _compare_events(g, o)


def test_annihilation_and_merges_mixed(oracle_lib):
    st = _state(3000, 4, 128.0, clustered=True, mass=(0.3, 2.5), anti_frac=0.4, spread=0.05)
    g, o = _pair(st, _params(), oracle_lib)
    g.update(DT)
    o.update(DT)
    _compare(g, o)
    ev = o.collision_events()
    assert (ev[:, 0] == 2).sum() > 50 and (ev[:, 0] == 1).sum() > 50
    _compare_events(g, o)


def test_merges_across_torus_edges(oracle_lib):
    # isolated pairs straddling the x edge, the y edge and the corner: the
    # reference finds the neighbour across an edge through the ghosts in its
    # tree (integrator.js:201-258), this backend through the wrapped images of
    # the query box; the merged body sits at the min-image COM, wrapped back
    W = 96.0
    g0 = np.random.default_rng(5)
    xs, ys = [], []
    for k in range(29):                         # pairs 3 apart: each meets only its partner
        t = 3.0 + 3.0 * k
        xs += [g0.uniform(0.05, 0.4), W - g0.uniform(0.05, 0.4)]; ys += [t, t + g0.uniform(-0.3, 0.3)]
        xs += [t, t + g0.uniform(-0.3, 0.3)]; ys += [g0.uniform(0.05, 0.4), W - g0.uniform(0.05, 0.4)]
    xs += [0.2, W - 0.3]; ys += [0.3, W - 0.2]  # across the corner
    for k in range(20):
        cx, cy = 10.0 + 3.5 * (k % 5), 20.0 + 4.0 * (k // 5)
        xs += [cx, cx + g0.uniform(-0.8, 0.8)]; ys += [cy, cy + g0.uniform(-0.8, 0.8)]
    n = len(xs)
    st = _state(n, 5, W, mass=(0.8, 1.2))
    st["x"], st["y"] = q(xs, 1024), q(ys, 1024)
    prm = _params(domain_w=W, domain_h=W, boundary_mode=A.BOUND_LOOP, topology=A.TORUS,
                  periodic_mode=A.PERIODIC_MIN_IMAGE)
    g, o = _pair(st, prm, oracle_lib)
    a, b = o.collision_pairs_ordered()
    across = int(np.sum((np.abs(st["x"][a] - st["x"][b]) > W / 2) | (np.abs(st["y"][a] - st["y"][b]) > W / 2)))
    print("%d pairs, %d across an edge" % (len(a), across))
    assert len(a) == n // 2 and across == 59
    g.update(DT)
    o.update(DT)
    _compare(g, o)
    assert o.count() == n // 2
    _compare_events(g, o)


def test_merged_bodies_retire_as_delayed_sources(oracle_lib):
    # a sparse moving system under gravity with signal delay: new overlaps every
    # update; the bodies removed after the first update carry two history
    # samples and act as retired sources (forces.js:128-153) in the next ones
    st = _state(2000, 6, 256.0, speed=0.5)
    prm = _params(domain_w=256.0, domain_h=256.0, toggles0=A.GRAVITY | A.RELATIVITY, signal_delay=A.DELAY_FOLLOW,
                  history_stride=1)
    g, o = _pair(st, prm, oracle_lib, slots=2)
    for k in range(4):
        assert g.update(0.25) == o.update(0.25), k
        _compare_events(g, o)
    _compare(g, o, tol=1e-4, spin_tol=2e-2)
    assert g.stats()["deadCount"] == o.dead_count()
    print("dead sources %d, %d particles left" % (o.dead_count(), o.count()))
    assert o.dead_count() > 400


def test_excitations_on_fields(oracle_lib):
    G = 64
    st = _state(2000, 7, 128.0, clustered=True)
    prm = _params(toggles0=A.HIGGS | A.AXION, field_grid=G)
    g, o = _pair(st, prm, oracle_lib)
    for w in (0, 1):
        base = np.ones(G * G) if w == 0 else np.zeros(G * G)
        g.set_field(w, base)
        o.field_set(w, base)
    g.update(DT)
    o.update(DT)
    assert len(o.collision_events()) > 100
    for w in (0, 1):
        dg, do = g.get_field(w, "field_dot"), o.field_get(w, "field_dot")
        err = np.sqrt(np.sum((dg - do) ** 2) / np.sum(do ** 2))
        print("field %d: fieldDot rel L2 %.3e" % (w, err))
        assert err < 1e-4


def test_gravitational_collapse_with_merges(oracle_lib):
    # direct-sum gravity pulls a cold cluster together; merges over several updates
    st = _state(800, 8, 64.0, clustered=True, mass=(0.2, 1.0), speed=0.02)
    prm = _params(domain_w=64.0, domain_h=64.0, toggles0=A.GRAVITY)
    g, o = _pair(st, prm, oracle_lib, slots=8)
    total = 0
    for k in range(4):
        assert g.update(0.25) == o.update(0.25), k
        total += len(o.collision_events())
        _compare_events(g, o)
    _compare(g, o, tol=2e-4, spin_tol=2e-2)
    print("events over 4 updates: %d, %d particles left" % (total, o.count()))
    assert total > 20


def test_merges_with_multipole_walk(oracle_lib):
    # Barnes-Hut with the analytic multipole terms: the merge pass builds its
    # tree with the far-field moments, and an update whose merge pass changes
    # nothing walks that same tree in its force pass
    # a sparse box: the initial overlaps merge in the first update, then
    # updates go by with no change at all
    st = _state(400, 1, 256.0)
    prm = _params(domain_w=256.0, domain_h=256.0, toggles0=A.GRAVITY | A.BARNES_HUT, far_field=A.FAR_MULTIPOLE,
                  theta=0.5)
    g, o = _pair(st, prm, oracle_lib)
    total, quiet = 0, 0
    for k in range(5):
        assert g.update(DT) == o.update(DT), k
        new = len(o.collision_events())           # the events of this update
        total += new
        quiet += new == 0
        _compare_events(g, o)
        _compare(g, o, tol=2e-4, spin_tol=2e-2)
    print("events over 5 updates: %d (%d updates without any), %d particles left" % (total, quiet, o.count()))
    assert total > 5 and quiet >= 2


@pytest.mark.parametrize("bh", [False, True])
def test_merges_under_one_pn(bh, oracle_lib):
    # the 1PN force stored for the next kick follows its particle through the
    # merge pass's re-sort and compaction
    st = _state(1000, 9, 64.0, clustered=True, mass=(0.2, 1.0), speed=0.05)
    st["charge"][:] = 0.0
    prm = _params(domain_w=64.0, domain_h=64.0,
                  toggles0=A.GRAVITY | A.RELATIVITY | A.ONE_PN | (A.BARNES_HUT if bh else 0), theta=0.5)
    g, o = _pair(st, prm, oracle_lib, slots=8)
    for k in range(3):
        assert g.update(0.1) == o.update(0.1), k
        _compare_events(g, o)
    _compare(g, o, tol=2e-4, spin_tol=2e-2)
    print("1PN bh=%d: %d particles left of 1000" % (bh, o.count()))
    assert o.count() < 990


def test_merge_slot_space_exhausted_leaves_state_unchanged(oracle_lib):
    """A merge pass whose products would need more slots than the capacity
    fails with GEON_EFULL before it changes any particle: every body is still
    alive with its mass, and the same call succeeds once the state is loaded
    into a handle with room."""
    from physsim_amd import HipPhysics, default_params, GeonError
    st = _state(2048, 71, 128.0, clustered=True, mass=(1.0, 4.0))
    prm = _params()
    n = len(st["x"])
    print("case: merge EFULL n=%d" % n, flush=True)
    g = HipPhysics(default_params(**prm), max_particles=n)      # no spare slot
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"], antimatter=st["anti"])
    with pytest.raises(GeonError) as ei:
        g.update(DT)
    assert ei.value.code == A.EFULL
    d = g.download()
    assert d["alive"].all() and g.getParticleCount() == n
    np.testing.assert_array_equal(d["mass"], f32(st["mass"]))
    g.destroy()
    g, o = _pair(st, prm, oracle_lib)
    assert g.update(DT) == o.update(DT)
    assert g.getParticleCount() < n
    _compare(g, o)


def test_merge_zero_dt_changes_nothing(oracle_lib):
    """update(0) takes no substep (integrator.js:810: dtRemain > EPSILON), so
    touching bodies neither merge nor annihilate."""
    st = _state(2048, 72, 128.0, clustered=True, anti_frac=0.2)
    g, o = _pair(st, _params(), oracle_lib)
    n = len(st["x"])
    assert g.update(0.0) == 0
    assert g.getParticleCount() == n
    d = g.download()
    np.testing.assert_array_equal(d["x"], f32(st["x"]))
    np.testing.assert_array_equal(d["mass"], f32(st["mass"]))
    assert o.update(0.0) == 0 and o.count() == n
