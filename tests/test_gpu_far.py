"""Far-field corrections of the gfx950 walk against the FP64 oracle: the
analytic multipole terms, the learned network's training samples, its BF16
gradient and forces, and what a network trained on the GPU gains over the
monopole walk at the same opening angle (open boxes: the Barnes-Hut walk
converges to the direct sum there)."""
import numpy as np
import pytest

from physsim_amd import _abi as A
from physsim_amd import workloads as W

pytestmark = pytest.mark.gpu

GC = A.GRAVITY | A.COULOMB | A.BARNES_HUT


def _gpu(st, prm, cap=None):
    from physsim_amd import HipPhysics, default_params
    n = len(st["x"])
    print("case: n=%d toggles=0x%x theta=%.2f far=%d" % (n, prm["toggles0"], prm.get("theta", 0.5),
                                                        prm.get("far_field", 0)), flush=True)
    g = HipPhysics(default_params(**prm), max_particles=cap or n)
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return g


def _orc(st, prm, oracle_lib):
    o = oracle_lib.Oracle(oracle_lib.default_params(**prm), len(st["x"]))
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    return o


def _rel_l2(ax, ay, bx, by):
    return float(np.sqrt(np.sum((ax - bx) ** 2 + (ay - by) ** 2) / np.sum(bx ** 2 + by ** 2)))


def _set(g, **kw):
    p = A.GeonParams.from_buffer_copy(g.params)
    for k, v in kw.items():
        setattr(p, k, v)
    g.set_params(p)


@pytest.mark.parametrize("theta", [0.5, 0.8])
def test_multipole_forces_match_oracle(theta, oracle_lib):
    st, prm, _ = W.config2(4096, open_box=True)
    prm = dict(prm, theta=theta, far_field=A.FAR_MULTIPOLE)
    g, o = _gpu(st, prm), _orc(st, prm, oracle_lib)
    g.compute_forces()
    o.compute_forces()
    fg, fo = g.forces(4096), o.forces()
    assert _rel_l2(fg["fx"], fg["fy"], fo["fx"], fo["fy"]) < 5e-5
    assert 0.5 * fg["pe"].sum() == pytest.approx(fo["pe_total"], rel=2e-5)
    idx = np.arange(0, 4096, 8, dtype=np.int32)
    fx, fy = o.direct_forces(idx)
    _set(g, far_field=A.FAR_NONE)
    g.compute_forces()
    fp = g.forces(4096)
    e_plain = _rel_l2(fp["fx"][idx], fp["fy"][idx], fx, fy)
    e_mp = _rel_l2(fg["fx"][idx], fg["fy"][idx], fx, fy)
    print("theta %.2f: monopole %.3e, multipole %.3e" % (theta, e_plain, e_mp))
    assert e_mp < e_plain / 4


def _by_observer(S, obs):
    out = {}
    order = np.lexsort((S[:, 2], S[:, 0], obs))
    S, obs = S[order], obs[order]
    cut = np.flatnonzero(np.diff(obs)) + 1
    for a, b in zip(np.concatenate([[0], cut]), np.concatenate([cut, [len(obs)]])):
        out[int(obs[a])] = S[a:b]
    return out


def test_far_samples_match_oracle(oracle_lib):
    st, prm = W.plummer_charged(2048, 0)
    prm = dict(prm, theta=0.5)
    g = _gpu(st, prm)
    ns = g.far_collect(1)
    S, obs = g.far_samples(observers=True)
    assert ns == len(S) > 0
    f = g.forces(2048)
    f2ref = float(np.mean(f["fx"].astype(np.float64) ** 2 + f["fy"].astype(np.float64) ** 2))
    pe2ref = float(np.mean(f["pe"].astype(np.float64) ** 2))
    o = _orc(st, prm, oracle_lib)
    So, obso = o.far_training_samples(np.arange(2048, dtype=np.int32), f2ref, pe2ref)
    assert abs(len(So) - len(S)) <= 0.002 * len(So)
    G, O = _by_observer(S, obs), _by_observer(So, obso)
    same = [k for k in O if k in G and len(G[k]) == len(O[k])]
    assert len(same) >= 0.99 * len(O)
    a = np.concatenate([G[k] for k in same])
    b = np.concatenate([O[k] for k in same])
    np.testing.assert_allclose(a[:, :10], b[:, :10], rtol=0, atol=2e-5 * np.abs(b[:, :10]).max())
    np.testing.assert_allclose(a[:, 10:16], b[:, 10:16], rtol=0, atol=1e-4 * np.abs(b[:, 10:16]).max())
    np.testing.assert_allclose(a[:, 16:], b[:, 16:], rtol=1e-4, atol=1e-9 * b[:, 16:].max())


def test_far_list_overflow_is_reported():
    # Each lane's queue holds far_list_cap / 64 nodes (whole 4-entry blocks):
    # a cap of 256 leaves 4 per lane, far fewer than theta 0.5 accepts, and the
    # call must say so rather than drop pairs.  The default cap then runs, and
    # its samples match a second collection of the same state.
    from physsim_amd.geon import GeonError
    st, prm = W.plummer_charged(2048, 3)
    g = _gpu(st, dict(prm, theta=0.5, far_list_cap=256))
    with pytest.raises(GeonError) as ei:
        g.far_collect(1)
    assert ei.value.code == A.EOVERFLOW and "far_list_cap" in str(ei.value)
    _set(g, far_list_cap=0)
    g.far_clear()
    n1 = g.far_collect(1)
    S1, o1 = g.far_samples(observers=True)
    g.far_clear()
    assert g.far_collect(1) == n1 > 0
    S2, o2 = g.far_samples(observers=True)
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_array_equal(S1, S2)


def _wscale(S):
    w = S[:, 16:].ravel()
    w = w[w > 0]
    return len(w) / w.sum()


def test_far_grad_matches_oracle(oracle_lib):
    st, prm = W.plummer_charged(2048, 1)
    prm = dict(prm, theta=0.5)
    g = _gpu(st, prm)
    g.far_collect(1)
    S = g.far_samples()
    g.far_train(0, seed=7)                       # initialises the model
    for k in range(2):
        model = g.far_get_model()
        rows = np.arange(0, min(len(S), 4096), dtype=np.int64)
        Lg, gg = g.far_grad(rows)
        Lo, go = oracle_lib.far_loss_grad(model, S[rows], _wscale(S))
        rel = np.linalg.norm(gg - go) / np.linalg.norm(go)
        cos = float(gg @ go / (np.linalg.norm(gg) * np.linalg.norm(go)))
        print("point %d: loss %.6e vs %.6e, grad rel %.3e, cos %.6f" % (k, Lg, Lo, rel, cos))
        assert Lg == pytest.approx(Lo, rel=2e-2)
        assert rel < 4e-2 and cos > 0.999
        g.far_train(30, batch=2048, lr=1e-3, lr_final=1e-3, seed=9)   # a second, trained point


def test_learned_forces_match_oracle(oracle_lib):
    st, prm = W.plummer_charged(2048, 2)
    prm = dict(prm, theta=0.5)
    g = _gpu(st, prm)
    g.far_collect(1)
    g.far_train(200, batch=4096, seed=5)
    model = g.far_get_model()
    assert np.all(np.isfinite(model))
    g.compute_forces()
    f0 = g.forces(2048)
    _set(g, far_field=A.FAR_LEARNED)
    g.compute_forces()
    f1 = g.forces(2048)
    lp = dict(prm, far_field=A.FAR_LEARNED)
    o = _orc(st, lp, oracle_lib)
    o.far_set_model(model)
    o.compute_forces()
    fo = o.forces()
    assert _rel_l2(f1["fx"], f1["fy"], fo["fx"], fo["fy"]) < 2e-4
    o0 = _orc(st, prm, oracle_lib)
    o0.compute_forces()
    fo0 = o0.forces()
    # the correction itself, BF16 network against the FP64 one: forces and potential
    assert _rel_l2(f1["fx"] - f0["fx"], f1["fy"] - f0["fy"], fo["fx"] - fo0["fx"], fo["fy"] - fo0["fy"]) < 3e-2
    dpe_g = 0.5 * (f1["pe"].sum() - f0["pe"].sum())
    dpe_o = fo["pe_total"] - fo0["pe_total"]
    assert dpe_g == pytest.approx(dpe_o, rel=3e-2)


def test_learned_correction_gain(oracle_lib):
    # Trained on the GPU from direct-sum residuals of 4 Plummer spheres, applied
    # to the (uniform) config-2 box: the network must cut the monopole error.
    from physsim_amd import HipPhysics, default_params
    theta = 0.5
    _, prm = W.plummer_charged(16384, 0)
    T = HipPhysics(default_params(**dict(prm, theta=theta)), max_particles=16384)
    for seed in range(4):
        st, _ = W.plummer_charged(16384, seed)
        T.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
        n = T.far_collect(4)
        print("seed %d: %d samples" % (seed, n), flush=True)
    loss = T.far_train(2000, batch=8192, lr=2e-3, lr_final=2e-5, seed=3)
    print("loss %.4e -> %.4e" % (loss[:20].mean(), loss[-20:].mean()), flush=True)
    assert loss[-20:].mean() < 0.2 * loss[:20].mean()
    model = T.far_get_model()
    st, prm2, _ = W.config2(16384, open_box=True)
    prm2 = dict(prm2, theta=theta, toggles0=GC)
    g = _gpu(st, prm2)
    g.compute_forces()
    fp = g.forces(16384)
    g.far_set_model(model)
    _set(g, far_field=A.FAR_LEARNED)
    g.compute_forces()
    fl = g.forces(16384)
    o = _orc(st, prm2, oracle_lib)
    idx = np.arange(0, 16384, 32, dtype=np.int32)
    fx, fy = o.direct_forces(idx)
    e_plain = _rel_l2(fp["fx"][idx], fp["fy"][idx], fx, fy)
    e_learn = _rel_l2(fl["fx"][idx], fl["fy"][idx], fx, fy)
    print("theta %.2f: monopole %.3e, learned %.3e (x%.1f)" % (theta, e_plain, e_learn, e_plain / e_learn))
    assert e_learn < e_plain / 4
