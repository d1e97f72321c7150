"""Pins of the oracle's non-orientable topologies (topology.js): minImage on
the Klein bottle and RP^2 against brute-force enumeration of each quotient
group's images, and wrapPosition's glide crossings against their closed form.
"""
import numpy as np
import pytest

from physsim_amd import _abi as A


def _sys(oracle_lib, topo, W=64.0, H=48.0, n=1, **kw):
    prm = oracle_lib.default_params(toggles0=A.GRAVITY, softening=0.5, boundary_mode=A.BOUND_LOOP,
                                    topology=topo, domain_w=W, domain_h=H, **kw)
    return oracle_lib.Oracle(prm, n)


def _orbit(topo, sx, sy, W, H, depth=3):
    """Images of (sx, sy) under the group that defines the surface: the torus
    translations; Klein: x + W and the glide (x, y) -> (W - x, y + H); RP^2:
    the glides (x, y) -> (W - x, y + H) and (x, y) -> (x + W, H - y).  Each
    generator with its inverse, words up to `depth` (enough near the domain)."""
    if topo == A.TORUS:
        gens = [lambda x, y: (x + W, y), lambda x, y: (x - W, y), lambda x, y: (x, y + H), lambda x, y: (x, y - H)]
    elif topo == A.KLEIN:
        gens = [lambda x, y: (x + W, y), lambda x, y: (x - W, y),
                lambda x, y: (W - x, y + H), lambda x, y: (W - x, y - H)]
    else:
        gens = [lambda x, y: (W - x, y + H), lambda x, y: (W - x, y - H),
                lambda x, y: (x + W, H - y), lambda x, y: (x - W, H - y)]
    pts = {(sx, sy)}
    frontier = [(sx, sy)]
    for _ in range(depth):
        nxt = []
        for (x, y) in frontier:
            for gfn in gens:
                q = tuple(round(v, 9) for v in gfn(x, y))
                if q not in pts:
                    pts.add(q)
                    nxt.append(q)
        frontier = nxt
    return np.array(sorted(pts))


@pytest.mark.parametrize("topo", [A.TORUS, A.KLEIN, A.RP2])
def test_min_image_is_nearest_image(oracle_lib, topo):
    W, H = 64.0, 48.0
    o = _sys(oracle_lib, topo, W, H)
    g = np.random.default_rng(topo + 3)
    for _ in range(300):
        ox, oy, sx, sy = g.uniform(0, W), g.uniform(0, H), g.uniform(0, W), g.uniform(0, H)
        d = o.min_image(ox, oy, sx, sy)
        imgs = _orbit(topo, sx, sy, W, H)
        dist = np.hypot(imgs[:, 0] - ox, imgs[:, 1] - oy)
        assert np.hypot(*d) == pytest.approx(dist.min(), abs=1e-9)
        # and the displacement points at an image of the source
        assert np.min(np.hypot(imgs[:, 0] - (ox + d[0]), imgs[:, 1] - (oy + d[1]))) < 1e-9


def test_klein_glide_wrap(oracle_lib):
    # crossing y = H: y -> y - H, x -> W - x, w_x, v_x, angw and the spin negate (topology.js:95-109)
    W, H = 64.0, 48.0
    o = _sys(oracle_lib, A.KLEIN, W, H, n=2)
    o.upload(np.array([10.0, 40.0]), np.array([47.9, 20.0]), np.array([0.05, 0.0]), np.array([0.4, 0.0]),
             np.array([1e-6, 1e-6]), np.zeros(2), np.array([0.3, 0.0]))
    o.update(1.0)
    d = o.download()
    i = int(np.flatnonzero(d["id"] == 0)[0])
    g = 1 / np.sqrt(1 + 0.05 ** 2 + 0.4 ** 2)            # relativity off here: v = w
    y = 47.9 + 0.4 - H
    assert d["y"][i] == pytest.approx(y, abs=1e-6)
    assert d["x"][i] == pytest.approx(W - (10.0 + 0.05), abs=1e-6)
    assert d["wx"][i] == pytest.approx(-0.05, rel=1e-6) and d["wy"][i] == pytest.approx(0.4, rel=1e-6)
    assert d["angw"][i] == pytest.approx(-0.3, rel=1e-6)


def test_rp2_glide_wraps(oracle_lib):
    # RP^2: crossing x = W mirrors y and negates w_y; crossing y = 0 mirrors x and negates w_x (topology.js:113-134)
    W, H = 64.0, 48.0
    o = _sys(oracle_lib, A.RP2, W, H, n=2)
    o.upload(np.array([63.8, 30.0]), np.array([0.1, 20.0]), np.array([0.5, 0.0]), np.array([-0.3, 0.0]),
             np.array([1e-6, 1e-6]), np.zeros(2), np.array([0.2, 0.0]))
    o.update(1.0)
    d = o.download()
    i = int(np.flatnonzero(d["id"] == 0)[0])
    x1, y1 = 63.8 + 0.5, 0.1 - 0.3                       # both edges crossed
    x2, y2 = x1 - W, H - y1                              # x-glide: y mirrored to H + 0.2, w_y and spin negate
    x3, y3 = W - x2, y2 - H                              # then the y-glide: x mirrored, w_x and spin negate
    assert d["x"][i] == pytest.approx(x3, abs=1e-6) and d["y"][i] == pytest.approx(y3, abs=1e-6)
    assert d["wx"][i] == pytest.approx(-0.5, rel=1e-6) and d["wy"][i] == pytest.approx(0.3, rel=1e-6)
    assert d["angw"][i] == pytest.approx(0.2, rel=1e-6)
