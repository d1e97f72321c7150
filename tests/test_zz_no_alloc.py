"""The allocation-free pair list (geon_collision_pairs_noalloc) against the
allocating one (geon_collision_pairs) through the C ABI: identical sorted
pairs on two calls in a row, the second with a smaller input; a cap that cuts
the list; and the clear error of a state with more pairs than its fixed
scratch holds."""
import numpy as np
import pytest

from physsim_amd import _abi as A

pytestmark = pytest.mark.gpu


def _state(n, side, seed):
    g = np.random.default_rng(seed)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    return dict(x=f32(g.uniform(1, side - 1, n)), y=f32(g.uniform(1, side - 1, n)), wx=np.zeros(n), wy=np.zeros(n),
                mass=f32(g.uniform(0.5, 4.0, n)), charge=np.zeros(n))


def _upload(g, st):
    g.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"])


@pytest.mark.parametrize("boundary", [A.BOUND_DESPAWN, A.BOUND_LOOP])
def test_noalloc_pairs_match_allocating_entry_point(boundary):
    from physsim_amd import HipPhysics, default_params
    side = 256.0
    prm = default_params(toggles0=A.GRAVITY | A.BARNES_HUT, boundary_mode=boundary, domain_w=side, domain_h=side)
    g = HipPhysics(prm, max_particles=40000)
    for n, seed in ((40000, 1), (9000, 2)):          # the second call's input is smaller
        print("case: noalloc pairs boundary=%d n=%d" % (boundary, n), flush=True)
        _upload(g, _state(n, side, seed))
        a0, b0 = g.collision_pairs()
        a1, b1, total = g.collision_pairs_noalloc()
        assert total == len(a0) > 1000
        np.testing.assert_array_equal(a1, a0)
        np.testing.assert_array_equal(b1, b0)
        assert (a1 < b1).all()
    a2, b2, total = g.collision_pairs_noalloc(cap=100)
    np.testing.assert_array_equal(a2, a0[:100])
    np.testing.assert_array_equal(b2, b0[:100])
    assert total == len(a0)
    g.destroy()


def test_noalloc_pairs_reject_more_than_scratch():
    from physsim_amd import HipPhysics, default_params, GeonError
    side, n = 512.0, 262144
    print("case: noalloc pairs overflow n=%d" % n, flush=True)
    g = HipPhysics(default_params(toggles0=A.GRAVITY | A.BARNES_HUT, domain_w=side, domain_h=side), max_particles=n)
    _upload(g, _state(n, side, 3))
    a, _ = g.collision_pairs()
    assert len(a) > (1 << 20)
    with pytest.raises(GeonError) as ei:
        g.collision_pairs_noalloc()
    assert ei.value.code == A.EOVERFLOW and "exceed its fixed scratch" in str(ei.value)
    g.destroy()
