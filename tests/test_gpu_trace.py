"""GPU closest-hit parity: the HIP traversal against the CPU oracle on the
seeded ray batches of the two plain-method scenes (bvh::hit, bvh.h:45-86).

The bar is bit-identical: same canonical primitive ID, same t bits, same
front_face for every ray (the GPU walks the same binned-SAH BVH in the same
order with the same double-precision arithmetic)."""
import numpy as np
import pytest

from conftest import scene_path

pytestmark = pytest.mark.gpu

N_RAYS = 1 << 20  # the seeded 1M-ray batch


@pytest.fixture(scope="module")
def mpt():
    import mypbr_amd
    return mypbr_amd


@pytest.mark.parametrize("name", ["cornell.scene", "weekend_cover.scene"])
def test_trace_matches_oracle_1m(mpt, name):
    from oracle_lib import OracleScene
    d = mpt.describe(scene_path(name))
    o = OracleScene(d.desc)
    rays, tt = o.ray_batch(d.camera, 1, N_RAYS)
    oid, ot, of = o.trace(rays, tt)
    s = mpt.Scene.from_desc(d.desc, d.camera)
    gid, gt, gf = s.trace_closest(rays, tt)
    print(f"{name}: {N_RAYS} rays, misses {(oid < 0).sum()}", flush=True)
    assert np.array_equal(gid, oid), f"{(gid != oid).sum()} ID mismatches"
    assert np.array_equal(gt.view(np.int64), ot.view(np.int64)), f"{(gt != ot).sum()} t mismatches"
    assert np.array_equal(gf, of)


def test_trace_edge_cases(mpt):
    """Empty batch, rays starting on a surface, zero-direction components,
    grazing rays along a wall and an interval that excludes everything."""
    from oracle_lib import OracleScene
    d = mpt.describe(scene_path("cornell.scene"))
    o = OracleScene(d.desc)
    s = mpt.Scene.from_desc(d.desc, d.camera)
    ids, t, f = s.trace_closest(np.zeros((0, 6)), np.zeros((0, 2)))
    assert ids.shape == (0,)
    inf = np.inf
    rays = np.array([
        [278, 278, -800, 0, 0, 1],        # straight down the axis
        [278, 0, 200, 0, 1, 0],           # from the floor plane
        [0, 100, 100, 1, 0, 0],           # from the red wall along +x
        [100, 0, 100, 1, 0, 0],           # grazing along the floor
        [278, 554, 278, 0, -1, 0],        # from the light plane down
        [278, 278, 278, 1e-300, 0, 1],    # tiny component
        [278, 278, 278, 0, 0, 0],         # zero direction
        [278, 278, -800, 0, 0, 1],        # empty interval
    ], dtype=np.float64)
    tt = np.array([[0.001, inf]] * 7 + [[5.0, 4.0]])
    oid, ot, of = o.trace(rays, tt)
    gid, gt, gf = s.trace_closest(rays, tt)
    assert np.array_equal(gid, oid)
    assert np.array_equal(gt.view(np.int64), ot.view(np.int64))
    assert np.array_equal(gf, of)
