"""The generated terrain-and-instances workload (mypbr_amd/workloads.py).

CPU: the generator's counts and determinism, and that the oracle renders a
small instance of it. GPU: a small instance (terrain mesh, instanced
icospheres, sun and sky environment as the only light, fog) renders like
the oracle at equal seed and its closest hits equal the oracle's bit for
bit; the full-size scene (2.09 M terrain triangles + 500 instances)
builds, renders, and traces like the oracle on a seeded ray batch.
"""
import numpy as np
import pytest

from mypbr_amd import workloads as W
from mypbr_amd.abi import camera_copy

SMALL = dict(grid=48, instances=12, subdiv=2, variants=3, env_size=(64, 32))


def test_generator_counts_and_determinism():
    p, f = W.icosphere(4)
    assert f.shape == (5120, 3) and np.allclose(np.linalg.norm(p, axis=1), 1)
    n = 1024
    assert 2 * (n - 1) ** 2 == 2_093_058
    _, tri, _ = W.terrain_mesh(64, 3)
    assert tri.shape == (2 * 63 * 63, 3)
    b1, c1 = W.scene(5, **SMALL)
    b2, c2 = W.scene(5, **SMALL)
    d1, k1 = b1.build()  # the keepalives own the arrays the descriptions point to
    d2, k2 = b2.build()
    assert d1.n_prims == d2.n_prims == 2 * 47 * 47 + 3 * 320 + 6
    assert d1.n_objects == 1 + 12 + 1
    assert np.array_equal(np.ctypeslib.as_array(d1.prim_data, (d1.n_prims * 9,)),
                          np.ctypeslib.as_array(d2.prim_data, (d2.n_prims * 9,)))
    assert list(c1.lookfrom) == list(c2.lookfrom)
    assert len(W.RESOLUTIONS) == 8


def test_sun_sky_environment():
    env = W.sun_sky(256, 128, sun_radiance=1000.0, sun_angle_deg=2.0)
    assert env.shape == (128, 256, 3) and np.all(env >= 0)
    assert env.max() >= 1000  # the disk
    assert env[:5].mean() < env.max() / 100  # the sky is dim beside it


def test_oracle_renders_a_small_instance():
    import oracle_lib as ol
    b, cam = W.scene(2, **SMALL)
    d, keep = b.build()
    r = ol.OracleScene(d).render(camera_copy(cam, image_width=24, samples_per_pixel=2), seed=1)["radiance"]
    assert np.all(np.isfinite(r)) and r.mean() > 0


@pytest.mark.gpu
def test_small_instance_matches_oracle():
    import mypbr_amd as m
    import oracle_lib as ol
    b, cam = W.scene(2, **SMALL)
    d, keep = b.build()
    o = ol.OracleScene(d)
    c = camera_copy(cam, image_width=48, samples_per_pixel=4)
    s = m.Scene.from_desc(d, cam, keepalive=keep)
    gpu = s.render(c, seed=9).radiance
    cpu = o.render(c, seed=9)["radiance"]
    close = np.all(np.abs(gpu - cpu) <= 1e-12 * np.maximum(np.abs(cpu), 1e-30), axis=-1).mean()
    rel = abs(gpu.mean() - cpu.mean()) / cpu.mean()
    print(f"pixels equal {close:.4f}, mean rel {rel:.2e}", flush=True)
    assert close >= 0.995 and rel < 1e-3
    rays, tt = o.ray_batch(cam, 5, 1 << 14)
    gi, gt, _ = s.trace_closest(rays, tt)
    ci, ct, _ = o.trace(rays, tt)
    assert np.array_equal(gi, ci) and np.array_equal(gt, ct)


@pytest.mark.gpu
def test_full_size_scene_builds_renders_and_traces_like_the_oracle():
    import mypbr_amd as m
    import oracle_lib as ol
    b, cam = W.scene(1001)
    d, keep = b.build()
    assert d.n_prims > 2_093_058
    s = m.Scene.from_desc(d, cam, keepalive=keep)
    f = s.render(camera_copy(cam, image_width=64, samples_per_pixel=2), seed=1)
    assert np.all(np.isfinite(f.radiance)) and f.radiance.mean() > 0
    o = ol.OracleScene(d)
    rays, tt = o.ray_batch(cam, 3, 4096)
    gi, gt, _ = s.trace_closest(rays, tt)
    ci, ct, _ = o.trace(rays, tt)
    assert np.array_equal(gi, ci) and np.array_equal(gt, ct)
    assert (gi >= 0).mean() > 0.5
