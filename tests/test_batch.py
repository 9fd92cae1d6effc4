"""Several scenes in one batch of launches (mpt_render_batch): the batched
wavefront form of the generated workload (8 terrain-and-instance scenes at
uneven resolutions, 32 spp, depth 12). It replaces rendering the scenes one
by one with camera::render (camera.h:79-105), whose per-image tile loop
(camera.h:92-138) becomes one pool of paths over all the images.

CPU: argument checks. GPU: a reduced batch (three small generated scenes,
two batches of passes) equals the per-scene renders bit for bit, scene k
keyed with scene index base + k, and matches the oracle at equal seed; the
full-size batch of 8 scenes renders every path, finite and bounded, and its
smallest scene equals that scene rendered alone bit for bit."""
import ctypes as C

import numpy as np
import pytest

from mypbr_amd import workloads as W
from mypbr_amd.abi import camera_copy

SMALL = dict(grid=48, instances=12, subdiv=2, variants=3, env_size=(64, 32))


def test_batch_rejects_bad_arguments():
    import mypbr_amd as m
    cam = m.Camera()
    for n in (0, 17):
        rc = m.lib.mpt_render_batch(None, C.byref(cam), n, None, None)
        assert rc == m.abi.MPT_E_INVALID
        assert "scenes" in m.lib.mpt_last_error().decode()


@pytest.mark.gpu
def test_reduced_batch_equals_single_renders_and_oracle():
    import mypbr_amd as m
    import oracle_lib as ol
    built = [W.scene(s, **SMALL) for s in (2, 3, 4)]
    sizes = [(40, 24), (32, 32), (24, 40)]
    scenes, cams, descs = [], [], []
    for (b, cam), (w, h) in zip(built, sizes):
        d, keep = b.build()
        c = camera_copy(cam, image_width=w, aspect_ratio=w / h, samples_per_pixel=4, max_depth=12)
        scenes.append(m.Scene.from_desc(d, c, keepalive=keep))
        cams.append(c)
        descs.append(d)
    print("batch of 3 small scenes, 4 spp, two batches of passes", flush=True)
    frames, stats = m.render_batch(scenes, cams, seed=9, scene_index=5, passes_in_flight=2)
    assert stats.paths == sum(w * h for w, h in sizes) * 4
    for k, (s, c, d) in enumerate(zip(scenes, cams, descs)):
        single = s.render(c, seed=9, scene_index=5 + k).radiance
        assert np.array_equal(frames[k].view(np.uint64), single.view(np.uint64)), k
        cpu = ol.OracleScene(d).render_sums(c, seed=9, passes=4, scene_index=5 + k)["radiance"] / 4
        close = np.all(np.abs(frames[k] - cpu) <= 1e-12 * np.maximum(np.abs(cpu), 1e-30), axis=-1).mean()
        print(f"scene {k}: pixels equal to the oracle {close:.4f}", flush=True)
        assert close >= 0.97
    # a tail in one launch from the first bounce on gives the same frames
    frames_t, _ = m.render_batch(scenes, cams, seed=9, scene_index=5, tail_paths=1 << 30)
    for k in range(3):
        assert np.array_equal(frames_t[k].view(np.uint64), frames[k].view(np.uint64)), k


@pytest.mark.gpu
def test_batch_refuses_scenes_with_different_lights():
    import mypbr_amd as m
    from conftest import scene_path
    b, cam = W.scene(2, **SMALL)
    d, keep = b.build()
    a = m.Scene.from_desc(d, cam, keepalive=keep)
    c = m.Scene.load(scene_path("cornell.scene"))
    with pytest.raises(m.MptError) as e:
        m.render_batch([a, c], [camera_copy(cam, image_width=16), camera_copy(c.camera, image_width=16)])
    assert e.value.code == m.abi.MPT_E_UNSUPPORTED


@pytest.mark.gpu
def test_full_batch_of_eight_generated_scenes():
    import mypbr_amd as m
    scenes, cams = [], []
    for k, (b, cam) in enumerate(W.batch(seed=1)):
        d, keep = b.build()
        scenes.append(m.Scene.from_desc(d, cam, keepalive=keep))
        cams.append(cam)
        print(f"built scene {k}: {cam.image_width}x{cam.height()}", flush=True)
    assert [(c.image_width, c.height()) for c in cams] == W.RESOLUTIONS
    assert all(c.samples_per_pixel == 32 and c.max_depth == 12 for c in cams)
    frames, stats = m.render_batch(scenes, cams, seed=2)
    pixels = sum(w * h for w, h in W.RESOLUTIONS)
    print(f"{stats.paths} paths in {stats.seconds:.2f} s, {stats.extension_rays} extension and "
          f"{stats.shadow_rays} shadow rays", flush=True)
    assert stats.paths == pixels * 32
    assert stats.extension_rays >= stats.paths and stats.shadow_rays > 0
    for k, f in enumerate(frames):
        assert np.all(np.isfinite(f)) and np.all(f >= 0), k
        assert 0.01 < f.mean() < 5.0, (k, f.mean())
    # the smallest scene alone gives its frame of the batch, bit for bit
    alone = scenes[0].render(cams[0], seed=2, scene_index=0).radiance
    assert np.array_equal(alone.view(np.uint64), frames[0].view(np.uint64))
