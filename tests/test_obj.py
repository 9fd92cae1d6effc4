"""Wavefront OBJ/MTL ingestion (obj_loader.h) and what it brings to the GPU
path: mesh lights (mesh.h:175-216) and interpolated vertex normals
(mesh.h:93-101).

The reference orders an OBJ file's parts by the address of each part's
material (a std::map keyed by shared_ptr, obj_loader.h:196), which depends
on the heap; this loader orders them by each material's first face. So
closest-hit IDs are compared with the reference's up to that order: every
part must map to exactly one reference part with the triangles inside in
the same order, and t and the front face must be bitwise equal.

CPU: the oracle against tests/golden/obj_room.npz (written by the compiled
reference through its own load_obj), the loader's semantics and errors.
GPU: the HIP path against the golden hits and converged means, and against
the oracle at equal seed."""
import collections
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden", "obj_room.npz")
SCENE = "obj_room.scene"


def _golden():
    g = np.load(GOLD)
    return g, json.loads(str(g["digest"]))


def _object_ranges(desc):
    return [(int(desc.object_id_base[k]), int(desc.object_count[k])) for k in range(desc.n_objects)]


def assert_ids_match_up_to_part_order(ids, ref_ids, desc):
    """Each object of ours maps to one reference offset, local indices kept."""
    assert np.array_equal(ids < 0, ref_ids < 0)
    ranges = _object_ranges(desc)
    starts = np.array([a for a, _ in ranges])
    hit = ids >= 0
    obj = np.searchsorted(starts, ids[hit], side="right") - 1
    local = ids[hit] - starts[obj]
    offsets = collections.defaultdict(set)
    for o, r in zip(obj, ref_ids[hit] - local):
        offsets[int(o)].add(int(r))
    assert all(len(v) == 1 for v in offsets.values()), offsets
    mapped = [next(iter(v)) for v in offsets.values()]
    assert len(set(mapped)) == len(mapped)


def test_loader_semantics():
    import mypbr_amd as m
    d = m.describe(scene_path(SCENE))
    D = d.desc
    # 7 parts of room.obj (one per material, in first-face order) + blob.obj three times
    assert D.n_objects == 10
    counts = [c for _, c in _object_ranges(D)]
    # floor 2, wall_red 2 + pentagon 3 (degenerate face skipped), wall_green 2, grey 4, lamp 2, gloss 10,
    # ball: 10 x 6 quads = 100 triangles after the 20 pole faces with a repeated vertex lose one triangle each
    assert counts[:7] == [2, 5, 2, 4, 2, 10, 100]
    assert counts[7:] == [168, 168, 168]
    assert D.n_lights == 1 and D.light_object[0] == 4  # the Ke part is a mesh light
    mats = {D.prim_material[int(D.object_first[k])]: k for k in range(D.n_objects)}
    kinds = [D.materials[i].type for i in sorted(mats)]
    assert kinds.count(4) == 1  # one emitter (MPT_MAT_DIFFUSE_LIGHT)
    gloss = D.materials[D.prim_material[int(D.object_first[5])]]
    assert gloss.type == 3 and abs(gloss.roughness - (2 / 82) ** 0.25) < 1e-15 and gloss.ior == 1.45
    grey = D.materials[D.prim_material[int(D.object_first[3])]]
    assert grey.type == 0
    # vertex normals: the ball's from vn, the smooth blobs' made from face areas; the walls have none
    nrm = np.ctypeslib.as_array(D.prim_normals, shape=(D.n_prims * 9,)).reshape(-1, 9)
    first = [int(D.object_first[k]) for k in range(D.n_objects)]
    assert np.all(np.isnan(nrm[first[0]:first[0] + 2, 0]))
    assert not np.any(np.isnan(nrm[first[6]:first[6] + 100, 0]))
    assert not np.any(np.isnan(nrm[first[7]:first[7] + 168, 0]))


def test_oracle_hits_match_reference():
    import mypbr_amd as m
    g, dg = _golden()
    d = m.describe(scene_path(SCENE))
    o = ol.OracleScene(d.desc)
    ids, t, f = o.trace(g["rays"], g["tt"])
    assert dg["ties"] == 0 and dg["unidentified"] == 0
    assert np.array_equal(t.view(np.int64), g["t"].view(np.int64))
    assert np.array_equal(f, g["front"])
    assert_ids_match_up_to_part_order(ids, g["id"], d.desc)


def test_oracle_1m_digest_matches_reference():
    import mypbr_amd as m
    g, dg = _golden()
    d = m.describe(scene_path(SCENE))
    o = ol.OracleScene(d.desc)
    rays, tt = o.ray_batch(d.camera, dg["seed"], dg["n"])
    assert np.array_equal(rays[g["index"]], g["rays"]), "the seeded batch generator changed"
    ids, t, f = o.trace(rays, tt)
    assert hashlib.sha256(t.astype("<f8").tobytes()).hexdigest() == dg["sha256_t"]
    assert hashlib.sha256(f.astype("u1").tobytes()).hexdigest() == dg["sha256_front"]
    assert int((ids == -1).sum()) == dg["misses"]


def test_oracle_bvh_equals_brute_force():
    import mypbr_amd as m
    d = m.describe(scene_path(SCENE))
    o = ol.OracleScene(d.desc)
    rays, tt = o.ray_batch(d.camera, 11, 20000)
    ids, t, _ = o.trace(rays, tt)
    lid, lt = o.trace_linear(rays, tt)
    assert np.array_equal(ids, lid) and np.array_equal(t, lt)


def test_oracle_render_converges_to_reference():
    """Mesh-light NEE (area-weighted triangle choice and pdf) and shading
    normals both enter this image; a wrong pdf biases it."""
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g, _ = _golden()
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    o = ol.OracleScene(d.desc)
    runs = np.stack([o.render(cam, seed=100 + k)["radiance"] for k in range(8)])
    render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, SCENE)


def _scene_with_obj(tmp_path, obj_text, mtl_text=None, line_tail="", extra_files=()):
    (tmp_path / "m.obj").write_text(obj_text)
    if mtl_text is not None:
        (tmp_path / "m.mtl").write_text(mtl_text)
    for name in extra_files:
        (tmp_path / name).write_bytes(b"\x89PNG not really")
    s = tmp_path / "s.scene"
    s.write_text("image_width 16\nsamples_per_pixel 1\ncamera lookfrom 0 0 3 lookat 0 0 0\n"
                 f"mesh m.obj {line_tail}\n")
    return str(s)


TRI = "v 0 0 0\nv 1 0 0\nv 0 1 0\n"


@pytest.mark.parametrize("obj,mtl,tail,files,ok", [
    (TRI + "f 1 2 3\n", None, "", (), True),
    (TRI + "f 1 2 4\n", None, "", (), False),                                  # index out of range
    (TRI + "f 1 x 3\n", None, "", (), False),                                  # bad face
    (TRI + "f -3 -2 -1\n", None, "", (), True),                                # relative indices
    # an image that cannot be decoded falls back to the plain color with a warning (obj_loader.h:83-87)
    ("mtllib m.mtl\nusemtl a\n" + TRI + "f 1 2 3\n", "newmtl a\nKd 1 0 0\nmap_Kd tex.png\n", "", ("tex.png",), True),
    ("mtllib m.mtl\nusemtl a\n" + TRI + "f 1 2 3\n", "newmtl a\nKd 1 0 0\nmap_Kd missing.png\n", "", (), True),
    ("mtllib m.mtl\nusemtl a\n" + TRI + "f 1 2 3\n", "newmtl a\nKd 1 0 0\nnorm n.png\n", "", ("n.png",), True),
    ("mtllib nowhere.mtl\n" + TRI + "f 1 2 3\n", None, "", (), True),          # a missing MTL only warns
    (TRI + "f 1 2 3\n", None, "replace a b", (), False),                       # glTF-only syntax
    ("v 0 0 0\nv 1 0 0\nv 2 0 0\nf 1 2 3\n", None, "", (), False),             # only degenerate faces
])
def test_loader_errors(tmp_path, obj, mtl, tail, files, ok):
    import mypbr_amd as m
    path = _scene_with_obj(tmp_path, obj, mtl, tail, files)
    if ok:
        m.describe(path)
    else:
        with pytest.raises(m.MptError):
            m.describe(path)


def test_normal_maps_off_skips_normal_map_lines(tmp_path):
    import mypbr_amd as m
    path = _scene_with_obj(tmp_path, "mtllib m.mtl\nusemtl a\n" + TRI + "f 1 2 3\n", "newmtl a\nKd 1 0 0\nnorm n.png\n",
                           "", ("n.png",))
    text = open(path).read()
    open(path, "w").write("normal_maps off\n" + text)
    m.describe(path)


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_hits_match_reference():
    import mypbr_amd as m
    g, dg = _golden()
    d = m.describe(scene_path(SCENE))
    rays, tt = ol.OracleScene(d.desc).ray_batch(d.camera, dg["seed"], dg["n"])
    s = m.Scene.from_desc(d.desc, d.camera)
    ids, t, f = s.trace_closest(rays, tt)
    assert hashlib.sha256(t.astype("<f8").tobytes()).hexdigest() == dg["sha256_t"]
    assert hashlib.sha256(f.astype("u1").tobytes()).hexdigest() == dg["sha256_front"]
    assert_ids_match_up_to_part_order(ids[g["index"]], g["id"], d.desc)


@pytest.mark.gpu
def test_gpu_render_matches_oracle():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=48, samples_per_pixel=8, max_depth=16, collect_features=1)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu["radiance"], 0.995, 1e-3)
    _compare(f.normal, cpu["normal"], 0.995, 1e-3)


@pytest.mark.gpu
def test_gpu_render_converges_to_reference():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g, _ = _golden()
    d = m.describe(scene_path(SCENE))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    s = m.Scene.from_desc(d.desc, d.camera)
    runs = np.stack([s.render(cam, seed=300 + k).radiance for k in range(8)])
    ratio, relmse = render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, SCENE)
    print(f"{SCENE}: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)
