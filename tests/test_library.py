"""The C ABI library on the host (no GPU calls): it loads, exports every
function include/mpt.h declares, flattens the .scene subset like the
reference loader (scene.h), and rejects bad input with an error code and a
message instead of falling back."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mypbr_amd as m
from conftest import ROOT, scene_path
from mypbr_amd import abi


def declared_functions():
    with open(os.path.join(ROOT, "include", "mpt.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(mpt_[a-z_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 15
    lib = C.CDLL(m.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert m.lib.mpt_abi_version() == 1


def test_oracle_is_not_linked_into_the_product():
    """The product library must not route through the CPU oracle."""
    with open(m.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"orc_trace" not in blob and b"orc_render" not in blob and b"liboracle" not in blob


@pytest.mark.parametrize("name,prims,objects,lights", [
    ("cornell.scene", 18, 8, 1),
    ("weekend_cover.scene", 484, 484, 0),
    ("test_lights.scene", 21, 11, 4),
    ("test_media.scene", 19, 9, 1),
])
def test_describe_counts(name, prims, objects, lights):
    d = m.describe(scene_path(name))
    assert (d.desc.n_prims, d.desc.n_objects, d.desc.n_lights) == (prims, objects, lights)


def test_cornell_flattening_matches_scene_file():
    """box() sides in the reference's order (quad.h:105-110) behind the
    rotate_y + translate instance, the light in the lights list."""
    d = m.describe(scene_path("cornell.scene"))
    kinds = [d.desc.object_kind[i] for i in range(8)]
    assert kinds == [abi.OBJ_PRIM] * 6 + [abi.OBJ_LIST] * 2
    assert [d.desc.object_xform[i] for i in range(8)] == [-1] * 6 + [0, 1]
    assert d.desc.light_object[0] == 2 and d.desc.light_strategy == 1
    pd = np.ctypeslib.as_array(d.desc.prim_data, shape=(18 * 9,)).reshape(18, 9)
    # first box side: front quad at z = max with edges dx, dy
    assert np.allclose(pd[6], [0, 0, 165, 165, 0, 0, 0, 330, 0])
    x = d.desc.xforms[0]
    c, s = np.cos(np.radians(15)), np.sin(np.radians(15))
    assert np.allclose(np.array(x.m[:]).reshape(3, 3), [[c, 0, s], [0, 1, 0], [-s, 0, c]])
    assert np.allclose(x.t[:], [265, 0, 295])
    cam = d.camera
    assert (cam.image_width, cam.samples_per_pixel, cam.max_depth, cam.use_sky) == (300, 200, 50, 0)
    assert tuple(cam.lookfrom) == (278, 278, -800) and cam.vfov == 40


def test_weekend_camera_and_materials():
    d = m.describe(scene_path("weekend_cover.scene"))
    cam = d.camera
    assert abs(cam.aspect_ratio - 16 / 9) < 1e-15 and cam.defocus_angle == 0.6 and cam.focus_dist == 10
    types = {d.desc.materials[i].type for i in range(d.desc.n_materials)}
    assert types == {abi.MAT_LAMBERTIAN, abi.MAT_METAL, abi.MAT_DIELECTRIC}


def write(tmp_path, text):
    p = tmp_path / "t.scene"
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("text,needle", [
    ("sphere 0 0 0 0\n", ":1: expected 'sphere"),
    ("material a lambertian 1 1\n", ":1: expected 'material <name> lambertian"),
    ("sphere 0 0 -1 0.5 nosuch\n", ":1: unknown material 'nosuch'"),
    ("material a lambertian 1 1 1\n\nquad 0 0 0 1 0 0 2 0 0 a\n", ":3: quad edges must not be parallel"),
    ("vox models/island.vox\n", ":1: cannot open"),
    ("mesh models/house.gltf\n", ":1: cannot read glTF file"),
    ("material a lambertian 1 1 1\nsphere 0 0 -1 0.5 a extra\n", "unexpected value 'extra'"),
    ("# nothing\n", "the scene has no objects"),
    ("material e emissive 1 1 1\nbox 0 0 0 1 1 1 e scale 1 2 1\n", "lights can only be moved"),
])
def test_parser_errors(tmp_path, text, needle):
    path = write(tmp_path, text)
    with pytest.raises(m.MptError) as e:
        m.describe(path)
    assert needle in str(e.value), str(e.value)
    assert path in str(e.value) or "no objects" in needle


def test_missing_file():
    with pytest.raises(m.MptError) as e:
        m.describe("/nonexistent/x.scene")
    assert e.value.code == abi.MPT_E_IO


def bad_desc(mutate):
    from scenebuild import DescBuilder
    b = DescBuilder()
    mat = b.lambertian((0.5, 0.5, 0.5))
    b.sphere((0, 0, -1), 0.5, mat)
    desc, keep = b.build()
    mutate(desc, keep)
    return desc, keep


@pytest.mark.parametrize("label,mutate,code", [
    ("negative prims", lambda d, k: setattr(d, "n_prims", -1), abi.MPT_E_INVALID),
    ("bad material index", lambda d, k: k["pm"].__setitem__(0, 7), abi.MPT_E_INVALID),
    ("bad object range", lambda d, k: k["oc"].__setitem__(0, 5), abi.MPT_E_INVALID),
    ("bad kind", lambda d, k: k["pk"].__setitem__(0, 9), abi.MPT_E_UNSUPPORTED),
    ("zero radius", lambda d, k: k["pd"].__setitem__(3, 0.0), abi.MPT_E_INVALID),
    ("light out of range", lambda d, k: (setattr(d, "n_lights", 1), setattr(d, "light_object", k["pk"].ctypes.data_as(C.POINTER(C.c_int32))), k["pk"].__setitem__(0, 0)) and None, abi.MPT_E_INVALID),
])
def test_scene_create_rejects_bad_descriptions(label, mutate, code):
    if label == "light out of range":
        from scenebuild import DescBuilder
        b = DescBuilder()
        mat = b.lambertian((0.5, 0.5, 0.5))
        b.sphere((0, 0, -1), 0.5, mat)
        b.light(5)
        desc, keep = b.build()
    else:
        desc, keep = bad_desc(mutate)
    h = C.c_void_p()
    rc = m.lib.mpt_scene_create(C.byref(desc), C.byref(h))
    assert rc == code, (label, rc, m.lib.mpt_last_error())
    assert m.lib.mpt_last_error()


def test_no_cpu_fallback_without_a_gpu(gpu_available):
    """A valid scene on a host without a GPU fails with a HIP error."""
    if gpu_available:
        pytest.skip("a GPU is present")
    d = m.describe(scene_path("cornell.scene"))
    with pytest.raises(m.MptError) as e:
        m.Scene.from_desc(d.desc, d.camera)
    assert e.value.code == abi.MPT_E_HIP


def test_image_height_and_table_entry_points():
    cam = abi.Camera()
    cam.image_width, cam.aspect_ratio = 1920, 16 / 9
    assert m.lib.mpt_image_height(C.byref(cam)) == 1080
    cam.image_width = 1
    assert m.lib.mpt_image_height(C.byref(cam)) == 1
    e, ea = m.ggx_albedo_table()
    assert e.shape == (32, 32) and ea.shape == (32,)
