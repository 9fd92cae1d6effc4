"""Voxel grids (voxel.h, vox_loader.h; the `vox` and `style` lines,
scene.h:612-645, 1040-1128): MagicaVoxel models walked cell by cell with a 3D
DDA over blocks of up to 8^3 cells in a BVH, palette materials, outlined
shapes, emissive voxels whose faces are sampled as lights.

CPU: the .vox reader (models, palette, MATL materials, the scene graph's
translations and rotation bytes, hidden nodes, version 150 without a graph)
checked against the format's rules computed here with numpy; the scene lines
and their errors; the oracle against the compiled reference on its own
loader and walk (tests/golden/voxel_island.npz, voxel_quads.npz: hit t and
front face, the surface record: u, v, tangents, palette or outline colour;
vox_palette.npz: MagicaVoxel's default palette as the reference holds it);
and, as the reference's voxel_test.cpp does, the grid in blocks of 64, 4 and
5 against the same model's visible faces as quads (built here, not by the
loader): same t, face side, normal and material colour. GPU: ray queries and
the grid-versus-faces check through the C-ABI, renders at equal seed against
the oracle, and the converged means against the reference's."""
import os

import numpy as np
import pytest

import oracle_lib as ol
from conftest import ROOT, scene_path

GOLD = os.path.join(ROOT, "tests", "golden")


def _golden(name):
    return np.load(os.path.join(GOLD, name.replace(".scene", ".npz")))


# ---------------------------------------------------------------- the .vox reader


def _xform(desc, k):
    x = desc.xforms[k]
    return np.array(x.m[:]).reshape(3, 3), np.array(x.t[:])


def _rot_byte(r):
    """The signed permutation of a ROTATION byte (the .vox format's
    extension text): bits 0-1 and 2-3 name the non-zero column of rows 0 and
    1, row 2 takes the remaining one, bits 4-6 negate rows 0-2."""
    c0, c1 = r & 3, (r >> 2) & 3
    m = np.zeros((3, 3))
    m[0, c0] = -1 if r & 16 else 1
    m[1, c1] = -1 if r & 32 else 1
    m[2, 3 - c0 - c1] = -1 if r & 64 else 1
    return m


def test_island_models_instances_and_palette():
    import mypbr_amd as m
    from mypbr_amd.abi import (MAT_DIELECTRIC, MAT_DIFFUSE_LIGHT, MAT_LAMBERTIAN, MAT_PRINCIPLED, OBJ_LIST,
                               OBJ_VOXEL, OBJF_LIGHT_ONLY)
    d = m.describe(scene_path("voxel_island.scene"))
    desc = d.desc
    assert desc.n_vox_models == 2
    dims = [(desc.vox_models[k].nx, desc.vox_models[k].ny, desc.vox_models[k].nz) for k in range(2)]
    assert sorted(dims) == [(5, 5, 6), (40, 40, 18)]  # in the order the scene graph places them
    objs = [o for o in range(desc.n_objects) if desc.object_kind[o] == OBJ_VOXEL]
    assert len(objs) == 4  # the terrain and three towers: the hidden node places nothing
    flags = [desc.object_flags[o] for o in range(desc.n_objects)]
    lit = [o for o in range(desc.n_objects) if flags[o] & OBJF_LIGHT_ONLY]
    assert len(lit) == 1 and desc.object_kind[lit[0]] == OBJ_LIST  # the lantern's faces
    assert lit[0] in [desc.light_object[k] for k in range(desc.n_lights)]
    # palette materials: glass, metal (and plastic), emitter, diffuse for _blend
    t0 = desc.vox_models[dims.index((40, 40, 18))]
    mt = {i: desc.materials[t0.material[i]].type for i in range(256) if t0.material[i] >= 0}
    assert mt[5] == MAT_DIELECTRIC and t0.kind[5] == 2
    assert mt[6] == MAT_PRINCIPLED and mt[7] == MAT_DIFFUSE_LIGHT and mt[3] == MAT_LAMBERTIAN
    assert t0.material[9] == -1  # no terrain cell uses the roof: no material for it
    assert t0.edge_width == 0.06
    # instance transforms: z-up to y-up, then the scene line's translation,
    # after the graph's translation and rotation and the model's centring
    zup = np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], float)
    want = {(9, -8, 9, 4), (-10, 6, 8, 1 | (2 << 2)), (-4, -12, 9, 4 | 32)}
    got = set()
    for o in objs:
        mm, tt = _xform(desc, desc.object_xform[o])
        for tx, ty, tz, r in want:
            R = _rot_byte(r)
            c = np.array([5 // 2, 5 // 2, 6 // 2])  # the tower's centre, rounded down
            m_exp = zup @ R
            t_exp = zup @ (np.array([tx, ty, tz]) - R @ c) + np.array([0, 0.01, 0])
            if np.allclose(mm, m_exp, atol=1e-12) and np.allclose(tt, t_exp, atol=1e-12):
                got.add((tx, ty, tz, r))
    assert got == want
    terrain = [o for o in objs if np.allclose(_xform(desc, desc.object_xform[o])[0], zup, atol=1e-12) and
               np.allclose(_xform(desc, desc.object_xform[o])[1], zup @ -np.array([20, 20, 9]) + [0, 0.01, 0])]
    assert len(terrain) == 1


def test_version_150_file_without_graph_or_palette():
    import mypbr_amd as m
    from mypbr_amd.abi import OBJ_VOXEL
    d = m.describe(scene_path("vox_palette.scene"))
    desc = d.desc
    assert desc.n_vox_models == 2
    objs = [o for o in range(desc.n_objects) if desc.object_kind[o] == OBJ_VOXEL]
    zup = np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], float)
    for o, (nx, ny, nz) in zip(objs, [(255, 1, 1), (2, 2, 3)]):
        mm, tt = _xform(desc, desc.object_xform[o])
        assert np.allclose(mm, zup)  # each model on z = 0, centred in x and y
        assert np.allclose(tt, zup @ (np.array([0, 0, nz // 2]) - np.array([nx // 2, ny // 2, nz // 2])))


@pytest.mark.parametrize("line,ok", [
    ("vox models/island.vox", True), ("vox models/island.vox quads", True),
    ("vox models/island.vox emission 2 edges 0.1 rotate_y 30", True), ("vox models/island.vox edges 0.2 1 0 0", True),
    ("vox models/island.vox edges 0.5", False), ("vox models/island.vox edges 0", False),
    ("vox models/island.vox emission -1", False), ("vox models/island.vox red", False),
    ("vox models/island.vox scale 1 2 1", False),  # emissive voxels need a uniform scale
    ("vox models/nosuch.vox", False), ("vox models/room.obj", False), ("vox", False),
])
def test_vox_lines(tmp_path, line, ok):
    import mypbr_amd as m
    s = tmp_path / "s.scene"
    s.write_text(f"image_width 8\nmaterial red lambertian 0.5 0 0\n{line.replace('models/', ROOT + '/scenes/models/')}\n")
    if ok:
        desc = m.describe(str(s)).desc
        assert desc.n_vox_models == (0 if "quads" in line else 2)  # quads: the faces, no grids
        assert desc.n_objects > (1000 if "quads" in line else 4)
    else:
        with pytest.raises(m.MptError):
            m.describe(str(s))


def test_style_lines(tmp_path):
    """style voxel (scene.h:612-645): ACES, bloom and vignette 0.25, sun at
    50/50 of intensity 4 and angle 3, sky atmosphere, the ground of grey 0.3,
    bounces 2/5/5, transparent shadows; later lines change any of it; style
    classic resets."""
    import mypbr_amd as m
    from mypbr_amd.abi import MAT_LAMBERTIAN
    s = tmp_path / "s.scene"
    s.write_text("image_width 8\nstyle voxel\nsphere 0 1 0 1\n")
    d = m.describe(str(s))
    assert d.desc.has_ground == 1 and d.desc.has_sun == 1
    g = d.desc.materials[d.desc.ground_material]
    assert g.type == MAT_LAMBERTIAN and list(d.desc.textures[g.texture].even) == [0.3, 0.3, 0.3]
    c = d.camera
    assert (c.max_diffuse, c.max_specular, c.max_scatter, c.transparent_shadows) == (2, 5, 5, 1)
    assert abs(d.desc.sun_angle_degrees - 3) < 1e-12
    s.write_text("image_width 8\nstyle voxel\nbounces diffuse 7\nstyle classic\nsphere 0 1 0 1\n")
    d = m.describe(str(s))
    assert d.desc.has_ground == 0 and d.desc.has_sun == 0
    assert (d.camera.max_diffuse, d.camera.max_specular, d.camera.transparent_shadows) == (0, 0, 0)
    with pytest.raises(m.MptError):
        s.write_text("image_width 8\nstyle fancy\n")
        m.describe(str(s))


# ---------------------------------------------------------------- oracle against the reference


@pytest.mark.parametrize("name", ["voxel_island.scene", "voxel_quads.scene"])
def test_oracle_hits_and_surfaces_match_reference(name):
    """The subset of the seeded 1M-ray batch: t and front face bitwise, and
    every hit's surface record (u, v, dp/du, dp/dv, the palette or outline
    colour, the shading normal) bitwise. The reference names a whole voxel
    model one primitive, so IDs are not compared."""
    import mypbr_amd as m
    g = _golden(name)
    d = m.describe(scene_path(name))
    o = ol.OracleScene(d.desc)
    ids, t, f = o.trace(g["rays"], g["tt"])
    hit = g["id"] != -1
    assert hit.sum() > 1000 and (~hit).sum() > 100
    assert np.array_equal(ids != -1, hit)
    assert np.array_equal(t[hit].view(np.int64), g["t"][hit].view(np.int64))
    assert np.array_equal(f[hit], g["front"][hit])
    got, want = o.surface(g["rays"], g["tt"]), g["surface"]
    assert np.array_equal(np.isnan(got), np.isnan(want))
    h = ~np.isnan(want[:, 0])
    assert np.array_equal(got[h].view(np.int64), want[h].view(np.int64))
    colours = np.unique(np.round(want[h, 8:11], 9), axis=0)
    assert len(colours) >= 5  # sand, grass, rock, the outline colour, ...


def test_default_palette_matches_reference():
    """A .vox file without an RGBA chunk takes MagicaVoxel's default palette;
    the oracle's surface colour of each of the 255 voxels (sRGB-decoded
    palette entry) equals the reference's bitwise."""
    import mypbr_amd as m
    g = np.load(os.path.join(GOLD, "vox_palette.npz"))
    d = m.describe(scene_path("vox_palette.scene"))
    o = ol.OracleScene(d.desc)
    got = o.surface(g["rays"], g["tt"])
    assert not np.isnan(g["surface"][:, 8]).any()
    assert np.array_equal(got[:, 8:11].view(np.int64), g["surface"][:, 8:11].view(np.int64))
    _, t, _ = o.trace(g["rays"], g["tt"])
    assert np.array_equal(t, g["trace_t"])
    # every entry once, but the ray above index 128 meets the file's other
    # model first, a column of index 200 standing at the origin
    assert len(np.unique(g["surface"][:, 8:11], axis=0)) == 254


def test_oracle_render_converges_to_reference():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = _golden("voxel_island.scene")
    d = m.describe(scene_path("voxel_island.scene"))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    runs = np.stack([ol.OracleScene(d.desc).render(cam, seed=200 + k)["radiance"] for k in range(8)])
    render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, "voxel_island.scene")


# ---------------------------------------------------------------- grid against faces


def _random_model(seed=3, shape=(11, 9, 13)):
    """voxel_test.cpp's model: empty, opaque, glass and emissive cells at
    random, and a solid 4^3 glass block in one corner."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    x = rng.random(shape)
    c = np.where(x < 0.55, 0, np.where(x < 0.8, 1 + (rng.random(shape) * 2).astype(int),
                                       np.where(x < 0.95, 3 + (rng.random(shape) * 2).astype(int), 5)))
    c[:4, :4, :4] = 3
    return c.astype(np.uint8)


def _grid_and_faces(block):
    """Two descriptions of one model: its voxel grid in blocks of `block`,
    and its visible faces as quads built here from the surface rule (a face
    between cells of different kinds: opaque towards empty or glass, glass
    towards empty), wound so that cross(u, v) points out of the cell."""
    from mypbr_amd.scenebuild import DescBuilder
    cells = _random_model()
    nz, ny, nx = cells.shape
    kinds = np.zeros(256, int)
    kinds[1:3] = 1
    kinds[3:5] = 2
    kinds[5] = 1

    def build(grid):
        b = DescBuilder()
        mats = [-1] * 256
        for i in (1, 2):
            mats[i] = b.lambertian(((i % 7) / 7.0 + 0.05, (i % 5) / 5.0 + 0.05, (i % 3) / 3.0 + 0.05))
        for i in (3, 4):
            mats[i] = b.dielectric(1.5 if i == 3 else 1.33)
        mats[5] = b.diffuse_light((4, 3, 2))
        if grid:
            mi = b.vox_model(cells, kinds, mats)
            first, count = b.voxel_blocks(mi, block)
            b.voxel(first, count)
            return b.build()
        kind_at = lambda x, y, z: kinds[cells[z, y, x]] if (0 <= x < nx and 0 <= y < ny and 0 <= z < nz) else 0  # noqa: E731
        first, n = len(b.prim_kind), 0
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    k = kinds[cells[z, y, x]]
                    if k == 0:
                        continue
                    for a in range(3):
                        for s in (1, -1):
                            nb = [x, y, z]
                            nb[a] += s
                            kn = kind_at(*nb)
                            if not (kn == 0 or (k == 1 and kn == 2)):
                                continue
                            q = [float(x), float(y), float(z)]
                            if s > 0:
                                q[a] += 1
                            e1, e2 = [0.0] * 3, [0.0] * 3
                            e1[(a + 1) % 3] = 1
                            e2[(a + 2) % 3] = 1
                            u, v = (e1, e2) if s > 0 else (e2, e1)
                            b._prim(1, mats[cells[z, y, x]], [*q, *u, *v])
                            n += 1
        b._obj(1, first, n)  # one list: every face tested, the closest kept
        return b.build()
    return build(True), build(False), cells.shape


def _grid_rays(shape, n, seed=7):
    """voxel_test.cpp's rays: from outside the box aimed at a point inside,
    or from inside in any direction; a third with a finite t max."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    target = rng.random((n, 3)) * [nx, ny, nz]
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    outside = np.arange(n) % 2 == 1
    origin = np.where(outside[:, None], target + 30 * u, target)
    direction = np.where(outside[:, None], target - origin, u * rng.uniform(0.5, 2, (n, 1)))
    tmax = np.where(np.arange(n) % 3 == 0, rng.uniform(0.1, 40, n), np.inf)
    return np.hstack([origin, direction]), np.stack([np.full(n, 0.001), tmax], 1)


@pytest.mark.parametrize("block", [64, 4, 5])
def test_oracle_grid_matches_its_faces(block):
    """t within 1e-9 relative (the walk's plane times against the quads'
    plane intersections), side, shading normal and colour equal."""
    (dg, kg), (df, kf), shape = _grid_and_faces(block)
    og, of = ol.OracleScene(dg), ol.OracleScene(df)
    rays, tt = _grid_rays(shape, 20000)
    ig, tg, fg = og.trace(rays, tt)
    i_f, tf, ff = of.trace(rays, tt)
    hg, hf = ig != -1, i_f != -1
    assert np.array_equal(hg, hf)
    assert hg.sum() > 5000
    assert np.all(np.abs(tg[hg] - tf[hg]) <= 1e-9 * np.maximum(1, tf[hg]))
    assert np.array_equal(fg[hg], ff[hg])
    sg, sf = og.surface(rays, tt), of.surface(rays, tt)
    assert np.array_equal(sg[hg][:, 8:14], sf[hg][:, 8:14])


@pytest.mark.parametrize("block", [4, 8])
@pytest.mark.gpu
def test_gpu_grid_matches_its_faces_and_the_oracle(block):
    import mypbr_amd as m
    (dg, kg), (df, kf), shape = _grid_and_faces(block)
    rays, tt = _grid_rays(shape, 200000, seed=11)
    sg, sf = m.Scene.from_desc(dg, keepalive=kg), m.Scene.from_desc(df, keepalive=kf)
    ig, tg, fg = sg.trace_closest(rays, tt)
    i_f, tf, ff = sf.trace_closest(rays, tt)
    hg, hf = ig != -1, i_f != -1
    assert np.array_equal(hg, hf) and hg.sum() > 50000
    assert np.all(np.abs(tg[hg] - tf[hg]) <= 1e-9 * np.maximum(1, tf[hg]))
    assert np.array_equal(fg[hg], ff[hg])
    # the GPU walk is the oracle's, bit for bit
    io, to, fo = ol.OracleScene(dg).trace(rays[:20000], tt[:20000])
    assert np.array_equal(ig[:20000], io) and np.array_equal(tg[:20000], to) and np.array_equal(fg[:20000], fo)


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["voxel_island.scene", "voxel_quads.scene"])
def test_gpu_hits_match_reference(name):
    import mypbr_amd as m
    g = _golden(name)
    d = m.describe(scene_path(name))
    ids, t, f = m.Scene.from_desc(d.desc, d.camera).trace_closest(g["rays"], g["tt"])
    hit = g["id"] != -1
    assert np.array_equal(ids != -1, hit)
    assert np.array_equal(t[hit].view(np.int64), g["t"][hit].view(np.int64))
    assert np.array_equal(f[hit], g["front"][hit])


@pytest.mark.gpu
def test_gpu_render_matches_oracle():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_gpu_render import _compare
    d = m.describe(scene_path("voxel_island.scene"))
    cam = camera_copy(d.camera, image_width=64, samples_per_pixel=8, max_depth=12, collect_features=1)
    cpu = ol.OracleScene(d.desc).render(cam, seed=7)
    f = m.Scene.from_desc(d.desc, d.camera).render(cam, seed=7)
    _compare(f.radiance, cpu["radiance"], 0.995, 1e-3)
    _compare(f.albedo, cpu["albedo"], 0.995, 1e-3)
    _compare(f.normal, cpu["normal"], 0.995, 1e-3)


@pytest.mark.gpu
def test_gpu_render_converges_to_reference():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    from test_oracle_golden import render_agrees
    g = _golden("voxel_island.scene")
    d = m.describe(scene_path("voxel_island.scene"))
    cam = camera_copy(d.camera, image_width=int(g["width"]), samples_per_pixel=int(g["spp"]) // 8,
                      max_depth=int(g["max_depth"]))
    s = m.Scene.from_desc(d.desc, d.camera)
    runs = np.stack([s.render(cam, seed=300 + k).radiance for k in range(8)])
    ratio, relmse = render_agrees(runs.mean(0), runs.std(0, ddof=1) / np.sqrt(8), g, "voxel_island.scene")
    print(f"voxel_island.scene: total ratio {ratio:.5f}, relMSE {relmse:.2e}", flush=True)
