"""Programmatic scene descriptions (mpt_scene_desc) for tests, benchmarks
and the generated workloads.

Builds the C ABI's flattened description from Python lists the way the
C++ scene API flattens objects (include/mpt/mpt.hpp): materials carry the
derived quantities the reference computes in its constructors
(roughness_to_alpha, Schlick F_avg, dielectric f0; material.h, microfacet.h).
"""
import ctypes as C
import math

import numpy as np

from . import abi


def roughness_to_alpha(r):  # microfacet.h:17-19
    return max(r * r, 1e-3)


class DescBuilder:
    def __init__(self):
        self.prim_kind, self.prim_mat, self.prim_data = [], [], []
        self.obj_kind, self.obj_first, self.obj_count, self.obj_xform = [], [], [], []
        self.obj_density, self.obj_mat, self.obj_bkind, self.obj_flags = [], [], [], []
        self.vox_models, self.vox_cells = [], []
        self.xforms, self.materials, self.textures = [], [], []
        self.lights = []
        self.light_strategy = 1
        self.uniform_fraction = 0.2
        self.env = None
        self.env_intensity = 1.0
        self.env_rotation = 0.0

    # ---------------------------------------------------------------- textures and materials
    def texture_solid(self, rgb):
        t = abi.Texture()
        t.type = abi.TEX_SOLID
        t.inv_scale = 0
        for k in range(3):
            t.even[k] = t.odd[k] = rgb[k]
        self.textures.append(t)
        return len(self.textures) - 1

    def texture_checker(self, scale, even, odd):
        t = abi.Texture()
        t.type = abi.TEX_CHECKER
        t.inv_scale = 1.0 / scale
        for k in range(3):
            t.even[k], t.odd[k] = even[k], odd[k]
        self.textures.append(t)
        return len(self.textures) - 1

    def _mat(self, type_, tex=-1):
        m = abi.Material()
        m.type = type_
        m.texture = tex
        m.spread = 180
        m.tan_half = math.inf
        m.normal_map = m.metallic_texture = m.roughness_texture = -1
        return m

    def _add_mat(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def lambertian(self, rgb=None, texture=None):
        tex = texture if texture is not None else self.texture_solid(rgb)
        return self._add_mat(self._mat(abi.MAT_LAMBERTIAN, tex))

    def metal(self, rgb, roughness, conductor=None):
        """conductor: a metal's name (Al, Ti, Cu, Au, Ag, Cr) for a measured
        metal; rgb is then unused, the scene build derives the head-on color
        from the index."""
        m = self._mat(abi.MAT_METAL)
        r = min(max(roughness, 0.0), 1.0)
        m.roughness, m.alpha = r, roughness_to_alpha(r)
        for k in range(3):
            m.albedo[k] = rgb[k]
            m.f_avg[k] = rgb[k] + (1 / 21) * (1 - rgb[k])  # f0 + (1 - f0) / 21, microfacet.h:102-107
        if conductor is not None:
            self._measure(m, conductor)
        return self._add_mat(m)

    @staticmethod
    def _measure(m, conductor):
        from . import lib  # the metal's index from the core's table (find_conductor)
        eta, k = (C.c_double * 3)(), (C.c_double * 3)()
        if lib.mpt_find_conductor(conductor.encode(), eta, k) != 0:
            raise ValueError(f"unknown conductor {conductor!r}")
        m.measured = 1
        for c in range(3):
            m.eta[c], m.k[c] = eta[c], k[c]

    def thin_glass(self, ior, tint=(1, 1, 1)):
        m = self._mat(abi.MAT_DIELECTRIC)
        m.ior, m.thin = ior, 1
        for k in range(3):
            m.albedo[k] = tint[k]
        return self._add_mat(m)

    def dielectric(self, ior, roughness=0.0, sigma_a=(0, 0, 0)):
        m = self._mat(abi.MAT_DIELECTRIC)
        r = min(max(roughness, 0.0), 1.0)
        m.ior, m.roughness, m.alpha = ior, r, roughness_to_alpha(r)
        for k in range(3):
            m.sigma_a[k] = sigma_a[k]
        return self._add_mat(m)

    def principled(self, rgb, metallic, roughness, ior=1.5, texture=None, conductor=None, coat=None,
                   coat_roughness=0.0, coat_tint=(1.0, 1.0, 1.0)):
        """coat: a clear coat of that weight (set_clearcoat, material.h:363-379),
        with its roughness and tint, as solid textures."""
        tex = texture if texture is not None else self.texture_solid(rgb)
        m = self._mat(abi.MAT_PRINCIPLED, tex)
        r = min(max(roughness, 0.0), 1.0)
        m.metallic = min(max(metallic, 0.0), 1.0)
        m.roughness, m.alpha, m.ior = r, roughness_to_alpha(r), ior
        r0 = (ior - 1) / (ior + 1)
        m.f0 = r0 * r0
        if conductor is not None:
            self._measure(m, conductor)
        if coat is not None:
            m.coat = 1
            m.coat_weight_texture = self.texture_solid((coat, coat, coat))
            m.coat_roughness_texture = self.texture_solid((coat_roughness,) * 3)
            m.coat_normal_map = -1
            for k in range(3):
                m.coat_tint[k] = coat_tint[k]
        return self._add_mat(m)

    def diffuse_light(self, rgb, spread=180.0):
        m = self._mat(abi.MAT_DIFFUSE_LIGHT)
        m.spread = spread
        half = spread * math.pi / 180.0 / 2
        m.tan_half = math.inf if spread >= 180 else math.tan(half)
        if half >= math.pi / 2 - 1e-9:
            share = 1.0
        else:
            s, c = math.sin(half), math.cos(half)
            share = s * s - (half - s * c) * c / s
        for k in range(3):
            m.emit[k] = rgb[k]
            m.avg_emission[k] = share * rgb[k]
        return self._add_mat(m)

    def phase(self, rgb, g=0.0):
        m = self._mat(abi.MAT_PHASE)
        m.g = g
        for k in range(3):
            m.albedo[k] = rgb[k]
        return self._add_mat(m)

    # ---------------------------------------------------------------- geometry
    def _prim(self, kind, mat, data):
        d = list(data) + [0.0] * (9 - len(data))
        self.prim_kind.append(kind)
        self.prim_mat.append(mat)
        self.prim_data.append(np.asarray(d, dtype=np.float64))
        return len(self.prim_kind) - 1

    def _obj(self, kind, first, count, xform=-1, density=0.0, mat=-1, bkind=abi.OBJ_PRIM):
        self.obj_kind.append(kind)
        self.obj_first.append(first)
        self.obj_count.append(count)
        self.obj_xform.append(xform)
        self.obj_density.append(density)
        self.obj_mat.append(mat)
        self.obj_bkind.append(bkind)
        self.obj_flags.append(0)
        return len(self.obj_kind) - 1

    def xform(self, m, t):
        """Affine transform p -> m p + t with its inverse (np.linalg.inv)."""
        m = np.asarray(m, dtype=np.float64).reshape(3, 3)
        t = np.asarray(t, dtype=np.float64)
        im = np.linalg.inv(m)
        x = abi.Xform()
        for i in range(9):
            x.m[i] = m.ravel()[i]
            x.inv_m[i] = im.ravel()[i]
        it = -(im @ t)
        for i in range(3):
            x.t[i] = t[i]
            x.inv_t[i] = it[i]
        self.xforms.append(x)
        return len(self.xforms) - 1

    def sphere(self, center, radius, mat, xform=-1):
        return self._obj(abi.OBJ_PRIM, self._prim(abi.PRIM_SPHERE, mat, [*center, radius]), 1, xform)

    def quad(self, Q, u, v, mat, xform=-1):
        return self._obj(abi.OBJ_PRIM, self._prim(abi.PRIM_QUAD, mat, [*Q, *u, *v]), 1, xform)

    def box_prims(self, a, b, mat):
        """The six quads of box() (quad.h:95-113); returns the first index."""
        lo = [min(a[k], b[k]) for k in range(3)]
        hi = [max(a[k], b[k]) for k in range(3)]
        dx, dy, dz = [hi[0] - lo[0], 0, 0], [0, hi[1] - lo[1], 0], [0, 0, hi[2] - lo[2]]
        neg = lambda v: [-x for x in v]  # noqa: E731
        sides = [([lo[0], lo[1], hi[2]], dx, dy), ([hi[0], lo[1], hi[2]], neg(dz), dy),
                 ([hi[0], lo[1], lo[2]], neg(dx), dy), ([lo[0], lo[1], lo[2]], dz, dy),
                 ([lo[0], hi[1], hi[2]], dx, neg(dz)), ([lo[0], lo[1], lo[2]], dx, dz)]
        first = None
        for Q, u, v in sides:
            p = self._prim(abi.PRIM_QUAD, mat, [*Q, *u, *v])
            first = p if first is None else first
        return first

    def box(self, a, b, mat, xform=-1):
        return self._obj(abi.OBJ_LIST, self.box_prims(a, b, mat), 6, xform)

    def mesh_prims(self, positions, indices, mat):
        """Triangles of an indexed mesh; returns (first, count)."""
        pos = np.asarray(positions, dtype=np.float64)
        idx = np.asarray(indices, dtype=np.int64).reshape(-1, 3)
        tri = pos[idx].reshape(-1, 9)
        first = len(self.prim_kind)
        n = tri.shape[0]
        self.prim_kind.extend([abi.PRIM_TRIANGLE] * n)
        self.prim_mat.extend([mat] * n)
        self.prim_data.append(np.ascontiguousarray(tri.ravel()))
        return first, n

    def mesh(self, first, count, xform=-1):
        return self._obj(abi.OBJ_MESH, first, count, xform)

    def medium_sphere(self, center, radius, density, phase_mat, xform=-1):
        p = self._prim(abi.PRIM_SPHERE, phase_mat, [*center, radius])
        return self._obj(abi.OBJ_MEDIUM, p, 1, xform, density, phase_mat, abi.OBJ_PRIM)

    def medium_box(self, a, b, density, phase_mat, xform=-1):
        p = self.box_prims(a, b, phase_mat)
        return self._obj(abi.OBJ_MEDIUM, p, 6, xform, density, phase_mat, abi.OBJ_LIST)

    def light(self, obj):
        self.lights.append(obj)

    def light_only(self, obj):
        """Marks obj as sampled by next event estimation but never hit (it
        stays out of the world BVH) and adds it to the lights."""
        self.obj_flags[obj] |= abi.OBJF_LIGHT_ONLY
        self.lights.append(obj)

    # ---------------------------------------------------------------- voxels (voxel.h)
    def vox_model(self, cells, kinds, materials, edge_material=-1, edge_width=0.0):
        """A voxel model: cells a (nz, ny, nx) array of palette indices (x
        fastest in memory), kinds[256] 0 empty / 1 opaque / 2 transparent,
        materials[256] material indices (-1 where no cell uses the entry).
        Returns its index."""
        c = np.ascontiguousarray(np.asarray(cells, dtype=np.uint8))
        assert c.ndim == 3
        m = abi.VoxModel()
        m.nz, m.ny, m.nx = c.shape
        m.cells = sum(x.size for x in self.vox_cells)
        m.edge_material = edge_material
        m.edge_width = edge_width
        for i in range(256):
            m.material[i] = int(materials[i])
            m.kind[i] = int(kinds[i])
        self.vox_cells.append(c.ravel())
        self.vox_models.append(m)
        return len(self.vox_models) - 1

    def voxel_blocks(self, model, block=8):
        """voxel_blocks (voxel.h:341-382): the model cut into blocks of up to
        block^3 cells, each shrunk to the box of its voxels, empty blocks
        skipped, as MPT_PRIM_VOXBLOCK primitives; returns (first, count)."""
        m = self.vox_models[model]
        off = m.cells
        cells = np.concatenate(self.vox_cells)[off:off + m.nx * m.ny * m.nz].reshape(m.nz, m.ny, m.nx)
        kinds = np.array([m.kind[i] for i in range(256)])
        solid = kinds[cells] != 0
        first, count = len(self.prim_kind), 0
        for z0 in range(0, m.nz, block):
            for y0 in range(0, m.ny, block):
                for x0 in range(0, m.nx, block):
                    sub = solid[z0:z0 + block, y0:y0 + block, x0:x0 + block]
                    if not sub.any():
                        continue
                    zz, yy, xx = np.nonzero(sub)
                    lo = [x0 + xx.min(), y0 + yy.min(), z0 + zz.min()]
                    hi = [x0 + xx.max() + 1, y0 + yy.max() + 1, z0 + zz.max() + 1]
                    self._prim(abi.PRIM_VOXBLOCK, -1, [*lo, *hi, model])
                    count += 1
        return first, count

    def voxel(self, first, count, xform=-1):
        return self._obj(abi.OBJ_VOXEL, first, count, xform)

    def environment(self, rgb_image, intensity=1.0, rotation=0.0, as_light=True):
        img = np.ascontiguousarray(rgb_image, dtype=np.float64)
        assert img.ndim == 3 and img.shape[2] == 3
        self.env = img
        self.env_intensity = intensity
        self.env_rotation = rotation
        if as_light:
            self.lights.append(-1)

    # ---------------------------------------------------------------- build
    def build(self):
        """Returns (SceneDesc, keepalive); keepalive owns the arrays."""
        keep = {}

        def arr(name, values, dtype, ctype):
            a = np.ascontiguousarray(np.asarray(values, dtype=dtype))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            keep[name] = a
            return a.ctypes.data_as(C.POINTER(ctype))

        d = abi.SceneDesc()
        d.n_prims = len(self.prim_kind)
        d.prim_kind = arr("pk", self.prim_kind, np.int32, C.c_int32)
        d.prim_material = arr("pm", self.prim_mat, np.int32, C.c_int32)
        d.prim_data = arr("pd", np.concatenate(self.prim_data) if self.prim_data else [], np.float64, C.c_double)
        d.n_objects = len(self.obj_kind)
        d.object_kind = arr("ok", self.obj_kind, np.int32, C.c_int32)
        d.object_first = arr("of", self.obj_first, np.int64, C.c_int64)
        d.object_count = arr("oc", self.obj_count, np.int64, C.c_int64)
        d.object_xform = arr("ox", self.obj_xform, np.int32, C.c_int32)
        ids = np.concatenate([[0], np.cumsum(self.obj_count)[:-1]]) if self.obj_count else []
        d.object_id_base = arr("oi", ids, np.int64, C.c_int64)
        d.object_density = arr("od", self.obj_density, np.float64, C.c_double)
        d.object_material = arr("om", self.obj_mat, np.int32, C.c_int32)
        d.object_boundary_kind = arr("ob", self.obj_bkind, np.int32, C.c_int32)
        if any(self.obj_flags):
            d.object_flags = arr("ofl", self.obj_flags, np.int32, C.c_int32)
        if self.vox_models:
            vm = (abi.VoxModel * len(self.vox_models))(*self.vox_models)
            keep["vm"] = vm
            d.n_vox_models = len(self.vox_models)
            d.vox_models = C.cast(vm, C.POINTER(abi.VoxModel))
            cells = np.concatenate(self.vox_cells)
            d.vox_cells = arr("vc", cells, np.uint8, C.c_uint8)
            d.n_vox_cells = cells.size
        xf = (abi.Xform * max(1, len(self.xforms)))(*self.xforms)
        mats = (abi.Material * max(1, len(self.materials)))(*self.materials)
        texs = (abi.Texture * max(1, len(self.textures)))(*self.textures)
        keep.update(xf=xf, mats=mats, texs=texs)
        d.n_xforms = len(self.xforms)
        d.xforms = C.cast(xf, C.POINTER(abi.Xform))
        d.n_materials = len(self.materials)
        d.materials = C.cast(mats, C.POINTER(abi.Material))
        d.n_textures = len(self.textures)
        d.textures = C.cast(texs, C.POINTER(abi.Texture))
        d.n_lights = len(self.lights)
        d.light_object = arr("lo", self.lights, np.int32, C.c_int32)
        d.light_strategy = self.light_strategy
        d.light_uniform_fraction = self.uniform_fraction
        if self.env is not None:
            keep["env"] = self.env
            d.env_height, d.env_width = self.env.shape[0], self.env.shape[1]
            d.env_rgb = self.env.ctypes.data_as(C.POINTER(C.c_double))
            d.env_intensity = self.env_intensity
            d.env_rotation_degrees = self.env_rotation
        return d, keep


def camera(width, aspect, spp, depth, lookfrom, lookat, vfov, vup=(0, 1, 0), background=None, aperture=0.0,
           focus=10.0):
    c = abi.Camera()
    c.aspect_ratio = aspect
    c.image_width = width
    c.samples_per_pixel = spp
    c.max_depth = depth
    c.vfov = vfov
    for k in range(3):
        c.lookfrom[k], c.lookat[k], c.vup[k] = lookfrom[k], lookat[k], vup[k]
    c.use_sky = 1
    if background is not None:
        c.use_sky = 0
        for k in range(3):
            c.background[k] = background[k]
    c.defocus_angle = aperture
    c.focus_dist = focus
    return c
