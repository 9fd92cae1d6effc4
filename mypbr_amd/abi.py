"""ctypes mirror of the C ABI structs in include/mpt.h.

Kept separate from the loader so that the oracle wrapper in tests/ can
reuse the same struct layouts without loading the GPU library.
"""
import ctypes as C

MPT_OK = 0
MPT_E_INVALID = -1
MPT_E_UNSUPPORTED = -2
MPT_E_HIP = -3
MPT_E_IO = -4
MPT_E_CAPACITY = -5

PRIM_SPHERE, PRIM_QUAD, PRIM_TRIANGLE, PRIM_VOXBLOCK = 0, 1, 2, 3
OBJ_PRIM, OBJ_LIST, OBJ_MESH, OBJ_MEDIUM, OBJ_VOXEL = 0, 1, 2, 3, 4
OBJF_LIGHT_ONLY = 1
VOX_EMPTY, VOX_OPAQUE, VOX_TRANSPARENT = 0, 1, 2
MAT_LAMBERTIAN, MAT_METAL, MAT_DIELECTRIC, MAT_PRINCIPLED, MAT_DIFFUSE_LIGHT, MAT_PHASE = range(6)
TEX_SOLID, TEX_CHECKER, TEX_IMAGE, TEX_SCALED, TEX_CHANNEL, TEX_UV_TRANSFORM, TEX_NORMAL_SCALE = range(7)
TEXF_LINEAR, TEXF_REPEAT, TEXF_HDR, TEXF_NO_ALPHA = 1, 2, 4, 8

D3 = C.c_double * 3


class Texture(C.Structure):
    _fields_ = [("type", C.c_int32), ("flags", C.c_int32), ("inv_scale", C.c_double),
                ("even", D3), ("odd", D3), ("width", C.c_int32), ("height", C.c_int32), ("offset", C.c_int64),
                ("inner", C.c_int32), ("channel", C.c_int32), ("scale", C.c_double), ("uv", C.c_double * 6)]


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("texture", C.c_int32), ("roughness", C.c_double),
                ("alpha", C.c_double), ("ior", C.c_double), ("metallic", C.c_double),
                ("f0", C.c_double), ("g", C.c_double), ("spread", C.c_double),
                ("tan_half", C.c_double), ("albedo", D3), ("f_avg", D3), ("emit", D3),
                ("sigma_a", D3), ("avg_emission", D3), ("normal_map", C.c_int32),
                ("metallic_texture", C.c_int32), ("roughness_texture", C.c_int32), ("uses_uv", C.c_int32),
                ("thin", C.c_int32), ("measured", C.c_int32), ("eta", D3), ("k", D3),
                ("coat", C.c_int32), ("coat_weight_texture", C.c_int32), ("coat_roughness_texture", C.c_int32),
                ("coat_normal_map", C.c_int32), ("glow", C.c_int32), ("glow_texture", C.c_int32),
                ("coat_tint", D3)]


class Xform(C.Structure):
    _fields_ = [("m", C.c_double * 9), ("t", D3), ("inv_m", C.c_double * 9), ("inv_t", D3)]


class VoxModel(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("edge_material", C.c_int32),
                ("cells", C.c_int64), ("edge_width", C.c_double), ("material", C.c_int32 * 256),
                ("kind", C.c_uint8 * 256)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_prims", C.c_int64),
        ("prim_kind", C.POINTER(C.c_int32)),
        ("prim_material", C.POINTER(C.c_int32)),
        ("prim_data", C.POINTER(C.c_double)),
        ("n_objects", C.c_int32),
        ("object_kind", C.POINTER(C.c_int32)),
        ("object_first", C.POINTER(C.c_int64)),
        ("object_count", C.POINTER(C.c_int64)),
        ("object_xform", C.POINTER(C.c_int32)),
        ("object_id_base", C.POINTER(C.c_int64)),
        ("object_density", C.POINTER(C.c_double)),
        ("object_material", C.POINTER(C.c_int32)),
        ("object_boundary_kind", C.POINTER(C.c_int32)),
        ("n_xforms", C.c_int32),
        ("xforms", C.POINTER(Xform)),
        ("n_materials", C.c_int32),
        ("materials", C.POINTER(Material)),
        ("n_textures", C.c_int32),
        ("textures", C.POINTER(Texture)),
        ("n_lights", C.c_int32),
        ("light_object", C.POINTER(C.c_int32)),
        ("light_strategy", C.c_int32),
        ("light_uniform_fraction", C.c_double),
        ("env_width", C.c_int32),
        ("env_height", C.c_int32),
        ("env_rgb", C.POINTER(C.c_double)),
        ("env_intensity", C.c_double),
        ("env_rotation_degrees", C.c_double),
        ("prim_normals", C.POINTER(C.c_double)),
        ("texels", C.POINTER(C.c_uint8)),
        ("n_texel_bytes", C.c_int64),
        ("texels_hdr", C.POINTER(C.c_float)),
        ("n_texels_hdr", C.c_int64),
        ("prim_uvs", C.POINTER(C.c_double)),
        ("has_ground", C.c_int32),
        ("ground_material", C.c_int32),
        ("ground_height", C.c_double),
        ("has_sun", C.c_int32),
        ("sun_pad", C.c_int32),
        ("sun_direction", D3),
        ("sun_angle_degrees", C.c_double),
        ("sun_irradiance", D3),
        ("has_light_bounds", C.c_int32),
        ("light_bounds_pad", C.c_int32),
        ("light_bounds", C.c_double * 6),
        ("n_fogs", C.c_int32),
        ("fog_pad", C.c_int32),
        ("fog_params", C.POINTER(C.c_double)),
        ("fog_material", C.POINTER(C.c_int32)),
        ("object_alpha", C.POINTER(C.c_int32)),
        ("object_flags", C.POINTER(C.c_int32)),
        ("n_vox_models", C.c_int32),
        ("vox_pad", C.c_int32),
        ("vox_models", C.POINTER(VoxModel)),
        ("vox_cells", C.POINTER(C.c_uint8)),
        ("n_vox_cells", C.c_int64),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("aspect_ratio", C.c_double),
        ("image_width", C.c_int32),
        ("samples_per_pixel", C.c_int32),
        ("max_depth", C.c_int32),
        ("max_diffuse", C.c_int32),
        ("max_specular", C.c_int32),
        ("max_scatter", C.c_int32),
        ("transparent_shadows", C.c_int32),
        ("vfov", C.c_double),
        ("lookfrom", D3), ("lookat", D3), ("vup", D3),
        ("use_sky", C.c_int32),
        ("collect_features", C.c_int32),
        ("background", D3),
        ("ortho_height", C.c_double),
        ("defocus_angle", C.c_double),
        ("focus_dist", C.c_double),
        ("blades", C.c_int32),
        ("sun_disk_hidden", C.c_int32),
        ("blade_rotation", C.c_double),
    ]

    def height(self):
        return max(1, int(self.image_width / self.aspect_ratio))


class RenderOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("first_pass", C.c_int32), ("save_every", C.c_int32),
                ("passes_in_flight", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("sort_by_material", C.c_int32), ("use_graph", C.c_int32),
                ("scene_index", C.c_int32), ("pass_count", C.c_int32), ("accumulate", C.c_int32),
                ("profile", C.c_int32), ("tail_paths", C.c_int32), ("guide_passes", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_int64), ("extension_rays", C.c_int64), ("shadow_rays", C.c_int64),
                ("seconds", C.c_double), ("stage_ms", C.c_double * 5), ("stage_launches", C.c_int64 * 5),
                ("node_bytes", C.c_int64 * 2), ("prim_bytes", C.c_int64 * 2),
                ("tail_rays", C.c_int64 * 2)]


PASS_CALLBACK = C.CFUNCTYPE(None, C.c_int32, C.c_void_p)


def camera_copy(cam, **overrides):
    """A copy of a Camera with fields overridden."""
    c = Camera()
    C.pointer(c)[0] = cam
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


class DenoiseOpts(C.Structure):
    """mpt_denoise_opts (denoiser.h:40-44)."""
    _fields_ = [("iterations", C.c_int32), ("pad", C.c_int32), ("sigma_luminance", C.c_double),
                ("sigma_normal", C.c_double), ("sigma_albedo", C.c_double)]


TONEMAPPERS = {"none": 0, "reinhard": 1, "aces": 2}


class Display(C.Structure):
    """mpt_display: display_settings + post_settings (tonemap.h:11-15,
    postprocess.h:12-21)."""
    _fields_ = [("exposure", C.c_double), ("tonemap", C.c_int32), ("pad", C.c_int32), ("bloom", C.c_double),
                ("bloom_threshold", C.c_double), ("bloom_radius", C.c_double), ("vignette", C.c_double)]


class NrcOpts(C.Structure):
    """mpt_nrc_opts."""
    _fields_ = [("query_depth", C.c_int32), ("train_stride", C.c_int32), ("train_steps", C.c_int32),
                ("infer", C.c_int32), ("train", C.c_int32), ("lr", C.c_float), ("box_min", C.c_double * 3),
                ("box_max", C.c_double * 3)]


class NrcStats(C.Structure):
    """mpt_nrc_stats."""
    _fields_ = [("queries", C.c_int64), ("train_samples", C.c_int64), ("loss", C.c_double)]


class AutotuneResult(C.Structure):
    """mpt_autotune_result."""
    _fields_ = [("passes_in_flight", C.c_int32), ("trials", C.c_int32), ("tail_paths", C.c_int64),
                ("seconds_per_pass", C.c_double), ("blocks_per_cu", C.c_int32), ("sort_by_material", C.c_int32),
                ("use_graph", C.c_int32), ("pad", C.c_int32)]
