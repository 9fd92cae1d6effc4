/*
 * mpt.h - C ABI of the MI355X wavefront path-tracing core.
 *
 * This is the drop-in boundary for the reference renderer's hot path,
 * camera::render (include/camera.h:79-105 of adriendjn/mypbr) and the
 * closest-hit query under it (bvh::hit, include/bvh.h:45-86). Everything is
 * plain C: opaque handles, plain pointers and sizes, no C++ or torch types.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative MPT_E_* code on
 *    error; mpt_last_error() then returns a message (thread-local).
 *  - Pointers to outputs are caller-allocated host memory unless the name
 *    says "dev" (device memory the caller owns, for callers that keep data
 *    in HBM).
 *  - Geometry and shading arithmetic is IEEE double, as in the reference
 *    (vec3.h:11); only the traversal bookkeeping uses integers.
 *  - Unsupported objects or materials make mpt_scene_create fail; nothing
 *    falls back to a CPU path.
 */
#ifndef MPT_H
#define MPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPT_ABI_VERSION 1

enum {
    MPT_OK = 0,
    MPT_E_INVALID = -1,     /* bad argument or malformed description */
    MPT_E_UNSUPPORTED = -2, /* object/material the GPU path does not model */
    MPT_E_HIP = -3,         /* a HIP runtime call failed */
    MPT_E_IO = -4,          /* a scene or data file could not be read */
    MPT_E_CAPACITY = -5     /* a queue or table would overflow */
};

/* ---------------------------------------------------------------------- */
/* Flattened scene description (host SoA handed to mpt_scene_create).     */
/* ---------------------------------------------------------------------- */

/* Primitive kinds (geometry a leaf tests). */
enum { MPT_PRIM_SPHERE = 0, MPT_PRIM_QUAD = 1, MPT_PRIM_TRIANGLE = 2,
       MPT_PRIM_VOXBLOCK = 3 /* a block of a voxel model (voxel_grid, voxel.h:168-337): prim_data lo xyz,
                                hi xyz (cells [lo, hi), whole numbers), the model's index, 2 unused */ };

/* Object kinds: one entry of the world list the top-level BVH is built on
 * (scene.h:406-420 adds each scene line as one such object). */
enum {
    MPT_OBJ_PRIM = 0,   /* a single primitive (sphere.h, quad.h) */
    MPT_OBJ_LIST = 1,   /* hittable_list tested linearly (hittable_list.h:25-39), e.g. box() */
    MPT_OBJ_MESH = 2,   /* triangle mesh with its own BVH (mesh.h:175-191) */
    MPT_OBJ_MEDIUM = 3, /* constant_medium over a list/sphere boundary (medium.h:60-88) */
    MPT_OBJ_VOXEL = 4   /* voxel_blocks (voxel.h:341-382): voxel blocks of one model with their own BVH */
};

/* object_flags bits. */
enum {
    MPT_OBJF_LIGHT_ONLY = 1  /* sampled as a light, never hit: not in the world BVH (the emissive faces
                                of a voxel model, which the model's grid is hit for; scene.h:1118-1128) */
};

/* A voxel model (vox_model, vox_loader.h:32-44) and its palette in the
 * scene's materials (voxel_palette, voxel.h:23-31): cells are palette
 * indices, x fastest, then y, then z, 0 = empty; kind[i] is 0 empty, 1
 * opaque, 2 transparent (voxel_kind, voxel.h:16-21); material[i] the
 * material of index i (-1 where no cell uses it). With edge_width > 0 the
 * outlines of its shapes are drawn with edge_material (voxel.h:290-311). */
typedef struct mpt_vox_model {
    int32_t nx, ny, nz;
    int32_t edge_material;         /* -1: no outlines */
    int64_t cells;                 /* offset of cell (0, 0, 0) in vox_cells */
    double edge_width;             /* in voxels, 0 or (0, 0.5) */
    int32_t material[256];
    uint8_t kind[256];
} mpt_vox_model;

/* Material types (material.h, medium.h). */
enum {
    MPT_MAT_LAMBERTIAN = 0,
    MPT_MAT_METAL = 1,
    MPT_MAT_DIELECTRIC = 2,
    MPT_MAT_PRINCIPLED = 3,
    MPT_MAT_DIFFUSE_LIGHT = 4,
    MPT_MAT_PHASE = 5
};

/* Texture types (texture.h:24-105). */
enum { MPT_TEX_SOLID = 0, MPT_TEX_CHECKER = 1, MPT_TEX_IMAGE = 2,
       /* wrappers of another texture (inner), texture.h:212-282: */
       MPT_TEX_SCALED = 3,       /* inner's value times even[3] (scaled_texture) */
       MPT_TEX_CHANNEL = 4,      /* grey of inner's channel `channel` times scale (channel_texture) */
       MPT_TEX_UV_TRANSFORM = 5, /* inner at (u, v) scaled, rotated, offset (uv_transform_texture) */
       MPT_TEX_NORMAL_SCALE = 6  /* inner's encoded normal with its tilt times scale (normal_scale_texture) */ };
/* Image texture flags: 8-bit values are data (k / 255), not sRGB; (u, v)
 * wraps around instead of clamping at the edge; the texels are linear
 * floats (a Radiance .hdr file) instead of bytes. */
/* MPT_TEXF_NO_ALPHA: the file had no alpha channel, so as an alpha mask the
 * texture's coverage is its first byte (the grey level), as
 * image_texture::alpha_value reads it (texture.h:130-136, 166-173). */
enum { MPT_TEXF_LINEAR = 1, MPT_TEXF_REPEAT = 2, MPT_TEXF_HDR = 4, MPT_TEXF_NO_ALPHA = 8 };

typedef struct mpt_texture {
    int32_t type;        /* MPT_TEX_* */
    int32_t flags;       /* image: MPT_TEXF_* */
    double inv_scale;    /* checker: 1 / cube side */
    double even[3];      /* solid: the color; checker: even cells */
    double odd[3];       /* checker: odd cells */
    int32_t width;       /* image, in texels */
    int32_t height;
    int64_t offset;      /* image: first texel in mpt_scene_desc.texels (4 bytes, RGBA, per texel) or
                            .texels_hdr (3 floats per texel), rows top first */
    /* Wrappers: the texture read, and their parameters. A chain is at most a
     * value wrapper (scaled, channel, normal scale) over uv transforms over a
     * plain texture, as glTF's texture views give it. */
    int32_t inner;
    int32_t channel;     /* channel: 0 red, 1 green, 2 blue */
    double scale;        /* channel, normal scale */
    double uv[6];        /* uv transform: offset u, offset v, scale u, scale v, cos and sin of the rotation */
} mpt_texture;

typedef struct mpt_material {
    int32_t type;        /* MPT_MAT_* */
    int32_t texture;     /* albedo / base color texture index */
    double roughness;    /* metal, dielectric, principled (clamped to [0,1]) */
    double alpha;        /* roughness_to_alpha(roughness), microfacet.h:17-19 */
    double ior;          /* dielectric index; principled ior */
    double metallic;     /* principled */
    double f0;           /* principled: dielectric f0 = ((ior-1)/(ior+1))^2 */
    double g;            /* phase: Henyey-Greenstein anisotropy */
    double spread;       /* diffuse_light: grid spread in degrees (180 = bare) */
    double tan_half;     /* diffuse_light: tan(spread/2) */
    double albedo[3];    /* metal albedo (Schlick f0), phase albedo */
    double f_avg[3];     /* metal: Schlick F_avg = f0 + (1-f0)/21 */
    double emit[3];      /* diffuse_light radiance */
    double sigma_a[3];   /* dielectric absorption (Beer-Lambert) */
    double avg_emission[3]; /* diffuse_light: emit * directional share */
    int32_t normal_map;        /* tangent-space normal map texture, -1 = none (normal_mapped, material.h:631-674;
                                  principled's own, material.h:389) */
    int32_t metallic_texture;  /* principled: metallic from this texture's red channel, -1 = the constant */
    int32_t roughness_texture; /* principled: roughness from this texture's red channel, -1 = the constant */
    int32_t uses_uv;           /* 1 when a texture above or the base texture is an image: hits need (u, v) */
    int32_t thin;              /* dielectric: a thin sheet (thin_dielectric, material.h:683-707): reflects
                                  2R / (1 + R) of the light, lets the rest through unbent, tinted by albedo */
    int32_t measured;          /* metal, principled: Fresnel of the complex index eta + i k (fresnel_conductor,
                                  microfacet.h:57-66) instead of Schlick's; albedo (metal) and f_avg hold
                                  its head-on value and hemispherical average (microfacet.h:109-121) */
    double eta[3];             /* measured conductor: real and imaginary index of refraction, RGB */
    double k[3];
    int32_t coat;              /* principled: 1 = a clear coat on top (set_clearcoat, material.h:326-345, 363-379):
                                  a colorless GGX layer of ior 1.5 that passes the rest to the base, twice */
    int32_t coat_weight_texture;     /* its weight: red channel, clamped to [0, 1] (a solid texture for a constant) */
    int32_t coat_roughness_texture;  /* its roughness: red channel, clamped to [0, 1] */
    int32_t coat_normal_map;         /* its own tangent-space normal map, -1: the unperturbed normal */
    /* emissive_surface (material.h:714-741): with glow set, the front face
     * glows with glow_texture's value while the material scatters as it is;
     * not a light. */
    int32_t glow;
    int32_t glow_texture;
    double coat_tint[3];       /* what the coat lets through head-on, in and out (set_clearcoat_color); 1: clear */
} mpt_material;

/* An affine transform with its stored inverse, as transform.h:120-123:
 * m (row-major 3x3), t, inv_m, inv_t. */
typedef struct mpt_xform {
    double m[9];
    double t[3];
    double inv_m[9];
    double inv_t[3];
} mpt_xform;

typedef struct mpt_scene_desc {
    /* Primitives. prim_data holds 9 doubles per primitive:
     *   sphere:   center xyz, radius, 5 unused
     *   quad:     Q xyz, u xyz, v xyz
     *   triangle: p0 xyz, p1 xyz, p2 xyz                                  */
    int64_t n_prims;
    const int32_t* prim_kind;      /* [n_prims] MPT_PRIM_* */
    const int32_t* prim_material;  /* [n_prims] material index (-1 = none) */
    const double* prim_data;       /* [n_prims * 9] */

    /* World objects, in world-list order. An object covers the primitive
     * range [first, first + count); several objects may share a range
     * (instances of one mesh). Canonical primitive IDs are id_base + k. */
    int32_t n_objects;
    const int32_t* object_kind;    /* [n_objects] MPT_OBJ_* */
    const int64_t* object_first;   /* [n_objects] */
    const int64_t* object_count;   /* [n_objects] */
    const int32_t* object_xform;   /* [n_objects] -1 or index into xforms */
    const int64_t* object_id_base; /* [n_objects] first canonical ID */
    const double* object_density;  /* [n_objects] medium density (else 0) */
    const int32_t* object_material;/* [n_objects] medium phase material (else -1) */
    const int32_t* object_boundary_kind; /* [n_objects] medium boundary: MPT_OBJ_PRIM or MPT_OBJ_LIST */

    int32_t n_xforms;
    const mpt_xform* xforms;

    int32_t n_materials;
    const mpt_material* materials;
    int32_t n_textures;
    const mpt_texture* textures;

    /* Lights for next event estimation, in light-list order
     * (light_list.h:33-134): an object index, -1 for the environment or -2
     * for the sun. */
    int32_t n_lights;
    const int32_t* light_object;
    int32_t light_strategy;        /* 0 = uniform, 1 = power (light_list.h:35-38), 2 = a plain
                                      hittable_list: pick min(int(xi n), n-1) (hittable_list.h:58-62) */
    double light_uniform_fraction; /* 0.2 by default */

    /* Environment (environment.h): an equirectangular RGB image, row 0 at +y,
     * or width = 0 for none. */
    int32_t env_width, env_height;
    const double* env_rgb;         /* [env_width * env_height * 3], linear RGB */
    double env_intensity;
    double env_rotation_degrees;

    /* Per-vertex shading normals of triangles (mesh_data::normals, mesh.h:14-19):
     * [n_prims * 9] as n0 xyz, n1 xyz, n2 xyz, or NULL for none. A triangle
     * whose first value is NaN has none (flat shading); values of other
     * primitive kinds are ignored. Interpolated at the hit with the
     * barycentrics, normalised and turned to the geometric normal's side
     * (mesh.h:93-101). */
    const double* prim_normals;

    /* Image texture storage: RGBA bytes and RGB floats, addressed by
     * mpt_texture.offset. May be NULL when no texture uses them. */
    const uint8_t* texels;
    int64_t n_texel_bytes;
    const float* texels_hdr;
    int64_t n_texels_hdr;
    /* Optional per-triangle texture coordinates, 6 per primitive (u, v at
     * p0, p1, p2); NaN marks a triangle without them, whose (u, v) are its
     * barycentrics b1, b2 (mesh.h:75-81). NULL: none. */
    const double* prim_uvs;

    /* With has_ground set, the infinite ground y = ground_height of material
     * ground_material (ground_plane, plane.h:10-42), tested beside the BVH
     * (main.cpp:70-74). A query hit on it reports the canonical ID one past
     * the last primitive. */
    int32_t has_ground;
    int32_t ground_material;
    double ground_height;

    /* With has_sun set, a distant light of small angular size (sun_light,
     * sun.h:23-80): radiance irradiance / omega inside the cone of half-angle
     * sun_angle_degrees / 2 around sun_direction, added to every escaping ray
     * (camera.h:526-527) and sampled uniformly in the cone by next event
     * estimation when a light_object entry is -2. */
    int32_t has_sun;
    int32_t sun_pad;
    double sun_direction[3];
    double sun_angle_degrees;      /* angular diameter, clamped below at 1e-3 (sun.h:29) */
    double sun_irradiance[3];

    /* With has_light_bounds set, the scene bounds light selection takes its
     * radius from (light_list::build(scene_bounds), light_list.h:45-58),
     * as (xmin, xmax, ymin, ymax, zmin, zmax); otherwise the world list's
     * bounding box, as scene.h:1248 passes it. */
    int32_t has_light_bounds;
    int32_t light_bounds_pad;
    double light_bounds[6];

    /* Height fog (height_fog, medium.h:98-155): n_fogs unbounded media,
     * tested beside the BVH after it and before the ground plane, in this
     * order (scene.h:656-666, 1237-1238). fog_params holds 3 doubles per fog:
     * density at y = base, scale height (the density falls by e per scale
     * height up), base; fog_material the phase (volume) material. */
    int32_t n_fogs;
    int32_t fog_pad;
    const double* fog_params;      /* [n_fogs * 3] */
    const int32_t* fog_material;   /* [n_fogs] */

    /* Alpha masks (mesh.h:24-32, 82-86): per object, an image texture whose
     * coverage (alpha byte / 255, nearest texel at the hit's (u, v)) cuts
     * holes in a mesh's triangles: a hit with coverage c < 1 is kept when a
     * uniform draw falls below c. -1: opaque. NULL: no object has one. Only
     * mesh objects may have one. */
    const int32_t* object_alpha;   /* [n_objects] */

    /* Per object, MPT_OBJF_* bits; NULL: all 0. */
    const int32_t* object_flags;   /* [n_objects] */

    /* Voxel models, read by MPT_PRIM_VOXBLOCK primitives and MPT_OBJ_VOXEL
     * objects (a block's model index names one). The grids are walked only by
     * the kernels that carry (u, v) and the ground plane. */
    int32_t n_vox_models;
    int32_t vox_pad;
    const mpt_vox_model* vox_models;  /* [n_vox_models] */
    const uint8_t* vox_cells;         /* every model's cells */
    int64_t n_vox_cells;
} mpt_scene_desc;

/* Camera and render settings: the public fields of the reference camera
 * (camera.h:21-56). */
typedef struct mpt_camera {
    double aspect_ratio;
    int32_t image_width;
    int32_t samples_per_pixel;
    int32_t max_depth;
    int32_t max_diffuse, max_specular, max_scatter;
    int32_t transparent_shadows;
    double vfov;
    double lookfrom[3], lookat[3], vup[3];
    int32_t use_sky;               /* gradient sky when no environment and use_sky */
    int32_t collect_features;      /* fill albedo/normal/variance */
    double background[3];
    double ortho_height;
    double defocus_angle;
    double focus_dist;
    int32_t blades;
    int32_t sun_disk_hidden;       /* !sun_disk_visible (camera.h:46): camera rays do not see the
                                      sun itself; zero keeps the reference default */
    double blade_rotation;
} mpt_camera;

/* Knobs of the GPU path that do not change the estimate. Zero means
 * "choose" (autotuned or default). */
typedef struct mpt_render_opts {
    uint64_t seed;                 /* counter-based RNG seed */
    int32_t first_pass;            /* pass index of the first sample rendered (resume) */
    int32_t save_every;            /* call on_pass every this many passes (0 = never) */
    int32_t passes_in_flight;      /* samples per pixel traced together (0 = auto) */
    int32_t blocks_per_cu;         /* 256-thread blocks per CU in the persistent stage grid, 1..64
                                    * (0 = auto: the tuned choice, else 16; 4, one resident
                                    * round, for scenes on the general-geometry kernels and for
                                    * mpt_render_batch; one resident round for scenes with quads,
                                    * lists or instances and no meshes or media) */
    int32_t sort_by_material;      /* 1 = on: each bounce is a wavefront of stage kernels, hits binned
                                    * into one queue per material and shaded by that material's
                                    * kernel; -1 = off: each batch runs in one kernel, a path per
                                    * thread through all its bounces, shading whatever it hits;
                                    * 0 = auto (the tuned choice, else on) */
    int32_t use_graph;             /* 1 = on: each bounce's kernels are captured once as a hipGraph
                                    * and replayed (per bounce index, while the camera and buffers
                                    * stay the same); -1 = off; 0 = auto (the tuned choice, else
                                    * off). Profiling and the radiance cache launch directly. */
    int32_t scene_index;           /* keys the RNG for batched multi-scene launches */
    int32_t pass_count;            /* passes to render in this call (0 = samples_per_pixel - first_pass) */
    int32_t accumulate;            /* 1: add to the previous call's sums (progressive render) */
    int32_t profile;               /* bit 0: time each stage with HIP events; bit 1: count BVH work */
    int32_t tail_paths;            /* once at most this many paths of a batch are left, the rest of the
                                    * batch runs in one launch, a path per thread (0 = auto, -1 = never) */
    int32_t guide_passes;          /* with collect_features: trace the albedo and normal guides
                                    * (camera::first_surface) only for passes with index below this, and
                                    * average them over those passes; the variance guide still takes
                                    * every pass. 0 = every pass, the reference's frame (camera.h:230-244).
                                    * For denoising inputs, whose guides converge long before the
                                    * radiance does. The radiance is unchanged bit for bit. */
} mpt_render_opts;

/* Per-render statistics. */
typedef struct mpt_render_stats {
    int64_t paths;                 /* pixels x passes rendered */
    int64_t extension_rays;        /* extension rays traced */
    int64_t shadow_rays;           /* shadow rays traced */
    double seconds;                /* wall time of the GPU part */
    /* with profile bit 0: summed kernel time and launches per stage,
     * [0] extend, [1] shade, [2] shadow, [3] camera + features + accumulate,
     * [4] the batch tails run as one launch each */
    double stage_ms[5];
    int64_t stage_launches[5];
    /* with profile bit 1: BVH node bytes and primitive bytes read by the
     * extension [0] and shadow [1] traversals (a root box counts 64 bytes,
     * a four-wide record 144, a slot re-tested on pop 32, a leaf's
     * double-precision box checked at its first hit 48) */
    int64_t node_bytes[2];
    int64_t prim_bytes[2];
    /* extension [0] and shadow [1] rays traced by the batch tails (included
     * in extension_rays and shadow_rays; their BVH work is not counted) */
    int64_t tail_rays[2];
} mpt_render_stats;

typedef struct mpt_scene mpt_scene;

/* Progress callback, invoked on the calling thread after every save_every
 * passes with the number of passes done (camera.h:99-101). The averaged
 * frame so far can be read with mpt_read_frame inside the callback. */
typedef void (*mpt_pass_callback)(int32_t done, void* user);

const char* mpt_last_error(void);
int32_t mpt_abi_version(void);

/* Device information for logs: name, CU count and HBM bytes. */
int32_t mpt_device_info(char* name, int32_t name_len, int32_t* cu_count, int64_t* hbm_bytes);

/* Device memory for callers that hand the _dev entry points their own
 * buffers (no reference counterpart: the reference is CPU-only). alloc
 * returns NULL on failure (mpt_last_error says why). */
void* mpt_device_alloc(size_t bytes);
void mpt_device_free(void* ptr);
int32_t mpt_copy_to_device(void* dst, const void* src, size_t bytes);
int32_t mpt_copy_to_host(void* dst, const void* src, size_t bytes);
int32_t mpt_device_synchronize(void);

/* Copies the description to HBM and builds BVHs, light CDFs and tables.
 * Replaces: bvh(list) construction (bvh.h:22-43, main.cpp:70), mesh's inner
 * BVH (mesh.h:177-179), light_list::build (light_list.h:45-96) and the
 * ggx_albedo_table (microfacet.h:265-338). */
int32_t mpt_scene_create(const mpt_scene_desc* desc, mpt_scene** out);

/* Loads a .scene file (the reference format, scene.h:42-243, for the subset
 * listed in the library's docs) and creates the scene; fills cam with the
 * file's camera settings when cam is not NULL. Replaces load_scene
 * (scene.h:244) followed by the BVH build (main.cpp:62-74). */
int32_t mpt_scene_load(const char* path, mpt_scene** out, mpt_camera* cam);

void mpt_scene_destroy(mpt_scene* scene);

/* Parses a .scene file and flattens it into *desc without touching the GPU
 * (load_scene, scene.h:244, then the world/lights flattening camera::render
 * does). The arrays *desc points to stay valid until mpt_desc_free(*holder).
 * cam, when not NULL, receives the file's camera settings. */
int32_t mpt_scene_describe(const char* path, mpt_scene_desc* desc, mpt_camera* cam, void** holder);
void mpt_desc_free(void* holder);

/* Counts of the created scene (for logs and tests). */
int32_t mpt_scene_counts(const mpt_scene* scene, int64_t* n_prims, int32_t* n_objects,
                         int64_t* n_tlas_nodes, int64_t* n_blas_nodes, int32_t* n_lights);

/* Closest-hit ray query (bvh::hit over the world, bvh.h:45-86), one result
 * per ray: the canonical primitive ID (-1 on a miss), t and front_face.
 * rays: n x 6 doubles (origin, direction, direction need not be unit);
 * tmin_tmax: n x 2 doubles (the interval ray_t). Media are not tested. */
int32_t mpt_trace_closest(mpt_scene* scene, const double* rays, const double* tmin_tmax, int64_t n,
                          int64_t* prim_id, double* t, uint8_t* front_face);

/* The same query on device-resident buffers (no host copies). */
int32_t mpt_trace_closest_dev(mpt_scene* scene, const double* rays_dev, const double* tmin_tmax_dev,
                              int64_t n, int64_t* prim_id_dev, double* t_dev, uint8_t* front_face_dev,
                              void* hip_stream);

/* The closest-hit query of mpt_trace_closest with grazing ties listed
 * separately. The reference's hit_record names no primitive, so its tests
 * identify a hit by testing each primitive alone (hittable::hit,
 * hittable.h:34-38, with the ray moved into instance space as
 * instance::hit does, instance.h:29-42) over the ray's whole interval;
 * primitives whose t is bitwise equal to the closest hit's are the
 * candidates. ties[i]: their number (0 on a miss, 1 for a unique hit, 2 or
 * more for a tie); tie_min[i] / tie_max[i]: their smallest and largest
 * canonical ID (-1 on a miss). prim_id is the primitive the BVH walk took,
 * one of the candidates. Media are not tested. Every primitive is tested
 * for every ray that hits: O(rays x primitives), for query batches. */
int32_t mpt_trace_ties(mpt_scene* scene, const double* rays, const double* tmin_tmax, int64_t n,
                       int64_t* prim_id, double* t, uint8_t* front_face, int32_t* ties, int64_t* tie_min,
                       int64_t* tie_max);

/* camera::render (camera.h:79-105): renders cam.samples_per_pixel passes
 * and writes the averaged frame (denoiser.h:14-20), row-major from the top
 * left, height = max(1, int(width / aspect)). radiance/albedo/normal are
 * width*height*3 doubles, variance width*height doubles; albedo, normal and
 * variance may be NULL and are only filled with cam->collect_features. */
int32_t mpt_render(mpt_scene* scene, const mpt_camera* cam, const mpt_render_opts* opts,
                   double* radiance, double* albedo, double* normal, double* variance,
                   mpt_pass_callback on_pass, void* user, mpt_render_stats* stats);

/* Several scenes rendered together: every stage launch (camera, extend,
 * shade, shadow, tail, accumulate) serves the paths of all n scenes (1 to
 * 16), scene k with cameras[k]. It replaces rendering the scenes one after
 * another with camera::render (camera.h:79-105, whose tile loop over one
 * image, camera.h:92-138, becomes one pool of paths over all the images).
 * Scene k's paths use scene index opts->scene_index + k, so its frame is the
 * frame mpt_render of that scene alone gives with that scene_index; read it
 * with mpt_read_frame(scenes[k], ...). Every scene renders the same number
 * of passes. Feature guides, save_every, the one-kernel schedule and scenes
 * with image textures, maps, thin glass, measured metals or the ground plane
 * are not supported here (MPT_E_UNSUPPORTED); the scenes must share their
 * kind of lights. stats covers the whole batch. */
int32_t mpt_render_batch(mpt_scene* const* scenes, const mpt_camera* cameras, int32_t n,
                         const mpt_render_opts* opts, mpt_render_stats* stats);

/* What mpt_autotune chose. */
typedef struct mpt_autotune_result {
    int32_t passes_in_flight;  /* samples per pixel traced together */
    int32_t trials;            /* timed trial renders */
    int64_t tail_paths;        /* live paths at which a batch's tail runs in one launch (-1: never) */
    double seconds_per_pass;   /* of the best trial */
    int32_t blocks_per_cu;     /* persistent grid: 256-thread blocks per CU */
    int32_t sort_by_material;  /* 1 = binned wavefront stages, -1 = one kernel per batch */
    int32_t use_graph;         /* 1 = bounces replayed as hipGraphs, -1 = launched directly */
    int32_t pad;
} mpt_autotune_result;

/* Times trial renders of trial_passes passes (8 if < 1) of this camera's
 * image: every combination of batch size (about 4, 8, 16 and 32 M paths in
 * flight) and batch-tail size, then, from the best, 8, 16 and 32 blocks per
 * CU, sorting by material off, and bounce graphs on. Keeps the fastest for
 * the scene (used by later renders of the same image size whose options
 * leave them at 0 = auto) and reports it. The image does not depend on any
 * of these settings. */
int32_t mpt_autotune(mpt_scene* scene, const mpt_camera* cam, int32_t trial_passes, mpt_autotune_result* out);

/* Installs settings found earlier (for instance read back from a cache of
 * mpt_autotune results) for renders of this camera's image size, without
 * timing anything. */
int32_t mpt_autotune_apply(mpt_scene* scene, const mpt_camera* cam, const mpt_autotune_result* settings);

/* A key for caching autotune results: the device, the scene's compiled
 * kernel profile (geometry, material types, lights, environment) and the
 * image size, as text in buf (NUL-terminated, truncated to len). */
int32_t mpt_autotune_key(const mpt_scene* scene, const mpt_camera* cam, char* buf, int32_t len);

/* Inside an on_pass callback: the frame averaged over the passes done. */
int32_t mpt_read_frame(mpt_scene* scene, double* radiance, double* albedo, double* normal,
                       double* variance);

/* mpt_read_frame into device buffers (any may be NULL): the averaged frame
 * of the last render computed on the GPU from its running sums, for callers
 * that keep render -> guides -> denoiser on the device. Guides only when the
 * last render collected them. stream may be NULL (then it synchronises). */
int32_t mpt_read_frame_dev(mpt_scene* scene, double* radiance_dev, double* albedo_dev, double* normal_dev,
                           double* variance_dev, void* hip_stream);

/* Image height camera::initialize computes (camera.h:143-144). */
int32_t mpt_image_height(const mpt_camera* cam);

/* The 32x32 E(mu, alpha) table and its 32 E_avg values the scene uses for
 * Kulla-Conty compensation (microfacet.h:265-338). */
int32_t mpt_ggx_albedo_table(double* e_1024, double* e_avg_32);

/* Light selection probabilities light_list::build assigned (light_list.h:45-96). */
int32_t mpt_scene_light_probabilities(const mpt_scene* scene, double* probabilities);

/* Host-only forms of the same computations on a description (no GPU): the
 * bounding box of the world list, as the BVH over it holds it
 * (bvh::bounding_box, bvh.h:88: the union of the objects' padded boxes,
 * aabb.h:83-88; the ground plane, beside the BVH, is not in it), as
 * (xmin, xmax, ymin, ymax, zmin, zmax); each light's power (hittable::power,
 * hittable.h:59; environment_light::power and sun_light::power with the
 * scene radius, environment.h:116-127, sun.h:66-71) before unknown powers
 * get the mean; and the selection probabilities light_list::build gives
 * (light_list.h:45-96), from scene_bounds (NULL: the description's own
 * light bounds or world box). powers and probabilities hold n_lights. */
int32_t mpt_desc_bounds(const mpt_scene_desc* desc, double* box6);
int32_t mpt_desc_light_powers(const mpt_scene_desc* desc, const double* scene_bounds6, double* powers);
int32_t mpt_desc_light_probabilities(const mpt_scene_desc* desc, const double* scene_bounds6, double* probabilities);

/* BSDF evaluation on the GPU for fixed directions (test hook for the
 * material::eval / pdf value pairs, material.h:22-72): for each of n queries
 * with material index mats[i], unit shading normal n, incoming ray direction
 * r_in and outgoing direction wi (9 doubles each), front_face flags, writes
 * eval (3 doubles) and the scatter pdf's value (1 double). */
int32_t mpt_bsdf_eval(mpt_scene* scene, const int32_t* mats, const double* n_rin_wi,
                      const uint8_t* front_face, int64_t n, double* eval_rgb, double* pdf);

/* Depth of the first surface seen through every pixel centre (no jitter or
 * defocus): t * |d| of the closest hit, 0 for the sky; h*w doubles. A guide
 * for the learned denoiser; the reference's frame has no depth
 * (denoiser.h:14-20). */
int32_t mpt_primary_depth(mpt_scene* scene, const mpt_camera* cam, double* depth);

/* mpt_primary_depth into a device buffer (h*w doubles), traced on the GPU
 * from the pixel centres; stream may be NULL (then it synchronises). */
int32_t mpt_primary_depth_dev(mpt_scene* scene, const mpt_camera* cam, double* depth_dev, void* hip_stream);

/* ---- classical denoiser: denoiser::run (denoiser.h:45-193) on the GPU ---- */

/* The denoiser's settings (denoiser.h:40-44). */
typedef struct mpt_denoise_opts {
    int32_t iterations;            /* a-trous passes, taps 2^it pixels apart (5) */
    int32_t pad;
    double sigma_luminance;        /* in standard deviations of the noise (4) */
    double sigma_normal;           /* (0.1) */
    double sigma_albedo;           /* (0.1) */
} mpt_denoise_opts;

/* Fills the reference's defaults. */
void mpt_denoise_defaults(mpt_denoise_opts* opts);

/* denoiser::run(frame, width, height): filters the frame's radiance guided by
 * its albedo, normal and variance (interleaved RGB doubles and one double per
 * pixel, row-major from the top-left, caller-owned host memory) into out.
 * opts NULL = the defaults. */
int32_t mpt_denoise(const double* radiance, const double* albedo, const double* normal, const double* variance,
                    int32_t width, int32_t height, const mpt_denoise_opts* opts, double* out);

/* The same on device buffers, enqueued on a HIP stream (NULL = the null
 * stream) without synchronising; workspace is device memory of
 * mpt_denoise_workspace_bytes(width, height) bytes. */
size_t mpt_denoise_workspace_bytes(int32_t width, int32_t height);
int32_t mpt_denoise_dev(const double* radiance, const double* albedo, const double* normal, const double* variance,
                        int32_t width, int32_t height, const mpt_denoise_opts* opts, double* out, void* workspace,
                        void* stream);

/* ---- display pipeline: postprocess.h, tonemap.h, color3.h, image_io.h ----
 * How radiance becomes a display image: bloom and vignetting
 * (postprocess(), postprocess.h:108-153), exposure and a tonemapper
 * (display_settings::apply, tonemap.h:28-37), the sRGB curve and byte
 * quantisation (to_byte, color3.h:30-33), and the writers of image_io.h
 * (.png/.ppm display images, .hdr/.pfm raw radiance). */

enum {
    MPT_TONEMAP_NONE = 0,     /* clamp (tonemap.h:10) */
    MPT_TONEMAP_REINHARD = 1, /* x / (1 + x) per channel */
    MPT_TONEMAP_ACES = 2      /* Hill's fit of the ACES RRT + ODT */
};

/* display_settings + post_settings (tonemap.h:11-15, postprocess.h:12-21). */
typedef struct mpt_display {
    double exposure;        /* stops: display images are scaled by 2^exposure (0) */
    int32_t tonemap;        /* MPT_TONEMAP_* (NONE) */
    int32_t pad;
    double bloom;           /* glow strength, 0 = off (0) */
    double bloom_threshold; /* luminance above which pixels glow (1) */
    double bloom_radius;    /* reach of the glow as a fraction of the width (0.04) */
    double vignette;        /* corner darkening, 0 = off (0) */
} mpt_display;

/* Fills the reference's defaults. */
void mpt_display_defaults(mpt_display* display);

/* The display settings a .scene file sets (its bloom, vignette, exposure
 * and tonemap lines, scene.h:667-688); the defaults where it sets none. */
int32_t mpt_scene_display(const char* path, mpt_display* display);

/* find_conductor (microfacet.h:76-96): the RGB complex index of refraction
 * (eta + i k) of a named metal (Al, Ti, Cu, Au, Ag, Cr, or their long names)
 * for mpt_material.measured; MPT_E_INVALID for an unknown name. */
int32_t mpt_find_conductor(const char* name, double* eta, double* k);

/* postprocess(pixels, width, height, post) (postprocess.h:108-153): bloom
 * and vignetting on linear radiance (interleaved RGB doubles, row-major
 * from the top-left, caller-owned host memory) into out. */
int32_t mpt_postprocess(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                        double* out);

/* The bytes write_png and write_ppm emit (image_io.h:16-37): with post set
 * and bloom or vignette on, the postprocessed image, then per channel
 * to_byte(display.apply(pixel)); out holds width * height * 3 bytes. */
int32_t mpt_display_bytes(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                          int32_t post, uint8_t* rgb);

/* The same two on device buffers, enqueued on a HIP stream (NULL = the null
 * stream) without synchronising; workspace is device memory of
 * mpt_display_workspace_bytes(width, height, display) bytes (0 when bloom
 * is off; then workspace may be NULL). */
size_t mpt_display_workspace_bytes(int32_t width, int32_t height, const mpt_display* display);
int32_t mpt_postprocess_dev(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                            double* out, void* workspace, void* stream);
int32_t mpt_display_bytes_dev(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                              int32_t post, uint8_t* rgb, void* workspace, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mpt_display_bytes_dev with the bloom pyramid in a fixed
 * array inside the library instead of a workspace argument. It calls no
 * hipMalloc, hipFree or allocating library routine. The array holds
 * mpt_display_noalloc_max_workspace() bytes (48 MiB: enough for 1920 x 1080
 * at the default bloom radius); a pyramid larger than that is refused with
 * MPT_E_CAPACITY. Each device has its own copy, held until the process
 * ends. Every call on a device shares that device's array: calls that
 * overlap on one device (on two streams, or from two threads) race, so
 * serialise them. Results are identical to mpt_display_bytes_dev's. */
size_t mpt_display_noalloc_max_workspace(void);
int32_t mpt_display_bytes_dev_noalloc(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                                      int32_t post, uint8_t* rgb, void* stream);

/* write_image(path, width, height, pixels, display) (image_io.h:81-106): the
 * format follows the extension. .png and .ppm are display images (bloom,
 * vignette, exposure, tonemap, sRGB bytes from the GPU); .hdr (Radiance
 * RGBE) and .pfm (little-endian floats, bottom row first) keep the raw
 * radiance. NULL path: PPM to stdout. */
int32_t mpt_write_image(const char* path, int32_t width, int32_t height, const double* radiance,
                        const mpt_display* display);

/* ---- learned denoiser: kernel-predicting CNN on BF16 matrix cores ----
 * No counterpart in the reference (it has only the classical filter above);
 * the learned mode the path tracer is compared against. */

/* C = A B^T with BF16 operands and FP32 accumulation on the matrix cores:
 * C[m][n] = (accumulate ? C[m][n] : 0) + sum_k A(m, k) B(n, k), where
 * A(m, k) = A[m * lda + k] (a_t = 0) or A[k * lda + m] (a_t = 1), and B
 * likewise. Device pointers; enqueued on `stream` (NULL = null stream). */
int32_t mpt_gemm_bf16(int32_t a_t, int32_t b_t, int32_t M, int32_t N, int32_t K, const uint16_t* A, int64_t lda,
                      const uint16_t* B, int64_t ldb, float* C, int64_t ldc, int32_t accumulate, void* stream);

typedef struct mpt_kpcnn mpt_kpcnn;

/* A network of `layers` 3x3 convolutions with `hidden` channels (a multiple
 * of 16), He-initialised from `seed`. Parameters are FP32: per layer the
 * weights [cout][9][cin] (tap = 3 * (dy + 1) + dx + 1) then the biases
 * [cout]; mpt_kpcnn_layout gives each layer's sizes and offsets. */
int32_t mpt_kpcnn_create(int32_t layers, int32_t hidden, uint64_t seed, mpt_kpcnn** out);
void mpt_kpcnn_destroy(mpt_kpcnn* net);
int64_t mpt_kpcnn_param_count(const mpt_kpcnn* net);
int32_t mpt_kpcnn_layout(const mpt_kpcnn* net, int32_t* cin, int32_t* cout, int64_t* weight_offset,
                         int64_t* bias_offset);
int32_t mpt_kpcnn_get_params(mpt_kpcnn* net, float* params, float* grads);
/* Training loss: 0 relMSE (o - t)^2 / (t^2 + 0.01) (the default); 1 the L1
 * distance of log(1 + x), which does not let a few dark pixels next to
 * bright ones dominate the gradient. */
int32_t mpt_kpcnn_set_loss(mpt_kpcnn* net, int32_t mode);
int32_t mpt_kpcnn_set_params(mpt_kpcnn* net, const float* params);

/* Runs the network on B frames of H x W (device arrays, double: radiance,
 * albedo, normal with 3 values per pixel, variance and depth with 1). out
 * (device, FP32, 3 per pixel, or NULL) receives the denoised radiance. With
 * a target (device, FP32 radiance, 3 per pixel) it also computes the relMSE
 * loss (o - t)^2 / (t^2 + 0.01) averaged over pixels and channels, written
 * to *loss (host), and the parameter gradients; lr > 0 then takes one Adam
 * step. */
int32_t mpt_kpcnn_run(mpt_kpcnn* net, const double* radiance, const double* albedo, const double* normal,
                      const double* variance, const double* depth, int32_t B, int32_t H, int32_t W,
                      const float* target, float lr, float* out, double* loss);

/* ---- neural radiance cache: hash-grid encoding + MLP on BF16 matrix cores ----
 * No counterpart in the reference; the learned mode of the path tracer
 * (the plain tracer, mpt_render, stays the fallback and the judge). */

typedef struct mpt_nrc mpt_nrc;

/* How a render uses the cache. */
typedef struct mpt_nrc_opts {
    int32_t query_depth;   /* paths are cut at the hit of this bounce (2) */
    int32_t train_stride;  /* one path in this many (hash-chosen) is traced to the end to train (256) */
    int32_t train_steps;   /* Adam steps per batch on its training paths (4) */
    int32_t infer;         /* 1: cut paths and add beta * cache; 0: trace every path (train only) */
    int32_t train;         /* 1: train after each batch */
    float lr;              /* Adam learning rate (1e-3) */
    double box_min[3];     /* the region the hash grid covers (all zero: the scene's bounds) */
    double box_max[3];
} mpt_nrc_opts;

typedef struct mpt_nrc_stats {
    int64_t queries;       /* paths that reached the query bounce */
    int64_t train_samples; /* training samples used (summed over batches) */
    double loss;           /* relative L2 loss of the last training step */
} mpt_nrc_stats;

void mpt_nrc_defaults(mpt_nrc_opts* opts);

/* A cache with its hash tables and MLP initialised from seed. */
int32_t mpt_nrc_create(uint64_t seed, mpt_nrc** out);
void mpt_nrc_destroy(mpt_nrc* cache);
/* MLP parameter count; *table_count receives the hash tables' (16 levels x 2^19 x 2). */
int64_t mpt_nrc_param_count(const mpt_nrc* cache, int64_t* table_count);
/* Grid resolution and dense-indexing flag of each of the 16 levels. */
int32_t mpt_nrc_levels(const mpt_nrc* cache, uint32_t* resolution, uint32_t* dense);
int32_t mpt_nrc_get_params(mpt_nrc* cache, float* mlp, float* mlp_grad, float* table, float* table_grad);
int32_t mpt_nrc_set_params(mpt_nrc* cache, const float* mlp, const float* table);
/* Predictions (3 per sample) for n samples: pos (3 per sample, in [0, 1]^3)
 * and other inputs (10: normal, direction towards the viewer, roughness,
 * albedo); device arrays. */
int32_t mpt_nrc_infer_dev(mpt_nrc* cache, const float* pos, const float* other, int64_t n, float* pred);
/* One training step towards target radiance (3 per sample): loss (host),
 * gradients, and with lr > 0 an Adam step. */
int32_t mpt_nrc_train_dev(mpt_nrc* cache, const float* pos, const float* other, const float* target, int64_t n,
                          float lr, double* loss);

/* camera::render with the cache: as mpt_render, and paths that reach the
 * hit of bounce query_depth (all but the training paths) end there with
 * radiance += beta * cache(hit); after each batch the cache trains on the
 * training paths' radiance from that vertex on. nopts NULL = defaults;
 * nstats may be NULL. */
int32_t mpt_render_nrc(mpt_scene* scene, const mpt_camera* cam, const mpt_render_opts* opts, mpt_nrc* cache,
                       const mpt_nrc_opts* nopts, double* radiance, double* albedo, double* normal, double* variance,
                       mpt_pass_callback on_pass, void* user, mpt_render_stats* stats, mpt_nrc_stats* nstats);

#ifdef __cplusplus
}
#endif

#endif /* MPT_H */
