// scene_host.hpp - host-side form of a flattened scene, after validation:
// primitives in the layout the kernels read, BVHs, light tables and the GGX
// albedo table. Built once per mpt_scene_create and then copied to HBM.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mpt.h"
#include "geom.hpp"

namespace mpt {

// One BVH node: 64 bytes, box as (xmin, xmax, ymin, ymax, zmin, zmax) in
// double like the reference node (bvh.h:93-98); offset is the first item of a
// leaf or the right child of an interior node; count is 0 for interior nodes.
struct bvh_node {
    double b[6];
    int32_t offset;
    int32_t count;
    int32_t axis;
    int32_t pad;
};
static_assert(sizeof(bvh_node) == 64, "bvh_node must be 64 bytes");

// The same tree re-laid for the GPU walk: one record per interior node
// holding both children's boxes (left = the reference's index + 1, right =
// its offset) and what each child is: a leaf's item range (count > 0) or the
// record of an interior child (count == 0). 128 bytes.
struct cnode {
    double b[12];      // left box (xmin, xmax, ymin, ymax, zmin, zmax), then right box
    int32_t off[2];    // leaf: first item; interior: cnode index
    int32_t cnt[2];    // leaf: item count; interior: 0
    int32_t axis;      // split axis of the parent (near-first order)
    int32_t pad[3];
};
static_assert(sizeof(cnode) == 128, "cnode must be 128 bytes");

// Four-wide single-precision records for the traversal's culls, 144 bytes.
// A cull against their boxes never rejects a box the double precision test
// accepts; the double boxes of the cnode records decide the leaves (the
// traversal checks a leaf's double box when the leaf first yields a hit).
// A record stands for one interior node N of the binary tree and holds, per
// child of N, either the child itself (a leaf) or the child's two children
// (an interior child is expanded), so up to four slots: left child's part
// first, then the right child's. Slot boxes are rounded outward. A leaf
// slot holds its item range and where its double box is (the cnode record
// and child it is in); an interior slot holds the record of that node.
// `info` gives what visit order needs: N's split axis (bits 0-1), the left
// and right children's split axes when expanded (bits 2-3, 4-5), and which
// children are expanded (bits 6, 7). The binary tree's near-first order is
// then N's near child's part first, each expanded part near child first.
struct qnode {
    float b[24];       // slot s: xmin, xmax, ymin, ymax, zmin, zmax at b[6 s]
    int32_t off[4];    // leaf slot: first item; interior slot: qnode index
    int32_t meta[4];   // leaf slot: (cnode record << 1 | child) | count << 28; interior slot: 0
    int32_t info;
    int32_t pad[3];
};
static_assert(sizeof(qnode) == 144, "qnode must be 144 bytes");

// Four-wide records of the pair tree whose root record is `root_record`,
// appended to out; returns the index of the root's record.
int quad_layout(const std::vector<cnode>& pairs, int root_record, std::vector<qnode>& out);

// Root of a re-laid tree: its box and what it is (as a cnode child); for an
// interior root, qoff is its four-wide record.
struct croot {
    double b[6];
    int32_t off;
    int32_t cnt;
    int32_t qoff;
    int32_t pad;
};
static_assert(sizeof(croot) == 64, "croot must be 64 bytes");

// Primitive record the kernels read: 16 doubles (128 bytes).
//   sphere:   [0..2] center, [3] radius
//   quad:     [0..2] Q, [3..5] u, [6..8] v, [9..11] unit normal, [12..14] w, [15] D
//   triangle: [0..2] p0, [3..5] e1 = p1 - p0, [6..8] e2 = p2 - p0,
//             [9..11] unit geometric normal, [12] area
constexpr int kPrimStride = 16;

// Object record (one world-list entry).
struct obj_rec {
    int32_t kind;        // MPT_OBJ_*
    int32_t xform;       // -1 or index
    int64_t first;       // first primitive (device order: BLAS leaf order for meshes)
    int64_t count;
    int64_t id_base;     // first canonical ID
    int32_t blas_root;   // meshes: index of the BVH root (blas node array / cblas_roots)
    int32_t material;    // media: phase material
    double density;      // media
    int32_t bkind;       // media: boundary kind (MPT_OBJ_PRIM or MPT_OBJ_LIST)
    int32_t light_table; // mesh lights: offset of its triangles in mlight_prim / mlight_cdf (else -1)
    int32_t alpha;       // meshes: alpha-mask image texture (mesh.h:82-86), -1 when opaque
    int32_t flags;       // MPT_OBJF_*
};

// A height fog (height_fog, medium.h:98-155): density at y = base, scale
// height, base and its phase material.
struct fog_rec {
    double density;
    double scale;
    double base;
    int32_t material;
    int32_t pad;
};

struct light_rec {
    int32_t object;      // -1 = environment
    int32_t pad;
    double probability;
};

struct host_scene {
    // primitives in device order
    std::vector<double> prim;            // n * kPrimStride
    std::vector<int32_t> prim_kind;
    std::vector<int32_t> prim_mat;
    std::vector<int64_t> prim_canon;     // canonical local index within its object range
    std::vector<double> prim_nrm;        // triangles' vertex normals, 9 per primitive (NaN: none); empty if none
    std::vector<double> prim_uv;         // triangles' texture coordinates, 6 per primitive (NaN: none); empty if none
    std::vector<obj_rec> objects;
    std::vector<bvh_node> tlas;
    std::vector<int32_t> tlas_items;     // object indices in leaf order
    std::vector<bvh_node> blas;          // all mesh BVHs back to back
    // the same trees in child-pair records for the device walk
    std::vector<cnode> ctlas;
    croot ctlas_root{};
    std::vector<cnode> cblas;
    std::vector<croot> cblas_roots;      // indexed by obj_rec.blas_root
    std::vector<qnode> qtlas, qblas;     // the same trees four-wide, single precision (qoff of the roots)
    std::vector<int32_t> blas_root_node; // cblas_roots index -> root in blas (for depths)
    std::vector<double> spheres;         // center xyz, radius of every primitive (spheres only meaningful)
    std::vector<xform> xforms;
    std::vector<mpt_material> materials;
    std::vector<mpt_texture> textures;
    std::vector<uint8_t> texels;         // image textures' RGBA bytes
    std::vector<float> texels_hdr;       // image textures' RGB floats
    std::vector<double> srgb;            // 256: linear value of each 8-bit sRGB code (texture.h:197-208)
    bool uses_uv = false;                // some material reads (u, v) or a tangent frame at its hits
    bool rich = false;                   // uses_uv, thin glass, a measured metal or the ground: the variant with them
    int32_t ground_mat = -1;             // the ground plane's material, -1 without one
    double ground_h = 0;
    std::vector<mpt_vox_model> vox;      // voxel models (their palettes in scene materials)
    std::vector<uint8_t> vox_cells;      // every model's cells
    std::vector<fog_rec> fogs;           // height fogs, tested after the BVH and before the ground
    // the sun (sun.h:26-34): its frame around the direction, the cone's
    // cosine and solid angle, and radiance = irradiance / solid angle
    bool has_sun = false;
    onb sun_frame{};
    double sun_cos_max = 1, sun_solid_angle = 0;
    v3 sun_irradiance{0, 0, 0}, sun_radiance{0, 0, 0};
    std::vector<light_rec> lights;
    std::vector<double> light_cdf;       // n_lights + 1
    // mesh lights (mesh.h:175-216): per light mesh, its triangles in the
    // mesh's own order (device primitive index) and the running sum of their
    // areas in that order
    std::vector<int64_t> mlight_prim;
    std::vector<double> mlight_cdf;
    int32_t light_strategy = 1;
    // environment (environment.h)
    int32_t env_w = 0, env_h = 0;
    std::vector<double> env_rgb;         // w*h*3
    std::vector<double> env_cond_cdf;    // h*(w+1)
    std::vector<double> env_marg_cdf;    // h+1
    std::vector<double> env_row_weight;  // h
    double env_total_weight = 0;
    double env_intensity = 1;
    double env_rotation = 0;             // turns
    // GGX albedo table (microfacet.h:265-338)
    std::vector<double> ggx_e;           // 32*32 [roughness][mu]
    std::vector<double> ggx_e_avg;       // 32
    box3 world_box;
};

// A measured metal's head-on color (metal) and hemispherical average Fresnel
// from its index (metal_fresnel::measured, microfacet.h:109-121).
void conductor_averages(mpt_material& m);

// Validates the description and builds the host scene. Returns MPT_OK or an
// error code with a message in err.
int build_host_scene(const mpt_scene_desc& d, host_scene& s, std::string& err);

// light_list::build (light_list.h:45-96) on a built host scene: each light's
// power with the radius of `bounds` (half its diagonal), and the selection
// probabilities and CDF from them (uniform_fraction as given); stores them
// in s.lights and s.light_cdf. power_out, when set, gets the raw powers.
void light_selection(host_scene& s, const box3& bounds, double uniform_fraction, std::vector<double>* power_out);

// BVH build of bvh.h:118-205 over items; reorders items into leaf order.
struct build_item {
    box3 box;
    v3 centroid;
    int64_t ref;
};
void build_bvh(std::vector<build_item>& items, std::vector<bvh_node>& nodes);

// Re-lays the tree rooted at nodes[root] as child-pair records appended to out.
void pair_layout(const std::vector<bvh_node>& nodes, int root, std::vector<cnode>& out, croot& r);

// ggx_albedo_table (microfacet.h:299-337).
void build_ggx_table(std::vector<double>& e, std::vector<double>& e_avg);

}  // namespace mpt
