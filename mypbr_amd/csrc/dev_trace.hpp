// dev_trace.hpp - device-side scene view and closest-hit traversal.
//
// Two-level BVH in the reference's layout: a top-level BVH over the world
// list (main.cpp:70) whose leaves are objects (a primitive, a linear list, a
// mesh with its own BVH, or a medium), each optionally behind an affine
// transform (instance.h:29-41). Traversal follows bvh::hit (bvh.h:45-86):
// near-first order by the split axis sign, leaf tests in double precision
// that shrink the interval on every hit, the leaves' slab test (aabb.h:40-54)
// in double precision. Identical leaf order and identical arithmetic give
// identical closest hits, ties included.
//
// Interior boxes are culled in single precision against four-wide records
// whose boxes are rounded outward, with the ray origin moved by a bound on
// the rounding error: such a cull never rejects a box the double test
// accepts, so it only adds visits. A leaf's double box is tested, with the
// interval the leaf was entered with, when one of its items first reports a
// hit; if the double test rejects the box, the reference never visits the
// leaf and its hits are dropped. Which leaves count, and in which order, is
// therefore the reference's.
//
// The traversal stack lives in LDS (a short stack of kLdsDepth entries per
// lane, strided by the block size so lanes hit distinct banks) and spills to
// a per-thread area in HBM beyond that.
#pragma once

#include <hip/hip_runtime.h>

#include "geom.hpp"
#include "rng.hpp"
#include "scene_host.hpp"

namespace mpt {

struct dscene {
    const double* prim;
    const int32_t* pkind;
    const int32_t* pmat;
    const int64_t* pcanon;
    const obj_rec* obj;
    const cnode* tlas;        // child-pair records of the top-level BVH
    const int32_t* titems;
    const cnode* blas;        // child-pair records of all mesh BVHs
    const qnode* qtlas;       // four-wide records of the same trees (root records in croot.qoff)
    const qnode* qblas;
    const croot* blas_roots;  // per mesh, indexed by obj_rec.blas_root
    const double* sph;        // center xyz, radius per primitive (spheres)
    const xform* xf;
    const mpt_material* mat;
    const mpt_texture* tex;
    const light_rec* lights;
    const double* lcdf;
    const double* vnrm;          // triangles' vertex normals, 9 per primitive (NaN: none), or null
    const double* vuv;           // triangles' texture coordinates, 6 per primitive (NaN: none), or null
    const uint8_t* texels;       // image textures: RGBA bytes
    const float* texels_hdr;     // image textures: RGB floats
    const double* srgb;          // 256: linear value of each 8-bit sRGB code
    const int64_t* mlight_prim;  // mesh lights: triangles in mesh order, and their area CDF
    const double* mlight_cdf;
    const double* env_rgb;
    const double* env_cond;
    const double* env_marg;
    const double* env_roww;
    const double* ggx_e;
    const double* ggx_eavg;
    double env_total;
    double env_intensity;
    double env_rotation;
    int32_t nobj;
    int32_t nlights;
    int32_t lstrategy;
    int32_t env_w, env_h;
    int32_t ground_mat;  // the ground plane's material, -1 without one (hits on it name object nobj)
    double ground_h;
    int32_t nfog;        // height fogs (hits on fog k name object nobj + 1 + k)
    const fog_rec* fog;
    // 1 when no material emits or lets a shadow ray through: a shadow ray then
    // only asks whether anything is in the way (an occlusion query)
    int32_t shadow_occlusion;
    int32_t nworld;      // objects in the top-level BVH (the others are only lights)
    const mpt_vox_model* vox;  // voxel models and their cells
    const uint8_t* vcells;
    int64_t ground_id;   // its canonical ID: one past the last primitive's
    croot tlas_root;
    // the sun (sun.h): on, its frame (w towards the sun), cone and radiance
    int32_t has_sun;
    onb sun_frame;
    double sun_cos_max, sun_solid_angle;
    v3 sun_radiance;
};

// Per-lane traversal stack: LDS for the first kLdsDepth entries, then HBM.
struct tstack {
    int32_t* lds;        // this block's LDS stack, [depth][block]
    int32_t* ovf;        // spill area, [depth][total threads]
    int32_t tid;         // lane index within the block
    int32_t bs;          // block size
    int64_t gtid;        // global thread index
    int64_t nthreads;    // total threads of the persistent grid
    int32_t lds_depth;
    __device__ __forceinline__ void put(int sp, int32_t v) {
        if (sp < lds_depth) lds[sp * bs + tid] = v;
        else ovf[int64_t(sp - lds_depth) * nthreads + gtid] = v;
    }
    __device__ __forceinline__ int32_t get(int sp) const {
        if (sp < lds_depth) return lds[sp * bs + tid];
        return ovf[int64_t(sp - lds_depth) * nthreads + gtid];
    }
};

// BVH work of one lane (counted only in the instrumented kernels): nodes
// visited and primitive bytes read (sphere 32, quad 128, triangle 72).
struct wcount {
    uint32_t nodes;
    uint32_t pbytes;
};

// Geometry features a scene uses; traversal kernels are compiled for the
// subset a scene needs so that unused primitive kinds and object kinds cost
// neither instructions nor registers.
enum : int {
    F_SPHERE = 1,
    F_QUAD = 2,
    F_TRI = 4,
    F_XFORM = 8,
    F_LIST = 16,
    F_MESH = 32,
    F_MEDIUM = 64,
    F_ALL = 127,
    // not geometry: hits also get (u, v) and dp/du, dp/dv, and shading has image
    // textures, texture and normal maps, thin glass and measured metals
    F_UV = 128,
};

struct hitinfo {
    double t;
    int32_t obj;    // object index, -1 = miss
    int32_t prim;   // device primitive index, -1 for a medium
};

// Slab test of aabb::hit (aabb.h:40-54) on a box given as (xmin, xmax, ymin, ymax, zmin, zmax).
// The early return per axis is kept: most boxes are missed on the first
// axis or two, and a branch the whole wave takes skips the rest.
__device__ __forceinline__ bool box_hit(const double* b, const v3& o, const v3& inv, double tmin, double tmax) {
    const double oo[3] = {o.x, o.y, o.z};
    const double ii[3] = {inv.x, inv.y, inv.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double t0 = (b[2 * a] - oo[a]) * ii[a];
        double t1 = (b[2 * a + 1] - oo[a]) * ii[a];
        if (t0 > t1) {
            double s = t0;
            t0 = t1;
            t1 = s;
        }
        if (t0 > tmin) tmin = t0;
        if (t1 < tmax) tmax = t1;
        if (tmax <= tmin) return false;
    }
    return true;
}

// A ray in single precision for the box culls. Per axis the origin is moved
// by delta = 2^-20 (|o| + the root box's largest coordinate magnitude) in
// both directions: o + delta for the planes at box minima, o - delta for the
// maxima. delta bounds (with a factor 4 to spare) what rounding the box, the
// origin, the reciprocal direction and the two operations to single
// precision can move a slab distance, in units of the inverse direction, so
// every single-precision interval contains the double one. An axis whose
// reciprocal direction is infinite behaves as in aabb::hit: the planes then
// give +-inf (or NaN, which the comparisons ignore).
struct fray {
    float olo[3], ohi[3], inv[3];
    float tmin;  // the interval's start, rounded down
};

// x rounded to single precision towards +inf / -inf.
__device__ __forceinline__ float f32_up(double x) {
    float f = float(x);
    if (double(f) < x) {  // f is finite here: step to the next float above
        const int32_t i = __float_as_int(f);
        f = f == 0.0f ? __int_as_float(1) : __int_as_float(i >= 0 ? i + 1 : i - 1);
    }
    return f;
}
__device__ __forceinline__ float f32_down(double x) {
    float f = float(x);
    if (double(f) > x) {
        const int32_t i = __float_as_int(f);
        f = f == 0.0f ? -__int_as_float(1) : __int_as_float(i >= 0 ? i - 1 : i + 1);
    }
    return f;
}

__device__ __forceinline__ fray make_fray(const v3& o, const v3& inv, double tmin, const double* root_box) {
    fray r;
    const double oo[3] = {o.x, o.y, o.z};
    const double ii[3] = {inv.x, inv.y, inv.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double bm = std::fmax(std::fabs(root_box[2 * a]), std::fabs(root_box[2 * a + 1]));
        const double delta = (std::fabs(oo[a]) + bm) * 0x1p-20;
        r.olo[a] = float(oo[a] + delta);
        r.ohi[a] = float(oo[a] - delta);
        r.inv[a] = float(ii[a]);
    }
    r.tmin = f32_down(tmin);
    return r;
}

// Conservative single-precision slab test: true whenever the double test of
// aabb::hit (aabb.h:40-54) on the unrounded box is true. tmax rounded up.
__device__ __forceinline__ bool fbox_hit(const float* b, const fray& r, float tmax) {
    float lo = r.tmin, hi = tmax;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float t0 = (b[2 * a] - r.olo[a]) * r.inv[a];
        float t1 = (b[2 * a + 1] - r.ohi[a]) * r.inv[a];
        if (t0 > t1) {
            const float t = t0;
            t0 = t1;
            t1 = t;
        }
        if (t0 > lo) lo = t0;
        if (t1 < hi) hi = t1;
    }
    return hi >= lo;
}

// The double-precision test of leaf e (child e & 1 of record e >> 1; e < 0:
// the root, already tested exactly) with the interval the leaf was entered
// with: the test the reference makes when it visits the leaf.
// The reciprocal direction is recomputed here rather than kept live through
// the walk: the check is rare and the registers are not.
template <bool CNT>
__device__ __forceinline__ bool leaf_exact(const cnode* nodes, int e, const v3& o, const v3& d, double tmin,
                                           double tmax_at_entry, wcount& wc) {
    if (e < 0) return true;
    if constexpr (CNT) wc.nodes += 48u;
    const v3 inv{1 / d.x, 1 / d.y, 1 / d.z};
    const double2* p = reinterpret_cast<const double2*>(nodes + (e >> 1)) + 3 * (e & 1);
    double b[6];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double2 v = p[k];
        b[2 * k] = v.x;
        b[2 * k + 1] = v.y;
    }
    return box_hit(b, o, inv, tmin, tmax_at_entry);
}

// Walk of bvh::hit (bvh.h:45-86) over four-wide records, two levels of the
// binary tree per record. A record's up-to-four slots are culled with the
// current interval; the slots that pass are taken in the binary tree's
// near-first order by the split axis sign (bvh.h:60-80; the node's near
// child's part first, an expanded child's near child first): the first is
// entered (or its leaf tested), the rest pushed and culled again, with the
// interval as it is then, when popped, as the reference tests a node when it
// pops it. A box the double test rejects with a larger tmax it rejects with
// a smaller one, and the culls are conservative, so every leaf the reference
// visits is reached, in the reference's order; leaf(off, cnt, sp, e) tests a
// leaf's items, shrinks tmax on hits, keeps a hit only if the leaf's
// double box test (leaf_exact) holds, and returns true to end the walk (an
// occlusion query's first hit). Against a binary walk over the same
// tree in 64-byte single-precision records this took the spheres scene from
// 1449 to 1721 M paths/s and the generated batch from 15.6 to 21.9 M.

__device__ __forceinline__ int neg_mask(const fray& fr) {
    return (fr.inv[0] < 0 ? 1 : 0) | (fr.inv[1] < 0 ? 2 : 0) | (fr.inv[2] < 0 ? 4 : 0);
}

// One record visit of the four-wide walk: record cur's slots culled with the
// interval up to tmax; every slot that passes but the first in visit order
// is pushed (farthest first); the first's (off, meta) are returned, or false
// when no slot passes.
template <bool CNT>
__device__ __forceinline__ bool visit_record(const qnode* qnodes, int cur, const fray& fr, int negmask, double tmax,
                                             tstack& st, int& sp, int32_t& off, int32_t& meta, wcount& wc) {
    const qnode* q = qnodes + cur;
    float b[24];
    {
        const float4* p = reinterpret_cast<const float4*>(q);
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const float4 v = p[k];
            b[4 * k] = v.x, b[4 * k + 1] = v.y, b[4 * k + 2] = v.z, b[4 * k + 3] = v.w;
        }
    }
    const int4 offs = reinterpret_cast<const int4*>(q)[6];
    const int4 metas = reinterpret_cast<const int4*>(q)[7];
    const int info = reinterpret_cast<const int*>(q)[32];
    if constexpr (CNT) wc.nodes += 144u;
    // visit order: slot positions, two bits each
    const int nN = (negmask >> (info & 3)) & 1;
    const int nL = (negmask >> ((info >> 2) & 3)) & 1;
    const int nR = (negmask >> ((info >> 4) & 3)) & 1;
    const bool eL = (info >> 6) & 1, eR = (info >> 7) & 1;
    const int ln = eL ? 2 : 1, rn = eR ? 2 : 1, rb = ln;
    const uint32_t lord = eL ? (nL ? 1u : 4u) : 0u;  // [1, 0] or [0, 1] or [0]
    const uint32_t rord = eR ? (nR ? uint32_t(rb + 1) | uint32_t(rb) << 2 : uint32_t(rb) | uint32_t(rb + 1) << 2)
                             : uint32_t(rb);
    const uint32_t ord = nN ? (rord | lord << (2 * rn)) : (lord | rord << (2 * ln));
    const int nslots = ln + rn;
    const float tf = f32_up(tmax);
    uint32_t hits = 0;
#pragma unroll
    for (int sl = 0; sl < 4; sl++)
        if (sl < nslots && fbox_hit(b + 6 * sl, fr, tf)) hits |= 1u << sl;
    // the first slot that passes is taken now; the others are pushed, farthest first
    int first = -1;
#pragma unroll
    for (int j = 3; j >= 0; j--) {
        const int sl = int(ord >> (2 * j)) & 3;
        if (j < nslots && ((hits >> sl) & 1)) {
            if (first >= 0) st.put(sp++, (cur << 2) | first);
            first = sl;
        }
    }
    if (first < 0) return false;
    off = first == 0 ? offs.x : first == 1 ? offs.y : first == 2 ? offs.z : offs.w;
    meta = first == 0 ? metas.x : first == 1 ? metas.y : first == 2 ? metas.z : metas.w;
    return true;
}

// A pushed slot e (record e >> 2, slot e & 3) culled again with the interval
// as it is now; its (off, meta) when it passes.
template <bool CNT>
__device__ __forceinline__ bool pop_slot(const qnode* qnodes, int e, const fray& fr, double tmax, int32_t& off,
                                         int32_t& meta, wcount& wc) {
    const qnode* pq = qnodes + (e >> 2);
    const int sl = e & 3;
    float bb[6];
    const float2* pb = reinterpret_cast<const float2*>(pq->b + 6 * sl);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float2 v = pb[k];
        bb[2 * k] = v.x;
        bb[2 * k + 1] = v.y;
    }
    off = pq->off[sl];
    meta = pq->meta[sl];
    if constexpr (CNT) wc.nodes += 32u;
    return fbox_hit(bb, fr, f32_up(tmax));
}

template <bool CNT, typename Leaf>
__device__ __forceinline__ void walk4(const qnode* qnodes, const croot& root, const v3& o, const v3& d, double tmin,
                                      double& tmax, tstack& st, int sp_base, wcount& wc, Leaf&& leaf) {
    if constexpr (CNT) wc.nodes += 64u;
    fray fr;
    {
        const v3 inv{1 / d.x, 1 / d.y, 1 / d.z};
        if (!box_hit(root.b, o, inv, tmin, tmax)) return;
        if (root.cnt > 0) {
            (void)leaf(root.off, root.cnt, sp_base, -1);
            return;
        }
        fr = make_fray(o, inv, tmin, root.b);
    }
    const int negmask = neg_mask(fr);
    int sp = sp_base;
    int cur = root.qoff;
    while (true) {
        int next = -1;
        int32_t off, meta;
        if (visit_record<CNT>(qnodes, cur, fr, negmask, tmax, st, sp, off, meta, wc)) {
            const int32_t cnt = int32_t(uint32_t(meta) >> 28);
            if (cnt > 0) {
                if (leaf(off, cnt, sp, meta & 0x0fffffff)) return;  // an occlusion query found its hit
            } else {
                next = off;
            }
        }
        while (next < 0) {
            if (sp == sp_base) return;
            const int e = st.get(--sp);
            if (!pop_slot<CNT>(qnodes, e, fr, tmax, off, meta, wc)) continue;
            const int32_t cnt = int32_t(uint32_t(meta) >> 28);
            if (cnt > 0) {
                if (leaf(off, cnt, sp, meta & 0x0fffffff)) return;
            } else {
                next = off;
            }
        }
        cur = next;
    }
}

__device__ __forceinline__ const double* prim_ptr(const dscene& S, int64_t p) { return S.prim + p * kPrimStride; }

// sphere::hit (sphere.h:23-41), t only.
__device__ __forceinline__ bool sphere_t(const double* q, const v3& o, const v3& d, double tmin, double tmax,
                                         double& t) {
    v3 c{q[0], q[1], q[2]};
    double radius = q[3];
    v3 oc = c - o;
    double a = length_squared(d);
    double h = dot(d, oc);
    double cc = length_squared(oc) - radius * radius;
    double disc = h * h - a * cc;
    if (disc < 0) return false;
    double sq = std::sqrt(disc);
    double root = (h - sq) / a;
    if (!(tmin < root && root < tmax)) {
        root = (h + sq) / a;
        if (!(tmin < root && root < tmax)) return false;
    }
    t = root;
    return true;
}

// quad::hit (quad.h:30-52), t only; the interval is inclusive.
__device__ __forceinline__ bool quad_t(const double* q, const v3& o, const v3& d, double tmin, double tmax,
                                       double& t) {
    v3 nrm{q[9], q[10], q[11]};
    double denom = dot(nrm, d);
    if (std::fabs(denom) < 1e-8) return false;
    double tt = (q[15] - dot(nrm, o)) / denom;
    if (!(tmin <= tt && tt <= tmax)) return false;
    v3 ip = o + tt * d;
    v3 Q{q[0], q[1], q[2]}, u{q[3], q[4], q[5]}, v{q[6], q[7], q[8]}, w{q[12], q[13], q[14]};
    v3 ph = ip - Q;
    double alpha = dot(w, cross(ph, v));
    double beta = dot(w, cross(u, ph));
    if (!(0 <= alpha && alpha <= 1) || !(0 <= beta && beta <= 1)) return false;
    t = tt;
    return true;
}

// triangle::hit (mesh.h:47-69), Moller-Trumbore, t only.
__device__ __forceinline__ bool tri_t(const double* q, const v3& o, const v3& d, double tmin, double tmax,
                                      double& t) {
    v3 p0{q[0], q[1], q[2]}, e1{q[3], q[4], q[5]}, e2{q[6], q[7], q[8]};
    v3 pvec = cross(d, e2);
    double det = dot(e1, pvec);
    if (std::fabs(det) < 1e-12) return false;
    double inv_det = 1 / det;
    v3 tvec = o - p0;
    double b1 = dot(tvec, pvec) * inv_det;
    if (b1 < 0 || b1 > 1) return false;
    v3 qvec = cross(tvec, e1);
    double b2 = dot(d, qvec) * inv_det;
    if (b2 < 0 || b1 + b2 > 1) return false;
    double tt = dot(e2, qvec) * inv_det;
    if (!(tmin < tt && tt < tmax)) return false;
    t = tt;
    return true;
}

template <bool CNT, int FEAT>
__device__ __forceinline__ bool prim_t(const dscene& S, int64_t p, const v3& o, const v3& d, double tmin,
                                       double tmax, double& t, wcount& wc) {
    constexpr int kinds = (FEAT & F_SPHERE ? 1 : 0) + (FEAT & F_QUAD ? 1 : 0) + (FEAT & F_TRI ? 1 : 0);
    if constexpr (kinds == 1 && (FEAT & F_SPHERE)) {
        if constexpr (CNT) wc.pbytes += 32u;
        return sphere_t(S.sph + 4 * p, o, d, tmin, tmax, t);
    }
    const double* q = prim_ptr(S, p);
    int k = S.pkind[p];
    if constexpr (CNT) wc.pbytes += k == MPT_PRIM_SPHERE ? 32u : (k == MPT_PRIM_QUAD ? 128u : 72u);
    if constexpr (kinds == 1) {
        if constexpr (FEAT & F_SPHERE) return sphere_t(q, o, d, tmin, tmax, t);
        if constexpr (FEAT & F_QUAD) return quad_t(q, o, d, tmin, tmax, t);
        return tri_t(q, o, d, tmin, tmax, t);
    }
    if ((FEAT & F_SPHERE) && k == MPT_PRIM_SPHERE) return sphere_t(q, o, d, tmin, tmax, t);
    if ((FEAT & F_QUAD) && k == MPT_PRIM_QUAD) return quad_t(q, o, d, tmin, tmax, t);
    if constexpr (FEAT & F_TRI) return tri_t(q, o, d, tmin, tmax, t);
    return false;
}

// hittable_list::hit (hittable_list.h:25-39) over [first, first + count).
template <bool CNT, int FEAT>
__device__ __forceinline__ bool list_t(const dscene& S, int64_t first, int64_t count, const v3& o, const v3& d,
                                       double tmin, double tmax, double& t, int32_t& which, wcount& wc) {
    bool any = false;
    double closest = tmax;
    for (int64_t k = 0; k < count; k++) {
        double tt;
        if (prim_t<CNT, FEAT>(S, first + k, o, d, tmin, closest, tt, wc)) {
            any = true;
            closest = tt;
            which = int32_t(first + k);
        }
    }
    t = closest;
    return any;
}

// Coverage of an alpha mask at (u, v): the nearest texel's alpha byte / 255,
// or its first byte for an image without alpha (image_texture::alpha_value,
// texture.h:130-136, 182-195).
__device__ __forceinline__ double mask_coverage(const dscene& S, int ti, double u, double v) {
    const mpt_texture& t = S.tex[ti];
    if (t.flags & MPT_TEXF_REPEAT) {
        u -= std::floor(u);
        v -= std::floor(v);
    } else {
        u = std::fmin(std::fmax(u, 0.0), 1.0);
        v = std::fmin(std::fmax(v, 0.0), 1.0);
    }
    v = 1.0 - v;  // image rows run down
    const int i = min(max(int(u * t.width), 0), t.width - 1);
    const int j = min(max(int(v * t.height), 0), t.height - 1);
    const int64_t k = t.offset + int64_t(j) * t.width + i;
    return double(S.texels[4 * k + ((t.flags & MPT_TEXF_NO_ALPHA) ? 0 : 3)]) / 255.0;
}

// triangle::hit's alpha test (mesh.h:72-86) for device primitive p of a mesh
// with mask ti, hit by the ray (o, d) (instance space): (u, v) from the
// triangle's texture coordinates, else its barycentrics; a coverage c < 1
// keeps the hit when the draw is below c. The draw is keyed by the ray's
// medium stream and the triangle (its mesh's first primitive plus its index
// in the mesh, shared by every instance), so it does not depend on the order
// in which the walk meets the triangles.
__device__ __forceinline__ bool alpha_keeps(const dscene& S, int ti, int64_t mesh_first, int64_t p, const v3& o,
                                            const v3& d, uint64_t mkey, uint32_t serial) {
    const double* q = prim_ptr(S, p);
    const v3 p0{q[0], q[1], q[2]}, e1{q[3], q[4], q[5]}, e2{q[6], q[7], q[8]};
    const v3 pvec = cross(d, e2);
    const double inv_det = 1 / dot(e1, pvec);
    const v3 tvec = o - p0;
    const double b1 = dot(tvec, pvec) * inv_det;
    const double b2 = dot(d, cross(tvec, e1)) * inv_det;
    double u = b1, v = b2;
    const double* t6 = S.vuv ? S.vuv + 6 * p : nullptr;
    if (t6 && !std::isnan(t6[0])) {
        const double b0 = 1 - b1 - b2;
        u = b0 * t6[0] + b1 * t6[2] + b2 * t6[4];
        v = b0 * t6[1] + b1 * t6[3] + b2 * t6[5];
    }
    const double c = mask_coverage(S, ti, u, v);
    if (!(c < 1)) return true;
    const uint32_t salt = 0x80000000u | uint32_t(mesh_first + S.pcanon[p]);
    return medium_draw(mkey, serial, salt) < c;
}

// Inner bvh::hit of a mesh; the stack continues above sp_base. alpha: the
// mesh's mask (-1: opaque), mesh_first its first primitive, (mkey, serial)
// the ray's medium stream for the mask's draws; masks read (u, v), so only
// the variants that carry it (F_UV) test them (the host puts every masked
// scene there). any_hit: stop at the first accepted hit.
// The triangles of a mesh leaf (off, cnt; e its child-pair record for the
// double test, made at the leaf's first hit): hits shrink tmax and set which
// and any; returns true when any_hit asks to stop at the first accepted hit.
template <bool CNT, int FEAT>
__device__ __forceinline__ bool mesh_leaf(const dscene& S, int off, int cnt, int e, const v3& o, const v3& d,
                                          double tmin, double& tmax, int32_t& which, bool& any, int alpha,
                                          int64_t mesh_first, uint64_t mkey, uint32_t serial, bool any_hit,
                                          wcount& wc) {
    const double tleaf = tmax;
    int ok = -1;  // the leaf's double test, made at its first hit
    for (int k = 0; k < cnt; k++) {
        double tt;
        if constexpr (CNT) wc.pbytes += 72u;
        if (!tri_t(prim_ptr(S, off + k), o, d, tmin, tmax, tt)) continue;
        if constexpr (FEAT & F_UV)
            if (alpha >= 0 && !alpha_keeps(S, alpha, mesh_first, off + k, o, d, mkey, serial)) continue;
        if (ok < 0) ok = leaf_exact<CNT>(S.blas, e, o, d, tmin, tleaf, wc) ? 1 : 0;
        if (!ok) break;
        any = true;
        tmax = tt;
        which = off + k;
        if (any_hit) return true;  // the first hit answers an occlusion query: stop the walk
    }
    return false;
}

template <bool CNT, int FEAT>
__device__ bool mesh_t(const dscene& S, int root, const v3& o, const v3& d, double tmin, double tmax, double& t,
                       int32_t& which, tstack& st, int sp_base, wcount& wc, int alpha, int64_t mesh_first,
                       uint64_t mkey, uint32_t serial, bool any_hit) {
    bool any = false;
    walk4<CNT>(S.qblas, S.blas_roots[root], o, d, tmin, tmax, st, sp_base, wc, [&](int off, int cnt, int, int e) {
        return mesh_leaf<CNT, FEAT>(S, off, cnt, e, o, d, tmin, tmax, which, any, alpha, mesh_first, mkey, serial,
                                    any_hit, wc);
    });
    t = tmax;
    return any;
}

// Hit of an object's boundary (a primitive or a list) for media.
template <bool CNT, int FEAT>
__device__ __forceinline__ bool boundary_t(const dscene& S, const obj_rec& r, const v3& o, const v3& d, double tmin,
                                           double tmax, double& t, wcount& wc) {
    int32_t w;
    if (r.bkind == MPT_OBJ_PRIM) return prim_t<CNT, FEAT>(S, r.first, o, d, tmin, tmax, t, wc);
    return list_t<CNT, FEAT>(S, r.first, r.count, o, d, tmin, tmax, t, w, wc);
}

// constant_medium::hit (medium.h:60-88), with the free-flight draw u.
// d is the boundary's (possibly instance-local) direction; world_len is the
// length of the world-space direction, which sets distances in the medium.
template <bool CNT, int FEAT>
__device__ bool medium_t(const dscene& S, const obj_rec& r, const v3& o, const v3& d, double world_len, double tmin,
                         double tmax, double u, double& t, wcount& wc) {
    double t1a, t2a;
    if (!boundary_t<CNT, FEAT>(S, r, o, d, -kInf, kInf, t1a, wc)) return false;
    if (!boundary_t<CNT, FEAT>(S, r, o, d, t1a + 0.0001, kInf, t2a, wc)) return false;
    double t0 = std::fmax(t1a, tmin);
    double t1 = std::fmin(t2a, tmax);
    if (t0 >= t1) return false;
    double ray_length = world_len;
    double distance_inside = (t1 - t0) * ray_length;
    double hit_distance = -std::log(1 - u) / r.density;
    if (hit_distance > distance_inside) return false;
    t = t0 + hit_distance / ray_length;
    return true;
}

// ---- voxel grids (voxel_grid, voxel.h:126-337)
// Cells of a block are addressed as three scalars, and axes picked with
// selects, so that nothing on the walk is an array indexed at run time
// (which would live in scratch memory).

__device__ __forceinline__ int isel(int x, int y, int z, int a) { return a == 0 ? x : (a == 1 ? y : z); }
__device__ __forceinline__ double dsel(const v3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// The palette index of cell (x, y, z) of a model, 0 outside it (voxel.h:276-281).
__device__ __forceinline__ int vox_at(const mpt_vox_model& m, const uint8_t* cells, int x, int y, int z) {
    if (unsigned(x) >= unsigned(m.nx) || unsigned(y) >= unsigned(m.ny) || unsigned(z) >= unsigned(m.nz)) return 0;
    return cells[(int64_t(z) * m.ny + y) * m.nx + x];
}

// Whether the face of a cell of kind k towards a neighbour of kind n is a
// surface (voxel_face_visible, voxel.h:103-111): 0 empty, 1 opaque, 2
// transparent.
__device__ __forceinline__ bool vox_face(int k, int n) { return k != 0 && (n == 0 || (k == 1 && n == 2)); }

// The face a walk stopped on: the plane perpendicular to axis, the outward
// normal's sign along it, the palette index and cell of the voxel it bounds.
struct vox_hit {
    int axis, sign, index;
    int x, y, z;
};

// voxel_grid::hit (voxel.h:193-267) over block q (lo xyz, hi xyz, model) for
// the ray (o, d) in model space: a 3D DDA (Amanatides and Woo 1987) from
// where the ray enters the block (or from tmin inside it), stopping at the
// first plane crossing where the two cells differ in kind (crossing,
// voxel.h:283-298), within tmax. The entry, the cell clamping and the plane
// times follow the reference's arithmetic, so t is bitwise the reference's.
// With REC the face is written to f.
template <bool REC>
__device__ bool vox_walk(const dscene& S, const double* q, const v3& o, const v3& d, double tmin, double tmax,
                         double& t_hit, vox_hit* f) {
    const int lo0 = int(q[0]), lo1 = int(q[1]), lo2 = int(q[2]);
    const int hi0 = int(q[3]), hi1 = int(q[4]), hi2 = int(q[5]);
    const mpt_vox_model& m = S.vox[int(q[6])];
    const uint8_t* cells = S.vcells + m.cells;
    const uint8_t* kind = m.kind;
    // where the ray's line enters and leaves the block
    double t_enter = -kInf, t_exit = kInf;
    int enter_axis = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double oa = dsel(o, a), da = dsel(d, a);
        const int la = a == 0 ? lo0 : (a == 1 ? lo1 : lo2), ha = a == 0 ? hi0 : (a == 1 ? hi1 : hi2);
        if (da == 0) {
            if (oa < la || oa > ha) return false;
            continue;
        }
        double t0 = (la - oa) / da, t1 = (ha - oa) / da;
        if (t0 > t1) {
            const double tt = t0;
            t0 = t1;
            t1 = tt;
        }
        if (t0 > t_enter) {
            t_enter = t0;
            enter_axis = a;
        }
        t_exit = std::fmin(t_exit, t1);
    }
    const bool outside = t_enter >= tmin;
    const double t_start = outside ? t_enter : tmin;
    if (t_start > std::fmin(t_exit, tmax)) return false;
    // the first cell and the DDA state per axis: step, next plane's t, t between planes
    int c[3], stp[3];
    double tn[3], td[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double oa = dsel(o, a), da = dsel(d, a);
        const int la = a == 0 ? lo0 : (a == 1 ? lo1 : lo2), ha = a == 0 ? hi0 : (a == 1 ? hi1 : hi2);
        const double sa = oa + t_start * da;  // ray::at
        int ca = int(std::floor(sa));
        ca = ca < la ? la : (ca > ha - 1 ? ha - 1 : ca);
        if (outside && a == enter_axis) ca = da > 0 ? la : ha - 1;  // the layer just inside the entry face
        c[a] = ca;
        if (da > 0) {
            stp[a] = 1;
            tn[a] = (ca + 1 - oa) / da;
            td[a] = 1 / da;
        } else if (da < 0) {
            stp[a] = -1;
            tn[a] = (ca - oa) / da;
            td[a] = -1 / da;
        } else {
            stp[a] = 0;
            tn[a] = td[a] = kInf;
        }
    }
    // crossing (voxel.h:283-298): into cell (x, y, z) of palette index to from
    // one of index from, across the plane of axis a at step s
    auto crossing = [&](double t, int a, int s, int from, int to, int x, int y, int z) {
        const int kf = kind[from], kt = kind[to];
        if (kt != 0 && vox_face(kt, kf)) {
            if (REC) *f = vox_hit{a, -s, to, x, y, z};
            t_hit = t;
            return true;
        }
        if (vox_face(kf, kt)) {
            if (REC) *f = vox_hit{a, s, from, a == 0 ? x - s : x, a == 1 ? y - s : y, a == 2 ? z - s : z};
            t_hit = t;
            return true;
        }
        return false;
    };
    int index = vox_at(m, cells, c[0], c[1], c[2]);
    if (outside) {  // from the cell behind the entry face
        const int s = isel(stp[0], stp[1], stp[2], enter_axis);
        const int bx = enter_axis == 0 ? c[0] - s : c[0], by = enter_axis == 1 ? c[1] - s : c[1];
        const int bz = enter_axis == 2 ? c[2] - s : c[2];
        if (crossing(t_enter, enter_axis, s, vox_at(m, cells, bx, by, bz), index, c[0], c[1], c[2])) return true;
    }
    // every step moves one cell towards the block's exit, so the walk ends;
    // a nearest plane without a step (a zero or NaN direction component, t
    // infinite) ends it too, as no hit: the reference would loop there
    while (true) {
        // the nearest plane ahead decides the next cell
        const int a = tn[0] < tn[1] ? (tn[0] < tn[2] ? 0 : 2) : (tn[1] < tn[2] ? 1 : 2);
        double t;
        int s;
        bool out;
        if (a == 0) {
            t = tn[0], s = stp[0];
            if (t > tmax || s == 0) return false;
            c[0] += s;
            tn[0] += td[0];
            out = c[0] < lo0 || c[0] >= hi0;
        } else if (a == 1) {
            t = tn[1], s = stp[1];
            if (t > tmax || s == 0) return false;
            c[1] += s;
            tn[1] += td[1];
            out = c[1] < lo1 || c[1] >= hi1;
        } else {
            t = tn[2], s = stp[2];
            if (t > tmax || s == 0) return false;
            c[2] += s;
            tn[2] += td[2];
            out = c[2] < lo2 || c[2] >= hi2;
        }
        const int next = vox_at(m, cells, c[0], c[1], c[2]);  // 0 outside the model
        if (crossing(t, a, s, index, next, c[0], c[1], c[2])) return true;
        if (out) return false;  // left the block
        index = next;
    }
}

// on_edge (voxel.h:300-322): whether a hit at p on the face of cell (x, y,
// z) whose outward normal runs along axis (sign) lies within the model's
// edge width of an outline: a convex edge (the face around the corner is a
// surface too) or a concave one (a voxel stands beside, in front).
__device__ __forceinline__ bool vox_on_edge(const dscene& S, const mpt_vox_model& m, const v3& p, const vox_hit& f) {
    const uint8_t* cells = S.vcells + m.cells;
    const int k = m.kind[vox_at(m, cells, f.x, f.y, f.z)];
    for (int side = 1; side <= 2; side++) {
        const int tangent = (f.axis + side) % 3;
        const double fr = dsel(p, tangent) - isel(f.x, f.y, f.z, tangent);
        for (int dir = -1; dir <= 1; dir += 2) {
            if ((dir > 0 ? 1 - fr : fr) >= m.edge_width) continue;
            const int bx = f.x + (tangent == 0 ? dir : 0), by = f.y + (tangent == 1 ? dir : 0);
            const int bz = f.z + (tangent == 2 ? dir : 0);
            const int ax = bx + (f.axis == 0 ? f.sign : 0), ay = by + (f.axis == 1 ? f.sign : 0);
            const int az = bz + (f.axis == 2 ? f.sign : 0);
            if (vox_face(k, m.kind[vox_at(m, cells, bx, by, bz)]) || m.kind[vox_at(m, cells, ax, ay, az)] != 0)
                return true;
        }
    }
    return false;
}

// voxel_blocks' BVH (voxel.h:341-382): the model's blocks, each walked by
// vox_walk; the stack continues above sp_base. any_hit: stop at the first
// accepted hit.
template <bool CNT, int FEAT>
__device__ bool voxel_t(const dscene& S, int root, const v3& o, const v3& d, double tmin, double tmax, double& t,
                        int32_t& which, tstack& st, int sp_base, wcount& wc, bool any_hit) {
    bool any = false;
    walk4<CNT>(S.qblas, S.blas_roots[root], o, d, tmin, tmax, st, sp_base, wc, [&](int off, int cnt, int, int e) {
        const double tleaf = tmax;
        int ok = -1;  // the leaf's double test, made at its first hit
        for (int k = 0; k < cnt; k++) {
            double tt;
            if constexpr (CNT) wc.pbytes += 56u;
            if (!vox_walk<false>(S, prim_ptr(S, off + k), o, d, tmin, tmax, tt, nullptr)) continue;
            if (ok < 0) ok = leaf_exact<CNT>(S.blas, e, o, d, tmin, tleaf, wc) ? 1 : 0;
            if (!ok) break;
            any = true;
            tmax = tt;
            which = off + k;
            if (any_hit) return true;
        }
        return false;
    });
    t = tmax;
    return any;
}

// One world object, the ray in world space. Returns the hit t and primitive.
template <bool CNT, int FEAT>
__device__ __forceinline__ bool object_t(const dscene& S, int oi, v3 o, v3 d, double tmin, double tmax, double& t,
                                         int32_t& prim, tstack& st, int sp, bool media, uint64_t mkey,
                                         uint32_t serial, wcount& wc, bool any_hit) {
    if constexpr (!(FEAT & (F_XFORM | F_LIST | F_MESH | F_MEDIUM))) {
        // primitives only: the object is its primitive
        prim = oi;
        return prim_t<CNT, FEAT>(S, oi, o, d, tmin, tmax, t, wc);
    }
    const obj_rec r = S.obj[oi];
    if ((FEAT & F_MEDIUM) && r.kind == MPT_OBJ_MEDIUM && !media) return false;
    const double world_len = (FEAT & F_MEDIUM) && r.kind == MPT_OBJ_MEDIUM ? length(d) : 0.0;
    if constexpr (FEAT & F_XFORM) {
        if (r.xform >= 0) {
            const xform& x = S.xf[r.xform];
            v3 lo = xf_inv_point(x, o);
            v3 ld = xf_inv_vector(x, d);
            o = lo;
            d = ld;
        }
    }
    if (r.kind == MPT_OBJ_PRIM) {
        prim = int32_t(r.first);
        return prim_t<CNT, FEAT>(S, r.first, o, d, tmin, tmax, t, wc);
    }
    if constexpr (FEAT & F_LIST)
        if (r.kind == MPT_OBJ_LIST) return list_t<CNT, FEAT>(S, r.first, r.count, o, d, tmin, tmax, t, prim, wc);
    if constexpr (FEAT & F_UV)  // voxel scenes run the variant that carries (u, v)
        if (r.kind == MPT_OBJ_VOXEL)
            return voxel_t<CNT, FEAT>(S, r.blas_root, o, d, tmin, tmax, t, prim, st, sp, wc, any_hit);
    if constexpr (FEAT & F_MESH)
        if (r.kind == MPT_OBJ_MESH)
            return mesh_t<CNT, FEAT>(S, r.blas_root, o, d, tmin, tmax, t, prim, st, sp, wc, r.alpha, r.first, mkey,
                                     serial, any_hit);
    if constexpr (FEAT & F_MEDIUM) {
        prim = -1;
        return medium_t<CNT, FEAT>(S, r, o, d, world_len, tmin, tmax, medium_draw(mkey, serial, uint32_t(oi)), t, wc);
    }
    return false;
}

// ground_plane::hit (plane.h:17-34), t only; the interval is exclusive. Only
// the kernels of scenes with rich shading (F_UV) carry it.
template <int FEAT>
__device__ __forceinline__ bool ground_t(const dscene& S, const v3& o, const v3& d, double tmin, double tmax,
                                         hitinfo& h) {
    if constexpr (FEAT & F_UV) {
        if (S.ground_mat < 0) return false;
        const double dy = d.y;
        if (std::fabs(dy) < 1e-12) return false;
        const double t = (S.ground_h - o.y) / dy;
        if (!(tmin < t && t < tmax)) return false;
        h.t = t;
        h.obj = S.nobj;
        h.prim = -1;
        return true;
    }
    return false;
}

// height_fog::hit (medium.h:116-146): the free flight of draw u along the
// ray, inverting the optical depth, which is exponential in t.
__device__ __forceinline__ bool fog_t(const fog_rec& f, const v3& o, const v3& d, double tmin, double tmax, double u,
                                      double& t) {
    const double len = length(d);
    const double k = d.y / f.scale;
    const double t0 = tmin;
    const double start = f.density * len * std::exp(-(o.y + t0 * d.y - f.base) / f.scale);
    if (!(start > 0)) return false;
    const double tau = -std::log(1 - u);
    double tt;
    if (std::fabs(k) < 1e-12) {
        tt = t0 + tau / start;
    } else {
        const double x = 1 - tau * k / start;
        if (x <= 0) return false;  // leaves upwards before the collision
        tt = t0 - std::log(x) / k;
    }
    if (!(tt < tmax) || !std::isfinite(tt)) return false;
    t = tt;
    return true;
}

// bvh::hit over the world with the meshes' own bvh::hit (bvh.h:45-86,
// instance.h, mesh.h) in one loop: a wave's lanes visit records of the world
// tree and of any mesh tree in the same step, where a walk nested in the
// world leaf's test would run each mesh tree only for the lanes that reached
// it together. Each tree is walked as walk4 walks it, in the same order: a
// world leaf's items are taken one per step; a mesh item enters the mesh's
// tree (its ray in instance space, its stack part above the world's) and the
// walk returns to the leaf's next item when that part is empty. A mesh's
// hits shrink the interval while its tree is walked, as in the nested walk;
// at its end the world leaf's double test (made at its first hit) accepts
// them or restores the interval and closes the leaf. tmax: in, the
// interval's end; out, the closest accepted hit's t.
// 0 builds the nested walk for scenes with meshes instead (for comparison).
#ifndef MPT_FLAT_WALK
#define MPT_FLAT_WALK 1
#endif
template <bool CNT, int FEAT>
__device__ __forceinline__ bool walk_world(const dscene& S, const v3& o, const v3& d, double tmin, double& tmax,
                                           hitinfo& h, tstack& st, bool media, uint64_t mkey, uint32_t serial,
                                           wcount& wc, bool any_hit) {
    if constexpr (CNT) wc.nodes += 64u;
    const croot& root = S.tlas_root;
    fray fr;
    int cur = -1;
    // the open world leaf: items [l_off + l_k, l_off + l_cnt), its child-pair
    // record l_e (-1: the root), the interval's end when it was entered and
    // its double test (-1: not made yet)
    int l_off = 0, l_cnt = 0, l_k = 0, l_e = -1, l_ok = -1;
    double l_tleaf = tmax;
    {
        const v3 inv{1 / d.x, 1 / d.y, 1 / d.z};
        if (!box_hit(root.b, o, inv, tmin, tmax)) return false;
        fr = make_fray(o, inv, tmin, root.b);
        if (root.cnt > 0) {
            l_off = root.off;
            l_cnt = root.cnt;
        } else {
            cur = root.qoff;
        }
    }
    int negmask = neg_mask(fr);
    bool any = false;
    // the mesh whose tree is being walked: object, stack base, instance-space
    // ray, hit so far, the interval's end when it was entered
    bool inmesh = false;
    int m_oi = 0, m_alpha = -1, bsp = 0;
    int64_t m_first = 0;
    int32_t m_prim = -1;
    bool m_any = false;
    double m_tw = tmax;
    v3 lo = o, ld = d;
    int sp = 0;
    // a mesh's hit goes to the world leaf's test; true: the walk stops
    auto mesh_end = [&]() -> bool {
        if (!m_any) return false;
        if (l_ok < 0) l_ok = leaf_exact<CNT>(S.tlas, l_e, o, d, tmin, l_tleaf, wc) ? 1 : 0;
        if (!l_ok) {
            tmax = m_tw;
            l_k = l_cnt;
            return false;
        }
        any = true;
        h.t = tmax;
        h.obj = m_oi;
        h.prim = m_prim;
        return any_hit;
    };
    while (true) {
        if (!inmesh && l_k < l_cnt) {  // the open world leaf's next item
            const int oi = S.titems[l_off + l_k];
            l_k++;
            const obj_rec r = S.obj[oi];
            if (r.kind == MPT_OBJ_MESH) {
                if constexpr (CNT) wc.nodes += 64u;
                v3 mo = o, md = d;
                if constexpr (FEAT & F_XFORM) {
                    if (r.xform >= 0) {
                        const xform& x = S.xf[r.xform];
                        mo = xf_inv_point(x, o);
                        md = xf_inv_vector(x, d);
                    }
                }
                const croot br = S.blas_roots[r.blas_root];
                const v3 minv{1 / md.x, 1 / md.y, 1 / md.z};
                if (!box_hit(br.b, mo, minv, tmin, tmax)) continue;
                m_oi = oi;
                m_alpha = r.alpha;
                m_first = r.first;
                m_prim = -1;
                m_any = false;
                m_tw = tmax;
                lo = mo;
                ld = md;
                if (br.cnt > 0) {  // a one-leaf tree
                    (void)mesh_leaf<CNT, FEAT>(S, br.off, br.cnt, -1, lo, ld, tmin, tmax, m_prim, m_any, m_alpha,
                                               m_first, mkey, serial, any_hit, wc);
                    if (mesh_end()) return true;
                    continue;
                }
                fr = make_fray(lo, minv, tmin, br.b);
                negmask = neg_mask(fr);
                cur = br.qoff;
                bsp = sp;
                inmesh = true;
                continue;
            }
            double tt;
            int32_t pr = -1;  // a mesh never gets here: the object test is compiled without them
            if (object_t<CNT, FEAT & ~F_MESH>(S, oi, o, d, tmin, tmax, tt, pr, st, sp, media, mkey, serial, wc,
                                              any_hit)) {
                if (l_ok < 0) l_ok = leaf_exact<CNT>(S.tlas, l_e, o, d, tmin, l_tleaf, wc) ? 1 : 0;
                if (!l_ok) {
                    l_k = l_cnt;
                    continue;
                }
                any = true;
                tmax = tt;
                h.t = tt;
                h.obj = oi;
                h.prim = pr;
                if (any_hit) return true;
            }
            continue;
        }
        // a record visit, then pops until an interior record (walk4's step);
        // ev: 0 walking on, 1 a world leaf was opened (its items come next),
        // 2 this tree's stack part is empty, 3 an occlusion query's first hit
        // in the mesh
        const qnode* qn = inmesh ? S.qblas : S.qtlas;
        int next = -1, ev = 0;
        auto take = [&](int32_t off, int32_t meta) {
            const int cnt = int(uint32_t(meta) >> 28);
            if (cnt == 0) {
                next = off;
            } else if (inmesh) {
                if (mesh_leaf<CNT, FEAT>(S, off, cnt, meta & 0x0fffffff, lo, ld, tmin, tmax, m_prim, m_any, m_alpha,
                                         m_first, mkey, serial, any_hit, wc))
                    ev = 3;
            } else {
                l_off = off;
                l_cnt = cnt;
                l_k = 0;
                l_e = meta & 0x0fffffff;
                l_ok = -1;
                l_tleaf = tmax;
                ev = 1;
            }
        };
        if (cur >= 0) {
            int32_t off, meta;
            if (visit_record<CNT>(qn, cur, fr, negmask, tmax, st, sp, off, meta, wc)) take(off, meta);
        }
        while (next < 0 && ev == 0) {
            if (sp == (inmesh ? bsp : 0)) {
                ev = 2;
                break;
            }
            int32_t off, meta;
            if (pop_slot<CNT>(qn, st.get(--sp), fr, tmax, off, meta, wc)) take(off, meta);
        }
        cur = next;
        if (ev >= 2) {
            if (!inmesh) break;  // the world's walk is done
            // the mesh's walk is done (or stopped at an occlusion query's
            // hit): back to the world leaf's items
            sp = bsp;
            inmesh = false;
            if (mesh_end()) return true;
            const v3 inv{1 / d.x, 1 / d.y, 1 / d.z};
            fr = make_fray(o, inv, tmin, root.b);
            negmask = neg_mask(fr);
        }
    }
    return any;
}

// bvh::hit over the world (bvh.h:45-86), then the unbounded objects beside
// it in order (main.cpp:70-74): the height fogs, the ground. media: test
// constant_medium objects and fogs with draws from (mkey, serial); otherwise
// they are skipped. any_hit: an occlusion query, which only asks whether a
// hit exists: the walk is the closest-hit walk up to its first accepted hit
// and stops there (one walk for both, the flag uniform across the wave). Every test the closest-hit walk makes before that hit is the
// same, and a hit it would accept later exists only if one was accepted
// first (intervals only shrink on hits; media, fogs and masks draw by key,
// not by order), so the answer is the closest-hit walk's; h is that first
// hit, not the closest.
template <bool CNT = false, int FEAT = F_ALL>
__device__ bool trace_closest(const dscene& S, const v3& o, const v3& d, double tmin, double tmax, hitinfo& h,
                              tstack& st, bool media, uint64_t mkey, uint32_t serial, wcount& wc,
                              bool any_hit = false) {
    h.obj = -1;
    h.prim = -1;
    h.t = tmax;
    bool any = false;
    if constexpr ((FEAT & F_MESH) && MPT_FLAT_WALK) {
        if (S.nworld > 0) any = walk_world<CNT, FEAT>(S, o, d, tmin, tmax, h, st, media, mkey, serial, wc, any_hit);
    } else if (S.nworld > 0) walk4<CNT>(S.qtlas, S.tlas_root, o, d, tmin, tmax, st, 0, wc, [&](int off, int cnt, int sp, int e) {
        const double tleaf = tmax;
        int ok = -1;  // the leaf's double test, made at its first hit
        for (int k = 0; k < cnt; k++) {
            int oi = S.titems[off + k];
            double tt;
            int32_t pr = -1;
            if (object_t<CNT, FEAT>(S, oi, o, d, tmin, tmax, tt, pr, st, sp, media, mkey, serial, wc, any_hit)) {
                if (ok < 0) ok = leaf_exact<CNT>(S.tlas, e, o, d, tmin, tleaf, wc) ? 1 : 0;
                if (!ok) break;
                any = true;
                tmax = tt;
                h.t = tt;
                h.obj = oi;
                h.prim = pr;
                if (any_hit) return true;
            }
        }
        return false;
    });
    if (any_hit && any) return true;
    // the objects beside the BVH, after it with the closest t so far
    // (main.cpp:70-74); only the variant that carries (u, v) has them
    if constexpr (FEAT & F_UV) {
        if (media)
            for (int k = 0; k < S.nfog; k++) {
                double tf;
                if (fog_t(S.fog[k], o, d, tmin, tmax, medium_draw(mkey, serial, uint32_t(S.nobj + 1 + k)), tf)) {
                    tmax = tf, h.t = tf, h.obj = S.nobj + 1 + k, h.prim = -1, any = true;
                }
            }
        if (ground_t<FEAT>(S, o, d, tmin, tmax, h)) any = true;
    }
    return any;
}

// Full hit record of a closest hit: point, face normal, front_face,
// material, as the primitive's hit() and instance::hit write them.
struct hitrec {
    v3 p;
    v3 normal;
    double t;
    int32_t mat;
    bool front;
    // texture coordinates and how p moves with them (hittable.h:14-24); set
    // by finish_hit when FEAT has F_UV, else zero
    double u = 0, v = 0;
    v3 dpdu{0, 0, 0}, dpdv{0, 0, 0};
};

// tmin: the interval's start the hit was found with (a voxel hit's face is
// found again by walking its block up to the hit).
template <int FEAT = F_ALL>
__device__ __forceinline__ void finish_hit(const dscene& S, const hitinfo& h, const v3& wo, const v3& wd, hitrec& rec,
                                           double tmin = 0.001) {
    if constexpr (FEAT & F_UV) {
        if (h.obj == S.nobj) {  // the ground (plane.h:23-33)
            v3 p = wo + h.t * wd;
            p.y = S.ground_h;  // exactly on the plane
            rec.t = h.t;
            rec.p = p;
            rec.u = p.x;
            rec.v = -p.z;
            rec.dpdu = v3{1, 0, 0};
            rec.dpdv = v3{0, 0, -1};
            rec.mat = S.ground_mat;
            const v3 up{0, 1, 0};
            rec.front = dot(wd, up) < 0;
            rec.normal = rec.front ? up : -up;
            return;
        }
        if (h.obj > S.nobj) {  // a height fog (medium.h:139-145)
            rec.t = h.t;
            rec.p = wo + h.t * wd;
            rec.normal = v3{1, 0, 0};
            rec.front = true;
            rec.mat = S.fog[h.obj - S.nobj - 1].material;
            return;
        }
    }
    const obj_rec r = S.obj[h.obj];
    rec.t = h.t;
    if (r.kind == MPT_OBJ_MEDIUM) {
        rec.p = wo + h.t * wd;
        rec.normal = v3{1, 0, 0};
        rec.front = true;
        rec.mat = r.material;
        return;
    }
    v3 o = wo, d = wd;
    const xform* x = r.xform >= 0 ? &S.xf[r.xform] : nullptr;
    if (x) {
        o = xf_inv_point(*x, wo);
        d = xf_inv_vector(*x, wd);
    }
    if constexpr (FEAT & F_UV) {
        if (r.kind == MPT_OBJ_VOXEL) {  // record (voxel.h:324-335) of the face the block's walk stops on
            const double* vq = prim_ptr(S, h.prim);
            vox_hit f{0, 1, 0, 0, 0, 0};
            double tt;
            vox_walk<true>(S, vq, o, d, tmin, h.t, tt, &f);
            const mpt_vox_model& m = S.vox[int(vq[6])];
            v3 p = o + h.t * d;
            const int b = (f.axis + 1) % 3, c = (f.axis + 2) % 3;
            const v3 outward{f.axis == 0 ? double(f.sign) : 0.0, f.axis == 1 ? double(f.sign) : 0.0,
                             f.axis == 2 ? double(f.sign) : 0.0};
            const v3 tb{b == 0 ? 1.0 : 0.0, b == 1 ? 1.0 : 0.0, b == 2 ? 1.0 : 0.0};
            const v3 tc{c == 0 ? 1.0 : 0.0, c == 1 ? 1.0 : 0.0, c == 2 ? 1.0 : 0.0};
            rec.dpdu = f.sign > 0 ? tb : tc;  // cross(dpdu, dpdv) = the outward normal
            rec.dpdv = f.sign > 0 ? tc : tb;
            const double pu = dot(p, rec.dpdu), pv = dot(p, rec.dpdv);
            rec.u = pu - std::floor(pu);
            rec.v = pv - std::floor(pv);
            rec.mat = m.edge_width > 0 && vox_on_edge(S, m, p, f) ? m.edge_material : m.material[f.index];
            rec.front = dot(d, outward) < 0;
            v3 n = rec.front ? outward : -outward;
            if (x) {  // instance.h:33-42
                p = xf_point(*x, p);
                n = unit_vector(xf_normal(*x, n));
                rec.dpdu = xf_vector(*x, rec.dpdu);
                rec.dpdv = xf_vector(*x, rec.dpdv);
            }
            rec.p = p;
            rec.normal = n;
            return;
        }
    }
    const double* q = prim_ptr(S, h.prim);
    int k = S.pkind[h.prim];
    v3 p = o + h.t * d;
    v3 outward;
    if (k == MPT_PRIM_SPHERE) {
        outward = (p - v3{q[0], q[1], q[2]}) / q[3];
    } else {
        outward = v3{q[9], q[10], q[11]};
    }
    rec.front = dot(d, outward) < 0;
    v3 n = rec.front ? outward : -outward;
    if constexpr (FEAT & F_UV) {
        if (k == MPT_PRIM_SPHERE) {  // sphere.h:46-48, 99-120, on the unit sphere's point
            const double r = q[3];
            const v3 un = outward * (r < 0 ? -1.0 : 1.0);
            const double theta = std::acos(std::fmin(std::fmax(-un.y, -1.0), 1.0));
            const double phi = std::atan2(-un.z, un.x) + kPi;
            rec.u = phi / (2 * kPi);
            rec.v = theta / kPi;
            const double ar = std::fabs(r);
            const double sin_theta = std::sqrt(un.x * un.x + un.z * un.z);
            if (sin_theta < 1e-8) {
                rec.dpdu = rec.dpdv = v3{0, 0, 0};
            } else {
                rec.dpdu = (2 * kPi * ar) * v3{un.z, 0, -un.x};
                rec.dpdv = (kPi * ar / sin_theta) * v3{-un.x * un.y, sin_theta * sin_theta, -un.y * un.z};
            }
        } else if (k == MPT_PRIM_QUAD) {  // quad.h:41-57
            const v3 Q{q[0], q[1], q[2]}, qu{q[3], q[4], q[5]}, qv{q[6], q[7], q[8]}, w{q[12], q[13], q[14]};
            const v3 ph = p - Q;
            rec.u = dot(w, cross(ph, qv));
            rec.v = dot(w, cross(qu, ph));
            rec.dpdu = qu;
            rec.dpdv = qv;
        } else {  // mesh.h:71-93, 149-166: barycentrics by the hit test's arithmetic
            const v3 p0{q[0], q[1], q[2]}, e1{q[3], q[4], q[5]}, e2{q[6], q[7], q[8]};
            const v3 pvec = cross(d, e2);
            const double inv_det = 1 / dot(e1, pvec);
            const v3 tvec = o - p0;
            const double b1 = dot(tvec, pvec) * inv_det;
            const double b2 = dot(d, cross(tvec, e1)) * inv_det;
            const double* t6 = S.vuv ? S.vuv + 6 * h.prim : nullptr;
            if (t6 && !std::isnan(t6[0])) {
                const double b0 = 1 - b1 - b2;
                rec.u = b0 * t6[0] + b1 * t6[2] + b2 * t6[4];
                rec.v = b0 * t6[1] + b1 * t6[3] + b2 * t6[5];
                const double du1 = t6[2] - t6[0], dv1 = t6[3] - t6[1];
                const double du2 = t6[4] - t6[0], dv2 = t6[5] - t6[1];
                const double det = du1 * dv2 - du2 * dv1;
                if (std::fabs(det) < 1e-12) {
                    rec.dpdu = rec.dpdv = v3{0, 0, 0};
                } else {
                    const double inv = 1 / det;
                    rec.dpdu = (dv2 * e1 - dv1 * e2) * inv;
                    rec.dpdv = (du1 * e2 - du2 * e1) * inv;
                }
            } else {
                rec.u = b1;
                rec.v = b2;
                rec.dpdu = e1;
                rec.dpdv = e2;
            }
        }
        if (x) {  // instance.h:38-39
            rec.dpdu = xf_vector(*x, rec.dpdu);
            rec.dpdv = xf_vector(*x, rec.dpdv);
        }
    }
    if ((FEAT & F_TRI) && S.vnrm && k == MPT_PRIM_TRIANGLE && !std::isnan(S.vnrm[9 * h.prim])) {
        // interpolated shading normal on the geometric normal's side (mesh.h:93-101);
        // the barycentrics by the arithmetic of the hit test
        const double* vn = S.vnrm + 9 * h.prim;
        const v3 p0{q[0], q[1], q[2]}, e1{q[3], q[4], q[5]}, e2{q[6], q[7], q[8]};
        const v3 pvec = cross(d, e2);
        const double inv_det = 1 / dot(e1, pvec);
        const v3 tvec = o - p0;
        const double b1 = dot(tvec, pvec) * inv_det;
        const double b2 = dot(d, cross(tvec, e1)) * inv_det;
        const double b0 = 1 - b1 - b2;
        v3 sh = b0 * v3{vn[0], vn[1], vn[2]} + b1 * v3{vn[3], vn[4], vn[5]} + b2 * v3{vn[6], vn[7], vn[8]};
        if (length_squared(sh) > 0) {
            sh = unit_vector(sh);
            n = dot(sh, n) < 0 ? -sh : sh;
        }
    }
    if (x) {
        p = xf_point(*x, p);
        n = unit_vector(xf_normal(*x, n));
    }
    rec.p = p;
    rec.normal = n;
    rec.mat = S.pmat[h.prim];
}

}  // namespace mpt
