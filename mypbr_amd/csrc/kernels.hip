// kernels.hip - wavefront path-tracing stages for gfx950 and the C ABI that
// drives them (include/mpt.h).
//
// One render is a sequence of batches; a batch traces `passes_in_flight`
// samples of every pixel at once. Per bounce the stages are
//   extend  closest-hit of every live path's ray (LDS short-stack BVH walk)
//   shade   miss/emission/MIS, BSDF scatter, NEE light sample, Russian
//           roulette; survivors and shadow rays are compacted into queues
//           with one atomic per wave (ballot + popcount prefix)
//   shadow  closest-hit of the shadow rays, emission found, MIS-weighted
//           light contribution added to the path
// and every batch ends with accumulate (per pixel, samples in pass order).
// This is ray_color (camera.h:356-461) cut at its traversal calls: each
// stage does, for every path of the batch, the part of one loop iteration
// between two world.hit() calls, in the same order.
//
// All stage kernels are persistent: a fixed grid sized to the device walks
// the queue with a wave-uniform stride, reading the queue length from HBM,
// so the per-bounce launches need no host round trip and are captured once
// into a hipGraph replayed for every bounce.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mpt.h"
#include "dev_shade.hpp"
#include "scene_host.hpp"

namespace mpt {

thread_local std::string g_last_error;

// the radiance cache's network (nrc.hip)
int nrc_infer(mpt_nrc* r, const float* pos, const float* other, int64_t n, float* pred, hipStream_t s);
int nrc_train(mpt_nrc* r, const float* pos, const float* other, const float* target, int64_t n, float lr,
              hipStream_t s, double* loss);

int set_error(int code, const std::string& m) {
    g_last_error = m;
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return set_error(MPT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

constexpr int kLdsDepth = 16;

// ------------------------------------------------------------------ camera
struct dcam {
    int w, h, spp, max_depth;
    int max_diffuse, max_specular, max_scatter, transparent_shadows;
    int sqrt_spp, blades, sky_mode, features;  // sky_mode: 0 env, 1 gradient, 2 constant
    int sun_disk_hidden, pad_;                 // camera rays do not see the sun itself (camera.h:46)
    double recip_sqrt_spp;
    v3 center, pixel00, du, dv, w_axis, disk_u, disk_v, background;
    double ortho_height, defocus_angle, focus_dist, blade_rotation;
};

// camera::initialize (camera.h:142-179)
static dcam init_camera(const mpt_camera& c, bool has_env) {
    dcam k{};
    k.w = c.image_width;
    int h = int(c.image_width / c.aspect_ratio);
    k.h = h < 1 ? 1 : h;
    k.spp = c.samples_per_pixel;
    k.max_depth = c.max_depth;
    k.max_diffuse = c.max_diffuse;
    k.max_specular = c.max_specular;
    k.max_scatter = c.max_scatter;
    k.transparent_shadows = c.transparent_shadows;
    k.sqrt_spp = int(std::sqrt(c.samples_per_pixel));
    k.recip_sqrt_spp = 1.0 / k.sqrt_spp;
    k.blades = c.blades;
    k.features = c.collect_features;
    k.sky_mode = has_env ? 0 : (c.use_sky ? 1 : 2);
    k.sun_disk_hidden = c.sun_disk_hidden != 0;
    k.background = v3{c.background[0], c.background[1], c.background[2]};
    k.ortho_height = c.ortho_height;
    k.defocus_angle = c.defocus_angle;
    k.focus_dist = c.focus_dist;
    k.blade_rotation = c.blade_rotation;
    v3 lookfrom{c.lookfrom[0], c.lookfrom[1], c.lookfrom[2]};
    v3 lookat{c.lookat[0], c.lookat[1], c.lookat[2]};
    v3 vup{c.vup[0], c.vup[1], c.vup[2]};
    k.center = lookfrom;
    double theta = deg2rad(c.vfov);
    double hh = std::tan(theta / 2);
    double vh = c.ortho_height > 0 ? c.ortho_height : 2 * hh * c.focus_dist;
    double vw = vh * (double(k.w) / k.h);
    v3 w = unit_vector(lookfrom - lookat);
    v3 u = unit_vector(cross(vup, w));
    v3 v = cross(w, u);
    v3 vu = vw * u;
    v3 vv = vh * -v;
    k.du = vu / k.w;
    k.dv = vv / k.h;
    v3 ul = k.center - (c.focus_dist * w) - vu / 2 - vv / 2;
    k.pixel00 = ul + 0.5 * (k.du + k.dv);
    double dr = c.focus_dist * std::tan(deg2rad(c.defocus_angle / 2));
    k.disk_u = u * dr;
    k.disk_v = v * dr;
    k.w_axis = w;
    return k;
}

// ENV: the scene has an environment map (sky mode 0); without one only the
// gradient and constant skies are compiled in. The sun is added to every
// escaping ray, to camera rays (depth 0) only while its disk is visible.
template <bool ENV = true>
__device__ __forceinline__ v3 escaped(const dscene& S, const dcam& C, const v3& dir, bool camera_ray) {  // camera.h:515-529
    v3 sky;
    if (ENV && C.sky_mode == 0) {
        sky = env_radiance(S, dir);
    } else if (C.sky_mode == 2) {
        sky = C.background;
    } else {
        v3 ud = unit_vector(dir);
        double a = 0.5 * (ud.y + 1.0);
        sky = (1.0 - a) * v3{1.0, 1.0, 1.0} + a * v3{0.5, 0.7, 1.0};
    }
    if (S.has_sun && (!C.sun_disk_hidden || !camera_ray)) sky = sky + sun_radiance(S, dir);
    return sky;
}

__device__ __forceinline__ double power_heuristic(double a, double b) {  // camera.h:343-347
    double a2 = a * a;
    double b2 = b * b;
    return (a2 + b2) > 0 ? a2 / (a2 + b2) : 0;
}

// ------------------------------------------------------------------ path state in HBM
// The double-precision fields of a path, grouped into records of 8 doubles
// (64 bytes) by which stage reads them together. With MPT_PATH_RECORDS (the
// default) a slot's record is contiguous (array of records), so the scattered
// slots a queue names each cost whole 64-byte segments instead of one
// partly used cache line per field; with 0 every field is its own array.
enum {
    // record 0: the ray (extend reads it; shade rewrites it), and the path's
    // flags and random-stream counter (the bits of one double: flags << 32 | counter)
    D_OX, D_OY, D_OZ, D_DX, D_DY, D_DZ, D_SPDF, D_FLAGCTR,
    // record 1: throughput, radiance, MIS weight of the pending shadow ray
    D_BX, D_BY, D_BZ, D_RX, D_RY, D_RZ, D_SW, D_PAD1,
    // record 2: the hit record shade reads (point, face normal, t, and the
    // bits of the hit material with the front face in bit 31)
    D_QX, D_QY, D_QZ, D_MX, D_MY, D_MZ, D_HT, D_HMAT,
    // record 3: previous vertex (MIS) and the shadow ray's origin
    D_PX, D_PY, D_PZ, D_SOX, D_SOY, D_SOZ, D_PAD3A, D_PAD3B,
    // record 4: the NEE sample (unnormalised direction to the light, f)
    D_TLX, D_TLY, D_TLZ, D_FX, D_FY, D_FZ, D_PAD4A, D_PAD4B,
    // record 5: throughput at the NEE vertex
    D_SBX, D_SBY, D_SBZ, D_PAD5A, D_PAD5B, D_PAD5C, D_PAD5D, D_PAD5E,
    // record 6: first-surface guides
    D_AX, D_AY, D_AZ, D_NX, D_NY, D_NZ, D_PAD6A, D_PAD6B,
    // record 7: texture coordinates and dp/du, dp/dv of the hit (allocated and
    // written only in scenes whose materials read them)
    D_U, D_V, D_DUX, D_DUY, D_DUZ, D_DVX, D_DVY, D_DVZ,
    D_COUNT
};
constexpr int kUvRecord = 7;
#ifndef MPT_PATH_RECORDS
#define MPT_PATH_RECORDS 1
#endif
constexpr int kRecDoubles = 8;
constexpr int64_t kFieldStride = MPT_PATH_RECORDS ? kRecDoubles : 1;  // doubles between slots of a field

// Flags (in D_FLAGCTR's high word): bit0 specular, bit1 seen_by_shadow_ray,
// 10-bit bounce counters from bits 2, 12, 22.
// One allocation holds every field: with records, record r of slot i at
// base + (r * cap + i) * 8; without, field f of slot i at base + f * cap + i.
// (A pointer per field would cost the stage kernels 128 scalar registers of
// arguments.)
struct pstate {
    double* base;
    int64_t cap;
    uint64_t* key;
};

__device__ __forceinline__ double as_double(uint64_t b) { return __longlong_as_double(static_cast<long long>(b)); }
__device__ __forceinline__ uint64_t as_bits(double v) { return static_cast<uint64_t>(__double_as_longlong(v)); }

// Field f of slot i.
__device__ __forceinline__ double& pd(const pstate& P, int f, uint32_t i) {
    if constexpr (MPT_PATH_RECORDS) return P.base[(int64_t(f / kRecDoubles) * P.cap + i) * kRecDoubles + f % kRecDoubles];
    return P.base[int64_t(f) * P.cap + i];
}
__device__ __forceinline__ v3 ld3(const pstate& P, int f, uint32_t i) { return v3{pd(P, f, i), pd(P, f + 1, i), pd(P, f + 2, i)}; }
__device__ __forceinline__ void st3(const pstate& P, int f, uint32_t i, const v3& v) {
    pd(P, f, i) = v.x;
    pd(P, f + 1, i) = v.y;
    pd(P, f + 2, i) = v.z;
}

// All 8 doubles of record r of slot i, written together: a record written
// in part leaves its 64-byte segment partly written, which costs the memory
// a read-modify-write (writing only 3 of record 3's doubles per bounce cost
// 15% of the shade kernels' time).
__device__ __forceinline__ void st_rec(const pstate& P, int r, uint32_t i, double a0, double a1, double a2, double a3,
                                       double a4, double a5, double a6, double a7) {
    if constexpr (MPT_PATH_RECORDS) {
        double2* q = reinterpret_cast<double2*>(&pd(P, r * kRecDoubles, i));
        q[0] = double2{a0, a1};
        q[1] = double2{a2, a3};
        q[2] = double2{a4, a5};
        q[3] = double2{a6, a7};
    } else {
        const double a[8] = {a0, a1, a2, a3, a4, a5, a6, a7};
        for (int k = 0; k < kRecDoubles; k++) pd(P, r * kRecDoubles + k, i) = a[k];
    }
}
__device__ __forceinline__ void st_rec(const pstate& P, int r, uint32_t i, const v3& a, const v3& b, double c,
                                       double d) {
    st_rec(P, r, i, a.x, a.y, a.z, b.x, b.y, b.z, c, d);
}

// A block's records r of its consecutive slots [s0, s0 + BS) staged through
// LDS: each thread fills its record there, then the block moves them to HBM
// as contiguous 16-byte chunks (slots beyond n are left alone).
template <int BS>
__device__ __forceinline__ void store_records(const pstate& P, int r, int64_t s0, int64_t n, double* lds) {
    __syncthreads();
    double2* dst = reinterpret_cast<double2*>(&pd(P, r * kRecDoubles, uint32_t(s0)));
    const double2* src = reinterpret_cast<const double2*>(lds);
    const int64_t slots = n - s0 < BS ? n - s0 : BS;
    for (int64_t c = threadIdx.x; c < slots * (kRecDoubles / 2); c += BS) dst[c] = src[c];
    __syncthreads();
}

struct launch_ctx {
    int32_t* ovf;
    int64_t nthreads;
};

// Shading classes: 0 = missed everything, 1 + t = a hit on material type t.
constexpr int kClasses = 7;
// Counters, each on its own 256-byte line so that atomics on different
// counters do not serialise on one cache line: counter k is word
// k * kCtrStride of the counter array. 0, 1: ping-pong queue lengths; 2:
// shadow queue; 3 .. 3 + kClasses - 1: class queues; kNrcQueue,
// kNrcTrain: the radiance cache's query queue and training paths. The
// 64-bit ray and work statistics follow at word kCtrWords.
constexpr int kNrcQueue = 10;
constexpr int kNrcTrain = 11;
constexpr int kCtrStride = 64;
constexpr int kCtrWords = 16 * kCtrStride;
__host__ __device__ __forceinline__ uint32_t* ctr(uint32_t* base, int k) { return base + k * kCtrStride; }

// Block-aggregated append to NQ queues: bit k of `bits` asks for `value` to
// be appended to queue k (counter ctr_of(k), entries q_of(k)). One atomic per
// queue per call for the whole block; lane offsets come from per-wave ballot
// popcounts and a prefix over the block's waves in LDS. Every thread of the
// block must make the call (block-uniform control flow).
template <int BS, int NQ, typename CtrOf, typename QOf>
__device__ __forceinline__ void block_append(uint32_t* lds, uint32_t bits, uint32_t value, CtrOf ctr_of, QOf q_of) {
    constexpr int W = BS / 64;
    const int wid = int(threadIdx.x) >> 6;
    const uint32_t lane = __lane_id();
    uint64_t mask[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        mask[k] = __ballot((bits >> k) & 1u);
        if (lane == 0) lds[k * W + wid] = uint32_t(__popcll(mask[k]));
    }
    __syncthreads();
    if (int(threadIdx.x) < NQ) {
        const int k = int(threadIdx.x);
        uint32_t run = 0;
        for (int w = 0; w < W; w++) {
            const uint32_t c = lds[k * W + w];
            lds[k * W + w] = run;
            run += c;
        }
        lds[NQ * W + k] = run ? atomicAdd(ctr_of(k), run) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        if ((bits >> k) & 1u) {
            const uint32_t below =
                __builtin_amdgcn_mbcnt_hi(uint32_t(mask[k] >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask[k]), 0u));
            q_of(k)[lds[NQ * W + k] + lds[k * W + wid] + below] = value;
        }
    }
    __syncthreads();
}
// geometry profile of scenes made of spheres and quads, in lists and instances
constexpr int kQuadProfile = F_SPHERE | F_QUAD | F_LIST | F_XFORM;

// Where the radiance cache cuts paths (depth < 0: never): at the hit of
// bounce `depth` every path is appended to q (counter cnt); only the
// training paths go on to be shaded.
struct nrc_route {
    int depth;
    int cut;          // 0: every path goes on (the cache only trains)
    uint32_t stride;
    uint32_t salt;
    uint32_t* q;
    uint32_t* cnt;
};

// Several scenes in one batch (mpt_render_batch): every stage launch serves
// the paths of all of them. Scene k owns slots [base[k], base[k + 1]) of the
// batch (its passes of this batch, pixel-major within a pass, as in a
// one-scene batch) and pixels [pbase[k], pbase[k + 1]) of the accumulation
// launch; a path's scene and camera are looked up from its slot.
constexpr int kMaxBatch = 16;
struct mbatch {
    const dscene* S;      // device arrays of n scenes and cameras
    const dcam* C;
    int32_t n;
    int32_t npix[kMaxBatch];
    uint32_t base[kMaxBatch + 1];
    uint32_t pbase[kMaxBatch + 1];
    double* acc[kMaxBatch];  // each scene's running sums (radiance, albedo, normal, l2)
    __device__ __forceinline__ int scene_of(uint32_t slot) const {
        int k = 0;
#pragma unroll
        for (int j = 1; j < kMaxBatch; j++) k += (j < n && slot >= base[j]) ? 1 : 0;
        return k;
    }
    __device__ __forceinline__ int scene_of_pixel(uint32_t g) const {
        int k = 0;
#pragma unroll
        for (int j = 1; j < kMaxBatch; j++) k += (j < n && g >= pbase[j]) ? 1 : 0;
        return k;
    }
};

// The scene and camera tables read through the constant address space: a
// load at a wave-uniform index then goes to scalar registers, as the fields
// of a kernel argument do.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(4))) const dscene const_dscene;
typedef __attribute__((address_space(4))) const dcam const_dcam;
#else  // the host pass only parses device code; address spaces are device-side
typedef const dscene const_dscene;
typedef const dcam const_dcam;
#endif

// f(S, C) for the scene of this lane's slot, run for the wave's scenes one
// at a time (usually one: scenes own contiguous slot ranges) with the scene
// index uniform across the lanes that run together. The index is taken from
// the lowest waiting lane (ballot, readlane) and the scene loaded before the
// lane test, so the compiler cannot substitute the lane's own index.
template <typename F>
__device__ __forceinline__ void with_scene(const mbatch& M, uint32_t slot, F&& f) {
    const int k = M.scene_of(slot);
    bool pending = true;
    while (true) {
        const uint64_t m = __ballot(pending);
        if (!m) break;
        const int k0 = __builtin_amdgcn_readlane(k, __builtin_ctzll(m));
        const dscene S = ((const_dscene*)M.S)[k0];
        const dcam C = ((const_dcam*)M.C)[k0];
        if (pending && k == k0) {
            f(S, C, k0);
            pending = false;
        }
    }
}

#define STACK_SETUP()                                                                   \
    __shared__ int32_t lds_stack[kLdsDepth * BS];                                       \
    tstack st;                                                                          \
    st.lds = lds_stack;                                                                 \
    st.ovf = L.ovf;                                                                     \
    st.tid = threadIdx.x;                                                               \
    st.bs = BS;                                                                         \
    st.gtid = int64_t(blockIdx.x) * BS + threadIdx.x;                                   \
    st.nthreads = L.nthreads;                                                           \
    st.lds_depth = kLdsDepth;

// ------------------------------------------------------------------ ray query
template <int BS, int FEAT>
__global__ __launch_bounds__(BS) void k_query(dscene S, launch_ctx L, const double* rays, const double* tt, int64_t n,
                                              int64_t* out_id, double* out_t, uint8_t* out_front) {
    STACK_SETUP();
    for (int64_t i = st.gtid; i < n; i += L.nthreads) {
        v3 o{rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
        v3 d{rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
        hitinfo h;
        wcount wc{0, 0};
        bool hit = trace_closest<false, FEAT>(S, o, d, tt[2 * i], tt[2 * i + 1], h, st, false, 0, 0, wc);
        if (hit) {
            hitrec rec;
            finish_hit<FEAT>(S, h, o, d, rec, tt[2 * i]);
            out_id[i] = h.obj == S.nobj ? S.ground_id : S.obj[h.obj].id_base + S.pcanon[h.prim];
            out_t[i] = h.t;
            out_front[i] = rec.front ? 1 : 0;
        } else {
            out_id[i] = -1;
            out_t[i] = 0;
            out_front[i] = 0;
        }
    }
}

// Grazing ties of closest hits found by k_query: every primitive is tested
// alone over the ray's whole interval (its object's transform applied, media
// left out, the ground plane last), and those whose t is bitwise equal to
// the closest hit's are counted with their smallest and largest canonical
// ID. One ray per thread over all primitives: a diagnostic for query
// batches, O(rays x primitives).
__global__ __launch_bounds__(256) void k_ties(dscene S, const double* rays, const double* tt, int64_t n,
                                              const int64_t* id, const double* tc, int32_t* ties, int64_t* tie_min,
                                              int64_t* tie_max) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        int32_t count = 0;
        int64_t lo = -1, hi = -1;
        if (id[i] >= 0) {
            const v3 o{rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
            const v3 d{rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
            const double tmin = tt[2 * i], tmax = tt[2 * i + 1];
            const uint64_t best = __double_as_longlong(tc[i]);
            auto take = [&](int64_t c) {
                lo = count == 0 ? c : (c < lo ? c : lo);
                hi = count == 0 ? c : (c > hi ? c : hi);
                count++;
            };
            for (int oi = 0; oi < S.nobj; oi++) {
                const obj_rec r = S.obj[oi];
                if (r.kind == MPT_OBJ_MEDIUM || (r.flags & MPT_OBJF_LIGHT_ONLY)) continue;
                v3 lo_o = o, lo_d = d;
                if (r.xform >= 0) {
                    lo_o = xf_inv_point(S.xf[r.xform], o);
                    lo_d = xf_inv_vector(S.xf[r.xform], d);
                }
                const int64_t cnt = r.kind == MPT_OBJ_PRIM ? 1 : r.count;
                for (int64_t k = 0; k < cnt; k++) {
                    const int64_t p = r.first + k;
                    double t;
                    wcount wc{0, 0};
                    const bool hit = r.kind == MPT_OBJ_VOXEL
                                         ? vox_walk<false>(S, prim_ptr(S, p), lo_o, lo_d, tmin, tmax, t, nullptr)
                                         : prim_t<false, F_ALL>(S, p, lo_o, lo_d, tmin, tmax, t, wc);
                    if (hit &&
                        uint64_t(__double_as_longlong(t)) == best)
                        take(r.id_base + S.pcanon[p]);
                }
            }
            hitinfo g;
            if (ground_t<F_ALL | F_UV>(S, o, d, tmin, tmax, g) && uint64_t(__double_as_longlong(g.t)) == best)
                take(S.ground_id);
        }
        ties[i] = count;
        tie_min[i] = lo;
        tie_max[i] = hi;
    }
}

// Depth of the first surface seen through each pixel centre (no jitter or
// defocus, the ray camera::get_ray makes for offset (0, 0) without the lens,
// camera.h:291-301): t * |d| of the closest hit, 0 for the sky.
template <int BS, int FEAT>
__global__ __launch_bounds__(BS) void k_depth(dscene S, launch_ctx L, v3 pixel00, v3 du, v3 dv, v3 center, v3 w_axis,
                                              double ortho, double focus, int w, int h, double* depth) {
    STACK_SETUP();
    for (int64_t k = st.gtid; k < int64_t(w) * h; k += L.nthreads) {
        const int i = int(k % w), j = int(k / w);
        const v3 ps = pixel00 + (double(i) * du) + (double(j) * dv);
        const v3 org = ortho > 0 ? ps + focus * w_axis : center;
        const v3 d = ps - org;
        hitinfo hi;
        wcount wc{0, 0};
        const bool hit = trace_closest<false, FEAT>(S, org, d, 0.001, kInf, hi, st, false, 0, 0, wc);
        depth[k] = hit ? hi.t * std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z) : 0.0;
    }
}

// camera::average (camera.h:223-246) of the running sums on the device.
// guide_samples: the passes whose albedo and normal guides are in the sums.
__global__ void k_average(const double* __restrict__ acc, int64_t npix, int samples, int guide_samples, int features,
                          double* radiance, double* albedo, double* normal, double* variance) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= npix) return;
    const double scale = 1.0 / samples;
    const double gscale = guide_samples == samples ? scale : (guide_samples > 0 ? 1.0 / guide_samples : 0.0);
    const double* ar = acc;
    const double* aa = ar + 3 * npix;
    const double* an = aa + 3 * npix;
    const double* al = an + 3 * npix;
    const double r0 = scale * ar[3 * k], r1 = scale * ar[3 * k + 1], r2 = scale * ar[3 * k + 2];
    if (radiance) radiance[3 * k] = r0, radiance[3 * k + 1] = r1, radiance[3 * k + 2] = r2;
    if (!features) return;
    if (albedo)
        for (int c = 0; c < 3; c++) albedo[3 * k + c] = gscale * aa[3 * k + c];
    if (normal)
        for (int c = 0; c < 3; c++) normal[3 * k + c] = gscale * an[3 * k + c];
    if (variance) {
        const double mean = luminance(v3{r0, r1, r2});
        const double spread = std::fmax(scale * al[k] - mean * mean, 0.0);
        variance[k] = samples > 1 ? spread / (samples - 1) : spread;
    }
}

// ------------------------------------------------------------------ camera rays
struct batch_args {
    uint64_t seed;
    uint32_t scene_index;
    int32_t first_pass;   // pass of slot 0
    int32_t npix;
    int32_t nslots;       // passes in this batch * npix
};

__device__ __forceinline__ v3 random_in_unit_disk(rstream& rs) {  // vec3.h:139-145
    while (true) {
        double x = rs.next(-1, 1);
        double y = rs.next(-1, 1);
        v3 p{x, y, 0};
        if (length_squared(p) < 1) return p;
    }
}

// The camera ray of slot `slot` (camera.h:286-302 get_ray), the `local`-th
// path of its scene's part of the batch; sets the slot's random stream,
// flags and first-queue entry.
__device__ __forceinline__ void camera_ray(const dcam& C, const pstate& P, const batch_args& B, uint32_t slot,
                                           uint32_t local, uint32_t* q0, v3& org_out, v3& dir_out, double& flagctr) {
    int pass = B.first_pass + int(local / B.npix);
    int pix = int(local % B.npix);
    int i = pix % C.w, j = pix / C.w;
    rstream rs{path_key(B.seed, B.scene_index, uint32_t(pass), uint32_t(pix)), 0};
    int strata = C.sqrt_spp * C.sqrt_spp;
    double ox, oy;
    if (pass >= strata) {  // camera.h:304-306
        ox = rs.next() - 0.5;
        oy = rs.next() - 0.5;
    } else {  // camera.h:205-206, 310-314
        int cell = (pass + i * 7919 + j * 104729) % strata;
        int si = cell % C.sqrt_spp, sj = cell / C.sqrt_spp;
        ox = ((si + rs.next()) * C.recip_sqrt_spp) - 0.5;
        oy = ((sj + rs.next()) * C.recip_sqrt_spp) - 0.5;
    }
    // camera.h:291-301
    v3 ps = C.pixel00 + ((i + ox) * C.du) + ((j + oy) * C.dv);
    v3 eye = C.ortho_height > 0 ? ps + C.focus_dist * C.w_axis : C.center;
    v3 org = eye;
    if (C.defocus_angle > 0) {
        v3 p;
        if (C.blades >= 3) {  // camera.h:329-339
            int k = int(rs.next() * C.blades);
            k = k < C.blades - 1 ? k : C.blades - 1;
            double a0 = deg2rad(C.blade_rotation) + 2 * kPi * k / C.blades;
            double a1 = a0 + 2 * kPi / C.blades;
            double s = rs.next(), t = rs.next();
            if (s + t > 1) {
                s = 1 - s;
                t = 1 - t;
            }
            p = s * v3{std::cos(a0), std::sin(a0), 0} + t * v3{std::cos(a1), std::sin(a1), 0};
        } else {
            p = random_in_unit_disk(rs);
        }
        org = eye + ((p.x * C.disk_u) + (p.y * C.disk_v));
    }
    org_out = org;
    dir_out = ps - org;
    P.key[slot] = rs.key;
    flagctr = as_double((uint64_t(1u) << 32) | rs.ctr);  // specular: camera rays see lights unweighted
    q0[slot] = slot;
}

template <bool MULTI = false>
__global__ __launch_bounds__(256) void k_camera(dcam C, pstate P, batch_args B, uint32_t* q0, uint32_t* cnt,
                                                int lights, mbatch M) {
    __shared__ double lds[256 * kRecDoubles];
    int64_t gt = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gt == 0) {
        *ctr(cnt, 0) = uint32_t(B.nslots);
        for (int k = 1; k < 3 + kClasses; k++) *ctr(cnt, k) = 0;
        *ctr(cnt, kNrcQueue) = 0;  // the radiance cache's query and training counts
        *ctr(cnt, kNrcTrain) = 0;
    }
    if constexpr (MPT_PATH_RECORDS) {  // records are written block-wide (store_records); every thread takes part
        const int64_t s0 = int64_t(blockIdx.x) * blockDim.x;
        v3 org{0, 0, 0}, dir{0, 0, 0};
        double fc = 0;
        if (gt < B.nslots) {
            if constexpr (MULTI) {  // slot gt is path gt - base[k] of scene k, whose key has scene index + k
                with_scene(M, uint32_t(gt), [&](const dscene&, const dcam& CC, int k) {
                    batch_args Bk = B;
                    Bk.scene_index = B.scene_index + uint32_t(k);
                    Bk.npix = M.npix[k];
                    camera_ray(CC, P, Bk, uint32_t(gt), uint32_t(gt) - M.base[k], q0, org, dir, fc);
                });
            } else {
                camera_ray(C, P, B, uint32_t(gt), uint32_t(gt), q0, org, dir, fc);
            }
        }
        double* my = lds + threadIdx.x * kRecDoubles;
        // record 0: ray, scatter pdf, flags and counter
        my[0] = org.x, my[1] = org.y, my[2] = org.z, my[3] = dir.x, my[4] = dir.y, my[5] = dir.z, my[6] = 0, my[7] = fc;
        if (s0 < B.nslots) store_records<256>(P, 0, s0, B.nslots, lds);
        // record 1: beta 1, radiance 0
        my[0] = 1, my[1] = 1, my[2] = 1, my[3] = 0, my[4] = 0, my[5] = 0, my[6] = 0, my[7] = 0;
        if (s0 < B.nslots) store_records<256>(P, 1, s0, B.nslots, lds);
        // record 3: previous vertex 0 (read only by next-event estimation's MIS weights)
        for (int k = 0; k < kRecDoubles; k++) my[k] = 0;
        if (lights && s0 < B.nslots) store_records<256>(P, 3, s0, B.nslots, lds);
        return;
    }
    if (gt >= B.nslots) return;
    const uint32_t slot = uint32_t(gt);
    v3 org, dir;
    double fc;
    camera_ray(C, P, B, slot, slot, q0, org, dir, fc);
    st3(P, D_OX, slot, org);
    st3(P, D_DX, slot, dir);
    pd(P, D_FLAGCTR, slot) = fc;
    st3(P, D_BX, slot, v3{1, 1, 1});
    st3(P, D_RX, slot, v3{0, 0, 0});
    if (lights) st3(P, D_PX, slot, v3{0, 0, 0});
    pd(P, D_SPDF, slot) = 0;
}

// camera::first_surface (camera.h:253-277) on the feature stream.
// FIRST: the camera rays were extended already and their hit records (or the
// miss mark) are in record 2, so the walk starts from them instead of tracing
// the camera ray again; valid where a ray's closest hit does not depend on
// the stream that traces it (no media, fogs or alpha masks).
template <int BS, int FEAT, bool FIRST = false>
__global__ __launch_bounds__(BS) void k_features(dscene S, launch_ctx L, pstate P, batch_args B) {
    STACK_SETUP();
    constexpr bool kUV = (FEAT & F_UV) != 0;
    for (int64_t s = st.gtid; s < B.nslots; s += L.nthreads) {
        uint32_t slot = uint32_t(s);
        v3 o = ld3(P, D_OX, slot), d = ld3(P, D_DX, slot);
        rstream rs{feature_key(P.key[slot]), 0};
        v3 tint{1, 1, 1};
        v3 albedo = tint, normal{0, 0, 0};
        bool done = false;
        for (int bounce = 0; bounce < 8 && !done; bounce++) {
            hitrec rec;
            if (FIRST && bounce == 0) {  // the extension's record of the camera ray's closest hit
                rec.t = pd(P, D_HT, slot);
                if (rec.t < 0) break;
                rec.p = ld3(P, D_QX, slot);
                rec.normal = ld3(P, D_MX, slot);
                const uint32_t hm = uint32_t(as_bits(pd(P, D_HMAT, slot)));
                rec.mat = int32_t(hm & 0x7fffffffu);
                rec.front = (hm >> 31) != 0;
                if (kUV && S.mat[rec.mat].uses_uv) {
                    rec.u = pd(P, D_U, slot);
                    rec.v = pd(P, D_V, slot);
                    rec.dpdu = ld3(P, D_DUX, slot);
                    rec.dpdv = ld3(P, D_DVX, slot);
                }
            } else {
                hitinfo h;
                wcount wc{0, 0};
                if (!trace_closest<false, FEAT>(S, o, d, 0.001, kInf, h, st, true, rs.key, uint32_t(bounce), wc)) break;
                finish_hit<FEAT>(S, h, o, d, rec);
            }
            const mpt_material& m = S.mat[rec.mat];
            tint = tint * mat_interior(m, d, rec);
            hitrec shaded = rec;  // the guide keeps the hit's own normal
            if (kUV && m.normal_map >= 0) shaded.normal = normal_from_map(S, m.normal_map, d, rec);
            scatter_out so = mat_scatter<-1, kUV>(S, m, d, shaded, rs, &rec.normal);
            if (kUV && so.skip && m.normal_map >= 0) so.skip_dir = unbend(so.skip_dir, shaded.normal, rec.normal);
            if (!so.scattered) {
                albedo = tint;
                normal = rec.normal;
                done = true;
            } else if (!so.skip) {
                albedo = tint * so.attenuation;
                normal = rec.normal;
                done = true;
            } else {
                tint = tint * so.attenuation;
                o = rec.p;
                d = so.skip_dir;
            }
        }
        if (!done) {
            albedo = tint;
            normal = v3{0, 0, 0};
        }
        st_rec(P, 6, slot, albedo, normal, 0.0, 0.0);
    }
}

// ------------------------------------------------------------------ extend
// Closest hit of every queued path, then binning by what shades the hit:
// class 0 is a miss, class 1 + t a surface or medium of material type t.
// Each class gets its own queue, appended with one atomic per wave.
// CNT: also count BVH work (nodes, primitive bytes) into work[0], work[1].
// Waves per SIMD the spheres-only traversal is compiled for (its register
// budget). With the single-precision culls the kernel needs about 90 VGPRs
// without spills: 5 waves (96 VGPRs) gave 1587 M paths/s on the 1080p
// Weekend cover, 6 waves (80 VGPRs, 13 spilled) 1575 M and 7 waves 1457 M.
#ifndef MPT_SPHERE_WAVES
#define MPT_SPHERE_WAVES 5
#endif
// Waves per SIMD of the general-geometry traversal kernels (meshes,
// instances, media). Uncapped they take about 200 VGPRs (255 with an
// environment in the shadow kernel) and run 1-2 waves per SIMD. With the
// world and mesh trees walked in one loop, the batch of 8 generated scenes
// ran 24.0 / 28.4 / 31.8 / 18.8 / 13.0 M paths/s at 2 / 3 / 4 / 5 / 6 waves
// (from 5 on, spills take over).
#ifndef MPT_GENERAL_WAVES
#define MPT_GENERAL_WAVES 4
#endif

// Closest hit of path p's ray and the hit record shade reads (point, face
// normal, material); returns the path's shading class.
template <bool CNT, int FEAT>
__device__ __forceinline__ int extend_one(const dscene& S, const pstate& P, int depth, uint32_t p, tstack& st,
                                          wcount& wc) {
    const v3 o = ld3(P, D_OX, p), d = ld3(P, D_DX, p);
    hitinfo h;
    trace_closest<CNT, FEAT>(S, o, d, 0.001, kInf, h, st, true, P.key[p], uint32_t(2 * depth), wc);
    if (h.obj < 0) {
        if (depth == 0) pd(P, D_HT, p) = -1.0;  // the camera ray's miss, for the guides' walk
        return 0;
    }
    // the hit record, every double of it (t first: it need not stay live through finish_hit)
    pd(P, D_HT, p) = h.t;
    hitrec rec;
    finish_hit<FEAT>(S, h, o, d, rec);
    st3(P, D_QX, p, rec.p);
    st3(P, D_MX, p, rec.normal);
    pd(P, D_HMAT, p) = as_double(uint32_t(rec.mat) | (rec.front ? 0x80000000u : 0u));
    const mpt_material& m = S.mat[rec.mat];
    if constexpr (FEAT & F_UV)
        if (m.uses_uv) st_rec(P, kUvRecord, p, rec.u, rec.v, rec.dpdu.x, rec.dpdu.y, rec.dpdu.z, rec.dpdv.x, rec.dpdv.y,
                              rec.dpdv.z);
    return 1 + m.type;
}

// Paths a batch traces to the end to train the radiance cache: one in
// `stride`, chosen by a hash of the slot and a per-batch salt.
__device__ __forceinline__ bool nrc_training(uint32_t slot, const nrc_route& R) {
    uint32_t h = (slot ^ R.salt) * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    return h % R.stride == 0;
}

template <int BS, bool CNT, int FEAT, bool MULTI = false>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(FEAT == F_SPHERE ? MPT_SPHERE_WAVES : (FEAT == F_ALL ? MPT_GENERAL_WAVES : 1)))) void k_extend(dscene S, launch_ctx L, pstate P, int depth, const uint32_t* q,
                                               const uint32_t* cnt_in, uint32_t* cnt_next, uint32_t* cnt_shadow,
                                               uint32_t* qcls, uint32_t* ccls, int64_t cap,
                                               unsigned long long* ray_stats, unsigned long long* work,
                                               nrc_route R, mbatch M) {
    STACK_SETUP();
    const uint32_t n = *cnt_in;
    if (st.gtid == 0) {  // the counters shade and shadow append to in this bounce
        *cnt_next = 0;
        *cnt_shadow = 0;
        atomicAdd(&ray_stats[0], (unsigned long long)n);
    }
    wcount wc{0, 0};
    // blocks walk the queue in block-sized steps; the class queues (and the
    // radiance cache's query queue) are appended once per block and step
    __shared__ uint32_t lds_app[(kClasses + 1) * (BS / 64 + 1)];
    const bool query = depth == R.depth;
    for (int64_t base = int64_t(blockIdx.x) * BS; base < n; base += int64_t(gridDim.x) * BS) {
        const int64_t i = base + threadIdx.x;
        int cls = -1;
        uint32_t p = 0;
        if (i < n) {
            p = q[i];
            if constexpr (MULTI)
                with_scene(M, p, [&](const dscene& SS, const dcam&, int) { cls = extend_one<CNT, FEAT>(SS, P, depth, p, st, wc); });
            else cls = extend_one<CNT, FEAT>(S, P, depth, p, st, wc);
        }
        uint32_t bits = cls >= 0 ? 1u << cls : 0u;
        if (query && cls > 0) {  // a hit at the cache's bounce: queried; shaded only if it trains
            bits |= 1u << kClasses;
            if (R.cut && !nrc_training(p, R)) bits = 1u << kClasses;
        }
        block_append<BS, kClasses + 1>(
            lds_app, bits, p, [&](int k) { return k < kClasses ? ccls + k * kCtrStride : R.cnt; },
            [&](int k) { return k < kClasses ? qcls + k * cap : R.q; });
    }
    if constexpr (CNT) {
        if (wc.nodes) atomicAdd(&work[0], (unsigned long long)wc.nodes);
        if (wc.pbytes) atomicAdd(&work[1], (unsigned long long)wc.pbytes);
    }
}

// ------------------------------------------------------------------ shade
// One kernel per shading class. MT is the material type (-1: the paths that
// missed); LIGHTS: the scene has lights, so non-specular vertices do next
// event estimation. Compiling each material apart keeps each kernel's
// register footprint to what that material needs.
template <int MT, int LK, bool ENV, bool RICH = false>
__device__ __forceinline__ void shade_one(const dscene& S, const dcam& C, const pstate& P, int depth, uint32_t p,
                                          bool& alive, bool& nee) {
    constexpr bool LIGHTS = LK != 0;
    v3 o = ld3(P, D_OX, p), d = ld3(P, D_DX, p);
    v3 beta = ld3(P, D_BX, p), rad = ld3(P, D_RX, p);
    // written at the end as whole records: the shadow ray's weight and the
    // scatter pdf (next-event estimation's MIS), and whether this vertex is
    // the previous vertex / shadow-ray origin of what follows
    double sw = 0, spdf_out = 0;
    bool vertex = false;
    v3 vertex_p{0, 0, 0};
    const uint64_t fc = as_bits(pd(P, D_FLAGCTR, p));
    const uint32_t fl = uint32_t(fc >> 32);
    uint32_t ctr_out = uint32_t(fc);
    bool specular = fl & 1u, seen = fl & 2u;
    int dcount = (fl >> 2) & 1023, scount = (fl >> 12) & 1023, vcount = (fl >> 22) & 1023;
    if constexpr (MT < 0) {  // camera.h:373-387
        v3 sky = escaped<ENV>(S, C, d, depth == 0);
        if (!seen) {
            if (!LIGHTS || specular) {
                rad = rad + beta * sky;
            } else {
                double lp = lights_pdf<LK>(S, ld3(P, D_PX, p), d);
                rad = rad + beta * sky * power_heuristic(pd(P, D_SPDF, p), lp);
            }
        }
        st_rec(P, 1, p, beta, rad, 0.0, 0.0);
        return;
    } else {
        hitrec rec;
        rec.t = pd(P, D_HT, p);
        rec.p = ld3(P, D_QX, p);
        rec.normal = ld3(P, D_MX, p);
        const uint32_t hm = uint32_t(as_bits(pd(P, D_HMAT, p)));
        rec.mat = int32_t(hm & 0x7fffffffu);
        rec.front = (hm >> 31) != 0;
        const mpt_material& m = S.mat[rec.mat];
        if (RICH && m.uses_uv) {
            rec.u = pd(P, D_U, p);
            rec.v = pd(P, D_V, p);
            rec.dpdu = ld3(P, D_DUX, p);
            rec.dpdv = ld3(P, D_DVX, p);
        }
        // the hit's own normal; a normal map bends the one the material shades with
        const v3 gn = rec.normal;
        if (RICH && m.normal_map >= 0) rec.normal = normal_from_map(S, m.normal_map, d, rec);
        if constexpr (MT == MPT_MAT_DIELECTRIC) beta = beta * mat_interior<MT>(m, d, rec);  // camera.h:396
        if constexpr (RICH && MT != MPT_MAT_DIFFUSE_LIGHT) {
            // emissive_surface::emitted (material.h:731-736) added as camera.h:399-407 adds
            // any emission; the surface then scatters as its inner material
            if (m.glow && rec.front && !seen) {
                const v3 e = tex_value(S, m.glow_texture, rec);
                if (e.x > 0 || e.y > 0 || e.z > 0) {
                    if (!LIGHTS || specular) {
                        rad = rad + beta * e;
                    } else {
                        double lp = lights_pdf<LK>(S, ld3(P, D_PX, p), d);
                        rad = rad + beta * e * power_heuristic(pd(P, D_SPDF, p), lp);
                    }
                }
            }
        }
        if constexpr (MT == MPT_MAT_DIFFUSE_LIGHT) {                                          // camera.h:399-407
            v3 e = mat_emitted<MT>(m, d, rec);
            if (!seen && (e.x > 0 || e.y > 0 || e.z > 0)) {
                if (!LIGHTS || specular) {
                    rad = rad + beta * e;
                } else {
                    double lp = lights_pdf<LK>(S, ld3(P, D_PX, p), d);
                    rad = rad + beta * e * power_heuristic(pd(P, D_SPDF, p), lp);
                }
            }
            st_rec(P, 1, p, beta, rad, 0.0, 0.0);
            return;  // a light scatters nothing (material.h:28-31)
        } else {
            rstream rs{P.key[p], ctr_out};
            scatter_out so = mat_scatter<MT, RICH>(S, m, d, rec, rs, &gn);  // camera.h:409-411
            if (RICH && so.skip && m.normal_map >= 0) so.skip_dir = unbend(so.skip_dir, rec.normal, gn);
            if (so.skip) {                                        // camera.h:413-423
                bool stop = C.max_specular > 0 && ++scount > C.max_specular;
                if (!stop) {
                    beta = beta * so.attenuation;
                    if (LIGHTS && C.transparent_shadows && mat_lets_shadows(m) && depth > 0 && !specular) seen = true;
                    o = rec.p;
                    d = so.skip_dir;
                    specular = true;
                    alive = true;
                }
            } else {  // camera.h:424-448
                constexpr bool volume = MT == MPT_MAT_PHASE;
                const int cap = volume ? C.max_scatter : C.max_diffuse;
                bool capped = false;
                if (cap > 0) capped = ++(volume ? vcount : dcount) >= cap;
                seen = false;
                if constexpr (LIGHTS) {  // sample_light, camera.h:466-506, up to the shadow ray
                    const bool last_bounce = depth == C.max_depth - 1 || capped;
                    v3 to_light = lights_random<LK>(S, rec.p, rs);
                    double light_pdf = lights_pdf<LK>(S, rec.p, to_light);
                    if (light_pdf > 0) {
                        v3 f = mat_eval<MT, RICH>(S, m, d, rec, gn, to_light);
                        if (!(f.x <= 0 && f.y <= 0 && f.z <= 0)) {
                            double weight =
                                last_bounce ? 1.0 : power_heuristic(light_pdf, density_value<MT>(so.dens, to_light));
                            st_rec(P, 4, p, to_light, f, 0.0, 0.0);
                            st_rec(P, 5, p, beta, v3{0, 0, 0}, 0.0, 0.0);
                            sw = weight / light_pdf;
                            vertex = true;
                            vertex_p = rec.p;
                            nee = true;
                        }
                    }
                }
                if (!capped) {
                    v3 dir = density_generate<MT>(so.dens, rs);
                    double spdf = density_value<MT>(so.dens, dir);
                    if (spdf > 0) {
                        beta = beta * mat_eval<MT, RICH>(S, m, d, rec, gn, dir) / spdf;
                        if constexpr (LIGHTS) {  // read only by the MIS weights of next-event estimation
                            spdf_out = spdf;
                            vertex = true;
                            vertex_p = rec.p;
                        }
                        o = rec.p;
                        d = dir;
                        specular = false;
                        alive = true;
                    }
                }
            }
            if (alive && depth >= 3) {  // camera.h:452-457
                double survive = std::fmin(std::fmax(beta.x, std::fmax(beta.y, beta.z)), 0.95);
                if (rs.next() >= survive) alive = false;
                else beta = beta / survive;
            }
            ctr_out = rs.ctr;
        }
    }
    if (depth + 1 >= C.max_depth) alive = false;
    // record 3: this vertex as the next one's previous vertex (MIS) and as
    // the shadow ray's origin; both are the hit point
    if (LIGHTS && vertex) st_rec(P, 3, p, vertex_p, vertex_p, 0.0, 0.0);
    st_rec(P, 1, p, beta, rad, sw, 0.0);  // a path that ends here keeps only its radiance (and a pending shadow ray)
    if (!alive) return;
    const uint32_t flags = (specular ? 1u : 0u) | (seen ? 2u : 0u) | (uint32_t(dcount) << 2) |
                           (uint32_t(scount) << 12) | (uint32_t(vcount) << 22);
    st_rec(P, 0, p, o, d, spdf_out, as_double((uint64_t(flags) << 32) | ctr_out));
}

// Waves per SIMD the shade kernels are compiled for (1: no register cap).
#ifndef MPT_SHADE_WAVES
#define MPT_SHADE_WAVES 1
#endif

template <int BS, int MT, int LK, bool ENV, bool RICH = false, bool MULTI = false>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(MPT_SHADE_WAVES))) void k_shade(dscene S, dcam C, pstate P, int depth, const uint32_t* q,
                                              const uint32_t* cnt_in, uint32_t* qnext, uint32_t* cnt_next,
                                              uint32_t* qshadow, uint32_t* cnt_shadow, mbatch M) {
    constexpr int NQ = LK != 0 ? 2 : 1;
    __shared__ uint32_t lds[NQ * (BS / 64 + 1)];
    const uint32_t n = *cnt_in;
    for (int64_t base = int64_t(blockIdx.x) * BS; base < n; base += int64_t(gridDim.x) * BS) {
        const int64_t i = base + threadIdx.x;
        bool alive = false, nee = false;
        uint32_t p = 0;
        if (i < n) {
            p = q[i];
            if constexpr (MULTI) {
                with_scene(M, p, [&](const dscene& SS, const dcam& CC, int) {
                    shade_one<MT, LK, ENV, RICH>(SS, CC, P, depth, p, alive, nee);
                });
            } else {
                shade_one<MT, LK, ENV, RICH>(S, C, P, depth, p, alive, nee);
            }
        }
        block_append<BS, NQ>(
            lds, (alive ? 1u : 0u) | (nee ? 2u : 0u), p, [&](int k) { return k ? cnt_shadow : cnt_next; },
            [&](int k) { return k ? qshadow : qnext; });
    }
}

// ------------------------------------------------------------------ shadow
// The shadow ray of path p's NEE sample (sample_light, camera.h:480-506):
// closest hit along the normalized direction, through surfaces that let
// shadows through when transparent shadows are on, emission found (or the
// sky), and the MIS-weighted contribution added to the path's radiance.
// OCC: compiled for scenes whose shadow rays are occlusion queries only (no
// emitter, glow or shadow-passing material), so the kernel carries neither
// the hit record nor the materials of the general loop.
template <bool CNT, int FEAT, bool ENV, bool OCC = false>
__device__ __forceinline__ void shadow_one(const dscene& S, const dcam& C, const pstate& P, int depth, uint32_t p,
                                           tstack& st, wcount& wc) {
    {
        v3 o = ld3(P, D_SOX, p);
        v3 to_light = ld3(P, D_TLX, p);
        v3 dir = unit_vector(to_light);
        v3 through{1, 1, 1};
        v3 emission;
        // nothing emits and nothing lets shadows through: whatever the ray hits
        // first gives no emission, so only whether it hits anything matters
        if constexpr (OCC) {
            hitinfo h;
            emission = trace_closest<CNT, FEAT>(S, o, dir, 0.001, kInf, h, st, true, P.key[p], uint32_t(2 * depth + 1),
                                                wc, true)
                           ? v3{0, 0, 0}
                           : escaped<ENV>(S, C, to_light, false);
        } else {
            const bool occlusion = S.shadow_occlusion != 0;
            for (int k = 0;; k++) {  // camera.h:488-502
                hitinfo h;
                const uint32_t serial = uint32_t(2 * depth + 1) | (uint32_t(k) << 20);
                if (!trace_closest<CNT, FEAT>(S, o, dir, 0.001, kInf, h, st, true, P.key[p], serial, wc, occlusion)) {
                    emission = escaped<ENV>(S, C, to_light, false);
                    break;
                }
                if (occlusion) {
                    emission = v3{0, 0, 0};
                    break;
                }
                hitrec rec;
                finish_hit<FEAT>(S, h, o, dir, rec);
                const mpt_material& m = S.mat[rec.mat];
                if (C.transparent_shadows && k < 16 && mat_lets_shadows(m)) {
                    through = through * mat_interior(m, dir, rec);
                    o = rec.p;
                    continue;
                }
                emission = mat_emitted(m, dir, rec);
                if constexpr ((FEAT & F_UV) != 0)  // a glowing surface in the way gives its glow (camera.h:500)
                    if (m.glow && rec.front) emission = emission + tex_value(S, m.glow_texture, rec);
                break;
            }
        }
        emission = through * emission;
        const double sw = pd(P, D_SW, p);
        v3 c = ld3(P, D_FX, p) * emission * sw;
        v3 rad = ld3(P, D_RX, p) + ld3(P, D_SBX, p) * c;
        st_rec(P, 1, p, ld3(P, D_BX, p), rad, sw, 0.0);
    }
}

template <int BS, bool CNT, int FEAT, bool ENV, bool MULTI = false, bool OCC = false>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(FEAT == F_ALL ? MPT_GENERAL_WAVES : 1))) void k_shadow(dscene S, dcam C, launch_ctx L, pstate P, int depth,
                                               const uint32_t* qshadow, const uint32_t* cnt_shadow, uint32_t* ccls,
                                               unsigned long long* ray_stats, unsigned long long* work, mbatch M) {
    STACK_SETUP();
    const uint32_t n = *cnt_shadow;
    if (st.gtid == 0) {
        atomicAdd(&ray_stats[1], (unsigned long long)n);
        for (int c = 0; c < kClasses; c++) ccls[c * kCtrStride] = 0;  // every shade kernel of this bounce has read them
    }
    wcount wc{0, 0};
    for (int64_t i = st.gtid; i < n; i += L.nthreads) {
        const uint32_t p = qshadow[i];
        if constexpr (MULTI) {
            with_scene(M, p, [&](const dscene& SS, const dcam& CC, int) {
                shadow_one<CNT, FEAT, ENV, OCC>(SS, CC, P, depth, p, st, wc);
            });
        } else {
            shadow_one<CNT, FEAT, ENV, OCC>(S, C, P, depth, p, st, wc);
        }
    }
    if constexpr (CNT) {
        if (wc.nodes) atomicAdd(&work[2], (unsigned long long)wc.nodes);
        if (wc.pbytes) atomicAdd(&work[3], (unsigned long long)wc.pbytes);
    }
}

// ------------------------------------------------------------------ finish
// The rest of a batch in one launch, for its tail: once few paths are left,
// per-bounce stage launches cost more than the work in them, so each
// remaining path runs all its bounces in one thread. A bounce is what the
// stage kernels do for the path, in the same order (extend, shade by
// material, shadow ray), so the results are the same.
// Waves per SIMD k_finish is compiled for: it holds every material's code,
// and without a cap its register footprint allows one wave per SIMD. The
// general-geometry variant is capped at 4: the batch of 8 generated scenes
// ran 32.2 / 33.2 / 33.6 M paths/s at 2 / 3 / 4 waves.
#ifndef MPT_FINISH_WAVES
#define MPT_FINISH_WAVES 2
#endif
#ifndef MPT_GENERAL_FINISH_WAVES
#define MPT_GENERAL_FINISH_WAVES 4
#endif

// Every remaining bounce of path p from depth0 on (one thread).
template <int FEAT, int LK, bool ENV>
__device__ __forceinline__ void finish_path(const dscene& S, const dcam& C, const pstate& P, int depth0, uint32_t p,
                                            tstack& st, wcount& wc, unsigned long long& ext,
                                            unsigned long long& shd) {
    constexpr bool kUV = (FEAT & F_UV) != 0;
    for (int depth = depth0; depth < C.max_depth; depth++) {
        const int cls = extend_one<false, FEAT>(S, P, depth, p, st, wc);
        ext++;
        bool alive = false, nee = false;
        switch (cls) {
            case 0: shade_one<-1, LK, ENV>(S, C, P, depth, p, alive, nee); break;
            case 1 + MPT_MAT_LAMBERTIAN:
                shade_one<MPT_MAT_LAMBERTIAN, LK, false, kUV>(S, C, P, depth, p, alive, nee);
                break;
            case 1 + MPT_MAT_METAL: shade_one<MPT_MAT_METAL, LK, false, kUV>(S, C, P, depth, p, alive, nee); break;
            case 1 + MPT_MAT_DIELECTRIC:
                shade_one<MPT_MAT_DIELECTRIC, LK, false, kUV>(S, C, P, depth, p, alive, nee);
                break;
            case 1 + MPT_MAT_PRINCIPLED:
                shade_one<MPT_MAT_PRINCIPLED, LK, false, kUV>(S, C, P, depth, p, alive, nee);
                break;
            case 1 + MPT_MAT_DIFFUSE_LIGHT:
                shade_one<MPT_MAT_DIFFUSE_LIGHT, LK, false>(S, C, P, depth, p, alive, nee);
                break;
            default: shade_one<MPT_MAT_PHASE, LK, false>(S, C, P, depth, p, alive, nee); break;
        }
        if constexpr (LK != 0) {
            if (nee) {
                shadow_one<false, FEAT, ENV>(S, C, P, depth, p, st, wc);
                shd++;
            }
        }
        if (!alive) break;
    }
}

template <int BS, int FEAT, int LK, bool ENV, bool MULTI = false>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(FEAT == F_ALL ? MPT_GENERAL_FINISH_WAVES : MPT_FINISH_WAVES))) void k_finish(
    dscene S, dcam C, launch_ctx L, pstate P, int depth0, const uint32_t* q, const uint32_t* cnt_in,
    unsigned long long* ray_stats, mbatch M) {
    STACK_SETUP();
    const uint32_t n = *cnt_in;
    unsigned long long ext = 0, shd = 0;
    wcount wc{0, 0};
    for (int64_t i = st.gtid; i < n; i += L.nthreads) {
        const uint32_t p = q[i];
        if constexpr (MULTI) {
            with_scene(M, p, [&](const dscene& SS, const dcam& CC, int) {
                finish_path<FEAT, LK, ENV>(SS, CC, P, depth0, p, st, wc, ext, shd);
            });
        } else {
            finish_path<FEAT, LK, ENV>(S, C, P, depth0, p, st, wc, ext, shd);
        }
    }
    // one atomic per wave for the ray counts
    for (int off = 32; off > 0; off >>= 1) {
        ext += __shfl_xor(ext, off);
        shd += __shfl_xor(shd, off);
    }
    if (__lane_id() == 0) {
        if (ext) {
            atomicAdd(&ray_stats[0], ext);
            atomicAdd(&ray_stats[6], ext);
        }
        if (shd) {
            atomicAdd(&ray_stats[1], shd);
            atomicAdd(&ray_stats[7], shd);
        }
    }
}

// ------------------------------------------------------------------ radiance cache
// Buffers of the cache's query and training samples (capacity: the path
// capacity for queries, tcap for training records).
struct nrc_bufs {
    float* pos;        // [cap][3] hit point in the cache's box, [0, 1]^3
    float* other;      // [cap][10] normal, direction towards the viewer, roughness, albedo
    float* pred;       // [cap][3] the cache's radiance
    uint32_t* tidx;    // [cap] query-queue index of each training path
    uint32_t* tcnt;    // their count
    uint32_t* tslot;   // [tcap] training records: path slot,
    double* trad;      //   radiance before the vertex,
    double* tbeta;     //   throughput at the vertex,
    float* tpos;       //   inputs,
    float* tother;
    float* ttarget;    //   and the path's radiance from the vertex on
    int64_t tcap;
    v3 box_min, inv_ext;
};

// The cache's inputs at every queried vertex; training paths' queue indices
// are appended to tidx.
template <int BS>
__global__ __launch_bounds__(BS) void k_nrc_prep(dscene S, pstate P, const uint32_t* q, const uint32_t* cnt,
                                                 nrc_route R, nrc_bufs B) {
    __shared__ uint32_t lds[BS / 64 + 1];
    const uint32_t n = *cnt;
    for (int64_t base = int64_t(blockIdx.x) * BS; base < n; base += int64_t(gridDim.x) * BS) {
        const int64_t i = base + threadIdx.x;
        bool train = false;
        if (i < n) {
            const uint32_t p = q[i];
            const v3 x = ld3(P, D_QX, p), nrm = ld3(P, D_MX, p), d = ld3(P, D_DX, p);
            const mpt_material& m = S.mat[uint32_t(as_bits(pd(P, D_HMAT, p))) & 0x7fffffffu];
            const v3 u = (x - B.box_min) * B.inv_ext;
            const v3 wo = -unit_vector(d);
            double rough = 1.0;
            v3 alb{1, 1, 1};
            const double tu = m.uses_uv ? pd(P, D_U, p) : 0.0, tv = m.uses_uv ? pd(P, D_V, p) : 0.0;
            switch (m.type) {
                case MPT_MAT_LAMBERTIAN: alb = tex_value(S, m.texture, tu, tv, x); break;
                case MPT_MAT_METAL: alb = v3{m.albedo[0], m.albedo[1], m.albedo[2]}, rough = m.roughness; break;
                case MPT_MAT_DIELECTRIC: rough = m.roughness; break;
                case MPT_MAT_PRINCIPLED: alb = tex_value(S, m.texture, tu, tv, x), rough = m.roughness; break;
                case MPT_MAT_PHASE: alb = v3{m.albedo[0], m.albedo[1], m.albedo[2]}; break;
                default: break;
            }
            float* ps = B.pos + 3 * i;
            ps[0] = float(u.x), ps[1] = float(u.y), ps[2] = float(u.z);
            float* o = B.other + 10 * i;
            o[0] = float(nrm.x), o[1] = float(nrm.y), o[2] = float(nrm.z);
            o[3] = float(wo.x), o[4] = float(wo.y), o[5] = float(wo.z);
            o[6] = float(rough);
            o[7] = float(alb.x), o[8] = float(alb.y), o[9] = float(alb.z);
            train = nrc_training(p, R);
        }
        block_append<BS, 1>(
            lds, train ? 1u : 0u, uint32_t(i), [&](int) { return B.tcnt; }, [&](int) { return B.tidx; });
    }
}

// Training records of the first tcap training paths: slot, radiance and throughput so far, inputs.
__global__ void k_nrc_record(pstate P, const uint32_t* q, nrc_bufs B) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = *B.tcnt < B.tcap ? int64_t(*B.tcnt) : B.tcap;
    if (j >= n) return;
    const uint32_t i = B.tidx[j];
    const uint32_t p = q[i];
    B.tslot[j] = p;
    const v3 r = ld3(P, D_RX, p), b = ld3(P, D_BX, p);
    B.trad[3 * j] = r.x, B.trad[3 * j + 1] = r.y, B.trad[3 * j + 2] = r.z;
    B.tbeta[3 * j] = b.x, B.tbeta[3 * j + 1] = b.y, B.tbeta[3 * j + 2] = b.z;
    for (int c = 0; c < 3; c++) B.tpos[3 * j + c] = B.pos[3 * int64_t(i) + c];
    for (int c = 0; c < 10; c++) B.tother[10 * j + c] = B.other[10 * int64_t(i) + c];
}

// Paths cut at the cache's bounce end here: radiance += beta * cache(x).
__global__ void k_nrc_apply(pstate P, const uint32_t* q, const uint32_t* cnt, nrc_route R, const float* pred) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= *cnt) return;
    const uint32_t p = q[i];
    if (!R.cut || nrc_training(p, R)) return;
    const v3 b = ld3(P, D_BX, p);
    const v3 c{std::fmax(double(pred[3 * i]), 0.0), std::fmax(double(pred[3 * i + 1]), 0.0),
               std::fmax(double(pred[3 * i + 2]), 0.0)};
    st_rec(P, 1, p, b, ld3(P, D_RX, p) + b * c, pd(P, D_SW, p), 0.0);
}

// Training targets once the batch is done: the radiance the path gathered
// from the vertex on, per unit of throughput there.
__global__ void k_nrc_collect(pstate P, nrc_bufs B) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = *B.tcnt < B.tcap ? int64_t(*B.tcnt) : B.tcap;
    if (j >= n) return;
    const v3 r = ld3(P, D_RX, B.tslot[j]);
    const double rr[3] = {r.x, r.y, r.z};
    for (int c = 0; c < 3; c++) {
        const double b = B.tbeta[3 * j + c];
        double t = b > 0 ? (rr[c] - B.trad[3 * j + c]) / b : 0.0;
        if (!std::isfinite(t)) t = 0.0;
        B.ttarget[3 * j + c] = float(t);
    }
}

// Zeroes n counters (stands in for the shadow stage when a scene has no lights).
__global__ void k_zero(uint32_t* c, int n) {
    if (int(threadIdx.x) < n) c[threadIdx.x * kCtrStride] = 0;
}

// ------------------------------------------------------------------ accumulate
// guide_passes: the batch's leading passes whose albedo and normal guides count.
__global__ void k_accumulate(pstate P, int npix, int passes, int features, int guide_passes, double* acc_rad,
                             double* acc_alb, double* acc_nrm, double* acc_l2) {
    int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    v3 a = v3{acc_rad[3 * pix], acc_rad[3 * pix + 1], acc_rad[3 * pix + 2]};
    v3 al{0, 0, 0}, nr{0, 0, 0};
    double l2 = 0;
    if (features) {
        al = v3{acc_alb[3 * pix], acc_alb[3 * pix + 1], acc_alb[3 * pix + 2]};
        nr = v3{acc_nrm[3 * pix], acc_nrm[3 * pix + 1], acc_nrm[3 * pix + 2]};
        l2 = acc_l2[pix];
    }
    for (int b = 0; b < passes; b++) {
        uint32_t s = uint32_t(b) * uint32_t(npix) + uint32_t(pix);
        v3 c = ld3(P, D_RX, s);
        if (!std::isfinite(c.x) || !std::isfinite(c.y) || !std::isfinite(c.z)) c = v3{0, 0, 0};  // camera.h:281-287
        a = a + c;
        if (features) {
            if (b < guide_passes) {
                al = al + ld3(P, D_AX, s);
                nr = nr + ld3(P, D_NX, s);
            }
            double l = luminance(c);
            l2 += l * l;
        }
    }
    acc_rad[3 * pix] = a.x, acc_rad[3 * pix + 1] = a.y, acc_rad[3 * pix + 2] = a.z;
    if (features) {
        acc_alb[3 * pix] = al.x, acc_alb[3 * pix + 1] = al.y, acc_alb[3 * pix + 2] = al.z;
        acc_nrm[3 * pix] = nr.x, acc_nrm[3 * pix + 1] = nr.y, acc_nrm[3 * pix + 2] = nr.z;
        acc_l2[pix] = l2;
    }
}

// The per-pixel sums of a batch of several scenes (features off): a thread
// per pixel of every scene, samples added in pass order as k_accumulate does.
__global__ void k_accumulate_multi(pstate P, int passes, mbatch M) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= M.pbase[M.n]) return;
    const int k = M.scene_of_pixel(g);
    const uint32_t pix = g - M.pbase[k];
    const uint32_t npix = uint32_t(M.npix[k]);
    double* acc_rad = M.acc[k];
    v3 a = v3{acc_rad[3 * pix], acc_rad[3 * pix + 1], acc_rad[3 * pix + 2]};
    for (int b = 0; b < passes; b++) {
        const uint32_t slot = M.base[k] + uint32_t(b) * npix + pix;
        v3 c = ld3(P, D_RX, slot);
        if (!std::isfinite(c.x) || !std::isfinite(c.y) || !std::isfinite(c.z)) c = v3{0, 0, 0};  // camera.h:281-287
        a = a + c;
    }
    acc_rad[3 * pix] = a.x, acc_rad[3 * pix + 1] = a.y, acc_rad[3 * pix + 2] = a.z;
}

// ------------------------------------------------------------------ BSDF test hook
__global__ void k_bsdf_eval(dscene S, const int32_t* mats, const double* q, const uint8_t* front, int64_t n,
                            double* eval, double* pdf) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    hitrec rec;
    rec.normal = v3{q[9 * i], q[9 * i + 1], q[9 * i + 2]};
    v3 rd{q[9 * i + 3], q[9 * i + 4], q[9 * i + 5]};
    v3 wi{q[9 * i + 6], q[9 * i + 7], q[9 * i + 8]};
    rec.p = v3{0, 0, 0};
    rec.t = 1;
    rec.front = front[i] != 0;
    rec.mat = mats[i];
    const mpt_material& m = S.mat[rec.mat];
    rstream rs{0, 0};
    scatter_out so = mat_scatter<-1, true>(S, m, rd, rec, rs);
    v3 e = mat_eval<-1, true>(S, m, rd, rec, rec.normal, wi);
    eval[3 * i] = e.x, eval[3 * i + 1] = e.y, eval[3 * i + 2] = e.z;
    pdf[i] = (so.scattered && !so.skip) ? density_value(so.dens, wi) : 0.0;
}

}  // namespace mpt

// ======================================================================== host side
using namespace mpt;

struct mpt_scene {
    host_scene hs;
    dscene ds{};
    std::vector<void*> allocs;
    int device = 0;
    int cu_count = 0;
    // persistent-grid launch configuration
    int grid = 0;
    int block = 256;
    int32_t* ovf = nullptr;
    int64_t ovf_threads = 0;
    int ovf_depth = 0;
    // render state
    pstate ps{};
    int64_t ps_capacity = 0;
    std::vector<void*> ps_allocs;
    uint32_t* queues = nullptr;   // [3][capacity]: two ping-pong + shadow
    uint32_t* counters = nullptr; // queue lengths [0..9], then 64-bit ray and work counters
    // kernels compiled for this scene: geometry profile, material types
    // present (bit t for type t), whether next event estimation runs
    int profile = 0;
    // the closest hit of a ray depends on the path's key (media and height
    // fogs draw free flights, alpha masks draw coverage): a path's first hit
    // cannot stand in for the guides' own
    bool keyed_hits = false;
    uint32_t mat_types = 0;
    bool has_lights = false;
    int light_profile = 0;  // compiled NEE variant: 1 one-primitive quads, 2 environment only, 3 all
    int32_t* pinned_counts = nullptr;  // host copies of queue lengths for the early exit
    double* acc = nullptr;        // radiance, albedo, normal (3 each) + l2
    int64_t acc_pix = 0;
    int last_done = 0;
    int guide_done = 0;           // passes whose albedo and normal guides are in the sums
    mpt_camera last_cam{};
    hipStream_t stream = nullptr;
    // mpt_autotune's choices for one image size (0: none)
    int64_t tuned_pix = 0;
    int tuned_pif = 0;
    int64_t tuned_tail = 0;
    int tuned_bpc = 0;    // blocks per CU
    int tuned_sort = 0;   // 1 binned stages, -1 one kernel per batch
    int tuned_graph = 0;  // 1 bounce graphs
    // per-bounce graphs (use_graph), by bounce index, valid while graph_key holds
    std::vector<hipGraphExec_t> graphs;
    std::vector<unsigned char> graph_key;
    // radiance cache buffers (mpt_render_nrc)
    std::vector<void*> nrc_allocs;
    int64_t nrc_capacity = 0;
    nrc_bufs nb{};
    uint32_t* nrc_queue = nullptr;
    // the scene and camera tables of batched renders (mpt_render_batch)
    void* batch_tab = nullptr;
    size_t batch_tab_bytes = 0;
    // the averaged frame on the device, staged for mpt_read_frame's copies
    double* frame_buf = nullptr;
    int64_t frame_buf_pix = 0;

    void drop_graphs() {
        for (hipGraphExec_t g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        graphs.clear();
        graph_key.clear();
    }

    ~mpt_scene() {
        drop_graphs();
        for (void* p : nrc_allocs) (void)hipFree(p);
        for (void* p : allocs) (void)hipFree(p);
        for (void* p : ps_allocs) (void)hipFree(p);
        if (ovf) (void)hipFree(ovf);
        if (queues) (void)hipFree(queues);
        if (counters) (void)hipFree(counters);
        if (acc) (void)hipFree(acc);
        if (batch_tab) (void)hipFree(batch_tab);
        if (frame_buf) (void)hipFree(frame_buf);
        if (stream) (void)hipStreamDestroy(stream);
        if (pinned_counts) (void)hipHostFree(pinned_counts);
    }
};

namespace {

template <typename T>
int upload(mpt_scene* s, const std::vector<T>& v, const T** out) {
    if (v.empty()) {
        *out = nullptr;
        return MPT_OK;
    }
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, v.size() * sizeof(T)));
    s->allocs.push_back(p);
    HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T*>(p);
    return MPT_OK;
}

int tree_depth(const std::vector<bvh_node>& nodes, int root) {
    if (nodes.empty()) return 0;
    // iterative depth of the subtree at root
    int best = 0;
    std::vector<std::pair<int, int>> todo{{root, 1}};
    while (!todo.empty()) {
        auto [i, dpt] = todo.back();
        todo.pop_back();
        best = std::max(best, dpt);
        const bvh_node& n = nodes[i];
        if (n.count == 0 && i + 1 < int(nodes.size())) {
            todo.push_back({i + 1, dpt + 1});
            todo.push_back({n.offset, dpt + 1});
        }
    }
    return best;
}

int ensure_ovf(mpt_scene* s) {
    int depth_tlas = tree_depth(s->hs.tlas, 0);
    int depth_blas = 0;
    for (const auto& o : s->hs.objects)
        if (o.kind == MPT_OBJ_MESH)
            depth_blas = std::max(depth_blas, tree_depth(s->hs.blas, s->hs.blas_root_node[o.blas_root]));
    // the walk pushes at most three entries per four-wide record (two levels)
    int need = 2 * (depth_tlas + depth_blas) + 4 - kLdsDepth;
    if (need < 1) need = 1;
    int64_t threads = int64_t(s->grid) * s->block;
    if (s->ovf && s->ovf_depth >= need && s->ovf_threads == threads) return MPT_OK;
    if (s->ovf) HIP_TRY(hipFree(s->ovf));
    s->ovf = nullptr;
    HIP_TRY(hipMalloc(&s->ovf, size_t(need) * threads * sizeof(int32_t)));
    s->ovf_depth = need;
    s->ovf_threads = threads;
    return MPT_OK;
}

constexpr int kDefaultBlocksPerCu = 16;

// Persistent stage kernels walk their queue with a grid-wide stride. The
// default grid is 16 blocks of 256 threads per CU, more than a CU holds at
// once: a second round of blocks starts as soon as the first ones run out of
// queue (2% faster than one round on the spheres scene). Scenes on the
// general-geometry kernels (and batches, which always run them) take one
// round instead, MPT_GENERAL_WAVES blocks per CU, as many as their 4-wave
// cap keeps resident: their walks are long and uneven, and a late round
// leaves a tail (the batch of 8 generated scenes: 36.6 M paths/s against
// 33.7 M with 16 blocks, 31.1 M with 6). Renders may set another size
// (blocks_per_cu) and the autotuner tries 8, 16 and 32 besides the default.
// Scenes with quads, lists or instances (and no meshes or media) also take one round, as many
// blocks as their (uncapped) traversal kernel keeps resident: the Cornell box
// ran 589 M paths/s with 3 blocks per CU against 498 M with 16.
int default_blocks_per_cu(const mpt_scene* s) {
    if (s->profile == 0) return MPT_GENERAL_WAVES;
    if (s->profile == 2) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void*>(&k_extend<256, false, kQuadProfile, false>), 256, 0) == hipSuccess &&
            nb > 0)
            return std::min(nb, kDefaultBlocksPerCu);
        (void)hipGetLastError();  // a failed query falls back to the default, leaving no error behind
    }
    return kDefaultBlocksPerCu;
}

int grid_for(mpt_scene* s) {
    s->block = 256;
    s->grid = s->cu_count * default_blocks_per_cu(s);
    return MPT_OK;
}

}  // namespace

extern "C" {

const char* mpt_last_error(void) { return g_last_error.c_str(); }
int32_t mpt_abi_version(void) { return MPT_ABI_VERSION; }

int32_t mpt_device_info(char* name, int32_t name_len, int32_t* cu_count, int64_t* hbm_bytes) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        if (prop.name[0]) std::snprintf(name, size_t(name_len), "%s (%s)", prop.name, prop.gcnArchName);
        else std::snprintf(name, size_t(name_len), "%s, %d CUs", prop.gcnArchName, prop.multiProcessorCount);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = int64_t(prop.totalGlobalMem);
    return MPT_OK;
}

void* mpt_device_alloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 1;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        set_error(MPT_E_HIP, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return nullptr;
    }
    return p;
}

void mpt_device_free(void* p) {
    if (p) (void)hipFree(p);
}

int32_t mpt_copy_to_device(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MPT_OK;
}

int32_t mpt_copy_to_host(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MPT_OK;
}

int32_t mpt_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return MPT_OK;
}

int32_t mpt_scene_create(const mpt_scene_desc* desc, mpt_scene** out) {
    if (!desc || !out) return set_error(MPT_E_INVALID, "null argument");
    *out = nullptr;
    auto s = std::make_unique<mpt_scene>();
    std::string err;
    int rc = build_host_scene(*desc, s->hs, err);
    if (rc != MPT_OK) return set_error(rc, err);
    HIP_TRY(hipGetDevice(&s->device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, s->device));
    s->cu_count = prop.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    host_scene& h = s->hs;
    dscene& d = s->ds;
    if ((rc = upload(s.get(), h.prim, &d.prim))) return rc;
    if ((rc = upload(s.get(), h.prim_kind, &d.pkind))) return rc;
    if ((rc = upload(s.get(), h.prim_mat, &d.pmat))) return rc;
    if ((rc = upload(s.get(), h.prim_canon, &d.pcanon))) return rc;
    if ((rc = upload(s.get(), h.objects, &d.obj))) return rc;
    if ((rc = upload(s.get(), h.ctlas, &d.tlas))) return rc;
    if ((rc = upload(s.get(), h.tlas_items, &d.titems))) return rc;
    if ((rc = upload(s.get(), h.cblas, &d.blas))) return rc;
    if ((rc = upload(s.get(), h.qtlas, &d.qtlas))) return rc;
    if ((rc = upload(s.get(), h.qblas, &d.qblas))) return rc;
    if ((rc = upload(s.get(), h.cblas_roots, &d.blas_roots))) return rc;
    if ((rc = upload(s.get(), h.spheres, &d.sph))) return rc;
    d.tlas_root = h.ctlas_root;
    if ((rc = upload(s.get(), h.xforms, &d.xf))) return rc;
    if ((rc = upload(s.get(), h.materials, &d.mat))) return rc;
    if ((rc = upload(s.get(), h.textures, &d.tex))) return rc;
    if ((rc = upload(s.get(), h.lights, &d.lights))) return rc;
    if ((rc = upload(s.get(), h.light_cdf, &d.lcdf))) return rc;
    if ((rc = upload(s.get(), h.mlight_prim, &d.mlight_prim))) return rc;
    if ((rc = upload(s.get(), h.prim_nrm, &d.vnrm))) return rc;
    if ((rc = upload(s.get(), h.prim_uv, &d.vuv))) return rc;
    if ((rc = upload(s.get(), h.texels, &d.texels))) return rc;
    if ((rc = upload(s.get(), h.texels_hdr, &d.texels_hdr))) return rc;
    if ((rc = upload(s.get(), h.srgb, &d.srgb))) return rc;
    if ((rc = upload(s.get(), h.mlight_cdf, &d.mlight_cdf))) return rc;
    if ((rc = upload(s.get(), h.env_rgb, &d.env_rgb))) return rc;
    if ((rc = upload(s.get(), h.env_cond_cdf, &d.env_cond))) return rc;
    if ((rc = upload(s.get(), h.env_marg_cdf, &d.env_marg))) return rc;
    if ((rc = upload(s.get(), h.env_row_weight, &d.env_roww))) return rc;
    if ((rc = upload(s.get(), h.ggx_e, &d.ggx_e))) return rc;
    if ((rc = upload(s.get(), h.ggx_e_avg, &d.ggx_eavg))) return rc;
    d.env_total = h.env_total_weight;
    d.env_intensity = h.env_intensity;
    d.env_rotation = h.env_rotation;
    d.nobj = int32_t(h.objects.size());
    d.nworld = int32_t(h.tlas_items.size());
    if ((rc = upload(s.get(), h.vox, &d.vox))) return rc;
    if ((rc = upload(s.get(), h.vox_cells, &d.vcells))) return rc;
    d.nlights = int32_t(h.lights.size());
    d.lstrategy = h.light_strategy;
    d.env_w = h.env_w;
    d.env_h = h.env_h;
    d.ground_mat = h.ground_mat;
    d.ground_h = h.ground_h;
    if ((rc = upload(s.get(), h.fogs, &d.fog))) return rc;
    d.nfog = int32_t(h.fogs.size());
    d.shadow_occlusion = 1;
    for (const mpt_material& m : h.materials)
        if (m.type == MPT_MAT_DIFFUSE_LIGHT || m.glow || mat_lets_shadows(m)) d.shadow_occlusion = 0;
    d.has_sun = h.has_sun;
    d.sun_frame = h.sun_frame;
    d.sun_cos_max = h.sun_cos_max;
    d.sun_solid_angle = h.sun_solid_angle;
    d.sun_radiance = h.sun_radiance;
    d.ground_id = 0;  // one past the last canonical ID of the world list
    for (const obj_rec& r : h.objects)
        if (!(r.flags & MPT_OBJF_LIGHT_ONLY)) d.ground_id = std::max(d.ground_id, r.id_base + r.count);
    // geometry profile: spheres only with object i = primitive i, quads and
    // spheres in lists and instances, or everything
    {
        bool prims_only = true, quads_ok = true;
        for (size_t o = 0; o < h.objects.size(); o++) {
            const obj_rec& r = h.objects[o];
            if (r.kind != MPT_OBJ_PRIM || r.xform >= 0 || r.first != int64_t(o) ||
                h.prim_kind[r.first] != MPT_PRIM_SPHERE)
                prims_only = false;
            if (r.kind == MPT_OBJ_MESH || r.kind == MPT_OBJ_MEDIUM) quads_ok = false;
        }
        for (int32_t k : h.prim_kind)
            if (k == MPT_PRIM_TRIANGLE) quads_ok = false;
        // 3: everything, hits carry (u, v) and tangents, shading has textures, maps, thin glass, measured
        // metals, and the objects beside the BVH (ground, height fogs) and alpha masks are tested
        s->profile = h.rich ? 3 : (prims_only && !h.objects.empty() ? 1 : (quads_ok ? 2 : 0));
        for (int32_t m : h.prim_mat)
            if (m >= 0) s->mat_types |= 1u << h.materials[m].type;
        for (const auto& r : h.objects)
            if (r.kind == MPT_OBJ_MEDIUM) s->mat_types |= 1u << MPT_MAT_PHASE;
        if (!h.fogs.empty()) s->mat_types |= 1u << MPT_MAT_PHASE;  // fog hits shade as a phase material
        if (h.ground_mat >= 0) s->mat_types |= 1u << h.materials[h.ground_mat].type;
        for (const mpt_vox_model& vm : h.vox) {  // voxel hits take their palette's (or the outline's) material
            for (int i = 0; i < 256; i++)
                if (vm.material[i] >= 0) s->mat_types |= 1u << h.materials[vm.material[i]].type;
            if (vm.edge_width > 0) s->mat_types |= 1u << h.materials[vm.edge_material].type;
        }
        s->keyed_hits = !h.fogs.empty();
        for (const auto& r : h.objects)
            if (r.kind == MPT_OBJ_MEDIUM || r.alpha != -1) s->keyed_hits = true;
        s->has_lights = !h.lights.empty();
        bool quads = true, env_only = true;
        for (const auto& L : h.lights) {
            if (L.object == -2) {  // the sun: only the general light code samples it
                quads = env_only = false;
                continue;
            }
            if (L.object < 0) {
                quads = false;
                continue;
            }
            env_only = false;
            const obj_rec& r = h.objects[L.object];
            if (r.kind != MPT_OBJ_PRIM || r.xform >= 0 || h.prim_kind[r.first] != MPT_PRIM_QUAD) quads = false;
        }
        s->light_profile = !s->has_lights ? 0 : (quads ? 1 : (env_only ? 2 : 3));
    }
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->pinned_counts), 8 * sizeof(int32_t), hipHostMallocDefault));
    if ((rc = grid_for(s.get()))) return rc;
    if ((rc = ensure_ovf(s.get()))) return rc;
    *out = s.release();
    return MPT_OK;
}

void mpt_scene_destroy(mpt_scene* scene) { delete scene; }

int32_t mpt_scene_counts(const mpt_scene* s, int64_t* n_prims, int32_t* n_objects, int64_t* n_tlas, int64_t* n_blas,
                         int32_t* n_lights) {
    if (!s) return set_error(MPT_E_INVALID, "null scene");
    if (n_prims) *n_prims = int64_t(s->hs.prim_kind.size());
    if (n_objects) *n_objects = int32_t(s->hs.objects.size());
    if (n_tlas) *n_tlas = int64_t(s->hs.tlas.size());
    if (n_blas) *n_blas = int64_t(s->hs.blas.size());
    if (n_lights) *n_lights = int32_t(s->hs.lights.size());
    return MPT_OK;
}

int32_t mpt_scene_light_probabilities(const mpt_scene* s, double* probabilities) {
    if (!s || !probabilities) return set_error(MPT_E_INVALID, "null argument");
    for (size_t k = 0; k < s->hs.lights.size(); k++) probabilities[k] = s->hs.lights[k].probability;
    return MPT_OK;
}

int32_t mpt_ggx_albedo_table(double* e_1024, double* e_avg_32) {
    std::vector<double> e, ea;
    build_ggx_table(e, ea);
    if (e_1024) std::memcpy(e_1024, e.data(), e.size() * sizeof(double));
    if (e_avg_32) std::memcpy(e_avg_32, ea.data(), ea.size() * sizeof(double));
    return MPT_OK;
}

int32_t mpt_image_height(const mpt_camera* cam) {
    if (!cam || !(cam->aspect_ratio > 0)) return set_error(MPT_E_INVALID, "bad camera");
    int h = int(cam->image_width / cam->aspect_ratio);
    return h < 1 ? 1 : h;
}

int32_t mpt_trace_closest_dev(mpt_scene* s, const double* rays, const double* tt, int64_t n, int64_t* id, double* t,
                              uint8_t* front, void* stream) {
    if (!s || n < 0) return set_error(MPT_E_INVALID, "bad argument");
    if (n == 0) return MPT_OK;
    if (!rays || !tt || !id || !t || !front) return set_error(MPT_E_INVALID, "null buffer");
    launch_ctx L{s->ovf, int64_t(s->grid) * s->block};
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
    if (s->profile == 1)
        hipLaunchKernelGGL((k_query<256, F_SPHERE>), dim3(s->grid), dim3(256), 0, st, s->ds, L, rays, tt, n, id, t, front);
    else if (s->profile == 2)
        hipLaunchKernelGGL((k_query<256, kQuadProfile>), dim3(s->grid), dim3(256), 0, st, s->ds, L, rays, tt, n, id, t,
                           front);
    else if (s->profile == 3)
        hipLaunchKernelGGL((k_query<256, F_ALL | F_UV>), dim3(s->grid), dim3(256), 0, st, s->ds, L, rays, tt, n, id, t,
                           front);
    else
        hipLaunchKernelGGL((k_query<256, F_ALL>), dim3(s->grid), dim3(256), 0, st, s->ds, L, rays, tt, n, id, t, front);
    HIP_TRY(hipGetLastError());
    if (!stream) HIP_TRY(hipStreamSynchronize(st));
    return MPT_OK;
}

int32_t mpt_trace_closest(mpt_scene* s, const double* rays, const double* tt, int64_t n, int64_t* id, double* t,
                          uint8_t* front) {
    if (!s || n < 0) return set_error(MPT_E_INVALID, "bad argument");
    if (n == 0) return MPT_OK;
    if (!rays || !tt || !id || !t || !front) return set_error(MPT_E_INVALID, "null buffer");
    double *dr = nullptr, *dt = nullptr, *dto = nullptr;
    int64_t* did = nullptr;
    uint8_t* df = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dr), (void)hipFree(dt), (void)hipFree(dto), (void)hipFree(did), (void)hipFree(df);
    };
    if (hipMalloc(&dr, n * 6 * sizeof(double)) != hipSuccess || hipMalloc(&dt, n * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc(&dto, n * sizeof(double)) != hipSuccess || hipMalloc(&did, n * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&df, n) != hipSuccess) {
        cleanup();
        return set_error(MPT_E_HIP, "hipMalloc failed for the ray batch");
    }
    int rc = MPT_OK;
    if (hipMemcpy(dr, rays, n * 6 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, tt, n * 2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return set_error(MPT_E_HIP, "hipMemcpy of the ray batch failed");
    }
    rc = mpt_trace_closest_dev(s, dr, dt, n, did, dto, df, nullptr);
    if (rc == MPT_OK) {
        if (hipMemcpy(id, did, n * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(t, dto, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(front, df, n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = set_error(MPT_E_HIP, "hipMemcpy of the results failed");
    }
    cleanup();
    return rc;
}

int32_t mpt_trace_ties(mpt_scene* s, const double* rays, const double* tt, int64_t n, int64_t* id, double* t,
                       uint8_t* front, int32_t* ties, int64_t* tie_min, int64_t* tie_max) {
    if (!s || n < 0) return set_error(MPT_E_INVALID, "bad argument");
    if (n == 0) return MPT_OK;
    if (!rays || !tt || !id || !t || !front || !ties || !tie_min || !tie_max)
        return set_error(MPT_E_INVALID, "null buffer");
    double *dr = nullptr, *dt = nullptr, *dto = nullptr;
    int64_t *did = nullptr, *dlo = nullptr, *dhi = nullptr;
    int32_t* dties = nullptr;
    uint8_t* df = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dr), (void)hipFree(dt), (void)hipFree(dto), (void)hipFree(did), (void)hipFree(df);
        (void)hipFree(dlo), (void)hipFree(dhi), (void)hipFree(dties);
    };
    if (hipMalloc(&dr, n * 6 * sizeof(double)) != hipSuccess || hipMalloc(&dt, n * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc(&dto, n * sizeof(double)) != hipSuccess || hipMalloc(&did, n * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&df, n) != hipSuccess || hipMalloc(&dlo, n * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&dhi, n * sizeof(int64_t)) != hipSuccess || hipMalloc(&dties, n * sizeof(int32_t)) != hipSuccess) {
        cleanup();
        return set_error(MPT_E_HIP, "hipMalloc failed for the ray batch");
    }
    if (hipMemcpy(dr, rays, n * 6 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, tt, n * 2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return set_error(MPT_E_HIP, "hipMemcpy of the ray batch failed");
    }
    int rc = mpt_trace_closest_dev(s, dr, dt, n, did, dto, df, nullptr);
    if (rc == MPT_OK) {
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, int64_t(s->grid));
        hipLaunchKernelGGL(k_ties, dim3(unsigned(blocks)), dim3(256), 0, s->stream, s->ds, dr, dt, n, did, dto, dties,
                           dlo, dhi);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess)
            rc = set_error(MPT_E_HIP, "tie listing kernel failed");
    }
    if (rc == MPT_OK) {
        if (hipMemcpy(id, did, n * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(t, dto, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(front, df, n, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ties, dties, n * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(tie_min, dlo, n * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(tie_max, dhi, n * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
            rc = set_error(MPT_E_HIP, "hipMemcpy of the results failed");
    }
    cleanup();
    return rc;
}

int32_t mpt_primary_depth(mpt_scene* s, const mpt_camera* cam, double* depth) {
    if (!s || !cam || !depth) return set_error(MPT_E_INVALID, "null argument");
    if (cam->image_width <= 0 || !(cam->aspect_ratio > 0)) return set_error(MPT_E_INVALID, "bad camera");
    const dcam C = init_camera(*cam, s->hs.env_w > 0);
    const int64_t n = int64_t(C.w) * C.h;
    std::vector<double> rays(6 * n), tt(2 * n), t(n);
    std::vector<int64_t> id(n);
    std::vector<uint8_t> front(n);
    for (int j = 0; j < C.h; j++)
        for (int i = 0; i < C.w; i++) {  // pixel centre, no defocus (camera.h:291-301)
            const int64_t k = int64_t(j) * C.w + i;
            v3 ps = C.pixel00 + (double(i) * C.du) + (double(j) * C.dv);
            v3 org = C.ortho_height > 0 ? ps + C.focus_dist * C.w_axis : C.center;
            v3 d = ps - org;
            rays[6 * k] = org.x, rays[6 * k + 1] = org.y, rays[6 * k + 2] = org.z;
            rays[6 * k + 3] = d.x, rays[6 * k + 4] = d.y, rays[6 * k + 5] = d.z;
            tt[2 * k] = 0.001;
            tt[2 * k + 1] = kInf;
        }
    int rc = mpt_trace_closest(s, rays.data(), tt.data(), n, id.data(), t.data(), front.data());
    if (rc) return rc;
    for (int64_t k = 0; k < n; k++) {
        const double dx = rays[6 * k + 3], dy = rays[6 * k + 4], dz = rays[6 * k + 5];
        depth[k] = id[k] < 0 ? 0.0 : t[k] * std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    return MPT_OK;
}

int32_t mpt_bsdf_eval(mpt_scene* s, const int32_t* mats, const double* q, const uint8_t* front, int64_t n,
                      double* eval, double* pdf) {
    if (!s || n < 0 || (n > 0 && (!mats || !q || !front || !eval || !pdf))) return set_error(MPT_E_INVALID, "bad argument");
    for (int64_t i = 0; i < n; i++)
        if (mats[i] < 0 || mats[i] >= int32_t(s->hs.materials.size())) return set_error(MPT_E_INVALID, "bad material index");
    if (n == 0) return MPT_OK;
    int32_t* dm = nullptr;
    double *dq = nullptr, *de = nullptr, *dp = nullptr;
    uint8_t* dfr = nullptr;
    HIP_TRY(hipMalloc(&dm, n * sizeof(int32_t)));
    HIP_TRY(hipMalloc(&dq, n * 9 * sizeof(double)));
    HIP_TRY(hipMalloc(&de, n * 3 * sizeof(double)));
    HIP_TRY(hipMalloc(&dp, n * sizeof(double)));
    HIP_TRY(hipMalloc(&dfr, n));
    HIP_TRY(hipMemcpy(dm, mats, n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dq, q, n * 9 * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dfr, front, n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_bsdf_eval, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s->stream, s->ds, dm, dq, dfr, n, de, dp);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(eval, de, n * 3 * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pdf, dp, n * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipFree(dm), (void)hipFree(dq), (void)hipFree(de), (void)hipFree(dp), (void)hipFree(dfr);
    return MPT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ render
namespace {

int ensure_paths(mpt_scene* s, int64_t cap) {
    if (s->ps_capacity >= cap) return MPT_OK;
    for (void* p : s->ps_allocs) (void)hipFree(p);
    s->ps_allocs.clear();
    if (s->queues) (void)hipFree(s->queues);
    s->queues = nullptr;
    s->ps_capacity = 0;
    auto alloc = [&](size_t bytes, void** out) -> int {
        HIP_TRY(hipMalloc(out, bytes));
        s->ps_allocs.push_back(*out);
        return MPT_OK;
    };
    int rc;
    double* base = nullptr;
    // the (u, v) record only where a material reads it
    const int fields = s->hs.uses_uv ? D_COUNT : kUvRecord * kRecDoubles;
    if ((rc = alloc(size_t(cap) * fields * sizeof(double), reinterpret_cast<void**>(&base)))) return rc;
    s->ps.base = base;  // fields beyond `fields` are never touched
    s->ps.cap = cap;
    if ((rc = alloc(size_t(cap) * sizeof(uint64_t), reinterpret_cast<void**>(&s->ps.key)))) return rc;
    HIP_TRY(hipMalloc(&s->queues, size_t(cap) * (3 + kClasses) * sizeof(uint32_t)));
    if (!s->counters) HIP_TRY(hipMalloc(&s->counters, kCtrWords * sizeof(uint32_t) + 16 * sizeof(uint64_t)));
    s->ps_capacity = cap;
    return MPT_OK;
}


}  // namespace

extern "C" {

int32_t mpt_read_frame(mpt_scene* s, double* radiance, double* albedo, double* normal, double* variance) {
    if (!s || !s->acc || s->last_done <= 0) return set_error(MPT_E_INVALID, "no frame rendered yet");
    // camera::average on the device (k_average), then only the requested
    // outputs cross PCIe
    const int64_t npix = s->acc_pix;
    const bool features = s->last_cam.collect_features != 0;
    if (s->frame_buf_pix < npix) {
        if (s->frame_buf) HIP_TRY(hipFree(s->frame_buf));
        s->frame_buf = nullptr;
        s->frame_buf_pix = 0;
        HIP_TRY(hipMalloc(&s->frame_buf, size_t(npix) * 10 * sizeof(double)));
        s->frame_buf_pix = npix;
    }
    double* fr = s->frame_buf;
    double* fa = fr + 3 * npix;
    double* fn = fa + 3 * npix;
    double* fv = fn + 3 * npix;
    hipLaunchKernelGGL(k_average, dim3(unsigned((npix + 255) / 256)), dim3(256), 0, s->stream, s->acc, npix,
                       s->last_done, s->guide_done, features ? 1 : 0, fr, fa, fn, fv);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (radiance) HIP_TRY(hipMemcpy(radiance, fr, size_t(npix) * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (features) {
        if (albedo) HIP_TRY(hipMemcpy(albedo, fa, size_t(npix) * 3 * sizeof(double), hipMemcpyDeviceToHost));
        if (normal) HIP_TRY(hipMemcpy(normal, fn, size_t(npix) * 3 * sizeof(double), hipMemcpyDeviceToHost));
        if (variance) HIP_TRY(hipMemcpy(variance, fv, size_t(npix) * sizeof(double), hipMemcpyDeviceToHost));
    }
    return MPT_OK;
}

int32_t mpt_read_frame_dev(mpt_scene* s, double* radiance, double* albedo, double* normal, double* variance,
                           void* stream) {
    if (!s || !s->acc || s->last_done <= 0) return set_error(MPT_E_INVALID, "no frame rendered yet");
    const bool features = s->last_cam.collect_features != 0;
    if (!features && (albedo || normal || variance))
        return set_error(MPT_E_INVALID, "the last render collected no guides");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
    const int64_t npix = s->acc_pix;
    hipLaunchKernelGGL(k_average, dim3(unsigned((npix + 255) / 256)), dim3(256), 0, st, s->acc, npix, s->last_done,
                       s->guide_done, features ? 1 : 0, radiance, albedo, normal, variance);
    HIP_TRY(hipGetLastError());
    if (!stream) HIP_TRY(hipStreamSynchronize(st));
    return MPT_OK;
}

int32_t mpt_primary_depth_dev(mpt_scene* s, const mpt_camera* cam, double* depth, void* stream) {
    if (!s || !cam || !depth) return set_error(MPT_E_INVALID, "null argument");
    if (cam->image_width <= 0 || !(cam->aspect_ratio > 0)) return set_error(MPT_E_INVALID, "bad camera");
    const dcam C = init_camera(*cam, s->hs.env_w > 0);
    launch_ctx L{s->ovf, int64_t(s->grid) * s->block};
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
#define DEPTH_ARGS s->ds, L, C.pixel00, C.du, C.dv, C.center, C.w_axis, C.ortho_height, C.focus_dist, C.w, C.h, depth
    if (s->profile == 1)
        hipLaunchKernelGGL((k_depth<256, F_SPHERE>), dim3(s->grid), dim3(256), 0, st, DEPTH_ARGS);
    else if (s->profile == 2)
        hipLaunchKernelGGL((k_depth<256, kQuadProfile>), dim3(s->grid), dim3(256), 0, st, DEPTH_ARGS);
    else if (s->profile == 3)
        hipLaunchKernelGGL((k_depth<256, F_ALL | F_UV>), dim3(s->grid), dim3(256), 0, st, DEPTH_ARGS);
    else
        hipLaunchKernelGGL((k_depth<256, F_ALL>), dim3(s->grid), dim3(256), 0, st, DEPTH_ARGS);
#undef DEPTH_ARGS
    HIP_TRY(hipGetLastError());
    if (!stream) HIP_TRY(hipStreamSynchronize(st));
    return MPT_OK;
}

}  // extern "C"

namespace {

// Query and training buffers of the radiance cache for `cap` paths.
int ensure_nrc(mpt_scene* s, int64_t cap) {
    if (s->nrc_capacity >= cap) return MPT_OK;
    for (void* p : s->nrc_allocs) (void)hipFree(p);
    s->nrc_allocs.clear();
    s->nrc_capacity = 0;
    auto get = [&](void** out, size_t bytes) -> int {
        HIP_TRY(hipMalloc(out, bytes));
        s->nrc_allocs.push_back(*out);
        return MPT_OK;
    };
    nrc_bufs& B = s->nb;
    B.tcap = std::max<int64_t>(1 << 16, cap / 16);
    int rc;
    if ((rc = get((void**)&s->nrc_queue, cap * sizeof(uint32_t))) || (rc = get((void**)&B.pos, cap * 3 * sizeof(float))) ||
        (rc = get((void**)&B.other, cap * 10 * sizeof(float))) || (rc = get((void**)&B.pred, cap * 3 * sizeof(float))) ||
        (rc = get((void**)&B.tidx, cap * sizeof(uint32_t))) || (rc = get((void**)&B.tslot, B.tcap * sizeof(uint32_t))) ||
        (rc = get((void**)&B.trad, B.tcap * 3 * sizeof(double))) || (rc = get((void**)&B.tbeta, B.tcap * 3 * sizeof(double))) ||
        (rc = get((void**)&B.tpos, B.tcap * 3 * sizeof(float))) || (rc = get((void**)&B.tother, B.tcap * 10 * sizeof(float))) ||
        (rc = get((void**)&B.ttarget, B.tcap * 3 * sizeof(float))))
        return rc;
    s->nrc_capacity = cap;
    return MPT_OK;
}

int render_impl(mpt_scene* s, const mpt_camera* cam, const mpt_render_opts* opts, double* radiance, double* albedo,
                double* normal, double* variance, mpt_pass_callback on_pass, void* user, mpt_render_stats* stats,
                mpt_nrc* nrc, const mpt_nrc_opts* nopts, mpt_nrc_stats* nstats) {
    if (!s || !cam) return set_error(MPT_E_INVALID, "null argument");
    if (cam->image_width < 1 || !(cam->aspect_ratio > 0) || cam->samples_per_pixel < 1 || cam->max_depth < 1 ||
        cam->max_depth > 1000)
        return set_error(MPT_E_INVALID, "bad camera settings (width, aspect, spp or max_depth out of range)");
    if (cam->max_diffuse < 0 || cam->max_diffuse > 1000 || cam->max_specular < 0 || cam->max_specular > 1000 ||
        cam->max_scatter < 0 || cam->max_scatter > 1000)
        return set_error(MPT_E_INVALID, "bounce caps must be in [0, 1000]");
    mpt_render_opts o{};
    if (opts) o = *opts;
    if (o.first_pass < 0 || o.pass_count < 0) return set_error(MPT_E_INVALID, "negative first_pass or pass_count");
    auto t0 = std::chrono::steady_clock::now();
    dcam C = init_camera(*cam, s->hs.env_w > 0);
    const int64_t npix64 = int64_t(C.w) * C.h;
    if (npix64 > (int64_t(1) << 28)) return set_error(MPT_E_CAPACITY, "image too large");
    const int npix = int(npix64);
    const int total = o.pass_count > 0 ? o.pass_count : C.spp - o.first_pass;
    if (total < 1) return set_error(MPT_E_INVALID, "no passes to render (first_pass >= samples_per_pixel)");
    int pif = o.passes_in_flight;
    if (pif <= 0 && s->tuned_pix == npix64 && s->tuned_pif > 0) pif = s->tuned_pif;  // mpt_autotune's choice
    if (pif <= 0) {
        // about 16M paths in flight: many times what the device holds at once,
        // so the tail of each bounce (few paths left) is a small share
        pif = int(std::max<int64_t>(1, (int64_t(16) << 20) / npix64));
    }
    pif = std::min(pif, total);
    if (int64_t(pif) * npix64 > (int64_t(1) << 31)) pif = int((int64_t(1) << 31) / npix64);
    int rc;
    if ((rc = ensure_paths(s, int64_t(pif) * npix))) return rc;
    const bool tuned = s->tuned_pix == npix64;
    int bpc = o.blocks_per_cu > 0 ? o.blocks_per_cu
                                  : (tuned && s->tuned_bpc > 0 ? s->tuned_bpc : default_blocks_per_cu(s));
    if (bpc > 64) return set_error(MPT_E_INVALID, "blocks_per_cu must be in [1, 64]");
    if (s->grid != s->cu_count * bpc) {  // the persistent grid, and the stack spill area sized for it
        s->grid = s->cu_count * bpc;
        if ((rc = ensure_ovf(s))) return rc;
    }
    const bool sorted = o.sort_by_material > 0 || (o.sort_by_material == 0 && !(tuned && s->tuned_sort < 0));
    if (!sorted && nrc) return set_error(MPT_E_INVALID, "the radiance cache needs sort_by_material (it cuts paths at a bounce)");
    const bool keep = o.accumulate && s->acc && s->acc_pix == npix && s->last_done > 0 &&
                      s->last_cam.collect_features == cam->collect_features;
    if (o.accumulate && !keep && s->last_done > 0)
        return set_error(MPT_E_INVALID, "accumulate: the previous render had another size or feature setting");
    if (s->acc_pix != npix || !s->acc) {
        if (s->acc) HIP_TRY(hipFree(s->acc));
        s->acc = nullptr;
        HIP_TRY(hipMalloc(&s->acc, size_t(npix) * 10 * sizeof(double)));
        s->acc_pix = npix;
    }
    if (!keep) {
        HIP_TRY(hipMemsetAsync(s->acc, 0, size_t(npix) * 10 * sizeof(double), s->stream));
        s->last_done = 0;
        s->guide_done = 0;
    }
    s->last_cam = *cam;
    const int done_before = s->last_done;
    if (o.guide_passes < 0) return set_error(MPT_E_INVALID, "guide_passes must be >= 0");

    const int64_t cap = s->ps_capacity;
    uint32_t* q[3] = {s->queues, s->queues + cap, s->queues + 2 * cap};
    uint32_t* qcls = s->queues + 3 * cap;  // kClasses queues of cap entries
    uint32_t* cnt = s->counters;           // counters 0, 1 ping-pong queue lengths, 2 shadow, 3.. classes
    uint32_t* ccls = ctr(cnt, 3);
    launch_ctx L{s->ovf, int64_t(s->grid) * s->block};
    const pstate P = s->ps;
    double* acc_rad = s->acc;
    double* acc_alb = acc_rad + 3 * size_t(npix);
    double* acc_nrm = acc_alb + 3 * size_t(npix);
    double* acc_l2 = acc_nrm + 3 * size_t(npix);
    unsigned long long* ray_stats = reinterpret_cast<unsigned long long*>(s->counters + kCtrWords);  // [0] ext, [1] shadow
    unsigned long long* work = ray_stats + 2;                                               // nodes/bytes ext, shadow
    HIP_TRY(hipMemsetAsync(ray_stats, 0, 8 * sizeof(unsigned long long), s->stream));
    const bool timing = (o.profile & 1) != 0;
    const bool counting = (o.profile & 2) != 0;
    const bool graphs = (o.use_graph > 0 || (o.use_graph == 0 && tuned && s->tuned_graph > 0)) && !timing && !counting &&
                        !nrc && sorted;

    // per-stage event timing (profile bit 0): events are recorded around
    // each launch and read back after each batch
    struct timed {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<timed> evs;
    size_t ev_used = 0;
    double stage_ms[5] = {0, 0, 0, 0, 0};
    int64_t stage_n[5] = {0, 0, 0, 0, 0};
    auto ev_begin = [&](int stage) -> int {
        if (!timing) return MPT_OK;
        if (ev_used == evs.size()) {
            timed t{stage, nullptr, nullptr};
            HIP_TRY(hipEventCreate(&t.a));
            HIP_TRY(hipEventCreate(&t.b));
            evs.push_back(t);
        }
        evs[ev_used].stage = stage;
        HIP_TRY(hipEventRecord(evs[ev_used].a, s->stream));
        return MPT_OK;
    };
    auto ev_end = [&]() -> int {
        if (!timing) return MPT_OK;
        HIP_TRY(hipEventRecord(evs[ev_used].b, s->stream));
        ev_used++;
        return MPT_OK;
    };
    auto ev_collect = [&]() -> int {
        if (!timing) return MPT_OK;
        HIP_TRY(hipStreamSynchronize(s->stream));
        for (size_t k = 0; k < ev_used; k++) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, evs[k].a, evs[k].b));
            stage_ms[evs[k].stage] += ms;
            stage_n[evs[k].stage]++;
        }
        ev_used = 0;
        return MPT_OK;
    };
    // a small ring of events marks when each bounce's queue length has been
    // copied to pinned host memory; the host looks two bounces back before
    // launching the next one and stops once no path is left
    constexpr int kRing = 4;
    hipEvent_t ring[kRing] = {};
    for (int k = 0; k < kRing; k++) HIP_TRY(hipEventCreateWithFlags(&ring[k], hipEventDisableTiming));
    auto ev_free = [&]() {
        for (auto& t : evs) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        for (auto& e : ring) (void)hipEventDestroy(e);
    };
#define EV(stage, launch)                            \
    do {                                             \
        if ((rc = ev_begin(stage))) {                \
            ev_free();                               \
            return rc;                               \
        }                                            \
        launch;                                      \
        if ((rc = ev_end())) {                       \
            ev_free();                               \
            return rc;                               \
        }                                            \
    } while (0)

    const dscene DS = s->ds;
    const bool has_env = s->hs.env_w > 0;
    // the radiance cache: where it cuts paths, its buffers
    nrc_route R{-1, 0, 1, 0, nullptr, nullptr};
    nrc_bufs NB{};
    if (nrc) {
        if (nopts->query_depth >= C.max_depth) return set_error(MPT_E_INVALID, "nrc: query_depth must be below max_depth");
        if ((rc = ensure_nrc(s, s->ps_capacity))) return rc;
        NB = s->nb;
        NB.tcnt = ctr(cnt, kNrcTrain);
        double lo[3], hi[3];
        const bool given = nopts->box_max[0] > nopts->box_min[0] && nopts->box_max[1] > nopts->box_min[1] &&
                           nopts->box_max[2] > nopts->box_min[2];
        for (int a = 0; a < 3; a++) {
            lo[a] = given ? nopts->box_min[a] : s->hs.world_box.ax[a].min;
            hi[a] = given ? nopts->box_max[a] : s->hs.world_box.ax[a].max;
        }
        NB.box_min = v3{lo[0], lo[1], lo[2]};
        NB.inv_ext = v3{1 / (hi[0] - lo[0]), 1 / (hi[1] - lo[1]), 1 / (hi[2] - lo[2])};
        R = nrc_route{nopts->query_depth, nopts->infer ? 1 : 0, uint32_t(nopts->train_stride), 0, s->nrc_queue,
                      ctr(cnt, kNrcQueue)};
        if (nstats) std::memset(nstats, 0, sizeof(*nstats));
    }
    const unsigned G = unsigned(s->grid);
    hipStream_t stream = s->stream;
    const mbatch MB{};  // one scene: the kernels read S and C
    auto launch_extend = [&](int depth, uint32_t* qin, uint32_t* cin, uint32_t* cout) {
#define EXT(FEAT)                                                                                                    \
    do {                                                                                                             \
        if (counting)                                                                                                \
            hipLaunchKernelGGL((k_extend<256, true, FEAT>), dim3(G), dim3(256), 0, stream, DS, L, P, depth, qin, cin, \
                               cout, ctr(cnt, 2), qcls, ccls, cap, ray_stats, work, R, MB);                              \
        else                                                                                                         \
            hipLaunchKernelGGL((k_extend<256, false, FEAT>), dim3(G), dim3(256), 0, stream, DS, L, P, depth, qin,    \
                               cin, cout, ctr(cnt, 2), qcls, ccls, cap, ray_stats, work, R, MB);                         \
    } while (0)
        if (s->profile == 1) EXT(F_SPHERE);
        else if (s->profile == 2) EXT(kQuadProfile);
        else if (s->profile == 3) EXT(F_ALL | F_UV);
        else EXT(F_ALL);
#undef EXT
    };
    auto launch_shadow = [&](int depth) {
#define SHD2(FEAT, ENV)                                                                                             \
    do {                                                                                                            \
        if (counting)                                                                                               \
            hipLaunchKernelGGL((k_shadow<256, true, FEAT, ENV>), dim3(G), dim3(256), 0, stream, DS, C, L, P, depth,  \
                               q[2], ctr(cnt, 2), ccls, ray_stats, work, MB);                                               \
        else                                                                                                        \
            hipLaunchKernelGGL((k_shadow<256, false, FEAT, ENV>), dim3(G), dim3(256), 0, stream, DS, C, L, P, depth, \
                               q[2], ctr(cnt, 2), ccls, ray_stats, work, MB);                                               \
    } while (0)
#define SHD(FEAT)                       \
    do {                                \
        if (has_env) SHD2(FEAT, true);  \
        else SHD2(FEAT, false);         \
    } while (0)
#define SHD_OCC(CNT, ENV)                                                                                          \
    hipLaunchKernelGGL((k_shadow<256, CNT, F_ALL, ENV, false, true>), dim3(G), dim3(256), 0, stream, DS, C, L, P, \
                       depth, q[2], ctr(cnt, 2), ccls, ray_stats, work, MB)
        if (s->profile == 1) SHD(F_SPHERE);
        else if (s->profile == 2) SHD(kQuadProfile);
        else if (s->profile == 3) SHD(F_ALL | F_UV);  // the ground occludes too
        else if (DS.shadow_occlusion) {  // occlusion queries only: the kernel without the general loop
            if (counting) {
                if (has_env) SHD_OCC(true, true);
                else SHD_OCC(true, false);
            } else {
                if (has_env) SHD_OCC(false, true);
                else SHD_OCC(false, false);
            }
        } else SHD(F_ALL);
#undef SHD_OCC
#undef SHD
#undef SHD2
    };
    auto launch_shade = [&](int cls, int depth, uint32_t* qout, uint32_t* cout) {
        const uint32_t* qc = qcls + cls * cap;
        const uint32_t* cc = ccls + cls * kCtrStride;
#define SHA3(MT, LK, ENV, RICH)                                                                                  \
    hipLaunchKernelGGL((k_shade<256, MT, LK, ENV, RICH>), dim3(G), dim3(256), 0, stream, DS, C, P, depth, qc, cc,    \
                       qout, cout, q[2], ctr(cnt, 2), MB)
    // in scenes with textures, maps, thin glass or measured metals the surface materials get the variant with them
#define SHA2(MT, LK, ENV)                                                                                           \
    do {                                                                                                            \
        constexpr bool kMapped = MT == MPT_MAT_LAMBERTIAN || MT == MPT_MAT_METAL || MT == MPT_MAT_DIELECTRIC ||     \
                                 MT == MPT_MAT_PRINCIPLED;                                                          \
        if (kMapped && s->profile == 3) SHA3(MT, LK, ENV, kMapped);                                                \
        else SHA3(MT, LK, ENV, false);                                                                              \
    } while (0)
#define SHA1(MT, LK)                                          \
    do {                                                      \
        if (MT < 0 && has_env) SHA2(MT, LK, (MT < 0));        \
        else SHA2(MT, LK, false);                             \
    } while (0)
#define SHA(MT)                                      \
    do {                                             \
        switch (s->light_profile) {                  \
            case 0: SHA1(MT, 0); break;              \
            case 1: SHA1(MT, L_QUAD); break;         \
            case 2: SHA1(MT, L_ENV); break;          \
            default: SHA1(MT, L_ALL); break;         \
        }                                            \
    } while (0)
        switch (cls) {
            case 0: SHA(-1); break;
            case 1 + MPT_MAT_LAMBERTIAN: SHA(MPT_MAT_LAMBERTIAN); break;
            case 1 + MPT_MAT_METAL: SHA(MPT_MAT_METAL); break;
            case 1 + MPT_MAT_DIELECTRIC: SHA(MPT_MAT_DIELECTRIC); break;
            case 1 + MPT_MAT_PRINCIPLED: SHA(MPT_MAT_PRINCIPLED); break;
            case 1 + MPT_MAT_DIFFUSE_LIGHT: SHA(MPT_MAT_DIFFUSE_LIGHT); break;
            default: SHA(MPT_MAT_PHASE); break;
        }
#undef SHA
#undef SHA1
#undef SHA2
#undef SHA3
    };

    // the tail of a batch runs as one k_finish launch once at most `tail`
    // paths are left; by default two paths per thread of the grid
    int64_t tail_default = 2 * L.nthreads;
    if (s->tuned_pix == npix64 && s->tuned_tail != 0) tail_default = s->tuned_tail < 0 ? 0 : s->tuned_tail;
    const int64_t tail = o.tail_paths < 0 ? 0 : (o.tail_paths > 0 ? o.tail_paths : tail_default);
    auto launch_finish = [&](int depth, uint32_t* qin, uint32_t* cin) {
#define FIN2(FEAT, LK)                                                                                        \
    do {                                                                                                      \
        if (has_env) hipLaunchKernelGGL((k_finish<256, FEAT, LK, true>), dim3(G), dim3(256), 0, stream, DS, C, \
                                        L, P, depth, qin, cin, ray_stats, MB);                                     \
        else hipLaunchKernelGGL((k_finish<256, FEAT, LK, false>), dim3(G), dim3(256), 0, stream, DS, C, L, P, \
                                depth, qin, cin, ray_stats, MB);                                                   \
    } while (0)
#define FIN(FEAT)                         \
    do {                                  \
        if (s->has_lights) FIN2(FEAT, L_ALL); \
        else FIN2(FEAT, 0);               \
    } while (0)
        if (s->profile == 1) FIN(F_SPHERE);
        else if (s->profile == 3) FIN(F_ALL | F_UV);
        else FIN(F_ALL);
#undef FIN
#undef FIN2
    };

    if (graphs) {  // graphs captured earlier hold kernel arguments: drop them if any has changed
        std::vector<unsigned char> key(sizeof(mpt_camera));
        std::memcpy(key.data(), cam, sizeof(mpt_camera));
        const void* ptrs[] = {s->queues, s->counters, s->ovf, P.base, P.key};
        const int64_t vals[] = {cap, int64_t(G), s->profile, s->light_profile, int64_t(s->mat_types), has_env,
                                s->has_lights};
        key.insert(key.end(), reinterpret_cast<const unsigned char*>(ptrs), reinterpret_cast<const unsigned char*>(ptrs) + sizeof(ptrs));
        key.insert(key.end(), reinterpret_cast<const unsigned char*>(vals), reinterpret_cast<const unsigned char*>(vals) + sizeof(vals));
        if (key != s->graph_key) {
            s->drop_graphs();
            s->graph_key = key;
        }
        if (int(s->graphs.size()) < C.max_depth) s->graphs.resize(size_t(C.max_depth), nullptr);
    }
    // one bounce's kernels: extend, shade per material class present, shadow
    // (or the class-counter reset); extend zeroes the next-queue and shadow
    // counters, shadow zeroes the class counters once every shade kernel read them
    auto launch_bounce = [&](int depth, uint32_t* qin, uint32_t* qout, uint32_t* cin, uint32_t* cout) {
        launch_extend(depth, qin, cin, cout);
        for (int cls = 0; cls < kClasses; cls++) {
            if (cls > 0 && !(s->mat_types & (1u << (cls - 1)))) continue;
            launch_shade(cls, depth, qout, cout);
        }
        if (s->has_lights) launch_shadow(depth);
        else hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, ccls, kClasses);
    };
    // the bounce as a graph: captured at its first use, replayed after
    auto graph_bounce = [&](int depth, uint32_t* qin, uint32_t* qout, uint32_t* cin, uint32_t* cout) -> int {
        hipGraphExec_t& ge = s->graphs[size_t(depth)];
        if (!ge) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            launch_bounce(depth, qin, qout, cin, cout);
            const hipError_t e = hipStreamEndCapture(stream, &g);
            if (e != hipSuccess || !g) return set_error(MPT_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                ge = nullptr;
                return set_error(MPT_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
            }
        }
        HIP_TRY(hipGraphLaunch(ge, stream));
        return MPT_OK;
    };

    int pass = 0;
    while (pass < total) {
        int nb = std::min(pif, total - pass);
        if (o.save_every > 0 && on_pass) {
            int done_abs = done_before + pass;
            int next_save = (done_abs / o.save_every + 1) * o.save_every;
            nb = std::min(nb, next_save - done_abs);
        }
        batch_args B{o.seed, uint32_t(o.scene_index), o.first_pass + pass, npix, nb * npix};
        R.salt = uint32_t(mix64(o.seed ^ (uint64_t(uint32_t(o.first_pass + pass)) << 32)));
        EV(3, hipLaunchKernelGGL(k_camera<false>, dim3(unsigned((B.nslots + 255) / 256)), dim3(256), 0, stream, C, P, B,
                                 q[0], cnt, int(s->has_lights), MB));
        // passes of this batch whose albedo and normal guides are traced (the leading ones)
        const int guide_nb = o.guide_passes > 0 ? std::max(0, std::min(nb, o.guide_passes - (o.first_pass + pass))) : nb;
        batch_args BG = B;
        BG.nslots = guide_nb * npix;  // slots run pass-major
        // the guides' walk, compiled for the scene's geometry as the extension kernels are
        auto launch_features = [&](auto first) {
            constexpr bool F = decltype(first)::value;
            if (s->profile == 1) EV(3, hipLaunchKernelGGL((k_features<256, F_SPHERE, F>), dim3(G), dim3(256), 0, stream, DS, L, P, BG));
            else if (s->profile == 2) EV(3, hipLaunchKernelGGL((k_features<256, kQuadProfile, F>), dim3(G), dim3(256), 0, stream, DS, L, P, BG));
            else if (s->profile == 3) EV(3, hipLaunchKernelGGL((k_features<256, F_ALL | F_UV, F>), dim3(G), dim3(256), 0, stream, DS, L, P, BG));
            else EV(3, hipLaunchKernelGGL((k_features<256, F_ALL, F>), dim3(G), dim3(256), 0, stream, DS, L, P, BG));
            return int(MPT_OK);  // EV returns its error, the events already freed
        };
        // where the camera rays' closest hits do not depend on the stream, the
        // guides start from the first extension's records (after it, before
        // shading rewrites the rays); else they trace the camera rays themselves
        const bool guides_after_extend = sorted && !graphs && !s->keyed_hits && C.max_depth > 0;
        if (C.features && guide_nb > 0 && !guides_after_extend && (rc = launch_features(std::false_type{}))) return rc;
        if (!sorted) EV(4, launch_finish(0, q[0], ctr(cnt, 0)));  // every bounce in one kernel, no binning
        for (int depth = 0; sorted && depth < C.max_depth; depth++) {
            uint32_t* qin = q[depth & 1];
            if (depth >= 2) {  // the queue length two bounces ago bounds this one's
                if (hipEventSynchronize(ring[(depth - 2) % kRing]) != hipSuccess) {
                    ev_free();
                    return set_error(MPT_E_HIP, "hipEventSynchronize failed");
                }
                const uint32_t left = uint32_t(s->pinned_counts[(depth - 2) % kRing]);
                if (left == 0) break;
                if (tail > 0 && left <= uint32_t(tail) && depth > R.depth) {
                    EV(4, launch_finish(depth, qin, ctr(cnt, depth & 1)));
                    break;
                }
            }
            uint32_t* qout = q[(depth + 1) & 1];
            uint32_t* cin = ctr(cnt, depth & 1);
            uint32_t* cout = ctr(cnt, (depth + 1) & 1);
            if (graphs) {
                if ((rc = graph_bounce(depth, qin, qout, cin, cout))) {
                    ev_free();
                    return rc;
                }
                HIP_TRY(hipMemcpyAsync(&s->pinned_counts[depth % kRing], cout, sizeof(int32_t), hipMemcpyDeviceToHost,
                                       stream));
                HIP_TRY(hipEventRecord(ring[depth % kRing], stream));
                continue;
            }
            // extend zeroes the next-queue and shadow counters of this bounce;
            // shadow zeroes the class counters once every shade kernel read them
            EV(0, launch_extend(depth, qin, cin, cout));
            if (depth == 0 && C.features && guide_nb > 0 && guides_after_extend &&
                (rc = launch_features(std::true_type{})))
                return rc;
            if (nrc && depth == R.depth) {  // the cache's inputs, then paths not training end with its radiance
                hipLaunchKernelGGL((k_nrc_prep<256>), dim3(G), dim3(256), 0, stream, DS, P, R.q, R.cnt, R, NB);
                hipLaunchKernelGGL(k_nrc_record, dim3(unsigned((NB.tcap + 255) / 256)), dim3(256), 0, stream, P, R.q,
                                   NB);
                uint32_t nq = 0;
                HIP_TRY(hipMemcpyAsync(&nq, R.cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
                HIP_TRY(hipStreamSynchronize(stream));
                if (nstats) nstats->queries += nq;
                if (R.cut && nq > 0) {
                    if ((rc = nrc_infer(nrc, NB.pos, NB.other, int64_t(nq), NB.pred, stream))) {
                        ev_free();
                        return rc;
                    }
                    hipLaunchKernelGGL(k_nrc_apply, dim3(unsigned((nq + 255) / 256)), dim3(256), 0, stream, P, R.q,
                                       R.cnt, R, NB.pred);
                }
            }
            for (int cls = 0; cls < kClasses; cls++) {
                if (cls > 0 && !(s->mat_types & (1u << (cls - 1)))) continue;
                EV(1, launch_shade(cls, depth, qout, cout));
            }
            if (s->has_lights) EV(2, launch_shadow(depth));
            else EV(2, hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, ccls, kClasses));
            HIP_TRY(hipMemcpyAsync(&s->pinned_counts[depth % kRing], cout, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipEventRecord(ring[depth % kRing], stream));
        }
        if (nrc && nopts->train && nopts->train_steps > 0) {  // train on this batch's path suffixes
            hipLaunchKernelGGL(k_nrc_collect, dim3(unsigned((NB.tcap + 255) / 256)), dim3(256), 0, stream, P, NB);
            uint32_t nt = 0;
            HIP_TRY(hipMemcpyAsync(&nt, NB.tcnt, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            const int64_t ntrain = std::min<int64_t>(nt, NB.tcap);
            double loss = 0;
            for (int k = 0; k < nopts->train_steps && ntrain > 0; k++) {
                if ((rc = nrc_train(nrc, NB.tpos, NB.tother, NB.ttarget, ntrain, nopts->lr, stream,
                                    k + 1 == nopts->train_steps ? &loss : nullptr))) {
                    ev_free();
                    return rc;
                }
            }
            if (nstats) {
                nstats->train_samples += ntrain;
                if (ntrain > 0) nstats->loss = loss;
            }
        }
        EV(3, hipLaunchKernelGGL(k_accumulate, dim3(unsigned((npix + 255) / 256)), dim3(256), 0, stream, P, npix, nb,
                                 C.features, guide_nb, acc_rad, acc_alb, acc_nrm, acc_l2));
        if (C.features) s->guide_done += guide_nb;
        if (hipGetLastError() != hipSuccess) {
            ev_free();
            return set_error(MPT_E_HIP, "kernel launch failed");
        }
        if ((rc = ev_collect())) {
            ev_free();
            return rc;
        }
        pass += nb;
        s->last_done = done_before + pass;
        if (on_pass && o.save_every > 0 && s->last_done % o.save_every == 0 && pass < total) {
            HIP_TRY(hipStreamSynchronize(stream));
            on_pass(s->last_done, user);
        }
    }
#undef EV
    ev_free();
    HIP_TRY(hipStreamSynchronize(stream));
    auto t1 = std::chrono::steady_clock::now();
    if (radiance || albedo || normal || variance) {
        if ((rc = mpt_read_frame(s, radiance, albedo, normal, variance))) return rc;
    }
    if (stats) {
        unsigned long long rs_host[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        HIP_TRY(hipMemcpy(rs_host, ray_stats, sizeof(rs_host), hipMemcpyDeviceToHost));
        std::memset(stats, 0, sizeof(*stats));
        stats->paths = npix64 * total;
        stats->extension_rays = int64_t(rs_host[0]);
        stats->shadow_rays = int64_t(rs_host[1]);
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        for (int k = 0; k < 5; k++) {
            stats->stage_ms[k] = stage_ms[k];
            stats->stage_launches[k] = stage_n[k];
        }
        stats->node_bytes[0] = int64_t(rs_host[2]);
        stats->prim_bytes[0] = int64_t(rs_host[3]);
        stats->node_bytes[1] = int64_t(rs_host[4]);
        stats->prim_bytes[1] = int64_t(rs_host[5]);
        stats->tail_rays[0] = int64_t(rs_host[6]);
        stats->tail_rays[1] = int64_t(rs_host[7]);
    }
    return MPT_OK;
}

}  // namespace

namespace {

// Several scenes in one batch of launches (mpt_render_batch). The first
// scene's path state, queues, counters, stack spill area and stream carry
// the batch; each scene keeps its own running sums. A batch holds `nb`
// passes of every scene; scene k's keys use scene index scene_index + k, so
// its paths are the ones a render of that scene alone would trace.
int render_batch_impl(mpt_scene* const* ss, const mpt_camera* cams, int n, const mpt_render_opts* opts,
                      mpt_render_stats* stats) {
    if (!ss || !cams || n < 1 || n > kMaxBatch)
        return set_error(MPT_E_INVALID, "render batch: 1 to " + std::to_string(kMaxBatch) + " scenes");
    if (!MPT_PATH_RECORDS) return set_error(MPT_E_UNSUPPORTED, "render batch: needs the path-record layout");
    mpt_render_opts o{};
    if (opts) o = *opts;
    if (o.first_pass < 0 || o.pass_count < 0) return set_error(MPT_E_INVALID, "negative first_pass or pass_count");
    if (o.save_every > 0) return set_error(MPT_E_UNSUPPORTED, "render batch: save_every is not supported");
    if (o.sort_by_material < 0) return set_error(MPT_E_UNSUPPORTED, "render batch: runs binned stages only");
    mpt_scene* s = ss[0];
    for (int k = 0; k < n; k++) {
        const mpt_scene* sk = ss[k];
        const mpt_camera* c = &cams[k];
        if (!sk) return set_error(MPT_E_INVALID, "render batch: null scene");
        for (int j = 0; j < k; j++)
            if (ss[j] == sk) return set_error(MPT_E_INVALID, "render batch: a scene appears twice");
        if (c->image_width < 1 || !(c->aspect_ratio > 0) || c->samples_per_pixel < 1 || c->max_depth < 1 ||
            c->max_depth > 1000 || c->max_diffuse < 0 || c->max_diffuse > 1000 || c->max_specular < 0 ||
            c->max_specular > 1000 || c->max_scatter < 0 || c->max_scatter > 1000)
            return set_error(MPT_E_INVALID, "render batch: bad camera settings for scene " + std::to_string(k));
        if (c->collect_features) return set_error(MPT_E_UNSUPPORTED, "render batch: feature guides are not supported");
        if (sk->profile == 3)
            return set_error(MPT_E_UNSUPPORTED, "render batch: scenes with image textures, maps, alpha masks, thin "
                                                "glass, measured metals, the ground plane or height fogs render one "
                                                "at a time");
        if (sk->device != s->device) return set_error(MPT_E_INVALID, "render batch: scenes on different devices");
        if (sk->light_profile != s->light_profile || (sk->hs.env_w > 0) != (s->hs.env_w > 0))
            return set_error(MPT_E_UNSUPPORTED, "render batch: scenes must share their light setup (same kind of "
                                                "lights, all with or all without an environment)");
    }
    auto t0 = std::chrono::steady_clock::now();
    std::vector<dcam> C(static_cast<size_t>(n));
    std::vector<int64_t> npix(static_cast<size_t>(n));
    int64_t pix_total = 0;
    int total = -1;
    uint32_t mat_types = 0;
    for (int k = 0; k < n; k++) {
        C[k] = init_camera(cams[k], ss[k]->hs.env_w > 0);
        npix[k] = int64_t(C[k].w) * C[k].h;
        pix_total += npix[k];
        const int tk = o.pass_count > 0 ? o.pass_count : C[k].spp - o.first_pass;
        if (tk < 1) return set_error(MPT_E_INVALID, "no passes to render (first_pass >= samples_per_pixel)");
        if (total >= 0 && tk != total)
            return set_error(MPT_E_INVALID, "render batch: every scene renders the same number of passes");
        total = tk;
        mat_types |= ss[k]->mat_types;
    }
    if (pix_total > (int64_t(1) << 28)) return set_error(MPT_E_CAPACITY, "render batch: too many pixels");
    int pif = o.passes_in_flight > 0 ? o.passes_in_flight : int(std::max<int64_t>(1, (int64_t(16) << 20) / pix_total));
    pif = std::min(pif, total);
    if (int64_t(pif) * pix_total > (int64_t(1) << 31)) pif = int((int64_t(1) << 31) / pix_total);
    int rc;
    if ((rc = ensure_paths(s, int64_t(pif) * pix_total))) return rc;
    const int bpc = o.blocks_per_cu > 0 ? o.blocks_per_cu : MPT_GENERAL_WAVES;  // the batch kernels' one round
    if (bpc > 64) return set_error(MPT_E_INVALID, "blocks_per_cu must be in [1, 64]");
    if (s->grid != s->cu_count * bpc) {
        s->grid = s->cu_count * bpc;
        if ((rc = ensure_ovf(s))) return rc;
    }
    for (int k = 1; k < n; k++) {  // the spill area must hold the deepest scene's stack
        const int64_t threads = int64_t(s->grid) * s->block;
        ss[k]->grid = s->grid;
        if ((rc = ensure_ovf(ss[k]))) return rc;
        if (ss[k]->ovf_depth > s->ovf_depth) {
            HIP_TRY(hipFree(s->ovf));
            s->ovf = nullptr;
            HIP_TRY(hipMalloc(&s->ovf, size_t(ss[k]->ovf_depth) * threads * sizeof(int32_t)));
            s->ovf_depth = ss[k]->ovf_depth;
            s->ovf_threads = threads;
        }
    }
    // each scene's sums, and the scene and camera tables the kernels read
    for (int k = 0; k < n; k++) {
        mpt_scene* sk = ss[k];
        const bool keep = o.accumulate && sk->acc && sk->acc_pix == npix[k] && sk->last_done > 0 &&
                          !sk->last_cam.collect_features;
        if (o.accumulate && !keep && sk->last_done > 0)
            return set_error(MPT_E_INVALID, "accumulate: the previous render had another size or feature setting");
        if (sk->acc_pix != npix[k] || !sk->acc) {
            if (sk->acc) HIP_TRY(hipFree(sk->acc));
            sk->acc = nullptr;
            HIP_TRY(hipMalloc(&sk->acc, size_t(npix[k]) * 10 * sizeof(double)));
            sk->acc_pix = npix[k];
        }
        if (!keep) {
            HIP_TRY(hipMemsetAsync(sk->acc, 0, size_t(npix[k]) * 10 * sizeof(double), s->stream));
            sk->last_done = 0;
        }
    }
    const size_t tab_bytes = size_t(n) * (sizeof(dscene) + sizeof(dcam));
    if (s->batch_tab_bytes < tab_bytes) {
        if (s->batch_tab) HIP_TRY(hipFree(s->batch_tab));
        s->batch_tab = nullptr;
        s->batch_tab_bytes = 0;
        HIP_TRY(hipMalloc(&s->batch_tab, tab_bytes));
        s->batch_tab_bytes = tab_bytes;
    }
    {
        std::vector<unsigned char> tab(tab_bytes);
        for (int k = 0; k < n; k++) std::memcpy(tab.data() + k * sizeof(dscene), &ss[k]->ds, sizeof(dscene));
        std::memcpy(tab.data() + n * sizeof(dscene), C.data(), n * sizeof(dcam));
        HIP_TRY(hipStreamSynchronize(s->stream));  // the table may still be read by the previous batch
        HIP_TRY(hipMemcpy(s->batch_tab, tab.data(), tab_bytes, hipMemcpyHostToDevice));
    }
    mbatch M{};
    M.S = static_cast<const dscene*>(s->batch_tab);
    M.C = reinterpret_cast<const dcam*>(static_cast<const unsigned char*>(s->batch_tab) + n * sizeof(dscene));
    M.n = n;
    for (int k = 0; k < n; k++) {
        M.npix[k] = int32_t(npix[k]);
        M.pbase[k + 1] = M.pbase[k] + uint32_t(npix[k]);
        M.acc[k] = ss[k]->acc;
    }

    const int64_t cap = s->ps_capacity;
    uint32_t* q[3] = {s->queues, s->queues + cap, s->queues + 2 * cap};
    uint32_t* qcls = s->queues + 3 * cap;
    uint32_t* cnt = s->counters;
    uint32_t* ccls = ctr(cnt, 3);
    launch_ctx L{s->ovf, int64_t(s->grid) * s->block};
    const pstate P = s->ps;
    unsigned long long* ray_stats = reinterpret_cast<unsigned long long*>(s->counters + kCtrWords);
    unsigned long long* work = ray_stats + 2;
    HIP_TRY(hipMemsetAsync(ray_stats, 0, 8 * sizeof(unsigned long long), s->stream));
    const bool timing = (o.profile & 1) != 0;
    const bool counting = (o.profile & 2) != 0;
    const bool has_env = s->hs.env_w > 0;
    const bool has_lights = s->has_lights;
    const int light_profile = s->light_profile;
    const unsigned G = unsigned(s->grid);
    hipStream_t stream = s->stream;
    const dscene DS = s->ds;  // unused by the batch kernels (they read M.S)
    bool all_occlusion = true;
    for (int k = 0; k < n; k++) all_occlusion = all_occlusion && ss[k]->ds.shadow_occlusion != 0;
    const dcam C0 = C[0];
    const nrc_route R{-1, 0, 1, 0, nullptr, nullptr};

    // per-stage event timing, as render_impl
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evs;
    size_t ev_used = 0;
    double stage_ms[5] = {0, 0, 0, 0, 0};
    int64_t stage_n[5] = {0, 0, 0, 0, 0};
    constexpr int kRing = 4;
    hipEvent_t ring[kRing] = {};
    auto ev_free = [&]() {
        for (auto& e : evs) {
            (void)hipEventDestroy(e.second.first);
            (void)hipEventDestroy(e.second.second);
        }
        for (auto& e : ring)
            if (e) (void)hipEventDestroy(e);
    };
    struct guard_t {
        std::function<void()> f;
        ~guard_t() { f(); }
    } guard{ev_free};
    for (int k = 0; k < kRing; k++) HIP_TRY(hipEventCreateWithFlags(&ring[k], hipEventDisableTiming));
    auto ev_begin = [&](int stage) -> int {
        if (!timing) return MPT_OK;
        if (ev_used == evs.size()) {
            hipEvent_t a = nullptr, b = nullptr;
            HIP_TRY(hipEventCreate(&a));
            HIP_TRY(hipEventCreate(&b));
            evs.push_back({stage, {a, b}});
        }
        evs[ev_used].first = stage;
        HIP_TRY(hipEventRecord(evs[ev_used].second.first, stream));
        return MPT_OK;
    };
    auto ev_end = [&]() -> int {
        if (!timing) return MPT_OK;
        HIP_TRY(hipEventRecord(evs[ev_used].second.second, stream));
        ev_used++;
        return MPT_OK;
    };
    auto ev_collect = [&]() -> int {
        if (!timing) return MPT_OK;
        HIP_TRY(hipStreamSynchronize(stream));
        for (size_t k = 0; k < ev_used; k++) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, evs[k].second.first, evs[k].second.second));
            stage_ms[evs[k].first] += ms;
            stage_n[evs[k].first]++;
        }
        ev_used = 0;
        return MPT_OK;
    };
#define BEV(stage, launch)                    \
    do {                                      \
        if ((rc = ev_begin(stage))) return rc; \
        launch;                               \
        if ((rc = ev_end())) return rc;       \
    } while (0)

    // the batch kernels: general geometry, the scenes' shared light profile
    auto launch_extend = [&](int depth, uint32_t* qin, uint32_t* cin, uint32_t* cout) {
        if (counting)
            hipLaunchKernelGGL((k_extend<256, true, F_ALL, true>), dim3(G), dim3(256), 0, stream, DS, L, P, depth, qin,
                               cin, cout, ctr(cnt, 2), qcls, ccls, cap, ray_stats, work, R, M);
        else
            hipLaunchKernelGGL((k_extend<256, false, F_ALL, true>), dim3(G), dim3(256), 0, stream, DS, L, P, depth, qin,
                               cin, cout, ctr(cnt, 2), qcls, ccls, cap, ray_stats, work, R, M);
    };
    auto launch_shadow = [&](int depth) {
#define BSHD(CNT, ENV, OCC)                                                                                          \
    hipLaunchKernelGGL((k_shadow<256, CNT, F_ALL, ENV, true, OCC>), dim3(G), dim3(256), 0, stream, DS, C0, L, P, depth, \
                       q[2], ctr(cnt, 2), ccls, ray_stats, work, M)
#define BSHD2(ENV, OCC)                    \
    do {                                   \
        if (counting) BSHD(true, ENV, OCC); \
        else BSHD(false, ENV, OCC);        \
    } while (0)
        // occlusion queries only when no scene of the batch has an emitter,
        // a glow or a material that lets shadows through
        if (all_occlusion) {
            if (has_env) BSHD2(true, true);
            else BSHD2(false, true);
        } else {
            if (has_env) BSHD2(true, false);
            else BSHD2(false, false);
        }
#undef BSHD2
#undef BSHD
    };
    auto launch_shade = [&](int cls, int depth, uint32_t* qout, uint32_t* cout) {
        const uint32_t* qc = qcls + cls * cap;
        const uint32_t* cc = ccls + cls * kCtrStride;
#define BSHA2(MT, LK, ENV)                                                                                    \
    hipLaunchKernelGGL((k_shade<256, MT, LK, ENV, false, true>), dim3(G), dim3(256), 0, stream, DS, C0, P, depth, \
                       qc, cc, qout, cout, q[2], ctr(cnt, 2), M)
#define BSHA1(MT, LK)                               \
    do {                                            \
        if (MT < 0 && has_env) BSHA2(MT, LK, (MT < 0)); \
        else BSHA2(MT, LK, false);                  \
    } while (0)
#define BSHA(MT)                                  \
    do {                                          \
        switch (light_profile) {                  \
            case 0: BSHA1(MT, 0); break;          \
            case 1: BSHA1(MT, L_QUAD); break;     \
            case 2: BSHA1(MT, L_ENV); break;      \
            default: BSHA1(MT, L_ALL); break;     \
        }                                         \
    } while (0)
        switch (cls) {
            case 0: BSHA(-1); break;
            case 1 + MPT_MAT_LAMBERTIAN: BSHA(MPT_MAT_LAMBERTIAN); break;
            case 1 + MPT_MAT_METAL: BSHA(MPT_MAT_METAL); break;
            case 1 + MPT_MAT_DIELECTRIC: BSHA(MPT_MAT_DIELECTRIC); break;
            case 1 + MPT_MAT_PRINCIPLED: BSHA(MPT_MAT_PRINCIPLED); break;
            case 1 + MPT_MAT_DIFFUSE_LIGHT: BSHA(MPT_MAT_DIFFUSE_LIGHT); break;
            default: BSHA(MPT_MAT_PHASE); break;
        }
#undef BSHA
#undef BSHA1
#undef BSHA2
    };
    auto launch_finish = [&](int depth, uint32_t* qin, uint32_t* cin) {
        if (has_lights) {
            if (has_env) hipLaunchKernelGGL((k_finish<256, F_ALL, L_ALL, true, true>), dim3(G), dim3(256), 0, stream, DS, C0, L, P, depth, qin, cin, ray_stats, M);
            else hipLaunchKernelGGL((k_finish<256, F_ALL, L_ALL, false, true>), dim3(G), dim3(256), 0, stream, DS, C0, L, P, depth, qin, cin, ray_stats, M);
        } else {
            if (has_env) hipLaunchKernelGGL((k_finish<256, F_ALL, 0, true, true>), dim3(G), dim3(256), 0, stream, DS, C0, L, P, depth, qin, cin, ray_stats, M);
            else hipLaunchKernelGGL((k_finish<256, F_ALL, 0, false, true>), dim3(G), dim3(256), 0, stream, DS, C0, L, P, depth, qin, cin, ray_stats, M);
        }
    };
    int max_depth = 0;
    for (int k = 0; k < n; k++) max_depth = std::max(max_depth, C[k].max_depth);
    const int64_t tail = o.tail_paths < 0 ? 0 : (o.tail_paths > 0 ? o.tail_paths : 2 * L.nthreads);

    int pass = 0;
    while (pass < total) {
        const int nb = std::min(pif, total - pass);
        for (int k = 0; k <= n; k++) M.base[k] = uint32_t(int64_t(nb) * M.pbase[k]);
        batch_args B{o.seed, uint32_t(o.scene_index), o.first_pass + pass, 0, int32_t(M.base[n])};
        BEV(3, hipLaunchKernelGGL(k_camera<true>, dim3(unsigned((B.nslots + 255) / 256)), dim3(256), 0, stream, C0, P,
                                  B, q[0], cnt, int(has_lights), M));
        for (int depth = 0; depth < max_depth; depth++) {
            uint32_t* qin = q[depth & 1];
            if (depth >= 2) {  // the queue length two bounces ago bounds this one's
                HIP_TRY(hipEventSynchronize(ring[(depth - 2) % kRing]));
                const uint32_t left = uint32_t(s->pinned_counts[(depth - 2) % kRing]);
                if (left == 0) break;
                if (tail > 0 && left <= uint32_t(tail)) {
                    BEV(4, launch_finish(depth, qin, ctr(cnt, depth & 1)));
                    break;
                }
            }
            uint32_t* qout = q[(depth + 1) & 1];
            uint32_t* cin = ctr(cnt, depth & 1);
            uint32_t* cout = ctr(cnt, (depth + 1) & 1);
            BEV(0, launch_extend(depth, qin, cin, cout));
            for (int cls = 0; cls < kClasses; cls++) {
                if (cls > 0 && !(mat_types & (1u << (cls - 1)))) continue;
                BEV(1, launch_shade(cls, depth, qout, cout));
            }
            if (has_lights) BEV(2, launch_shadow(depth));
            else BEV(2, hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, ccls, kClasses));
            HIP_TRY(hipMemcpyAsync(&s->pinned_counts[depth % kRing], cout, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipEventRecord(ring[depth % kRing], stream));
        }
        BEV(3, hipLaunchKernelGGL(k_accumulate_multi, dim3(unsigned((M.pbase[n] + 255) / 256)), dim3(256), 0, stream, P,
                                  nb, M));
        if (hipGetLastError() != hipSuccess) return set_error(MPT_E_HIP, "kernel launch failed");
        if ((rc = ev_collect())) return rc;
        pass += nb;
    }
#undef BEV
    HIP_TRY(hipStreamSynchronize(stream));
    auto t1 = std::chrono::steady_clock::now();
    for (int k = 0; k < n; k++) {
        ss[k]->last_done += total;
        ss[k]->guide_done = ss[k]->last_done;
        ss[k]->last_cam = cams[k];
    }
    if (stats) {
        unsigned long long rs_host[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        HIP_TRY(hipMemcpy(rs_host, ray_stats, sizeof(rs_host), hipMemcpyDeviceToHost));
        std::memset(stats, 0, sizeof(*stats));
        stats->paths = pix_total * total;
        stats->extension_rays = int64_t(rs_host[0]);
        stats->shadow_rays = int64_t(rs_host[1]);
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        for (int k = 0; k < 5; k++) {
            stats->stage_ms[k] = stage_ms[k];
            stats->stage_launches[k] = stage_n[k];
        }
        stats->node_bytes[0] = int64_t(rs_host[2]);
        stats->prim_bytes[0] = int64_t(rs_host[3]);
        stats->node_bytes[1] = int64_t(rs_host[4]);
        stats->prim_bytes[1] = int64_t(rs_host[5]);
        stats->tail_rays[0] = int64_t(rs_host[6]);
        stats->tail_rays[1] = int64_t(rs_host[7]);
    }
    return MPT_OK;
}

}  // namespace

extern "C" {

int32_t mpt_render(mpt_scene* s, const mpt_camera* cam, const mpt_render_opts* opts, double* radiance, double* albedo,
                   double* normal, double* variance, mpt_pass_callback on_pass, void* user, mpt_render_stats* stats) {
    return render_impl(s, cam, opts, radiance, albedo, normal, variance, on_pass, user, stats, nullptr, nullptr,
                       nullptr);
}

int32_t mpt_render_batch(mpt_scene* const* scenes, const mpt_camera* cameras, int32_t n, const mpt_render_opts* opts,
                         mpt_render_stats* stats) {
    return render_batch_impl(scenes, cameras, n, opts, stats);
}

void mpt_nrc_defaults(mpt_nrc_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->query_depth = 2;
    o->train_stride = 256;
    o->train_steps = 4;
    o->infer = 1;
    o->train = 1;
    o->lr = 1e-3f;
}

int32_t mpt_render_nrc(mpt_scene* s, const mpt_camera* cam, const mpt_render_opts* opts, mpt_nrc* nrc,
                       const mpt_nrc_opts* nopts, double* radiance, double* albedo, double* normal, double* variance,
                       mpt_pass_callback on_pass, void* user, mpt_render_stats* stats, mpt_nrc_stats* nstats) {
    if (!nrc) return set_error(MPT_E_INVALID, "null radiance cache");
    mpt_nrc_opts o;
    if (nopts) o = *nopts;
    else mpt_nrc_defaults(&o);
    if (o.query_depth < 1 || o.train_stride < 1 || o.train_steps < 0 || !(o.lr >= 0))
        return set_error(MPT_E_INVALID, "nrc: query_depth >= 1, train_stride >= 1, train_steps >= 0, lr >= 0");
    return render_impl(s, cam, opts, radiance, albedo, normal, variance, on_pass, user, stats, nrc, &o, nstats);
}

int32_t mpt_autotune(mpt_scene* s, const mpt_camera* cam, int32_t trial_passes, mpt_autotune_result* out) {
    if (!s || !cam || !out) return set_error(MPT_E_INVALID, "null argument");
    if (trial_passes < 1) trial_passes = 8;
    const dcam C = init_camera(*cam, s->hs.env_w > 0);
    const int64_t npix = int64_t(C.w) * C.h;
    // candidates: paths in flight (passes per batch) around 4, 8, 16, 32 M
    // paths, and the batch tail's size in paths per thread of the default grid
    std::vector<int> pifs;
    for (int64_t target : {int64_t(4) << 20, int64_t(8) << 20, int64_t(16) << 20, int64_t(32) << 20}) {
        int v = int(std::max<int64_t>(1, target / npix));
        v = std::min(v, trial_passes);
        if (int64_t(v) * npix > (int64_t(1) << 31)) continue;
        if (std::find(pifs.begin(), pifs.end(), v) == pifs.end()) pifs.push_back(v);
    }
    const int bpc0 = default_blocks_per_cu(s);
    const int64_t nthreads = int64_t(s->cu_count) * bpc0 * 256;
    const int64_t tails[] = {-1, nthreads / 2, nthreads, 2 * nthreads, 4 * nthreads};
    mpt_render_opts o{};
    o.seed = 0x5EED;
    o.pass_count = trial_passes;
    o.blocks_per_cu = bpc0;
    o.sort_by_material = 1;
    o.use_graph = -1;
    s->tuned_pix = 0;
    mpt_autotune_result best{};
    best.seconds_per_pass = 1e300;
    out->trials = 0;
    auto trial = [&]() -> int {
        auto t0 = std::chrono::steady_clock::now();
        int rc = render_impl(s, cam, &o, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr);
        if (rc) return rc;
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / trial_passes;
        out->trials++;
        if (dt < best.seconds_per_pass) {
            best.seconds_per_pass = dt;
            best.passes_in_flight = o.passes_in_flight;
            best.tail_paths = o.tail_paths;
            best.blocks_per_cu = o.blocks_per_cu;
            best.sort_by_material = o.sort_by_material;
            best.use_graph = o.use_graph;
        }
        return MPT_OK;
    };
    // a warm-up render sizes the buffers for the largest batch
    o.passes_in_flight = *std::max_element(pifs.begin(), pifs.end());
    int rc = render_impl(s, cam, &o, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr);
    if (rc) return rc;
    for (int pif : pifs)
        for (int64_t tail : tails) {
            o.passes_in_flight = pif;
            o.tail_paths = tail < 0 ? -1 : int32_t(std::min<int64_t>(tail, INT32_MAX));
            if ((rc = trial())) return rc;
        }
    // from the best: the grid, then no binning, then graphs
    o.passes_in_flight = best.passes_in_flight;
    o.tail_paths = int32_t(best.tail_paths);
    for (int bpc : {8, 16, 32}) {
        if (bpc == bpc0) continue;
        o.blocks_per_cu = bpc;
        if ((rc = trial())) return rc;
    }
    o.blocks_per_cu = best.blocks_per_cu;
    o.sort_by_material = -1;
    if ((rc = trial())) return rc;
    o.sort_by_material = 1;
    o.use_graph = 1;
    if ((rc = trial())) return rc;  // the first render captures; time the replays
    if ((rc = trial())) return rc;
    best.trials = out->trials;
    *out = best;
    return mpt_autotune_apply(s, cam, out);
}

int32_t mpt_autotune_apply(mpt_scene* s, const mpt_camera* cam, const mpt_autotune_result* r) {
    if (!s || !cam || !r) return set_error(MPT_E_INVALID, "null argument");
    if (r->passes_in_flight < 1 || r->blocks_per_cu < 1 || r->blocks_per_cu > 64)
        return set_error(MPT_E_INVALID, "autotune settings out of range");
    const dcam C = init_camera(*cam, s->hs.env_w > 0);
    s->tuned_pix = int64_t(C.w) * C.h;
    s->tuned_pif = r->passes_in_flight;
    s->tuned_tail = r->tail_paths < 0 ? -1 : r->tail_paths;
    s->tuned_bpc = r->blocks_per_cu;
    s->tuned_sort = r->sort_by_material < 0 ? -1 : 1;
    s->tuned_graph = r->use_graph > 0 ? 1 : -1;
    return MPT_OK;
}

int32_t mpt_autotune_key(const mpt_scene* s, const mpt_camera* cam, char* buf, int32_t len) {
    if (!s || !cam || !buf || len < 1) return set_error(MPT_E_INVALID, "null argument");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, s->device));
    const dcam C = init_camera(*cam, s->hs.env_w > 0);
    std::snprintf(buf, size_t(len), "%s/%dcu/geometry%d/materials%x/lights%d/env%d/%dx%d", prop.gcnArchName,
                  s->cu_count, s->profile, s->mat_types, s->light_profile, s->hs.env_w > 0 ? 1 : 0, C.w, C.h);
    return MPT_OK;
}

}  // extern "C"

