// geom.hpp - double-precision vector, interval and box arithmetic shared by
// the host (scene flattening, BVH build, tables) and the device kernels.
//
// Every operation keeps the operand order of the reference's math core so
// that results are bit-identical when compiled without FP contraction:
//   vec3 ops            vec3.h:75-113  (note v / t is (1 / t) * v)
//   interval            interval.h:8-48
//   aabb                aabb.h:7-88    (padding 1e-4, slab test)
//   onb                 onb.h:7-32
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIP__)
#define MPT_HD __host__ __device__ __forceinline__
#else
#define MPT_HD inline
#endif

namespace mpt {

constexpr double kPi = 3.1415926535897932385;  // rtweekend.h:11
constexpr double kInf = __builtin_huge_val();

struct v3 {
    double x, y, z;
};

MPT_HD v3 mk(double x, double y, double z) { return v3{x, y, z}; }
MPT_HD double at(const v3& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
MPT_HD v3 operator+(const v3& a, const v3& b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MPT_HD v3 operator-(const v3& a, const v3& b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MPT_HD v3 operator-(const v3& a) { return v3{-a.x, -a.y, -a.z}; }
MPT_HD v3 operator*(const v3& a, const v3& b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
MPT_HD v3 operator*(double t, const v3& a) { return v3{t * a.x, t * a.y, t * a.z}; }
MPT_HD v3 operator*(const v3& a, double t) { return v3{t * a.x, t * a.y, t * a.z}; }
// The reference divides by multiplying with the reciprocal (vec3.h:95-97).
MPT_HD v3 operator/(const v3& a, double t) { return (1 / t) * a; }
MPT_HD double dot(const v3& a, const v3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MPT_HD v3 cross(const v3& u, const v3& v) {
    return v3{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
MPT_HD double length_squared(const v3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
MPT_HD double length(const v3& a) { return std::sqrt(length_squared(a)); }
MPT_HD v3 unit_vector(const v3& a) { return a / length(a); }
MPT_HD bool near_zero(const v3& a) {
    const double s = 1e-8;
    return std::fabs(a.x) < s && std::fabs(a.y) < s && std::fabs(a.z) < s;
}
MPT_HD v3 reflect(const v3& v, const v3& n) { return v - 2 * dot(v, n) * n; }  // vec3.h:148-150
MPT_HD v3 refract(const v3& uv, const v3& n, double etai_over_etat) {      // vec3.h:154-159
    double cos_theta = std::fmin(dot(-uv, n), 1.0);
    v3 r_out_perp = etai_over_etat * (uv + cos_theta * n);
    v3 r_out_parallel = -std::sqrt(std::fabs(1.0 - length_squared(r_out_perp))) * n;
    return r_out_perp + r_out_parallel;
}
MPT_HD double luminance(const v3& c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
MPT_HD double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
MPT_HD double deg2rad(double d) { return d * kPi / 180.0; }
// sin and cos of one angle; the device computes both from one argument reduction.
MPT_HD void sin_cos(double x, double& s, double& c) {
#if defined(__HIP_DEVICE_COMPILE__)
    sincos(x, &s, &c);
#else
    s = std::sin(x);
    c = std::cos(x);
#endif
}
// sin and cos of 2 pi u for u in [0, 1): the reference's sin(2 * pi * u). The
// device evaluates sin(pi * 2u) with a bounded-argument reduction (2u is
// exact), which differs from rounding 2 pi u first by about an ulp; the host
// keeps the reference's expression (the GGX table must match it bit for bit).
MPT_HD void sin_cos_2pi(double u, double& s, double& c) {
#if defined(__HIP_DEVICE_COMPILE__)
    sincospi(2 * u, &s, &c);
#else
    double phi = 2 * kPi * u;
    s = std::sin(phi);
    c = std::cos(phi);
#endif
}

// interval.h: default-constructed intervals are empty (+inf, -inf).
struct span {
    double min, max;
};
MPT_HD span iv_empty() { return span{kInf, -kInf}; }
MPT_HD span iv_merge(const span& a, const span& b) {
    return span{a.min <= b.min ? a.min : b.min, a.max >= b.max ? a.max : b.max};
}
MPT_HD double iv_size(const span& a) { return a.max - a.min; }
MPT_HD span iv_expand(const span& a, double delta) {
    double padding = delta / 2;
    return span{a.min - padding, a.max + padding};
}

struct box3 {
    span ax[3];
};
MPT_HD box3 box_empty() { return box3{{iv_empty(), iv_empty(), iv_empty()}}; }
// aabb(interval x, y, z) and aabb(point a, point b) pad thin sides (aabb.h:83-88).
MPT_HD void box_pad(box3& b) {
    const double delta = 0.0001;
    for (int k = 0; k < 3; k++)
        if (iv_size(b.ax[k]) < delta) b.ax[k] = iv_expand(b.ax[k], delta);
}
MPT_HD box3 box_points(const v3& a, const v3& b) {
    box3 r;
    r.ax[0] = a.x <= b.x ? span{a.x, b.x} : span{b.x, a.x};
    r.ax[1] = a.y <= b.y ? span{a.y, b.y} : span{b.y, a.y};
    r.ax[2] = a.z <= b.z ? span{a.z, b.z} : span{b.z, a.z};
    box_pad(r);
    return r;
}
MPT_HD box3 box_merge(const box3& a, const box3& b) {
    return box3{{iv_merge(a.ax[0], b.ax[0]), iv_merge(a.ax[1], b.ax[1]), iv_merge(a.ax[2], b.ax[2])}};
}
MPT_HD v3 box_centroid(const box3& b) {
    return v3{(b.ax[0].min + b.ax[0].max) / 2, (b.ax[1].min + b.ax[1].max) / 2,
              (b.ax[2].min + b.ax[2].max) / 2};
}
MPT_HD double box_area(const box3& b) {
    double dx = iv_size(b.ax[0]), dy = iv_size(b.ax[1]), dz = iv_size(b.ax[2]);
    return 2 * (dx * dy + dy * dz + dz * dx);
}
MPT_HD int box_longest_axis(const box3& b) {
    double sx = iv_size(b.ax[0]), sy = iv_size(b.ax[1]), sz = iv_size(b.ax[2]);
    if (sx > sy) return sx > sz ? 0 : 2;
    return sy > sz ? 1 : 2;
}

// Orthonormal basis around n (onb.h:7-32).
struct onb {
    v3 u, v, w;
};
MPT_HD onb make_onb(const v3& n) {
    onb b;
    b.w = unit_vector(n);
    v3 a = (std::fabs(b.w.x) > 0.9) ? v3{0, 1, 0} : v3{1, 0, 0};
    b.v = unit_vector(cross(b.w, a));
    b.u = cross(b.w, b.v);
    return b;
}
MPT_HD v3 onb_transform(const onb& b, const v3& a) { return (a.x * b.u) + (a.y * b.v) + (a.z * b.w); }
MPT_HD v3 onb_to_local(const onb& b, const v3& a) { return v3{dot(a, b.u), dot(a, b.v), dot(a, b.w)}; }

// Affine transform with stored inverse (transform.h). m and inv_m are row-major.
struct xform {
    double m[9];
    double t[3];
    double im[9];
    double it[3];
};
MPT_HD v3 mat_mul(const double* a, const v3& v) {
    return v3{a[0] * v.x + a[1] * v.y + a[2] * v.z, a[3] * v.x + a[4] * v.y + a[5] * v.z,
              a[6] * v.x + a[7] * v.y + a[8] * v.z};
}
MPT_HD v3 xf_vector(const xform& x, const v3& v) { return mat_mul(x.m, v); }
MPT_HD v3 xf_point(const xform& x, const v3& p) { return xf_vector(x, p) + v3{x.t[0], x.t[1], x.t[2]}; }
MPT_HD v3 xf_inv_vector(const xform& x, const v3& v) { return mat_mul(x.im, v); }
MPT_HD v3 xf_inv_point(const xform& x, const v3& p) {
    return xf_inv_vector(x, p) + v3{x.it[0], x.it[1], x.it[2]};
}
// Normals transform by the inverse transpose (transform.h:73-77).
MPT_HD v3 xf_normal(const xform& x, const v3& n) {
    const double* a = x.im;
    return v3{a[0] * n.x + a[3] * n.y + a[6] * n.z, a[1] * n.x + a[4] * n.y + a[7] * n.z,
              a[2] * n.x + a[5] * n.y + a[8] * n.z};
}

}  // namespace mpt
