// oracle.cpp - CPU restatement of the reference renderer's hot path.
//
// TEST INFRASTRUCTURE. This is the checker the GPU path is compared with,
// and the CPU baseline the benchmark times beside it. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
// product library never does.
//
// It takes the same flattened scene description as the GPU core
// (include/mpt.h, mpt_scene_desc) and restates, in double precision and in
// the reference's object-oriented shape, what adriendjn/mypbr computes:
//   bvh build + hit         include/bvh.h:22-205, aabb.h:40-88
//   sphere / quad / tri     sphere.h:23-132, quad.h:13-79, mesh.h:47-139
//   instance, list          instance.h:29-56, hittable_list.h:25-62
//   constant_medium         medium.h:60-88
//   materials               material.h:74-264, 282-308, 346-615, 631-674, medium.h:13-39
//   textures                texture.h:24-208 (solid, checker, image)
//   microfacet / pdfs       microfacet.h:17-470, pdf.h:21-134
//   light_list, environment light_list.h:45-127, environment.h:66-195
//   camera                  camera.h:79-529
// Random numbers come from the counter-based stream the GPU path uses (a
// path's i-th draw is SplitMix64(key + i*gamma)), drawn in the same order,
// so a seeded CPU render and a seeded GPU render trace the same paths.
//
// Pinning (what checks this restatement, besides the GPU comparison):
//   - golden vectors written by the compiled reference (tests/golden):
//     closest-hit IDs and t on seeded ray batches, the GGX albedo table,
//     light selection probabilities, material eval/pdf values;
//   - closed forms: white furnace, pdf normalisation, Beer-Lambert slab;
//   - brute force: BVH closest hit == linear scan over all primitives.
// Built with -ffp-contract=off so its arithmetic is the reference's.
#include <algorithm>
#include <complex>
#include <atomic>
#include <chrono>
#include <cmath>
#include <functional>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../include/mpt.h"

namespace orc {

const double pi = 3.1415926535897932385;
const double infinity = __builtin_huge_val();

// ------------------------------------------------------------------ math (vec3.h, onb.h)
struct vec3 {
    double e[3];
    vec3() : e{0, 0, 0} {}
    vec3(double x, double y, double z) : e{x, y, z} {}
    double x() const { return e[0]; }
    double y() const { return e[1]; }
    double z() const { return e[2]; }
    double operator[](int i) const { return e[i]; }
    double& operator[](int i) { return e[i]; }
    vec3 operator-() const { return vec3(-e[0], -e[1], -e[2]); }
    double length_squared() const { return e[0] * e[0] + e[1] * e[1] + e[2] * e[2]; }
    double length() const { return std::sqrt(length_squared()); }
};
using point3 = vec3;
using color = vec3;
inline vec3 operator+(const vec3& a, const vec3& b) { return vec3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline vec3 operator-(const vec3& a, const vec3& b) { return vec3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline vec3 operator*(const vec3& a, const vec3& b) { return vec3(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
inline vec3 operator*(double t, const vec3& a) { return vec3(t * a[0], t * a[1], t * a[2]); }
inline vec3 operator*(const vec3& a, double t) { return vec3(t * a[0], t * a[1], t * a[2]); }
inline vec3 operator/(const vec3& a, double t) { return (1 / t) * a; }
inline double dot(const vec3& a, const vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline vec3 cross(const vec3& u, const vec3& v) {
    return vec3(u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]);
}
inline vec3 unit_vector(const vec3& v) { return v / v.length(); }
inline vec3 reflect(const vec3& v, const vec3& n) { return v - 2 * dot(v, n) * n; }
inline vec3 refract(const vec3& uv, const vec3& n, double etai_over_etat) {
    double cos_theta = std::fmin(dot(-uv, n), 1.0);
    vec3 perp = etai_over_etat * (uv + cos_theta * n);
    vec3 par = -std::sqrt(std::fabs(1.0 - perp.length_squared())) * n;
    return perp + par;
}
inline double luminance(const color& c) { return 0.2126 * c.x() + 0.7152 * c.y() + 0.0722 * c.z(); }
inline double dclamp(double x, double lo, double hi) { return x < lo ? lo : (hi < x ? hi : x); }

struct onb {
    vec3 axis[3];
    explicit onb(const vec3& n) {
        axis[2] = unit_vector(n);
        vec3 a = (std::fabs(axis[2].x()) > 0.9) ? vec3(0, 1, 0) : vec3(1, 0, 0);
        axis[1] = unit_vector(cross(axis[2], a));
        axis[0] = cross(axis[2], axis[1]);
    }
    const vec3& w() const { return axis[2]; }
    vec3 transform(const vec3& v) const { return (v[0] * axis[0]) + (v[1] * axis[1]) + (v[2] * axis[2]); }
    vec3 to_local(const vec3& v) const { return vec3(dot(v, axis[0]), dot(v, axis[1]), dot(v, axis[2])); }
};

// ------------------------------------------------------------------ random numbers
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
const uint64_t gamma64 = 0x9E3779B97F4A7C15ull;
inline double to_unit(uint64_t x) { return double(x >> 11) * 0x1.0p-53; }
inline uint64_t key_of(uint64_t seed, uint32_t scene, uint32_t sample, uint32_t pixel) {
    uint64_t s = mix64(seed + uint64_t(scene) * gamma64);
    return mix64(s ^ ((uint64_t(sample) << 32) | pixel));
}
struct sampler {
    uint64_t key = 0;
    uint32_t count = 0;
    double next() { return to_unit(mix64(key + uint64_t(++count) * gamma64)); }
    double next(double lo, double hi) { return lo + (hi - lo) * next(); }
};
inline double medium_u(uint64_t key, uint32_t serial, uint32_t obj) {
    uint64_t k = mix64(key ^ 0x6A09E667F3BCC909ull);
    return to_unit(mix64(k + ((uint64_t(serial) << 32) | obj) * gamma64));
}

// ------------------------------------------------------------------ rays, intervals, boxes
struct ray {
    point3 orig;
    vec3 dir;
    ray() {}
    ray(const point3& o, const vec3& d) : orig(o), dir(d) {}
    point3 at(double t) const { return orig + t * dir; }
};
struct interval {
    double min = +infinity, max = -infinity;
    interval() {}
    interval(double a, double b) : min(a), max(b) {}
    interval(const interval& a, const interval& b) : min(a.min <= b.min ? a.min : b.min), max(a.max >= b.max ? a.max : b.max) {}
    double size() const { return max - min; }
    bool contains(double x) const { return min <= x && x <= max; }
    bool surrounds(double x) const { return min < x && x < max; }
    interval expand(double d) const { return interval(min - d / 2, max + d / 2); }
};
struct aabb {
    interval ax[3];
    aabb() {}
    aabb(const point3& a, const point3& b) {
        for (int k = 0; k < 3; k++) ax[k] = a[k] <= b[k] ? interval(a[k], b[k]) : interval(b[k], a[k]);
        for (int k = 0; k < 3; k++)
            if (ax[k].size() < 0.0001) ax[k] = ax[k].expand(0.0001);
    }
    aabb(const aabb& a, const aabb& b) {
        for (int k = 0; k < 3; k++) ax[k] = interval(a.ax[k], b.ax[k]);
    }
    bool hit(const point3& o, const vec3& inv, interval t) const {
        for (int a = 0; a < 3; a++) {
            double t0 = (ax[a].min - o[a]) * inv[a];
            double t1 = (ax[a].max - o[a]) * inv[a];
            if (t0 > t1) std::swap(t0, t1);
            if (t0 > t.min) t.min = t0;
            if (t1 < t.max) t.max = t1;
            if (t.max <= t.min) return false;
        }
        return true;
    }
    point3 centroid() const {
        return point3((ax[0].min + ax[0].max) / 2, (ax[1].min + ax[1].max) / 2, (ax[2].min + ax[2].max) / 2);
    }
    double area() const {
        double dx = ax[0].size(), dy = ax[1].size(), dz = ax[2].size();
        return 2 * (dx * dy + dy * dz + dz * dx);
    }
    int longest_axis() const {
        if (ax[0].size() > ax[1].size()) return ax[0].size() > ax[2].size() ? 0 : 2;
        return ax[1].size() > ax[2].size() ? 1 : 2;
    }
};

// ------------------------------------------------------------------ hits
struct material;
struct hit_record {
    point3 p;
    vec3 normal;
    const material* mat = nullptr;
    double t = 0;
    bool front_face = false;
    int64_t id = -1;
    double u = 0, v = 0;  // texture coordinates and dp/du, dp/dv (hittable.h:14-24)
    vec3 dpdu, dpdv;
    void set_face_normal(const ray& r, const vec3& outward) {
        front_face = dot(r.dir, outward) < 0;
        normal = front_face ? outward : -outward;
    }
};

// Draws a hit() makes (constant_medium) come from the ray's own stream.
struct hit_ctx {
    uint64_t key = 0;
    uint32_t serial = 0;
    bool media = true;
};

struct hittable {
    virtual ~hittable() = default;
    virtual bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const = 0;
    virtual aabb bounding_box() const = 0;
    virtual double pdf_value(const point3&, const vec3&) const { return 0; }
    virtual vec3 random(const point3&, sampler&) const { return vec3(1, 0, 0); }
    virtual double power() const { return 0; }
};

// ------------------------------------------------------------------ materials
struct pdf {
    virtual ~pdf() = default;
    virtual double value(const vec3& d) const = 0;
    virtual vec3 generate(sampler& s) const = 0;
};

struct scatter_record {
    color attenuation;
    std::unique_ptr<pdf> pdf_ptr;
    bool skip_pdf = false;
    ray skip_pdf_ray;
};

struct ggx_table {  // microfacet.h:265-338, built once per scene
    double e[32][32];
    double e_avg[32];
};

struct material {
    const ggx_table* table = nullptr;
    virtual ~material() = default;
    virtual bool scatter(const ray&, const hit_record&, scatter_record&, sampler&) const { return false; }
    virtual color eval(const ray&, const hit_record&, const vec3&) const { return color(0, 0, 0); }
    virtual color emitted(const ray&, const hit_record&) const { return color(0, 0, 0); }
    virtual color interior_transmittance(const ray&, const hit_record&) const { return color(1, 1, 1); }
    virtual color average_emission() const { return color(0, 0, 0); }
    virtual bool lets_shadows_through() const { return false; }
    virtual bool is_volume() const { return false; }
};

// microfacet.h
double roughness_to_alpha(double r) { return std::max(r * r, 1e-3); }
double fresnel_schlick(double c, double f0) { return f0 + (1 - f0) * std::pow(1 - dclamp(c, 0.0, 1.0), 5); }
color fresnel_schlick(double c, const color& f0) {
    return color(fresnel_schlick(c, f0.x()), fresnel_schlick(c, f0.y()), fresnel_schlick(c, f0.z()));
}
double fresnel_dielectric(double cos_i, double eta) {
    cos_i = dclamp(cos_i, -1.0, 1.0);
    if (cos_i < 0) {
        eta = 1 / eta;
        cos_i = -cos_i;
    }
    double sin2_i = 1 - cos_i * cos_i;
    double sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1) return 1;
    double cos_t = std::sqrt(1 - sin2_t);
    double r_parl = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    double r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return (r_parl * r_parl + r_perp * r_perp) / 2;
}
struct ggx {
    double alpha;
    explicit ggx(double a) : alpha(a) {}
    double D(const vec3& h) const {
        if (h.z() <= 0) return 0;
        double a2 = alpha * alpha;
        double d = h.z() * h.z() * (a2 - 1) + 1;
        return a2 / (pi * d * d);
    }
    double lambda(const vec3& w) const {
        double cos2 = w.z() * w.z();
        if (cos2 <= 0) return infinity;
        double tan2 = std::fmax(0.0, 1 - cos2) / cos2;
        return (-1 + std::sqrt(1 + alpha * alpha * tan2)) / 2;
    }
    double G1(const vec3& w) const { return 1 / (1 + lambda(w)); }
    double G(const vec3& wo, const vec3& wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    double visible_pdf(const vec3& wo, const vec3& h) const {
        if (wo.z() <= 0) return 0;
        return G1(wo) / wo.z() * D(h) * std::fabs(dot(wo, h));
    }
    vec3 sample_visible(const vec3& wo, double u1, double u2) const {
        vec3 vh = unit_vector(vec3(alpha * wo.x(), alpha * wo.y(), wo.z()));
        double lensq = vh.x() * vh.x() + vh.y() * vh.y();
        vec3 t1 = lensq > 0 ? vec3(-vh.y(), vh.x(), 0) / std::sqrt(lensq) : vec3(1, 0, 0);
        vec3 t2 = cross(vh, t1);
        double r = std::sqrt(u1);
        double phi = 2 * pi * u2;
        double p1 = r * std::cos(phi);
        double p2 = r * std::sin(phi);
        double s = 0.5 * (1 + vh.z());
        p2 = (1 - s) * std::sqrt(std::fmax(0.0, 1 - p1 * p1)) + s * p2;
        vec3 nh = p1 * t1 + p2 * t2 + std::sqrt(std::fmax(0.0, 1 - p1 * p1 - p2 * p2)) * vh;
        return unit_vector(vec3(alpha * nh.x(), alpha * nh.y(), std::fmax(1e-6, nh.z())));
    }
};

void build_table(ggx_table& t) {
    const int size = 32, strata = 64;
    for (int j = 0; j < size; j++) {
        ggx dist(roughness_to_alpha(double(j) / (size - 1)));
        for (int i = 0; i < size; i++) {
            double mu = std::fmax(double(i) / (size - 1), 1e-3);
            vec3 wo(std::sqrt(1 - mu * mu), 0, mu);
            double sum = 0;
            for (int a = 0; a < strata; a++)
                for (int b = 0; b < strata; b++) {
                    vec3 h = dist.sample_visible(wo, (a + 0.5) / strata, (b + 0.5) / strata);
                    vec3 wi = -wo + 2 * dot(wo, h) * h;
                    if (wi.z() > 0) sum += dist.G(wo, wi) / dist.G1(wo);
                }
            t.e[j][i] = std::fmin(sum / (strata * strata), 1.0);
        }
        double integral = 0, h = 1.0 / (size - 1);
        for (int i = 0; i + 1 < size; i++) {
            double a = i * h, b = (i + 1) * h;
            integral += h * (t.e[j][i] * (2 * a + b) + t.e[j][i + 1] * (a + 2 * b)) / 6;
        }
        t.e_avg[j] = std::fmin(2 * integral, 1.0);
    }
}
double table_E(const ggx_table& t, double c, double alpha) {
    double x = dclamp(c, 0.0, 1.0) * 31, y = dclamp(std::sqrt(alpha), 0.0, 1.0) * 31;
    int i = std::min(int(x), 30), j = std::min(int(y), 30);
    double tx = x - i, ty = y - j;
    double lo = (1 - tx) * t.e[j][i] + tx * t.e[j][i + 1];
    double hi = (1 - tx) * t.e[j + 1][i] + tx * t.e[j + 1][i + 1];
    return (1 - ty) * lo + ty * hi;
}
double table_E_avg(const ggx_table& t, double alpha) {
    double y = dclamp(std::sqrt(alpha), 0.0, 1.0) * 31;
    int j = std::min(int(y), 30);
    double ty = y - j;
    return (1 - ty) * t.e_avg[j] + ty * t.e_avg[j + 1];
}
double ms_eval(const ggx_table& t, double co, double ci, double alpha) {
    if (co <= 0 || ci <= 0) return 0;
    double missing = 1 - table_E_avg(t, alpha);
    if (missing < 1e-6) return 0;
    return (1 - table_E(t, co, alpha)) * (1 - table_E(t, ci, alpha)) / (pi * missing) * ci;
}
double ms_albedo(const ggx_table& t, double co, double alpha) {
    if (co <= 0 || 1 - table_E_avg(t, alpha) < 1e-6) return 0;
    return 1 - table_E(t, co, alpha);
}
color ms_fresnel_avg(const ggx_table& t, const color& f_avg, double alpha) {
    double ea = table_E_avg(t, alpha);
    color f;
    for (int k = 0; k < 3; k++) f[k] = f_avg[k] * f_avg[k] * ea / (1 - f_avg[k] * (1 - ea));
    return f;
}
// fresnel_conductor (microfacet.h:57-66)
double fresnel_conductor(double cos_i, double eta, double k) {
    cos_i = std::clamp(cos_i, 0.0, 1.0);
    std::complex<double> n(eta, k);
    std::complex<double> sin2_t = (1 - cos_i * cos_i) / (n * n);
    std::complex<double> cos_t = std::sqrt(1.0 - sin2_t);
    std::complex<double> r_parl = (n * cos_i - cos_t) / (n * cos_i + cos_t);
    std::complex<double> r_perp = (cos_i - n * cos_t) / (cos_i + n * cos_t);
    return (std::norm(r_parl) + std::norm(r_perp)) / 2;
}

// metal_fresnel (microfacet.h:98-140): Schlick through f0, or a measured metal
struct metal_fresnel {
    bool measured = false;
    color eta, k, f0{1, 1, 1}, f_avg{1, 1, 1};
    static metal_fresnel schlick(const color& f0) {
        metal_fresnel f;
        f.f0 = f0;
        f.f_avg = f0 + (color(1, 1, 1) - f0) / 21;
        return f;
    }
    static metal_fresnel from_ior(const color& eta, const color& k) {
        metal_fresnel f;
        f.measured = true;
        f.eta = eta;
        f.k = k;
        f.f0 = f.at(1);
        const int n = 512;
        color sum(0, 0, 0);
        for (int i = 0; i < n; i++) {
            double mu = (i + 0.5) / n;
            sum = sum + f.at(mu) * mu;
        }
        f.f_avg = 2 * sum / n;
        return f;
    }
    color at(double c) const {
        if (!measured) return fresnel_schlick(c, f0);
        return color(fresnel_conductor(c, eta.x(), k.x()), fresnel_conductor(c, eta.y(), k.y()),
                     fresnel_conductor(c, eta.z(), k.z()));
    }
};

color ms_fresnel(const ggx_table& t, const color& f0, double alpha) {
    return ms_fresnel_avg(t, f0 + (color(1, 1, 1) - f0) / 21, alpha);
}
double average(const color& c) { return (c.x() + c.y() + c.z()) / 3; }

// pdf.h
struct cosine_pdf : pdf {
    onb uvw;
    explicit cosine_pdf(const vec3& w) : uvw(w) {}
    double value(const vec3& d) const override { return std::fmax(0, dot(unit_vector(d), uvw.w()) / pi); }
    vec3 generate(sampler& s) const override {
        double r1 = s.next();
        double r2 = s.next();
        double phi = 2 * pi * r1;
        return uvw.transform(vec3(std::cos(phi) * std::sqrt(r2), std::sin(phi) * std::sqrt(r2), std::sqrt(1 - r2)));
    }
};
struct hg_pdf : pdf {
    onb uvw;
    double g;
    hg_pdf(const vec3& w, double g) : uvw(w), g(g) {}
    static double phase(double c, double g) {
        double denom = 1 + g * g - 2 * g * c;
        return (1 - g * g) / (4 * pi * denom * std::sqrt(denom));
    }
    double value(const vec3& d) const override { return phase(dot(unit_vector(d), uvw.w()), g); }
    vec3 generate(sampler& s) const override {
        double xi = s.next();
        double c;
        if (std::fabs(g) < 1e-3) c = 1 - 2 * xi;
        else {
            double q = (1 - g * g) / (1 - g + 2 * g * xi);
            c = (1 + g * g - q * q) / (2 * g);
        }
        c = dclamp(c, -1.0, 1.0);
        double st = std::sqrt(1 - c * c);
        double phi = 2 * pi * s.next();
        return uvw.transform(vec3(std::cos(phi) * st, std::sin(phi) * st, c));
    }
};
struct ggx_reflection_pdf : pdf {
    onb frame;
    vec3 wo;
    ggx dist;
    ggx_reflection_pdf(const vec3& n, const vec3& wo_world, double alpha)
        : frame(n), wo(unit_vector(frame.to_local(wo_world))), dist(alpha) {}
    double value(const vec3& d) const override {
        if (d.length_squared() == 0) return 0;
        vec3 wi = unit_vector(frame.to_local(d));
        if (wi.z() <= 0 || wo.z() <= 0) return 0;
        vec3 h = unit_vector(wo + wi);
        return dist.visible_pdf(wo, h) / (4 * std::fabs(dot(wo, h)));
    }
    vec3 generate(sampler& s) const override {
        if (wo.z() <= 0) return vec3(0, 0, 0);
        double u1 = s.next();
        double u2 = s.next();
        vec3 h = dist.sample_visible(wo, u1, u2);
        vec3 wi = -wo + 2 * dot(wo, h) * h;
        if (wi.z() <= 0) return vec3(0, 0, 0);
        return frame.transform(wi);
    }
};
struct mixture_pdf : pdf {
    std::unique_ptr<pdf> p0, p1;
    double w;
    mixture_pdf(std::unique_ptr<pdf> a, std::unique_ptr<pdf> b, double w) : p0(std::move(a)), p1(std::move(b)), w(w) {}
    double value(const vec3& d) const override { return w * p0->value(d) + (1 - w) * p1->value(d); }
    vec3 generate(sampler& s) const override { return s.next() < w ? p0->generate(s) : p1->generate(s); }
};
struct rough_glass {  // microfacet.h:384-470
    onb frame;
    vec3 wo;
    double eta;
    ggx dist;
    rough_glass(const vec3& n, const vec3& wo_world, double eta, double alpha)
        : frame(n), wo(unit_vector(frame.to_local(wo_world))), eta(eta), dist(alpha) {}
    bool half_vector(const vec3& wiw, vec3& wi, vec3& wm, bool& refl) const {
        if (wiw.length_squared() == 0 || wo.z() <= 0) return false;
        wi = unit_vector(frame.to_local(wiw));
        if (wi.z() == 0) return false;
        refl = wi.z() > 0;
        wm = refl ? wo + wi : wi * eta + wo;
        if (wm.length_squared() == 0) return false;
        wm = unit_vector(wm);
        if (wm.z() < 0) wm = -wm;
        if (dot(wm, wi) * wi.z() < 0 || dot(wm, wo) * wo.z() < 0) return false;
        return true;
    }
    double eval(const vec3& wiw) const {
        vec3 wi, wm;
        bool refl;
        if (!half_vector(wiw, wi, wm, refl)) return 0;
        double F = fresnel_dielectric(dot(wo, wm), eta);
        if (refl) return F * dist.D(wm) * dist.G(wo, wi) / (4 * wo.z());
        double q = dot(wi, wm) + dot(wo, wm) / eta;
        double denom = q * q;
        double ft = (1 - F) * dist.D(wm) * dist.G(wo, wi) * std::fabs(dot(wi, wm) * dot(wo, wm) / (wi.z() * wo.z() * denom));
        ft /= eta * eta;
        return ft * std::fabs(wi.z());
    }
    double pdf_value(const vec3& wiw) const {
        vec3 wi, wm;
        bool refl;
        if (!half_vector(wiw, wi, wm, refl)) return 0;
        double F = fresnel_dielectric(dot(wo, wm), eta);
        if (refl) return F * dist.visible_pdf(wo, wm) / (4 * std::fabs(dot(wo, wm)));
        double q = dot(wi, wm) + dot(wo, wm) / eta;
        double denom = q * q;
        return (1 - F) * dist.visible_pdf(wo, wm) * std::fabs(dot(wi, wm)) / denom;
    }
    vec3 sample(sampler& s) const {
        if (wo.z() <= 0) return vec3(0, 0, 0);
        double u1 = s.next();
        double u2 = s.next();
        vec3 wm = dist.sample_visible(wo, u1, u2);
        double F = fresnel_dielectric(dot(wo, wm), eta);
        vec3 wi;
        if (s.next() < F) {
            wi = -wo + 2 * dot(wo, wm) * wm;
            if (wi.z() <= 0) return vec3(0, 0, 0);
        } else {
            wi = refract(-wo, wm, 1 / eta);
            if (wi.z() >= 0) return vec3(0, 0, 0);
        }
        return frame.transform(wi);
    }
};
struct rough_glass_pdf : pdf {
    rough_glass g;
    explicit rough_glass_pdf(const rough_glass& g) : g(g) {}
    double value(const vec3& d) const override { return g.pdf_value(d); }
    vec3 generate(sampler& s) const override { return g.sample(s); }
};

// solid_color, checker_texture and image_texture (texture.h:24-205)
struct texture_t {
    int type;
    double inv_scale;
    color even, odd;
    int width = 0, height = 0, flags = 0;
    const uint8_t* ldr = nullptr;       // RGBA bytes of this image
    const float* hdr = nullptr;         // or its RGB floats
    const double* srgb = nullptr;       // 256 linear values of the sRGB codes
    // wrappers (texture.h:212-282): what they read and their parameters
    std::shared_ptr<const texture_t> inner;
    int channel = 0;
    double scale = 1;
    double offset_u = 0, offset_v = 0, scale_u = 1, scale_v = 1, cos_r = 1, sin_r = 0;
    color value(double u, double v, const point3& p) const {
        switch (type) {
            case MPT_TEX_SCALED: return even * inner->value(u, v, p);
            case MPT_TEX_CHANNEL: {
                const double x = scale * inner->value(u, v, p)[channel];
                return color(x, x, x);
            }
            case MPT_TEX_UV_TRANSFORM: {
                const double su = scale_u * u, sv = scale_v * (1 - v);
                const double tu = cos_r * su + sin_r * sv + offset_u;
                const double tv = -sin_r * su + cos_r * sv + offset_v;
                return inner->value(tu, 1 - tv, p);
            }
            case MPT_TEX_NORMAL_SCALE: {
                const color c = inner->value(u, v, p);
                return color(0.5 + scale * (c.x() - 0.5), 0.5 + scale * (c.y() - 0.5), c.z());
            }
            default: break;
        }
        if (type == MPT_TEX_SOLID) return even;
        if (type == MPT_TEX_CHECKER) {
            int x = int(std::floor(inv_scale * p.x())), y = int(std::floor(inv_scale * p.y())),
                z = int(std::floor(inv_scale * p.z()));
            return (x + y + z) % 2 == 0 ? even : odd;
        }
        size_t k = pixel_index(u, v);
        if (hdr) return color(hdr[3 * k], hdr[3 * k + 1], hdr[3 * k + 2]);
        if (flags & MPT_TEXF_LINEAR) return color(ldr[4 * k], ldr[4 * k + 1], ldr[4 * k + 2]) / 255.0;
        return color(srgb[ldr[4 * k]], srgb[ldr[4 * k + 1]], srgb[ldr[4 * k + 2]]);
    }
    // image_texture::alpha_value (texture.h:130-136): the alpha byte, or the
    // first byte of an image without an alpha channel, over 255
    double alpha_value(double u, double v) const {
        return ldr[4 * pixel_index(u, v) + ((flags & MPT_TEXF_NO_ALPHA) ? 0 : 3)] / 255.0;
    }
    size_t pixel_index(double u, double v) const {  // texture.h:182-195
        if (flags & MPT_TEXF_REPEAT) {
            u -= std::floor(u);
            v -= std::floor(v);
        } else {
            u = std::fmin(std::fmax(u, 0.0), 1.0);
            v = std::fmin(std::fmax(v, 0.0), 1.0);
        }
        v = 1.0 - v;
        int i = std::clamp(int(u * width), 0, width - 1);
        int j = std::clamp(int(v * height), 0, height - 1);
        return size_t(j) * width + i;
    }
};

// normal_from_map (material.h:282-308)
vec3 normal_from_map(const ray& r, const hit_record& rec, const texture_t& map) {
    const double min_cos_o = 0.01;
    vec3 n = rec.front_face ? rec.normal : -rec.normal;
    vec3 t = rec.dpdu - dot(rec.dpdu, n) * n;
    if (t.length_squared() <= 1e-12 * rec.dpdu.length_squared()) return rec.normal;
    t = unit_vector(t);
    vec3 b = cross(n, t);
    if (dot(b, rec.dpdv) < 0) b = -b;
    color c = map.value(rec.u, rec.v, rec.p);
    double nz = std::fmax(2 * c.z() - 1, 1e-3);
    vec3 m = unit_vector((2 * c.x() - 1) * t + (2 * c.y() - 1) * b + nz * n);
    if (!rec.front_face) m = -m;
    vec3 wo = unit_vector(-r.dir);
    double cos_o = dot(m, wo);
    if (cos_o < min_cos_o) m = unit_vector(m + (min_cos_o - cos_o) * wo);
    return m;
}

struct lambertian : material {  // material.h:74-101
    texture_t tex;
    bool scatter(const ray&, const hit_record& rec, scatter_record& s, sampler&) const override {
        s.attenuation = tex.value(rec.u, rec.v, rec.p);
        s.pdf_ptr = std::make_unique<cosine_pdf>(rec.normal);
        s.skip_pdf = false;
        return true;
    }
    color eval(const ray&, const hit_record& rec, const vec3& d) const override {
        double c = dot(rec.normal, unit_vector(d));
        return tex.value(rec.u, rec.v, rec.p) * (c < 0 ? 0 : c / pi);
    }
};

struct metal : material {  // material.h:109-173
    color albedo, f_avg;
    double roughness, alpha;
    metal_fresnel fresnel;
    bool scatter(const ray& r, const hit_record& rec, scatter_record& s, sampler&) const override {
        if (roughness == 0) {
            s.attenuation = fresnel.measured ? fresnel.at(dot(rec.normal, unit_vector(-r.dir))) : albedo;
            s.skip_pdf = true;
            s.skip_pdf_ray = ray(rec.p, reflect(unit_vector(r.dir), rec.normal));
            return true;
        }
        s.attenuation = albedo;
        auto spec = std::make_unique<ggx_reflection_pdf>(rec.normal, -r.dir, alpha);
        s.skip_pdf = false;
        double cos_o = std::fmax(0.0, dot(rec.normal, unit_vector(-r.dir)));
        double ms_w = ms_albedo(*table, cos_o, alpha) * average(ms_fresnel_avg(*table, f_avg, alpha));
        double ss_w = table_E(*table, cos_o, alpha) * average(fresnel.at(cos_o));
        if (ms_w <= 0) {
            s.pdf_ptr = std::move(spec);
            return true;
        }
        double p_spec = dclamp(ss_w / (ss_w + ms_w), 0.1, 1.0);
        s.pdf_ptr = std::make_unique<mixture_pdf>(std::move(spec), std::make_unique<cosine_pdf>(rec.normal), p_spec);
        return true;
    }
    color eval(const ray& r, const hit_record& rec, const vec3& d) const override {
        onb f(rec.normal);
        vec3 wo = unit_vector(f.to_local(-r.dir));
        vec3 wi = unit_vector(f.to_local(d));
        if (wo.z() <= 0 || wi.z() <= 0) return color(0, 0, 0);
        vec3 h = unit_vector(wo + wi);
        ggx dist(alpha);
        return fresnel.at(dot(wi, h)) * (dist.D(h) * dist.G(wo, wi) / (4 * wo.z())) +
               ms_fresnel_avg(*table, f_avg, alpha) * ms_eval(*table, wo.z(), wi.z(), alpha);
    }
};

struct thin_dielectric : material {  // material.h:683-707
    double ior;
    color tint;
    bool scatter(const ray& r, const hit_record& rec, scatter_record& s, sampler& smp) const override {
        vec3 ud = unit_vector(r.dir);
        double cos_theta = std::fmin(dot(-ud, rec.normal), 1.0);
        double f = fresnel_dielectric(cos_theta, ior);
        double reflectance = 2 * f / (1 + f);
        s.skip_pdf = true;
        if (smp.next() < reflectance) {
            s.attenuation = color(1, 1, 1);
            s.skip_pdf_ray = ray(rec.p, reflect(ud, rec.normal));
        } else {
            s.attenuation = tint;
            s.skip_pdf_ray = ray(rec.p, ud);
        }
        return true;
    }
};

struct dielectric : material {  // material.h:184-264
    double ior, roughness, alpha;
    color sigma_a;
    color interior_transmittance(const ray& r, const hit_record& rec) const override {
        if (rec.front_face || (sigma_a.x() == 0 && sigma_a.y() == 0 && sigma_a.z() == 0)) return color(1, 1, 1);
        double d = rec.t * r.dir.length();
        return color(std::exp(-sigma_a.x() * d), std::exp(-sigma_a.y() * d), std::exp(-sigma_a.z() * d));
    }
    bool lets_shadows_through() const override { return roughness == 0; }
    rough_glass bsdf(const ray& r, const hit_record& rec) const {
        return rough_glass(rec.normal, -r.dir, rec.front_face ? ior : 1.0 / ior, alpha);
    }
    bool scatter(const ray& r, const hit_record& rec, scatter_record& s, sampler& smp) const override {
        if (roughness > 0) {
            s.attenuation = color(1, 1, 1);
            s.pdf_ptr = std::make_unique<rough_glass_pdf>(bsdf(r, rec));
            s.skip_pdf = false;
            return true;
        }
        s.attenuation = color(1, 1, 1);
        s.skip_pdf = true;
        double ri = rec.front_face ? (1.0 / ior) : ior;
        vec3 ud = unit_vector(r.dir);
        double cos_theta = std::fmin(dot(-ud, rec.normal), 1.0);
        double sin_theta = std::sqrt(1.0 - cos_theta * cos_theta);
        bool cannot_refract = ri * sin_theta > 1.0;
        vec3 d;
        if (cannot_refract || fresnel_dielectric(cos_theta, 1.0 / ri) > smp.next()) d = reflect(ud, rec.normal);
        else d = refract(ud, rec.normal, ri);
        s.skip_pdf_ray = ray(rec.p, d);
        return true;
    }
    color eval(const ray& r, const hit_record& rec, const vec3& d) const override {
        if (roughness == 0) return color(0, 0, 0);
        double f = bsdf(r, rec).eval(d);
        return color(f, f, f);
    }
};

struct principled : material {  // material.h:346-552 (base layers and clear coat)
    texture_t base;
    double metallic, alpha, f0;
    // clear coat (material.h:326-345, 460-527): weight and roughness from the
    // red channels of their textures, an optional normal map of its own, a tint
    bool has_coat = false, has_coat_normal_map = false;
    texture_t coat_map, coat_roughness_map, coat_normal_map;
    color coat_tint = color(1, 1, 1);
    static constexpr double coat_f0 = 0.04;
    double coat_at(const hit_record& rec) const {
        return has_coat ? dclamp(coat_map.value(rec.u, rec.v, rec.p).x(), 0.0, 1.0) : 0.0;
    }
    double coat_alpha_at(const hit_record& rec) const {
        return roughness_to_alpha(dclamp(coat_roughness_map.value(rec.u, rec.v, rec.p).x(), 0.0, 1.0));
    }
    vec3 coat_normal(const ray& r, const hit_record& rec) const {
        return has_coat_normal_map ? normal_from_map(r, rec, coat_normal_map) : rec.normal;
    }
    color coat_transmittance(double w, double cos_o, double cos_i) const {
        if (coat_tint.x() == 1 && coat_tint.y() == 1 && coat_tint.z() == 1) return color(1, 1, 1);
        auto cos_t = [](double c) {
            c = dclamp(c, 0.0, 1.0);
            return std::sqrt(1 - (1 - c * c) / (1.5 * 1.5));
        };
        double e = (1 / cos_t(cos_o) + 1 / cos_t(cos_i)) / 2;
        color t(std::pow(coat_tint.x(), e), std::pow(coat_tint.y(), e), std::pow(coat_tint.z(), e));
        return (1 - w) * color(1, 1, 1) + w * t;
    }
    // metallic and roughness maps (red channel) and the base's normal map, when present
    bool has_metallic_map = false, has_roughness_map = false, has_normal_map = false;
    texture_t metallic_map, roughness_map, normal_map;
    bool measured = false;
    metal_fresnel conductor;  // set_conductor (material.h:381-385)
    color specular_fresnel(double c, const color& b, double met) const {  // material.h:537-543
        if (!measured) return fresnel_schlick(c, spec_f0(b, met));
        double fd = fresnel_schlick(c, f0);
        return (1 - met) * color(fd, fd, fd) + met * conductor.at(c);
    }
    color ms_fresnel_of(const color& b, double met, double a) const {  // material.h:545-551
        if (!measured) return ms_fresnel(*table, spec_f0(b, met), a);
        double fd = f0 + (1 - f0) / 21;
        return ms_fresnel_avg(*table, (1 - met) * color(fd, fd, fd) + met * conductor.f_avg, a);
    }
    double metallic_at(const hit_record& rec) const {  // material.h:508-515
        return has_metallic_map ? dclamp(metallic_map.value(rec.u, rec.v, rec.p).x(), 0.0, 1.0) : metallic;
    }
    double alpha_at(const hit_record& rec) const {
        return has_roughness_map ? roughness_to_alpha(dclamp(roughness_map.value(rec.u, rec.v, rec.p).x(), 0.0, 1.0))
                                 : alpha;
    }
    vec3 base_normal(const ray& r, const hit_record& rec) const {
        return has_normal_map ? normal_from_map(r, rec, normal_map) : rec.normal;
    }
    color spec_f0(const color& b, double met) const { return (1 - met) * color(f0, f0, f0) + met * b; }
    bool scatter(const ray& r, const hit_record& rec, scatter_record& s, sampler&) const override {
        vec3 wo = unit_vector(-r.dir);
        vec3 n = base_normal(r, rec);
        double met = metallic_at(rec), a = alpha_at(rec);
        color b = base.value(rec.u, rec.v, rec.p);
        double cos_o = std::fmax(0.0, dot(n, wo));
        double spec_w = average(specular_fresnel(cos_o, b, met));
        double diff_w = (1 - met) * average(b) * (1 - fresnel_schlick(cos_o, f0));
        diff_w += ms_albedo(*table, cos_o, a) * average(ms_fresnel_of(b, met, a));
        double p_spec = spec_w + diff_w > 0 ? spec_w / (spec_w + diff_w) : 1;
        p_spec = dclamp(p_spec, 0.1, 1.0);
        s.pdf_ptr = std::make_unique<mixture_pdf>(std::make_unique<ggx_reflection_pdf>(n, wo, a),
                                                  std::make_unique<cosine_pdf>(n), p_spec);
        s.attenuation = b;
        s.skip_pdf = false;
        double w = coat_at(rec);
        if (w > 0) {  // the coat's sharp lobe gets its share of the samples (material.h:409-419)
            vec3 n_coat = coat_normal(r, rec);
            double coat_weight = w * fresnel_schlick(dot(n_coat, wo), coat_f0);
            double below = (1 - coat_weight) * (spec_w + diff_w) * average(coat_transmittance(w, 1, 1));
            double p_coat = dclamp(coat_weight / (coat_weight + below), 0.1, 0.9);
            s.pdf_ptr = std::make_unique<mixture_pdf>(std::make_unique<ggx_reflection_pdf>(n_coat, wo, coat_alpha_at(rec)),
                                                      std::move(s.pdf_ptr), p_coat);
        }
        return true;
    }
    // the base layers with normal n (material.h:486-506)
    color base_eval(const hit_record& rec, const vec3& n, const vec3& wo_w, const vec3& wi_w) const {
        onb f(n);
        vec3 wo = unit_vector(f.to_local(wo_w)), wi = unit_vector(f.to_local(wi_w));
        if (wo.z() <= 0 || wi.z() <= 0) return color(0, 0, 0);
        double met = metallic_at(rec), a = alpha_at(rec);
        color b = base.value(rec.u, rec.v, rec.p);
        vec3 h = unit_vector(wo + wi);
        ggx dist(a);
        color specular = specular_fresnel(dot(wi, h), b, met) * (dist.D(h) * dist.G(wo, wi) / (4 * wo.z())) +
                         ms_fresnel_of(b, met, a) * ms_eval(*table, wo.z(), wi.z(), a);
        double transmitted = (1 - fresnel_schlick(wi.z(), f0)) * (1 - fresnel_schlick(wo.z(), f0));
        color diffuse = ((1 - met) * transmitted * wi.z() / pi) * b;
        return specular + diffuse;
    }
    color eval(const ray& r, const hit_record& rec, const vec3& d) const override {  // material.h:423-456
        if (dot(d, rec.normal) <= 0) return color(0, 0, 0);  // the unbent normal
        vec3 wo_w = unit_vector(-r.dir), wi_w = unit_vector(d);
        color below = base_eval(rec, base_normal(r, rec), wo_w, wi_w);
        double w = coat_at(rec);
        if (w <= 0) return below;
        vec3 n_coat = coat_normal(r, rec);
        double cos_o = dot(n_coat, wo_w), cos_i = dot(n_coat, wi_w);
        double through = (1 - w * fresnel_schlick(cos_o, coat_f0)) * (1 - w * fresnel_schlick(cos_i, coat_f0));
        color result = through * coat_transmittance(w, cos_o, cos_i) * below;
        if (cos_o > 0 && cos_i > 0) {
            onb frame(n_coat);
            vec3 lo = unit_vector(frame.to_local(wo_w)), li = unit_vector(frame.to_local(wi_w));
            vec3 h = unit_vector(lo + li);
            double ca = coat_alpha_at(rec);
            ggx dist(ca);
            double coat = fresnel_schlick(dot(li, h), coat_f0) * dist.D(h) * dist.G(lo, li) / (4 * lo.z());
            coat += (ms_fresnel(*table, color(coat_f0, coat_f0, coat_f0), ca) * ms_eval(*table, lo.z(), li.z(), ca)).x();
            result = result + color(w * coat, w * coat, w * coat);
        }
        return result;
    }
};

struct normal_mapped : material {  // material.h:631-674
    std::unique_ptr<material> inner;
    texture_t map;
    hit_record perturbed(const ray& r, const hit_record& rec) const {
        hit_record shaded = rec;
        shaded.normal = normal_from_map(r, rec, map);
        return shaded;
    }
    bool scatter(const ray& r, const hit_record& rec, scatter_record& s, sampler& smp) const override {
        hit_record shaded = perturbed(r, rec);
        if (!inner->scatter(r, shaded, s, smp)) return false;
        if (s.skip_pdf) {
            vec3 d = s.skip_pdf_ray.dir;
            if (dot(d, shaded.normal) > 0 && dot(d, rec.normal) < 0)
                s.skip_pdf_ray = ray(s.skip_pdf_ray.orig, d - 2 * dot(d, rec.normal) * rec.normal);
        }
        return true;
    }
    color eval(const ray& r, const hit_record& rec, const vec3& d) const override {
        if (dot(d, rec.normal) <= 0) return color(0, 0, 0);
        return inner->eval(r, perturbed(r, rec), d);
    }
    color emitted(const ray& r, const hit_record& rec) const override { return inner->emitted(r, rec); }
};

// emissive_surface (material.h:714-741): the inner material's scattering,
// and a glow of the texture's value on the front face; not a light.
struct glowing : material {
    std::unique_ptr<material> inner;
    texture_t glow;
    bool scatter(const ray& r, const hit_record& rec, scatter_record& s, sampler& smp) const override {
        return inner->scatter(r, rec, s, smp);
    }
    color eval(const ray& r, const hit_record& rec, const vec3& d) const override { return inner->eval(r, rec, d); }
    color emitted(const ray&, const hit_record& rec) const override {
        if (!rec.front_face) return color(0, 0, 0);
        return glow.value(rec.u, rec.v, rec.p);
    }
    color interior_transmittance(const ray& r, const hit_record& rec) const override {
        return inner->interior_transmittance(r, rec);
    }
    color average_emission() const override { return inner->average_emission(); }
    bool lets_shadows_through() const override { return inner->lets_shadows_through(); }
    bool is_volume() const override { return inner->is_volume(); }
};

struct diffuse_light : material {  // material.h:568-615
    color emit, avg;
    double spread, tan_half;
    color emitted(const ray& r, const hit_record& rec) const override {
        if (!rec.front_face) return color(0, 0, 0);
        if (spread >= 180) return emit;
        double c = -dot(unit_vector(r.dir), rec.normal);
        double gf;
        if (c <= 0) gf = 0;
        else gf = std::fmax(0.0, 1 - (std::sqrt(std::fmax(0.0, 1 - c * c)) / c) / tan_half);
        return emit * gf;
    }
    color average_emission() const override { return avg; }
};

struct phase_material : material {  // medium.h:13-39
    color albedo;
    double g;
    bool is_volume() const override { return true; }
    bool scatter(const ray& r, const hit_record&, scatter_record& s, sampler&) const override {
        s.attenuation = albedo;
        s.pdf_ptr = std::make_unique<hg_pdf>(r.dir, g);
        s.skip_pdf = false;
        return true;
    }
    color eval(const ray& r, const hit_record&, const vec3& d) const override {
        return albedo * hg_pdf::phase(dot(unit_vector(r.dir), unit_vector(d)), g);
    }
};

// ------------------------------------------------------------------ geometry
struct sphere : hittable {  // sphere.h
    point3 center;
    double radius;
    const material* mat;
    int64_t id;
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx&) const override {
        vec3 oc = center - r.orig;
        double a = r.dir.length_squared();
        double h = dot(r.dir, oc);
        double c = oc.length_squared() - radius * radius;
        double disc = h * h - a * c;
        if (disc < 0) return false;
        double sq = std::sqrt(disc);
        double root = (h - sq) / a;
        if (!t.surrounds(root)) {
            root = (h + sq) / a;
            if (!t.surrounds(root)) return false;
        }
        rec.t = root;
        rec.p = r.at(root);
        vec3 outward = (rec.p - center) / radius;
        rec.set_face_normal(r, outward);
        // get_sphere_uv and get_sphere_tangents (sphere.h:46-48, 99-120)
        vec3 un = outward * (radius < 0 ? -1.0 : 1.0);
        double theta = std::acos(std::clamp(-un.y(), -1.0, 1.0));
        double phi = std::atan2(-un.z(), un.x()) + pi;
        rec.u = phi / (2 * pi);
        rec.v = theta / pi;
        double ar = std::fabs(radius);
        double sin_theta = std::sqrt(un.x() * un.x() + un.z() * un.z());
        if (sin_theta < 1e-8) {
            rec.dpdu = rec.dpdv = vec3(0, 0, 0);
        } else {
            rec.dpdu = (2 * pi * ar) * vec3(un.z(), 0, -un.x());
            rec.dpdv = (pi * ar / sin_theta) * vec3(-un.x() * un.y(), sin_theta * sin_theta, -un.y() * un.z());
        }
        rec.mat = mat;
        rec.id = id;
        return true;
    }
    aabb bounding_box() const override {
        double r = std::fabs(radius);
        vec3 rv(r, r, r);
        return aabb(center - rv, center + rv);
    }
    double pdf_value(const point3& o, const vec3& d) const override {
        hit_record rec;
        if (!hit(ray(o, d), interval(0.001, infinity), rec, hit_ctx{})) return 0;
        double r = std::fabs(radius);
        double dist2 = (center - o).length_squared();
        if (dist2 <= r * r) {
            double hd2 = rec.t * rec.t * d.length_squared();
            double cosine = std::fabs(dot(unit_vector(d), (rec.p - center) / r));
            return hd2 / (cosine * 4 * pi * r * r);
        }
        double cmax = std::sqrt(1 - r * r / dist2);
        return 1 / (2 * pi * (1 - cmax));
    }
    vec3 random(const point3& o, sampler& s) const override {
        double r = std::fabs(radius);
        vec3 dir = center - o;
        double dist2 = dir.length_squared();
        if (dist2 <= r * r) {
            while (true) {
                double x = s.next(-1, 1);
                double y = s.next(-1, 1);
                double z = s.next(-1, 1);
                vec3 p(x, y, z);
                double l = p.length_squared();
                if (1e-160 < l && l <= 1) return center + r * (p / std::sqrt(l)) - o;
            }
        }
        onb uvw(dir);
        double r1 = s.next();
        double r2 = s.next();
        double z = 1 + r2 * (std::sqrt(1 - r * r / dist2) - 1);
        double phi = 2 * pi * r1;
        return uvw.transform(vec3(std::cos(phi) * std::sqrt(1 - z * z), std::sin(phi) * std::sqrt(1 - z * z), z));
    }
    double power() const override { return pi * 4 * pi * radius * radius * luminance(mat->average_emission()); }
};

struct quad : hittable {  // quad.h
    point3 Q;
    vec3 u, v, w, normal;
    double D, area;
    const material* mat;
    int64_t id;
    aabb box;
    quad(const point3& Q, const vec3& u, const vec3& v, const material* m, int64_t id) : Q(Q), u(u), v(v), mat(m), id(id) {
        vec3 n = cross(u, v);
        normal = unit_vector(n);
        D = dot(normal, Q);
        w = n / dot(n, n);
        area = n.length();
        box = aabb(aabb(Q, Q + u + v), aabb(Q + u, Q + v));
    }
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx&) const override {
        double denom = dot(normal, r.dir);
        if (std::fabs(denom) < 1e-8) return false;
        double tt = (D - dot(normal, r.orig)) / denom;
        if (!t.contains(tt)) return false;
        point3 ip = r.at(tt);
        vec3 ph = ip - Q;
        double alpha = dot(w, cross(ph, v));
        double beta = dot(w, cross(u, ph));
        interval unit(0, 1);
        if (!unit.contains(alpha) || !unit.contains(beta)) return false;
        rec.t = tt;
        rec.p = ip;
        rec.u = alpha;  // quad.h:53-57
        rec.v = beta;
        rec.dpdu = u;
        rec.dpdv = v;
        rec.mat = mat;
        rec.id = id;
        rec.set_face_normal(r, normal);
        return true;
    }
    aabb bounding_box() const override { return box; }
    double pdf_value(const point3& o, const vec3& d) const override {
        hit_record rec;
        if (!hit(ray(o, d), interval(0.001, infinity), rec, hit_ctx{})) return 0;
        double dist2 = rec.t * rec.t * d.length_squared();
        double cosine = std::fabs(dot(d, rec.normal) / d.length());
        return dist2 / (cosine * area);
    }
    vec3 random(const point3& o, sampler& s) const override {
        double a = s.next();
        double b = s.next();
        return (Q + (a * u) + (b * v)) - o;
    }
    double power() const override { return pi * area * luminance(mat->average_emission()); }
};

struct triangle : hittable {  // mesh.h:28-166 (vertex normals, uvs, alpha masks)
    point3 p0, p1, p2;
    const material* mat;
    int64_t id;
    double area;
    bool has_normals = false;
    vec3 n0, n1, n2;
    bool has_uvs = false;
    vec3 t0, t1, t2;  // (u, v, 0) at each vertex
    // the mesh's alpha mask (mesh.h:82-86) and its first primitive: the draw
    // of the mask test is keyed by mesh_first + id (the triangle's index in
    // its mesh), as on the GPU
    const texture_t* alpha = nullptr;
    int64_t mesh_first = 0;
    triangle(const point3& a, const point3& b, const point3& c, const material* m, int64_t id)
        : p0(a), p1(b), p2(c), mat(m), id(id) {
        area = 0.5 * cross(p1 - p0, p2 - p0).length();
    }
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const override {
        vec3 e1 = p1 - p0, e2 = p2 - p0;
        vec3 pvec = cross(r.dir, e2);
        double det = dot(e1, pvec);
        if (std::fabs(det) < 1e-12) return false;
        double inv_det = 1 / det;
        vec3 tvec = r.orig - p0;
        double b1 = dot(tvec, pvec) * inv_det;
        if (b1 < 0 || b1 > 1) return false;
        vec3 qvec = cross(tvec, e1);
        double b2 = dot(r.dir, qvec) * inv_det;
        if (b2 < 0 || b1 + b2 > 1) return false;
        double tt = dot(e2, qvec) * inv_det;
        if (!t.surrounds(tt)) return false;
        double b0 = 1 - b1 - b2;
        if (alpha) {  // checked before rec is touched (mesh.h:82-86)
            double u = b1, v = b2;
            if (has_uvs) {
                vec3 uv = b0 * t0 + b1 * t1 + b2 * t2;
                u = uv.x();
                v = uv.y();
            }
            double coverage = alpha->alpha_value(u, v);
            if (coverage < 1 && medium_u(c.key, c.serial, 0x80000000u | uint32_t(mesh_first + id)) >= coverage)
                return false;
        }
        rec.t = tt;
        rec.p = r.at(tt);
        rec.mat = mat;
        rec.id = id;
        rec.u = b1;
        rec.v = b2;
        if (has_uvs) {  // mesh.h:75-81
            vec3 uv = b0 * t0 + b1 * t1 + b2 * t2;
            rec.u = uv.x();
            rec.v = uv.y();
        }
        if (!has_uvs) {  // set_tangents, mesh.h:149-166
            rec.dpdu = e1;
            rec.dpdv = e2;
        } else {
            vec3 duv1 = t1 - t0, duv2 = t2 - t0;
            double det = duv1.x() * duv2.y() - duv2.x() * duv1.y();
            if (std::fabs(det) < 1e-12) {
                rec.dpdu = rec.dpdv = vec3(0, 0, 0);
            } else {
                double inv = 1 / det;
                rec.dpdu = (duv2.y() * e1 - duv1.y() * e2) * inv;
                rec.dpdv = (duv1.x() * e2 - duv2.x() * e1) * inv;
            }
        }
        rec.set_face_normal(r, unit_vector(cross(e1, e2)));
        if (has_normals) {  // mesh.h:93-101
            double b0 = 1 - b1 - b2;
            vec3 shading = b0 * n0 + b1 * n1 + b2 * n2;
            if (shading.length_squared() > 0) {
                shading = unit_vector(shading);
                rec.normal = dot(shading, rec.normal) < 0 ? -shading : shading;
            }
        }
        return true;
    }
    aabb bounding_box() const override { return aabb(aabb(p0, p1), aabb(p2, p2)); }
    double pdf_value(const point3& o, const vec3& d) const override {
        hit_record rec;
        if (!hit(ray(o, d), interval(0.001, infinity), rec, hit_ctx{})) return 0;
        vec3 g = unit_vector(cross(p1 - p0, p2 - p0));
        double dist2 = rec.t * rec.t * d.length_squared();
        double cosine = std::fabs(dot(d, g) / d.length());
        return dist2 / (cosine * area);
    }
    vec3 random(const point3& o, sampler& s) const override {
        double r1 = s.next(), r2 = s.next();
        if (r1 + r2 > 1) {
            r1 = 1 - r1;
            r2 = 1 - r2;
        }
        return (p0 + r1 * (p1 - p0) + r2 * (p2 - p0)) - o;
    }
    double power() const override { return mat ? pi * area * luminance(mat->average_emission()) : 0; }
};

struct hlist : hittable {  // hittable_list.h
    std::vector<const hittable*> objects;
    aabb box;
    void add(const hittable* h) {
        objects.push_back(h);
        box = aabb(box, h->bounding_box());
    }
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const override {
        hit_record tmp;
        bool any = false;
        double closest = t.max;
        for (auto* o : objects)
            if (o->hit(r, interval(t.min, closest), tmp, c)) {
                any = true;
                closest = tmp.t;
                rec = tmp;
            }
        return any;
    }
    aabb bounding_box() const override { return box; }
    double pdf_value(const point3& o, const vec3& d) const override {
        if (objects.empty()) return 0.0;
        double weight = 1.0 / objects.size(), sum = 0.0;
        for (auto* h : objects) sum += weight * h->pdf_value(o, d);
        return sum;
    }
    vec3 random(const point3& o, sampler& s) const override {
        int n = int(objects.size());
        return objects[std::min(int(s.next() * n), n - 1)]->random(o, s);
    }
    double power() const override {
        double sum = 0;
        for (auto* h : objects) sum += h->power();
        return sum;
    }
};

struct bvh : hittable {  // bvh.h
    struct node {
        aabb box;
        int offset = 0, count = 0, axis = 0;
    };
    struct item {
        aabb box;
        point3 centroid;
        const hittable* obj;
    };
    std::vector<const hittable*> prims;
    std::vector<node> nodes;
    explicit bvh(const std::vector<const hittable*>& objs) {
        if (objs.empty()) {
            nodes.push_back(node());
            return;
        }
        std::vector<item> items;
        for (auto* o : objs) {
            aabb b = o->bounding_box();
            items.push_back({b, b.centroid(), o});
        }
        build(items, 0, items.size(), 0);
        for (auto& it : items) prims.push_back(it.obj);
    }
    int make_leaf(int index, size_t begin, size_t count, std::vector<item>& items, int axis, int depth) {
        if (count > 4) {
            size_t mid = begin + count / 2;
            nodes[index].axis = axis;
            build(items, begin, mid, depth + 1);
            int right = build(items, mid, begin + count, depth + 1);
            nodes[index].offset = right;
            return index;
        }
        nodes[index].offset = int(begin);
        nodes[index].count = int(count);
        return index;
    }
    int build(std::vector<item>& items, size_t begin, size_t end, int depth) {
        int index = int(nodes.size());
        nodes.push_back(node());
        aabb bounds, cb;
        for (size_t k = begin; k < end; k++) {
            bounds = aabb(bounds, items[k].box);
            cb = aabb(cb, aabb(items[k].centroid, items[k].centroid));
        }
        nodes[index].box = bounds;
        size_t count = end - begin;
        int axis = cb.longest_axis();
        interval ext = cb.ax[axis];
        if (count == 1 || ext.size() <= 0 || depth >= 64) return make_leaf(index, begin, count, items, axis, depth);
        struct bin {
            aabb box;
            int count = 0;
        } bins[16];
        auto bin_of = [&](const point3& c) { return std::min(int(16 * (c[axis] - ext.min) / ext.size()), 15); };
        for (size_t k = begin; k < end; k++) {
            bin& b = bins[bin_of(items[k].centroid)];
            b.box = aabb(b.box, items[k].box);
            b.count++;
        }
        double best = infinity;
        int split = -1;
        for (int s = 0; s < 15; s++) {
            aabb l, r;
            int lc = 0, rc = 0;
            for (int b = 0; b <= s; b++) {
                l = aabb(l, bins[b].box);
                lc += bins[b].count;
            }
            for (int b = s + 1; b < 16; b++) {
                r = aabb(r, bins[b].box);
                rc += bins[b].count;
            }
            if (lc == 0 || rc == 0) continue;
            double cost = l.area() * lc + r.area() * rc;
            if (cost < best) {
                best = cost;
                split = s;
            }
        }
        double leaf_cost = double(count), split_cost = 0.125 + best / bounds.area();
        if (split < 0 || (count <= 4 && leaf_cost <= split_cost)) return make_leaf(index, begin, count, items, axis, depth);
        auto mid_it = std::partition(items.begin() + begin, items.begin() + end,
                                     [&](const item& it) { return bin_of(it.centroid) <= split; });
        size_t mid = size_t(mid_it - items.begin());
        nodes[index].axis = axis;
        build(items, begin, mid, depth + 1);
        int right = build(items, mid, end, depth + 1);
        nodes[index].offset = right;
        return index;
    }
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const override {
        if (prims.empty()) return false;
        vec3 inv(1 / r.dir.x(), 1 / r.dir.y(), 1 / r.dir.z());
        bool neg[3] = {inv.x() < 0, inv.y() < 0, inv.z() < 0};
        bool any = false;
        int stack[256];
        int sp = 0, cur = 0;
        while (true) {
            const node& n = nodes[cur];
            if (n.box.hit(r.orig, inv, t)) {
                if (n.count > 0) {
                    for (int k = 0; k < n.count; k++)
                        if (prims[n.offset + k]->hit(r, t, rec, c)) {
                            any = true;
                            t.max = rec.t;
                        }
                    if (sp == 0) break;
                    cur = stack[--sp];
                } else if (neg[n.axis]) {
                    stack[sp++] = cur + 1;
                    cur = n.offset;
                } else {
                    stack[sp++] = n.offset;
                    cur = cur + 1;
                }
            } else {
                if (sp == 0) break;
                cur = stack[--sp];
            }
        }
        return any;
    }
    aabb bounding_box() const override { return nodes[0].box; }
};

struct affine {  // transform.h
    double m[9], t[3], im[9], it[3];
    vec3 mul(const double* a, const vec3& v) const {
        return vec3(a[0] * v[0] + a[1] * v[1] + a[2] * v[2], a[3] * v[0] + a[4] * v[1] + a[5] * v[2],
                    a[6] * v[0] + a[7] * v[1] + a[8] * v[2]);
    }
    vec3 vector(const vec3& v) const { return mul(m, v); }
    point3 point(const point3& p) const { return vector(p) + vec3(t[0], t[1], t[2]); }
    vec3 inv_vector(const vec3& v) const { return mul(im, v); }
    point3 inv_point(const point3& p) const { return inv_vector(p) + vec3(it[0], it[1], it[2]); }
    vec3 normal(const vec3& n) const {
        return vec3(im[0] * n[0] + im[3] * n[1] + im[6] * n[2], im[1] * n[0] + im[4] * n[1] + im[7] * n[2],
                    im[2] * n[0] + im[5] * n[1] + im[8] * n[2]);
    }
    double area_scale() const {
        double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
        return std::pow(std::fabs(det), 2.0 / 3.0);
    }
};

struct instance : hittable {  // instance.h
    const hittable* object;
    affine x;
    aabb box;
    instance(const hittable* o, const affine& a) : object(o), x(a) {
        aabb b = o->bounding_box();
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    point3 c(i ? b.ax[0].max : b.ax[0].min, j ? b.ax[1].max : b.ax[1].min, k ? b.ax[2].max : b.ax[2].min);
                    point3 p = x.point(c);
                    box = aabb(box, aabb(p, p));
                }
    }
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const override {
        ray local(x.inv_point(r.orig), x.inv_vector(r.dir));
        if (!object->hit(local, t, rec, c)) return false;
        rec.p = x.point(rec.p);
        rec.normal = unit_vector(x.normal(rec.normal));
        rec.dpdu = x.vector(rec.dpdu);  // instance.h:38-39
        rec.dpdv = x.vector(rec.dpdv);
        return true;
    }
    aabb bounding_box() const override { return box; }
    double pdf_value(const point3& o, const vec3& d) const override { return object->pdf_value(x.inv_point(o), x.inv_vector(d)); }
    vec3 random(const point3& o, sampler& s) const override { return x.vector(object->random(x.inv_point(o), s)); }
    double power() const override { return object->power() * x.area_scale(); }
};

// A triangle mesh (mesh.h:175-233): hits through its own BVH; as a light it
// picks triangles by area in the mesh's own order.
struct mesh_obj : hittable {
    const bvh* tree;
    std::vector<const hittable*> tris;
    std::vector<double> area_cdf;
    double total_area = 0;
    mesh_obj(const bvh* t, std::vector<const hittable*> v) : tree(t), tris(std::move(v)) {
        for (auto* h : tris) {
            total_area += static_cast<const triangle*>(h)->area;
            area_cdf.push_back(total_area);
        }
    }
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const override { return tree->hit(r, t, rec, c); }
    aabb bounding_box() const override { return tree->bounding_box(); }
    double pdf_value(const point3& o, const vec3& d) const override {
        if (total_area <= 0) return 0;
        double sum = 0;
        for (auto* h : tris) sum += (static_cast<const triangle*>(h)->area / total_area) * h->pdf_value(o, d);
        return sum;
    }
    vec3 random(const point3& o, sampler& s) const override {
        double x = s.next() * total_area;
        size_t i = size_t(std::upper_bound(area_cdf.begin(), area_cdf.end(), x) - area_cdf.begin());
        return tris[std::min(i, tris.size() - 1)]->random(o, s);
    }
    double power() const override {
        double sum = 0;
        for (auto* h : tris) sum += h->power();
        return sum;
    }
};

// A block of a voxel model (voxel_grid, voxel.h:168-337): the cells [lo, hi)
// walked along the ray with a 3D DDA (Amanatides and Woo 1987); a surface is
// a plane crossing between cells of different kinds (voxel.h:103-111). The
// model's whole grid is looked up, so a block boundary is no surface.
struct vox_block : hittable {
    const mpt_vox_model* m = nullptr;
    const uint8_t* cells = nullptr;
    const material* mats[256] = {};
    const material* edge = nullptr;
    double edge_width = 0;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    int64_t id = 0;

    int at(const int c[3]) const {  // palette index, 0 outside the model
        if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= m->nx || c[1] >= m->ny || c[2] >= m->nz) return 0;
        return cells[(size_t(c[2]) * m->ny + c[1]) * m->nx + c[0]];
    }
    static bool face(int k, int n) { return k != 0 && (n == 0 || (k == 1 && n == 2)); }

    bool hit(const ray& r, interval ray_t, hit_record& rec, const hit_ctx&) const override {
        const point3& o = r.orig;
        const vec3& d = r.dir;
        double t_enter = -infinity, t_exit = infinity;
        int enter_axis = 0;
        for (int a = 0; a < 3; a++) {
            if (d[a] == 0) {
                if (o[a] < lo[a] || o[a] > hi[a]) return false;
                continue;
            }
            double t0 = (lo[a] - o[a]) / d[a], t1 = (hi[a] - o[a]) / d[a];
            if (t0 > t1) std::swap(t0, t1);
            if (t0 > t_enter) {
                t_enter = t0;
                enter_axis = a;
            }
            t_exit = std::fmin(t_exit, t1);
        }
        const bool outside = t_enter >= ray_t.min;
        const double t_start = outside ? t_enter : ray_t.min;
        if (t_start > std::fmin(t_exit, ray_t.max)) return false;
        const point3 start = r.at(t_start);
        int c[3], step[3];
        double t_next[3], t_delta[3];
        for (int a = 0; a < 3; a++) {
            c[a] = std::clamp(int(std::floor(start[a])), lo[a], hi[a] - 1);
            if (outside && a == enter_axis) c[a] = d[a] > 0 ? lo[a] : hi[a] - 1;
            if (d[a] > 0) {
                step[a] = 1, t_next[a] = (c[a] + 1 - o[a]) / d[a], t_delta[a] = 1 / d[a];
            } else if (d[a] < 0) {
                step[a] = -1, t_next[a] = (c[a] - o[a]) / d[a], t_delta[a] = -1 / d[a];
            } else {
                step[a] = 0, t_next[a] = t_delta[a] = infinity;
            }
        }
        int index = at(c);
        if (outside) {
            int behind[3] = {c[0], c[1], c[2]};
            behind[enter_axis] -= step[enter_axis];
            if (crossing(r, t_enter, enter_axis, step[enter_axis], at(behind), index, c, rec)) return true;
        }
        for (;;) {
            const int a = t_next[0] < t_next[1] ? (t_next[0] < t_next[2] ? 0 : 2) : (t_next[1] < t_next[2] ? 1 : 2);
            const double t = t_next[a];
            if (t > ray_t.max || step[a] == 0) return false;  // no plane ahead: the reference would loop
            c[a] += step[a];
            t_next[a] += t_delta[a];
            const int next = at(c);
            if (crossing(r, t, a, step[a], index, next, c, rec)) return true;
            if (c[a] < lo[a] || c[a] >= hi[a]) return false;
            index = next;
        }
    }
    // into cell to_cell (index to) from index from across the plane of axis at t
    bool crossing(const ray& r, double t, int axis, int step, int from, int to, const int to_cell[3],
                  hit_record& rec) const {
        const int kf = m->kind[from], kt = m->kind[to];
        if (kt != 0 && face(kt, kf)) return record(r, t, axis, -step, to, to_cell, rec);
        if (face(kf, kt)) {
            int from_cell[3] = {to_cell[0], to_cell[1], to_cell[2]};
            from_cell[axis] -= step;
            return record(r, t, axis, step, from, from_cell, rec);
        }
        return false;
    }
    bool on_edge(const point3& p, int axis, int sign, const int cell[3]) const {
        const int k = m->kind[at(cell)];
        for (int side = 1; side <= 2; side++) {
            const int tg = (axis + side) % 3;
            const double f = p[tg] - cell[tg];
            for (int dir : {-1, 1}) {
                if ((dir > 0 ? 1 - f : f) >= edge_width) continue;
                int beside[3] = {cell[0], cell[1], cell[2]};
                beside[tg] += dir;
                int above[3] = {beside[0], beside[1], beside[2]};
                above[axis] += sign;
                if (face(k, m->kind[at(beside)]) || m->kind[at(above)] != 0) return true;
            }
        }
        return false;
    }
    bool record(const ray& r, double t, int axis, int sign, int index, const int cell[3], hit_record& rec) const {
        rec.t = t;
        rec.p = r.at(t);
        vec3 outward(0, 0, 0), tb(0, 0, 0), tc(0, 0, 0);
        outward[axis] = sign;
        tb[(axis + 1) % 3] = 1;
        tc[(axis + 2) % 3] = 1;
        rec.dpdu = sign > 0 ? tb : tc;
        rec.dpdv = sign > 0 ? tc : tb;
        rec.u = dot(rec.p, rec.dpdu) - std::floor(dot(rec.p, rec.dpdu));
        rec.v = dot(rec.p, rec.dpdv) - std::floor(dot(rec.p, rec.dpdv));
        rec.mat = edge_width > 0 && on_edge(rec.p, axis, sign, cell) ? edge : mats[index];
        rec.set_face_normal(r, outward);
        rec.id = id;
        return true;
    }
    aabb bounding_box() const override {
        return aabb(point3(lo[0], lo[1], lo[2]), point3(hi[0], hi[1], hi[2]));
    }
};

// A world-list object that owns its ID range (lists of an instanced box are
// re-identified per instance by offsetting the child IDs).
struct idshift : hittable {
    const hittable* inner;
    int64_t shift;
    idshift(const hittable* h, int64_t s) : inner(h), shift(s) {}
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const override {
        if (!inner->hit(r, t, rec, c)) return false;
        rec.id += shift;
        return true;
    }
    aabb bounding_box() const override { return inner->bounding_box(); }
    double pdf_value(const point3& o, const vec3& d) const override { return inner->pdf_value(o, d); }
    vec3 random(const point3& o, sampler& s) const override { return inner->random(o, s); }
    double power() const override { return inner->power(); }
};

struct medium : hittable {  // constant_medium, medium.h:57-96
    const hittable* boundary;
    double density;
    const material* phase;
    uint32_t index;  // world-list position, keys the free-flight draw
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const override {
        if (!c.media) return false;
        hit_record r1, r2;
        if (!boundary->hit(r, interval(-infinity, infinity), r1, c)) return false;
        if (!boundary->hit(r, interval(r1.t + 0.0001, infinity), r2, c)) return false;
        double t0 = std::fmax(r1.t, t.min), t1 = std::fmin(r2.t, t.max);
        if (t0 >= t1) return false;
        double len = r.dir.length();
        double inside = (t1 - t0) * len;
        double hd = -std::log(1 - medium_u(c.key, c.serial, index)) / density;
        if (hd > inside) return false;
        rec.t = t0 + hd / len;
        rec.p = r.at(rec.t);
        rec.normal = vec3(1, 0, 0);
        rec.front_face = true;
        rec.mat = phase;
        rec.id = -2;
        return true;
    }
    aabb bounding_box() const override { return boundary->bounding_box(); }
};

// A distant light of small angular size (sun.h:23-80): never hit; radiance
// E / omega inside the cone, sampled uniformly in it.
struct sun_light : hittable {
    onb frame;
    double cos_max, solid_angle;
    color irradiance, radiance_in_cone;
    sun_light(const vec3& direction, double angular_diameter, const color& e) : frame(direction) {
        double half = std::fmax(angular_diameter, 1e-3) * pi / 180.0 / 2;  // sun.h:28-33
        cos_max = std::cos(half);
        solid_angle = 2 * pi * (1 - cos_max);
        irradiance = e;
        radiance_in_cone = e / solid_angle;
    }
    bool hit(const ray&, interval, hit_record&, const hit_ctx&) const override { return false; }
    aabb bounding_box() const override { return aabb(); }
    bool in_cone(const vec3& d) const { return dot(unit_vector(d), frame.w()) >= cos_max; }  // sun.h:76-78
    color radiance(const vec3& d) const { return in_cone(d) ? radiance_in_cone : color(0, 0, 0); }
    double pdf_value(const point3&, const vec3& d) const override { return in_cone(d) ? 1 / solid_angle : 0; }
    vec3 random(const point3&, sampler& s) const override {  // sun.h:56-62
        double cos_theta = 1 - s.next() * (1 - cos_max);
        double sin_theta = std::sqrt(std::fmax(0.0, 1 - cos_theta * cos_theta));
        double phi = 2 * pi * s.next();
        return frame.transform(vec3(std::cos(phi) * sin_theta, std::sin(phi) * sin_theta, cos_theta));
    }
    double power(double radius) const { return pi * radius * radius * luminance(irradiance); }  // sun.h:66-71
};

struct environment : hittable {  // environment.h
    int width = 0, height = 0;
    std::vector<color> pixels;
    double intensity = 1, rotation = 0;
    std::vector<double> cond, marg, row_weight;
    double total = 0;
    bool hit(const ray&, interval, hit_record&, const hit_ctx&) const override { return false; }
    aabb bounding_box() const override { return aabb(); }
    double weight(int i, int j) const { return luminance(pixels[size_t(j) * width + i]) * std::sin(pi * (j + 0.5) / height); }
    void build() {
        cond.assign(size_t(height) * (width + 1), 0.0);
        marg.assign(height + 1, 0.0);
        row_weight.assign(height, 0.0);
        for (int j = 0; j < height; j++) {
            double* c = &cond[size_t(j) * (width + 1)];
            for (int i = 0; i < width; i++) c[i + 1] = c[i] + std::fmax(weight(i, j), 0.0);
            row_weight[j] = c[width];
            for (int i = 1; i <= width; i++) c[i] = row_weight[j] > 0 ? c[i] / row_weight[j] : double(i) / width;
            marg[j + 1] = marg[j] + row_weight[j];
        }
        total = marg[height];
        if (total <= 0) {
            for (int j = 0; j < height; j++) {
                row_weight[j] = std::sin(pi * (j + 0.5) / height);
                marg[j + 1] = marg[j] + row_weight[j];
            }
            total = marg[height];
        }
        for (int j = 1; j <= height; j++) marg[j] /= total;
    }
    void angles(const vec3& dir, double& u, double& theta) const {
        vec3 d = unit_vector(dir);
        theta = std::acos(dclamp(d.y(), -1.0, 1.0));
        u = 0.5 + std::atan2(d.x(), -d.z()) / (2 * pi) + rotation;
        u -= std::floor(u);
    }
    color radiance(const vec3& dir) const {
        double u, theta;
        angles(dir, u, theta);
        double x = u * width - 0.5;
        double y = dclamp(theta / pi * height - 0.5, 0.0, height - 1.0);
        double x0 = std::floor(x), y0 = std::floor(y), fx = x - x0, fy = y - y0;
        int i0 = (int(x0) + width) % width, i1 = (i0 + 1) % width;
        int j0 = int(y0), j1 = std::min(j0 + 1, height - 1);
        auto at = [&](int i, int j) { return pixels[size_t(j) * width + i]; };
        color top = (1 - fx) * at(i0, j0) + fx * at(i1, j0);
        color bottom = (1 - fx) * at(i0, j1) + fx * at(i1, j1);
        return intensity * ((1 - fy) * top + fy * bottom);
    }
    double pdf_value(const point3&, const vec3& dir) const override {
        double u, theta;
        angles(dir, u, theta);
        double st = std::sin(theta);
        int i = std::clamp(int(u * width), 0, width - 1), j = std::clamp(int(theta / pi * height), 0, height - 1);
        if (st <= 0) return 0;
        const double* c = &cond[size_t(j) * (width + 1)];
        double pix = row_weight[j] / total * (c[i + 1] - c[i]) * width * height;
        return pix / (2 * pi * pi * st);
    }
    static double sample_cdf(const double* cdf, int n, double xi) {
        int k = int(std::upper_bound(cdf, cdf + n + 1, xi) - cdf) - 1;
        k = std::clamp(k, 0, n - 1);
        double span = cdf[k + 1] - cdf[k];
        double f = span > 0 ? (xi - cdf[k]) / span : 0.5;
        return k + dclamp(f, 0.0, 1.0 - 1e-9);
    }
    vec3 random(const point3&, sampler& s) const override {
        double y = sample_cdf(marg.data(), height, s.next());
        int j = std::min(int(y), height - 1);
        double x = sample_cdf(&cond[size_t(j) * (width + 1)], width, s.next());
        double theta = pi * y / height, phi = 2 * pi * (x / width - 0.5 - rotation);
        double st = std::sin(theta);
        return vec3(st * std::sin(phi), std::cos(theta), -st * std::cos(phi));
    }
    double power(double radius) const {
        double sum = 0;
        for (int j = 0; j < height; j++)
            for (int i = 0; i < width; i++) sum += std::fmax(weight(i, j), 0.0);
        double integral = intensity * sum * (2 * pi / width) * (pi / height);
        return pi * radius * radius * integral;
    }
};

// The infinite ground y = height (plane.h:10-42).
struct ground_plane : hittable {
    double height;
    const material* mat;
    int64_t id;
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx&) const override {
        double dy = r.dir.y();
        if (std::fabs(dy) < 1e-12) return false;
        double tt = (height - r.orig.y()) / dy;
        if (!t.surrounds(tt)) return false;
        rec.t = tt;
        rec.p = r.at(tt);
        rec.p[1] = height;  // exactly on the plane
        rec.u = rec.p.x();
        rec.v = -rec.p.z();
        rec.dpdu = vec3(1, 0, 0);
        rec.dpdv = vec3(0, 0, -1);
        rec.mat = mat;
        rec.id = id;
        rec.set_face_normal(r, vec3(0, 1, 0));
        return true;
    }
    aabb bounding_box() const override { return aabb(); }  // never in a BVH
};

// height_fog (medium.h:98-155): density * exp(-(y - base) / scale) over all
// of space; the free flight inverts the closed-form optical depth along the
// ray. Drawn from the medium stream with object index `index` (the world
// list's size plus one plus the fog's position), as the GPU draws it.
struct height_fog : hittable {
    double density, scale, base;
    const material* phase;
    uint32_t index;
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const override {
        if (!c.media) return false;
        double length = r.dir.length();
        double k = r.dir.y() / scale;
        double t0 = t.min;
        double start = density * length * std::exp(-(r.orig.y() + t0 * r.dir.y() - base) / scale);
        if (!(start > 0)) return false;
        double tau = -std::log(1 - medium_u(c.key, c.serial, index));
        double tt;
        if (std::fabs(k) < 1e-12) {
            tt = t0 + tau / start;
        } else {
            double x = 1 - tau * k / start;
            if (x <= 0) return false;
            tt = t0 - std::log(x) / k;
        }
        if (!(tt < t.max) || !std::isfinite(tt)) return false;
        rec.t = tt;
        rec.p = r.at(tt);
        rec.normal = vec3(1, 0, 0);
        rec.front_face = true;
        rec.mat = phase;
        rec.id = -2;
        return true;
    }
    aabb bounding_box() const override { return aabb(); }  // never in a BVH
};

// ------------------------------------------------------------------ scene
struct scene {
    std::vector<std::unique_ptr<material>> mats;
    std::vector<std::unique_ptr<hittable>> owned;
    std::vector<const hittable*> world;      // world-list objects (with their IDs)
    std::vector<const hittable*> items;      // those the BVH holds (not the lights that are only sampled)
    std::unique_ptr<bvh> tree;
    std::vector<const hittable*> lights;     // nullptr entries are the environment
    std::vector<double> prob, cdf;
    int strategy = 1;
    std::unique_ptr<environment> env;
    std::unique_ptr<sun_light> sun;    // also in lights, when the scene has one
    ggx_table table;
    aabb world_box;
    std::vector<const hittable*> fogs; // beside the BVH, before the ground (scene::unbounded)
    std::vector<std::unique_ptr<texture_t>> masks;  // meshes' alpha masks
    const hittable* ground = nullptr;  // beside the BVH (scene::unbounded)
    std::vector<uint8_t> texels;    // image textures, copied from the description
    std::vector<float> texels_hdr;
    std::vector<double> srgb;       // image_texture::srgb_table (texture.h:197-208)

    // the BVH, then the objects beside it, as one hittable_list (main.cpp:70-74)
    bool hit(const ray& r, interval t, hit_record& rec, const hit_ctx& c) const {
        bool any = tree->hit(r, t, rec, c);
        for (const hittable* f : fogs) {
            hit_record g;
            if (f->hit(r, interval(t.min, any ? rec.t : t.max), g, c)) {
                rec = g;
                any = true;
            }
        }
        if (ground) {
            hit_record g;
            if (ground->hit(r, interval(t.min, any ? rec.t : t.max), g, c)) {
                rec = g;
                any = true;
            }
        }
        return any;
    }

    double lights_pdf(const point3& o, const vec3& d) const {
        if (lights.empty()) return 0.0;
        double sum = 0;
        for (size_t k = 0; k < lights.size(); k++)
            sum += prob[k] * (lights[k] ? lights[k]->pdf_value(o, d) : env->pdf_value(o, d));
        return sum;
    }
    vec3 lights_random(const point3& o, sampler& s) const {
        double xi = s.next();
        size_t n = lights.size(), k;
        if (strategy == 2) k = std::min(size_t(xi * n), n - 1);
        else {
            k = size_t(std::upper_bound(cdf.begin(), cdf.end(), xi) - cdf.begin());
            k = std::clamp(k, size_t(1), n) - 1;
        }
        return lights[k] ? lights[k]->random(o, s) : env->random(o, s);
    }
};

template <typename T>
T* own(scene& s, std::unique_ptr<T> p) {
    T* r = p.get();
    s.owned.push_back(std::move(p));
    return r;
}

int build_scene(const mpt_scene_desc& d, scene& s, std::string& err) {
    build_table(s.table);
    if (d.n_texel_bytes > 0) s.texels.assign(d.texels, d.texels + d.n_texel_bytes);
    if (d.n_texels_hdr > 0) s.texels_hdr.assign(d.texels_hdr, d.texels_hdr + d.n_texels_hdr);
    s.srgb.resize(256);
    for (int k = 0; k < 256; k++) {
        double encoded = k / 255.0;
        s.srgb[k] = encoded <= 0.04045 ? encoded / 12.92 : std::pow((encoded + 0.055) / 1.055, 2.4);
    }
    std::function<texture_t(int)> tex = [&](int ti) {
        texture_t t{};
        const mpt_texture& x = d.textures[ti];
        t.type = x.type;
        t.inv_scale = x.inv_scale;
        t.even = color(x.even[0], x.even[1], x.even[2]);
        t.odd = color(x.odd[0], x.odd[1], x.odd[2]);
        if (x.type >= MPT_TEX_SCALED && x.type <= MPT_TEX_NORMAL_SCALE) {
            t.inner = std::make_shared<texture_t>(tex(x.inner));
            t.channel = x.channel;
            t.scale = x.scale;
            t.offset_u = x.uv[0], t.offset_v = x.uv[1], t.scale_u = x.uv[2], t.scale_v = x.uv[3];
            t.cos_r = x.uv[4], t.sin_r = x.uv[5];
        }
        if (x.type == MPT_TEX_IMAGE) {
            t.width = x.width;
            t.height = x.height;
            t.flags = x.flags;
            if (x.flags & MPT_TEXF_HDR) t.hdr = s.texels_hdr.data() + 3 * x.offset;
            else t.ldr = s.texels.data() + 4 * x.offset;
            t.srgb = s.srgb.data();
        }
        return t;
    };
    for (int k = 0; k < d.n_materials; k++) {
        const mpt_material& m = d.materials[k];
        std::unique_ptr<material> mp;
        switch (m.type) {
            case MPT_MAT_LAMBERTIAN: {
                auto p = std::make_unique<lambertian>();
                p->tex = tex(m.texture);
                mp = std::move(p);
                break;
            }
            case MPT_MAT_METAL: {
                auto p = std::make_unique<metal>();
                if (m.measured) {  // the measured metal's own head-on color and average (material.h:117-121)
                    p->fresnel = metal_fresnel::from_ior(color(m.eta[0], m.eta[1], m.eta[2]), color(m.k[0], m.k[1], m.k[2]));
                    p->albedo = p->fresnel.f0;
                } else {
                    p->albedo = color(m.albedo[0], m.albedo[1], m.albedo[2]);
                    p->fresnel = metal_fresnel::schlick(p->albedo);
                }
                p->f_avg = p->fresnel.f_avg;
                p->roughness = m.roughness;
                p->alpha = roughness_to_alpha(m.roughness);
                mp = std::move(p);
                break;
            }
            case MPT_MAT_DIELECTRIC: {
                if (m.thin) {
                    auto t = std::make_unique<thin_dielectric>();
                    t->ior = m.ior;
                    t->tint = color(m.albedo[0], m.albedo[1], m.albedo[2]);
                    mp = std::move(t);
                    break;
                }
                auto p = std::make_unique<dielectric>();
                p->ior = m.ior;
                p->roughness = m.roughness;
                p->alpha = roughness_to_alpha(m.roughness);
                p->sigma_a = color(m.sigma_a[0], m.sigma_a[1], m.sigma_a[2]);
                mp = std::move(p);
                break;
            }
            case MPT_MAT_PRINCIPLED: {
                auto p = std::make_unique<principled>();
                p->base = tex(m.texture);
                p->metallic = dclamp(m.metallic, 0.0, 1.0);
                p->alpha = roughness_to_alpha(dclamp(m.roughness, 0.0, 1.0));
                if ((p->has_metallic_map = m.metallic_texture >= 0)) p->metallic_map = tex(m.metallic_texture);
                if ((p->has_roughness_map = m.roughness_texture >= 0)) p->roughness_map = tex(m.roughness_texture);
                if ((p->has_normal_map = m.normal_map >= 0)) p->normal_map = tex(m.normal_map);
                if ((p->measured = m.measured != 0))
                    p->conductor = metal_fresnel::from_ior(color(m.eta[0], m.eta[1], m.eta[2]), color(m.k[0], m.k[1], m.k[2]));
                double r0 = (m.ior - 1) / (m.ior + 1);
                p->f0 = r0 * r0;
                if ((p->has_coat = m.coat != 0)) {
                    p->coat_map = tex(m.coat_weight_texture);
                    p->coat_roughness_map = tex(m.coat_roughness_texture);
                    if ((p->has_coat_normal_map = m.coat_normal_map >= 0)) p->coat_normal_map = tex(m.coat_normal_map);
                    p->coat_tint = color(m.coat_tint[0], m.coat_tint[1], m.coat_tint[2]);
                }
                mp = std::move(p);
                break;
            }
            case MPT_MAT_DIFFUSE_LIGHT: {
                auto p = std::make_unique<diffuse_light>();
                p->emit = color(m.emit[0], m.emit[1], m.emit[2]);
                p->spread = m.spread;
                double half = m.spread * pi / 180.0 / 2;
                p->tan_half = m.spread >= 180 ? infinity : std::tan(half);
                double share = 1;
                if (!(half >= pi / 2 - 1e-9)) {
                    double sn = std::sin(half), cs = std::cos(half);
                    share = sn * sn - (half - sn * cs) * cs / sn;
                }
                p->avg = p->emit * share;
                mp = std::move(p);
                break;
            }
            case MPT_MAT_PHASE: {
                auto p = std::make_unique<phase_material>();
                p->albedo = color(m.albedo[0], m.albedo[1], m.albedo[2]);
                p->g = m.g;
                mp = std::move(p);
                break;
            }
            default:
                err = "unknown material type";
                return MPT_E_UNSUPPORTED;
        }
        mp->table = &s.table;
        if (m.normal_map >= 0 && m.type != MPT_MAT_PRINCIPLED) {  // the scene file's normal_mapped wrapper
            auto w = std::make_unique<normal_mapped>();
            w->inner = std::move(mp);
            w->map = tex(m.normal_map);
            w->table = &s.table;
            mp = std::move(w);
        }
        if (m.glow && m.glow_texture >= 0) {
            auto w = std::make_unique<glowing>();
            w->inner = std::move(mp);
            w->glow = tex(m.glow_texture);
            w->table = &s.table;
            mp = std::move(w);
        }
        s.mats.push_back(std::move(mp));
    }
    auto matp = [&](int i) -> const material* { return i >= 0 ? s.mats[i].get() : nullptr; };
    auto prim = [&](int64_t i, int64_t id) -> const hittable* {
        const double* q = d.prim_data + 9 * i;
        const material* m = matp(d.prim_material[i]);
        switch (d.prim_kind[i]) {
            case MPT_PRIM_SPHERE: {
                auto p = std::make_unique<sphere>();
                p->center = point3(q[0], q[1], q[2]);
                p->radius = q[3];
                p->mat = m;
                p->id = id;
                return own(s, std::move(p));
            }
            case MPT_PRIM_QUAD:
                return own(s, std::make_unique<quad>(point3(q[0], q[1], q[2]), vec3(q[3], q[4], q[5]), vec3(q[6], q[7], q[8]),
                                                     m, id));
            default: {
                auto t = std::make_unique<triangle>(point3(q[0], q[1], q[2]), point3(q[3], q[4], q[5]),
                                                    point3(q[6], q[7], q[8]), m, id);
                const double* vn = d.prim_normals ? d.prim_normals + 9 * i : nullptr;
                if (vn && !std::isnan(vn[0])) {
                    t->has_normals = true;
                    t->n0 = vec3(vn[0], vn[1], vn[2]);
                    t->n1 = vec3(vn[3], vn[4], vn[5]);
                    t->n2 = vec3(vn[6], vn[7], vn[8]);
                }
                const double* vt = d.prim_uvs ? d.prim_uvs + 6 * i : nullptr;
                if (vt && !std::isnan(vt[0])) {
                    t->has_uvs = true;
                    t->t0 = vec3(vt[0], vt[1], 0);
                    t->t1 = vec3(vt[2], vt[3], 0);
                    t->t2 = vec3(vt[4], vt[5], 0);
                }
                return own(s, std::move(t));
            }
        }
    };
    std::vector<std::pair<int64_t, const hittable*>> meshes;  // (first, mesh) of meshes already built
    std::vector<const hittable*> world_items;                   // the objects the BVH is built over
    for (int o = 0; o < d.n_objects; o++) {
        int64_t first = d.object_first[o], count = d.object_count[o];
        int64_t idb = d.object_id_base ? d.object_id_base[o] : 0;
        int xi = d.object_xform ? d.object_xform[o] : -1;
        const hittable* h = nullptr;
        switch (d.object_kind[o]) {
            case MPT_OBJ_PRIM:
                h = prim(first, idb);
                break;
            case MPT_OBJ_LIST: {
                auto l = std::make_unique<hlist>();
                for (int64_t k = 0; k < count; k++) l->add(prim(first + k, idb + k));
                h = own(s, std::move(l));
                break;
            }
            case MPT_OBJ_MESH: {
                const hittable* tree = nullptr;
                for (auto& mm : meshes)
                    if (mm.first == first) tree = mm.second;
                int64_t shift = idb;
                if (!tree) {
                    std::vector<const hittable*> tris;
                    const int32_t am = d.object_alpha ? d.object_alpha[o] : -1;
                    const texture_t* mask = nullptr;
                    if (am >= 0) {
                        s.masks.push_back(std::make_unique<texture_t>(tex(am)));
                        mask = s.masks.back().get();
                    }
                    for (int64_t k = 0; k < count; k++) {
                        const hittable* t = prim(first + k, k);
                        if (mask) {
                            auto* tri = const_cast<triangle*>(static_cast<const triangle*>(t));
                            tri->alpha = mask;
                            tri->mesh_first = first;
                        }
                        tris.push_back(t);
                    }
                    const bvh* b = own(s, std::make_unique<bvh>(tris));
                    tree = own(s, std::make_unique<mesh_obj>(b, tris));
                    meshes.push_back({first, tree});
                }
                h = own(s, std::make_unique<idshift>(tree, shift));
                break;
            }
            case MPT_OBJ_VOXEL: {  // voxel_blocks (voxel.h:341-382): its blocks in a BVH
                const hittable* tree = nullptr;
                for (auto& mm : meshes)
                    if (mm.first == first) tree = mm.second;
                if (!tree) {
                    std::vector<const hittable*> blocks;
                    for (int64_t k = 0; k < count; k++) {
                        const double* q = d.prim_data + 9 * (first + k);
                        if (d.prim_kind[first + k] != MPT_PRIM_VOXBLOCK || q[6] < 0 || q[6] >= d.n_vox_models) {
                            err = "voxel objects hold voxel blocks of known models";
                            return MPT_E_INVALID;
                        }
                        auto vb = std::make_unique<vox_block>();
                        vb->m = &d.vox_models[int(q[6])];
                        vb->cells = d.vox_cells + vb->m->cells;
                        for (int i = 0; i < 256; i++) vb->mats[i] = matp(vb->m->material[i]);
                        vb->edge_width = vb->m->edge_width;
                        vb->edge = vb->edge_width > 0 ? matp(vb->m->edge_material) : nullptr;
                        for (int a = 0; a < 3; a++) vb->lo[a] = int(q[a]), vb->hi[a] = int(q[3 + a]);
                        vb->id = k;
                        blocks.push_back(own(s, std::move(vb)));
                    }
                    tree = own(s, std::make_unique<bvh>(blocks));
                    meshes.push_back({first, tree});
                }
                h = own(s, std::make_unique<idshift>(tree, idb));
                break;
            }
            case MPT_OBJ_MEDIUM: {
                const hittable* b;
                if (d.object_boundary_kind[o] == MPT_OBJ_PRIM) b = prim(first, idb);
                else {
                    auto l = std::make_unique<hlist>();
                    for (int64_t k = 0; k < count; k++) l->add(prim(first + k, idb + k));
                    b = own(s, std::move(l));
                }
                if (xi >= 0) {
                    const mpt_xform& x = d.xforms[xi];
                    affine a;
                    std::memcpy(a.m, x.m, sizeof(a.m));
                    std::memcpy(a.t, x.t, sizeof(a.t));
                    std::memcpy(a.im, x.inv_m, sizeof(a.im));
                    std::memcpy(a.it, x.inv_t, sizeof(a.it));
                    b = own(s, std::make_unique<instance>(b, a));
                    xi = -1;
                }
                auto m = std::make_unique<medium>();
                m->boundary = b;
                m->density = d.object_density[o];
                m->phase = matp(d.object_material[o]);
                m->index = uint32_t(o);
                h = own(s, std::move(m));
                break;
            }
            default:
                err = "unknown object kind";
                return MPT_E_UNSUPPORTED;
        }
        if (xi >= 0) {
            const mpt_xform& x = d.xforms[xi];
            affine a;
            std::memcpy(a.m, x.m, sizeof(a.m));
            std::memcpy(a.t, x.t, sizeof(a.t));
            std::memcpy(a.im, x.inv_m, sizeof(a.im));
            std::memcpy(a.it, x.inv_t, sizeof(a.it));
            h = own(s, std::make_unique<instance>(h, a));
        }
        s.world.push_back(h);
        if (!(d.object_flags && (d.object_flags[o] & MPT_OBJF_LIGHT_ONLY))) {  // sampled only, never hit
            world_items.push_back(h);
            s.world_box = aabb(s.world_box, h->bounding_box());
        }
    }
    s.tree = std::make_unique<bvh>(world_items);
    s.items = world_items;
    for (int k = 0; k < d.n_fogs; k++) {
        auto f = std::make_unique<height_fog>();
        f->density = d.fog_params[3 * k];
        f->scale = d.fog_params[3 * k + 1];
        f->base = d.fog_params[3 * k + 2];
        f->phase = matp(d.fog_material[k]);
        f->index = uint32_t(d.n_objects + 1 + k);
        s.fogs.push_back(own(s, std::move(f)));
    }
    if (d.has_ground) {
        auto g = std::make_unique<ground_plane>();
        g->height = d.ground_height;
        g->mat = matp(d.ground_material);
        g->id = 0;  // one past the last canonical ID
        for (int o = 0; o < d.n_objects; o++)
            if (!(d.object_flags && (d.object_flags[o] & MPT_OBJF_LIGHT_ONLY)))
                g->id = std::max(g->id, d.object_id_base[o] + d.object_count[o]);
        s.ground = own(s, std::move(g));
    }

    if (d.env_width > 0) {
        s.env = std::make_unique<environment>();
        s.env->width = d.env_width;
        s.env->height = d.env_height;
        s.env->intensity = d.env_intensity;
        s.env->rotation = d.env_rotation_degrees / 360;
        s.env->pixels.resize(size_t(d.env_width) * d.env_height);
        for (size_t k = 0; k < s.env->pixels.size(); k++)
            s.env->pixels[k] = color(d.env_rgb[3 * k], d.env_rgb[3 * k + 1], d.env_rgb[3 * k + 2]);
        s.env->build();
    }
    // light_list::build (light_list.h:45-96)
    size_t n = size_t(d.n_lights);
    s.strategy = d.light_strategy;
    if (d.has_sun)
        s.sun = std::make_unique<sun_light>(vec3(d.sun_direction[0], d.sun_direction[1], d.sun_direction[2]),
                                            d.sun_angle_degrees,
                                            color(d.sun_irradiance[0], d.sun_irradiance[1], d.sun_irradiance[2]));
    for (size_t k = 0; k < n; k++) {
        const int ob = d.light_object[k];
        if (ob == -2 && !s.sun) return MPT_E_INVALID;  // a sun light without a sun
        s.lights.push_back(ob == -2 ? s.sun.get() : ob < 0 ? nullptr : s.world[ob]);
    }
    s.prob.assign(n, n > 0 ? 1.0 / n : 0.0);
    if (n > 0 && s.strategy == 1) {
        const aabb& b = s.world_box;
        double radius = 0;
        if (b.ax[0].size() > 0 && b.ax[1].size() > 0 && b.ax[2].size() > 0)
            radius = 0.5 * std::sqrt(b.ax[0].size() * b.ax[0].size() + b.ax[1].size() * b.ax[1].size() +
                                     b.ax[2].size() * b.ax[2].size());
        if (!std::isfinite(radius) || radius <= 0) radius = 1;
        std::vector<double> power(n);
        double known_sum = 0;
        int known = 0;
        for (size_t k = 0; k < n; k++) {
            power[k] = d.light_object[k] == -2 ? s.sun->power(radius)
                       : s.lights[k]               ? s.lights[k]->power()
                                                   : s.env->power(radius);
            if (std::isfinite(power[k]) && power[k] > 0) {
                known_sum += power[k];
                known++;
            }
        }
        double fallback = known > 0 ? known_sum / known : 1.0, total = 0;
        for (auto& p : power) {
            if (!std::isfinite(p) || p <= 0) p = fallback;
            total += p;
        }
        double uf = d.light_uniform_fraction;
        for (size_t k = 0; k < n; k++) s.prob[k] = (1 - uf) * power[k] / total + uf / n;
    }
    s.cdf.assign(n + 1, 0.0);
    for (size_t k = 0; k < n; k++) s.cdf[k + 1] = s.cdf[k] + s.prob[k];
    if (n > 0) s.cdf[n] = 1;
    return MPT_OK;
}

// ------------------------------------------------------------------ camera (camera.h)
struct cam {
    int w, h, spp, max_depth, max_diffuse, max_specular, max_scatter;
    bool transparent_shadows, features, use_sky, has_env, sun_disk_visible;
    int sqrt_spp, blades;
    double recip, ortho, defocus, focus, blade_rot;
    vec3 center, p00, du, dv, w_axis, disk_u, disk_v;
    color background;
    explicit cam(const mpt_camera& c, bool env) {
        w = c.image_width;
        h = std::max(1, int(c.image_width / c.aspect_ratio));
        spp = c.samples_per_pixel;
        max_depth = c.max_depth;
        max_diffuse = c.max_diffuse;
        max_specular = c.max_specular;
        max_scatter = c.max_scatter;
        transparent_shadows = c.transparent_shadows;
        features = c.collect_features;
        use_sky = c.use_sky;
        has_env = env;
        sun_disk_visible = !c.sun_disk_hidden;
        background = color(c.background[0], c.background[1], c.background[2]);
        sqrt_spp = int(std::sqrt(spp));
        recip = 1.0 / sqrt_spp;
        blades = c.blades;
        ortho = c.ortho_height;
        defocus = c.defocus_angle;
        focus = c.focus_dist;
        blade_rot = c.blade_rotation;
        point3 lf(c.lookfrom[0], c.lookfrom[1], c.lookfrom[2]), la(c.lookat[0], c.lookat[1], c.lookat[2]);
        vec3 vup(c.vup[0], c.vup[1], c.vup[2]);
        center = lf;
        double theta = c.vfov * pi / 180.0;
        double hh = std::tan(theta / 2);
        double vh = ortho > 0 ? ortho : 2 * hh * focus;
        double vw = vh * (double(w) / h);
        w_axis = unit_vector(lf - la);
        vec3 u = unit_vector(cross(vup, w_axis));
        vec3 v = cross(w_axis, u);
        vec3 vu = vw * u, vv = vh * -v;
        du = vu / w;
        dv = vv / h;
        point3 ul = center - (focus * w_axis) - vu / 2 - vv / 2;
        p00 = ul + 0.5 * (du + dv);
        double dr = focus * std::tan((defocus / 2) * pi / 180.0);
        disk_u = u * dr;
        disk_v = v * dr;
    }
    ray get_ray(int i, int j, int pass, sampler& s) const {
        int strata = sqrt_spp * sqrt_spp;
        double ox, oy;
        if (pass >= strata) {
            ox = s.next() - 0.5;
            oy = s.next() - 0.5;
        } else {
            int cell = (pass + i * 7919 + j * 104729) % strata;
            ox = ((cell % sqrt_spp + s.next()) * recip) - 0.5;
            oy = ((cell / sqrt_spp + s.next()) * recip) - 0.5;
        }
        point3 ps = p00 + ((i + ox) * du) + ((j + oy) * dv);
        point3 eye = ortho > 0 ? ps + focus * w_axis : center;
        point3 org = eye;
        if (defocus > 0) {
            vec3 p;
            if (blades >= 3) {
                int k = std::min(int(s.next() * blades), blades - 1);
                double a0 = blade_rot * pi / 180.0 + 2 * pi * k / blades, a1 = a0 + 2 * pi / blades;
                double a = s.next(), b = s.next();
                if (a + b > 1) {
                    a = 1 - a;
                    b = 1 - b;
                }
                p = a * vec3(std::cos(a0), std::sin(a0), 0) + b * vec3(std::cos(a1), std::sin(a1), 0);
            } else {
                while (true) {
                    double x = s.next(-1, 1);
                    double y = s.next(-1, 1);
                    p = vec3(x, y, 0);
                    if (p.length_squared() < 1) break;
                }
            }
            org = eye + ((p[0] * disk_u) + (p[1] * disk_v));
        }
        return ray(org, ps - org);
    }
};

struct counters {
    uint64_t extension = 0, shadow = 0;
};

// camera::escaped_radiance (camera.h:515-529): the sky, plus the sun for
// every ray but a camera ray while the sun's disk is hidden
color escaped(const scene& S, const cam& C, const vec3& d, bool camera_ray) {
    color sky;
    if (C.has_env) {
        sky = S.env->radiance(d);
    } else if (!C.use_sky) {
        sky = C.background;
    } else {
        vec3 u = unit_vector(d);
        double a = 0.5 * (u.y() + 1.0);
        sky = (1.0 - a) * color(1.0, 1.0, 1.0) + a * color(0.5, 0.7, 1.0);
    }
    if (S.sun && (C.sun_disk_visible || !camera_ray)) sky = sky + S.sun->radiance(d);
    return sky;
}
double power_heuristic(double a, double b) {
    double a2 = a * a, b2 = b * b;
    return (a2 + b2) > 0 ? a2 / (a2 + b2) : 0;
}

color sample_light(const scene& S, const cam& C, const ray& r_in, const hit_record& rec, const pdf& bsdf, bool last,
                   sampler& s, int depth, counters& cnt) {
    vec3 to_light = S.lights_random(rec.p, s);
    double light_pdf = S.lights_pdf(rec.p, to_light);
    if (light_pdf <= 0) return color(0, 0, 0);
    color f = rec.mat->eval(r_in, rec, to_light);
    if (f.x() <= 0 && f.y() <= 0 && f.z() <= 0) return color(0, 0, 0);
    double weight = last ? 1.0 : power_heuristic(light_pdf, bsdf.value(to_light));
    ray shadow(rec.p, unit_vector(to_light));
    color emission, through(1, 1, 1);
    for (int k = 0;; k++) {
        hit_record lr;
        cnt.shadow++;
        hit_ctx hc{s.key, uint32_t(2 * depth + 1) | (uint32_t(k) << 20), true};
        if (!S.hit(shadow, interval(0.001, infinity), lr, hc)) {
            emission = escaped(S, C, to_light, false);
            break;
        }
        if (C.transparent_shadows && k < 16 && lr.mat->lets_shadows_through()) {
            through = through * lr.mat->interior_transmittance(shadow, lr);
            shadow = ray(lr.p, shadow.dir);
            continue;
        }
        emission = lr.mat->emitted(shadow, lr);
        break;
    }
    emission = through * emission;
    return f * emission * (weight / light_pdf);
}

color ray_color(const scene& S, const cam& C, const ray& camera_ray, sampler& s, counters& cnt) {  // camera.h:356-461
    color radiance(0, 0, 0), beta(1, 1, 1);
    ray r = camera_ray;
    bool specular = true, seen = false;
    double scatter_pdf = 0;
    point3 prev;
    int dc = 0, sc = 0, vc = 0;
    const bool no_lights = S.lights.empty();
    for (int depth = 0; depth < C.max_depth; depth++) {
        hit_record rec;
        cnt.extension++;
        hit_ctx hc{s.key, uint32_t(2 * depth), true};
        if (!S.hit(r, interval(0.001, infinity), rec, hc)) {
            color sky = escaped(S, C, r.dir, depth == 0);
            if (seen) break;
            if (specular || no_lights) radiance = radiance + beta * sky;
            else radiance = radiance + beta * sky * power_heuristic(scatter_pdf, S.lights_pdf(prev, r.dir));
            break;
        }
        beta = beta * rec.mat->interior_transmittance(r, rec);
        color e = rec.mat->emitted(r, rec);
        if (!seen && (e.x() > 0 || e.y() > 0 || e.z() > 0)) {
            if (specular || no_lights) radiance = radiance + beta * e;
            else radiance = radiance + beta * e * power_heuristic(scatter_pdf, S.lights_pdf(prev, r.dir));
        }
        scatter_record srec;
        if (!rec.mat->scatter(r, rec, srec, s)) break;
        if (srec.skip_pdf) {
            if (C.max_specular > 0 && ++sc > C.max_specular) break;
            beta = beta * srec.attenuation;
            r = srec.skip_pdf_ray;
            if (C.transparent_shadows && rec.mat->lets_shadows_through() && !no_lights && depth > 0 && !specular) seen = true;
            specular = true;
        } else {
            bool volume = rec.mat->is_volume();
            int& count = volume ? vc : dc;
            int cap = volume ? C.max_scatter : C.max_diffuse;
            bool capped = cap > 0 && ++count >= cap;
            bool last = depth == C.max_depth - 1 || capped;
            seen = false;
            if (!no_lights) radiance = radiance + beta * sample_light(S, C, r, rec, *srec.pdf_ptr, last, s, depth, cnt);
            if (capped) break;
            vec3 dir = srec.pdf_ptr->generate(s);
            scatter_pdf = srec.pdf_ptr->value(dir);
            if (scatter_pdf <= 0) break;
            beta = beta * rec.mat->eval(r, rec, dir) / scatter_pdf;
            prev = rec.p;
            r = ray(rec.p, dir);
            specular = false;
        }
        if (depth >= 3) {
            double survive = std::fmin(std::fmax(beta.x(), std::fmax(beta.y(), beta.z())), 0.95);
            if (s.next() >= survive) break;
            beta = beta / survive;
        }
    }
    return radiance;
}

void first_surface(const scene& S, const ray& cr, uint64_t key, color& albedo, color& normal) {  // camera.h:253-277
    ray r = cr;
    color tint(1, 1, 1);
    sampler s;
    s.key = mix64(key ^ 0x3C6EF372FE94F82Bull);
    for (int b = 0; b < 8; b++) {
        hit_record rec;
        hit_ctx hc{s.key, uint32_t(b), true};
        if (!S.hit(r, interval(0.001, infinity), rec, hc)) break;
        tint = tint * rec.mat->interior_transmittance(r, rec);
        scatter_record srec;
        if (!rec.mat->scatter(r, rec, srec, s)) {
            albedo = tint;
            normal = rec.normal;
            return;
        }
        if (!srec.skip_pdf) {
            albedo = tint * srec.attenuation;
            normal = rec.normal;
            return;
        }
        tint = tint * srec.attenuation;
        r = srec.skip_pdf_ray;
    }
    albedo = tint;
    normal = color(0, 0, 0);
}

// ------------------------------------------------------------------ display (postprocess.h, tonemap.h, color3.h)
// Restatement of postprocess() (postprocess.h:108-153), display_settings::
// apply (tonemap.h:28-67) and to_byte (color3.h:17-33), whole images at a
// time. Pinned against the compiled reference (tests/test_display.py) and
// by closed forms: no bloom or vignette is the identity, a uniform image
// blurs to itself, the vignette factor at the centre and corners.
struct display_img {
    int w = 0, h = 0;
    std::vector<color> px;
    color at(int i, int j) const {  // postprocess.h:28-32
        i = std::clamp(i, 0, w - 1);
        j = std::clamp(j, 0, h - 1);
        return px[size_t(j) * w + i];
    }
    color sample(double x, double y) const {  // postprocess.h:35-42
        x -= 0.5;
        y -= 0.5;
        const int i = int(std::floor(x)), j = int(std::floor(y));
        const double fx = x - i, fy = y - j;
        return (1 - fy) * ((1 - fx) * at(i, j) + fx * at(i + 1, j)) + fy * ((1 - fx) * at(i, j + 1) + fx * at(i + 1, j + 1));
    }
};

display_img half_size(const display_img& in) {  // postprocess.h:47-59
    display_img out;
    out.w = std::max(1, in.w / 2);
    out.h = std::max(1, in.h / 2);
    out.px.resize(size_t(out.w) * out.h);
    for (int j = 0; j < out.h; j++)
        for (int i = 0; i < out.w; i++)
            out.px[size_t(j) * out.w + i] =
                0.25 * (in.at(2 * i, 2 * j) + in.at(2 * i + 1, 2 * j) + in.at(2 * i, 2 * j + 1) + in.at(2 * i + 1, 2 * j + 1));
    return out;
}

display_img gauss_blur(const display_img& in, double sigma) {  // postprocess.h:62-89
    const int r = int(std::ceil(3 * sigma));
    std::vector<double> k(2 * r + 1);
    double sum = 0;
    for (int t = -r; t <= r; t++) sum += k[t + r] = std::exp(-0.5 * t * t / (sigma * sigma));
    for (double& x : k) x /= sum;
    display_img tmp = in, out = in;
    for (int j = 0; j < in.h; j++)
        for (int i = 0; i < in.w; i++) {
            color c(0, 0, 0);
            for (int t = -r; t <= r; t++) c = c + k[t + r] * in.at(i + t, j);
            tmp.px[size_t(j) * in.w + i] = c;
        }
    for (int j = 0; j < in.h; j++)
        for (int i = 0; i < in.w; i++) {
            color c(0, 0, 0);
            for (int t = -r; t <= r; t++) c = c + k[t + r] * tmp.at(i, j + t);
            out.px[size_t(j) * in.w + i] = c;
        }
    return out;
}

std::vector<color> post_image(const std::vector<color>& pixels, int width, int height, double bloom, double threshold,
                              double radius, double vignette) {
    std::vector<color> out = pixels;
    if (bloom > 0) {
        display_img bright;
        bright.w = width;
        bright.h = height;
        bright.px.resize(pixels.size());
        for (size_t k = 0; k < pixels.size(); k++) {
            const double l = luminance(pixels[k]);
            bright.px[k] = (std::isfinite(l) && l > threshold) ? pixels[k] * ((l - threshold) / l) : color(0, 0, 0);
        }
        const double reach = std::fmax(radius * width, 1.0);
        const int levels = std::clamp(int(std::ceil(std::log2(reach))), 1, 10);
        std::vector<display_img> blurred;
        display_img level = bright;
        for (int k = 0; k < levels && level.w > 1 && level.h > 1; k++) {
            level = half_size(level);
            blurred.push_back(gauss_blur(level, 1.5));
        }
        for (int j = 0; j < height; j++)
            for (int i = 0; i < width; i++) {
                color glow(0, 0, 0);
                for (const auto& b : blurred) {
                    const double sx = double(b.w) / width, sy = double(b.h) / height;
                    glow = glow + b.sample((i + 0.5) * sx, (j + 0.5) * sy);
                }
                if (!blurred.empty()) {
                    color& o = out[size_t(j) * width + i];
                    o = o + bloom / blurred.size() * glow;
                }
            }
    }
    if (vignette > 0) {
        const double cx = 0.5 * width, cy = 0.5 * height;
        const double corner2 = cx * cx + cy * cy;
        for (int j = 0; j < height; j++)
            for (int i = 0; i < width; i++) {
                const double dx = i + 0.5 - cx, dy = j + 0.5 - cy;
                const double f = 1 + vignette * (dx * dx + dy * dy) / corner2;
                out[size_t(j) * width + i] = out[size_t(j) * width + i] / (f * f);
            }
    }
    return out;
}

double aces_curve(double x) {  // tonemap.h:63-67
    const double a = x * (x + 0.0245786) - 0.000090537;
    const double b = x * (0.983729 * x + 0.4329510) + 0.238081;
    return a / b;
}
double unit_clamp(double x) { return x < 0 ? 0 : (x > 1 ? 1 : x); }

color display_apply(const color& radiance, double exposure, int op) {  // tonemap.h:28-37
    const color c = std::pow(2.0, exposure) * radiance;
    if (op == 1) return color(c.x() / (1 + c.x()), c.y() / (1 + c.y()), c.z() / (1 + c.z()));
    if (op == 2) {
        color v(0.59719 * c.x() + 0.35458 * c.y() + 0.04823 * c.z(), 0.07600 * c.x() + 0.90834 * c.y() + 0.01566 * c.z(),
                0.02840 * c.x() + 0.13383 * c.y() + 0.83777 * c.z());
        v = color(aces_curve(v.x()), aces_curve(v.y()), aces_curve(v.z()));
        v = color(1.60475 * v.x() + -0.53108 * v.y() + -0.07367 * v.z(),
                  -0.10208 * v.x() + 1.10813 * v.y() + -0.00605 * v.z(),
                  -0.00327 * v.x() + -0.07276 * v.y() + 1.07602 * v.z());
        return color(unit_clamp(v.x()), unit_clamp(v.y()), unit_clamp(v.z()));
    }
    return c;
}

unsigned char srgb_byte(double x) {  // color3.h:17-33
    double s;
    if (x <= 0) s = 0;
    else if (x <= 0.0031308) s = 12.92 * x;
    else s = 1.055 * std::pow(x, 1.0 / 2.4) - 0.055;
    s = s < 0.000 ? 0.000 : (s > 0.999 ? 0.999 : s);
    return (unsigned char)(256 * s);
}

// ------------------------------------------------------------------ denoiser (denoiser.h:38-193)
// Edge-avoiding a-trous filter on lighting = radiance / albedo, with SVGF's
// variance-scaled luminance weight; a restatement of denoiser::run.
struct atrous {
    int iterations = 5;
    double sigma_luminance = 4, sigma_normal = 0.1, sigma_albedo = 0.1;
    static constexpr double eps = 1e-3;  // denoiser.h:95

    std::vector<color> run(const std::vector<color>& rad, const std::vector<color>& alb, const std::vector<color>& nrm,
                           const std::vector<double>& variance, int w, int h) const {
        const size_t n = size_t(w) * h;
        const color e(eps, eps, eps);
        std::vector<color> light(n);
        std::vector<double> var(n);
        for (size_t k = 0; k < n; k++) {  // denoiser.h:48-55
            color a = alb[k] + e;
            light[k] = color(rad[k].x() / a.x(), rad[k].y() / a.y(), rad[k].z() / a.z());
            double la = luminance(a);
            var[k] = variance[k] / (la * la);
        }
        clamp_fireflies(light, var, w, h);
        std::vector<color> light2(n);
        std::vector<double> var2(n);
        const double inv_n = 1 / (sigma_normal * sigma_normal);
        const double inv_a = 1 / (sigma_albedo * sigma_albedo);
        for (int it = 0; it < iterations; it++) {
            std::vector<double> bvar = blur(var, w, h);
            for (int j = 0; j < h; j++)
                for (int i = 0; i < w; i++) pass(i, j, 1 << it, w, h, alb, nrm, light, var, bvar, inv_n, inv_a, light2, var2);
            light.swap(light2);
            var.swap(var2);
        }
        std::vector<color> out(n);
        for (size_t k = 0; k < n; k++) out[k] = light[k] * (alb[k] + e);
        return out;
    }

    // denoiser.h:98-124: a pixel brighter than its 8 neighbours whose
    // deviation exceeds half its luminance drops to the brightest neighbour.
    static void clamp_fireflies(std::vector<color>& light, std::vector<double>& var, int w, int h) {
        const std::vector<color> src = light;
        for (int j = 1; j < h - 1; j++)
            for (int i = 1; i < w - 1; i++) {
                const size_t p = size_t(j) * w + i;
                const double lum = luminance(src[p]);
                if (lum <= 0 || std::sqrt(var[p]) < 0.5 * lum) continue;
                double top = 0;
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++)
                        if (dx || dy) top = std::fmax(top, luminance(src[p + dy * w + dx]));
                if (lum > top) {
                    const double f = top / lum;
                    light[p] = f * src[p];
                    var[p] *= f * f;
                }
            }
    }

    // denoiser.h:169-188: normalised 3x3 binomial blur, clipped at the borders.
    static std::vector<double> blur(const std::vector<double>& v, int w, int h) {
        const double k3[3] = {0.25, 0.5, 0.25};
        std::vector<double> out(v.size());
        for (int j = 0; j < h; j++)
            for (int i = 0; i < w; i++) {
                double acc = 0, wsum = 0;
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        const int x = i + dx, y = j + dy;
                        if (x < 0 || x >= w || y < 0 || y >= h) continue;
                        const double wt = k3[dx + 1] * k3[dy + 1];
                        acc += wt * v[size_t(y) * w + x];
                        wsum += wt;
                    }
                out[size_t(j) * w + i] = acc / wsum;
            }
        return out;
    }

    // denoiser.h:126-167: one 5x5 tap pass at spacing `step`.
    void pass(int i, int j, int step, int w, int h, const std::vector<color>& alb, const std::vector<color>& nrm,
              const std::vector<color>& light, const std::vector<double>& var, const std::vector<double>& bvar,
              double inv_n, double inv_a, std::vector<color>& out, std::vector<double>& out_var) const {
        const double k5[5] = {1.0 / 16, 1.0 / 4, 3.0 / 8, 1.0 / 4, 1.0 / 16};
        const size_t p = size_t(j) * w + i;
        const double lp = luminance(light[p]);
        const double scale = sigma_luminance * std::sqrt(std::fmax(bvar[p], 0.0)) + 1e-10;
        color acc(0, 0, 0);
        double wsum = 0, vsum = 0;
        for (int dy = -2; dy <= 2; dy++) {
            const int y = j + dy * step;
            if (y < 0 || y >= h) continue;
            for (int dx = -2; dx <= 2; dx++) {
                const int x = i + dx * step;
                if (x < 0 || x >= w) continue;
                const size_t q = size_t(y) * w + x;
                const double wn = std::exp(-(nrm[p] - nrm[q]).length_squared() * inv_n);
                const double wa = std::exp(-(alb[p] - alb[q]).length_squared() * inv_a);
                const double wl = std::exp(-std::fabs(lp - luminance(light[q])) / scale);
                const double wt = k5[dx + 2] * k5[dy + 2] * wn * wa * wl;
                acc = acc + wt * light[q];
                wsum += wt;
                vsum += wt * wt * var[q];
            }
        }
        out[p] = acc / wsum;
        out_var[p] = vsum / (wsum * wsum);
    }
};

}  // namespace orc

// ======================================================================== C ABI
using namespace orc;

struct orc_scene {
    scene s;
};

static thread_local std::string orc_err;

extern "C" {

const char* orc_last_error(void) { return orc_err.c_str(); }

// denoiser::run on interleaved RGB doubles (denoiser.h:45-78).
int32_t orc_denoise(const double* rad, const double* alb, const double* nrm, const double* var, int32_t w, int32_t h,
                    int32_t iterations, double sigma_l, double sigma_n, double sigma_a, double* out) {
    const size_t n = size_t(w) * h;
    std::vector<color> r(n), a(n), nn(n);
    std::vector<double> v(var, var + n);
    for (size_t k = 0; k < n; k++) {
        r[k] = color(rad[3 * k], rad[3 * k + 1], rad[3 * k + 2]);
        a[k] = color(alb[3 * k], alb[3 * k + 1], alb[3 * k + 2]);
        nn[k] = color(nrm[3 * k], nrm[3 * k + 1], nrm[3 * k + 2]);
    }
    atrous f;
    f.iterations = iterations;
    f.sigma_luminance = sigma_l;
    f.sigma_normal = sigma_n;
    f.sigma_albedo = sigma_a;
    std::vector<color> o = f.run(r, a, nn, v, w, h);
    for (size_t k = 0; k < n; k++)
        for (int c = 0; c < 3; c++) out[3 * k + c] = o[k][c];
    return 0;
}

int32_t orc_scene_create(const mpt_scene_desc* d, orc_scene** out) {
    auto p = std::make_unique<orc_scene>();
    std::string err;
    int rc = build_scene(*d, p->s, err);
    if (rc != MPT_OK) {
        orc_err = err;
        return rc;
    }
    *out = p.release();
    return MPT_OK;
}

void orc_scene_destroy(orc_scene* s) { delete s; }

// Closest hit for n rays (media skipped): canonical ID, t, front face.
int32_t orc_trace(orc_scene* S, const double* rays, const double* tt, int64_t n, int64_t* id, double* t, uint8_t* front,
                  int32_t threads) {
    auto work = [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; i++) {
            ray r(point3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]), vec3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]));
            hit_record rec;
            hit_ctx hc{0, 0, false};
            if (S->s.hit(r, interval(tt[2 * i], tt[2 * i + 1]), rec, hc)) {
                id[i] = rec.id;
                t[i] = rec.t;
                front[i] = rec.front_face;
            } else {
                id[i] = -1;
                t[i] = 0;
                front[i] = 0;
            }
        }
    };
    int nt = std::max(1, threads);
    std::vector<std::thread> pool;
    int64_t chunk = (n + nt - 1) / nt;
    for (int k = 1; k < nt; k++) pool.emplace_back(work, std::min(n, k * chunk), std::min(n, (k + 1) * chunk));
    work(0, std::min(n, chunk));
    for (auto& th : pool) th.join();
    return MPT_OK;
}

// What a hit hands its material, per ray (16 doubles, NaN on a miss): u, v,
// dp/du, dp/dv, the base texture's value (lambertian, principled), the
// shading normal after a normal map, principled's metallic and alpha.
int32_t orc_surface(orc_scene* S, const double* rays, const double* tt, int64_t n, double* out) {
    for (int64_t i = 0; i < n; i++) {
        double* o = out + 16 * i;
        for (int k = 0; k < 16; k++) o[k] = std::nan("");
        ray r(point3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]), vec3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]));
        hit_record rec;
        hit_ctx hc{0, 0, false};
        if (!S->s.hit(r, interval(tt[2 * i], tt[2 * i + 1]), rec, hc)) continue;
        const material* m = rec.mat;
        vec3 shading = rec.normal;
        if (auto nm = dynamic_cast<const normal_mapped*>(m)) {
            shading = nm->perturbed(r, rec).normal;
            m = nm->inner.get();
        }
        color base(0, 0, 0);
        double metallic = 0, alpha = 0;
        if (auto l = dynamic_cast<const lambertian*>(m)) {
            base = l->tex.value(rec.u, rec.v, rec.p);
        } else if (auto p = dynamic_cast<const principled*>(m)) {
            base = p->base.value(rec.u, rec.v, rec.p);
            metallic = p->metallic_at(rec);
            alpha = p->alpha_at(rec);
            shading = p->base_normal(r, rec);
        }
        const double v[16] = {rec.u, rec.v, rec.dpdu.x(), rec.dpdu.y(), rec.dpdu.z(), rec.dpdv.x(), rec.dpdv.y(), rec.dpdv.z(),
                              base.x(), base.y(), base.z(), shading.x(), shading.y(), shading.z(), metallic, alpha};
        std::memcpy(o, v, sizeof(v));
    }
    return MPT_OK;
}

// Linear scan over every world object: the closest hit any correct BVH
// must return (brute-force check of the BVH).
int32_t orc_trace_linear(orc_scene* S, const double* rays, const double* tt, int64_t n, int64_t* id, double* t) {
    for (int64_t i = 0; i < n; i++) {
        ray r(point3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]), vec3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]));
        hit_record best;
        bool any = false;
        double closest = tt[2 * i + 1];
        hit_ctx hc{0, 0, false};
        std::vector<const hittable*> all(S->s.items);
        if (S->s.ground) all.push_back(S->s.ground);
        for (auto* o : all) {
            hit_record rec;
            if (o->hit(r, interval(tt[2 * i], closest), rec, hc)) {
                any = true;
                closest = rec.t;
                best = rec;
            }
        }
        id[i] = any ? best.id : -1;
        t[i] = any ? best.t : 0;
    }
    return MPT_OK;
}

// Grazing ties, listed as the reference's test harness lists them: after the
// closest hit, every primitive of the world is tested alone over the ray's
// whole interval (the instance transforms above it applied, medium
// boundaries left out); those whose t is bitwise equal to the closest hit's
// are the candidates. ties: their number (0 on a miss); tie_min/tie_max:
// their smallest and largest canonical ID (-1 on a miss).
struct leaf_prim {
    const hittable* prim;
    std::vector<const affine*> xf;  // outermost first
    int64_t shift;
};

static void expand_leaves(const hittable* h, std::vector<const affine*> xf, int64_t shift, std::vector<leaf_prim>& out) {
    if (auto l = dynamic_cast<const hlist*>(h)) {
        for (auto* o : l->objects) expand_leaves(o, xf, shift, out);
    } else if (auto in = dynamic_cast<const instance*>(h)) {
        xf.push_back(&in->x);
        expand_leaves(in->object, xf, shift, out);
    } else if (auto m = dynamic_cast<const mesh_obj*>(h)) {
        for (auto* t : m->tris) out.push_back({t, xf, shift});
    } else if (auto s = dynamic_cast<const idshift*>(h)) {
        expand_leaves(s->inner, xf, shift + s->shift, out);
    } else if (auto b = dynamic_cast<const bvh*>(h)) {  // a voxel model's blocks
        for (auto* o : b->prims) expand_leaves(o, xf, shift, out);
    } else if (dynamic_cast<const medium*>(h)) {
        return;
    } else {
        out.push_back({h, xf, shift});
    }
}

int32_t orc_trace_ties(orc_scene* S, const double* rays, const double* tt, int64_t n, int64_t* id, double* t,
                       int32_t* ties, int64_t* tie_min, int64_t* tie_max) {
    std::vector<leaf_prim> leaves;
    for (auto* o : S->s.items) expand_leaves(o, {}, 0, leaves);
    if (S->s.ground) leaves.push_back({S->s.ground, {}, 0});
    for (int64_t i = 0; i < n; i++) {
        ray r(point3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]), vec3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]));
        interval iv(tt[2 * i], tt[2 * i + 1]);
        hit_record rec;
        hit_ctx hc{0, 0, false};
        ties[i] = 0, tie_min[i] = -1, tie_max[i] = -1;
        if (!S->s.hit(r, iv, rec, hc)) {
            id[i] = -1, t[i] = 0;
            continue;
        }
        id[i] = rec.id;
        t[i] = rec.t;
        for (const leaf_prim& lp : leaves) {
            ray local = r;
            for (const affine* x : lp.xf) local = ray(x->inv_point(local.orig), x->inv_vector(local.dir));
            hit_record one;
            if (!lp.prim->hit(local, iv, one, hc) || std::memcmp(&one.t, &rec.t, sizeof(double)) != 0) continue;
            const int64_t c = one.id + lp.shift;
            tie_min[i] = ties[i] == 0 ? c : std::min(tie_min[i], c);
            tie_max[i] = ties[i] == 0 ? c : std::max(tie_max[i], c);
            ties[i]++;
        }
    }
    return MPT_OK;
}

int32_t orc_light_probabilities(orc_scene* S, double* out) {
    for (size_t k = 0; k < S->s.prob.size(); k++) out[k] = S->s.prob[k];
    return MPT_OK;
}

int32_t orc_ggx_table(double* e, double* eavg) {
    ggx_table t;
    build_table(t);
    std::memcpy(e, t.e, sizeof(t.e));
    std::memcpy(eavg, t.e_avg, sizeof(t.e_avg));
    return MPT_OK;
}

// material eval and scatter-pdf value for fixed directions (as mpt_bsdf_eval).
int32_t orc_bsdf_eval(orc_scene* S, const int32_t* mats, const double* q, const uint8_t* front, int64_t n, double* eval,
                      double* pdfv) {
    for (int64_t i = 0; i < n; i++) {
        hit_record rec;
        rec.normal = vec3(q[9 * i], q[9 * i + 1], q[9 * i + 2]);
        ray r(point3(0, 0, 0), vec3(q[9 * i + 3], q[9 * i + 4], q[9 * i + 5]));
        vec3 wi(q[9 * i + 6], q[9 * i + 7], q[9 * i + 8]);
        rec.p = point3(0, 0, 0);
        rec.t = 1;
        rec.front_face = front[i] != 0;
        rec.mat = S->s.mats[mats[i]].get();
        sampler smp;
        scatter_record sr;
        bool sc = rec.mat->scatter(r, rec, sr, smp);
        color e = rec.mat->eval(r, rec, wi);
        eval[3 * i] = e.x(), eval[3 * i + 1] = e.y(), eval[3 * i + 2] = e.z();
        pdfv[i] = (sc && !sr.skip_pdf) ? sr.pdf_ptr->value(wi) : 0.0;
    }
    return MPT_OK;
}

// camera::render on the CPU: passes [first_pass, first_pass + passes) of
// the camera, tile-parallel over `threads` workers (camera.h:92-139);
// accumulates sums (not averages) into acc_* (3 doubles per pixel; l2 one).
int32_t orc_render(orc_scene* S, const mpt_camera* c, uint64_t seed, uint32_t scene_index, int32_t first_pass,
                   int32_t passes, int32_t threads, double* acc_rad, double* acc_alb, double* acc_nrm, double* acc_l2,
                   int64_t* ext_rays, int64_t* shadow_rays, double* seconds) {
    auto t0 = std::chrono::steady_clock::now();
    cam C(*c, S->s.env != nullptr);
    const int tile = 16;
    int tx = (C.w + tile - 1) / tile, ty = (C.h + tile - 1) / tile, tiles = tx * ty;
    int nt = std::max(1, threads);
    std::atomic<uint64_t> ext{0}, sh{0};
    for (int pass = first_pass; pass < first_pass + passes; pass++) {
        std::atomic<int> next{0};
        auto worker = [&]() {
            counters cnt;
            while (true) {
                int tl = next.fetch_add(1);
                if (tl >= tiles) break;
                int i0 = (tl % tx) * tile, j0 = (tl / tx) * tile;
                int i1 = std::min(i0 + tile, C.w), j1 = std::min(j0 + tile, C.h);
                for (int j = j0; j < j1; j++)
                    for (int i = i0; i < i1; i++) {
                        size_t k = size_t(j) * C.w + i;
                        sampler s;
                        s.key = key_of(seed, scene_index, uint32_t(pass), uint32_t(k));
                        ray r = C.get_ray(i, j, pass, s);
                        color col = ray_color(S->s, C, r, s, cnt);
                        for (int q = 0; q < 3; q++)
                            if (!std::isfinite(col[q])) {
                                col = color(0, 0, 0);
                                break;
                            }
                        for (int q = 0; q < 3; q++) acc_rad[3 * k + q] += col[q];
                        if (C.features) {
                            color a, nn;
                            first_surface(S->s, r, s.key, a, nn);
                            for (int q = 0; q < 3; q++) {
                                acc_alb[3 * k + q] += a[q];
                                acc_nrm[3 * k + q] += nn[q];
                            }
                            double l = luminance(col);
                            acc_l2[k] += l * l;
                        }
                    }
            }
            ext += cnt.extension;
            sh += cnt.shadow;
        };
        std::vector<std::thread> pool;
        for (int k = 1; k < std::min(nt, tiles); k++) pool.emplace_back(worker);
        worker();
        for (auto& th : pool) th.join();
    }
    if (ext_rays) *ext_rays = int64_t(ext.load());
    if (shadow_rays) *shadow_rays = int64_t(sh.load());
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return MPT_OK;
}

// The seeded ray batch of the closest-hit oracle: half camera rays of the
// camera (random jitter in each pixel), half cosine-distributed rays leaving
// the points those camera rays hit (the first bounce). Rays 6 doubles,
// intervals [0.001, inf).
int32_t orc_make_ray_batch(orc_scene* S, const mpt_camera* c, uint64_t seed, int64_t n, double* rays, double* tt) {
    cam C(*c, S->s.env != nullptr);
    int64_t half = n / 2;
    sampler s;
    s.key = mix64(seed ^ 0x5851F42D4C957F2Dull);
    auto put = [&](int64_t i, const ray& r) {
        rays[6 * i] = r.orig.x(), rays[6 * i + 1] = r.orig.y(), rays[6 * i + 2] = r.orig.z();
        rays[6 * i + 3] = r.dir.x(), rays[6 * i + 4] = r.dir.y(), rays[6 * i + 5] = r.dir.z();
        tt[2 * i] = 0.001, tt[2 * i + 1] = infinity;
    };
    for (int64_t i = 0; i < half; i++) {
        int px = std::min(int(s.next() * C.w), C.w - 1), py = std::min(int(s.next() * C.h), C.h - 1);
        ray r = C.get_ray(px, py, C.spp, s);  // pass >= strata: uniform jitter
        put(i, r);
    }
    for (int64_t i = half; i < n; i++) {
        const int64_t src = half > 0 ? (i - half) % half : 0;
        ray cr(point3(rays[6 * src], rays[6 * src + 1], rays[6 * src + 2]),
               vec3(rays[6 * src + 3], rays[6 * src + 4], rays[6 * src + 5]));
        hit_record rec;
        hit_ctx hc{0, 0, false};
        ray r = cr;
        if (S->s.hit(cr, interval(0.001, infinity), rec, hc)) {
            cosine_pdf cp(rec.normal);
            r = ray(rec.p, cp.generate(s));
        }
        put(i, r);
    }
    return MPT_OK;
}


int32_t orc_postprocess(const double* rad, int32_t w, int32_t h, double bloom, double threshold, double radius,
                        double vignette, double* out) {
    std::vector<color> px(size_t(w) * h);
    for (size_t k = 0; k < px.size(); k++) px[k] = color(rad[3 * k], rad[3 * k + 1], rad[3 * k + 2]);
    const std::vector<color> o = post_image(px, w, h, bloom, threshold, radius, vignette);
    for (size_t k = 0; k < o.size(); k++)
        for (int c = 0; c < 3; c++) out[3 * k + c] = o[k][c];
    return MPT_OK;
}

// The bytes write_png emits (image_io.h:25-37, 81-92): postprocess when
// post is set (and bloom or vignette is on), then apply and to_byte.
int32_t orc_display_bytes(const double* rad, int32_t w, int32_t h, double exposure, int32_t op, double bloom,
                          double threshold, double radius, double vignette, int32_t post, uint8_t* out) {
    std::vector<color> px(size_t(w) * h);
    for (size_t k = 0; k < px.size(); k++) px[k] = color(rad[3 * k], rad[3 * k + 1], rad[3 * k + 2]);
    if (post && (bloom > 0 || vignette > 0)) px = post_image(px, w, h, bloom, threshold, radius, vignette);
    for (size_t k = 0; k < px.size(); k++) {
        const color c = display_apply(px[k], exposure, op);
        for (int q = 0; q < 3; q++) out[3 * k + q] = srgb_byte(c[q]);
    }
    return MPT_OK;
}

// to_byte alone, for direct checks of the sRGB quantisation.
int32_t orc_to_byte(const double* x, int64_t n, uint8_t* out) {
    for (int64_t k = 0; k < n; k++) out[k] = srgb_byte(x[k]);
    return MPT_OK;
}

}  // extern "C"
