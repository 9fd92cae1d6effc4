// dev_shade.hpp - device-side materials, scatter densities, lights and the
// environment, in the reference's operand order (material.h, pdf.h,
// microfacet.h, light_list.h, environment.h, quad.h, sphere.h,
// instance.h). The reference allocates a pdf object per scatter
// (material.h:82, 134, 148); here a scatter density is a small tagged record
// in registers (lobe kind, frame normal, wo, alpha, eta, mixture weight).
#pragma once

#include "dev_trace.hpp"
#include "microfacet.hpp"

namespace mpt {

// A path's random stream: key and draw counter.
struct rstream {
    uint64_t key;
    uint32_t ctr;
    __device__ __forceinline__ double next() { return draw(key, ++ctr); }
    __device__ __forceinline__ double next(double lo, double hi) { return lo + (hi - lo) * next(); }
};

// texture::value(u, v, p) (texture.h:24-56, 119-127) of a plain texture:
// solid, the 3D checker, or the image texel nearest to (u, v)
// (texture.h:182-195). IMG false: the scene has no image textures, and their
// code is left out.
template <bool IMG = true>
__device__ __forceinline__ v3 tex_plain(const dscene& S, const mpt_texture& t, double u, double v, const v3& p) {
    if (t.type == MPT_TEX_SOLID) return v3{t.even[0], t.even[1], t.even[2]};
    if (!IMG || t.type == MPT_TEX_CHECKER) {
        int x = int(std::floor(t.inv_scale * p.x));
        int y = int(std::floor(t.inv_scale * p.y));
        int z = int(std::floor(t.inv_scale * p.z));
        bool even = (x + y + z) % 2 == 0;
        return even ? v3{t.even[0], t.even[1], t.even[2]} : v3{t.odd[0], t.odd[1], t.odd[2]};
    }
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
    if (t.flags & MPT_TEXF_HDR) {
        const float* c = S.texels_hdr + 3 * k;
        return v3{double(c[0]), double(c[1]), double(c[2])};
    }
    const uchar4 c = *reinterpret_cast<const uchar4*>(S.texels + 4 * k);
    if (t.flags & MPT_TEXF_LINEAR) return v3{double(c.x), double(c.y), double(c.z)} / 255.0;
    return v3{S.srgb[c.x], S.srgb[c.y], S.srgb[c.z]};
}

// A texture's value, through its wrappers (texture.h:212-282): a value
// wrapper (scaled, one channel, a normal's tilt) over uv transforms over a
// plain texture, the chain the host checked.
template <bool IMG = true>
__device__ __forceinline__ v3 tex_value(const dscene& S, int ti, double u, double v, const v3& p) {
    const mpt_texture& t = S.tex[ti];
    if (!IMG || t.type < MPT_TEX_SCALED) return tex_plain<IMG>(S, t, u, v, p);
    const mpt_texture* x = t.type == MPT_TEX_UV_TRANSFORM ? &t : &S.tex[t.inner];
    while (x->type == MPT_TEX_UV_TRANSFORM) {  // scale, rotate, offset in glTF's v-down coordinates
        const double su = x->uv[2] * u, sv = x->uv[3] * (1 - v);
        const double tu = x->uv[4] * su + x->uv[5] * sv + x->uv[0];
        const double tv = -x->uv[5] * su + x->uv[4] * sv + x->uv[1];
        u = tu;
        v = 1 - tv;
        x = &S.tex[x->inner];
    }
    const v3 c = tex_plain<IMG>(S, *x, u, v, p);
    if (t.type == MPT_TEX_SCALED) return v3{t.even[0] * c.x, t.even[1] * c.y, t.even[2] * c.z};
    if (t.type == MPT_TEX_CHANNEL) {
        const double g = t.scale * (t.channel == 0 ? c.x : (t.channel == 1 ? c.y : c.z));
        return v3{g, g, g};
    }
    if (t.type == MPT_TEX_NORMAL_SCALE) return v3{0.5 + t.scale * (c.x - 0.5), 0.5 + t.scale * (c.y - 0.5), c.z};
    return c;
}
template <bool IMG = true>
__device__ __forceinline__ v3 tex_value(const dscene& S, int ti, const hitrec& rec) {
    return tex_value<IMG>(S, ti, rec.u, rec.v, rec.p);
}

// normal_from_map (material.h:282-308): the tangent-space normal map's normal
// in the frame of the outward normal and dp/du, back on the incoming ray's
// side and tilted towards the viewer if it faces away (cosine at least 0.01).
// A hit without a tangent keeps its normal.
__device__ __forceinline__ v3 normal_from_map(const dscene& S, int ti, const v3& rd, const hitrec& rec) {
    constexpr double min_cos_o = 0.01;
    const v3 n = rec.front ? rec.normal : -rec.normal;
    v3 t = rec.dpdu - dot(rec.dpdu, n) * n;
    if (length_squared(t) <= 1e-12 * length_squared(rec.dpdu)) return rec.normal;
    t = unit_vector(t);
    v3 b = cross(n, t);
    if (dot(b, rec.dpdv) < 0) b = -b;  // mirrored (u, v)
    const v3 c = tex_value(S, ti, rec);
    const double nz = std::fmax(2 * c.z - 1, 1e-3);
    v3 m = unit_vector((2 * c.x - 1) * t + (2 * c.y - 1) * b + nz * n);
    if (!rec.front) m = -m;
    const v3 wo = unit_vector(-rd);
    const double cos_o = dot(m, wo);
    if (cos_o < min_cos_o) m = unit_vector(m + (min_cos_o - cos_o) * wo);
    return m;
}

// normal_mapped::scatter's repair (material.h:640-645): a mirror direction on
// the shading normal's side that went below the surface is reflected about
// the hit's own normal gn.
__device__ __forceinline__ v3 unbend(const v3& dir, const v3& shading, const v3& gn) {
    if (dot(dir, shading) > 0 && dot(dir, gn) < 0) return dir - 2 * dot(dir, gn) * gn;
    return dir;
}

// principled's metallic and GGX alpha at a hit: from the red channel of their
// maps, clamped to [0, 1] (material.h:508-515), or the constants.
struct pparams {
    double metallic, alpha;
};
template <bool RICH>
__device__ __forceinline__ pparams principled_at(const dscene& S, const mpt_material& m, const hitrec& rec) {
    pparams q{m.metallic, m.alpha};
    if (!RICH) return q;
    if (m.metallic_texture >= 0) q.metallic = std::fmin(std::fmax(tex_value(S, m.metallic_texture, rec).x, 0.0), 1.0);
    if (m.roughness_texture >= 0)
        q.alpha = roughness_to_alpha(std::fmin(std::fmax(tex_value(S, m.roughness_texture, rec).x, 0.0), 1.0));
    return q;
}

__device__ __forceinline__ double avg3(const v3& c) { return (c.x + c.y + c.z) / 3; }
__device__ __forceinline__ v3 m3(const double* a) { return v3{a[0], a[1], a[2]}; }

// ---------------------------------------------------------------- densities
enum { LOBE_COSINE = 0, LOBE_GGX = 1, LOBE_MIX = 2, LOBE_ROUGH_GLASS = 3, LOBE_HG = 4, LOBE_COAT = 5 };

struct sdensity {
    int kind;
    v3 n;          // frame normal (cosine/GGX/glass), HG propagation direction
    v3 wo_world;   // GGX / rough glass: direction towards the viewer (world)
    double alpha;
    double eta;    // rough glass
    double w0;     // mixture weight of the GGX lobe; HG g
    // LOBE_COAT (principled with a clear coat, material.h:409-419): the coat's
    // GGX lobe (normal n2, alpha2) with weight w1 against the base's mixture
    v3 n2;
    double alpha2;
    double w1;
};

__device__ __forceinline__ double cosine_value(const v3& n, const v3& dir) {  // pdf.h:41-44
    onb f = make_onb(n);
    double c = dot(unit_vector(dir), f.w);
    return std::fmax(0, c / kPi);
}
__device__ __forceinline__ v3 cosine_generate(const v3& n, rstream& rs) {  // pdf.h:46-48, vec3.h:126-136
    onb f = make_onb(n);
    double r1 = rs.next();
    double r2 = rs.next();
    double sp, cp;
    sin_cos_2pi(r1, sp, cp);  // phi = 2 pi r1
    double x = cp * std::sqrt(r2);
    double y = sp * std::sqrt(r2);
    double z = std::sqrt(1 - r2);
    return onb_transform(f, v3{x, y, z});
}
__device__ __forceinline__ double ggx_value(const v3& n, const v3& wo_world, double alpha, const v3& dir) {
    // ggx_reflection_pdf::value (microfacet.h:220-228)
    if (length_squared(dir) == 0) return 0;
    onb f = make_onb(n);
    v3 wo = unit_vector(onb_to_local(f, wo_world));
    v3 wi = unit_vector(onb_to_local(f, dir));
    if (wi.z <= 0 || wo.z <= 0) return 0;
    v3 h = unit_vector(wo + wi);
    return ggx_visible_pdf(alpha, wo, h) / (4 * std::fabs(dot(wo, h)));
}
__device__ __forceinline__ v3 ggx_generate(const v3& n, const v3& wo_world, double alpha, rstream& rs) {
    onb f = make_onb(n);
    v3 wo = unit_vector(onb_to_local(f, wo_world));
    if (wo.z <= 0) return v3{0, 0, 0};
    double u1 = rs.next();
    double u2 = rs.next();
    v3 h = ggx_sample_visible(alpha, wo, u1, u2);
    v3 wi = -wo + 2 * dot(wo, h) * h;
    if (wi.z <= 0) return v3{0, 0, 0};
    return onb_transform(f, wi);
}

// rough_dielectric_bsdf (microfacet.h:384-470)
struct rglass {
    onb f;
    v3 wo;
    double eta, alpha;
};
__device__ __forceinline__ rglass make_rglass(const v3& n, const v3& wo_world, double eta, double alpha) {
    rglass g;
    g.f = make_onb(n);
    g.wo = unit_vector(onb_to_local(g.f, wo_world));
    g.eta = eta;
    g.alpha = alpha;
    return g;
}
__device__ __forceinline__ bool rglass_half(const rglass& g, const v3& wi_world, v3& wi, v3& wm, bool& refl) {
    if (length_squared(wi_world) == 0 || g.wo.z <= 0) return false;
    wi = unit_vector(onb_to_local(g.f, wi_world));
    if (wi.z == 0) return false;
    refl = wi.z > 0;
    wm = refl ? g.wo + wi : wi * g.eta + g.wo;
    if (length_squared(wm) == 0) return false;
    wm = unit_vector(wm);
    if (wm.z < 0) wm = -wm;
    if (dot(wm, wi) * wi.z < 0 || dot(wm, g.wo) * g.wo.z < 0) return false;
    return true;
}
__device__ __forceinline__ double rglass_eval(const rglass& g, const v3& wi_world) {
    v3 wi, wm;
    bool refl;
    if (!rglass_half(g, wi_world, wi, wm, refl)) return 0;
    double F = fresnel_dielectric(dot(g.wo, wm), g.eta);
    if (refl) return F * ggx_D(g.alpha, wm) * ggx_G(g.alpha, g.wo, wi) / (4 * g.wo.z);
    double s = dot(wi, wm) + dot(g.wo, wm) / g.eta;
    double denom = s * s;
    double ft = (1 - F) * ggx_D(g.alpha, wm) * ggx_G(g.alpha, g.wo, wi) *
                std::fabs(dot(wi, wm) * dot(g.wo, wm) / (wi.z * g.wo.z * denom));
    ft /= g.eta * g.eta;
    return ft * std::fabs(wi.z);
}
__device__ __forceinline__ double rglass_pdf(const rglass& g, const v3& wi_world) {
    v3 wi, wm;
    bool refl;
    if (!rglass_half(g, wi_world, wi, wm, refl)) return 0;
    double F = fresnel_dielectric(dot(g.wo, wm), g.eta);
    if (refl) return F * ggx_visible_pdf(g.alpha, g.wo, wm) / (4 * std::fabs(dot(g.wo, wm)));
    double s = dot(wi, wm) + dot(g.wo, wm) / g.eta;
    double denom = s * s;
    return (1 - F) * ggx_visible_pdf(g.alpha, g.wo, wm) * std::fabs(dot(wi, wm)) / denom;
}
__device__ __forceinline__ v3 rglass_sample(const rglass& g, rstream& rs) {
    if (g.wo.z <= 0) return v3{0, 0, 0};
    double u1 = rs.next();
    double u2 = rs.next();
    v3 wm = ggx_sample_visible(g.alpha, g.wo, u1, u2);
    double F = fresnel_dielectric(dot(g.wo, wm), g.eta);
    v3 wi;
    if (rs.next() < F) {
        wi = -g.wo + 2 * dot(g.wo, wm) * wm;
        if (wi.z <= 0) return v3{0, 0, 0};
    } else {
        wi = refract(-g.wo, wm, 1 / g.eta);
        if (wi.z >= 0) return v3{0, 0, 0};
    }
    return onb_transform(g.f, wi);
}

__device__ __forceinline__ double hg_phase(double c, double g) {  // pdf.h:85-88
    double denom = 1 + g * g - 2 * g * c;
    return (1 - g * g) / (4 * kPi * denom * std::sqrt(denom));
}

__device__ __forceinline__ double hg_value(const sdensity& s, const v3& dir) {
    onb f = make_onb(s.n);
    return hg_phase(dot(unit_vector(dir), f.w), s.w0);
}

// The scatter density's value; MT (a material type known at compile time,
// or -1) limits the lobes the code handles to those of that material.
template <int MT = -1>
__device__ __forceinline__ double density_value(const sdensity& s, const v3& dir) {
    if constexpr (MT == MPT_MAT_LAMBERTIAN) return cosine_value(s.n, dir);
    if constexpr (MT == MPT_MAT_METAL) {
        if (s.kind == LOBE_GGX) return ggx_value(s.n, s.wo_world, s.alpha, dir);
        return s.w0 * ggx_value(s.n, s.wo_world, s.alpha, dir) + (1 - s.w0) * cosine_value(s.n, dir);
    }
    if constexpr (MT == MPT_MAT_PRINCIPLED) {
        const double base = s.w0 * ggx_value(s.n, s.wo_world, s.alpha, dir) + (1 - s.w0) * cosine_value(s.n, dir);
        if (s.kind != LOBE_COAT) return base;
        return s.w1 * ggx_value(s.n2, s.wo_world, s.alpha2, dir) + (1 - s.w1) * base;  // pdf.h:120-122
    }
    if constexpr (MT == MPT_MAT_DIELECTRIC) return rglass_pdf(make_rglass(s.n, s.wo_world, s.eta, s.alpha), dir);
    if constexpr (MT == MPT_MAT_PHASE) return hg_value(s, dir);
    if constexpr (MT < 0) {
        switch (s.kind) {
            case LOBE_COSINE: return cosine_value(s.n, dir);
            case LOBE_GGX: return ggx_value(s.n, s.wo_world, s.alpha, dir);
            case LOBE_MIX:
                return s.w0 * ggx_value(s.n, s.wo_world, s.alpha, dir) + (1 - s.w0) * cosine_value(s.n, dir);
            case LOBE_ROUGH_GLASS: return rglass_pdf(make_rglass(s.n, s.wo_world, s.eta, s.alpha), dir);
            case LOBE_COAT:
                return s.w1 * ggx_value(s.n2, s.wo_world, s.alpha2, dir) +
                       (1 - s.w1) * (s.w0 * ggx_value(s.n, s.wo_world, s.alpha, dir) + (1 - s.w0) * cosine_value(s.n, dir));
            default: return hg_value(s, dir);
        }
    }
    return 0;
}

__device__ __forceinline__ v3 hg_generate(const sdensity& s, rstream& rs) {  // pdf.h:67-80
    onb f = make_onb(s.n);
    double g = s.w0;
    double xi = rs.next();
    double c;
    if (std::fabs(g) < 1e-3) {
        c = 1 - 2 * xi;
    } else {
        double sq = (1 - g * g) / (1 - g + 2 * g * xi);
        c = (1 + g * g - sq * sq) / (2 * g);
    }
    c = clampd(c, -1.0, 1.0);
    double st = std::sqrt(1 - c * c);
    double sp, cp;
    sin_cos_2pi(rs.next(), sp, cp);  // phi = 2 pi xi
    return onb_transform(f, v3{cp * st, sp * st, c});
}

template <int MT = -1>
__device__ __forceinline__ v3 density_generate(const sdensity& s, rstream& rs) {
    if constexpr (MT == MPT_MAT_LAMBERTIAN) return cosine_generate(s.n, rs);
    if constexpr (MT == MPT_MAT_METAL) {
        if (s.kind == LOBE_GGX) return ggx_generate(s.n, s.wo_world, s.alpha, rs);
        if (rs.next() < s.w0) return ggx_generate(s.n, s.wo_world, s.alpha, rs);
        return cosine_generate(s.n, rs);
    }
    if constexpr (MT == MPT_MAT_PRINCIPLED) {
        if (s.kind == LOBE_COAT && rs.next() < s.w1) return ggx_generate(s.n2, s.wo_world, s.alpha2, rs);  // pdf.h:124-129
        if (rs.next() < s.w0) return ggx_generate(s.n, s.wo_world, s.alpha, rs);
        return cosine_generate(s.n, rs);
    }
    if constexpr (MT == MPT_MAT_DIELECTRIC) return rglass_sample(make_rglass(s.n, s.wo_world, s.eta, s.alpha), rs);
    if constexpr (MT == MPT_MAT_PHASE) return hg_generate(s, rs);
    if constexpr (MT < 0) {
        switch (s.kind) {
            case LOBE_COSINE: return cosine_generate(s.n, rs);
            case LOBE_GGX: return ggx_generate(s.n, s.wo_world, s.alpha, rs);
            case LOBE_MIX:
                if (rs.next() < s.w0) return ggx_generate(s.n, s.wo_world, s.alpha, rs);
                return cosine_generate(s.n, rs);
            case LOBE_ROUGH_GLASS: return rglass_sample(make_rglass(s.n, s.wo_world, s.eta, s.alpha), rs);
            case LOBE_COAT:
                if (rs.next() < s.w1) return ggx_generate(s.n2, s.wo_world, s.alpha2, rs);
                if (rs.next() < s.w0) return ggx_generate(s.n, s.wo_world, s.alpha, rs);
                return cosine_generate(s.n, rs);
            default: return hg_generate(s, rs);
        }
    }
    return v3{0, 0, 0};
}

// ---------------------------------------------------------------- materials
struct scatter_out {
    bool scattered;
    bool skip;
    v3 attenuation;
    v3 skip_dir;
    sdensity dens;
};

__device__ __forceinline__ v3 principled_f0(const mpt_material& m, double metallic, const v3& base) {
    return (1 - metallic) * v3{m.f0, m.f0, m.f0} + metallic * base;
}

// ---------------------------------------------------------------- measured metals
// fresnel_conductor (microfacet.h:57-66) with the complex arithmetic written
// out: products as (ac - bd, ad + bc); quotients by Smith's method, in the
// branch order of the C runtime's complex division; the square root by the
// half-angle identity on hypot, taking the root of the larger of d +- Re z
// first so that nothing cancels.
struct cpx {
    double re, im;
};
__device__ __forceinline__ cpx cmul(cpx a, cpx b) { return cpx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cpx cdiv(cpx a, cpx b) {
    if (std::fabs(b.re) < std::fabs(b.im)) {
        const double ratio = b.re / b.im;
        const double denom = (b.re * ratio) + b.im;
        return cpx{((a.re * ratio) + a.im) / denom, ((a.im * ratio) - a.re) / denom};
    }
    const double ratio = b.im / b.re;
    const double denom = (b.im * ratio) + b.re;
    return cpx{((a.im * ratio) + a.re) / denom, (a.im - (a.re * ratio)) / denom};
}
__device__ __forceinline__ cpx csqrt_(cpx z) {
    if (z.im == 0) {
        if (z.re < 0) return cpx{0, std::copysign(std::sqrt(-z.re), z.im)};
        return cpx{std::fabs(std::sqrt(z.re)), std::copysign(0.0, z.im)};
    }
    if (z.re == 0) {
        const double r = std::sqrt(0.5 * std::fabs(z.im));
        return cpx{r, std::copysign(r, z.im)};
    }
    const double d = hypot(z.re, z.im);
    double r, s;
    if (z.re > 0) {
        r = std::sqrt(0.5 * (d + z.re));
        s = 0.5 * (z.im / r);
    } else {
        s = std::sqrt(0.5 * (d - z.re));
        r = std::fabs(0.5 * (z.im / s));
    }
    return cpx{r, std::copysign(s, z.im)};
}
__device__ __forceinline__ double conductor_fresnel(double cos_i, double eta, double k) {
    cos_i = std::fmin(std::fmax(cos_i, 0.0), 1.0);
    const cpx n{eta, k};
    const cpx sin2_t = cdiv(cpx{1 - cos_i * cos_i, 0}, cmul(n, n));
    const cpx cos_t = csqrt_(cpx{1.0 - sin2_t.re, 0.0 - sin2_t.im});
    const cpx nc{n.re * cos_i, n.im * cos_i};
    const cpx r_parl = cdiv(cpx{nc.re - cos_t.re, nc.im - cos_t.im}, cpx{nc.re + cos_t.re, nc.im + cos_t.im});
    const cpx nt = cmul(n, cos_t);
    const cpx r_perp = cdiv(cpx{cos_i - nt.re, 0.0 - nt.im}, cpx{cos_i + nt.re, 0.0 + nt.im});
    return ((r_parl.re * r_parl.re + r_parl.im * r_parl.im) + (r_perp.re * r_perp.re + r_perp.im * r_perp.im)) / 2;
}
__device__ __forceinline__ v3 conductor_at(const mpt_material& m, double c) {
    return v3{conductor_fresnel(c, m.eta[0], m.k[0]), conductor_fresnel(c, m.eta[1], m.k[1]),
              conductor_fresnel(c, m.eta[2], m.k[2])};
}
// metal_fresnel::at (microfacet.h:123-129): Schlick through albedo, or measured
template <bool RICH>
__device__ __forceinline__ v3 metal_fresnel(const mpt_material& m, double c) {
    if (RICH && m.measured) return conductor_at(m, c);
    return fresnel_schlick3(c, m3(m.albedo));
}
// principled::specular_fresnel and ms_fresnel (material.h:537-551)
template <bool RICH>
__device__ __forceinline__ v3 principled_fresnel(const mpt_material& m, double c, const v3& base, double metallic) {
    if (RICH && m.measured) {
        const double fd = fresnel_schlick(c, m.f0);
        return (1 - metallic) * v3{fd, fd, fd} + metallic * conductor_at(m, c);
    }
    return fresnel_schlick3(c, principled_f0(m, metallic, base));
}
template <bool RICH>
__device__ __forceinline__ v3 principled_ms_fresnel(const dscene& S, const mpt_material& m, const v3& base,
                                                    double metallic, double alpha) {
    if (RICH && m.measured) {
        const double fd = m.f0 + (1 - m.f0) / 21;
        return ggx_ms_fresnel_avg(S.ggx_eavg, (1 - metallic) * v3{fd, fd, fd} + metallic * m3(m.f_avg), alpha);
    }
    return ggx_ms_fresnel(S.ggx_eavg, principled_f0(m, metallic, base), alpha);
}

// material::scatter (material.h:79-85, 123-150, 210-242, 387-421; medium.h:21-27).
// principled's clear coat (material.h:460-527): weight and GGX alpha at the
// hit from the red channels of their textures, clamped to [0, 1]; and what a
// coat of weight w lets through to the base and back for cosines cos_o and
// cos_i to its normal (coat_transmittance): Beer-Lambert through the tinted
// layer along the refracted directions into ior 1.5, blended by w.
constexpr double kCoatF0 = 0.04;
__device__ __forceinline__ double coat_weight(const dscene& S, const mpt_material& m, const hitrec& rec) {
    return std::fmin(std::fmax(tex_value(S, m.coat_weight_texture, rec).x, 0.0), 1.0);
}
__device__ __forceinline__ double coat_alpha(const dscene& S, const mpt_material& m, const hitrec& rec) {
    return roughness_to_alpha(std::fmin(std::fmax(tex_value(S, m.coat_roughness_texture, rec).x, 0.0), 1.0));
}
__device__ __forceinline__ v3 coat_transmittance(const mpt_material& m, double w, double cos_o, double cos_i) {
    if (m.coat_tint[0] == 1 && m.coat_tint[1] == 1 && m.coat_tint[2] == 1) return v3{1, 1, 1};
    auto cos_t = [](double c) {
        c = std::fmin(std::fmax(c, 0.0), 1.0);
        return std::sqrt(1 - (1 - c * c) / (1.5 * 1.5));
    };
    const double e = (1 / cos_t(cos_o) + 1 / cos_t(cos_i)) / 2;
    const v3 t{std::pow(m.coat_tint[0], e), std::pow(m.coat_tint[1], e), std::pow(m.coat_tint[2], e)};
    return (1 - w) * v3{1, 1, 1} + w * t;
}

// MT: the material type when known at compile time (-1: dispatch on m.type).
// RICH: the scene has image textures, texture or normal maps (hits carry
// (u, v)), thin glass or measured metals; without it their code is left out.
// gn: the hit's own normal when rec carries a normal map's (null: rec.normal),
// which a clear coat without a map of its own keeps.
template <int MT = -1, bool RICH = false>
__device__ __forceinline__ scatter_out mat_scatter(const dscene& S, const mpt_material& m, const v3& rd,
                                                   const hitrec& rec, rstream& rs, const v3* gn = nullptr) {
    scatter_out o;
    o.scattered = true;
    o.skip = false;
    o.attenuation = v3{1, 1, 1};
    o.skip_dir = v3{0, 0, 0};
    o.dens.kind = LOBE_COSINE;
    o.dens.n = rec.normal;
    o.dens.wo_world = v3{0, 0, 0};
    o.dens.alpha = 0;
    o.dens.eta = 1;
    o.dens.w0 = 1;
    switch (MT >= 0 ? MT : m.type) {
        case MPT_MAT_LAMBERTIAN:
            o.attenuation = tex_value<RICH>(S, m.texture, rec);
            break;
        case MPT_MAT_METAL: {
            v3 albedo = m3(m.albedo);
            if (m.roughness == 0) {  // a Schlick mirror reflects albedo at every angle
                o.attenuation = RICH && m.measured ? conductor_at(m, dot(rec.normal, unit_vector(-rd))) : albedo;
                o.skip = true;
                o.skip_dir = reflect(unit_vector(rd), rec.normal);
                break;
            }
            o.attenuation = albedo;
            o.dens.kind = LOBE_GGX;
            o.dens.wo_world = -rd;
            o.dens.alpha = m.alpha;
            double cos_o = std::fmax(0.0, dot(rec.normal, unit_vector(-rd)));
            double ms_w = ggx_ms_albedo(S.ggx_e, S.ggx_eavg, cos_o, m.alpha) *
                          avg3(ggx_ms_fresnel_avg(S.ggx_eavg, m3(m.f_avg), m.alpha));
            double ss_w = ggx_E(S.ggx_e, cos_o, m.alpha) * avg3(metal_fresnel<RICH>(m, cos_o));
            if (ms_w <= 0) break;
            o.dens.kind = LOBE_MIX;
            o.dens.w0 = clampd(ss_w / (ss_w + ms_w), 0.1, 1.0);
            break;
        }
        case MPT_MAT_DIELECTRIC: {
            if (RICH && m.thin) {  // thin_dielectric::scatter (material.h:687-702)
                v3 ud = unit_vector(rd);
                double cos_theta = std::fmin(dot(-ud, rec.normal), 1.0);
                double r = fresnel_dielectric(cos_theta, m.ior);
                double reflectance = 2 * r / (1 + r);
                o.skip = true;
                if (rs.next() < reflectance) {
                    o.skip_dir = reflect(ud, rec.normal);
                } else {
                    o.attenuation = m3(m.albedo);
                    o.skip_dir = ud;
                }
                break;
            }
            if (m.roughness > 0) {
                o.dens.kind = LOBE_ROUGH_GLASS;
                o.dens.wo_world = -rd;
                o.dens.eta = rec.front ? m.ior : 1.0 / m.ior;
                o.dens.alpha = m.alpha;
                break;
            }
            o.skip = true;
            double ri = rec.front ? (1.0 / m.ior) : m.ior;
            v3 ud = unit_vector(rd);
            double cos_theta = std::fmin(dot(-ud, rec.normal), 1.0);
            double sin_theta = std::sqrt(1.0 - cos_theta * cos_theta);
            bool cannot_refract = ri * sin_theta > 1.0;
            if (cannot_refract || fresnel_dielectric(cos_theta, 1.0 / ri) > rs.next())
                o.skip_dir = reflect(ud, rec.normal);
            else
                o.skip_dir = refract(ud, rec.normal, ri);
            break;
        }
        case MPT_MAT_PRINCIPLED: {
            v3 wo = unit_vector(-rd);
            const pparams pp = principled_at<RICH>(S, m, rec);
            v3 base = tex_value<RICH>(S, m.texture, rec);
            double metallic = pp.metallic;
            double alpha = pp.alpha;
            double cos_o = std::fmax(0.0, dot(rec.normal, wo));
            double spec_w = avg3(principled_fresnel<RICH>(m, cos_o, base, metallic));
            double diff_w = (1 - metallic) * avg3(base) * (1 - fresnel_schlick(cos_o, m.f0));
            diff_w += ggx_ms_albedo(S.ggx_e, S.ggx_eavg, cos_o, alpha) *
                      avg3(principled_ms_fresnel<RICH>(S, m, base, metallic, alpha));
            double p_spec = spec_w + diff_w > 0 ? spec_w / (spec_w + diff_w) : 1;
            o.dens.kind = LOBE_MIX;
            o.dens.w0 = clampd(p_spec, 0.1, 1.0);
            o.dens.wo_world = wo;
            o.dens.alpha = alpha;
            o.attenuation = base;
            if constexpr (RICH) {
                if (m.coat) {  // material.h:409-419
                    const double w = coat_weight(S, m, rec);
                    if (w > 0) {
                        hitrec geo = rec;
                        if (gn) geo.normal = *gn;
                        const v3 n_coat = m.coat_normal_map >= 0 ? normal_from_map(S, m.coat_normal_map, rd, geo)
                                                                 : geo.normal;
                        const double cw = w * fresnel_schlick(dot(n_coat, wo), kCoatF0);
                        const double below = (1 - cw) * (spec_w + diff_w) * avg3(coat_transmittance(m, w, 1, 1));
                        o.dens.kind = LOBE_COAT;
                        o.dens.n2 = n_coat;
                        o.dens.alpha2 = coat_alpha(S, m, rec);
                        o.dens.w1 = clampd(cw / (cw + below), 0.1, 0.9);
                    }
                }
            }
            break;
        }
        case MPT_MAT_PHASE:
            o.attenuation = m3(m.albedo);
            o.dens.kind = LOBE_HG;
            o.dens.n = rd;
            o.dens.w0 = m.g;
            break;
        default:
            o.scattered = false;
            break;
    }
    return o;
}

// material::eval: f * |cos| for a non-specular direction. rec.normal is the
// shading normal (a normal map's, when the material has one) and gn the
// hit's own: no light arrives from below gn (material.h:426, 651).
template <int MT = -1, bool RICH = false>
__device__ __forceinline__ v3 mat_eval(const dscene& S, const mpt_material& m, const v3& rd, const hitrec& rec,
                                       const v3& gn, const v3& dir) {
    if (RICH && m.normal_map >= 0 && dot(dir, gn) <= 0) return v3{0, 0, 0};
    switch (MT >= 0 ? MT : m.type) {
        case MPT_MAT_LAMBERTIAN: {  // material.h:88-97
            double c = dot(rec.normal, unit_vector(dir));
            double pdf = c < 0 ? 0 : c / kPi;
            return tex_value<RICH>(S, m.texture, rec) * pdf;
        }
        case MPT_MAT_METAL: {  // material.h:154-164
            onb f = make_onb(rec.normal);
            v3 wo = unit_vector(onb_to_local(f, -rd));
            v3 wi = unit_vector(onb_to_local(f, dir));
            if (wo.z <= 0 || wi.z <= 0) return v3{0, 0, 0};
            v3 h = unit_vector(wo + wi);
            v3 spec = metal_fresnel<RICH>(m, dot(wi, h)) * (ggx_D(m.alpha, h) * ggx_G(m.alpha, wo, wi) / (4 * wo.z));
            return spec + ggx_ms_fresnel_avg(S.ggx_eavg, m3(m.f_avg), m.alpha) *
                              ggx_ms_eval(S.ggx_e, S.ggx_eavg, wo.z, wi.z, m.alpha);
        }
        case MPT_MAT_DIELECTRIC: {  // material.h:244-249
            if (m.roughness == 0) return v3{0, 0, 0};
            double eta = rec.front ? m.ior : 1.0 / m.ior;
            double f = rglass_eval(make_rglass(rec.normal, -rd, eta, m.alpha), dir);
            return v3{f, f, f};
        }
        case MPT_MAT_PRINCIPLED: {  // material.h:423-456, 486-506
            if (dot(dir, gn) <= 0) return v3{0, 0, 0};
            v3 wo_w = unit_vector(-rd);
            v3 wi_w = unit_vector(dir);
            onb f = make_onb(rec.normal);
            v3 wo = unit_vector(onb_to_local(f, wo_w));
            v3 wi = unit_vector(onb_to_local(f, wi_w));
            v3 below{0, 0, 0};
            if (wo.z > 0 && wi.z > 0) {
                const pparams pp = principled_at<RICH>(S, m, rec);
                double metallic = pp.metallic, alpha = pp.alpha;
                v3 base = tex_value<RICH>(S, m.texture, rec);
                v3 h = unit_vector(wo + wi);
                v3 specular = principled_fresnel<RICH>(m, dot(wi, h), base, metallic) *
                                  (ggx_D(alpha, h) * ggx_G(alpha, wo, wi) / (4 * wo.z)) +
                              principled_ms_fresnel<RICH>(S, m, base, metallic, alpha) *
                                  ggx_ms_eval(S.ggx_e, S.ggx_eavg, wo.z, wi.z, alpha);
                double transmitted = (1 - fresnel_schlick(wi.z, m.f0)) * (1 - fresnel_schlick(wo.z, m.f0));
                v3 diffuse = ((1 - metallic) * transmitted * wi.z / kPi) * base;
                below = specular + diffuse;
            }
            if constexpr (RICH) {
                if (m.coat) {
                    const double w = coat_weight(S, m, rec);
                    if (!(w <= 0)) {  // the coat: a colorless GGX reflection, and what it lets through twice
                        hitrec geo = rec;
                        geo.normal = gn;
                        const v3 n_coat = m.coat_normal_map >= 0 ? normal_from_map(S, m.coat_normal_map, rd, geo) : gn;
                        const double cos_o = dot(n_coat, wo_w), cos_i = dot(n_coat, wi_w);
                        const double through = (1 - w * fresnel_schlick(cos_o, kCoatF0)) *
                                               (1 - w * fresnel_schlick(cos_i, kCoatF0));
                        v3 result = (through * coat_transmittance(m, w, cos_o, cos_i)) * below;
                        if (cos_o > 0 && cos_i > 0) {
                            const onb cf = make_onb(n_coat);
                            const v3 lo = unit_vector(onb_to_local(cf, wo_w)), li = unit_vector(onb_to_local(cf, wi_w));
                            const v3 h = unit_vector(lo + li);
                            const double ca = coat_alpha(S, m, rec);
                            double coat = fresnel_schlick(dot(li, h), kCoatF0) * ggx_D(ca, h) * ggx_G(ca, lo, li) /
                                          (4 * lo.z);
                            coat += ggx_ms_fresnel(S.ggx_eavg, v3{kCoatF0, kCoatF0, kCoatF0}, ca).x *
                                    ggx_ms_eval(S.ggx_e, S.ggx_eavg, lo.z, li.z, ca);
                            result = result + v3{w * coat, w * coat, w * coat};
                        }
                        return result;
                    }
                }
            }
            return below;
        }
        case MPT_MAT_PHASE: {  // medium.h:33-36
            double c = dot(unit_vector(rd), unit_vector(dir));
            return m3(m.albedo) * hg_phase(c, m.g);
        }
        default:
            return v3{0, 0, 0};
    }
}

template <int MT = -1>
__device__ __forceinline__ v3 mat_emitted(const mpt_material& m, const v3& rd, const hitrec& rec) {
    if ((MT >= 0 ? MT : m.type) != MPT_MAT_DIFFUSE_LIGHT || !rec.front) return v3{0, 0, 0};  // material.h:578-584
    v3 e = m3(m.emit);
    if (m.spread >= 180) return e;
    double c = -dot(unit_vector(rd), rec.normal);
    double gf;
    if (c <= 0) gf = 0;
    else {
        double tan_theta = std::sqrt(std::fmax(0.0, 1 - c * c)) / c;
        gf = std::fmax(0.0, 1 - tan_theta / m.tan_half);
    }
    return e * gf;
}

template <int MT = -1>
__device__ __forceinline__ v3 mat_interior(const mpt_material& m, const v3& rd, const hitrec& rec) {
    if ((MT >= 0 ? MT : m.type) != MPT_MAT_DIELECTRIC) return v3{1, 1, 1};  // material.h:201-206
    if (rec.front || (m.sigma_a[0] == 0 && m.sigma_a[1] == 0 && m.sigma_a[2] == 0)) return v3{1, 1, 1};
    double dd = rec.t * length(rd);
    return v3{std::exp(-m.sigma_a[0] * dd), std::exp(-m.sigma_a[1] * dd), std::exp(-m.sigma_a[2] * dd)};
}

__host__ __device__ inline bool mat_lets_shadows(const mpt_material& m) {
    return m.type == MPT_MAT_DIELECTRIC && m.roughness == 0 && !m.thin;  // thin_dielectric keeps the default
}

// ---------------------------------------------------------------- environment
__device__ __forceinline__ void env_angles(const dscene& S, const v3& dir, double& u, double& theta) {
    v3 d = unit_vector(dir);
    theta = std::acos(clampd(d.y, -1.0, 1.0));
    u = 0.5 + std::atan2(d.x, -d.z) / (2 * kPi) + S.env_rotation;
    u -= std::floor(u);
}

__device__ v3 env_radiance(const dscene& S, const v3& dir) {  // environment.h:66-79
    double u, theta;
    env_angles(S, dir, u, theta);
    const int w = S.env_w, h = S.env_h;
    double x = u * w - 0.5;
    double y = clampd(theta / kPi * h - 0.5, 0.0, h - 1.0);
    double x0 = std::floor(x), y0 = std::floor(y);
    double fx = x - x0, fy = y - y0;
    int i0 = (int(x0) + w) % w, i1 = (i0 + 1) % w;
    int j0 = int(y0), j1 = j0 + 1 < h - 1 ? j0 + 1 : h - 1;
    auto px = [&](int i, int j) { return m3(&S.env_rgb[(size_t(j) * w + i) * 3]); };
    v3 top = (1 - fx) * px(i0, j0) + fx * px(i1, j0);
    v3 bottom = (1 - fx) * px(i0, j1) + fx * px(i1, j1);
    return S.env_intensity * ((1 - fy) * top + fy * bottom);
}

__device__ double env_pdf(const dscene& S, const v3& dir) {  // environment.h:84-92
    double u, theta;
    env_angles(S, dir, u, theta);
    double sin_theta = std::sin(theta);
    const int w = S.env_w, h = S.env_h;
    int i = int(u * w);
    i = i < 0 ? 0 : (i > w - 1 ? w - 1 : i);
    int j = int(theta / kPi * h);
    j = j < 0 ? 0 : (j > h - 1 ? h - 1 : j);
    if (sin_theta <= 0) return 0;
    double row = S.env_roww[j] / S.env_total;
    const double* cdf = &S.env_cond[size_t(j) * (w + 1)];
    double column = cdf[i + 1] - cdf[i];
    double pix = row * column * w * h;
    return pix / (2 * kPi * kPi * sin_theta);
}

__device__ __forceinline__ double sample_cdf(const double* cdf, int n, double xi) {  // environment.h:180-186
    // upper_bound over cdf[0..n]
    int lo = 0, hi = n + 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (xi < cdf[mid]) hi = mid;
        else lo = mid + 1;
    }
    int k = lo - 1;
    k = k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
    double span = cdf[k + 1] - cdf[k];
    double fraction = span > 0 ? (xi - cdf[k]) / span : 0.5;
    return k + clampd(fraction, 0.0, 1.0 - 1e-9);
}

__device__ v3 env_random(const dscene& S, rstream& rs) {  // environment.h:94-107
    const int w = S.env_w, h = S.env_h;
    double y = sample_cdf(S.env_marg, h, rs.next());
    int j = int(y) < h - 1 ? int(y) : h - 1;
    double x = sample_cdf(&S.env_cond[size_t(j) * (w + 1)], w, rs.next());
    double theta = kPi * y / h;
    double phi = 2 * kPi * (x / w - 0.5 - S.env_rotation);
    double st, ct, sp, cp;
    sin_cos(theta, st, ct);
    sin_cos(phi, sp, cp);
    return v3{st * sp, ct, -st * cp};
}

// ---------------------------------------------------------------- sun
__device__ __forceinline__ bool sun_in_cone(const dscene& S, const v3& dir) {  // sun.h:76-78
    return dot(unit_vector(dir), S.sun_frame.w) >= S.sun_cos_max;
}
__device__ __forceinline__ v3 sun_radiance(const dscene& S, const v3& dir) {  // sun.h:46-48
    return sun_in_cone(S, dir) ? S.sun_radiance : v3{0, 0, 0};
}
__device__ __forceinline__ double sun_pdf(const dscene& S, const v3& dir) {  // sun.h:50-53
    return sun_in_cone(S, dir) ? 1 / S.sun_solid_angle : 0;
}
__device__ __forceinline__ v3 sun_random(const dscene& S, rstream& rs) {  // sun.h:56-62: cos(theta) uniform in the cone
    double cos_theta = 1 - rs.next() * (1 - S.sun_cos_max);
    double sin_theta = std::sqrt(std::fmax(0.0, 1 - cos_theta * cos_theta));
    double phi = 2 * kPi * rs.next();
    double sp, cp;
    sin_cos(phi, sp, cp);
    return onb_transform(S.sun_frame, v3{cp * sin_theta, sp * sin_theta, cos_theta});
}

// ---------------------------------------------------------------- lights
// Light kinds a scene's light list holds; the NEE code is compiled for the
// kinds present (LK bit set), like the traversal for geometry.
enum : int {
    L_ENV = 1,
    L_SPHERE = 2,
    L_QUAD = 4,
    L_TRI = 8,
    L_LIST = 16,
    L_XFORM = 32,
    L_MESH = 64,
    L_ALL = 127,
};

template <int LK>
__device__ __forceinline__ double prim_light_pdf(const dscene& S, int64_t p, const v3& o, const v3& d) {
    const double* q = prim_ptr(S, p);
    int k = S.pkind[p];
    double t;
    constexpr bool only_quads = (LK & (L_SPHERE | L_QUAD | L_TRI)) == L_QUAD;
    if ((LK & L_QUAD) && (only_quads || k == MPT_PRIM_QUAD)) {  // quad.h:65-74
        if (!quad_t(q, o, d, 0.001, kInf, t)) return 0;
        double dist2 = t * t * length_squared(d);
        v3 nrm{q[9], q[10], q[11]};
        v3 n = dot(d, nrm) < 0 ? nrm : -nrm;
        double cosine = std::fabs(dot(d, n) / length(d));
        double area = length(cross(v3{q[3], q[4], q[5]}, v3{q[6], q[7], q[8]}));
        return dist2 / (cosine * area);
    }
    if ((LK & L_SPHERE) && k == MPT_PRIM_SPHERE) {  // sphere.h:56-73
        if (!sphere_t(q, o, d, 0.001, kInf, t)) return 0;
        v3 c{q[0], q[1], q[2]};
        double r = std::fabs(q[3]);
        double dist2 = length_squared(c - o);
        if (dist2 <= r * r) {
            double hd2 = t * t * length_squared(d);
            v3 p = o + t * d;
            double cosine = std::fabs(dot(unit_vector(d), (p - c) / r));
            return hd2 / (cosine * 4 * kPi * r * r);
        }
        double cos_theta_max = std::sqrt(1 - r * r / dist2);
        double solid_angle = 2 * kPi * (1 - cos_theta_max);
        return 1 / solid_angle;
    }
    if constexpr (!(LK & L_TRI)) return 0;
    // triangle (mesh.h:115-124)
    if (!tri_t(q, o, d, 0.001, kInf, t)) return 0;
    double dist2 = t * t * length_squared(d);
    v3 g{q[9], q[10], q[11]};
    double cosine = std::fabs(dot(d, g) / length(d));
    return dist2 / (cosine * q[12]);
}

__device__ __forceinline__ v3 random_unit_vector(rstream& rs) {  // vec3.h:116-123
    while (true) {
        double x = rs.next(-1, 1);
        double y = rs.next(-1, 1);
        double z = rs.next(-1, 1);
        v3 p{x, y, z};
        double lensq = length_squared(p);
        if (1e-160 < lensq && lensq <= 1) return p / std::sqrt(lensq);
    }
}

template <int LK>
__device__ __forceinline__ v3 prim_light_random(const dscene& S, int64_t p, const v3& o, rstream& rs) {
    const double* q = prim_ptr(S, p);
    int k = S.pkind[p];
    constexpr bool only_quads = (LK & (L_SPHERE | L_QUAD | L_TRI)) == L_QUAD;
    if ((LK & L_QUAD) && (only_quads || k == MPT_PRIM_QUAD)) {  // quad.h:76-79
        double a = rs.next();
        double b = rs.next();
        v3 pt = v3{q[0], q[1], q[2]} + (a * v3{q[3], q[4], q[5]}) + (b * v3{q[6], q[7], q[8]});
        return pt - o;
    }
    if ((LK & L_SPHERE) && k == MPT_PRIM_SPHERE) {  // sphere.h:75-83, 122-132
        v3 c{q[0], q[1], q[2]};
        double r = std::fabs(q[3]);
        v3 dir = c - o;
        double dist2 = length_squared(dir);
        if (dist2 <= r * r) return c + r * random_unit_vector(rs) - o;
        onb f = make_onb(dir);
        double r1 = rs.next();
        double r2 = rs.next();
        double z = 1 + r2 * (std::sqrt(1 - r * r / dist2) - 1);
        double sp, cp;
        sin_cos_2pi(r1, sp, cp);  // phi = 2 pi r1
        double x = cp * std::sqrt(1 - z * z);
        double y = sp * std::sqrt(1 - z * z);
        return onb_transform(f, v3{x, y, z});
    }
    if constexpr (!(LK & L_TRI)) return v3{0, 0, 0};
    // triangle (mesh.h:127-139)
    double r1 = rs.next(), r2 = rs.next();
    if (r1 + r2 > 1) {
        r1 = 1 - r1;
        r2 = 1 - r2;
    }
    v3 p0{q[0], q[1], q[2]};
    v3 pt = p0 + r1 * v3{q[3], q[4], q[5]} + r2 * v3{q[6], q[7], q[8]};
    return pt - o;
}

// hittable::pdf_value of light object ob (instance.h:47-50, hittable_list.h:46-56).
template <int LK>
__device__ __forceinline__ double object_light_pdf(const dscene& S, int ob, v3 o, v3 d) {
    const obj_rec r = S.obj[ob];
    if constexpr (LK & L_XFORM) {
        if (r.xform >= 0) {
            const xform& x = S.xf[r.xform];
            v3 lo = xf_inv_point(x, o);
            v3 ld = xf_inv_vector(x, d);
            o = lo;
            d = ld;
        }
    }
    if (!(LK & (L_LIST | L_MESH)) || r.kind == MPT_OBJ_PRIM) return prim_light_pdf<LK>(S, r.first, o, d);
    if ((LK & L_MESH) && r.kind == MPT_OBJ_MESH) {  // mesh::pdf_value (mesh.h:197-206): area-weighted over its triangles
        const int64_t* tri = S.mlight_prim + r.light_table;
        const double total = S.mlight_cdf[r.light_table + r.count - 1];
        if (total <= 0) return 0;
        double sum = 0;
        for (int64_t k = 0; k < r.count; k++)
            sum += (prim_ptr(S, tri[k])[12] / total) * prim_light_pdf<LK>(S, tri[k], o, d);
        return sum;
    }
    double weight = 1.0 / double(r.count);
    double sum = 0.0;
    for (int64_t k = 0; k < r.count; k++) sum += weight * prim_light_pdf<LK>(S, r.first + k, o, d);
    return sum;
}

template <int LK>
__device__ __forceinline__ v3 object_light_random(const dscene& S, int ob, const v3& o, rstream& rs) {
    const obj_rec r = S.obj[ob];
    v3 lo = o;
    const xform* x = nullptr;
    if constexpr (LK & L_XFORM) {
        x = r.xform >= 0 ? &S.xf[r.xform] : nullptr;
        if (x) lo = xf_inv_point(*x, o);
    }
    v3 res;
    if (!(LK & (L_LIST | L_MESH)) || r.kind == MPT_OBJ_PRIM) {
        res = prim_light_random<LK>(S, r.first, lo, rs);
    } else if ((LK & L_MESH) && r.kind == MPT_OBJ_MESH) {  // mesh::random (mesh.h:208-213)
        const double* cdf = S.mlight_cdf + r.light_table;
        const int64_t n = r.count;
        const double x = rs.next() * cdf[n - 1];
        int64_t lo_i = 0, hi_i = n;  // upper_bound: the first running sum above x
        while (lo_i < hi_i) {
            const int64_t mid = (lo_i + hi_i) >> 1;
            if (x < cdf[mid]) hi_i = mid;
            else lo_i = mid + 1;
        }
        res = prim_light_random<LK>(S, S.mlight_prim[r.light_table + (lo_i < n - 1 ? lo_i : n - 1)], lo, rs);
    } else {
        int n = int(r.count);
        int idx = int(rs.next() * n);
        idx = idx < n - 1 ? idx : n - 1;
        res = prim_light_random<LK>(S, r.first + idx, lo, rs);
    }
    if constexpr (LK & L_XFORM)
        if (x) res = xf_vector(*x, res);
    return res;
}

// light_list::pdf_value (light_list.h:106-113): sum of p_k pdf_k.
template <int LK = L_ALL>
__device__ __forceinline__ double lights_pdf(const dscene& S, const v3& o, const v3& d) {
    if (S.nlights == 0) return 0.0;
    double sum = 0;
    for (int k = 0; k < S.nlights; k++) {
        const light_rec L = S.lights[k];
        double pk;
        if constexpr (LK == L_ENV) pk = env_pdf(S, d);
        else if constexpr (!(LK & L_ENV)) pk = object_light_pdf<LK>(S, L.object, o, d);
        else pk = L.object == -2 ? sun_pdf(S, d) : (L.object < 0 ? env_pdf(S, d) : object_light_pdf<LK>(S, L.object, o, d));
        sum += L.probability * pk;
    }
    return sum;
}

// light_list::random (light_list.h:115-127), or hittable_list::random for an
// unbuilt list (strategy 2).
template <int LK = L_ALL>
__device__ __forceinline__ v3 lights_random(const dscene& S, const v3& o, rstream& rs) {
    double xi = rs.next();
    int n = S.nlights;
    int k;
    if (S.lstrategy == 2) {
        k = int(xi * n);
        k = k < n - 1 ? k : n - 1;
    } else {
        int lo = 0, hi = n + 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (xi < S.lcdf[mid]) hi = mid;
            else lo = mid + 1;
        }
        k = lo;
        k = k < 1 ? 1 : (k > n ? n : k);
        k -= 1;
    }
    const light_rec L = S.lights[k];
    if constexpr (LK == L_ENV) return env_random(S, rs);
    else if constexpr (!(LK & L_ENV)) return object_light_random<LK>(S, L.object, o, rs);
    else {
        if (L.object == -2) return sun_random(S, rs);
        if (L.object < 0) return env_random(S, rs);
        return object_light_random<LK>(S, L.object, o, rs);
    }
}

}  // namespace mpt
