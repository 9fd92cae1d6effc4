// ref_harness.cpp - a C entry layer over the reference renderer itself,
// compiled from the reference's own headers where they lie (see
// oracle/Makefile, target `ref`) into oracle/_ref/libmpt_ref.so.
//
// TEST INFRASTRUCTURE. It is built only where the reference sources are
// (their headers are compiled in place); the built library then travels
// with the tests. It writes the golden vectors under tests/golden, pins the
// CPU restatement (oracle.cpp), is compared with the GPU in tests where it
// is present, and is the CPU baseline bench.py times (the timing builds,
// Makefile target ref_timing). The product library never loads it.
// Nothing from the reference is copied here; the harness only calls it.
//
// The reference's hit_record has no primitive ID (hittable.h:12-32), so a
// hit is identified afterwards: every primitive of the world, in canonical
// order (world list order, lists expanded in place, box() sides in their
// order, instances and meshes descended), is tested alone with the same ray
// and interval; the ones whose t is bitwise equal to the BVH's are the
// candidates. One candidate: its ID. Several: a tie, reported separately.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

// The reference keeps geometry members private; the harness reads them.
#define private public
#include "scene.h"
#undef private
#include "image_io.h"

namespace {

struct canon_prim {
    const hittable* prim;
    const instance* inst;  // nullptr when not instanced
    bool medium;           // boundary of a medium: never a ray-query answer
};

struct ref_scene {
    scene sc;
    std::shared_ptr<bvh> tree;
    std::shared_ptr<hittable_list> everything;  // the BVH and the unbounded objects (main.cpp:70-74)
    std::vector<canon_prim> canon;
    const hittable& world() const {
        if (sc.unbounded.empty()) return *tree;
        return *everything;
    }
};

void expand(const hittable* h, const instance* inst, bool medium, std::vector<canon_prim>& out) {
    if (auto l = dynamic_cast<const hittable_list*>(h)) {
        for (const auto& o : l->objects) expand(o.get(), inst, medium, out);
    } else if (auto in = dynamic_cast<const instance*>(h)) {
        expand(in->object.get(), in, medium, out);
    } else if (auto m = dynamic_cast<const mesh*>(h)) {
        for (const auto& t : m->triangles) out.push_back({t.get(), inst, medium});
    } else if (auto cm = dynamic_cast<const constant_medium*>(h)) {
        expand(cm->boundary.get(), inst, true, out);
    } else {
        out.push_back({h, inst, medium});
    }
}

}  // namespace

extern "C" {

int ref_load(const char* path, void** out) {
    auto r = new ref_scene;
    std::clog.setstate(std::ios::failbit);
    if (!load_scene(path, r->sc)) {
        delete r;
        return -1;
    }
    r->tree = std::make_shared<bvh>(r->sc.world);  // main.cpp:70
    r->everything = std::make_shared<hittable_list>(r->tree);
    for (const auto& o : r->sc.unbounded.objects) r->everything->add(o);
    for (const auto& o : r->sc.world.objects) expand(o.get(), nullptr, false, r->canon);
    for (const auto& o : r->sc.unbounded.objects) expand(o.get(), nullptr, false, r->canon);  // IDs after the world's
    *out = r;
    return 0;
}

void ref_free(void* h) { delete static_cast<ref_scene*>(h); }

int64_t ref_canon_count(void* h) { return int64_t(static_cast<ref_scene*>(h)->canon.size()); }

// Closest hit of n rays through the reference BVH. id: canonical ID or -1
// (miss); ties: number of canonical primitives with the same t (0 on a
// miss, 1 for a unique hit); tie_min/tie_max: smallest/largest candidate ID.
int ref_trace(void* h, const double* rays, const double* tt, int64_t n, int64_t* id, double* t, uint8_t* front,
              int32_t* ties, int64_t* tie_min, int64_t* tie_max, int threads) {
    auto* r = static_cast<ref_scene*>(h);
    auto work = [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; i++) {
            ray rr(point3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]), vec3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]));
            interval iv(tt[2 * i], tt[2 * i + 1]);
            hit_record rec;
            if (!r->world().hit(rr, iv, rec)) {
                id[i] = -1, t[i] = 0, front[i] = 0, ties[i] = 0, tie_min[i] = -1, tie_max[i] = -1;
                continue;
            }
            t[i] = rec.t;
            front[i] = rec.front_face;
            int32_t nc = 0;
            int64_t lo = -1, hi = -1;
            for (size_t k = 0; k < r->canon.size(); k++) {
                const canon_prim& c = r->canon[k];
                if (c.medium) continue;
                ray local = rr;
                if (c.inst)
                    local = ray(c.inst->xform.apply_inverse_point(rr.origin()), c.inst->xform.apply_inverse_vector(rr.direction()));
                hit_record one;
                if (c.prim->hit(local, iv, one) && std::memcmp(&one.t, &rec.t, sizeof(double)) == 0) {
                    if (nc == 0) lo = int64_t(k);
                    hi = int64_t(k);
                    nc++;
                }
            }
            ties[i] = nc;
            tie_min[i] = lo;
            tie_max[i] = hi;
            id[i] = nc == 1 ? lo : (nc == 0 ? -2 : lo);
        }
    };
    int nt = std::max(1, threads);
    std::vector<std::thread> pool;
    int64_t chunk = (n + nt - 1) / nt;
    for (int k = 1; k < nt; k++) pool.emplace_back(work, std::min(n, k * chunk), std::min(n, (k + 1) * chunk));
    work(0, std::min(n, chunk));
    for (auto& th : pool) th.join();
    return 0;
}

int ref_light_count(void* h) { return int(static_cast<ref_scene*>(h)->sc.lights.objects.size()); }

int ref_light_probabilities(void* h, double* out) {
    auto* r = static_cast<ref_scene*>(h);
    for (size_t k = 0; k < r->sc.lights.objects.size(); k++) out[k] = r->sc.lights.probability(k);
    return 0;
}

// The scene camera's environment (an image file, a uniform sky or the baked
// atmosphere) looked up by environment_light::radiance (environment.h:66-79)
// along n directions; returns -1 without one.
int ref_env_radiance(void* h, const double* dirs, int64_t n, double* out) {
    auto* r = static_cast<ref_scene*>(h);
    if (!r->sc.cam.environment) return -1;
    for (int64_t k = 0; k < n; k++) {
        color c = r->sc.cam.environment->radiance(vec3(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]));
        for (int i = 0; i < 3; i++) out[3 * k + i] = c[i];
    }
    return 0;
}

int ref_ggx_table(double* e, double* eavg) {
    const auto& t = ggx_albedo_table::get();
    std::memcpy(e, t.e, sizeof(t.e));
    std::memcpy(eavg, t.e_avg, sizeof(t.e_avg));
    return 0;
}

// Material eval and scatter-pdf value for fixed directions. spec per query
// (12 doubles): type (0 lambertian, 1 metal, 2 dielectric, 3 principled,
// 5 phase), r, g, b, roughness, ior, metallic, g (phase), sigma_a rgb, and a
// measured metal for metal and principled (0 none, 1.. Al Ti Cu Au Ag Cr).
// q: normal, ray direction, wi (9 doubles); front flags.
// material::eval and the scatter pdf's value of the materials spec rows
// describe (stride doubles each: type, r, g, b, roughness, ior, metallic, g,
// sigma_a (3), measured metal; with stride 17 also principled's clear coat:
// weight (0: none), roughness, tint (3)), at the queries' normals and
// directions.
int ref_bsdf_eval_rows(const double* spec, int stride, const double* q, const uint8_t* front, int64_t n, double* eval,
                       double* pdfv) {
    for (int64_t i = 0; i < n; i++) {
        const double* s = spec + stride * i;
        int type = int(s[0]);
        color c(s[1], s[2], s[3]);
        std::shared_ptr<material> m;
        static const char* metals[] = {"Al", "Ti", "Cu", "Au", "Ag", "Cr"};
        conductor_ior ior{};
        const int measured = int(s[11]);
        if (measured > 0) find_conductor(metals[measured - 1], ior);
        if (type == 0) m = std::make_shared<lambertian>(c);
        else if (type == 1) m = measured ? std::make_shared<metal>(ior, s[4]) : std::make_shared<metal>(c, s[4]);
        else if (type == 2) m = std::make_shared<dielectric>(s[5], s[4], color(s[8], s[9], s[10]));
        else if (type == 3) {
            auto p = std::make_shared<principled>(std::make_shared<solid_color>(c), s[6], s[4], s[5]);
            if (measured) p->set_conductor(ior);
            if (stride >= 17 && s[12] > 0) {
                p->set_clearcoat(std::make_shared<solid_color>(color(s[12], s[12], s[12])),
                                 std::make_shared<solid_color>(color(s[13], s[13], s[13])));
                p->set_clearcoat_color(color(s[14], s[15], s[16]));
            }
            m = p;
        } else m = std::make_shared<phase_material>(c, s[7]);
        hit_record rec{};
        rec.normal = vec3(q[9 * i], q[9 * i + 1], q[9 * i + 2]);
        rec.p = point3(0, 0, 0);
        rec.t = 1;
        rec.u = rec.v = 0;
        rec.front_face = front[i] != 0;
        rec.mat = m.get();
        ray r(point3(0, 0, 0), vec3(q[9 * i + 3], q[9 * i + 4], q[9 * i + 5]));
        vec3 wi(q[9 * i + 6], q[9 * i + 7], q[9 * i + 8]);
        color e = m->eval(r, rec, wi);
        eval[3 * i] = e.x(), eval[3 * i + 1] = e.y(), eval[3 * i + 2] = e.z();
        scatter_record sr;
        pdfv[i] = 0;
        if (type != 2 || s[4] > 0) {  // smooth glass would draw a random number and is specular
            if (m->scatter(r, rec, sr) && !sr.skip_pdf) pdfv[i] = sr.pdf_ptr->value(wi);
        }
    }
    return 0;
}

int ref_bsdf_eval(const double* spec, const double* q, const uint8_t* front, int64_t n, double* eval, double* pdfv) {
    return ref_bsdf_eval_rows(spec, 12, q, front, n, eval, pdfv);
}

// camera::render of the loaded scene with width/spp/depth overridden;
// radiance is w*h*3 doubles (the reference's own unseeded RNG).
int ref_render(void* h, int width, int spp, int max_depth, int threads, double* radiance) {
    auto* r = static_cast<ref_scene*>(h);
    camera cam = r->sc.cam;
    cam.image_width = width;
    cam.samples_per_pixel = spp;
    cam.max_depth = max_depth;
    cam.threads = threads;
    hittable_list everything(r->tree);
    for (const auto& o : r->sc.unbounded.objects) everything.add(o);
    const hittable& world = r->sc.unbounded.empty() ? static_cast<const hittable&>(*r->tree) : everything;
    frame f = cam.render(world, r->sc.lights);
    for (size_t k = 0; k < f.radiance.size(); k++)
        for (int c = 0; c < 3; c++) radiance[3 * k + c] = f.radiance[k][c];
    return 0;
}

// denoiser::run (denoiser.h:45-78) on interleaved RGB doubles.
int ref_denoise(const double* rad, const double* alb, const double* nrm, const double* var, int w, int h,
                int iterations, double sigma_l, double sigma_n, double sigma_a, double* out) {
    const size_t n = size_t(w) * h;
    frame f;
    for (size_t k = 0; k < n; k++) {
        f.radiance.push_back(color(rad[3 * k], rad[3 * k + 1], rad[3 * k + 2]));
        f.albedo.push_back(color(alb[3 * k], alb[3 * k + 1], alb[3 * k + 2]));
        f.normal.push_back(color(nrm[3 * k], nrm[3 * k + 1], nrm[3 * k + 2]));
        f.variance.push_back(var[k]);
    }
    denoiser d;
    d.iterations = iterations;
    d.sigma_luminance = sigma_l;
    d.sigma_normal = sigma_n;
    d.sigma_albedo = sigma_a;
    d.threads = 1;
    std::vector<color> o = d.run(f, w, h);
    for (size_t k = 0; k < n; k++)
        for (int c = 0; c < 3; c++) out[3 * k + c] = o[k][c];
    return 0;
}


// Display settings from the harness's flat arguments.
static display_settings make_display(double exposure, int op, double bloom, double threshold, double radius,
                                     double vignette) {
    display_settings d;
    d.exposure = exposure;
    d.op = op == 1 ? tonemapper::reinhard : (op == 2 ? tonemapper::aces : tonemapper::none);
    d.post.bloom = bloom;
    d.post.bloom_threshold = threshold;
    d.post.bloom_radius = radius;
    d.post.vignette = vignette;
    return d;
}

static std::vector<color> to_colors(const double* rad, size_t n) {
    std::vector<color> px(n);
    for (size_t k = 0; k < n; k++) px[k] = color(rad[3 * k], rad[3 * k + 1], rad[3 * k + 2]);
    return px;
}

int ref_postprocess(const double* rad, int w, int h, double bloom, double threshold, double radius, double vignette,
                    double* out) {
    const display_settings d = make_display(0, 0, bloom, threshold, radius, vignette);
    const std::vector<color> o = postprocess(to_colors(rad, size_t(w) * h), w, h, d.post);
    for (size_t k = 0; k < o.size(); k++)
        for (int c = 0; c < 3; c++) out[3 * k + c] = o[k][c];
    return 0;
}

// The bytes write_png builds (image_io.h:25-37 after write_image's postprocess, 81-92).
int ref_display_bytes(const double* rad, int w, int h, double exposure, int op, double bloom, double threshold,
                      double radius, double vignette, int post, uint8_t* out) {
    const display_settings d = make_display(exposure, op, bloom, threshold, radius, vignette);
    std::vector<color> px = to_colors(rad, size_t(w) * h);
    if (post && d.post.active()) px = postprocess(px, w, h, d.post);
    for (size_t k = 0; k < px.size(); k++) {
        const color c = d.apply(px[k]);
        for (int q = 0; q < 3; q++) out[3 * k + q] = to_byte(c[q]);
    }
    return 0;
}

int ref_write_image(const char* path, int w, int h, const double* rad, double exposure, int op, double bloom,
                    double threshold, double radius, double vignette) {
    const display_settings d = make_display(exposure, op, bloom, threshold, radius, vignette);
    return write_image(path ? path : "", w, h, to_colors(rad, size_t(w) * h), d) ? 0 : -1;
}

// A .scene file's display settings: exposure, op, bloom, threshold, radius, vignette.
int ref_scene_display(const char* path, double* out) {
    scene sc;
    if (!load_scene(path, sc)) return -1;
    const display_settings& d = sc.display;
    out[0] = d.exposure;
    out[1] = d.op == tonemapper::reinhard ? 1 : (d.op == tonemapper::aces ? 2 : 0);
    out[2] = d.post.bloom;
    out[3] = d.post.bloom_threshold;
    out[4] = d.post.bloom_radius;
    out[5] = d.post.vignette;
    return 0;
}


// stb_image's decoding of an image as floats (environment_light's loader,
// environment.h:30-41): width, height and up to cap floats of RGB.
int ref_load_image_f(const char* path, int* w, int* h, float* out, int64_t cap) {
    int c;
    float* data = stbi_loadf(path, w, h, &c, 3);
    if (!data) return -1;
    const int64_t n = int64_t(*w) * *h * 3;
    for (int64_t k = 0; k < n && k < cap; k++) out[k] = data[k];
    stbi_image_free(data);
    return 0;
}

// What a hit hands its material, per ray (16 doubles): rec.u, rec.v,
// rec.dpdu, rec.dpdv (world space), the base texture's value there
// (lambertian's albedo, principled's base color; 0 otherwise), the normal
// the material shades with (normal_from_map's through normal_mapped or
// principled's own map; rec.normal without one), principled's metallic_at
// and alpha_at (0 otherwise). A miss leaves its row NaN.
int ref_surface(void* h, const double* rays, const double* tt, int64_t n, double* out) {
    auto* r = static_cast<ref_scene*>(h);
    for (int64_t i = 0; i < n; i++) {
        double* o = out + 16 * i;
        for (int k = 0; k < 16; k++) o[k] = std::nan("");
        ray rr(point3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]), vec3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]));
        hit_record rec;
        if (!r->world().hit(rr, interval(tt[2 * i], tt[2 * i + 1]), rec)) continue;
        const material* m = rec.mat;
        vec3 shading = rec.normal;
        if (auto nm = dynamic_cast<const normal_mapped*>(m)) {
            shading = nm->perturbed(rr, rec).normal;
            m = nm->inner.get();
        }
        color base(0, 0, 0);
        double metallic = 0, alpha = 0;
        if (auto l = dynamic_cast<const lambertian*>(m)) {
            base = l->tex->value(rec.u, rec.v, rec.p);
        } else if (auto p = dynamic_cast<const principled*>(m)) {
            base = p->base_color->value(rec.u, rec.v, rec.p);
            metallic = p->metallic_at(rec);
            alpha = p->alpha_at(rec);
            if (p->normal_map) shading = normal_from_map(rr, rec, *p->normal_map);
        }
        const double v[16] = {rec.u, rec.v, rec.dpdu.x(), rec.dpdu.y(), rec.dpdu.z(), rec.dpdv.x(), rec.dpdv.y(), rec.dpdv.z(),
                              base.x(), base.y(), base.z(), shading.x(), shading.y(), shading.z(), metallic, alpha};
        std::memcpy(o, v, sizeof(v));
    }
    return 0;
}

// stbi_load(path, ..., 4): what image_texture keeps of an 8-bit file
// (texture.h:81-84), and the file's own channel count.
int ref_load_image_u8(const char* path, int* w, int* h, int* channels, uint8_t* out, int64_t cap) {
    unsigned char* data = stbi_load(path, w, h, channels, 4);
    if (!data) return -1;
    const int64_t n = int64_t(*w) * *h * 4;
    for (int64_t k = 0; k < n && k < cap; k++) out[k] = data[k];
    stbi_image_free(data);
    return 0;
}

// stbi_write_jpg: baseline JPEG files for the decoder's tests (quality above
// 90 keeps full-resolution chroma, at or below it chroma is subsampled 2x2).
int ref_write_jpg(const char* path, int w, int h, int comp, const uint8_t* data, int quality) {
    return stbi_write_jpg(path, w, h, comp, data, quality) ? 0 : -1;
}

}  // extern "C"
