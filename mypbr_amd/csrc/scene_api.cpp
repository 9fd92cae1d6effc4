// scene_api.cpp - the C++ scene API (include/mpt/mpt.hpp): transforms,
// materials, flattening of world and lights into the C ABI description,
// camera::render on the GPU, and the .scene loader (scene.h:244-1250, for
// the subset of keywords the GPU path models).
#include "../../include/mpt/mpt.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <tuple>

#include "geom.hpp"
#include "image_decode.hpp"
#include "scene_host.hpp"

namespace mpt {

namespace {
constexpr double kPiApi = 3.1415926535897932385;
double degrees_to_radians(double d) { return d * kPiApi / 180.0; }
double clamp01(double x) { return x < 0 ? 0 : (x > 1 ? 1 : x); }
double r2a(double r) {  // roughness_to_alpha, microfacet.h:17-19
    double a = r * r;
    return a > 1e-3 ? a : 1e-3;
}
}  // namespace

// ------------------------------------------------------------------ transform (transform.h)
transform::transform() : m{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, t(0, 0, 0), inv_m{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, inv_t(0, 0, 0) {}

transform transform::translate(const vec3& offset) {
    transform x;
    x.t = offset;
    x.inv_t = -offset;
    return x;
}

transform transform::scale(const vec3& s) {
    transform x;
    for (int k = 0; k < 3; k++) {
        x.m[k][k] = s[k];
        x.inv_m[k][k] = 1 / s[k];
    }
    return x;
}

transform transform::rotate(double degrees, const vec3& axis) {
    vec3 a = unit_vector(axis);
    double theta = degrees_to_radians(degrees);
    double c = std::cos(theta), s = std::sin(theta), C = 1 - c;
    transform x;
    double r[3][3] = {
        {a.x() * a.x() * C + c, a.x() * a.y() * C - a.z() * s, a.x() * a.z() * C + a.y() * s},
        {a.y() * a.x() * C + a.z() * s, a.y() * a.y() * C + c, a.y() * a.z() * C - a.x() * s},
        {a.z() * a.x() * C - a.y() * s, a.z() * a.y() * C + a.x() * s, a.z() * a.z() * C + c},
    };
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            x.m[i][j] = r[i][j];
            x.inv_m[j][i] = r[i][j];
        }
    return x;
}

transform transform::orthogonal(const double r[3][3]) {
    transform x;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            x.m[i][j] = r[i][j];
            x.inv_m[j][i] = r[i][j];
        }
    return x;
}

static void m_mul(const double a[3][3], const double b[3][3], double out[3][3]) {
    double r[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i][j] = r[i][j];
}
static vec3 mul(const double a[3][3], const vec3& v) {
    return vec3(a[0][0] * v[0] + a[0][1] * v[1] + a[0][2] * v[2], a[1][0] * v[0] + a[1][1] * v[1] + a[1][2] * v[2],
                a[2][0] * v[0] + a[2][1] * v[1] + a[2][2] * v[2]);
}

transform transform::then_after(const transform& first) const {
    transform x;
    m_mul(m, first.m, x.m);
    x.t = apply_vector(first.t) + t;
    m_mul(first.inv_m, inv_m, x.inv_m);
    x.inv_t = first.apply_inverse_vector(inv_t) + first.inv_t;
    return x;
}
point3 transform::apply_point(const point3& p) const { return apply_vector(p) + t; }
vec3 transform::apply_vector(const vec3& v) const { return mul(m, v); }
point3 transform::apply_inverse_point(const point3& p) const { return apply_inverse_vector(p) + inv_t; }
vec3 transform::apply_inverse_vector(const vec3& v) const { return mul(inv_m, v); }

bool transform::is_similarity() const {
    vec3 c0(m[0][0], m[1][0], m[2][0]), c1(m[0][1], m[1][1], m[2][1]), c2(m[0][2], m[1][2], m[2][2]);
    double l = c0.length_squared();
    auto close = [l](double x, double y) { return std::fabs(x - y) <= 1e-9 * std::fmax(1.0, l); };
    return close(c1.length_squared(), l) && close(c2.length_squared(), l) && close(dot(c0, c1), 0) &&
           close(dot(c1, c2), 0) && close(dot(c0, c2), 0);
}

double transform::area_scale() const {
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    return std::pow(std::fabs(det), 2.0 / 3.0);
}

mpt_xform transform::to_abi() const {
    mpt_xform x;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            x.m[3 * i + j] = m[i][j];
            x.inv_m[3 * i + j] = inv_m[i][j];
        }
    for (int i = 0; i < 3; i++) {
        x.t[i] = t[i];
        x.inv_t[i] = inv_t[i];
    }
    return x;
}

// ------------------------------------------------------------------ textures and materials
mpt_texture solid_color::to_abi() const {
    mpt_texture t{};
    t.type = MPT_TEX_SOLID;
    for (int k = 0; k < 3; k++) t.even[k] = t.odd[k] = albedo[k];
    return t;
}

checker_texture::checker_texture(double scale, std::shared_ptr<texture> e, std::shared_ptr<texture> o)
    : inv_scale(1.0 / scale), even(std::move(e)), odd(std::move(o)) {}

mpt_texture checker_texture::to_abi() const {
    auto se = std::dynamic_pointer_cast<solid_color>(even);
    auto so = std::dynamic_pointer_cast<solid_color>(odd);
    if (!se || !so) throw std::runtime_error("checker_texture: only solid colors are supported as cells");
    mpt_texture t{};
    t.type = MPT_TEX_CHECKER;
    t.inv_scale = inv_scale;
    for (int k = 0; k < 3; k++) {
        t.even[k] = se->albedo[k];
        t.odd[k] = so->albedo[k];
    }
    return t;
}

namespace {
bool read_hdr(const std::string& path, int& w, int& h, std::vector<color>& px);
}

// image_texture's constructor (texture.h:74-87): a Radiance file by its
// signature (stbi_is_hdr), else PNG or JPEG.
image_texture::image_texture(const std::string& path, bool rep, bool lin) : repeat(rep), linear(lin) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        error = "cannot open '" + path + "'";
        return;
    }
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    auto starts = [&](const char* sig) {
        const size_t n = std::strlen(sig);
        return bytes.size() >= n && std::memcmp(bytes.data(), sig, n) == 0;
    };
    if (starts("#?RADIANCE\n") || starts("#?RGBE\n")) {
        std::vector<color> px;
        int w = 0, h = 0;
        if (!read_hdr(path, w, h, px)) {
            error = "cannot read the Radiance file '" + path + "'";
            return;
        }
        hdr.resize(px.size() * 3);
        for (size_t i = 0; i < px.size(); i++)
            for (int k = 0; k < 3; k++) hdr[3 * i + k] = float(px[i][k]);  // exact: the decoder's values are floats
        width = w;
        height = h;
        channels = 3;
        return;
    }
    ldr_image im;
    if (!decode_ldr_image(bytes, im, error)) {
        error = path + ": " + error;
        return;
    }
    width = im.width;
    height = im.height;
    channels = im.channels;
    ldr = std::move(im.rgba);
}

image_texture::image_texture(const uint8_t* data, size_t size, bool rep, bool lin) : repeat(rep), linear(lin) {
    const std::vector<uint8_t> bytes(data, data + size);
    ldr_image im;
    if (!decode_ldr_image(bytes, im, error)) return;
    width = im.width;
    height = im.height;
    channels = im.channels;
    ldr = std::move(im.rgba);
}

image_texture::image_texture(double coverage) : repeat(false), linear(false) {
    const auto a = static_cast<uint8_t>(std::lround(255 * std::clamp(coverage, 0.0, 1.0)));
    width = height = 1;
    channels = 4;
    ldr = {255, 255, 255, a};
}

mpt_texture scaled_texture::to_abi() const {
    mpt_texture t{};
    t.type = MPT_TEX_SCALED;
    for (int k = 0; k < 3; k++) t.even[k] = scale[k];
    return t;
}
mpt_texture channel_texture::to_abi() const {
    mpt_texture t{};
    t.type = MPT_TEX_CHANNEL;
    t.channel = channel;
    t.scale = scale;
    return t;
}
uv_transform_texture::uv_transform_texture(std::shared_ptr<texture> in, double ou, double ov, double rotation, double su,
                                           double sv)
    : inner(std::move(in)), offset_u(ou), offset_v(ov), scale_u(su), scale_v(sv), cos_r(std::cos(rotation)),
      sin_r(std::sin(rotation)) {}
mpt_texture uv_transform_texture::to_abi() const {
    mpt_texture t{};
    t.type = MPT_TEX_UV_TRANSFORM;
    const double uv[6] = {offset_u, offset_v, scale_u, scale_v, cos_r, sin_r};
    for (int k = 0; k < 6; k++) t.uv[k] = uv[k];
    return t;
}
mpt_texture normal_scale_texture::to_abi() const {
    mpt_texture t{};
    t.type = MPT_TEX_NORMAL_SCALE;
    t.scale = scale;
    return t;
}

bool image_texture::fully_opaque() const {
    if (!hdr.empty()) return true;
    const size_t mask = (channels == 2 || channels == 4) ? 3 : 0;
    for (size_t k = mask; k < ldr.size(); k += 4)
        if (ldr[k] < 255) return false;
    return true;
}

mpt_texture image_texture::to_abi() const {
    mpt_texture t{};
    t.type = MPT_TEX_IMAGE;
    t.flags = (linear ? MPT_TEXF_LINEAR : 0) | (repeat ? MPT_TEXF_REPEAT : 0) | (hdr.empty() ? 0 : MPT_TEXF_HDR) |
              (channels == 2 || channels == 4 ? 0 : MPT_TEXF_NO_ALPHA);
    t.width = width;
    t.height = height;
    t.offset = 0;
    return t;
}

static mpt_material blank(int type, int32_t ti) {
    mpt_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = type;
    m.texture = ti;
    m.spread = 180;
    m.tan_half = __builtin_huge_val();
    m.normal_map = m.metallic_texture = m.roughness_texture = -1;
    m.coat = 0;
    m.coat_weight_texture = m.coat_roughness_texture = m.coat_normal_map = -1;
    m.coat_tint[0] = m.coat_tint[1] = m.coat_tint[2] = 1;
    return m;
}

mpt_material lambertian::to_abi(int32_t ti) const { return blank(MPT_MAT_LAMBERTIAN, ti); }

// The table of find_conductor (microfacet.h:76-96): RGB (eta, k) of six
// metals, data that has to equal the reference's for the same reflectance.
bool find_conductor(const std::string& name, conductor_ior& ior) {
    struct entry {
        const char* names[3];
        double eta[3], k[3];
    };
    static const entry table[] = {
        {{"Al", "aluminium", "aluminum"}, {1.6575, 0.880405, 0.521244}, {9.22381, 6.2695, 4.837}},
        {{"Ti", "titanium", "titanium"}, {3.2754, 2.37975, 2.02469}, {4.57233, 3.22453, 2.72781}},
        {{"Cu", "copper", "copper"}, {0.201005, 0.92375, 1.10222}, {3.91326, 2.45305, 2.14209}},
        {{"Au", "gold", "gold"}, {0.143036, 0.375307, 1.44205}, {3.983, 2.38556, 1.60336}},
        {{"Ag", "silver", "silver"}, {0.155276, 0.116728, 0.138388}, {4.82835, 3.12222, 2.1469}},
        {{"Cr", "chromium", "chrome"}, {4.36967, 2.91651, 1.6549}, {5.20655, 4.23131, 3.7547}},
    };
    for (const entry& e : table)
        for (const char* n : e.names)
            if (name == n) {
                ior.eta = color(e.eta[0], e.eta[1], e.eta[2]);
                ior.k = color(e.k[0], e.k[1], e.k[2]);
                return true;
            }
    return false;
}


// A measured metal's index; the scene build derives its head-on color and
// hemispherical average from it.
static void set_conductor(mpt_material& m, const conductor_ior& ior) {
    m.measured = 1;
    for (int c = 0; c < 3; c++) {
        m.eta[c] = ior.eta[c];
        m.k[c] = ior.k[c];
    }
}

metal::metal(const color& a, double r) : albedo(a), roughness(std::clamp(r, 0.0, 1.0)) {}
metal::metal(const conductor_ior& i, double r) : metal(color(1, 1, 1), r) {
    measured = true;
    ior = i;
}
mpt_material metal::to_abi(int32_t ti) const {
    mpt_material m = blank(MPT_MAT_METAL, ti);
    m.roughness = roughness;
    m.alpha = r2a(roughness);
    color f_avg = albedo + (color(1, 1, 1) - albedo) / 21;  // metal_fresnel::schlick, microfacet.h:102-107
    for (int k = 0; k < 3; k++) {
        m.albedo[k] = albedo[k];
        m.f_avg[k] = f_avg[k];
    }
    if (measured) set_conductor(m, ior);
    return m;
}

mpt_material thin_dielectric::to_abi(int32_t ti) const {
    mpt_material m = blank(MPT_MAT_DIELECTRIC, ti);
    m.ior = ior;
    m.thin = 1;
    for (int k = 0; k < 3; k++) m.albedo[k] = tint[k];
    return m;
}

dielectric::dielectric(double ri, double r, const color& absorption)
    : refraction_index(ri), roughness(std::clamp(r, 0.0, 1.0)), sigma_a(absorption) {}
color dielectric::absorption_for(const color& tint, double distance) {
    return color(-std::log(tint.x()), -std::log(tint.y()), -std::log(tint.z())) / distance;
}
mpt_material dielectric::to_abi(int32_t ti) const {
    mpt_material m = blank(MPT_MAT_DIELECTRIC, ti);
    m.ior = refraction_index;
    m.roughness = roughness;
    m.alpha = r2a(roughness);
    for (int k = 0; k < 3; k++) m.sigma_a[k] = sigma_a[k];
    return m;
}

principled::principled(std::shared_ptr<texture> base, double met, double rough, double i)
    : base_color(std::move(base)), metallic(met), roughness(rough), ior(i) {}
void principled::set_clearcoat(std::shared_ptr<texture> weight, std::shared_ptr<texture> roughness,
                               std::shared_ptr<texture> coat_normals) {
    coat_map = std::move(weight);
    coat_roughness_map = roughness ? std::move(roughness) : std::make_shared<solid_color>(color(0, 0, 0));
    coat_normal_map = std::move(coat_normals);
}

mpt_material principled::to_abi(int32_t ti) const {
    mpt_material m = blank(MPT_MAT_PRINCIPLED, ti);
    m.metallic = clamp01(metallic);
    m.roughness = clamp01(roughness);
    m.alpha = r2a(m.roughness);
    m.ior = ior;
    double r0 = (ior - 1) / (ior + 1);
    m.f0 = r0 * r0;
    if (measured) set_conductor(m, conductor);
    if (coat_map) {
        m.coat = 1;
        for (int k = 0; k < 3; k++) m.coat_tint[k] = coat_tint[k];
    }
    return m;
}

diffuse_light::diffuse_light(const color& e, double s) : emit(e), spread(s) {}
color diffuse_light::average_emission() const {  // material.h:586-615
    double half = degrees_to_radians(spread) / 2;
    double share;
    if (half >= kPiApi / 2 - 1e-9) share = 1;
    else {
        double s = std::sin(half), c = std::cos(half);
        share = s * s - (half - s * c) * c / s;
    }
    return emit * share;
}
mpt_material diffuse_light::to_abi(int32_t ti) const {
    mpt_material m = blank(MPT_MAT_DIFFUSE_LIGHT, ti);
    m.spread = spread;
    double half = degrees_to_radians(spread) / 2;
    m.tan_half = spread >= 180 ? __builtin_huge_val() : std::tan(half);
    color avg = average_emission();
    for (int k = 0; k < 3; k++) {
        m.emit[k] = emit[k];
        m.avg_emission[k] = avg[k];
    }
    return m;
}

mpt_material phase_material::to_abi(int32_t ti) const {
    mpt_material m = blank(MPT_MAT_PHASE, ti);
    m.g = g;
    for (int k = 0; k < 3; k++) m.albedo[k] = albedo[k];
    return m;
}

std::shared_ptr<hittable_list> box(const point3& a, const point3& b, std::shared_ptr<material> mat) {
    auto sides = std::make_shared<hittable_list>();
    auto mn = point3(std::fmin(a.x(), b.x()), std::fmin(a.y(), b.y()), std::fmin(a.z(), b.z()));
    auto mx = point3(std::fmax(a.x(), b.x()), std::fmax(a.y(), b.y()), std::fmax(a.z(), b.z()));
    auto dx = vec3(mx.x() - mn.x(), 0, 0);
    auto dy = vec3(0, mx.y() - mn.y(), 0);
    auto dz = vec3(0, 0, mx.z() - mn.z());
    sides->add(std::make_shared<quad>(point3(mn.x(), mn.y(), mx.z()), dx, dy, mat));   // front
    sides->add(std::make_shared<quad>(point3(mx.x(), mn.y(), mx.z()), -dz, dy, mat));  // right
    sides->add(std::make_shared<quad>(point3(mx.x(), mn.y(), mn.z()), -dx, dy, mat));  // back
    sides->add(std::make_shared<quad>(point3(mn.x(), mn.y(), mn.z()), dz, dy, mat));   // left
    sides->add(std::make_shared<quad>(point3(mn.x(), mx.y(), mx.z()), dx, -dz, mat));  // top
    sides->add(std::make_shared<quad>(point3(mn.x(), mn.y(), mn.z()), dx, dz, mat));   // bottom
    return sides;
}

// ------------------------------------------------------------------ flattening
struct flat_builder::impl {
    std::vector<int32_t> prim_kind, prim_mat;
    std::vector<double> prim_data;
    std::vector<double> prim_nrm;  // 9 per primitive once any triangle has vertex normals
    std::vector<int32_t> obj_kind, obj_xform, obj_mat, obj_bkind, obj_alpha, obj_flags;
    bool any_alpha = false, any_flags = false;
    std::vector<mpt_vox_model> vox_models;
    std::vector<uint8_t> vox_cells;
    std::map<std::tuple<const void*, const void*, const void*, double>, int32_t> vox_index;
    std::vector<int64_t> obj_first, obj_count, obj_id;
    std::vector<double> obj_density;
    std::vector<mpt_xform> xforms;
    std::vector<mpt_material> materials;
    std::vector<mpt_texture> textures;
    std::vector<uint8_t> texels;     // image textures' RGBA bytes
    std::vector<float> texels_hdr;   // image textures' RGB floats
    std::vector<double> prim_uv;     // 6 per primitive once any triangle has texture coordinates
    bool has_ground = false;
    int32_t ground_mat = -1;
    double ground_height = 0;
    std::vector<double> fog_params;   // 3 per height fog
    std::vector<int32_t> fog_material;
    std::vector<int32_t> light_object;
    std::vector<double> env_rgb;
    std::map<const material*, int32_t> mat_index;
    std::map<const texture*, int32_t> tex_index;
    std::map<const hittable*, int32_t> object_of;
    int64_t next_id = 0;
    int32_t last_object = -1;
    mpt_scene_desc d{};
};

flat_builder::flat_builder() : p(new impl) {}
flat_builder::~flat_builder() { delete p; }

int32_t flat_builder::texture_index(const std::shared_ptr<texture>& t) {
    if (!t) return -1;
    auto jt = p->tex_index.find(t.get());
    if (jt != p->tex_index.end()) return jt->second;
    mpt_texture a = t->to_abi();
    a.inner = -1;
    if (auto in = t->inner_texture()) a.inner = texture_index(in);  // placed before the wrapper
    if (auto im = dynamic_cast<const image_texture*>(t.get())) {
        if (!im->loaded()) throw std::runtime_error("image texture not loaded: " + im->error);
        if (!im->hdr.empty()) {
            a.offset = int64_t(p->texels_hdr.size() / 3);
            p->texels_hdr.insert(p->texels_hdr.end(), im->hdr.begin(), im->hdr.end());
        } else {
            a.offset = int64_t(p->texels.size() / 4);
            p->texels.insert(p->texels.end(), im->ldr.begin(), im->ldr.end());
        }
    }
    const int32_t ti = int32_t(p->textures.size());
    p->textures.push_back(a);
    p->tex_index[t.get()] = ti;
    return ti;
}

int32_t flat_builder::material_index(const std::shared_ptr<material>& m) {
    if (!m) throw std::runtime_error("object without a material");
    auto it = p->mat_index.find(m.get());
    if (it != p->mat_index.end()) return it->second;
    auto image = [](const std::shared_ptr<texture>& t) { return dynamic_cast<const image_texture*>(t.get()) != nullptr; };
    mpt_material a = m->to_abi(texture_index(m->base_texture()));
    a.normal_map = texture_index(m->normal_map_texture());
    a.metallic_texture = texture_index(m->metallic_texture());
    a.roughness_texture = texture_index(m->roughness_texture());
    if (a.coat) {
        a.coat_weight_texture = texture_index(m->coat_weight_texture());
        auto cr = m->coat_roughness_texture();
        a.coat_roughness_texture = texture_index(cr ? cr : std::make_shared<solid_color>(color(0, 0, 0)));
        a.coat_normal_map = texture_index(m->coat_normal_texture());
    }
    if (auto g = m->glow_texture()) {
        a.glow = 1;
        a.glow_texture = texture_index(g);
    } else {
        a.glow = 0;
        a.glow_texture = -1;
    }
    a.uses_uv = (a.normal_map >= 0 || a.coat_normal_map >= 0 || image(m->base_texture()) ||
                 image(m->metallic_texture()) || image(m->roughness_texture()) ||
                 (a.coat && (image(m->coat_weight_texture()) || image(m->coat_roughness_texture())))) ? 1 : 0;
    int32_t mi = int32_t(p->materials.size());
    p->materials.push_back(a);
    p->mat_index[m.get()] = mi;
    return mi;
}

int32_t flat_builder::xform_index(const transform& x) {
    p->xforms.push_back(x.to_abi());
    return int32_t(p->xforms.size() - 1);
}

int64_t flat_builder::add_prim(int32_t kind, int32_t mat, const double* data9) {
    p->prim_kind.push_back(kind);
    p->prim_mat.push_back(mat);
    p->prim_data.insert(p->prim_data.end(), data9, data9 + 9);
    if (!p->prim_nrm.empty()) {
        p->prim_nrm.push_back(std::nan(""));
        p->prim_nrm.insert(p->prim_nrm.end(), 8, 0.0);
    }
    if (!p->prim_uv.empty()) {
        p->prim_uv.push_back(std::nan(""));
        p->prim_uv.insert(p->prim_uv.end(), 5, 0.0);
    }
    return int64_t(p->prim_kind.size() - 1);
}

void flat_builder::set_ground(double height, int32_t material) {
    if (p->has_ground) throw std::runtime_error("only one ground plane is supported");
    p->has_ground = true;
    p->ground_height = height;
    p->ground_mat = material;
}

void flat_builder::add_fog(double density, double scale_height, double base, int32_t material) {
    p->fog_params.insert(p->fog_params.end(), {density, scale_height, base});
    p->fog_material.push_back(material);
}

void height_fog::flatten(flat_builder& b, int32_t xform) const {
    if (xform >= 0) throw std::runtime_error("a height fog cannot be instanced");
    if (!std::dynamic_pointer_cast<phase_material>(phase)) throw std::runtime_error("a height fog needs a volume material");
    b.add_fog(density, scale_height, base, b.material_index(phase));
}

aabb height_fog::bounding_box() const {
    const double inf = std::numeric_limits<double>::infinity();
    return aabb(interval(-inf, inf), interval(-inf, inf), interval(-inf, inf));
}

void ground_plane::flatten(flat_builder& b, int32_t xform) const {
    if (xform >= 0) throw std::runtime_error("a ground plane cannot be instanced");
    b.set_ground(height, b.material_index(mat));
}

void flat_builder::set_prim_uvs(int64_t prim, const double* uv6) {
    if (p->prim_uv.empty()) {  // every primitive so far has none
        p->prim_uv.assign(6 * p->prim_kind.size(), 0.0);
        for (size_t k = 0; k < p->prim_kind.size(); k++) p->prim_uv[6 * k] = std::nan("");
    }
    std::copy(uv6, uv6 + 6, p->prim_uv.begin() + 6 * prim);
}

void flat_builder::set_prim_normals(int64_t prim, const double* n9) {
    if (p->prim_nrm.empty()) {  // every primitive so far has none
        p->prim_nrm.assign(9 * p->prim_kind.size(), 0.0);
        for (size_t k = 0; k < p->prim_kind.size(); k++) p->prim_nrm[9 * k] = std::nan("");
    }
    std::copy(n9, n9 + 9, p->prim_nrm.begin() + 9 * prim);
}

int32_t flat_builder::add_object(int32_t kind, int64_t first, int64_t count, int32_t xform, const hittable* owner) {
    p->obj_kind.push_back(kind);
    p->obj_first.push_back(first);
    p->obj_count.push_back(count);
    p->obj_xform.push_back(xform);
    p->obj_id.push_back(p->next_id);
    p->obj_density.push_back(0);
    p->obj_mat.push_back(-1);
    p->obj_bkind.push_back(MPT_OBJ_PRIM);
    p->obj_alpha.push_back(-1);
    p->obj_flags.push_back(0);
    p->next_id += count;
    int32_t idx = int32_t(p->obj_kind.size() - 1);
    p->object_of[owner] = idx;
    p->last_object = idx;
    return idx;
}

static void sphere_data(const sphere& s, double* d) {
    std::memset(d, 0, 9 * sizeof(double));
    d[0] = s.center[0], d[1] = s.center[1], d[2] = s.center[2], d[3] = s.radius;
}
static void quad_data(const quad& q, double* d) {
    for (int k = 0; k < 3; k++) d[k] = q.Q[k], d[3 + k] = q.u[k], d[6 + k] = q.v[k];
}

void sphere::flatten(flat_builder& b, int32_t xform) const {
    double d[9];
    sphere_data(*this, d);
    int64_t pi = b.add_prim(MPT_PRIM_SPHERE, b.material_index(mat), d);
    b.add_object(MPT_OBJ_PRIM, pi, 1, xform, this);
}

void quad::flatten(flat_builder& b, int32_t xform) const {
    double d[9];
    quad_data(*this, d);
    int64_t pi = b.add_prim(MPT_PRIM_QUAD, b.material_index(mat), d);
    b.add_object(MPT_OBJ_PRIM, pi, 1, xform, this);
}

// Appends a list's children (spheres and quads) as primitives; returns the first.
static int64_t list_prims(flat_builder& b, const hittable_list& l) {
    int64_t first = -1;
    for (const auto& o : l.objects) {
        double d[9];
        int64_t pi;
        if (auto s = dynamic_cast<const sphere*>(o.get())) {
            sphere_data(*s, d);
            pi = b.add_prim(MPT_PRIM_SPHERE, b.material_index(s->mat), d);
        } else if (auto q = dynamic_cast<const quad*>(o.get())) {
            quad_data(*q, d);
            pi = b.add_prim(MPT_PRIM_QUAD, b.material_index(q->mat), d);
        } else {
            throw std::runtime_error("a hittable_list object may only hold spheres and quads on the GPU path");
        }
        if (first < 0) first = pi;
    }
    return first;
}

void hittable_list::flatten(flat_builder& b, int32_t xform) const {
    if (objects.empty()) throw std::runtime_error("empty hittable_list in the world");
    int64_t first = list_prims(b, *this);
    b.add_object(MPT_OBJ_LIST, first, int64_t(objects.size()), xform, this);
}

void bvh::flatten(flat_builder& b, int32_t xform) const {
    if (xform >= 0) throw std::runtime_error("instances of a bvh are not supported on the GPU path");
    for (const auto& o : objects) o->flatten(b, -1);
}

void mesh::flatten(flat_builder& b, int32_t xform) const {
    if (indices.empty() || indices.size() % 3 != 0) throw std::runtime_error("mesh needs index triples");
    auto it = b.mesh_ranges.find(this);
    int64_t first, count = int64_t(indices.size() / 3);
    if (it == b.mesh_ranges.end()) {
        int32_t mi = b.material_index(mat);
        first = -1;
        for (size_t k = 0; k < indices.size(); k += 3) {
            double d[9];
            for (int c = 0; c < 3; c++) {
                int32_t vi = indices[k + c];
                if (vi < 0 || size_t(vi) >= positions.size()) throw std::runtime_error("mesh index out of range");
                for (int a = 0; a < 3; a++) d[3 * c + a] = positions[vi][a];
            }
            int64_t pi = b.add_prim(MPT_PRIM_TRIANGLE, mi, d);
            if (first < 0) first = pi;
            if (!normal_indices.empty()) {
                if (normal_indices.size() != indices.size()) throw std::runtime_error("mesh normal indices must parallel its indices");
                double n[9];
                bool all = true;
                for (int c = 0; c < 3; c++) {
                    int32_t ni = normal_indices[k + c];
                    if (ni < 0 || size_t(ni) >= normals.size()) {
                        all = false;
                        break;
                    }
                    for (int a = 0; a < 3; a++) n[3 * c + a] = normals[ni][a];
                }
                if (all) b.set_prim_normals(pi, n);
            }
            if (!uv_indices.empty()) {
                if (uv_indices.size() != indices.size()) throw std::runtime_error("mesh uv indices must parallel its indices");
                double t[6];
                bool all = true;
                for (int c = 0; c < 3; c++) {
                    int32_t ti = uv_indices[k + c];
                    if (ti < 0 || size_t(ti) >= uvs.size()) {
                        all = false;
                        break;
                    }
                    t[2 * c] = uvs[ti][0];
                    t[2 * c + 1] = uvs[ti][1];
                }
                if (all) b.set_prim_uvs(pi, t);
            }
        }
        b.mesh_ranges[this] = {first, count};
    } else {
        first = it->second.first;
    }
    const int32_t oi = b.add_object(MPT_OBJ_MESH, first, count, xform, this);
    if (alpha_mask && !alpha_mask->fully_opaque()) {
        if (!alpha_mask->hdr.empty()) throw std::runtime_error("an alpha mask must be an 8-bit image");
        b.p->obj_alpha[size_t(oi)] = b.texture_index(alpha_mask);
        b.p->any_alpha = true;
    }
}

void instance::flatten(flat_builder& b, int32_t xform) const {
    if (xform >= 0) throw std::runtime_error("nested instances are not supported on the GPU path");
    if (dynamic_cast<const constant_medium*>(object.get()))
        throw std::runtime_error("transform a medium's boundary, not the medium (medium.h:55-56)");
    int32_t x = b.xform_index(this->xform);
    object->flatten(b, x);
    b.p->object_of[this] = b.p->last_object;
}

void constant_medium::flatten(flat_builder& b, int32_t xform) const {
    if (xform >= 0) throw std::runtime_error("instances of a medium are not supported");
    const hittable* bd = boundary.get();
    int32_t bx = -1;
    if (auto inst = dynamic_cast<const instance*>(bd)) {
        bx = b.xform_index(inst->xform);
        bd = inst->object.get();
    }
    int32_t mi = b.material_index(phase);
    int64_t first, count;
    int32_t bkind;
    if (auto s = dynamic_cast<const sphere*>(bd)) {
        double d[9];
        sphere_data(*s, d);
        first = b.add_prim(MPT_PRIM_SPHERE, s->mat ? b.material_index(s->mat) : mi, d);
        count = 1;
        bkind = MPT_OBJ_PRIM;
    } else if (auto l = dynamic_cast<const hittable_list*>(bd)) {
        first = list_prims(b, *l);
        count = int64_t(l->objects.size());
        bkind = MPT_OBJ_LIST;
    } else {
        throw std::runtime_error("a medium's boundary must be a sphere or a box on the GPU path");
    }
    int32_t oi = b.add_object(MPT_OBJ_MEDIUM, first, count, bx, this);
    b.p->obj_density[oi] = density;
    b.p->obj_mat[oi] = mi;
    b.p->obj_bkind[oi] = bkind;
}

void environment_light::flatten(flat_builder&, int32_t) const {
    // never hit (environment.h:53-56): not part of the world's geometry
}

// ------------------------------------------------------------------ sky atmosphere
// Single-scattering sky over a spherical Earth (atmosphere.h:26-165): a
// viewer 200 m above the ground looks along d; 32 midpoint steps to the top
// of the air (or the ground) gather sunlight scattered by air (Rayleigh) and
// haze (Mie, Henyey-Greenstein phase), each dimmed by the optical depth back
// to the viewer and, towards the sun, by a 16-step march to the top. The
// densities fall off with height (scale heights 8 km and 1.2 km; ozone a tent
// around 25 km). The sums keep the reference's operand order, so the baked
// image equals the reference's bit for bit.
namespace {
struct sky_air {
    double rayleigh = 1, mie = 1, mie_g = 0.8, ozone = 1, intensity = 3;  // atmosphere.h:13-24
    static constexpr double r_ground = 6360e3, r_top = 6460e3, r_view = 6360e3 + 200;

    // extinction (or, with scatter_only, scattering) coefficients at height h, per channel
    void coefficients(double h, double ext[3]) const {
        static const double br[3] = {5.802e-6, 13.558e-6, 33.1e-6}, bo[3] = {0.650e-6, 1.881e-6, 0.085e-6};
        const double rr = std::exp(-h / 8000), rm = std::exp(-h / 1200);
        const double ro = std::fmax(0.0, 1 - std::fabs(h - 25000) / 15000);
        for (int c = 0; c < 3; c++) ext[c] = rr * (rayleigh * br[c]) + rm * (mie * 4.40e-6) + ro * (ozone * bo[c]);
    }
    static double to_top(const vec3& o, const vec3& d) {
        double b = dot(o, d), c = o.length_squared() - r_top * r_top;
        return -b + std::sqrt(std::fmax(0.0, b * b - c));
    }
    static double to_ground(const vec3& o, const vec3& d) {  // -1 when the ray misses the ground
        double b = dot(o, d), c = o.length_squared() - r_ground * r_ground;
        double disc = b * b - c;
        if (disc < 0) return -1;
        double t = -b - std::sqrt(disc);
        return t > 0 ? t : -1;
    }
    color transmittance(const vec3& p, const vec3& d) const {  // atmosphere.h:146-160
        if (to_ground(p, d) > 0) return color(0, 0, 0);
        const double dt = to_top(p, d) / 16;
        double tau[3] = {0, 0, 0}, ext[3];
        for (int k = 0; k < 16; k++) {
            coefficients((p + (k + 0.5) * dt * d).length() - r_ground, ext);
            for (int c = 0; c < 3; c++) tau[c] += dt * ext[c];
        }
        return color(std::exp(-1.0 * tau[0]), std::exp(-1.0 * tau[1]), std::exp(-1.0 * tau[2]));
    }
    color radiance(const vec3& direction, const vec3& sun_direction, const color& e) const {  // atmosphere.h:58-95
        static const double br[3] = {5.802e-6, 13.558e-6, 33.1e-6};
        const vec3 d = unit_vector(direction), s = unit_vector(sun_direction);
        const vec3 o(0, r_view, 0);
        double t_end = to_top(o, d);
        const double tg = to_ground(o, d);
        if (tg > 0) t_end = std::fmin(t_end, tg);
        const double mu = dot(d, s);
        const double phase_r = 3 / (16 * kPiApi) * (1 + mu * mu);
        const double g = mie_g, denom = 1 + g * g - 2 * g * mu;
        const double phase_m = (1 - g * g) / (4 * kPiApi * denom * std::sqrt(denom));
        const double dt = t_end / 32;
        double tau[3] = {0, 0, 0}, sum[3] = {0, 0, 0}, ext[3];
        for (int k = 0; k < 32; k++) {
            const vec3 p = o + (k + 0.5) * dt * d;
            const double h = p.length() - r_ground;
            coefficients(h, ext);
            const double rr = std::exp(-h / 8000), rm = std::exp(-h / 1200);
            for (int c = 0; c < 3; c++) tau[c] += 0.5 * dt * ext[c];  // midpoint: half a step to p
            const color sun_t = transmittance(p, s);
            for (int c = 0; c < 3; c++) {
                const double scatter = rr * phase_r * (rayleigh * br[c]) + rm * phase_m * (mie * 3.996e-6);
                sum[c] += std::exp(-1.0 * tau[c]) * sun_t[c] * scatter * dt;
                tau[c] += 0.5 * dt * ext[c];
            }
        }
        return color(intensity * sum[0] * e[0], intensity * sum[1] * e[1], intensity * sum[2] * e[2]);
    }
    std::vector<color> bake(int w, int h, const vec3& sun_direction, const color& e) const {  // atmosphere.h:100-114
        std::vector<color> px(size_t(w) * h);
        for (int j = 0; j < h; j++) {
            const double theta = kPiApi * (j + 0.5) / h;
            for (int i = 0; i < w; i++) {
                const double phi = 2 * kPiApi * ((i + 0.5) / w - 0.5);
                const vec3 d(std::sin(theta) * std::sin(phi), std::cos(theta), -std::sin(theta) * std::cos(phi));
                px[size_t(j) * w + i] = radiance(d, sun_direction, e);
            }
        }
        return px;
    }
};
}  // namespace

namespace {
sky_air air_of(const atmosphere_settings& s) {
    sky_air a;
    a.rayleigh = s.rayleigh;
    a.mie = s.mie;
    a.mie_g = s.mie_g;
    a.ozone = s.ozone;
    a.intensity = s.intensity;
    return a;
}
}  // namespace

color atmosphere::sun_transmittance(const vec3& direction) const {  // atmosphere.h:52-56
    return air_of(settings).transmittance(vec3(0, sky_air::r_view, 0), unit_vector(direction));
}
color atmosphere::radiance(const vec3& direction, const vec3& sun_direction, const color& sun_irradiance) const {
    return air_of(settings).radiance(direction, sun_direction, sun_irradiance);
}
std::vector<color> atmosphere::bake(int width, int height, const vec3& sun_direction, const color& sun_irradiance) const {
    if (width < 1 || height < 1) throw std::runtime_error("atmosphere::bake: width and height must be positive");
    return air_of(settings).bake(width, height, sun_direction, sun_irradiance);
}

void sun_light::flatten(flat_builder&, int32_t) const {
    // never hit (sun.h:38-41): not part of the world's geometry
}

// ------------------------------------------------------------------ boxes, powers, light selection
namespace {
// A non-owning handle on an object the caller keeps alive (for the one-object
// descriptions below).
std::shared_ptr<hittable> borrowed(const hittable* h) {
    return std::shared_ptr<hittable>(std::shared_ptr<hittable>(), const_cast<hittable*>(h));
}
}  // namespace

aabb hittable::bounding_box() const {
    flat_builder b;
    b.add_world(*this);  // a list is expanded in place, as a world list
    double box[6];
    if (mpt_desc_bounds(&b.desc(), box) != MPT_OK)
        throw std::runtime_error(std::string("bounding_box: ") + mpt_last_error());
    return aabb(interval(box[0], box[1]), interval(box[2], box[3]), interval(box[4], box[5]));
}

aabb ground_plane::bounding_box() const {
    const double inf = std::numeric_limits<double>::infinity();
    return aabb(interval(-inf, inf), interval(height, height), interval(-inf, inf));
}

double hittable::power() const {
    if (auto l = dynamic_cast<const hittable_list*>(this)) {  // hittable_list::power (hittable_list.h:65-70)
        double sum = 0;
        for (const auto& o : l->objects) sum += o->power();
        return sum;
    }
    hittable_list w;
    w.add(borrowed(this));
    flat_builder b;
    try {
        b.add_world(w);
        b.add_lights(w, nullptr);
    } catch (const std::exception&) {
        return 0;  // nothing the light code can weigh (a sky, a sun, a plane)
    }
    double pw = 0;
    if (mpt_desc_light_powers(&b.desc(), nullptr, &pw) != MPT_OK) return 0;  // media are never lights
    return pw;
}

double environment_light::power(double scene_radius) const {  // environment.h:116-127
    if (pixels.empty()) return 0;
    double sum = 0;
    for (int j = 0; j < height; j++) {
        const double sin_theta = std::sin(kPi * (j + 0.5) / height);
        for (int i = 0; i < width; i++) {
            const color& c = pixels[size_t(j) * width + i];
            sum += std::fmax(luminance(v3{c[0], c[1], c[2]}) * sin_theta, 0.0);
        }
    }
    const double integral = intensity * sum * (2 * kPi / width) * (kPi / height);
    return kPi * scene_radius * scene_radius * integral;
}

double sun_light::power(double scene_radius) const {  // sun.h:66-71
    return kPi * scene_radius * scene_radius * luminance(v3{irradiance[0], irradiance[1], irradiance[2]});
}

void light_list::build(const aabb& scene_bounds) {
    built = true;
    has_bounds = true;
    bounds = scene_bounds;
    const size_t n = objects.size();
    probabilities_.assign(n, n ? 1.0 / double(n) : 0.0);
    cdf_.assign(n + 1, 0.0);
    if (n == 0) return;
    // the lights alone as the world (the sky and the sun are lights only), so
    // that the core's light code weighs them with these bounds
    hittable_list w;
    for (const auto& o : objects)
        if (!dynamic_cast<const environment_light*>(o.get()) && !dynamic_cast<const sun_light*>(o.get())) w.add(o);
    flat_builder b;
    b.add_world(w);
    b.add_lights(*this, nullptr);
    const double b6[6] = {scene_bounds.x.min, scene_bounds.x.max, scene_bounds.y.min,
                          scene_bounds.y.max, scene_bounds.z.min, scene_bounds.z.max};
    if (mpt_desc_light_probabilities(&b.desc(), b6, probabilities_.data()) != MPT_OK)
        throw std::runtime_error(std::string("light_list::build: ") + mpt_last_error());
    for (size_t k = 0; k < n; k++) cdf_[k + 1] = cdf_[k] + probabilities_[k];
    cdf_[n] = 1;  // light_list.h:92: no rounding gap at the top
}

double light_list::probability(size_t k) const {
    if (built && has_bounds && probabilities_.size() == objects.size()) return probabilities_[k];
    return 1.0 / double(objects.size());
}

size_t light_list::pick(double xi) const {
    const size_t n = objects.size();
    if (!(built && has_bounds && cdf_.size() == n + 1)) return std::min(size_t(xi * double(n)), n - 1);
    const auto it = std::upper_bound(cdf_.begin(), cdf_.end(), xi);
    const size_t k = size_t(it - cdf_.begin());
    return std::clamp(k, size_t(1), n) - 1;
}

void flat_builder::add_world(const hittable& world) {
    if (auto l = dynamic_cast<const hittable_list*>(&world)) {
        for (const auto& o : l->objects) o->flatten(*this, -1);
    } else {
        world.flatten(*this, -1);
    }
}

void flat_builder::add_lights(const hittable_list& lights, const std::shared_ptr<environment_light>& env) {
    const environment_light* envp = env.get();
    for (const auto& o : lights.objects) {
        if (auto e = dynamic_cast<const environment_light*>(o.get())) {
            if (envp && envp != e) throw std::runtime_error("only one environment is allowed");
            envp = e;
            p->light_object.push_back(-1);
            continue;
        }
        if (auto sun = dynamic_cast<const sun_light*>(o.get())) {
            if (p->d.has_sun) throw std::runtime_error("only one sun is allowed");
            p->d.has_sun = 1;
            for (int k = 0; k < 3; k++) {
                p->d.sun_direction[k] = sun->direction[k];
                p->d.sun_irradiance[k] = sun->irradiance[k];
            }
            p->d.sun_angle_degrees = sun->angular_diameter;
            p->light_object.push_back(-2);
            continue;
        }
        auto it = p->object_of.find(o.get());
        if (it == p->object_of.end()) {
            // sampled only, never hit: the emissive faces of a voxel model, whose
            // grid the rays hit instead (scene.h:1118-1128)
            const size_t before = p->obj_kind.size();
            o->flatten(*this, -1);
            if (p->obj_kind.size() != before + 1)
                throw std::runtime_error("a light outside the world must flatten to one object");
            p->obj_flags[before] = MPT_OBJF_LIGHT_ONLY;
            p->any_flags = true;
            it = p->object_of.find(o.get());
        }
        p->light_object.push_back(it->second);
    }
    if (envp) {
        if (envp->width < 1 || envp->height < 1 || envp->pixels.size() != size_t(envp->width) * envp->height)
            throw std::runtime_error("environment_light: pixel count does not match its size");
        p->env_rgb.resize(envp->pixels.size() * 3);
        for (size_t k = 0; k < envp->pixels.size(); k++)
            for (int c = 0; c < 3; c++) p->env_rgb[3 * k + c] = envp->pixels[k][c];
        p->d.env_width = envp->width;
        p->d.env_height = envp->height;
        p->d.env_intensity = envp->intensity;
        p->d.env_rotation_degrees = envp->rotation_degrees;
    }
    auto ll = dynamic_cast<const light_list*>(&lights);
    if (ll && ll->built) {
        p->d.light_strategy = ll->strategy == light_list::selection::power ? 1 : 0;
        p->d.light_uniform_fraction = ll->uniform_fraction;
        p->d.has_light_bounds = ll->has_bounds ? 1 : 0;
        if (ll->has_bounds) {
            const aabb& lb = ll->bounds;
            const double v[6] = {lb.x.min, lb.x.max, lb.y.min, lb.y.max, lb.z.min, lb.z.max};
            std::copy(v, v + 6, p->d.light_bounds);
        }
    } else {
        p->d.light_strategy = 2;  // plain hittable_list: uniform pick (hittable_list.h:58-62)
        p->d.light_uniform_fraction = 0;
    }
}

// ------------------------------------------------------------------ studio
std::shared_ptr<mesh> make_cyclorama(const cyclorama_settings& s, std::shared_ptr<material> mat) {
    // the profile in the (y, z) plane, from the floor's front to the wall's top;
    // the curve's centre is (y = radius, z = -back) and its normals point at it
    std::vector<double> ys, zs, nys, nzs;
    auto add = [&](double y, double z, double ny, double nz) {
        ys.push_back(y), zs.push_back(z), nys.push_back(ny), nzs.push_back(nz);
    };
    add(0, s.front, 1, 0);
    for (int i = 0; i <= s.segments; i++) {
        const double theta = (kPiApi / 2) * i / s.segments;
        add(s.radius * (1 - std::cos(theta)), -s.back - s.radius * std::sin(theta), std::cos(theta), std::sin(theta));
    }
    if (s.height > s.radius) add(s.height, -s.back - s.radius, 0, 1);
    std::vector<double> arc(ys.size(), 0);
    for (size_t i = 1; i < ys.size(); i++) arc[i] = arc[i - 1] + std::hypot(ys[i] - ys[i - 1], zs[i] - zs[i - 1]);
    std::vector<point3> positions;
    auto m = std::make_shared<mesh>(positions, std::vector<int32_t>(), mat);
    for (size_t i = 0; i < ys.size(); i++)
        for (int side = 0; side < 2; side++) {
            m->positions.push_back(point3(side == 0 ? -s.width / 2 : s.width / 2, ys[i], zs[i]));
            m->normals.push_back(vec3(0, nys[i], nzs[i]));
            m->uvs.push_back(vec3(side, arc[i] / arc.back(), 0));
        }
    // two triangles per strip between rows i and i + 1, counterclockwise seen from the front
    for (int i = 0; i + 1 < int(ys.size()); i++) {
        const int l0 = 2 * i, r0 = 2 * i + 1, l1 = 2 * i + 2, r1 = 2 * i + 3;
        for (int v : {l0, r0, r1, l0, r1, l1}) {
            m->indices.push_back(v);
            m->normal_indices.push_back(v);
            m->uv_indices.push_back(v);
        }
    }
    return m;
}

std::shared_ptr<quad> make_softbox(const point3& center, const point3& aim, double width, double height,
                                   double roll_degrees, std::shared_ptr<material> mat) {
    const vec3 dir = unit_vector(aim - center);
    const vec3 up = std::fabs(dir.y()) > 0.999 ? vec3(0, 0, 1) : vec3(0, 1, 0);
    const vec3 u = unit_vector(cross(up, dir));
    const vec3 v = cross(dir, u);  // cross(u, v) = dir: the front faces the aim point
    const double roll = degrees_to_radians(roll_degrees);
    const vec3 ur = std::cos(roll) * u + std::sin(roll) * v;
    const vec3 vr = -std::sin(roll) * u + std::cos(roll) * v;
    const point3 corner = center - (width / 2) * ur - (height / 2) * vr;
    return std::make_shared<quad>(corner, width * ur, height * vr, std::move(mat));
}

// ------------------------------------------------------------------ voxels
namespace {

// sRGB-encoded [0, 1] to linear (color3.h:26-30: the IEC 61966-2-1 curve).
double srgb_decode(double encoded) {
    return encoded <= 0.04045 ? encoded / 12.92 : std::pow((encoded + 0.055) / 1.055, 2.4);
}

// MagicaVoxel's default palette, for files without an RGBA chunk
// (vox_loader.h:62-81), as its structure gives it: index 0 empty; 1-215 the
// 6x6x6 cube of levels ff, cc, 99, 66, 33, 00 (blue fastest, then green, then
// red) without black; then ramps of ten levels (ee, dd, bb, aa, 88, 77, 55,
// 44, 22, 11) in red, green, blue and grey. Colours are in [0, 1], sRGB as stored.
void default_vox_palette(color out[256]) {
    static const int cube[6] = {0xff, 0xcc, 0x99, 0x66, 0x33, 0x00};
    static const int ramp[10] = {0xee, 0xdd, 0xbb, 0xaa, 0x88, 0x77, 0x55, 0x44, 0x22, 0x11};
    out[0] = color(0, 0, 0);
    for (int i = 0; i < 215; i++)
        out[1 + i] = color(cube[i / 36] / 255.0, cube[(i / 6) % 6] / 255.0, cube[i % 6] / 255.0);
    for (int j = 0; j < 10; j++) {
        const double v = ramp[j] / 255.0;
        out[216 + j] = color(v, 0, 0);
        out[226 + j] = color(0, v, 0);
        out[236 + j] = color(0, 0, v);
        out[246 + j] = color(v, v, v);
    }
}

// Little-endian fields of one chunk's content; a read past its end marks the
// chunk corrupt.
struct vox_reader {
    const uint8_t* p;
    const uint8_t* end;
    bool ok = true;
    int32_t i32() {
        if (end - p < 4) {
            ok = false;
            p = end;
            return 0;
        }
        const uint32_t v = uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24;
        p += 4;
        return int32_t(v);
    }
    std::string str() {
        const int32_t n = i32();
        if (n < 0 || n > end - p) {
            ok = false;
            p = end;
            return std::string();
        }
        std::string r(reinterpret_cast<const char*>(p), size_t(n));
        p += n;
        return r;
    }
    std::map<std::string, std::string> dict() {
        std::map<std::string, std::string> d;
        const int32_t n = i32();
        for (int32_t k = 0; k < n && ok; k++) {
            std::string key = str();
            d[key] = str();
        }
        return d;
    }
};

struct vox_node {
    char type = 0;  // 'T' transform, 'G' group, 'S' shape
    bool hidden = false;
    std::vector<int> children;
    int model = -1;  // a shape's (first frame) model
    transform xform;  // a transform's (first frame)
};

}  // namespace

bool load_vox(const std::string& path, vox_file& out, std::string& error) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        error = "cannot open '" + path + "'";
        return false;
    }
    const std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (bytes.size() < 20 || std::memcmp(bytes.data(), "VOX ", 4) != 0) {
        error = "'" + path + "' is not a MagicaVoxel .vox file";
        return false;
    }
    default_vox_palette(out.palette);
    out.models.clear();
    out.instances.clear();
    for (auto& m : out.materials) m.clear();
    const uint8_t* p = bytes.data() + 8;  // past "VOX " and the version
    const uint8_t* end = bytes.data() + bytes.size();
    if (end - p < 12 || std::memcmp(p, "MAIN", 4) != 0) {
        error = "'" + path + "' has no MAIN chunk";
        return false;
    }
    {
        vox_reader h{p + 4, p + 12};
        p += 12 + std::max(h.i32(), 0);  // MAIN's own content; its children follow as one flat run
    }
    std::map<int, vox_node> nodes;
    int sx = 0, sy = 0, sz = 0;
    while (end - p >= 12) {
        const std::string id(reinterpret_cast<const char*>(p), 4);
        vox_reader h{p + 4, p + 12};
        const int32_t content = h.i32(), children = h.i32();
        if (content < 0 || children < 0 || int64_t(content) + children > end - p - 12) {
            error = "'" + path + "' is truncated or corrupt (chunk " + id + ")";
            return false;
        }
        const uint8_t* body = p + 12;
        vox_reader in{body, body + content};
        p += 12 + int64_t(content) + children;
        if (id == "SIZE") {
            sx = in.i32(), sy = in.i32(), sz = in.i32();
            if (sx <= 0 || sy <= 0 || sz <= 0 || sx > 2048 || sy > 2048 || sz > 2048) {
                error = "'" + path + "' has a model of invalid size";
                return false;
            }
        } else if (id == "XYZI") {
            if (sx <= 0) {
                error = "'" + path + "' has voxels before their SIZE";
                return false;
            }
            vox_model m;
            m.nx = sx, m.ny = sy, m.nz = sz;
            m.cells.assign(size_t(sx) * sy * sz, 0);
            const int32_t n = in.i32();
            if (n < 0 || int64_t(n) * 4 > in.end - in.p) {
                error = "'" + path + "' has a truncated voxel list";
                return false;
            }
            for (int32_t k = 0; k < n; k++) {
                const uint8_t* v = in.p + 4 * int64_t(k);  // x, y, z, palette index
                if (v[0] < sx && v[1] < sy && v[2] < sz) m.cells[(size_t(v[2]) * sy + v[1]) * sx + v[0]] = v[3];
            }
            out.models.push_back(std::move(m));
            sx = sy = sz = 0;
        } else if (id == "RGBA") {
            if (content < 1024) {
                error = "'" + path + "' has a short palette";
                return false;
            }
            for (int k = 0; k < 255; k++)  // entry k is palette index k + 1
                out.palette[k + 1] = color(body[4 * k] / 255.0, body[4 * k + 1] / 255.0, body[4 * k + 2] / 255.0);
        } else if (id == "MATL") {
            const int index = in.i32();
            auto props = in.dict();
            if (in.ok && index >= 0 && index < 256) out.materials[index] = props;
        } else if (id == "nTRN") {
            vox_node& t = nodes[in.i32()];
            auto attr = in.dict();
            t.type = 'T';
            t.hidden = attr["_hidden"] == "1";
            t.children.push_back(in.i32());
            in.i32();  // reserved
            in.i32();  // layer
            if (in.i32() > 0) {  // frames: the first one places the child
                auto frame = in.dict();
                transform rotation;
                if (frame.count("_r")) {
                    // a signed permutation: bits 0-1 and 2-3 the columns of rows 0 and
                    // 1 (row 2 takes the other), bits 4-6 the rows' signs
                    int r = 4;
                    try {
                        r = std::stoi(frame["_r"]);
                    } catch (...) {
                        r = 4;
                    }
                    const int c0 = r & 3, c1 = (r >> 2) & 3, c2 = 3 - c0 - c1;
                    if (c0 > 2 || c1 > 2 || c0 == c1) {
                        error = "'" + path + "' has an invalid rotation in an nTRN chunk";
                        return false;
                    }
                    double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    m[0][c0] = (r & 16) ? -1 : 1;
                    m[1][c1] = (r & 32) ? -1 : 1;
                    m[2][c2] = (r & 64) ? -1 : 1;
                    rotation = transform::orthogonal(m);
                }
                vec3 offset(0, 0, 0);
                if (frame.count("_t")) {
                    std::istringstream xyz(frame["_t"]);
                    xyz >> offset[0] >> offset[1] >> offset[2];
                }
                t.xform = transform::translate(offset).then_after(rotation);
            }
        } else if (id == "nGRP") {
            vox_node& g = nodes[in.i32()];
            auto attr = in.dict();
            g.type = 'G';
            g.hidden = attr["_hidden"] == "1";
            const int32_t n = in.i32();
            for (int32_t k = 0; k < n && in.ok; k++) g.children.push_back(in.i32());
        } else if (id == "nSHP") {
            vox_node& sh = nodes[in.i32()];
            in.dict();
            sh.type = 'S';
            if (in.i32() > 0) {
                sh.model = in.i32();
                in.dict();
            }
        }
        if (!in.ok) {
            error = "'" + path + "' has a corrupt " + id + " chunk";
            return false;
        }
    }
    if (out.models.empty()) {
        error = "'" + path + "' has no voxels";
        return false;
    }
    // a model's cells sit around its centre, rounded down, as MagicaVoxel places them
    auto centred = [&](int k) {
        const vox_model& m = out.models[size_t(k)];
        return transform::translate(-vec3(m.nx / 2, m.ny / 2, m.nz / 2));
    };
    if (!nodes.count(0)) {  // no scene graph (version 150): each model on z = 0, centred in x and y
        for (int k = 0; k < int(out.models.size()); k++)
            out.instances.push_back({k, transform::translate(vec3(0, 0, out.models[size_t(k)].nz / 2)).then_after(centred(k))});
        return true;
    }
    // the graph from node 0, transforms composed on the way down; a stack (so
    // children are taken last first) with a depth cap for cyclic files
    struct item {
        int id, depth;
        transform xform;
    };
    std::vector<item> stack{{0, 0, transform()}};
    while (!stack.empty()) {
        const item it = stack.back();
        stack.pop_back();
        auto nt = nodes.find(it.id);
        if (nt == nodes.end() || nt->second.hidden || it.depth > 64) continue;
        const vox_node& n = nt->second;
        const transform here = n.type == 'T' ? it.xform.then_after(n.xform) : it.xform;
        for (int child : n.children) stack.push_back({child, it.depth + 1, here});
        if (n.model >= 0 && n.model < int(out.models.size())) out.instances.push_back({n.model, here.then_after(centred(n.model))});
    }
    return true;
}

bool voxel_face_visible(voxel_kind k, voxel_kind n) {
    if (k == voxel_kind::empty) return false;
    return n == voxel_kind::empty || (k == voxel_kind::opaque && n == voxel_kind::transparent);
}

voxel_palette make_voxel_palette(const vox_file& vox, double emission_scale) {
    voxel_palette pal;
    bool warned = false;
    auto number = [](const std::map<std::string, std::string>& props, const char* key, double fallback) {
        auto it = props.find(key);
        if (it == props.end()) return fallback;
        try {
            return std::stod(it->second);
        } catch (...) {
            return fallback;
        }
    };
    for (int k = 0; k < 256; k++) {
        const color& c = vox.palette[k];
        const color albedo(srgb_decode(c[0]), srgb_decode(c[1]), srgb_decode(c[2]));
        const auto& props = vox.materials[k];
        auto ti = props.find("_type");
        const std::string type = ti == props.end() ? "_diffuse" : ti->second;
        const double roughness = std::clamp(number(props, "_rough", 0.1), 0.0, 1.0);
        double ior = number(props, "_ri", 1 + number(props, "_ior", 0.5));
        if (ior <= 1) ior = 1.5;
        pal.kinds[k] = k == 0 ? voxel_kind::empty : voxel_kind::opaque;
        pal.emissive[k] = false;
        if (type == "_metal") {
            const double metallic = number(props, "_plastic", 0) > 0 ? 0 : std::clamp(number(props, "_metal", 1), 0.0, 1.0);
            pal.materials[k] = std::make_shared<principled>(std::make_shared<solid_color>(albedo), metallic, roughness, ior);
        } else if (type == "_glass") {
            // light crossing one voxel keeps mix(white, colour, _att) of itself
            const double att = std::clamp(number(props, "_att", 0.5), 0.0, 1.0);
            color tint = (1 - att) * color(1, 1, 1) + att * albedo;
            for (int i = 0; i < 3; i++) tint[i] = std::clamp(tint[i], 1e-3, 1.0);
            pal.materials[k] = std::make_shared<dielectric>(ior, roughness <= 0.05 ? 0 : roughness,
                                                            dielectric::absorption_for(tint, 1));
            pal.kinds[k] = voxel_kind::transparent;
        } else if (type == "_emit") {
            const double strength = number(props, "_emit", 1) * std::pow(10.0, number(props, "_flux", 0)) * emission_scale;
            pal.materials[k] = std::make_shared<diffuse_light>(strength * albedo);
            pal.emissive[k] = strength > 0;
        } else {
            if (type != "_diffuse" && !warned && k > 0) {
                std::fprintf(stderr, "Warning: .vox material type %s is rendered as diffuse\n", type.c_str());
                warned = true;
            }
            pal.materials[k] = std::make_shared<lambertian>(albedo);
        }
    }
    return pal;
}

void voxel_model_faces(const vox_model& m, const voxel_palette& pal, hittable_list& faces, hittable_list& emitters) {
    // the six sides in the order +x, -x, +y, -y, +z, -z; a + side's corner is
    // one step along its axis, and its edges (u, v) are the next two axes in
    // turn, swapped on the - side, so that cross(u, v) points out of the cell
    for (int z = 0; z < m.nz; z++)
        for (int y = 0; y < m.ny; y++)
            for (int x = 0; x < m.nx; x++) {
                const uint8_t index = m.at(x, y, z);
                const voxel_kind k = pal.kind(index);
                if (k == voxel_kind::empty) continue;
                for (int side = 0; side < 6; side++) {
                    const int a = side / 2, s = side % 2 == 0 ? 1 : -1;
                    int nb[3] = {x, y, z};
                    nb[a] += s;
                    if (!voxel_face_visible(k, pal.kind(m.at(nb[0], nb[1], nb[2])))) continue;
                    point3 Q(x, y, z);
                    if (s > 0) Q[a] += 1;
                    vec3 e1(0, 0, 0), e2(0, 0, 0);
                    e1[(a + 1) % 3] = 1;
                    e2[(a + 2) % 3] = 1;
                    auto face = std::make_shared<quad>(Q, s > 0 ? e1 : e2, s > 0 ? e2 : e1, pal.materials[index]);
                    faces.add(face);
                    if (pal.emissive[index]) emitters.add(face);
                }
            }
}

voxel_grid::voxel_grid(std::shared_ptr<const vox_model> model, std::shared_ptr<const voxel_palette> palette,
                       const int lo_[3], const int hi_[3], std::shared_ptr<material> edge_material, double edge_width)
    : model(std::move(model)), palette(std::move(palette)), edge_material(std::move(edge_material)),
      edge_width(this->edge_material ? edge_width : 0) {
    for (int a = 0; a < 3; a++) lo[a] = lo_[a], hi[a] = hi_[a];
}

int32_t flat_builder::vox_model_index(const vox_model& m, const voxel_palette& pal,
                                      const std::shared_ptr<material>& edge_material, double edge_width) {
    const auto key = std::make_tuple(static_cast<const void*>(&m), static_cast<const void*>(&pal),
                                     static_cast<const void*>(edge_material.get()), edge_width);
    auto it = p->vox_index.find(key);
    if (it != p->vox_index.end()) return it->second;
    if (m.nx < 1 || m.ny < 1 || m.nz < 1 || m.cells.size() != size_t(m.nx) * m.ny * m.nz)
        throw std::runtime_error("a voxel model's cells do not match its size");
    mpt_vox_model v;
    std::memset(&v, 0, sizeof(v));
    v.nx = m.nx, v.ny = m.ny, v.nz = m.nz;
    v.cells = int64_t(p->vox_cells.size());
    v.edge_width = edge_material ? edge_width : 0;
    v.edge_material = v.edge_width > 0 ? material_index(edge_material) : -1;
    bool used[256] = {};
    for (uint8_t c : m.cells) used[c] = true;
    for (int i = 0; i < 256; i++) {
        v.kind[i] = uint8_t(pal.kinds[size_t(i)]);
        // materials only for the indices cells use: an unused emitter or glass entry
        // would change which shading kernels and shadow queries the scene gets
        v.material[i] = used[i] && pal.kinds[size_t(i)] != voxel_kind::empty ? material_index(pal.materials[size_t(i)]) : -1;
    }
    p->vox_cells.insert(p->vox_cells.end(), m.cells.begin(), m.cells.end());
    p->vox_models.push_back(v);
    const int32_t idx = int32_t(p->vox_models.size() - 1);
    p->vox_index[key] = idx;
    return idx;
}

// The blocks as MPT_PRIM_VOXBLOCK primitives (once per set) and one voxel
// object over them, behind the instance transform if any.
static void flatten_blocks(flat_builder& b, int32_t xform, const void* key,
                           const std::vector<std::shared_ptr<voxel_grid>>& blocks, const hittable* owner) {
    if (blocks.empty()) throw std::runtime_error("a voxel model without voxels");
    auto it = b.mesh_ranges.find(key);
    if (it == b.mesh_ranges.end()) {
        int64_t first = -1;
        for (const auto& g : blocks) {
            const int32_t mi = b.vox_model_index(*g->model, *g->palette, g->edge_material, g->edge_width);
            double d[9] = {double(g->lo[0]), double(g->lo[1]), double(g->lo[2]), double(g->hi[0]), double(g->hi[1]),
                           double(g->hi[2]), double(mi), 0, 0};
            const int64_t pi = b.add_prim(MPT_PRIM_VOXBLOCK, -1, d);
            if (first < 0) first = pi;
        }
        it = b.mesh_ranges.emplace(key, std::make_pair(first, int64_t(blocks.size()))).first;
    }
    b.add_object(MPT_OBJ_VOXEL, it->second.first, it->second.second, xform, owner);
}

void voxel_grid::flatten(flat_builder& b, int32_t xform) const {
    std::vector<std::shared_ptr<voxel_grid>> one{std::make_shared<voxel_grid>(*this)};
    flatten_blocks(b, xform, this, one, this);
}

void voxel_block_set::flatten(flat_builder& b, int32_t xform) const { flatten_blocks(b, xform, this, blocks, this); }

std::shared_ptr<hittable> voxel_blocks(std::shared_ptr<const vox_model> model,
                                       std::shared_ptr<const voxel_palette> palette, int block,
                                       std::shared_ptr<material> edge_material, double edge_width) {
    if (block < 1) throw std::invalid_argument("voxel_blocks: block must be at least 1");
    auto set = std::make_shared<voxel_block_set>();
    for (int z0 = 0; z0 < model->nz; z0 += block)
        for (int y0 = 0; y0 < model->ny; y0 += block)
            for (int x0 = 0; x0 < model->nx; x0 += block) {
                // the box of the block's voxels: every surface lies on a voxel's faces
                int lo[3] = {x0 + block, y0 + block, z0 + block}, hi[3] = {x0, y0, z0};
                for (int z = z0; z < std::min(z0 + block, model->nz); z++)
                    for (int y = y0; y < std::min(y0 + block, model->ny); y++)
                        for (int x = x0; x < std::min(x0 + block, model->nx); x++)
                            if (palette->kind(model->at(x, y, z)) != voxel_kind::empty) {
                                lo[0] = std::min(lo[0], x), hi[0] = std::max(hi[0], x + 1);
                                lo[1] = std::min(lo[1], y), hi[1] = std::max(hi[1], y + 1);
                                lo[2] = std::min(lo[2], z), hi[2] = std::max(hi[2], z + 1);
                            }
                if (hi[0] > lo[0])
                    set->blocks.push_back(std::make_shared<voxel_grid>(model, palette, lo, hi, edge_material, edge_width));
            }
    return set;
}

const mpt_scene_desc& flat_builder::desc() {
    mpt_scene_desc& d = p->d;
    d.n_prims = int64_t(p->prim_kind.size());
    d.prim_kind = p->prim_kind.data();
    d.prim_material = p->prim_mat.data();
    d.prim_data = p->prim_data.data();
    d.prim_normals = p->prim_nrm.empty() ? nullptr : p->prim_nrm.data();
    d.n_objects = int32_t(p->obj_kind.size());
    d.object_kind = p->obj_kind.data();
    d.object_first = p->obj_first.data();
    d.object_count = p->obj_count.data();
    d.object_xform = p->obj_xform.data();
    d.object_id_base = p->obj_id.data();
    d.object_density = p->obj_density.data();
    d.object_material = p->obj_mat.data();
    d.object_boundary_kind = p->obj_bkind.data();
    d.n_xforms = int32_t(p->xforms.size());
    d.xforms = p->xforms.data();
    d.n_materials = int32_t(p->materials.size());
    d.materials = p->materials.data();
    d.n_textures = int32_t(p->textures.size());
    d.textures = p->textures.data();
    d.texels = p->texels.empty() ? nullptr : p->texels.data();
    d.n_texel_bytes = int64_t(p->texels.size());
    d.texels_hdr = p->texels_hdr.empty() ? nullptr : p->texels_hdr.data();
    d.n_texels_hdr = int64_t(p->texels_hdr.size());
    d.prim_uvs = p->prim_uv.empty() ? nullptr : p->prim_uv.data();
    d.has_ground = p->has_ground ? 1 : 0;
    d.ground_material = p->ground_mat;
    d.ground_height = p->ground_height;
    d.object_alpha = p->any_alpha ? p->obj_alpha.data() : nullptr;
    d.n_fogs = int32_t(p->fog_material.size());
    d.fog_params = p->fog_params.empty() ? nullptr : p->fog_params.data();
    d.fog_material = p->fog_material.empty() ? nullptr : p->fog_material.data();
    d.n_lights = int32_t(p->light_object.size());
    d.light_object = p->light_object.data();
    d.env_rgb = p->env_rgb.empty() ? nullptr : p->env_rgb.data();
    d.object_flags = p->any_flags ? p->obj_flags.data() : nullptr;
    d.n_vox_models = int32_t(p->vox_models.size());
    d.vox_models = p->vox_models.empty() ? nullptr : p->vox_models.data();
    d.vox_cells = p->vox_cells.empty() ? nullptr : p->vox_cells.data();
    d.n_vox_cells = int64_t(p->vox_cells.size());
    return d;
}

// ------------------------------------------------------------------ camera
mpt_camera camera::to_abi() const {
    mpt_camera c;
    std::memset(&c, 0, sizeof(c));
    c.aspect_ratio = aspect_ratio;
    c.image_width = image_width;
    c.samples_per_pixel = samples_per_pixel;
    c.max_depth = max_depth;
    c.max_diffuse = max_diffuse;
    c.max_specular = max_specular;
    c.max_scatter = max_scatter;
    c.transparent_shadows = transparent_shadows;
    c.vfov = vfov;
    for (int k = 0; k < 3; k++) {
        c.lookfrom[k] = lookfrom[k];
        c.lookat[k] = lookat[k];
        c.vup[k] = vup[k];
        c.background[k] = background[k];
    }
    c.use_sky = use_sky;
    c.collect_features = collect_features;
    c.ortho_height = ortho_height;
    c.defocus_angle = defocus_angle;
    c.focus_dist = focus_dist;
    c.blades = blades;
    c.sun_disk_hidden = !sun_disk_visible;
    c.blade_rotation = blade_rotation;
    return c;
}

gpu_scene::gpu_scene(const hittable& world, const hittable_list& lights, const std::shared_ptr<environment_light>& env) {
    flat_builder b;
    b.add_world(world);
    b.add_lights(lights, env);
    int rc = mpt_scene_create(&b.desc(), &handle);
    if (rc != MPT_OK) throw std::runtime_error(std::string("mpt_scene_create: ") + mpt_last_error());
}
gpu_scene::~gpu_scene() { mpt_scene_destroy(handle); }

namespace {
struct pass_ctx {
    const camera::pass_callback* cb;
    mpt_scene* scene;
    size_t npix;
    bool features;
};
frame to_frame(const std::vector<double>& r, const std::vector<double>& a, const std::vector<double>& n,
               const std::vector<double>& v, size_t npix, bool features) {
    frame f;
    f.radiance.resize(npix);
    for (size_t k = 0; k < npix; k++) f.radiance[k] = color(r[3 * k], r[3 * k + 1], r[3 * k + 2]);
    if (!features) return f;
    f.albedo.resize(npix);
    f.normal.resize(npix);
    f.variance.assign(v.begin(), v.end());
    for (size_t k = 0; k < npix; k++) {
        f.albedo[k] = color(a[3 * k], a[3 * k + 1], a[3 * k + 2]);
        f.normal[k] = color(n[3 * k], n[3 * k + 1], n[3 * k + 2]);
    }
    return f;
}
void pass_trampoline(int32_t done, void* user) {
    auto* c = static_cast<pass_ctx*>(user);
    std::vector<double> r(3 * c->npix), a, n, v;
    if (c->features) {
        a.resize(3 * c->npix);
        n.resize(3 * c->npix);
        v.resize(c->npix);
    }
    if (mpt_read_frame(c->scene, r.data(), c->features ? a.data() : nullptr, c->features ? n.data() : nullptr,
                       c->features ? v.data() : nullptr) == MPT_OK)
        (*c->cb)(done, to_frame(r, a, n, v, c->npix, c->features));
}
}  // namespace

frame camera::render(const hittable& world, const hittable_list& lights, int save_every, const pass_callback& on_pass) {
    image_height = int(image_width / aspect_ratio);
    image_height = image_height < 1 ? 1 : image_height;
    // camera::sun is added to every escaping ray (camera.h:526-527) and the
    // lights sample it (camera.h:45); the core models the two together, so
    // they must name the same sun
    const sun_light* listed = nullptr;
    for (const auto& o : lights.objects)
        if (auto s = dynamic_cast<const sun_light*>(o.get())) listed = s;
    if (sun && listed != sun.get())
        throw std::runtime_error("camera::sun must also be in the lights, and be the only sun there (camera.h:45)");
    if (!sun && listed)
        throw std::runtime_error("a sun_light in the lights needs camera::sun set to it: escaping rays add "
                                 "camera::sun (camera.h:526-527)");
    gpu_scene gs(world, lights, environment);
    mpt_camera c = to_abi();
    mpt_render_opts o{};
    o.seed = seed;
    o.save_every = save_every;
    size_t npix = size_t(image_width) * image_height;
    std::vector<double> r(3 * npix), a, n, v;
    if (collect_features) {
        a.resize(3 * npix);
        n.resize(3 * npix);
        v.resize(npix);
    }
    pass_ctx ctx{&on_pass, gs.get(), npix, collect_features};
    int rc = mpt_render(gs.get(), &c, &o, r.data(), collect_features ? a.data() : nullptr,
                        collect_features ? n.data() : nullptr, collect_features ? v.data() : nullptr,
                        on_pass ? pass_trampoline : nullptr, &ctx, nullptr);
    if (rc != MPT_OK) throw std::runtime_error(std::string("mpt_render: ") + mpt_last_error());
    return to_frame(r, a, n, v, npix, collect_features);
}

std::vector<color> denoiser::run(const frame& f, int width, int height) const {
    const size_t n = size_t(width > 0 ? width : 0) * size_t(height > 0 ? height : 0);
    if (f.radiance.size() != n || f.albedo.size() != n || f.normal.size() != n || f.variance.size() != n)
        throw std::runtime_error("denoiser::run: the frame needs radiance, albedo, normal and variance for every pixel"
                                 " (render with collect_features)");
    std::vector<double> r(3 * n), a(3 * n), nn(3 * n), out(3 * n);
    for (size_t k = 0; k < n; k++)
        for (int c = 0; c < 3; c++) {
            r[3 * k + c] = f.radiance[k][c];
            a[3 * k + c] = f.albedo[k][c];
            nn[3 * k + c] = f.normal[k][c];
        }
    mpt_denoise_opts o;
    o.iterations = iterations;
    o.pad = 0;
    o.sigma_luminance = sigma_luminance;
    o.sigma_normal = sigma_normal;
    o.sigma_albedo = sigma_albedo;
    if (mpt_denoise(r.data(), a.data(), nn.data(), f.variance.data(), width, height, &o, out.data()) != MPT_OK)
        throw std::runtime_error(std::string("mpt_denoise: ") + mpt_last_error());
    std::vector<color> res(n);
    for (size_t k = 0; k < n; k++) res[k] = color(out[3 * k], out[3 * k + 1], out[3 * k + 2]);
    return res;
}

// ------------------------------------------------------------------ display
bool display_settings::set_tonemapper(const std::string& name) {  // tonemap.h:18-25
    if (name == "none") op = tonemapper::none;
    else if (name == "reinhard") op = tonemapper::reinhard;
    else if (name == "aces") op = tonemapper::aces;
    else return false;
    return true;
}

mpt_display display_settings::to_abi() const {
    mpt_display d;
    mpt_display_defaults(&d);
    d.exposure = exposure;
    d.tonemap = op == tonemapper::reinhard ? MPT_TONEMAP_REINHARD : (op == tonemapper::aces ? MPT_TONEMAP_ACES : MPT_TONEMAP_NONE);
    d.bloom = post.bloom;
    d.bloom_threshold = post.bloom_threshold;
    d.bloom_radius = post.bloom_radius;
    d.vignette = post.vignette;
    return d;
}

namespace {
std::vector<double> flatten_pixels(const std::vector<color>& pixels) {
    std::vector<double> r(3 * pixels.size());
    for (size_t k = 0; k < pixels.size(); k++)
        for (int c = 0; c < 3; c++) r[3 * k + c] = pixels[k][c];
    return r;
}
}  // namespace

std::vector<color> postprocess(const std::vector<color>& pixels, int width, int height, const post_settings& post) {
    if (pixels.size() != size_t(width > 0 ? width : 0) * size_t(height > 0 ? height : 0))
        throw std::runtime_error("postprocess: the image needs width * height pixels");
    display_settings d;
    d.post = post;
    const mpt_display abi = d.to_abi();
    std::vector<double> r = flatten_pixels(pixels), out(r.size());
    if (mpt_postprocess(r.data(), width, height, &abi, out.data()) != MPT_OK)
        throw std::runtime_error(std::string("mpt_postprocess: ") + mpt_last_error());
    std::vector<color> res(pixels.size());
    for (size_t k = 0; k < res.size(); k++) res[k] = color(out[3 * k], out[3 * k + 1], out[3 * k + 2]);
    return res;
}

bool write_image(const std::string& path, int width, int height, const std::vector<color>& pixels,
                 const display_settings& display) {
    if (pixels.size() != size_t(width > 0 ? width : 0) * size_t(height > 0 ? height : 0)) {
        std::fprintf(stderr, "Error: write_image needs width * height pixels\n");
        return false;
    }
    const std::vector<double> r = flatten_pixels(pixels);
    const mpt_display abi = display.to_abi();
    if (mpt_write_image(path.empty() ? nullptr : path.c_str(), width, height, r.data(), &abi) != MPT_OK) {
        std::fprintf(stderr, "Error: %s\n", mpt_last_error());
        return false;
    }
    return true;
}

// ------------------------------------------------------------------ OBJ / MTL (obj_loader.h)
namespace {

// load_mtl (obj_loader.h:61-168): Kd -> lambertian, a non-black Ks -> a
// principled layer with roughness from the Phong exponent Ns and Ni, Ke ->
// an emitter; map_Kd replaces Kd, norm (or map_Bump, bump, map_Disp, disp
// naming a color image) bends the normal. Textures wrap around; one that
// cannot be loaded is skipped with a warning and the plain color stays. The
// alpha mask (obj_loader.h:169-172): map_d when it is not fully opaque, else
// without map_d the diffuse texture when it has a non-opaque alpha channel.
bool load_mtl(const std::string& path, std::map<std::string, std::shared_ptr<material>>& result,
              std::map<std::string, std::shared_ptr<image_texture>>& masks, std::string& error,
              bool use_normal_maps) {
    std::ifstream file(path);
    if (!file) {
        error = "cannot open '" + path + "'";
        return false;
    }
    auto slash = path.find_last_of('/');
    std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    // images shared between materials are loaded once per decoding (sRGB colors, linear normals)
    std::map<std::pair<std::string, bool>, std::shared_ptr<image_texture>> images;
    auto load_image = [&](const std::string& name, bool linear) -> std::shared_ptr<image_texture> {
        const std::string full = name[0] == '/' ? name : dir + name;
        auto it = images.find({full, linear});
        if (it != images.end()) return it->second;
        auto image = std::make_shared<image_texture>(full, true, linear);
        if (!image->loaded()) {
            std::fprintf(stderr, "Warning: %s: cannot load texture '%s'\n", path.c_str(), full.c_str());
            image = nullptr;
        }
        images[{full, linear}] = image;
        return image;
    };
    struct entry {
        color kd{0.8, 0.8, 0.8}, ks{0, 0, 0}, ke{0, 0, 0};
        double ns = 0, ni = 1.5;
        std::shared_ptr<image_texture> map_kd, map_d, normal_map;
    };
    std::vector<std::pair<std::string, entry>> entries;
    std::string line;
    while (std::getline(file, line)) {
        std::istringstream in(line);
        std::string keyword;
        if (!(in >> keyword) || keyword[0] == '#') continue;
        if (keyword == "newmtl") {
            std::string name;
            in >> name;
            entries.push_back({name, entry()});
            continue;
        }
        if (entries.empty()) continue;
        entry& e = entries.back().second;
        auto read_color = [&](color& c) {  // one value is grey
            double r, g, b;
            if (in >> r) {
                c = color(r, r, r);
                if (in >> g >> b) c = color(r, g, b);
            }
        };
        auto texture_file = [&]() {  // options come first: the file is the last word
            std::string word, last;
            while (in >> word) last = word;
            return last;
        };
        if (keyword == "Kd") read_color(e.kd);
        else if (keyword == "Ks") read_color(e.ks);
        else if (keyword == "Ke") read_color(e.ke);
        else if (keyword == "Ns") in >> e.ns;
        else if (keyword == "Ni") in >> e.ni;
        else if (keyword == "map_Kd") {
            std::string f = texture_file();
            if (!f.empty()) e.map_kd = load_image(f, false);
        } else if (keyword == "map_d") {
            std::string f = texture_file();
            if (!f.empty()) e.map_d = load_image(f, false);
        } else if (keyword == "norm" || keyword == "map_Bump" || keyword == "map_bump" || keyword == "bump" ||
                   keyword == "map_Disp" || keyword == "disp") {
            std::string f = texture_file();
            if (f.empty() || !use_normal_maps) continue;
            auto image = load_image(f, true);
            if (image && keyword != "norm" && image->greyscale()) {
                std::fprintf(stderr, "Warning: %s: '%s' is a height map; only normal maps are supported\n", path.c_str(),
                             f.c_str());
                continue;
            }
            if (image) e.normal_map = image;
        }
    }
    for (auto& [name, e] : entries) {
        std::shared_ptr<texture> base =
            e.map_kd ? std::static_pointer_cast<texture>(e.map_kd) : std::make_shared<solid_color>(e.kd);
        std::shared_ptr<material> m;
        if (e.ke.length_squared() > 0) {
            m = std::make_shared<diffuse_light>(e.ke);
        } else if (e.ks.length_squared() > 0) {
            // a Phong exponent n spreads like a microfacet alpha of sqrt(2 / (n + 2)); roughness is sqrt(alpha)
            double roughness = std::sqrt(std::sqrt(2 / (std::fmax(e.ns, 0.0) + 2)));
            double ior = e.ni > 1 ? e.ni : 1.5;
            m = std::make_shared<principled>(base, 0.0, roughness, ior);
        } else {
            m = std::make_shared<lambertian>(base);
        }
        if (e.normal_map && e.ke.length_squared() == 0) m = std::make_shared<normal_mapped>(m, e.normal_map);
        std::shared_ptr<image_texture> mask;
        if (e.map_d && !e.map_d->fully_opaque())
            mask = e.map_d;
        else if (!e.map_d && e.map_kd && (e.map_kd->channels == 2 || e.map_kd->channels == 4) &&
                 !e.map_kd->fully_opaque())
            mask = e.map_kd;
        result[name] = m;
        masks[name] = mask;
    }
    return true;
}

}  // namespace

bool load_obj(const std::string& path, const std::map<std::string, std::shared_ptr<material>>& materials,
              std::shared_ptr<material> default_mat, bool smooth, std::vector<obj_part>& parts, std::string& error,
              bool use_normal_maps) {
    std::ifstream file(path);
    if (!file) {
        error = "cannot open '" + path + "'";
        return false;
    }
    std::vector<point3> positions;
    std::vector<vec3> normals;
    std::vector<vec3> uvs;  // (u, v, 0)
    struct face {
        int v[3], n[3], t[3];
    };
    // faces per material and alpha mask (obj_loader.h:196-200), in the order of their first face
    struct group {
        std::shared_ptr<material> mat;
        std::shared_ptr<image_texture> mask;
        std::vector<face> faces;
    };
    std::vector<group> groups;
    std::map<std::pair<const material*, const image_texture*>, size_t> group_of;
    std::shared_ptr<material> current = default_mat;
    std::shared_ptr<image_texture> current_mask;
    auto slash = path.find_last_of('/');
    std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    std::map<std::string, std::shared_ptr<material>> file_materials;
    std::map<std::string, std::shared_ptr<image_texture>> file_masks;
    auto argument = [](const char* q) {  // the rest of the line without surrounding blanks
        std::string arg(q);
        arg.erase(0, arg.find_first_not_of(" \t"));
        arg.erase(arg.find_last_not_of(" \t\r") + 1);
        return arg;
    };
    auto resolve = [](long index, size_t count) -> int {  // 1-based, or negative from the end; -1 if absent
        if (index > 0 && size_t(index) <= count) return int(index - 1);
        if (index < 0 && size_t(-index) <= count) return int(long(count) + index);
        return -1;
    };
    std::string line;
    int line_number = 0;
    while (std::getline(file, line)) {
        line_number++;
        const char* q = line.c_str();
        while (*q == ' ' || *q == '\t') q++;
        auto read_vec = [&](const char* r, int n) {
            double x[3] = {0, 0, 0};
            char* end;
            for (int k = 0; k < n; k++) {
                x[k] = std::strtod(r, &end);
                r = end;
            }
            return vec3(x[0], x[1], x[2]);
        };
        const size_t at = size_t(q - line.c_str());
        if (q[0] == 'v' && (q[1] == ' ' || q[1] == '\t')) {
            positions.push_back(read_vec(q + 2, 3));
        } else if (q[0] == 'v' && q[1] == 'n') {
            normals.push_back(read_vec(q + 2, 3));
        } else if (q[0] == 'v' && q[1] == 't') {
            uvs.push_back(read_vec(q + 2, 2));
        } else if (q[0] == 'f' && (q[1] == ' ' || q[1] == '\t')) {
            std::vector<int> vs, ns, ts;
            const char* r = q + 2;
            while (true) {
                while (*r == ' ' || *r == '\t' || *r == '\r') r++;
                if (*r == '\0') break;
                char* end;
                long v = std::strtol(r, &end, 10), t = 0, n = 0;
                if (end == r) {
                    error = path + ":" + std::to_string(line_number) + ": bad face";
                    return false;
                }
                r = end;
                if (*r == '/') {
                    r++;
                    if (*r != '/') {
                        t = std::strtol(r, &end, 10);
                        r = end;
                    }
                    if (*r == '/') {
                        r++;
                        n = std::strtol(r, &end, 10);
                        r = end;
                    }
                }
                int vi = resolve(v, positions.size());
                if (vi < 0) {
                    error = path + ":" + std::to_string(line_number) + ": vertex index out of range";
                    return false;
                }
                vs.push_back(vi);
                ts.push_back(resolve(t, uvs.size()));
                ns.push_back(resolve(n, normals.size()));
            }
            for (size_t k = 1; k + 1 < vs.size(); k++) {  // a fan; degenerate triangles are skipped
                const size_t idx[3] = {0, k, k + 1};
                face f;
                for (int c = 0; c < 3; c++) {
                    f.v[c] = vs[idx[c]];
                    f.n[c] = ns[idx[c]];
                    f.t[c] = ts[idx[c]];
                }
                const point3& p0 = positions[f.v[0]];
                if (cross(positions[f.v[1]] - p0, positions[f.v[2]] - p0).near_zero()) continue;
                auto g = group_of.find({current.get(), current_mask.get()});
                if (g == group_of.end()) {
                    g = group_of.emplace(std::make_pair(current.get(), current_mask.get()), groups.size()).first;
                    groups.push_back({current, current_mask, {}});
                }
                groups[g->second].faces.push_back(f);
            }
        } else if (line.compare(at, 6, "usemtl") == 0) {
            std::string name = argument(q + 6);
            current_mask = nullptr;  // a scene material or the default has no mask
            if (auto it = materials.find(name); it != materials.end()) {
                current = it->second;
            } else if (auto jt = file_materials.find(name); jt != file_materials.end()) {
                current = jt->second;
                current_mask = file_masks[name];
            } else {
                current = default_mat;
            }
        } else if (line.compare(at, 6, "mtllib") == 0) {
            std::istringstream names(argument(q + 6));
            std::string name, mtl_error;
            while (names >> name) {
                const std::string full = name[0] == '/' ? name : dir + name;
                if (!load_mtl(full, file_materials, file_masks, mtl_error, use_normal_maps)) {
                    if (mtl_error.find("cannot open") == 0)
                        std::fprintf(stderr, "Warning: %s: %s\n", path.c_str(), mtl_error.c_str());
                    else {
                        error = mtl_error;
                        return false;
                    }
                }
            }
        }
    }
    if (smooth && normals.empty()) {  // area-weighted vertex normals (obj_loader.h:311-324)
        normals.assign(positions.size(), vec3(0, 0, 0));
        for (auto& gr : groups)
            for (auto& f : gr.faces) {
                const point3& p0 = positions[f.v[0]];
                vec3 n = cross(positions[f.v[1]] - p0, positions[f.v[2]] - p0);
                for (int c = 0; c < 3; c++) {
                    normals[f.v[c]] = normals[f.v[c]] + n;
                    f.n[c] = f.v[c];
                }
            }
    }
    for (auto& gr : groups) {
        const auto& mat = gr.mat;
        std::vector<int32_t> idx, nidx, tidx;
        bool any_normals = false, any_uvs = false;
        for (const auto& f : gr.faces)
            for (int c = 0; c < 3; c++) {
                idx.push_back(f.v[c]);
                nidx.push_back(f.n[c]);
                tidx.push_back(f.t[c]);
                any_normals |= f.n[c] >= 0;
                any_uvs |= f.t[c] >= 0;
            }
        auto m = std::make_shared<mesh>(positions, std::move(idx), mat);
        if (any_normals) {
            m->normals = normals;
            m->normal_indices = std::move(nidx);
        }
        if (any_uvs) {  // a triangle with all three uses them (mesh.h:75-81)
            m->uvs = uvs;
            m->uv_indices = std::move(tidx);
        }
        m->alpha_mask = gr.mask;
        parts.push_back({mat, m});
    }
    if (parts.empty()) {
        error = "'" + path + "' has no faces";
        return false;
    }
    return true;
}

// ------------------------------------------------------------------ .scene loader
namespace {

// Reads a PFM image (the reference's own output format, image_io.h:49-62)
// into rows from the top.
bool read_pfm(const std::string& path, int& w, int& h, std::vector<color>& px) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::string magic;
    double scale;
    f >> magic >> w >> h >> scale;
    f.get();
    if (magic != "PF" || w <= 0 || h <= 0) return false;
    std::vector<float> data(size_t(w) * h * 3);
    f.read(reinterpret_cast<char*>(data.data()), std::streamsize(data.size() * sizeof(float)));
    if (!f) return false;
    px.resize(size_t(w) * h);
    for (int j = 0; j < h; j++)  // PFM stores the bottom row first
        for (int i = 0; i < w; i++) {
            const float* s = &data[(size_t(h - 1 - j) * w + i) * 3];
            px[size_t(j) * w + i] = color(s[0], s[1], s[2]);
        }
    return true;
}

// Radiance RGBE (.hdr), as environment_light reads it through stb_image
// 2.x (environment.h:30-41): the "#?RADIANCE" / "#?RGBE" header with
// FORMAT=32-bit_rle_rgbe, a "-Y h +X w" resolution line, scanlines flat or
// new-style run-length encoded (per component: count > 128 repeats the next
// byte count - 128 times, else count literal bytes). A pixel is
// float(m) * 2^(e - 136) per channel in float arithmetic, 0 when e = 0.
bool read_hdr(const std::string& path, int& w, int& h, std::vector<color>& px) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::string line;
    if (!std::getline(f, line) || (line != "#?RADIANCE" && line != "#?RGBE")) return false;
    bool format = false;
    while (std::getline(f, line) && !line.empty())
        if (line == "FORMAT=32-bit_rle_rgbe") format = true;
    if (!format || !std::getline(f, line)) return false;
    char ya[3] = {0}, xa[3] = {0};
    if (std::sscanf(line.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || std::string(ya) != "-Y" ||
        std::string(xa) != "+X" || w <= 0 || h <= 0)
        return false;
    std::vector<unsigned char> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<unsigned char> rgbe(size_t(w) * h * 4);
    size_t at = 0;
    auto flat_rest = [&](size_t from_pixel) {  // the rest of the image as plain RGBE
        const size_t need = (size_t(w) * h - from_pixel) * 4;
        if (at + need > data.size()) return false;
        std::copy(data.begin() + long(at), data.begin() + long(at + need), rgbe.begin() + long(from_pixel * 4));
        return true;
    };
    if (w < 8 || w >= 32768) {
        if (!flat_rest(0)) return false;
    } else {
        for (int j = 0; j < h; j++) {
            if (at + 4 > data.size()) return false;
            const unsigned char c1 = data[at], c2 = data[at + 1], len = data[at + 2];
            if (c1 != 2 || c2 != 2 || (len & 0x80)) {  // not run-length encoded: the file is flat from here
                if (!flat_rest(size_t(j) * w)) return false;
                break;
            }
            if (((int(len) << 8) | data[at + 3]) != w) return false;
            at += 4;
            for (int c = 0; c < 4; c++) {
                int i = 0;
                while (i < w) {
                    if (at >= data.size()) return false;
                    int count = data[at++];
                    if (count > 128) {
                        count -= 128;
                        if (count > w - i || at >= data.size()) return false;
                        const unsigned char v = data[at++];
                        for (int k = 0; k < count; k++) rgbe[(size_t(j) * w + i++) * 4 + c] = v;
                    } else {
                        if (count == 0 || count > w - i || at + size_t(count) > data.size()) return false;
                        for (int k = 0; k < count; k++) rgbe[(size_t(j) * w + i++) * 4 + c] = data[at++];
                    }
                }
            }
        }
    }
    px.resize(size_t(w) * h);
    for (size_t k = 0; k < px.size(); k++) {
        const unsigned char* e = &rgbe[4 * k];
        if (e[3] == 0) {
            px[k] = color(0, 0, 0);
            continue;
        }
        const float f1 = float(std::ldexp(1.0f, int(e[3]) - (128 + 8)));
        px[k] = color(float(e[0]) * f1, float(e[1]) * f1, float(e[2]) * f1);
    }
    return true;
}

}  // namespace

bool load_scene(const std::string& path, scene& result, std::string* err_out) {
    auto& world = result.world;
    auto& lights = result.lights;
    auto& cam = result.cam;
    std::string errbuf;
    std::string& err = err_out ? *err_out : errbuf;

    std::ifstream file(path);
    if (!file) {
        err = "cannot open scene file '" + path + "'";
        return false;
    }
    int line_number = 0;
    auto fail = [&](const std::string& message) {
        err = path + ":" + std::to_string(line_number) + ": " + message;
        return false;
    };
    std::map<std::string, std::shared_ptr<material>> materials;
    std::map<std::string, std::shared_ptr<texture>> textures;
    auto slash = path.find_last_of('/');
    std::string scene_dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    auto default_material = std::make_shared<lambertian>(color(0.5, 0.5, 0.5));
    bool use_normal_maps = true;
    // the ground (scene.h:437-439, 596-611), made at the end of the file
    bool ground = false;
    double ground_height = 0;
    std::shared_ptr<material> ground_material = default_material;
    // meshes loaded once per (file, smooth, material, normal maps) (scene.h:1167-1192)
    std::map<std::tuple<std::string, bool, const material*, bool>, std::vector<obj_part>> mesh_cache;

    auto read_color_or_texture = [&](std::istringstream& in, const std::string& first, std::shared_ptr<texture>& tex,
                                     std::string& error) {
        if (first == "texture") {
            std::string name;
            if (!(in >> name)) return false;
            auto it = textures.find(name);
            if (it == textures.end()) {
                error = "unknown texture '" + name + "'";
                return false;
            }
            tex = it->second;
            return true;
        }
        std::istringstream rgb(first);
        double r, g, b;
        if (!(rgb >> r) || !(in >> g >> b)) return false;
        tex = std::make_shared<solid_color>(color(r, g, b));
        return true;
    };

    auto read_tail = [&](std::istringstream& in, std::shared_ptr<material>& mat, transform& xform, bool& transformed,
                         std::string& error) {
        mat = default_material;
        xform = transform();
        transformed = false;
        auto is_transform = [](const std::string& t) {
            return t == "translate" || t == "scale" || t == "rotate" || t == "rotate_x" || t == "rotate_y" ||
                   t == "rotate_z";
        };
        std::string token;
        if (!(in >> token)) return true;
        if (!is_transform(token)) {
            auto it = materials.find(token);
            if (it == materials.end()) {
                error = "unknown material '" + token + "'";
                return false;
            }
            mat = it->second;
            if (!(in >> token)) return true;
        }
        do {
            transform step;
            bool ok = true;
            if (token == "translate") {
                vec3 d;
                ok = bool(in >> d[0] >> d[1] >> d[2]);
                step = transform::translate(d);
            } else if (token == "rotate_x" || token == "rotate_y" || token == "rotate_z") {
                double degrees;
                ok = bool(in >> degrees);
                vec3 axis(token == "rotate_x", token == "rotate_y", token == "rotate_z");
                step = transform::rotate(degrees, axis);
            } else if (token == "rotate") {
                double degrees;
                vec3 axis;
                ok = bool(in >> degrees >> axis[0] >> axis[1] >> axis[2]) && axis.length_squared() > 0;
                if (ok) step = transform::rotate(degrees, axis);
            } else if (token == "scale") {
                vec3 f;
                ok = bool(in >> f[0]);
                if (ok) {
                    std::streampos mark = in.tellg();
                    if (in >> f[1]) {
                        ok = bool(in >> f[2]);
                    } else {
                        in.clear();
                        in.seekg(mark);
                        f = vec3(f[0], f[0], f[0]);
                    }
                }
                ok = ok && f[0] != 0 && f[1] != 0 && f[2] != 0;
                step = transform::scale(f);
            } else {
                error = "unexpected value '" + token + "' (expected a transform)";
                return false;
            }
            if (!ok) {
                error = "bad values for '" + token + "'";
                return false;
            }
            xform = step.then_after(xform);
            transformed = true;
        } while (in >> token);
        return true;
    };

    auto add_object = [&](std::shared_ptr<hittable> object, const std::shared_ptr<material>& mat,
                          const transform& xform, bool transformed, std::string& error) {
        bool emissive = bool(std::dynamic_pointer_cast<diffuse_light>(mat));
        if (transformed) {
            if (emissive && !xform.is_similarity()) {
                error = "lights can only be moved, rotated and scaled uniformly";
                return false;
            }
            object = std::make_shared<instance>(object, xform);
        }
        world.add(object);
        if (emissive) lights.add(object);
        return true;
    };

    bool sky_uniform = false;
    color sky_radiance;
    // the sun line's settings, applied at the end of the file (scene.h:422-432)
    bool sun_on = false, sun_disk = true, sky_atmosphere = false;
    sky_air air;
    double sun_elevation = 45, sun_azimuth = 45, sun_intensity = 3, sun_angle = 4;
    color sun_tint(1, 1, 1);
    double fog_density = 0;
    std::shared_ptr<material> fog_material;
    std::string line;
    while (std::getline(file, line)) {
        line_number++;
        auto comment = line.find('#');
        if (comment != std::string::npos) line.erase(comment);
        std::istringstream in(line);
        std::string keyword;
        if (!(in >> keyword)) continue;

        if (keyword == "image_width") {
            if (!(in >> cam.image_width) || cam.image_width < 1) return fail("expected 'image_width <positive integer>'");
        } else if (keyword == "aspect_ratio") {
            double num, den = 1.0;
            if (!(in >> num)) return fail("expected 'aspect_ratio <w>/<h>' or 'aspect_ratio <number>'");
            if (in.peek() == '/') {
                in.get();
                if (!(in >> den)) return fail("expected 'aspect_ratio <w>/<h>'");
            }
            if (num <= 0 || den <= 0) return fail("aspect_ratio must be positive");
            cam.aspect_ratio = num / den;
        } else if (keyword == "samples_per_pixel") {
            if (!(in >> cam.samples_per_pixel) || cam.samples_per_pixel < 1)
                return fail("expected 'samples_per_pixel <positive integer>'");
        } else if (keyword == "max_depth") {
            if (!(in >> cam.max_depth) || cam.max_depth < 1) return fail("expected 'max_depth <positive integer>'");
        } else if (keyword == "camera") {
            std::string key;
            while (in >> key) {
                bool ok = true;
                if (key == "lookfrom") ok = bool(in >> cam.lookfrom[0] >> cam.lookfrom[1] >> cam.lookfrom[2]);
                else if (key == "lookat") ok = bool(in >> cam.lookat[0] >> cam.lookat[1] >> cam.lookat[2]);
                else if (key == "vup") ok = bool(in >> cam.vup[0] >> cam.vup[1] >> cam.vup[2]);
                else if (key == "vfov") ok = bool(in >> cam.vfov) && cam.vfov > 0 && cam.vfov < 180;
                else if (key == "aperture") ok = bool(in >> cam.defocus_angle) && cam.defocus_angle >= 0;
                else if (key == "focus") ok = bool(in >> cam.focus_dist) && cam.focus_dist > 0;
                else if (key == "ortho") ok = bool(in >> cam.ortho_height) && cam.ortho_height >= 0;
                else if (key == "blades") ok = bool(in >> cam.blades) && (cam.blades == 0 || cam.blades >= 3);
                else if (key == "blade_rotation") ok = bool(in >> cam.blade_rotation);
                else return fail("unknown camera setting '" + key + "'");
                if (!ok) return fail("bad value for camera setting '" + key + "'");
            }
        } else if (keyword == "vfov") {
            if (!(in >> cam.vfov) || cam.vfov <= 0 || cam.vfov >= 180) return fail("expected 'vfov <degrees between 0 and 180>'");
        } else if (keyword == "aperture") {
            if (!(in >> cam.defocus_angle) || cam.defocus_angle < 0) return fail("expected 'aperture <defocus angle in degrees>'");
        } else if (keyword == "focus") {
            if (!(in >> cam.focus_dist) || cam.focus_dist <= 0) return fail("expected 'focus <positive distance>'");
        } else if (keyword == "background") {
            double r, g, b;
            if (!(in >> r >> g >> b)) return fail("expected 'background <r> <g> <b>'");
            cam.use_sky = false;
            cam.background = color(r, g, b);
        } else if (keyword == "light_selection") {
            std::string mode;
            if (!(in >> mode) || (mode != "power" && mode != "uniform")) return fail("expected 'light_selection power|uniform'");
            lights.strategy = mode == "power" ? light_list::selection::power : light_list::selection::uniform;
        } else if (keyword == "environment") {
            std::string fname, key;
            double intensity = 1, rotation = 0;
            const char* usage = "expected 'environment <file.hdr|file.pfm> [intensity <s>] [rotate <degrees>]'";
            if (!(in >> fname)) return fail(usage);
            while (in >> key) {
                bool ok = (key == "intensity" && in >> intensity && intensity >= 0) || (key == "rotate" && in >> rotation);
                if (!ok) return fail(usage);
            }
            if (fname[0] != '/') fname = scene_dir + fname;
            int w = 0, h = 0;
            std::vector<color> px;
            const bool hdr = fname.size() >= 4 && fname.compare(fname.size() - 4, 4, ".hdr") == 0;
            if (!(hdr ? read_hdr(fname, w, h, px) : read_pfm(fname, w, h, px)))
                return fail("cannot load environment map '" + fname + "' (the GPU path reads .hdr and .pfm images)");
            if (cam.environment) return fail("only one environment is allowed");
            cam.environment = std::make_shared<environment_light>(w, h, std::move(px), intensity, rotation);
            lights.add(cam.environment);
        } else if (keyword == "sun") {  // scene.h:543-562
            const char* usage = "expected 'sun elevation <deg> azimuth <deg> [intensity <E>] [color <r g b>] "
                                "[angle <deg>] [disk on|off]' or 'sun off'";
            std::string key;
            sun_on = true;
            while (in >> key) {
                bool ok = true;
                if (key == "off") sun_on = false;
                else if (key == "elevation") ok = bool(in >> sun_elevation) && sun_elevation >= -90 && sun_elevation <= 90;
                else if (key == "azimuth") ok = bool(in >> sun_azimuth);
                else if (key == "intensity") ok = bool(in >> sun_intensity) && sun_intensity >= 0;
                else if (key == "color") ok = bool(in >> sun_tint[0] >> sun_tint[1] >> sun_tint[2]);
                else if (key == "angle") ok = bool(in >> sun_angle) && sun_angle > 0 && sun_angle < 90;
                else if (key == "disk") {
                    std::string value;
                    ok = bool(in >> value) && (value == "on" || value == "off");
                    sun_disk = value == "on";
                } else ok = false;
                if (!ok) return fail(usage);
            }
        } else if (keyword == "sky") {
            std::string type;
            if (!(in >> type)) return fail("expected 'sky uniform <r g b>' or 'sky gradient'");
            if (type == "uniform") {
                if (!(in >> sky_radiance[0] >> sky_radiance[1] >> sky_radiance[2])) return fail("expected 'sky uniform <r g b>'");
                sky_uniform = true;
                cam.use_sky = true;
                sky_atmosphere = false;
            } else if (type == "atmosphere") {  // scene.h:574-588
                const char* usage = "expected 'sky atmosphere [rayleigh <x>] [mie <x>] [mie_g <g>] [ozone <x>] [intensity <s>]'";
                std::string key;
                while (in >> key) {
                    bool ok = (key == "rayleigh" && in >> air.rayleigh && air.rayleigh >= 0) ||
                              (key == "mie" && in >> air.mie && air.mie >= 0) ||
                              (key == "mie_g" && in >> air.mie_g && air.mie_g > -1 && air.mie_g < 1) ||
                              (key == "ozone" && in >> air.ozone && air.ozone >= 0) ||
                              (key == "intensity" && in >> air.intensity && air.intensity >= 0);
                    if (!ok) return fail(usage);
                }
                sky_atmosphere = true;
                sky_uniform = false;
                cam.use_sky = true;
            } else if (type == "gradient") {
                sky_uniform = false;
                sky_atmosphere = false;
                cam.use_sky = true;
            } else {
                return fail("unknown sky type '" + type + "' (expected uniform, atmosphere or gradient)");
            }
        } else if (keyword == "bounces") {
            std::string key;
            while (in >> key) {
                bool ok = (key == "diffuse" && in >> cam.max_diffuse && cam.max_diffuse >= 0) ||
                          (key == "specular" && in >> cam.max_specular && cam.max_specular >= 0) ||
                          (key == "scatter" && in >> cam.max_scatter && cam.max_scatter >= 0);
                if (!ok) return fail("expected 'bounces [diffuse <n>] [specular <n>] [scatter <n>]'");
            }
        } else if (keyword == "transparent_shadows") {
            std::string value;
            if (!(in >> value) || (value != "on" && value != "off")) return fail("expected 'transparent_shadows on|off'");
            cam.transparent_shadows = value == "on";
        } else if (keyword == "bloom") {  // scene.h:667-678
            auto& post = result.display.post;
            const char* usage = "expected 'bloom <strength> [threshold <radiance>] [radius <fraction of width>]'";
            std::string key;
            if (!(in >> post.bloom) || post.bloom < 0) return fail(usage);
            while (in >> key) {
                bool ok = (key == "threshold" && in >> post.bloom_threshold && post.bloom_threshold >= 0) ||
                          (key == "radius" && in >> post.bloom_radius && post.bloom_radius > 0);
                if (!ok) return fail(usage);
            }
        } else if (keyword == "vignette") {  // scene.h:679-681
            if (!(in >> result.display.post.vignette) || result.display.post.vignette < 0)
                return fail("expected 'vignette <strength>'");
        } else if (keyword == "exposure") {  // scene.h:682-684
            if (!(in >> result.display.exposure)) return fail("expected 'exposure <stops>'");
        } else if (keyword == "tonemap") {  // scene.h:685-688
            std::string name;
            if (!(in >> name) || !result.display.set_tonemapper(name))
                return fail("expected 'tonemap none|reinhard|aces'");
        } else if (keyword == "ground") {  // scene.h:596-611
            std::string key;
            ground = true;
            while (in >> key) {
                if (key == "off") {
                    ground = false;
                } else if (key == "height") {
                    if (!(in >> ground_height)) return fail("expected 'ground [height <y>] [material]'");
                } else {
                    auto it = materials.find(key);
                    if (it == materials.end()) return fail("unknown material '" + key + "'");
                    ground_material = it->second;
                }
            }
        } else if (keyword == "normal_maps") {  // scene.h:689-693
            std::string state;
            if (!(in >> state) || (state != "on" && state != "off")) return fail("expected 'normal_maps on|off'");
            use_normal_maps = state == "on";
        } else if (keyword == "cyclorama") {  // scene.h:935-968
            const char* usage = "expected 'cyclorama [width <w>] [front <f>] [back <b>] [height <h>] [radius <r>]"
                                " [segments <n>] [material] [transforms]'";
            cyclorama_settings cyc;
            std::string key, error;
            std::streampos mark = in.tellg();
            while (in >> key) {  // settings first, as key value pairs
                bool ok = true;
                if (key == "width") ok = bool(in >> cyc.width) && cyc.width > 0;
                else if (key == "front") ok = bool(in >> cyc.front);
                else if (key == "back") ok = bool(in >> cyc.back);
                else if (key == "height") ok = bool(in >> cyc.height) && cyc.height > 0;
                else if (key == "radius") ok = bool(in >> cyc.radius) && cyc.radius > 0;
                else if (key == "segments") ok = bool(in >> cyc.segments) && cyc.segments > 0;
                else break;
                if (!ok) return fail(usage);
                mark = in.tellg();
            }
            if (cyc.front < -cyc.back) return fail("a cyclorama's front must be in front of its back (front >= -back)");
            in.clear();
            in.seekg(mark);
            std::shared_ptr<material> mat;
            transform xform;
            bool transformed;
            if (!read_tail(in, mat, xform, transformed, error)) return fail(error);
            if (std::dynamic_pointer_cast<diffuse_light>(mat))
                return fail("a cyclorama cannot be a light (mesh light sampling is linear in triangles)");
            if (!add_object(make_cyclorama(cyc, mat), mat, xform, transformed, error)) return fail(error);
        } else if (keyword == "softbox") {  // scene.h:969-1001
            const char* usage = "expected 'softbox center <x y z> aim <x y z> size <width> <height> [roll <degrees>]"
                                " <emissive material> [transforms]'";
            point3 center, aim;
            double width = 0, height = 0, roll = 0;
            bool has_center = false, has_aim = false;
            std::string key, error;
            std::streampos mark = in.tellg();
            while (in >> key) {
                bool ok = true;
                if (key == "center") ok = has_center = bool(in >> center[0] >> center[1] >> center[2]);
                else if (key == "aim") ok = has_aim = bool(in >> aim[0] >> aim[1] >> aim[2]);
                else if (key == "size") ok = bool(in >> width >> height) && width > 0 && height > 0;
                else if (key == "roll") ok = bool(in >> roll);
                else break;
                if (!ok) return fail(usage);
                mark = in.tellg();
            }
            if (!has_center || !has_aim || width <= 0 || (aim - center).near_zero()) return fail(usage);
            in.clear();
            in.seekg(mark);
            std::shared_ptr<material> mat;
            transform xform;
            bool transformed;
            if (!read_tail(in, mat, xform, transformed, error)) return fail(error);
            if (!std::dynamic_pointer_cast<diffuse_light>(mat)) return fail("a softbox needs an emissive material");
            if (!add_object(make_softbox(center, aim, width, height, roll, mat), mat, xform, transformed, error))
                return fail(error);
        } else if (keyword == "style") {  // scene.h:612-645: a preset of other lines, which later lines can change
            std::string name;
            if (!(in >> name) || (name != "voxel" && name != "classic"))
                return fail("expected 'style voxel' or 'style classic'");
            const bool voxel = name == "voxel";
            auto& post = result.display.post;
            result.display.op = voxel ? tonemapper::aces : tonemapper::none;
            post = post_settings();
            sun_on = false, sun_disk = true;
            sun_elevation = 45, sun_azimuth = 45, sun_intensity = 3, sun_angle = 4;
            sun_tint = color(1, 1, 1);
            sky_uniform = false;
            sky_atmosphere = voxel;
            air = sky_air();
            cam.use_sky = true;
            ground = voxel;
            ground_material = voxel ? std::make_shared<lambertian>(color(0.3, 0.3, 0.3)) : default_material;
            cam.max_diffuse = voxel ? 2 : 0;
            cam.max_specular = voxel ? 5 : 0;
            cam.max_scatter = voxel ? 5 : 0;
            cam.transparent_shadows = voxel;
            if (voxel) {
                post.bloom = 0.25;
                post.vignette = 0.25;
                sun_on = true;
                sun_elevation = 50, sun_azimuth = 50, sun_intensity = 4, sun_angle = 3;
            }
        } else if (keyword == "vox") {  // scene.h:1040-1128, MagicaVoxel models
            const char* usage = "expected 'vox <file.vox> [quads] [emission <scale>] [edges <width> [<r g b>]] [transforms]'";
            std::string file, error, token;
            if (!(in >> file)) return fail(usage);
            if (file[0] != '/') file = scene_dir + file;
            // options first, then the transforms
            bool quads = false;
            double emission = 1, edge_width = 0;
            color edge_color(0, 0, 0);
            std::vector<std::string> rest;
            while (in >> token) rest.push_back(token);
            size_t k = 0;
            for (; k < rest.size(); k++) {
                if (rest[k] == "quads") {
                    quads = true;
                } else if (rest[k] == "emission" && k + 1 < rest.size()) {
                    std::istringstream number(rest[++k]);
                    if (!(number >> emission) || emission < 0) return fail(usage);
                } else if (rest[k] == "edges" && k + 1 < rest.size()) {
                    std::istringstream number(rest[++k]);
                    if (!(number >> edge_width) || edge_width <= 0 || edge_width >= 0.5) return fail(usage);
                    if (k + 3 < rest.size()) {  // an optional colour
                        std::istringstream rgb(rest[k + 1] + ' ' + rest[k + 2] + ' ' + rest[k + 3]);
                        color c;
                        if (rgb >> c[0] >> c[1] >> c[2]) {
                            edge_color = c;
                            k += 3;
                        }
                    }
                } else {
                    break;
                }
            }
            std::string tail_text;
            for (; k < rest.size(); k++) tail_text += rest[k] + ' ';
            std::istringstream tail(tail_text);
            std::shared_ptr<material> mat;
            transform xform;
            bool transformed;
            if (!read_tail(tail, mat, xform, transformed, error)) return fail(error);
            if (mat != default_material) return fail("vox models take their materials from the file's palette");
            vox_file vox;
            if (!load_vox(file, vox, error)) return fail(error);
            auto palette = std::make_shared<voxel_palette>(make_voxel_palette(vox, emission));
            // per model: a grid of blocks (or with "quads" its visible faces), shared
            // by the instances that place it; emissive faces are lights either way,
            // and with a grid only sampled: the grid is what rays hit
            std::vector<std::shared_ptr<hittable>> geometry(vox.models.size());
            std::vector<std::shared_ptr<hittable_list>> emitters(vox.models.size());
            std::vector<std::shared_ptr<hittable_list>> face_lists(vox.models.size());
            for (size_t m = 0; m < vox.models.size(); m++) {
                face_lists[m] = std::make_shared<hittable_list>();
                emitters[m] = std::make_shared<hittable_list>();
                voxel_model_faces(vox.models[m], *palette, *face_lists[m], *emitters[m]);
                if (!quads)
                    geometry[m] = voxel_blocks(std::make_shared<vox_model>(vox.models[m]), palette, 8,
                                               edge_width > 0 ? std::make_shared<lambertian>(edge_color) : nullptr,
                                               edge_width);
            }
            // MagicaVoxel is z-up: its +z becomes +y (and +y becomes -z)
            const transform z_up = transform::rotate(-90, vec3(1, 0, 0));
            for (const auto& placed : vox.instances) {
                const transform to_world = xform.then_after(z_up.then_after(placed.xform));
                if (quads) {  // each face behind the instance's transform (the GPU path instances no BVH)
                    for (const auto& face : face_lists[size_t(placed.model)]->objects)
                        world.add(std::make_shared<instance>(face, to_world));
                } else {
                    world.add(std::make_shared<instance>(geometry[size_t(placed.model)], to_world));
                }
                if (!emitters[size_t(placed.model)]->empty()) {
                    if (!to_world.is_similarity())
                        return fail("a vox model with emissive voxels can only be scaled uniformly");
                    lights.add(std::make_shared<instance>(emitters[size_t(placed.model)], to_world));
                }
            }
        } else if (keyword == "mesh") {  // scene.h:1129-1197, Wavefront OBJ files
            std::string file, error;
            if (!(in >> file)) return fail("expected 'mesh <file.obj> [material] [transforms]'");
            if (file[0] != '/') file = scene_dir + file;
            std::string rest, first;
            std::getline(in, rest);
            std::istringstream peek(rest);
            const bool smooth = (peek >> first) && first == "smooth";
            std::istringstream tail(smooth ? rest.substr(rest.find("smooth") + 6) : rest);
            auto ends_with = [&](const std::string& suffix) {
                return file.size() >= suffix.size() && file.compare(file.size() - suffix.size(), suffix.size(), suffix) == 0;
            };
            const bool gltf = ends_with(".gltf") || ends_with(".glb");
            // then any number of "replace <material in the file> <scene material>" (glTF)
            std::map<std::string, std::shared_ptr<material>> replace;
            std::string replace_key, token;
            std::streampos mark = tail.tellg();
            while (tail >> token && token == "replace") {
                std::string from, to;
                if (!(tail >> from >> to)) return fail("expected 'replace <material in the file> <scene material>'");
                auto it = materials.find(to);
                if (it == materials.end()) return fail("unknown material '" + to + "'");
                replace[from] = it->second;
                replace_key += from + '\n' + to + '\n';
                mark = tail.tellg();
            }
            tail.clear();
            tail.seekg(mark);
            std::shared_ptr<material> mat;
            transform xform;
            bool transformed;
            if (!read_tail(tail, mat, xform, transformed, error)) return fail(error);
            if (!replace.empty() && mat != default_material)
                return fail("'replace' and a material for the whole mesh can't be used together");
            if (!gltf && !replace.empty())
                return fail("'replace' is for glTF files (OBJ files take scene materials by usemtl name)");
            auto key = std::make_tuple(file + '\n' + replace_key, smooth, mat.get(), use_normal_maps);
            auto cached = mesh_cache.find(key);
            if (cached == mesh_cache.end()) {
                std::vector<obj_part> parts;
                // glTF files bring their own materials: one named on the line replaces them all
                if (gltf ? !load_gltf(file, mat == default_material ? nullptr : mat, parts, error, use_normal_maps, replace)
                         : !load_obj(file, materials, mat, smooth, parts, error, use_normal_maps))
                    return fail(error);
                cached = mesh_cache.emplace(key, std::move(parts)).first;
            }
            for (const auto& part : cached->second)
                if (!add_object(part.geometry, part.mat, xform, transformed, error)) return fail(error);
        } else if (keyword == "texture") {
            std::string name, type;
            if (!(in >> name >> type)) return fail("expected 'texture <name> <type> <parameters...>'");
            std::shared_ptr<texture> tex;
            if (type == "solid") {
                double r, g, b;
                if (!(in >> r >> g >> b)) return fail("expected 'texture <name> solid <r> <g> <b>'");
                tex = std::make_shared<solid_color>(color(r, g, b));
            } else if (type == "checker") {
                double scale, r0, g0, b0, r1, g1, b1;
                if (!(in >> scale >> r0 >> g0 >> b0 >> r1 >> g1 >> b1) || scale <= 0)
                    return fail("expected 'texture <name> checker <scale> <r g b> <r g b>'");
                tex = std::make_shared<checker_texture>(scale, std::make_shared<solid_color>(color(r0, g0, b0)),
                                                        std::make_shared<solid_color>(color(r1, g1, b1)));
            } else if (type == "image") {  // scene.h:712-727
                std::string file, option;
                if (!(in >> file)) return fail("expected 'texture <name> image <path> [linear]'");
                bool linear = false;
                if (in >> option) {
                    if (option != "linear") return fail("expected 'texture <name> image <path> [linear]'");
                    linear = true;
                }
                if (file[0] != '/') file = scene_dir + file;
                auto image = std::make_shared<image_texture>(file, false, linear);
                if (!image->loaded()) return fail("cannot load image '" + file + "': " + image->error);
                tex = image;
            } else {
                return fail("unknown texture type '" + type + "'");
            }
            textures[name] = tex;
        } else if (keyword == "material") {
            std::string name, type;
            if (!(in >> name >> type)) return fail("expected 'material <name> <type> <parameters...>'");
            std::shared_ptr<material> mat;
            std::shared_ptr<texture> normal_map;  // lambertian and metal: an optional "normal_map <texture>" last
            auto read_texture_name = [&](std::shared_ptr<texture>& tex, std::string& error) {
                std::string tname;
                if (!(in >> tname)) {
                    error = "expected a texture name";
                    return false;
                }
                auto it = textures.find(tname);
                if (it == textures.end()) {
                    error = "unknown texture '" + tname + "'";
                    return false;
                }
                tex = it->second;
                return true;
            };
            auto read_normal_map = [&](std::string& error) {  // scene.h:740-749
                std::string key;
                if (!(in >> key)) return true;
                if (key != "normal_map") {
                    error = "unexpected value '" + key + "' (expected 'normal_map <texture>')";
                    return false;
                }
                return read_texture_name(normal_map, error);
            };
            // "<number>" in [0, 1] or "texture <name>" (scene.h:308-322)
            auto read_value_or_texture = [&](double& value, std::shared_ptr<texture>& tex, std::string& error) {
                std::string first;
                if (!(in >> first)) return false;
                if (first == "texture") return read_texture_name(tex, error);
                std::istringstream number(first);
                double x;
                if (!(number >> x) || x < 0 || x > 1) return false;
                value = x;
                tex = nullptr;
                return true;
            };
            if (type == "lambertian") {
                std::shared_ptr<texture> tex;
                std::string first, error;
                if (!(in >> first) || !read_color_or_texture(in, first, tex, error))
                    return fail(error.empty() ? "expected 'material <name> lambertian <r> <g> <b>'" : error);
                if (!read_normal_map(error)) return fail(error);
                mat = std::make_shared<lambertian>(tex);
            } else if (type == "metal") {  // scene.h:758-780
                const char* usage = "expected 'material <name> metal <r> <g> <b> <roughness>'"
                                    " or '... metal <Al|Ti|Cu|Au|Ag|Cr> <roughness>'";
                double r = 0, g = 0, b = 0, roughness;
                std::string first, error;
                conductor_ior ior;
                if (!(in >> first)) return fail(usage);
                const bool measured = find_conductor(first, ior);
                if (measured) {
                    if (!(in >> roughness) || roughness < 0) return fail(usage);
                } else {
                    std::istringstream number(first);
                    if (!(number >> r) || !(in >> g >> b >> roughness) || roughness < 0) return fail(usage);
                }
                if (!read_normal_map(error)) return fail(error);
                mat = measured ? std::make_shared<metal>(ior, roughness) : std::make_shared<metal>(color(r, g, b), roughness);
            } else if (type == "thin_glass") {  // scene.h:810-823
                const char* usage = "expected 'material <name> thin_glass <index of refraction> [color <r> <g> <b>]'";
                double ior;
                color tint(1, 1, 1);
                std::string key;
                if (!(in >> ior) || ior <= 0) return fail(usage);
                if (in >> key) {
                    double r, g, b;
                    if (key != "color" || !(in >> r >> g >> b) || r < 0 || g < 0 || b < 0 || r > 1 || g > 1 || b > 1)
                        return fail(usage);
                    tint = color(r, g, b);
                }
                mat = std::make_shared<thin_dielectric>(ior, tint);
            } else if (type == "dielectric") {
                const char* usage = "expected 'material <name> dielectric <ior> [roughness <r>] [color <r> <g> <b> [distance <d>]]'";
                double ior, roughness = 0, distance = 1;
                color tint(1, 1, 1);
                std::string key;
                if (!(in >> ior) || ior <= 0) return fail(usage);
                while (in >> key) {
                    bool ok = true;
                    if (key == "roughness") ok = bool(in >> roughness) && roughness >= 0;
                    else if (key == "color") {
                        double r, g, b;
                        ok = bool(in >> r >> g >> b) && r > 0 && g > 0 && b > 0 && r <= 1 && g <= 1 && b <= 1;
                        tint = color(r, g, b);
                    } else if (key == "distance") ok = bool(in >> distance) && distance > 0;
                    else ok = false;
                    if (!ok) return fail(usage);
                }
                mat = std::make_shared<dielectric>(ior, roughness, dielectric::absorption_for(tint, distance));
            } else if (type == "principled") {
                std::shared_ptr<texture> base = std::make_shared<solid_color>(color(0.8, 0.8, 0.8));
                std::shared_ptr<texture> metallic_map, roughness_map, base_normals;
                std::shared_ptr<texture> coat, coat_roughness, coat_normals;  // no coat unless given
                color coat_tint(1, 1, 1);
                bool coat_tinted = false;
                double metallic = 0, roughness = 0.5, ior = 1.5;
                bool measured = false;
                // "<number in [0, 1]>" or "texture <name>" as a texture (scene.h:310-322)
                auto value_texture = [&](std::shared_ptr<texture>& tex, std::string& err) {
                    std::string first;
                    if (!(in >> first)) return false;
                    if (first == "texture") return read_texture_name(tex, err);
                    std::istringstream number(first);
                    double x;
                    if (!(number >> x) || x < 0 || x > 1) return false;
                    tex = std::make_shared<solid_color>(color(x, x, x));
                    return true;
                };
                conductor_ior ior_metal;
                std::string key, error;
                while (in >> key) {
                    bool ok = true;
                    if (key == "texture") ok = read_color_or_texture(in, key, base, error);
                    else if (key == "color") {
                        std::string first;
                        ok = bool(in >> first) && read_color_or_texture(in, first, base, error);
                    } else if (key == "metallic") ok = read_value_or_texture(metallic, metallic_map, error);
                    else if (key == "roughness") ok = read_value_or_texture(roughness, roughness_map, error);
                    else if (key == "ior") ok = bool(in >> ior) && ior > 0;
                    else if (key == "normal_map") ok = read_texture_name(base_normals, error);
                    else if (key == "clearcoat") ok = value_texture(coat, error);
                    else if (key == "clearcoat_roughness") ok = value_texture(coat_roughness, error);
                    else if (key == "clearcoat_normal_map") ok = read_texture_name(coat_normals, error);
                    else if (key == "clearcoat_color") {
                        double r, g, b;  // 0 would need infinite absorption, > 1 would add energy
                        ok = bool(in >> r >> g >> b) && r > 0 && g > 0 && b > 0 && r <= 1 && g <= 1 && b <= 1;
                        coat_tint = color(r, g, b);
                        coat_tinted = true;
                    } else if (key == "conductor") {
                        std::string metal_name;
                        ok = bool(in >> metal_name) && (measured = find_conductor(metal_name, ior_metal));
                        if (!ok) error = "unknown conductor '" + metal_name + "' (known: Al Ti Cu Au Ag Cr)";
                    } else return fail("principled setting '" + key + "' is not modelled by the GPU path");
                    if (!ok) return fail(error.empty() ? "bad value for principled setting '" + key + "'" : error);
                }
                auto pm = std::make_shared<principled>(base, metallic, roughness, ior);
                pm->metallic_map = metallic_map;
                pm->roughness_map = roughness_map;
                pm->measured = measured;
                pm->conductor = ior_metal;
                if (use_normal_maps) pm->normal_map = base_normals;  // principled bends its own normals
                if ((coat_roughness || coat_normals || coat_tinted) && !coat)
                    return fail("clearcoat_roughness, clearcoat_normal_map and clearcoat_color need 'clearcoat <weight>'");
                if (coat) {
                    pm->set_clearcoat(coat, coat_roughness, use_normal_maps ? coat_normals : nullptr);
                    pm->set_clearcoat_color(coat_tint);
                }
                mat = pm;
            } else if (type == "volume") {
                double r, g, b, aniso = 0;
                std::string key;
                if (!(in >> r >> g >> b)) return fail("expected 'material <name> volume <r> <g> <b> [g <anisotropy>]'");
                if (in >> key) {
                    if (key != "g" || !(in >> aniso) || aniso <= -1 || aniso >= 1)
                        return fail("expected 'material <name> volume <r> <g> <b> [g <anisotropy between -1 and 1>]'");
                }
                mat = std::make_shared<phase_material>(color(r, g, b), aniso);
            } else if (type == "emissive") {
                double r, g, b, spread = 180;
                std::string key;
                if (!(in >> r >> g >> b)) return fail("expected 'material <name> emissive <r> <g> <b> [spread <degrees>]'");
                if (in >> key && (key != "spread" || !(in >> spread) || spread <= 0 || spread > 180))
                    return fail("expected 'material <name> emissive <r> <g> <b> [spread <degrees, 0 to 180>]'");
                mat = std::make_shared<diffuse_light>(color(r, g, b), spread);
            } else {
                return fail("material type '" + type + "' is not modelled by the GPU path");
            }
            if (normal_map && use_normal_maps) mat = std::make_shared<normal_mapped>(mat, normal_map);
            materials[name] = mat;
        } else if (keyword == "sphere" || keyword == "quad" || keyword == "box") {
            std::shared_ptr<hittable> object;
            std::shared_ptr<material> mat;
            transform xform;
            bool transformed;
            std::string error;
            if (keyword == "sphere") {
                double x, y, z, radius;
                if (!(in >> x >> y >> z >> radius) || radius == 0)
                    return fail("expected 'sphere <x> <y> <z> <non-zero radius> [material] [transforms]'");
                if (!read_tail(in, mat, xform, transformed, error)) return fail(error);
                object = std::make_shared<sphere>(point3(x, y, z), radius, mat);
            } else if (keyword == "quad") {
                point3 Q;
                vec3 u, v;
                if (!(in >> Q[0] >> Q[1] >> Q[2] >> u[0] >> u[1] >> u[2] >> v[0] >> v[1] >> v[2]))
                    return fail("expected 'quad <Qx Qy Qz> <ux uy uz> <vx vy vz> [material] [transforms]'");
                if (cross(u, v).near_zero()) return fail("quad edges must not be parallel");
                if (!read_tail(in, mat, xform, transformed, error)) return fail(error);
                object = std::make_shared<quad>(Q, u, v, mat);
            } else {
                point3 a, b;
                if (!(in >> a[0] >> a[1] >> a[2] >> b[0] >> b[1] >> b[2]))
                    return fail("expected 'box <x0 y0 z0> <x1 y1 z1> [material] [transforms]'");
                if (!read_tail(in, mat, xform, transformed, error)) return fail(error);
                object = box(a, b, mat);
            }
            if (!add_object(object, mat, xform, transformed, error)) return fail(error);
        } else if (keyword == "medium") {
            double density;
            std::string shape, error;
            std::shared_ptr<hittable> boundary;
            std::shared_ptr<material> mat;
            transform xform;
            bool transformed;
            if (!(in >> density >> shape) || density <= 0)
                return fail("expected 'medium <positive density> sphere|box <...> <material> [transforms]'");
            if (shape == "sphere") {
                double x, y, z, radius;
                if (!(in >> x >> y >> z >> radius) || radius <= 0)
                    return fail("expected 'medium <density> sphere <x> <y> <z> <positive radius> <material>'");
                boundary = std::make_shared<sphere>(point3(x, y, z), radius, default_material);
            } else if (shape == "box") {
                point3 a, b;
                if (!(in >> a[0] >> a[1] >> a[2] >> b[0] >> b[1] >> b[2]))
                    return fail("expected 'medium <density> box <x0 y0 z0> <x1 y1 z1> <material>'");
                boundary = box(a, b, default_material);
            } else {
                return fail("a medium's boundary must be 'sphere' or 'box'");
            }
            if (!read_tail(in, mat, xform, transformed, error)) return fail(error);
            if (!std::dynamic_pointer_cast<phase_material>(mat)) return fail("a medium needs a volume material");
            if (transformed) boundary = std::make_shared<instance>(boundary, xform);
            world.add(std::make_shared<constant_medium>(boundary, density, mat));
        } else if (keyword == "height_fog") {  // scene.h:656-666
            const char* usage = "expected 'height_fog <density> <scale height> <volume material> [base <y>]'";
            double density, scale, base = 0;
            std::string name, key;
            if (!(in >> density >> scale >> name) || density <= 0 || scale <= 0) return fail(usage);
            if (in >> key && (key != "base" || !(in >> base))) return fail(usage);
            auto it = materials.find(name);
            if (it == materials.end() || !std::dynamic_pointer_cast<phase_material>(it->second))
                return fail("expected a volume material, got '" + name + "'");
            result.unbounded.add(std::make_shared<height_fog>(density, scale, base, it->second));
        } else if (keyword == "fog") {
            std::string name;
            if (!(in >> fog_density >> name) || fog_density <= 0) return fail("expected 'fog <positive density> <volume material>'");
            auto it = materials.find(name);
            if (it == materials.end() || !std::dynamic_pointer_cast<phase_material>(it->second))
                return fail("expected a volume material, got '" + name + "'");
            fog_material = it->second;
        } else {
            return fail("keyword '" + keyword + "' is not modelled by the GPU path");
        }
        std::string extra;
        if (in >> extra) return fail("unexpected value '" + extra + "'");
    }

    if (world.empty()) return fail("the scene has no objects");
    if (ground) result.unbounded.add(std::make_shared<ground_plane>(ground_height, ground_material));  // scene.h:1237-1238
    if (sky_uniform) {
        if (cam.environment) return fail("'sky uniform' and 'environment' can't be used together");
        cam.environment = std::make_shared<environment_light>(8, 4, std::vector<color>(32, sky_radiance));
        lights.add(cam.environment);
    }
    // scene.h:1217-1235: the sky bake, then the sun after the environment in the light list
    double el = degrees_to_radians(sun_elevation), az = degrees_to_radians(sun_azimuth);
    vec3 sun_dir(std::cos(el) * std::sin(az), std::sin(el), -std::cos(el) * std::cos(az));
    color sun_e = sun_intensity * sun_tint;
    if (sky_atmosphere) {
        // lit by the sun's settings even with "sun off"; the sun's own light
        // then reaches the ground through the air
        if (cam.environment) return fail("'sky atmosphere' and 'environment' can't be used together");
        cam.environment = std::make_shared<environment_light>(256, 128, air.bake(256, 128, sun_dir, sun_e));
        lights.add(cam.environment);
        color tr = air.transmittance(vec3(0, sky_air::r_view, 0), unit_vector(sun_dir));
        sun_e = color(sun_e[0] * tr[0], sun_e[1] * tr[1], sun_e[2] * tr[2]);
    }
    if (sun_on) {
        cam.sun = std::make_shared<sun_light>(sun_dir, sun_angle, sun_e);
        cam.sun_disk_visible = sun_disk;
        lights.add(cam.sun);
    }
    if (fog_material) {
        // the world's bounding box: flatten what is there and let the core compute it
        flat_builder fb;
        try {
            fb.add_world(world);
        } catch (const std::exception& e) {
            return fail(e.what());
        }
        host_scene hs;
        std::string e2;
        if (build_host_scene(fb.desc(), hs, e2) != MPT_OK) return fail(e2);
        const box3& wb = hs.world_box;
        point3 lo(wb.ax[0].min, wb.ax[1].min, wb.ax[2].min);
        point3 hi(wb.ax[0].max, wb.ax[1].max, wb.ax[2].max);
        world.add(std::make_shared<constant_medium>(box(lo, hi, default_material), fog_density, fog_material));
    }
    if ((cam.lookfrom - cam.lookat).near_zero() || cross(cam.vup, cam.lookfrom - cam.lookat).near_zero())
        return fail("camera lookfrom must differ from lookat, and vup must not point along the view");
    try {
        lights.build(world.bounding_box());  // scene.h:1248
    } catch (const std::exception& e) {
        return fail(e.what());
    }
    return true;
}

}  // namespace mpt

// ------------------------------------------------------------------ C ABI: mpt_scene_load
namespace mpt {
int set_error(int code, const std::string& m);
}

extern "C" int32_t mpt_scene_load(const char* path, mpt_scene** out, mpt_camera* cam) {
    if (!path || !out) return mpt::set_error(MPT_E_INVALID, "null argument");
    *out = nullptr;
    mpt::scene sc;
    std::string err;
    if (!mpt::load_scene(path, sc, &err)) return mpt::set_error(err.find("cannot open") == 0 ? MPT_E_IO : MPT_E_UNSUPPORTED, err);
    try {
        mpt::flat_builder b;
        b.add_world(sc.world);
        b.add_world(sc.unbounded);
        b.add_lights(sc.lights, sc.cam.environment);
        int rc = mpt_scene_create(&b.desc(), out);
        if (rc != MPT_OK) return rc;
    } catch (const std::exception& e) {
        return mpt::set_error(MPT_E_UNSUPPORTED, e.what());
    }
    if (cam) *cam = sc.cam.to_abi();
    return MPT_OK;
}

extern "C" int32_t mpt_scene_describe(const char* path, mpt_scene_desc* desc, mpt_camera* cam, void** holder) {
    if (!path || !desc || !holder) return mpt::set_error(MPT_E_INVALID, "null argument");
    *holder = nullptr;
    mpt::scene sc;
    std::string err;
    if (!mpt::load_scene(path, sc, &err)) return mpt::set_error(err.find("cannot open") == 0 ? MPT_E_IO : MPT_E_UNSUPPORTED, err);
    auto* b = new mpt::flat_builder;
    try {
        b->add_world(sc.world);
        b->add_world(sc.unbounded);
        b->add_lights(sc.lights, sc.cam.environment);
        *desc = b->desc();
    } catch (const std::exception& e) {
        delete b;
        return mpt::set_error(MPT_E_UNSUPPORTED, e.what());
    }
    if (cam) *cam = sc.cam.to_abi();
    *holder = b;
    return MPT_OK;
}

extern "C" void mpt_desc_free(void* holder) { delete static_cast<mpt::flat_builder*>(holder); }

namespace {
// The host scene of a description (validation, boxes, BVHs, light tables), without the GPU.
int host_of(const mpt_scene_desc* d, mpt::host_scene& hs) {
    if (!d) return mpt::set_error(MPT_E_INVALID, "null description");
    std::string err;
    const int rc = mpt::build_host_scene(*d, hs, err);
    return rc == MPT_OK ? MPT_OK : mpt::set_error(rc, err);
}
mpt::box3 bounds_of(const mpt::host_scene& hs, const mpt_scene_desc* d, const double* b6) {
    mpt::box3 b = hs.world_box;
    const double* src = b6 ? b6 : (d->has_light_bounds ? d->light_bounds : nullptr);
    if (src)
        for (int a = 0; a < 3; a++) b.ax[a] = mpt::span{src[2 * a], src[2 * a + 1]};
    return b;
}
}  // namespace

extern "C" int32_t mpt_desc_bounds(const mpt_scene_desc* desc, double* box6) {
    if (!box6) return mpt::set_error(MPT_E_INVALID, "null argument");
    mpt::host_scene hs;
    int rc = host_of(desc, hs);
    if (rc) return rc;
    for (int a = 0; a < 3; a++) {
        box6[2 * a] = hs.world_box.ax[a].min;
        box6[2 * a + 1] = hs.world_box.ax[a].max;
    }
    return MPT_OK;
}

extern "C" int32_t mpt_desc_light_powers(const mpt_scene_desc* desc, const double* scene_bounds6, double* powers) {
    mpt::host_scene hs;
    int rc = host_of(desc, hs);
    if (rc) return rc;
    if (!powers && !hs.lights.empty()) return mpt::set_error(MPT_E_INVALID, "null argument");
    std::vector<double> pw;
    mpt::light_selection(hs, bounds_of(hs, desc, scene_bounds6), desc->light_uniform_fraction, &pw);
    for (size_t k = 0; k < pw.size(); k++) powers[k] = pw[k];
    return MPT_OK;
}

extern "C" int32_t mpt_desc_light_probabilities(const mpt_scene_desc* desc, const double* scene_bounds6,
                                                double* probabilities) {
    mpt::host_scene hs;
    int rc = host_of(desc, hs);
    if (rc) return rc;
    if (!probabilities && !hs.lights.empty()) return mpt::set_error(MPT_E_INVALID, "null argument");
    mpt::light_selection(hs, bounds_of(hs, desc, scene_bounds6), desc->light_uniform_fraction, nullptr);
    for (size_t k = 0; k < hs.lights.size(); k++) probabilities[k] = hs.lights[k].probability;
    return MPT_OK;
}

extern "C" int32_t mpt_find_conductor(const char* name, double* eta, double* k) {
    mpt::conductor_ior ior;
    if (!name || !eta || !k || !mpt::find_conductor(name, ior)) return MPT_E_INVALID;
    for (int c = 0; c < 3; c++) {
        eta[c] = ior.eta[c];
        k[c] = ior.k[c];
    }
    return MPT_OK;
}

extern "C" int32_t mpt_scene_display(const char* path, mpt_display* display) {
    if (!path || !display) return mpt::set_error(MPT_E_INVALID, "null argument");
    mpt::scene sc;
    std::string err;
    if (!mpt::load_scene(path, sc, &err)) return mpt::set_error(err.find("cannot open") == 0 ? MPT_E_IO : MPT_E_UNSUPPORTED, err);
    *display = sc.display.to_abi();
    return MPT_OK;
}
