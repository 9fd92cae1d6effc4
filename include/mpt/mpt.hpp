// mpt.hpp - C++ host API of the MI355X path-tracing core, mirroring the
// reference renderer's scene API so that code written against it switches
// over by changing the include and namespace:
//
//   reference (adriendjn/mypbr)                this header
//   hittable / hit_record  hittable.h:12-60    mpt::hittable (flattened, never hit on the host)
//   hittable_list          hittable_list.h     mpt::hittable_list
//   bvh                    bvh.h:20-43         mpt::bvh (the BVH itself is built on upload)
//   sphere, quad, box()    sphere.h, quad.h    mpt::sphere, mpt::quad, mpt::box()
//   mesh / triangle        mesh.h:175-233      mpt::mesh (positions + index triples)
//   instance, transform    instance.h, transform.h
//   constant_medium        medium.h:57-96      mpt::constant_medium
//   material classes       material.h, medium.h (lambertian, metal, dielectric,
//                                                 principled, diffuse_light, phase_material)
//   texture classes        texture.h:24-56     mpt::solid_color, mpt::checker_texture
//   environment_light      environment.h       mpt::environment_light (in-memory pixels)
//   light_list             light_list.h:33-134 mpt::light_list
//   camera::render         camera.h:79-105     mpt::camera::render -> mpt_render (include/mpt.h)
//   frame                  denoiser.h:14-20    mpt::frame
//   denoiser               denoiser.h:38-193   mpt::denoiser -> mpt_denoise (include/mpt.h)
//   scene / load_scene     scene.h:32-244      mpt::scene / mpt::load_scene
//   display_settings,      tonemap.h:11-37,    mpt::display_settings, mpt::post_settings,
//   post_settings,         postprocess.h       mpt::postprocess -> mpt_postprocess,
//   postprocess(), write_image()  image_io.h   mpt::write_image -> mpt_write_image
//
// Objects are plain descriptions; camera::render flattens world and lights
// into the C ABI's structure-of-arrays description, uploads it once per
// (world, lights) pair and renders on the GPU. Types the GPU path does not
// model make render() throw std::runtime_error instead of falling back.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mpt.h"

namespace mpt {

class vec3 {
public:
    double v[3];
    vec3() : v{0, 0, 0} {}
    vec3(double x, double y, double z) : v{x, y, z} {}
    double x() const { return v[0]; }
    double y() const { return v[1]; }
    double z() const { return v[2]; }
    double operator[](int i) const { return v[i]; }
    double& operator[](int i) { return v[i]; }
    vec3 operator-() const { return vec3(-v[0], -v[1], -v[2]); }
    double length_squared() const { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
    double length() const { return std::sqrt(length_squared()); }
    bool near_zero() const {
        const double s = 1e-8;
        return std::fabs(v[0]) < s && std::fabs(v[1]) < s && std::fabs(v[2]) < s;
    }
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
inline vec3 unit_vector(const vec3& a) { return a / a.length(); }

// An affine transform p -> M p + t with its inverse (transform.h).
class transform {
public:
    transform();
    static transform translate(const vec3& offset);
    static transform scale(const vec3& s);
    static transform rotate(double degrees, const vec3& axis);
    static transform orthogonal(const double r[3][3]);
    transform then_after(const transform& first) const;
    point3 apply_point(const point3& p) const;
    vec3 apply_vector(const vec3& v) const;
    point3 apply_inverse_point(const point3& p) const;
    vec3 apply_inverse_vector(const vec3& v) const;
    bool is_similarity() const;
    double area_scale() const;
    mpt_xform to_abi() const;

private:
    double m[3][3];
    vec3 t;
    double inv_m[3][3];
    vec3 inv_t;
};

// ------------------------------------------------------------------ textures
class texture {
public:
    virtual ~texture() = default;
    virtual mpt_texture to_abi() const = 0;
    virtual std::shared_ptr<texture> inner_texture() const { return nullptr; }  // wrappers: what they read
};
class solid_color : public texture {
public:
    explicit solid_color(const color& albedo) : albedo(albedo) {}
    mpt_texture to_abi() const override;
    color albedo;
};
class checker_texture : public texture {
public:
    checker_texture(double scale, std::shared_ptr<texture> even, std::shared_ptr<texture> odd);
    mpt_texture to_abi() const override;
    double inv_scale;
    std::shared_ptr<texture> even, odd;
};
// An image over the (u, v) unit square, v = 0 at the bottom row, looked up at
// the nearest texel (texture.h:58-205). PNG and JPEG files are 8-bit: sRGB
// encoded colors, or data taken as k / 255 with linear set; Radiance .hdr
// files are linear floats. repeat wraps (u, v) around instead of clamping.
class image_texture : public texture {
public:
    explicit image_texture(const std::string& path, bool repeat = false, bool linear = false);
    // a PNG or JPEG file's bytes in memory (an image inside a glTF file)
    image_texture(const uint8_t* data, size_t size, bool repeat, bool linear);
    // a 1 x 1 white image whose alpha is the coverage rounded to a byte: a
    // constant mask (texture.h:98-106)
    explicit image_texture(double coverage);
    bool has_alpha_channel() const { return hdr.empty() && (channels == 2 || channels == 4); }
    bool loaded() const { return width > 0 && height > 0; }
    // a grey (1- or 2-channel) file: a height map rather than a normal map (texture.h:108-110)
    bool greyscale() const { return hdr.empty() && channels <= 2; }
    // every texel's coverage (alpha, or the grey level without alpha) is 1 (texture.h:101-105, 166-173)
    bool fully_opaque() const;
    mpt_texture to_abi() const override;  // offset 0: the builder places the texels
    int width = 0, height = 0;
    int channels = 0;           // of the file (1 grey, 2 grey + alpha, 3, 4)
    bool repeat, linear;
    std::string error;          // why loading failed
    std::vector<uint8_t> ldr;   // RGBA bytes, rows top first
    std::vector<float> hdr;     // RGB floats, rows top first
};

// Wrappers of another texture (texture.h:212-282). The core reads at most
// one value wrapper (scaled, channel, normal scale) over uv transforms over
// a plain texture: the chains glTF's texture views give.
class scaled_texture : public texture {  // inner times a colour
public:
    scaled_texture(std::shared_ptr<texture> inner, const color& scale) : inner(std::move(inner)), scale(scale) {}
    mpt_texture to_abi() const override;
    std::shared_ptr<texture> inner_texture() const override { return inner; }
    std::shared_ptr<texture> inner;
    color scale;
};
class channel_texture : public texture {  // grey of one channel (0 red, 1 green, 2 blue) times a factor
public:
    channel_texture(std::shared_ptr<texture> inner, int channel, double scale)
        : inner(std::move(inner)), channel(channel), scale(scale) {}
    mpt_texture to_abi() const override;
    std::shared_ptr<texture> inner_texture() const override { return inner; }
    std::shared_ptr<texture> inner;
    int channel;
    double scale;
};
class uv_transform_texture : public texture {  // KHR_texture_transform: scale, rotate, offset
public:
    uv_transform_texture(std::shared_ptr<texture> inner, double offset_u, double offset_v, double rotation,
                         double scale_u, double scale_v);
    mpt_texture to_abi() const override;
    std::shared_ptr<texture> inner_texture() const override { return inner; }
    std::shared_ptr<texture> inner;
    double offset_u, offset_v, scale_u, scale_v, cos_r, sin_r;
};
class normal_scale_texture : public texture {  // a normal map's tilt times a factor
public:
    normal_scale_texture(std::shared_ptr<texture> inner, double scale) : inner(std::move(inner)), scale(scale) {}
    mpt_texture to_abi() const override;
    std::shared_ptr<texture> inner_texture() const override { return inner; }
    std::shared_ptr<texture> inner;
    double scale;
};

// ------------------------------------------------------------------ materials
class material {
public:
    virtual ~material() = default;
    // Fills the ABI record; texture is the index already assigned to the
    // material's texture (or -1).
    virtual mpt_material to_abi(int32_t texture_index) const = 0;
    virtual std::shared_ptr<texture> base_texture() const { return nullptr; }
    // Tangent-space normal map and principled's metallic / roughness maps
    // (null: none, or the material's constant).
    virtual std::shared_ptr<texture> normal_map_texture() const { return nullptr; }
    virtual std::shared_ptr<texture> metallic_texture() const { return nullptr; }
    virtual std::shared_ptr<texture> roughness_texture() const { return nullptr; }
    // principled's clear coat: weight, roughness and its own normal map (null: none)
    virtual std::shared_ptr<texture> coat_weight_texture() const { return nullptr; }
    virtual std::shared_ptr<texture> coat_roughness_texture() const { return nullptr; }
    virtual std::shared_ptr<texture> coat_normal_texture() const { return nullptr; }
    virtual color average_emission() const { return color(0, 0, 0); }
    // emissive_surface's glow (null: none)
    virtual std::shared_ptr<texture> glow_texture() const { return nullptr; }
};
// A material whose shading normal is bent by a tangent-space normal map
// (normal_mapped, material.h:631-674); for lambertian and metal.
class normal_mapped : public material {
public:
    normal_mapped(std::shared_ptr<material> inner, std::shared_ptr<texture> normal_map)
        : inner(std::move(inner)), normal_map(std::move(normal_map)) {}
    mpt_material to_abi(int32_t ti) const override { return inner->to_abi(ti); }
    std::shared_ptr<texture> base_texture() const override { return inner->base_texture(); }
    std::shared_ptr<texture> normal_map_texture() const override { return normal_map; }
    color average_emission() const override { return inner->average_emission(); }
    std::shared_ptr<material> inner;
    std::shared_ptr<texture> normal_map;
};
class lambertian : public material {
public:
    explicit lambertian(const color& albedo) : tex(std::make_shared<solid_color>(albedo)) {}
    explicit lambertian(std::shared_ptr<texture> t) : tex(std::move(t)) {}
    mpt_material to_abi(int32_t ti) const override;
    std::shared_ptr<texture> base_texture() const override { return tex; }
    std::shared_ptr<texture> tex;
};
// A measured metal's complex index of refraction, RGB (microfacet.h:68-96).
struct conductor_ior {
    color eta, k;
};
// The named metals of the reference's table (Al, Ti, Cu, Au, Ag, Cr and their
// long names); false for any other name.
bool find_conductor(const std::string& name, conductor_ior& ior);

class metal : public material {
public:
    metal(const color& albedo, double roughness);
    metal(const conductor_ior& ior, double roughness);  // exact Fresnel, also on the mirror
    mpt_material to_abi(int32_t ti) const override;
    color albedo;
    double roughness;
    bool measured = false;
    conductor_ior ior{};
};
// A thin sheet of glass (thin_dielectric, material.h:683-707): light is
// reflected (2R / (1 + R) of it) or goes straight through, tinted.
class thin_dielectric : public material {
public:
    explicit thin_dielectric(double ior, const color& tint = color(1, 1, 1)) : ior(ior), tint(tint) {}
    mpt_material to_abi(int32_t ti) const override;
    double ior;
    color tint;
};
class dielectric : public material {
public:
    dielectric(double refraction_index, double roughness = 0, const color& absorption = color(0, 0, 0));
    static color absorption_for(const color& tint, double distance);
    mpt_material to_abi(int32_t ti) const override;
    double refraction_index, roughness;
    color sigma_a;
};
class principled : public material {
public:
    principled(std::shared_ptr<texture> base_color, double metallic, double roughness, double ior);
    mpt_material to_abi(int32_t ti) const override;
    std::shared_ptr<texture> base_texture() const override { return base_color; }
    std::shared_ptr<texture> normal_map_texture() const override { return normal_map; }
    std::shared_ptr<texture> metallic_texture() const override { return metallic_map; }
    std::shared_ptr<texture> roughness_texture() const override { return roughness_map; }
    std::shared_ptr<texture> base_color;
    double metallic, roughness, ior;
    // read from the red channel at each hit when set (material.h:508-515)
    std::shared_ptr<texture> metallic_map, roughness_map;
    std::shared_ptr<texture> normal_map;  // bends the base layer (material.h:389)
    // the metallic part as a measured metal (set_conductor, material.h:381-385)
    bool measured = false;
    conductor_ior conductor{};
    // a clear coat of weight w and roughness (red channels, clamped to [0, 1])
    // and an optional normal map of its own (set_clearcoat, material.h:363-371);
    // a null weight removes it
    void set_clearcoat(std::shared_ptr<texture> weight, std::shared_ptr<texture> roughness,
                       std::shared_ptr<texture> coat_normals = nullptr);
    // what the coat lets through head-on, in and back out (set_clearcoat_color, material.h:373-379)
    void set_clearcoat_color(const color& tint) { coat_tint = tint; }
    std::shared_ptr<texture> coat_weight_texture() const override { return coat_map; }
    std::shared_ptr<texture> coat_roughness_texture() const override { return coat_map ? coat_roughness_map : nullptr; }
    std::shared_ptr<texture> coat_normal_texture() const override { return coat_map ? coat_normal_map : nullptr; }
    std::shared_ptr<texture> coat_map, coat_roughness_map, coat_normal_map;
    color coat_tint{1, 1, 1};
};
class diffuse_light : public material {
public:
    explicit diffuse_light(const color& emit, double spread_degrees = 180);
    mpt_material to_abi(int32_t ti) const override;
    color average_emission() const override;
    color emit;
    double spread;
};
// A material whose front face glows with a texture while it scatters as its
// inner material (emissive_surface, material.h:714-741): not a light.
class emissive_surface : public material {
public:
    emissive_surface(std::shared_ptr<material> inner, std::shared_ptr<texture> emission)
        : inner(std::move(inner)), emission(std::move(emission)) {}
    mpt_material to_abi(int32_t ti) const override { return inner->to_abi(ti); }
    std::shared_ptr<texture> base_texture() const override { return inner->base_texture(); }
    std::shared_ptr<texture> normal_map_texture() const override { return inner->normal_map_texture(); }
    std::shared_ptr<texture> metallic_texture() const override { return inner->metallic_texture(); }
    std::shared_ptr<texture> roughness_texture() const override { return inner->roughness_texture(); }
    std::shared_ptr<texture> coat_weight_texture() const override { return inner->coat_weight_texture(); }
    std::shared_ptr<texture> coat_roughness_texture() const override { return inner->coat_roughness_texture(); }
    std::shared_ptr<texture> coat_normal_texture() const override { return inner->coat_normal_texture(); }
    std::shared_ptr<texture> glow_texture() const override { return emission; }
    std::shared_ptr<material> inner;
    std::shared_ptr<texture> emission;
};
class phase_material : public material {
public:
    explicit phase_material(const color& albedo, double g = 0) : albedo(albedo), g(g) {}
    mpt_material to_abi(int32_t ti) const override;
    color albedo;
    double g;
};

// ------------------------------------------------------------------ geometry
class flat_builder;

// interval (interval.h): default-constructed intervals are empty (+inf, -inf).
struct interval {
    double min = std::numeric_limits<double>::infinity();
    double max = -std::numeric_limits<double>::infinity();
    interval() = default;
    interval(double min, double max) : min(min), max(max) {}
    double size() const { return max - min; }
    bool contains(double x) const { return min <= x && x <= max; }
};

// aabb (aabb.h): three intervals; default-constructed boxes are empty.
struct aabb {
    interval x, y, z;
    aabb() = default;
    aabb(const interval& x, const interval& y, const interval& z) : x(x), y(y), z(z) {}
    const interval& axis_interval(int n) const { return n == 1 ? y : n == 2 ? z : x; }
};

class hittable {
public:
    virtual ~hittable() = default;
    // Appends this object to the world list of the description; xform is the
    // index of the instance transform it sits behind (-1 for none).
    virtual void flatten(flat_builder& b, int32_t xform) const = 0;
    // hittable::bounding_box (hittable.h:40): the box the core's BVH build
    // gives this object (aabb.h padding included); for a list, the union.
    virtual aabb bounding_box() const;
    // hittable::power (hittable.h:54-59): luminance of the emitted power,
    // pi * area * luminance(average emission) for emitters (sums for lists
    // and meshes, times the area scale behind an instance), 0 otherwise.
    // The core's own light-power code computes it.
    virtual double power() const;
};

class sphere : public hittable {
public:
    sphere(const point3& center, double radius, std::shared_ptr<material> mat)
        : center(center), radius(radius), mat(std::move(mat)) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    point3 center;
    double radius;
    std::shared_ptr<material> mat;
};

class quad : public hittable {
public:
    quad(const point3& Q, const vec3& u, const vec3& v, std::shared_ptr<material> mat)
        : Q(Q), u(u), v(v), mat(std::move(mat)) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    point3 Q;
    vec3 u, v;
    std::shared_ptr<material> mat;
};

class hittable_list : public hittable {
public:
    std::vector<std::shared_ptr<hittable>> objects;
    hittable_list() = default;
    explicit hittable_list(std::shared_ptr<hittable> object) { add(std::move(object)); }
    void clear() { objects.clear(); }
    void add(std::shared_ptr<hittable> object) { objects.push_back(std::move(object)); }
    bool empty() const { return objects.empty(); }
    void flatten(flat_builder& b, int32_t xform) const override;
};

// A BVH over a list. The GPU path builds its own binned-SAH BVH over the
// world when the scene is uploaded, so this only marks the list.
class bvh : public hittable {
public:
    explicit bvh(const hittable_list& list) : objects(list.objects) {}
    explicit bvh(const std::vector<std::shared_ptr<hittable>>& objs) : objects(objs) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    std::vector<std::shared_ptr<hittable>> objects;
};

// Returns the six sides of the box with opposite corners a and b (quad.h:95-113).
std::shared_ptr<hittable_list> box(const point3& a, const point3& b, std::shared_ptr<material> mat);

class mesh : public hittable {
public:
    // positions and counter-clockwise index triples; one material.
    mesh(std::vector<point3> positions, std::vector<int32_t> indices, std::shared_ptr<material> mat)
        : positions(std::move(positions)), indices(std::move(indices)), mat(std::move(mat)) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    size_t size() const { return indices.size() / 3; }
    std::vector<point3> positions;
    std::vector<int32_t> indices;
    std::shared_ptr<material> mat;
    // optional per-vertex shading normals (mesh_data::normals, mesh.h:14-19):
    // normal_indices parallels indices; a triangle with any -1 is flat
    std::vector<vec3> normals;
    std::vector<int32_t> normal_indices;
    // optional per-vertex texture coordinates (mesh_data::uvs, x = u, y = v):
    // uv_indices parallels indices; a triangle with any -1 has none
    std::vector<vec3> uvs;
    std::vector<int32_t> uv_indices;
    // optional alpha mask (mesh.h:24-32): its coverage at a hit's (u, v) cuts
    // holes in the triangles; null or fully opaque: none
    std::shared_ptr<image_texture> alpha_mask;
};

// A part of an OBJ file that uses one material (obj_loader.h:19-23).
struct obj_part {
    std::shared_ptr<material> mat;
    std::shared_ptr<mesh> geometry;
};

// load_obj (obj_loader.h:170-340): v, vn, vt, f (fans; v, v/t, v//n, v/t/n;
// negative indices), usemtl, mtllib with load_mtl's mapping (Kd, Ks/Ns/Ni,
// Ke; map_Kd and normal maps as repeating image textures; a texture that
// cannot be loaded falls back to the plain color with a warning, as in the
// reference). Alpha masks make it fail: the GPU path does not model them.
// Parts come in the order of each material's first face.
bool load_obj(const std::string& path, const std::map<std::string, std::shared_ptr<material>>& materials,
              std::shared_ptr<material> default_mat, bool smooth, std::vector<obj_part>& parts, std::string& error,
              bool use_normal_maps = true);

// load_gltf (gltf_loader.h:21-399): a glTF 2.0 file (.gltf with its buffers
// and images, or .glb) as one mesh part per material and coverage mask, the
// node tree flattened into world-space vertices. override_mat, when set,
// replaces every material; replace swaps materials by their name in the file.
// Returns false with a message in error when the file cannot be read or has
// no triangles.
bool load_gltf(const std::string& path, std::shared_ptr<material> override_mat, std::vector<obj_part>& parts,
               std::string& error, bool use_normal_maps = true,
               const std::map<std::string, std::shared_ptr<material>>& replace = {});

// Photo studio pieces (studio.h). The shape of a cyclorama in its own frame:
// floor y = 0 from z = front back to z = -back, then a quarter circle of the
// radius up into a wall at z = -back - radius, height high, width wide
// about x = 0.
struct cyclorama_settings {
    double width = 20, front = 10, back = 4, height = 8, radius = 3;
    int segments = 32;  // flat strips along the curve
};
// A cyclorama as one mesh (make_cyclorama, studio.h:39-87): profile vertices
// carry the circle's true normal, uv runs along x and along the profile by
// arc length.
std::shared_ptr<mesh> make_cyclorama(const cyclorama_settings& s, std::shared_ptr<material> mat);
// A width x height rectangle light centred on center whose front faces aim,
// its width horizontal, turned by roll degrees about the aiming axis
// (make_softbox, studio.h:89-101).
std::shared_ptr<quad> make_softbox(const point3& center, const point3& aim, double width, double height,
                                   double roll_degrees, std::shared_ptr<material> mat);

class instance : public hittable {
public:
    instance(std::shared_ptr<hittable> object, const transform& xform) : object(std::move(object)), xform(xform) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    std::shared_ptr<hittable> object;
    transform xform;
};

class constant_medium : public hittable {
public:
    constant_medium(std::shared_ptr<hittable> boundary, double density, std::shared_ptr<material> phase)
        : boundary(std::move(boundary)), density(density), phase(std::move(phase)) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    std::shared_ptr<hittable> boundary;
    double density;
    std::shared_ptr<material> phase;
};

// Equirectangular sky from pixels in memory, row 0 at +y (environment.h:44-50).
class environment_light : public hittable {
public:
    environment_light(int width, int height, std::vector<color> pixels, double intensity = 1, double rotation_degrees = 0)
        : width(width), height(height), pixels(std::move(pixels)), intensity(intensity), rotation_degrees(rotation_degrees) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    aabb bounding_box() const override { return aabb(); }  // environment.h:59
    double power() const override { return 0; }           // needs the radius: power(scene_radius)
    // environment_light::power (environment.h:116-127): pi R^2 times the
    // luminance of the radiance integrated over the sphere.
    double power(double scene_radius) const;
    int width, height;
    std::vector<color> pixels;
    double intensity, rotation_degrees;
};

// A distant light of small angular size (sun_light, sun.h:23-80): never hit,
// added to escaping rays, sampled uniformly inside its cone. angular_diameter
// in degrees; irradiance on a surface facing it.
class sun_light : public hittable {
public:
    sun_light(const vec3& direction, double angular_diameter, const color& irradiance)
        : direction(direction), angular_diameter(angular_diameter), irradiance(irradiance) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    aabb bounding_box() const override { return aabb(); }  // sun.h:41
    double power() const override { return 0; }
    // sun_light::power (sun.h:66-71): pi R^2 luminance(irradiance).
    double power(double scene_radius) const;
    vec3 direction;
    double angular_diameter;
    color irradiance;
};

// atmosphere_settings and atmosphere (atmosphere.h:36-165): the
// single-scattering sky a `sky atmosphere` line bakes, for scenes built in
// memory. bake() gives the pixels of an equirectangular environment_light
// (row 0 at +y) lit by a sun of the given direction and irradiance above
// the air; sun_transmittance() is the share of that irradiance reaching the
// ground (0 below the horizon), by which the loader dims the sun. The values
// are the loader's, bit for bit.
struct atmosphere_settings {
    double rayleigh = 1;   // air density, relative to Earth's
    double mie = 1;        // aerosol (haze) density
    double mie_g = 0.8;    // forward scattering of the haze
    double ozone = 1;      // ozone density
    double intensity = 3;  // overall scale of the sky's radiance
};
class atmosphere {
public:
    explicit atmosphere(const atmosphere_settings& s) : settings(s) {}
    color sun_transmittance(const vec3& direction) const;
    color radiance(const vec3& direction, const vec3& sun_direction, const color& sun_irradiance) const;
    std::vector<color> bake(int width, int height, const vec3& sun_direction, const color& sun_irradiance) const;
    atmosphere_settings settings;
};

// light_list (light_list.h:33-134): picks light k with probability p_k,
// in proportion to its power with a uniform_fraction share kept uniform.
// build(scene_bounds) computes the probabilities with the core's light
// code (the radius of scene_bounds sets the environment's and the sun's
// power) and renders use the same bounds; until it is called, selection is
// uniform as in the reference. build() without bounds uses the world's own
// box, as scene.h:1248 passes it.
class light_list : public hittable_list {
public:
    enum class selection { uniform, power };
    selection strategy = selection::power;
    double uniform_fraction = 0.2;
    void build(const aabb& scene_bounds);
    void build() {
        built = true;
        has_bounds = false;
    }
    // Probability that the sampler picks light k (light_list.h:98-103).
    double probability(size_t k) const;
    // The light a uniform xi in [0, 1) selects (light_list.h:118-127).
    size_t pick(double xi) const;
    bool built = false;
    bool has_bounds = false;
    aabb bounds;
private:
    std::vector<double> probabilities_, cdf_;
};

// ------------------------------------------------------------------ camera
struct frame {  // denoiser.h:14-20
    std::vector<color> radiance, albedo, normal;
    std::vector<double> variance;
};

class camera {
public:
    double aspect_ratio = 1.0;
    int image_width = 100;
    int samples_per_pixel = 10;
    int max_depth = 10;
    int max_diffuse = 0, max_specular = 0, max_scatter = 0;
    bool transparent_shadows = false;
    double vfov = 90;
    point3 lookfrom = point3(0, 0, 0);
    point3 lookat = point3(0, 0, -1);
    vec3 vup = vec3(0, 1, 0);
    bool use_sky = true;
    color background = color(0, 0, 0);
    std::shared_ptr<environment_light> environment;
    std::shared_ptr<sun_light> sun;  // added to escaping rays; must also be in the lights (camera.h:45)
    bool sun_disk_visible = true;    // whether camera rays see the sun itself (camera.h:46)
    double ortho_height = 0;
    double defocus_angle = 0;
    double focus_dist = 10;
    int blades = 0;
    double blade_rotation = 0;
    int threads = 0;  // kept for source compatibility; the GPU ignores it
    bool collect_features = false;
    uint64_t seed = 1;  // the counter-based RNG's seed

    int height() const { return image_height; }
    using pass_callback = std::function<void(int, const frame&)>;

    // camera::render (camera.h:79-105) on the GPU.
    frame render(const hittable& world, const hittable_list& lights, int save_every = 0,
                 const pass_callback& on_pass = nullptr);

    mpt_camera to_abi() const;

private:
    int image_height = 1;
};

// denoiser (denoiser.h:38-193): the edge-avoiding a-trous filter, run on the
// GPU (mpt_denoise). run() needs a frame rendered with collect_features.
class denoiser {
public:
    int iterations = 5;
    double sigma_luminance = 4;
    double sigma_normal = 0.1;
    double sigma_albedo = 0.1;
    int threads = 0;  // kept for source compatibility; the GPU ignores it

    std::vector<color> run(const frame& f, int width, int height) const;
};

// A flattened scene and the GPU scene made from it.
class gpu_scene {
public:
    gpu_scene(const hittable& world, const hittable_list& lights, const std::shared_ptr<environment_light>& env);
    ~gpu_scene();
    gpu_scene(const gpu_scene&) = delete;
    gpu_scene& operator=(const gpu_scene&) = delete;
    mpt_scene* get() const { return handle; }

private:
    mpt_scene* handle = nullptr;
};

// post_settings (postprocess.h:12-21): bloom and vignetting of display images.
class post_settings {
public:
    double bloom = 0;            // glow strength (0 = off)
    double bloom_threshold = 1;  // luminance above which pixels glow
    double bloom_radius = 0.04;  // reach of the glow, as a fraction of the width
    double vignette = 0;         // corner darkening (0 = off)

    bool active() const { return bloom > 0 || vignette > 0; }
};

enum class tonemapper { none, reinhard, aces };  // tonemap.h:10

// display_settings (tonemap.h:11-37): how radiance becomes PNG/PPM pixels.
class display_settings {
public:
    double exposure = 0;              // stops: radiance is scaled by 2^exposure
    tonemapper op = tonemapper::none; // none clamps
    post_settings post;

    // Parses "none", "reinhard" or "aces"; returns false if unknown.
    bool set_tonemapper(const std::string& name);
    mpt_display to_abi() const;
};

// postprocess(pixels, width, height, post) (postprocess.h:108-153) on the GPU.
std::vector<color> postprocess(const std::vector<color>& pixels, int width, int height, const post_settings& post);

// write_image(path, width, height, pixels, display) (image_io.h:81-106):
// .png/.ppm display images (bytes from the GPU), .hdr/.pfm raw radiance; an
// empty path writes PPM to stdout. Returns false and prints the error.
bool write_image(const std::string& path, int width, int height, const std::vector<color>& pixels,
                 const display_settings& display);

// The infinite horizontal plane y = height (plane.h:10-42). It has no finite
// box: it goes in scene::unbounded and is tested beside the BVH, not in it;
// (u, v) = (x, -z), so textures repeat over it and normal maps work.
class ground_plane : public hittable {
public:
    ground_plane(double height, std::shared_ptr<material> mat) : height(height), mat(std::move(mat)) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    aabb bounding_box() const override;  // plane.h:36: the whole plane y = height
    double height;
    std::shared_ptr<material> mat;
};

// Fog whose density falls off with height, density(y) = density *
// exp(-(y - base) / scale_height), filling all of space (height_fog,
// medium.h:98-155): free flights drawn by inverting the closed-form optical
// depth along the ray. No finite box: it goes in scene::unbounded.
class height_fog : public hittable {
public:
    height_fog(double density, double scale_height, double base, std::shared_ptr<material> phase)
        : density(density), scale_height(scale_height), base(base), phase(std::move(phase)) {}
    void flatten(flat_builder& b, int32_t xform) const override;
    aabb bounding_box() const override;  // medium.h:151: the universe
    double density, scale_height, base;
    std::shared_ptr<material> phase;
};

// ------------------------------------------------------------------ voxels
// MagicaVoxel .vox files (vox_loader.h) and voxel grids walked with a 3D DDA
// (voxel.h). One voxel is one unit; cell (x, y, z) covers [x, x+1] x
// [y, y+1] x [z, z+1] of its model; files are z-up.

// A dense grid of palette indices, x fastest, then y, then z; 0 is empty.
struct vox_model {
    int nx = 0, ny = 0, nz = 0;
    std::vector<uint8_t> cells;
    uint8_t at(int x, int y, int z) const {  // 0 outside the model
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0;
        return cells[(size_t(z) * ny + y) * nx + x];
    }
};
// A model placed by the file's scene graph: xform takes model coordinates to
// the file's (still z-up) world.
struct vox_instance {
    int model = 0;
    transform xform;
};
struct vox_file {
    std::vector<vox_model> models;
    std::vector<vox_instance> instances;
    color palette[256];                                 // sRGB-encoded, in [0, 1], as stored
    std::map<std::string, std::string> materials[256];  // MATL properties per palette index
};
// load_vox (vox_loader.h:176-327): versions 150 and 200; SIZE/XYZI models,
// RGBA (MagicaVoxel's default palette without one), MATL, and the
// nTRN/nGRP/nSHP scene graph (first animation frame; hidden nodes skipped).
bool load_vox(const std::string& path, vox_file& out, std::string& error);

// How a palette index takes part in surfaces (voxel.h:16-21).
enum class voxel_kind : uint8_t { empty, opaque, transparent };
// A surface is where two neighbouring cells differ in kind (voxel.h:103-111).
bool voxel_face_visible(voxel_kind k, voxel_kind n);

// The 256 palette entries as materials (make_voxel_palette, voxel.h:49-99):
// _diffuse lambertian, _metal principled, _glass tinted dielectric
// (transparent), _emit diffuse light; colours sRGB-decoded.
struct voxel_palette {
    std::array<std::shared_ptr<material>, 256> materials;
    std::array<voxel_kind, 256> kinds;
    std::array<bool, 256> emissive;
    voxel_kind kind(uint8_t index) const { return kinds[index]; }
};
voxel_palette make_voxel_palette(const vox_file& vox, double emission_scale);

// The visible faces of a model as quads with their cells' materials; the
// faces of emissive cells also go to emitters (voxel_model_faces, voxel.h:113-152).
void voxel_model_faces(const vox_model& m, const voxel_palette& pal, hittable_list& faces, hittable_list& emitters);

// The cells [lo, hi) of a model, walked cell by cell with a 3D DDA
// (voxel_grid, voxel.h:168-337); with an edge material the outlines of the
// model's shapes are drawn with it, edge_width wide (in voxels).
class voxel_grid : public hittable {
public:
    voxel_grid(std::shared_ptr<const vox_model> model, std::shared_ptr<const voxel_palette> palette, const int lo[3],
               const int hi[3], std::shared_ptr<material> edge_material = nullptr, double edge_width = 0);
    void flatten(flat_builder& b, int32_t xform) const override;
    std::shared_ptr<const vox_model> model;
    std::shared_ptr<const voxel_palette> palette;
    std::shared_ptr<material> edge_material;
    double edge_width;
    int lo[3], hi[3];
};
// A model as voxel_grid blocks of up to block^3 cells, each shrunk to the
// box of its voxels, blocks without voxels skipped, in one BVH (voxel_blocks,
// voxel.h:339-382). On the GPU path the blocks are one object with its own
// BVH, instanced by instance like a mesh.
class voxel_block_set : public hittable {
public:
    void flatten(flat_builder& b, int32_t xform) const override;
    std::vector<std::shared_ptr<voxel_grid>> blocks;
};
std::shared_ptr<hittable> voxel_blocks(std::shared_ptr<const vox_model> model,
                                       std::shared_ptr<const voxel_palette> palette, int block = 8,
                                       std::shared_ptr<material> edge_material = nullptr, double edge_width = 0);

struct scene {  // scene.h:32-40
    hittable_list world;
    hittable_list unbounded;  // objects without a finite box (height fogs, then a ground plane)
    light_list lights;
    camera cam;
    display_settings display;  // how radiance becomes PNG/PPM pixels
};

// Loads a .scene file (scene.h:42-243) for the subset the GPU path models.
// Returns false and writes "path:line: message" to err on failure.
bool load_scene(const std::string& path, scene& result, std::string* err = nullptr);

// Flattens world + lights (+ environment) into the ABI description; the
// arrays live in the builder.
class flat_builder {
public:
    flat_builder();
    ~flat_builder();
    void add_world(const hittable& world);
    void add_lights(const hittable_list& lights, const std::shared_ptr<environment_light>& env);
    const mpt_scene_desc& desc();

    // used by hittable::flatten
    int32_t material_index(const std::shared_ptr<material>& m);
    int32_t texture_index(const std::shared_ptr<texture>& t);  // -1 for null; image texels placed once
    int32_t xform_index(const transform& x);
    int64_t add_prim(int32_t kind, int32_t mat, const double* data9);
    void set_prim_normals(int64_t prim, const double* n9);  // a triangle's three vertex normals
    void set_prim_uvs(int64_t prim, const double* uv6);     // a triangle's three texture coordinates
    void set_ground(double height, int32_t material);       // the one ground plane
    void add_fog(double density, double scale_height, double base, int32_t material);  // a height fog
    int32_t add_object(int32_t kind, int64_t first, int64_t count, int32_t xform, const hittable* owner);
    // a voxel model with its palette and edges in the description's tables (once per combination)
    int32_t vox_model_index(const vox_model& m, const voxel_palette& pal, const std::shared_ptr<material>& edge_material,
                            double edge_width);
    std::map<const void*, std::pair<int64_t, int64_t>> mesh_ranges;
    struct impl;
    impl* p;
};

}  // namespace mpt
