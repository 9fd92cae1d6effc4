// scene_host.cpp - validation of a flattened scene description and the
// one-off host work before upload: primitive records, object boxes, the
// binned-SAH BVHs, light selection probabilities, the environment's
// sampling tables and the GGX albedo table.
#include "scene_host.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <utility>

#include "microfacet.hpp"

namespace mpt {

namespace {

constexpr int kBins = 16;        // bvh.h:110
constexpr int kMaxLeaf = 4;      // bvh.h:111
constexpr int kMaxSahDepth = 64; // bvh.h:115

struct builder {
    std::vector<build_item>& items;
    std::vector<bvh_node>& nodes;

    int make_leaf(int index, size_t begin, size_t count, int axis, int depth) {  // bvh.h:190-205
        if (count > size_t(kMaxLeaf)) {
            size_t mid = begin + count / 2;
            nodes[index].axis = axis;
            build(begin, mid, depth + 1);
            int right = build(mid, begin + count, depth + 1);
            nodes[index].offset = right;
            return index;
        }
        nodes[index].offset = int(begin);
        nodes[index].count = int(count);
        return index;
    }

    int build(size_t begin, size_t end, int depth) {  // bvh.h:118-188
        int index = int(nodes.size());
        nodes.push_back(bvh_node{});
        nodes[index].count = 0;
        nodes[index].offset = 0;
        nodes[index].axis = 0;

        box3 bounds = box_empty(), cbounds = box_empty();
        for (size_t k = begin; k < end; k++) {
            bounds = box_merge(bounds, items[k].box);
            // aabb(c, c) pads the degenerate point box before the merge
            cbounds = box_merge(cbounds, box_points(items[k].centroid, items[k].centroid));
        }
        for (int a = 0; a < 3; a++) {
            nodes[index].b[2 * a] = bounds.ax[a].min;
            nodes[index].b[2 * a + 1] = bounds.ax[a].max;
        }

        size_t count = end - begin;
        int axis = box_longest_axis(cbounds);
        const span extent = cbounds.ax[axis];

        if (count == 1 || iv_size(extent) <= 0 || depth >= kMaxSahDepth)
            return make_leaf(index, begin, count, axis, depth);

        struct bin {
            box3 box;
            int count;
        } bins[kBins];
        for (auto& b : bins) {
            b.box = box_empty();
            b.count = 0;
        }
        auto bin_of = [&](const v3& c) {
            int b = int(kBins * (at(c, axis) - extent.min) / iv_size(extent));
            return std::min(b, kBins - 1);
        };
        for (size_t k = begin; k < end; k++) {
            auto& b = bins[bin_of(items[k].centroid)];
            b.box = box_merge(b.box, items[k].box);
            b.count++;
        }

        double best_cost = kInf;
        int best_split = -1;
        for (int split = 0; split < kBins - 1; split++) {
            box3 left = box_empty(), right = box_empty();
            int lc = 0, rc = 0;
            for (int b = 0; b <= split; b++) {
                left = box_merge(left, bins[b].box);
                lc += bins[b].count;
            }
            for (int b = split + 1; b < kBins; b++) {
                right = box_merge(right, bins[b].box);
                rc += bins[b].count;
            }
            if (lc == 0 || rc == 0) continue;
            double cost = box_area(left) * lc + box_area(right) * rc;
            if (cost < best_cost) {
                best_cost = cost;
                best_split = split;
            }
        }

        double leaf_cost = double(count);
        double split_cost = 0.125 + best_cost / box_area(bounds);
        if (best_split < 0 || (count <= size_t(kMaxLeaf) && leaf_cost <= split_cost))
            return make_leaf(index, begin, count, axis, depth);

        auto mid_it = std::partition(items.begin() + begin, items.begin() + end,
                                     [&](const build_item& it) { return bin_of(it.centroid) <= best_split; });
        size_t mid = size_t(mid_it - items.begin());

        nodes[index].axis = axis;
        build(begin, mid, depth + 1);
        int right = build(mid, end, depth + 1);
        nodes[index].offset = right;
        return index;
    }
};

// Boxes of the primitive kinds (sphere.h:14-19, quad.h:17-22, mesh.h:40-41).
box3 prim_box(int kind, const double* p) {
    if (kind == MPT_PRIM_SPHERE) {
        v3 c{p[0], p[1], p[2]};
        double r = std::fabs(p[3]);
        v3 rv{r, r, r};
        return box_points(c - rv, c + rv);
    }
    if (kind == MPT_PRIM_QUAD) {
        v3 Q{p[0], p[1], p[2]}, u{p[3], p[4], p[5]}, v{p[6], p[7], p[8]};
        return box_merge(box_points(Q, Q + u + v), box_points(Q + u, Q + v));
    }
    if (kind == MPT_PRIM_VOXBLOCK) {  // the block's cells (voxel.h:187)
        box3 b;
        for (int a = 0; a < 3; a++) b.ax[a] = span{p[a], p[3 + a]};
        return b;
    }
    v3 p0{p[0], p[1], p[2]}, p1{p[3], p[4], p[5]}, p2{p[6], p[7], p[8]};
    return box_merge(box_points(p0, p1), box_points(p2, p2));
}

// instance bounding box: the eight transformed corners (instance.h:13-23).
box3 instance_box(const xform& x, const box3& b) {
    box3 r = box_empty();
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
                v3 c{i ? b.ax[0].max : b.ax[0].min, j ? b.ax[1].max : b.ax[1].min, k ? b.ax[2].max : b.ax[2].min};
                v3 p = xf_point(x, c);
                r = box_merge(r, box_points(p, p));
            }
    return r;
}

double area_scale(const xform& x) {  // transform.h:92-97
    const double* m = x.m;
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                 m[2] * (m[3] * m[7] - m[4] * m[6]);
    return std::pow(std::fabs(det), 2.0 / 3.0);
}

}  // namespace

void build_bvh(std::vector<build_item>& items, std::vector<bvh_node>& nodes) {
    nodes.clear();
    if (items.empty()) {
        bvh_node n{};
        box3 e = box_empty();
        for (int a = 0; a < 3; a++) {
            n.b[2 * a] = e.ax[a].min;
            n.b[2 * a + 1] = e.ax[a].max;
        }
        nodes.push_back(n);
        return;
    }
    nodes.reserve(2 * items.size());
    builder b{items, nodes};
    b.build(0, items.size(), 0);
}

void pair_layout(const std::vector<bvh_node>& nodes, int root, std::vector<cnode>& out, croot& r) {
    const bvh_node& R = nodes[root];
    for (int k = 0; k < 6; k++) r.b[k] = R.b[k];
    r.qoff = r.pad = 0;
    if (R.count > 0) {
        r.off = R.offset;
        r.cnt = R.count;
        return;
    }
    r.cnt = 0;
    r.off = int32_t(out.size());
    // depth-first: a record's index is fixed before its children are made
    std::vector<std::pair<int, int>> todo{{root, int(out.size())}};
    out.push_back(cnode{});
    while (!todo.empty()) {
        auto [n, idx] = todo.back();
        todo.pop_back();
        const int child[2] = {n + 1, nodes[n].offset};
        cnode c{};
        c.axis = nodes[n].axis;
        for (int w = 0; w < 2; w++) {
            const bvh_node& ch = nodes[child[w]];
            for (int k = 0; k < 6; k++) c.b[6 * w + k] = ch.b[k];
            if (ch.count > 0) {
                c.off[w] = ch.offset;
                c.cnt[w] = ch.count;
            } else {
                c.off[w] = int32_t(out.size());
                c.cnt[w] = 0;
                out.push_back(cnode{});
                todo.push_back({child[w], c.off[w]});
            }
        }
        out[idx] = c;
    }
}

int quad_layout(const std::vector<cnode>& pairs, int root_record, std::vector<qnode>& out) {
    const int root = int(out.size());
    out.push_back(qnode{});
    std::vector<std::pair<int, int>> todo{{root_record, root}};  // (pair record, its qnode)
    while (!todo.empty()) {
        auto [r, qi] = todo.back();
        todo.pop_back();
        const cnode& c = pairs[r];
        qnode q{};
        int slot = 0;
        q.info = c.axis & 3;
        // slot from child w of pair record pr
        auto put = [&](int pr, int w) {
            const cnode& p = pairs[pr];
            for (int k = 0; k < 6; k++) {
                const double v = p.b[6 * w + k];
                float f = float(v);
                if ((k & 1) ? double(f) < v : double(f) > v) f = std::nextafterf(f, (k & 1) ? INFINITY : -INFINITY);
                q.b[6 * slot + k] = f;
            }
            if (p.cnt[w] > 0) {
                q.off[slot] = p.off[w];
                q.meta[slot] = ((pr << 1) | w) | (p.cnt[w] << 28);
            } else {
                q.off[slot] = int32_t(out.size());
                q.meta[slot] = 0;
                out.push_back(qnode{});
                todo.push_back({p.off[w], q.off[slot]});
            }
            slot++;
        };
        for (int w = 0; w < 2; w++) {
            if (c.cnt[w] > 0) {
                put(r, w);  // a leaf child: one slot
            } else {
                const int cr = c.off[w];  // an interior child: its two children
                q.info |= (pairs[cr].axis & 3) << (2 + 2 * w);
                q.info |= 1 << (6 + w);
                put(cr, 0);
                put(cr, 1);
            }
        }
        for (int k = slot; k < 4; k++) q.meta[k] = 0;
        out[qi] = q;
    }
    return root;
}

void build_ggx_table(std::vector<double>& e, std::vector<double>& e_avg) {  // microfacet.h:299-337
    const int size = kGgxTableSize;
    const int strata = 64;
    e.assign(size * size, 0.0);
    e_avg.assign(size, 0.0);
    for (int j = 0; j < size; j++) {
        double roughness = double(j) / (size - 1);
        double alpha = roughness_to_alpha(roughness);
        for (int i = 0; i < size; i++) {
            double mu = std::fmax(double(i) / (size - 1), 1e-3);
            v3 wo{std::sqrt(1 - mu * mu), 0, mu};
            double sum = 0;
            for (int a = 0; a < strata; a++)
                for (int b = 0; b < strata; b++) {
                    v3 h = ggx_sample_visible(alpha, wo, (a + 0.5) / strata, (b + 0.5) / strata);
                    v3 wi = -wo + 2 * dot(wo, h) * h;
                    if (wi.z > 0) sum += ggx_G(alpha, wo, wi) / ggx_G1(alpha, wo);
                }
            e[j * size + i] = std::fmin(sum / (strata * strata), 1.0);
        }
        double integral = 0;
        double h = 1.0 / (size - 1);
        for (int i = 0; i + 1 < size; i++) {
            double a = i * h, b = (i + 1) * h;
            integral += h * (e[j * size + i] * (2 * a + b) + e[j * size + i + 1] * (a + 2 * b)) / 6;
        }
        e_avg[j] = std::fmin(2 * integral, 1.0);
    }
}

namespace {
// fresnel_conductor (microfacet.h:57-66): unpolarised Fresnel reflectance of a
// complex index eta + i k, in std::complex arithmetic as the reference's.
double conductor_fresnel(double cos_i, double eta, double k) {
    cos_i = std::clamp(cos_i, 0.0, 1.0);
    std::complex<double> n(eta, k);
    std::complex<double> sin2_t = (1 - cos_i * cos_i) / (n * n);
    std::complex<double> cos_t = std::sqrt(1.0 - sin2_t);
    std::complex<double> r_parl = (n * cos_i - cos_t) / (n * cos_i + cos_t);
    std::complex<double> r_perp = (cos_i - n * cos_t) / (cos_i + n * cos_t);
    return (std::norm(r_parl) + std::norm(r_perp)) / 2;
}
}  // namespace

void conductor_averages(mpt_material& m) {
    auto at = [&](double c, int ch) { return conductor_fresnel(c, m.eta[ch], m.k[ch]); };
    const int n = 512;  // midpoint rule (F is smooth)
    double sum[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        const double mu = (i + 0.5) / n;
        for (int c = 0; c < 3; c++) sum[c] = sum[c] + at(mu, c) * mu;
    }
    for (int c = 0; c < 3; c++) {
        if (m.type == MPT_MAT_METAL) m.albedo[c] = at(1, c);
        m.f_avg[c] = (1 / double(n)) * (2 * sum[c]);
    }
}

void light_selection(host_scene& s, const box3& bounds, double uniform_fraction, std::vector<double>* power_out) {
    const int nl = int(s.lights.size());
    if (power_out) power_out->assign(size_t(nl), 0.0);
    if (nl > 0 && (s.light_strategy == 1 || power_out)) {
        double radius = 0;
        const box3& sb = bounds;
        double sx = iv_size(sb.ax[0]), sy = iv_size(sb.ax[1]), sz = iv_size(sb.ax[2]);
        if (sx > 0 && sy > 0 && sz > 0) radius = 0.5 * std::sqrt(sx * sx + sy * sy + sz * sz);
        if (!std::isfinite(radius) || radius <= 0) radius = 1;
        auto prim_power = [&](int64_t p) {
            int m = s.prim_mat[p];
            if (m < 0) return 0.0;
            double lum = luminance(v3{s.materials[m].avg_emission[0], s.materials[m].avg_emission[1],
                                      s.materials[m].avg_emission[2]});
            const double* q = &s.prim[size_t(p) * kPrimStride];
            if (s.prim_kind[p] == MPT_PRIM_SPHERE) return kPi * 4 * kPi * q[3] * q[3] * lum;
            if (s.prim_kind[p] == MPT_PRIM_QUAD) {
                v3 n = cross(v3{q[3], q[4], q[5]}, v3{q[6], q[7], q[8]});
                return kPi * length(n) * lum;
            }
            return kPi * q[12] * lum;
        };
        std::vector<double> power(nl);
        double known_sum = 0;
        int known = 0;
        for (int k = 0; k < nl; k++) {
            int ob = s.lights[k].object;
            if (ob == -2) {  // sun_light::power (sun.h:66-71)
                power[k] = kPi * radius * radius * luminance(s.sun_irradiance);
            } else if (ob < 0) {  // environment_light::power (environment.h:116-127)
                double sum = 0;
                for (int j = 0; j < s.env_h; j++)
                    for (int i = 0; i < s.env_w; i++) {
                        double sin_theta = std::sin(kPi * (j + 0.5) / s.env_h);
                        const double* p = &s.env_rgb[(size_t(j) * s.env_w + i) * 3];
                        sum += std::fmax(luminance(v3{p[0], p[1], p[2]}) * sin_theta, 0.0);
                    }
                double integral = s.env_intensity * sum * (2 * kPi / s.env_w) * (kPi / s.env_h);
                power[k] = kPi * radius * radius * integral;
            } else {
                const obj_rec& r = s.objects[ob];
                double pw = 0;
                if (r.kind == MPT_OBJ_PRIM) pw = prim_power(r.first);
                else if (r.kind == MPT_OBJ_MESH)  // mesh::power, in the mesh's triangle order
                    for (int64_t i = 0; i < r.count; i++) pw += prim_power(s.mlight_prim[size_t(r.light_table + i)]);
                else
                    for (int64_t i = 0; i < r.count; i++) pw += prim_power(r.first + i);
                if (r.xform >= 0) pw = pw * area_scale(s.xforms[r.xform]);
                power[k] = pw;
            }
            if (std::isfinite(power[k]) && power[k] > 0) {
                known_sum += power[k];
                known++;
            }
        }
        if (power_out) *power_out = power;
        double fallback = known > 0 ? known_sum / known : 1.0;
        double total = 0;
        for (auto& p : power) {
            if (!std::isfinite(p) || p <= 0) p = fallback;
            total += p;
        }
        const double uf = uniform_fraction;
        if (s.light_strategy == 1)
            for (int k = 0; k < nl; k++) s.lights[k].probability = (1 - uf) * power[k] / total + uf / double(nl);
    }
    s.light_cdf.assign(nl + 1, 0.0);
    for (int k = 0; k < nl; k++) s.light_cdf[k + 1] = s.light_cdf[k] + s.lights[k].probability;
    if (nl > 0) s.light_cdf[nl] = 1;

}

int build_host_scene(const mpt_scene_desc& d, host_scene& s, std::string& err) {
    auto fail = [&](int code, const std::string& m) {
        err = m;
        return code;
    };
    if (d.n_prims < 0 || d.n_objects < 0 || d.n_xforms < 0 || d.n_materials < 0 || d.n_textures < 0 ||
        d.n_lights < 0)
        return fail(MPT_E_INVALID, "negative count in scene description");
    if (d.n_prims > 0 && (!d.prim_kind || !d.prim_material || !d.prim_data))
        return fail(MPT_E_INVALID, "primitive arrays missing");
    if (d.n_objects > 0 && (!d.object_kind || !d.object_first || !d.object_count))
        return fail(MPT_E_INVALID, "object arrays missing");
    if (d.n_prims >= (int64_t(1) << 31))
        return fail(MPT_E_CAPACITY, "more than 2^31 primitives");

    // materials and textures
    s.textures.assign(d.textures, d.textures + d.n_textures);
    if (d.n_texel_bytes < 0 || d.n_texels_hdr < 0 || (d.n_texel_bytes % 4) != 0 || (d.n_texels_hdr % 3) != 0 ||
        (d.n_texel_bytes > 0 && !d.texels) || (d.n_texels_hdr > 0 && !d.texels_hdr))
        return fail(MPT_E_INVALID, "bad texel arrays");
    for (size_t k = 0; k < s.textures.size(); k++) {
        const mpt_texture& t = s.textures[k];
        if (t.type == MPT_TEX_IMAGE) {
            const bool hdr = t.flags & MPT_TEXF_HDR;
            const int64_t have = hdr ? d.n_texels_hdr / 3 : d.n_texel_bytes / 4;
            if (t.width < 1 || t.height < 1 || t.offset < 0 || t.offset + int64_t(t.width) * t.height > have)
                return fail(MPT_E_INVALID, "image texture " + std::to_string(k) + " lies outside the texel arrays");
        } else if (t.type >= MPT_TEX_SCALED && t.type <= MPT_TEX_NORMAL_SCALE) {
            // a value wrapper over uv transforms over a plain texture, at most four deep
            int32_t cur = int32_t(k), depth = 0;
            bool value_seen = false;
            while (s.textures[size_t(cur)].type >= MPT_TEX_SCALED) {
                const mpt_texture& w = s.textures[size_t(cur)];
                const bool value = w.type != MPT_TEX_UV_TRANSFORM;
                if ((value && (value_seen || depth > 0)) || ++depth > 4 || w.inner < 0 || w.inner >= d.n_textures)
                    return fail(MPT_E_INVALID, "texture " + std::to_string(k) + ": a wrapper chain must be one value "
                                               "wrapper over uv transforms over a texture, at most 4 deep");
                if (w.type == MPT_TEX_CHANNEL && (w.channel < 0 || w.channel > 2))
                    return fail(MPT_E_INVALID, "texture " + std::to_string(k) + ": channel must be 0, 1 or 2");
                value_seen = value_seen || value;
                cur = w.inner;
            }
        } else if (t.type != MPT_TEX_SOLID && t.type != MPT_TEX_CHECKER) {
            return fail(MPT_E_UNSUPPORTED, "unsupported texture type " + std::to_string(t.type));
        }
    }
    s.texels.assign(d.texels, d.texels + d.n_texel_bytes);
    s.texels_hdr.assign(d.texels_hdr, d.texels_hdr + d.n_texels_hdr);
    s.srgb.resize(256);
    for (int k = 0; k < 256; k++) {  // texture.h:197-208
        const double encoded = k / 255.0;
        s.srgb[size_t(k)] = encoded <= 0.04045 ? encoded / 12.92 : std::pow((encoded + 0.055) / 1.055, 2.4);
    }
    s.materials.assign(d.materials, d.materials + d.n_materials);
    for (auto& m : s.materials)
        if (m.measured) conductor_averages(m);
    s.uses_uv = false;
    s.rich = false;
    s.ground_mat = -1;
    if (d.has_ground) {
        if (d.ground_material < 0 || d.ground_material >= d.n_materials)
            return fail(MPT_E_INVALID, "the ground plane's material does not exist");
        if (!std::isfinite(d.ground_height)) return fail(MPT_E_INVALID, "the ground plane's height is not finite");
        s.ground_mat = d.ground_material;
        s.ground_h = d.ground_height;
        s.rich = true;  // the ground is tested in the kernels of the variant that also carries (u, v)
    }
    s.fogs.clear();
    if (d.n_fogs < 0 || (d.n_fogs > 0 && (!d.fog_params || !d.fog_material)))
        return fail(MPT_E_INVALID, "height fogs without their parameters");
    for (int32_t k = 0; k < d.n_fogs; k++) {
        const double* q = d.fog_params + 3 * k;
        const int32_t m = d.fog_material[k];
        if (!(q[0] > 0) || !(q[1] > 0) || !std::isfinite(q[0]) || !std::isfinite(q[1]) || !std::isfinite(q[2]))
            return fail(MPT_E_INVALID, "a height fog needs a positive density and scale height");
        if (m < 0 || m >= d.n_materials || d.materials[m].type != MPT_MAT_PHASE)
            return fail(MPT_E_INVALID, "a height fog needs a volume material");
        s.fogs.push_back(fog_rec{q[0], q[1], q[2], m, 0});
        s.rich = true;  // tested beside the BVH, as the ground is, by the variant that carries (u, v)
    }
    auto tex_ok = [&](int32_t t) { return t >= -1 && t < d.n_textures; };
    auto is_image = [&](int32_t t) {  // an image, or wrappers of one: read at (u, v)
        if (t < 0) return false;
        while (s.textures[size_t(t)].type >= MPT_TEX_SCALED) t = s.textures[size_t(t)].inner;
        return s.textures[size_t(t)].type == MPT_TEX_IMAGE;
    };
    for (size_t k = 0; k < s.materials.size(); k++) {
        auto& m = s.materials[k];
        if (m.type < MPT_MAT_LAMBERTIAN || m.type > MPT_MAT_PHASE)
            return fail(MPT_E_UNSUPPORTED, "unsupported material type " + std::to_string(m.type));
        bool needs_tex = m.type == MPT_MAT_LAMBERTIAN || m.type == MPT_MAT_PRINCIPLED;
        if (needs_tex && (m.texture < 0 || m.texture >= d.n_textures))
            return fail(MPT_E_INVALID, "material " + std::to_string(k) + " has no valid texture");
        if (!tex_ok(m.normal_map) || !tex_ok(m.metallic_texture) || !tex_ok(m.roughness_texture))
            return fail(MPT_E_INVALID, "material " + std::to_string(k) + " names a texture that does not exist");
        const bool mapped = m.type == MPT_MAT_LAMBERTIAN || m.type == MPT_MAT_METAL || m.type == MPT_MAT_PRINCIPLED;
        if (m.normal_map >= 0 && !mapped)
            return fail(MPT_E_UNSUPPORTED, "material " + std::to_string(k) + ": normal maps apply to lambertian, metal and principled");
        if ((m.metallic_texture >= 0 || m.roughness_texture >= 0) && m.type != MPT_MAT_PRINCIPLED)
            return fail(MPT_E_UNSUPPORTED, "material " + std::to_string(k) + ": metallic and roughness maps are principled's");
        if (m.coat) {
            if (m.type != MPT_MAT_PRINCIPLED)
                return fail(MPT_E_INVALID, "material " + std::to_string(k) + ": only principled takes a clear coat");
            if (m.coat_weight_texture < 0 || m.coat_weight_texture >= d.n_textures || m.coat_roughness_texture < 0 ||
                m.coat_roughness_texture >= d.n_textures || !tex_ok(m.coat_normal_map))
                return fail(MPT_E_INVALID, "material " + std::to_string(k) + ": the clear coat names a texture that does not exist");
            for (int c = 0; c < 3; c++)
                if (!(m.coat_tint[c] > 0 && m.coat_tint[c] <= 1))
                    return fail(MPT_E_INVALID, "material " + std::to_string(k) + ": the coat's tint must be in (0, 1]");
        } else {
            m.coat_weight_texture = m.coat_roughness_texture = m.coat_normal_map = -1;
        }
        if (m.glow) {
            if (m.type == MPT_MAT_DIFFUSE_LIGHT || m.type == MPT_MAT_PHASE)
                return fail(MPT_E_INVALID, "material " + std::to_string(k) + ": lights and volumes take no glow");
            if (m.glow_texture < 0 || m.glow_texture >= d.n_textures)
                return fail(MPT_E_INVALID, "material " + std::to_string(k) + ": its glow names a texture that does not exist");
        } else {
            m.glow_texture = -1;
        }
        // recomputed here: whatever the caller set, hits of this material need (u, v) exactly when
        m.uses_uv = (m.normal_map >= 0 || (needs_tex && is_image(m.texture)) || is_image(m.metallic_texture) ||
                     is_image(m.roughness_texture) || m.coat_normal_map >= 0 || is_image(m.coat_weight_texture) ||
                     is_image(m.coat_roughness_texture) || is_image(m.glow_texture)) ? 1 : 0;
        if (m.uses_uv) s.uses_uv = true;
        if (m.thin && m.type != MPT_MAT_DIELECTRIC)
            return fail(MPT_E_INVALID, "material " + std::to_string(k) + ": only a dielectric can be thin");
        if (m.measured && m.type != MPT_MAT_METAL && m.type != MPT_MAT_PRINCIPLED)
            return fail(MPT_E_INVALID, "material " + std::to_string(k) + ": only metal and principled can be measured");
        if (m.uses_uv || m.thin || m.measured || m.coat || m.glow) s.rich = true;
    }
    s.xforms.resize(d.n_xforms);
    for (int k = 0; k < d.n_xforms; k++) {
        const mpt_xform& x = d.xforms[k];
        xform& y = s.xforms[k];
        for (int i = 0; i < 9; i++) {
            y.m[i] = x.m[i];
            y.im[i] = x.inv_m[i];
        }
        for (int i = 0; i < 3; i++) {
            y.t[i] = x.t[i];
            y.it[i] = x.inv_t[i];
        }
    }

    // voxel models (vox_loader.h:32-44, voxel.h:23-31)
    s.vox.clear();
    s.vox_cells.clear();
    if (d.n_vox_models < 0 || (d.n_vox_models > 0 && (!d.vox_models || !d.vox_cells || d.n_vox_cells < 0)))
        return fail(MPT_E_INVALID, "voxel models without their cells");
    for (int32_t k = 0; k < d.n_vox_models; k++) {
        const mpt_vox_model& m = d.vox_models[k];
        if (m.nx < 1 || m.ny < 1 || m.nz < 1 || m.nx > 2048 || m.ny > 2048 || m.nz > 2048)
            return fail(MPT_E_INVALID, "a voxel model's size must be 1 to 2048 cells a side");
        if (m.cells < 0 || m.cells + int64_t(m.nx) * m.ny * m.nz > d.n_vox_cells)
            return fail(MPT_E_INVALID, "a voxel model's cells lie outside vox_cells");
        if (m.kind[0] != 0) return fail(MPT_E_INVALID, "palette index 0 of a voxel model must be empty");
        for (int i = 0; i < 256; i++) {
            if (m.kind[i] > 2) return fail(MPT_E_INVALID, "a voxel palette kind must be 0, 1 or 2");
            if (m.material[i] < -1 || m.material[i] >= d.n_materials)
                return fail(MPT_E_INVALID, "a voxel palette entry names a bad material");
        }
        const uint8_t* c = d.vox_cells + m.cells;
        for (int64_t i = 0, n = int64_t(m.nx) * m.ny * m.nz; i < n; i++)
            if (m.kind[c[i]] != 0 && m.material[c[i]] < 0)
                return fail(MPT_E_INVALID, "a voxel cell's palette entry has no material");
        if (!(m.edge_width >= 0 && m.edge_width < 0.5))
            return fail(MPT_E_INVALID, "a voxel model's edge width must be in [0, 0.5)");
        if (m.edge_width > 0 && (m.edge_material < 0 || m.edge_material >= d.n_materials))
            return fail(MPT_E_INVALID, "a voxel model with edges needs an edge material");
        mpt_vox_model h = m;
        h.cells = int64_t(s.vox_cells.size());
        s.vox_cells.insert(s.vox_cells.end(), c, c + int64_t(m.nx) * m.ny * m.nz);
        s.vox.push_back(h);
    }

    // primitives in description order first
    const int64_t np = d.n_prims;
    s.prim.assign(size_t(np) * kPrimStride, 0.0);
    s.prim_kind.assign(d.prim_kind, d.prim_kind + np);
    s.prim_mat.assign(d.prim_material, d.prim_material + np);
    s.prim_canon.resize(np);
    std::vector<box3> pbox(np);
    for (int64_t i = 0; i < np; i++) {
        const double* src = d.prim_data + 9 * i;
        double* dst = &s.prim[size_t(i) * kPrimStride];
        int kind = s.prim_kind[i];
        int mat = s.prim_mat[i];
        if (mat < -1 || mat >= d.n_materials)
            return fail(MPT_E_INVALID, "primitive " + std::to_string(i) + " has a bad material index");
        if (kind == MPT_PRIM_SPHERE) {
            for (int k = 0; k < 4; k++) dst[k] = src[k];
            if (src[3] == 0) return fail(MPT_E_INVALID, "sphere with zero radius");
        } else if (kind == MPT_PRIM_QUAD) {
            v3 Q{src[0], src[1], src[2]}, u{src[3], src[4], src[5]}, v{src[6], src[7], src[8]};
            v3 n = cross(u, v);
            v3 normal = unit_vector(n);
            double D = dot(normal, Q);
            v3 w = n / dot(n, n);
            for (int k = 0; k < 9; k++) dst[k] = src[k];
            dst[9] = normal.x, dst[10] = normal.y, dst[11] = normal.z;
            dst[12] = w.x, dst[13] = w.y, dst[14] = w.z;
            dst[15] = D;
        } else if (kind == MPT_PRIM_TRIANGLE) {
            v3 p0{src[0], src[1], src[2]}, p1{src[3], src[4], src[5]}, p2{src[6], src[7], src[8]};
            v3 e1 = p1 - p0, e2 = p2 - p0;
            v3 g = unit_vector(cross(e1, e2));
            dst[0] = p0.x, dst[1] = p0.y, dst[2] = p0.z;
            dst[3] = e1.x, dst[4] = e1.y, dst[5] = e1.z;
            dst[6] = e2.x, dst[7] = e2.y, dst[8] = e2.z;
            dst[9] = g.x, dst[10] = g.y, dst[11] = g.z;
            dst[12] = 0.5 * length(cross(p1 - p0, p2 - p0));
        } else if (kind == MPT_PRIM_VOXBLOCK) {
            const int32_t model = int32_t(src[6]);
            if (!(src[6] >= 0 && src[6] < d.n_vox_models && double(model) == src[6]))
                return fail(MPT_E_INVALID, "a voxel block names a bad model");
            const int32_t dims[3] = {s.vox[model].nx, s.vox[model].ny, s.vox[model].nz};
            for (int a = 0; a < 3; a++)
                if (!(src[a] >= 0 && src[a] < src[3 + a] && src[3 + a] <= dims[a]) || src[a] != std::floor(src[a]) ||
                    src[3 + a] != std::floor(src[3 + a]))
                    return fail(MPT_E_INVALID, "a voxel block must be whole cells inside its model");
            for (int k = 0; k < 7; k++) dst[k] = src[k];
            if (mat != -1) return fail(MPT_E_INVALID, "a voxel block takes its materials from its palette");
        } else {
            return fail(MPT_E_UNSUPPORTED, "unsupported primitive kind " + std::to_string(kind));
        }
        pbox[i] = prim_box(kind, src);
        s.prim_canon[i] = -1;
    }
    s.prim_nrm.clear();
    if (d.prim_normals) {
        bool any = false;
        for (int64_t i = 0; i < np && !any; i++)
            any = s.prim_kind[i] == MPT_PRIM_TRIANGLE && !std::isnan(d.prim_normals[9 * i]);
        if (any) {
            s.prim_nrm.assign(d.prim_normals, d.prim_normals + 9 * np);
            for (int64_t i = 0; i < np; i++)
                if (s.prim_kind[i] != MPT_PRIM_TRIANGLE) s.prim_nrm[size_t(9 * i)] = std::nan("");
        }
    }
    s.prim_uv.clear();
    if (d.prim_uvs) {
        bool any = false;
        for (int64_t i = 0; i < np && !any; i++)
            any = s.prim_kind[i] == MPT_PRIM_TRIANGLE && !std::isnan(d.prim_uvs[6 * i]);
        if (any) {
            s.prim_uv.assign(d.prim_uvs, d.prim_uvs + 6 * np);
            for (int64_t i = 0; i < np; i++)
                if (s.prim_kind[i] != MPT_PRIM_TRIANGLE) s.prim_uv[size_t(6 * i)] = std::nan("");
        }
    }

    // objects
    const int no = d.n_objects;
    s.objects.resize(no);
    std::vector<box3> obox(no);
    std::map<std::pair<int64_t, int64_t>, int32_t> mesh_roots;
    s.blas.clear();
    int64_t next_id = 0;
    for (int o = 0; o < no; o++) {
        obj_rec& r = s.objects[o];
        r.kind = d.object_kind[o];
        r.first = d.object_first[o];
        r.count = d.object_count[o];
        r.xform = d.object_xform ? d.object_xform[o] : -1;
        r.id_base = d.object_id_base ? d.object_id_base[o] : next_id;
        r.blas_root = -1;
        r.material = d.object_material ? d.object_material[o] : -1;
        r.density = d.object_density ? d.object_density[o] : 0;
        r.bkind = d.object_boundary_kind ? d.object_boundary_kind[o] : MPT_OBJ_PRIM;
        r.light_table = -1;
        r.alpha = d.object_alpha ? d.object_alpha[o] : -1;
        r.flags = d.object_flags ? d.object_flags[o] : 0;
        if (r.flags & ~int32_t(MPT_OBJF_LIGHT_ONLY)) return fail(MPT_E_INVALID, "unknown object flags");
        if (r.kind == MPT_OBJ_VOXEL) s.rich = true;  // grids are walked by the variant that carries (u, v)
        next_id = r.id_base + r.count;
        if (r.alpha != -1) {
            s.rich = true;  // a mask is read at the hit's (u, v): the variant that carries it tests masks
            if (r.kind != MPT_OBJ_MESH) return fail(MPT_E_INVALID, "only meshes take alpha masks");
            if (r.alpha < 0 || r.alpha >= d.n_textures || d.textures[r.alpha].type != MPT_TEX_IMAGE ||
                (d.textures[r.alpha].flags & MPT_TEXF_HDR))
                return fail(MPT_E_INVALID, "an alpha mask must be an 8-bit image texture");
        }
        if (r.first < 0 || r.count < 1 || r.first + r.count > np)
            return fail(MPT_E_INVALID, "object " + std::to_string(o) + " has a bad primitive range");
        if (r.xform < -1 || r.xform >= d.n_xforms)
            return fail(MPT_E_INVALID, "object " + std::to_string(o) + " has a bad transform index");
        box3 b = box_empty();
        switch (r.kind) {
            case MPT_OBJ_PRIM:
                if (r.count != 1) return fail(MPT_E_INVALID, "single-primitive object with count != 1");
                b = pbox[r.first];
                break;
            case MPT_OBJ_LIST:
                for (int64_t k = 0; k < r.count; k++) {
                    if (s.prim_kind[r.first + k] == MPT_PRIM_TRIANGLE || s.prim_kind[r.first + k] == MPT_PRIM_VOXBLOCK)
                        return fail(MPT_E_UNSUPPORTED, "triangles and voxel blocks must be in a mesh or voxel object");
                    b = box_merge(b, pbox[r.first + k]);
                }
                break;
            case MPT_OBJ_MESH:
            case MPT_OBJ_VOXEL: {  // a BVH of triangles (mesh.h:175-191) or of voxel blocks (voxel.h:379)
                const int32_t want = r.kind == MPT_OBJ_MESH ? MPT_PRIM_TRIANGLE : MPT_PRIM_VOXBLOCK;
                for (int64_t k = 0; k < r.count; k++)
                    if (s.prim_kind[r.first + k] != want)
                        return fail(MPT_E_INVALID, r.kind == MPT_OBJ_MESH ? "mesh objects hold triangles only"
                                                                        : "voxel objects hold voxel blocks only");
                if (r.first < 0 || r.count < 1 || r.first + r.count > np)
                    return fail(MPT_E_INVALID, "object " + std::to_string(o) + " has a bad primitive range");
                auto key = std::make_pair(r.first, r.count);
                auto it = mesh_roots.find(key);
                if (it == mesh_roots.end()) {
                    std::vector<build_item> items(r.count);
                    for (int64_t k = 0; k < r.count; k++) {
                        items[k].box = pbox[r.first + k];
                        items[k].centroid = box_centroid(items[k].box);
                        items[k].ref = k;
                    }
                    std::vector<bvh_node> nodes;
                    build_bvh(items, nodes);
                    int32_t root = int32_t(s.blas.size());
                    // leaf offsets become absolute positions in the primitive array
                    for (auto& n : nodes) {
                        if (n.count > 0) n.offset += int32_t(r.first);
                        else n.offset += root;
                    }
                    s.blas.insert(s.blas.end(), nodes.begin(), nodes.end());
                    // reorder the mesh's primitives into leaf order
                    std::vector<double> tmp(size_t(r.count) * kPrimStride);
                    std::vector<box3> tbox(r.count);
                    for (int64_t k = 0; k < r.count; k++) {
                        int64_t src = r.first + items[k].ref;
                        std::copy(&s.prim[size_t(src) * kPrimStride], &s.prim[size_t(src) * kPrimStride] + kPrimStride,
                                  &tmp[size_t(k) * kPrimStride]);
                        tbox[k] = pbox[src];
                    }
                    std::vector<int32_t> tmat(r.count);
                    for (int64_t k = 0; k < r.count; k++) tmat[k] = s.prim_mat[r.first + items[k].ref];
                    if (!s.prim_nrm.empty()) {  // vertex normals follow their triangles
                        std::vector<double> tn(size_t(r.count) * 9);
                        for (int64_t k = 0; k < r.count; k++)
                            std::copy(&s.prim_nrm[size_t(r.first + items[k].ref) * 9],
                                      &s.prim_nrm[size_t(r.first + items[k].ref) * 9] + 9, &tn[size_t(k) * 9]);
                        std::copy(tn.begin(), tn.end(), &s.prim_nrm[size_t(r.first) * 9]);
                    }
                    if (!s.prim_uv.empty()) {  // and so do texture coordinates
                        std::vector<double> tu(size_t(r.count) * 6);
                        for (int64_t k = 0; k < r.count; k++)
                            std::copy(&s.prim_uv[size_t(r.first + items[k].ref) * 6],
                                      &s.prim_uv[size_t(r.first + items[k].ref) * 6] + 6, &tu[size_t(k) * 6]);
                        std::copy(tu.begin(), tu.end(), &s.prim_uv[size_t(r.first) * 6]);
                    }
                    for (int64_t k = 0; k < r.count; k++) {
                        std::copy(&tmp[size_t(k) * kPrimStride], &tmp[size_t(k) * kPrimStride] + kPrimStride,
                                  &s.prim[size_t(r.first + k) * kPrimStride]);
                        pbox[r.first + k] = tbox[k];
                        s.prim_mat[r.first + k] = tmat[k];
                        s.prim_canon[r.first + k] = items[k].ref;
                    }
                    croot cr;
                    pair_layout(s.blas, root, s.cblas, cr);
                    s.cblas_roots.push_back(cr);
                    s.blas_root_node.push_back(root);
                    it = mesh_roots.emplace(key, int32_t(s.cblas_roots.size() - 1)).first;
                }
                r.blas_root = it->second;
                const bvh_node& rn = s.blas[s.blas_root_node[r.blas_root]];
                for (int a = 0; a < 3; a++) b.ax[a] = span{rn.b[2 * a], rn.b[2 * a + 1]};
                break;
            }
            case MPT_OBJ_MEDIUM:
                if (r.material < 0 || r.material >= d.n_materials || s.materials[r.material].type != MPT_MAT_PHASE)
                    return fail(MPT_E_INVALID, "medium needs a volume (phase) material");
                if (!(r.density > 0)) return fail(MPT_E_INVALID, "medium density must be positive");
                if (r.bkind == MPT_OBJ_PRIM && r.count != 1)
                    return fail(MPT_E_INVALID, "medium boundary prim with count != 1");
                for (int64_t k = 0; k < r.count; k++) {
                    if (s.prim_kind[r.first + k] == MPT_PRIM_TRIANGLE || s.prim_kind[r.first + k] == MPT_PRIM_VOXBLOCK)
                        return fail(MPT_E_UNSUPPORTED, "mesh and voxel boundaries of media are not supported");
                    b = box_merge(b, pbox[r.first + k]);
                }
                break;
            default:
                return fail(MPT_E_UNSUPPORTED, "unsupported object kind " + std::to_string(r.kind));
        }
        if (r.xform >= 0) b = instance_box(s.xforms[r.xform], b);
        obox[o] = b;
    }
    for (int64_t i = 0; i < np; i++)
        if (s.prim_canon[i] < 0) s.prim_canon[i] = 0;  // fixed below for non-mesh ranges
    for (int o = 0; o < no; o++) {
        const obj_rec& r = s.objects[o];
        if (r.kind != MPT_OBJ_MESH && r.kind != MPT_OBJ_VOXEL)
            for (int64_t k = 0; k < r.count; k++) s.prim_canon[r.first + k] = k;
    }

    // top-level BVH over the world list (main.cpp:70, bvh.h:24-43); objects
    // that are only lights stay out of it
    std::vector<build_item> items;
    for (int o = 0; o < no; o++) {
        if (s.objects[o].flags & MPT_OBJF_LIGHT_ONLY) continue;
        build_item it;
        it.box = obox[o];
        it.centroid = box_centroid(obox[o]);
        it.ref = o;
        items.push_back(it);
    }
    const int nw = int(items.size());
    s.ctlas.clear();
    s.tlas.clear();
    if (nw > 0) {  // an empty world has no tree: the walk never starts
        build_bvh(items, s.tlas);
        pair_layout(s.tlas, 0, s.ctlas, s.ctlas_root);
    }
    s.qtlas.clear();
    s.qblas.clear();
    if (nw > 0 && s.ctlas_root.cnt == 0) s.ctlas_root.qoff = quad_layout(s.ctlas, s.ctlas_root.off, s.qtlas);
    for (croot& r : s.cblas_roots)
        if (r.cnt == 0) r.qoff = quad_layout(s.cblas, r.off, s.qblas);
    s.tlas_items.resize(nw);
    for (int k = 0; k < nw; k++) s.tlas_items[k] = int32_t(items[k].ref);
    s.world_box = box_empty();
    for (int k = 0; k < nw; k++) s.world_box = box_merge(s.world_box, obox[items[k].ref]);

    // environment (environment.h:140-177)
    s.env_w = d.env_width;
    s.env_h = d.env_height;
    s.env_intensity = d.env_intensity;
    s.env_rotation = d.env_rotation_degrees / 360;
    if (s.env_w > 0) {
        if (s.env_h <= 0 || !d.env_rgb) return fail(MPT_E_INVALID, "environment image missing");
        const int w = s.env_w, h = s.env_h;
        s.env_rgb.resize(size_t(w) * h * 3);
        for (size_t k = 0; k < s.env_rgb.size(); k++) s.env_rgb[k] = double(d.env_rgb[k]);
        auto weight = [&](int i, int j) {
            double sin_theta = std::sin(kPi * (j + 0.5) / h);
            const double* p = &s.env_rgb[(size_t(j) * w + i) * 3];
            return luminance(v3{p[0], p[1], p[2]}) * sin_theta;
        };
        s.env_cond_cdf.assign(size_t(h) * (w + 1), 0.0);
        s.env_marg_cdf.assign(h + 1, 0.0);
        s.env_row_weight.assign(h, 0.0);
        for (int j = 0; j < h; j++) {
            double* cdf = &s.env_cond_cdf[size_t(j) * (w + 1)];
            for (int i = 0; i < w; i++) cdf[i + 1] = cdf[i] + std::fmax(weight(i, j), 0.0);
            s.env_row_weight[j] = cdf[w];
            for (int i = 1; i <= w; i++) cdf[i] = s.env_row_weight[j] > 0 ? cdf[i] / s.env_row_weight[j] : double(i) / w;
            s.env_marg_cdf[j + 1] = s.env_marg_cdf[j] + s.env_row_weight[j];
        }
        s.env_total_weight = s.env_marg_cdf[h];
        if (s.env_total_weight <= 0) {
            for (int j = 0; j < h; j++) {
                s.env_row_weight[j] = std::sin(kPi * (j + 0.5) / h);
                s.env_marg_cdf[j + 1] = s.env_marg_cdf[j] + s.env_row_weight[j];
            }
            s.env_total_weight = s.env_marg_cdf[h];
        }
        for (int j = 1; j <= h; j++) s.env_marg_cdf[j] /= s.env_total_weight;
    }

    // the sun (sun_light's constructor, sun.h:26-34)
    s.has_sun = d.has_sun != 0;
    if (s.has_sun) {
        v3 dir{d.sun_direction[0], d.sun_direction[1], d.sun_direction[2]};
        if (!(length_squared(dir) > 0) || !std::isfinite(length_squared(dir)))
            return fail(MPT_E_INVALID, "sun direction must be finite and non-zero");
        s.sun_frame = make_onb(dir);
        double half = std::fmax(d.sun_angle_degrees, 1e-3) * kPi / 180.0 / 2;
        s.sun_cos_max = std::cos(half);
        s.sun_solid_angle = 2 * kPi * (1 - s.sun_cos_max);
        s.sun_irradiance = v3{d.sun_irradiance[0], d.sun_irradiance[1], d.sun_irradiance[2]};
        s.sun_radiance = s.sun_irradiance / s.sun_solid_angle;
    }

    // light selection (light_list.h:45-96)
    const int nl = d.n_lights;
    s.light_strategy = d.light_strategy;
    s.lights.resize(nl);
    for (int k = 0; k < nl; k++) {
        int ob = d.light_object[k];
        if (ob < -2 || ob >= no) return fail(MPT_E_INVALID, "light refers to a bad object");
        if (ob == -2 && !s.has_sun) return fail(MPT_E_INVALID, "sun light without a sun");
        if (ob == -1 && s.env_w <= 0) return fail(MPT_E_INVALID, "environment light without an environment");
        if (ob >= 0) {
            const obj_rec& r = s.objects[ob];
            if (r.kind == MPT_OBJ_MEDIUM) return fail(MPT_E_UNSUPPORTED, "media cannot be lights");
            if (r.kind == MPT_OBJ_VOXEL)
                return fail(MPT_E_UNSUPPORTED, "a voxel grid cannot be a light: its emissive faces are (scene.h:1118-1128)");
            if (r.kind == MPT_OBJ_MESH && s.objects[ob].light_table < 0) {
                // the mesh's triangles in its own order and their area CDF
                // (mesh.h:180-186); instances of one mesh share the table
                int table = -1;
                for (int o2 = 0; o2 < no && table < 0; o2++)
                    if (s.objects[o2].kind == MPT_OBJ_MESH && s.objects[o2].first == r.first && s.objects[o2].light_table >= 0)
                        table = s.objects[o2].light_table;
                if (table < 0) {
                    table = int32_t(s.mlight_prim.size());
                    s.mlight_prim.resize(size_t(table + r.count));
                    s.mlight_cdf.resize(size_t(table + r.count));
                    for (int64_t k = 0; k < r.count; k++) s.mlight_prim[size_t(table + s.prim_canon[r.first + k])] = r.first + k;
                    double total = 0;
                    for (int64_t k = 0; k < r.count; k++) {
                        total += s.prim[size_t(s.mlight_prim[size_t(table + k)]) * kPrimStride + 12];
                        s.mlight_cdf[size_t(table + k)] = total;
                    }
                }
                s.objects[ob].light_table = table;
            }
        }
        s.lights[k].object = ob;
        s.lights[k].pad = 0;
        s.lights[k].probability = nl > 0 ? 1.0 / nl : 0.0;
    }
    box3 lb = s.world_box;
    if (d.has_light_bounds)
        for (int a = 0; a < 3; a++) lb.ax[a] = span{d.light_bounds[2 * a], d.light_bounds[2 * a + 1]};
    light_selection(s, lb, d.light_uniform_fraction, nullptr);

    s.spheres.assign(size_t(np) * 4, 0.0);
    for (int64_t i = 0; i < np; i++)
        for (int k = 0; k < 4; k++) s.spheres[size_t(i) * 4 + k] = s.prim[size_t(i) * kPrimStride + k];
    build_ggx_table(s.ggx_e, s.ggx_e_avg);
    return MPT_OK;
}

}  // namespace mpt
