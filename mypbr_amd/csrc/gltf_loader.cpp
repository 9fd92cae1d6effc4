// gltf_loader.cpp - glTF 2.0 models (.gltf with its buffers and images, or a
// single .glb) as mesh parts with this core's materials, the mapping of the
// reference's load_gltf (gltf_loader.h:21-57). The reference reads files with
// the vendored cgltf 1.14; this is a reader of its own from the glTF 2.0
// specification (Khronos), with cgltf's arithmetic where results depend on it:
// a node's local matrix from translation, rotation and scale and the chain to
// the root are composed in single precision as cgltf_node_transform_local /
// cgltf_node_transform_world compute them (the published quaternion matrix,
// then each parent's 3x3 part and translation), and normalized integer
// attributes are divided in single precision.
//
// Supported: triangle primitives (indexed or not) with POSITION, NORMAL and
// TEXCOORD_0 (float, or normalized unsigned byte / short); indices of
// unsigned byte, short or int; buffers from .bin files, data: URIs (base64)
// or the GLB binary chunk; images from buffer views (PNG, JPEG) or files
// (percent-decoded URIs); nodes by matrix or TRS, the default scene (or every
// root); pbrMetallicRoughness, normalTexture, emissive factor and texture,
// alphaMode MASK / BLEND, KHR_materials_ior, _transmission, _clearcoat,
// _emissive_strength and KHR_texture_transform. What the reference leaves
// out (sheen, iridescence, volume, specular, anisotropy, unlit, other
// texture coordinate sets, sparse accessors) is left out and listed once.

#include "../../include/mpt/mpt.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <set>
#include <sstream>

namespace mpt {
namespace {

// ------------------------------------------------------------------ JSON
struct jv {
    enum kind_t { null_v, bool_v, num_v, str_v, arr_v, obj_v } kind = null_v;
    bool b = false;
    double n = 0;
    std::string s;
    std::vector<jv> a;
    std::vector<std::pair<std::string, jv>> o;

    const jv* get(const char* key) const {
        if (kind != obj_v) return nullptr;
        for (const auto& kv : o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    double num(const char* key, double fallback) const {
        const jv* v = get(key);
        return v && v->kind == num_v ? v->n : fallback;
    }
    int idx(const char* key) const {  // an index, -1 when absent
        const jv* v = get(key);
        return v && v->kind == num_v ? int(v->n) : -1;
    }
    std::string str(const char* key) const {
        const jv* v = get(key);
        return v && v->kind == str_v ? v->s : std::string();
    }
    size_t size() const { return kind == arr_v ? a.size() : 0; }
};

struct json_parser {
    const char* p;
    const char* end;
    std::string err;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    bool fail(const char* what) {
        if (err.empty()) err = what;
        return false;
    }
    static void put_utf8(std::string& out, uint32_t c) {
        if (c < 0x80) {
            out += char(c);
        } else if (c < 0x800) {
            out += char(0xC0 | (c >> 6));
            out += char(0x80 | (c & 63));
        } else if (c < 0x10000) {
            out += char(0xE0 | (c >> 12));
            out += char(0x80 | ((c >> 6) & 63));
            out += char(0x80 | (c & 63));
        } else {
            out += char(0xF0 | (c >> 18));
            out += char(0x80 | ((c >> 12) & 63));
            out += char(0x80 | ((c >> 6) & 63));
            out += char(0x80 | (c & 63));
        }
    }
    bool hex4(uint32_t& c) {
        if (end - p < 4) return fail("truncated \\u escape");
        c = 0;
        for (int k = 0; k < 4; k++) {
            const char h = *p++;
            c <<= 4;
            if (h >= '0' && h <= '9') c |= uint32_t(h - '0');
            else if (h >= 'a' && h <= 'f') c |= uint32_t(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= uint32_t(h - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& out) {
        p++;  // the opening quote
        while (p < end && *p != '"') {
            if (*p != '\\') {
                out += *p++;
                continue;
            }
            if (++p >= end) return fail("truncated escape");
            const char e = *p++;
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t c;
                    if (!hex4(c)) return false;
                    if (c >= 0xD800 && c < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        uint32_t lo;
                        if (!hex4(lo)) return false;
                        c = 0x10000 + ((c - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, c);
                    break;
                }
                default: return fail("bad escape");
            }
        }
        if (p >= end) return fail("unterminated string");
        p++;
        return true;
    }
    bool value(jv& v, int depth) {
        if (depth > 256) return fail("nested too deeply");
        ws();
        if (p >= end) return fail("unexpected end");
        const char c = *p;
        if (c == '{') {
            v.kind = jv::obj_v;
            p++;
            ws();
            if (p < end && *p == '}') {
                p++;
                return true;
            }
            while (true) {
                ws();
                if (p >= end || *p != '"') return fail("expected a key");
                std::string key;
                if (!string(key)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                p++;
                v.o.emplace_back(std::move(key), jv());
                if (!value(v.o.back().second, depth + 1)) return false;
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == '}') {
                    p++;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = jv::arr_v;
            p++;
            ws();
            if (p < end && *p == ']') {
                p++;
                return true;
            }
            while (true) {
                v.a.emplace_back();
                if (!value(v.a.back(), depth + 1)) return false;
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == ']') {
                    p++;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = jv::str_v;
            return string(v.s);
        }
        auto word = [&](const char* w) {
            const size_t n = std::strlen(w);
            if (size_t(end - p) >= n && std::memcmp(p, w, n) == 0) {
                p += n;
                return true;
            }
            return false;
        };
        if (word("true")) {
            v.kind = jv::bool_v;
            v.b = true;
            return true;
        }
        if (word("false")) {
            v.kind = jv::bool_v;
            return true;
        }
        if (word("null")) return true;
        // a number: strtod on a bounded copy
        const char* q = p;
        while (q < end && (std::strchr("+-0123456789.eE", *q) != nullptr)) q++;
        if (q == p) return fail("unexpected character");
        const std::string num(p, q);
        char* stop = nullptr;
        v.n = std::strtod(num.c_str(), &stop);
        if (stop != num.c_str() + num.size()) return fail("bad number");
        v.kind = jv::num_v;
        p = q;
        return true;
    }
};

bool base64(const std::string& in, size_t from, std::vector<uint8_t>& out) {
    uint32_t acc = 0;
    int bits = 0;
    for (size_t k = from; k < in.size(); k++) {
        const char c = in[k];
        int v;
        if (c >= 'A' && c <= 'Z') v = c - 'A';
        else if (c >= 'a' && c <= 'z') v = c - 'a' + 26;
        else if (c >= '0' && c <= '9') v = c - '0' + 52;
        else if (c == '+' || c == '-') v = 62;
        else if (c == '/' || c == '_') v = 63;
        else if (c == '=') break;
        else if (c == ' ' || c == '\n' || c == '\r') continue;
        else return false;
        acc = (acc << 6) | uint32_t(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(uint8_t((acc >> bits) & 0xFF));
        }
    }
    return true;
}

std::string percent_decode(const std::string& s) {  // %20 and friends
    std::string out;
    for (size_t k = 0; k < s.size(); k++) {
        if (s[k] == '%' && k + 2 < s.size() && std::isxdigit(uint8_t(s[k + 1])) && std::isxdigit(uint8_t(s[k + 2]))) {
            out += char(std::stoi(s.substr(k + 1, 2), nullptr, 16));
            k += 2;
        } else {
            out += s[k];
        }
    }
    return out;
}

// ------------------------------------------------------------------ the file
struct gltf_file {
    jv doc;
    std::vector<std::vector<uint8_t>> buffers;
    std::string dir;

    const jv& arr(const char* key) const {
        static const jv none;
        const jv* v = doc.get(key);
        return v && v->kind == jv::arr_v ? *v : none;
    }
    // a buffer view's bytes, or null when out of range
    const uint8_t* view(int vi, size_t& size) const {
        const jv& views = arr("bufferViews");
        if (vi < 0 || size_t(vi) >= views.size()) return nullptr;
        const jv& v = views.a[size_t(vi)];
        const int b = v.idx("buffer");
        const size_t off = size_t(v.num("byteOffset", 0)), len = size_t(v.num("byteLength", 0));
        if (b < 0 || size_t(b) >= buffers.size() || off + len > buffers[size_t(b)].size()) return nullptr;
        size = len;
        return buffers[size_t(b)].data() + off;
    }
};

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

bool open_gltf(const std::string& path, gltf_file& g, std::string& error) {
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) {
        error = "cannot read glTF file '" + path + "'";
        return false;
    }
    auto slash = path.find_last_of('/');
    g.dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    auto u32 = [&](size_t at) {
        return uint32_t(bytes[at]) | uint32_t(bytes[at + 1]) << 8 | uint32_t(bytes[at + 2]) << 16 |
               uint32_t(bytes[at + 3]) << 24;
    };
    const char* json_begin = reinterpret_cast<const char*>(bytes.data());
    const char* json_end = json_begin + bytes.size();
    std::vector<uint8_t> bin;
    bool have_bin = false;
    if (bytes.size() >= 12 && std::memcmp(bytes.data(), "glTF", 4) == 0) {  // GLB: header, JSON chunk, BIN chunk
        if (u32(4) != 2) {
            error = "'" + path + "' is a GLB file of an unsupported version";
            return false;
        }
        size_t at = 12;
        bool have_json = false;
        while (at + 8 <= bytes.size()) {
            const uint32_t len = u32(at), type = u32(at + 4);
            if (at + 8 + size_t(len) > bytes.size()) {
                error = "'" + path + "' has a truncated GLB chunk";
                return false;
            }
            if (type == 0x4E4F534Au && !have_json) {  // "JSON"
                json_begin = reinterpret_cast<const char*>(bytes.data() + at + 8);
                json_end = json_begin + len;
                have_json = true;
            } else if (type == 0x004E4942u && !have_bin) {  // "BIN\0"
                bin.assign(bytes.begin() + long(at + 8), bytes.begin() + long(at + 8 + len));
                have_bin = true;
            }
            at += 8 + size_t(len);
        }
        if (!have_json) {
            error = "'" + path + "' has no JSON chunk";
            return false;
        }
    }
    json_parser jp{json_begin, json_end, {}};
    if (!jp.value(g.doc, 0) || g.doc.kind != jv::obj_v) {
        error = "cannot read glTF file '" + path + "'" + (jp.err.empty() ? "" : ": " + jp.err);
        return false;
    }
    const jv& bufs = g.arr("buffers");
    for (size_t k = 0; k < bufs.size(); k++) {
        const jv& b = bufs.a[k];
        const std::string uri = b.str("uri");
        std::vector<uint8_t> data;
        if (uri.empty()) {
            if (!have_bin || k != 0) {
                error = "cannot load the buffers of '" + path + "'";
                return false;
            }
            data = bin;
        } else if (uri.rfind("data:", 0) == 0) {
            const size_t comma = uri.find(',');
            if (comma == std::string::npos || uri.find(";base64") == std::string::npos || comma < uri.find(";base64") ||
                !base64(uri, comma + 1, data)) {
                error = "cannot load the buffers of '" + path + "'";
                return false;
            }
        } else {
            const std::string file = percent_decode(uri);
            if (!read_file(file[0] == '/' ? file : g.dir + file, data)) {
                error = "cannot load the buffers of '" + path + "'";
                return false;
            }
        }
        const size_t need = size_t(b.num("byteLength", 0));
        if (data.size() < need) {
            error = "cannot load the buffers of '" + path + "'";
            return false;
        }
        g.buffers.push_back(std::move(data));
    }
    return true;
}

// An accessor's elements as floats (normalized integers divided in single
// precision, as the specification's decoding gives them), or its indices.
struct accessor {
    const uint8_t* base = nullptr;
    size_t count = 0, stride = 0;
    int comp = 0;  // componentType
    int ncomp = 0;
    bool normalized = false;

    static int comp_size(int c) { return c == 5120 || c == 5121 ? 1 : c == 5122 || c == 5123 ? 2 : 4; }
    float read_comp(const uint8_t* p) const {
        switch (comp) {
            case 5126: {
                float f;
                std::memcpy(&f, p, 4);
                return f;
            }
            case 5121: return normalized ? float(p[0]) / 255.0f : float(p[0]);
            case 5123: {
                uint16_t v;
                std::memcpy(&v, p, 2);
                return normalized ? float(v) / 65535.0f : float(v);
            }
            case 5120: {
                const int8_t v = int8_t(p[0]);
                return normalized ? float(v) / 127.0f : float(v);
            }
            case 5122: {
                int16_t v;
                std::memcpy(&v, p, 2);
                return normalized ? float(v) / 32767.0f : float(v);
            }
            default: {
                uint32_t v;
                std::memcpy(&v, p, 4);
                return float(v);
            }
        }
    }
    void read_float(size_t k, float* out, int n) const {
        const uint8_t* p = base + k * stride;
        for (int c = 0; c < n; c++) out[c] = c < ncomp ? read_comp(p + size_t(c) * size_t(comp_size(comp))) : 0.0f;
    }
    size_t read_index(size_t k) const {
        const uint8_t* p = base + k * stride;
        if (comp == 5121) return p[0];
        if (comp == 5123) {
            uint16_t v;
            std::memcpy(&v, p, 2);
            return v;
        }
        uint32_t v;
        std::memcpy(&v, p, 4);
        return v;
    }
};

bool get_accessor(const gltf_file& g, int ai, accessor& out, std::set<std::string>& ignored) {
    const jv& acc = g.arr("accessors");
    if (ai < 0 || size_t(ai) >= acc.size()) return false;
    const jv& a = acc.a[size_t(ai)];
    if (a.get("sparse")) {
        ignored.insert("sparse accessors");
        return false;
    }
    static const std::pair<const char*, int> types[] = {{"SCALAR", 1}, {"VEC2", 2}, {"VEC3", 3}, {"VEC4", 4}};
    out.ncomp = 0;
    for (const auto& t : types)
        if (a.str("type") == t.first) out.ncomp = t.second;
    out.comp = a.idx("componentType");
    out.count = size_t(a.num("count", 0));
    const jv* norm = a.get("normalized");
    out.normalized = norm && norm->kind == jv::bool_v && norm->b;
    if (out.ncomp == 0 || (out.comp < 5120 || out.comp > 5126 || out.comp == 5124)) return false;
    const size_t elem = size_t(accessor::comp_size(out.comp)) * size_t(out.ncomp);
    const int vi = a.idx("bufferView");
    if (vi < 0) return false;  // all zeros: no geometry to take
    size_t size = 0;
    const uint8_t* view = g.view(vi, size);
    const jv& bv = g.arr("bufferViews").a[size_t(vi)];
    out.stride = size_t(bv.num("byteStride", 0));
    if (out.stride == 0) out.stride = elem;
    const size_t off = size_t(a.num("byteOffset", 0));
    if (!view || (out.count > 0 && off + (out.count - 1) * out.stride + elem > size)) return false;
    out.base = view + off;
    return true;
}

// cgltf_node_transform_local: a node's matrix (column-major), or its TRS
// composed in single precision from the unit quaternion's rotation matrix.
void node_local(const jv& node, float* lm) {
    const jv* m = node.get("matrix");
    if (m && m->size() == 16) {
        for (int k = 0; k < 16; k++) lm[k] = float(m->a[size_t(k)].n);
        return;
    }
    float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
    if (const jv* v = node.get("translation"))
        for (size_t k = 0; k < 3 && k < v->size(); k++) t[k] = float(v->a[k].n);
    if (const jv* v = node.get("rotation"))
        for (size_t k = 0; k < 4 && k < v->size(); k++) q[k] = float(v->a[k].n);
    if (const jv* v = node.get("scale"))
        for (size_t k = 0; k < 3 && k < v->size(); k++) s[k] = float(v->a[k].n);
    const float qx = q[0], qy = q[1], qz = q[2], qw = q[3];
    lm[0] = (1 - 2 * qy * qy - 2 * qz * qz) * s[0];
    lm[1] = (2 * qx * qy + 2 * qz * qw) * s[0];
    lm[2] = (2 * qx * qz - 2 * qy * qw) * s[0];
    lm[3] = 0.f;
    lm[4] = (2 * qx * qy - 2 * qz * qw) * s[1];
    lm[5] = (1 - 2 * qx * qx - 2 * qz * qz) * s[1];
    lm[6] = (2 * qy * qz + 2 * qx * qw) * s[1];
    lm[7] = 0.f;
    lm[8] = (2 * qx * qz + 2 * qy * qw) * s[2];
    lm[9] = (2 * qy * qz - 2 * qx * qw) * s[2];
    lm[10] = (1 - 2 * qx * qx - 2 * qy * qy) * s[2];
    lm[11] = 0.f;
    lm[12] = t[0];
    lm[13] = t[1];
    lm[14] = t[2];
    lm[15] = 1.f;
}

// cgltf_node_transform_world: the local matrix, then each ancestor's 3x3
// part applied to its columns and the ancestor's translation added.
void node_world(const gltf_file& g, const std::vector<int>& parent, int ni, float* lm) {
    const jv& nodes = g.arr("nodes");
    node_local(nodes.a[size_t(ni)], lm);
    int depth = 0;
    for (int pi = parent[size_t(ni)]; pi >= 0 && depth < 1024; pi = parent[size_t(pi)], depth++) {
        float pm[16];
        node_local(nodes.a[size_t(pi)], pm);
        for (int i = 0; i < 4; ++i) {
            const float l0 = lm[i * 4 + 0], l1 = lm[i * 4 + 1], l2 = lm[i * 4 + 2];
            const float r0 = l0 * pm[0] + l1 * pm[4] + l2 * pm[8];
            const float r1 = l0 * pm[1] + l1 * pm[5] + l2 * pm[9];
            const float r2 = l0 * pm[2] + l1 * pm[6] + l2 * pm[10];
            lm[i * 4 + 0] = r0;
            lm[i * 4 + 1] = r1;
            lm[i * 4 + 2] = r2;
        }
        lm[12] += pm[12];
        lm[13] += pm[13];
        lm[14] += pm[14];
    }
}

}  // namespace

bool load_gltf(const std::string& path, std::shared_ptr<material> override_mat, std::vector<obj_part>& parts,
               std::string& error, bool use_normal_maps, const std::map<std::string, std::shared_ptr<material>>& replace) {
    gltf_file g;
    if (!open_gltf(path, g, error)) return false;
    std::set<std::string> ignored;  // left out, listed once
    const jv& textures = g.arr("textures");
    const jv& images = g.arr("images");
    const jv& samplers = g.arr("samplers");
    const jv& materials = g.arr("materials");

    // images decoded once per use: colours as sRGB, data (normals, metallic-roughness) linear
    std::map<std::pair<int, bool>, std::shared_ptr<image_texture>> image_cache;
    auto load_image = [&](int ti, bool linear) -> std::shared_ptr<image_texture> {
        if (ti < 0 || size_t(ti) >= textures.size()) return nullptr;
        const jv& t = textures.a[size_t(ti)];
        const int ii = t.idx("source");
        if (ii < 0 || size_t(ii) >= images.size()) return nullptr;
        const auto key = std::make_pair(ti, linear);
        auto it = image_cache.find(key);
        if (it != image_cache.end()) return it->second;
        // samplers repeat unless both directions clamp (33071 = CLAMP_TO_EDGE)
        const int si = t.idx("sampler");
        bool repeat = true;
        if (si >= 0 && size_t(si) < samplers.size()) {
            const jv& smp = samplers.a[size_t(si)];
            repeat = !(smp.num("wrapS", 10497) == 33071 && smp.num("wrapT", 10497) == 33071);
        }
        const jv& im = images.a[size_t(ii)];
        std::shared_ptr<image_texture> result;
        const int vi = im.idx("bufferView");
        const std::string uri = im.str("uri");
        if (vi >= 0) {
            size_t size = 0;
            if (const uint8_t* data = g.view(vi, size)) result = std::make_shared<image_texture>(data, size, repeat, linear);
        } else if (!uri.empty() && uri.rfind("data:", 0) != 0) {
            const std::string file = percent_decode(uri);
            result = std::make_shared<image_texture>(file[0] == '/' ? file : g.dir + file, repeat, linear);
        }
        if (!result || !result->loaded()) {
            const std::string name = !uri.empty() ? uri : (!im.str("name").empty() ? im.str("name") : "embedded");
            std::fprintf(stderr, "Warning: %s: cannot load image '%s'\n", path.c_str(), name.c_str());
            result = nullptr;
        }
        image_cache[key] = result;
        return result;
    };
    // a texture view: the image, moved by KHR_texture_transform
    auto view_texture = [&](const jv* info, bool linear) -> std::shared_ptr<texture> {
        if (!info) return nullptr;
        const jv* ext = info->get("extensions");
        const jv* tr = ext ? ext->get("KHR_texture_transform") : nullptr;
        int texcoord = info->idx("texCoord");
        if (texcoord < 0) texcoord = 0;
        if (tr && tr->get("texCoord")) texcoord = tr->idx("texCoord");
        if (texcoord != 0) {
            ignored.insert("second texture coordinates");
            return nullptr;
        }
        std::shared_ptr<texture> result = load_image(info->idx("index"), linear);
        if (result && tr) {
            double off[2] = {0, 0}, sc[2] = {1, 1};
            if (const jv* o = tr->get("offset"))
                for (size_t k = 0; k < 2 && k < o->size(); k++) off[k] = double(float(o->a[k].n));
            if (const jv* o = tr->get("scale"))
                for (size_t k = 0; k < 2 && k < o->size(); k++) sc[k] = double(float(o->a[k].n));
            result = std::make_shared<uv_transform_texture>(result, off[0], off[1], double(float(tr->num("rotation", 0))),
                                                            sc[0], sc[1]);
        }
        return result;
    };
    // a name to replace that the file does not have is a typo, not a no-op
    for (const auto& kv : replace) {
        bool found = false;
        for (size_t k = 0; k < materials.size() && !found; k++) found = materials.a[k].str("name") == kv.first;
        if (!found) {
            error = "'" + path + "' has no material named '" + kv.first + "'";
            return false;
        }
    }

    using surface = std::pair<std::shared_ptr<material>, std::shared_ptr<image_texture>>;
    std::map<int, surface> converted;
    auto f32 = [](double x) { return double(float(x)); };  // glTF numbers as single precision, as cgltf keeps them
    auto convert = [&](int mi) -> surface {
        auto it = converted.find(mi);
        if (it != converted.end()) return it->second;
        surface result;
        if (mi < 0 || size_t(mi) >= materials.size()) {  // the specification's default: white, metallic 1, roughness 1
            result.first = override_mat ? override_mat
                                        : std::make_shared<principled>(std::make_shared<solid_color>(color(1, 1, 1)),
                                                                       1.0, 1.0, 1.5);
            return converted[mi] = result;
        }
        const jv& m = materials.a[size_t(mi)];
        if (!m.str("name").empty()) {
            auto r = replace.find(m.str("name"));
            if (r != replace.end()) {
                result.first = r->second;
                return converted[mi] = result;
            }
        }
        static const jv empty_obj = [] {
            jv e;
            e.kind = jv::obj_v;
            return e;
        }();
        const jv* pbr_p = m.get("pbrMetallicRoughness");
        const jv& pbr = pbr_p ? *pbr_p : empty_obj;
        const jv* ext_p = m.get("extensions");
        const jv& ext = ext_p ? *ext_p : empty_obj;
        double bcf[4] = {1, 1, 1, 1};
        if (const jv* v = pbr.get("baseColorFactor"))
            for (size_t k = 0; k < 4 && k < v->size(); k++) bcf[k] = f32(v->a[k].n);
        const color base_factor(bcf[0], bcf[1], bcf[2]);
        const jv* ior_ext = ext.get("KHR_materials_ior");
        const double ior = ior_ext ? f32(ior_ext->num("ior", 1.5)) : 1.5;
        const jv* bct = pbr.get("baseColorTexture");
        std::shared_ptr<texture> base_image = view_texture(bct, false);
        const std::string mode = m.str("alphaMode").empty() ? "OPAQUE" : m.str("alphaMode");
        const bool bct_transform = bct && bct->get("extensions") && bct->get("extensions")->get("KHR_texture_transform");
        if (mode != "OPAQUE" && !bct_transform) {
            auto image = load_image(bct ? bct->idx("index") : -1, false);
            double alpha = bcf[3];
            if (image && image->has_alpha_channel() && !image->fully_opaque()) {
                result.second = image;
                if (alpha < 1) ignored.insert("alpha factor times an alpha texture (the texture is used alone)");
            } else if (!image && alpha < 1) {
                // a constant coverage: MASK cuts it at alphaCutoff, BLEND keeps it
                if (mode == "MASK") alpha = alpha >= f32(m.num("alphaCutoff", 0.5)) ? 1 : 0;
                if (alpha < 1) result.second = std::make_shared<image_texture>(alpha);
            }
        }
        if (override_mat) {
            result.first = override_mat;
            return converted[mi] = result;
        }
        for (const char* e : {"KHR_materials_sheen", "KHR_materials_iridescence", "KHR_materials_volume",
                              "KHR_materials_specular", "KHR_materials_anisotropy", "KHR_materials_unlit"})
            if (ext.get(e)) ignored.insert(std::string(e).substr(14));
        std::shared_ptr<texture> base =
            base_image ? std::shared_ptr<texture>(std::make_shared<scaled_texture>(base_image, base_factor))
                       : std::make_shared<solid_color>(base_factor);
        const double metallic_factor = f32(pbr.num("metallicFactor", 1)), roughness_factor = f32(pbr.num("roughnessFactor", 1));
        std::shared_ptr<material> mat;
        const jv* tr_ext = ext.get("KHR_materials_transmission");
        if (tr_ext && f32(tr_ext->num("transmissionFactor", 0)) > 0) {
            if (f32(tr_ext->num("transmissionFactor", 0)) < 1 || roughness_factor > 0)
                ignored.insert("partial or rough transmission");
            mat = std::make_shared<thin_dielectric>(ior, base_factor);
        } else {
            std::shared_ptr<texture> metallic, roughness;
            if (auto mr = view_texture(pbr.get("metallicRoughnessTexture"), true)) {
                metallic = std::make_shared<channel_texture>(mr, 2, metallic_factor);
                roughness = std::make_shared<channel_texture>(mr, 1, roughness_factor);
            } else {
                metallic = std::make_shared<solid_color>(color(metallic_factor, metallic_factor, metallic_factor));
                roughness = std::make_shared<solid_color>(color(roughness_factor, roughness_factor, roughness_factor));
            }
            auto p = std::make_shared<principled>(base, 0.0, 0.0, ior);
            p->metallic_map = metallic;
            p->roughness_map = roughness;
            auto normal_view = [&](const jv* info) -> std::shared_ptr<texture> {  // with its scale
                if (!use_normal_maps || !info) return nullptr;
                auto normals = view_texture(info, true);
                const double scale = f32(info->num("scale", 1));
                if (normals && scale != 1) normals = std::make_shared<normal_scale_texture>(normals, scale);
                return normals;
            };
            p->normal_map = normal_view(m.get("normalTexture"));
            const jv* cc = ext.get("KHR_materials_clearcoat");
            if (cc && f32(cc->num("clearcoatFactor", 0)) > 0) {
                // weight in red, roughness in green, each times its factor; without a
                // clearcoatNormalTexture the coat is not bent
                const double cf = f32(cc->num("clearcoatFactor", 0)), crf = f32(cc->num("clearcoatRoughnessFactor", 0));
                std::shared_ptr<texture> weight, coat_roughness;
                if (auto t = view_texture(cc->get("clearcoatTexture"), true))
                    weight = std::make_shared<channel_texture>(t, 0, cf);
                else
                    weight = std::make_shared<solid_color>(color(1, 1, 1) * cf);
                if (auto t = view_texture(cc->get("clearcoatRoughnessTexture"), true))
                    coat_roughness = std::make_shared<channel_texture>(t, 1, crf);
                else
                    coat_roughness = std::make_shared<solid_color>(color(1, 1, 1) * crf);
                p->set_clearcoat(weight, coat_roughness, normal_view(cc->get("clearcoatNormalTexture")));
            }
            mat = p;
        }
        const jv* es = ext.get("KHR_materials_emissive_strength");
        const double strength = es ? f32(es->num("emissiveStrength", 1)) : 1;
        double ef[3] = {0, 0, 0};
        if (const jv* v = m.get("emissiveFactor"))
            for (size_t k = 0; k < 3 && k < v->size(); k++) ef[k] = f32(v->a[k].n);
        const color emission = strength * color(ef[0], ef[1], ef[2]);
        if (emission.length_squared() > 0) {
            if (const jv* et = m.get("emissiveTexture")) {
                if (auto image = view_texture(et, false))
                    mat = std::make_shared<emissive_surface>(mat, std::make_shared<scaled_texture>(image, emission));
            } else {
                mat = std::make_shared<diffuse_light>(emission);  // a light of the scene
            }
        }
        result.first = mat;
        return converted[mi] = result;
    };

    // the shared vertex arrays and the faces of each surface, in the order the
    // surfaces are first met
    std::vector<point3> positions;
    std::vector<vec3> normals, uvs;
    struct face {
        int v[3], n[3], t[3];
    };
    std::vector<surface> order;
    std::map<surface, std::vector<face>> faces_by_surface;

    const jv& meshes = g.arr("meshes");
    auto add_primitive = [&](const jv& prim, const float* m) {
        if (prim.num("mode", 4) != 4) {
            ignored.insert("non-triangle primitives");
            return;
        }
        const jv* ext = prim.get("extensions");
        if (ext && ext->get("KHR_draco_mesh_compression")) {
            ignored.insert("Draco-compressed meshes");
            return;
        }
        const jv* attrs = prim.get("attributes");
        if (!attrs) return;
        accessor pa, na, ta, ia;
        if (!get_accessor(g, attrs->idx("POSITION"), pa, ignored) || pa.ncomp != 3) return;
        const bool has_n = get_accessor(g, attrs->idx("NORMAL"), na, ignored) && na.count == pa.count;
        const bool has_t = get_accessor(g, attrs->idx("TEXCOORD_0"), ta, ignored) && ta.count == pa.count;
        // positions by the world matrix (column-major), normals by its inverse
        // transpose: the cofactor matrix, whose scale the normalisation drops
        const double a00 = m[0], a01 = m[4], a02 = m[8];
        const double a10 = m[1], a11 = m[5], a12 = m[9];
        const double a20 = m[2], a21 = m[6], a22 = m[10];
        const double c00 = a11 * a22 - a12 * a21, c01 = a12 * a20 - a10 * a22, c02 = a10 * a21 - a11 * a20;
        const double c10 = a02 * a21 - a01 * a22, c11 = a00 * a22 - a02 * a20, c12 = a01 * a20 - a00 * a21;
        const double c20 = a01 * a12 - a02 * a11, c21 = a02 * a10 - a00 * a12, c22 = a00 * a11 - a01 * a10;
        const double det = a00 * c00 + a01 * c01 + a02 * c02;
        const double sign = det < 0 ? -1 : 1;
        const int v_base = int(positions.size()), n_base = int(normals.size()), t_base = int(uvs.size());
        for (size_t k = 0; k < pa.count; k++) {
            float p[3];
            pa.read_float(k, p, 3);
            positions.push_back(point3(a00 * p[0] + a01 * p[1] + a02 * p[2] + m[12], a10 * p[0] + a11 * p[1] + a12 * p[2] + m[13],
                                       a20 * p[0] + a21 * p[1] + a22 * p[2] + m[14]));
        }
        if (has_n)
            for (size_t k = 0; k < pa.count; k++) {
                float n[3];
                na.read_float(k, n, 3);
                normals.push_back(sign * vec3(c00 * n[0] + c01 * n[1] + c02 * n[2], c10 * n[0] + c11 * n[1] + c12 * n[2],
                                              c20 * n[0] + c21 * n[1] + c22 * n[2]));
            }
        if (has_t)
            for (size_t k = 0; k < pa.count; k++) {
                float t[2];
                ta.read_float(k, t, 2);
                uvs.push_back(vec3(t[0], 1 - t[1], 0));  // glTF's v runs down the image
            }
        const surface surf = convert(prim.idx("material"));
        if (!faces_by_surface.count(surf)) order.push_back(surf);
        auto& faces = faces_by_surface[surf];
        const bool indexed = get_accessor(g, prim.idx("indices"), ia, ignored) && ia.ncomp == 1;
        const size_t index_count = indexed ? ia.count : pa.count;
        for (size_t k = 0; k + 2 < index_count; k += 3) {
            int idx[3];
            for (int c = 0; c < 3; c++) idx[c] = int(indexed ? ia.read_index(k + size_t(c)) : k + size_t(c));
            if (det < 0) std::swap(idx[1], idx[2]);  // a mirroring matrix flips the winding
            if (idx[0] >= int(pa.count) || idx[1] >= int(pa.count) || idx[2] >= int(pa.count) || idx[0] < 0 ||
                idx[1] < 0 || idx[2] < 0)
                continue;
            face f;
            for (int c = 0; c < 3; c++) {
                f.v[c] = v_base + idx[c];
                f.n[c] = has_n ? n_base + idx[c] : -1;
                f.t[c] = has_t ? t_base + idx[c] : -1;
            }
            const point3& p0 = positions[size_t(f.v[0])];
            if (cross(positions[size_t(f.v[1])] - p0, positions[size_t(f.v[2])] - p0).near_zero()) continue;  // degenerate
            faces.push_back(f);
        }
    };

    // the default scene's node tree (or every root), children pushed and taken last first
    const jv& nodes = g.arr("nodes");
    std::vector<int> parent(nodes.size(), -1);
    for (size_t k = 0; k < nodes.size(); k++)
        if (const jv* ch = nodes.a[k].get("children"))
            for (const jv& c : ch->a)
                if (c.kind == jv::num_v && c.n >= 0 && size_t(c.n) < nodes.size()) parent[size_t(c.n)] = int(k);
    std::vector<int> stack;
    const jv& scenes = g.arr("scenes");
    int si = g.doc.idx("scene");
    if (si < 0 && scenes.size() > 0) si = 0;
    if (si >= 0 && size_t(si) < scenes.size()) {
        if (const jv* roots = scenes.a[size_t(si)].get("nodes"))
            for (const jv& r : roots->a)
                if (r.kind == jv::num_v && r.n >= 0 && size_t(r.n) < nodes.size()) stack.push_back(int(r.n));
    } else {
        for (size_t k = 0; k < nodes.size(); k++)
            if (parent[k] < 0) stack.push_back(int(k));
    }
    size_t visits = 0;
    while (!stack.empty() && visits++ < 1000000) {
        const int ni = stack.back();
        stack.pop_back();
        const jv& node = nodes.a[size_t(ni)];
        if (const jv* ch = node.get("children"))
            for (const jv& c : ch->a)
                if (c.kind == jv::num_v && c.n >= 0 && size_t(c.n) < nodes.size()) stack.push_back(int(c.n));
        const int mi = node.idx("mesh");
        if (mi < 0 || size_t(mi) >= meshes.size()) continue;
        if (node.get("skin") || meshes.a[size_t(mi)].get("weights")) ignored.insert("skins and morph targets (the rest pose is used)");
        float matrix[16];
        node_world(g, parent, ni, matrix);
        if (const jv* prims = meshes.a[size_t(mi)].get("primitives"))
            for (const jv& prim : prims->a) add_primitive(prim, matrix);
    }

    for (const surface& surf : order) {
        const auto& faces = faces_by_surface[surf];
        if (faces.empty()) continue;
        // one mesh per surface over the shared vertex arrays
        auto me = std::make_shared<mesh>(positions, std::vector<int32_t>(), surf.first);
        me->indices.reserve(faces.size() * 3);
        bool any_n = false, any_t = false;
        for (const face& f : faces) {
            any_n = any_n || f.n[0] >= 0;
            any_t = any_t || f.t[0] >= 0;
        }
        if (any_n) me->normals = normals;
        if (any_t) me->uvs = uvs;
        for (const face& f : faces)
            for (int c = 0; c < 3; c++) {
                me->indices.push_back(f.v[c]);
                if (any_n) me->normal_indices.push_back(f.n[c]);
                if (any_t) me->uv_indices.push_back(f.t[c]);
            }
        me->alpha_mask = surf.second;
        parts.push_back({surf.first, me});
    }
    if (!ignored.empty()) {
        std::string list;
        for (const auto& f : ignored) list += (list.empty() ? " " : ", ") + f;
        std::fprintf(stderr, "Note: %s: not supported, left out:%s\n", path.c_str(), list.c_str());
    }
    if (parts.empty()) {
        error = "'" + path + "' has no triangles";
        return false;
    }
    return true;
}

}  // namespace mpt
