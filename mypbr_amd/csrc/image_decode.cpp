// PNG and baseline JPEG decoders for image textures. The reference reads
// textures with stb_image (texture.h:67-105, stbi_load with 4 channels); these
// give the same RGBA bytes. PNG is lossless, so any correct decoder agrees;
// its channel bookkeeping (tRNS, palettes, 16-bit samples) follows what
// stbi_load reports. JPEG is lossy in its integer arithmetic: the inverse DCT
// (the Loeffler-Ligtenberg-Moschytz factorisation of IJG's "islow" method, in
// 12-bit fixed point with rounding after each pass), the chroma upsampling
// (a 3:1 triangle filter) and the YCbCr conversion (20-bit fixed point) are
// restated step for step so that every byte equals stb_image's.
#include "image_decode.hpp"

#include <zlib.h>

#include <algorithm>
#include <array>
#include <cstring>

namespace mpt {
namespace {

// ------------------------------------------------------------------ PNG

uint32_t be32(const uint8_t* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

bool fail(std::string& err, const std::string& m) {
    err = m;
    return false;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

}  // namespace

bool decode_png(const std::vector<uint8_t>& f, ldr_image& out, std::string& err) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return fail(err, "PNG: bad signature");
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1;
    bool first = true, have_trans = false;
    std::array<uint8_t, 1024> pal{};
    size_t pal_len = 0;
    bool pal_alpha = false;
    uint16_t key[3] = {0, 0, 0};
    std::vector<uint8_t> idat;
    bool ended = false;
    while (!ended) {
        if (pos + 12 > f.size()) return fail(err, "PNG: truncated chunk");
        const uint32_t len = be32(&f[pos]);
        const uint32_t type = be32(&f[pos + 4]);
        if (len > f.size() - pos - 12) return fail(err, "PNG: truncated chunk");
        const uint8_t* d = &f[pos + 8];
        auto is = [&](const char* t) { return type == be32(reinterpret_cast<const uint8_t*>(t)); };
        if (first && !is("IHDR")) return fail(err, "PNG: first chunk is not IHDR");
        if (is("IHDR")) {
            if (!first || len != 13) return fail(err, "PNG: bad IHDR");
            first = false;
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            if (w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24)) return fail(err, "PNG: bad image size");
            const bool ok_depth = depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
            if (!ok_depth || (ctype != 0 && ctype != 2 && ctype != 3 && ctype != 4 && ctype != 6) ||
                (ctype == 3 && depth == 16) || ((ctype == 2 || ctype == 4 || ctype == 6) && depth < 8))
                return fail(err, "PNG: bad color type or bit depth");
            if (d[10] != 0 || d[11] != 0) return fail(err, "PNG: bad compression or filter method");
            if (d[12] != 0) return fail(err, "PNG: interlaced images are not supported");
        } else if (is("PLTE")) {
            if (len % 3 != 0 || len / 3 > 256 || len == 0) return fail(err, "PNG: bad PLTE");
            pal_len = len / 3;
            for (size_t i = 0; i < pal_len; i++) {
                pal[4 * i] = d[3 * i];
                pal[4 * i + 1] = d[3 * i + 1];
                pal[4 * i + 2] = d[3 * i + 2];
                pal[4 * i + 3] = 255;
            }
        } else if (is("tRNS")) {
            if (!idat.empty()) return fail(err, "PNG: tRNS after IDAT");
            if (ctype == 3) {
                if (pal_len == 0) return fail(err, "PNG: tRNS before PLTE");
                if (len > pal_len) return fail(err, "PNG: bad tRNS length");
                for (uint32_t i = 0; i < len; i++) pal[4 * i + 3] = d[i];
                pal_alpha = true;
            } else {
                const int n = ctype == 0 ? 1 : (ctype == 2 ? 3 : 0);
                if (n == 0) return fail(err, "PNG: tRNS on an image with alpha");
                if (len != uint32_t(2 * n)) return fail(err, "PNG: bad tRNS length");
                for (int k = 0; k < n; k++) key[k] = uint16_t((d[2 * k] << 8) | d[2 * k + 1]);
                have_trans = true;
            }
        } else if (is("IDAT")) {
            if (ctype == 3 && pal_len == 0) return fail(err, "PNG: no PLTE before IDAT");
            idat.insert(idat.end(), d, d + len);
        } else if (is("IEND")) {
            ended = true;
        } else if (!(f[pos + 4] & 0x20)) {
            return fail(err, "PNG: unknown critical chunk");
        }
        pos += 12 + size_t(len);
    }
    if (idat.empty()) return fail(err, "PNG: no IDAT");

    const int samples = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4;
    const size_t row_bytes = (size_t(w) * samples * depth + 7) / 8;
    const int bpp = std::max(1, samples * depth / 8);  // filter distance in bytes
    std::vector<uint8_t> raw(size_t(h) * (row_bytes + 1));
    {
        z_stream zs{};
        if (inflateInit(&zs) != Z_OK) return fail(err, "PNG: zlib init failed");
        zs.next_in = idat.data();
        zs.avail_in = uInt(idat.size());
        zs.next_out = raw.data();
        zs.avail_out = uInt(raw.size());
        const int rc = inflate(&zs, Z_FINISH);
        const size_t got = raw.size() - zs.avail_out;
        inflateEnd(&zs);
        if ((rc != Z_STREAM_END && rc != Z_BUF_ERROR && rc != Z_OK) || got < raw.size())
            return fail(err, "PNG: corrupt or short image data");
    }
    // undo the row filters in place (the row before the first reads as zeros)
    std::vector<uint8_t> px(size_t(h) * row_bytes);
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t filter = raw[y * (row_bytes + 1)];
        const uint8_t* src = &raw[y * (row_bytes + 1) + 1];
        uint8_t* cur = &px[y * row_bytes];
        const uint8_t* up = y ? &px[(y - 1) * row_bytes] : nullptr;
        if (filter > 4) return fail(err, "PNG: bad filter type");
        for (size_t i = 0; i < row_bytes; i++) {
            const int a = i >= size_t(bpp) ? cur[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= size_t(bpp)) ? up[i - bpp] : 0;
            int v = src[i];
            switch (filter) {
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: break;
            }
            cur[i] = uint8_t(v);
        }
    }
    // samples to RGBA bytes: 16-bit samples keep their high byte, grey below
    // 8 bits is scaled to 0..255, a tRNS key makes its color transparent
    static const int grey_scale[17] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1};
    auto sample = [&](const uint8_t* row, size_t x, int k) -> int {
        const size_t idx = x * samples + k;
        if (depth == 16) return (row[2 * idx] << 8) | row[2 * idx + 1];
        if (depth == 8) return row[idx];
        const size_t bit = idx * depth;
        return (row[bit >> 3] >> (8 - depth - int(bit & 7))) & ((1 << depth) - 1);
    };
    auto to8 = [&](int v) { return depth == 16 ? v >> 8 : v; };
    uint8_t key8[3];
    for (int k = 0; k < 3; k++) key8[k] = uint8_t((key[k] & 255) * grey_scale[depth]);
    out.width = int(w);
    out.height = int(h);
    out.rgba.assign(size_t(w) * h * 4, 0);
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t* row = &px[y * row_bytes];
        uint8_t* o = &out.rgba[size_t(y) * w * 4];
        for (uint32_t x = 0; x < w; x++, o += 4) {
            if (ctype == 3) {
                const int i = sample(row, x, 0);
                if (size_t(i) >= pal_len) return fail(err, "PNG: palette index out of range");
                std::memcpy(o, &pal[4 * i], 4);
            } else if (ctype == 0) {
                const int v = sample(row, x, 0);
                const int g = depth < 8 ? v * grey_scale[depth] : to8(v);
                o[0] = o[1] = o[2] = uint8_t(g);
                const bool clear = have_trans && (depth == 16 ? v == key[0] : uint8_t(g) == key8[0]);
                o[3] = clear ? 0 : 255;
            } else if (ctype == 2) {
                int v[3];
                for (int k = 0; k < 3; k++) v[k] = sample(row, x, k);
                for (int k = 0; k < 3; k++) o[k] = uint8_t(to8(v[k]));
                bool clear = have_trans;
                for (int k = 0; k < 3 && clear; k++) clear = depth == 16 ? v[k] == key[k] : o[k] == key8[k];
                o[3] = clear ? 0 : 255;
            } else if (ctype == 4) {
                o[0] = o[1] = o[2] = uint8_t(to8(sample(row, x, 0)));
                o[3] = uint8_t(to8(sample(row, x, 1)));
            } else {
                for (int k = 0; k < 4; k++) o[k] = uint8_t(to8(sample(row, x, k)));
            }
        }
    }
    out.channels = ctype == 3 ? (pal_alpha ? 4 : 3) : samples + (have_trans ? 1 : 0);
    return true;
}

// ------------------------------------------------------------------ JPEG

namespace {

// zigzag position -> natural (row-major) index; positions past 63, which a
// corrupt run length can reach, land on 63
const uint8_t kNatural[80] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33,
                              40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36,
                              29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54,
                              47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct huffman {
    bool defined = false;
    uint8_t vals[256] = {};
    int maxcode[18] = {};  // largest code of each length, -1 if none
    int offset[17] = {};   // vals index of the first code of each length, minus that code
    int16_t fast[512];     // 9-bit prefix -> vals index, -1 if the code is longer
    uint8_t fast_len[512];
    bool build(const uint8_t counts[16], const uint8_t* v, int n) {
        std::memcpy(vals, v, size_t(n));
        std::fill(fast, fast + 512, int16_t(-1));
        int code = 0, k = 0;
        for (int len = 1; len <= 16; len++) {
            offset[len] = k - code;
            for (int i = 0; i < counts[len - 1]; i++, k++, code++) {
                if (len <= 9) {
                    const int lo = code << (9 - len), hi = (code + 1) << (9 - len);
                    for (int j = lo; j < hi; j++) {
                        fast[j] = int16_t(k);
                        fast_len[j] = uint8_t(len);
                    }
                }
            }
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            if (code > (1 << len)) return false;  // more codes than the length allows
            code <<= 1;
        }
        defined = true;
        return true;
    }
};

// Entropy-coded segment reader: bytes 0xFF 0x00 are a data 0xFF; a marker ends
// the data, after which zero bits are fed.
struct bits {
    const uint8_t* p;
    const uint8_t* end;
    uint32_t acc = 0;
    int n = 0;
    bool at_marker = false;
    void fill() {
        while (n <= 24) {
            uint32_t b = 0;
            if (!at_marker && p < end) {
                b = *p++;
                if (b == 0xFF) {
                    if (p < end && *p == 0x00) {
                        ++p;
                    } else {
                        at_marker = true;
                        --p;
                        b = 0;
                    }
                }
            }
            acc |= b << (24 - n);
            n += 8;
        }
    }
    uint32_t get(int k) {
        if (k == 0) return 0;
        fill();
        const uint32_t v = acc >> (32 - k);
        acc <<= k;
        n -= k;
        return v;
    }
    int decode(const huffman& h) {
        fill();
        const uint32_t look = acc >> 23;
        if (h.fast[look] >= 0) {
            const int len = h.fast_len[look];
            acc <<= len;
            n -= len;
            return h.vals[h.fast[look]];
        }
        for (int len = 10; len <= 16; len++) {
            const int code = int(acc >> (32 - len));
            if (code <= h.maxcode[len]) {
                acc <<= len;
                n -= len;
                return h.vals[h.offset[len] + code];
            }
        }
        return -1;
    }
    void restart() {
        acc = 0;
        n = 0;
        at_marker = false;
    }
};

// a k-bit magnitude category value to its signed coefficient (JPEG F.2.2.1)
int extend(uint32_t v, int k) { return v < (1u << (k - 1)) ? int(v) - (1 << k) + 1 : int(v); }

// 12-bit fixed-point constant, rounded as (int)(x * 4096 + 0.5)
constexpr int fix12(float x) { return int(double(x * 4096.0f) + 0.5); }

// One 8-point inverse DCT, LLM factorisation. in[k * stride] are the inputs;
// out[0..7] come back scaled by 4096 * sqrt(8) (before the caller's rounding).
inline void idct8(const int* in, int stride, int out[8]) {
    // even part: inputs 0, 2, 4, 6
    const int r = (in[2 * stride] + in[6 * stride]) * fix12(0.5411961f);
    const int e2 = r + in[6 * stride] * fix12(-1.847759065f);
    const int e3 = r + in[2 * stride] * fix12(0.765366865f);
    const int e0 = (in[0] + in[4 * stride]) * 4096;
    const int e1 = (in[0] - in[4 * stride]) * 4096;
    const int a0 = e0 + e3, a3 = e0 - e3, a1 = e1 + e2, a2 = e1 - e2;
    // odd part: inputs 7, 5, 3, 1
    int o0 = in[7 * stride], o1 = in[5 * stride], o2 = in[3 * stride], o3 = in[1 * stride];
    int z3 = o0 + o2, z4 = o1 + o3;
    const int z1 = o0 + o3, z2 = o1 + o2;
    const int z5 = (z3 + z4) * fix12(1.175875602f);
    o0 *= fix12(0.298631336f);
    o1 *= fix12(2.053119869f);
    o2 *= fix12(3.072711026f);
    o3 *= fix12(1.501321110f);
    const int y1 = z5 + z1 * fix12(-0.899976223f);
    const int y2 = z5 + z2 * fix12(-2.562915447f);
    z3 *= fix12(-1.961570560f);
    z4 *= fix12(-0.390180644f);
    o3 += y1 + z4;
    o2 += y2 + z3;
    o1 += y2 + z4;
    o0 += y1 + z3;
    out[0] = a0 + o3;
    out[7] = a0 - o3;
    out[1] = a1 + o2;
    out[6] = a1 - o2;
    out[2] = a2 + o1;
    out[5] = a2 - o1;
    out[3] = a3 + o0;
    out[4] = a3 - o0;
}

inline uint8_t clamp_byte(int v) { return uint8_t(v < 0 ? 0 : (v > 255 ? 255 : v)); }

// Dequantised coefficients (natural order) to 8x8 samples. Columns first,
// rounded to 2 extra bits; then rows, rounded, level-shifted by 128, clamped.
void idct_block(const int16_t* c, uint8_t* dst, int stride) {
    int tmp[64];
    for (int x = 0; x < 8; x++) {
        int in[8], o[8];
        for (int k = 0; k < 8; k++) in[k] = c[8 * k + x];
        idct8(in, 1, o);
        for (int k = 0; k < 8; k++) tmp[8 * k + x] = (o[k] + 512) >> 10;
    }
    for (int y = 0; y < 8; y++) {
        int o[8];
        idct8(tmp + 8 * y, 1, o);
        for (int k = 0; k < 8; k++) dst[y * stride + k] = clamp_byte((o[k] + 65536 + (128 << 17)) >> 17);
    }
}

struct component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;
    int dc_pred = 0;
    int x = 0, y = 0;    // size in samples
    int w2 = 0, h2 = 0;  // plane size (whole MCUs)
    std::vector<uint8_t> plane;
};

// One output row of a component upsampled by (hs, vs) from the rows near and
// far (the nearer and the other neighbouring source row).
void upsample_row(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int hs, int vs) {
    if (hs == 1 && vs == 1) {
        std::memcpy(out, near, size_t(w));
    } else if (hs == 1 && vs == 2) {
        for (int i = 0; i < w; i++) out[i] = uint8_t((3 * near[i] + far[i] + 2) >> 2);
    } else if (hs == 2 && vs == 1) {
        if (w == 1) {
            out[0] = out[1] = near[0];
            return;
        }
        out[0] = near[0];
        out[1] = uint8_t((near[0] * 3 + near[1] + 2) >> 2);
        int i = 1;
        for (; i < w - 1; i++) {
            const int n = 3 * near[i] + 2;
            out[2 * i] = uint8_t((n + near[i - 1]) >> 2);
            out[2 * i + 1] = uint8_t((n + near[i + 1]) >> 2);
        }
        out[2 * i] = uint8_t((near[w - 2] * 3 + near[w - 1] + 2) >> 2);
        out[2 * i + 1] = near[w - 1];
    } else if (hs == 2 && vs == 2) {
        if (w == 1) {
            out[0] = out[1] = uint8_t((3 * near[0] + far[0] + 2) >> 2);
            return;
        }
        int t1 = 3 * near[0] + far[0];
        out[0] = uint8_t((t1 + 2) >> 2);
        for (int i = 1; i < w; i++) {
            const int t0 = t1;
            t1 = 3 * near[i] + far[i];
            out[2 * i - 1] = uint8_t((3 * t0 + t1 + 8) >> 4);
            out[2 * i] = uint8_t((3 * t1 + t0 + 8) >> 4);
        }
        out[2 * w - 1] = uint8_t((t1 + 2) >> 2);
    } else {
        for (int i = 0; i < w; i++)
            for (int j = 0; j < hs; j++) out[i * hs + j] = near[i];
    }
}

// YCbCr to RGB in 20-bit fixed point; green's Cb term keeps only its high 16 bits.
void ycc_to_rgb(uint8_t* o, int y, int cb, int cr) {
    auto fix = [](float x) { return int(x * 4096.0f + 0.5f) << 8; };
    const int yf = (y << 20) + (1 << 19);
    cr -= 128;
    cb -= 128;
    const int r = yf + cr * fix(1.40200f);
    const uint32_t gcb = uint32_t(cb * -fix(0.34414f)) & 0xffff0000u;
    const int g = int(uint32_t(yf + cr * -fix(0.71414f)) + gcb);
    const int b = yf + cb * fix(1.77200f);
    o[0] = clamp_byte(r >> 20);
    o[1] = clamp_byte(g >> 20);
    o[2] = clamp_byte(b >> 20);
    o[3] = 255;
}

}  // namespace

bool decode_jpeg(const std::vector<uint8_t>& f, ldr_image& out, std::string& err) {
    if (f.size() < 4 || f[0] != 0xFF || f[1] != 0xD8) return fail(err, "JPEG: bad signature");
    const uint8_t* p = f.data() + 2;
    const uint8_t* end = f.data() + f.size();
    uint16_t dq[4][64] = {};
    bool dq_defined[4] = {};
    huffman hdc[4], hac[4];
    std::vector<component> comp;
    int W = 0, H = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0, restart = 0;
    bool jfif = false, frame = false, any_scan = false;
    int adobe_transform = -1;
    auto u16 = [](const uint8_t* q) { return (q[0] << 8) | q[1]; };
    for (;;) {
        // next marker (fill bytes 0xFF allowed)
        while (p < end && *p != 0xFF) ++p;
        while (p < end && *p == 0xFF) ++p;
        if (p >= end) break;
        const int m = *p++;
        if (m == 0xD9) break;                       // EOI
        if (m == 0x00) continue;                    // a data 0xFF left after a scan
        if (m >= 0xD0 && m <= 0xD7) continue;       // stray RST
        if (m == 0x01) continue;                    // TEM
        if (end - p < 2) return fail(err, "JPEG: truncated marker");
        const int L = u16(p);
        if (L < 2 || L > end - p) return fail(err, "JPEG: bad segment length");
        const uint8_t* s = p + 2;
        const uint8_t* se = p + L;
        p += L;
        if (m == 0xC4) {  // DHT
            while (s < se) {
                if (se - s < 17) return fail(err, "JPEG: bad DHT");
                const int tc = *s >> 4, th = *s & 15;
                if (tc > 1 || th > 3) return fail(err, "JPEG: bad DHT");
                int n = 0;
                for (int i = 0; i < 16; i++) n += s[1 + i];
                if (n > 256 || se - s < 17 + n) return fail(err, "JPEG: bad DHT");
                if (!(tc ? hac[th] : hdc[th]).build(s + 1, s + 17, n)) return fail(err, "JPEG: bad Huffman table");
                s += 17 + n;
            }
        } else if (m == 0xDB) {  // DQT
            while (s < se) {
                const int pq = *s >> 4, tq = *s & 15;
                if (pq > 1 || tq > 3 || se - s < 1 + 64 * (pq + 1)) return fail(err, "JPEG: bad DQT");
                for (int k = 0; k < 64; k++)
                    dq[tq][kNatural[k]] = uint16_t(pq ? u16(s + 1 + 2 * k) : s[1 + k]);
                dq_defined[tq] = true;
                s += 1 + 64 * (pq + 1);
            }
        } else if (m == 0xDD) {  // DRI
            if (L != 4) return fail(err, "JPEG: bad DRI");
            restart = u16(s);
        } else if (m == 0xE0) {
            if (se - s >= 5 && std::memcmp(s, "JFIF\0", 5) == 0) jfif = true;
        } else if (m == 0xEE) {
            if (se - s >= 12 && std::memcmp(s, "Adobe", 5) == 0) adobe_transform = s[11];
        } else if (m == 0xC0 || m == 0xC1) {  // baseline / extended sequential, Huffman
            if (frame) return fail(err, "JPEG: more than one frame");
            if (se - s < 6 || s[0] != 8) return fail(err, "JPEG: only 8-bit samples are supported");
            H = u16(s + 1);
            W = u16(s + 3);
            const int nf = s[5];
            if (W == 0 || H == 0) return fail(err, "JPEG: bad image size");
            if (nf != 1 && nf != 3) return fail(err, "JPEG: only grey and 3-channel images are supported");
            if (se - s < 6 + 3 * nf) return fail(err, "JPEG: bad SOF");
            comp.resize(size_t(nf));
            for (int i = 0; i < nf; i++) {
                component& c = comp[size_t(i)];
                c.id = s[6 + 3 * i];
                c.h = s[7 + 3 * i] >> 4;
                c.v = s[7 + 3 * i] & 15;
                c.tq = s[8 + 3 * i];
                if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return fail(err, "JPEG: bad SOF");
                hmax = std::max(hmax, c.h);
                vmax = std::max(vmax, c.v);
            }
            mcux = (W + 8 * hmax - 1) / (8 * hmax);
            mcuy = (H + 8 * vmax - 1) / (8 * vmax);
            for (component& c : comp) {
                if (hmax % c.h || vmax % c.v) return fail(err, "JPEG: unsupported sampling factors");
                c.x = (W * c.h + hmax - 1) / hmax;
                c.y = (H * c.v + vmax - 1) / vmax;
                c.w2 = mcux * c.h * 8;
                c.h2 = mcuy * c.v * 8;
                c.plane.assign(size_t(c.w2) * c.h2, 0);
            }
            frame = true;
        } else if (m == 0xC2 || m == 0xC3 || (m >= 0xC5 && m <= 0xCF && m != 0xC8 && m != 0xCC)) {
            return fail(err, "JPEG: progressive, lossless and arithmetic-coded images are not supported");
        } else if (m == 0xDA) {  // SOS
            if (!frame) return fail(err, "JPEG: scan before frame");
            const int ns = s[0];
            if (ns < 1 || ns > int(comp.size()) || se - s < 4 + 2 * ns) return fail(err, "JPEG: bad SOS");
            std::vector<component*> sc;
            for (int i = 0; i < ns; i++) {
                component* c = nullptr;
                for (component& k : comp)
                    if (k.id == s[1 + 2 * i]) c = &k;
                if (!c) return fail(err, "JPEG: scan names an unknown component");
                c->td = s[2 + 2 * i] >> 4;
                c->ta = s[2 + 2 * i] & 15;
                if (c->td > 3 || c->ta > 3 || !hdc[c->td].defined || !hac[c->ta].defined || !dq_defined[c->tq])
                    return fail(err, "JPEG: scan uses an undefined table");
                c->dc_pred = 0;
                sc.push_back(c);
            }
            bits br{p, end};
            int16_t blk[64];
            auto block = [&](component& c, int bx, int by) -> bool {
                std::memset(blk, 0, sizeof(blk));
                const uint16_t* q = dq[c.tq];
                const int t = br.decode(hdc[c.td]);
                if (t < 0 || t > 15) return fail(err, "JPEG: bad Huffman code");
                const int diff = t ? extend(br.get(t), t) : 0;
                c.dc_pred += diff;
                blk[0] = int16_t(c.dc_pred * q[0]);
                for (int k = 1; k < 64;) {
                    const int rs = br.decode(hac[c.ta]);
                    if (rs < 0) return fail(err, "JPEG: bad Huffman code");
                    const int sz = rs & 15, run = rs >> 4;
                    if (sz == 0) {
                        if (rs != 0xF0) break;  // end of block
                        k += 16;
                    } else {
                        k += run;
                        const int z = kNatural[k++];
                        blk[z] = int16_t(extend(br.get(sz), sz) * q[z]);
                    }
                }
                idct_block(blk, &c.plane[size_t(by) * 8 * c.w2 + size_t(bx) * 8], c.w2);
                return true;
            };
            int todo = restart ? restart : -1;
            auto next_unit = [&]() -> bool {  // true: go on with the scan
                if (todo < 0 || --todo > 0) return true;
                br.fill();
                if (!br.at_marker || br.p + 1 >= end || br.p[1] < 0xD0 || br.p[1] > 0xD7) return false;
                br.p += 2;
                br.restart();
                for (component* c : sc) c->dc_pred = 0;
                todo = restart;
                return true;
            };
            bool go = true;
            if (ns == 1) {
                component& c = *sc[0];
                const int bw = (c.x + 7) / 8, bh = (c.y + 7) / 8;
                for (int by = 0; by < bh && go; by++)
                    for (int bx = 0; bx < bw && go; bx++) {
                        if (!block(c, bx, by)) return false;
                        go = next_unit();
                    }
            } else {
                for (int my = 0; my < mcuy && go; my++)
                    for (int mx = 0; mx < mcux && go; mx++) {
                        for (component* c : sc)
                            for (int v = 0; v < c->v; v++)
                                for (int h = 0; h < c->h; h++)
                                    if (!block(*c, mx * c->h + h, my * c->v + v)) return false;
                        go = next_unit();
                    }
            }
            p = br.p;  // the markers scan on from here
            any_scan = true;
        }
    }
    if (!frame || !any_scan) return fail(err, "JPEG: no image data");

    // upsample each component to full size and convert to RGBA
    out.width = W;
    out.height = H;
    out.channels = int(comp.size());
    out.rgba.assign(size_t(W) * H * 4, 0);
    const int nc = int(comp.size());
    bool is_rgb = false;
    if (nc == 3)
        is_rgb = (comp[0].id == 'R' && comp[1].id == 'G' && comp[2].id == 'B') || (adobe_transform == 0 && !jfif);
    struct resampler {
        int hs, vs, ystep, ypos, wlo;
        const uint8_t *line0, *line1;
    };
    std::vector<resampler> rs(static_cast<size_t>(nc));
    std::vector<std::vector<uint8_t>> lines(static_cast<size_t>(nc), std::vector<uint8_t>(size_t(W) + 3 + 4 * 8));
    for (int k = 0; k < nc; k++) {
        resampler& r = rs[size_t(k)];
        r.hs = hmax / comp[size_t(k)].h;
        r.vs = vmax / comp[size_t(k)].v;
        r.ystep = r.vs >> 1;
        r.ypos = 0;
        r.wlo = (W + r.hs - 1) / r.hs;
        r.line0 = r.line1 = comp[size_t(k)].plane.data();
    }
    const uint8_t* row[3] = {nullptr, nullptr, nullptr};
    for (int y = 0; y < H; y++) {
        for (int k = 0; k < nc; k++) {
            resampler& r = rs[size_t(k)];
            const bool bottom = r.ystep >= (r.vs >> 1);
            uint8_t* lb = lines[size_t(k)].data();
            upsample_row(lb, bottom ? r.line1 : r.line0, bottom ? r.line0 : r.line1, r.wlo, r.hs, r.vs);
            row[k] = lb;
            if (++r.ystep >= r.vs) {
                r.ystep = 0;
                r.line0 = r.line1;
                if (++r.ypos < comp[size_t(k)].y) r.line1 += comp[size_t(k)].w2;
            }
        }
        uint8_t* o = &out.rgba[size_t(y) * W * 4];
        for (int x = 0; x < W; x++, o += 4) {
            if (nc == 1) {
                o[0] = o[1] = o[2] = row[0][x];
                o[3] = 255;
            } else if (is_rgb) {
                o[0] = row[0][x];
                o[1] = row[1][x];
                o[2] = row[2][x];
                o[3] = 255;
            } else {
                ycc_to_rgb(o, row[0][x], row[1][x], row[2][x]);
            }
        }
    }
    return true;
}

bool decode_ldr_image(const std::vector<uint8_t>& bytes, ldr_image& out, std::string& err) {
    if (bytes.size() >= 8 && bytes[0] == 0x89 && bytes[1] == 'P') return decode_png(bytes, out, err);
    if (bytes.size() >= 2 && bytes[0] == 0xFF && bytes[1] == 0xD8) return decode_jpeg(bytes, out, err);
    return fail(err, "unsupported image format (PNG and JPEG are read)");
}

}  // namespace mpt
