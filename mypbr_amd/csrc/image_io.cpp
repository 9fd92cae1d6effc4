// image_io.cpp - write_image (image_io.h:81-106): display images from the
// GPU's bytes, raw radiance as Radiance RGBE or PFM.
//
// .png and .ppm: bloom/vignette (when on), exposure, tonemap and the sRGB
// bytes all come from mpt_display_bytes, i.e. the GPU. The PPM text is the
// reference's byte for byte (P3, one pixel per line). The PNG holds the same
// pixel bytes; its zlib stream is zlib's, not stb_image_write's, so the file
// bytes differ while every decoder reads the same image.
// .hdr: the floats of the radiance (image_io.h:57-63) as RGBE pixels in
// Greg Ward's Radiance format, which stb_image_write also writes; mantissas
// by frexp as in Ward's float2rgbe. Scanlines are stored flat (uncompressed),
// which the format allows and every reader accepts.
// .pfm: the reference's bytes exactly (image_io.h:66-79).
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpt.h"

namespace mpt {
int set_error(int code, const std::string& m);
}

namespace {

bool ends_with(const std::string& s, const std::string& suffix) {  // image_io.h:81-83
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

struct file_closer {
    FILE* f;
    ~file_closer() {
        if (f) std::fclose(f);
    }
};

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back(uint8_t(x >> 24));
    v.push_back(uint8_t(x >> 16));
    v.push_back(uint8_t(x >> 8));
    v.push_back(uint8_t(x));
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* data, size_t n) {
    put_be32(out, uint32_t(n));
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    if (n) out.insert(out.end(), data, data + n);
    put_be32(out, uint32_t(crc32(0L, out.data() + at, uInt(4 + n))));
}

int write_png(const std::string& path, int w, int h, const std::vector<uint8_t>& rgb) {
    // every scanline with filter type 0, then one zlib stream
    std::vector<uint8_t> raw;
    raw.reserve(size_t(h) * (1 + 3 * size_t(w)));
    for (int j = 0; j < h; j++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb.begin() + size_t(j) * 3 * w, rgb.begin() + size_t(j + 1) * 3 * w);
    }
    uLongf zn = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), uLong(raw.size()), 6) != Z_OK)
        return mpt::set_error(MPT_E_IO, "write_image: zlib failed for '" + path + "'");
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, uint32_t(w));
    put_be32(ihdr, uint32_t(h));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, no interlace
    png_chunk(out, "IHDR", ihdr.data(), ihdr.size());
    png_chunk(out, "IDAT", z.data(), zn);
    png_chunk(out, "IEND", nullptr, 0);
    file_closer f{std::fopen(path.c_str(), "wb")};
    if (!f.f || std::fwrite(out.data(), 1, out.size(), f.f) != out.size())
        return mpt::set_error(MPT_E_IO, "write_image: cannot write '" + path + "'");
    return MPT_OK;
}

// write_ppm (image_io.h:16-22) with write_color's format (color3.h:35-40).
int write_ppm(FILE* f, int w, int h, const std::vector<uint8_t>& rgb) {
    if (std::fprintf(f, "P3\n%d %d\n255\n", w, h) < 0) return -1;
    for (size_t k = 0; k < size_t(w) * h; k++)
        if (std::fprintf(f, "%d %d %d\n", int(rgb[3 * k]), int(rgb[3 * k + 1]), int(rgb[3 * k + 2])) < 0) return -1;
    return 0;
}

// Ward's float2rgbe on the radiance's floats, in float arithmetic as
// stb_image_write 1.16 computes it: e = frexp's exponent of the largest
// channel, each channel times frexp's mantissa * 256 / largest. Negative
// channels, which the tracer does not produce, are written as 0.
void to_rgbe(const float* c, uint8_t* e) {
    const float v = c[0] > (c[1] > c[2] ? c[1] : c[2]) ? c[0] : (c[1] > c[2] ? c[1] : c[2]);
    if (!(v >= 1e-32f)) {
        e[0] = e[1] = e[2] = e[3] = 0;
        return;
    }
    int ex;
    const float m = float(std::frexp(v, &ex)) * 256.0f / v;
    for (int k = 0; k < 3; k++) e[k] = uint8_t(c[k] > 0 ? c[k] * m : 0.0f);
    e[3] = uint8_t(ex + 128);
}

int write_hdr(const std::string& path, int w, int h, const double* rad) {
    file_closer f{std::fopen(path.c_str(), "wb")};
    if (!f.f) return mpt::set_error(MPT_E_IO, "write_image: cannot write '" + path + "'");
    std::fprintf(f.f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
    std::vector<uint8_t> line(size_t(w) * 4);
    for (int j = 0; j < h; j++) {
        for (int i = 0; i < w; i++) {
            const double* p = rad + 3 * (size_t(j) * w + i);
            const float c[3] = {float(p[0]), float(p[1]), float(p[2])};
            to_rgbe(c, &line[4 * size_t(i)]);
        }
        if (std::fwrite(line.data(), 1, line.size(), f.f) != line.size())
            return mpt::set_error(MPT_E_IO, "write_image: cannot write '" + path + "'");
    }
    return MPT_OK;
}

int write_pfm(const std::string& path, int w, int h, const double* rad) {  // image_io.h:66-79
    file_closer f{std::fopen(path.c_str(), "wb")};
    if (!f.f) return mpt::set_error(MPT_E_IO, "write_image: cannot write '" + path + "'");
    std::fprintf(f.f, "PF\n%d %d\n-1.0\n", w, h);
    std::vector<float> row(size_t(w) * 3);
    for (int j = h - 1; j >= 0; j--) {
        for (size_t k = 0; k < row.size(); k++) row[k] = float(rad[size_t(j) * 3 * w + k]);
        if (std::fwrite(row.data(), sizeof(float), row.size(), f.f) != row.size())
            return mpt::set_error(MPT_E_IO, "write_image: cannot write '" + path + "'");
    }
    return MPT_OK;
}

}  // namespace

extern "C" int32_t mpt_write_image(const char* path_c, int32_t width, int32_t height, const double* radiance,
                                   const mpt_display* display) {
    if (!radiance || width <= 0 || height <= 0) return mpt::set_error(MPT_E_INVALID, "write_image: empty image");
    const std::string path = path_c ? path_c : "";
    const bool shown = path.empty() || ends_with(path, ".png") || ends_with(path, ".ppm");
    if (shown) {
        std::vector<uint8_t> rgb(size_t(width) * height * 3);
        int rc = mpt_display_bytes(radiance, width, height, display, 1, rgb.data());
        if (rc) return rc;
        if (ends_with(path, ".png")) return write_png(path, width, height, rgb);
        if (path.empty()) return write_ppm(stdout, width, height, rgb) == 0 ? MPT_OK
                                                                           : mpt::set_error(MPT_E_IO, "write_image: stdout");
        file_closer f{std::fopen(path.c_str(), "wb")};
        if (!f.f || write_ppm(f.f, width, height, rgb) != 0)
            return mpt::set_error(MPT_E_IO, "write_image: cannot write '" + path + "'");
        return MPT_OK;
    }
    if (ends_with(path, ".hdr")) return write_hdr(path, width, height, radiance);
    if (ends_with(path, ".pfm")) return write_pfm(path, width, height, radiance);
    return mpt::set_error(MPT_E_INVALID,
                          "write_image: unsupported output format '" + path + "' (use .png, .ppm, .hdr or .pfm)");
}
