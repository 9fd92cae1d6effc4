// PNG and baseline JPEG decoding for image textures (texture.h:67-105 loads
// them through stb_image with four channels requested). The decoders give the
// bytes stb_image's stbi_load(path, &w, &h, &channels, 4) gives: RGBA, rows
// top first, and channels = the file's own channel count.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mpt {

struct ldr_image {
    int width = 0, height = 0;
    int channels = 0;           // 1 grey, 2 grey + alpha, 3 RGB, 4 RGBA (as stored in the file)
    std::vector<uint8_t> rgba;  // width * height * 4
};

// By the file's signature: PNG or JPEG. False with a message in err.
bool decode_ldr_image(const std::vector<uint8_t>& bytes, ldr_image& out, std::string& err);
bool decode_png(const std::vector<uint8_t>& bytes, ldr_image& out, std::string& err);
bool decode_jpeg(const std::vector<uint8_t>& bytes, ldr_image& out, std::string& err);

}  // namespace mpt
