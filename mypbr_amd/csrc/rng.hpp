// rng.hpp - counter-based random numbers for the wavefront integrator.
//
// The reference draws from a thread_local, unseeded std::mt19937
// (rtweekend.h:21-25), so its renders cannot be reproduced. Here every path
// owns a 64-bit key derived from (seed, scene, sample, pixel), and its i-th
// draw is the SplitMix64 output mix(key + i * gamma). A path's numbers thus
// depend only on how many it has drawn, never on which thread or wave ran it,
// so a batched GPU launch and a sequential CPU loop see the same sequence.
//
// Draws that the reference makes inside hit() (the free-flight distance of
// constant_medium, medium.h:77) come from a second stream keyed by the ray's
// serial number within the path and the object tested, so that they do not
// depend on BVH traversal order.
#pragma once

#include <cstdint>

#include "geom.hpp"

namespace mpt {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

MPT_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform double in [0, 1) from the top 53 bits.
MPT_HD double u01(uint64_t x) { return double(x >> 11) * 0x1.0p-53; }

// Key of the path for (seed, scene, sample, pixel).
MPT_HD uint64_t path_key(uint64_t seed, uint32_t scene, uint32_t sample, uint32_t pixel) {
    uint64_t s = mix64(seed + uint64_t(scene) * kGamma);
    return mix64(s ^ ((uint64_t(sample) << 32) | uint64_t(pixel)));
}

// Key of the separate stream that fills the denoiser guides.
MPT_HD uint64_t feature_key(uint64_t key) { return mix64(key ^ 0x3C6EF372FE94F82Bull); }

// The counter-th draw (counter starts at 1) of the path with this key.
MPT_HD double draw(uint64_t key, uint32_t counter) { return u01(mix64(key + uint64_t(counter) * kGamma)); }

// Free-flight draw of object obj for the ray with this serial number.
MPT_HD double medium_draw(uint64_t key, uint32_t ray_serial, uint32_t obj) {
    uint64_t k = mix64(key ^ 0x6A09E667F3BCC909ull);
    return u01(mix64(k + ((uint64_t(ray_serial) << 32) | uint64_t(obj)) * kGamma));
}

}  // namespace mpt
