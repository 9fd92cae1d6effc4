// display.hip - the display pipeline on the GPU: bloom, vignetting,
// exposure, tonemapping and sRGB bytes.
//
// postprocess() (postprocess.h:108-153) keeps the part of each pixel above
// the luminance threshold, builds a pyramid of half-size images of it
// (downsample, postprocess.h:47-59), blurs every level with a separable
// Gaussian of sigma 1.5 (postprocess.h:62-89) and adds back the average of
// the blurred levels, sampled bilinearly at each full-size pixel, times the
// bloom strength; vignetting divides by (1 + k r^2)^2. display_settings::
// apply (tonemap.h:28-37) then scales by 2^exposure and tonemaps, and
// to_byte (color3.h:30-33) encodes with the sRGB curve and quantises.
//
// Kernels, all in double precision in the reference's operation order:
//   k_down      the first level, a thread per output pixel: the radiance's
//               2 x 2 blocks with the threshold applied on the fly, averaged
//               (a pure stream over the full-size image).
//   k_level     one launch per level: a block makes a 32 x 16 tile of the
//               blurred level in LDS: the downsampled window with a 5-pixel
//               halo (for the first level, the window read from k_down's
//               output), the horizontal pass, the vertical pass. It also
//               stores the unblurred tile, the source of the next level.
//   k_compose   a block per 32 x 8 tile of full-size pixels: first the texels
//               of every level that the tile's bilinear samples touch are
//               staged in LDS (at most 18 x 6 per level, since a level is at
//               most half the width and height), then a thread per pixel
//               adds the glow, vignettes, and writes either the
//               postprocessed radiance or, fused, the exposed, tonemapped
//               sRGB byte.
// The byte is found by comparing with the 255 radiance thresholds at which
// to_byte steps up, computed once on the host from the same formula and the
// host's libm (the one the reference's own writer calls), so the bytes are
// the reference's even where a GPU pow would round differently.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpt.h"
#include "geom.hpp"

namespace mpt {

int set_error(int code, const std::string& m);

namespace {

constexpr int kMaxLevels = 10;  // postprocess.h:124
constexpr int kTaps = 11;       // blur radius ceil(3 * 1.5) = 5 (postprocess.h:63)
constexpr int kR = 5;
constexpr int kTX = 32, kTY = 16;  // k_level tile
constexpr int kWX = kTX + 2 * kR, kWY = kTY + 2 * kR;
constexpr int kCX = 32, kCY = 8;      // k_compose tile
constexpr int kCWX = 20, kCWY = 8;    // its per-level texel window (18 x 6 needed)

struct level_desc {
    double* level;    // unblurred (the next level's source)
    double* blurred;
    int32_t w, h;
};

struct pyramid {
    level_desc lv[kMaxLevels];
    double sx[kMaxLevels], sy[kMaxLevels];  // double(b.w) / width, double(b.h) / height
    int32_t count;
};

struct blur_weights {
    double w[kTaps];
};

// 0 < T[1] <= ... <= T[255]: the smallest radiance to_byte maps to k or above.
struct byte_thresholds {
    double t[256];
};

struct display_args {
    double scale;       // 2^exposure (std::pow on the host, as tonemap.h:29)
    int32_t tonemap;
    int32_t vignette_on;
    double vignette;
    double cx, cy, corner2;
    double bloom_over_levels;  // post.bloom / blurred.size()
};

__device__ __forceinline__ v3 ld(const double* b, int64_t k) { return v3{b[3 * k], b[3 * k + 1], b[3 * k + 2]}; }
__device__ __forceinline__ void st(double* b, int64_t k, const v3& v) {
    b[3 * k] = v.x;
    b[3 * k + 1] = v.y;
    b[3 * k + 2] = v.z;
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The part of a pixel above the threshold (postprocess.h:113-119).
__device__ __forceinline__ v3 bright(const v3& p, double threshold) {
    const double l = luminance(p);
    return (__builtin_isfinite(l) && l > threshold) ? p * ((l - threshold) / l) : v3{0, 0, 0};
}

// downsample (postprocess.h:47-59) of src (sw x sh) at output pixel (i, j);
// bright_src: src is the radiance and the threshold applies to each pixel read.
__device__ __forceinline__ v3 down_at(const double* src, int sw, int sh, bool bright_src, double threshold, int i,
                                      int j) {
    auto at = [&](int a, int b) {  // image::at (postprocess.h:28-32)
        const v3 p = ld(src, int64_t(clampi(b, 0, sh - 1)) * sw + clampi(a, 0, sw - 1));
        return bright_src ? bright(p, threshold) : p;
    };
    return 0.25 * (at(2 * i, 2 * j) + at(2 * i + 1, 2 * j) + at(2 * i, 2 * j + 1) + at(2 * i + 1, 2 * j + 1));
}

// The first level, unblurred: a thread per pixel.
__global__ __launch_bounds__(256) void k_down(const double* rad, int sw, int sh, double threshold, double* dst, int w,
                                              int h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= w) return;
    st(dst, int64_t(j) * w + i, down_at(rad, sw, sh, true, threshold, i, j));
}

// One pyramid level. With src set, the level is downsampled from src (the
// level above, sw x sh) on the fly and its unblurred tile stored; with src
// null the unblurred level is out.level already.
__global__ __launch_bounds__(256) void k_level(const double* src, int sw, int sh, level_desc out, blur_weights bw) {
    __shared__ double dwin[kWY][kWX][3];   // downsampled window, clamped coordinates
    __shared__ double hwin[kWY][kTX][3];   // after the horizontal pass
    const int x0 = blockIdx.x * kTX, y0 = blockIdx.y * kTY;
    const int w = out.w, h = out.h;
    for (int k = threadIdx.x; k < kWX * kWY; k += blockDim.x) {
        const int c = k % kWX, r = k / kWX;
        const int i = clampi(x0 - kR + c, 0, w - 1), j = clampi(y0 - kR + r, 0, h - 1);
        const v3 d = src ? down_at(src, sw, sh, false, 0.0, i, j) : ld(out.level, int64_t(j) * w + i);
        dwin[r][c][0] = d.x, dwin[r][c][1] = d.y, dwin[r][c][2] = d.z;
        // the unblurred level, for the next one
        if (src && c >= kR && c < kR + kTX && r >= kR && r < kR + kTY) {
            const int gi = x0 - kR + c, gj = y0 - kR + r;
            if (gi < w && gj < h) st(out.level, int64_t(gj) * w + gi, d);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kTX * kWY; k += blockDim.x) {  // postprocess.h:72-79
        const int c = k % kTX, r = k / kTX;
        v3 s{0, 0, 0};
#pragma unroll
        for (int t = 0; t < kTaps; t++) {
            const v3 v{dwin[r][c + t][0], dwin[r][c + t][1], dwin[r][c + t][2]};
            s = s + bw.w[t] * v;
        }
        hwin[r][c][0] = s.x, hwin[r][c][1] = s.y, hwin[r][c][2] = s.z;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kTX * kTY; k += blockDim.x) {  // postprocess.h:80-87
        const int c = k % kTX, r = k / kTX;
        const int gi = x0 + c, gj = y0 + r;
        if (gi >= w || gj >= h) continue;
        v3 s{0, 0, 0};
#pragma unroll
        for (int t = 0; t < kTaps; t++) {
            // tmp.at(i, j + t): rows clamp to the level, as the window's rows do
            const int rr = clampi(gj - kR + t, 0, h - 1) - (y0 - kR);
            const v3 v{hwin[rr][c][0], hwin[rr][c][1], hwin[rr][c][2]};
            s = s + bw.w[t] * v;
        }
        st(out.blurred, int64_t(gj) * w + gi, s);
    }
}

__device__ __forceinline__ double rrt_odt_fit(double x) {  // tonemap.h:63-67
    const double a = x * (x + 0.0245786) - 0.000090537;
    const double b = x * (0.983729 * x + 0.4329510) + 0.238081;
    return a / b;
}
__device__ __forceinline__ double clamp01(double x) { return x < 0 ? 0 : (x > 1 ? 1 : x); }

// display_settings::apply (tonemap.h:28-37) after the exposure scale.
__device__ __forceinline__ v3 tonemap(v3 c, int op) {
    if (op == MPT_TONEMAP_REINHARD) return v3{c.x / (1 + c.x), c.y / (1 + c.y), c.z / (1 + c.z)};
    if (op == MPT_TONEMAP_ACES) {  // tonemap.h:46-57
        v3 v{0.59719 * c.x + 0.35458 * c.y + 0.04823 * c.z, 0.07600 * c.x + 0.90834 * c.y + 0.01566 * c.z,
             0.02840 * c.x + 0.13383 * c.y + 0.83777 * c.z};
        v = v3{rrt_odt_fit(v.x), rrt_odt_fit(v.y), rrt_odt_fit(v.z)};
        v = v3{1.60475 * v.x + -0.53108 * v.y + -0.07367 * v.z, -0.10208 * v.x + 1.10813 * v.y + -0.00605 * v.z,
               -0.00327 * v.x + -0.07276 * v.y + 1.07602 * v.z};
        return v3{clamp01(v.x), clamp01(v.y), clamp01(v.z)};
    }
    return c;
}

// to_byte by the thresholds: the number of T[k] (k = 1..255) at or below x.
// A NaN compares false everywhere and gives 0, as the reference's cast does
// on x86-64.
__device__ __forceinline__ uint32_t to_byte(const double* t, double x) {
    uint32_t b = 0;
#pragma unroll
    for (uint32_t step = 128; step >= 1; step >>= 1)
        if (t[b + step] <= x) b += step;
    return b;
}

template <bool BYTES>
__global__ __launch_bounds__(256) void k_compose(const double* rad, int width, int height, pyramid py,
                                                 display_args a, byte_thresholds bt, double* out, uint8_t* rgb) {
    __shared__ double thr[256];
    __shared__ double win[kMaxLevels][kCWY][kCWX][3];
    __shared__ int org[kMaxLevels][2];
    if constexpr (BYTES) thr[threadIdx.x] = bt.t[threadIdx.x];
    const int x0 = blockIdx.x * kCX, y0 = blockIdx.y * kCY;
    // the sample coordinates of sample() (postprocess.h:35-42), as the pixels compute them
    auto sx_of = [&](int k, int i) { return (i + 0.5) * py.sx[k] - 0.5; };
    auto sy_of = [&](int k, int j) { return (j + 0.5) * py.sy[k] - 0.5; };
    for (int k = 0; k < py.count; k++) {  // stage each level's window of texels (clamped as image::at)
        const level_desc L = py.lv[k];
        const int wx0 = int(std::floor(sx_of(k, x0))), wy0 = int(std::floor(sy_of(k, y0)));
        const int nx = int(std::floor(sx_of(k, x0 + kCX - 1))) - wx0 + 2;
        const int ny = int(std::floor(sy_of(k, y0 + kCY - 1))) - wy0 + 2;
        if (threadIdx.x == 0) org[k][0] = wx0, org[k][1] = wy0;
        for (int t = threadIdx.x; t < nx * ny; t += blockDim.x) {
            const int c = t % nx, r = t / nx;
            const v3 v = ld(L.blurred, int64_t(clampi(wy0 + r, 0, L.h - 1)) * L.w + clampi(wx0 + c, 0, L.w - 1));
            win[k][r][c][0] = v.x, win[k][r][c][1] = v.y, win[k][r][c][2] = v.z;
        }
    }
    __syncthreads();
    const int i = x0 + int(threadIdx.x) % kCX, j = y0 + int(threadIdx.x) / kCX;
    if (i >= width || j >= height) return;
    const int64_t p = int64_t(j) * width + i;
    v3 c = ld(rad, p);
    if (py.count > 0) {  // postprocess.h:128-136
        v3 glow{0, 0, 0};
        for (int k = 0; k < py.count; k++) {
            const double x = sx_of(k, i), y = sy_of(k, j);
            const int ia = int(std::floor(x)), ja = int(std::floor(y));
            const double fx = x - ia, fy = y - ja;
            const int cc = ia - org[k][0], rr = ja - org[k][1];
            auto at = [&](int r, int q) { return v3{win[k][r][q][0], win[k][r][q][1], win[k][r][q][2]}; };
            glow = glow + ((1 - fy) * ((1 - fx) * at(rr, cc) + fx * at(rr, cc + 1)) +
                           fy * ((1 - fx) * at(rr + 1, cc) + fx * at(rr + 1, cc + 1)));
        }
        c = c + a.bloom_over_levels * glow;
    }
    if (a.vignette_on) {  // postprocess.h:139-150
        const double dx = i + 0.5 - a.cx, dy = j + 0.5 - a.cy;
        const double f = 1 + a.vignette * (dx * dx + dy * dy) / a.corner2;
        c = c / (f * f);
    }
    if constexpr (BYTES) {
        c = tonemap(a.scale * c, a.tonemap);
        rgb[3 * p] = uint8_t(to_byte(thr, c.x));
        rgb[3 * p + 1] = uint8_t(to_byte(thr, c.y));
        rgb[3 * p + 2] = uint8_t(to_byte(thr, c.z));
    } else {
        st(out, p, c);
    }
}

// ---------------------------------------------------------------- host side

// to_byte (color3.h:17-33) on the host, for the thresholds.
uint8_t host_to_byte(double x) {
    double s;
    if (x <= 0) s = 0;
    else if (x <= 0.0031308) s = 12.92 * x;
    else s = 1.055 * std::pow(x, 1.0 / 2.4) - 0.055;
    if (s < 0.000) s = 0.000;  // interval(0.000, 0.999).clamp (interval.h:33-37)
    else if (s > 0.999) s = 0.999;
    return uint8_t(256 * s);
}

double from_bits(uint64_t b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

const byte_thresholds& thresholds() {
    static byte_thresholds t;
    static std::once_flag once;
    std::call_once(once, [] {
        t.t[0] = -kInf;
        const uint64_t inf_bits = 0x7FF0000000000000ull;  // positive doubles order as their bits
        for (int k = 1; k < 256; k++) {
            uint64_t lo = 0, hi = inf_bits;  // host_to_byte(lo) < k <= host_to_byte(hi)
            while (hi - lo > 1) {
                const uint64_t mid = lo + (hi - lo) / 2;
                if (host_to_byte(from_bits(mid)) >= k) hi = mid;
                else lo = mid;
            }
            t.t[k] = from_bits(hi);
        }
    });
    return t;
}

blur_weights gaussian_weights() {  // blur(), postprocess.h:62-69, sigma 1.5
    const double sigma = 1.5;
    blur_weights bw;
    double sum = 0;
    for (int k = -kR; k <= kR; k++) sum += bw.w[k + kR] = std::exp(-0.5 * k * k / (sigma * sigma));
    for (double& k : bw.w) k /= sum;
    return bw;
}

// The pyramid's level sizes (postprocess.h:122-130).
std::vector<std::pair<int, int>> level_sizes(int width, int height, const mpt_display& d) {
    std::vector<std::pair<int, int>> out;
    if (!(d.bloom > 0)) return out;
    const double reach = std::fmax(d.bloom_radius * width, 1.0);
    const int levels = std::clamp(int(std::ceil(std::log2(reach))), 1, kMaxLevels);
    int w = width, h = height;
    for (int k = 0; k < levels && w > 1 && h > 1; k++) {
        w = std::max(1, w / 2);
        h = std::max(1, h / 2);
        out.push_back({w, h});
    }
    return out;
}

int check(int width, int height, const mpt_display* d) {
    if (width <= 0 || height <= 0) return set_error(MPT_E_INVALID, "display: the image is empty");
    if (!d) return MPT_OK;
    if (d->tonemap < MPT_TONEMAP_NONE || d->tonemap > MPT_TONEMAP_ACES)
        return set_error(MPT_E_INVALID, "display: unknown tonemapper");
    if (!(d->bloom >= 0) || !(d->bloom_threshold >= 0) || !(d->bloom_radius > 0) || !(d->vignette >= 0))
        return set_error(MPT_E_INVALID, "display: bloom and vignette need bloom >= 0, threshold >= 0, radius > 0, "
                                        "vignette >= 0 (scene.h:667-681)");
    return MPT_OK;
}

#define DP_TRY(x)                                                                                        \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return set_error(MPT_E_HIP, std::string("display: ") + hipGetErrorString(e_)); \
    } while (0)

// The whole pipeline on device buffers. post: apply bloom/vignette.
int run(const double* rad, int width, int height, const mpt_display* display, bool post, double* out, uint8_t* rgb,
        void* workspace, hipStream_t s) {
    mpt_display d;
    if (display) d = *display;
    else mpt_display_defaults(&d);
    pyramid py{};
    display_args a{};
    a.scale = std::pow(2.0, d.exposure);
    a.tonemap = d.tonemap;
    if (post) {
        const auto sizes = level_sizes(width, height, d);
        py.count = int32_t(sizes.size());
        double* ws = static_cast<double*>(workspace);
        if (py.count && !ws) return set_error(MPT_E_INVALID, "display: bloom needs a workspace");
        for (int k = 0; k < py.count; k++) {
            const int64_t n = int64_t(sizes[k].first) * sizes[k].second;
            py.lv[k].w = sizes[k].first;
            py.lv[k].h = sizes[k].second;
            py.lv[k].blurred = ws;
            py.lv[k].level = ws + 3 * n;
            ws += 6 * n;
            py.sx[k] = double(py.lv[k].w) / width;
            py.sy[k] = double(py.lv[k].h) / height;
        }
        for (int k = 0; k < py.count; k++)
            if (!(py.sx[k] <= 0.5 && py.sy[k] <= 0.5))  // a level is at most half the image: the windows fit
                return set_error(MPT_E_INVALID, "display: pyramid level larger than half the image");
        if (py.count) a.bloom_over_levels = d.bloom / double(py.count);
        a.vignette_on = d.vignette > 0;
        a.vignette = d.vignette;
        a.cx = 0.5 * width;
        a.cy = 0.5 * height;
        a.corner2 = a.cx * a.cx + a.cy * a.cy;
        const blur_weights bw = gaussian_weights();
        for (int k = 0; k < py.count; k++) {
            level_desc L = py.lv[k];
            if (k == 0) {  // the first level streams from the radiance, then blurs from its own buffer
                hipLaunchKernelGGL(k_down, dim3(unsigned((L.w + 255) / 256), unsigned(L.h)), dim3(256), 0, s, rad, width,
                                   height, d.bloom_threshold, L.level, L.w, L.h);
            }
            const dim3 g(unsigned((L.w + kTX - 1) / kTX), unsigned((L.h + kTY - 1) / kTY));
            hipLaunchKernelGGL(k_level, g, dim3(256), 0, s, k ? py.lv[k - 1].level : nullptr, k ? py.lv[k - 1].w : 0,
                               k ? py.lv[k - 1].h : 0, L, bw);
        }
    }
    const dim3 g(unsigned((width + kCX - 1) / kCX), unsigned((height + kCY - 1) / kCY));
    if (rgb) hipLaunchKernelGGL(k_compose<true>, g, dim3(256), 0, s, rad, width, height, py, a, thresholds(), out, rgb);
    else hipLaunchKernelGGL(k_compose<false>, g, dim3(256), 0, s, rad, width, height, py, a, byte_thresholds{}, out, rgb);
    DP_TRY(hipGetLastError());
    return MPT_OK;
}

// Host buffers in and out around run().
int run_host(const double* rad, int width, int height, const mpt_display* display, bool post, double* out,
             uint8_t* rgb) {
    const size_t n = size_t(width) * height;
    const size_t img = n * 3 * sizeof(double);
    const size_t ws = post ? mpt_display_workspace_bytes(width, height, display) : 0;
    const size_t out_bytes = rgb ? n * 3 : img;
    char* buf = nullptr;
    DP_TRY(hipMalloc(&buf, img + ws + out_bytes));
    auto done = [&](int code) {
        (void)hipFree(buf);
        return code;
    };
    double* d_rad = reinterpret_cast<double*>(buf);
    void* d_ws = ws ? buf + img : nullptr;
    char* d_out = buf + img + ws;
    if (hipMemcpy(d_rad, rad, img, hipMemcpyHostToDevice) != hipSuccess)
        return done(set_error(MPT_E_HIP, "display: copy to the device failed"));
    int rc = run(d_rad, width, height, display, post, rgb ? nullptr : reinterpret_cast<double*>(d_out),
                 rgb ? reinterpret_cast<uint8_t*>(d_out) : nullptr, d_ws, nullptr);
    if (rc) return done(rc);
    if (hipMemcpy(rgb ? static_cast<void*>(rgb) : static_cast<void*>(out), d_out, out_bytes, hipMemcpyDeviceToHost) !=
        hipSuccess)
        return done(set_error(MPT_E_HIP, "display: copy from the device failed"));
    return done(MPT_OK);
}

// The allocation-free entry point's bloom pyramid: a fixed array in device
// memory, one copy per device, held until the process ends. 48 MiB covers
// the pyramid of a 1920 x 1080 image at the default bloom radius (33.2 MB,
// 7 levels) with room to spare. Host code never passes `scratch` itself to a
// kernel: in host code the name stands for a host-side shadow, so every use
// goes through the device address hipGetSymbolAddress returns for it.
constexpr size_t kScratchBytes = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's scratch array, looked up once
// per device (with that device current) and kept.
int scratch_for_current_device(void** out) {
    int dev = 0;
    DP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(MPT_E_INVALID, "display: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(MPT_E_HIP, std::string("display: hipGetSymbolAddress of the scratch array: ") +
                                            hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MPT_OK;
}

}  // namespace
}  // namespace mpt

using namespace mpt;

extern "C" {

void mpt_display_defaults(mpt_display* d) {
    if (!d) return;
    d->exposure = 0;  // tonemap.h:13-15, postprocess.h:13-16
    d->tonemap = MPT_TONEMAP_NONE;
    d->pad = 0;
    d->bloom = 0;
    d->bloom_threshold = 1;
    d->bloom_radius = 0.04;
    d->vignette = 0;
}

size_t mpt_display_workspace_bytes(int32_t width, int32_t height, const mpt_display* display) {
    if (width <= 0 || height <= 0) return 0;
    mpt_display d;
    if (display) d = *display;
    else mpt_display_defaults(&d);
    const auto sizes = level_sizes(width, height, d);
    size_t doubles = 0;
    for (size_t k = 0; k < sizes.size(); k++) doubles += 6 * size_t(sizes[k].first) * sizes[k].second;
    return doubles * sizeof(double);
}

int32_t mpt_postprocess_dev(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                            double* out, void* workspace, void* stream) {
    int rc = check(width, height, display);
    if (rc) return rc;
    if (!radiance || !out) return set_error(MPT_E_INVALID, "postprocess: null buffer");
    return run(radiance, width, height, display, true, out, nullptr, workspace, static_cast<hipStream_t>(stream));
}

int32_t mpt_display_bytes_dev(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                              int32_t post, uint8_t* rgb, void* workspace, void* stream) {
    int rc = check(width, height, display);
    if (rc) return rc;
    if (!radiance || !rgb) return set_error(MPT_E_INVALID, "display bytes: null buffer");
    return run(radiance, width, height, display, post != 0, nullptr, rgb, workspace, static_cast<hipStream_t>(stream));
}

size_t mpt_display_noalloc_max_workspace(void) { return kScratchBytes; }

int32_t mpt_display_bytes_dev_noalloc(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                                      int32_t post, uint8_t* rgb, void* stream) {
    int rc = check(width, height, display);
    if (rc) return rc;
    if (!radiance || !rgb) return set_error(MPT_E_INVALID, "display bytes: null buffer");
    const size_t need = post ? mpt_display_workspace_bytes(width, height, display) : 0;
    if (need > kScratchBytes)
        return set_error(MPT_E_CAPACITY, "display bytes (no allocation): the bloom pyramid of a " +
                                             std::to_string(width) + " x " + std::to_string(height) + " image needs " +
                                             std::to_string(need) + " bytes, more than the fixed scratch of " +
                                             std::to_string(kScratchBytes) + " bytes");
    void* ws = nullptr;
    if (need > 0 && (rc = scratch_for_current_device(&ws))) return rc;
    // the pyramid kernels write every texel before any kernel reads it, so the
    // array needs no clearing between calls
    return run(radiance, width, height, display, post != 0, nullptr, rgb, ws, static_cast<hipStream_t>(stream));
}

int32_t mpt_postprocess(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                        double* out) {
    int rc = check(width, height, display);
    if (rc) return rc;
    if (!radiance || !out) return set_error(MPT_E_INVALID, "postprocess: null buffer");
    return run_host(radiance, width, height, display, true, out, nullptr);
}

int32_t mpt_display_bytes(const double* radiance, int32_t width, int32_t height, const mpt_display* display,
                          int32_t post, uint8_t* rgb) {
    int rc = check(width, height, display);
    if (rc) return rc;
    if (!radiance || !rgb) return set_error(MPT_E_INVALID, "display bytes: null buffer");
    return run_host(radiance, width, height, display, post != 0, nullptr, rgb);
}

}  // extern "C"
