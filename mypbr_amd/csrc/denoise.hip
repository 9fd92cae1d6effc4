// denoise.hip - the classical edge-avoiding a-trous denoiser on the GPU.
//
// The filter of denoiser::run (denoiser.h:45-193): radiance is divided by
// albedo (+1e-3) into lighting, its variance by the albedo's luminance
// squared; fireflies are clamped to their brightest neighbour
// (denoiser.h:98-124); then `iterations` passes of a 5x5 B3-spline kernel
// with taps 2^it pixels apart, each tap weighted by normal, albedo and
// luminance similarity, the luminance scale taken from a 3x3 blur of the
// variance (denoiser.h:126-193); the variance is carried as
// sum(w^2 var) / (sum w)^2; finally the albedo is multiplied back.
//
// Every stage is one kernel over the pixels, a thread per pixel, in double
// precision with the reference's operation order; the taps are read
// straight from HBM through the caches (rows of a block are contiguous, and
// at every step a tap row is a contiguous run of pixels). All buffers are
// interleaved RGB doubles, row-major from the top-left, as frame holds them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/mpt.h"
#include "geom.hpp"

namespace mpt {

int set_error(int code, const std::string& m);

namespace {

constexpr double kAlbedoEps = 1e-3;  // denoiser.h:95

__device__ __forceinline__ v3 ld(const double* b, int64_t k) { return v3{b[3 * k], b[3 * k + 1], b[3 * k + 2]}; }
__device__ __forceinline__ void st(double* b, int64_t k, const v3& v) {
    b[3 * k] = v.x;
    b[3 * k + 1] = v.y;
    b[3 * k + 2] = v.z;
}

// lighting = radiance / (albedo + eps), var = variance / luminance(albedo + eps)^2 (denoiser.h:48-55)
__global__ void k_dn_prepare(int64_t n, const double* rad, const double* alb, const double* variance, double* light,
                             double* var) {
    int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    v3 a = ld(alb, k) + v3{kAlbedoEps, kAlbedoEps, kAlbedoEps};
    v3 c = ld(rad, k);
    st(light, k, v3{c.x / a.x, c.y / a.y, c.z / a.z});
    double la = luminance(a);
    var[k] = variance[k] / (la * la);
}

// remove_fireflies (denoiser.h:98-124): reads src, writes light and var in place.
__global__ void k_dn_fireflies(int w, int h, const double* src, double* light, double* var) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= w) return;
    int64_t p = int64_t(j) * w + i;
    v3 s = ld(src, p);
    st(light, p, s);
    if (i < 1 || i >= w - 1 || j < 1 || j >= h - 1) return;
    double lum = luminance(s);
    if (lum <= 0 || std::sqrt(var[p]) < 0.5 * lum) return;
    double brightest = 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
            if (dx != 0 || dy != 0) brightest = std::fmax(brightest, luminance(ld(src, p + dy * w + dx)));
    if (lum > brightest) {
        double scale = brightest / lum;
        st(light, p, scale * s);
        var[p] *= scale * scale;
    }
}

// blur3x3 (denoiser.h:169-188)
__global__ void k_dn_blur(int w, int h, const double* v, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= w) return;
    const double kern[3] = {0.25, 0.5, 0.25};
    double sum = 0, weight = 0;
    for (int dy = -1; dy <= 1; dy++) {
        for (int dx = -1; dx <= 1; dx++) {
            int x = i + dx, y = j + dy;
            if (x < 0 || x >= w || y < 0 || y >= h) continue;
            double wt = kern[dx + 1] * kern[dy + 1];
            sum += wt * v[int64_t(y) * w + x];
            weight += wt;
        }
    }
    out[int64_t(j) * w + i] = sum / weight;
}

// filter_pixel (denoiser.h:126-167) at tap spacing `step`.
__global__ void k_dn_filter(int w, int h, int step, double sigma_l, double inv_normal, double inv_albedo,
                            const double* alb, const double* nrm, const double* light, const double* var,
                            const double* bvar, double* out, double* out_var) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= w) return;
    const double kern[5] = {1.0 / 16, 1.0 / 4, 3.0 / 8, 1.0 / 4, 1.0 / 16};
    int64_t p = int64_t(j) * w + i;
    const v3 np = ld(nrm, p), ap = ld(alb, p);
    double lum_p = luminance(ld(light, p));
    double lum_scale = sigma_l * std::sqrt(std::fmax(bvar[p], 0.0)) + 1e-10;
    v3 sum{0, 0, 0};
    double weight_sum = 0;
    double var_sum = 0;
    for (int dy = -2; dy <= 2; dy++) {
        int y = j + dy * step;
        if (y < 0 || y >= h) continue;
        for (int dx = -2; dx <= 2; dx++) {
            int x = i + dx * step;
            if (x < 0 || x >= w) continue;
            int64_t q = int64_t(y) * w + x;
            const v3 lq = ld(light, q);
            double w_normal = std::exp(-length_squared(np - ld(nrm, q)) * inv_normal);
            double w_albedo = std::exp(-length_squared(ap - ld(alb, q)) * inv_albedo);
            double w_lum = std::exp(-std::fabs(lum_p - luminance(lq)) / lum_scale);
            double wt = kern[dx + 2] * kern[dy + 2] * w_normal * w_albedo * w_lum;
            sum = sum + wt * lq;
            weight_sum += wt;
            var_sum += wt * wt * var[q];
        }
    }
    st(out, p, sum / weight_sum);
    out_var[p] = var_sum / (weight_sum * weight_sum);
}

// out = lighting * (albedo + eps) (denoiser.h:75-77)
__global__ void k_dn_compose(int64_t n, const double* light, const double* alb, double* out) {
    int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    st(out, k, ld(light, k) * (ld(alb, k) + v3{kAlbedoEps, kAlbedoEps, kAlbedoEps}));
}

int check_opts(const mpt_denoise_opts* o, int w, int h) {
    if (w <= 0 || h <= 0) return set_error(MPT_E_INVALID, "denoise: width and height must be positive");
    if (o && (o->iterations < 0 || o->iterations > 30))
        return set_error(MPT_E_INVALID, "denoise: iterations must be in [0, 30]");
    if (o && !(o->sigma_luminance > 0 && o->sigma_normal > 0 && o->sigma_albedo > 0))
        return set_error(MPT_E_INVALID, "denoise: the sigmas must be positive");
    return MPT_OK;
}

#define DN_TRY(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return set_error(MPT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace
}  // namespace mpt

using namespace mpt;

extern "C" {

void mpt_denoise_defaults(mpt_denoise_opts* o) {
    if (!o) return;
    o->iterations = 5;  // denoiser.h:40-44
    o->pad = 0;
    o->sigma_luminance = 4;
    o->sigma_normal = 0.1;
    o->sigma_albedo = 0.1;
}

size_t mpt_denoise_workspace_bytes(int32_t width, int32_t height) {
    const size_t n = size_t(width > 0 ? width : 0) * size_t(height > 0 ? height : 0);
    return n * (3 * 3 + 3) * sizeof(double);  // lighting x2 + source, variance x2 + blurred
}

int32_t mpt_denoise_dev(const double* radiance, const double* albedo, const double* normal, const double* variance,
                        int32_t width, int32_t height, const mpt_denoise_opts* opts, double* out, void* workspace,
                        void* stream) {
    int rc = check_opts(opts, width, height);
    if (rc) return rc;
    if (!radiance || !albedo || !normal || !variance || !out || !workspace)
        return set_error(MPT_E_INVALID, "denoise: null buffer");
    mpt_denoise_opts o;
    if (opts) o = *opts;
    else mpt_denoise_defaults(&o);
    const int64_t n = int64_t(width) * height;
    double* ws = static_cast<double*>(workspace);
    double* light = ws;
    double* light2 = light + 3 * n;  // also the firefly pass's source
    double* var = light2 + 3 * n;
    double* var2 = var + n;
    double* bvar = var2 + n;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned g1 = unsigned((n + 255) / 256);
    const dim3 g2(unsigned((width + 127) / 128), unsigned(height));
    hipLaunchKernelGGL(k_dn_prepare, dim3(g1), dim3(256), 0, s, n, radiance, albedo, variance, light2, var);
    hipLaunchKernelGGL(k_dn_fireflies, g2, dim3(128), 0, s, width, height, light2, light, var);
    const double inv_normal = 1 / (o.sigma_normal * o.sigma_normal);
    const double inv_albedo = 1 / (o.sigma_albedo * o.sigma_albedo);
    for (int it = 0; it < o.iterations; it++) {
        hipLaunchKernelGGL(k_dn_blur, g2, dim3(128), 0, s, width, height, var, bvar);
        hipLaunchKernelGGL(k_dn_filter, g2, dim3(128), 0, s, width, height, 1 << it, o.sigma_luminance, inv_normal,
                           inv_albedo, albedo, normal, light, var, bvar, light2, var2);
        std::swap(light, light2);
        std::swap(var, var2);
    }
    hipLaunchKernelGGL(k_dn_compose, dim3(g1), dim3(256), 0, s, n, light, albedo, out);
    DN_TRY(hipGetLastError());
    return MPT_OK;
}

int32_t mpt_denoise(const double* radiance, const double* albedo, const double* normal, const double* variance,
                    int32_t width, int32_t height, const mpt_denoise_opts* opts, double* out) {
    int rc = check_opts(opts, width, height);
    if (rc) return rc;
    if (!radiance || !albedo || !normal || !variance || !out) return set_error(MPT_E_INVALID, "denoise: null buffer");
    const size_t n = size_t(width) * height;
    const size_t img = n * 3 * sizeof(double);
    char* buf = nullptr;
    const size_t ws = mpt_denoise_workspace_bytes(width, height);
    DN_TRY(hipMalloc(&buf, 4 * img + n * sizeof(double) + ws));
    double* d_rad = reinterpret_cast<double*>(buf);
    double* d_alb = d_rad + 3 * n;
    double* d_nrm = d_alb + 3 * n;
    double* d_out = d_nrm + 3 * n;
    double* d_var = d_out + 3 * n;
    void* d_ws = d_var + n;
    auto fail = [&](int code) {
        (void)hipFree(buf);
        return code;
    };
    if (hipMemcpy(d_rad, radiance, img, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_alb, albedo, img, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_nrm, normal, img, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_var, variance, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(MPT_E_HIP, "denoise: copy to the device failed"));
    rc = mpt_denoise_dev(d_rad, d_alb, d_nrm, d_var, width, height, opts, d_out, d_ws, nullptr);
    if (rc) return fail(rc);
    if (hipMemcpy(out, d_out, img, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(set_error(MPT_E_HIP, "denoise: copy from the device failed"));
    return fail(MPT_OK);
}

}  // extern "C"
