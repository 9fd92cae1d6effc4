// learn.hip - the learned denoiser: a kernel-predicting CNN (KPCNN) trained
// and run on the GPU with hand-written BF16 matrix-core kernels.
//
// The network reads, per pixel, the noisy render's demodulated lighting
// (radiance / (albedo + 1e-3), as log(1 + x)), albedo, normal, log depth and
// log relative variance (kFeat = 16 channels, zero-padded), runs `layers`
// 3x3 convolutions (ReLU between them, `hidden` channels), and the last one
// predicts 25 logits per pixel: the softmax of those weights a 5x5
// neighbourhood of the noisy lighting (clamped at the border), which is then
// multiplied back by the albedo (Bako et al. 2017, kernel prediction).
//
// Every convolution is an im2col + GEMM on the matrix cores:
//   forward    Y[p][co]   = sum_k Xcol[p][k] W[co][k]          (k = tap * Cin + ci)
//   weights    dW[co][k]  = sum_p G[p][co] Xcol[p][k]
//   inputs     dXcol[p][k]= sum_co G[p][co] W[co][k], then col2im over the taps
// with one GEMM kernel, k_gemm, on v_mfma_f32_16x16x16_bf16: BF16 operands,
// FP32 accumulation. Activations are stored BF16 (NHWC), weights kept in
// FP32 (Adam state too) with a BF16 copy for the GEMMs; the kernel
// application, loss and their gradients are FP32.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "nn_common.hpp"

namespace mpt {
namespace nn {

// ------------------------------------------------------------------ conv helpers
// im2col of X [B][H][W][Cin] (BF16) into Xcol [B*H*W][9*Cin]: tap t = (dy + 1) * 3 + (dx + 1),
// zero outside the image.
__global__ void k_im2col(int64_t npix, int H, int W, int Cin, const bf16* __restrict__ X, bf16* __restrict__ Xcol) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // one (pixel, tap) pair
    if (e >= npix * 9) return;
    const int64_t p = e / 9;
    const int t = int(e % 9);
    const int x = int(p % W), y = int((p / W) % H);
    const int64_t b = p / (int64_t(W) * H);
    const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    // Cin is a multiple of 8: whole 16-byte groups
    uint4* dst = reinterpret_cast<uint4*>(Xcol + p * 9 * Cin + t * Cin);
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) {
        for (int c = 0; c < Cin / 8; c++) dst[c] = uint4{0, 0, 0, 0};
        return;
    }
    const uint4* src = reinterpret_cast<const uint4*>(X + ((b * H + yy) * W + xx) * Cin);
    for (int c = 0; c < Cin / 8; c++) dst[c] = src[c];
}

// dX[p][c] = sum over taps of dXcol[q][t * Cin + c] where q is the pixel whose
// tap t reads p (q = p - offset(t)); FP32.
__global__ void k_col2im(int64_t npix, int H, int W, int Cin, const float* __restrict__ dXcol, float* __restrict__ dX) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // one (pixel, channel) pair
    if (e >= npix * Cin) return;
    const int64_t p = e / Cin;
    const int c = int(e % Cin);
    const int x = int(p % W), y = int((p / W) % H);
    const int64_t b = p / (int64_t(W) * H);
    float s = 0.f;
    for (int t = 0; t < 9; t++) {
        const int qy = y - (t / 3 - 1), qx = x - (t % 3 - 1);
        if (qy < 0 || qy >= H || qx < 0 || qx >= W) continue;
        s += dXcol[((b * H + qy) * W + qx) * 9 * Cin + t * Cin + c];
    }
    dX[e] = s;
}

// ------------------------------------------------------------------ features and kernel application
constexpr int kFeat = 16;   // input channels (10 used)
constexpr int kTaps = 25;   // 5x5 predicted kernel
constexpr int kOut = 32;    // logits channels of the last layer (25 used)
constexpr float kAlbEps = 1e-3f;

// Per-pixel inputs from a frame in double (radiance, albedo, normal: 3 per
// pixel; variance, depth: 1): features (BF16, kFeat channels), the noisy
// lighting and the albedo + eps (FP32, 3 each) that the kernel application
// uses.
__global__ void k_features(int64_t npix, const double* __restrict__ rad, const double* __restrict__ alb,
                           const double* __restrict__ nrm, const double* __restrict__ var,
                           const double* __restrict__ depth, bf16* __restrict__ feat, float* __restrict__ light,
                           float* __restrict__ albe) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    float f[kFeat];
    for (int c = 0; c < kFeat; c++) f[c] = 0.f;
    float lum = 0.f;
    for (int c = 0; c < 3; c++) {
        const float a = float(alb[3 * p + c]) + kAlbEps;
        const float l = float(rad[3 * p + c]) / a;
        light[3 * p + c] = l;
        albe[3 * p + c] = a;
        f[c] = std::log1p(std::fmax(l, 0.f));
        f[3 + c] = float(alb[3 * p + c]);
        f[6 + c] = float(nrm[3 * p + c]);
    }
    lum = 0.2126f * float(rad[3 * p]) + 0.7152f * float(rad[3 * p + 1]) + 0.0722f * float(rad[3 * p + 2]);
    f[9] = std::log1p(std::fmax(float(depth[p]), 0.f));
    f[10] = std::log1p(std::fmax(float(var[p]), 0.f) / (lum * lum + 1e-2f));
    for (int c = 0; c < kFeat; c++) feat[p * kFeat + c] = to_bf16(f[c]);
}

// Softmax of the 25 logits, the weighted 5x5 sum of the noisy lighting
// (coordinates clamped to the image), times albedo + eps. With a target:
// loss += relMSE terms (o - t)^2 / (t^2 + 0.01) / (3 * npix) and the
// gradient of the loss with respect to the logits (FP32, kOut per pixel).
// Loss modes: 0 relMSE (o - t)^2 / (t^2 + 0.01); 1 L1 of log(1 + x),
// |log(1 + max(o, 0)) - log(1 + max(t, 0))|; both averaged over pixels and channels.
__global__ void k_apply(int64_t npix, int H, int W, const float* __restrict__ logits, const float* __restrict__ light,
                        const float* __restrict__ albe, const float* __restrict__ target, float* __restrict__ out,
                        float* __restrict__ dlogits, double* __restrict__ loss, int mode) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double lsum = 0;
    if (p < npix) {
        const int x = int(p % W), y = int((p / W) % H);
        const int64_t base = p - (int64_t(y) * W + x);
        const float* lg = logits + p * kOut;
        float mx = lg[0];
        for (int k = 1; k < kTaps; k++) mx = std::fmax(mx, lg[k]);
        float w[kTaps];
        float s = 0.f;
        for (int k = 0; k < kTaps; k++) {
            w[k] = std::exp(lg[k] - mx);
            s += w[k];
        }
        const float inv = 1.f / s;
        float o[3] = {0.f, 0.f, 0.f};
        int64_t q[kTaps];
        for (int k = 0; k < kTaps; k++) {
            w[k] *= inv;
            int yy = y + k / 5 - 2, xx = x + k % 5 - 2;
            yy = yy < 0 ? 0 : (yy >= H ? H - 1 : yy);
            xx = xx < 0 ? 0 : (xx >= W ? W - 1 : xx);
            q[k] = base + int64_t(yy) * W + xx;
            for (int c = 0; c < 3; c++) o[c] += w[k] * light[3 * q[k] + c];
        }
        float dl[3];
        for (int c = 0; c < 3; c++) {
            const float a = albe[3 * p + c];
            const float rad = o[c] * a;
            if (out) out[3 * p + c] = rad;
            if (target) {
                const float t = target[3 * p + c];
                if (mode == 1) {
                    const float ro = std::fmax(rad, 0.f);
                    const float d = std::log1p(ro) - std::log1p(std::fmax(t, 0.f));
                    lsum += double(std::fabs(d));
                    const float sg = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
                    dl[c] = rad > 0.f ? sg / (1.f + ro) / float(3 * npix) * a : 0.f;
                } else {
                    const float den = t * t + 1e-2f;
                    const float d = rad - t;
                    lsum += double(d * d / den);
                    dl[c] = 2.f * d / den / float(3 * npix) * a;  // d loss / d o[c]
                }
            }
        }
        if (dlogits) {
            float gw[kTaps];
            float dot = 0.f;
            for (int k = 0; k < kTaps; k++) {
                gw[k] = dl[0] * light[3 * q[k]] + dl[1] * light[3 * q[k] + 1] + dl[2] * light[3 * q[k] + 2];
                dot += w[k] * gw[k];
            }
            float* dg = dlogits + p * kOut;
            for (int k = 0; k < kTaps; k++) dg[k] = w[k] * (gw[k] - dot);
            for (int k = kTaps; k < kOut; k++) dg[k] = 0.f;
        }
    }
    if (loss) {  // one atomic per wave
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if ((threadIdx.x & 63) == 0 && lsum != 0) atomicAdd(loss, lsum);
    }
}

}  // namespace nn
}  // namespace mpt

// ======================================================================== host side
using namespace mpt;
using namespace mpt::nn;

struct mpt_kpcnn {
    int layers = 6, hidden = 64;
    std::vector<int> cin, cout;   // per layer (Cin padded to 16)
    std::vector<int64_t> woff, boff;  // offsets into the parameter vector
    int64_t nparams = 0;
    float* w = nullptr;    // FP32 master parameters (weights then biases per layer)
    bf16* wb = nullptr;    // BF16 copy
    float* g = nullptr;    // gradients
    float* m = nullptr;    // Adam moments
    float* v = nullptr;
    int64_t step = 0;
    int loss_mode = 0;
    // activation workspace for a given pixel count
    int64_t cap = 0;
    std::vector<bf16*> act;  // act[0] = features, act[l + 1] = output of layer l (BF16), last unused
    bf16* xcol = nullptr;
    float* dcol = nullptr;
    float* logits = nullptr;
    float* dA = nullptr;   // FP32 gradient w.r.t. the current layer's output
    float* dB = nullptr;
    bf16* G = nullptr;
    float* light = nullptr;
    float* albe = nullptr;
    double* loss = nullptr;
    hipStream_t stream = nullptr;
    ~mpt_kpcnn() {
        for (void* p : {(void*)w, (void*)wb, (void*)g, (void*)m, (void*)v}) (void)hipFree(p);
        release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void release() {
        for (bf16* a : act) (void)hipFree(a);
        act.clear();
        for (void* p : {(void*)xcol, (void*)dcol, (void*)logits, (void*)dA, (void*)dB, (void*)G, (void*)light,
                        (void*)albe, (void*)loss})
            (void)hipFree(p);
        xcol = nullptr, dcol = nullptr, logits = nullptr, dA = nullptr, dB = nullptr, G = nullptr, light = nullptr;
        albe = nullptr, loss = nullptr;
        cap = 0;
    }
    int64_t wcount(int l) const { return int64_t(cout[l]) * 9 * cin[l]; }
};

namespace {

int ensure(mpt_kpcnn* k, int64_t npix) {
    if (k->cap >= npix) return MPT_OK;
    k->release();
    int maxc = kFeat, maxk = 0;
    for (int l = 0; l < k->layers; l++) {
        maxc = std::max(maxc, std::max(k->cin[l], k->cout[l]));
        maxk = std::max(maxk, 9 * k->cin[l]);
    }
    k->act.assign(k->layers + 1, nullptr);
    for (int l = 0; l <= k->layers; l++) {
        const int c = l == 0 ? kFeat : k->cout[l - 1];
        NN_TRY(hipMalloc(&k->act[l], size_t(npix) * c * sizeof(bf16)));
    }
    NN_TRY(hipMalloc(&k->xcol, size_t(npix) * maxk * sizeof(bf16)));
    NN_TRY(hipMalloc(&k->dcol, size_t(npix) * maxk * sizeof(float)));
    NN_TRY(hipMalloc(&k->logits, size_t(npix) * kOut * sizeof(float)));
    NN_TRY(hipMalloc(&k->dA, size_t(npix) * maxc * sizeof(float)));
    NN_TRY(hipMalloc(&k->dB, size_t(npix) * maxc * sizeof(float)));
    NN_TRY(hipMalloc(&k->G, size_t(npix) * maxc * sizeof(bf16)));
    NN_TRY(hipMalloc(&k->light, size_t(npix) * 3 * sizeof(float)));
    NN_TRY(hipMalloc(&k->albe, size_t(npix) * 3 * sizeof(float)));
    NN_TRY(hipMalloc(&k->loss, sizeof(double)));
    k->cap = npix;
    return MPT_OK;
}

// Forward through every layer from act[0]; logits in k->logits.
int forward(mpt_kpcnn* k, int B, int H, int W) {
    const int64_t npix = int64_t(B) * H * W;
    hipStream_t s = k->stream;
    for (int l = 0; l < k->layers; l++) {
        const int ci = k->cin[l], co = k->cout[l];
        hipLaunchKernelGGL(k_im2col, dim3(blocks(npix * 9)), dim3(256), 0, s, npix, H, W, ci, k->act[l], k->xcol);
        const bool last = l == k->layers - 1;
        int rc = gemm(false, false, last ? 3 : 2, int(npix), co, 9 * ci, k->xcol, 9 * ci, k->wb + k->woff[l], 9 * ci,
                      last ? (void*)k->logits : (void*)k->act[l + 1], co, k->w + k->boff[l], false, s);
        if (rc) return rc;
    }
    return MPT_OK;
}

// Backward from dlogits (in k->dA, FP32 [npix][kOut]); gradients into k->g.
int backward(mpt_kpcnn* k, int B, int H, int W) {
    const int64_t npix = int64_t(B) * H * W;
    hipStream_t s = k->stream;
    float* dY = k->dA;
    float* dX = k->dB;
    for (int l = k->layers - 1; l >= 0; l--) {
        const int ci = k->cin[l], co = k->cout[l];
        const bool last = l == k->layers - 1;
        hipLaunchKernelGGL(k_relu_grad, dim3(blocks(npix * co)), dim3(256), 0, s, npix * co, dY, k->act[l + 1],
                           last ? 0 : 1, k->G);
        hipLaunchKernelGGL(k_bias_grad, dim3(co), dim3(256), 0, s, npix, co, k->G, k->g + k->boff[l]);
        hipLaunchKernelGGL(k_im2col, dim3(blocks(npix * 9)), dim3(256), 0, s, npix, H, W, ci, k->act[l], k->xcol);
        // dW[co][k] = sum_p G[p][co] Xcol[p][k], K = pixels split over blocks
        NN_TRY(hipMemsetAsync(k->g + k->woff[l], 0, k->wcount(l) * sizeof(float), s));
        int rc = gemm(true, true, 4, co, 9 * ci, int(npix), k->G, co, k->xcol, 9 * ci, k->g + k->woff[l], 9 * ci,
                      nullptr, false, s);
        if (rc) return rc;
        if (l == 0) break;
        // dXcol[p][k] = sum_co G[p][co] W[co][k]
        rc = gemm(false, true, 0, int(npix), 9 * ci, co, k->G, co, k->wb + k->woff[l], 9 * ci, k->dcol, 9 * ci,
                  nullptr, false, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_col2im, dim3(blocks(npix * ci)), dim3(256), 0, s, npix, H, W, ci, k->dcol, dX);
        std::swap(dY, dX);
    }
    return MPT_OK;
}

int prepare_inputs(mpt_kpcnn* k, const double* rad, const double* alb, const double* nrm, const double* var,
                   const double* depth, int64_t npix) {
    hipLaunchKernelGGL(k_features, dim3(blocks(npix)), dim3(256), 0, k->stream, npix, rad, alb, nrm, var, depth,
                       k->act[0], k->light, k->albe);
    return MPT_OK;
}

int check_shape(int B, int H, int W) {
    if (B <= 0 || H <= 0 || W <= 0) return set_error(MPT_E_INVALID, "kpcnn: batch and image sizes must be positive");
    if (int64_t(B) * H * W > (int64_t(1) << 28)) return set_error(MPT_E_INVALID, "kpcnn: too many pixels");
    return MPT_OK;
}

}  // namespace

extern "C" {

int32_t mpt_gemm_bf16(int32_t a_t, int32_t b_t, int32_t M, int32_t N, int32_t K, const uint16_t* A, int64_t lda,
                      const uint16_t* B, int64_t ldb, float* C, int64_t ldc, int32_t accumulate, void* stream) {
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return set_error(MPT_E_INVALID, "gemm: bad arguments");
    int rc = gemm(a_t != 0, b_t != 0, 0, M, N, K, A, lda, B, ldb, C, ldc, nullptr, accumulate != 0,
                  static_cast<hipStream_t>(stream));
    if (rc) return rc;
    NN_TRY(hipGetLastError());
    return MPT_OK;
}

int32_t mpt_kpcnn_create(int32_t layers, int32_t hidden, uint64_t seed, mpt_kpcnn** out) {
    if (!out) return set_error(MPT_E_INVALID, "kpcnn: null output");
    if (layers < 2 || layers > 16 || hidden < 16 || hidden > 256 || hidden % 16)
        return set_error(MPT_E_INVALID, "kpcnn: layers in [2, 16], hidden a multiple of 16 in [16, 256]");
    auto k = std::make_unique<mpt_kpcnn>();
    k->layers = layers;
    k->hidden = hidden;
    for (int l = 0; l < layers; l++) {
        k->cin.push_back(l == 0 ? kFeat : hidden);
        k->cout.push_back(l == layers - 1 ? kOut : hidden);
    }
    int64_t off = 0;
    for (int l = 0; l < layers; l++) {
        k->woff.push_back(off);
        off += k->wcount(l);
        k->boff.push_back(off);
        off += k->cout[l];
    }
    k->nparams = off;
    // He-uniform initialisation from a SplitMix64 stream; the last layer
    // starts small so the first kernels are near uniform.
    std::vector<float> h(off, 0.f);
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * 0x1.0p-53;
    };
    for (int l = 0; l < layers; l++) {
        const int fan_in = 9 * (l == 0 ? 11 : hidden);
        const double lim = std::sqrt(6.0 / fan_in) * (l == layers - 1 ? 0.1 : 1.0);
        for (int co = 0; co < k->cout[l]; co++)
            for (int t = 0; t < 9; t++)
                for (int ci = 0; ci < k->cin[l]; ci++) {
                    const bool used = (l > 0 || ci < 11) && (l < layers - 1 || co < kTaps);
                    const double u = next();
                    h[k->woff[l] + (int64_t(co) * 9 + t) * k->cin[l] + ci] = used ? float((2 * u - 1) * lim) : 0.f;
                }
    }
    NN_TRY(hipStreamCreate(&k->stream));
    NN_TRY(hipMalloc(&k->w, off * sizeof(float)));
    NN_TRY(hipMalloc(&k->wb, off * sizeof(bf16)));
    NN_TRY(hipMalloc(&k->g, off * sizeof(float)));
    NN_TRY(hipMalloc(&k->m, off * sizeof(float)));
    NN_TRY(hipMalloc(&k->v, off * sizeof(float)));
    NN_TRY(hipMemcpy(k->w, h.data(), off * sizeof(float), hipMemcpyHostToDevice));
    NN_TRY(hipMemset(k->m, 0, off * sizeof(float)));
    NN_TRY(hipMemset(k->v, 0, off * sizeof(float)));
    NN_TRY(hipMemset(k->g, 0, off * sizeof(float)));
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(off)), dim3(256), 0, k->stream, off, k->w, k->wb);
    NN_TRY(hipStreamSynchronize(k->stream));
    *out = k.release();
    return MPT_OK;
}

void mpt_kpcnn_destroy(mpt_kpcnn* k) { delete k; }

int32_t mpt_kpcnn_set_loss(mpt_kpcnn* k, int32_t mode) {
    if (!k || mode < 0 || mode > 1) return set_error(MPT_E_INVALID, "kpcnn: loss mode 0 (relMSE) or 1 (log L1)");
    k->loss_mode = mode;
    return MPT_OK;
}

int64_t mpt_kpcnn_param_count(const mpt_kpcnn* k) { return k ? k->nparams : 0; }

int32_t mpt_kpcnn_layout(const mpt_kpcnn* k, int32_t* cin, int32_t* cout, int64_t* woff, int64_t* boff) {
    if (!k) return set_error(MPT_E_INVALID, "kpcnn: null model");
    for (int l = 0; l < k->layers; l++) {
        if (cin) cin[l] = k->cin[l];
        if (cout) cout[l] = k->cout[l];
        if (woff) woff[l] = k->woff[l];
        if (boff) boff[l] = k->boff[l];
    }
    return MPT_OK;
}

int32_t mpt_kpcnn_get_params(mpt_kpcnn* k, float* params, float* grads) {
    if (!k) return set_error(MPT_E_INVALID, "kpcnn: null model");
    NN_TRY(hipStreamSynchronize(k->stream));
    if (params) NN_TRY(hipMemcpy(params, k->w, k->nparams * sizeof(float), hipMemcpyDeviceToHost));
    if (grads) NN_TRY(hipMemcpy(grads, k->g, k->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return MPT_OK;
}

int32_t mpt_kpcnn_set_params(mpt_kpcnn* k, const float* params) {
    if (!k || !params) return set_error(MPT_E_INVALID, "kpcnn: null argument");
    NN_TRY(hipMemcpy(k->w, params, k->nparams * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(k->nparams)), dim3(256), 0, k->stream, k->nparams, k->w, k->wb);
    NN_TRY(hipStreamSynchronize(k->stream));
    return MPT_OK;
}

// One forward (and with a target, the loss, backward and, with lr > 0, an
// Adam step) over B images of H x W. Inputs are device arrays in double:
// radiance, albedo, normal (3 per pixel), variance, depth (1 per pixel);
// target radiance in FP32 (3 per pixel) or NULL; out radiance (FP32, 3 per
// pixel) or NULL. loss (host) receives the mean relMSE when a target is given.
int32_t mpt_kpcnn_run(mpt_kpcnn* k, const double* rad, const double* alb, const double* nrm, const double* var,
                      const double* depth, int32_t B, int32_t H, int32_t W, const float* target, float lr,
                      float* out, double* loss) {
    if (!k || !rad || !alb || !nrm || !var || !depth) return set_error(MPT_E_INVALID, "kpcnn: null argument");
    int rc = check_shape(B, H, W);
    if (rc) return rc;
    const int64_t npix = int64_t(B) * H * W;
    if ((rc = ensure(k, npix))) return rc;
    hipStream_t s = k->stream;
    if ((rc = prepare_inputs(k, rad, alb, nrm, var, depth, npix))) return rc;
    if ((rc = forward(k, B, H, W))) return rc;
    NN_TRY(hipMemsetAsync(k->loss, 0, sizeof(double), s));
    hipLaunchKernelGGL(k_apply, dim3(blocks(npix)), dim3(256), 0, s, npix, H, W, k->logits, k->light, k->albe,
                       target, out, target ? k->dA : nullptr, target ? k->loss : nullptr, k->loss_mode);
    if (target) {
        if ((rc = backward(k, B, H, W))) return rc;
        if (lr > 0) {
            k->step++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - std::pow(b1, float(k->step)), c2 = 1.f - std::pow(b2, float(k->step));
            hipLaunchKernelGGL(k_adam, dim3(blocks(k->nparams)), dim3(256), 0, s, k->nparams, k->w, k->wb, k->g,
                               k->m, k->v, lr, b1, b2, 1e-8f, c1, c2);
        }
    }
    NN_TRY(hipGetLastError());
    if (loss) {
        double l = 0;
        NN_TRY(hipMemcpyAsync(&l, k->loss, sizeof(double), hipMemcpyDeviceToHost, s));
        NN_TRY(hipStreamSynchronize(s));
        *loss = l / double(3 * npix);
    } else {
        NN_TRY(hipStreamSynchronize(s));
    }
    return MPT_OK;
}

}  // extern "C"
