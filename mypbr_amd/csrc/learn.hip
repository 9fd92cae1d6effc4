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
// Every convolution is an implicit GEMM on the matrix cores (k_conv3x3:
// the input tile and its halo staged in LDS, the nine taps read at shifted
// offsets, no im2col buffer):
//   forward    Y[p][co]   = sum_{t,ci} X[p + off(t)][ci] W[co][t][ci]
//   inputs     dX[p][ci]  = sum_{t,co} G[p + off(t)][co] W[co][8 - t][ci]  (the same kernel)
//   weights    dW[co][t][ci] = sum_p G[p][co] X[p + off(t)][ci]            (k_conv3x3_wgrad)
// on v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation.
// Activations are stored BF16 (NHWC), weights kept in FP32 (Adam state too)
// with a BF16 copy for the products; the kernel application, loss and their
// gradients are FP32.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "nn_common.hpp"

namespace mpt {
namespace nn {

// ------------------------------------------------------------------ implicit-GEMM convolution
// Y[p][co] = sum over taps t and channels ci of X[p + off(t)][ci] W[co][t * CIN + ci]
// (off(t) = (t / 3 - 1, t % 3 - 1) rows and columns, zero outside the image),
// computed on the matrix cores without an im2col buffer: a block stages its
// weights [COUT][9 CIN] in LDS once, then walks tiles of kTileH rows x
// kTileW columns of one image; each tile's input halo ((kTileH + 2) x
// (kTileW + 2) pixels x CIN) is staged in LDS, and the nine taps read it at
// shifted offsets. Wave w computes tile row w: 64 pixels = 4 fragments of 16
// against COUT / 16 fragments of 16 output channels, 32 k per MFMA step.
// The next tile's halo is loaded into registers while the current one is
// multiplied. One kernel serves the forward pass (EPI 2: + bias, ReLU,
// BF16; EPI 3: + bias, FP32 logits) and the input gradient (EPI 0: FP32),
// which is the same convolution of the output gradient with the weights
// flipped in the taps and transposed in the channels (k_flip_weights).
constexpr int kTileW = 64;
constexpr int kTileH = 4;
constexpr int kHaloW = kTileW + 2;
constexpr int kHaloH = kTileH + 2;
constexpr int kMaxCin = 64;  // input and output channels of one convolution
constexpr int kHaloChunks = (kHaloH * kHaloW * kMaxCin / 8 + 255) / 256;  // 16-byte loads per thread

struct conv_shape {
    int B, H, W, CIN;
    int tiles_x, tiles_y;
    __host__ __device__ int64_t ntiles() const { return int64_t(B) * tiles_x * tiles_y; }
};

__device__ __forceinline__ void halo_load(const conv_shape& S, const bf16* __restrict__ X, int64_t tile,
                                          uint4 (&pre)[kHaloChunks]) {
    const int tx = int(tile % S.tiles_x);
    const int ty = int((tile / S.tiles_x) % S.tiles_y);
    const int64_t b = tile / (int64_t(S.tiles_x) * S.tiles_y);
    const int c8n = S.CIN / 8;
    const int n = kHaloH * kHaloW * c8n;
#pragma unroll
    for (int r = 0; r < kHaloChunks; r++) {
        const int e = int(threadIdx.x) + 256 * r;
        uint4 v{0, 0, 0, 0};
        if (e < n) {
            const int c8 = e % c8n, pix = e / c8n;
            const int y = ty * kTileH + pix / kHaloW - 1, x = tx * kTileW + pix % kHaloW - 1;
            if (y >= 0 && y < S.H && x >= 0 && x < S.W)
                v = *reinterpret_cast<const uint4*>(X + ((b * S.H + y) * S.W + x) * S.CIN + c8 * 8);
        }
        pre[r] = v;
    }
}

template <int NF, int EPI>
__global__ __launch_bounds__(256) void k_conv3x3(conv_shape S, const bf16* __restrict__ X, const bf16* __restrict__ Wt,
                                                 const float* __restrict__ bias, void* __restrict__ Yv) {
    constexpr int COUT = 16 * NF;
    constexpr int kMaxKP = 9 * kMaxCin + 8;
    __shared__ __attribute__((aligned(16))) bf16 lw[COUT * kMaxKP];
    __shared__ __attribute__((aligned(16))) bf16 lx[kHaloH * kHaloW * (kMaxCin + 8)];
    const int CIN = S.CIN;
    const int K = 9 * CIN;
    const int KP = (K + 31) / 32 * 32 + 8;  // weight row stride in LDS (the pad k read as zero)
    const int XS = CIN + 8;                 // halo pixel stride in LDS
    for (int e = threadIdx.x; e < COUT * (KP / 8); e += 256) {
        const int co = e / (KP / 8), k8 = e % (KP / 8) * 8;
        uint4 v{0, 0, 0, 0};
        if (k8 < K) v = *reinterpret_cast<const uint4*>(Wt + int64_t(co) * K + k8);
        *reinterpret_cast<uint4*>(lw + co * KP + k8) = v;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int c8n = CIN / 8;
    const int nchunks = kHaloH * kHaloW * c8n;
    const int64_t ntiles = S.ntiles();
    uint4 pre[kHaloChunks];
    if (int64_t(blockIdx.x) < ntiles) halo_load(S, X, blockIdx.x, pre);
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();  // the previous tile's reads of the halo are done (and the weights are staged)
#pragma unroll
        for (int r = 0; r < kHaloChunks; r++) {
            const int e = int(threadIdx.x) + 256 * r;
            if (e < nchunks) *reinterpret_cast<uint4*>(lx + (e / c8n) * XS + (e % c8n) * 8) = pre[r];
        }
        __syncthreads();
        if (tile + gridDim.x < ntiles) halo_load(S, X, tile + gridDim.x, pre);
        float4_t acc[4][NF];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < NF; j++) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < K; kc += 32) {
            const int k = kc + 8 * g;
            const int t = k / CIN, ci = k - t * CIN;
            const bool tap = t < 9;
            const int dy = tap ? t / 3 : 0, dx = tap ? t % 3 : 0;
            bf16x8_t a[4], b[NF];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint4 u = *reinterpret_cast<const uint4*>(lx + ((w + dy) * kHaloW + 16 * i + r16 + dx) * XS +
                                                          (tap ? ci : 0));
                if (!tap) u = uint4{0, 0, 0, 0};
                a[i] = as_bf16x8(u);
            }
#pragma unroll
            for (int j = 0; j < NF; j++) b[j] = as_bf16x8(*reinterpret_cast<const uint4*>(lw + (16 * j + r16) * KP + k));
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < NF; j++) acc[i][j] = mfma32(a[i], b[j], acc[i][j]);
        }
        const int tx = int(tile % S.tiles_x);
        const int ty = int((tile / S.tiles_x) % S.tiles_y);
        const int64_t bimg = tile / (int64_t(S.tiles_x) * S.tiles_y);
        const int y = ty * kTileH + w;
        if (y >= S.H) continue;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int x = tx * kTileW + 16 * i + 4 * g + q;
                if (x >= S.W) continue;
                const int64_t p = (bimg * S.H + y) * S.W + x;
#pragma unroll
                for (int j = 0; j < NF; j++) {
                    const int co = 16 * j + r16;
                    float v = acc[i][j][q];
                    if constexpr (EPI == 0) {
                        static_cast<float*>(Yv)[p * COUT + co] = v;
                    } else if constexpr (EPI == 3) {
                        static_cast<float*>(Yv)[p * COUT + co] = v + bias[co];
                    } else {
                        v += bias[co];
                        if (!(v > 0.f)) v = 0.f;
                        static_cast<bf16*>(Yv)[p * COUT + co] = to_bf16(v);
                    }
                }
            }
    }
}

// Wf[ci][t][co] = W[co][8 - t][ci]: the weights of the input-gradient
// convolution dX[p][ci] = sum_t sum_co G[p - off(t)][co] W[co][t][ci]
// = sum_t' sum_co G[p + off(t')][co] W[co][8 - t'][ci].
__global__ void k_flip_weights(int COUT, int CIN, const bf16* __restrict__ W, bf16* __restrict__ Wf) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // index into Wf
    if (e >= int64_t(COUT) * 9 * CIN) return;
    const int co = int(e % COUT);
    const int t = int((e / COUT) % 9);
    const int ci = int(e / (int64_t(COUT) * 9));
    Wf[e] = W[(int64_t(co) * 9 + (8 - t)) * CIN + ci];
}

// dW[co][t * CIN + ci] += sum_p G[p][co] X[p + off(t)][ci]: the weight
// gradient as a GEMM over pixels (M = COUT, N = CIN, K = pixels), one tap per
// blockIdx.y. A block walks chunks of 256 pixels, stages the chunk's output
// gradient and the tap-shifted inputs transposed ([channel][pixel]) in LDS,
// so that a lane's eight consecutive pixels are one 16-byte read, and adds
// its partial sums into dW atomically at the end. FPW: 16 x 16 fragments of
// dW per wave.
template <int FPW>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(conv_shape S, int COUT, const bf16* __restrict__ G,
                                                       const bf16* __restrict__ X, float* __restrict__ dW) {
    constexpr int kChunk = 256, kRow = kChunk + 8;
    __shared__ __attribute__((aligned(16))) bf16 gt[kMaxCin * kRow];
    __shared__ __attribute__((aligned(16))) bf16 xt[kMaxCin * kRow];
    const int CIN = S.CIN;
    const int t = blockIdx.y, dy = t / 3 - 1, dx = t % 3 - 1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int nfr = CIN / 16, nfrags = (COUT / 16) * nfr;
    const int64_t npix = int64_t(S.B) * S.H * S.W;
    float4_t acc[FPW];
#pragma unroll
    for (int q = 0; q < FPW; q++) acc[q] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int64_t p0 = int64_t(blockIdx.x) * kChunk; p0 < npix; p0 += int64_t(gridDim.x) * kChunk) {
        __syncthreads();
        for (int e = threadIdx.x; e < kChunk * (COUT / 8); e += 256) {
            const int px = e % kChunk, c8 = e / kChunk;
            const int64_t p = p0 + px;
            uint4 v{0, 0, 0, 0};
            if (p < npix) v = *reinterpret_cast<const uint4*>(G + p * COUT + c8 * 8);
            const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int q = 0; q < 8; q++) gt[(c8 * 8 + q) * kRow + px] = h[q];
        }
        for (int e = threadIdx.x; e < kChunk * (CIN / 8); e += 256) {
            const int px = e % kChunk, c8 = e / kChunk;
            const int64_t p = p0 + px;
            uint4 v{0, 0, 0, 0};
            if (p < npix) {
                const int x = int(p % S.W) + dx, y = int((p / S.W) % S.H) + dy;
                if (x >= 0 && x < S.W && y >= 0 && y < S.H)
                    v = *reinterpret_cast<const uint4*>(X + (p + int64_t(dy) * S.W + dx) * CIN + c8 * 8);
            }
            const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int q = 0; q < 8; q++) xt[(c8 * 8 + q) * kRow + px] = h[q];
        }
        __syncthreads();
        for (int kc = 0; kc < kChunk; kc += 32) {
#pragma unroll
            for (int q = 0; q < FPW; q++) {
                const int f = w + 4 * q;
                if (f >= nfrags) break;
                const int mi = f / nfr, ni = f % nfr;
                const bf16x8_t a = as_bf16x8(*reinterpret_cast<const uint4*>(gt + (16 * mi + r16) * kRow + kc + 8 * g));
                const bf16x8_t b = as_bf16x8(*reinterpret_cast<const uint4*>(xt + (16 * ni + r16) * kRow + kc + 8 * g));
                acc[q] = mfma32(a, b, acc[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < FPW; q++) {
        const int f = w + 4 * q;
        if (f >= nfrags) break;
        const int mi = f / nfr, ni = f % nfr;
        const int ci = 16 * ni + r16;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int co = 16 * mi + 4 * g + r;
            atomicAdd(dW + (int64_t(co) * 9 + t) * CIN + ci, acc[q][r]);
        }
    }
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
    bf16* wflip = nullptr;   // one layer's weights flipped for its input gradient
    float* logits = nullptr;
    float* dA = nullptr;   // FP32 gradient w.r.t. the current layer's output
    float* dB = nullptr;
    bf16* G = nullptr;
    float* light = nullptr;
    float* albe = nullptr;
    double* loss = nullptr;
    hipStream_t stream = nullptr;
    int cu_count = 256;
    ~mpt_kpcnn() {
        for (void* p : {(void*)w, (void*)wb, (void*)g, (void*)m, (void*)v}) (void)hipFree(p);
        release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void release() {
        for (bf16* a : act) (void)hipFree(a);
        act.clear();
        for (void* p : {(void*)wflip, (void*)logits, (void*)dA, (void*)dB, (void*)G, (void*)light, (void*)albe,
                        (void*)loss})
            (void)hipFree(p);
        wflip = nullptr, logits = nullptr, dA = nullptr, dB = nullptr, G = nullptr, light = nullptr;
        albe = nullptr, loss = nullptr;
        cap = 0;
    }
    int64_t wcount(int l) const { return int64_t(cout[l]) * 9 * cin[l]; }
};

namespace {

int ensure(mpt_kpcnn* k, int64_t npix) {
    if (k->cap >= npix) return MPT_OK;
    k->release();
    int maxc = kFeat;
    int64_t maxw = 0;
    for (int l = 0; l < k->layers; l++) {
        maxc = std::max(maxc, std::max(k->cin[l], k->cout[l]));
        maxw = std::max(maxw, k->wcount(l));
    }
    k->act.assign(k->layers + 1, nullptr);
    for (int l = 0; l <= k->layers; l++) {
        const int c = l == 0 ? kFeat : k->cout[l - 1];
        NN_TRY(hipMalloc(&k->act[l], size_t(npix) * c * sizeof(bf16)));
    }
    NN_TRY(hipMalloc(&k->wflip, size_t(maxw) * sizeof(bf16)));
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

conv_shape shape_of(int B, int H, int W, int cin) {
    return conv_shape{B, H, W, cin, (W + kTileW - 1) / kTileW, (H + kTileH - 1) / kTileH};
}

// One implicit-GEMM convolution (k_conv3x3) of cout output channels; a
// block per CU walks the tiles (each block holds ~130 KB of LDS).
int conv(mpt_kpcnn* k, const conv_shape& S, int cout, int epi, const bf16* X, const bf16* Wt, const float* bias,
         void* Y) {
    const unsigned grid = unsigned(std::max<int64_t>(1, std::min<int64_t>(S.ntiles(), k->cu_count)));
    hipStream_t s = k->stream;
#define CONV_CASE(NF, E)                                                                                     \
    if (cout == 16 * NF && epi == E) {                                                                       \
        hipLaunchKernelGGL((k_conv3x3<NF, E>), dim3(grid), dim3(256), 0, s, S, X, Wt, bias, Y);              \
        return MPT_OK;                                                                                       \
    }
    CONV_CASE(1, 0) CONV_CASE(2, 0) CONV_CASE(3, 0) CONV_CASE(4, 0)
    CONV_CASE(1, 2) CONV_CASE(2, 2) CONV_CASE(3, 2) CONV_CASE(4, 2)
    CONV_CASE(2, 3)
#undef CONV_CASE
    return set_error(MPT_E_INVALID, "kpcnn: unsupported convolution width");
}

// dW of one layer: k_conv3x3_wgrad over every pixel, one tap per grid row.
int conv_wgrad(mpt_kpcnn* k, const conv_shape& S, int cout, const bf16* G, const bf16* X, float* dW) {
    const int64_t npix = int64_t(S.B) * S.H * S.W;
    const int64_t chunks = (npix + 255) / 256;
    const unsigned gx = unsigned(std::max<int64_t>(1, std::min<int64_t>(chunks, (4 * int64_t(k->cu_count) + 8) / 9)));
    const int fpw = ((cout / 16) * (S.CIN / 16) + 3) / 4;
    hipStream_t s = k->stream;
    if (fpw <= 1)
        hipLaunchKernelGGL((k_conv3x3_wgrad<1>), dim3(gx, 9), dim3(256), 0, s, S, cout, G, X, dW);
    else if (fpw <= 2)
        hipLaunchKernelGGL((k_conv3x3_wgrad<2>), dim3(gx, 9), dim3(256), 0, s, S, cout, G, X, dW);
    else if (fpw <= 4)
        hipLaunchKernelGGL((k_conv3x3_wgrad<4>), dim3(gx, 9), dim3(256), 0, s, S, cout, G, X, dW);
    else
        return set_error(MPT_E_INVALID, "kpcnn: unsupported convolution width");
    return MPT_OK;
}

// Forward through every layer from act[0]; logits in k->logits.
int forward(mpt_kpcnn* k, int B, int H, int W) {
    for (int l = 0; l < k->layers; l++) {
        const bool last = l == k->layers - 1;
        int rc = conv(k, shape_of(B, H, W, k->cin[l]), k->cout[l], last ? 3 : 2, k->act[l], k->wb + k->woff[l],
                      k->w + k->boff[l], last ? (void*)k->logits : (void*)k->act[l + 1]);
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
        // weights and bias of the layer are contiguous in the gradient vector: zero both
        NN_TRY(hipMemsetAsync(k->g + k->woff[l], 0, (k->wcount(l) + co) * sizeof(float), s));
        const int bs = (256 / co) * co;
        const unsigned gb = unsigned(std::max<int64_t>(1, std::min<int64_t>(int64_t(k->cu_count) * 8,
                                                                             (npix * co + bs - 1) / bs)));
        hipLaunchKernelGGL(k_relu_grad, dim3(gb), dim3(bs), 0, s, npix * co, co, dY, k->act[l + 1], last ? 0 : 1,
                           k->G, k->g + k->boff[l]);
        int rc = conv_wgrad(k, shape_of(B, H, W, ci), co, k->G, k->act[l], k->g + k->woff[l]);
        if (rc) return rc;
        if (l == 0) break;
        // dX = the convolution of G with the flipped, transposed weights
        hipLaunchKernelGGL(k_flip_weights, dim3(blocks(k->wcount(l))), dim3(256), 0, s, co, ci, k->wb + k->woff[l],
                           k->wflip);
        rc = conv(k, shape_of(B, H, W, co), ci, 0, k->G, k->wflip, nullptr, dX);
        if (rc) return rc;
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
    if (layers < 2 || layers > 16 || hidden < 16 || hidden > kMaxCin || hidden % 16)
        return set_error(MPT_E_INVALID, "kpcnn: layers in [2, 16], hidden a multiple of 16 in [16, 64]");
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
    int dev = 0;
    NN_TRY(hipGetDevice(&dev));
    NN_TRY(hipDeviceGetAttribute(&k->cu_count, hipDeviceAttributeMultiprocessorCount, dev));
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
