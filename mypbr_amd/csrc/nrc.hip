// nrc.hip - the neural radiance cache: a multiresolution hash-grid encoding
// of position (Mueller et al. 2022) and a small MLP (Mueller et al. 2021),
// trained online on the path tracer's own path-suffix radiance and queried
// where paths are cut short (include/mpt.h, mpt_nrc_*; the renderer side is
// in kernels.hip).
//
// Encoding: kLevels = 16 levels of kF = 2 features each, tables of T = 2^19
// FP16 entries per level (dense indexing where the level's grid fits),
// resolutions growing geometrically from 16 to 2048 over the scene box;
// trilinear interpolation of the 8 corners. The MLP input row (BF16, kIn =
// 48 columns) is the 32 encoded features, the shading normal, the direction
// towards the viewer, the roughness and the albedo (42 used); 5 hidden
// layers of 64 with ReLU; 3 outputs (of kOutN = 16). Every dense layer is
// the BF16 MFMA GEMM of nn_common.hpp with FP32 accumulation; the loss is
// the relative L2 (y - t)^2 / (y^2 + 0.01) with y held constant in the
// denominator; Adam on FP32 master copies of the MLP and the tables.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "nn_common.hpp"

namespace mpt {
namespace nrc {

using nn::bf16;
using nn::from_bf16;
using nn::to_bf16;

constexpr int kLevels = 16;
constexpr int kF = 2;
constexpr int kEnc = kLevels * kF;  // 32
constexpr int kOther = 10;          // normal 3, direction 3, roughness 1, albedo 3
constexpr int kIn = 48;
constexpr int kHidden = 64;
constexpr int kLayers = 6;          // 48 -> 64, 4 x (64 -> 64), 64 -> 16
constexpr int kOutN = 16;
constexpr int kLogT = 19;
constexpr uint32_t kT = 1u << kLogT;
constexpr double kNmin = 16, kNmax = 2048;

struct level_info {
    uint32_t res;    // grid cells per axis
    uint32_t dense;  // 1: (res + 1)^3 <= T, index the corners directly
};

__host__ __device__ inline uint32_t corner_index(const level_info& L, uint32_t x, uint32_t y, uint32_t z) {
    if (L.dense) return x + (L.res + 1) * (y + (L.res + 1) * z);
    return (x * 1u ^ y * 2654435761u ^ z * 805459861u) & (kT - 1);  // spatial hash (Teschner et al. 2003)
}

struct levels_t {
    level_info l[kLevels];
};

// Encoded features of one sample at one level: the 8 corners and their
// trilinear weights.
__device__ __forceinline__ void corners(const levels_t& LV, int lv, const float* pos, uint32_t idx[8], float w[8]) {
    const level_info L = LV.l[lv];
    float f[3];
    uint32_t c[3];
    for (int a = 0; a < 3; a++) {
        float x = pos[a] * float(L.res);
        x = x < 0.f ? 0.f : (x > float(L.res) ? float(L.res) : x);
        float fl = std::floor(x);
        if (fl >= float(L.res)) fl = float(L.res) - 1.f;
        c[a] = uint32_t(fl);
        f[a] = x - fl;
    }
    for (int k = 0; k < 8; k++) {
        const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        idx[k] = corner_index(L, c[0] + dx, c[1] + dy, c[2] + dz);
        w[k] = (dx ? f[0] : 1.f - f[0]) * (dy ? f[1] : 1.f - f[1]) * (dz ? f[2] : 1.f - f[2]);
    }
}

// X[s] = [encoding (32), other (10), zeros] in BF16; one thread per sample.
__global__ void k_encode(int64_t n, levels_t LV, const float* __restrict__ pos, const float* __restrict__ other,
                         const _Float16* __restrict__ table, bf16* __restrict__ X) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float p[3] = {pos[3 * s], pos[3 * s + 1], pos[3 * s + 2]};
    bf16* row = X + s * kIn;
    for (int lv = 0; lv < kLevels; lv++) {
        uint32_t idx[8];
        float w[8];
        corners(LV, lv, p, idx, w);
        float e0 = 0.f, e1 = 0.f;
        const _Float16* t = table + int64_t(lv) * kT * kF;
        for (int k = 0; k < 8; k++) {
            e0 += w[k] * float(t[int64_t(idx[k]) * kF]);
            e1 += w[k] * float(t[int64_t(idx[k]) * kF + 1]);
        }
        row[kF * lv] = to_bf16(e0);
        row[kF * lv + 1] = to_bf16(e1);
    }
    for (int c = 0; c < kOther; c++) row[kEnc + c] = to_bf16(other[s * kOther + c]);
    for (int c = kEnc + kOther; c < kIn; c++) row[c] = 0;
}

// Table gradients: dtable[lv][idx][f] += w * dX[s][2 lv + f].
__global__ void k_encode_backward(int64_t n, levels_t LV, const float* __restrict__ pos, const float* __restrict__ dX,
                                  float* __restrict__ gtable) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // one (sample, level)
    if (e >= n * kLevels) return;
    const int64_t s = e / kLevels;
    const int lv = int(e % kLevels);
    const float p[3] = {pos[3 * s], pos[3 * s + 1], pos[3 * s + 2]};
    uint32_t idx[8];
    float w[8];
    corners(LV, lv, p, idx, w);
    const float g0 = dX[s * kIn + kF * lv], g1 = dX[s * kIn + kF * lv + 1];
    float* t = gtable + int64_t(lv) * kT * kF;
    for (int k = 0; k < 8; k++) {
        if (g0 != 0.f) atomicAdd(t + int64_t(idx[k]) * kF, w[k] * g0);
        if (g1 != 0.f) atomicAdd(t + int64_t(idx[k]) * kF + 1, w[k] * g1);
    }
}

// Relative L2 loss and its gradient with respect to the outputs.
__global__ void k_loss(int64_t n, const float* __restrict__ y, const float* __restrict__ target, float* __restrict__ dy,
                       double* __restrict__ loss) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double l = 0;
    if (s < n) {
        for (int c = 0; c < kOutN; c++) {
            float g = 0.f;
            if (c < 3) {
                const float yc = y[s * kOutN + c], t = target[3 * s + c];
                const float den = yc * yc + 1e-2f;
                const float d = yc - t;
                l += double(d * d / den);
                g = 2.f * d / den / float(3 * n);
            }
            dy[s * kOutN + c] = g;
        }
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0 && l != 0) atomicAdd(loss, l);
}

__global__ void k_pred(int64_t n, const float* __restrict__ y, float* __restrict__ pred) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    for (int c = 0; c < 3; c++) pred[3 * s + c] = y[s * kOutN + c];
}

__global__ void k_table_adam(int64_t n, float* __restrict__ w, _Float16* __restrict__ wh, const float* __restrict__ g,
                             float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
                             float c1, float c2) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    if (gi == 0.f && m[e] == 0.f) return;  // untouched entry with no momentum: nothing moves
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    const float wi = w[e] - lr * (mi / c1) / (std::sqrt(vi / c2) + eps);
    w[e] = wi;
    wh[e] = _Float16(wi);
}

__global__ void k_init_table(int64_t n, uint64_t seed, float* __restrict__ w, _Float16* __restrict__ wh) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    uint64_t z = seed + uint64_t(e) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float u = float(double(z >> 11) * 0x1.0p-53);
    const float v = (2.f * u - 1.f) * 1e-4f;  // instant-ngp's initialisation range
    w[e] = v;
    wh[e] = _Float16(v);
}

}  // namespace nrc
}  // namespace mpt

using namespace mpt;
using namespace mpt::nrc;
using mpt::nn::blocks;

struct mpt_nrc {
    levels_t LV{};
    int64_t nparams = 0;            // MLP parameters
    int cin[kLayers], cout[kLayers];
    int64_t woff[kLayers], boff[kLayers];
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16* wb = nullptr;
    int64_t ntable = int64_t(kLevels) * kT * kF;
    float *tw = nullptr, *tg = nullptr, *tm = nullptr, *tv = nullptr;
    _Float16* th = nullptr;
    int64_t step = 0;
    // workspace
    int64_t cap = 0;
    bf16* act[kLayers + 1] = {};
    float* y = nullptr;
    float* dA = nullptr;
    float* dB = nullptr;
    bf16* G = nullptr;
    double* loss = nullptr;
    ~mpt_nrc() {
        for (void* p : {(void*)w, (void*)g, (void*)m, (void*)v, (void*)wb, (void*)tw, (void*)tg, (void*)tm, (void*)tv,
                        (void*)th})
            (void)hipFree(p);
        release();
    }
    void release() {
        for (auto& a : act) {
            (void)hipFree(a);
            a = nullptr;
        }
        for (void* p : {(void*)y, (void*)dA, (void*)dB, (void*)G, (void*)loss}) (void)hipFree(p);
        y = nullptr, dA = nullptr, dB = nullptr, G = nullptr, loss = nullptr;
        cap = 0;
    }
};

namespace {

int ensure(mpt_nrc* r, int64_t n) {
    if (r->cap >= n) return MPT_OK;
    r->release();
    for (int l = 0; l <= kLayers; l++) {
        const int c = l == 0 ? kIn : (l == kLayers ? kOutN : kHidden);
        NN_TRY(hipMalloc(&r->act[l], size_t(n) * c * sizeof(bf16)));
    }
    NN_TRY(hipMalloc(&r->y, size_t(n) * kOutN * sizeof(float)));
    NN_TRY(hipMalloc(&r->dA, size_t(n) * kHidden * sizeof(float)));
    NN_TRY(hipMalloc(&r->dB, size_t(n) * kHidden * sizeof(float)));
    NN_TRY(hipMalloc(&r->G, size_t(n) * kHidden * sizeof(bf16)));
    NN_TRY(hipMalloc(&r->loss, sizeof(double)));
    r->cap = n;
    return MPT_OK;
}

int forward(mpt_nrc* r, int64_t n, const float* pos, const float* other, hipStream_t s) {
    hipLaunchKernelGGL(k_encode, dim3(blocks(n)), dim3(256), 0, s, n, r->LV, pos, other, r->th, r->act[0]);
    for (int l = 0; l < kLayers; l++) {
        const bool last = l == kLayers - 1;
        int rc = nn::gemm(false, false, last ? 3 : 2, int(n), r->cout[l], r->cin[l], r->act[l], r->cin[l],
                          r->wb + r->woff[l], r->cin[l], last ? (void*)r->y : (void*)r->act[l + 1], r->cout[l],
                          r->w + r->boff[l], false, s);
        if (rc) return rc;
    }
    return MPT_OK;
}

// Backward from dy (FP32 [n][kOutN] in r->dA): MLP gradients into r->g,
// table gradients into r->tg (both zeroed here first).
int backward(mpt_nrc* r, int64_t n, const float* pos, hipStream_t s) {
    NN_TRY(hipMemsetAsync(r->tg, 0, r->ntable * sizeof(float), s));
    float* dY = r->dA;
    float* dX = r->dB;
    for (int l = kLayers - 1; l >= 0; l--) {
        const int ci = r->cin[l], co = r->cout[l];
        const bool last = l == kLayers - 1;
        NN_TRY(hipMemsetAsync(r->g + r->boff[l], 0, size_t(co) * sizeof(float), s));
        const int bs = (256 / co) * co;
        const unsigned gb = unsigned(std::max<int64_t>(1, std::min<int64_t>(2048, (n * co + bs - 1) / bs)));
        hipLaunchKernelGGL(nn::k_relu_grad, dim3(gb), dim3(bs), 0, s, n * co, co, dY, r->act[l + 1], last ? 0 : 1,
                           r->G, r->g + r->boff[l]);
        NN_TRY(hipMemsetAsync(r->g + r->woff[l], 0, size_t(co) * ci * sizeof(float), s));
        int rc = nn::gemm(true, true, 4, co, ci, int(n), r->G, co, r->act[l], ci, r->g + r->woff[l], ci, nullptr,
                          false, s);
        if (rc) return rc;
        // dX[s][k] = sum_co G[s][co] W[co][k] (for layer 0 only the encoding columns matter)
        rc = nn::gemm(false, true, 0, int(n), ci, co, r->G, co, r->wb + r->woff[l], ci, dX, ci, nullptr, false, s);
        if (rc) return rc;
        std::swap(dY, dX);
    }
    hipLaunchKernelGGL(k_encode_backward, dim3(blocks(n * kLevels)), dim3(256), 0, s, n, r->LV, pos, dY, r->tg);
    return MPT_OK;
}

}  // namespace

namespace mpt {

// Used by the renderer (kernels.hip): predictions (FP32, 3 per sample) for
// n samples given normalised positions and the other inputs.
int nrc_infer(mpt_nrc* r, const float* pos, const float* other, int64_t n, float* pred, hipStream_t s) {
    if (n <= 0) return MPT_OK;
    int rc = ensure(r, n);
    if (rc) return rc;
    if ((rc = forward(r, n, pos, other, s))) return rc;
    hipLaunchKernelGGL(k_pred, dim3(blocks(n)), dim3(256), 0, s, n, r->y, pred);
    NN_TRY(hipGetLastError());
    return MPT_OK;
}

// One training step on n samples: loss, gradients, and with lr > 0 Adam.
int nrc_train(mpt_nrc* r, const float* pos, const float* other, const float* target, int64_t n, float lr,
              hipStream_t s, double* loss) {
    if (n <= 0) {
        if (loss) *loss = 0;
        return MPT_OK;
    }
    int rc = ensure(r, n);
    if (rc) return rc;
    if ((rc = forward(r, n, pos, other, s))) return rc;
    NN_TRY(hipMemsetAsync(r->loss, 0, sizeof(double), s));
    hipLaunchKernelGGL(k_loss, dim3(blocks(n)), dim3(256), 0, s, n, r->y, target, r->dA, r->loss);
    if ((rc = backward(r, n, pos, s))) return rc;
    if (lr > 0) {
        r->step++;
        const float b1 = 0.9f, b2 = 0.99f;
        const float c1 = 1.f - std::pow(b1, float(r->step)), c2 = 1.f - std::pow(b2, float(r->step));
        hipLaunchKernelGGL(nn::k_adam, dim3(blocks(r->nparams)), dim3(256), 0, s, r->nparams, r->w, r->wb, r->g, r->m,
                           r->v, lr, b1, b2, 1e-8f, c1, c2);
        hipLaunchKernelGGL(k_table_adam, dim3(blocks(r->ntable)), dim3(256), 0, s, r->ntable, r->tw, r->th, r->tg,
                           r->tm, r->tv, lr, b1, b2, 1e-15f, c1, c2);
    }
    NN_TRY(hipGetLastError());
    if (loss) {
        double l = 0;
        NN_TRY(hipMemcpyAsync(&l, r->loss, sizeof(double), hipMemcpyDeviceToHost, s));
        NN_TRY(hipStreamSynchronize(s));
        *loss = l / double(3 * n);
    }
    return MPT_OK;
}

}  // namespace mpt

extern "C" {

int32_t mpt_nrc_create(uint64_t seed, mpt_nrc** out) {
    if (!out) return set_error(MPT_E_INVALID, "nrc: null output");
    auto r = std::make_unique<mpt_nrc>();
    const double b = std::exp((std::log(kNmax) - std::log(kNmin)) / (kLevels - 1));
    for (int lv = 0; lv < kLevels; lv++) {
        const uint32_t res = uint32_t(std::floor(kNmin * std::pow(b, lv)));
        r->LV.l[lv].res = res;
        r->LV.l[lv].dense = uint64_t(res + 1) * (res + 1) * (res + 1) <= kT ? 1u : 0u;
    }
    int64_t off = 0;
    for (int l = 0; l < kLayers; l++) {
        r->cin[l] = l == 0 ? kIn : kHidden;
        r->cout[l] = l == kLayers - 1 ? kOutN : kHidden;
        r->woff[l] = off;
        off += int64_t(r->cin[l]) * r->cout[l];
        r->boff[l] = off;
        off += r->cout[l];
    }
    r->nparams = off;
    std::vector<float> h(off, 0.f);
    uint64_t st = seed ^ 0x5DEECE66Dull;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * 0x1.0p-53;
    };
    for (int l = 0; l < kLayers; l++) {
        const int fan_in = l == 0 ? kEnc + kOther : kHidden;
        const double lim = std::sqrt(6.0 / fan_in);
        for (int co = 0; co < r->cout[l]; co++)
            for (int ci = 0; ci < r->cin[l]; ci++) {
                const bool used = (l > 0 || ci < kEnc + kOther) && (l < kLayers - 1 || co < 3);
                const double u = next();
                h[r->woff[l] + int64_t(co) * r->cin[l] + ci] = used ? float((2 * u - 1) * lim) : 0.f;
            }
    }
    NN_TRY(hipMalloc(&r->w, off * sizeof(float)));
    NN_TRY(hipMalloc(&r->wb, off * sizeof(bf16)));
    NN_TRY(hipMalloc(&r->g, off * sizeof(float)));
    NN_TRY(hipMalloc(&r->m, off * sizeof(float)));
    NN_TRY(hipMalloc(&r->v, off * sizeof(float)));
    NN_TRY(hipMemcpy(r->w, h.data(), off * sizeof(float), hipMemcpyHostToDevice));
    NN_TRY(hipMemset(r->g, 0, off * sizeof(float)));
    NN_TRY(hipMemset(r->m, 0, off * sizeof(float)));
    NN_TRY(hipMemset(r->v, 0, off * sizeof(float)));
    hipLaunchKernelGGL(nn::k_to_bf16, dim3(blocks(off)), dim3(256), 0, nullptr, off, r->w, r->wb);
    const int64_t nt = r->ntable;
    NN_TRY(hipMalloc(&r->tw, nt * sizeof(float)));
    NN_TRY(hipMalloc(&r->tg, nt * sizeof(float)));
    NN_TRY(hipMalloc(&r->tm, nt * sizeof(float)));
    NN_TRY(hipMalloc(&r->tv, nt * sizeof(float)));
    NN_TRY(hipMalloc(&r->th, nt * sizeof(_Float16)));
    NN_TRY(hipMemset(r->tg, 0, nt * sizeof(float)));
    NN_TRY(hipMemset(r->tm, 0, nt * sizeof(float)));
    NN_TRY(hipMemset(r->tv, 0, nt * sizeof(float)));
    hipLaunchKernelGGL(k_init_table, dim3(blocks(nt)), dim3(256), 0, nullptr, nt, seed, r->tw, r->th);
    NN_TRY(hipDeviceSynchronize());
    *out = r.release();
    return MPT_OK;
}

void mpt_nrc_destroy(mpt_nrc* r) { delete r; }

int64_t mpt_nrc_param_count(const mpt_nrc* r, int64_t* table_count) {
    if (table_count) *table_count = r ? r->ntable : 0;
    return r ? r->nparams : 0;
}

int32_t mpt_nrc_levels(const mpt_nrc* r, uint32_t* res, uint32_t* dense) {
    if (!r) return set_error(MPT_E_INVALID, "nrc: null cache");
    for (int lv = 0; lv < kLevels; lv++) {
        if (res) res[lv] = r->LV.l[lv].res;
        if (dense) dense[lv] = r->LV.l[lv].dense;
    }
    return MPT_OK;
}

int32_t mpt_nrc_get_params(mpt_nrc* r, float* mlp, float* mlp_grad, float* table, float* table_grad) {
    if (!r) return set_error(MPT_E_INVALID, "nrc: null cache");
    NN_TRY(hipDeviceSynchronize());
    if (mlp) NN_TRY(hipMemcpy(mlp, r->w, r->nparams * sizeof(float), hipMemcpyDeviceToHost));
    if (mlp_grad) NN_TRY(hipMemcpy(mlp_grad, r->g, r->nparams * sizeof(float), hipMemcpyDeviceToHost));
    if (table) NN_TRY(hipMemcpy(table, r->tw, r->ntable * sizeof(float), hipMemcpyDeviceToHost));
    if (table_grad) NN_TRY(hipMemcpy(table_grad, r->tg, r->ntable * sizeof(float), hipMemcpyDeviceToHost));
    return MPT_OK;
}

int32_t mpt_nrc_set_params(mpt_nrc* r, const float* mlp, const float* table) {
    if (!r) return set_error(MPT_E_INVALID, "nrc: null cache");
    if (mlp) {
        NN_TRY(hipMemcpy(r->w, mlp, r->nparams * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(nn::k_to_bf16, dim3(blocks(r->nparams)), dim3(256), 0, nullptr, r->nparams, r->w, r->wb);
    }
    if (table) {
        std::vector<_Float16> h(r->ntable);
        for (int64_t i = 0; i < r->ntable; i++) h[i] = _Float16(table[i]);
        NN_TRY(hipMemcpy(r->tw, table, r->ntable * sizeof(float), hipMemcpyHostToDevice));
        NN_TRY(hipMemcpy(r->th, h.data(), r->ntable * sizeof(_Float16), hipMemcpyHostToDevice));
    }
    NN_TRY(hipDeviceSynchronize());
    return MPT_OK;
}

int32_t mpt_nrc_infer_dev(mpt_nrc* r, const float* pos, const float* other, int64_t n, float* pred) {
    if (!r || n < 0 || (n > 0 && (!pos || !other || !pred))) return set_error(MPT_E_INVALID, "nrc: bad argument");
    int rc = nrc_infer(r, pos, other, n, pred, nullptr);
    if (rc) return rc;
    NN_TRY(hipDeviceSynchronize());
    return MPT_OK;
}

int32_t mpt_nrc_train_dev(mpt_nrc* r, const float* pos, const float* other, const float* target, int64_t n, float lr,
                          double* loss) {
    if (!r || n < 0 || (n > 0 && (!pos || !other || !target))) return set_error(MPT_E_INVALID, "nrc: bad argument");
    int rc = nrc_train(r, pos, other, target, n, lr, nullptr, loss);
    if (rc) return rc;
    NN_TRY(hipDeviceSynchronize());
    return MPT_OK;
}

}  // extern "C"
