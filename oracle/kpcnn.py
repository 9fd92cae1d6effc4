"""CPU restatement of the learned kernel-predicting denoiser (numpy).

TEST INFRASTRUCTURE: the checker for mypbr_amd/csrc/learn.hip. Only tests/
import it. The network has no counterpart in the reference renderer; this
module states what the GPU kernels compute, and is itself pinned by
finite differences of its own FP64 forward pass (tests/test_kpcnn.py).

With bf16=True it rounds where the GPU rounds (features, the convolution
inputs and weights, hidden activations after ReLU, the layer gradients fed
to the GEMMs), accumulating in float64 where the GPU accumulates in FP32;
with bf16=False everything is float64.
"""
import numpy as np

K_FEAT = 16
K_TAPS = 25
K_OUT = 32
ALB_EPS = 1e-3


def bf16(x):
    """Round to bfloat16 (nearest even) and back, elementwise."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    nan &= (u & 0x7FFFFF) != 0
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(rad, alb, nrm, var, depth):
    """Per-pixel inputs (learn.hip k_features): (npix, 16) feature rows, the
    noisy lighting and albedo + eps (npix, 3), all from float32 arithmetic."""
    f32 = np.float32
    rad = rad.reshape(-1, 3).astype(f32)
    alb = alb.reshape(-1, 3).astype(f32)
    a = alb + f32(ALB_EPS)
    light = rad / a
    n = rad.shape[0]
    f = np.zeros((n, K_FEAT), dtype=f32)
    f[:, 0:3] = np.log1p(np.maximum(light, f32(0)))
    f[:, 3:6] = alb
    f[:, 6:9] = nrm.reshape(-1, 3).astype(f32)
    lum = f32(0.2126) * rad[:, 0] + f32(0.7152) * rad[:, 1] + f32(0.0722) * rad[:, 2]
    f[:, 9] = np.log1p(np.maximum(depth.reshape(-1).astype(f32), f32(0)))
    f[:, 10] = np.log1p(np.maximum(var.reshape(-1).astype(f32), f32(0)) / (lum * lum + f32(1e-2)))
    return f.astype(np.float64), light.astype(np.float64), a.astype(np.float64)


def im2col(x, B, H, W):
    """x (B*H*W, C) -> (B*H*W, 9*C), tap t = 3 * (dy + 1) + dx + 1, zero outside."""
    C = x.shape[1]
    xi = x.reshape(B, H, W, C)
    pad = np.zeros((B, H + 2, W + 2, C))
    pad[:, 1:H + 1, 1:W + 1] = xi
    cols = [pad[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    return np.concatenate(cols, axis=-1).reshape(B * H * W, 9 * C)


def col2im(dcol, B, H, W, C):
    """Adjoint of im2col: sums each tap's column gradient back onto the pixel it read."""
    d = dcol.reshape(B, H, W, 9, C)
    out = np.zeros((B, H + 2, W + 2, C))
    t = 0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            out[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W] += d[:, :, :, t]
            t += 1
    return out[:, 1:H + 1, 1:W + 1].reshape(B * H * W, C)


class Net:
    """The network's parameters in the GPU's layout: per layer W [cout][9][cin] then b [cout]."""

    def __init__(self, params, cin, cout, woff, boff):
        self.p = np.asarray(params, dtype=np.float64)
        self.cin, self.cout, self.woff, self.boff = list(cin), list(cout), list(woff), list(boff)

    def layer(self, l):
        ci, co = self.cin[l], self.cout[l]
        w = self.p[self.woff[l]:self.woff[l] + co * 9 * ci].reshape(co, 9 * ci)
        b = self.p[self.boff[l]:self.boff[l] + co]
        return w, b

    @property
    def layers(self):
        return len(self.cin)


def apply_kernel(logits, light, albe, B, H, W):
    """Softmax over the 25 logits, weighted 5x5 sum of lighting with clamped
    coordinates, times albedo + eps. Returns radiance, weights, neighbour indices."""
    lg = logits[:, :K_TAPS]
    w = np.exp(lg - lg.max(axis=1, keepdims=True))
    w /= w.sum(axis=1, keepdims=True)
    npix = B * H * W
    p = np.arange(npix)
    x = p % W
    y = (p // W) % H
    base = p - (y * W + x)
    q = np.zeros((npix, K_TAPS), dtype=np.int64)
    for k in range(K_TAPS):
        yy = np.clip(y + k // 5 - 2, 0, H - 1)
        xx = np.clip(x + k % 5 - 2, 0, W - 1)
        q[:, k] = base + yy * W + xx
    o = np.einsum("pk,pkc->pc", w, light[q])
    return o * albe, w, q


def loss_and_grad(logits, light, albe, target, B, H, W, mode=0):
    """Mode 0: relMSE (o - t)^2 / (t^2 + 0.01); mode 1: |log(1 + max(o, 0)) -
    log(1 + max(t, 0))|; averaged over pixels and channels, and the gradient
    with respect to the logits (npix, 32)."""
    rad, w, q = apply_kernel(logits, light, albe, B, H, W)
    npix = rad.shape[0]
    if mode == 1:
        ro = np.maximum(rad, 0)
        d = np.log1p(ro) - np.log1p(np.maximum(target, 0))
        loss = float(np.abs(d).sum() / (3 * npix))
        dl = np.where(rad > 0, np.sign(d) / (1 + ro), 0.0) / (3 * npix) * albe
    else:
        den = target * target + 1e-2
        d = rad - target
        loss = float((d * d / den).sum() / (3 * npix))
        dl = 2 * d / den / (3 * npix) * albe          # d loss / d o
    gw = np.einsum("pc,pkc->pk", dl, light[q])        # d loss / d w_k
    dlog = np.zeros((npix, K_OUT))
    dlog[:, :K_TAPS] = w * (gw - (w * gw).sum(axis=1, keepdims=True))
    return loss, dlog, rad


def forward(net, feat, B, H, W, round_bf16=True):
    """Activations of every layer (the first is the features) and the logits."""
    r = bf16 if round_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    acts = [r(feat)]
    for l in range(net.layers):
        w, b = net.layer(l)
        y = im2col(acts[-1], B, H, W) @ r(w).T + b
        if l == net.layers - 1:
            return acts, y
        acts.append(r(np.maximum(y, 0.0)))


def gradients(net, frame, target, B, H, W, round_bf16=True, mode=0):
    """Loss and the gradient of every parameter, in the parameter layout."""
    r = bf16 if round_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    feat, light, albe = frame
    acts, logits = forward(net, feat, B, H, W, round_bf16)
    loss, dy, _ = loss_and_grad(logits, light, albe, target, B, H, W, mode)
    grad = np.zeros_like(net.p)
    for l in range(net.layers - 1, -1, -1):
        w, _ = net.layer(l)
        ci, co = net.cin[l], net.cout[l]
        g = dy if l == net.layers - 1 else dy * (acts[l + 1] > 0)
        g = r(g)
        grad[net.boff[l]:net.boff[l] + co] = g.sum(axis=0)
        xcol = im2col(acts[l], B, H, W)
        grad[net.woff[l]:net.woff[l] + co * 9 * ci] = (g.T @ xcol).reshape(-1)
        if l:
            dy = col2im(g @ r(w), B, H, W, ci)
    return loss, grad
