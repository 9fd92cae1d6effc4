"""CPU restatement of the neural radiance cache (numpy).

TEST INFRASTRUCTURE: the checker for mypbr_amd/csrc/nrc.hip. Only tests/
import it. The cache has no counterpart in the reference renderer; this
module states what the GPU computes and is pinned by finite differences of
its own float64 pass (tests/test_nrc.py).

Encoding: 16 levels x 2 features, tables of 2^19 entries per level, dense
indexing where (res + 1)^3 fits, else the spatial hash
x * 1 ^ y * 2654435761 ^ z * 805459861 (mod 2^32, masked to 2^19);
trilinear weights in float32. MLP: 48 -> 64 (x5, ReLU) -> 16 outputs (3
used). Loss: relative L2 (y - t)^2 / (y^2 + 0.01), y constant in the
denominator, averaged over samples and the 3 channels.
"""
import numpy as np

from kpcnn import bf16

LEVELS = 16
F = 2
LOG_T = 19
T = 1 << LOG_T
ENC = LEVELS * F
OTHER = 10
IN = 48
HIDDEN = 64
OUT = 16
LAYERS = 6


def layout():
    cin = [IN] + [HIDDEN] * (LAYERS - 1)
    cout = [HIDDEN] * (LAYERS - 1) + [OUT]
    woff, boff, off = [], [], 0
    for ci, co in zip(cin, cout):
        woff.append(off)
        off += ci * co
        boff.append(off)
        off += co
    return cin, cout, woff, boff, off


def level_table(nmin=16.0, nmax=2048.0):
    b = np.exp((np.log(nmax) - np.log(nmin)) / (LEVELS - 1))
    res = [int(np.floor(nmin * b ** lv)) for lv in range(LEVELS)]
    dense = [int((r + 1) ** 3 <= T) for r in res]
    return res, dense


def corners(pos, res, dense):
    """Corner indices (n, 8) and float32 weights (n, 8) at one level."""
    f32 = np.float32
    x = pos.astype(f32) * f32(res)
    x = np.clip(x, f32(0), f32(res))
    fl = np.floor(x)
    fl = np.where(fl >= f32(res), f32(res - 1), fl)
    c = fl.astype(np.uint32)
    fr = (x - fl).astype(f32)
    idx = np.zeros((pos.shape[0], 8), dtype=np.uint32)
    w = np.zeros((pos.shape[0], 8), dtype=f32)
    for k in range(8):
        d = [(k >> a) & 1 for a in range(3)]
        cx, cy, cz = (c[:, a] + np.uint32(d[a]) for a in range(3))
        if dense:
            r1 = np.uint32(res + 1)
            idx[:, k] = cx + r1 * (cy + r1 * cz)
        else:
            idx[:, k] = (cx * np.uint32(1) ^ cy * np.uint32(2654435761) ^ cz * np.uint32(805459861)) & np.uint32(T - 1)
        wx = fr[:, 0] if d[0] else f32(1) - fr[:, 0]
        wy = fr[:, 1] if d[1] else f32(1) - fr[:, 1]
        wz = fr[:, 2] if d[2] else f32(1) - fr[:, 2]
        w[:, k] = (wx * wy) * wz
    return idx, w


def encode(pos, other, table, levels, round_bf16=True):
    """MLP input rows (n, 48): the encoding from the FP16 table, then other."""
    res, dense = levels
    n = pos.shape[0]
    r = bf16 if round_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    tab = table.reshape(LEVELS, T, F)
    if round_bf16:
        tab = tab.astype(np.float16).astype(np.float32)
    x = np.zeros((n, IN))
    for lv in range(LEVELS):
        idx, w = corners(pos, res[lv], dense[lv])
        t = tab[lv]
        for f in range(F):
            acc = np.zeros(n, dtype=np.float32 if round_bf16 else np.float64)
            for k in range(8):
                acc = acc + (w[:, k] * t[idx[:, k], f]).astype(acc.dtype)
            x[:, F * lv + f] = acc
    x[:, ENC:ENC + OTHER] = other
    return r(x)


def forward(params, x, round_bf16=True):
    cin, cout, woff, boff, _ = layout()
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    acts = [x]
    for l in range(LAYERS):
        w = params[woff[l]:woff[l] + cout[l] * cin[l]].reshape(cout[l], cin[l])
        b = params[boff[l]:boff[l] + cout[l]]
        y = acts[-1] @ r(w).T + b
        if l == LAYERS - 1:
            return acts, y
        acts.append(r(np.maximum(y, 0.0)))


def loss_and_grad(y, target, den=None):
    """The loss and its gradient with the denominator held constant; `den`
    fixes the denominator (for finite differences of that gradient)."""
    n = y.shape[0]
    yc = y[:, :3]
    if den is None:
        den = yc * yc + 1e-2
    d = yc - target
    loss = float((d * d / den).sum() / (3 * n))
    dy = np.zeros_like(y)
    dy[:, :3] = 2 * d / den / (3 * n)
    return loss, dy


def gradients(params, table, pos, other, target, levels, round_bf16=True, den=None):
    """Loss, MLP parameter gradients and table gradients (flat, like the GPU's),
    and the outputs. `den`: a fixed loss denominator (see loss_and_grad)."""
    cin, cout, woff, boff, _ = layout()
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    x = encode(pos, other, table, levels, round_bf16)
    acts, y = forward(params, x, round_bf16)
    loss, dy = loss_and_grad(y, target, den)
    g = np.zeros_like(params, dtype=np.float64)
    for l in range(LAYERS - 1, -1, -1):
        w = params[woff[l]:woff[l] + cout[l] * cin[l]].reshape(cout[l], cin[l])
        gl = dy if l == LAYERS - 1 else dy * (acts[l + 1] > 0)
        gl = r(gl)
        g[boff[l]:boff[l] + cout[l]] = gl.sum(axis=0)
        g[woff[l]:woff[l] + cout[l] * cin[l]] = (gl.T @ acts[l]).reshape(-1)
        dy = gl @ r(w)
    gt = np.zeros((LEVELS, T, F))
    res, dense = levels
    for lv in range(LEVELS):
        idx, w = corners(pos, res[lv], dense[lv])
        for f in range(F):
            for k in range(8):
                np.add.at(gt[lv, :, f], idx[:, k], w[:, k].astype(np.float64) * dy[:, F * lv + f])
    return loss, g, gt.reshape(-1), y
