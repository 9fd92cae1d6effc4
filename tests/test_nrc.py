"""The neural radiance cache (include/mpt.h, mpt_nrc_*, mpt_render_nrc).

CPU: the numpy restatement (oracle/nrc.py) is pinned by finite differences
of its float64 pass, for MLP and hash-table parameters, and by the
encoding's closed forms (trilinear reproduction of a linear field, the
dense index). GPU: predictions, loss and gradients against the oracle with
the GPU's roundings; in a render, a cache that only trains leaves the image
bit-identical to the plain tracer, and online training brings the cached
render's error below the plain tracer's at equal samples per pixel (the
equal-time comparison at 1080p is measured by bench.py).
"""
import os
import sys
import time

import numpy as np
import pytest

from conftest import ROOT, scene_path

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import nrc as nr  # noqa: E402


def samples(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.standard_normal((n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    other = np.concatenate([nrm, wo, rng.uniform(0, 1, (n, 1)), rng.uniform(0, 1, (n, 3))], axis=1).astype(np.float32)
    target = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    return pos, other, target


def random_params(seed, scale=0.2):
    _, _, _, _, n = nr.layout()
    return np.random.default_rng(seed).uniform(-scale, scale, n)


# ---------------------------------------------------------------- CPU

def test_levels_and_dense_indexing():
    res, dense = nr.level_table()
    assert res[0] == 16 and res[-1] in (2047, 2048) and all(a < b for a, b in zip(res, res[1:]))
    assert dense[:3] == [1, 1, 1] and dense[-1] == 0
    # a dense level's corner indices are the grid coordinates
    pos = np.array([[0.5, 0.25, 0.75]], dtype=np.float32)
    idx, w = nr.corners(pos, 16, 1)
    assert idx[0, 0] == 8 + 17 * (4 + 17 * 12)
    assert abs(w.sum() - 1) < 1e-6


def test_encoding_reproduces_a_linear_field():
    # table entries at a dense level set to a linear function of the corner
    # position: trilinear interpolation returns that function exactly
    res, dense = nr.level_table()
    table = np.zeros((nr.LEVELS, nr.T, nr.F))
    r = res[0]
    g = np.arange(r + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lin = (X + 2 * Y - 3 * Z) / r
    table[0, (X + (r + 1) * (Y + (r + 1) * Z)).ravel(), 0] = lin.ravel()
    pos = np.random.default_rng(1).uniform(0, 1, (50, 3)).astype(np.float32)
    x = nr.encode(pos, np.zeros((50, nr.OTHER)), table.reshape(-1), (res, dense), round_bf16=False)
    want = pos[:, 0] + 2 * pos[:, 1] - 3 * pos[:, 2]
    assert np.allclose(x[:, 0], want, atol=1e-5)


def test_oracle_gradients_equal_finite_differences():
    levels = nr.level_table()
    pos, other, target = samples(40, 2)
    params = random_params(3)
    table = np.random.default_rng(4).uniform(-0.5, 0.5, nr.LEVELS * nr.T * nr.F)
    loss, g, gt, y = nr.gradients(params, table, pos, other, target, levels, round_bf16=False)
    # the loss's denominator is held constant in its gradient: hold it for the differences too
    den = y[:, :3] ** 2 + 1e-2

    def L(p, t):
        return nr.gradients(p, t, pos, other, target, levels, False, den)[0]
    rng = np.random.default_rng(5)
    cin, cout, woff, boff, n = nr.layout()
    picks = list(rng.integers(0, n, 40))
    worst = 0.0
    for i in picks:
        h = 1e-6
        p2 = params.copy()
        p2[i] += h
        lp = L(p2, table)
        p2[i] -= 2 * h
        lm = L(p2, table)
        fd = (lp - lm) / (2 * h)
        if abs(fd) > 1e-9 or abs(g[i]) > 1e-9:
            worst = max(worst, abs(fd - g[i]) / max(abs(fd), abs(g[i])))
    nz = np.flatnonzero(np.abs(gt) > 1e-5)  # entries whose difference quotient is well above rounding
    for i in rng.choice(nz, 20, replace=False):
        h = 1e-5
        t2 = table.copy()
        t2[i] += h
        lp = L(params, t2)
        t2[i] -= 2 * h
        lm = L(params, t2)
        fd = (lp - lm) / (2 * h)
        worst = max(worst, abs(fd - gt[i]) / max(abs(fd), abs(gt[i])))
    assert worst < 1e-4, worst


# ---------------------------------------------------------------- GPU

class _Cache:
    def __init__(self, seed):
        import ctypes as C
        import mypbr_amd as m
        self.m, self.C = m, C
        self.h = C.c_void_p()
        m._check(m.lib.mpt_nrc_create(seed, C.byref(self.h)))
        nt = C.c_int64()
        self.n = m.lib.mpt_nrc_param_count(self.h, C.byref(nt))
        self.nt = nt.value
        res = (C.c_uint32 * 16)()
        dense = (C.c_uint32 * 16)()
        m._check(m.lib.mpt_nrc_levels(self.h, res, dense))
        self.levels = (list(res), list(dense))

    def fp(self, a):
        return a.ctypes.data_as(self.C.POINTER(self.C.c_float)) if a is not None else None

    def set(self, mlp=None, table=None):
        mlp = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        table = None if table is None else np.ascontiguousarray(table, np.float32)
        self.m._check(self.m.lib.mpt_nrc_set_params(self.h, self.fp(mlp), self.fp(table)))

    def get(self):
        out = [np.zeros(self.n, np.float32), np.zeros(self.n, np.float32), np.zeros(self.nt, np.float32),
               np.zeros(self.nt, np.float32)]
        self.m._check(self.m.lib.mpt_nrc_get_params(self.h, *[self.fp(a) for a in out]))
        return out

    def __del__(self):
        self.m.lib.mpt_nrc_destroy(self.h)


@pytest.mark.gpu
def test_gpu_prediction_loss_and_gradients_match_oracle():
    import ctypes as C
    import mypbr_amd as m
    c = _Cache(7)
    assert c.levels == tuple(nr.level_table()) or c.levels[0] == nr.level_table()[0]
    params = random_params(8).astype(np.float32)
    table = np.random.default_rng(9).uniform(-0.5, 0.5, c.nt).astype(np.float32)
    c.set(params, table)
    pos, other, target = samples(500, 10)
    dp, do, dt = m.DeviceArray.of(pos), m.DeviceArray.of(other), m.DeviceArray.of(target)
    pred = m.DeviceArray((500, 3), np.float32)
    m._check(m.lib.mpt_nrc_infer_dev(c.h, dp.handle, do.handle, 500, pred.handle))
    loss = C.c_double()
    m._check(m.lib.mpt_nrc_train_dev(c.h, dp.handle, do.handle, dt.handle, 500, 0.0, C.byref(loss)))
    _, g, _, tg = c.get()
    oloss, og, otg, oy = nr.gradients(params.astype(np.float64), table.astype(np.float64), pos, other, target,
                                      c.levels, round_bf16=True)
    got = pred.numpy()
    perr = np.max(np.abs(got - oy[:, :3])) / np.max(np.abs(oy[:, :3]))
    print(f"pred rel {perr:.2e} loss gpu {loss.value:.6e} oracle {oloss:.6e}", flush=True)
    assert perr < 1e-4
    assert abs(loss.value - oloss) / oloss < 1e-3
    cin, cout, woff, boff, _ = nr.layout()
    for l in range(nr.LAYERS):
        a, b = woff[l], boff[l] + cout[l]
        err = np.linalg.norm(g[a:b] - og[a:b]) / max(np.linalg.norm(og[a:b]), 1e-12)
        print(f"layer {l}: rel err {err:.2e}", flush=True)
        assert err < 2e-2
    err = np.linalg.norm(tg - otg) / np.linalg.norm(otg)
    print(f"tables: rel err {err:.2e}, touched {np.count_nonzero(otg)}", flush=True)
    assert err < 2e-2


def _render(m, s, cam, seed, nrc=None, nopts=None, passes=None):
    import ctypes as C
    from mypbr_amd.abi import RenderOpts, RenderStats
    w, h = cam.image_width, cam.height()
    rad = np.zeros((h, w, 3))
    o = RenderOpts(seed=seed, pass_count=passes or 0)
    st = RenderStats()
    if nrc is None:
        m._check(m.lib.mpt_render(s._h, C.byref(cam), C.byref(o), m._dp(rad), None, None, None,
                                  m.abi.PASS_CALLBACK(), None, C.byref(st)))
        return rad, None
    ns = m.abi.NrcStats()
    m._check(m.lib.mpt_render_nrc(s._h, C.byref(cam), C.byref(o), nrc.h, C.byref(nopts), m._dp(rad), None, None,
                                  None, m.abi.PASS_CALLBACK(), None, C.byref(st), C.byref(ns)))
    return rad, ns


def _opts(m, **kw):
    o = m.abi.NrcOpts()
    m.lib.mpt_nrc_defaults(m.C.byref(o))
    for k, v in kw.items():
        if k in ("box_min", "box_max"):
            getattr(o, k)[:] = v
        else:
            setattr(o, k, v)
    return o


@pytest.mark.gpu
def test_training_only_leaves_the_image_unchanged():
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    d = m.describe(scene_path("cornell.scene"))
    s = m.Scene.from_desc(d.desc, d.camera)
    cam = camera_copy(d.camera, image_width=48, samples_per_pixel=4, max_depth=8)
    plain, _ = _render(m, s, cam, 3)
    c = _Cache(1)
    img, ns = _render(m, s, cam, 3, c, _opts(m, infer=0, train_stride=4))
    assert np.array_equal(img, plain)
    assert ns.queries > 0 and ns.train_samples > 0 and np.isfinite(ns.loss)


def rel_mse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


@pytest.mark.gpu
def test_cache_learns_online_and_beats_the_plain_tracer_at_equal_spp():
    """Online training on the renders' own training paths: the error of the
    cached render (paths cut at bounce 2) falls well below the untrained
    cache's, and below the plain tracer's at the same samples per pixel."""
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    d = m.describe(scene_path("cornell.scene"))
    s = m.Scene.from_desc(d.desc, d.camera)
    ref, _ = _render(m, s, camera_copy(d.camera, image_width=160, samples_per_pixel=2048, max_depth=16), 99)
    box = dict(box_min=(0, 0, -1), box_max=(556, 556, 556))
    cam = camera_copy(d.camera, image_width=160, samples_per_pixel=4, max_depth=16)
    c = _Cache(5)
    untrained, _ = _render(m, s, cam, 7, c, _opts(m, train=0, train_stride=16, **box))
    for k in range(20):  # frames: each renders 4 spp and trains on its training paths
        _render(m, s, cam, 1000 + k, c, _opts(m, train_stride=16, **box))
    t0 = time.perf_counter()
    cached, _ = _render(m, s, cam, 7, c, _opts(m, train=0, train_stride=16, **box))
    t_nrc = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain, _ = _render(m, s, cam, 7)
    t_plain = time.perf_counter() - t0
    e0, e_nrc, e_plain = rel_mse(untrained, ref), rel_mse(cached, ref), rel_mse(plain, ref)
    print(f"4 spp relMSE: untrained cache {e0:.4f}, trained cache {e_nrc:.4f} ({t_nrc * 1e3:.1f} ms), "
          f"plain {e_plain:.4f} ({t_plain * 1e3:.1f} ms)", flush=True)
    assert e_nrc < 0.5 * e0
    assert e_nrc < e_plain
