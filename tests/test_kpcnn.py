"""The learned kernel-predicting denoiser (include/mpt.h, mpt_kpcnn_*).

CPU: the numpy restatement (oracle/kpcnn.py) is pinned by finite
differences of its own float64 forward pass, by the bf16 rounding rule and
by closed forms of the kernel application. GPU: the BF16 matrix-core GEMM
against numpy, the network's loss and every parameter gradient against the
oracle computed with the same roundings, and training on the tracer's own
renders, which must beat the noisy input it starts from on a held-out view.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, scene_path

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import kpcnn as kp  # noqa: E402


def layout(layers, hidden):
    cin = [kp.K_FEAT] + [hidden] * (layers - 1)
    cout = [hidden] * (layers - 1) + [kp.K_OUT]
    woff, boff, off = [], [], 0
    for ci, co in zip(cin, cout):
        woff.append(off)
        off += co * 9 * ci
        boff.append(off)
        off += co
    return cin, cout, woff, boff, off


def random_frames(B, H, W, seed):
    rng = np.random.default_rng(seed)
    frames = []
    for _ in range(B):
        alb = rng.uniform(0.05, 0.95, (H, W, 3))
        nrm = rng.standard_normal((H, W, 3))
        nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
        rad = alb * rng.uniform(0.1, 2.0, (H, W, 1)) * rng.uniform(0.5, 1.5, (H, W, 3))
        frames.append(dict(radiance=rad, albedo=alb, normal=nrm, variance=rng.exponential(0.05, (H, W)),
                           depth=rng.uniform(0, 20, (H, W))))
    target = np.stack([f["radiance"] * rng.uniform(0.8, 1.2, f["radiance"].shape) for f in frames])
    return frames, target


def oracle_inputs(frames):
    cat = {k: np.stack([f[k] for f in frames]) for k in ("radiance", "albedo", "normal", "variance", "depth")}
    return kp.features(cat["radiance"], cat["albedo"], cat["normal"], cat["variance"], cat["depth"])


def random_net(layers, hidden, seed, scale=0.3):
    cin, cout, woff, boff, n = layout(layers, hidden)
    p = np.random.default_rng(seed).uniform(-scale, scale, n)
    return kp.Net(p, cin, cout, woff, boff)


# ---------------------------------------------------------------- CPU: the oracle itself

def test_bf16_rounding_rule():
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -9, 1 + 2.0 ** -7 + 2.0 ** -8, -2.5, 0.0, 3.0e38])
    want = np.array([1.0, 1.0, 1 + 2.0 ** -7, 1 + 2.0 ** -6, -2.5, 0.0, 226 / 128 * 2.0 ** 127])
    assert np.array_equal(kp.bf16(x), want)


def test_kernel_application_closed_forms():
    B, H, W = 1, 9, 9
    rng = np.random.default_rng(2)
    light = rng.uniform(0, 1, (B * H * W, 3))
    albe = np.ones((B * H * W, 3))
    p = 4 * W + 4  # centre: the whole 5x5 neighbourhood is inside
    rad, w, q = kp.apply_kernel(np.zeros((B * H * W, kp.K_OUT)), light, albe, B, H, W)
    box = light.reshape(H, W, 3)[2:7, 2:7].reshape(-1, 3).mean(0)
    assert np.allclose(rad[p], box, rtol=1e-14)
    logits = np.zeros((B * H * W, kp.K_OUT))
    logits[:, 7] = 200.0  # tap 7 = (dy, dx) = (-1, 0)
    rad, _, _ = kp.apply_kernel(logits, light, albe, B, H, W)
    assert np.allclose(rad[p], light[p - W], rtol=1e-12)


def test_oracle_backward_equals_finite_differences():
    B, H, W = 2, 6, 5
    frames, target = random_frames(B, H, W, 3)
    inp = oracle_inputs(frames)
    tgt = target.reshape(-1, 3)
    net = random_net(3, 16, 4)
    for mode in (0, 1):
        _finite_differences(net, inp, tgt, B, H, W, mode)


def _finite_differences(net, inp, tgt, B, H, W, mode):
    loss, grad = kp.gradients(net, inp, tgt, B, H, W, round_bf16=False, mode=mode)
    rng = np.random.default_rng(5)
    # parameters that matter: used feature channels / logits of each layer
    picks = []
    for l in range(net.layers):
        lo = net.woff[l]
        hi = net.boff[l] + net.cout[l]
        picks += list(rng.integers(lo, hi, 12))
    worst = 0.0
    for i in picks:
        h = 1e-6 * max(1.0, abs(net.p[i]))
        saved = net.p[i]
        net.p[i] = saved + h
        lp, _ = kp.gradients(net, inp, tgt, B, H, W, round_bf16=False, mode=mode)
        net.p[i] = saved - h
        lm, _ = kp.gradients(net, inp, tgt, B, H, W, round_bf16=False, mode=mode)
        net.p[i] = saved
        fd = (lp - lm) / (2 * h)
        err = abs(fd - grad[i]) / max(abs(fd), abs(grad[i]), 1e-9)
        if abs(fd) > 1e-9 or abs(grad[i]) > 1e-9:
            worst = max(worst, err)
    assert worst < 1e-4, worst


def test_planted_mistake_in_oracle_is_caught():
    B, H, W = 1, 5, 5
    frames, target = random_frames(B, H, W, 6)
    inp = oracle_inputs(frames)
    net = random_net(2, 16, 7)
    _, good = kp.gradients(net, inp, target.reshape(-1, 3), B, H, W, round_bf16=False)
    orig = kp.col2im
    try:  # a transposed tap order in the adjoint
        kp.col2im = lambda d, B_, H_, W_, C_: orig(d.reshape(-1, 9, C_)[:, ::-1].reshape(d.shape), B_, H_, W_, C_)
        _, bad = kp.gradients(net, inp, target.reshape(-1, 3), B, H, W, round_bf16=False)
    finally:
        kp.col2im = orig
    lo = net.woff[0]
    assert not np.allclose(good[lo:net.boff[0]], bad[lo:net.boff[0]])


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("at,bt", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_bf16_matches_numpy(at, bt):
    import mypbr_amd as m
    rng = np.random.default_rng(at * 2 + bt)
    M, N, K = 100, 72, 160
    a = kp.bf16(rng.standard_normal((M, K)))
    b = kp.bf16(rng.standard_normal((N, K)))
    A = a.T if at else a
    Bm = b.T if bt else b

    def dev(x):  # the bf16 bit patterns
        return m.DeviceArray.of((np.ascontiguousarray(x, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16))
    dA, dB = dev(A), dev(Bm)
    c = m.DeviceArray((M, N), np.float32)
    rc = m.lib.mpt_gemm_bf16(at, bt, M, N, K, dA.handle, A.shape[1], dB.handle, Bm.shape[1], c.handle, N, 0, None)
    assert rc == 0
    m.synchronize()
    want = a @ b.T
    got = c.numpy()
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gradients_match_oracle(mode):
    import mypbr_amd as m
    from mypbr_amd.learn import KPCNN, to_device
    B, H, W = 2, 12, 10
    frames, target = random_frames(B, H, W, 11)
    net = KPCNN(layers=3, hidden=16, seed=3)
    cin, cout, woff, boff = net.layout()
    assert (cin, cout) == layout(3, 16)[:2]
    p = np.random.default_rng(12).uniform(-0.3, 0.3, net.nparams).astype(np.float32)
    net.set_params(p)
    net.set_loss(mode)
    dev = to_device(frames)
    tgt = m.DeviceArray.of(target, np.float32)
    loss = net.run(dev, target=tgt)
    g = net.grads().astype(np.float64)
    onet = kp.Net(p.astype(np.float64), cin, cout, woff, boff)
    oloss, og = kp.gradients(onet, oracle_inputs(frames), target.reshape(-1, 3), B, H, W, round_bf16=True,
                             mode=mode)
    print(f"loss gpu {loss:.6e} oracle {oloss:.6e}", flush=True)
    assert abs(loss - oloss) / oloss < 1e-3
    for l in range(3):
        for lo, hi, what in [(woff[l], boff[l], "w"), (boff[l], boff[l] + cout[l], "b")]:
            err = np.linalg.norm(g[lo:hi] - og[lo:hi]) / max(np.linalg.norm(og[lo:hi]), 1e-12)
            print(f"layer {l} {what}: rel err {err:.2e}", flush=True)
            assert err < 2e-2, (l, what, err)


def rel_mse(x, ref):
    """relMSE with epsilon 0.01: mean over pixels and channels of (x - r)^2 / (r^2 + 0.01)."""
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def views(n, seed):
    """Camera positions around the spheres scene (the file's camera is at (13, 2, 3))."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(0, 2 * np.pi)
        r = rng.uniform(8, 14)
        out.append(((r * np.cos(a), rng.uniform(1.0, 4.0), r * np.sin(a)), (0, rng.uniform(0, 1), 0)))
    return out


def render_views(m, s, base, cams, spp_noisy, spp_ref, width, seed):
    from mypbr_amd.abi import camera_copy
    frames, refs = [], []
    for k, (lf, la) in enumerate(cams):
        cam = camera_copy(base, image_width=width, max_depth=12, collect_features=1)
        cam.lookfrom[:] = lf
        cam.lookat[:] = la
        cam.samples_per_pixel = spp_noisy
        f = s.render(cam, seed=seed + k)
        frames.append(dict(radiance=f.radiance, albedo=f.albedo, normal=f.normal, variance=f.variance,
                           depth=s.primary_depth(cam)))
        cam.samples_per_pixel = spp_ref
        cam.collect_features = 0
        refs.append(s.render(cam, seed=10_000 + seed + k).radiance)
    return frames, np.stack(refs)


@pytest.mark.gpu
def test_training_on_own_renders_beats_the_noisy_input():
    import mypbr_amd as m
    from mypbr_amd.learn import KPCNN, to_device
    d = m.describe(scene_path("weekend_cover.scene"))
    s = m.Scene.from_desc(d.desc, d.camera)
    # 24 views: with 12 the held-out error sat near the bar from one run of
    # the (atomically summed, so run-dependent) training to the next
    train, train_ref = render_views(m, s, d.camera, views(24, 1), 4, 256, 96, 100)
    test, test_ref = render_views(m, s, d.camera, views(2, 2), 4, 256, 96, 500)
    net = KPCNN(layers=6, hidden=32, seed=7)
    dev = to_device(train)
    tgt = m.DeviceArray.of(train_ref, np.float32)
    first = None
    for it in range(250):
        loss = net.run(dev, target=tgt, lr=2e-3)
        first = loss if first is None else first
    print(f"train loss {first:.4f} -> {loss:.4f}", flush=True)
    assert loss < first
    out = net.denoise(test)
    noisy = np.stack([f["radiance"] for f in test])
    classical = np.stack([m.denoise(f) for f in test])
    e_noisy, e_net, e_cls = rel_mse(noisy, test_ref), rel_mse(out, test_ref), rel_mse(classical, test_ref)
    print(f"held-out relMSE: noisy {e_noisy:.4f}, classical {e_cls:.4f}, learned {e_net:.4f}", flush=True)
    assert e_net < 0.5 * e_noisy
