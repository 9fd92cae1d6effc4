"""Image error and the learned-versus-plain comparison at equal time.

relMSE is the mean over pixels and channels of (x - r)^2 / (r^2 + eps)
with eps = 0.01, against a reference rendered by the plain GPU tracer at
high spp. SSIM is computed on tone-mapped luminance l / (1 + l) with 8 x 8
windows (Wang et al. 2004 constants for a unit range).
"""
import ctypes as C
import time

import numpy as np

from . import Scene, _check, _dp, abi, denoise, lib
from .abi import camera_copy

EPS = 1e-2


def rel_mse(x, ref, eps=EPS):
    return float(np.mean((np.asarray(x, np.float64) - ref) ** 2 / (ref ** 2 + eps)))


def _luma(x):
    l = 0.2126 * x[..., 0] + 0.7152 * x[..., 1] + 0.0722 * x[..., 2]
    l = np.maximum(l, 0)
    return l / (1 + l)


def ssim(x, ref, win=8):
    a, b = _luma(np.asarray(x, np.float64)), _luma(ref)
    h, w = (a.shape[0] // win) * win, (a.shape[1] // win) * win
    a = a[:h, :w].reshape(h // win, win, w // win, win)
    b = b[:h, :w].reshape(h // win, win, w // win, win)
    ma, mb = a.mean((1, 3)), b.mean((1, 3))
    va, vb = a.var((1, 3)), b.var((1, 3))
    cov = ((a - ma[:, None, :, None]) * (b - mb[:, None, :, None])).mean((1, 3))
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    s = ((2 * ma * mb + c1) * (2 * cov + c2)) / ((ma ** 2 + mb ** 2 + c1) * (va + vb + c2))
    return float(s.mean())


def render(scene, cam, spp, seed, features=False):
    """Frame (radiance, and guides with features) and its wall time."""
    c = camera_copy(cam, samples_per_pixel=spp, collect_features=int(features))
    t0 = time.perf_counter()
    f = scene.render(c, seed=seed)
    return f, time.perf_counter() - t0


def views(n, seed, radius=(8, 14), height=(1.0, 4.0)):
    """Camera positions around the origin (the spheres scene's centre)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(0, 2 * np.pi)
        r = rng.uniform(*radius)
        out.append(((r * np.cos(a), rng.uniform(*height), r * np.sin(a)), (0, rng.uniform(0, 1), 0)))
    return out


def view_camera(base, lookfrom, lookat, width):
    c = camera_copy(base, image_width=width)
    c.lookfrom[:] = lookfrom
    c.lookat[:] = lookat
    return c


def frame_dict(scene, f, cam):
    return dict(radiance=f.radiance, albedo=f.albedo, normal=f.normal, variance=f.variance,
                depth=scene.primary_depth(cam))


def train_kpcnn(scene, base_cam, n_views=8, width=1920, spp_in=4, spp_ref=256, steps=600, lr=5e-3, seed=1,
                hidden=32, crop=128, crops_per_view=8, loss=1):
    """A kernel-predicting denoiser trained on the tracer's own renders of
    n_views camera positions (inputs spp_in with guides, targets spp_ref),
    at the pixel scale it will run at: random crop x crop windows of
    width-wide views, all of them in every step."""
    from .learn import KPCNN, to_device
    from . import DeviceArray
    rng = np.random.default_rng(seed)
    frames, refs = [], []
    for k, (lf, la) in enumerate(views(n_views, seed)):
        cam = view_camera(base_cam, lf, la, width)
        f, _ = render(scene, cam, spp_in, 100 + k, features=True)
        full = frame_dict(scene, f, cam)
        ref = render(scene, cam, spp_ref, 10_000 + k)[0].radiance
        h, w = ref.shape[:2]
        for _ in range(crops_per_view if crop else 1):
            if crop:
                y0, x0 = rng.integers(0, h - crop + 1), rng.integers(0, w - crop + 1)
                sl = (slice(y0, y0 + crop), slice(x0, x0 + crop))
            else:
                sl = (slice(None), slice(None))
            frames.append({key: v[sl] for key, v in full.items()})
            refs.append(ref[sl])
    net = KPCNN(layers=6, hidden=hidden, seed=seed)
    net.set_loss(loss)
    dev = to_device(frames)
    tgt = DeviceArray.of(np.stack(refs), np.float32)
    t0 = time.perf_counter()
    losses = [net.run(dev, target=tgt, lr=lr) for _ in range(steps)]
    return net, {"views": n_views, "width": width, "crop": crop, "crops": len(frames), "spp_in": spp_in,
                 "spp_target": spp_ref, "steps": steps, "lr": lr, "hidden": hidden, "loss": ["relmse", "log_l1"][loss],
                 "loss_first": losses[0],
                 "loss_last": losses[-1], "train_s": time.perf_counter() - t0}


class Cache:
    """A neural radiance cache (mpt_nrc) and renders with it."""

    def __init__(self, seed=1):
        self.h = C.c_void_p()
        _check(lib.mpt_nrc_create(seed, C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.mpt_nrc_destroy(self.h)
            self.h = None

    def render(self, scene, cam, spp, seed, train=True, box=None, stride=256):
        c = camera_copy(cam, samples_per_pixel=spp)
        o = abi.NrcOpts()
        lib.mpt_nrc_defaults(C.byref(o))
        o.train = int(train)
        o.train_stride = stride
        if box:
            o.box_min[:] = box[0]
            o.box_max[:] = box[1]
        rad = np.zeros((c.height(), c.image_width, 3))
        ro = abi.RenderOpts(seed=seed)
        st = abi.RenderStats()
        ns = abi.NrcStats()
        t0 = time.perf_counter()
        _check(lib.mpt_render_nrc(scene._h, C.byref(c), C.byref(ro), self.h, C.byref(o), _dp(rad), None, None, None,
                                  abi.PASS_CALLBACK(), None, C.byref(st), C.byref(ns)))
        return rad, time.perf_counter() - t0, ns


def equal_time_study(scene_path, width=1920, ref_spp=4096, plain_spp=(4, 16, 64, 256, 1024), box=None, log=print):
    """Plain tracer at several spp against the learned forms (kernel-predicting
    denoiser on 4 and 16 spp, radiance cache) and the classical filter, all
    against one ref_spp render of the plain tracer; times are wall clock."""
    from . import describe
    d = describe(scene_path)
    scene = Scene.from_desc(d.desc, d.camera)
    cam = camera_copy(d.camera, image_width=width)
    ref, t_ref = render(scene, cam, ref_spp, 987_654)
    ref = ref.radiance
    log(f"reference {ref_spp} spp: {t_ref:.2f} s")
    out = {"reference_spp": ref_spp, "epsilon": EPS, "plain": [], "classical": [], "kpcnn": [], "nrc": []}
    for spp in plain_spp:
        f, t = render(scene, cam, spp, 1)
        out["plain"].append({"spp": spp, "seconds": t, "relmse": rel_mse(f.radiance, ref), "ssim": ssim(f.radiance, ref)})
        log(f"plain {spp} spp: {t * 1e3:.1f} ms relMSE {out['plain'][-1]['relmse']:.5f}")
    net, info = train_kpcnn(scene, d.camera)
    out["kpcnn_training"] = info
    depth_t0 = time.perf_counter()
    depth = scene.primary_depth(cam)
    t_depth = time.perf_counter() - depth_t0
    for spp in (4, 16):
        f, t = render(scene, cam, spp, 2, features=True)
        fr = dict(radiance=f.radiance, albedo=f.albedo, normal=f.normal, variance=f.variance, depth=depth)
        t0 = time.perf_counter()
        den = net.denoise([fr])[0]
        t_net = time.perf_counter() - t0
        out["kpcnn"].append({"spp": spp, "seconds": t + t_depth + t_net, "network_s": t_net,
                             "relmse": rel_mse(den, ref), "ssim": ssim(den, ref)})
        t0 = time.perf_counter()
        cl = denoise(fr)
        t_cl = time.perf_counter() - t0
        out["classical"].append({"spp": spp, "seconds": t + t_cl, "relmse": rel_mse(cl, ref), "ssim": ssim(cl, ref)})
        log(f"{spp} spp: kpcnn relMSE {out['kpcnn'][-1]['relmse']:.5f}, classical {out['classical'][-1]['relmse']:.5f}")
    cache = Cache(3)
    t0 = time.perf_counter()
    for k in range(16):  # online training frames of 4 spp
        cache.render(scene, cam, 4, 5000 + k, box=box)
    out["nrc_training_s"] = time.perf_counter() - t0
    for spp in (4, 16, 64):
        rad, t, ns = cache.render(scene, cam, spp, 3, box=box)
        out["nrc"].append({"spp": spp, "seconds": t, "relmse": rel_mse(rad, ref), "ssim": ssim(rad, ref),
                           "queries": ns.queries})
        log(f"nrc {spp} spp: {t * 1e3:.1f} ms relMSE {out['nrc'][-1]['relmse']:.5f}")
    # the target: the learned forms reaching the plain tracer's 1024-spp error in a quarter of its time
    p1024 = next((p for p in out["plain"] if p["spp"] == 1024), None)
    if p1024:
        learned = out["kpcnn"] + out["nrc"]
        hits = [x for x in learned if x["relmse"] <= p1024["relmse"]]
        best = min(hits, key=lambda x: x["seconds"]) if hits else None
        out["target"] = {"plain_1024_relmse": p1024["relmse"], "plain_1024_seconds": p1024["seconds"],
                         "learned_time_ratio": best["seconds"] / p1024["seconds"] if best else None,
                         "met": bool(best and best["seconds"] <= p1024["seconds"] / 4)}
    return out
