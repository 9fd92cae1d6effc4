"""Image error and the learned-versus-plain comparison at equal time.

relMSE is the mean over pixels and channels of (x - r)^2 / (r^2 + eps)
with eps = 0.01, against a reference rendered by the plain GPU tracer at
high spp. SSIM is computed on tone-mapped luminance l / (1 + l) with 8 x 8
windows (Wang et al. 2004 constants for a unit range).
"""
import ctypes as C
import time

import numpy as np

from . import Scene, _check, _dp, abi, denoise, lib, synchronize
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


# the denoiser's albedo and normal guides come from the first GUIDE_PASSES passes
# (mpt_render_opts.guide_passes; 0: all). On the spheres scene, whose camera
# has a defocus disk, guides from 16 or 64 passes keep enough edge noise to
# cost the trained network most of its gain at 192 spp (relMSE 0.00050 and
# 0.00043 against 0.00026 with every pass), so every pass traces them
GUIDE_PASSES = 0


def render(scene, cam, spp, seed, features=False):
    """Frame (radiance, and guides with features) and its wall time."""
    c = camera_copy(cam, samples_per_pixel=spp, collect_features=int(features))
    t0 = time.perf_counter()
    f = scene.render(c, seed=seed, guide_passes=GUIDE_PASSES if features else 0)
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


def crop_camera(base, lookfrom, lookat, full_width, crop, cx, cy):
    """A crop x crop camera that sees, at the same pixel scale, the window of
    the full_width-wide view (lookfrom -> lookat, base's vfov and aspect)
    centred on pixel (cx, cy): aimed at that pixel, its vertical field of
    view shrunk to crop rows of the full view's. Off-axis windows differ from
    a true crop only by the small change of perspective."""
    full_h = max(1, int(full_width / base.aspect_ratio))
    lf, la = np.array(lookfrom, float), np.array(lookat, float)
    w = lf - la
    w /= np.linalg.norm(w)
    u = np.cross(np.array(base.vup[:]), w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    half_h = np.tan(np.radians(base.vfov) / 2)
    half_w = half_h * full_width / full_h
    px = (cx + 0.5) / full_width * 2 - 1
    py = (cy + 0.5) / full_h * 2 - 1
    d = -w + px * half_w * u - py * half_h * v
    c = camera_copy(base, image_width=crop, aspect_ratio=1.0)
    c.vfov = float(np.degrees(2 * np.arctan(half_h * crop / full_h)))
    c.lookfrom[:] = lf
    c.lookat[:] = lf + d
    return c


def kpcnn_batches(scene, base_cam, n_views=16, width=1920, spp_in=(32, 64, 128, 192), spp_ref=4096, seed=1,
                  crop=256, crops_per_view=4, batch=16):
    """The denoiser's training set on the device: n_views x crops_per_view
    crop x crop windows of width-wide views (random camera positions, windows
    at the full view's pixel scale), inputs at each spp of spp_in with their
    guides, targets at spp_ref; grouped into batches of `batch` windows at one
    input spp, the spp level cycling fastest. Returns (batches, info)."""
    from .learn import to_device
    from . import DeviceArray
    rng = np.random.default_rng(seed)
    full_h = max(1, int(width / base_cam.aspect_ratio))
    cams = []
    for lf, la in views(n_views, seed):
        for _ in range(crops_per_view):
            cx, cy = rng.integers(crop // 2, width - crop // 2), rng.integers(crop // 2, full_h - crop // 2)
            cams.append(crop_camera(base_cam, lf, la, width, crop, cx, cy))
    t0 = time.perf_counter()
    refs = np.stack([render(scene, c, spp_ref, 10_000 + k)[0].radiance for k, c in enumerate(cams)])
    t_ref = time.perf_counter() - t0
    groups = [list(range(g, min(g + batch, len(cams)))) for g in range(0, len(cams), batch)]
    batches = []  # (device inputs, device targets), spp level fastest
    inputs = {spp: [frame_dict(scene, render(scene, c, spp, 100 * spp + k, features=True)[0], c)
                    for k, c in enumerate(cams)] for spp in spp_in}
    for g in groups:
        tgt = DeviceArray.of(refs[g], np.float32)
        for spp in spp_in:
            batches.append((to_device([inputs[spp][k] for k in g]), tgt))
    return batches, {"views": n_views, "width": width, "crop": crop, "crops": len(cams), "batch": batch,
                     "spp_in": list(spp_in), "spp_target": spp_ref, "target_render_s": t_ref}


def fit_kpcnn(batches, steps=3000, lr=2e-3, lr_end=1e-4, seed=1, layers=6, hidden=64, relmse_share=1 / 3):
    """Adam over the batches in turn; the learning rate falls from lr to
    lr_end on a cosine. The loss is the L1 of log(1 + x) (robust to the
    inputs' fireflies), then for the last relmse_share of the steps relMSE
    itself, the metric it is judged by. On the spheres scene (16 views x 4
    windows, 192-spp inputs) 3000 steps left relMSE 0.00019 where 2000 left
    0.00026 (two seeds within 1%), and 8 or 10 layers did no better than 6.
    Returns (net, info)."""
    from .learn import KPCNN
    net = KPCNN(layers=layers, hidden=hidden, seed=seed)
    switch = int(steps * (1 - relmse_share))
    t0 = time.perf_counter()
    losses = []
    for i in range(steps):
        if i in (0, switch):
            net.set_loss(1 if i < switch else 0)
        rate = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * i / steps))
        dev, tgt = batches[i % len(batches)]
        losses.append(net.run(dev, target=tgt, lr=float(rate)))
    train_s = time.perf_counter() - t0
    return net, {"steps": steps, "lr": lr, "hidden": hidden, "layers": layers, "parameters": int(net.nparams),
                 "loss": "log_l1, then relmse for the last %.2f" % relmse_share, "lr_end": lr_end,
                 "loss_first": losses[0], "loss_last_log_l1": losses[switch - 1], "loss_first_relmse": losses[switch],
                 "loss_last": losses[-1], "train_s": train_s, "step_ms": 1e3 * train_s / steps}


def train_kpcnn(scene, base_cam, n_views=16, width=1920, spp_in=(32, 64, 128, 192), spp_ref=4096, steps=3000,
                lr=2e-3, lr_end=1e-4, seed=1, layers=6, hidden=64, crop=256, crops_per_view=4, batch=16,
                relmse_share=1 / 3, log=None):
    """A kernel-predicting denoiser trained on the tracer's own renders
    (kpcnn_batches, then fit_kpcnn)."""
    batches, info = kpcnn_batches(scene, base_cam, n_views, width, spp_in, spp_ref, seed, crop, crops_per_view, batch)
    net, fit = fit_kpcnn(batches, steps, lr, lr_end, seed, layers, hidden, relmse_share)
    info.update(fit)
    if log:
        log(f"kpcnn: {net.nparams} parameters, {steps} steps of {batch} of {info['crops']} {crop}^2 crops in "
            f"{fit['train_s']:.1f} s, log-L1 loss {fit['loss_first']:.4f} -> {fit['loss_last_log_l1']:.4f}, relMSE "
            f"loss {fit['loss_first_relmse']:.5f} -> {fit['loss_last']:.5f}")
    return net, info


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


def equal_time_study(scene_path, width=1920, ref_spp=4096, plain_spp=(4, 16, 64, 128, 192, 256, 1024),
                     kpcnn_spp=(4, 16, 64, 128, 160, 192, 224), nrc_frames=256, box=None, log=print):
    """Plain tracer at several spp against the learned forms (kernel-predicting
    denoiser on kpcnn_spp inputs, radiance cache) and the classical filter,
    all against one ref_spp render of the plain tracer; times are wall clock
    from the render's start to the image in host memory."""
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
    net, info = train_kpcnn(scene, d.camera, width=width, log=log)
    out["kpcnn_training"] = info
    for spp in kpcnn_spp:
        # render -> averaged frame and guides -> depth -> network, all in HBM; the
        # denoised image is read back as the plain tracer's frame is
        c = camera_copy(cam, samples_per_pixel=spp, collect_features=1)
        synchronize()
        t0 = time.perf_counter()
        scene.render(c, seed=2, read=False, guide_passes=GUIDE_PASSES)
        dev = scene.frame_dev(c)
        t1 = time.perf_counter()
        den = net.denoise_dev(dev).numpy()
        t2 = time.perf_counter()
        out["kpcnn"].append({"spp": spp, "seconds": t2 - t0, "network_s": t2 - t1, "relmse": rel_mse(den, ref),
                             "ssim": ssim(den, ref)})
        msg = f"{spp} spp: kpcnn {1e3 * (t2 - t0):.1f} ms relMSE {out['kpcnn'][-1]['relmse']:.5f}"
        if spp <= 16:  # the classical filter on the same frame (host arrays in, as mpt_denoise takes them)
            fr = {k: v.numpy() for k, v in dev.items()}
            t3 = time.perf_counter()
            cl = denoise(fr)
            t_cl = time.perf_counter() - t3
            out["classical"].append({"spp": spp, "seconds": (t1 - t0) + t_cl, "relmse": rel_mse(cl, ref),
                                     "ssim": ssim(cl, ref)})
            msg += f", classical relMSE {out['classical'][-1]['relmse']:.5f}"
        log(msg)
    # the radiance cache online, as it would run in a viewer: nrc_frames frames
    # of 1 spp, each training it on about 65,536 path suffixes of its own
    # paths, then frames at a few spp with the trained cache
    cache = Cache(3)
    npix = cam.image_width * cam.height()
    t0 = time.perf_counter()
    for k in range(nrc_frames):
        cache.render(scene, cam, 1, 5000 + k, box=box, stride=max(1, npix // 65536))
    out["nrc_training"] = {"frames": nrc_frames, "spp": 1, "samples_per_frame": npix // max(1, npix // 65536),
                           "seconds": time.perf_counter() - t0}
    for spp in (1, 4, 16, 64):
        rad, t, ns = cache.render(scene, cam, spp, 3, box=box, stride=max(1, npix * spp // 65536))
        out["nrc"].append({"spp": spp, "seconds": t, "relmse": rel_mse(rad, ref), "ssim": ssim(rad, ref),
                           "queries": ns.queries})
        log(f"nrc {spp} spp: {t * 1e3:.1f} ms relMSE {out['nrc'][-1]['relmse']:.5f}")
    # the target: the learned forms reaching the plain tracer's 1024-spp error in a quarter of its time
    p1024 = next((p for p in out["plain"] if p["spp"] == 1024), None)
    if p1024:
        learned = [dict(x, method="kpcnn") for x in out["kpcnn"]] + [dict(x, method="nrc") for x in out["nrc"]]
        hits = [x for x in learned if x["relmse"] <= p1024["relmse"]]
        best = min(hits, key=lambda x: x["seconds"]) if hits else None
        quarter = [x for x in learned if x["seconds"] <= p1024["seconds"] / 4]
        closest = min(quarter, key=lambda x: x["relmse"]) if quarter else None
        out["target"] = {"plain_1024_relmse": p1024["relmse"], "plain_1024_seconds": p1024["seconds"],
                         "learned_time_ratio": best["seconds"] / p1024["seconds"] if best else None,
                         "met": bool(best and best["seconds"] <= p1024["seconds"] / 4),
                         "best_within_quarter_time": None if closest is None else
                         {"method": closest["method"], "spp": closest["spp"], "seconds": closest["seconds"],
                          "relmse": closest["relmse"], "relmse_over_target": closest["relmse"] / p1024["relmse"]}}
    return out
