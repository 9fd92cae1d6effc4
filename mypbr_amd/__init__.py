"""mypbr_amd - MI355X wavefront path-tracing core (Python host side).

A thin ctypes layer over libmpt.so, the C ABI in include/mpt.h. It mirrors
the reference renderer's render entry (camera::render, include/camera.h:79)
for Python callers, tests and the benchmark:

    scene = Scene.load("scenes/cornell.scene")      # load_scene + bvh (scene.h:244, main.cpp:70)
    frame = scene.render(seed=1)                    # camera::render -> Frame(radiance, ...)
    ids, t, front = scene.trace_closest(rays, tt)   # bvh::hit for a batch of rays

Every compute call goes to the HIP kernels in libmpt.so; if the library is
missing this module raises at import time instead of falling back.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import Camera, RenderOpts, RenderStats, SceneDesc

HERE = os.path.dirname(os.path.abspath(__file__))
# MPT_LIBRARY selects another build of the same library (tuning experiments)
LIB_PATH = os.environ.get("MPT_LIBRARY", os.path.join(HERE, "libmpt.so"))

__all__ = ["Scene", "Frame", "render_batch", "MptError", "lib", "abi", "Camera", "RenderOpts", "device_info", "denoise", "DeviceArray", "synchronize",
           "ggx_albedo_table", "describe"]


class MptError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mpt error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run python -c 'import __graft_entry__ as g; g.build()' "
                          "(or python mypbr_amd/build.py)")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "mpt_last_error": (C.c_char_p, []),
        "mpt_abi_version": (C.c_int32, []),
        "mpt_device_info": (C.c_int32, [C.c_char_p, C.c_int32, P(C.c_int32), P(C.c_int64)]),
        "mpt_scene_create": (C.c_int32, [P(SceneDesc), P(vp)]),
        "mpt_scene_load": (C.c_int32, [C.c_char_p, P(vp), P(Camera)]),
        "mpt_scene_destroy": (None, [vp]),
        "mpt_scene_describe": (C.c_int32, [C.c_char_p, P(SceneDesc), P(Camera), P(vp)]),
        "mpt_desc_free": (None, [vp]),
        "mpt_scene_counts": (C.c_int32, [vp, P(C.c_int64), P(C.c_int32), P(C.c_int64), P(C.c_int64), P(C.c_int32)]),
        "mpt_trace_closest": (C.c_int32, [vp, P(C.c_double), P(C.c_double), C.c_int64, P(C.c_int64),
                                          P(C.c_double), P(C.c_uint8)]),
        "mpt_trace_closest_dev": (C.c_int32, [vp, vp, vp, C.c_int64, vp, vp, vp, vp]),
        "mpt_trace_ties": (C.c_int32, [vp, P(C.c_double), P(C.c_double), C.c_int64, P(C.c_int64), P(C.c_double),
                                       P(C.c_uint8), P(C.c_int32), P(C.c_int64), P(C.c_int64)]),
        "mpt_render": (C.c_int32, [vp, P(Camera), P(RenderOpts), P(C.c_double), P(C.c_double), P(C.c_double),
                                   P(C.c_double), abi.PASS_CALLBACK, vp, P(RenderStats)]),
        "mpt_read_frame": (C.c_int32, [vp, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)]),
        "mpt_read_frame_dev": (C.c_int32, [vp, vp, vp, vp, vp, vp]),
        "mpt_primary_depth_dev": (C.c_int32, [vp, P(Camera), vp, vp]),
        "mpt_render_batch": (C.c_int32, [P(vp), P(Camera), C.c_int32, P(RenderOpts), P(RenderStats)]),
        "mpt_image_height": (C.c_int32, [P(Camera)]),
        "mpt_ggx_albedo_table": (C.c_int32, [P(C.c_double), P(C.c_double)]),
        "mpt_scene_light_probabilities": (C.c_int32, [vp, P(C.c_double)]),
        "mpt_desc_bounds": (C.c_int32, [P(SceneDesc), P(C.c_double)]),
        "mpt_desc_light_powers": (C.c_int32, [P(SceneDesc), P(C.c_double), P(C.c_double)]),
        "mpt_desc_light_probabilities": (C.c_int32, [P(SceneDesc), P(C.c_double), P(C.c_double)]),
        "mpt_bsdf_eval": (C.c_int32, [vp, P(C.c_int32), P(C.c_double), P(C.c_uint8), C.c_int64, P(C.c_double),
                                      P(C.c_double)]),
        "mpt_primary_depth": (C.c_int32, [vp, P(Camera), P(C.c_double)]),
        "mpt_autotune": (C.c_int32, [vp, P(Camera), C.c_int32, P(abi.AutotuneResult)]),
        "mpt_autotune_apply": (C.c_int32, [vp, P(Camera), P(abi.AutotuneResult)]),
        "mpt_autotune_key": (C.c_int32, [vp, P(Camera), C.c_char_p, C.c_int32]),
        "mpt_device_alloc": (vp, [C.c_size_t]),
        "mpt_device_free": (None, [vp]),
        "mpt_copy_to_device": (C.c_int32, [vp, vp, C.c_size_t]),
        "mpt_copy_to_host": (C.c_int32, [vp, vp, C.c_size_t]),
        "mpt_device_synchronize": (C.c_int32, []),
        "mpt_gemm_bf16": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, C.c_int64, vp,
                                      C.c_int64, vp, C.c_int64, C.c_int32, vp]),
        "mpt_kpcnn_create": (C.c_int32, [C.c_int32, C.c_int32, C.c_uint64, P(vp)]),
        "mpt_kpcnn_destroy": (None, [vp]),
        "mpt_kpcnn_param_count": (C.c_int64, [vp]),
        "mpt_kpcnn_layout": (C.c_int32, [vp, P(C.c_int32), P(C.c_int32), P(C.c_int64), P(C.c_int64)]),
        "mpt_kpcnn_get_params": (C.c_int32, [vp, P(C.c_float), P(C.c_float)]),
        "mpt_kpcnn_set_params": (C.c_int32, [vp, P(C.c_float)]),
        "mpt_kpcnn_set_loss": (C.c_int32, [vp, C.c_int32]),
        "mpt_kpcnn_run": (C.c_int32, [vp] * 6 + [C.c_int32] * 3 + [vp, C.c_float, vp, P(C.c_double)]),
        "mpt_nrc_defaults": (None, [P(abi.NrcOpts)]),
        "mpt_nrc_create": (C.c_int32, [C.c_uint64, P(vp)]),
        "mpt_nrc_destroy": (None, [vp]),
        "mpt_nrc_param_count": (C.c_int64, [vp, P(C.c_int64)]),
        "mpt_nrc_levels": (C.c_int32, [vp, P(C.c_uint32), P(C.c_uint32)]),
        "mpt_nrc_get_params": (C.c_int32, [vp] + [P(C.c_float)] * 4),
        "mpt_nrc_set_params": (C.c_int32, [vp, P(C.c_float), P(C.c_float)]),
        "mpt_nrc_infer_dev": (C.c_int32, [vp, vp, vp, C.c_int64, vp]),
        "mpt_nrc_train_dev": (C.c_int32, [vp, vp, vp, vp, C.c_int64, C.c_float, P(C.c_double)]),
        "mpt_render_nrc": (C.c_int32, [vp, P(Camera), P(RenderOpts), vp, P(abi.NrcOpts)] + [P(C.c_double)] * 4 +
                           [abi.PASS_CALLBACK, vp, P(RenderStats), P(abi.NrcStats)]),
        "mpt_denoise_defaults": (None, [P(abi.DenoiseOpts)]),
        "mpt_denoise": (C.c_int32, [P(C.c_double)] * 4 + [C.c_int32, C.c_int32, P(abi.DenoiseOpts),
                                                           P(C.c_double)]),
        "mpt_denoise_workspace_bytes": (C.c_size_t, [C.c_int32, C.c_int32]),
        "mpt_denoise_dev": (C.c_int32, [vp] * 4 + [C.c_int32, C.c_int32, P(abi.DenoiseOpts), vp, vp, vp]),
        "mpt_display_defaults": (None, [P(abi.Display)]),
        "mpt_scene_display": (C.c_int32, [C.c_char_p, P(abi.Display)]),
        "mpt_find_conductor": (C.c_int32, [C.c_char_p, P(C.c_double), P(C.c_double)]),
        "mpt_postprocess": (C.c_int32, [P(C.c_double), C.c_int32, C.c_int32, P(abi.Display), P(C.c_double)]),
        "mpt_display_bytes": (C.c_int32, [P(C.c_double), C.c_int32, C.c_int32, P(abi.Display), C.c_int32,
                                          P(C.c_uint8)]),
        "mpt_display_workspace_bytes": (C.c_size_t, [C.c_int32, C.c_int32, P(abi.Display)]),
        "mpt_postprocess_dev": (C.c_int32, [vp, C.c_int32, C.c_int32, P(abi.Display), vp, vp, vp]),
        "mpt_display_bytes_dev": (C.c_int32, [vp, C.c_int32, C.c_int32, P(abi.Display), C.c_int32, vp, vp, vp]),
        "mpt_display_noalloc_max_workspace": (C.c_size_t, []),
        "mpt_display_bytes_dev_noalloc": (C.c_int32, [vp, C.c_int32, C.c_int32, P(abi.Display), C.c_int32, vp, vp]),
        "mpt_write_image": (C.c_int32, [C.c_char_p, C.c_int32, C.c_int32, P(C.c_double), P(abi.Display)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _check(rc):
    if rc != abi.MPT_OK:
        raise MptError(rc, lib.mpt_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def device_info():
    name = C.create_string_buffer(256)
    cu = C.c_int32()
    mem = C.c_int64()
    _check(lib.mpt_device_info(name, 256, C.byref(cu), C.byref(mem)))
    return {"name": name.value.decode(), "cu_count": cu.value, "hbm_bytes": mem.value}


def ggx_albedo_table():
    e = np.zeros(1024)
    ea = np.zeros(32)
    _check(lib.mpt_ggx_albedo_table(_dp(e), _dp(ea)))
    return e.reshape(32, 32), ea


class DeviceArray:
    """A numpy-shaped buffer in HBM (mpt_device_alloc), for the entry points
    that take device pointers."""

    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib.mpt_device_alloc(self.nbytes)
        if not self.ptr:
            raise MptError(-3, lib.mpt_last_error().decode())

    @classmethod
    def of(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        _check(lib.mpt_copy_to_device(d.ptr, a.ctypes.data, d.nbytes))
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib.mpt_copy_to_host(out.ctypes.data, self.ptr, self.nbytes))
        return out

    @property
    def handle(self):
        return C.c_void_p(self.ptr)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib.mpt_device_free(self.ptr)
            self.ptr = None


def synchronize():
    _check(lib.mpt_device_synchronize())


def denoise(frame, iterations=5, sigma_luminance=4.0, sigma_normal=0.1, sigma_albedo=0.1):
    """denoiser::run (denoiser.h:45-78) on the GPU. `frame` is a Frame
    rendered with features, or a dict with radiance, albedo, normal (h, w, 3)
    and variance (h, w) float64 arrays. Returns the filtered radiance."""
    get = (lambda k: getattr(frame, k)) if isinstance(frame, Frame) else (lambda k: frame[k])
    raw = [get(k) for k in ("radiance", "albedo", "normal", "variance")]
    if any(a is None for a in raw):
        raise ValueError("denoise needs the frame's albedo, normal and variance (render with features)")
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in raw]
    h, w = arrs[0].shape[:2]
    if arrs[0].shape != (h, w, 3) or arrs[3].shape != (h, w) or any(a.shape != (h, w, 3) for a in arrs[1:3]):
        raise ValueError("denoise needs radiance, albedo, normal (h, w, 3) and variance (h, w)")
    o = abi.DenoiseOpts(iterations=iterations, pad=0, sigma_luminance=sigma_luminance, sigma_normal=sigma_normal,
                        sigma_albedo=sigma_albedo)
    out = np.zeros((h, w, 3))
    _check(lib.mpt_denoise(*[_dp(a) for a in arrs], w, h, C.byref(o), _dp(out)))
    return out


def display_settings(exposure=0.0, tonemap="none", bloom=0.0, bloom_threshold=1.0, bloom_radius=0.04, vignette=0.0):
    """display_settings with its post_settings (tonemap.h:11-37,
    postprocess.h:12-21) as an mpt_display."""
    d = abi.Display()
    lib.mpt_display_defaults(C.byref(d))
    if tonemap not in abi.TONEMAPPERS:
        raise ValueError("tonemap must be none, reinhard or aces")
    d.exposure, d.tonemap = exposure, abi.TONEMAPPERS[tonemap]
    d.bloom, d.bloom_threshold, d.bloom_radius, d.vignette = bloom, bloom_threshold, bloom_radius, vignette
    return d


def scene_display(path):
    """The display settings a .scene file sets (scene.h:667-688)."""
    d = abi.Display()
    _check(lib.mpt_scene_display(os.fsencode(path), C.byref(d)))
    return d


def _image(radiance):
    a = np.ascontiguousarray(radiance, dtype=np.float64)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("an image is an (h, w, 3) array of linear radiance")
    return a


def postprocess(radiance, display):
    """postprocess(pixels, width, height, post) (postprocess.h:108-153) on the
    GPU: bloom and vignetting of an (h, w, 3) radiance image."""
    a = _image(radiance)
    out = np.zeros_like(a)
    _check(lib.mpt_postprocess(_dp(a), a.shape[1], a.shape[0], C.byref(display), _dp(out)))
    return out


def display_bytes(radiance, display, post=True):
    """The sRGB bytes write_png writes (image_io.h:25-37): postprocess (when
    post and bloom/vignette are on), exposure, tonemap, to_byte. (h, w, 3) uint8."""
    a = _image(radiance)
    out = np.zeros(a.shape, dtype=np.uint8)
    _check(lib.mpt_display_bytes(_dp(a), a.shape[1], a.shape[0], C.byref(display), int(post),
                                 out.ctypes.data_as(C.POINTER(C.c_uint8))))
    return out


def write_image(path, radiance, display=None):
    """write_image (image_io.h:81-106): .png/.ppm display images, .hdr/.pfm raw radiance."""
    a = _image(radiance)
    d = display if display is not None else display_settings()
    _check(lib.mpt_write_image(os.fsencode(path) if path else None, a.shape[1], a.shape[0], _dp(a), C.byref(d)))


class Description:
    """A scene flattened on the host (no GPU): the mpt_scene_desc plus the
    camera of the file. Keeps the arrays alive until closed."""

    def __init__(self, path):
        self.desc = SceneDesc()
        self.camera = Camera()
        self._holder = C.c_void_p()
        _check(lib.mpt_scene_describe(os.fsencode(path), C.byref(self.desc), C.byref(self.camera),
                                      C.byref(self._holder)))

    def close(self):
        if self._holder:
            lib.mpt_desc_free(self._holder)
            self._holder = C.c_void_p()

    def __del__(self):
        self.close()


def describe(path):
    return Description(path)


def render_batch(scenes, cameras=None, seed=1, first_pass=0, scene_index=0, pass_count=0, accumulate=False,
                 passes_in_flight=0, profile=0, tail_paths=0, blocks_per_cu=0, read=True):
    """mpt_render_batch: the scenes rendered together, every stage launch
    serving all of them; scene k keyed with scene index scene_index + k.
    Returns (frames, stats): frames[k] is scene k's averaged radiance (or
    None with read=False; read it later with Scene.read_frame)."""
    scenes = list(scenes)
    cams = [abi.camera_copy(c if c is not None else s.camera)
            for s, c in zip(scenes, cameras if cameras is not None else [None] * len(scenes))]
    n = len(scenes)
    handles = (C.c_void_p * n)(*[s._h for s in scenes])
    cam_arr = (Camera * n)(*cams)
    opts = RenderOpts(seed=seed, first_pass=first_pass, scene_index=scene_index, pass_count=pass_count,
                      accumulate=int(bool(accumulate)), passes_in_flight=passes_in_flight, profile=profile,
                      tail_paths=tail_paths, blocks_per_cu=blocks_per_cu)
    stats = RenderStats()
    _check(lib.mpt_render_batch(handles, cam_arr, n, C.byref(opts), C.byref(stats)))
    frames = [s.read_frame(c) for s, c in zip(scenes, cams)] if read else [None] * n
    return frames, stats


class Frame:
    """The averaged frame camera::render returns (denoiser.h:14-20)."""

    def __init__(self, radiance, albedo=None, normal=None, variance=None, stats=None):
        self.radiance, self.albedo, self.normal, self.variance = radiance, albedo, normal, variance
        self.stats = stats


class Scene:
    """A scene resident in HBM (mpt_scene)."""

    def __init__(self, handle, camera=None, keepalive=None):
        self._h = handle
        self.camera = camera
        self._keepalive = keepalive

    @classmethod
    def load(cls, path):
        h = C.c_void_p()
        cam = Camera()
        _check(lib.mpt_scene_load(os.fsencode(path), C.byref(h), C.byref(cam)))
        return cls(h, cam)

    @classmethod
    def from_desc(cls, desc, camera=None, keepalive=None):
        h = C.c_void_p()
        _check(lib.mpt_scene_create(C.byref(desc), C.byref(h)))
        return cls(h, camera, keepalive)

    def close(self):
        if self._h:
            lib.mpt_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def counts(self):
        np_, no, nt, nb, nl = C.c_int64(), C.c_int32(), C.c_int64(), C.c_int64(), C.c_int32()
        _check(lib.mpt_scene_counts(self._h, C.byref(np_), C.byref(no), C.byref(nt), C.byref(nb), C.byref(nl)))
        return {"prims": np_.value, "objects": no.value, "tlas_nodes": nt.value, "blas_nodes": nb.value,
                "lights": nl.value}

    def light_probabilities(self):
        n = self.counts()["lights"]
        out = np.zeros(max(n, 1))
        _check(lib.mpt_scene_light_probabilities(self._h, _dp(out)))
        return out[:n]

    def trace_closest(self, rays, tt):
        """bvh::hit for each ray: (canonical id or -1, t, front_face)."""
        rays = np.ascontiguousarray(rays, dtype=np.float64)
        tt = np.ascontiguousarray(tt, dtype=np.float64)
        n = rays.shape[0]
        assert rays.shape == (n, 6) and tt.shape == (n, 2)
        ids = np.zeros(n, dtype=np.int64)
        t = np.zeros(n)
        front = np.zeros(n, dtype=np.uint8)
        _check(lib.mpt_trace_closest(self._h, _dp(rays), _dp(tt), n, ids.ctypes.data_as(C.POINTER(C.c_int64)),
                                     _dp(t), front.ctypes.data_as(C.POINTER(C.c_uint8))))
        return ids, t, front

    def trace_ties(self, rays, tt):
        """trace_closest plus the grazing ties: a dict of id, t, front, ties
        (candidates with the closest t bitwise), tie_min and tie_max."""
        rays = np.ascontiguousarray(rays, dtype=np.float64)
        tt = np.ascontiguousarray(tt, dtype=np.float64)
        n = rays.shape[0]
        assert rays.shape == (n, 6) and tt.shape == (n, 2)
        out = {"id": np.zeros(n, dtype=np.int64), "t": np.zeros(n), "front": np.zeros(n, dtype=np.uint8),
               "ties": np.zeros(n, dtype=np.int32), "tie_min": np.zeros(n, dtype=np.int64),
               "tie_max": np.zeros(n, dtype=np.int64)}
        i64 = C.POINTER(C.c_int64)
        _check(lib.mpt_trace_ties(self._h, _dp(rays), _dp(tt), n, out["id"].ctypes.data_as(i64), _dp(out["t"]),
                                  out["front"].ctypes.data_as(C.POINTER(C.c_uint8)),
                                  out["ties"].ctypes.data_as(C.POINTER(C.c_int32)), out["tie_min"].ctypes.data_as(i64),
                                  out["tie_max"].ctypes.data_as(i64)))
        return out

    def bsdf_eval(self, mats, q, front):
        mats = np.ascontiguousarray(mats, dtype=np.int32)
        q = np.ascontiguousarray(q, dtype=np.float64)
        front = np.ascontiguousarray(front, dtype=np.uint8)
        n = mats.shape[0]
        ev = np.zeros((n, 3))
        pdf = np.zeros(n)
        _check(lib.mpt_bsdf_eval(self._h, mats.ctypes.data_as(C.POINTER(C.c_int32)), _dp(q),
                                 front.ctypes.data_as(C.POINTER(C.c_uint8)), n, _dp(ev), _dp(pdf)))
        return ev, pdf

    def render(self, camera=None, seed=1, features=None, passes_in_flight=0, save_every=0, on_pass=None,
               first_pass=0, scene_index=0, pass_count=0, accumulate=False, profile=0, read=True, tail_paths=0,
               blocks_per_cu=0, sort_by_material=0, use_graph=0, guide_passes=0):
        cam = abi.camera_copy(camera if camera is not None else self.camera)
        if features is not None:
            cam.collect_features = int(bool(features))
        w, h = cam.image_width, cam.height()
        rad = np.zeros((h, w, 3))
        feats = bool(cam.collect_features)
        alb = np.zeros((h, w, 3)) if feats else None
        nrm = np.zeros((h, w, 3)) if feats else None
        var = np.zeros((h, w)) if feats else None
        opts = RenderOpts(seed=seed, first_pass=first_pass, save_every=save_every,
                          passes_in_flight=passes_in_flight, scene_index=scene_index, pass_count=pass_count,
                          accumulate=int(bool(accumulate)), profile=profile, tail_paths=tail_paths,
                          blocks_per_cu=blocks_per_cu, sort_by_material=sort_by_material, use_graph=use_graph,
                          guide_passes=guide_passes)
        stats = RenderStats()
        cb = abi.PASS_CALLBACK(lambda done, user: on_pass(done, self)) if on_pass else abi.PASS_CALLBACK()
        if not read:  # leave the frame in HBM (read it later with read_frame)
            _check(lib.mpt_render(self._h, C.byref(cam), C.byref(opts), None, None, None, None, cb, None,
                                  C.byref(stats)))
            return Frame(None, stats=stats)
        _check(lib.mpt_render(self._h, C.byref(cam), C.byref(opts), _dp(rad),
                              _dp(alb) if feats else None, _dp(nrm) if feats else None,
                              _dp(var) if feats else None, cb, None, C.byref(stats)))
        return Frame(rad, alb, nrm, var, stats)

    _TUNED = ("passes_in_flight", "tail_paths", "blocks_per_cu", "sort_by_material", "use_graph")

    def autotune(self, camera=None, trial_passes=8, cache=None):
        """mpt_autotune: times batch size, tail size, grid, material binning and
        bounce graphs for this image size and keeps the fastest. With `cache` (a
        JSON file path) results are stored there keyed by device, kernel profile
        and image size, and a stored result is applied without timing."""
        import json
        cam = abi.camera_copy(camera if camera is not None else self.camera)
        key = None
        if cache:
            buf = C.create_string_buffer(512)
            _check(lib.mpt_autotune_key(self._h, C.byref(cam), buf, 512))
            key = buf.value.decode()
            table = json.load(open(cache)) if os.path.exists(cache) else {}
            if key in table:
                r = abi.AutotuneResult(**{k: table[key][k] for k in self._TUNED})
                _check(lib.mpt_autotune_apply(self._h, C.byref(cam), C.byref(r)))
                return dict(table[key], cached=True)
        r = abi.AutotuneResult()
        _check(lib.mpt_autotune(self._h, C.byref(cam), trial_passes, C.byref(r)))
        out = {k: getattr(r, k) for k in self._TUNED}
        out.update(trials=r.trials, seconds_per_pass=r.seconds_per_pass)
        if cache:
            table = json.load(open(cache)) if os.path.exists(cache) else {}
            table[key] = out
            tmp = cache + ".tmp"
            with open(tmp, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return dict(out, cached=False)

    def primary_depth(self, camera=None):
        """Depth of the first surface through every pixel centre (0 for the sky)."""
        cam = abi.camera_copy(camera if camera is not None else self.camera)
        d = np.zeros((cam.height(), cam.image_width))
        _check(lib.mpt_primary_depth(self._h, C.byref(cam), _dp(d)))
        return d

    def primary_depth_dev(self, camera=None):
        """primary_depth traced on the GPU into a DeviceArray (h, w)."""
        cam = abi.camera_copy(camera if camera is not None else self.camera)
        d = DeviceArray((cam.height(), cam.image_width), np.float64)
        _check(lib.mpt_primary_depth_dev(self._h, C.byref(cam), d.handle, None))
        return d

    def frame_dev(self, camera, features=True, depth=True):
        """The last render's averaged frame as DeviceArrays (radiance, and with
        features albedo, normal, variance; with depth the primary depth), the
        inputs of the learned denoiser without a host round trip."""
        w, h = camera.image_width, camera.height()
        out = {"radiance": DeviceArray((h, w, 3), np.float64)}
        if features:
            out.update(albedo=DeviceArray((h, w, 3), np.float64), normal=DeviceArray((h, w, 3), np.float64),
                       variance=DeviceArray((h, w), np.float64))
        _check(lib.mpt_read_frame_dev(self._h, out["radiance"].handle, out["albedo"].handle if features else None,
                                      out["normal"].handle if features else None,
                                      out["variance"].handle if features else None, None))
        if depth:
            out["depth"] = self.primary_depth_dev(camera)
        return out

    def read_frame(self, camera):
        w, h = camera.image_width, camera.height()
        rad = np.zeros((h, w, 3))
        _check(lib.mpt_read_frame(self._h, _dp(rad), None, None, None))
        return rad
