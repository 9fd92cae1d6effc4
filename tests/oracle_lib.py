"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, the compiled reference harness (oracle/_ref/libmpt_ref.so).

Test infrastructure: the checker, never the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

from mypbr_amd.abi import Camera, SceneDesc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libmpt_ref.so")

P = C.POINTER
vp = C.c_void_p
dp = P(C.c_double)


def _d(a):
    return a.ctypes.data_as(dp)


def _i64(a):
    return a.ctypes.data_as(P(C.c_int64))


def _u8(a):
    return a.ctypes.data_as(P(C.c_uint8))


def _i32(a):
    return a.ctypes.data_as(P(C.c_int32))


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_scene_create": (C.c_int32, [P(SceneDesc), P(vp)]),
            "orc_scene_destroy": (None, [vp]),
            "orc_trace": (C.c_int32, [vp, dp, dp, C.c_int64, P(C.c_int64), dp, P(C.c_uint8), C.c_int32]),
            "orc_trace_linear": (C.c_int32, [vp, dp, dp, C.c_int64, P(C.c_int64), dp]),
            "orc_trace_ties": (C.c_int32, [vp, dp, dp, C.c_int64, P(C.c_int64), dp, P(C.c_int32), P(C.c_int64),
                                           P(C.c_int64)]),
            "orc_surface": (C.c_int32, [vp, dp, dp, C.c_int64, dp]),
            "orc_light_probabilities": (C.c_int32, [vp, dp]),
            "orc_ggx_table": (C.c_int32, [dp, dp]),
            "orc_bsdf_eval": (C.c_int32, [vp, P(C.c_int32), dp, P(C.c_uint8), C.c_int64, dp, dp]),
            "orc_render": (C.c_int32, [vp, P(Camera), C.c_uint64, C.c_uint32, C.c_int32, C.c_int32, C.c_int32,
                                       dp, dp, dp, dp, P(C.c_int64), P(C.c_int64), dp]),
            "orc_make_ray_batch": (C.c_int32, [vp, P(Camera), C.c_uint64, C.c_int64, dp, dp]),
            "orc_denoise": (C.c_int32, [dp, dp, dp, dp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                        C.c_double, dp]),
            "orc_postprocess": (C.c_int32, [dp, C.c_int32, C.c_int32] + [C.c_double] * 4 + [dp]),
            "orc_display_bytes": (C.c_int32, [dp, C.c_int32, C.c_int32, C.c_double, C.c_int32] + [C.c_double] * 4 +
                                  [C.c_int32, P(C.c_uint8)]),
            "orc_to_byte": (C.c_int32, [dp, C.c_int64, P(C.c_uint8)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _oracle = lib
    return _oracle


class OracleScene:
    def __init__(self, desc):
        self.lib = oracle()
        self.h = vp()
        rc = self.lib.orc_scene_create(C.byref(desc), C.byref(self.h))
        if rc != 0:
            raise RuntimeError("orc_scene_create: " + self.lib.orc_last_error().decode())

    def close(self):
        if self.h:
            self.lib.orc_scene_destroy(self.h)
            self.h = vp()

    def __del__(self):
        self.close()

    def trace(self, rays, tt, threads=8):
        n = rays.shape[0]
        ids = np.zeros(n, dtype=np.int64)
        t = np.zeros(n)
        f = np.zeros(n, dtype=np.uint8)
        self.lib.orc_trace(self.h, _d(rays), _d(tt), n, _i64(ids), _d(t), _u8(f), threads)
        return ids, t, f

    def surface(self, rays, tt):
        """Per ray: u, v, dp/du, dp/dv, base texture value, shading normal,
        metallic, alpha (16 doubles; NaN on a miss)."""
        out = np.zeros((rays.shape[0], 16))
        self.lib.orc_surface(self.h, _d(rays), _d(tt), rays.shape[0], _d(out))
        return out

    def trace_linear(self, rays, tt):
        n = rays.shape[0]
        ids = np.zeros(n, dtype=np.int64)
        t = np.zeros(n)
        self.lib.orc_trace_linear(self.h, _d(rays), _d(tt), n, _i64(ids), _d(t))
        return ids, t

    def trace_ties(self, rays, tt):
        """Closest hit plus the candidates with its t bitwise (ties, tie_min, tie_max)."""
        n = rays.shape[0]
        out = {"id": np.zeros(n, dtype=np.int64), "t": np.zeros(n), "ties": np.zeros(n, dtype=np.int32),
               "tie_min": np.zeros(n, dtype=np.int64), "tie_max": np.zeros(n, dtype=np.int64)}
        self.lib.orc_trace_ties(self.h, _d(rays), _d(tt), n, _i64(out["id"]), _d(out["t"]), _i32(out["ties"]),
                                _i64(out["tie_min"]), _i64(out["tie_max"]))
        return out

    def light_probabilities(self, n):
        out = np.zeros(max(n, 1))
        self.lib.orc_light_probabilities(self.h, _d(out))
        return out[:n]

    def bsdf_eval(self, mats, q, front):
        mats = np.ascontiguousarray(mats, dtype=np.int32)
        q = np.ascontiguousarray(q, dtype=np.float64)
        front = np.ascontiguousarray(front, dtype=np.uint8)
        n = mats.shape[0]
        ev = np.zeros((n, 3))
        pdf = np.zeros(n)
        self.lib.orc_bsdf_eval(self.h, _i32(mats), _d(q), _u8(front), n, _d(ev), _d(pdf))
        return ev, pdf

    def ray_batch(self, cam, seed, n):
        rays = np.zeros((n, 6))
        tt = np.zeros((n, 2))
        self.lib.orc_make_ray_batch(self.h, C.byref(cam), seed, n, _d(rays), _d(tt))
        return rays, tt

    def render_sums(self, cam, seed=1, first_pass=0, passes=None, threads=8, scene_index=0):
        """Sums (not averages) over passes, as the accumulator of camera.h:61-65."""
        w, h = cam.image_width, cam.height()
        passes = cam.samples_per_pixel if passes is None else passes
        rad = np.zeros((h, w, 3))
        alb = np.zeros((h, w, 3))
        nrm = np.zeros((h, w, 3))
        l2 = np.zeros((h, w))
        ext = C.c_int64()
        sh = C.c_int64()
        secs = C.c_double()
        self.lib.orc_render(self.h, C.byref(cam), seed, scene_index, first_pass, passes, threads, _d(rad), _d(alb),
                            _d(nrm), _d(l2), C.byref(ext), C.byref(sh), C.byref(secs))
        return {"radiance": rad, "albedo": alb, "normal": nrm, "l2": l2, "extension_rays": ext.value,
                "shadow_rays": sh.value, "seconds": secs.value}

    def render(self, cam, seed=1, threads=8):
        """camera::render's averaged frame (camera.h:223-246)."""
        s = self.render_sums(cam, seed=seed, threads=threads)
        n = cam.samples_per_pixel
        scale = 1.0 / n
        out = {"radiance": scale * s["radiance"], "stats": s}
        if cam.collect_features:
            out["albedo"] = scale * s["albedo"]
            out["normal"] = scale * s["normal"]
            rad = out["radiance"]
            mean = 0.2126 * rad[..., 0] + 0.7152 * rad[..., 1] + 0.0722 * rad[..., 2]
            spread = np.fmax(scale * s["l2"] - mean * mean, 0.0)
            out["variance"] = spread / (n - 1) if n > 1 else spread
        return out


_ref = {}
# timing-only builds of the reference harness with the reference's Release flags (oracle/Makefile ref_timing)
TIMING_REFS = [("x86-64-v4", os.path.join(ROOT, "oracle", "_ref", "libmpt_ref_time_v4.so"),
                ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")),
               ("x86-64-v3", os.path.join(ROOT, "oracle", "_ref", "libmpt_ref_time_v3.so"),
                ("avx2", "fma", "bmi2"))]
TIMING_FLAGS = "-O3 -DNDEBUG -std=c++17 -mtune=znver3, gcc default FP contraction"


def reference_available():
    return os.path.exists(REF)


def _cpu_flags():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def timing_reference():
    """(library, -march level, flags) of the fastest timing build of the
    reference the host can run, or None when none is built."""
    have = _cpu_flags()
    for march, path, need in TIMING_REFS:
        if os.path.exists(path) and all(f in have for f in need):
            return reference(path), march, f"-march={march} {TIMING_FLAGS}"
    return None


def cpu_description():
    """CPU model, logical CPUs and physical cores of this host (from /proc/cpuinfo)."""
    model, logical, cores = "unknown", 0, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                    logical += 1
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "logical_cpus": logical, "physical_cores": len(cores) or logical}


def reference(path=REF):
    if path not in _ref:
        lib = C.CDLL(path)
        sig = {
            "ref_load": (C.c_int, [C.c_char_p, P(vp)]),
            "ref_free": (None, [vp]),
            "ref_canon_count": (C.c_int64, [vp]),
            "ref_surface": (C.c_int, [vp, dp, dp, C.c_int64, dp]),
            "ref_load_image_u8": (C.c_int, [C.c_char_p, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_uint8), C.c_int64]),
            "ref_write_jpg": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, P(C.c_uint8), C.c_int]),
            "ref_trace": (C.c_int, [vp, dp, dp, C.c_int64, P(C.c_int64), dp, P(C.c_uint8), P(C.c_int32),
                                    P(C.c_int64), P(C.c_int64), C.c_int]),
            "ref_light_count": (C.c_int, [vp]),
            "ref_light_probabilities": (C.c_int, [vp, dp]),
            "ref_env_radiance": (C.c_int, [vp, dp, C.c_int64, dp]),
            "ref_ggx_table": (C.c_int, [dp, dp]),
            "ref_bsdf_eval": (C.c_int, [dp, dp, P(C.c_uint8), C.c_int64, dp, dp]),
            "ref_render": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp]),
            "ref_denoise": (C.c_int, [dp, dp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      dp]),
            "ref_postprocess": (C.c_int, [dp, C.c_int, C.c_int] + [C.c_double] * 4 + [dp]),
            "ref_display_bytes": (C.c_int, [dp, C.c_int, C.c_int, C.c_double, C.c_int] + [C.c_double] * 4 +
                                  [C.c_int, P(C.c_uint8)]),
            "ref_write_image": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, C.c_double, C.c_int] + [C.c_double] * 4),
            "ref_scene_display": (C.c_int, [C.c_char_p, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _ref[path] = lib
    return _ref[path]


class RefScene:
    def __init__(self, path, lib=None):
        self.lib = lib or reference()
        self.h = vp()
        if self.lib.ref_load(os.fsencode(path), C.byref(self.h)) != 0:
            raise RuntimeError("reference load_scene failed for " + path)

    def close(self):
        if self.h:
            self.lib.ref_free(self.h)
            self.h = vp()

    def __del__(self):
        self.close()

    def surface(self, rays, tt):
        out = np.zeros((rays.shape[0], 16))
        self.lib.ref_surface(self.h, _d(rays), _d(tt), rays.shape[0], _d(out))
        return out

    def trace(self, rays, tt, threads=8):
        n = rays.shape[0]
        ids = np.zeros(n, dtype=np.int64)
        t = np.zeros(n)
        f = np.zeros(n, dtype=np.uint8)
        ties = np.zeros(n, dtype=np.int32)
        lo = np.zeros(n, dtype=np.int64)
        hi = np.zeros(n, dtype=np.int64)
        self.lib.ref_trace(self.h, _d(rays), _d(tt), n, _i64(ids), _d(t), _u8(f), _i32(ties), _i64(lo), _i64(hi),
                           threads)
        return {"id": ids, "t": t, "front": f, "ties": ties, "tie_min": lo, "tie_max": hi}

    def env_radiance(self, dirs):
        dirs = np.ascontiguousarray(dirs, dtype=np.float64)
        out = np.zeros((len(dirs), 3))
        if self.lib.ref_env_radiance(self.h, _d(dirs), len(dirs), _d(out)) != 0:
            raise RuntimeError("the reference scene has no environment")
        return out

    def light_probabilities(self):
        n = self.lib.ref_light_count(self.h)
        out = np.zeros(max(n, 1))
        self.lib.ref_light_probabilities(self.h, _d(out))
        return out[:n]

    def render(self, width, spp, max_depth, threads=8):
        # height from the scene's own aspect ratio, as camera::initialize
        return None


def _denoise(fn, frame, settings):
    rad = np.ascontiguousarray(frame["radiance"], dtype=np.float64)
    h, w = rad.shape[:2]
    out = np.zeros((h, w, 3))
    it, sl, sn, sa = settings
    rc = fn(_d(rad), _d(np.ascontiguousarray(frame["albedo"], dtype=np.float64)),
            _d(np.ascontiguousarray(frame["normal"], dtype=np.float64)),
            _d(np.ascontiguousarray(frame["variance"], dtype=np.float64)), w, h, it, sl, sn, sa, _d(out))
    if rc != 0:
        raise RuntimeError("denoise failed")
    return out


def oracle_denoise(frame, settings=(5, 4.0, 0.1, 0.1)):
    """The oracle's restatement of denoiser::run (denoiser.h:45-78)."""
    return _denoise(oracle().orc_denoise, frame, settings)


def reference_denoise(frame, settings=(5, 4.0, 0.1, 0.1)):
    """The reference's own denoiser::run, through the compiled harness."""
    return _denoise(reference().ref_denoise, frame, settings)


# ---- display pipeline (postprocess.h, tonemap.h, color3.h)

def _img(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def oracle_postprocess(rad, setting):
    _, _, bloom, thr, radius, vig = setting
    a = _img(rad)
    out = np.zeros_like(a)
    oracle().orc_postprocess(_d(a), a.shape[1], a.shape[0], bloom, thr, radius, vig, _d(out))
    return out


def oracle_display_bytes(rad, setting, post=True):
    exposure, op, bloom, thr, radius, vig = setting
    a = _img(rad)
    out = np.zeros(a.shape, dtype=np.uint8)
    oracle().orc_display_bytes(_d(a), a.shape[1], a.shape[0], exposure, op, bloom, thr, radius, vig, int(post), _u8(out))
    return out


def oracle_to_byte(x):
    x = _img(x).ravel()
    out = np.zeros(x.shape, dtype=np.uint8)
    oracle().orc_to_byte(_d(x), x.size, _u8(out))
    return out


def reference_postprocess(rad, setting):
    _, _, bloom, thr, radius, vig = setting
    a = _img(rad)
    out = np.zeros_like(a)
    reference().ref_postprocess(_d(a), a.shape[1], a.shape[0], bloom, thr, radius, vig, _d(out))
    return out


def reference_display_bytes(rad, setting, post=True):
    exposure, op, bloom, thr, radius, vig = setting
    a = _img(rad)
    out = np.zeros(a.shape, dtype=np.uint8)
    reference().ref_display_bytes(_d(a), a.shape[1], a.shape[0], exposure, op, bloom, thr, radius, vig, int(post),
                                  _u8(out))
    return out


def reference_write_image(path, rad, setting):
    exposure, op, bloom, thr, radius, vig = setting
    a = _img(rad)
    return reference().ref_write_image(os.fsencode(path), a.shape[1], a.shape[0], _d(a), exposure, op, bloom, thr,
                                       radius, vig)


def reference_scene_display(path):
    out = np.zeros(6)
    if reference().ref_scene_display(os.fsencode(path), _d(out)) != 0:
        return None
    return out


def reference_image_u8(path):
    """stb_image's stbi_load(path, ..., 4) through the compiled reference:
    RGBA bytes (h, w, 4) and the file's channel count."""
    R = reference()
    w, h, c = C.c_int(), C.c_int(), C.c_int()
    buf = np.zeros(4096 * 4096 * 4, np.uint8)
    if R.ref_load_image_u8(os.fsencode(path), C.byref(w), C.byref(h), C.byref(c), buf.ctypes.data_as(P(C.c_uint8)),
                           buf.size) != 0:
        raise RuntimeError("stbi_load failed for " + path)
    return buf[: w.value * h.value * 4].reshape(h.value, w.value, 4).copy(), c.value


def reference_write_jpg(path, img, quality):
    """stbi_write_jpg of an (h, w, comp) uint8 image through the compiled reference."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape[:2]
    comp = 1 if img.ndim == 2 else img.shape[2]
    if reference().ref_write_jpg(os.fsencode(path), w, h, comp, img.ctypes.data_as(P(C.c_uint8)), quality) != 0:
        raise RuntimeError("stbi_write_jpg failed")
