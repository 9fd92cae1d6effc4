"""The learned denoiser (include/mpt.h, mpt_kpcnn_*): a kernel-predicting
CNN whose forward and backward passes run on the GPU's BF16 matrix cores
(mypbr_amd/csrc/learn.hip). Device buffers are libmpt allocations
(DeviceArray)."""
import ctypes as C

import numpy as np

from . import DeviceArray, _check, lib

FRAME_KEYS = ("radiance", "albedo", "normal", "variance", "depth")


def _vp(t):
    return t.handle if t is not None else None


def to_device(frames):
    """Stacks frames (dicts of numpy arrays: radiance, albedo, normal (h, w, 3),
    variance, depth (h, w)) into float64 device arrays (B, h, w[, 3])."""
    return {k: DeviceArray.of(np.stack([np.asarray(f[k], dtype=np.float64) for f in frames])) for k in FRAME_KEYS}


class KPCNN:
    """mpt_kpcnn: `layers` 3x3 convolutions of `hidden` channels predicting a
    5x5 kernel per pixel."""

    def __init__(self, layers=6, hidden=64, seed=1):
        self._h = C.c_void_p()
        _check(lib.mpt_kpcnn_create(layers, hidden, seed, C.byref(self._h)))
        self.layers, self.hidden = layers, hidden
        self.nparams = lib.mpt_kpcnn_param_count(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib.mpt_kpcnn_destroy(self._h)
            self._h = None

    def set_loss(self, mode):
        """0: relMSE (o - t)^2 / (t^2 + 0.01); 1: L1 of log(1 + x)."""
        _check(lib.mpt_kpcnn_set_loss(self._h, int(mode)))

    def layout(self):
        n = self.layers
        cin = (C.c_int32 * n)()
        cout = (C.c_int32 * n)()
        wo = (C.c_int64 * n)()
        bo = (C.c_int64 * n)()
        _check(lib.mpt_kpcnn_layout(self._h, cin, cout, wo, bo))
        return list(cin), list(cout), list(wo), list(bo)

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(lib.mpt_kpcnn_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), None))
        return p

    def grads(self):
        g = np.zeros(self.nparams, dtype=np.float32)
        _check(lib.mpt_kpcnn_get_params(self._h, None, g.ctypes.data_as(C.POINTER(C.c_float))))
        return g

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib.mpt_kpcnn_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def run(self, dev, target=None, lr=0.0, out=None):
        """One pass over the device batch `dev` (from to_device). With a target
        (float32 DeviceArray (B, h, w, 3)) returns the loss and leaves the
        gradients (and with lr > 0 takes an Adam step)."""
        B, H, W = dev["radiance"].shape[:3]
        loss = C.c_double(0)
        _check(lib.mpt_kpcnn_run(self._h, *[_vp(dev[k]) for k in FRAME_KEYS], B, H, W, _vp(target), float(lr),
                                 _vp(out), C.byref(loss)))
        return loss.value

    def denoise_dev(self, dev):
        """Denoised radiance (float32 DeviceArray (h, w, 3)) of one frame that
        is already in HBM (Scene.frame_dev): no host round trip."""
        h, w = dev["radiance"].shape[:2]
        out = DeviceArray((h, w, 3), np.float32)
        loss = C.c_double(0)
        _check(lib.mpt_kpcnn_run(self._h, *[_vp(dev[k]) for k in FRAME_KEYS], 1, h, w, None, 0.0, _vp(out),
                                 C.byref(loss)))
        return out

    def denoise(self, frames):
        """Denoised radiance (numpy, (B, h, w, 3) float32) of host frames."""
        dev = to_device(frames)
        out = DeviceArray(dev["radiance"].shape, np.float32)
        self.run(dev, out=out)
        return out.numpy()
