"""The allocation-free display entry point (mpt_display_bytes_dev_noalloc):
the bloom pyramid in the library's fixed device array instead of a
workspace argument. Its bytes must equal mpt_display_bytes_dev's (and the
oracle's) call after call: the second call is smaller than the first, so a
kernel that read texels the first call left behind would show. Sizes whose
pyramid exceeds the fixed array are refused on the host before any GPU
work."""
import ctypes as C

import numpy as np
import pytest

import display_cases as dc
import oracle_lib as ol


def _display(m, st):
    names = {0: "none", 1: "reinhard", 2: "aces"}
    return m.display_settings(exposure=st[0], tonemap=names[st[1]], bloom=st[2], bloom_threshold=st[3],
                              bloom_radius=st[4], vignette=st[5])


def test_noalloc_refuses_pyramids_beyond_its_array():
    import mypbr_amd as m
    limit = m.lib.mpt_display_noalloc_max_workspace()
    assert limit == 48 << 20
    d = _display(m, dc.SETTINGS[2])
    w, h = 4096, 2304
    assert m.lib.mpt_display_workspace_bytes(w, h, C.byref(d)) > limit
    one = (C.c_uint8 * 1)()
    rc = m.lib.mpt_display_bytes_dev_noalloc(C.cast(one, C.c_void_p), w, h, C.byref(d), 1,
                                             C.cast(one, C.c_void_p), None)
    assert rc == m.abi.MPT_E_CAPACITY
    assert "fixed scratch" in m.lib.mpt_last_error().decode()
    # 1920 x 1080 at the default radius fits
    assert m.lib.mpt_display_workspace_bytes(1920, 1080, C.byref(d)) <= limit


@pytest.mark.gpu
def test_gpu_noalloc_matches_workspace_entry_point():
    import mypbr_amd as m
    for si in (2, 0, 3):  # bloom + vignette + ACES; bloom only; no bloom (no pyramid)
        st = dc.SETTINGS[si]
        d = _display(m, st)
        for w, h, seed in ((1920, 1080, 41), (333, 187, 42)):  # the second call is smaller
            print(f"noalloc display: settings {si}, {w}x{h}", flush=True)
            img = dc.sky_sun(w, h, seed=seed)
            src = m.DeviceArray.of(img)
            ws_bytes = m.lib.mpt_display_workspace_bytes(w, h, C.byref(d))
            ws = m.DeviceArray((max(ws_bytes, 1),), np.uint8)
            ref = m.DeviceArray((h, w, 3), np.uint8)
            got = m.DeviceArray((h, w, 3), np.uint8)
            m._check(m.lib.mpt_display_bytes_dev(src.handle, w, h, C.byref(d), 1, ref.handle,
                                                 ws.handle if ws_bytes else None, None))
            m._check(m.lib.mpt_display_bytes_dev_noalloc(src.handle, w, h, C.byref(d), 1, got.handle, None))
            m.synchronize()
            a, b = got.numpy(), ref.numpy()
            assert np.array_equal(a, b), (si, w, h, int(np.count_nonzero(a != b)))
            assert np.array_equal(a, ol.oracle_display_bytes(img, st)), (si, w, h)
