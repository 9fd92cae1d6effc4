"""The display pipeline: postprocess (postprocess.h:108-153),
display_settings::apply (tonemap.h:28-67), to_byte (color3.h:17-33) and
write_image (image_io.h:16-106).

CPU: the oracle's restatement against the reference's own outputs
(tests/golden/display.npz, written by the compiled reference) and, where it
is built, against the compiled reference on fresh images; closed forms; the
premise of the GPU's byte encoding (to_byte steps at 255 thresholds); the
.scene display lines; the raw-radiance writers (.pfm, .hdr), which run on
the host. GPU: the HIP kernels against the oracle and the golden outputs,
bit for bit (every transcendental the pipeline uses is evaluated on the
host), at fixture sizes and at 1080p; the display-image writers (.ppm,
.png) and the device entry points."""
import os
import struct
import zlib

import numpy as np
import pytest

import display_cases as dc
import oracle_lib as ol
from conftest import ROOT, scene_path

GOLDEN = os.path.join(ROOT, "tests", "golden", "display.npz")


def _golden():
    g = np.load(GOLDEN)
    settings = [tuple(x) for x in g["settings"]]
    return g, [(s[0], int(s[1]), s[2], s[3], s[4], s[5]) for s in settings]


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _display(m, st):
    names = {0: "none", 1: "reinhard", 2: "aces"}
    return m.display_settings(exposure=st[0], tonemap=names[st[1]], bloom=st[2], bloom_threshold=st[3],
                              bloom_radius=st[4], vignette=st[5])


# ---------------------------------------------------------------- CPU: the oracle

@pytest.mark.parametrize("name", list(dc.CASES))
def test_oracle_matches_reference_outputs(name):
    g, settings = _golden()
    img = g[f"{name}/radiance"]
    assert settings == dc.SETTINGS
    for si, st in enumerate(settings):
        assert np.array_equal(_bits(ol.oracle_postprocess(img, st)), _bits(g[f"{name}/post{si}"])), (name, si)
        assert np.array_equal(ol.oracle_display_bytes(img, st), g[f"{name}/bytes{si}"]), (name, si)
    assert np.array_equal(ol.oracle_display_bytes(img, settings[2], post=False), g[f"{name}/bytes_nopost"])


@pytest.mark.skipif(not ol.reference_available(), reason="the compiled reference (oracle/_ref) is not built")
def test_oracle_matches_reference_on_fresh_images():
    rng = np.random.default_rng(11)
    imgs = [dc.sky_sun(150, 91, seed=21), rng.lognormal(-0.5, 2.0, (33, 70, 3)), dc.heavy(5, 9, seed=22)]
    for img in imgs:
        for st in dc.SETTINGS:
            assert np.array_equal(_bits(ol.oracle_postprocess(img, st)), _bits(ol.reference_postprocess(img, st)))
            assert np.array_equal(ol.oracle_display_bytes(img, st), ol.reference_display_bytes(img, st))


def test_closed_forms():
    rng = np.random.default_rng(2)
    img = rng.lognormal(0, 1, (17, 29, 3))
    # no bloom, no vignette: the identity
    assert np.array_equal(_bits(ol.oracle_postprocess(img, (0, 0, 0.0, 1.0, 0.04, 0.0))), _bits(img))
    # a uniform image: every level and its blur are the bright part, so the glow is bloom * bright
    c = np.array([3.0, 2.0, 1.0])
    uni = np.broadcast_to(c, (32, 48, 3)).copy()
    l = 0.2126 * 3 + 0.7152 * 2 + 0.0722 * 1
    out = ol.oracle_postprocess(uni, (0, 0, 0.5, 1.0, 0.2, 0.0))
    np.testing.assert_allclose(out, uni + 0.5 * c * (l - 1.0) / l, rtol=1e-12)
    # vignette: 1 at the exact centre of an odd-sized image, 1 / (1 + k)^2 towards the corners
    ones = np.ones((9, 11, 3))
    v = ol.oracle_postprocess(ones, (0, 0, 0.0, 1.0, 0.04, 0.6))
    assert np.array_equal(v[4, 5], [1.0, 1.0, 1.0])
    cx, cy = 5.5, 4.5
    r2 = (0.5 - cx) ** 2 + (0.5 - cy) ** 2
    np.testing.assert_allclose(v[0, 0], 1 / (1 + 0.6 * r2 / (cx * cx + cy * cy)) ** 2, rtol=1e-14)
    # to_byte's fixed points (color3.h:17-33)
    assert list(ol.oracle_to_byte(np.array([-1.0, 0.0, 0.0031308, 1.0, 1e300, np.nan]))) == [0, 0, 10, 255, 255, 0]


def _thresholds(to_byte):
    """T[k], k = 1..255: the smallest positive double to_byte maps to k or above (bisection on the bit patterns)."""
    lo = np.zeros(255, dtype=np.uint64)
    hi = np.full(255, np.float64(np.inf).view(np.uint64), dtype=np.uint64)
    k = np.arange(1, 256)
    while True:
        open_ = hi - lo > 1
        if not open_.any():
            return hi.view(np.float64)
        mid = lo + (hi - lo) // np.uint64(2)
        up = to_byte(mid.view(np.float64)).astype(int) >= k
        hi = np.where(open_ & up, mid, hi)
        lo = np.where(open_ & ~up, mid, lo)


def test_byte_thresholds_reproduce_to_byte():
    """The GPU encodes a byte by counting the thresholds at or below the value;
    that equals to_byte if to_byte never steps down. Checked on a million random
    values and on every double within 2000 ulps of each threshold."""
    t = _thresholds(ol.oracle_to_byte)
    assert np.all(np.diff(t) > 0)
    rng = np.random.default_rng(9)
    xs = [rng.uniform(-0.1, 1.3, 1_000_000), rng.lognormal(-4, 3, 200_000)]
    bits = t.view(np.int64)
    near = (bits[:, None] + np.arange(-2000, 2001)[None, :]).ravel().view(np.float64)
    xs.append(near)
    for x in xs:
        lut = np.searchsorted(t, x, side="right")
        assert np.array_equal(lut.astype(np.uint8), ol.oracle_to_byte(x))


# ---------------------------------------------------------------- CPU: .scene display lines

def _scene_with(tmp_path, lines):
    base = open(scene_path("cornell.scene")).read()
    p = tmp_path / "s.scene"
    p.write_text(base + "\n" + "\n".join(lines) + "\n")
    return str(p)


DISPLAY_LINES = [
    (["bloom 0.3"], (0, 0, 0.3, 1, 0.04, 0)),
    (["bloom 0.25 threshold 2 radius 0.1", "vignette 0.4"], (0, 0, 0.25, 2, 0.1, 0.4)),
    (["exposure -1.5", "tonemap aces"], (-1.5, 2, 0, 1, 0.04, 0)),
    (["tonemap reinhard", "bloom 1 radius 0.5 threshold 0"], (0, 1, 1, 0, 0.5, 0)),
]


@pytest.mark.parametrize("lines,expect", DISPLAY_LINES)
def test_scene_display_lines(tmp_path, lines, expect):
    import mypbr_amd as m
    path = _scene_with(tmp_path, lines)
    d = m.scene_display(path)
    got = (d.exposure, d.tonemap, d.bloom, d.bloom_threshold, d.bloom_radius, d.vignette)
    assert got == expect
    if ol.reference_available():
        assert tuple(ol.reference_scene_display(path)) == got


@pytest.mark.parametrize("line", ["bloom -1", "bloom 0.3 radius 0", "bloom 0.3 size 2", "vignette -0.1",
                                  "tonemap filmic", "exposure"])
def test_scene_display_errors(tmp_path, line):
    import mypbr_amd as m
    path = _scene_with(tmp_path, [line])
    with pytest.raises(m.MptError):
        m.scene_display(path)
    if ol.reference_available():
        assert ol.reference_scene_display(path) is None


# ---------------------------------------------------------------- CPU: raw-radiance writers

def _read_pfm(data):
    parts = data.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = map(int, parts[1].split())
    assert float(parts[2]) < 0  # little-endian
    return np.frombuffer(parts[3], dtype="<f4").reshape(h, w, 3)[::-1]


def _read_hdr(data):
    """Radiance RGBE: header lines, an empty line, the resolution line, then
    flat or new-style run-length scanlines; returns float32 (h, w, 3)."""
    pos = data.index(b"\n\n") + 2
    end = data.index(b"\n", pos)
    f = data[pos:end].split()
    assert f[0] == b"-Y" and f[2] == b"+X"
    h, w = int(f[1]), int(f[3])
    buf = data[end + 1:]
    rgbe = np.zeros((h, w, 4), dtype=np.uint8)
    p = 0
    for j in range(h):
        if 8 <= w < 32768 and buf[p] == 2 and buf[p + 1] == 2 and (buf[p + 2] << 8 | buf[p + 3]) == w:
            p += 4
            for c in range(4):
                i = 0
                while i < w:
                    n = buf[p]
                    p += 1
                    if n > 128:
                        rgbe[j, i:i + n - 128, c] = buf[p]
                        p += 1
                        i += n - 128
                    else:
                        rgbe[j, i:i + n, c] = np.frombuffer(buf[p:p + n], dtype=np.uint8)
                        p += n
                        i += n
        else:
            rgbe[j] = np.frombuffer(buf[p:p + 4 * w], dtype=np.uint8).reshape(w, 4)
            p += 4 * w
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(np.float32(1), e - 136).astype(np.float32), np.float32(0))
    return rgbe[..., :3].astype(np.float32) * scale[..., None]


def test_pfm_and_hdr_writers(tmp_path):
    import mypbr_amd as m
    img = dc.sky_sun(40, 23, seed=8)
    img[0, 0] = 0.0  # a black pixel: RGBE 0
    m.write_image(str(tmp_path / "a.pfm"), img)
    m.write_image(str(tmp_path / "a.hdr"), img)
    pfm = (tmp_path / "a.pfm").read_bytes()
    assert pfm.startswith(b"PF\n40 23\n-1.0\n")
    assert np.array_equal(_read_pfm(pfm), img.astype(np.float32))
    hdr = _read_hdr((tmp_path / "a.hdr").read_bytes())
    # RGBE keeps 8 bits of mantissa per channel relative to the largest
    mx = img.astype(np.float32).max(axis=2, keepdims=True)
    assert np.all(np.abs(hdr - img.astype(np.float32)) <= mx / 128 + 1e-30)
    if ol.reference_available():
        st = dc.SETTINGS[0]
        assert ol.reference_write_image(str(tmp_path / "r.pfm"), img, st) == 0
        assert ol.reference_write_image(str(tmp_path / "r.hdr"), img, st) == 0
        assert (tmp_path / "r.pfm").read_bytes() == pfm
        assert np.array_equal(_read_hdr((tmp_path / "r.hdr").read_bytes()), hdr)


def test_writer_rejects_unknown_format(tmp_path):
    import mypbr_amd as m
    with pytest.raises(m.MptError):
        m.write_image(str(tmp_path / "a.jpg"), np.ones((2, 2, 3)))


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("name", list(dc.CASES))
def test_gpu_matches_golden(name):
    import mypbr_amd as m
    g, settings = _golden()
    img = g[f"{name}/radiance"]
    for si, st in enumerate(settings):
        d = _display(m, st)
        assert np.array_equal(_bits(m.postprocess(img, d)), _bits(g[f"{name}/post{si}"])), (name, si)
        assert np.array_equal(m.display_bytes(img, d), g[f"{name}/bytes{si}"]), (name, si)
    assert np.array_equal(m.display_bytes(img, _display(m, settings[2]), post=False), g[f"{name}/bytes_nopost"])


@pytest.mark.gpu
@pytest.mark.parametrize("si", [1, 2, 5])
def test_gpu_matches_oracle_at_1080p(si):
    import mypbr_amd as m
    img = dc.sky_sun(1920, 1080, seed=31)
    st = dc.SETTINGS[si]
    d = _display(m, st)
    assert np.array_equal(_bits(m.postprocess(img, d)), _bits(ol.oracle_postprocess(img, st)))
    assert np.array_equal(m.display_bytes(img, d), ol.oracle_display_bytes(img, st))


@pytest.mark.gpu
def test_gpu_odd_sizes_and_many_levels():
    import mypbr_amd as m
    rng = np.random.default_rng(12)
    for h, w in [(1, 1), (1, 300), (300, 1), (3, 2), (257, 131), (33, 1025)]:
        img = rng.lognormal(0.0, 1.5, (h, w, 3))
        for st in dc.SETTINGS:
            d = _display(m, st)
            assert np.array_equal(_bits(m.postprocess(img, d)), _bits(ol.oracle_postprocess(img, st))), (h, w, st)
            assert np.array_equal(m.display_bytes(img, d), ol.oracle_display_bytes(img, st)), (h, w, st)


@pytest.mark.gpu
def test_gpu_pyramid_splits():
    """Pyramids whose first level is large and the rest small, and a single
    large level: every split between the per-level kernels and the one-block
    tail kernel."""
    import mypbr_amd as m
    rng = np.random.default_rng(15)
    cases = [((300, 400), (0.0, 0, 0.4, 0.5, 0.04, 0.0)), ((1080, 1920), (0.0, 2, 0.3, 1.0, 0.001, 0.2)),
             ((540, 960), (1.0, 1, 0.6, 0.8, 0.3, 0.0))]
    for (h, w), st in cases:
        img = rng.lognormal(-0.5, 1.5, (h, w, 3))
        d = _display(m, st)
        assert np.array_equal(_bits(m.postprocess(img, d)), _bits(ol.oracle_postprocess(img, st))), (h, w, st)
        assert np.array_equal(m.display_bytes(img, d), ol.oracle_display_bytes(img, st)), (h, w, st)


def _png_pixels(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat = 8, b""
    w = h = 0
    while p < len(data):
        n = struct.unpack(">I", data[p:p + 4])[0]
        typ, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        assert zlib.crc32(data[p + 4:p + 8 + n]) == struct.unpack(">I", data[p + 8 + n:p + 12 + n])[0]
        if typ == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
            assert body[8:13] == bytes([8, 2, 0, 0, 0])
        elif typ == b"IDAT":
            idat += body
        p += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(raw[:, 0] == 0)  # filter type None on every row
    return raw[:, 1:].reshape(h, w, 3)


@pytest.mark.gpu
def test_gpu_display_image_writers(tmp_path):
    import mypbr_amd as m
    img = dc.sky_sun(70, 41, seed=13)
    st = dc.SETTINGS[5]
    d = _display(m, st)
    expect = ol.oracle_display_bytes(img, st)
    m.write_image(str(tmp_path / "a.ppm"), img, d)
    lines = ["P3", "70 41", "255"] + [f"{r} {g} {b}" for r, g, b in expect.reshape(-1, 3)]
    assert (tmp_path / "a.ppm").read_text() == "\n".join(lines) + "\n"
    m.write_image(str(tmp_path / "a.png"), img, d)
    assert np.array_equal(_png_pixels((tmp_path / "a.png").read_bytes()), expect)
    if ol.reference_available():
        assert ol.reference_write_image(str(tmp_path / "r.ppm"), img, st) == 0
        assert (tmp_path / "r.ppm").read_bytes() == (tmp_path / "a.ppm").read_bytes()


@pytest.mark.gpu
def test_gpu_device_entry_points():
    import ctypes as C
    import mypbr_amd as m
    img = dc.sky_sun(300, 170, seed=14)
    st = dc.SETTINGS[2]
    d = _display(m, st)
    h, w = img.shape[:2]
    src = m.DeviceArray.of(img)
    ws_bytes = m.lib.mpt_display_workspace_bytes(w, h, C.byref(d))
    assert ws_bytes > 0
    ws = m.DeviceArray((ws_bytes,), np.uint8)
    out = m.DeviceArray((h, w, 3), np.float64)
    rgb = m.DeviceArray((h, w, 3), np.uint8)
    m._check(m.lib.mpt_postprocess_dev(src.handle, w, h, C.byref(d), out.handle, ws.handle, None))
    m._check(m.lib.mpt_display_bytes_dev(src.handle, w, h, C.byref(d), 1, rgb.handle, ws.handle, None))
    m.synchronize()
    assert np.array_equal(_bits(out.numpy()), _bits(ol.oracle_postprocess(img, st)))
    assert np.array_equal(rgb.numpy(), ol.oracle_display_bytes(img, st))
    # without bloom no workspace is needed
    d0 = _display(m, dc.SETTINGS[3])
    assert m.lib.mpt_display_workspace_bytes(w, h, C.byref(d0)) == 0
    m._check(m.lib.mpt_display_bytes_dev(src.handle, w, h, C.byref(d0), 1, rgb.handle, None, None))
    m.synchronize()
    assert np.array_equal(rgb.numpy(), ol.oracle_display_bytes(img, dc.SETTINGS[3]))
