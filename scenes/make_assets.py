"""Writes the test assets under scenes/models (run from the repository root:
python scenes/make_assets.py).

room.obj / room.mtl  a room of MTL materials: lambertian walls (Kd, one as
                     a single grey value), a glossy box (Ks/Ns/Ni ->
                     principled) indexed with negative indices, an emissive
                     ceiling panel (Ke) written as one quad face, a pentagon
                     (fan triangulation), a degenerate face (skipped), and a
                     faceted sphere with vertex normals (v//n) whose faces
                     use v/t/n with texture coordinates;
blob.obj             a low-poly sphere with no normals and a scene material
                     (usemtl names a material of the .scene file), for
                     `mesh ... smooth` and instancing;
sky_rle.hdr          a 256 x 128 sun-and-sky environment in Radiance RGBE with
                     new-style run-length encoded scanlines (runs and
                     literals), written here from the published format;
bumps_normal.png     a 64 x 64 tangent-space normal map of round bumps (RGB,
                     8-bit), its rows written with all five PNG filters in turn;
tiles_rough.png      a 48 x 48 roughness map of tiles, 16-bit grey;
tiles_metal.png      a 48 x 48 metallic map, a 2-bit palette image;
crate.obj / .mtl     a box whose MTL materials take map_Kd (the JPEG below) and a
                     norm normal map, with texture coordinates beyond [0, 1]
                     (the textures repeat), mirrored on one side, and one
                     face without them (its (u, v) are the barycentrics);
leaves.obj / .mtl    alpha-masked quads: "leaf" takes map_d leaf_mask.png (grey +
                     alpha, a leaf outline with a half-transparent rim), "fence"
                     a diffuse map with an alpha channel and no map_d (the
                     diffuse texture's alpha is the mask), "board" stays opaque;
leaf_mask.png        64 x 64 grey + alpha (PNG color type 4);
fence_rgba.png       32 x 32 RGBA slats with transparent gaps (PNG color type 6);
island.vox           a MagicaVoxel (version 200) island: a 40 x 40 x 18 terrain
                     model (sand, grass, rock, a glass pond, a metal post,
                     an emissive lantern) and a 5 x 5 x 6 tower placed three
                     times by the scene graph (nTRN translations and ROTATION
                     bytes, one mirrored, one hidden), an RGBA palette and
                     MATL materials (_glass, _metal, _emit);
plain150.vox         a version-150 file with two models (a row of 255 voxels of
                     palette indices 1-255, and a small column), no scene graph
                     and no RGBA chunk (MagicaVoxel's default palette);
globe.jpg            a 128 x 64 continents-and-ocean color map, baseline JPEG
                     with 2x2-subsampled chroma. JPEG needs an encoder: this
                     one is written by stb_image_write through the compiled
                     reference harness (oracle/_ref), so it is only rewritten
                     where that is built.
"""
import math
import os

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models")


def sphere(nu, nv, r, c):
    verts, faces = [], []
    for j in range(nv + 1):
        th = math.pi * j / nv
        for i in range(nu):
            ph = 2 * math.pi * i / nu
            verts.append((c[0] + r * math.sin(th) * math.cos(ph), c[1] + r * math.cos(th),
                          c[2] + r * math.sin(th) * math.sin(ph)))
    for j in range(nv):
        for i in range(nu):
            a = j * nu + i
            b = j * nu + (i + 1) % nu
            faces.append((a, a + nu, b + nu, b))  # quads (degenerate at the poles: skipped or fanned)
    return verts, faces


def fmt(x):
    return f"{x:.6f}"


def room():
    L = []
    L.append("# test room for the OBJ loader")
    L.append("mtllib room.mtl")
    L.append("o room")
    # floor / walls: 8 corners of a 4 x 3 x 4 box around the origin, floor at y = 0
    c = [(-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2), (-2, 3, -2), (2, 3, -2), (2, 3, 2), (-2, 3, 2)]
    for p in c:
        L.append("v " + " ".join(fmt(x) for x in p))
    L.append("usemtl floor")
    L.append("f 1 4 3 2")           # floor, facing up
    L.append("usemtl wall_red")
    L.append("f 1 5 8 4")           # left wall
    L.append("usemtl wall_green")
    L.append("f 2 3 7 6")           # right wall
    L.append("usemtl grey")
    L.append("f 1 2 6 5")           # back wall
    L.append("f 5 6 7 8")           # ceiling
    # ceiling light panel, slightly below the ceiling, facing down, one quad face
    for p in [(-0.5, 2.99, -0.5), (0.5, 2.99, -0.5), (0.5, 2.99, 0.5), (-0.5, 2.99, 0.5)]:
        L.append("v " + " ".join(fmt(x) for x in p))
    L.append("usemtl lamp")
    L.append("f 9 10 11 12")
    # glossy box on the floor, faces with negative (relative) indices
    x0, x1, z0, z1, h = -1.4, -0.4, -1.2, -0.2, 1.1
    box = [(x0, 0, z0), (x1, 0, z0), (x1, 0, z1), (x0, 0, z1), (x0, h, z0), (x1, h, z0), (x1, h, z1), (x0, h, z1)]
    for p in box:
        L.append("v " + " ".join(fmt(x) for x in p))
    L.append("usemtl gloss")
    for f in [(1, 2, 6, 5), (2, 3, 7, 6), (3, 4, 8, 7), (4, 1, 5, 8), (5, 6, 7, 8)]:
        L.append("f " + " ".join(str(k - 9) for k in f))
    # a pentagon on the back wall and a degenerate face
    for k in range(5):
        a = 2 * math.pi * k / 5
        L.append(f"v {fmt(1.0 + 0.5 * math.cos(a))} {fmt(1.8 + 0.5 * math.sin(a))} {fmt(-1.99)}")
    L.append("usemtl wall_red")
    L.append("f -5 -4 -3 -2 -1")
    L.append("f 1 1 2")
    # faceted sphere with vertex normals and texture coordinates
    verts, faces = sphere(10, 6, 0.6, (0.9, 0.6, 0.6))
    base_v = 12 + 8 + 5
    for p in verts:
        L.append("v " + " ".join(fmt(x) for x in p))
    for p in verts:
        n = [(p[k] - (0.9, 0.6, 0.6)[k]) / 0.6 for k in range(3)]
        L.append("vn " + " ".join(fmt(x) for x in n))
    for k in range(len(verts)):
        L.append(f"vt {fmt((k % 10) / 10)} {fmt((k // 10) / 6)}")
    L.append("usemtl ball")
    for f in faces:
        L.append("f " + " ".join(f"{base_v + a + 1}/{a + 1}/{a + 1}" for a in f))
    open(os.path.join(HERE, "room.obj"), "w").write("\n".join(L) + "\n")
    M = ["# materials of room.obj",
         "newmtl floor", "Kd 0.7 0.65 0.6", "illum 2",
         "newmtl wall_red", "Kd 0.63 0.07 0.05",
         "newmtl wall_green", "Kd 0.14 0.45 0.09",
         "newmtl grey", "Kd 0.6",
         "newmtl lamp", "Kd 0 0 0", "Ke 12 11 9",
         "newmtl gloss", "Kd 0.2 0.3 0.6", "Ks 0.5 0.5 0.5", "Ns 80", "Ni 1.45",
         "newmtl ball", "Kd 0.8 0.6 0.2", "Ks 0.04", "Ns 10"]
    open(os.path.join(HERE, "room.mtl"), "w").write("\n".join(M) + "\n")


def blob():
    verts, faces = sphere(12, 8, 0.35, (0, 0.35, 0))
    L = ["# low-poly sphere without normals; its material comes from the scene"]
    for p in verts:
        L.append("v " + " ".join(fmt(x) for x in p))
    L.append("usemtl jade")
    for f in faces:
        L.append("f " + " ".join(str(a + 1) for a in f))
    open(os.path.join(HERE, "blob.obj"), "w").write("\n".join(L) + "\n")


def rgbe(c):
    """Ward's float2rgbe on float32 channels."""
    import numpy as np
    c = np.asarray(c, dtype=np.float32)
    v = c.max()
    if v < 1e-32:
        return (0, 0, 0, 0)
    m, e = math.frexp(float(v))
    scale = np.float32(m) * np.float32(256) / v
    return tuple(int(np.float32(x) * scale) for x in c) + (e + 128,)


def rle_channel(vals):
    """New-style RLE of one scanline channel: runs of 3+ equal bytes as
    (128 + n, byte), everything else as literals of up to 128 bytes."""
    out, i, n = [], 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += [128 + (j - i), vals[i]]
            i = j
            continue
        k = i
        while k < n and k - i < 128:
            if k + 2 < n and vals[k] == vals[k + 1] == vals[k + 2]:
                break
            k += 1
        out += [k - i] + list(vals[i:k])
        i = k
    return out


def sky_hdr(w=256, h=128):
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from mypbr_amd.workloads import sun_sky
    img = sun_sky(w, h, sun_radiance=40.0, sun_angle_deg=4.0)
    img[: h // 8, : w // 4] = 0.0  # a black patch: runs of zero bytes
    data = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for j in range(h):
        px = [rgbe(img[j, i]) for i in range(w)]
        data += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            data += bytes(rle_channel([p[c] for p in px]))
    open(os.path.join(HERE, "sky_rle.hdr"), "wb").write(bytes(data))


def crate():
    # one material: parts of several materials come in heap order in the reference (a map keyed by pointers)
    mtl = ["newmtl crate_wood", "Kd 0.8 0.8 0.8", "Ks 0.3 0.3 0.3", "Ns 60", "map_Kd -bm 1 globe.jpg",
           "norm bumps_normal.png"]
    open(os.path.join(HERE, "crate.mtl"), "w").write("\n".join(mtl) + "\n")
    s = 0.45
    v = [(-s, 0, -s), (s, 0, -s), (s, 2 * s, -s), (-s, 2 * s, -s), (-s, 0, s), (s, 0, s), (s, 2 * s, s), (-s, 2 * s, s)]
    L = ["# a box of textured MTL materials", "mtllib crate.mtl"]
    L += ["v " + " ".join(fmt(x) for x in p) for p in v]
    # texture coordinates: a 2 x 2 repeat, an offset one outside [0, 1], and a mirrored one
    vts = [(0, 0), (2, 0), (2, 2), (0, 2), (-0.5, 0.25), (0.75, 0.25), (0.75, 1.5), (-0.5, 1.5), (1, 0), (0, 0), (0, 1),
           (1, 1)]
    L += [f"vt {fmt(a)} {fmt(b)}" for a, b in vts]
    L += ["vn 0 0 1", "vn 0 0 -1"]
    L.append("usemtl crate_wood")
    L.append("f 5/1/1 6/2/1 7/3/1 8/4/1")      # front (+z): repeated
    L.append("f 2/5/2 1/6/2 4/7/2 3/8/2")      # back: offset outside [0, 1]
    L.append("f 1/9 5/10 8/11 4/12")           # left: mirrored (u runs the other way)
    L.append("f 6 2 3 7")                      # right: no texture coordinates
    L.append("f 8/1 7/2 3/3 4/4")              # top
    L.append("f 1/1 2/2 6/3 5/4")              # bottom
    open(os.path.join(HERE, "crate.obj"), "w").write("\n".join(L) + "\n")


def png_write(path, rows, color_type, depth, palette=None):
    """PNG from rows of samples (lists of ints), filters 0-4 in turn by row
    (PNG spec, section 9: each filter's difference from the raw bytes)."""
    import struct
    import zlib
    samples = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[color_type]

    def raw(row):  # samples to bytes, big-endian 16-bit, or packed low bit depths
        if depth == 16:
            return b"".join(struct.pack(">H", v) for v in row)
        if depth == 8:
            return bytes(row)
        per = 8 // depth
        out = bytearray()
        for i in range(0, len(row), per):
            b = 0
            for k in range(per):
                v = row[i + k] if i + k < len(row) else 0
                b |= v << (8 - depth * (k + 1))
            out.append(b)
        return bytes(out)

    bpp = max(1, samples * depth // 8)
    prev = None
    body = bytearray()
    for y, row in enumerate(rows):
        cur = raw(row)
        up = prev if prev is not None else bytes(len(cur))
        f = y % 5
        enc = bytearray()
        for i, x in enumerate(cur):
            a = cur[i - bpp] if i >= bpp else 0
            b = up[i]
            c = up[i - bpp] if i >= bpp else 0
            if f == 0:
                pred = 0
            elif f == 1:
                pred = a
            elif f == 2:
                pred = b
            elif f == 3:
                pred = (a + b) >> 1
            else:
                pp = a + b - c
                pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
                pred = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            enc.append((x - pred) & 255)
        body += bytes([f]) + enc
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    w = len(rows[0]) // samples
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, len(rows), depth, color_type, 0, 0, 0))
    if palette is not None:
        out += chunk(b"PLTE", bytes(v for rgb in palette for v in rgb))
    out += chunk(b"IDAT", zlib.compress(bytes(body), 9)) + chunk(b"IEND", b"")
    open(path, "wb").write(out)


def texture_maps():
    n = 64
    rows = []
    for j in range(n):
        row = []
        for i in range(n):
            # bumps of radius 0.4 on an 8 x 8 cell grid; the normal as a color
            cx, cy = (i % 8 + 0.5) / 8 - 0.5, (j % 8 + 0.5) / 8 - 0.5
            r2 = (cx * cx + cy * cy) / 0.16
            nx, ny = (cx / 0.4, -cy / 0.4) if r2 < 1 else (0.0, 0.0)
            nz = math.sqrt(max(0.0, 1 - nx * nx - ny * ny))
            row += [int(round(255 * (0.5 + 0.5 * v))) for v in (nx, ny, nz)]
        rows.append(row)
    png_write(os.path.join(HERE, "bumps_normal.png"), rows, 2, 8)
    m = 48
    rough = [[(12000 if ((i // 12) + (j // 12)) % 2 else 52000) + 37 * ((i * 7 + j * 3) % 11) for i in range(m)]
             for j in range(m)]
    png_write(os.path.join(HERE, "tiles_rough.png"), rough, 0, 16)
    metal = [[(0 if (i // 6) % 2 else 3) if (j // 6) % 2 else (1 if (i // 6) % 2 else 2) for i in range(m)]
             for j in range(m)]
    png_write(os.path.join(HERE, "tiles_metal.png"), metal, 3, 2, palette=[(0, 0, 0), (80, 80, 80), (180, 0, 0), (255, 255, 255)])


def leaves():
    n = 64
    mask = []
    for j in range(n):
        row = []
        for i in range(n):
            x, y = (i + 0.5) / n - 0.5, (j + 0.5) / n - 0.5
            r = (x / 0.45) ** 2 + (y / 0.3) ** 2  # an ellipse: opaque inside, a soft rim, clear outside
            a = 255 if r < 0.7 else (128 if r < 1 else (40 if r < 1.15 else 0))
            row += [200, a]
        mask.append(row)
    png_write(os.path.join(HERE, "leaf_mask.png"), mask, 4, 8)
    m = 32
    fence = []
    for j in range(m):
        row = []
        for i in range(m):
            slat = (i // 4) % 2 == 0
            row += [150, 110, 70, 255 if slat else (96 if (j // 8) % 2 else 0)]
        fence.append(row)
    png_write(os.path.join(HERE, "fence_rgba.png"), fence, 6, 8)
    with open(os.path.join(HERE, "leaves.mtl"), "w") as f:
        f.write("# alpha-masked materials (see scenes/make_assets.py)\n"
                "newmtl leaf\nKd 0.25 0.55 0.15\nmap_d leaf_mask.png\n\n"
                "newmtl fence\nKd 1 1 1\nmap_Kd fence_rgba.png\n\n"
                "newmtl board\nKd 0.6 0.6 0.55\n")
    v, vt, faces = [], [], []

    def quad(c, u, w, mat, uv_scale=1.0):
        base = len(v)
        for a, b in ((0, 0), (1, 0), (1, 1), (0, 1)):
            v.append([c[k] + a * u[k] + b * w[k] for k in range(3)])
            vt.append([a * uv_scale, b * uv_scale])
        faces.append((mat, base))

    quad((-1.6, 0.2, 0.0), (1.2, 0, 0.3), (0, 1.0, 0), "leaf")
    quad((-0.9, 0.5, 0.6), (1.1, 0.1, -0.2), (0, 0.9, 0.1), "leaf")
    quad((0.3, 0.0, -0.2), (1.4, 0, 0), (0, 1.4, 0), "fence", 2.0)
    quad((-2.0, 0.0, -1.0), (4.0, 0, 0), (0, 2.0, 0), "board")
    with open(os.path.join(HERE, "leaves.obj"), "w") as f:
        f.write("# two leaves, a fence panel and an opaque board behind them (see scenes/make_assets.py)\n"
                "mtllib leaves.mtl\n")
        for p in v:
            f.write("v %s %s %s\n" % tuple(fmt(x) for x in p))
        for t in vt:
            f.write("vt %s %s\n" % tuple(fmt(x) for x in t))
        cur = None
        for mat, b in faces:
            if mat != cur:
                f.write("usemtl %s\n" % mat)
                cur = mat
            f.write("f %d/%d %d/%d %d/%d %d/%d\n" % (b + 1, b + 1, b + 2, b + 2, b + 3, b + 3, b + 4, b + 4))


def globe_jpg(w=128, h=64):
    import ctypes as C
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib = os.path.join(root, "oracle", "_ref", "libmpt_ref.so")
    if not os.path.exists(lib):
        print("oracle/_ref is not built: globe.jpg kept as it is")
        return
    import numpy as np
    img = np.zeros((h, w, 3), np.uint8)
    for j in range(h):
        for i in range(w):
            lon, lat = 2 * math.pi * i / w, math.pi * (j / h - 0.5)
            land = math.sin(3 * lon) * math.cos(2 * lat) + 0.5 * math.sin(5 * lon + 1.3) * math.sin(3 * lat + 0.4)
            if abs(lat) > 1.25:
                img[j, i] = (235, 240, 245)
            elif land > 0.35:
                img[j, i] = (60 + int(80 * (land - 0.35)), 120, 40)
            else:
                img[j, i] = (20, 50 + int(30 * (1 + land)), 140)
    R = C.CDLL(lib)
    R.ref_write_jpg.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_int]
    assert R.ref_write_jpg(os.fsencode(os.path.join(HERE, "globe.jpg")), w, h, 3,
                           img.ctypes.data_as(C.POINTER(C.c_uint8)), 85) == 0
    del sys


def _chunk(cid, content=b"", children=b""):
    import struct
    return cid + struct.pack("<ii", len(content), len(children)) + content + children


def _vstr(x):
    import struct
    b = x.encode()
    return struct.pack("<i", len(b)) + b


def _vdict(d):
    import struct
    return struct.pack("<i", len(d)) + b"".join(_vstr(k) + _vstr(v) for k, v in d.items())


def vox_write(path, models, palette=None, materials=None, graph=None, version=200):
    """A .vox file from the published format (ephtracy/voxel-model):
    models is a list of ((nx, ny, nz), [(x, y, z, index), ...]); palette 256
    RGBA tuples (entry k is palette index k + 1) or None for none; materials
    {index: {key: value}}; graph a list of (node id, chunk id, content bytes)
    for nTRN / nGRP / nSHP, or None."""
    import struct
    body = b""
    for (nx, ny, nz), vox in models:
        body += _chunk(b"SIZE", struct.pack("<iii", nx, ny, nz))
        body += _chunk(b"XYZI", struct.pack("<i", len(vox)) + b"".join(bytes(v) for v in vox))
    for _, cid, content in graph or []:
        body += _chunk(cid, content)
    if palette is not None:
        body += _chunk(b"RGBA", b"".join(bytes(c) for c in palette))
    for idx, props in sorted((materials or {}).items()):
        body += _chunk(b"MATL", struct.pack("<i", idx) + _vdict(props))
    with open(path, "wb") as f:
        f.write(b"VOX " + struct.pack("<i", version) + _chunk(b"MAIN", b"", body))


def _ntrn(nid, child, t=None, r=None, hidden=False):
    import struct
    frame = {}
    if r is not None:
        frame["_r"] = str(r)
    if t is not None:
        frame["_t"] = " ".join(str(v) for v in t)
    attrs = {"_hidden": "1"} if hidden else {}
    return (nid, b"nTRN", struct.pack("<i", nid) + _vdict(attrs) + struct.pack("<iiii", child, -1, 0, 1) + _vdict(frame))


def _ngrp(nid, children):
    import struct
    return (nid, b"nGRP", struct.pack("<i", nid) + _vdict({}) + struct.pack("<i", len(children)) +
            b"".join(struct.pack("<i", c) for c in children))


def _nshp(nid, model):
    import struct
    return (nid, b"nSHP", struct.pack("<i", nid) + _vdict({}) + struct.pack("<ii", 1, model) + _vdict({}))


def island():
    """scenes/models/island.vox and plain150.vox (see the module text)."""
    # palette: 1 sand, 2 grass, 3 rock, 4 dark rock, 5 glass water, 6 metal,
    # 7 lantern (emissive), 8 wood, 9 roof, 10 tinted glass window
    pal = [(0, 0, 0, 255)] * 256
    cols = {1: (218, 196, 140), 2: (86, 150, 62), 3: (128, 124, 118), 4: (84, 80, 78), 5: (70, 140, 200),
            6: (200, 200, 210), 7: (255, 200, 120), 8: (120, 80, 50), 9: (170, 60, 50), 10: (180, 220, 255)}
    for k, c in cols.items():
        pal[k - 1] = (*c, 255)  # entry k - 1 is palette index k
    mats = {5: {"_type": "_glass", "_rough": "0.0", "_ior": "0.33", "_att": "0.6"},
            6: {"_type": "_metal", "_rough": "0.25", "_metal": "1"},
            7: {"_type": "_emit", "_emit": "2", "_flux": "1"},
            9: {"_type": "_metal", "_plastic": "1", "_rough": "0.5"},
            10: {"_type": "_glass", "_rough": "0.2", "_ri": "1.45"},
            3: {"_type": "_blend"}}  # an unsupported type: diffuse, with a warning
    nx, ny, nz = 40, 40, 18
    vox = []
    for y in range(ny):
        for x in range(nx):
            r = math.hypot(x - 19.5, y - 19.5)
            h = int(10 - 0.45 * r + 2.5 * math.sin(0.45 * x) * math.cos(0.35 * y))
            for z in range(max(h, 0)):
                top = z == h - 1
                idx = 3 if z < h - 3 else (4 if r > 16 else (2 if top and r < 13 else 1))
                if 9 <= x <= 14 and 22 <= y <= 27 and z >= h - 2 and r < 14:
                    idx = 5  # a glass pond sunk into the top
                vox.append((x, y, z, idx))
    for z in range(8, 14):  # a metal post with a lantern on top
        vox.append((20, 12, z, 6))
    vox.append((20, 12, 14, 7))
    vox.append((21, 12, 14, 7))
    tower = []
    for z in range(6):
        for y in range(5):
            for x in range(5):
                edge = x in (0, 4) or y in (0, 4)
                if z == 5:
                    tower.append((x, y, z, 9))
                elif edge:
                    tower.append((x, y, z, 10 if (z == 3 and x == 2) else 8))
    graph = [_ntrn(0, 1), _ngrp(1, [2, 4, 6, 8, 10]),
             _ntrn(2, 3), _nshp(3, 0),
             _ntrn(4, 5, t=(9, -8, 9), r=4), _nshp(5, 1),               # the tower, upright
             _ntrn(6, 7, t=(-10, 6, 8), r=1 | (2 << 2)), _nshp(7, 1),   # turned about z (a signed permutation)
             _ntrn(8, 9, t=(0, 0, 20), r=4 | 16, hidden=True), _nshp(9, 1),  # hidden: not placed
             _ntrn(10, 11, t=(-4, -12, 9), r=4 | 32), _nshp(11, 1)]    # mirrored in y
    vox_write(os.path.join(HERE, "island.vox"), [((nx, ny, nz), vox), ((5, 5, 6), tower)], pal, mats, graph)
    small = [((255, 1, 1), [(x, 0, 0, x + 1) for x in range(255)]),  # every palette index in a row
             ((2, 2, 3), [(1, 1, z, 200) for z in range(3)])]
    vox_write(os.path.join(HERE, "plain150.vox"), small, None, None, None, version=150)


if __name__ == "__main__":
    os.makedirs(HERE, exist_ok=True)
    room()
    blob()
    sky_hdr()
    texture_maps()
    globe_jpg()
    crate()
    leaves()
    island()
