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
props.gltf / .bin    a glTF 2.0 model: a node tree of matrices and TRS (a
                     rotated parent, a mirroring node, a mesh placed twice);
                     a box with normals and texture coordinates (u16
                     indices), a quad (u8 indices, normalized u16 texture
                     coordinates), an unindexed pyramid without normals; the
                     materials of the reference's mapping: base colour
                     texture with KHR_texture_transform and clamping,
                     metallic-roughness and scaled normal textures, thin
                     glass (KHR_materials_transmission, _ior), clear coat,
                     an emissive light (emissive_strength), an emissive
                     texture (emissive_surface), MASK by a texture's alpha
                     (0 or 255: no random draw decides a hit), and a
                     primitive with no material; one image inside the .bin,
                     one a file whose name needs percent-decoding ("props
                     grid.png"), one buffer a data: URI;
props_box.glb        a binary glTF: a box with an embedded PNG texture, and a
                     box of BLEND coverage by a constant alpha (0.46);
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


def _png_bytes(rows, color_type):
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t.png")
        png_write(path, rows, color_type, 8)
        return open(path, "rb").read()


def gltf_props():
    """scenes/models/props.gltf, props.bin, props grid.png, props_box.glb
    (see the module text)."""
    import base64
    import json
    import struct
    import numpy as np

    def box(h):
        """24 vertices (4 per side) with normals and uvs, 36 indices."""
        pos, nrm, uv, idx = [], [], [], []
        for a in range(3):
            for s in (1, -1):
                n = [0.0, 0.0, 0.0]
                n[a] = s
                b, c = (a + 1) % 3, (a + 2) % 3
                base = len(pos)
                for (i, j) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    p = [0.0, 0.0, 0.0]
                    p[a] = s * h
                    p[b] = (2 * i - 1) * h
                    p[c] = (2 * j - 1) * h * s
                    pos.append(p)
                    nrm.append(n)
                    uv.append([i * 1.5, j * 1.5])  # past 1: repeats (or clamps)
                idx += [base, base + 1, base + 2, base, base + 2, base + 3]
        return (np.array(pos, np.float32), np.array(nrm, np.float32), np.array(uv, np.float32),
                np.array(idx, np.uint16))

    blob = bytearray()
    views, accessors = [], []

    def add_view(data, target=None, stride=None):
        while len(blob) % 4:
            blob.append(0)
        v = {"buffer": 0, "byteOffset": len(blob), "byteLength": len(data)}
        if target:
            v["target"] = target
        if stride:
            v["byteStride"] = stride
        blob.extend(data)
        views.append(v)
        return len(views) - 1

    def add_acc(arr, type_, comp, target=None, normalized=False, minmax=False):
        vi = add_view(arr.tobytes(), target)
        a = {"bufferView": vi, "componentType": comp, "count": int(arr.shape[0]), "type": type_}
        if normalized:
            a["normalized"] = True
        if minmax:
            a["min"] = [float(x) for x in arr.min(0)]
            a["max"] = [float(x) for x in arr.max(0)]
        accessors.append(a)
        return len(accessors) - 1

    P, N, T, I = box(0.5)
    box_prim = {"attributes": {"POSITION": add_acc(P, "VEC3", 5126, 34962, minmax=True),
                               "NORMAL": add_acc(N, "VEC3", 5126, 34962),
                               "TEXCOORD_0": add_acc(T, "VEC2", 5126, 34962)},
                "indices": add_acc(I, "SCALAR", 5123, 34963)}
    qp = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    qt = np.array([[0, 0], [65535, 0], [65535, 65535], [0, 65535]], np.uint16)  # normalized u16
    qi = np.array([0, 2, 1, 0, 3, 2], np.uint8)
    quad = {"POSITION": add_acc(qp, "VEC3", 5126, 34962, minmax=True), "TEXCOORD_0": add_acc(qt, "VEC2", 5123, 34962, True)}
    quad_idx = add_acc(qi, "SCALAR", 5121, 34963)
    apex = [0, 0.8, 0]
    corners = [[-0.5, 0, -0.5], [0.5, 0, -0.5], [0.5, 0, 0.5], [-0.5, 0, 0.5]]
    tris = []
    for k in range(4):
        tris += [corners[k], corners[(k + 1) % 4], apex]
    pyr = {"POSITION": add_acc(np.array(tris, np.float32), "VEC3", 5126, 34962, minmax=True)}
    pyr2 = {"POSITION": add_acc(np.array(tris, np.float32) * [0.6, 0.6, 0.6] + [1.1, 0, 0], "VEC3", 5126, 34962,
                                minmax=True)}  # a second, smaller pyramid beside it
    # images: a checker in the .bin (sRGB base colour), a metallic-roughness map
    # (G roughness, B metal), a normal map of ridges, an RGBA leaf mask, a glow panel
    checker = [[c for i in range(16) for c in ((230, 60, 40) if (i // 4 + j // 4) % 2 else (240, 230, 200))]
               for j in range(16)]
    mr = [[c for i in range(16) for c in (0, 40 + 12 * j, 255 if i < 8 else 0)] for j in range(16)]
    ridges = [[c for i in range(16) for c in (int(128 + 90 * math.sin(i * 0.8)), 128, 230)] for j in range(16)]
    leaf = [[c for i in range(16) for c in (60, 160, 50, 255 if (i - 7.5) ** 2 + (j - 7.5) ** 2 < 40 else 0)]
            for j in range(16)]
    panel = [[c for i in range(8) for c in ((255, 255, 255) if (i + j) % 3 else (20, 80, 255))] for j in range(8)]
    img_checker = add_view(_png_bytes(checker, 2))
    png_write(os.path.join(HERE, "props grid.png"), mr, 2, 8)
    img_ridges = add_view(_png_bytes(ridges, 2))
    img_leaf = add_view(_png_bytes(leaf, 6))
    img_panel = add_view(_png_bytes(panel, 2))
    images = [{"bufferView": img_checker, "mimeType": "image/png"}, {"uri": "props%20grid.png"},
              {"bufferView": img_ridges, "mimeType": "image/png"}, {"bufferView": img_leaf, "mimeType": "image/png"},
              {"bufferView": img_panel, "mimeType": "image/png"}]
    samplers = [{}, {"wrapS": 33071, "wrapT": 33071}]
    textures = [{"source": 0, "sampler": 1}, {"source": 1}, {"source": 2}, {"source": 3}, {"source": 4}]
    tt = {"KHR_texture_transform": {"offset": [0.25, -0.1], "rotation": 0.4, "scale": [2.0, 1.5]}}
    materials = [
        {"name": "painted", "pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.8, 0.7, 1],
                                                     "baseColorTexture": {"index": 0, "extensions": tt},
                                                     "metallicFactor": 0, "roughnessFactor": 0.6}},
        {"name": "metal_rough", "pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.8, 0.85, 1],
                                                         "metallicRoughnessTexture": {"index": 1},
                                                         "metallicFactor": 1, "roughnessFactor": 0.9},
         "normalTexture": {"index": 2, "scale": 0.5}},
        {"name": "glass", "pbrMetallicRoughness": {"baseColorFactor": [0.8, 1.0, 0.9, 1], "roughnessFactor": 0},
         "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1},
                        "KHR_materials_ior": {"ior": 1.45}}},
        {"name": "coated", "pbrMetallicRoughness": {"baseColorFactor": [0.1, 0.3, 0.8, 1], "metallicFactor": 0,
                                                    "roughnessFactor": 0.5},
         "extensions": {"KHR_materials_clearcoat": {"clearcoatFactor": 0.8, "clearcoatRoughnessFactor": 0.1}}},
        {"name": "lamp", "emissiveFactor": [1, 0.8, 0.5],
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4}}},
        {"name": "screen", "pbrMetallicRoughness": {"baseColorFactor": [0.05, 0.05, 0.05, 1], "metallicFactor": 0},
         "emissiveFactor": [1, 1, 1], "emissiveTexture": {"index": 4}},
        {"name": "leafy", "alphaMode": "MASK", "pbrMetallicRoughness": {"baseColorTexture": {"index": 3},
                                                                      "metallicFactor": 0}},
        {"name": "ghost", "alphaMode": "BLEND", "pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.9, 1.0, 0.46],
                                                                       "metallicFactor": 0}},
        {"name": "sheeny", "pbrMetallicRoughness": {"baseColorFactor": [0.7, 0.2, 0.6, 1], "metallicFactor": 0},
         "extensions": {"KHR_materials_sheen": {"sheenColorFactor": [1, 1, 1]}}},
    ]
    meshes = [
        {"name": "painted box", "primitives": [dict(box_prim, material=0)]},
        {"name": "metal box", "primitives": [dict(box_prim, material=1)]},
        {"name": "glass box", "primitives": [dict(box_prim, material=2)]},
        {"name": "coated box", "primitives": [dict(box_prim, material=3)]},
        {"name": "lamp", "primitives": [{"attributes": quad, "indices": quad_idx, "material": 4}]},
        {"name": "screen", "primitives": [{"attributes": quad, "indices": quad_idx, "material": 5}]},
        {"name": "leaf", "primitives": [{"attributes": quad, "indices": quad_idx, "material": 6}]},
        {"name": "ghost", "primitives": [dict(box_prim, material=7)]},
        {"name": "pyramid", "primitives": [{"attributes": pyr}, {"attributes": pyr2, "material": 8}]},
    ]
    s2 = math.sqrt(0.5)
    nodes = [
        {"name": "root", "children": [1, 2, 3, 4, 5, 6, 7, 8, 9], "translation": [0, 0.5, 0],
         "rotation": [0, math.sin(0.15), 0, math.cos(0.15)], "scale": [1.0, 1.0, 1.0]},
        {"mesh": 0, "translation": [-2.2, 0, 0]},
        {"mesh": 1, "translation": [-0.8, 0, 0], "rotation": [0, 0.3826834, 0, 0.9238795]},
        {"mesh": 2, "translation": [0.7, 0, 0], "scale": [0.8, 1.2, 0.8]},
        {"mesh": 3, "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 2.1, 0, 0, 1]},
        {"mesh": 4, "translation": [0, 2.4, 0], "rotation": [1, 0, 0, 0], "scale": [0.6, 1, 0.6]},  # facing down
        {"mesh": 5, "translation": [0, 0.6, -2.0], "rotation": [s2, 0, 0, s2], "scale": [1.6, 1, 0.9]},
        {"mesh": 6, "translation": [1.4, 0.2, 1.4], "rotation": [s2, 0, 0, s2], "scale": [0.5, 1, 0.5]},
        {"mesh": 8, "translation": [0.0, -0.5, 1.6], "scale": [-1, 1, 1]},  # mirroring: the winding flips
        {"mesh": 0, "translation": [2.0, 0.6, 1.8], "scale": [0.4, 0.4, 0.4]},  # the painted box again
    ]
    # one more buffer as a data: URI: a second copy of the quad's positions
    extra = qp.tobytes()
    doc = {"asset": {"version": "2.0", "generator": "scenes/make_assets.py"}, "scene": 0,
           "scenes": [{"nodes": [0]}], "nodes": nodes, "meshes": meshes, "materials": materials,
           "textures": textures, "images": images, "samplers": samplers, "accessors": accessors,
           "bufferViews": views,
           "buffers": [{"uri": "props.bin", "byteLength": len(blob)},
                       {"uri": "data:application/octet-stream;base64," + base64.b64encode(extra).decode(),
                        "byteLength": len(extra)}],
           "extensionsUsed": ["KHR_texture_transform", "KHR_materials_transmission", "KHR_materials_ior",
                              "KHR_materials_clearcoat", "KHR_materials_emissive_strength", "KHR_materials_sheen"]}
    accessors.append({"bufferView": len(views), "componentType": 5126, "count": 4, "type": "VEC3"})
    views.append({"buffer": 1, "byteOffset": 0, "byteLength": len(extra)})
    doc["meshes"].append({"name": "floor tile", "primitives": [{"attributes": {"POSITION": len(accessors) - 1},
                                                               "indices": quad_idx, "material": 1}]})
    nodes[0]["children"].append(len(nodes))
    nodes.append({"mesh": len(doc["meshes"]) - 1, "translation": [0, -0.49, 0], "scale": [3, 1, 3]})
    open(os.path.join(HERE, "props.bin"), "wb").write(bytes(blob))
    with open(os.path.join(HERE, "props.gltf"), "w") as f:
        json.dump(doc, f, indent=1)

    # the binary container: one box with an embedded texture
    blob = bytearray()
    views, accessors = [], []
    P, N, T, I = box(0.7)
    prim = {"attributes": {"POSITION": add_acc(P, "VEC3", 5126, 34962, minmax=True),
                           "NORMAL": add_acc(N, "VEC3", 5126, 34962), "TEXCOORD_0": add_acc(T, "VEC2", 5126, 34962)},
            "indices": add_acc(I, "SCALAR", 5123, 34963), "material": 0}
    iv = add_view(_png_bytes(checker, 2))
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0, 1]}],
           "nodes": [{"mesh": 0, "rotation": [0, 0.2588190, 0, 0.9659258]},
                     {"mesh": 1, "translation": [0, 0, 1.6], "scale": [0.6, 0.6, 0.6]}],
           "meshes": [{"primitives": [prim]}, {"primitives": [dict(prim, material=1)]}],
           "materials": [{"name": "crate", "pbrMetallicRoughness": {"baseColorTexture": {"index": 0},
                                                                    "metallicFactor": 0.0, "roughnessFactor": 0.8}},
                         {"name": "ghost", "alphaMode": "BLEND",
                          "pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.9, 1.0, 0.46], "metallicFactor": 0}}],
           "textures": [{"source": 0}], "images": [{"bufferView": iv, "mimeType": "image/png"}],
           "accessors": accessors, "bufferViews": views, "buffers": [{"byteLength": 0}]}
    while len(blob) % 4:
        blob.append(0)
    doc["buffers"][0]["byteLength"] = len(blob)
    js = json.dumps(doc).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + bytes(blob)
    with open(os.path.join(HERE, "props_box.glb"), "wb") as f:
        f.write(b"glTF" + struct.pack("<II", 2, 12 + len(body)) + body)


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
    gltf_props()
