"""Seeded generator of the batched wavefront workload: a batch of
terrain-and-instances scenes of uneven resolutions.

Each scene (seeded):
  terrain     an fBm heightfield triangulated on an n x n grid (1024: 2 * 1023^2
              = 2,093,058 triangles), one mesh;
  instances   icospheres of subdivision 4 (5,120 triangles) under random affine
              transforms (rotation, anisotropic scale, translation onto the
              terrain); the sphere mesh is shared, in `variants` copies that
              carry different materials;
  materials   principled GGX: roughness U[0.05, 1], metallic Bernoulli(0.3),
              base colour U[0.1, 0.9]^3;
  lighting    an analytic sun + sky lat-long environment (2048 x 1024), the
              only light, importance-sampled;
  medium      homogeneous fog over the scene box, sigma_t 0.02, albedo 0.9.
The batch is 8 such scenes at 256^2, 512^2, 640x360, 768^2, 1024x576,
1280x720, 1600x900 and 1920x1080, 32 spp, max depth 12.
"""
import math

import numpy as np

from .scenebuild import DescBuilder, camera

RESOLUTIONS = [(256, 256), (512, 512), (640, 360), (768, 768), (1024, 576), (1280, 720), (1600, 900), (1920, 1080)]
TERRAIN_SIZE = 100.0


def icosphere(subdiv):
    """Unit icosphere: (positions, triangle indices) with 20 * 4^subdiv faces."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        mid = {}
        nf = []

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(verts), np.array(f, dtype=np.int64)


def fbm_heightfield(n, seed, octaves=6, amplitude=8.0):
    """n x n heights: octaves of lattice noise, cubic-spline interpolated."""
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(seed)
    h = np.zeros((n, n))
    amp = amplitude
    cells = 4
    for _ in range(octaves):
        lat = rng.standard_normal((cells + 1, cells + 1))
        g = np.linspace(0, cells, n)
        Y, X = np.meshgrid(g, g, indexing="ij")
        h += amp * map_coordinates(lat, [Y, X], order=3, mode="nearest")
        amp *= 0.5
        cells *= 2
    return h


def terrain_mesh(n, seed):
    """Grid positions (n^2, 3) over [-50, 50]^2 and 2 (n - 1)^2 triangles."""
    h = fbm_heightfield(n, seed)
    xs = np.linspace(-TERRAIN_SIZE / 2, TERRAIN_SIZE / 2, n)
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    pos = np.stack([X, h, Z], axis=-1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    a = (j * n + i).ravel()
    b, c, d = a + 1, a + n, a + n + 1
    tris = np.stack([np.stack([a, c, b], 1), np.stack([b, c, d], 1)], 1).reshape(-1, 3)
    return pos, tris, h


def sun_sky(width=2048, height=1024, sun_dir=(0.4, 0.55, -0.73), sun_radiance=2000.0, sun_angle_deg=0.53):
    """Lat-long environment, row 0 at +y: an analytic sky gradient (zenith
    blue to horizon white, dim ground) plus a sun disk, float32 values."""
    s = np.asarray(sun_dir, dtype=np.float64)
    s /= np.linalg.norm(s)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * math.pi                      # polar angle from +y
    phi = (u - 0.5) * 2 * math.pi            # azimuth, u = 0.5 + atan2(x, -z) / 2pi
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    x = st * np.sin(phi)[None, :]
    z = -st * np.cos(phi)[None, :]
    y = np.broadcast_to(ct, x.shape)
    up = np.clip(y, 0, 1)
    zenith = np.array([0.25, 0.45, 0.95])
    horizon = np.array([0.9, 0.92, 0.95])
    ground = np.array([0.25, 0.22, 0.2])
    t = up[..., None] ** 0.5
    sky = np.where(y[..., None] > 0, (1 - t) * horizon + t * zenith, ground * (0.5 + 0.5 * (1 + y[..., None])))
    cosg = x * s[0] + y * s[1] + z * s[2]
    halo = np.exp((cosg - 1) * 60.0)[..., None] * np.array([1.0, 0.85, 0.6]) * 3.0
    disk = (cosg > math.cos(math.radians(sun_angle_deg))) * sun_radiance
    img = sky + halo + disk[..., None] * np.array([1.0, 0.95, 0.85])
    return img.astype(np.float32).astype(np.float64)


def random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def scene(seed, grid=1024, instances=500, subdiv=4, variants=16, env_size=(2048, 1024), fog=0.02,
          width=256, height=256, spp=32, depth=12):
    """One seeded scene: (DescBuilder, camera)."""
    rng = np.random.default_rng(seed)
    b = DescBuilder()

    def principled():
        return b.principled(tuple(rng.uniform(0.1, 0.9, 3)), float(rng.uniform() < 0.3),
                            float(rng.uniform(0.05, 1.0)), 1.5)
    tpos, ttri, hf = terrain_mesh(grid, seed)
    tf, tn = b.mesh_prims(tpos, ttri, principled())
    b.mesh(tf, tn)
    spos, stri = icosphere(subdiv)
    meshes = [b.mesh_prims(spos, stri, principled()) for _ in range(variants)]
    n = hf.shape[0]
    for _ in range(instances):
        first, count = meshes[rng.integers(variants)]
        r = random_rotation(rng)
        sc = rng.uniform(0.4, 2.0) * rng.uniform(0.7, 1.3, 3)
        m = r @ np.diag(sc)
        px, pz = rng.uniform(-TERRAIN_SIZE * 0.45, TERRAIN_SIZE * 0.45, 2)
        gi = int((px / TERRAIN_SIZE + 0.5) * (n - 1))
        gj = int((pz / TERRAIN_SIZE + 0.5) * (n - 1))
        py = hf[gj, gi] + 0.6 * sc.max()
        b.mesh(first, count, b.xform(m, (px, py, pz)))
    if env_size:
        b.environment(sun_sky(*env_size), as_light=True)
    if fog > 0:
        lo = (-TERRAIN_SIZE / 2 - 1, float(hf.min()) - 1, -TERRAIN_SIZE / 2 - 1)
        hi = (TERRAIN_SIZE / 2 + 1, float(hf.max()) + 20, TERRAIN_SIZE / 2 + 1)
        b.medium_box(lo, hi, fog, b.phase((0.9, 0.9, 0.9)))
    a = rng.uniform(0, 2 * math.pi)
    look_from = (38 * math.cos(a), float(hf.max()) + 12, 38 * math.sin(a))
    cam = camera(width, width / height, spp, depth, look_from, (0, float(np.median(hf)), 0), 50)
    return b, cam


def batch(seed=1, **kw):
    """The 8 scenes of the batch with their resolutions (a generator)."""
    for k, (w, h) in enumerate(RESOLUTIONS):
        yield scene(seed * 1000 + k, width=w, height=h, **kw)
