"""Deterministic synthetic scenes for the benchmark configurations.

Every generator returns triangles in the reference's ``basic_triangle<3, float>``
layout (``math/primitive.h:14-20``, ``math/triangle.h:34-36``): ``geom_id``,
``prim_id``, then ``v1``, ``e1 = v2 - v1``, ``e2 = v3 - v1`` as float32, 44 bytes
per triangle.  ``prim_id`` is the triangle's index; ``geom_id`` selects the
material.  All randomness comes from seeded numpy generators, so a scene is a
pure function of its arguments.
"""
from __future__ import annotations

import numpy as np

TRI_DTYPE = np.dtype([
    ("geom_id", "<u4"),
    ("prim_id", "<u4"),
    ("v1", "<f4", (3,)),
    ("e1", "<f4", (3,)),
    ("e2", "<f4", (3,)),
])
assert TRI_DTYPE.itemsize == 44

MAT_DTYPE = np.dtype([
    ("type", "<i4"),
    ("cd", "<f4", (3,)),
    ("kd", "<f4"),
    ("ce", "<f4", (3,)),
    ("ls", "<f4"),
])
assert MAT_DTYPE.itemsize == 36


def make_triangles(a: np.ndarray, b: np.ndarray, c: np.ndarray, geom_id=0) -> np.ndarray:
    """Pack corner arrays (n, 3) float32 into the 44-byte triangle layout."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    c = np.asarray(c, np.float32)
    n = a.shape[0]
    t = np.zeros(n, TRI_DTYPE)
    t["v1"] = a
    t["e1"] = b - a
    t["e2"] = c - a
    t["prim_id"] = np.arange(n, dtype=np.uint32)
    t["geom_id"] = geom_id
    return t


def concat(*parts: np.ndarray) -> np.ndarray:
    """Concatenate triangle arrays and renumber prim_id to the global index."""
    t = np.concatenate(parts)
    t["prim_id"] = np.arange(t.shape[0], dtype=np.uint32)
    return t


def room(half: float = 1.0) -> np.ndarray:
    """12 axis-aligned quads (24 triangles): the six walls of the cube
    [-half, half]^3, each wall split into two quads."""
    h = np.float32(half)
    quads = []
    for axis in range(3):
        u, v = [k for k in range(3) if k != axis]
        for side in (-h, h):
            for half_idx in range(2):
                lo = -h + half_idx * h
                hi = lo + h
                corners = []
                for (pu, pv) in ((lo, -h), (hi, -h), (hi, h), (lo, h)):
                    p = [0.0, 0.0, 0.0]
                    p[axis] = side
                    p[u] = pu
                    p[v] = pv
                    corners.append(p)
                quads.append(corners)
    q = np.asarray(quads, np.float32)
    a = np.concatenate([q[:, 0], q[:, 0]])
    b = np.concatenate([q[:, 1], q[:, 2]])
    c = np.concatenate([q[:, 2], q[:, 3]])
    return make_triangles(a, b, c)


def _icosahedron():
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([
        [-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0],
        [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
        [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1],
    ], dtype=np.float64)
    f = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, f


def icosphere_mesh(level: int):
    """Unit icosphere: (vertices float64 (V,3), faces int64 (20*4^level, 3))."""
    v, f = _icosahedron()
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1),
            np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1),
            np.stack([m01, m12, m20], 1),
        ])
    return v, f


def icosphere(level: int, radius: float, center=(0.0, 0.0, 0.0), jitter: float = 0.0,
              rng: np.random.Generator | None = None) -> np.ndarray:
    """Icosphere with 20*4^level triangles; each shared vertex is jittered once by
    U(-jitter, jitter) per axis, so the mesh stays watertight."""
    v, f = icosphere_mesh(level)
    v = v * radius + np.asarray(center, np.float64)
    if jitter:
        v = v + rng.uniform(-jitter, jitter, size=v.shape)
    v = v.astype(np.float32)
    return make_triangles(v[f[:, 0]], v[f[:, 1]], v[f[:, 2]])


def _value_noise(x: np.ndarray, z: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """Smooth value noise on the integer lattice (quintic fade), float64."""
    xi = np.floor(x).astype(np.int64)
    zi = np.floor(z).astype(np.int64)
    xf = x - xi
    zf = z - zi
    fx = xf * xf * xf * (xf * (xf * 6 - 15) + 10)
    fz = zf * zf * zf * (zf * (zf * 6 - 15) + 10)
    m = perm.shape[0] - 1

    def h(ix, iz):
        return perm[(perm[ix & m] + iz) & m] / float(m) * 2.0 - 1.0

    a = h(xi, zi)
    b = h(xi + 1, zi)
    c = h(xi, zi + 1)
    d = h(xi + 1, zi + 1)
    return (a * (1 - fx) + b * fx) * (1 - fz) + (c * (1 - fx) + d * fx) * fz


def heightfield(cells: int = 1448, octaves: int = 6, gain: float = 0.5, lacunarity: float = 2.0,
                seed: int = 1, extent: float = 1.0, amplitude: float = 0.25) -> np.ndarray:
    """cells x cells grid over [-extent, extent]^2 in x/z, y from fBm value noise;
    2 * cells^2 triangles."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(4096).astype(np.int64)
    g = np.linspace(-extent, extent, cells + 1, dtype=np.float64)
    X, Z = np.meshgrid(g, g, indexing="xy")
    Y = np.zeros_like(X)
    freq, amp = 4.0, 1.0
    for _ in range(octaves):
        Y += amp * _value_noise(X * freq + 17.3, Z * freq + 91.7, perm)
        freq *= lacunarity
        amp *= gain
    Y *= amplitude
    P = np.stack([X, Y, Z], axis=-1).astype(np.float32)
    p00 = P[:-1, :-1].reshape(-1, 3)
    p10 = P[:-1, 1:].reshape(-1, 3)
    p01 = P[1:, :-1].reshape(-1, 3)
    p11 = P[1:, 1:].reshape(-1, 3)
    a = np.concatenate([p00, p10])
    b = np.concatenate([p10, p11])
    c = np.concatenate([p01, p01])
    # interleave the two triangles of each cell so neighbours stay adjacent
    n = p00.shape[0]
    order = np.empty(2 * n, np.int64)
    order[0::2] = np.arange(n)
    order[1::2] = np.arange(n) + n
    return make_triangles(a[order], b[order], c[order])


def icosphere_cloud(count: int, level: int, rmin: float, rmax: float, lo, hi, seed: int) -> np.ndarray:
    """`count` icospheres with log-uniform radius in [rmin, rmax] and centres
    uniform in the box [lo, hi]."""
    rng = np.random.default_rng(seed)
    v, f = icosphere_mesh(level)
    radii = np.exp(rng.uniform(np.log(rmin), np.log(rmax), size=count))
    centers = rng.uniform(np.asarray(lo, np.float64), np.asarray(hi, np.float64), size=(count, 3))
    verts = (v[None, :, :] * radii[:, None, None] + centers[:, None, :]).astype(np.float32)
    tri = verts[:, f]  # (count, F, 3 corners, 3)
    tri = tri.reshape(-1, 3, 3)
    return make_triangles(tri[:, 0], tri[:, 1], tri[:, 2])


def random_clusters(n: int, seed: int, clusters: int = 16, sigma: float = 0.25,
                    emin: float = 1e-3, emax: float = 1e-1) -> np.ndarray:
    """n random triangles around Gaussian cluster centres, edge lengths
    log-uniform in [emin, emax]."""
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, 1.0, size=(clusters, 3))
    which = rng.integers(0, clusters, size=n)
    p = centers[which] + rng.normal(0.0, sigma, size=(n, 3))
    length = np.exp(rng.uniform(np.log(emin), np.log(emax), size=(n, 1)))
    d1 = rng.normal(size=(n, 3))
    d2 = rng.normal(size=(n, 3))
    d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    a = p.astype(np.float32)
    b = (p + d1 * length).astype(np.float32)
    c = (p + d2 * length).astype(np.float32)
    return make_triangles(a, b, c)


def lambert_materials(n: int, seed: int, lo: float = 0.2, hi: float = 0.8,
                      emissive_fraction: float = 0.0, emission: float = 4.0) -> np.ndarray:
    """n materials: matte with RGB albedo U(lo, hi), kd = 1; a seeded
    `emissive_fraction` of them emissive with ce = (1, 1, 1), ls = `emission`."""
    rng = np.random.default_rng(seed)
    m = np.zeros(n, MAT_DTYPE)
    m["type"] = 0
    m["cd"] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    m["kd"] = 1.0
    if emissive_fraction > 0.0:
        k = int(round(n * emissive_fraction))
        idx = rng.choice(n, size=k, replace=False)
        m["type"][idx] = 1
        m["ce"][idx] = 1.0
        m["ls"][idx] = emission
    return m


# ---------------------------------------------------------------------------
# benchmark configurations
# ---------------------------------------------------------------------------

def config0_scene(level: int = 5, seed: int = 0) -> np.ndarray:
    """Room (24 triangles) plus an icosphere of radius 0.3, vertices jittered by
    U(+-0.002): 20,504 triangles at level 5."""
    rng = np.random.default_rng(seed)
    return concat(room(1.0), icosphere(level, 0.3, jitter=0.002, rng=rng))


# area_light<float, basic_triangle<3, float>> (area_light.h): cl, kl, triangle
LIGHT_DTYPE = np.dtype([("cl", "<f4", (3,)), ("kl", "<f4"), ("geometry", TRI_DTYPE)])
assert LIGHT_DTYPE.itemsize == 60


def lights_from_emissive(tris: np.ndarray, mats: np.ndarray) -> np.ndarray:
    """One area light per triangle whose material is emissive, in triangle
    order, with cl = ce and kl = ls (as the reference's viewer builds its
    light list from the emissive triangles, src/viewer/viewer.cpp:3028-3083)."""
    m = mats[np.minimum(tris["geom_id"], mats.size - 1)]
    sel = m["type"] == 1
    out = np.zeros(int(sel.sum()), LIGHT_DTYPE)
    out["geometry"] = tris[sel]
    out["cl"] = m["ce"][sel]
    out["kl"] = m["ls"][sel]
    return out


def lit_room_scene(level: int = 4, light_half: float = 0.25, emission: float = 8.0, seed: int = 0):
    """The config-0 room and icosphere lit by a square ceiling light of half
    size `light_half` just below the ceiling (two emissive triangles facing
    down).  Materials: 6 matte walls/sphere colours (geom 0-5), the light
    (geom 6).  Returns (tris, mats, lights)."""
    base = config0_scene(level=level, seed=seed)
    base["geom_id"] = np.arange(base.size) % 6
    h, y = np.float32(light_half), np.float32(0.98)
    a = np.array([[-h, y, -h], [-h, y, -h]], np.float32)
    b = np.array([[h, y, h], [-h, y, h]], np.float32)
    c = np.array([[h, y, -h], [h, y, h]], np.float32)
    light = make_triangles(a, b, c, geom_id=6)
    tris = concat(base, light)
    mats = lambert_materials(7, seed=seed + 5)
    mats["type"][6] = 1
    mats["ce"][6] = (1.0, 0.9, 0.75)
    mats["ls"][6] = emission
    return tris, mats, lights_from_emissive(tris, mats)


def config1_scene(cells: int = 1448, spheres: int = 2000, sphere_level: int = 3) -> np.ndarray:
    """fBm heightfield (2 * cells^2 triangles, seed 1) plus `spheres` level-3
    icospheres (radius log-U[0.005, 0.05], seed 2) over the heightfield box:
    6,753,408 triangles at the default sizes.  geom_id = prim_id (one material
    per triangle)."""
    hf = heightfield(cells, seed=1)
    parts = [hf]
    if spheres:
        v = np.concatenate([hf["v1"], hf["v1"] + hf["e1"], hf["v1"] + hf["e2"]])
        lo, hi = v.min(0), v.max(0)
        hi = hi.copy()
        hi[1] += 0.1
        parts.append(icosphere_cloud(spheres, sphere_level, 0.005, 0.05, lo, hi, seed=2))
    t = concat(*parts)
    t["geom_id"] = t["prim_id"]
    return t


def config1_materials(n: int) -> np.ndarray:
    """Lambertian albedo U(0.2, 0.8) per triangle, 1% emissive (seed 3)."""
    return lambert_materials(n, seed=3, emissive_fraction=0.01, emission=4.0)


def config1_camera():
    """Pinhole camera over the heightfield: eye, center, up, fovy (rad), aspect."""
    return (0.0, 0.55, 1.45), (0.0, -0.05, 0.0), (0.0, 1.0, 0.0), np.float32(45.0 * np.pi / 180.0), 1920.0 / 1080.0


def config0_camera(width: int = 512, height: int = 512):
    """Camera inside the room looking at the icosphere."""
    return (0.2, 0.3, 0.95), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), np.float32(45.0 * np.pi / 180.0), width / height


RESOLUTIONS = ((256, 256), (512, 512), (1024, 1024), (1920, 1080))


def config2_scenes(count: int = 64, nmin: int = 1_000, nmax: int = 2_000_000, seed: int = 4,
                   resolutions=RESOLUTIONS):
    """`count` independent scenes of seeded random-triangle clusters, sizes
    log-uniform in [nmin, nmax], each with its own camera and resolution.
    Yields dicts with tris, mats (8 Lambertian materials, one emissive),
    camera (eye, center, up, fovy, aspect) and (width, height)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count)).astype(np.int64)
    for i, n in enumerate(sizes):
        tris = random_clusters(int(n), seed=seed * 1000 + i, clusters=int(rng.integers(4, 32)), sigma=0.3)
        tris["geom_id"] = rng.integers(0, 8, size=tris.size)
        mats = lambert_materials(8, seed=seed * 1000 + i, emissive_fraction=0.125)
        w, h = resolutions[int(rng.integers(0, len(resolutions)))]
        v = tris["v1"]
        c = np.median(v, axis=0)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = c + d * float(rng.uniform(2.5, 4.0))
        yield dict(tris=tris, mats=mats, camera=(tuple(eye), tuple(c), (0.0, 1.0, 0.0), np.float32(0.8), w / h),
                   size=(w, h))


def config3_scene(level: int = 8, num_materials: int = 256, num_emissive: int = 8, seed: int = 0):
    """Room plus an icosphere of level `level` (1,310,720 triangles at 8) with
    `num_materials` materials: albedo U(0.1, 0.9) (seed 5); the last
    `num_emissive` materials are emissive, ce U(0.5, 1.5), ls = 1.  Room quads
    take materials 0..11 except the two ceiling quads, which take the first two
    emissive materials; sphere triangles take a material per latitude/longitude
    cell, six cells emissive.  Returns (tris, mats)."""
    rng = np.random.default_rng(seed)
    room_t = room(1.0)
    sph = icosphere(level, 0.3, jitter=0.002, rng=rng)
    nm, ne = num_materials, num_emissive
    first_e = nm - ne
    # room: 12 quads, triangles q and q + 12 belong to quad q
    q = np.arange(24) % 12
    ceiling = []
    for k in range(12):
        ys = np.concatenate([room_t["v1"][[k, k + 12], 1]])
        if np.all(ys > 0.999):
            ceiling.append(k)
    room_m = q.copy()
    for j, k in enumerate(ceiling[:2]):
        room_m[q == k] = first_e + j
    room_t["geom_id"] = room_m
    # sphere: latitude/longitude cells over the centroid direction
    c = sph["v1"] + (sph["e1"] + sph["e2"]) / 3.0
    theta = np.arccos(np.clip(c[:, 1] / np.linalg.norm(c, axis=1), -1, 1))
    phi = np.arctan2(c[:, 2], c[:, 0]) + np.pi
    ncell = first_e - 12
    nlat = max(1, int(round(np.sqrt(ncell / 2))))
    nlon = ncell // nlat
    cell = np.minimum((theta / np.pi * nlat).astype(np.int64), nlat - 1) * nlon + \
        np.minimum((phi / (2 * np.pi) * nlon).astype(np.int64), nlon - 1)
    sph_m = 12 + cell
    emissive_cells = rng.choice(nlat * nlon, size=ne - 2, replace=False)
    for j, ec in enumerate(emissive_cells):
        sph_m[cell == ec] = first_e + 2 + j
    sph["geom_id"] = sph_m
    tris = concat(room_t, sph)
    mats = lambert_materials(nm, seed=5, lo=0.1, hi=0.9)
    erng = np.random.default_rng(6)
    mats["type"][first_e:] = 1
    mats["ce"][first_e:] = erng.uniform(0.5, 1.5, size=(ne, 3))
    mats["ls"][first_e:] = 1.0
    return tris, mats


def displace(tris: np.ndarray, frame: int, amplitude: float = 0.01, seed: int = 0) -> np.ndarray:
    """Per-frame sinusoidal vertex displacement (dynamic geometry): every
    corner moves by amplitude * sin(k . p + frame * w + phase) along a seeded
    direction, so shared corners move together."""
    rng = np.random.default_rng(seed)
    k = rng.normal(size=3) * 8.0
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    wfreq = float(rng.uniform(0.5, 1.5))
    out = tris.copy()
    v1 = tris["v1"].astype(np.float64)
    v2 = v1 + tris["e1"]
    v3 = v1 + tris["e2"]

    def move(p):
        return p + amplitude * np.sin(p @ k + frame * wfreq)[:, None] * direction[None, :]

    a, b, c = move(v1).astype(np.float32), move(v2).astype(np.float32), move(v3).astype(np.float32)
    out["v1"] = a
    out["e1"] = b - a
    out["e2"] = c - a
    return out


def config4_scenes(count: int = 32, nmin: int = 50_000, nmax: int = 1_000_000, seed: int = 7):
    """`count` scenes of random-triangle clusters with 50k-1M triangles and 64
    materials each (one in eight emissive), camera per scene, 256 x 256."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count)).astype(np.int64)
    for i, n in enumerate(sizes):
        tris = random_clusters(int(n), seed=seed * 1000 + i, clusters=int(rng.integers(4, 16)), sigma=0.3)
        tris["geom_id"] = rng.integers(0, 64, size=tris.size)
        mats = lambert_materials(64, seed=seed * 1000 + i, lo=0.1, hi=0.9, emissive_fraction=0.125, emission=2.0)
        c = np.median(tris["v1"], axis=0)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        yield dict(tris=tris, mats=mats, camera=(tuple(c + d * 3.0), tuple(c), (0.0, 1.0, 0.0), np.float32(0.8), 1.0),
                   size=(256, 256))


RAY_DTYPE = np.dtype([("ori", "<f4", (3,)), ("dir", "<f4", (3,)), ("tmin", "<f4"), ("tmax", "<f4")])

# instance records of a two-level BVH: bottom-level index, id, mat4x3 as 12
# floats column by column (three columns of the 3x3 part, then translation)
INSTANCE_DTYPE = np.dtype([("blas", "<u4"), ("inst_id", "<i4"), ("transform", "<f4", (12,))])


def random_transform(rng, scale_range=(0.5, 1.5), shear: float = 0.2, extent: float = 2.0) -> np.ndarray:
    """A random affine mat4x3 (12 floats): rotation x per-axis scale (one in
    four mirrored) plus a small shear, and a translation in [-extent, extent]^3."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    rot = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                    [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                    [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    s = rng.uniform(*scale_range, size=3)
    if rng.uniform() < 0.25:
        s[0] = -s[0]
    m = rot @ np.diag(s) + rng.uniform(-shear, shear, size=(3, 3))
    t = rng.uniform(-extent, extent, size=3)
    return np.concatenate([m[:, 0], m[:, 1], m[:, 2], t]).astype(np.float32)


def instanced_scene(num_instances: int = 24, seed: int = 11, sphere_level: int = 2, patch_cells: int = 10,
                    soup: int = 120):
    """Three bottom levels (an icosphere, a height-field patch, a triangle
    soup) and `num_instances` instances of them under random affine
    transforms -> (list of triangle arrays, instances)."""
    rng = np.random.default_rng(seed)
    b0 = icosphere(sphere_level, 0.5)
    b1 = heightfield(cells=patch_cells, seed=seed, extent=0.6, amplitude=0.2)
    c = rng.uniform(-0.5, 0.5, size=(soup, 3))
    e1 = rng.normal(scale=0.08, size=(soup, 3))
    e2 = rng.normal(scale=0.08, size=(soup, 3))
    b2 = make_triangles(c, c + e1, c + e2)
    blas = [b0, b1, b2]
    for k, b in enumerate(blas):
        b["geom_id"] = k
    inst = np.zeros(num_instances, INSTANCE_DTYPE)
    inst["blas"] = rng.integers(0, len(blas), size=num_instances)
    inst["inst_id"] = 100 + np.arange(num_instances)
    for i in range(num_instances):
        inst["transform"][i] = random_transform(rng)
    return blas, inst


def instanced_rays(n: int, seed: int, extent: float = 3.0) -> np.ndarray:
    """Rays from points in [-extent, extent]^3 in uniform directions (unit
    length), tmax = FLT_MAX."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY_DTYPE)
    r["ori"] = rng.uniform(-extent, extent, size=(n, 3))
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmax"] = np.finfo(np.float32).max
    return r


def config1_instanced(cells: int = 1448, spheres: int = 2000, sphere_level: int = 3):
    """The C1 scene as a two-level BVH: the height field as one identity
    instance, and one unit level-3 icosphere instanced `spheres` times with the
    radii and centres config1_scene draws (same seeds) -> (bottom-level
    triangle arrays, instances).  World vertices equal config1_scene's up to
    float rounding (radius x vertex + centre is evaluated per ray here)."""
    hf = heightfield(cells, seed=1)
    v = np.concatenate([hf["v1"], hf["v1"] + hf["e1"], hf["v1"] + hf["e2"]])
    lo, hi = v.min(0), v.max(0)
    hi = hi.copy()
    hi[1] += 0.1
    rng = np.random.default_rng(2)  # the draws of icosphere_cloud(..., seed=2)
    uv, f = icosphere_mesh(sphere_level)
    radii = np.exp(rng.uniform(np.log(0.005), np.log(0.05), size=spheres))
    centers = rng.uniform(np.asarray(lo, np.float64), np.asarray(hi, np.float64), size=(spheres, 3))
    unit = uv.astype(np.float32)[f]
    sphere = make_triangles(unit[:, 0], unit[:, 1], unit[:, 2], geom_id=1)
    inst = np.zeros(spheres + 1, INSTANCE_DTYPE)
    inst["transform"][0] = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
    inst["blas"][1:] = 1
    inst["inst_id"] = np.arange(spheres + 1)
    t = np.zeros((spheres, 12), np.float32)
    t[:, 0] = t[:, 4] = t[:, 8] = radii
    t[:, 9:12] = centers
    inst["transform"][1:] = t
    return [hf, sphere], inst
