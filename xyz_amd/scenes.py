"""Deterministic synthetic scenes and ray batches for the workloads this package
is measured on.

Every generator returns plain numpy float32/uint32 arrays; triangle order is the
caller order that defines primitive ids.  Winding: the renderer's geometric
normal is cross(v2 - v0, v1 - v0) (xyz/geom/triangle.cpp:20), so triangles whose
outside matters (glass, lights) are emitted with that normal pointing outward.
"""
import math

import numpy as np

LAMBERT, MIRROR, GLASS, METAL, LIGHT = 0, 1, 2, 3, 4


# --------------------------------------------------------------- helpers
def _orient(v0, v1, v2, outward):
    """Swap v1/v2 where cross(v2-v0, v1-v0) does not point along `outward`."""
    n = np.cross(v2 - v0, v1 - v0)
    flip = (n * outward).sum(axis=1) < 0
    a, b = v1.copy(), v2.copy()
    a[flip], b[flip] = v2[flip], v1[flip]
    return v0, a, b


def quad(p, u, v, outward):
    """Two triangles spanning p, p+u, p+u+v, p+v with normal along `outward`."""
    p, u, v = (np.asarray(x, np.float64) for x in (p, u, v))
    a, b, c, d = p, p + u, p + u + v, p + v
    v0 = np.array([a, a], np.float64)
    v1 = np.array([b, c], np.float64)
    v2 = np.array([c, d], np.float64)
    out = np.tile(np.asarray(outward, np.float64), (2, 1))
    return _orient(v0, v1, v2, out)


def box(center, size, rot=None, tess=1):
    """Axis box (size per axis) rotated by 3x3 `rot`, each face split into
    tess x tess x 2 triangles, outward normals."""
    center = np.asarray(center, np.float64)
    h = np.asarray(size, np.float64) / 2
    rot = np.eye(3) if rot is None else np.asarray(rot, np.float64)
    ii, jj = np.meshgrid(np.arange(tess), np.arange(tess), indexing="ij")
    ii, jj = ii.reshape(-1, 1), jj.reshape(-1, 1)
    V0, V1, V2 = [], [], []
    for ax in range(3):
        for sgn in (-1.0, 1.0):
            n = np.zeros(3)
            n[ax] = sgn
            a1, a2 = (ax + 1) % 3, (ax + 2) % 3
            e_u = np.zeros(3)
            e_u[a1] = 2 * h[a1] / tess
            e_v = np.zeros(3)
            e_v[a2] = 2 * h[a2] / tess
            base = np.zeros(3)
            base[ax] = sgn * h[ax]
            base[a1] = -h[a1]
            base[a2] = -h[a2]
            p = base + ii * e_u + jj * e_v
            a, b, c, d = p, p + e_u, p + e_u + e_v, p + e_v
            v0 = np.concatenate([a, a])
            v1 = np.concatenate([b, c])
            v2 = np.concatenate([c, d])
            v0, v1, v2 = _orient(v0, v1, v2, np.tile(n, (len(v0), 1)))
            V0.append(v0)
            V1.append(v1)
            V2.append(v2)
    V0, V1, V2 = (np.concatenate(x) for x in (V0, V1, V2))
    V0, V1, V2 = (x @ rot.T + center for x in (V0, V1, V2))
    return V0, V1, V2


def boxes_y(centers, sizes, yaw_deg):
    """Many boxes at once: centres (N,3), sizes (N,3), rotation about y (N,) in
    degrees; 12 triangles per box in box order, outward normals."""
    unit = box((0, 0, 0), (1, 1, 1))  # 12 triangles of the unit cube
    centers = np.asarray(centers, np.float64)
    sizes = np.asarray(sizes, np.float64)
    a = np.radians(np.asarray(yaw_deg, np.float64))
    c, s = np.cos(a), np.sin(a)
    out = []
    for v in unit:
        p = v[None, :, :] * sizes[:, None, :]  # (N, 12, 3)
        x = c[:, None] * p[..., 0] + s[:, None] * p[..., 2]
        z = -s[:, None] * p[..., 0] + c[:, None] * p[..., 2]
        q = np.stack([x, p[..., 1], z], -1) + centers[:, None, :]
        out.append(q.reshape(-1, 3))
    return out[0], out[1], out[2]


def rot_y(deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


# -------------------------------------------------------- icosphere + noise
def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def icosphere(level):
    """Unit icosphere subdivided `level` times (20 * 4**level faces), shared vertices."""
    v, f = _icosahedron()
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (len(v) + 1)
        b = uniq % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
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


def geodesic_sphere(freq):
    """Unit class-I geodesic sphere: each icosahedron face split into freq**2
    triangles (20 * freq**2 faces); vertices on shared edges are computed from
    the same corner pair in the same order, so neighbouring faces meet exactly."""
    v, f = _icosahedron()
    n = int(freq)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    mask = (ii + jj) <= n
    I, J = ii[mask], jj[mask]
    idx = -np.ones((n + 1, n + 1), np.int64)
    idx[I, J] = np.arange(len(I))
    up_i, up_j = np.nonzero((ii + jj) <= n - 1)
    dn_i, dn_j = np.nonzero((ii + jj) <= n - 2)
    tri = np.concatenate([
        np.stack([idx[up_i, up_j], idx[up_i + 1, up_j], idx[up_i, up_j + 1]], 1),
        np.stack([idx[dn_i + 1, dn_j], idx[dn_i + 1, dn_j + 1], idx[dn_i, dn_j + 1]], 1)])
    npts = len(I)
    V, F = [], []
    for fi, (a, b, c) in enumerate(f):
        A, B, C = v[a], v[b], v[c]
        P = (np.outer(n - I - J, A) + np.outer(I, B) + np.outer(J, C)) / n
        P /= np.linalg.norm(P, axis=1, keepdims=True)
        V.append(P)
        F.append(tri + fi * npts)
    return np.concatenate(V), np.concatenate(F)


def _lattice(ix, iy, iz, seed):
    """Seeded integer hash of lattice points -> uniform in [-1, 1]."""
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
         ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ np.uint64(seed) * np.uint64(0x9E3779B9))
    h &= np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x846CA68B)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    return h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def value_noise(p, freq, seed):
    q = p * freq
    i = np.floor(q)
    fr = q - i
    w = fr * fr * (3 - 2 * fr)
    i = i.astype(np.int64) + (1 << 20)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c = _lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out += c * wx * wy * wz
    return out


def noise3(p, seed):
    """3-octave value noise (frequencies 4, 8, 16; amplitudes 1, 1/2, 1/4) in [-1, 1]."""
    n = (value_noise(p, 4.0, seed) + 0.5 * value_noise(p, 8.0, seed + 1)
         + 0.25 * value_noise(p, 16.0, seed + 2))
    return n / 1.75


def displaced_sphere(v, f, amplitude=0.05, seed=0, center=(0, 0, 0), radius=1.0):
    v = v * (1.0 + amplitude * noise3(v, seed))[:, None] * radius + np.asarray(center)
    v0, v1, v2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    out = (v0 + v1 + v2) / 3 - np.asarray(center)
    return _orient(v0, v1, v2, out)


def _pack(parts):
    v0 = np.concatenate([p[0] for p in parts]).astype(np.float32)
    v1 = np.concatenate([p[1] for p in parts]).astype(np.float32)
    v2 = np.concatenate([p[2] for p in parts]).astype(np.float32)
    return v0, v1, v2


# ------------------------------------------------------------ workloads
def cornell_box(textured=False, tex_res=64, tex_seed=0):
    """Cornell box: unit cube [0,1]^3 with the z=0 face open, 5 Lambert walls,
    a 0.3^3 box rotated -18 deg and a 0.3x0.6x0.3 box rotated +18 deg about y,
    and a 0.25x0.25 quad light at y=0.999 (L_e = 15).  36 triangles.

    With `textured`, each wall gets a tex_res^2 RGB albedo texture with texels
    uniform in [0.1, 0.9] and UVs spanning the wall."""
    white, red, green = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)
    parts, bs = [], []
    uvs = []
    walls = [  # origin, u, v, inward normal, bsdf
        ((0, 0, 0), (1, 0, 0), (0, 0, 1), (0, 1, 0), 0),   # floor
        ((0, 1, 0), (1, 0, 0), (0, 0, 1), (0, -1, 0), 0),  # ceiling
        ((0, 0, 1), (1, 0, 0), (0, 1, 0), (0, 0, -1), 0),  # back
        ((0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0), 1),   # left x=0 (red)
        ((1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), 2),  # right x=1 (green)
    ]
    for wi, (p, u, v, n, b) in enumerate(walls):
        v0, v1, v2 = quad(p, u, v, n)
        parts.append((v0, v1, v2))
        bsdf = (3 + wi) if textured else b
        bs += [bsdf, bsdf]
        # uv of each vertex: position relative to the wall frame
        for tv in (v0, v1, v2):
            pass
        P, U, Vv = (np.asarray(x, np.float64) for x in (p, u, v))
        tri_uv = []
        for t in range(2):
            row = []
            for vert in (v0[t], v1[t], v2[t]):
                d = vert - P
                row += [float(d @ U), float(d @ Vv)]
            tri_uv.append(row)
        uvs += tri_uv
    for center, size, ang in (((0.30, 0.15, 0.30), (0.3, 0.3, 0.3), -18.0),
                              ((0.68, 0.30, 0.66), (0.3, 0.6, 0.3), 18.0)):
        parts.append(box(center, size, rot_y(ang)))
        bs += [0] * 12
        uvs += [[0.0] * 6] * 12
    lv = quad((0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25), (0, -1, 0))
    parts.append(lv)
    light_bsdf = 3 if not textured else 8
    bs += [light_bsdf, light_bsdf]
    uvs += [[0.0] * 6] * 2
    v0, v1, v2 = _pack(parts)
    bsdfs = [dict(type=LAMBERT, rgb=white), dict(type=LAMBERT, rgb=red),
             dict(type=LAMBERT, rgb=green)]
    textures = None
    if textured:
        rng = np.random.default_rng(tex_seed)
        textures = [rng.uniform(0.1, 0.9, size=(tex_res, tex_res, 3)).astype(np.float32)
                    for _ in range(5)]
        for wi in range(5):
            bsdfs.append(dict(type=LAMBERT, rgb=(1.0, 1.0, 1.0), texture=wi))
    bsdfs.append(dict(type=LIGHT, rgb=(15.0, 15.0, 15.0)))
    if not textured:
        bsdfs = bsdfs[:3] + [bsdfs[-1]]
    camera = np.array([0.5, 0.5, -1.4, 0.5, 0.5, 0.5, 0, 1, 0, 40.0], np.float32)
    return dict(v0=v0, v1=v1, v2=v2, tri_bsdf=np.array(bs, np.uint32), bsdfs=bsdfs,
                uv=np.array(uvs, np.float32), textures=textures, camera=camera)


def closest_hit_mesh(level=8, n_boxes=64, seed=1):
    """Icosphere of `level` (20*4**level tris) radially displaced by 0.05 x
    3-octave value noise, plus `n_boxes` random boxes (12 tris each) with
    centres in [-2,2]^3, extents in [0.05, 0.3] and random rotations."""
    v, f = icosphere(level)
    parts = [displaced_sphere(v, f, 0.05, seed)]
    rng = np.random.default_rng(seed)
    for _ in range(n_boxes):
        c = rng.uniform(-2, 2, 3)
        s = rng.uniform(0.05, 0.3, 3)
        parts.append(box(c, s, random_rotation(rng)))
    v0, v1, v2 = _pack(parts)
    return dict(v0=v0, v1=v1, v2=v2)


def pinhole_rays(eye, at, fov_deg, W, H, up=(0, 1, 0)):
    """One ray per pixel centre of a W x H pinhole camera (row 0 = bottom).
    Directions follow the renderer's pinhole (film 0.025, V = 0.0125/tan(fov/2))."""
    eye = np.asarray(eye, np.float32)
    at = np.asarray(at, np.float32)
    up = np.asarray(up, np.float32)

    def nrm(a):
        return a * (np.float32(1) / np.sqrt((a * a).sum(-1, keepdims=True, dtype=np.float32)))
    back = nrm(eye - at)
    right = nrm(np.cross(up, back).astype(np.float32))
    upv = np.cross(back, right).astype(np.float32)
    F = np.float32(0.025)
    A = np.float32(W) / np.float32(H)
    V = np.float32(0.0125 / math.tan(math.radians(fov_deg) / 2))
    xs = ((np.arange(W, dtype=np.float32) + np.float32(0.5)) / np.float32(W) - np.float32(0.5)) * F * A
    ys = ((np.arange(H, dtype=np.float32) + np.float32(0.5)) / np.float32(H) - np.float32(0.5)) * F
    d = (xs[None, :, None] * right[None, None, :] + ys[:, None, None] * upv[None, None, :]
         - V * back[None, None, :]).astype(np.float32)
    d = nrm(d.reshape(-1, 3)).astype(np.float32)
    o = np.broadcast_to(eye, d.shape).astype(np.float32)
    return np.ascontiguousarray(o), np.ascontiguousarray(d)


def incoherent_rays(n, bmin, bmax, seed=2):
    """Origins uniform in [bmin, bmax], directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(bmin, bmax, size=(n, 3)).astype(np.float32)
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(np.maximum(0, 1 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
    return o, d


def closest_hit_rays(mesh, W=4096, H=2048, n_incoherent=None, seed=2):
    """Coherent primary rays from a W x H pinhole at (0,0,-3.5) looking at the
    origin (40 deg), plus as many incoherent rays inside the mesh's AABB."""
    oc, dc = pinhole_rays((0, 0, -3.5), (0, 0, 0), 40.0, W, H)
    lo = np.minimum(np.minimum(mesh["v0"], mesh["v1"]), mesh["v2"]).min(0)
    hi = np.maximum(np.maximum(mesh["v0"], mesh["v1"]), mesh["v2"]).max(0)
    n_inc = W * H if n_incoherent is None else n_incoherent
    oi, di = incoherent_rays(n_inc, lo, hi, seed)
    return (oc, dc), (oi, di)


def _material(rng):
    """70 % Lambert (albedo uniform in [0.1, 0.9]^3), 20 % mirror, 10 % glass."""
    u = rng.random()
    if u < 0.7:
        return dict(type=LAMBERT, rgb=tuple(float(x) for x in rng.uniform(0.1, 0.9, 3)))
    if u < 0.9:
        return dict(type=MIRROR)
    return dict(type=GLASS, ior=1.5)


def _assemble(objects, bsdfs):
    """objects: list of ((v0, v1, v2), bsdf index)."""
    v0 = np.concatenate([o[0][0] for o in objects]).astype(np.float32)
    v1 = np.concatenate([o[0][1] for o in objects]).astype(np.float32)
    v2 = np.concatenate([o[0][2] for o in objects]).astype(np.float32)
    tb = np.concatenate([np.full(len(o[0][0]), o[1], np.uint32) for o in objects])
    return dict(v0=v0, v1=v1, v2=v2, tri_bsdf=tb, bsdfs=bsdfs)


def production_scene(sphere_freq=274, n_boxes=512, box_tess=9, seed=3):
    """Displaced geodesic sphere (20*freq^2 tris; 1,501,520 at freq 274) of radius
    1.5 resting on a 2-triangle ground, 512 boxes with faces split into
    box_tess^2 x 2 triangles (972 each), 4 quad lights with emission uniform in
    [5, 20].  Each object draws its material: 70 % Lambert, 20 % mirror, 10 % glass.
    Camera: pinhole, 40 deg vertical field of view."""
    rng = np.random.default_rng(seed)
    bsdfs, objects = [], []

    def add_mat(m):
        bsdfs.append(m)
        return len(bsdfs) - 1

    v, f = geodesic_sphere(sphere_freq)
    objects.append((displaced_sphere(v, f, 0.05, seed, center=(0, 1.5, 0), radius=1.5),
                    add_mat(_material(rng))))
    objects.append((quad((-12, 0, -12), (24, 0, 0), (0, 0, 24), (0, 1, 0)),
                    add_mat(dict(type=LAMBERT, rgb=(0.5, 0.5, 0.5)))))
    placed = 0
    while placed < n_boxes:
        c = rng.uniform(-10, 10, 2)
        if np.hypot(c[0], c[1]) < 2.2:
            continue
        s = rng.uniform(0.2, 0.7, 3)
        objects.append((box((c[0], s[1] / 2, c[1]), s, rot_y(rng.uniform(0, 90)), tess=box_tess),
                        add_mat(_material(rng))))
        placed += 1
    for i in range(4):
        ang = i * np.pi / 2 + 0.4
        p = np.array([5 * np.cos(ang), rng.uniform(5, 7), 5 * np.sin(ang)])
        sz = rng.uniform(1.0, 2.0)
        le = float(rng.uniform(5, 20))
        objects.append((quad(p - np.array([sz / 2, 0, sz / 2]), (sz, 0, 0), (0, 0, sz), (0, -1, 0)),
                        add_mat(dict(type=LIGHT, rgb=(le, le, le)))))
    sc = _assemble(objects, bsdfs)
    sc["camera"] = np.array([0.0, 3.0, -7.5, 0.0, 1.2, 0.0, 0, 1, 0, 40.0], np.float32)
    return sc


def batch_scene(seed):
    """One job of the batched workload: a displaced sphere plus boxes with about
    exp(U[ln 1e3, ln 5e5]) triangles, materials as production_scene, two lights,
    resolution from {128^2, 256^2, 512^2, 1024^2}, spp from {16, 64, 256}."""
    rng = np.random.default_rng(1000 + seed)
    target = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
    res = int(rng.choice([128, 256, 512, 1024]))
    spp = int(rng.choice([16, 64, 256]))
    bsdfs, objects = [], []

    def add_mat(m):
        bsdfs.append(m)
        return len(bsdfs) - 1

    freq = max(2, int(round(np.sqrt(0.8 * target / 20))))
    v, f = geodesic_sphere(freq)
    objects.append((displaced_sphere(v, f, 0.05, seed, center=(0, 1, 0), radius=1.0),
                    add_mat(_material(rng))))
    objects.append((quad((-6, 0, -6), (12, 0, 0), (0, 0, 12), (0, 1, 0)),
                    add_mat(dict(type=LAMBERT, rgb=(0.5, 0.5, 0.5)))))
    n_boxes = max(1, int(0.2 * target / 12))
    c = rng.uniform(-5, 5, (n_boxes, 2))
    r = np.hypot(c[:, 0], c[:, 1])
    near = r < 1.3
    c[near] = c[near] / np.maximum(r[near], 1e-3)[:, None] * 1.5
    sz = rng.uniform(0.1, 0.5, (n_boxes, 3))
    yaw = rng.uniform(0, 90, n_boxes)
    mats = [add_mat(_material(rng)) for _ in range(n_boxes)]
    bv = boxes_y(np.stack([c[:, 0], sz[:, 1] / 2, c[:, 1]], 1), sz, yaw)
    for k in range(n_boxes):
        objects.append(((bv[0][12 * k:12 * k + 12], bv[1][12 * k:12 * k + 12],
                         bv[2][12 * k:12 * k + 12]), mats[k]))
    for i in range(2):
        p = np.array([(-2.5 if i == 0 else 2.0), 4.0, (-1.0 if i == 0 else 1.5)])
        le = float(rng.uniform(5, 20))
        objects.append((quad(p - 0.5, (1.0, 0, 0), (0, 0, 1.0), (0, -1, 0)),
                        add_mat(dict(type=LIGHT, rgb=(le, le, le)))))
    sc = _assemble(objects, bsdfs)
    sc["camera"] = np.array([0.0, 2.0, -5.0, 0.0, 0.8, 0.0, 0, 1, 0, 40.0], np.float32)
    sc["width"] = sc["height"] = res
    sc["spp"] = spp
    return sc


# ------------------------------------------- importance-driven path tracing
def mirror_ceiling_scene(albedo=(0.2, 0.5, 0.8), light_size=0.02, extent=1000.0):
    """Mirror floor (y = 0, facing up) under a Lambert ceiling (y = 1, facing
    down) of `albedo`, both 2 * extent wide, and a small square light at
    y = 0.5 facing down.  Light paths bounce between mirror and ceiling, so the
    ceiling holds an importon for every visit."""
    e = extent
    parts = [quad((-e, 0, -e), (2 * e, 0, 0), (0, 0, 2 * e), (0, 1, 0)),
             quad((-e, 1, -e), (2 * e, 0, 0), (0, 0, 2 * e), (0, -1, 0)),
             quad((-light_size / 2, 0.5, -light_size / 2), (light_size, 0, 0),
                  (0, 0, light_size), (0, -1, 0))]
    v0, v1, v2 = _pack(parts)
    bsdfs = [dict(type=MIRROR), dict(type=LAMBERT, rgb=albedo), dict(type=LIGHT, rgb=(5, 5, 5))]
    return dict(v0=v0, v1=v1, v2=v2, tri_bsdf=np.array([0, 0, 1, 1, 2, 2], np.uint32),
                bsdfs=bsdfs, light=(-light_size / 2, light_size / 2, 0.5),
                camera=np.array([0, 0.5, -3, 0, 0.5, 0, 0, 1, 0, 40.0], np.float32))


def mirror_wall_scene(wall="mirror", floor_albedo=(0.5, 0.5, 0.5)):
    """Lambert floor (y = 0, 2 x 2, facing up), a 0.4-wide square light at
    y = 1 facing down between x = 0.45 and 0.85, and a wall at x = 1 (1.5 high,
    facing -x) that is a mirror (`wall="mirror"`), black Lambert ("black") or
    a light ("light", for finding the pixels that see it).

    No path returns to the floor after leaving it upwards (the wall keeps a
    ray's y direction), so every path from the camera has at most one Lambert
    vertex.  `camera` sees the floor and a few pixels of the wall;
    `mirror_camera` sees the light only in the mirror."""
    parts = [quad((-1, 0, -1), (2, 0, 0), (0, 0, 2), (0, 1, 0)),
             quad((0.45, 1, -0.2), (0.4, 0, 0), (0, 0, 0.4), (0, -1, 0)),
             quad((1, 0, -1), (0, 1.5, 0), (0, 0, 2), (-1, 0, 0))]
    v0, v1, v2 = _pack(parts)
    light = dict(type=LIGHT, rgb=(5.0, 4.0, 3.0))
    walls = dict(mirror=dict(type=MIRROR), black=dict(type=LAMBERT, rgb=(0.0, 0.0, 0.0)),
                 light=light)
    bsdfs = [dict(type=LAMBERT, rgb=floor_albedo), light, walls[wall]]
    return dict(v0=v0, v1=v1, v2=v2, tri_bsdf=np.array([0, 0, 1, 1, 2, 2], np.uint32),
                bsdfs=bsdfs,
                camera=np.array([0, 1.3, -1.2, 0, 0, 0.3, 0, 1, 0, 40.0], np.float32),
                mirror_camera=np.array([-0.6, 0.7, -2.2, 0.5, 0.25, 0, 0, 1, 0, 55.0],
                                       np.float32))


def glass_slab_scene(ior=1.5, extent=100.0):
    """Square light (0.5 wide, y = 2, facing down) above a glass slab between
    y = 0.9 and 1.0 (2 * extent wide) over a black Lambert floor at y = 0."""
    e = extent
    light = quad((-0.25, 2.0, -0.25), (0.5, 0, 0), (0, 0, 0.5), (0, -1, 0))
    slab = box((0, 0.95, 0), (2 * e, 0.1, 2 * e))
    floor = quad((-10 * e, 0, -10 * e), (20 * e, 0, 0), (0, 0, 20 * e), (0, 1, 0))
    v0, v1, v2 = _pack([light, slab, floor])
    tb = np.concatenate([np.full(2, 2), np.full(len(slab[0]), 1), np.full(2, 0)]).astype(np.uint32)
    bsdfs = [dict(type=LAMBERT, rgb=(0.0, 0.0, 0.0)), dict(type=GLASS, ior=ior),
             dict(type=LIGHT, rgb=(5, 5, 5))]
    return dict(v0=v0, v1=v1, v2=v2, tri_bsdf=tb, bsdfs=bsdfs,
                camera=np.array([0, 1.5, -4, 0, 0.5, 0, 0, 1, 0, 40.0], np.float32))


def caustic_scene(sphere_level=3):
    """Caustics: a smooth glass sphere (radius 0.5) and a mirror box on a
    Lambert floor with a Lambert back wall, under a small bright light and a
    larger dim one.  The camera sees the floor directly, through the sphere and
    in the mirror."""
    v, f = icosphere(sphere_level)
    sph = displaced_sphere(v, f, 0.0, 0, center=(0, 0.5, 0), radius=0.5)
    objects = [
        (quad((-3, 0, -3), (6, 0, 0), (0, 0, 6), (0, 1, 0)), 0),
        (quad((-3, 0, 2), (6, 0, 0), (0, 3, 0), (0, 0, -1)), 0),
        (sph, 1),
        (box((1.1, 0.5, 0.6), (0.1, 1.0, 1.2), rot_y(-30)), 2),
        (quad((-0.15, 2.2, -0.15), (0.3, 0, 0), (0, 0, 0.3), (0, -1, 0)), 3),
        (quad((-2.5, 2.8, -1.5), (1.0, 0, 0), (0, 0, 1.0), (0, -1, 0)), 4),
    ]
    bsdfs = [dict(type=LAMBERT, rgb=(0.7, 0.65, 0.6)), dict(type=GLASS, ior=1.5),
             dict(type=MIRROR), dict(type=LIGHT, rgb=(60, 55, 50)),
             dict(type=LIGHT, rgb=(1.5, 1.5, 1.5))]
    parts = [o[0] for o in objects]
    v0, v1, v2 = _pack(parts)
    tb = np.concatenate([np.full(len(p[0]), m) for p, m in objects]).astype(np.uint32)
    # smooth normals on the sphere: the radial direction
    ng = np.cross(v2 - v0, v1 - v0).astype(np.float64)
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    n0, n1, n2 = ng.copy(), ng.copy(), ng.copy()
    a = len(objects[0][0][0]) + len(objects[1][0][0])
    b = a + len(sph[0])
    for vv, nn in ((v0, n0), (v1, n1), (v2, n2)):
        r = vv[a:b].astype(np.float64) - np.array([0, 0.5, 0])
        nn[a:b] = r / np.linalg.norm(r, axis=1, keepdims=True)
    return dict(v0=v0, v1=v1, v2=v2, tri_bsdf=tb, bsdfs=bsdfs,
                normals=tuple(x.astype(np.float32) for x in (n0, n1, n2)),
                camera=np.array([0.0, 1.6, -2.6, 0.0, 0.3, 0.2, 0, 1, 0, 45.0], np.float32))
