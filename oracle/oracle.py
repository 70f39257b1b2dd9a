"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by the test suite, the smoke check and the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def build():
    """Compile liboracle.so from oracle/xyz_oracle.cpp (make)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_scene_create.restype = _vp
        L.orc_scene_create.argtypes = [ctypes.c_uint32, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                       _f32p, _u32p, ctypes.c_uint32, _u32p, _f32p, _f32p, _f32p,
                                       _i32p, ctypes.c_uint32, _u32p, _f32p, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.c_int]
        L.orc_scene_destroy.argtypes = [_vp]
        L.orc_light_area.restype = ctypes.c_float
        L.orc_light_area.argtypes = [_vp]
        L.orc_tree_counts.argtypes = [_vp, _u64p, _f32p]
        L.orc_tree_dump.argtypes = [_vp, _u32p, _f32p, _u32p, _u32p, _u32p]
        L.orc_closest_hit.restype = ctypes.c_double
        L.orc_closest_hit.argtypes = [_vp, ctypes.c_uint64, _f32p, _f32p, _i32p, _f32p, _f32p,
                                      _f32p, _u8p, _u32p, _u32p, ctypes.c_int]
        L.orc_brute_force.argtypes = [_vp, ctypes.c_uint64, _f32p, _f32p, _i32p, _f32p, ctypes.c_int]
        L.orc_intersect_triangle.restype = ctypes.c_int
        L.orc_intersect_triangle.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_float,
                                             ctypes.c_float, _f32p]
        L.orc_philox.argtypes = [_u32p, _u32p, _u32p]
        L.orc_cos_sin.argtypes = [ctypes.c_uint64, _f32p, _f32p, _f32p]
        L.orc_pow.argtypes = [ctypes.c_uint64, _f32p, _f32p, _f32p]
        L.orc_glossy_half.argtypes = [ctypes.c_uint64, _f32p, _f32p, ctypes.c_float, _f32p]
        L.orc_shading_basis.argtypes = [_vp, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_int, _f32p]
        L.orc_sample_light.restype = ctypes.c_int
        L.orc_sample_light.argtypes = [_vp, ctypes.c_float, ctypes.c_float, _f32p]
        L.orc_primary_ray.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float, _f32p, _f32p]
        L.orc_primary_ray.restype = None
        L.orc_primary_dir.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                      ctypes.c_float, _f32p]
        L.orc_render_bdpt.restype = ctypes.c_double
        L.orc_render_bdpt.argtypes = [_vp, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint64, _f32p, _f32p, _u64p, ctypes.c_int]
        L.orc_render.restype = ctypes.c_double
        L.orc_render.argtypes = [_vp, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                 ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _f32p, _u64p,
                                 ctypes.c_int]
        L.orc_render_backward.restype = ctypes.c_double
        L.orc_render_backward.argtypes = [_vp, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64, _f32p,
                                          _f64p, _f64p, ctypes.c_int]
        L.orc_importon_map_create.restype = _vp
        L.orc_importon_map_create.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_uint64]
        L.orc_importon_count.restype = ctypes.c_uint64
        L.orc_importon_count.argtypes = [_vp]
        L.orc_importon_export.argtypes = [_vp, _f32p]
        L.orc_importon_map_destroy.argtypes = [_vp]
        L.orc_guided_sample.argtypes = [_f32p, _f32p, ctypes.c_uint64, _f32p, _f32p]
        L.orc_importance_density.argtypes = [_f32p, _f32p, ctypes.c_uint64, _f32p, _f32p]
        L.orc_azimuth_octant.argtypes = [ctypes.c_uint64, _f32p, _f32p, _i32p]
        L.orc_importon_map_from.restype = _vp
        L.orc_importon_map_from.argtypes = [_f32p, ctypes.c_uint64]
        L.orc_importance_query.argtypes = [_vp, _f32p, ctypes.c_uint32, _f32p, _f32p, _f32p]
        L.orc_render_idpt.restype = ctypes.c_double
        L.orc_render_idpt.argtypes = [_vp, _vp, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64,
                                      ctypes.c_uint32, ctypes.c_uint32, _f32p, _u64p, ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _c(a, dtype):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=dtype)


class OracleScene:
    """CPU restatement of xyz::Scene (kd-tree or brute force) for one triangle set.

    Triangles are given as float32 arrays (n, 3) per vertex, in caller order;
    that order defines primitive ids.
    """

    def __init__(self, v0, v1, v2, tri_bsdf=None, bsdfs=None, normals=None, uv=None,
                 textures=None, max_bucket=16, max_depth=16, use_tree=True):
        L = lib()
        self._keep = []
        v0, v1, v2 = (_c(a, np.float32).reshape(-1, 3) for a in (v0, v1, v2))
        n = v0.shape[0]
        self.n_tris = n
        if tri_bsdf is None:
            tri_bsdf = np.zeros(n, np.uint32)
        tri_bsdf = _c(tri_bsdf, np.uint32)
        if bsdfs is None:
            bsdfs = [dict(type=0, rgb=(0.5, 0.5, 0.5))]
        nb = len(bsdfs)
        btype = np.array([b["type"] for b in bsdfs], np.uint32)
        brgb = np.array([b.get("rgb", (0, 0, 0)) for b in bsdfs], np.float32).reshape(-1)
        bshin = np.array([b.get("shininess", 0.0) for b in bsdfs], np.float32)
        bior = np.array([b.get("ior", 1.5) for b in bsdfs], np.float32)
        btex = np.array([b.get("texture", -1) for b in bsdfs], np.int32)
        n0 = n1 = n2 = None
        if normals is not None:
            n0, n1, n2 = (_c(a, np.float32).reshape(-1, 3) for a in normals)
        uvp = None if uv is None else _c(uv, np.float32).reshape(-1, 6)
        ntex = 0
        tex_wh = None
        texels = None
        if textures:
            ntex = len(textures)
            tex_wh = np.array([[t.shape[1], t.shape[0]] for t in textures], np.uint32).reshape(-1)
            texels = np.concatenate([_c(t, np.float32).reshape(-1) for t in textures])
        self._keep += [v0, v1, v2, tri_bsdf, btype, brgb, bshin, bior, btex, n0, n1, n2, uvp,
                       tex_wh, texels]
        self.h = L.orc_scene_create(n, _p(v0, _f32p), _p(v1, _f32p), _p(v2, _f32p), _p(n0, _f32p),
                                    _p(n1, _f32p), _p(n2, _f32p), _p(uvp, _f32p),
                                    _p(tri_bsdf, _u32p), nb, _p(btype, _u32p), _p(brgb, _f32p),
                                    _p(bshin, _f32p), _p(bior, _f32p), _p(btex, _i32p), ntex,
                                    _p(tex_wh, _u32p), _p(texels, _f32p), max_bucket, max_depth,
                                    1 if use_tree else 0)
        self.n_bsdfs = nb
        self.n_texels = 0 if texels is None else texels.size

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    @property
    def light_area(self):
        return lib().orc_light_area(self.h)

    def tree_counts(self):
        c = np.zeros(3, np.uint64)
        b = np.zeros(6, np.float32)
        lib().orc_tree_counts(self.h, _p(c, _u64p), _p(b, _f32p))
        return dict(nodes=int(c[0]), leaves=int(c[1]), refs=int(c[2]), bmin=b[:3], bmax=b[3:])

    def tree_dump(self):
        """Pre-order dump: kind (0..2 axis, 3 leaf, 4 null), split, leaf ranges, refs."""
        c = self.tree_counts()
        nn, nr = c["nodes"], c["refs"]
        kind = np.zeros(nn, np.uint32)
        split = np.zeros(nn, np.float32)
        loff = np.zeros(nn, np.uint32)
        lcnt = np.zeros(nn, np.uint32)
        refs = np.zeros(max(nr, 1), np.uint32)
        lib().orc_tree_dump(self.h, _p(kind, _u32p), _p(split, _f32p), _p(loff, _u32p),
                            _p(lcnt, _u32p), _p(refs, _u32p))
        return dict(kind=kind, split=split, leaf_off=loff, leaf_cnt=lcnt, refs=refs[:nr])

    def closest_hit(self, org, dir, threads=0, counters=False):
        org = _c(org, np.float32).reshape(-1, 3)
        dir = _c(dir, np.float32).reshape(-1, 3)
        n = org.shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        f = np.empty(n, np.float32)
        g = np.empty(n, np.float32)
        back = np.empty(n, np.uint8)
        cn = np.empty(n, np.uint32) if counters else None
        ct = np.empty(n, np.uint32) if counters else None
        secs = lib().orc_closest_hit(self.h, n, _p(org, _f32p), _p(dir, _f32p), _p(prim, _i32p),
                                     _p(t, _f32p), _p(f, _f32p), _p(g, _f32p), _p(back, _u8p),
                                     _p(cn, _u32p), _p(ct, _u32p), threads)
        out = dict(prim=prim, t=t, f=f, g=g, back=back, seconds=secs)
        if counters:
            out["nodes"] = cn
            out["tris"] = ct
        return out

    def brute_force(self, org, dir, threads=0):
        org = _c(org, np.float32).reshape(-1, 3)
        dir = _c(dir, np.float32).reshape(-1, 3)
        n = org.shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        lib().orc_brute_force(self.h, n, _p(org, _f32p), _p(dir, _f32p), _p(prim, _i32p),
                              _p(t, _f32p), threads)
        return dict(prim=prim, t=t)

    def shading_basis(self, tri, f, g, back):
        out = np.zeros(12, np.float32)
        lib().orc_shading_basis(self.h, tri, f, g, int(back), _p(out, _f32p))
        return out.reshape(4, 3)

    def sample_light(self, u, v):
        out = np.zeros(9, np.float32)
        tri = lib().orc_sample_light(self.h, u, v, _p(out, _f32p))
        return tri, out.reshape(3, 3)

    def render(self, camera, W, H, spp, max_depth, rr_start=0, rr_threshold=0.2, seed=0,
               sample0=0, rows=None, threads=0, film=None, stats=False):
        """Sum of `spp` sample colours per pixel, (H, W, 3) float32, row 0 = bottom."""
        cam = camera12(camera)
        if film is None:
            film = np.zeros((H, W, 3), np.float32)
        y0, y1 = (0, H) if rows is None else rows
        st = np.zeros(4, np.uint64)
        secs = lib().orc_render(self.h, _p(cam, _f32p), W, H, spp, sample0, max_depth, rr_start,
                                rr_threshold, seed, y0, y1, _p(film, _f32p), _p(st, _u64p),
                                threads)
        if stats:
            return film, dict(seconds=secs, queries=int(st[0]), nodes=int(st[1]), tris=int(st[2]),
                              vertices=int(st[3]))
        return film

    def render_bdpt(self, camera, W, H, spp, max_depth, seed=0, sample0=0, threads=0,
                    stats=False):
        """Bidirectional path tracing: (own, splat) films of per-pixel sums, (H, W, 3)
        float32, row 0 = bottom; the image is (own + splat) / spp."""
        cam = camera12(camera)
        own = np.zeros((H, W, 3), np.float32)
        splat = np.zeros((H, W, 3), np.float32)
        st = np.zeros(3, np.uint64)
        secs = lib().orc_render_bdpt(self.h, _p(cam, _f32p), W, H, spp, sample0, max_depth, seed,
                                     _p(own, _f32p), _p(splat, _f32p), _p(st, _u64p), threads)
        if secs < 0:
            raise ValueError("max_depth above 16")
        if stats:
            return own, splat, dict(seconds=secs, queries=int(st[0]), nodes=int(st[1]),
                                    tris=int(st[2]))
        return own, splat

    def importon_map(self, n_emit=100000, max_depth=16, seed=0):
        """Importons of `n_emit` light paths (Scene::CreateLightImportanceMap)."""
        return OracleImportonMap(self, n_emit, max_depth, seed)

    def render_idpt(self, imap, camera, W, H, spp, max_depth, rr_start=0, rr_threshold=0.2,
                    seed=0, sample0=0, rows=None, threads=0, film=None, stats=False):
        """Importance-driven path tracing guided by `imap`: per-pixel sums, (H, W, 3)."""
        cam = camera12(camera)
        if film is None:
            film = np.zeros((H, W, 3), np.float32)
        y0, y1 = (0, H) if rows is None else rows
        st = np.zeros(3, np.uint64)
        secs = lib().orc_render_idpt(self.h, imap.h, _p(cam, _f32p), W, H, spp, sample0,
                                     max_depth, rr_start, rr_threshold, seed, y0, y1,
                                     _p(film, _f32p), _p(st, _u64p), threads)
        if stats:
            return film, dict(seconds=secs, queries=int(st[0]), nodes=int(st[1]), tris=int(st[2]))
        return film

    def render_backward(self, camera, W, H, spp, max_depth, d_film, rr_start=0,
                        rr_threshold=0.2, seed=0, sample0=0, threads=0):
        cam = camera12(camera)
        d_film = _c(d_film, np.float32).reshape(H, W, 3)
        d_tex = np.zeros(max(self.n_texels, 1), np.float64)
        d_emit = np.zeros(self.n_bsdfs * 3, np.float64)
        lib().orc_render_backward(self.h, _p(cam, _f32p), W, H, spp, sample0, max_depth, rr_start,
                                  rr_threshold, seed, _p(d_film, _f32p), _p(d_tex, _f64p),
                                  _p(d_emit, _f64p), threads)
        return d_tex[:self.n_texels], d_emit.reshape(-1, 3)


class OracleImportonMap:
    """Importons (position, incoming direction) in path order, then vertex order."""

    def __init__(self, scene, n_emit, max_depth, seed):
        self.scene = scene  # keeps the scene alive
        self.h = lib().orc_importon_map_create(scene.h, n_emit, max_depth, seed)

    @classmethod
    def from_importons(cls, data):
        """Map of caller importons, an (n, 6) array of position, incoming direction."""
        m = cls.__new__(cls)
        m.scene = None
        data = _c(data, np.float32).reshape(-1, 6)
        m.h = lib().orc_importon_map_from(_p(data, _f32p), len(data))
        return m

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_importon_map_destroy(self.h)
            self.h = None

    def __len__(self):
        return int(lib().orc_importon_count(self.h))

    def export(self):
        out = np.zeros((len(self), 6), np.float32)
        if len(out):
            lib().orc_importon_export(self.h, _p(out, _f32p))
        return out

    def query(self, pos, ns, t, b, dirs=None):
        """ImportanceQuery at a vertex: (cdf[17], density of each row of dirs)."""
        basis = np.concatenate([np.asarray(x, np.float32).reshape(3) for x in (pos, ns, t, b)])
        dirs = np.zeros((0, 3), np.float32) if dirs is None else _c(dirs, np.float32).reshape(-1, 3)
        cdf = np.zeros(17, np.float32)
        dens = np.zeros(len(dirs), np.float32)
        lib().orc_importance_query(self.h, _p(basis, _f32p), len(dirs), _p(dirs, _f32p),
                                   _p(cdf, _f32p), _p(dens, _f32p))
        return cdf, dens


def _basis12(ns, t, b, pos=(0, 0, 0)):
    return np.concatenate([np.asarray(x, np.float32).reshape(3) for x in (pos, ns, t, b)])


def guided_sample(cdf, ns, t, b, u):
    """Guided Lambert direction for each row of u (bin, cos^2 theta, azimuth
    numbers): (dirs (n, 3), bin density (n,), density per solid angle (n,))."""
    cdf = _c(cdf, np.float32).reshape(17)
    basis = _basis12(ns, t, b)
    u = _c(u, np.float32).reshape(-1, 3)
    out = np.zeros((len(u), 5), np.float32)
    lib().orc_guided_sample(_p(cdf, _f32p), _p(basis, _f32p), len(u), _p(u, _f32p),
                            _p(out, _f32p))
    return out[:, :3], out[:, 3], out[:, 4]


def importance_density(cdf, ns, t, b, dirs):
    """ImportanceQuery::GetDensity of each incoming direction for a given cdf."""
    cdf = _c(cdf, np.float32).reshape(17)
    basis = _basis12(ns, t, b)
    dirs = _c(dirs, np.float32).reshape(-1, 3)
    out = np.zeros(len(dirs), np.float32)
    lib().orc_importance_density(_p(cdf, _f32p), _p(basis, _f32p), len(dirs), _p(dirs, _f32p),
                                 _p(out, _f32p))
    return out


def azimuth_octant(x, y):
    """Azimuth octant of the importance histogram for arrays x, y."""
    x = _c(x, np.float32).ravel()
    y = _c(y, np.float32).ravel()
    out = np.zeros(len(x), np.int32)
    lib().orc_azimuth_octant(len(x), _p(x, _f32p), _p(y, _f32p), _p(out, _i32p))
    return out


def intersect_triangle(v0, v1, v2, o, d, t_min=1e-5, t_max=np.inf):
    a = [np.ascontiguousarray(x, np.float32) for x in (v0, v1, v2, o, d)]
    out = np.zeros(4, np.float32)
    ok = lib().orc_intersect_triangle(*[_p(x, _f32p) for x in a], np.float32(t_min),
                                      np.float32(t_max), _p(out, _f32p))
    return bool(ok), out


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(_p(c, _u32p), _p(k, _u32p), _p(out, _u32p))
    return out


def cos_sin(x):
    x = np.ascontiguousarray(x, np.float32)
    c = np.empty_like(x)
    s = np.empty_like(x)
    lib().orc_cos_sin(x.size, _p(x, _f32p), _p(c, _f32p), _p(s, _f32p))
    return c, s


def pow32(x, y):
    """The oracle's deterministic float pow (Metal lobe) of arrays x, y."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float32), x.shape), np.float32)
    out = np.empty_like(x)
    lib().orc_pow(x.size, _p(x, _f32p), _p(y, _f32p), _p(out, _f32p))
    return out


def glossy_half(mu_t, mu_p, shininess):
    """The Metal sampler's half vectors (n, 3) for uniforms mu_t, mu_p in the
    basis t = +x, n = +y, b = +z."""
    mu_t = np.ascontiguousarray(mu_t, np.float32)
    mu_p = np.ascontiguousarray(mu_p, np.float32)
    out = np.empty((mu_t.size, 3), np.float32)
    lib().orc_glossy_half(mu_t.size, _p(mu_t, _f32p), _p(mu_p, _f32p), np.float32(shininess),
                          _p(out, _f32p))
    return out


def camera12(camera):
    """eye[3], at[3], up[3], fov_y_deg[, focal_length, f_number] -> 12 float32
    (focal_length 0 = pinhole)."""
    c = np.zeros(12, np.float32)
    a = np.asarray(camera, np.float32).reshape(-1)
    if a.size not in (10, 12):
        raise ValueError("camera: 10 or 12 floats")
    c[:a.size] = a
    return c


def primary_ray(camera, W, H, s, t, ls=0.0, lt=0.0):
    """(origin, direction) of the primary ray through film point (s, t) and,
    for a thin lens, lens numbers (ls, lt)."""
    cam = camera12(camera)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().orc_primary_ray(_p(cam, _f32p), W, H, ctypes.c_float(s), ctypes.c_float(t),
                          ctypes.c_float(ls), ctypes.c_float(lt), _p(o, _f32p), _p(d, _f32p))
    return o, d


def primary_dir(camera, W, H, s, t):
    cam = camera12(camera)
    out = np.zeros(3, np.float32)
    lib().orc_primary_dir(_p(cam, _f32p), W, H, s, t, _p(out, _f32p))
    return out


def max_threads():
    return lib().orc_max_threads()
