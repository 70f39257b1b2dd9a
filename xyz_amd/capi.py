"""ctypes binding of the C ABI declared in include/xyz_hip.h.

This is the host-side mirror of the reference's interfaces for the hot path:
`KDTree` (xyz/geom/kdtree.hpp:15-51: Build + FindIntersection), `Scene`
(xyz/core/scene.hpp: FindIntersection, EvalutePathTracing through `render`),
materials as `Bsdf` records (xyz/bsdf/bsdf.hpp:21-28) and the progressive
`PathTracer` loop (xyz/integrator/integrator.cpp:88-162).

Device buffers are xyz_amd.hip.DeviceArray objects (plain HIP allocations);
every computation runs in libxyz_hip.so.
The library must be built (xyz_amd.build.build()); nothing falls back to a
CPU implementation.
"""
import ctypes
import os

import numpy as np

from . import build as _build

XYZ_OK, XYZ_E_INVALID, XYZ_E_NOMEM, XYZ_E_DEVICE, XYZ_E_UNSUPPORTED = 0, -1, -2, -3, -4
LAMBERT, MIRROR, GLASS, METAL, LIGHT = 0, 1, 2, 3, 4

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


class BsdfDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("rgb", ctypes.c_float * 3),
                ("shininess", ctypes.c_float), ("ior", ctypes.c_float),
                ("texture", ctypes.c_int32)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("n_tris", ctypes.c_uint32), ("v0", _f32p), ("v1", _f32p), ("v2", _f32p),
                ("n0", _f32p), ("n1", _f32p), ("n2", _f32p), ("uv", _f32p),
                ("bsdf", _u32p), ("n_bsdfs", ctypes.c_uint32),
                ("bsdfs", ctypes.POINTER(BsdfDesc)), ("n_textures", ctypes.c_uint32),
                ("tex_wh", _u32p), ("texels", _f32p), ("kd_max_bucket", ctypes.c_uint32),
                ("kd_max_depth", ctypes.c_uint32), ("kd_builder", ctypes.c_uint32)]


KD_HOST, KD_DEVICE = 0, 1


class SceneInfo(ctypes.Structure):
    _fields_ = [("n_tris", ctypes.c_uint64), ("n_nodes", ctypes.c_uint64),
                ("n_leaves", ctypes.c_uint64), ("n_leaf_refs", ctypes.c_uint64),
                ("depth", ctypes.c_uint32), ("max_leaf", ctypes.c_uint32),
                ("bmin", ctypes.c_float * 3), ("bmax", ctypes.c_float * 3),
                ("build_seconds", ctypes.c_double), ("device_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return dict(n_tris=self.n_tris, n_nodes=self.n_nodes, n_leaves=self.n_leaves,
                    n_leaf_refs=self.n_leaf_refs, depth=self.depth, max_leaf=self.max_leaf,
                    bmin=list(self.bmin), bmax=list(self.bmax),
                    build_seconds=self.build_seconds, device_bytes=self.device_bytes)


class Camera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("at", ctypes.c_float * 3),
                ("up", ctypes.c_float * 3), ("fov_y_deg", ctypes.c_float),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("focal_length", ctypes.c_float), ("f_number", ctypes.c_float)]


class RenderParams(ctypes.Structure):
    _fields_ = [("spp", ctypes.c_uint32), ("sample_offset", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("rr_start", ctypes.c_uint32),
                ("rr_threshold", ctypes.c_float), ("engine", ctypes.c_uint32),
                ("seed", ctypes.c_uint64)]


class RenderStats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_uint64), ("queries", ctypes.c_uint64),
                ("seconds", ctypes.c_double)]


class MqoInfo(ctypes.Structure):
    _fields_ = [("n_tris", ctypes.c_uint32), ("n_bsdfs", ctypes.c_uint32),
                ("n_objects", ctypes.c_uint32), ("n_smooth_objects", ctypes.c_uint32),
                ("n_lights", ctypes.c_uint32), ("has_camera", ctypes.c_uint32),
                ("eye", ctypes.c_float * 3), ("at", ctypes.c_float * 3),
                ("light_power", ctypes.c_float * 3), ("surface_area", ctypes.c_float),
                ("light_area", ctypes.c_float)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        for k in ("eye", "at", "light_power"):
            d[k] = list(d[k])
        return d


class ImportonInfo(ctypes.Structure):
    _fields_ = [("n_importons", ctypes.c_uint64), ("n_emit", ctypes.c_uint32),
                ("grid", ctypes.c_uint32 * 3), ("cell", ctypes.c_float),
                ("radius", ctypes.c_float), ("seconds", ctypes.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["grid"] = list(d["grid"])
        return d


class BatchTuning(ctypes.Structure):
    _fields_ = [("tile_pixels", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("block_threads", ctypes.c_uint32), ("use_graph", ctypes.c_uint32),
                ("engine", ctypes.c_uint32)]


# Every symbol declared in include/xyz_hip.h.
EXPORTS = ["xyz_kdtree_build", "xyz_kdtree_build_device", "xyz_kdtree_info", "xyz_kdtree_export", "xyz_kdtree_destroy",
           "xyz_mqo_load", "xyz_mqo_load_file", "xyz_mqo_describe", "xyz_mqo_destroy",
           "xyz_scene_create", "xyz_scene_get_info", "xyz_scene_destroy", "xyz_closest_hit",
           "xyz_closest_hit_work", "xyz_closest_hit_noalloc", "xyz_check_reciprocal", "xyz_eval_pow",
           "xyz_importon_map_create", "xyz_importon_map_from_host", "xyz_importon_map_info",
           "xyz_importon_map_export", "xyz_importance_query", "xyz_importon_map_destroy",
           "xyz_render", "xyz_render_idpt", "xyz_render_bdpt", "xyz_render_backward", "xyz_render_batch",
           "xyz_batch_autotune",
           "xyz_last_error", "xyz_version"]

_lib = None


class XyzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"xyz error {code}: {msg}")
        self.code = code


def lib():
    """Load libxyz_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = os.environ.get("XYZ_HIP_LIB", _build.LIB)  # developer override of the build
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run xyz_amd.build.build() first")
        L = ctypes.CDLL(path)
        L.xyz_kdtree_build.argtypes = [ctypes.c_uint32, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.xyz_kdtree_build_device.argtypes = L.xyz_kdtree_build.argtypes
        L.xyz_kdtree_info.argtypes = [_vp, ctypes.POINTER(SceneInfo)]
        L.xyz_kdtree_export.argtypes = [_vp, _u32p, _u32p]
        L.xyz_kdtree_destroy.argtypes = [_vp]
        L.xyz_kdtree_destroy.restype = None
        L.xyz_mqo_load.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(_vp)]
        L.xyz_mqo_load_file.argtypes = [ctypes.c_char_p, ctypes.POINTER(_vp)]
        L.xyz_mqo_describe.argtypes = [_vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(MqoInfo)]
        L.xyz_mqo_destroy.argtypes = [_vp]
        L.xyz_mqo_destroy.restype = None
        L.xyz_scene_create.argtypes = [ctypes.POINTER(SceneDesc), ctypes.POINTER(_vp)]
        L.xyz_scene_get_info.argtypes = [_vp, ctypes.POINTER(SceneInfo)]
        L.xyz_scene_destroy.argtypes = [_vp]
        L.xyz_scene_destroy.restype = None
        L.xyz_closest_hit.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp, _vp]
        L.xyz_closest_hit_work.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp]
        L.xyz_closest_hit_noalloc.argtypes = L.xyz_closest_hit.argtypes
        L.xyz_check_reciprocal.argtypes = [ctypes.POINTER(ctypes.c_ulonglong), _vp]
        L.xyz_eval_pow.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, _vp]
        L.xyz_render.argtypes = [_vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), _vp,
                                 ctypes.POINTER(RenderStats), _vp]
        L.xyz_importon_map_create.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_uint64, ctypes.POINTER(_vp)]
        L.xyz_importon_map_from_host.argtypes = [_vp, ctypes.c_uint64, _f32p, ctypes.POINTER(_vp)]
        L.xyz_importon_map_info.argtypes = [_vp, ctypes.POINTER(ImportonInfo)]
        L.xyz_importon_map_export.argtypes = [_vp, _f32p]
        L.xyz_importance_query.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, _vp]
        L.xyz_importon_map_destroy.argtypes = [_vp]
        L.xyz_importon_map_destroy.restype = None
        L.xyz_render_idpt.argtypes = [_vp, _vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams),
                                      _vp, ctypes.POINTER(RenderStats), _vp]
        L.xyz_render_bdpt.argtypes = [_vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams),
                                      _vp, _vp, ctypes.POINTER(RenderStats), _vp]
        L.xyz_render_backward.argtypes = [_vp, ctypes.POINTER(Camera),
                                          ctypes.POINTER(RenderParams), _vp, _vp, _vp, _vp]
        L.xyz_render_batch.argtypes = [ctypes.c_uint32, ctypes.POINTER(_vp),
                                       ctypes.POINTER(Camera), ctypes.POINTER(RenderParams),
                                       ctypes.POINTER(_vp), ctypes.POINTER(BatchTuning),
                                       ctypes.POINTER(RenderStats), _vp]
        L.xyz_batch_autotune.argtypes = [ctypes.c_uint32, ctypes.POINTER(_vp),
                                         ctypes.POINTER(Camera), ctypes.POINTER(RenderParams),
                                         ctypes.POINTER(_vp), ctypes.POINTER(BatchTuning), _vp]
        L.xyz_last_error.restype = ctypes.c_char_p
        L.xyz_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != XYZ_OK:
        raise XyzError(rc, lib().xyz_last_error().decode())
    return rc


def _np(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _dptr(t):
    """Device pointer of a DeviceArray (xyz_amd.hip) or None."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    """hipStream_t as void*: None selects the null stream."""
    if stream is None:
        return ctypes.c_void_p(None)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.value if isinstance(stream, ctypes.c_void_p) else stream)


class KDTree:
    """SAH kd-tree (reference KDTree::Build), flattened to the device layout.

    device=False builds on the host (no GPU needed), device=True on the current
    GPU; both give the same tree."""

    def __init__(self, v0, v1, v2, max_bucket=16, max_depth=16, device=False):
        v0, v1, v2 = (_np(a, np.float32).reshape(-1, 3) for a in (v0, v1, v2))
        h = _vp()
        fn = lib().xyz_kdtree_build_device if device else lib().xyz_kdtree_build
        check(fn(v0.shape[0], _ptr(v0, _f32p), _ptr(v1, _f32p), _ptr(v2, _f32p), max_bucket,
                 max_depth, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().xyz_kdtree_destroy(self.h)
            self.h = None

    def info(self):
        i = SceneInfo()
        check(lib().xyz_kdtree_info(self.h, ctypes.byref(i)))
        return i.as_dict()

    def export(self):
        """(nodes uint32 (n_nodes, 2), leaf_prims uint32)."""
        i = self.info()
        nodes = np.zeros((i["n_nodes"], 2), np.uint32)
        prims = np.zeros(max(i["n_leaf_refs"], 1), np.uint32)
        check(lib().xyz_kdtree_export(self.h, _ptr(nodes, _u32p), _ptr(prims, _u32p)))
        return nodes, prims[:i["n_leaf_refs"]]


def bsdf_array(bsdfs):
    arr = (BsdfDesc * max(len(bsdfs), 1))()
    for i, b in enumerate(bsdfs):
        arr[i].type = b["type"]
        rgb = b.get("rgb", (0.0, 0.0, 0.0))
        for k in range(3):
            arr[i].rgb[k] = rgb[k]
        arr[i].shininess = b.get("shininess", 0.0)
        arr[i].ior = b.get("ior", 1.5)
        arr[i].texture = b.get("texture", -1)
    return arr


class Scene:
    """Device-resident scene: kd-tree, leaf-ordered triangle records, materials.

    Mirrors xyz::Scene for the hot path: closest_hit() is FindIntersection over a
    batch, render() is the progressive path-tracing pass.
    """

    def __init__(self, v0, v1, v2, tri_bsdf=None, bsdfs=None, normals=None, uv=None,
                 textures=None, max_bucket=16, max_depth=16, kd_builder=KD_HOST):
        v0, v1, v2 = (_np(a, np.float32).reshape(-1, 3) for a in (v0, v1, v2))
        n = v0.shape[0]
        if bsdfs is None:
            bsdfs = [dict(type=LAMBERT, rgb=(0.5, 0.5, 0.5))]
        keep = [v0, v1, v2]
        d = SceneDesc()
        d.n_tris = n
        d.v0, d.v1, d.v2 = (_ptr(a, _f32p) for a in (v0, v1, v2))
        if normals is not None:
            ns = [_np(a, np.float32).reshape(-1, 3) for a in normals]
            keep += ns
            d.n0, d.n1, d.n2 = (_ptr(a, _f32p) for a in ns)
        if uv is not None:
            uvn = _np(uv, np.float32).reshape(-1, 6)
            keep.append(uvn)
            d.uv = _ptr(uvn, _f32p)
        if tri_bsdf is not None:
            tb = _np(tri_bsdf, np.uint32)
            keep.append(tb)
            d.bsdf = _ptr(tb, _u32p)
        ba = bsdf_array(bsdfs)
        keep.append(ba)
        d.n_bsdfs = len(bsdfs)
        d.bsdfs = ba
        self.n_texels = 0
        if textures:
            wh = np.array([[t.shape[1], t.shape[0]] for t in textures], np.uint32).reshape(-1)
            tx = np.concatenate([_np(t, np.float32).reshape(-1) for t in textures])
            keep += [wh, tx]
            d.n_textures = len(textures)
            d.tex_wh = _ptr(wh, _u32p)
            d.texels = _ptr(tx, _f32p)
            self.n_texels = tx.size
        d.kd_max_bucket = max_bucket
        d.kd_max_depth = max_depth
        d.kd_builder = kd_builder
        h = _vp()
        check(lib().xyz_scene_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.n_tris = n
        self.n_bsdfs = len(bsdfs)

    @classmethod
    def from_dict(cls, sc, **kw):
        return cls(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc.get("tri_bsdf"),
                   bsdfs=sc.get("bsdfs"), normals=sc.get("normals"), uv=sc.get("uv"),
                   textures=sc.get("textures"), **kw)

    def __del__(self):
        if getattr(self, "h", None):
            lib().xyz_scene_destroy(self.h)
            self.h = None

    def info(self):
        i = SceneInfo()
        check(lib().xyz_scene_get_info(self.h, ctypes.byref(i)))
        return i.as_dict()

    def closest_hit(self, org, dir, prim, t=None, f=None, g=None, back=None, counters=None,
                    stream=None):
        """FindIntersection for a batch (device arrays: org/dir float32 (n,3),
        prim int32 (n), optional t/f/g float32, back uint8, counters uint32 (2n):
        nodes visited in [0, n), triangles tested in [n, 2n))."""
        n = org.shape[0]
        check(lib().xyz_closest_hit(self.h, n, _dptr(org), _dptr(dir), _dptr(prim), _dptr(t),
                                    _dptr(f), _dptr(g), _dptr(back), _dptr(counters),
                                    _stream(stream)))

    def closest_hit_work(self, org, dir, prim, work, t=None, stream=None):
        """closest_hit that also adds the launch's totals to `work` (device
        uint64 (4,)): nodes visited, reference triangle tests, triangle tests
        executed, cluster box tests executed."""
        n = org.shape[0]
        check(lib().xyz_closest_hit_work(self.h, n, _dptr(org), _dptr(dir), _dptr(prim), _dptr(t),
                                         _dptr(work), _stream(stream)))

    def render(self, camera, film, spp, max_depth, rr_start=0, rr_threshold=0.2, seed=0,
               sample_offset=0, stream=None):
        """Adds the per-pixel SUM of `spp` samples to `film` (device float32 (H, W, 3),
        row 0 = bottom)."""
        cam = make_camera(camera, film.shape[1], film.shape[0])
        p = make_params(spp, max_depth, rr_start, rr_threshold, seed, sample_offset)
        st = RenderStats()
        check(lib().xyz_render(self.h, ctypes.byref(cam), ctypes.byref(p), _dptr(film),
                               ctypes.byref(st), _stream(stream)))
        return dict(paths=st.paths, queries=st.queries, seconds=st.seconds)

    def importon_map(self, n_emit=100000, max_depth=16, seed=0):
        """Importon map of `n_emit` light paths traced on the device."""
        return ImportonMap(self, n_emit=n_emit, max_depth=max_depth, seed=seed)

    def render_idpt(self, imap, camera, film, spp, max_depth, rr_start=0, rr_threshold=0.2,
                    seed=0, sample_offset=0, stream=None):
        """Importance-driven path tracing guided by `imap` (an ImportonMap of this
        scene); film contract of render()."""
        cam = make_camera(camera, film.shape[1], film.shape[0])
        p = make_params(spp, max_depth, rr_start, rr_threshold, seed, sample_offset)
        st = RenderStats()
        check(lib().xyz_render_idpt(self.h, imap.h, ctypes.byref(cam), ctypes.byref(p),
                                    _dptr(film), ctypes.byref(st), _stream(stream)))
        return dict(paths=st.paths, queries=st.queries, seconds=st.seconds)

    def render_bdpt(self, camera, film, spp, max_depth, seed=0, sample_offset=0, splat=None,
                    stream=None, engine=0):
        """Bidirectional path tracing: adds the per-pixel SUM of `spp` samples' own
        colours to `film` and their light-tracing contributions to `splat` (or to
        `film` when splat is None); device float32 (H, W, 3), row 0 = bottom.
        engine: 0 by the scene's size, 1 per-thread queries, 2 queued queries."""
        cam = make_camera(camera, film.shape[1], film.shape[0])
        p = make_params(spp, max_depth, 0, 0.2, seed, sample_offset)
        p.engine = engine
        st = RenderStats()
        check(lib().xyz_render_bdpt(self.h, ctypes.byref(cam), ctypes.byref(p), _dptr(film),
                                    _dptr(splat), ctypes.byref(st), _stream(stream)))
        return dict(paths=st.paths, queries=st.queries, seconds=st.seconds)

    def render_backward(self, camera, d_film, d_texels, d_emission, spp, max_depth, rr_start=0,
                        rr_threshold=0.2, seed=0, sample_offset=0, stream=None):
        cam = make_camera(camera, d_film.shape[1], d_film.shape[0])
        p = make_params(spp, max_depth, rr_start, rr_threshold, seed, sample_offset)
        check(lib().xyz_render_backward(self.h, ctypes.byref(cam), ctypes.byref(p),
                                        _dptr(d_film), _dptr(d_texels), _dptr(d_emission),
                                        _stream(stream)))


class ImportonMap:
    """Importons (position, incoming direction) of a scene, on the device with
    a grid for the 0.1-radius query (reference Scene::CreateLightImportanceMap)."""

    def __init__(self, scene, n_emit=100000, max_depth=16, seed=0, importons=None):
        self.scene = scene  # the map refers to the scene
        h = _vp()
        if importons is None:
            check(lib().xyz_importon_map_create(scene.h, n_emit, max_depth, seed, ctypes.byref(h)))
        else:
            data = _np(importons, np.float32).reshape(-1, 6)
            check(lib().xyz_importon_map_from_host(scene.h, len(data), _ptr(data, _f32p),
                                                   ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().xyz_importon_map_destroy(self.h)
            self.h = None

    def info(self):
        i = ImportonInfo()
        check(lib().xyz_importon_map_info(self.h, ctypes.byref(i)))
        return i.as_dict()

    def __len__(self):
        return int(self.info()["n_importons"])

    def export(self):
        out = np.zeros((len(self), 6), np.float32)
        check(lib().xyz_importon_map_export(self.h, _ptr(out, _f32p)))
        return out

    def query(self, basis, cdf, stream=None):
        """cdf (device float32 (n, 17)) of the importance histogram at each row of
        basis (device float32 (n, 12): position, shading normal, tangent, binormal)."""
        check(lib().xyz_importance_query(self.h, basis.shape[0], _dptr(basis), _dptr(cdf),
                                         _stream(stream)))


def check_reciprocal(stream=None):
    """Number of binary32 values with 1e-6 <= |x| < 2^126 for which the
    triangle test's fast reciprocal differs from the IEEE quotient on the
    current device (0 expected)."""
    n = ctypes.c_ulonglong(0)
    check(lib().xyz_check_reciprocal(ctypes.byref(n), _stream(stream)))
    return int(n.value)


def eval_pow(x, y, out, stream=None):
    """out = the device's Metal-lobe pow of device float32 arrays x, y."""
    n = int(np.prod(x.shape))
    check(lib().xyz_eval_pow(n, _dptr(x), _dptr(y), _dptr(out), _stream(stream)))


def make_camera(camera, W, H):
    """camera: eye[3], at[3], up[3], fov_y_deg[, focal_length, f_number]
    (focal_length 0 or absent: pinhole; > 0: thin lens, fov unused)."""
    c = Camera()
    cam = np.asarray(camera, np.float32).reshape(-1)
    if cam.size not in (10, 12):
        raise ValueError("camera: 10 or 12 floats")
    for k in range(3):
        c.eye[k] = cam[k]
        c.at[k] = cam[3 + k]
        c.up[k] = cam[6 + k]
    c.fov_y_deg = cam[9]
    c.width = W
    c.height = H
    if cam.size == 12:
        c.focal_length = cam[10]
        c.f_number = cam[11]
    return c


def make_params(spp, max_depth, rr_start=0, rr_threshold=0.2, seed=0, sample_offset=0):
    p = RenderParams()
    p.spp = spp
    p.sample_offset = sample_offset
    p.max_depth = max_depth
    p.rr_start = rr_start
    p.rr_threshold = rr_threshold
    p.seed = seed
    return p


def render_batch(scenes, cameras, films, spp, max_depth, rr_start=0, rr_threshold=0.2, seeds=None,
                 tuning=None, stream=None):
    n = len(scenes)
    hs = (_vp * n)(*[s.h for s in scenes])
    cams = (Camera * n)(*[make_camera(c, f.shape[1], f.shape[0]) for c, f in zip(cameras, films)])
    spps = spp if isinstance(spp, (list, tuple)) else [spp] * n
    seeds = seeds if seeds is not None else list(range(n))
    ps = (RenderParams * n)(*[make_params(spps[i], max_depth, rr_start, rr_threshold, seeds[i])
                              for i in range(n)])
    fs = (_vp * n)(*[f.data_ptr() for f in films])
    st = RenderStats()
    tp = ctypes.byref(tuning) if tuning is not None else None
    check(lib().xyz_render_batch(n, hs, cams, ps, fs, tp, ctypes.byref(st), _stream(stream)))
    return dict(paths=st.paths, queries=st.queries, seconds=st.seconds)


def batch_autotune(scenes, cameras, films, spp, max_depth, rr_start=0, rr_threshold=0.2,
                   seeds=None, stream=None):
    n = len(scenes)
    hs = (_vp * n)(*[s.h for s in scenes])
    cams = (Camera * n)(*[make_camera(c, f.shape[1], f.shape[0]) for c, f in zip(cameras, films)])
    spps = spp if isinstance(spp, (list, tuple)) else [spp] * n
    seeds = seeds if seeds is not None else list(range(n))
    ps = (RenderParams * n)(*[make_params(spps[i], max_depth, rr_start, rr_threshold, seeds[i])
                              for i in range(n)])
    fs = (_vp * n)(*[f.data_ptr() for f in films])
    best = BatchTuning()
    check(lib().xyz_batch_autotune(n, hs, cams, ps, fs, ctypes.byref(best), _stream(stream)))
    return best
