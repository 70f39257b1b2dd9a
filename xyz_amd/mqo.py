"""Metasequoia (.mqo) scenes through the C ABI (xyz_mqo_load / xyz_mqo_describe).

Mirrors the reference's Scene::Load (xyz/core/scene_mqo.cpp:547-777): the
loader itself is native (xyz_amd/csrc/mqo_loader.cpp); this module copies its
arrays out and builds a device Scene and the reference's default thin-lens
camera from them."""
import ctypes

import numpy as np

from . import capi

# ThinLensCamera defaults (xyz/core/thinlenscamera.cpp:21-28): f = 35 mm, F1.4
FOCAL_LENGTH = 0.035
F_NUMBER = 1.4


class MqoScene:
    """A parsed Metasequoia document: triangles in file order, materials as
    xyz_bsdf_desc dicts, the camera of its Scene chunk."""

    def __init__(self, text=None, path=None):
        if (text is None) == (path is None):
            raise ValueError("give text or path")
        L = capi.lib()
        h = capi._vp()
        if path is not None:
            capi.check(L.xyz_mqo_load_file(str(path).encode(), ctypes.byref(h)))
        else:
            raw = text.encode() if isinstance(text, str) else bytes(text)
            capi.check(L.xyz_mqo_load(raw, len(raw), ctypes.byref(h)))
        try:
            d = capi.SceneDesc()
            info = capi.MqoInfo()
            capi.check(L.xyz_mqo_describe(h, ctypes.byref(d), ctypes.byref(info)))
            self.info = info.as_dict()
            n = d.n_tris

            def arr(p, count, dtype):
                if count == 0 or not p:
                    return np.zeros(0, dtype)
                return np.ctypeslib.as_array(p, shape=(count,)).astype(dtype, copy=True)

            self.v0, self.v1, self.v2, self.n0, self.n1, self.n2 = (
                arr(getattr(d, k), 3 * n, np.float32).reshape(-1, 3)
                for k in ("v0", "v1", "v2", "n0", "n1", "n2"))
            self.uv = arr(d.uv, 6 * n, np.float32).reshape(-1, 6)
            self.tri_bsdf = arr(d.bsdf, n, np.uint32)
            self.bsdfs = [dict(type=int(b.type), rgb=tuple(float(x) for x in b.rgb),
                               shininess=float(b.shininess), ior=float(b.ior),
                               texture=int(b.texture))
                          for b in (d.bsdfs[i] for i in range(d.n_bsdfs))]
        finally:
            L.xyz_mqo_destroy(h)

    @property
    def n_tris(self):
        return int(self.info["n_tris"])

    def scene(self, **kw):
        """Device scene (kd-tree built with the reference's rules)."""
        return capi.Scene(self.v0, self.v1, self.v2, tri_bsdf=self.tri_bsdf, bsdfs=self.bsdfs,
                          normals=(self.n0, self.n1, self.n2), uv=self.uv, **kw)

    def camera(self, eye=None, at=None, pinhole_fov_deg=None):
        """12 floats for capi.make_camera: the reference's thin lens at the
        document's eye / at (or the given ones); pinhole_fov_deg selects a
        pinhole instead."""
        e = self.info["eye"] if eye is None else list(eye)
        a = self.info["at"] if at is None else list(at)
        if pinhole_fov_deg is not None:
            return [*e, *a, 0.0, 1.0, 0.0, float(pinhole_fov_deg), 0.0, 0.0]
        return [*e, *a, 0.0, 1.0, 0.0, 0.0, FOCAL_LENGTH, F_NUMBER]


def load(path):
    return MqoScene(path=path)


def loads(text):
    return MqoScene(text=text)
