"""The allocation-free closest hit (xyz_closest_hit_noalloc) against
xyz_closest_hit, through the C ABI: two calls in a row, the second on fewer
rays, results bit for bit (ids, t, f, g, back side and the per-ray counters);
rays beyond the maximum are rejected with an error."""
import ctypes

import numpy as np
import pytest

from xyz_amd import scenes as S

pytestmark = pytest.mark.gpu
F = np.float32


def _bufs(n):
    from xyz_amd.hip import DeviceArray
    return dict(prim=DeviceArray(n, np.int32), t=DeviceArray(n, np.float32),
                f=DeviceArray(n, np.float32), g=DeviceArray(n, np.float32),
                back=DeviceArray(n, np.uint8), cnt=DeviceArray(2 * n, np.uint32))


def _call(L, fn, scene, o, d, b):
    vp = ctypes.c_void_p
    rc = getattr(L, fn)(scene.h, ctypes.c_uint32(o.shape[0]), vp(o.data_ptr()), vp(d.data_ptr()),
                        vp(b["prim"].data_ptr()), vp(b["t"].data_ptr()), vp(b["f"].data_ptr()),
                        vp(b["g"].data_ptr()), vp(b["back"].data_ptr()), vp(b["cnt"].data_ptr()),
                        None)
    return rc


def test_noalloc_equals_closest_hit(capi_mod, device):
    from xyz_amd.hip import DeviceArray, synchronize
    L = capi_mod.lib()
    m = S.closest_hit_mesh(5, 16, seed=3)
    scene = capi_mod.Scene(m["v0"], m["v1"], m["v2"])
    (oc, dc), (oi, di) = S.closest_hit_rays(m, 256, 128, n_incoherent=40000, seed=5)
    o_all = np.concatenate([oc, oi]).astype(F)
    d_all = np.concatenate([dc, di]).astype(F)
    for n in (len(o_all), 5003):  # the second call smaller
        print(f"[no-alloc] {n} rays", flush=True)
        o = DeviceArray.from_numpy(o_all[:n])
        d = DeviceArray.from_numpy(d_all[:n])
        a, b = _bufs(n), _bufs(n)
        assert _call(L, "xyz_closest_hit", scene, o, d, a) == 0
        assert _call(L, "xyz_closest_hit_noalloc", scene, o, d, b) == 0, L.xyz_last_error()
        synchronize()
        for k in ("prim", "t", "f", "g", "back", "cnt"):
            x, y = a[k].numpy(), b[k].numpy()
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (n, k)
        assert (a["prim"].numpy() >= 0).mean() > 0.05  # rays that hit something


def test_noalloc_rejects_null_buffers(capi_mod, device):
    L = capi_mod.lib()
    assert L.xyz_closest_hit_noalloc(None, 0, *([None] * 9)) == capi_mod.XYZ_E_INVALID
