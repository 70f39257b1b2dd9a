"""The device SAH build (xyz_kdtree_build_device) returns the host builder's
tree: the same nodes, split values, leaf sizes and leaf references in the same
order, on meshes that exercise ties, small buckets, the depth limit, straddling
triangles, costs that overflow to infinity (no finite split: a leaf) and the
closest-hit workload mesh.  The host builder itself is pinned to the oracle's
tree by test_kdtree_host.py; the reference's rules are xyz/geom/kdtree.cpp:207-357."""
import numpy as np
import pytest

from xyz_amd import scenes as S

pytestmark = pytest.mark.gpu


def _same_tree(capi_mod, v0, v1, v2, **kw):
    print(f"[kd device build] n={len(v0)} {kw}", flush=True)
    host = capi_mod.KDTree(v0, v1, v2, **kw)
    dev = capi_mod.KDTree(v0, v1, v2, device=True, **kw)
    hn, hp = host.export()
    dn, dp = dev.export()
    assert hn.shape == dn.shape, (hn.shape, dn.shape)
    assert np.array_equal(hn[:, 0], dn[:, 0])
    leaf = (hn[:, 0] & 3) == 3
    assert np.array_equal(hn[leaf, 1], dn[leaf, 1])
    # split planes: equal as binary32 values (a split at 0 may carry either sign)
    assert np.array_equal(hn[~leaf, 1].view(np.float32), dn[~leaf, 1].view(np.float32))
    assert np.array_equal(hp, dp)
    hi, di = host.info(), dev.info()
    for k in ("n_nodes", "n_leaves", "n_leaf_refs", "depth", "max_leaf", "bmin", "bmax"):
        assert hi[k] == di[k], k
    return hi, di


@pytest.mark.parametrize("level,boxes,seed", [(2, 8, 0), (4, 64, 1), (5, 64, 2), (6, 16, 3)])
def test_device_build_matches_host(capi_mod, device, level, boxes, seed):
    m = S.closest_hit_mesh(level, boxes, seed=seed)
    hi, _ = _same_tree(capi_mod, m["v0"], m["v1"], m["v2"])
    assert hi["n_nodes"] > 100


def test_device_build_ties_small_buckets_depth_limit(capi_mod, device):
    rng = np.random.default_rng(3)
    tri = (np.round(rng.uniform(-1, 1, (5000, 3, 3)) * 4) / 4).astype(np.float32)  # many equal events
    for kw in (dict(max_bucket=2, max_depth=10), dict(max_bucket=1, max_depth=20),
               dict(max_bucket=16, max_depth=3), dict(max_bucket=4, max_depth=30)):
        _same_tree(capi_mod, tri[:, 0], tri[:, 1], tri[:, 2], **kw)


def test_device_build_degenerate_inputs(capi_mod, device):
    rng = np.random.default_rng(5)
    # all triangles in the plane z = 0, and signed zeros among the coordinates
    tri = rng.uniform(-1, 1, (3000, 3, 3)).astype(np.float32)
    tri[..., 2] = 0.0
    tri[::7, :, 2] = -0.0
    tri[::5, 0, 0] = -0.0
    _same_tree(capi_mod, tri[:, 0], tri[:, 1], tri[:, 2], max_bucket=4)
    # identical triangles: no split separates them
    one = np.tile(rng.uniform(-1, 1, (1, 3, 3)).astype(np.float32), (200, 1, 1))
    _same_tree(capi_mod, one[:, 0], one[:, 1], one[:, 2])
    # long thin triangles straddling many splits (references multiply)
    a = rng.uniform(-1, 1, (4000, 3)).astype(np.float32)
    b = a + rng.normal(0, 0.8, (4000, 3)).astype(np.float32)
    c = a + rng.normal(0, 0.01, (4000, 3)).astype(np.float32)
    hi, _ = _same_tree(capi_mod, a, b, c, max_bucket=2, max_depth=24)
    assert hi["n_leaf_refs"] > 2 * 4000
    # coordinates whose surface areas overflow: every cost is +inf, the root is a leaf
    big = (rng.uniform(-1, 1, (100, 3, 3)) * 3e19).astype(np.float32)
    hi, _ = _same_tree(capi_mod, big[:, 0], big[:, 1], big[:, 2])
    assert hi["n_nodes"] == 1 and hi["max_leaf"] == 100


def test_device_build_small_and_invalid(capi_mod, device):
    z = np.zeros((0, 3), np.float32)
    t = capi_mod.KDTree(z, z, z, device=True)
    nodes, prims = t.export()
    assert nodes.shape == (1, 2) and nodes[0, 0] == 3 and prims.size == 0
    c = S.cornell_box()
    _same_tree(capi_mod, c["v0"], c["v1"], c["v2"])
    _same_tree(capi_mod, c["v0"][:16], c["v1"][:16], c["v2"][:16])  # one leaf of 16
    _same_tree(capi_mod, c["v0"][:17], c["v1"][:17], c["v2"][:17])
    bad = np.array([[0, 0, 0]], np.float32)
    nan = np.array([[np.nan, 0, 0]], np.float32)
    with pytest.raises(capi_mod.XyzError) as ei:
        capi_mod.KDTree(bad, nan, bad, device=True)
    assert ei.value.code == capi_mod.XYZ_E_INVALID and "non-finite" in str(ei.value)


def test_device_build_workload_mesh_and_scene(capi_mod, device):
    """The 1.3M-triangle closest-hit mesh: the same tree, and a scene created
    with the device builder gives bit-identical closest hits."""
    from xyz_amd import hip
    m = S.closest_hit_mesh(8, 64, seed=1)
    hi, di = _same_tree(capi_mod, m["v0"], m["v1"], m["v2"])
    print(f"host build {hi['build_seconds']:.3f}s, device build {di['build_seconds']:.3f}s",
          flush=True)
    sh = capi_mod.Scene(m["v0"], m["v1"], m["v2"])
    sd = capi_mod.Scene(m["v0"], m["v1"], m["v2"], kd_builder=capi_mod.KD_DEVICE)
    (oc, dc), (oi, di_) = S.closest_hit_rays(m, 256, 128, seed=2)
    o = hip.DeviceArray.from_numpy(np.concatenate([oc, oi]))
    d = hip.DeviceArray.from_numpy(np.concatenate([dc, di_]))
    n = o.shape[0]
    res = []
    for sc in (sh, sd):
        p = hip.DeviceArray(n, np.int32)
        t = hip.DeviceArray(n, np.float32)
        sc.closest_hit(o, d, p, t)
        hip.synchronize()
        res.append((p.numpy(), t.numpy()))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1].view(np.uint32), res[1][1].view(np.uint32))
    assert (res[0][0] >= 0).mean() > 0.2
    with pytest.raises(capi_mod.XyzError):
        capi_mod.Scene(m["v0"][:10], m["v1"][:10], m["v2"][:10], kd_builder=7)


def test_device_build_root_box_signed_zero_ties(capi_mod, device):
    """200,000 triangles (the root box is accumulated in parts on host threads)
    whose y coordinates are all +0 or -0: the box's y bounds are ties between
    the two zeros, which the host builder breaks by triangle order."""
    rng = np.random.default_rng(7)
    n = 200000
    centre = rng.uniform(-1, 1, (n, 1, 3))
    tri = (centre + rng.uniform(-0.005, 0.005, (n, 3, 3))).astype(np.float32)
    tri[..., 1] = np.where(rng.uniform(size=(n, 3)) < 0.5, np.float32(0.0), np.float32(-0.0))
    hi, di = _same_tree(capi_mod, tri[:, 0], tri[:, 1], tri[:, 2])
    for k in ("bmin", "bmax"):
        assert np.array_equal(np.array(hi[k], np.float32).view(np.uint32),
                              np.array(di[k], np.float32).view(np.uint32)), k
