"""Closest-hit parity on the GPU: the HIP kernel against the oracle, bit for bit
(primitive id, t, f, g, back side, and the per-ray node / triangle counts)."""
import numpy as np
import pytest

from xyz_amd import scenes as S

pytestmark = pytest.mark.gpu
F = np.float32


def _run_gpu(capi_mod, scene, o, d, counters=True):
    from xyz_amd.hip import DeviceArray
    n = len(o)
    to = DeviceArray.from_numpy(np.ascontiguousarray(o, F))
    td = DeviceArray.from_numpy(np.ascontiguousarray(d, F))
    prim = DeviceArray(n, np.int32)
    t = DeviceArray(n, np.float32)
    f = DeviceArray(n, np.float32)
    g = DeviceArray(n, np.float32)
    back = DeviceArray(n, np.uint8)
    cnt = DeviceArray(2 * n, np.uint32) if counters else None
    scene.closest_hit(to, td, prim, t, f, g, back, cnt)
    out = dict(prim=prim.numpy(), t=t.numpy(), f=f.numpy(), g=g.numpy(), back=back.numpy())
    if counters:
        c = cnt.numpy()
        out["nodes"], out["tris"] = c[:n], c[n:]
    return out


def _assert_bit_exact(gpu, orc, counters=True):
    """Bit-identical results.  A ray with NaN components can hit with a NaN t,
    f or g; IEEE 754 leaves the sign and payload of a NaN result open (x86 and
    the GPU make different ones), so NaNs are compared as NaN."""
    assert np.array_equal(gpu["prim"], orc["prim"])
    hit = orc["prim"] >= 0
    for k in ("t", "f", "g"):
        a, b = gpu[k][hit], orc[k][hit]
        nan = np.isnan(b)
        assert np.array_equal(np.isnan(a), nan), k
        assert np.array_equal(a[~nan].view(np.uint32), b[~nan].view(np.uint32)), k
    assert np.all(np.isinf(gpu["t"][~hit]))
    assert np.array_equal(gpu["back"][hit], orc["back"][hit])
    if counters:
        assert np.array_equal(gpu["nodes"], orc["nodes"])
        assert np.array_equal(gpu["tris"], orc["tris"])


def _pair(oracle_mod, capi_mod, v0, v1, v2, **kw):
    return (oracle_mod.OracleScene(v0, v1, v2, **kw),
            capi_mod.Scene(v0, v1, v2, **kw))


def test_corner_cases(oracle_mod, capi_mod, device):
    t0 = np.array([[0, 0, 0], [1, 1, 0.5], [0, 1, 1]], F)
    t1 = t0 + np.array([2, 0, 0], F)
    v0, v1, v2 = np.stack([t0[0], t1[0]]), np.stack([t0[1], t1[1]]), np.stack([t0[2], t1[2]])
    osc, gsc = _pair(oracle_mod, capi_mod, v0, v1, v2, max_bucket=1)
    s = F(osc.tree_dump()["split"][0])
    o = np.array([[s, 0.5, 0.5], [s, 0.5, 0.5], [s, 0.5, -1.0], [-1, 0.5, 0.5],
                  [5, 0.5, 0.5], [0.5, 0.5, 0.5], [2.5, 0.6, 0.5], [s, 0.2, 0.2]], F)
    d = np.array([[1, 0, 0], [-1, 0, 0], [0, 0, 1], [1, 0, 0], [-1, 0, 0], [0, 1, 0],
                  [0, 0, -1], [0, 0, 0]], F)
    orc = osc.closest_hit(o, d, counters=True)
    assert orc["prim"][0] == -1 and orc["prim"][1] == 0  # split-plane quirk
    _assert_bit_exact(_run_gpu(capi_mod, gsc, o, d), orc)


def test_ties_and_empty(oracle_mod, capi_mod, device):
    t = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], F)
    v = [np.stack([t[i]] * 3) for i in range(3)]
    osc, gsc = _pair(oracle_mod, capi_mod, *v)
    o = np.array([[0.2, 0.2, 1], [0.2, 0.2, -1]], F)
    d = np.array([[0, 0, -1], [0, 0, 1]], F)
    g = _run_gpu(capi_mod, gsc, o, d)
    assert g["prim"].tolist() == [2, 2]
    _assert_bit_exact(g, osc.closest_hit(o, d, counters=True))
    z = np.zeros((0, 3), F)
    gz = capi_mod.Scene(z, z, z)
    r = _run_gpu(capi_mod, gz, o, d)
    assert r["prim"].tolist() == [-1, -1]


@pytest.mark.parametrize("level,seed", [(3, 0), (5, 1)])
def test_random_meshes(oracle_mod, capi_mod, device, level, seed):
    m = S.closest_hit_mesh(level, 64, seed=seed)
    osc, gsc = _pair(oracle_mod, capi_mod, m["v0"], m["v1"], m["v2"])
    (oc, dc), (oi, di) = S.closest_hit_rays(m, 256, 128, n_incoherent=50000, seed=seed)
    for o, d in ((oc, dc), (oi, di)):
        _assert_bit_exact(_run_gpu(capi_mod, gsc, o, d), osc.closest_hit(o, d, counters=True))


def test_small_buckets_deep_tree(oracle_mod, capi_mod, device):
    m = S.closest_hit_mesh(4, 64, seed=9)
    osc, gsc = _pair(oracle_mod, capi_mod, m["v0"], m["v1"], m["v2"], max_bucket=2, max_depth=24)
    (oc, dc), (oi, di) = S.closest_hit_rays(m, 128, 64, n_incoherent=20000, seed=3)
    for o, d in ((oc, dc), (oi, di)):
        _assert_bit_exact(_run_gpu(capi_mod, gsc, o, d), osc.closest_hit(o, d, counters=True))


@pytest.mark.parametrize("max_depth", [1, 3])
def test_big_leaves_and_nan_rays(oracle_mod, capi_mod, device, max_depth):
    """Shallow trees whose leaves hold hundreds of triangles: a leaf then spans
    several cooperative steps (16 clusters of 16 records each per step), so its
    best hit is carried from step to step; rays with NaN components take the
    sequential path."""
    m = S.closest_hit_mesh(4, 64, seed=5)
    osc, gsc = _pair(oracle_mod, capi_mod, m["v0"], m["v1"], m["v2"], max_depth=max_depth)
    biggest = int(osc.tree_dump()["leaf_cnt"].max())
    assert biggest > 256
    (oc, dc), (oi, di) = S.closest_hit_rays(m, 96, 64, n_incoherent=8000, seed=4)
    rng = np.random.default_rng(6)
    on = oi[:64].copy()
    dn = di[:64].copy()
    on[:16, rng.integers(0, 3, 16)] = np.nan
    dn[16:32, rng.integers(0, 3, 16)] = np.nan
    dn[32:48] = 0.0
    print(f"[gpu closest hit] max_depth {max_depth}, {len(m['v0'])} triangles, "
          f"largest leaf {biggest}", flush=True)
    for o, d in ((oc, dc), (oi, di), (on, dn)):
        _assert_bit_exact(_run_gpu(capi_mod, gsc, o, d), osc.closest_hit(o, d, counters=True))


def test_fast_reciprocal_is_exact(capi_mod, device):
    """The triangle test takes 1/det from the hardware reciprocal and one
    Newton step when 1e-6 <= |det| < 2^126: over every binary32 value of that
    range the result equals the IEEE quotient."""
    print("[gpu closest hit] exhaustive reciprocal check", flush=True)
    assert capi_mod.check_reciprocal() == 0


@pytest.fixture(scope="module")
def full_mesh_pair(oracle_mod, capi_mod, device):
    """The 1,311,488-triangle closest-hit workload mesh (icosphere level 8 + 64 boxes)."""
    m = S.closest_hit_mesh(8, 64, seed=1)
    osc = oracle_mod.OracleScene(m["v0"], m["v1"], m["v2"])
    gsc = capi_mod.Scene(m["v0"], m["v1"], m["v2"])
    return m, osc, gsc


def test_full_mesh_subsets_bit_exact(full_mesh_pair, capi_mod):
    m, osc, gsc = full_mesh_pair
    (oc, dc), (oi, di) = S.closest_hit_rays(m, 4096, 2048, n_incoherent=262144, seed=2)
    sub = slice(0, None, 29)  # ~290K coherent rays spread over the whole image
    for o, d in ((oc[sub], dc[sub]), (oi, di)):
        _assert_bit_exact(_run_gpu(capi_mod, gsc, o, d), osc.closest_hit(o, d, counters=True))


def _recompute_hits(m, o, d, prim):
    """numpy float32 restatement of the triangle test for the reported primitive
    (same operation order as xyz/geom/triangle.cpp:127-165)."""
    v0 = m["v0"][prim]
    e1 = (m["v1"][prim] - v0).astype(F)
    e2 = (m["v2"][prim] - v0).astype(F)

    def cross(a, b):
        return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                         a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1).astype(F)

    def dot(a, b):
        return ((a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]).astype(F)
    n = cross(e2, e1)
    det = dot(d, n)
    inv = (F(1) / det).astype(F)
    vmo = (v0 - o).astype(F)
    t = (inv * dot(vmo, n)).astype(F)
    q = cross(d, vmo)
    f = (inv * dot(e2, q)).astype(F)
    g = (inv * -dot(e1, q)).astype(F)
    return t, f, g, (det > 0).astype(np.uint8)


def test_full_size_properties(full_mesh_pair, capi_mod):
    """All 16,777,216 rays of the workload: every reported hit re-derives to the
    same bits, the run is deterministic, and hit statistics are sane."""
    m, osc, gsc = full_mesh_pair
    (oc, dc), (oi, di) = S.closest_hit_rays(m, 4096, 2048, seed=2)
    o = np.concatenate([oc, oi])
    d = np.concatenate([dc, di])
    a = _run_gpu(capi_mod, gsc, o, d, counters=False)
    b = _run_gpu(capi_mod, gsc, o, d, counters=False)
    for k in ("prim", "t", "f", "g", "back"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
    hit = a["prim"] >= 0
    # coherent rays cover the sphere; incoherent origins fill the whole box, so
    # most of those rays escape
    assert hit[:len(oc)].mean() > 0.2 and hit[len(oc):].mean() > 0.1
    p = a["prim"][hit]
    assert p.max() < len(m["v0"])
    t, f, g, back = _recompute_hits(m, o[hit], d[hit], p)
    assert np.array_equal(t.view(np.uint32), a["t"][hit].view(np.uint32))
    assert np.array_equal(f.view(np.uint32), a["f"][hit].view(np.uint32))
    assert np.array_equal(g.view(np.uint32), a["g"][hit].view(np.uint32))
    assert np.array_equal(back, a["back"][hit])
    assert np.all(a["t"][hit] >= F(1e-5))
    assert np.all((a["f"][hit] >= 0) & (a["g"][hit] >= 0) & (a["f"][hit] + a["g"][hit] <= 1))


def test_work_totals_match_counters(full_mesh_pair, capi_mod):
    """xyz_closest_hit_work returns the same hits and, as launch totals, the sums
    of the per-ray reference counters; culling executes fewer triangle tests
    than the reference performs."""
    from xyz_amd.hip import DeviceArray
    m, osc, gsc = full_mesh_pair
    (oc, dc), (oi, di) = S.closest_hit_rays(m, 256, 128, n_incoherent=32768, seed=5)
    o = np.concatenate([oc, oi])
    d = np.concatenate([dc, di])
    a = _run_gpu(capi_mod, gsc, o, d, counters=True)
    do, dd = DeviceArray.from_numpy(o), DeviceArray.from_numpy(d)
    prim = DeviceArray(len(o), np.int32)
    work = DeviceArray.zeros(4, np.uint64)
    gsc.closest_hit_work(do, dd, prim, work)
    nodes, ref_tests, tests, boxes = (int(x) for x in work.numpy())
    assert np.array_equal(prim.numpy(), a["prim"])
    assert nodes == int(a["nodes"].astype(np.int64).sum())
    assert ref_tests == int(a["tris"].astype(np.int64).sum())
    assert 0 < tests < ref_tests and boxes > 0
