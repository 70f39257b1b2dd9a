"""Host logic of the library (no GPU): the product's SAH builder must produce the
oracle's tree exactly, and the C ABI must export every declared symbol."""
import os
import re

import numpy as np
import pytest

from xyz_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _product_dump(tree):
    nodes, prims = tree.export()
    w0, w1 = nodes[:, 0], nodes[:, 1]
    kind = (w0 & 3).astype(np.uint32)
    cnt = np.where(kind == 3, w0 >> 2, 0)
    kind = np.where((kind == 3) & (cnt == 0), 4, kind)
    split = np.where(kind < 3, w1, 0).view(np.float32)
    # interior nodes: left child = i + 1, right child stored
    return kind, split, cnt, prims, nodes


def _compare(oracle_mod, capi_mod, v0, v1, v2, **kw):
    osc = oracle_mod.OracleScene(v0, v1, v2, **kw)
    od = osc.tree_dump()
    tree = capi_mod.KDTree(v0, v1, v2, **kw)
    kind, split, cnt, prims, nodes = _product_dump(tree)
    assert np.array_equal(od["kind"], kind)
    inner = kind < 3
    assert np.array_equal(od["split"][inner], split[inner])
    assert np.array_equal(od["leaf_cnt"], cnt)
    assert np.array_equal(od["refs"], prims)
    # pre-order: each interior node's right child follows its left subtree
    info = tree.info()
    assert info["n_nodes"] == len(kind) and info["n_leaf_refs"] == len(prims)
    oc = osc.tree_counts()
    assert np.array_equal(np.array(info["bmin"], np.float32), oc["bmin"])
    assert np.array_equal(np.array(info["bmax"], np.float32), oc["bmax"])
    return info


@pytest.mark.parametrize("level,boxes,seed", [(2, 8, 0), (4, 64, 1), (5, 64, 2)])
def test_builder_matches_oracle(oracle_mod, capi_mod, level, boxes, seed):
    m = S.closest_hit_mesh(level, boxes, seed=seed)
    _compare(oracle_mod, capi_mod, m["v0"], m["v1"], m["v2"])


def test_builder_matches_oracle_small_buckets_and_ties(oracle_mod, capi_mod):
    rng = np.random.default_rng(3)
    tri = np.round(rng.uniform(-1, 1, (500, 3, 3)) * 4) / 4  # many equal events
    tri = tri.astype(np.float32)
    _compare(oracle_mod, capi_mod, tri[:, 0], tri[:, 1], tri[:, 2], max_bucket=2, max_depth=10)
    _compare(oracle_mod, capi_mod, tri[:, 0], tri[:, 1], tri[:, 2], max_bucket=1, max_depth=20)


def test_builder_cornell_box(oracle_mod, capi_mod):
    c = S.cornell_box()
    info = _compare(oracle_mod, capi_mod, c["v0"], c["v1"], c["v2"])
    assert info["n_tris"] == 36


def test_builder_empty_and_invalid(capi_mod):
    z = np.zeros((0, 3), np.float32)
    t = capi_mod.KDTree(z, z, z)
    nodes, prims = t.export()
    assert nodes.shape == (1, 2) and nodes[0, 0] == 3 and prims.size == 0
    bad = np.array([[0, 0, 0]], np.float32)
    nan = np.array([[np.nan, 0, 0]], np.float32)
    with pytest.raises(capi_mod.XyzError) as ei:
        capi_mod.KDTree(bad, nan, bad)
    assert ei.value.code == capi_mod.XYZ_E_INVALID
    assert "non-finite" in str(ei.value)


def test_library_exports_every_declared_symbol(capi_mod):
    header = open(os.path.join(ROOT, "include", "xyz_hip.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    declared = set(re.findall(r"\b(xyz_[a-z_]+)\s*\(", header))
    assert declared == set(capi_mod.EXPORTS)
    L = capi_mod.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert capi_mod.lib().xyz_version().decode().startswith("xyz-hip")


def test_check_reciprocal_rejects_null_without_a_device(capi_mod):
    """The self-check's argument test runs before any HIP call."""
    L = capi_mod.lib()
    assert L.xyz_check_reciprocal(None, None) == capi_mod.XYZ_E_INVALID
    assert "NULL" in L.xyz_last_error().decode()
