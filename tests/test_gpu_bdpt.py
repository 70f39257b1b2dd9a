"""Bidirectional path tracing on the GPU (xyz_render_bdpt) against the oracle
on the same seeds.

Each pixel's own colours are summed in sample order by one thread, so the own
film is compared bit for bit; light-tracing contributions reach their pixels
through float atomics, whose order changes the last bits, so the splat film is
compared to 1e-5 relative.  METAL's lobe uses the deterministic pow that the
oracle evaluates too, so metal scenes are compared the same way.  Both forms of
the renderer are compared: engine 1 (each thread traces its connections'
visibility queries) and engine 2 (the queries queued and traced together by the
cooperative closest-hit kernel, then replayed in order); engine 0 chooses by
the scene's size."""
import os

import numpy as np
import pytest

from xyz_amd import scenes as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def gpu_bdpt(scene, cam, W, H, spp, D, seed, sample_offset=0, split=True, engine=0):
    from xyz_amd.hip import DeviceArray
    own = DeviceArray.zeros((H, W, 3), np.float32)
    spl = DeviceArray.zeros((H, W, 3), np.float32) if split else None
    st = scene.render_bdpt(cam, own, spp, D, seed=seed, sample_offset=sample_offset, splat=spl,
                           engine=engine)
    return own.numpy(), (spl.numpy() if split else None), st


def scenes_of(oracle_mod, capi_mod, sc):
    o = oracle_mod.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc.get("tri_bsdf"),
                               bsdfs=sc.get("bsdfs"), normals=sc.get("normals"), uv=sc.get("uv"),
                               textures=sc.get("textures"))
    return o, capi_mod.Scene.from_dict(sc)


def assert_films(own, spl, ref_own, ref_spl):
    diff = own.view(np.uint32) != ref_own.view(np.uint32)
    assert not diff.any(), f"{diff.any(-1).mean():.4%} own pixels differ"
    assert np.allclose(spl, ref_spl, rtol=1e-5, atol=1e-7 * max(1.0, float(ref_spl.max())))


ENGINES = [1, 2]


@pytest.mark.parametrize("engine", ENGINES)
def test_cornell_box_matches_oracle(oracle_mod, capi_mod, device, engine):
    sc = S.cornell_box()
    o, g = scenes_of(oracle_mod, capi_mod, sc)
    print(f"[gpu bdpt] cornell 40x32 8 spp D 8 engine {engine}", flush=True)
    ref_own, ref_spl = o.render_bdpt(sc["camera"], 40, 32, 8, 8, seed=7)
    own, spl, st = gpu_bdpt(g, sc["camera"], 40, 32, 8, 8, seed=7, engine=engine)
    assert_films(own, spl, ref_own, ref_spl)
    assert st["paths"] == 40 * 32 * 8 and st["queries"] > st["paths"]
    assert spl.sum() > 0.05 * own.sum()
    # both forms make the same queries
    _, _, st1 = gpu_bdpt(g, sc["camera"], 40, 32, 8, 8, seed=7, engine=1)
    assert st["queries"] == st1["queries"]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("smooth", [False, True])
def test_glass_mirror_two_lights_matches_oracle(oracle_mod, capi_mod, device, smooth, engine):
    from test_gpu_render import _specular_scene
    sc = _specular_scene(smooth)
    o, g = scenes_of(oracle_mod, capi_mod, sc)
    print(f"[gpu bdpt] specular scene smooth={smooth} engine {engine}", flush=True)
    ref_own, ref_spl = o.render_bdpt(sc["camera"], 32, 24, 8, 12, seed=2)
    own, spl, _ = gpu_bdpt(g, sc["camera"], 32, 24, 8, 12, seed=2, engine=engine)
    assert_films(own, spl, ref_own, ref_spl)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("camera", ["camera", "mirror_camera"])
def test_mirror_wall_matches_oracle(oracle_mod, capi_mod, device, camera, engine):
    """Caustics through a mirror by light tracing, and eye subpaths that end
    at the mirror as the reference's do."""
    sc = S.mirror_wall_scene("mirror")
    o, g = scenes_of(oracle_mod, capi_mod, sc)
    print(f"[gpu bdpt] mirror wall {camera} 48x36 16 spp D 8 engine {engine}", flush=True)
    ref_own, ref_spl = o.render_bdpt(sc[camera], 48, 36, 16, 8, seed=5)
    own, spl, _ = gpu_bdpt(g, sc[camera], 48, 36, 16, 8, seed=5, engine=engine)
    assert_films(own, spl, ref_own, ref_spl)
    assert spl.sum() > 0.1 * own.sum()


def test_large_scene_takes_the_queued_form(oracle_mod, capi_mod, device):
    """A 221,114-triangle scene of the batch workload: engine 0 takes the
    queued form (at least 300,000 kd leaf references) and equals engine 2 bit
    for bit; a band of rows equals the oracle, and the per-thread form gives
    the same own film."""
    sc = S.batch_scene(39)
    assert len(sc["v0"]) == 221114
    o, g = scenes_of(oracle_mod, capi_mod, sc)
    assert g.info()["n_leaf_refs"] >= 300000
    W, H, spp, D = 48, 32, 4, 8
    print(f"[gpu bdpt] batch scene 39 {W}x{H} {spp} spp D {D}", flush=True)
    own0, spl0, st0 = gpu_bdpt(g, sc["camera"], W, H, spp, D, seed=3, engine=0)
    own2, spl2, st2 = gpu_bdpt(g, sc["camera"], W, H, spp, D, seed=3, engine=2)
    own1, spl1, st1 = gpu_bdpt(g, sc["camera"], W, H, spp, D, seed=3, engine=1)
    assert np.array_equal(own0.view(np.uint32), own2.view(np.uint32))
    assert np.array_equal(own1.view(np.uint32), own2.view(np.uint32))
    assert st0["queries"] == st1["queries"] == st2["queries"]
    assert np.allclose(spl1, spl2, rtol=1e-5, atol=1e-7 * max(1.0, float(spl2.max())))
    ref_own, ref_spl = o.render_bdpt(sc["camera"], W, H, spp, D, seed=3)
    assert_films(own2, spl2, ref_own, ref_spl)
    assert own2.sum() > 0


@pytest.mark.parametrize("engine", ENGINES)
def test_mqo_scene_with_metal(oracle_mod, capi_mod, device, engine):
    from xyz_amd import mqo
    m = mqo.load(os.path.join(ROOT, "tests", "golden", "cornell_box.mqo"))
    sc = dict(v0=m.v0, v1=m.v1, v2=m.v2, tri_bsdf=m.tri_bsdf, bsdfs=m.bsdfs,
              normals=(m.n0, m.n1, m.n2), uv=m.uv)
    o, g = scenes_of(oracle_mod, capi_mod, sc)
    cam = m.camera(pinhole_fov_deg=40.0)
    print(f"[gpu bdpt] mqo scene engine {engine}", flush=True)
    ref_own, ref_spl = o.render_bdpt(cam, 32, 24, 16, 15, seed=4)
    own, spl, _ = gpu_bdpt(g, cam, 32, 24, 16, 15, seed=4, engine=engine)
    assert_films(own, spl, ref_own, ref_spl)


def test_one_film_and_passes(capi_mod, device):
    sc = S.cornell_box()
    g = capi_mod.Scene.from_dict(sc)
    own, spl, _ = gpu_bdpt(g, sc["camera"], 24, 24, 4, 8, seed=5)
    both, _, _ = gpu_bdpt(g, sc["camera"], 24, 24, 4, 8, seed=5, split=False)
    assert np.allclose(both, own + spl, rtol=1e-5, atol=1e-7)
    # samples [0, 2) and [2, 4) own colours add up to samples [0, 4)
    a, _, _ = gpu_bdpt(g, sc["camera"], 24, 24, 2, 8, seed=5)
    b, _, _ = gpu_bdpt(g, sc["camera"], 24, 24, 2, 8, seed=5, sample_offset=2)
    assert np.allclose(a + b, own, rtol=1e-5, atol=1e-7)


def test_bdpt_equals_pt_for_direct_light(capi_mod, device):
    from test_oracle_bdpt import floor_under_light
    sc = floor_under_light()
    g = capi_mod.Scene.from_dict(sc)
    from xyz_amd.hip import DeviceArray
    W, H, spp, chunks = 64, 48, 256, 8
    bd, pt = [], []
    for k in range(chunks):
        own, spl, _ = gpu_bdpt(g, sc["camera"], W, H, spp, 8, seed=11, sample_offset=k * spp)
        bd.append((own + spl).astype(np.float64).sum((0, 1)) / (spp * W * H))
        film = DeviceArray.zeros((H, W, 3), np.float32)
        g.render(sc["camera"], film, spp, 8, rr_start=100, seed=12, sample_offset=k * spp)
        pt.append(film.numpy().astype(np.float64).sum((0, 1)) / (spp * W * H))
    bd, pt = np.array(bd), np.array(pt)
    z = (bd.mean(0) - pt.mean(0)) / np.sqrt(bd.var(0, ddof=1) / chunks + pt.var(0, ddof=1) / chunks)
    print(f"bdpt {bd.mean(0)}, pt {pt.mean(0)}, z {z}")
    assert np.all(np.abs(z) < 4.5)


def test_bdpt_errors(capi_mod, device):
    sc = S.cornell_box()
    g = capi_mod.Scene.from_dict(sc)
    from xyz_amd.hip import DeviceArray
    film = DeviceArray.zeros((8, 8, 3), np.float32)
    with pytest.raises(capi_mod.XyzError) as ei:
        g.render_bdpt(sc["camera"], film, 1, 17)
    assert ei.value.code == capi_mod.XYZ_E_UNSUPPORTED
    with pytest.raises(capi_mod.XyzError) as ei:
        g.render_bdpt(sc["camera"], film, 1, 8, engine=3)
    assert ei.value.code == capi_mod.XYZ_E_INVALID
    assert not film.numpy().any()
