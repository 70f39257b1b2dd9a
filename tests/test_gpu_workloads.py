"""The production-size workloads on the GPU against the oracle.

* The 1920x1080 frame of the ~2M-triangle production scene (displaced
  geodesic sphere, 512 tessellated boxes, ground plane; 70 % Lambert, 20 %
  mirror, 10 % glass; four lights) at depth 16 with Russian roulette after
  depth 4 (xyz/core/scene_pathtracing.cpp:11-144 with Glass, glass.cpp:104-165,
  and Mirror, mirror.cpp:56-105): a band of 64 rows at 1 spp bit-identical to
  the oracle; at the full frame, passes of 1 spp add up to one 4-spp render,
  and to one 33-spp render that spans two waves of path slots.
* Jobs of the 64-scene batched workload (the smallest, a mid-size and the
  largest by triangle count), at their own resolution with 4 spp: both batched
  engines give the films of one xyz_render per job bit for bit
  (xyz/integrator/integrator.cpp:119-128 over many scenes), and a row band of
  each equals the oracle."""
import numpy as np
import pytest

from xyz_amd import scenes as S

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def production():
    return S.production_scene()


def test_production_frame_band_bit_exact(oracle_mod, capi_mod, device, production):
    from xyz_amd.hip import DeviceArray
    sc = production
    W, H, D, R = 1920, 1080, 16, 4
    print(f"[gpu workloads] production scene {len(sc['v0'])} triangles", flush=True)
    g = capi_mod.Scene.from_dict(sc)
    film = DeviceArray.zeros((H, W, 3), np.float32)
    st = g.render(sc["camera"], film, 1, D, rr_start=R, seed=1)
    got = film.numpy()
    assert st["paths"] == W * H
    print("[gpu workloads] oracle kd build (2M triangles)", flush=True)
    o = oracle_mod.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc["tri_bsdf"],
                               bsdfs=sc["bsdfs"])
    rows = (540, 604)
    ref = o.render(sc["camera"], W, H, 1, D, rr_start=R, seed=1, rows=rows)
    band = slice(*rows)
    assert _same(got[band], ref[band]), (
        f"{(got[band].view(np.uint32) != ref[band].view(np.uint32)).any(-1).mean():.4%} pixels differ")
    # the band sees every material
    assert got[band].max() > 0 and np.isfinite(got).all()


def test_production_frame_passes_add_up(capi_mod, device, production):
    """Sample s of a pixel depends only on (seed, pixel, s): four 1-spp passes
    at sample offsets 0..3 sum, in sample order, to one 4-spp render, over the
    whole frame."""
    from xyz_amd.hip import DeviceArray
    sc = production
    W, H, D, R = 1920, 1080, 16, 4
    g = capi_mod.Scene.from_dict(sc)
    one = DeviceArray.zeros((H, W, 3), np.float32)
    g.render(sc["camera"], one, 4, D, rr_start=R, seed=7)
    acc = DeviceArray.zeros((H, W, 3), np.float32)
    for s in range(4):
        g.render(sc["camera"], acc, 1, D, rr_start=R, seed=7, sample_offset=s)
    assert _same(one.numpy(), acc.numpy())


def test_production_frame_across_two_waves(capi_mod, device, production):
    """33 spp of the 1920x1080 frame are 68.4M work items, more than one wave of
    2^26 path slots holds: the wave boundary falls inside sample 32, so some
    pixels get that sample from the first wave and the rest from the second.
    The film still equals 33 one-spp passes, bit for bit."""
    from xyz_amd.hip import DeviceArray
    sc = production
    W, H, D, R, spp = 1920, 1080, 16, 4, 33
    assert W * H * spp > 1 << 26
    g = capi_mod.Scene.from_dict(sc)
    one = DeviceArray.zeros((H, W, 3), np.float32)
    st = g.render(sc["camera"], one, spp, D, rr_start=R, seed=3)
    assert st["paths"] == W * H * spp
    acc = DeviceArray.zeros((H, W, 3), np.float32)
    for s in range(spp):
        g.render(sc["camera"], acc, 1, D, rr_start=R, seed=3, sample_offset=s)
    assert _same(one.numpy(), acc.numpy())


@pytest.mark.parametrize("seed", [4, 21, 60])  # 918, 10,046 and 455,322 triangles
def test_batch_jobs_match_sequential_and_oracle(oracle_mod, capi_mod, device, seed):
    from xyz_amd.hip import DeviceArray
    D, R, spp = 8, 3, 4
    j = S.batch_scene(seed)
    W = H = j["width"]
    print(f"[gpu workloads] batch job {seed}: {len(j['v0'])} triangles, {W}x{H}", flush=True)
    # the job with two small neighbours, so a wave holds several scenes
    others = [S.cornell_box(), S.cornell_box(textured=True, tex_res=8)]
    scenes = [capi_mod.Scene.from_dict(x) for x in [j] + others]
    cams = [j["camera"]] + [x["camera"] for x in others]
    sizes = [(W, H), (64, 48), (40, 40)]
    seeds = [seed, 11, 12]
    seq = []
    for sc, cam, (w, h), sd in zip(scenes, cams, sizes, seeds):
        f = DeviceArray.zeros((h, w, 3), np.float32)
        sc.render(cam, f, spp, D, rr_start=R, seed=sd)
        seq.append(f.numpy())
    for engine, use_graph in ((1, 0), (1, 1), (0, 1)):
        films = [DeviceArray.zeros((h, w, 3), np.float32) for (w, h) in sizes]
        t = capi_mod.BatchTuning(64, 2, 128, use_graph, engine)
        capi_mod.render_batch(scenes, cams, films, [spp] * 3, D, rr_start=R, seeds=seeds, tuning=t)
        for f, s in zip(films, seq):
            assert _same(f.numpy(), s), (engine, use_graph)
    o = oracle_mod.OracleScene(j["v0"], j["v1"], j["v2"], tri_bsdf=j["tri_bsdf"], bsdfs=j["bsdfs"])
    rows = (H // 2, H // 2 + 16)
    ref = o.render(j["camera"], W, H, spp, D, rr_start=R, seed=seed, rows=rows)
    assert _same(seq[0][rows[0]:rows[1]], ref[rows[0]:rows[1]])
