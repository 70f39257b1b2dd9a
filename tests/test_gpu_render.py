"""Path-tracing parity on the GPU: the wavefront renderer (and the batched
megakernel) against the oracle integrator on the same seeds.

Both sides evaluate the same binary32 operations in the same order with the
same counter-based random numbers, so films are compared bit for bit.  The
transcendentals of the path (cos/sin of the azimuth, the METAL lobe's pow)
are evaluated in double by fixed series on both sides, so METAL films are
compared bit for bit too."""
import math

import numpy as np
import pytest

from xyz_amd import scenes as S

pytestmark = pytest.mark.gpu
F = np.float32
LAMBERT, MIRROR, GLASS, METAL, LIGHT = 0, 1, 2, 3, 4


def gpu_render(capi_mod, scene, cam, W, H, spp, D, rr_start=0, seed=0, sample_offset=0):
    from xyz_amd.hip import DeviceArray
    film = DeviceArray.zeros((H, W, 3), np.float32)
    st = scene.render(cam, film, spp, D, rr_start=rr_start, seed=seed, sample_offset=sample_offset)
    return film.numpy(), st


def both(oracle_mod, capi_mod, sc, **kw):
    o = oracle_mod.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc.get("tri_bsdf"),
                               bsdfs=sc.get("bsdfs"), normals=sc.get("normals"), uv=sc.get("uv"),
                               textures=sc.get("textures"), **kw)
    g = capi_mod.Scene.from_dict(sc, **kw)
    return o, g


def _assert_same(gpu, orc):
    diff = gpu.view(np.uint32) != orc.view(np.uint32)
    assert not diff.any(), (f"{diff.any(-1).mean():.4%} pixels differ; max abs "
                            f"{np.abs(gpu - orc).max()}")


def test_cornell_box_bit_exact(oracle_mod, capi_mod, device):
    sc = S.cornell_box()
    o, g = both(oracle_mod, capi_mod, sc)
    cam = sc["camera"]
    ref = o.render(cam, 64, 64, 16, max_depth=8, rr_start=3, seed=7)
    got, st = gpu_render(capi_mod, g, cam, 64, 64, 16, 8, rr_start=3, seed=7)
    _assert_same(got, ref)
    assert st["paths"] == 64 * 64 * 16


def test_cornell_box_reference_depth_and_rr(oracle_mod, capi_mod, device):
    # the reference's own settings: depth cap 16 (D = 15), RR at every vertex
    sc = S.cornell_box()
    o, g = both(oracle_mod, capi_mod, sc)
    ref = o.render(sc["camera"], 48, 40, 8, max_depth=15, rr_start=0, seed=99)
    got, _ = gpu_render(capi_mod, g, sc["camera"], 48, 40, 8, 15, rr_start=0, seed=99)
    _assert_same(got, ref)


def test_progressive_passes_equal_one_pass(capi_mod, device):
    sc = S.cornell_box()
    g = capi_mod.Scene.from_dict(sc)
    from xyz_amd.hip import DeviceArray
    a = DeviceArray.zeros((32, 32, 3), np.float32)
    g.render(sc["camera"], a, 12, 8, rr_start=3, seed=1)
    b = DeviceArray.zeros((32, 32, 3), np.float32)
    g.render(sc["camera"], b, 5, 8, rr_start=3, seed=1)
    g.render(sc["camera"], b, 7, 8, rr_start=3, seed=1, sample_offset=5)
    assert np.array_equal(a.numpy().view(np.uint32), b.numpy().view(np.uint32))


@pytest.mark.parametrize("budget", [None, "0"])
def test_workspace_reused_across_streams(capi_mod, device, monkeypatch, budget):
    """Wave buffers of 32 MiB and more (about 220 bytes per path slot) come from
    blocks the library keeps for later calls.  Renders issued on two streams
    with no synchronisation between them give the films of the same renders
    made one at a time: within the cache's budget each stream takes blocks of
    its own; with a budget of 0 every render reuses the block the one before
    released on the other stream (after an event wait), and the third, which
    needs a larger block, frees the idle smaller ones first."""
    import ctypes
    if budget is not None:
        monkeypatch.setenv("XYZ_WORKSPACE_CACHE_BYTES", budget)
    from xyz_amd import hip
    from xyz_amd.hip import DeviceArray
    sc = S.cornell_box()
    g = capi_mod.Scene.from_dict(sc)
    cam = sc["camera"]
    # (W, H, spp, seed): 0.5M, 0.4M, 1.3M and 33k path slots
    jobs = [(256, 256, 8, 1), (192, 160, 12, 2), (320, 256, 16, 3), (128, 128, 2, 4)]
    refs = [gpu_render(capi_mod, g, cam, W, H, spp, 8, rr_start=3, seed=sd)[0]
            for W, H, spp, sd in jobs]
    R = hip.rt()
    streams = [ctypes.c_void_p(), ctypes.c_void_p()]
    for s in streams:
        assert R.hipStreamCreate(ctypes.byref(s)) == 0
    try:
        films = [DeviceArray.zeros((H, W, 3), np.float32) for W, H, _, _ in jobs]
        hip.synchronize()
        for k, ((W, H, spp, sd), film) in enumerate(zip(jobs, films)):
            c = capi_mod.make_camera(cam, W, H)
            p = capi_mod.make_params(spp, 8, 3, 0.2, sd, 0)
            capi_mod.check(capi_mod.lib().xyz_render(g.h, ctypes.byref(c), ctypes.byref(p),
                                                     ctypes.c_void_p(film.data_ptr()), None,
                                                     streams[k % 2]))
        for s in streams:
            assert R.hipStreamSynchronize(s) == 0
        for f, r in zip(films, refs):
            _assert_same(f.numpy(), r)
    finally:
        for s in streams:
            R.hipStreamDestroy(s)


def test_render_captured_in_a_graph(capi_mod, device):
    """xyz_render (no stats) captured on a stream into a hipGraph: its
    workspaces come from the stream-ordered pool inside the graph.  Each of
    three replays, with eager renders and synchronisations between them, adds
    the same film as one eager render, bit for bit."""
    import ctypes
    from xyz_amd import hip
    from xyz_amd.hip import DeviceArray
    R = hip.rt()
    vp = ctypes.c_void_p
    R.hipStreamBeginCapture.argtypes = [vp, ctypes.c_int]
    R.hipStreamEndCapture.argtypes = [vp, ctypes.POINTER(vp)]
    R.hipGraphInstantiate.argtypes = [ctypes.POINTER(vp), vp, vp, ctypes.c_char_p, ctypes.c_size_t]
    R.hipGraphLaunch.argtypes = [vp, vp]
    R.hipGraphExecDestroy.argtypes = [vp]
    R.hipGraphDestroy.argtypes = [vp]
    sc = S.cornell_box()
    g = capi_mod.Scene.from_dict(sc)
    cam = sc["camera"]
    W, H, spp = 256, 256, 4  # 262,144 path slots: wave buffers above 32 MiB
    one, _ = gpu_render(capi_mod, g, cam, W, H, spp, 8, rr_start=3, seed=5)
    film = DeviceArray.zeros((H, W, 3), np.float32)
    st = vp()
    assert R.hipStreamCreate(ctypes.byref(st)) == 0
    graph, exe = vp(), vp()
    try:
        hip.synchronize()
        assert R.hipStreamBeginCapture(st, 0) == 0  # global capture mode
        c = capi_mod.make_camera(cam, W, H)
        p = capi_mod.make_params(spp, 8, 3, 0.2, 5, 0)
        rc = capi_mod.lib().xyz_render(g.h, ctypes.byref(c), ctypes.byref(p), ctypes.c_void_p(film.data_ptr()),
                                       None, st)
        assert R.hipStreamEndCapture(st, ctypes.byref(graph)) == 0
        capi_mod.check(rc)
        assert R.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
        for k in range(3):
            film.zero()
            hip.synchronize()
            assert R.hipGraphLaunch(exe, st) == 0
            assert R.hipStreamSynchronize(st) == 0
            _assert_same(film.numpy(), one)
            # eager work between the replays
            other, _ = gpu_render(capi_mod, g, cam, 64, 48, 2, 8, rr_start=3, seed=k)
            assert other.sum() > 0
    finally:
        if exe:
            R.hipGraphExecDestroy(exe)
        if graph:
            R.hipGraphDestroy(graph)
        R.hipStreamDestroy(st)


def test_textured_cornell_box_bit_exact(oracle_mod, capi_mod, device):
    sc = S.cornell_box(textured=True, tex_res=16)
    o, g = both(oracle_mod, capi_mod, sc)
    ref = o.render(sc["camera"], 40, 40, 8, max_depth=6, rr_start=3, seed=4)
    got, _ = gpu_render(capi_mod, g, sc["camera"], 40, 40, 8, 6, rr_start=3, seed=4)
    _assert_same(got, ref)


def _specular_scene(smooth=False):
    """Glass sphere and mirror box on a Lambert floor under two lights of
    different emission."""
    v, f = S.icosphere(2)
    sph = S.displaced_sphere(v, f, 0.0, 0, center=(0, 0.6, 0), radius=0.5)
    parts = [S.quad((-3, 0, -3), (6, 0, 0), (0, 0, 6), (0, 1, 0)), sph,
             S.box((1.0, 0.4, 0.3), (0.5, 0.8, 0.5), S.rot_y(25)),
             S.quad((-1.0, 2.5, -0.5), (0.6, 0, 0), (0, 0, 0.6), (0, -1, 0)),
             S.quad((0.8, 2.2, 0.5), (0.4, 0, 0), (0, 0, 0.4), (0, -1, 0))]
    v0 = np.concatenate([p[0] for p in parts]).astype(F)
    v1 = np.concatenate([p[1] for p in parts]).astype(F)
    v2 = np.concatenate([p[2] for p in parts]).astype(F)
    n = [len(p[0]) for p in parts]
    bs = np.concatenate([np.full(n[0], 0), np.full(n[1], 1), np.full(n[2], 2),
                         np.full(n[3], 3), np.full(n[4], 4)]).astype(np.uint32)
    bsdfs = [dict(type=LAMBERT, rgb=(0.6, 0.55, 0.5)), dict(type=GLASS, ior=1.5),
             dict(type=MIRROR), dict(type=LIGHT, rgb=(12, 10, 8)), dict(type=LIGHT, rgb=(3, 6, 9))]
    sc = dict(v0=v0, v1=v1, v2=v2, tri_bsdf=bs, bsdfs=bsdfs,
              camera=np.array([0.3, 1.2, -3.0, 0.2, 0.5, 0, 0, 1, 0, 45.0], F))
    if smooth:
        # per-vertex normals: the radial direction for the sphere, flat elsewhere
        ng = np.cross(v2 - v0, v1 - v0)
        ng /= np.linalg.norm(ng, axis=1, keepdims=True)
        n0, n1, n2 = ng.copy(), ng.copy(), ng.copy()
        a, b = n[0], n[0] + n[1]
        for vv, nn in ((v0, n0), (v1, n1), (v2, n2)):
            r = vv[a:b] - np.array([0, 0.6, 0], F)
            nn[a:b] = r / np.linalg.norm(r, axis=1, keepdims=True)
        sc["normals"] = (n0.astype(F), n1.astype(F), n2.astype(F))
    return sc


@pytest.mark.parametrize("smooth", [False, True])
def test_glass_mirror_two_lights_bit_exact(oracle_mod, capi_mod, device, smooth):
    sc = _specular_scene(smooth)
    o, g = both(oracle_mod, capi_mod, sc)
    ref = o.render(sc["camera"], 48, 32, 8, max_depth=12, rr_start=4, seed=2)
    got, _ = gpu_render(capi_mod, g, sc["camera"], 48, 32, 8, 12, rr_start=4, seed=2)
    _assert_same(got, ref)


@pytest.mark.parametrize("shininess", [10.0, 40.0, 3.5])
def test_metal_bit_exact(oracle_mod, capi_mod, device, shininess):
    sc = _specular_scene()
    sc["bsdfs"][2] = dict(type=METAL, rgb=(0.9, 0.8, 0.6), shininess=shininess)
    o, g = both(oracle_mod, capi_mod, sc)
    W, H, spp = 32, 24, 16
    ref = o.render(sc["camera"], W, H, spp, max_depth=8, rr_start=3, seed=3)
    got, _ = gpu_render(capi_mod, g, sc["camera"], W, H, spp, 8, rr_start=3, seed=3)
    _assert_same(got, ref)


def test_device_pow_equals_oracle(oracle_mod, capi_mod, device):
    """The device's Metal-lobe pow gives the oracle's bits (the oracle's pow
    is pinned against float64 pow in test_oracle_math.py)."""
    from xyz_amd.hip import DeviceArray
    from pow_cases import pow_inputs, pow_special
    x, y = pow_inputs(np.random.default_rng(11))
    xs, ys = pow_special()
    x = np.concatenate([x, xs]).astype(F)
    y = np.concatenate([y, ys]).astype(F)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y)
    out = DeviceArray(x.shape, np.float32)
    capi_mod.eval_pow(dx, dy, out)
    got = out.numpy()
    ref = oracle_mod.pow32(x, y)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def test_batch_equals_sequential(oracle_mod, capi_mod, device):
    from xyz_amd.hip import DeviceArray
    jobs = [(S.cornell_box(), 32, 32, 4), (_specular_scene(), 24, 16, 6),
            (S.cornell_box(textured=True, tex_res=8), 16, 24, 3)]
    scenes = [capi_mod.Scene.from_dict(j[0]) for j in jobs]
    cams = [j[0]["camera"] for j in jobs]
    seq = []
    for sc, cam, (_, W, H, spp) in zip(scenes, cams, jobs):
        f = DeviceArray.zeros((H, W, 3), np.float32)
        sc.render(cam, f, spp, 8, rr_start=3, seed=11)
        seq.append(f.numpy())
    for engine, use_graph in ((0, 0), (0, 1), (1, 0), (1, 1)):
        films = [DeviceArray.zeros((j[2], j[1], 3), np.float32) for j in jobs]
        t = capi_mod.BatchTuning(64, 2, 128, use_graph, engine)
        capi_mod.render_batch(scenes, cams, films, [j[3] for j in jobs], 8, rr_start=3,
                              seeds=[11, 11, 11], tuning=t)
        for f, s in zip(films, seq):
            assert np.array_equal(f.numpy().view(np.uint32), s.view(np.uint32))
    # graph replays of both engines: each replay adds the same samples again
    # (film += sum); eager renders and synchronisations between the replays
    for engine in (0, 1):
        films = [DeviceArray.zeros((j[2], j[1], 3), np.float32) for j in jobs]
        t = capi_mod.BatchTuning(64, 2, 128, 1, engine)
        for rep in range(3):
            capi_mod.render_batch(scenes, cams, films, [j[3] for j in jobs], 8, rr_start=3,
                                  seeds=[11, 11, 11], tuning=t)
            for f, s in zip(films, seq):
                assert np.allclose(f.numpy(), s * (rep + 1), rtol=1e-5, atol=1e-6), (engine, rep)
            other = DeviceArray.zeros((16, 16, 3), np.float32)
            scenes[0].render(cams[0], other, 2, 8, rr_start=3, seed=rep)
            other.numpy()


def test_render_errors(capi_mod, device):
    from xyz_amd.hip import DeviceArray
    sc = S.cornell_box()
    g = capi_mod.Scene.from_dict(sc)
    film = DeviceArray.zeros((8, 8, 3), np.float32)
    with pytest.raises(capi_mod.XyzError):
        g.render(sc["camera"], film, 1, 0)  # max_depth 0
    bad = sc["camera"].copy()
    bad[9] = 0.0
    with pytest.raises(capi_mod.XyzError):
        g.render(bad, film, 1, 4)  # fov 0
    g.render(sc["camera"], film, 0, 4)  # zero samples: nothing to do
    assert not film.numpy().any()
