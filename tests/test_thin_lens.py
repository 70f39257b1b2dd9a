"""Thin-lens camera (the reference's default model, xyz/core/thinlenscamera.cpp:
4-14 SampleDisk, 71-90 Reset, 159-205 lens origin and direction).

CPU: the oracle's primary rays are pinned by the geometry of a thin lens.
Every ray for one film point passes through the same point of the focus plane,
and the ray from the lens centre is the pinhole ray with V = L f / (L - f).
GPU: renders with the thin lens are compared bit for bit with the oracle."""
import math

import numpy as np
import pytest

from xyz_amd import scenes as S

F = np.float32


def _basis(eye, at):
    back = (eye - at) / np.linalg.norm(eye - at)
    right = np.cross([0, 1, 0], back)
    right /= np.linalg.norm(right)
    up = np.cross(back, right)
    return back, right, up


@pytest.mark.parametrize("s,t", [(0.5, 0.5), (0.1, 0.8), (0.93, 0.07)])
def test_rays_meet_on_the_focus_plane(oracle_mod, s, t):
    eye = np.array([0.3, 1.2, -2.0])
    at = np.array([0.1, 0.9, 0.7])
    f, fnum = 0.035, 1.4
    W, H = 64, 48
    cam = [*eye, *at, 0, 1, 0, 40.0, f, fnum]
    L = np.linalg.norm(eye - at)
    V = L * f / (L - f)
    R = f / (2 * fnum)
    FP = 0.025 * L / V
    back, right, up = _basis(eye, at)
    target = eye + (s - 0.5) * FP * (W / H) * right + (t - 0.5) * FP * up - L * back
    rng = np.random.default_rng(3)
    for ls, lt in [(0.0, 0.0), (1.0, 0.0), (1.0, 0.25), (0.5, 0.6)] + [tuple(x) for x in rng.random((20, 2))]:
        o, d = oracle_mod.primary_ray(cam, W, H, s, t, ls, lt)
        o = o.astype(np.float64)
        d = d.astype(np.float64)
        # the origin lies on the lens disk: radius R sqrt(ls), angle 2 pi lt
        lens = np.array([(o - eye) @ right, (o - eye) @ up])
        assert abs((o - eye) @ back) < 1e-6
        assert np.allclose(lens, R * math.sqrt(ls) * np.array([math.cos(2 * math.pi * lt),
                                                              math.sin(2 * math.pi * lt)]),
                           atol=2e-7)
        assert abs(np.linalg.norm(d) - 1) < 1e-6
        lam = ((eye - L * back - o) @ back) / (d @ back)
        p = o + lam * d
        assert np.linalg.norm(p - target) < 2e-6 * L


def test_lens_centre_ray_is_the_pinhole_ray(oracle_mod):
    eye = np.array([0.5, 0.5, -1.4])
    at = np.array([0.5, 0.5, 0.5])
    f = 0.035
    L = np.linalg.norm(eye - at)
    V = L * f / (L - f)
    fov = math.degrees(2 * math.atan(0.0125 / V))
    thin = [*eye, *at, 0, 1, 0, 0.0, f, 2.0]
    pin = [*eye, *at, 0, 1, 0, fov]
    for s, t in [(0.5, 0.5), (0.0, 1.0), (0.25, 0.6)]:
        o, d = oracle_mod.primary_ray(thin, 32, 32, s, t, 0.0, 0.0)
        po, pd = oracle_mod.primary_ray(pin, 32, 32, s, t)
        assert np.array_equal(o, po)
        assert np.allclose(d, pd, atol=3e-7)


@pytest.mark.gpu
def test_thin_lens_needs_valid_parameters(capi_mod, device):
    from xyz_amd.capi import XyzError
    from xyz_amd.hip import DeviceArray
    sc = S.cornell_box()
    g = capi_mod.Scene.from_dict(sc)
    film = DeviceArray.zeros((4, 4, 3), np.float32)
    # the focus distance 1.9 is not beyond a 2 m focal length; f_number 0
    for bad in (list(sc["camera"][:10]) + [2.0, 1.4], list(sc["camera"][:10]) + [0.035, 0.0]):
        with pytest.raises(XyzError):
            g.render(bad, film, 1, 4)


@pytest.mark.gpu
def test_thin_lens_render_bit_exact(oracle_mod, capi_mod, device):
    from xyz_amd.hip import DeviceArray
    sc = S.cornell_box()
    cam = list(sc["camera"][:10]) + [0.035, 1.4]
    o = oracle_mod.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc["tri_bsdf"],
                               bsdfs=sc["bsdfs"])
    g = capi_mod.Scene.from_dict(sc)
    ref = o.render(cam, 40, 32, 8, 8, rr_start=3, seed=5)
    film = DeviceArray.zeros((32, 40, 3), np.float32)
    g.render(cam, film, 8, 8, rr_start=3, seed=5)
    got = film.numpy()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # the batched renderer draws the same lens positions
    films = [DeviceArray.zeros((32, 40, 3), np.float32)]
    capi_mod.render_batch([g], [cam], films, [8], 8, rr_start=3, seeds=[5],
                          tuning=capi_mod.BatchTuning(64, 2, 128, 0))
    assert np.array_equal(films[0].numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("fov", [20.0, 40.0, 75.0])
def test_pinhole_field_of_view(oracle_mod, fov):
    # xyz/core/thinlenscamera.cpp:98-124 with V = 0.0125 / tan(fov / 2): the
    # centre ray is -back, the top edge is fov/2 above it, the right edge is
    # atan(aspect tan(fov/2)) to its right
    eye = np.array([0.3, 1.2, -2.0])
    at = np.array([0.1, 0.9, 0.7])
    W, H = 64, 48
    cam = [*eye, *at, 0, 1, 0, fov]
    back, right, up = _basis(eye, at)
    o, d = oracle_mod.primary_ray(cam, W, H, 0.5, 0.5)
    assert np.array_equal(o, np.float32(eye))
    assert np.allclose(d, -back, atol=2e-7)
    _, top = oracle_mod.primary_ray(cam, W, H, 0.5, 1.0)
    _, rgt = oracle_mod.primary_ray(cam, W, H, 1.0, 0.5)
    half = math.radians(fov) / 2
    assert abs(math.acos(float(np.clip(top @ -back, -1, 1))) - half) < 2e-6
    assert float(top @ up) > 0
    assert abs(math.acos(float(np.clip(rgt @ -back, -1, 1))) - math.atan(W / H * math.tan(half))) < 2e-6
    assert float(rgt @ right) > 0
