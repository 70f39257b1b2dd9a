"""Pins the oracle's path tracer (xyz/core/scene_pathtracing.cpp:11-144 and the
BSDFs) with scenes whose answer is known in closed form.  The reference's
random-number header is absent from its tree, so nothing in the integrator can
be pinned bit for bit against the reference; these checks pin its mean."""
import math

import numpy as np
import pytest

from xyz_amd import scenes as S

F = np.float32
LAMBERT, MIRROR, GLASS, METAL, LIGHT = 0, 1, 2, 3, 4


def _parts(*parts):
    v0 = np.concatenate([p[0] for p in parts]).astype(F)
    v1 = np.concatenate([p[1] for p in parts]).astype(F)
    v2 = np.concatenate([p[2] for p in parts]).astype(F)
    return v0, v1, v2


def _mean_and_err(film, spp):
    px = film.reshape(-1, 3) / spp
    m = px.mean(0)
    e = px.std(0) / math.sqrt(len(px))
    return m, e


def test_mirror_reflects_light_exactly(oracle_mod):
    # camera looking +z -> mirror plane through the origin with normal (-1,0,1)/sqrt2
    # reflects +z into +x -> light at x = 3 facing -x: every sample = L_e
    u = np.array([1, 0, 1]) / math.sqrt(2)
    mirror = S.quad(-u - np.array([0, 1, 0]), 2 * u, (0, 2, 0), (-1, 0, 1))
    light = S.quad((3, -1, -1), (0, 0, 2), (0, 2, 0), (-1, 0, 0))
    v0, v1, v2 = _parts(mirror, light)
    bs = np.array([0, 0, 1, 1], np.uint32)
    bsdfs = [dict(type=MIRROR), dict(type=LIGHT, rgb=(15, 7, 3))]
    sc = oracle_mod.OracleScene(v0, v1, v2, tri_bsdf=bs, bsdfs=bsdfs)
    cam = [0, 0, -3, 0, 0, 0, 0, 1, 0, 2.0]
    film = sc.render(cam, 8, 8, 4, max_depth=8)
    px = film.reshape(-1, 3) / 4
    assert np.allclose(px, [15, 7, 3], rtol=1e-5), px[:3]


def test_metal_high_shininess_is_a_fresnel_mirror(oracle_mod):
    # the 45-degree mirror scene with Metal of shininess 1e6: the lobe is a
    # near-delta, NEE's MIS weight vanishes and the implicit hit's is ~1, so a
    # sample carries L_e * albedo * F(cos 45), F the reference's
    # FakeFresnelTerm 0.9 + 0.1 (1 - c)^5 (xyz/stdafx.hpp:44-47); the lobe's
    # width leaves about 1e-3 (cos of the outgoing side below the incoming)
    u = np.array([1, 0, 1]) / math.sqrt(2)
    metal = S.quad(-u - np.array([0, 1, 0]), 2 * u, (0, 2, 0), (-1, 0, 1))
    light = S.quad((3, -1, -1), (0, 0, 2), (0, 2, 0), (-1, 0, 0))
    v0, v1, v2 = _parts(metal, light)
    bs = np.array([0, 0, 1, 1], np.uint32)
    alb = np.array([0.9, 0.8, 0.7])
    bsdfs = [dict(type=METAL, rgb=tuple(alb), shininess=1e6), dict(type=LIGHT, rgb=(15, 7, 3))]
    sc = oracle_mod.OracleScene(v0, v1, v2, tri_bsdf=bs, bsdfs=bsdfs)
    cam = [0, 0, -3, 0, 0, 0, 0, 1, 0, 2.0]
    film = sc.render(cam, 8, 8, 64, max_depth=8, seed=1)
    px = film.reshape(-1, 3) / 64
    c = math.cos(math.pi / 4)
    want = np.array([15, 7, 3]) * alb * (0.9 + 0.1 * (1 - c) ** 5)
    assert np.all(np.abs(px.mean(0) / want - 1) < 3e-3), (px.mean(0), want)
    assert np.all(np.abs(px / want - 1) < 0.02)


def test_glass_slab_normal_incidence(oracle_mod):
    # slab 0.2 thick between camera and light; F(0) = ((1.5-1)/(1.5+1))^2 = 0.04;
    # infinite internal bounces: T = (1-F)^2 / (1-F^2) = (1-F)/(1+F)
    front = S.quad((-2, -2, 0.0), (4, 0, 0), (0, 4, 0), (0, 0, -1))
    back = S.quad((-2, -2, 0.2), (4, 0, 0), (0, 4, 0), (0, 0, 1))
    light = S.quad((-5, -5, 2.0), (10, 0, 0), (0, 10, 0), (0, 0, -1))
    v0, v1, v2 = _parts(front, back, light)
    bs = np.array([0, 0, 0, 0, 1, 1], np.uint32)
    bsdfs = [dict(type=GLASS, ior=1.5), dict(type=LIGHT, rgb=(10, 10, 10))]
    sc = oracle_mod.OracleScene(v0, v1, v2, tri_bsdf=bs, bsdfs=bsdfs)
    cam = [0, 0, -3, 0, 0, 0, 0, 1, 0, 1.0]
    spp = 256
    film = sc.render(cam, 32, 32, spp, max_depth=32, seed=3)
    m, e = _mean_and_err(film, spp)
    Fr = ((1.5 - 1) / (1.5 + 1)) ** 2
    expect = 10 * (1 - Fr) / (1 + Fr)
    assert abs(m[1] - expect) < 5 * e[1] + 1e-3, (m, e, expect)


def _rect_form_factor(a, b, h):
    """Differential area to a parallel a x b rectangle whose corner is above it."""
    X, Y = a / h, b / h
    return (X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
            + Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))) / (2 * math.pi)


@pytest.mark.parametrize("rr_start", [0, 100])
def test_direct_lighting_under_rectangle(oracle_mod, rr_start):
    # floor y=0 (albedo 0.5) under a 0.5 x 0.5 light at height 1 (L_e = 4).
    # D = 2: NEE at the floor + emission through BSDF sampling, MIS-weighted.
    rho, Le, s, h = 0.5, 4.0, 0.5, 1.0
    floor = S.quad((-20, 0, -20), (40, 0, 0), (0, 0, 40), (0, 1, 0))
    light = S.quad((-s / 2, h, -s / 2), (s, 0, 0), (0, 0, s), (0, -1, 0))
    v0, v1, v2 = _parts(floor, light)
    bs = np.array([0, 0, 1, 1], np.uint32)
    bsdfs = [dict(type=LAMBERT, rgb=(rho, rho, rho)), dict(type=LIGHT, rgb=(Le, Le, Le))]
    sc = oracle_mod.OracleScene(v0, v1, v2, tri_bsdf=bs, bsdfs=bsdfs)
    cam = [0, 0.5, -0.5, 0, 0, 0, 0, 1, 0, 0.2]  # 45 deg view: footprint +-2 mm
    spp = 64
    film = sc.render(cam, 32, 32, spp, max_depth=2, rr_start=rr_start, seed=11)
    m, e = _mean_and_err(film, spp)
    expect = rho * Le * 4 * _rect_form_factor(s / 2, s / 2, h)
    assert abs(m[0] - expect) < 5 * e[0] + 1e-4 * expect, (m, e, expect)


def test_direct_lighting_depth_one_sees_only_emitters(oracle_mod):
    floor = S.quad((-20, 0, -20), (40, 0, 0), (0, 0, 40), (0, 1, 0))
    light = S.quad((-0.25, 1, -0.25), (0.5, 0, 0), (0, 0, 0.5), (0, -1, 0))
    v0, v1, v2 = _parts(floor, light)
    bs = np.array([0, 0, 1, 1], np.uint32)
    bsdfs = [dict(type=LAMBERT, rgb=(0.5, 0.5, 0.5)), dict(type=LIGHT, rgb=(4, 4, 4))]
    sc = oracle_mod.OracleScene(v0, v1, v2, tri_bsdf=bs, bsdfs=bsdfs)
    # looking up at the light from below: every sample sees the emitter -> L_e
    film = sc.render([0, 0.5, 0, 0, 1, 0.001, 0, 0, 1, 5.0], 8, 8, 2, max_depth=1)
    assert np.allclose(film.reshape(-1, 3) / 2, 4.0, rtol=1e-5)
    # looking at the floor: depth 1 stops before NEE -> 0
    film = sc.render([0, 0.3, -3, 0, 0, 0, 0, 1, 0, 0.2], 8, 8, 2, max_depth=1)
    assert np.all(film == 0)


def _integrating_sphere(level, rho, Le, cap_y):
    v, f = S.icosphere(level)
    v0, v1, v2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    c = (v0 + v1 + v2) / 3
    v0, v1, v2 = S._orient(v0, v1, v2, -c)  # normals point inward
    light = c[:, 1] > cap_y
    bs = light.astype(np.uint32)
    area = 0.5 * np.linalg.norm(np.cross(v1 - v0, v2 - v0), axis=1)
    fL = area[light].sum() / area.sum()
    bsdfs = [dict(type=LAMBERT, rgb=(rho, rho, rho)), dict(type=LIGHT, rgb=(Le, Le, Le))]
    return (v0.astype(F), v1.astype(F), v2.astype(F)), bs, bsdfs, fL


@pytest.mark.parametrize("D,rr_start", [(3, 100), (6, 0)])
def test_integrating_sphere_series(oracle_mod, D, rr_start):
    """Inside a Lambertian sphere every wall point sees every patch with form
    factor dA/A_sphere, so the wall radiance after n bounces is
    rho*Le*f*(rho*(1-f))**(n-1) (f = emitter area fraction); a depth cap D keeps
    n = 1 .. D-1.  Russian roulette must not change the mean."""
    rho, Le = 0.6, 1.0
    (v0, v1, v2), bs, bsdfs, fL = _integrating_sphere(5, rho, Le, 0.9)
    sc = oracle_mod.OracleScene(v0, v1, v2, tri_bsdf=bs, bsdfs=bsdfs)
    cam = [0, -0.2, 0, 1, -0.2, 0, 0, 1, 0, 30.0]  # looks at the equator
    spp = 16
    film = sc.render(cam, 64, 64, spp, max_depth=D, rr_start=rr_start, seed=5)
    m, e = _mean_and_err(film, spp)
    q = rho * (1 - fL)
    expect = rho * Le * fL * sum(q ** (n - 1) for n in range(1, D))
    # the faceted sphere's form factors are uniform only to ~0.5 %
    assert abs(m[0] - expect) < 5 * e[0] + 0.01 * expect, (m, e, expect)
