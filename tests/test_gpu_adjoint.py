"""Path-replay adjoint on the GPU against the oracle's reference gradient and
against central finite differences of the GPU primal (common random numbers)."""
import numpy as np
import pytest

from xyz_amd import scenes as S

pytestmark = pytest.mark.gpu
F = np.float32


def _scene(tex_res, textures=None, light_scale=1.0):
    sc = S.cornell_box(textured=True, tex_res=tex_res)
    if textures is not None:
        sc["textures"] = textures
    sc["bsdfs"][-1] = dict(type=4, rgb=tuple(15.0 * light_scale for _ in range(3)))
    return sc


def _loss_weights(W, H, seed=0):
    return np.random.default_rng(seed).uniform(-1, 1, (H, W, 3)).astype(F)


def _gpu_grad(capi_mod, sc, w, spp, D, rr_start, seed):
    from xyz_amd.hip import DeviceArray
    g = capi_mod.Scene.from_dict(sc)
    dt = DeviceArray.zeros(g.n_texels, np.float32)
    de = DeviceArray.zeros(g.n_bsdfs * 3, np.float32)
    g.render_backward(sc["camera"], DeviceArray.from_numpy(w), dt, de, spp, D, rr_start=rr_start,
                      seed=seed)
    return dt.numpy(), de.numpy().reshape(-1, 3)


def _gpu_loss(capi_mod, sc, w, W, H, spp, D, rr_start, seed):
    from xyz_amd.hip import DeviceArray
    g = capi_mod.Scene.from_dict(sc)
    film = DeviceArray.zeros((H, W, 3), np.float32)
    g.render(sc["camera"], film, spp, D, rr_start=rr_start, seed=seed)
    return float((film.numpy().astype(np.float64) * w).sum())


def test_adjoint_matches_oracle_gradient(oracle_mod, capi_mod, device):
    W = H = 32
    spp, D = 16, 6
    sc = _scene(8)
    w = _loss_weights(W, H)
    gt, ge = _gpu_grad(capi_mod, sc, w, spp, D, 3, 5)
    o = oracle_mod.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc["tri_bsdf"],
                               bsdfs=sc["bsdfs"], uv=sc["uv"], textures=sc["textures"])
    ot, oe = o.render_backward(sc["camera"], W, H, spp, D, w, rr_start=3, seed=5)
    # identical paths; only the float atomics' summation order differs
    assert np.linalg.norm(gt - ot) <= 1e-4 * np.linalg.norm(ot)
    assert np.allclose(ge, oe, rtol=1e-4, atol=1e-6 * np.abs(oe).max())
    assert np.abs(ot).max() > 0 and np.abs(oe[-1]).max() > 0


def test_adjoint_matches_central_differences(capi_mod, device):
    """Without Russian roulette a sample's colour is a polynomial in the albedos
    (directions do not depend on them), so with common random numbers the
    central difference matches the path-replay gradient up to O(h^2)."""
    W = H = 24
    spp, D, rr_off, seed = 8, 6, 100, 9
    sc = _scene(4)
    w = _loss_weights(W, H, 1)
    gt, ge = _gpu_grad(capi_mod, sc, w, spp, D, rr_off, seed)
    rng = np.random.default_rng(3)
    order = np.argsort(-np.abs(gt))
    picks = list(order[:6]) + list(rng.choice(len(gt), 6, replace=False))
    h = 1e-3
    for idx in picks:
        t = [x.copy() for x in sc["textures"]]
        flat_sizes = [x.size for x in t]
        k, local = 0, int(idx)
        while local >= flat_sizes[k]:
            local -= flat_sizes[k]
            k += 1
        plus = [x.copy() for x in t]
        minus = [x.copy() for x in t]
        plus[k].reshape(-1)[local] += h
        minus[k].reshape(-1)[local] -= h
        lp = _gpu_loss(capi_mod, _scene(4, plus), w, W, H, spp, D, rr_off, seed)
        lm = _gpu_loss(capi_mod, _scene(4, minus), w, W, H, spp, D, rr_off, seed)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gt[idx]) <= 0.02 * abs(gt[idx]) + 2e-3 * np.abs(gt).max(), (idx, fd, gt[idx])
    # emission: the loss is linear in L_e
    lp = _gpu_loss(capi_mod, _scene(4, light_scale=1 + 1e-3), w, W, H, spp, D, rr_off, seed)
    lm = _gpu_loss(capi_mod, _scene(4, light_scale=1 - 1e-3), w, W, H, spp, D, rr_off, seed)
    fd = (lp - lm) / (2 * 15.0 * 1e-3)
    assert abs(fd - ge[-1].sum()) <= 0.01 * abs(ge[-1].sum())
