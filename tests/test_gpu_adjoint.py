"""Adjoint (xyz_render_backward) on the GPU against the oracle's reference
gradient and against central finite differences of the GPU primal with
common random numbers, at the inverse-rendering workload's texture size
(five walls with 64 x 64 RGB albedo textures: 61,440 parameters)."""
import numpy as np
import pytest

from xyz_amd import scenes as S

pytestmark = pytest.mark.gpu
F = np.float32


def _scene(tex_res, textures=None, le=(15.0, 15.0, 15.0)):
    sc = S.cornell_box(textured=True, tex_res=tex_res)
    if textures is not None:
        sc["textures"] = textures
    sc["bsdfs"][-1] = dict(type=4, rgb=tuple(le))
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


def _oracle_grad(oracle_mod, sc, W, H, spp, D, w, rr_start, seed):
    o = oracle_mod.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc["tri_bsdf"],
                               bsdfs=sc["bsdfs"], uv=sc["uv"], textures=sc["textures"])
    return o.render_backward(sc["camera"], W, H, spp, D, w, rr_start=rr_start, seed=seed)


@pytest.mark.parametrize("tex_res,W,spp,rr_start", [(8, 32, 16, 3), (64, 64, 16, 3), (64, 48, 8, 0)])
def test_adjoint_matches_oracle_gradient(oracle_mod, capi_mod, device, tex_res, W, spp, rr_start):
    """Same paths on both sides; only the order of the float sums differs.  At
    64 x 64 textures a workgroup's paths touch more texels than its LDS table
    holds, so deposits also take the table's overflow path to global memory."""
    H, D = W, 6
    sc = _scene(tex_res)
    w = _loss_weights(W, H)
    print(f"[gpu adjoint] tex {tex_res} film {W}x{H} spp {spp}", flush=True)
    gt, ge = _gpu_grad(capi_mod, sc, w, spp, D, rr_start, 5)
    ot, oe = _oracle_grad(oracle_mod, sc, W, H, spp, D, w, rr_start, 5)
    assert np.linalg.norm(gt - ot) <= 1e-4 * np.linalg.norm(ot)
    assert np.allclose(ge, oe, rtol=1e-4, atol=1e-6 * np.abs(oe).max())
    assert np.abs(ot).max() > 0 and np.abs(oe[-1]).max() > 0
    if tex_res == 64:
        assert np.count_nonzero(ot) > 4096  # the paths reach many distinct texels


def _fd_check(capi_mod, sc_of, tex, le, w, W, H, spp, D, seed, picks, gt, h=1e-3):
    sizes = [t.size for t in tex]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    bad = []
    for idx in picks:
        k = int(np.searchsorted(offs, idx, side="right") - 1)
        local = int(idx - offs[k])
        plus = [t.copy() for t in tex]
        minus = [t.copy() for t in tex]
        plus[k].reshape(-1)[local] += h
        minus[k].reshape(-1)[local] -= h
        lp = _gpu_loss(capi_mod, sc_of(plus, le), w, W, H, spp, D, 100, seed)
        lm = _gpu_loss(capi_mod, sc_of(minus, le), w, W, H, spp, D, 100, seed)
        fd = (lp - lm) / (2 * h)
        if not abs(fd - gt[idx]) <= 0.05 * abs(gt[idx]) + 2e-3 * np.abs(gt).max():
            bad.append((int(idx), fd, float(gt[idx])))
    return bad


def test_adjoint_matches_central_differences_64_texels(capi_mod, device):
    """Without Russian roulette a sample's colour is a polynomial in the albedos
    (directions do not depend on them), so with common random numbers the
    central difference (h = 1e-3) matches the gradient up to O(h^2) and the
    float rounding of the two films.  64 texels of the 64 x 64 textures: the
    16 with the largest gradients and 48 drawn at random."""
    W = H = 32
    spp, D, seed = 16, 6, 9
    tex = _scene(64)["textures"]
    w = _loss_weights(W, H, 1)
    gt, ge = _gpu_grad(capi_mod, _scene(64, tex), w, spp, D, 100, seed)
    rng = np.random.default_rng(3)
    order = np.argsort(-np.abs(gt))
    picks = list(order[:16]) + list(rng.choice(len(gt), 48, replace=False))
    print("[gpu adjoint] FD on 64 texels", flush=True)
    bad = _fd_check(capi_mod, lambda t, le: _scene(64, t, le), tex, (15.0, 15.0, 15.0), w, W, H,
                    spp, D, seed, picks, gt)
    assert not bad, bad
    # emission: the loss is linear in L_e
    for c in range(3):
        lp, lm = [15.0] * 3, [15.0] * 3
        lp[c] += 1e-2
        lm[c] -= 1e-2
        fd = (_gpu_loss(capi_mod, _scene(64, tex, lp), w, W, H, spp, D, 100, seed)
              - _gpu_loss(capi_mod, _scene(64, tex, lm), w, W, H, spp, D, 100, seed)) / 2e-2
        assert abs(fd - ge[-1, c]) <= 0.01 * abs(ge[-1, c]), (c, fd, ge[-1, c])


def test_adjoint_at_zero_albedo_and_dark_light_channel(oracle_mod, capi_mod, device):
    """Albedo channels at 0 and a light with a dark channel: their derivatives
    are the radiance arriving there (not 0); checked against the oracle and
    against central differences."""
    W = H = 32
    spp, D, seed = 16, 6, 13
    tex = [t.copy() for t in _scene(8)["textures"]]
    tex[0][:, :, 0] = 0.0         # floor: no red
    tex[2][3:6, 2:5, 2] = 0.0     # back wall: a patch without blue
    le = (15.0, 0.0, 12.0)
    sc = _scene(8, tex, le)
    w = _loss_weights(W, H, 2)
    gt, ge = _gpu_grad(capi_mod, sc, w, spp, D, 100, seed)
    ot, oe = _oracle_grad(oracle_mod, sc, W, H, spp, D, w, 100, seed)
    assert np.linalg.norm(gt - ot) <= 1e-4 * np.linalg.norm(ot)
    assert np.allclose(ge, oe, rtol=1e-4, atol=1e-6 * np.abs(oe).max())
    sizes = [t.size for t in tex]
    floor_red = np.arange(0, sizes[0], 3)
    assert np.abs(gt[floor_red]).max() > 1e-2 * np.abs(gt).max()
    assert abs(ge[-1, 1]) > 0.05 * abs(ge[-1, 0])
    order = floor_red[np.argsort(-np.abs(gt[floor_red]))]
    picks = list(order[:6])
    bad = _fd_check(capi_mod, lambda t, l: _scene(8, t, l), tex, le, w, W, H, spp, D, seed, picks, gt)
    assert not bad, bad
    lp, lm = list(le), list(le)
    lp[1] += 1e-2
    lm[1] -= 1e-2
    fd = (_gpu_loss(capi_mod, _scene(8, tex, lp), w, W, H, spp, D, 100, seed)
          - _gpu_loss(capi_mod, _scene(8, tex, lm), w, W, H, spp, D, 100, seed)) / 2e-2
    assert abs(fd - ge[-1, 1]) <= 0.01 * abs(ge[-1, 1]), (fd, ge[-1, 1])
