"""A Metasequoia scene on the GPU: tests/golden/cornell_box.mqo, loaded by the
native loader (xyz_mqo_load), rendered by xyz_render through the reference's
thin-lens camera, against the oracle integrator on the same scene and seeds;
and the .ren front-end end to end (tests/golden/cornell_box.ren)."""
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_mqo_scene_renders_like_the_oracle(oracle_mod, capi_mod, device):
    from xyz_amd import mqo
    from xyz_amd.hip import DeviceArray
    m = mqo.load(os.path.join(GOLDEN, "cornell_box.mqo"))
    print(f"[gpu mqo] {m.n_tris} triangles, {len(m.bsdfs)} materials", flush=True)
    g = m.scene()
    o = oracle_mod.OracleScene(m.v0, m.v1, m.v2, tri_bsdf=m.tri_bsdf, bsdfs=m.bsdfs,
                               normals=(m.n0, m.n1, m.n2), uv=m.uv)
    cam = m.camera()  # the document's eye / at, thin lens f = 35 mm, F1.4
    W, H, spp = 48, 36, 16
    ref = o.render(cam, W, H, spp, max_depth=15, rr_start=0, seed=5)
    film = DeviceArray.zeros((H, W, 3), np.float32)
    g.render(cam, film, spp, 15, rr_start=0, seed=5)
    got = film.numpy()
    assert np.isfinite(got).all() and got.mean() > 0
    diff = got.view(np.uint32) != ref.view(np.uint32)
    assert not diff.any(), f"{diff.any(-1).mean():.4%} pixels differ"


def test_ren_script_renders_snapshots(capi_mod, device, tmp_path):
    from xyz_amd import film, shell
    logs = []
    sh = shell.run_script(os.path.join(GOLDEN, "cornell_box.ren"), mqo_dir=GOLDEN,
                          out_dir=str(tmp_path), ren_dir=str(tmp_path), log=logs.append)
    print("\n".join(logs), flush=True)
    assert sh.width == 64 and sh.height == 48 and sh.rendering_time == 2
    assert len(sh.outputs) >= 2
    spps = []
    for path in sh.outputs:
        name = os.path.basename(path)
        mm = re.match(r"(\d\d)h(\d\d)m(\d\d)s (\d{6})spp (\d{12})\.pfm$", name)
        assert mm, name
        spps.append(int(mm.group(4)))
        assert int(mm.group(5)) > 0  # closest-hit queries so far
    assert spps == sorted(spps) and spps[-1] > spps[0]
    img = film.load_pfm(sh.outputs[-1])
    assert img.shape == (48, 64, 3) and np.isfinite(img).all() and img.mean() > 0


@pytest.mark.parametrize("rep,sub", [(1, "IDPT"), (8, "BPT")])
def test_ren_script_other_integrators(capi_mod, device, tmp_path, rep, sub):
    from xyz_amd import film, shell
    sh = shell.Shell(mqo_dir=GOLDEN, out_dir=str(tmp_path), ren_dir=str(tmp_path),
                     log=lambda *a: None)
    sh.run(["load cornell_box", "set dim 32 24", "set rt 1 s", "set it 1 s", f"set rep {rep}",
            "ren"])
    assert sh.outputs
    assert all(os.path.basename(os.path.dirname(p)) == sub for p in sh.outputs)
    img = film.load_pfm(sh.outputs[-1])
    assert img.shape == (24, 32, 3) and np.isfinite(img).all() and img.mean() > 0
