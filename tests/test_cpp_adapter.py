"""The header-only C++ adapter (include/xyz_hip.hpp) compiles against the C ABI
and its host-only part runs; the GPU part runs on a device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "adapter_example.cpp")
BIN = os.path.join(ROOT, "xyz_amd", "adapter_example")


def compile_example(out):
    subprocess.check_call([
        "g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
        "-I", "/opt/rocm/include", SRC, "-o", out, "-L", os.path.join(ROOT, "xyz_amd"),
        "-lxyz_hip", "-L", "/opt/rocm/lib", "-lamdhip64",
        "-Wl,-rpath," + os.path.join(ROOT, "xyz_amd"), "-Wl,-rpath,/opt/rocm/lib"])


def test_adapter_compiles_and_builds_tree(capi_mod, tmp_path):
    out = str(tmp_path / "adapter_example")
    compile_example(out)
    r = subprocess.run([out, "host"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "refs" in r.stdout


@pytest.mark.gpu
def test_adapter_on_device(capi_mod, device):
    if not os.path.exists(BIN):
        pytest.skip("adapter example not built (run __graft_entry__.build())")
    r = subprocess.run([BIN, "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "film sum" in r.stdout


@pytest.mark.gpu
def test_film_files_of_a_gpu_render_match_the_oracle(oracle_mod, capi_mod, device, tmp_path):
    """The adapter renders on the device (Scene::Render) and writes the film
    through the reference's ImageFilm writers (include/xyz_hip.hpp:
    SaveAsPfm / SaveAsPfmRgb / SaveAsFloat / SaveAsHdr, xyz/core/imagefilm.cpp:
    108-241, xyz/core/imagefilm_hdr.cpp).  The oracle renders the same scene;
    the Python writers (pinned byte for byte against the C++ ones and against
    hand-derived bytes in test_film.py) write its film: every file is
    byte-identical, and the PFM decodes to the oracle's film / spp with
    non-finite values as 0."""
    import numpy as np
    from xyz_amd import film as FL
    if not os.path.exists(BIN):
        pytest.skip("adapter example not built (run __graft_entry__.build())")
    pre = str(tmp_path / "gpu")
    r = subprocess.run([BIN, "films", pre], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    F = np.float32
    v0 = np.array([0, 0, 1, 0, 0, 1, 0.2, 2, 0.2, 0.2, 2, 0.2], F).reshape(4, 3)
    v1 = np.array([1, 0, 1, 1, 1, 1, 0.8, 2, 0.8, 0.2, 2, 0.8], F).reshape(4, 3)
    v2 = np.array([1, 1, 1, 0, 1, 1, 0.8, 2, 0.2, 0.8, 2, 0.8], F).reshape(4, 3)
    o = oracle_mod.OracleScene(v0, v1, v2, tri_bsdf=np.array([0, 0, 1, 1], np.uint32),
                               bsdfs=[dict(type=0, rgb=(0.5, 0.5, 0.5)),
                                      dict(type=4, rgb=(10, 10, 10))])
    ref = o.render([0.5, 0.5, -2, 0.5, 0.5, 1, 0, 1, 0, 40.0], 48, 40, 8, 4, rr_start=0, seed=3)
    assert ref.sum() > 0
    py = str(tmp_path / "orc")
    FL.save_pfm(py + ".pfm", ref, 8)
    FL.save_pfm_rgb(py + ".rgb.pfm", ref, 8)
    FL.save_float(py + ".float", ref, 8)
    FL.save_hdr(py + ".hdr", ref, 8)
    for ext in (".pfm", ".rgb.pfm", ".float", ".hdr"):
        assert open(pre + ext, "rb").read() == open(py + ext, "rb").read(), ext
    img = FL.load_pfm(pre + ".pfm")
    want = (ref / F(8)).astype(F)
    want[~np.isfinite(want)] = 0
    assert np.array_equal(img.reshape(want.shape).view(np.uint32), want.view(np.uint32))
