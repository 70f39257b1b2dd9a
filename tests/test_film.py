"""Film files (ImageFilm::SaveAsPfm / SaveAsPfmRgb / SaveAsFloat / SaveAsHdr,
xyz/core/imagefilm.cpp:108-241, xyz/core/imagefilm_hdr.cpp): the C++ adapter's
writers (include/xyz_hip.hpp) and the Python ones (xyz_amd/film.py) write
byte-identical files; the files decode to the film's values within each
format's precision; RGBE and run-length cases derived by hand."""
import math
import os
import subprocess

import numpy as np
import pytest

from xyz_amd import film as FL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F = np.float32


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("film") / "film_writer")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
                           "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                           os.path.join(ROOT, "tests", "cpp", "film_writer.cpp"), "-o", out])
    return out


def _film(h, w, seed):
    rng = np.random.default_rng(seed)
    f = rng.gamma(0.7, 2.0, (h, w, 3)).astype(F) * F(37)
    f[rng.random((h, w)) < 0.05] *= F(-0.3)           # negative sums
    f[rng.random((h, w)) < 0.05] = 0                   # black
    f[: h // 3, : w // 2] = f[0, 0]                    # long runs
    f[h // 2, 3] = [np.nan, 1, 2]
    f[h // 2, 5] = [np.inf, 0, 0]
    f[h - 1, w - 1] = [1e-35, 1e-36, 0]                # below the RGBE threshold
    return f


@pytest.mark.parametrize("h,w,spp", [(21, 37, 64), (5, 7, 3), (3, 300, 1)])
def test_cpp_and_python_writers_agree(writer, tmp_path, h, w, spp):
    film = _film(h, w, h * w)
    raw = tmp_path / "film.raw"
    film.tofile(raw)
    prefix = str(tmp_path / "cpp")
    subprocess.check_call([writer, str(raw), str(w), str(h), str(spp), prefix])
    py = str(tmp_path / "py")
    FL.save_pfm(py + ".pfm", film, spp)
    FL.save_pfm_rgb(py + ".rgb.pfm", film, spp)
    FL.save_float(py + ".float", film, spp)
    FL.save_hdr(py + ".hdr", film, spp)
    for ext in (".pfm", ".rgb.pfm", ".float", ".hdr"):
        a = open(prefix + ext, "rb").read()
        b = open(py + ext, "rb").read()
        assert a == b, ext


def test_files_decode_to_the_film(tmp_path):
    h, w, spp = 17, 40, 16
    film = np.abs(_film(h, w, 5))
    film[~np.isfinite(film)] = 1
    mean = film.reshape(-1, 3) * (F(1) / F(spp))
    FL.save_pfm(tmp_path / "a.pfm", film, spp)
    assert np.array_equal(FL.load_pfm(tmp_path / "a.pfm").reshape(-1, 3), mean)
    FL.save_float(tmp_path / "a.float", film, spp)
    rgb = FL.ccir601_xyz_to_rgb(mean)
    got = np.fromfile(tmp_path / "a.float", "<f4").reshape(-1, 3)
    assert np.array_equal(got, np.maximum(rgb, 0))
    FL.save_hdr(tmp_path / "a.hdr", film, spp)
    dec = FL.load_hdr(tmp_path / "a.hdr")[::-1].reshape(-1, 3)   # top row first in the file
    d = rgb.max(1)
    pos = (rgb >= 0).all(1) & (d > 1e-30)
    # RGBE keeps 8 bits of mantissa relative to the largest component
    err = np.abs(dec[pos] - rgb[pos]).max(1) / d[pos]
    assert err.max() <= 2.0 ** -7


def test_ccir601_matrix():
    m = FL.ccir601_xyz_to_rgb(np.eye(3, dtype=F))
    want = np.array([[1.9104, -0.5338, -0.2891], [-0.9844, 1.9985, -0.0279],
                     [0.0585, -0.1187, 0.9017]], F)
    assert np.array_equal(m.T, want)


def test_rgbe_known_values():
    # d = 1: frexp -> (0.5, 1), f = 127.99995 -> 127, exponent byte 129
    assert FL.rgbe_encode([[1, 1, 1]]).tolist() == [[127, 127, 127, 129]]
    # d = 0.5: frexp -> (0.5, 0), f = 255.9999 -> 127, 63, 0, exponent byte 128
    assert FL.rgbe_encode([[0.5, 0.25, 0]]).tolist() == [[127, 63, 0, 128]]
    assert FL.rgbe_encode([[1e-33, 0, 0]]).tolist() == [[0, 0, 0, 0]]
    # decoding (m + 0.5) 2^(e - 136): 127.5 / 128 for d = 1
    assert math.isclose((127 + 0.5) * 2.0 ** (129 - 136), 127.5 / 128)


def test_run_length_packets():
    px = np.tile(np.array([[10, 20, 30, 129]], np.uint8), (8, 1))
    assert FL.rgbe_line(px) == bytes([2, 2, 0, 8, 136, 10, 136, 20, 136, 30, 136, 129])
    assert FL.rgbe_line(px[:7]) == px[:7].tobytes()   # short lines are flat
    # a b b before a run of five c: the three go out as literals, then the run
    line = np.zeros((8, 4), np.uint8)
    line[:, 0] = [1, 2, 2, 3, 3, 3, 3, 3]
    enc = FL.rgbe_line(line)
    assert enc[4:] == bytes([3, 1, 2, 2, 133, 3]) + bytes([136, 0]) * 3
    # b b before a run of six c: a short run packet (2 + 128) for the two
    line[:, 0] = [2, 2, 3, 3, 3, 3, 3, 3]
    assert FL.rgbe_line(line)[4:] == bytes([130, 2, 134, 3]) + bytes([136, 0]) * 3
    # more than 128 distinct bytes split into literal packets of 128
    wide = np.zeros((300, 4), np.uint8)
    wide[:, 0] = np.arange(300) % 251
    enc = FL.rgbe_line(wide)
    assert enc[4] == 128 and enc[4 + 129] == 128 and enc[4 + 258] == 44
