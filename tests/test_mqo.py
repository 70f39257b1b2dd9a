"""Metasequoia loader (xyz_mqo_load, the reference's Scene::Load,
xyz/core/scene_mqo.cpp:547-777) and the .ren front-end (xyz/core/shell.cpp),
on the CPU.

The native loader is compared with the oracle restatement (oracle/mqo_oracle.py)
on tests/golden/cornell_box.mqo, bit for bit; the oracle itself is pinned by
closed forms that do not come from either implementation: the reference
camera's default eye, rotations of the Scene chunk, smooth normals of an
octahedron and a cube, the CCIR 601-1 row sums, the material-flag table and
quad areas.  The spectral tables it restates are checked against published
colorimetry and physics: the Macbeth chart's Munsell luminances and measured
hues, the CIE 1931 functions' equal integrals and peaks, and the mercury
lines."""
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
LAMBERT, MIRROR, GLASS, METAL, LIGHT = 0, 1, 2, 3, 4
HEADER = "Metasequoia Document\nFormat Text Ver 1.0\n\n"


@pytest.fixture(scope="module")
def mqo(capi_mod):
    from xyz_amd import mqo
    return mqo


@pytest.fixture(scope="module")
def orc():
    from oracle import mqo_oracle
    return mqo_oracle


def material_line(name, col=(1, 1, 1, 1), emi=0.0, spc=0.0, power=5.0):
    return (f'\t"{name}" shader(3) col({col[0]:.3f} {col[1]:.3f} {col[2]:.3f} {col[3]:.3f}) '
            f"dif(0.800) amb(0.600) emi({emi:.3f}) spc({spc:.3f}) power({power:.2f})")


def document(materials, objects, scene=None):
    out = HEADER
    if scene is not None:
        out += "Scene {\n" + "".join(f"\t{l}\n" for l in scene) + "}\n"
    out += f"Material {len(materials)} {{\n" + "\n".join(materials) + "\n}\n"
    for name, shading, verts, faces in objects:
        out += (f'Object "{name}" {{\n\tdepth 0\n\tvisible 15\n\tshading {shading}\n'
                f"\tvertex {len(verts)} {{\n")
        out += "".join("\t\t" + " ".join(f"{x:.4f}" for x in v) + "\n" for v in verts)
        out += f"\t}}\n\tface {len(faces)} {{\n"
        out += "".join(f"\t\t3 V({a} {b} {c}) M({m})\n" for a, b, c, m in faces)
        out += "\t}\n}\n"
    return out + "Eof\n"


def test_fixture_matches_oracle_bit_for_bit(mqo, orc):
    path = os.path.join(GOLDEN, "cornell_box.mqo")
    m = mqo.load(path)
    with open(path, newline="") as f:
        o = orc.load(f.read())
    assert m.n_tris == len(o["bsdf"]) == 65
    for k, name in enumerate(("v0", "v1", "v2")):
        assert np.array_equal(getattr(m, name), o["v"][:, k])
    for k, name in enumerate(("n0", "n1", "n2")):
        a, b = getattr(m, name), o["n"][:, k]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    assert np.array_equal(m.uv, o["uv"])
    assert np.array_equal(m.tri_bsdf, o["bsdf"])
    assert len(m.bsdfs) == len(o["bsdfs"]) == 10
    for b, ob in zip(m.bsdfs, o["bsdfs"]):
        assert b["type"] == ob["type"]
        assert np.array_equal(np.float32(b["rgb"]), np.float32(ob["rgb"]))
        assert np.float32(b["shininess"]) == ob["shininess"]
    i = m.info
    assert i["n_objects"] == o["n_objects"] == 7  # the invisible one is counted
    assert i["n_smooth_objects"] == 3
    assert i["n_lights"] == o["n_lights"] == 2
    assert np.float32(i["surface_area"]) == o["surface_area"]
    assert np.float32(i["light_area"]) == o["light_area"]
    assert np.array_equal(np.float32(i["light_power"]), np.float32(o["light_power"]))
    assert np.array_equal(np.float32(i["eye"]), np.float32(o["eye"]))
    assert np.array_equal(np.float32(i["at"]), np.float32(o["at"]))


def test_fixture_material_table_and_quirks(mqo):
    m = mqo.load(os.path.join(GOLDEN, "cornell_box.mqo"))
    types = [b["type"] for b in m.bsdfs]
    # white (spectral), red (gamma), green (spectral), lamp, mirror, glass, gold,
    # orange-yellow, wood (texture), swallowed
    assert types == [LAMBERT, LAMBERT, LAMBERT, LIGHT, MIRROR, GLASS, METAL, LAMBERT, LAMBERT,
                     LAMBERT]
    assert m.bsdfs[6]["shininess"] == 30.0
    # "mcc-orange-yellow" contains "mcc-orange", which is matched first
    o = mqo.loads(document([material_line("x mcc-orange")], []))
    assert m.bsdfs[7]["rgb"] == o.bsdfs[0]["rgb"]
    # tex("...") reads the file up to the next ')', which swallows the next
    # material's name and shader: it gets the defaults, spc 0 -> black Lambert
    assert m.bsdfs[9]["rgb"] == (0.0, 0.0, 0.0)
    # the quad face becomes the degenerate triangle (0, 0, 0) with material 0
    room_last = 10
    assert np.array_equal(m.v0[room_last], m.v1[room_last])
    assert np.array_equal(m.v0[room_last], m.v2[room_last])
    assert m.tri_bsdf[room_last] == 0
    # light emission L_e = Phi / (pi * A_light), Phi = the mercury arc lamp in XYZ
    phi = np.array(m.info["light_power"], np.float64)
    le = phi / (math.pi * m.info["light_area"])
    assert np.allclose(m.bsdfs[3]["rgb"], le, rtol=1e-6)
    # metres: the room is the unit cube, the light a 0.25 m square at y = 0.999
    assert m.v0[:10].min() == 0.0 and m.v0[:10].max() == 1.0
    assert abs(m.info["light_area"] - 0.0625) < 1e-6
    assert np.allclose(m.v0[11:13, 1], 0.999)


def test_camera_closed_forms(mqo, orc):
    mats = [material_line("w")]
    # the default Scene chunk reproduces ThinLensCamera's default eye (0, 1, 4)
    for text in (document(mats, [], scene=["pos 0.0 0.0 400.0", "lookat 0.0 100.0 0.0",
                                           "head 0.0", "pich 0.0"]),
                 document(mats, [], scene=["amb 0.25 0.25 0.25"])):
        m = mqo.loads(text)
        assert m.info["has_camera"] == 1
        assert np.allclose(m.info["eye"], (0, 1, 4), atol=1e-6)
        assert np.allclose(m.info["at"], (0, 1, 0), atol=0)
        o = orc.load(text)
        assert np.array_equal(np.float32(m.info["eye"]), np.float32(o["eye"]))
    # no Scene chunk: the camera keeps its defaults
    m = mqo.loads(document(mats, []))
    assert m.info["has_camera"] == 0 and m.info["eye"] == [0.0, 1.0, 4.0]
    # head = 90 degrees turns the offset (0, 0, 4 m) to (4 m, 0, 0)
    m = mqo.loads(document(mats, [], scene=[f"pos 0 0 400", "lookat 0 100 0",
                                            f"head {math.pi / 2:.7f}", "pich 0"]))
    assert np.allclose(m.info["eye"], (4, 1, 0), atol=2e-6)
    # pitch = 90 degrees lifts it to (0, 4 m, 0) above the gazing point
    m = mqo.loads(document(mats, [], scene=[f"pos 0 0 400", "lookat 0 100 0",
                                            "head 0", f"pich {math.pi / 2:.7f}"]))
    assert np.allclose(m.info["eye"], (0, 5, 0), atol=2e-6)


def octahedron(r=10.0):
    P = [(r, 0, 0), (-r, 0, 0), (0, r, 0), (0, -r, 0), (0, 0, r), (0, 0, -r)]
    F = []
    for sx in (0, 1):
        for sy in (2, 3):
            for sz in (4, 5):  # cross(e2, e1) outward
                odd = (sx == 1) ^ (sy == 3) ^ (sz == 5)
                F.append((sx, sy, sz, 0) if odd else (sx, sz, sy, 0))
    return P, F


def test_smooth_normals_closed_forms(mqo, orc):
    mats = [material_line("w", spc=1.0, power=0.0)]
    # octahedron: four equal 60-degree corners at each vertex, so the smooth
    # normal is the vertex direction
    P, Fc = octahedron()
    m = mqo.loads(document(mats, [("o", 1, P, Fc)]))
    for k, name in enumerate(("n0", "n1", "n2")):
        n = getattr(m, name)
        for t, face in enumerate(Fc):
            expect = np.array(P[face[k]], np.float64) / 10.0
            assert np.allclose(n[t], expect, atol=2e-7), (t, name)
    # the weight of a corner is its interior angle / pi: with eb = p2 - p0 and
    # ec = p0 - p1, 1 - acos(eb.ec / |eb||ec|) / pi = angle(p1 - p0, p2 - p0) / pi.
    # A vertex shared by a 90-degree corner (normal -z) and a 45-degree corner
    # (normal +y) gets normalize(0.5 * (0, 0, -1) + 0.25 * (0, 1, 0)).
    V = [(0, 0, 0), (10, 0, 0), (0, 10, 0), (10, 0, 10), (0, 0, 10)]
    faces = [(0, 1, 2, 0), (0, 3, 4, 0)]
    m = mqo.loads(document(mats, [("fan", 1, V, faces)]))
    e = np.array([0.0, 0.25, -0.5]) / math.hypot(0.25, 0.5)
    assert np.allclose(m.n0[0], e, atol=1e-6) and np.allclose(m.n0[1], e, atol=1e-6)
    # a cube split along one diagonal per face: a corner on the diagonal has two
    # 45-degree corners, one off it a 90-degree corner, both weigh 1/2, so every
    # corner normal is the diagonal direction (into the cube: cross(e2, e1) of
    # these quads points inward)
    V = [(x, y, z) for x in (0, 10) for y in (0, 10) for z in (0, 10)]
    vid = {v: i for i, v in enumerate(V)}
    quads = [((0, 0, 0), (10, 0, 0), (10, 0, 10), (0, 0, 10), (0, -1, 0)),
             ((0, 10, 0), (0, 10, 10), (10, 10, 10), (10, 10, 0), (0, 1, 0)),
             ((0, 0, 0), (0, 0, 10), (0, 10, 10), (0, 10, 0), (-1, 0, 0)),
             ((10, 0, 0), (10, 10, 0), (10, 10, 10), (10, 0, 10), (1, 0, 0)),
             ((0, 0, 0), (0, 10, 0), (10, 10, 0), (10, 0, 0), (0, 0, -1)),
             ((0, 0, 10), (10, 0, 10), (10, 10, 10), (0, 10, 10), (0, 0, 1))]
    faces = []
    for a, b, c, d, _ in quads:
        a, b, c, d = (vid[v] for v in (a, b, c, d))
        faces += [(a, b, c, 0), (a, c, d, 0)]
    m = mqo.loads(document(mats, [("c", 1, V, faces)]))
    o = orc.load(document(mats, [("c", 1, V, faces)]))
    for t, (a, b, c, _) in enumerate(faces):
        for k, v in enumerate((a, b, c)):
            e = (np.array([5.0, 5.0, 5.0]) - np.array(V[v])) / (5.0 * math.sqrt(3.0))
            got = getattr(m, ("n0", "n1", "n2")[k])[t]
            assert np.allclose(got, e, atol=1e-6), (t, k, got, e)
            assert np.array_equal(got.view(np.uint32), o["n"][t, k].view(np.uint32))
    # flat shading: every corner carries the face normal of cross(e2, e1)
    m = mqo.loads(document(mats, [("c", 0, V, faces)]))
    for t in range(len(faces)):
        inward = -np.array(quads[t // 2][4], np.float64)
        for name in ("n0", "n1", "n2"):
            assert np.allclose(getattr(m, name)[t], inward, atol=1e-7), (t, name)


def test_colours_closed_forms(mqo):
    # white through pow(., 2.2) and the CCIR 601-1 matrix: the row sums, clamped
    m = mqo.loads(document([material_line("w", spc=1.0, power=0.0)], []))
    x, y, z = m.bsdfs[0]["rgb"]
    assert x == pytest.approx(0.6069 + 0.1735 + 0.2003, abs=1e-6)
    assert y == pytest.approx(1.0, abs=1e-6)
    assert z == 1.0  # 1.1823 clamped
    # spc multiplies the colour: spc(0) makes a black Lambert; 0.5^2.2 per channel
    m = mqo.loads(document([material_line("k", spc=0.0), material_line("h", (0.5, 0, 0, 1),
                                                                         spc=1.0, power=0.0)], []))
    assert m.bsdfs[0]["rgb"] == (0.0, 0.0, 0.0)
    r = 0.5 ** 2.2
    assert np.allclose(m.bsdfs[1]["rgb"], (0.6069 * r, 0.2989 * r, 0.0), rtol=1e-6)
    # spectral colours are linear in the spectrum: mcc-neutral-35 is the flat
    # 0.09 spectrum except its first two samples (0.08), mcc-black the flat 0.03
    m = mqo.loads(document([material_line("mcc-black"), material_line("mcc-white")], []))
    blk = np.array(m.bsdfs[0]["rgb"])
    assert np.all(blk > 0) and np.all(blk < 0.04)
    wht = np.array(m.bsdfs[1]["rgb"])
    assert np.all(wht > 0.8) and np.all(wht <= 1.0)


def test_material_flag_table(mqo):
    mats = [material_line("light", emi=1.0),
            material_line("glass", (1, 1, 1, 0.5)),
            material_line("glass-light", (1, 1, 1, 0.5), emi=1.0),
            material_line("mirror", spc=1.0, power=100.0),
            material_line("metal", spc=1.0, power=20.0),
            material_line("lambert-spc1-power0", spc=1.0, power=0.0),
            material_line("lambert-spc-half", spc=0.5, power=100.0)]
    m = mqo.loads(document(mats, []))
    assert [b["type"] for b in m.bsdfs] == [LIGHT, GLASS, LIGHT, MIRROR, METAL, LAMBERT, LAMBERT]
    # no light triangle: L_e stays 0
    assert m.bsdfs[0]["rgb"] == (0.0, 0.0, 0.0)
    # Material 0: one mcc-white Lambert
    m = mqo.loads(HEADER + "Material 0 {\n}\nEof\n")
    assert [b["type"] for b in m.bsdfs] == [LAMBERT] and m.bsdfs[0]["rgb"][1] > 0.8


def test_areas_and_ids(mqo):
    # a 2 m x 3 m quad (cm in the file) in two triangles; ids are file order
    V = [(0, 0, 0), (200, 0, 0), (200, 0, 300), (0, 0, 300)]
    faces = [(0, 1, 2, 0), (0, 2, 3, 1)]
    m = mqo.loads(document([material_line("a", spc=1.0, power=0.0),
                            material_line("l", emi=1.0)], [("q", 0, V, faces)]))
    assert m.info["surface_area"] == pytest.approx(6.0, rel=1e-6)
    assert m.info["light_area"] == pytest.approx(3.0, rel=1e-6)
    assert list(m.tri_bsdf) == [0, 1]
    assert np.allclose(m.v2[0], (2.0, 0.0, 3.0))


def test_errors(mqo, capi_mod):
    with pytest.raises(capi_mod.XyzError, match="not a Metasequoia"):
        mqo.loads("hello\n")
    with pytest.raises(capi_mod.XyzError, match="unsupported format"):
        mqo.loads("Metasequoia Document\nFormat Text Ver 1.1\n")
    mats = [material_line("w")]
    with pytest.raises(capi_mod.XyzError, match="vertex index"):
        mqo.loads(document(mats, [("o", 0, [(0, 0, 0)] * 3, [(0, 1, 7, 0)])]))
    with pytest.raises(capi_mod.XyzError, match="material index"):
        mqo.loads(document(mats, [("o", 0, [(0, 0, 0), (1, 0, 0), (0, 1, 0)], [(0, 1, 2, 3)])]))
    with pytest.raises(capi_mod.XyzError, match="cannot open"):
        mqo.load(os.path.join(GOLDEN, "missing.mqo"))


def test_shell_commands_and_save(tmp_path):
    from xyz_amd import shell
    assert shell.stream2time("1 h 30 min 5 s".split()) == 5405
    assert shell.stream2time("2 days 1 hour".split()) == 2 * 86400 + 3600
    assert shell.stream2time("10 parsecs".split()) == 0
    assert shell.time2string(5405) == "1 hour 30 minutes 5 seconds"
    assert shell.time2string(86400 + 1) == "1 day 1 second"
    assert shell.time2string(0) == " 0 s"
    sh = shell.Shell(mqo_dir=GOLDEN, out_dir=str(tmp_path / "out"), ren_dir=str(tmp_path),
                     log=lambda *a: None)
    script = ["rem a comment", "=begin", "set dim 1 1", "=end", "", "set eye 1 2 3",
              "set at 4 5 6", "set dim 320 240", "set it 30 s", "set rt 10 s", "set rep 0",
              "bogus"]
    assert sh.run(script) == "undefined"
    assert (sh.eye, sh.at, sh.width, sh.height) == ([1, 2, 3], [4, 5, 6], 320, 240)
    # rt below it lowers it; it above rt raises rt (shell.cpp:227-246)
    assert (sh.rendering_time, sh.interval_time) == (10, 10)
    sh.scene_name = "cornell_box"
    sh.save("again")
    text = (tmp_path / "again.ren").read_text()
    assert "set dim 320 240" in text and "set rt  10 seconds" in text and text.endswith("ren\n")
    sh2 = shell.Shell(log=lambda *a: None)
    sh2.run([l for l in text.splitlines() if not l.startswith(("load", "ren"))])
    assert (sh2.eye, sh2.width, sh2.rendering_time, sh2.algorithm) == ([1, 2, 3], 320, 10, 0)
    assert sh2.run(["quit", "set dim 2 2"]) == "quit" and sh2.width == 320


# McCamy, Marcus and Davidson, "A color-rendition chart", J. Appl. Photogr.
# Eng. 2 (1976) 95-99, table 1: CIE 1931 (x, y) of the 18 chromatic patches
# under illuminant C (white point 0.3101, 0.3163) and their Y.
_MCCAMY_1976 = {
    "mcc-dark-skin": (0.4002, 0.3504, 10.05), "mcc-light-skin": (0.3773, 0.3446, 35.82),
    "mcc-blue-sky": (0.2470, 0.2514, 19.33), "mcc-foliage": (0.3372, 0.4220, 13.29),
    "mcc-blue-flower": (0.2651, 0.2400, 24.27), "mcc-bluish-green": (0.2608, 0.3430, 43.06),
    "mcc-orange": (0.5060, 0.4070, 30.05), "mcc-purplish-blue": (0.2110, 0.1750, 11.97),
    "mcc-moderate-red": (0.4533, 0.3058, 19.77), "mcc-purple": (0.2845, 0.2020, 6.56),
    "mcc-yellow-green": (0.3800, 0.4887, 44.29), "mcc-orange-yellow": (0.4729, 0.4375, 43.06),
    "mcc-blue": (0.1866, 0.1285, 6.11), "mcc-green": (0.3046, 0.4782, 23.39),
    "mcc-red": (0.5385, 0.3129, 12.00), "mcc-yellow": (0.4480, 0.4703, 59.10),
    "mcc-magenta": (0.3635, 0.2325, 19.77), "mcc-cyan": (0.1958, 0.2519, 19.77),
}
# Munsell values of the neutral series (N 9.5, 8, 6.5, 5, 3.5, 2)
_NEUTRAL_V = {"mcc-white": 9.5, "mcc-neutral-80": 8.0, "mcc-neutral-65": 6.5,
              "mcc-neutral-50": 5.0, "mcc-neutral-35": 3.5, "mcc-black": 2.0}


def _munsell_y(v):
    """Luminance factor of Munsell value v (Newhall, Nickerson and Judd 1943)."""
    return 1.2219 * v - 0.23111 * v ** 2 + 0.23951 * v ** 3 - 0.021009 * v ** 4 + 0.0008404 * v ** 5


def test_spectral_tables_against_published_colorimetry(mqo, orc):
    """The restated spectral tables, checked against colorimetry that comes
    from neither implementation.  CIE 1931: the three matching functions have
    equal integrals (equal-energy white is x = y = 1/3), which holds on the
    400-700 nm window to 0.3 %, and y-bar peaks at 555 nm.  Macbeth chart,
    under equal energy: the neutral series' Y equals the Munsell renotation's
    luminance factor of its value within 2 (in %) with achromatic
    chromaticity; each chromatic patch's hue angle about the white point is
    within 25 degrees of the published one and its Y within 7 (the published
    figures are under illuminant C over 380-780 nm).  A swapped or misnamed
    row moves a hue by far more.  The native loader gives the oracle's XYZ for
    all 24 names (with the reference's first-substring lookup)."""
    cmf = np.array(orc._CMF, np.float64)
    s = cmf.sum(1)
    assert np.all(np.abs(s / s.mean() - 1) < 0.005)
    assert 15 <= int(np.argmax(cmf[1])) <= 16  # y-bar: 550 or 560 nm
    assert int(np.argmax(cmf[0])) == 20 and int(np.argmax(cmf[0][:10])) == 4  # x-bar: 600, 440
    assert int(np.argmax(cmf[2])) == 5  # z-bar: 450 nm
    xw, yw = s[0] / s.sum(), s[1] / s.sum()
    table = dict(orc.MACBETH)
    assert len(table) == 24 and set(table) == set(_MCCAMY_1976) | set(_NEUTRAL_V)
    for name, spd in table.items():
        X, Y, Z = (cmf * np.array(spd)).sum(1) / s[1] * 100
        x, y = X / (X + Y + Z), Y / (X + Y + Z)
        if name in _NEUTRAL_V:
            assert abs(Y - _munsell_y(_NEUTRAL_V[name])) < 2.0, name
            assert abs(x - xw) < 0.005 and abs(y - yw) < 0.005, name
        else:
            px, py, pY = _MCCAMY_1976[name]
            hue = math.degrees(math.atan2(y - yw, x - xw))
            phue = math.degrees(math.atan2(py - 0.3163, px - 0.3101))
            assert abs((hue - phue + 180) % 360 - 180) < 25, name
            assert abs(Y - pY) < 7, name
    names = [k for k, _ in orc.MACBETH]
    m = mqo.loads(document([material_line(k) for k in names], []))
    for i, k in enumerate(names):
        want = np.array(orc.material_xyz(k, (1, 1, 1, 1), 0.0), np.float32)
        got = np.array(m.bsdfs[i]["rgb"], np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k


def test_mercury_lamp_table_peaks_at_the_mercury_lines(orc):
    """The mercury arc lamp's spectrum (the reference's light power) has its
    prominent peaks in the 10 nm bins of mercury's visible emission lines,
    404.7, 435.8, 546.1 and 577.0/579.1 nm: 410, 440, 550 and 580 nm."""
    v = np.array(orc.MERCURY)
    assert len(v) == 31
    peaks = [400 + 10 * i for i in range(1, 30)
             if v[i] > v[i - 1] and v[i] > v[i + 1] and v[i] > 1.5 * np.median(v)]
    assert peaks == [410, 440, 550, 580]
