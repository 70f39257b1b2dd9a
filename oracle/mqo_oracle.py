"""CPU restatement of the reference's Metasequoia loader, Scene::Load
(xyz/core/scene_mqo.cpp:9-777) -- TEST INFRASTRUCTURE ONLY.

Only tests/ may use this module: it is the checker of the library's native
loader (xyz_amd/csrc/mqo_loader.cpp, reached through xyz_mqo_load), never part
of the product path.  Written independently of the native loader, in Python
over a small model of the C++ stream operations the reference uses (token
reads, getline to a delimiter, eof), with binary32 arithmetic through numpy
float32 scalars.

What pins it (tests/test_mqo.py): closed forms that do not come from this
code -- the default Scene chunk gives the reference camera's default eye
(0, 1, 4) (xyz/core/thinlenscamera.cpp:21-28); smooth normals weigh each
corner by its interior angle / pi (an octahedron's are its vertex directions, a
triangulated cube's its diagonals, a 90 + 45 degree fan's a 2:1 blend); white
through the CCIR 601-1 matrix gives the
matrix row sums; the material-flag table of scene_mqo.cpp:126-235; areas of
axis-aligned quads.  The spectral tables (Macbeth patches, CIE 1931 colour
matching functions, mercury arc lamp) are data: only their use (spd2xyz as a
float dot product, linear in the spectrum) is pinned, the values are "parity
unpinned" beyond the published tables they restate."""
import math
import re

import numpy as np

F = np.float32
M_1_PI = 0.318309886183790671538

# spectral data, 400-700 nm every 10 nm (libhi/hi/image/spectrum.cpp:270-533)
_CMF = [
    [0.014, 0.044, 0.134, 0.284, 0.348, 0.336, 0.291, 0.195, 0.096, 0.032, 0.005, 0.009,
     0.063, 0.166, 0.290, 0.433, 0.595, 0.762, 0.916, 1.026, 1.062, 1.003, 0.854, 0.642,
     0.448, 0.284, 0.165, 0.087, 0.047, 0.023, 0.011],
    [0.000, 0.001, 0.004, 0.012, 0.023, 0.038, 0.060, 0.091, 0.139, 0.208, 0.323, 0.503,
     0.710, 0.862, 0.954, 0.995, 0.995, 0.952, 0.870, 0.757, 0.631, 0.503, 0.381, 0.265,
     0.175, 0.107, 0.061, 0.032, 0.017, 0.008, 0.004],
    [0.068, 0.207, 0.646, 1.386, 1.747, 1.772, 1.669, 1.288, 0.813, 0.465, 0.272, 0.158,
     0.078, 0.042, 0.020, 0.009, 0.004, 0.002, 0.002, 0.001, 0.001, 0.000, 0.000, 0.000,
     0.000, 0.000, 0.000, 0.000, 0.000, 0.000, 0.000],
]
CMF = [[F(v / 9.80328) for v in row] for row in _CMF]
MERCURY = [46.2, 59.5, 46.2, 45.7, 71.9, 40.9, 30.2, 26.6, 11.7, 26.6, 22.2, 18.7, 19.5,
           23.9, 51.5, 72.8, 51.5, 45.3, 58.6, 47.1, 19.8, 24.9, 23.9, 23.1, 23.2, 23.3,
           23.1, 21.2, 17.7, 21.3, 25.9]
MACBETH = [  # matching order of scene_mqo.cpp:142-209
    ("mcc-dark-skin", "0.05 0.05 0.04 0.04 0.04 0.04 0.04 0.04 0.04 0.05 0.05 0.05 0.06 0.06 0.06 0.07 0.08 0.10 0.11 0.13 0.13 0.14 0.15 0.16 0.18 0.19 0.21 0.23 0.25 0.27 0.29"),
    ("mcc-light-skin", "0.21 0.21 0.22 0.23 0.24 0.26 0.29 0.31 0.33 0.34 0.35 0.34 0.31 0.30 0.31 0.31 0.32 0.38 0.47 0.53 0.57 0.60 0.62 0.64 0.66 0.68 0.70 0.73 0.77 0.80 0.83"),
    ("mcc-blue-sky", "0.33 0.35 0.35 0.35 0.34 0.34 0.32 0.30 0.29 0.27 0.25 0.23 0.22 0.20 0.20 0.18 0.17 0.17 0.16 0.15 0.15 0.14 0.13 0.12 0.12 0.11 0.11 0.11 0.10 0.10 0.10"),
    ("mcc-foliage", "0.03 0.03 0.03 0.03 0.03 0.04 0.04 0.04 0.05 0.05 0.08 0.13 0.17 0.17 0.14 0.12 0.11 0.10 0.09 0.08 0.08 0.08 0.08 0.08 0.08 0.08 0.08 0.10 0.13 0.16 0.18"),
    ("mcc-blue-flower", "0.43 0.46 0.46 0.46 0.45 0.44 0.42 0.40 0.38 0.36 0.33 0.28 0.24 0.22 0.22 0.21 0.20 0.21 0.23 0.24 0.25 0.25 0.25 0.27 0.33 0.40 0.47 0.52 0.55 0.56 0.57"),
    ("mcc-bluish-green", "0.32 0.34 0.35 0.37 0.38 0.42 0.48 0.54 0.59 0.61 0.60 0.58 0.57 0.53 0.49 0.44 0.40 0.36 0.31 0.27 0.24 0.23 0.23 0.22 0.21 0.22 0.22 0.24 0.25 0.26 0.26"),
    ("mcc-orange", "0.05 0.04 0.05 0.04 0.04 0.05 0.05 0.05 0.05 0.06 0.07 0.09 0.12 0.18 0.23 0.30 0.39 0.46 0.52 0.55 0.56 0.58 0.59 0.60 0.61 0.61 0.61 0.62 0.62 0.63 0.64"),
    ("mcc-purplish-blue", "0.30 0.33 0.35 0.37 0.38 0.37 0.34 0.30 0.24 0.20 0.16 0.14 0.12 0.11 0.10 0.09 0.08 0.08 0.08 0.08 0.08 0.08 0.08 0.09 0.10 0.11 0.11 0.11 0.10 0.10 0.11"),
    ("mcc-moderate-red", "0.15 0.15 0.14 0.14 0.13 0.13 0.13 0.13 0.11 0.11 0.10 0.10 0.09 0.10 0.10 0.11 0.12 0.17 0.29 0.42 0.51 0.56 0.59 0.60 0.60 0.60 0.60 0.60 0.60 0.61 0.61"),
    ("mcc-purple", "0.20 0.19 0.18 0.16 0.13 0.11 0.09 0.08 0.07 0.06 0.05 0.05 0.05 0.04 0.04 0.04 0.04 0.04 0.04 0.05 0.06 0.08 0.11 0.14 0.17 0.19 0.23 0.26 0.31 0.36 0.40"),
    ("mcc-yellow-green", "0.05 0.05 0.05 0.06 0.07 0.08 0.10 0.13 0.19 0.27 0.38 0.47 0.53 0.54 0.52 0.50 0.48 0.45 0.41 0.37 0.35 0.34 0.33 0.33 0.32 0.32 0.33 0.35 0.35 0.37 0.38"),
    ("mcc-orange-yellow", "0.05 0.05 0.05 0.05 0.05 0.06 0.07 0.08 0.10 0.12 0.14 0.18 0.26 0.35 0.44 0.50 0.55 0.59 0.61 0.62 0.62 0.63 0.64 0.65 0.65 0.66 0.66 0.66 0.67 0.68 0.68"),
    ("mcc-blue", "0.20 0.23 0.26 0.29 0.32 0.32 0.28 0.22 0.16 0.12 0.09 0.07 0.06 0.05 0.05 0.04 0.04 0.04 0.04 0.04 0.04 0.04 0.04 0.04 0.04 0.05 0.05 0.05 0.05 0.05 0.06"),
    ("mcc-green", "0.06 0.06 0.06 0.07 0.07 0.08 0.10 0.12 0.15 0.18 0.23 0.31 0.36 0.36 0.33 0.29 0.25 0.21 0.17 0.13 0.10 0.09 0.08 0.08 0.07 0.07 0.07 0.08 0.09 0.09 0.09"),
    ("mcc-red", "0.05 0.05 0.05 0.05 0.05 0.05 0.05 0.04 0.04 0.04 0.04 0.04 0.04 0.04 0.04 0.05 0.05 0.07 0.11 0.20 0.33 0.49 0.61 0.67 0.71 0.72 0.73 0.74 0.75 0.76 0.77"),
    ("mcc-yellow", "0.04 0.04 0.04 0.04 0.05 0.05 0.06 0.10 0.16 0.26 0.37 0.47 0.56 0.61 0.64 0.67 0.70 0.73 0.75 0.75 0.75 0.76 0.77 0.78 0.78 0.78 0.79 0.79 0.79 0.80 0.80"),
    ("mcc-magenta", "0.36 0.37 0.36 0.34 0.31 0.28 0.25 0.22 0.20 0.18 0.15 0.13 0.11 0.10 0.11 0.10 0.11 0.14 0.22 0.31 0.41 0.52 0.60 0.67 0.70 0.72 0.73 0.74 0.75 0.76 0.77"),
    ("mcc-cyan", "0.26 0.27 0.30 0.33 0.35 0.39 0.43 0.44 0.45 0.43 0.39 0.34 0.29 0.24 0.19 0.15 0.12 0.10 0.09 0.08 0.07 0.07 0.07 0.07 0.07 0.07 0.08 0.07 0.07 0.07 0.07"),
    ("mcc-white", "0.67 0.82 0.87 0.88 0.88 0.90 0.90 0.90 0.91 0.92 0.91 0.91 0.92 0.92 0.91 0.91 0.92 0.93 0.93 0.92 0.91 0.92 0.92 0.92 0.92 0.92 0.91 0.92 0.92 0.93 0.92"),
    ("mcc-neutral-80", "0.53 0.58 0.60 0.60 0.60 0.60 0.60 0.60 0.60 0.60 0.60 0.60 0.61 0.61 0.60 0.60 0.61 0.61 0.61 0.61 0.60 0.60 0.60 0.60 0.60 0.59 0.59 0.59 0.59 0.59 0.59"),
    ("mcc-neutral-65", "0.34 0.35 0.36 0.36 0.36 0.36 0.36 0.36 0.36 0.36 0.36 0.36 0.37 0.37 0.36 0.36 0.37 0.37 0.37 0.36 0.36 0.36 0.36 0.36 0.36 0.35 0.35 0.35 0.35 0.35 0.35"),
    ("mcc-neutral-50", "0.19 0.19 0.19 0.20 0.20 0.20 0.20 0.19 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.20 0.19 0.19 0.20 0.20 0.19"),
    ("mcc-neutral-35", "0.08 0.08 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09 0.09"),
    ("mcc-black", " ".join(["0.03"] * 31)),
]
MACBETH = [(k, [float(x) for x in v.split()]) for k, v in MACBETH]

LAMBERT, MIRROR, GLASS, METAL, LIGHT = 0, 1, 2, 3, 4


def spd2xyz(spd):
    """hi::spd2xyz: per channel, sum += a[i] * b[i] in binary32 (spectrum.cpp:478-485)."""
    out = []
    for k in range(3):
        s = F(0)
        for i in range(31):
            s = F(s + F(F(spd[i]) * CMF[k][i]))
        out.append(s)
    return out


def clamp01(v):
    return [F(0) if x < 0 else (F(1) if x > 1 else x) for x in v]


def ccir601_rgb2xyz(r, g, b):
    """libhi/hi/image/color_space.hpp:58-68 in binary32, left to right."""
    def row(a, b_, c):
        return F(F(F(F(a) * r) + F(F(b_) * g)) + F(F(c) * b))
    return [row(0.6069, 0.1735, 0.2003), row(0.2989, 0.5866, 0.1145), row(0.0000, 0.0661, 1.1162)]


# --------------------------------------------------------------- C++ streams
class Stream:
    """The subset of std::istream the reference's loader uses."""

    def __init__(self, text):
        self.s = text
        self.p = 0
        self.eof = False
        self.fail = False

    def getline(self, delim="\n", text_mode=True):
        if self.p >= len(self.s):
            self.eof = self.fail = True
            return ""
        j = self.s.find(delim, self.p)
        if j < 0:
            out = self.s[self.p:]
            self.p = len(self.s)
            self.eof = True
        else:
            out = self.s[self.p:j]
            self.p = j + 1
        if text_mode and delim == "\n" and out.endswith("\r"):
            out = out[:-1]
        return out

    def _skip_ws(self):
        while self.p < len(self.s) and self.s[self.p] in " \t\r\n\v\f":
            self.p += 1

    def token(self):
        self._skip_ws()
        if self.p >= len(self.s):
            self.eof = self.fail = True
            return None
        j = self.p
        while j < len(self.s) and self.s[j] not in " \t\r\n\v\f":
            j += 1
        out = self.s[self.p:j]
        self.p = j
        if j >= len(self.s):
            self.eof = True
        return out

    _num = {"f": re.compile(r"[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)"),
            "i": re.compile(r"[+-]?\d+")}

    def number(self, kind, default):
        """operator>> for float ("f") or integer ("i"); on failure the target keeps
        `default` here (the reference's targets are initialised)."""
        if self.fail:
            return default
        self._skip_ws()
        m = self._num[kind].match(self.s, self.p)
        if not m:
            self.fail = True
            if self.p >= len(self.s):
                self.eof = True
            return default
        self.p = m.end()
        if self.p >= len(self.s):
            self.eof = True
        return F(float(m.group(0))) if kind == "f" else int(m.group(0))


def trim(s):
    return s.strip(" \t\r\n")


def getstring(st):
    st.getline('"', False)
    return st.getline('"', False)


def skip_block(st):
    indent = 0
    while True:
        line = trim(st.getline())
        if st.fail and not line:
            return False
        if line.endswith("{"):
            indent += 1
        elif line.endswith("}"):
            indent -= 1
        if indent <= 0:
            return True


# --------------------------------------------------------------- loader
def material_xyz(name, col, spc):
    for key, spd in MACBETH:
        if key in name:
            return clamp01(spd2xyz(spd))
    # powf(x, 2.2f): evaluated in double and rounded once to binary32
    rgb = [F(math.pow(float(F(F(col[k]) * F(spc))), float(F(2.2)))) for k in range(3)]
    return clamp01(ccir601_rgb2xyz(*rgb))


def load_materials(st, count):
    out = []
    for _ in range(count):
        line = st.getline()
        lin = Stream(line)
        name = getstring(lin)
        col = [F(1)] * 4
        emi, spc, power = F(0), F(0), F(5)
        while not lin.eof:
            tag = trim(lin.getline("(", False))
            if not tag:
                break
            if tag in ("shader", "dif", "amb", "emi", "spc", "power", "col"):
                d = Stream(lin.getline(")", False))
                if tag == "col":
                    col = [d.number("f", col[k]) for k in range(4)]
                elif tag == "emi":
                    emi = d.number("f", emi)
                elif tag == "spc":
                    spc = d.number("f", spc)
                elif tag == "power":
                    power = d.number("f", power)
            elif tag in ("tex", "aplane", "bump"):
                getstring(lin)
                st.getline(")", False)
        if emi == 1:
            out.append(dict(type=LIGHT, rgb=[F(0)] * 3, shininess=F(0)))
        elif col[3] == 1:
            if spc == 1 and power > 0:
                if power == 100:
                    out.append(dict(type=MIRROR, rgb=[F(0)] * 3, shininess=F(0)))
                else:
                    out.append(dict(type=METAL, rgb=material_xyz(name, col, spc), shininess=power))
            else:
                out.append(dict(type=LAMBERT, rgb=material_xyz(name, col, spc), shininess=F(0)))
        else:
            out.append(dict(type=GLASS, rgb=[F(0)] * 3, shininess=F(0)))
    return out


def v3(*a):
    return np.array(a, F)


def dot(a, b):
    return F(F(F(a[0] * b[0]) + F(a[1] * b[1])) + F(a[2] * b[2]))


def cross(a, b):
    return v3(F(a[1] * b[2]) - F(a[2] * b[1]), F(a[2] * b[0]) - F(a[0] * b[2]),
              F(a[0] * b[1]) - F(a[1] * b[0]))


def normalize(a):
    with np.errstate(divide="ignore", invalid="ignore"):
        r = F(F(1) / F(np.sqrt(dot(a, a))))
        return (a * r).astype(F)


def length(a):
    return F(np.sqrt(dot(a, a)))


def load_object(st, vertices, scene):
    vn = len(vertices)
    vertex_count = triangle_count = 0
    visible, shading = 0, 1
    faces = []
    while True:
        tok = st.token()
        if tok is None:
            raise ValueError("object chunk not closed")
        if tok == "visible":
            visible = st.number("i", visible)
        elif tok == "shading":
            shading = st.number("i", shading)
        elif tok in ("vertex", "face"):
            if not visible:
                skip_block(st)
                continue
            count = max(0, st.number("i", 0))
            st.getline()
            for _ in range(count):
                ln = Stream(st.getline())
                if tok == "vertex":
                    p = v3(*[ln.number("f", F(0)) for _ in range(3)])
                    vertices.append((p * F(0.01)).astype(F))
                else:
                    ln = Stream(ln.s.replace("(", " ").replace(")", " "))
                    f = dict(v=[0, 0, 0], m=0, uv=[F(0), F(0), F(1), F(0), F(0), F(1)])
                    if ln.number("i", 0) == 3:
                        for _k in range(3):
                            t = ln.token()
                            if t == "V":
                                f["v"] = [ln.number("i", 0) for _ in range(3)]
                            elif t == "M":
                                f["m"] = ln.number("i", 0)
                            elif t == "UV":
                                f["uv"] = [ln.number("f", f["uv"][q]) for q in range(6)]
                    faces.append(f)
            st.getline()
            if tok == "vertex":
                vertex_count = count
            else:
                triangle_count = count
        elif tok == "}":
            break
        else:
            skip_block(st)
    scene["n_objects"] += 1
    if not visible:
        return
    V = vertices
    normals = None
    if shading:
        normals = [v3(0, 0, 0) for _ in range(vertex_count)]
        for f in faces:
            i0, i1, i2 = f["v"]
            p0, p1, p2 = V[vn + i0], V[vn + i1], V[vn + i2]
            ea, eb, ec = (p1 - p2).astype(F), (p2 - p0).astype(F), (p0 - p1).astype(F)
            a, b, c = length(ea), length(eb), length(ec)

            def w(x):
                with np.errstate(divide="ignore", invalid="ignore"):
                    val = 1.0 - M_1_PI * float(F(math.acos(float(F(x))))) if abs(float(x)) <= 1 else float("nan")
                return F(val if 0.0 < val else 0.0)
            with np.errstate(divide="ignore", invalid="ignore"):
                aa = w(F(dot(eb, ec) / F(b * c)))
                ab = w(F(dot(ec, ea) / F(c * a)))
                ac = w(F(dot(ea, eb) / F(a * b)))
            n = normalize(cross(-eb, ec))
            for idx, wt in ((i0, aa), (i1, ab), (i2, ac)):
                normals[idx] = (normals[idx] + (n * wt).astype(F)).astype(F)
        normals = [normalize(n) for n in normals]
    area_sum = F(0)
    for f in faces:
        i0, i1, i2 = f["v"]
        p = [V[vn + i0], V[vn + i1], V[vn + i2]]
        e1, e2 = (p[1] - p[0]).astype(F), (p[2] - p[0]).astype(F)
        if shading:
            nn = [normals[i0], normals[i1], normals[i2]]
        else:
            g = normalize(cross(e2, e1))
            nn = [g, g, g]
        scene["v"].append(p)
        scene["n"].append(nn)
        scene["uv"].append(f["uv"])
        scene["bsdf"].append(f["m"])
        area = F(length(cross(e2, e1)) * F(0.5))
        area_sum = F(area_sum + area)
        if scene["bsdfs"][f["m"]]["type"] == LIGHT:
            scene["n_lights"] += 1
            scene["light_area"] = F(scene["light_area"] + area)
    scene["surface_area"] = F(scene["surface_area"] + area_sum)


def load_camera(st):
    pos, look = [F(0), F(0), F(400)], [F(0), F(100), F(0)]
    head = pitch = F(0)
    indent = 0
    while True:
        line = trim(st.getline())
        if line.endswith("{"):
            indent += 1
        elif line.endswith("}"):
            indent -= 1
        elif line:
            d = Stream(line)
            t = d.token()
            if t == "pos":
                pos = [d.number("f", pos[k]) for k in range(3)]
            elif t == "lookat":
                look = [d.number("f", look[k]) for k in range(3)]
            elif t == "head":
                head = d.number("f", head)
            elif t == "pich":
                pitch = d.number("f", pitch)
        if indent <= 0:
            break
    s, c = F(np.sin(pitch)), F(np.cos(pitch))
    v = list(pos)
    pos[1] = F(F(c * v[1]) + F(s * v[2]))
    pos[2] = F(F(s * v[1]) - F(c * v[2]))
    head = F(float(head) - math.pi)
    s, c = F(np.sin(head)), F(np.cos(head))
    v = list(pos)
    pos[2] = F(F(c * v[2]) + F(s * v[0]))
    pos[0] = F(F(s * v[2]) - F(c * v[0]))
    pos = [F(x * F(0.01)) for x in pos]
    look = [F(x * F(0.01)) for x in look]
    return [F(look[k] + pos[k]) for k in range(3)], look


def load(text):
    """Returns a dict: v (n,3,3), n (n,3,3), uv (n,6), bsdf (n,), bsdfs, eye, at,
    has_camera, n_objects, n_lights, surface_area, light_area, light_power."""
    st = Stream(text)
    if st.getline() != "Metasequoia Document" or st.getline() != "Format Text Ver 1.0":
        raise ValueError("not a Metasequoia text document")
    scene = dict(v=[], n=[], uv=[], bsdf=[], bsdfs=[], eye=[F(0), F(1), F(4)],
                 at=[F(0), F(1), F(0)], has_camera=0, n_objects=0, n_lights=0,
                 surface_area=F(0), light_area=F(0), light_power=spd2xyz(MERCURY))
    vertices = []
    while True:
        tag = st.token()
        if tag is None:
            break
        if tag == "Scene":
            scene["eye"], scene["at"] = load_camera(st)
            scene["has_camera"] = 1
        if tag == "Material":
            count = st.number("i", 0)
            if count > 0:
                st.getline()
                scene["bsdfs"] += load_materials(st, count)
                st.getline()
            else:
                scene["bsdfs"].append(dict(type=LAMBERT, rgb=clamp01(spd2xyz(MACBETH[18][1])),
                                           shininess=F(0)))
                skip_block(st)
        elif tag == "Object":
            st.getline()
            load_object(st, vertices, scene)
            st.getline()
        elif tag == "Eof":
            break
    if scene["light_area"] > 0:
        for b in scene["bsdfs"]:
            if b["type"] == LIGHT:
                b["rgb"] = [F(float(p) / (math.pi * float(scene["light_area"])))
                            for p in scene["light_power"]]
    n = len(scene["bsdf"])
    scene["v"] = np.array(scene["v"], F).reshape(n, 3, 3)
    scene["n"] = np.array(scene["n"], F).reshape(n, 3, 3)
    scene["uv"] = np.array(scene["uv"], F).reshape(n, 6)
    scene["bsdf"] = np.array(scene["bsdf"], np.uint32)
    return scene
