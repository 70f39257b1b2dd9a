"""Writes tests/golden/cornell_box.mqo and cornell_box.ren: a Cornell box in
Metasequoia text format (centimetres) that exercises the loader's material
mapping (spectral and gamma colours, light, mirror, glass, metal), smooth and
flat objects, an invisible object, a quad face, a texture reference that
swallows the following material, and the Scene chunk's camera.

Run from the repository root: python tests/golden/make_mqo_fixtures.py"""
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))

MATERIALS = [
    '"mcc-white wall" shader(3) col(1.000 1.000 1.000 1.000) dif(0.800) amb(0.600) emi(0.000) spc(0.000) power(5.00)',
    '"red" shader(3) col(0.900 0.200 0.200 1.000) dif(0.800) amb(0.600) emi(0.000) spc(1.000) power(0.00)',
    '"green mcc-green" shader(3) col(0.200 0.900 0.200 1.000) dif(0.800) amb(0.600) emi(0.000) spc(0.000) power(5.00)',
    '"lamp" shader(3) col(1.000 1.000 1.000 1.000) dif(0.800) amb(0.600) emi(1.000) spc(0.000) power(5.00)',
    '"mirror" shader(3) col(1.000 1.000 1.000 1.000) dif(0.800) amb(0.600) emi(0.000) spc(1.000) power(100.00)',
    '"glass" shader(3) col(1.000 1.000 1.000 0.300) dif(0.800) amb(0.600) emi(0.000) spc(0.000) power(5.00)',
    '"gold" shader(3) col(1.000 0.800 0.300 1.000) dif(0.800) amb(0.600) emi(0.000) spc(1.000) power(30.00)',
    '"mcc-orange-yellow" shader(3) col(1.000 1.000 1.000 1.000) dif(0.800) amb(0.600) emi(0.000) spc(0.000) power(5.00)',
    '"wood" shader(3) col(0.600 0.400 0.200 1.000) dif(0.800) amb(0.600) emi(0.000) spc(1.000) power(0.00) tex("wood.png")',
    '"swallowed" shader(3) col(0.100 0.900 0.100 1.000) dif(0.800) amb(0.600) emi(0.000) spc(1.000) power(0.00)',
]
WHITE, RED, GREEN, LAMP, MIRROR, GLASS, GOLD, ORANGE, WOOD, SWALLOWED = range(10)


def fmt(v):
    return " ".join(f"{x:.4f}" for x in v)


class Obj:
    def __init__(self, name, shading=0, visible=15):
        self.name, self.shading, self.visible = name, shading, visible
        self.v, self.f = [], []

    def vert(self, p):
        self.v.append(p)
        return len(self.v) - 1

    def tri(self, a, b, c, m, uv=(0, 0, 1, 0, 0, 1)):
        self.f.append(f"3 V({a} {b} {c}) M({m}) UV({fmt(uv)})")

    def quad(self, a, b, c, d, m):
        # (a, b, c) and (a, c, d)
        self.tri(a, b, c, m)
        self.tri(a, c, d, m, (0, 0, 1, 1, 0, 1))

    def text(self):
        out = [f'Object "{self.name}" {{', "\tdepth 0", "\tfolding 0",
               "\tscale 1.000000 1.000000 1.000000", "\trotation 0.000000 0.000000 0.000000",
               "\ttranslation 0.000000 0.000000 0.000000", f"\tvisible {self.visible}",
               "\tlocking 0", f"\tshading {self.shading}", "\tfacet 59.5",
               "\tcolor 0.898 0.498 0.698", "\tcolor_type 0", f"\tvertex {len(self.v)} {{"]
        out += [f"\t\t{fmt(p)}" for p in self.v]
        out += ["\t}", f"\tface {len(self.f)} {{"]
        out += [f"\t\t{f}" for f in self.f]
        out += ["\t}", "}"]
        return out


def box(obj, center, size, angle_deg, m):
    """Axis box rotated about y; faces outward (normal = cross(e2, e1))."""
    c, s = math.cos(math.radians(angle_deg)), math.sin(math.radians(angle_deg))
    idx = {}
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                x, y, z = ((i - 0.5) * size[0], j * size[1], (k - 0.5) * size[2])
                idx[i, j, k] = obj.vert((center[0] + c * x + s * z, center[1] + y,
                                         center[2] - s * x + c * z))
    V = idx
    # listed so that (a, b, c) has cross(e2, e1) pointing into the box; the
    # quads are emitted reversed, facing out
    faces = [((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)),  # bottom (y-)
             ((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)),  # top (y+)
             ((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)),  # x-
             ((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)),  # x+
             ((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)),  # z-
             ((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))]  # z+
    for a, b, cc, d in faces:
        obj.quad(V[a], V[d], V[cc], V[b], m)


def octahedron(obj, center, r, m):
    P = [(r, 0, 0), (-r, 0, 0), (0, r, 0), (0, -r, 0), (0, 0, r), (0, 0, -r)]
    ids = [obj.vert((center[0] + p[0], center[1] + p[1], center[2] + p[2])) for p in P]
    for sx in (0, 1):
        for sy in (2, 3):
            for sz in (4, 5):
                a, b, c = ids[sx], ids[sy], ids[sz]
                # outward normal: cross(e2, e1) must point along (sign x, sign y, sign z)
                odd = (sx == 1) ^ (sy == 3) ^ (sz == 5)
                if odd:
                    obj.tri(a, b, c, m)
                else:
                    obj.tri(a, c, b, m)


def cornell():
    room = Obj("room")
    P = [(0, 0, 0), (100, 0, 0), (100, 0, 100), (0, 0, 100),
         (0, 100, 0), (100, 100, 0), (100, 100, 100), (0, 100, 100)]
    v = [room.vert(p) for p in P]
    # inward-facing walls: floor +y, ceiling -y, back -z, left +x, right -x
    room.quad(v[0], v[1], v[2], v[3], WHITE)      # floor: cross(e2, e1) = +y
    room.quad(v[4], v[7], v[6], v[5], WHITE)      # ceiling
    room.quad(v[3], v[2], v[6], v[7], ORANGE)     # back
    room.quad(v[0], v[3], v[7], v[4], RED)        # left x = 0
    room.quad(v[1], v[5], v[6], v[2], GREEN)      # right x = 100
    room.f.append(f"4 V({v[0]} {v[1]} {v[2]} {v[3]}) M({WHITE})")  # a quad face
    light = Obj("light")
    L = [light.vert(p) for p in ((37.5, 99.9, 37.5), (62.5, 99.9, 37.5),
                                 (62.5, 99.9, 62.5), (37.5, 99.9, 62.5))]
    light.quad(L[0], L[3], L[2], L[1], LAMP)      # faces -y
    boxes = Obj("boxes")
    box(boxes, (33, 0, 35), (30, 30, 30), -18, WOOD)
    box(boxes, (67, 0, 67), (30, 60, 30), 18, SWALLOWED)
    ball = Obj("ball", shading=1)
    octahedron(ball, (30, 45, 30), 12, MIRROR)
    gem = Obj("gem", shading=1)
    octahedron(gem, (70, 75, 45), 8, GLASS)
    cube = Obj("smooth cube", shading=1)
    box(cube, (50, 10, 15), (12, 12, 12), 0, GOLD)
    hidden = Obj("hidden", visible=0)
    box(hidden, (50, 50, 50), (90, 90, 90), 0, WHITE)
    lines = ["Metasequoia Document", "Format Text Ver 1.0", "",
             "Scene {", "\tpos 0.0000 0.0000 -150.0000", "\tlookat 50.0000 50.0000 50.0000",
             "\thead 0.0000", "\tpich 0.0000", "\tortho 0", "\tzoom2 5.0000",
             "\tamb 0.250 0.250 0.250", "}",
             f"Material {len(MATERIALS)} {{"] + [f"\t{m}" for m in MATERIALS] + ["}"]
    for o in (room, light, boxes, ball, gem, cube, hidden):
        lines += o.text()
    lines.append("Eof")
    return "\r\n".join(lines) + "\r\n"


REN = """rem Cornell box of cornell_box.mqo through the path tracer
=begin
Renders for two seconds with a snapshot every second.
=end
load cornell_box

set eye 0.5 0.5 -1.4
set at  0.5 0.5 0.5

set dim 64 48
set it  1 s
set rt  2 s

set rep 0
ren
"""

if __name__ == "__main__":
    with open(os.path.join(HERE, "cornell_box.mqo"), "w", newline="") as f:
        f.write(cornell())
    with open(os.path.join(HERE, "cornell_box.ren"), "w") as f:
        f.write(REN)
    print("wrote cornell_box.mqo, cornell_box.ren")
