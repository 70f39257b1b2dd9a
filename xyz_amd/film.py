"""Film files of the reference renderer's ImageFilm (xyz/core/imagefilm.cpp:108-241,
xyz/core/imagefilm_hdr.cpp), from a film of per-pixel sums.

A film is float32 (H, W, 3) with row 0 = bottom, as xyz_render fills it; every
writer first scales by 1 / spp.  Like ImageFilm, the .float and .hdr writers
treat the film as CIE XYZ and convert it to RGB with the CCIR 601-1 matrix
(libhi/hi/image/color_space.hpp:70-79); the PFM writers write values as they
are.  The arithmetic is float32, in the reference's operation order, so the
files are byte-identical to include/xyz_hip.hpp's writers."""
import numpy as np

F = np.float32


def _scaled(film, spp):
    film = np.asarray(film, F).reshape(-1, 3)
    return film * (F(1.0) / F(spp))


def ccir601_xyz_to_rgb(xyz):
    """hi::CCIR601_1_XYZ2RGB in float32: rows of (x, y, z) -> (r, g, b)."""
    x, y, z = (np.asarray(xyz, F)[..., k] for k in range(3))
    r = (F(1.9104) * x - F(0.5338) * y) - F(0.2891) * z
    g = (F(-0.9844) * x + F(1.9985) * y) - F(0.0279) * z
    b = (F(0.0585) * x - F(0.1187) * y) + F(0.9017) * z
    return np.stack([r, g, b], -1).astype(F)


def save_pfm(path, film, spp):
    """ImageFilm::SaveAsPfm: rows bottom up, non-finite pixels as 0."""
    h, w = np.asarray(film).shape[:2]
    v = _scaled(film, spp)
    v[~np.isfinite(v).all(1)] = 0
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1\n".encode())
        f.write(v.astype("<f4").tobytes())


def save_pfm_rgb(path, film, spp):
    """ImageFilm::SaveAsPfmRgb: the PFM layout, values unchecked."""
    h, w = np.asarray(film).shape[:2]
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1\n".encode())
        f.write(_scaled(film, spp).astype("<f4").tobytes())


def save_float(path, film, spp):
    """ImageFilm::SaveAsFloat: headerless RGB float32, rows bottom up,
    negative components clamped to 0."""
    rgb = ccir601_xyz_to_rgb(_scaled(film, spp))
    rgb[rgb < 0] = 0
    with open(path, "wb") as f:
        f.write(rgb.astype("<f4").tobytes())


def rgbe_encode(rgb):
    """RGBE pixels (imagefilm_hdr.cpp:12-28): the largest component d gives
    the exponent; mantissas are rgb * float(frexp(d) * 255.9999 / d), cast to
    int and wrapped to a byte (a negative component is undefined in the
    reference's cast)."""
    rgb = np.asarray(rgb, F).reshape(-1, 3)
    d = np.maximum(np.maximum(rgb[:, 0], rgb[:, 1]), rgb[:, 2])
    out = np.zeros((len(rgb), 4), np.uint8)
    ok = d.astype(np.float64) > 1e-32
    m, e = np.frexp(d[ok])
    f = (m.astype(np.float64) * 255.9999 / d[ok].astype(np.float64)).astype(F)
    q = (rgb[ok] * f[:, None]).astype(np.float32)
    out[ok, :3] = (np.trunc(q).astype(np.int64) & 0xFF).astype(np.uint8)
    out[ok, 3] = (e + 128).astype(np.uint8)
    return out


def rgbe_line(px):
    """One scanline (4 bytes per pixel), run-length encoded as
    imagefilm_hdr.cpp:44-109 does."""
    size = len(px)
    if size < 8 or size > 0x7FFF:
        return px.tobytes()
    out = bytearray([2, 2, (size >> 8) & 0xFF, size & 0xFF])
    for c in range(4):
        a = px[:, c].tolist()
        j = 0
        count = 0
        while j < size:
            begin = j
            while begin < size:
                count = 1
                while count < 127 and begin + count < size and a[begin + count] == a[begin]:
                    count += 1
                if count >= 4:
                    break
                begin += count
            if 1 < begin - j < 4:
                k = j + 1
                while True:
                    same = a[k] == a[j]
                    k += 1
                    if not same:
                        break
                    if k == begin:
                        out += bytes([begin - j + 128, a[j]])
                        j = begin
                        break
            while j < begin:
                n = min(begin - j, 128)
                out.append(n)
                out += bytes(a[j:j + n])
                j += n
            if count >= 4:
                out += bytes([count + 128, a[begin]])
            else:
                count = 0
            j += count
    return bytes(out)


def save_hdr(path, film, spp):
    """ImageFilm::SaveAsHdr: Radiance RGBE, "-Y H +X W", top row first,
    XYZ -> RGB, non-finite pixels as 0."""
    h, w = np.asarray(film).shape[:2]
    rgb = ccir601_xyz_to_rgb(_scaled(film, spp)).reshape(h, w, 3)
    rgb[~np.isfinite(rgb).all(-1)] = 0
    out = bytearray(b"#?RADIANCE\n# Made with Mikado Renderer\nFORMAT=32-bit_rle_rgbe\n"
                    b"EXPOSURE=1.0000000000000\n\n")
    out += f"-Y {h} +X {w}\n".encode()
    for y in range(h):
        out += rgbe_line(rgbe_encode(rgb[h - 1 - y]))
    with open(path, "wb") as f:
        f.write(bytes(out))


def load_pfm(path):
    """(H, W, 3) float32, row 0 = bottom (little-endian PFM)."""
    with open(path, "rb") as f:
        data = f.read()
    head, rest = data.split(b"\n", 1)
    dims, rest = rest.split(b"\n", 1)
    scale, rest = rest.split(b"\n", 1)
    if head != b"PF" or float(scale) >= 0:
        raise ValueError("not a little-endian colour PFM")
    w, h = (int(x) for x in dims.split())
    return np.frombuffer(rest, "<f4", count=w * h * 3).reshape(h, w, 3).copy()


def load_hdr(path):
    """Decodes a Radiance RGBE file (flat or run-length encoded scanlines) to
    (H, W, 3) float32 RGB, row 0 = top as stored; pixel values
    (m + 0.5) * 2^(e - 136), 0 where e = 0."""
    with open(path, "rb") as f:
        data = f.read()
    pos = data.index(b"\n\n") + 2
    end = data.index(b"\n", pos)
    parts = data[pos:end].split()
    h, w = int(parts[1]), int(parts[3])
    pos = end + 1
    img = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if 8 <= w <= 0x7FFF and data[pos] == 2 and data[pos + 1] == 2:
            if (data[pos + 2] << 8 | data[pos + 3]) != w:
                raise ValueError("bad scanline")
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = data[pos]
                    pos += 1
                    if n > 128:
                        img[y, x:x + n - 128, c] = data[pos]
                        pos += 1
                        x += n - 128
                    else:
                        img[y, x:x + n, c] = np.frombuffer(data[pos:pos + n], np.uint8)
                        pos += n
                        x += n
        else:
            img[y] = np.frombuffer(data[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
    e = img[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return ((img[..., :3] + 0.5) * scale[..., None]).astype(F)
