"""Image error metrics and image files.

mse / rel_mse follow the reference's experiment scripts
(SDK/data/Scripts/parse_logs.py:43-50): NaN pixels of the image count as 0,
mse = mean((ref - img)^2), rel_mse = mean((ref - img)^2 / (ref^2 + 1e-2)),
over all pixels and the three colour channels.
Writers: binary PPM (8-bit sRGB), PFM (float RGB) and uncompressed OpenEXR
(float RGBA scanlines) -- the reference writes EXR/PNG through tinyexr
(sutil.cpp:556-651); these are minimal writers of the published formats.
"""
import struct

import numpy as np


def _rgb(img):
    a = np.array(np.asarray(img)[..., :3], np.float64)
    a[np.isnan(a)] = 0.0
    return a


def mse(img, ref):
    a, b = _rgb(img), np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((b - a) ** 2))


def rel_mse(img, ref, eps=1e-2):
    a, b = _rgb(img), np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((b - a) ** 2 / (b * b + eps)))


def write_ppm(path, rgba8):
    """rgba8: (H, W, 4) uint8, row 0 at the bottom (the sampler's y axis)."""
    a = np.asarray(rgba8, np.uint8)[::-1, :, :3]
    h, w = a.shape[:2]
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(a).tobytes())


def write_pfm(path, rgba):
    """rgba: (H, W, >=3) float; PFM stores rows bottom to top, little endian."""
    a = np.asarray(rgba, np.float32)[..., :3]
    h, w = a.shape[:2]
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(a).astype("<f4").tobytes())


def write_exr(path, rgba):
    """Uncompressed scanline OpenEXR with FLOAT channels A, B, G, R."""
    a = np.asarray(rgba, np.float32)
    if a.shape[-1] == 3:
        a = np.concatenate([a, np.ones(a.shape[:2] + (1,), np.float32)], axis=-1)
    a = a[::-1]  # EXR rows run top to bottom
    h, w = a.shape[:2]

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data

    chans = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1) for c in "ABGR") + b"\0"
    header = (struct.pack("<ii", 20000630, 2) +
              attr("channels", "chlist", chans) +
              attr("compression", "compression", b"\0") +
              attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1)) +
              attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1)) +
              attr("lineOrder", "lineOrder", b"\0") +
              attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
              attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)) +
              attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    line_bytes = 4 * 4 * w
    table_start = len(header)
    first = table_start + 8 * h
    offsets = struct.pack("<%dQ" % h, *[first + y * (8 + line_bytes) for y in range(h)])
    body = bytearray()
    for y in range(h):
        body += struct.pack("<ii", y, line_bytes)
        for c in (3, 2, 1, 0):  # A, B, G, R channel order
            body += np.ascontiguousarray(a[y, :, c]).astype("<f4").tobytes()
    with open(path, "wb") as f:
        f.write(header + offsets + bytes(body))


def read_exr(path):
    """Reader for the files write_exr produces (uncompressed FLOAT ABGR)."""
    data = open(path, "rb").read()
    magic, ver = struct.unpack_from("<ii", data, 0)
    if magic != 20000630:
        raise ValueError("not an OpenEXR file")
    p = 8
    attrs = {}
    while data[p] != 0:
        e = data.index(b"\0", p)
        name = data[p:e].decode()
        t = data.index(b"\0", e + 1)
        size = struct.unpack_from("<i", data, t + 1)[0]
        attrs[name] = data[t + 5:t + 5 + size]
        p = t + 5 + size
    p += 1
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offs = struct.unpack_from("<%dQ" % h, data, p)
    out = np.zeros((h, w, 4), np.float32)
    for y in range(h):
        q = offs[y] + 8
        for c in (3, 2, 1, 0):
            out[y, :, c] = np.frombuffer(data, "<f4", w, q)
            q += 4 * w
    return out[::-1]
