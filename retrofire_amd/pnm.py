"""PPM output of rendered frames (util/pnm.rs:271-312): `save_ppm` /
`write_ppm` write the P6 format (binary 8-bit RGB) with the reference's
one-line header `P6 {w} {h} 255\\n` (pnm.rs:171-180), pixels row by row as
Color4/Color3 into [u8; 3] Rgb888 (pixfmt.rs:111-115). `load_ppm` reads P6
files (pnm.rs:183-270, binary pixmap only; `#` comments skipped)."""
from __future__ import annotations

import io

import numpy as np


def write_ppm(out: io.RawIOBase, color: np.ndarray) -> None:
    """`color` is an (h, w, 4) or (h, w, 3) uint8 array (Buf2<Color4> or
    Buf2<Color3>); alpha is dropped."""
    c = np.asarray(color)
    if c.dtype != np.uint8 or c.ndim != 3 or c.shape[2] not in (3, 4):
        raise ValueError("write_ppm takes an (h, w, 3|4) uint8 image")
    h, w = c.shape[:2]
    out.write(f"P6 {w} {h} 255\n".encode())
    out.write(np.ascontiguousarray(c[..., :3]).tobytes())


def save_ppm(path, color: np.ndarray) -> None:
    """Writes `color` to `path` (overwriting it) as P6."""
    with open(path, "wb") as f:
        write_ppm(f, color)


def load_ppm(path) -> np.ndarray:
    """Reads a P6 file into an (h, w, 3) uint8 array."""
    data = open(path, "rb").read()
    fields, i = [], 0
    while len(fields) < 4:
        if data[i:i + 1] == b"#":
            while data[i:i + 1] not in (b"\n", b""):
                i += 1
        elif data[i:i + 1].isspace():
            i += 1
        else:
            j = i
            while j < len(data) and not data[j:j + 1].isspace():
                j += 1
            fields.append(data[i:j])
            i = j
    i += 1  # the single whitespace byte after the maximum value
    if fields[0] != b"P6" or fields[3] != b"255":
        raise ValueError(f"unsupported PNM: {fields[0]!r}, max {fields[3]!r}")
    w, h = int(fields[1]), int(fields[2])
    px = np.frombuffer(data[i:i + w * h * 3], dtype=np.uint8)
    if px.size < w * h * 3:
        raise ValueError("unexpected end of PNM data")
    return px.reshape(h, w, 3).copy()
