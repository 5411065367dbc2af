"""A small TIFF / BigTIFF reader written from the TIFF 6.0 and BigTIFF
specifications, used to check exported COG files independently of the
writer that made them: it walks the IFD chain, decodes every tag, and
reassembles each level's raster from its float32 tiles.
"""
from __future__ import annotations

import struct

import numpy as np

# TIFF field type → (struct code, size)
_TYPES = {1: ("B", 1), 2: ("c", 1), 3: ("H", 2), 4: ("I", 4), 5: ("II", 8), 11: ("f", 4), 12: ("d", 8),
          16: ("Q", 8)}


def parse(buf: bytes, pointer_tags=(324, 325)) -> dict:
    """The IFD chain of `buf`.  Tags in `pointer_tags` are read through their
    value field as an offset even when the values would fit inline: the
    reference's COG writer stores TileOffsets / TileByteCounts that way for
    single-tile levels too (export.rs:1675-1677 reads them so), where the
    specification would store the value inline."""
    if buf[:2] != b"II":
        raise ValueError("not a little-endian TIFF")
    version = struct.unpack_from("<H", buf, 2)[0]
    if version == 42:
        big = False
        first = struct.unpack_from("<I", buf, 4)[0]
    elif version == 43:
        big = True
        if struct.unpack_from("<HH", buf, 4) != (8, 0):
            raise ValueError("bad BigTIFF header")
        first = struct.unpack_from("<Q", buf, 8)[0]
    else:
        raise ValueError(f"bad TIFF version {version}")
    ifds = []
    off = first
    while off:
        here = off
        if big:
            n = struct.unpack_from("<Q", buf, off)[0]
            p = off + 8
        else:
            n = struct.unpack_from("<H", buf, off)[0]
            p = off + 2
        tags = {}
        order = []
        for _ in range(n):
            if big:
                tag, typ, count = struct.unpack_from("<HHQ", buf, p)
                vbytes, width = buf[p + 12:p + 20], 8
                p += 20
            else:
                tag, typ, count = struct.unpack_from("<HHI", buf, p)
                vbytes, width = buf[p + 8:p + 12], 4
                p += 12
            code, size = _TYPES[typ]
            nbytes = size * count
            if nbytes <= width and tag not in pointer_tags:
                raw = vbytes[:nbytes]
            else:
                ptr = struct.unpack_from("<Q" if big else "<I", vbytes)[0]
                raw = buf[ptr:ptr + nbytes]
            if typ == 5:
                vals = [struct.unpack_from("<II", raw, 8 * i) for i in range(count)]
            else:
                vals = list(struct.unpack_from("<" + code * count, raw))
            tags[tag] = (typ, vals)
            order.append(tag)
        off = struct.unpack_from("<Q" if big else "<I", buf, p)[0]
        ifds.append({"offset": here, "tags": tags, "order": order})
    return {"bigtiff": big, "ifds": ifds}


def level_raster(buf: bytes, ifd: dict, padded: bool = False) -> np.ndarray:
    """The float32 raster of one IFD from its tiles (with padding if asked)."""
    t = ifd["tags"]
    cols, rows = t[256][1][0], t[257][1][0]
    tw, th = t[322][1][0], t[323][1][0]
    offs, counts = t[324][1], t[325][1]
    nx, ny = -(-cols // tw), -(-rows // th)
    assert len(offs) == nx * ny == len(counts)
    out = np.empty((ny * th, nx * tw), np.float32)
    for i, (o, n) in enumerate(zip(offs, counts)):
        assert n == tw * th * 4
        ty, tx = divmod(i, nx)
        out[ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw] = np.frombuffer(buf, "<f4", tw * th, o).reshape(th, tw)
    return out if padded else out[:rows, :cols]


def strip_raster(buf: bytes, ifd: dict) -> np.ndarray:
    """The raster of a single-band striped IFD (8-bit unsigned or 32-bit
    float samples)."""
    t = ifd["tags"]
    cols, rows = t[256][1][0], t[257][1][0]
    bits, fmt = t[258][1][0], t[339][1][0]
    dt = {(8, 1): "u1", (32, 3): "<f4"}[(bits, fmt)]
    per = t[278][1][0]
    offs, counts = t[273][1], t[279][1]
    out = np.empty((rows, cols), dt)
    for i, (o, n) in enumerate(zip(offs, counts)):
        r0, r1 = i * per, min((i + 1) * per, rows)
        assert n == (r1 - r0) * cols * (bits // 8)
        out[r0:r1] = np.frombuffer(buf, dt, (r1 - r0) * cols, o).reshape(r1 - r0, cols)
    return out
