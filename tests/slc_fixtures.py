"""Writers of synthetic Sentinel-1-style SLC measurement TIFFs for the
reader's tests: classic TIFF, uncompressed CInt16 (SampleFormat 5, 32 bits
per sample), one row per strip, the rows stored contiguously after the
strip tables — the layout SlcReader::open accepts (slc_reader.rs:268-461) —
with switches that break one property at a time for the error cases.
"""
from __future__ import annotations

import struct

import numpy as np


def slc_tiff_bytes(data: np.ndarray, le: bool = True, offsets_type: int = 4, magic: int = 42,
                   overrides: dict = None, drop: tuple = (), strip_count: int = None, gap_last: int = 0,
                   data_gap: int = 0) -> bytes:
    """data: (height, width, 2) int16.  overrides: {tag: (type, value)} for
    scalar tags; drop: tags left out; strip_count: StripOffsets count other
    than the height; gap_last: bytes added to the last strip offset (a
    non-contiguous table); data_gap: padding between the tables and row 0."""
    h, w = data.shape[:2]
    e = "<" if le else ">"
    n_strips = h if strip_count is None else strip_count
    tags = {256: (4, w), 257: (4, h), 258: (3, 32), 259: (3, 1), 277: (3, 1), 278: (4, 1), 339: (3, 5)}
    tags.update(overrides or {})
    entries = [t for t in sorted(set(tags) | {273, 279}) if t not in drop]
    ifd_size = 2 + 12 * len(entries) + 4
    each = 2 if offsets_type == 3 else 4
    table = 8 + ifd_size
    counts_at = table + (n_strips * each if n_strips > 1 else 0)
    data_at = counts_at + (n_strips * 4 if n_strips > 1 else 0) + data_gap
    row = w * 4
    offs = [data_at + k * row for k in range(n_strips)]
    if n_strips > 1 and gap_last:
        offs[-1] += gap_last
    out = bytearray(b"II" if le else b"MM")
    out += struct.pack(e + "HI", magic, 8)
    out += struct.pack(e + "H", len(entries))
    for t in entries:
        if t == 273:
            if n_strips == 1:
                field = struct.pack(e + "I", offs[0]) if offsets_type == 4 else struct.pack(e + "HH", offs[0], 0)
            else:
                field = struct.pack(e + "I", table)
            out += struct.pack(e + "HHI", 273, offsets_type, n_strips) + field
        elif t == 279:
            field = struct.pack(e + "I", row if n_strips == 1 else counts_at)
            out += struct.pack(e + "HHI", 279, 4, n_strips) + field
        else:
            typ, v = tags[t]
            field = struct.pack(e + "HH", v, 0) if typ == 3 else struct.pack(e + "I", v)
            out += struct.pack(e + "HHI", t, typ, 1) + field
    out += struct.pack(e + "I", 0)
    if n_strips > 1:
        out += b"".join(struct.pack(e + ("H" if each == 2 else "I"), o) for o in offs)
        out += b"".join(struct.pack(e + "I", row) for _ in offs)
    out += b"\0" * data_gap
    assert len(out) == data_at
    out += np.ascontiguousarray(data, dtype=e + "i2").tobytes()
    return bytes(out)


def write_slc_tiff(path, data: np.ndarray, **kw) -> None:
    with open(path, "wb") as f:
        f.write(slc_tiff_bytes(data, **kw))


def minimal_tiff(entries) -> bytes:
    """The reference's make_minimal_tiff (slc_reader.rs:663-700): a header, one
    IFD of inline (tag, type, count, value) entries sorted by tag, no data."""
    out = bytearray(b"II") + struct.pack("<HI", 42, 8) + struct.pack("<H", len(entries))
    for tag, typ, count, val in sorted(entries):
        field = struct.pack("<HH", val, 0) if typ == 3 else struct.pack("<I", val)
        out += struct.pack("<HHI", tag, typ, count) + field
    out += struct.pack("<I", 0) + b"\0"
    return bytes(out)


def srtm_field(lat_sw: int, lon_sw: int, voids=()) -> np.ndarray:
    """A 3601 x 3601 SRTM 1" tile of a smooth synthetic height field sampled on
    its grid (row 0 = the north edge), so neighbouring tiles agree on their
    shared edges; `voids` = (row, col) cells set to -32768."""
    r = np.arange(3601)
    lat = (lat_sw + 1) - r / 3600.0
    lon = lon_sw + r / 3600.0
    h = 800.0 + 600.0 * np.sin(np.radians(lat[:, None] * 37.0)) * np.cos(np.radians(lon[None, :] * 53.0))
    t = np.round(h).astype(np.int16)
    for rc in voids:
        t[rc] = -32768
    return t


def write_hgt(path, tile: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(tile.astype(">i2").tobytes())


def mosaic(tiles: dict) -> tuple:
    """{(lat_sw, lon_sw): tile} → (raster float32 with NaN voids/holes, north, west)."""
    lats = [k[0] for k in tiles]
    lons = [k[1] for k in tiles]
    north, south, west, east = max(lats) + 1, min(lats), min(lons), max(lons) + 1
    out = np.full(((north - south) * 3600 + 1, (east - west) * 3600 + 1), np.nan, np.float32)
    for (la, lo), t in tiles.items():
        f = t.astype(np.float32)
        f[t == -32768] = np.nan
        r0, c0 = (north - (la + 1)) * 3600, (lo - west) * 3600
        out[r0:r0 + 3601, c0:c0 + 3601] = f
    return out, north, west
