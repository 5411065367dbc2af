"""Writers of synthetic Sentinel-1-style SLC measurement TIFFs for the
reader's tests: classic TIFF, uncompressed CInt16 (SampleFormat 5, 32 bits
per sample), one row per strip, the rows stored contiguously after the
strip tables — the layout SlcReader::open accepts (slc_reader.rs:268-461) —
with switches that break one property at a time for the error cases.
"""
from __future__ import annotations

import struct
import zlib

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
    """{(lat_sw, lon_sw): tile} → (raster float32 with NaN voids/holes, north, west).
    Where tiles overlap (shared edges) the first tile in dict order wins."""
    lats = [k[0] for k in tiles]
    lons = [k[1] for k in tiles]
    north, south, west, east = max(lats) + 1, min(lats), min(lons), max(lons) + 1
    out = np.full(((north - south) * 3600 + 1, (east - west) * 3600 + 1), np.nan, np.float32)
    for (la, lo), t in reversed(list(tiles.items())):
        f = t.astype(np.float32)
        f[t == -32768] = np.nan
        r0, c0 = (north - (la + 1)) * 3600, (lo - west) * 3600
        out[r0:r0 + 3601, c0:c0 + 3601] = f
    return out, north, west


# ── GeoTIFF DEM tiles (the Copernicus GLO-30 case) ─────────────────────────
def _lzw_encode(data: bytes) -> bytes:
    """TIFF LZW: MSB-first codes of 9-12 bits, a clear code first and before
    the table fills, end of information last.  The encoder widens its codes
    when its next free code reaches 512 / 1024 / 2048 (the decoder, one entry
    behind, at 511 / 1023 / 2047)."""
    out, acc, nbits = bytearray(), 0, 0

    def put(code, width):
        nonlocal acc, nbits
        acc = (acc << width) | code
        nbits += width
        while nbits >= 8:
            nbits -= 8
            out.append((acc >> nbits) & 0xFF)

    table = {bytes([i]): i for i in range(256)}
    nxt, width = 258, 9
    put(256, width)
    w = b""
    for ch in data:
        wc = w + bytes([ch])
        if wc in table:
            w = wc
            continue
        put(table[w], width)
        table[wc] = nxt
        nxt += 1
        if nxt == 512:
            width = 10
        elif nxt == 1024:
            width = 11
        elif nxt == 2048:
            width = 12
        if nxt == 4094:
            put(256, width)
            table = {bytes([i]): i for i in range(256)}
            nxt, width = 258, 9
        w = bytes([ch])
    if w:
        put(table[w], width)
        nxt += 1
        if nxt == 512:
            width = 10
        elif nxt == 1024:
            width = 11
        elif nxt == 2048:
            width = 12
    put(257, width)
    if nbits:
        out.append((acc << (8 - nbits)) & 0xFF)
    return bytes(out)


def _packbits(data: bytes) -> bytes:
    out, i = bytearray(), 0
    while i < len(data):
        j = i
        while j + 1 < len(data) and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j > i:
            out += bytes([(257 - (j - i + 1)) & 0xFF, data[i]])
            i = j + 1
        else:
            k = i
            while k < len(data) and k - i < 128 and not (k + 1 < len(data) and data[k + 1] == data[k]):
                k += 1
            k = max(k, i + 1)
            out += bytes([k - i - 1]) + data[i:k]
            i = k
    return bytes(out)


def geotiff_dem_bytes(a: np.ndarray, tile=None, compression: str = "deflate", predictor: int = 1, le: bool = True,
                      big: bool = False, scale=None, tiepoint=None, nodata: str = None, rows_per_strip: int = 7) -> bytes:
    """A single-band DEM GeoTIFF of `a` (float32 / int16 / uint16 / other
    dtypes for the unsupported-type cases), strips or (tw, th) tiles."""
    h, w = a.shape
    bps = a.dtype.itemsize
    fmt = {"f": 3, "i": 2, "u": 1}[a.dtype.kind]
    comp = {"none": 1, "lzw": 5, "deflate": 8, "packbits": 32773}[compression]
    e = "<" if le else ">"
    bw, bh = (tile if tile else (w, rows_per_strip))
    nx, ny = -(-w // bw), -(-h // bh)
    blocks = []
    for by in range(ny):
        for bx in range(nx):
            rows = bh if tile else min(bh, h - by * bh)
            blk = np.zeros((rows, bw), a.dtype)
            src = a[by * bh:by * bh + rows, bx * bw:bx * bw + bw]
            blk[:src.shape[0], :src.shape[1]] = src
            if predictor == 3:  # byte planes, most significant first, then byte differences
                be = blk.astype(blk.dtype.newbyteorder(">")).view(np.uint8).reshape(rows, bw, bps)
                planes = np.ascontiguousarray(be.transpose(0, 2, 1)).reshape(rows, bw * bps)
                diff = planes.copy()
                diff[:, 1:] = (planes[:, 1:].astype(np.int16) - planes[:, :-1].astype(np.int16)) & 0xFF
                raw = diff.astype(np.uint8).tobytes()
            else:
                x = blk
                if predictor == 2:
                    iv = x.view(np.dtype(f"u{bps}"))
                    d = iv.copy()
                    d[:, 1:] = iv[:, 1:] - iv[:, :-1]
                    x = d.view(a.dtype)
                raw = x.astype(x.dtype.newbyteorder(e)).tobytes()
            blocks.append(zlib.compress(raw) if comp == 8 else _lzw_encode(raw) if comp == 5
                          else _packbits(raw) if comp == 32773 else raw)
    tags = [(256, 4, [w]), (257, 4, [h]), (258, 3, [bps * 8]), (259, 3, [comp]), (262, 3, [1]), (277, 3, [1]),
            (284, 3, [1]), (317, 3, [predictor]), (339, 3, [fmt])]
    if tile:
        tags += [(322, 4, [bw]), (323, 4, [bh]), (324, 16 if big else 4, None), (325, 16 if big else 4, None)]
    else:
        tags += [(273, 16 if big else 4, None), (278, 4, [bh]), (279, 16 if big else 4, None)]
    if scale is not None:
        tags.append((33550, 12, list(scale)))
    if tiepoint is not None:
        tags.append((33922, 12, list(tiepoint)))
    if nodata is not None:
        tags.append((42113, 2, nodata.encode() + b"\0"))
    tags.sort(key=lambda t: t[0])
    head = 16 if big else 8
    ent, cnt, nxt = (20, 8, 8) if big else (12, 2, 4)
    ifd_len = cnt + ent * len(tags) + nxt
    fmt_ch = {3: "H", 4: "I", 12: "d", 16: "Q"}
    extra = bytearray()
    data_at = head + ifd_len
    # out-of-line values first, then the blocks
    offs_tag = 324 if tile else 273
    cnt_tag = 325 if tile else 279
    counts = [len(b) for b in blocks]
    offsets_placeholder = {}
    vals = {}
    for tag, typ, v in tags:
        if tag in (offs_tag, cnt_tag):
            continue
        vals[tag] = struct.pack(e + "".join(fmt_ch[typ] for _ in v), *v) if typ != 2 else v
    nb = len(blocks)
    per = 8 if big else 4
    arr_len = nb * per
    fieldsz = 8 if big else 4
    layout = {}
    cur = data_at
    for tag, typ, v in tags:
        n = arr_len if tag in (offs_tag, cnt_tag) else len(vals[tag])
        if n > fieldsz:
            layout[tag] = cur
            cur += n
    block_at = cur
    offs = []
    for b in blocks:
        offs.append(block_at)
        block_at += len(b)
    vals[offs_tag] = struct.pack(e + ("Q" if big else "I") * nb, *offs)
    vals[cnt_tag] = struct.pack(e + ("Q" if big else "I") * nb, *counts)
    out = bytearray(b"II" if le else b"MM")
    out += struct.pack(e + "HHHQ", 43, 8, 0, head) if big else struct.pack(e + "HI", 42, head)
    out += struct.pack(e + ("Q" if big else "H"), len(tags))
    for tag, typ, v in tags:
        payload = vals[tag]
        count = (len(payload) // {2: 1, 3: 2, 4: 4, 12: 8, 16: 8}[typ])
        out += struct.pack(e + "HH" + ("Q" if big else "I"), tag, typ, count)
        if tag in layout:
            out += struct.pack(e + ("Q" if big else "I"), layout[tag])
        else:
            out += payload + b"\0" * (fieldsz - len(payload))
    out += b"\0" * nxt
    for tag, typ, v in tags:
        if tag in layout:
            assert len(out) == layout[tag]
            out += vals[tag]
    assert len(out) == offs[0] if offs else True
    for b in blocks:
        out += b
    return bytes(out)
