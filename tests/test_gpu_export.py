"""GPU parity of the Cloud-Optimised GeoTIFF export (export.rs:329-760)
through the C ABI: the file the device path writes equals the oracle's byte
for byte (header, IFDs, overview means, tile packing and padding), on odd
and even sizes, both CRS families, several tile sizes, both load paths of
the pack kernel (16-byte vector rows and the scalar fallback) and rasters
holding NaNs with payloads, signed zeros, infinities and subnormals.
"""
import numpy as np
import pytest

import oracle as O
import tiffread
from sardine_amd import abi, api, synthetic
from sardine_amd.workload import DeviceScene

pytestmark = pytest.mark.gpu

GT_WGS = [8.0, 1e-4, 0.0, 52.0, 0.0, -1e-4]
GT_UTM = [500000.0, 10.0, 0.0, 5.2e6, 0.0, -10.0]


def _raster(rows, cols, seed, special=True):
    rng = np.random.default_rng(seed)
    a = (rng.normal(0, 2, (rows, cols)) * 10.0 ** rng.integers(-20, 20, (rows, cols))).astype(np.float32)
    if special and rows * cols >= 16:
        u = a.view(np.uint32).reshape(-1)
        idx = rng.choice(rows * cols, size=min(rows * cols // 4, 4000), replace=False)
        pats = np.array([0x7FC00000, 0x7FC00001, 0xFFC00000, 0x7F800001, 0x00000000, 0x80000000, 0x7F800000,
                         0xFF800000, 0x00000001, 0x007FFFFF, 0x80400000], np.uint32)
        u[idx] = pats[rng.integers(0, len(pats), idx.size)]
        if rows >= 4 and cols >= 4:
            a[:2, :2] = np.nan  # one all-NaN patch
            a[2:4, 2:4] = np.float32(1e-40)  # a subnormal patch
    return a


def _first_diff(a: bytes, b: bytes):
    if len(a) != len(b):
        return f"sizes {len(a)} vs {len(b)}"
    x = np.frombuffer(a, np.uint8)
    y = np.frombuffer(b, np.uint8)
    i = int(np.flatnonzero(x != y)[0])
    return f"first difference at byte {i} of {len(a)}"


def _compare(tmp_path, data, gt, crs=abi.CRS_WGS84, zone=0, ts=512, dev=None):
    import torch

    rows, cols = data.shape
    print(f"cog parity rows={rows} cols={cols} ts={ts} crs={crs}", flush=True)
    ref_path, out_path = tmp_path / "ref.tif", tmp_path / "gpu.tif"
    assert O.write_cog(ref_path, data, gt, crs=crs, utm_zone=zone, tile_size=ts) == 0
    t = dev if dev is not None else torch.from_numpy(data).to("cuda")
    api.write_cog_with_crs(out_path, t, cols, rows, gt, crs=crs, tile_size=ts, utm_zone=zone)
    a, b = out_path.read_bytes(), ref_path.read_bytes()
    assert a == b, _first_diff(a, b)
    return a


@pytest.mark.parametrize("rows,cols,ts,crs,zone", [
    (1, 1, 64, abi.CRS_WGS84, 0), (2, 2, 64, abi.CRS_WGS84, 0), (64, 64, 64, abi.CRS_WGS84, 0),
    (130, 129, 64, abi.CRS_UTM_SOUTH, 17), (200, 300, 64, abi.CRS_UTM_NORTH, 33),
    (777, 1000, 128, abi.CRS_WGS84, 0), (700, 90, 128, abi.CRS_UTM_NORTH, 60), (3001, 2500, 256, abi.CRS_WGS84, 0),
    (1030, 4099, 1024, abi.CRS_UTM_SOUTH, 1)])
def test_cog_file_bytes_match_oracle(cuda, tmp_path, rows, cols, ts, crs, zone):
    data = _raster(rows, cols, rows * 31 + cols)
    _compare(tmp_path, data, GT_UTM if crs else GT_WGS, crs, zone, ts)


def test_cog_reference_cases(cuda, tmp_path):
    # the rasters of export.rs:1605-1695 (1024x512 ramp at ts 512, 64x64 ramp at ts 64)
    d = (np.arange(512 * 1024, dtype=np.float32) * np.float32(0.001)).reshape(512, 1024)
    buf = _compare(tmp_path, d, GT_WGS, ts=512)
    assert buf[:4] == b"II*\x00" and int.from_bytes(buf[8:10], "little") == 17
    d = np.arange(64 * 64, dtype=np.float32).reshape(64, 64)
    buf = _compare(tmp_path, d, [0.0, 0.001, 0.0, 1.0, 0.0, -0.001], ts=64)
    t = tiffread.parse(buf)
    full = tiffread.level_raster(buf, t["ifds"][-1])
    assert full[0, 0] == 0.0 and full[1, 0] == 64.0 and full[63, 63] == 4095.0
    ov = tiffread.level_raster(buf, t["ifds"][0])
    assert ov.shape == (32, 32) and ov[0, 0] == np.float32((0 + 1 + 64 + 65) / 4)


def test_cog_scalar_and_vector_load_paths(cuda, tmp_path):
    """Rows whose 4-column groups are 16-byte aligned take the vector loads;
    a raster starting one float into its buffer takes the scalar ones."""
    import torch

    data = _raster(333, 512, 5)
    _compare(tmp_path, data, GT_WGS, ts=64)
    buf = torch.empty(333 * 512 + 1, dtype=torch.float32, device=cuda)
    view = buf[1:].view(333, 512)
    view.copy_(torch.from_numpy(data).to(cuda))
    assert view.data_ptr() % 16 != 0
    _compare(tmp_path, data, GT_WGS, ts=64, dev=view)


def test_cog_pack_and_header_compose_the_file(cuda, tmp_path):
    import torch

    data = _raster(515, 700, 11)
    ref_path = tmp_path / "ref.tif"
    assert O.write_cog(ref_path, data, GT_UTM, crs=abi.CRS_UTM_NORTH, utm_zone=31, tile_size=128) == 0
    ref = ref_path.read_bytes()
    lay = api.cog_layout(700, 515, 128, abi.CRS_UTM_NORTH)
    head = api.cog_header(lay, GT_UTM, abi.CRS_UTM_NORTH, 31)
    tiles = api.cog_pack(torch.from_numpy(data).to(cuda), lay)
    torch.cuda.synchronize()
    body = tiles.cpu().numpy().tobytes()
    assert head == ref[:lay.header_bytes]
    assert body == ref[lay.header_bytes:], _first_diff(body, ref[lay.header_bytes:])
    # a second pack into the same buffer (every byte is rewritten)
    tiles.fill_(1.0)
    api.cog_pack(torch.from_numpy(data).to(cuda), lay, tiles=tiles)
    assert tiles.cpu().numpy().tobytes() == body


def test_cog_of_geocoded_db_product(cuda, tmp_path):
    """The export as the pipeline uses it: a geocoded sigma0 raster, turned
    into dB on the device, written as a COG; the oracle writes the same
    raster (the device's own dB values) and the files are identical."""
    sc = synthetic.scene_small_plain(4, lines=512, samples=1024, out_rows=300, out_cols=420, dem_side=512)
    ds = DeviceScene.build(sc, cuda, host_raw=True)
    img = ds.terrain_correction(ds.prepare_merged(), ds.tc_config())
    api.to_db_inplace(img.data, 0.0)
    dev = img.data.reshape(img.rows, img.cols).contiguous()
    host = dev.cpu().numpy()
    assert np.isfinite(host).any()
    _compare(tmp_path, host, list(img.geotransform), ts=64, dev=dev)


def test_cog_write_errors(cuda, tmp_path):
    import torch

    d = torch.ones(4, dtype=torch.float32, device=cuda)
    with pytest.raises(Exception) as e:
        api.write_cog_with_crs(tmp_path / "a.tif", d, 2, 2, GT_WGS, tile_size=100)
    assert e.value.status == abi.ERR_EXPORT_INVALID_TILE_SIZE
    with pytest.raises(Exception) as e:
        api.write_cog_with_crs(tmp_path / "a.tif", d, 2, 2, [8.0, 0.0, 0.0, 52.0, 0.0, -1e-4], tile_size=512)
    assert e.value.status == abi.ERR_EXPORT_INVALID_GEOTRANSFORM
    with pytest.raises(Exception) as e:
        api.write_cog_with_crs(tmp_path / "no_such_dir" / "a.tif", d, 2, 2, GT_WGS, tile_size=64)
    assert e.value.status == abi.ERR_EXPORT_IO


# ── striped GeoTIFF / BigTIFF writers (export.rs:139-284, 762-1219) ──────────
@pytest.mark.parametrize("rows,cols,kind,crs,zone", [
    (2, 2, "f32", abi.CRS_WGS84, 0), (64, 5, "f32", abi.CRS_UTM_NORTH, 32), (65, 7, "u8", abi.CRS_UTM_SOUTH, 1),
    (200, 333, "none", 0, 0), (1, 1, "u8", abi.CRS_WGS84, 0), (3, 0, "f32", abi.CRS_WGS84, 0),
    (1000, 1537, "f32", abi.CRS_UTM_NORTH, 60)])
def test_geotiff_file_bytes_match_oracle(cuda, tmp_path, rows, cols, kind, crs, zone):
    import torch

    print(f"geotiff parity rows={rows} cols={cols} kind={kind} crs={crs}", flush=True)
    rng = np.random.default_rng(rows * 3 + cols)
    data = rng.integers(0, 256, (rows, cols)).astype(np.uint8) if kind == "u8" else _raster(rows, cols, rows + cols)
    gt = GT_UTM if crs else GT_WGS
    ref_path, out_path = tmp_path / "ref.tif", tmp_path / "gpu.tif"
    dev = torch.from_numpy(data).to(cuda)
    if kind == "none":
        assert O.write_geotiff(ref_path, data, None) == 0
        api.write_tiff_no_crs(out_path, dev, cols, rows)
    elif kind == "u8":
        assert O.write_geotiff(ref_path, data, gt, crs=crs, utm_zone=zone) == 0
        api.write_geotiff_u8_with_crs(out_path, dev, cols, rows, gt, crs=crs, utm_zone=zone)
    else:
        assert O.write_geotiff(ref_path, data, gt, crs=crs, utm_zone=zone) == 0
        api.write_geotiff_with_crs(out_path, dev, cols, rows, gt, crs=crs, utm_zone=zone)
    a, b = out_path.read_bytes(), ref_path.read_bytes()
    assert a == b, _first_diff(a, b)


def test_geotiff_of_geocoded_product_and_mask(cuda, tmp_path):
    """The pipeline's default export (run.rs:443-451): the geocoded dB raster
    and its uint8 quality mask as striped GeoTIFFs."""
    sc = synthetic.scene_small_plain(6, lines=512, samples=1024, out_rows=260, out_cols=310, dem_side=512)
    ds = DeviceScene.build(sc, cuda, host_raw=True)
    img = ds.terrain_correction(ds.prepare_merged(), ds.tc_config(compute_lia=True))
    api.to_db_inplace(img.data, 0.0)
    gt = list(img.geotransform)
    for name, dev in (("db", img.data.reshape(img.rows, img.cols).contiguous()),
                      ("mask", img.mask.reshape(img.rows, img.cols).contiguous())):
        host = dev.cpu().numpy()
        ref_path, out_path = tmp_path / f"{name}_ref.tif", tmp_path / f"{name}.tif"
        assert O.write_geotiff(ref_path, host, gt) == 0
        if name == "db":
            api.write_geotiff_with_crs(out_path, dev, img.cols, img.rows, gt)
        else:
            api.write_geotiff_u8_with_crs(out_path, dev, img.cols, img.rows, gt)
        a, b = out_path.read_bytes(), ref_path.read_bytes()
        assert a == b, (name, _first_diff(a, b))


def test_bigtiff_file_matches_oracle(cuda):
    """A raster above the 4 GiB classic-TIFF limit (32768 x 32768 float32):
    the device path's BigTIFF equals the oracle's (compared by digest; the
    files live in /dev/shm and are removed)."""
    import hashlib
    import os
    import pathlib

    import torch

    rows = cols = 32768
    shm = pathlib.Path("/dev/shm")
    if not shm.is_dir():
        pytest.skip("no /dev/shm")
    print("bigtiff parity 32768x32768 f32", flush=True)
    dev = torch.arange(rows * cols, dtype=torch.int32, device=cuda).mul_(2654435761 % (1 << 31)).view(
        torch.float32).reshape(rows, cols)
    ref_path, out_path = shm / f"sd_big_ref_{os.getpid()}.tif", shm / f"sd_big_gpu_{os.getpid()}.tif"

    def digest(p):
        h = hashlib.sha256()
        with open(p, "rb") as f:
            while True:
                b = f.read(64 << 20)
                if not b:
                    return h.hexdigest()
                h.update(b)

    try:
        api.write_geotiff_with_crs(out_path, dev, cols, rows, GT_WGS)
        host = dev.cpu().numpy()
        del dev
        assert O.write_geotiff(ref_path, host, GT_WGS) == 0
        del host
        with open(out_path, "rb") as f:
            head = f.read(16)
        assert head[:4] == b"II+\x00"
        assert out_path.stat().st_size == ref_path.stat().st_size
        assert digest(out_path) == digest(ref_path)
    finally:
        for p in (ref_path, out_path):
            if p.exists():
                p.unlink()
