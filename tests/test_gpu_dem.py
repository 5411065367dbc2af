"""GPU parity of the SRTM mosaic loader (DemMosaic::load_srtm_files,
dem.rs:628-640, as one device raster): the assembled raster equals the
tiles' values at their places (big-endian decode, voids and holes NaN), and
the terrain correction's DEM lookup on it gives the per-tile heights."""
import pathlib
import zlib

import numpy as np
import pytest

import oracle as O
import slc_fixtures as F
from sardine_amd import abi, api

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", [((50, 10), (50, 11), (51, 10), (51, 11)), ((50, 10), (51, 11)), ((-6, -120),)])
def test_srtm_mosaic_raster(cuda, tmp_path, layout):
    print(f"srtm mosaic {layout}", flush=True)
    tiles = {}
    # voids in the first tile: one inside, one a sample away from the corner
    # it shares with its neighbours (on a shared edge the per-tile lookup
    # moves on to the next tile when a tap is void, which one joined raster
    # cannot do; real SRTM tiles agree on their shared edges)
    for la, lo in sorted(layout):
        tiles[(la, lo)] = F.srtm_field(la, lo, [(7, 9), (3599, 3599)] if (la, lo) == min(layout) else [])
        ns, ew = ("N" if la >= 0 else "S"), ("E" if lo >= 0 else "W")
        F.write_hgt(tmp_path / f"{ns}{abs(la):02d}{ew}{abs(lo):03d}.hgt", tiles[(la, lo)])
    paths = sorted(str(p) for p in tmp_path.glob("*.hgt"))
    assert [pathlib.Path(p).stem for p in paths] == [
        f"{'N' if la >= 0 else 'S'}{abs(la):02d}{'E' if lo >= 0 else 'W'}{abs(lo):03d}" for la, lo in sorted(layout)]
    dem = api.load_srtm_mosaic(paths, device=cuda)
    want, north, west = F.mosaic(tiles)
    got = dem.data.cpu().numpy()
    assert (dem.origin_lat_deg, dem.origin_lon_deg, dem.pixel_deg) == (north, west, 1.0 / 3600.0)
    assert got.shape == want.shape and np.array_equal(got, want, equal_nan=True)
    # the device raster under the TC's DEM rule vs the per-tile lookup
    rng = np.random.default_rng(len(layout))
    la0, lo0 = min(layout)
    lat = rng.uniform(la0 + 0.001, la0 + 0.999, 300)
    lon = rng.uniform(lo0 + 0.001, lo0 + 0.999, 300)
    # points around the voids too
    lat = np.concatenate([lat, la0 + 1 - (7 + rng.uniform(-1, 1, 20)) / 3600, la0 + 1 - (3599 + rng.uniform(-1, 1, 20)) / 3600])
    lon = np.concatenate([lon, lo0 + (9 + rng.uniform(-1, 1, 20)) / 3600, lo0 + (3599 + rng.uniform(-1, 1, 20)) / 3600])
    st, ref, ok = O.srtm_sample(paths, lat, lon)
    spec = abi.DemRaster(0, got.shape[1], got.shape[0], north, west, 1.0 / 3600.0)
    single = np.array([np.nan if (v := O.dem_elevation(got, spec, a, b)) is None else v for a, b in zip(lat, lon)],
                      np.float32)
    assert st == 0 and np.array_equal(np.isfinite(single), ok.astype(bool))
    assert (~ok.astype(bool)).sum() >= 2  # the voids were hit
    # heights: the joined raster's row coordinate (north - lat) / pix and the
    # tile's (1 - frac) * 3600 differ in the last bit of the f64 weight, so
    # the f32 heights agree to 1 ulp
    both = np.isfinite(single) & ok.astype(bool)
    ulps = np.abs(single[both].view(np.int32).astype(np.int64) - ref[both].astype(np.float32).view(np.int32))
    assert ulps.max(initial=0) <= 1


def _expect(a, nodata):
    """Glo30Tile::load's sample rules (dem.rs:317-342): non-finite floats and
    the nodata value (cast to the sample type, saturating) become NaN."""
    f = a.astype(np.float32)
    if a.dtype == np.float32:
        bad = ~np.isfinite(a)
        if nodata is not None:
            bad |= a == np.float32(nodata)
    else:
        bad = np.zeros(a.shape, bool)
        if nodata is not None:
            nd = float(nodata)
            info = np.iinfo(a.dtype)
            cast = 0 if np.isnan(nd) else int(np.clip(np.trunc(nd), info.min, info.max))
            bad = a == cast
    f[bad] = np.nan
    return f


CASES = [(dt, comp, pred, tile, le, big)
         for dt in ("f4", "i2", "u2")
         for comp, pred in (("none", 1), ("deflate", 1), ("deflate", 2), ("lzw", 2), ("packbits", 1), ("deflate", 3),
                            ("lzw", 3))
         if not (pred == 3 and dt != "f4")
         for tile, le, big in ((None, True, False), ((64, 32), True, True), ((32, 32), False, False))]


@pytest.mark.parametrize("dt,comp,pred,tile,le,big", CASES)
def test_glo30_tile_decode(cuda, tmp_path, dt, comp, pred, tile, le, big):
    rng = np.random.default_rng(zlib.crc32(f"{dt}/{comp}/{pred}".encode()))
    shape = (131, 97)
    if dt == "f4":
        a = rng.normal(800, 300, shape).astype(np.float32)
        a[3, 4], a[5, 6], a[7, 8] = np.nan, np.inf, -32767.0
        nodata = "-32767"
    else:
        a = rng.integers(-400 if dt == "i2" else 0, 4000, shape).astype(dt)
        a[10, 10] = -32768 if dt == "i2" else 65535
        nodata = " -32768.7 " if dt == "i2" else "1e9"  # trimmed; saturating casts: -32768 and 65535
    p = tmp_path / "N46E007.tif"
    p.write_bytes(F.geotiff_dem_bytes(a, tile=tile, compression=comp, predictor=pred, le=le, big=big,
                                      scale=(1 / 3600, 1 / 3600, 0), tiepoint=(0, 0, 0, 7.0, 47.0, 0), nodata=nodata,
                                      rows_per_strip=9))
    dem = api.load_glo30_tile(p, device=cuda)
    got = dem.data.cpu().numpy()
    want = _expect(a, nodata)
    assert np.array_equal(got, want, equal_nan=True)
    assert (dem.origin_lat_deg, dem.origin_lon_deg) == (47.0 - 0.5 / 3600, 7.0 + 0.5 / 3600)


def test_srtm_back_to_back_loads(cuda, tmp_path):
    """Mosaic loads issued back to back on one stream, with no
    synchronisation between them, each equal their own tiles: the pinned
    chunks a load reads through are refilled by the next load only after the
    copies out of them have run."""
    import torch

    layouts = [((10, 20), (10, 21)), ((-3, 100),), ((40, -5), (41, -5))]
    paths, wants = [], []
    for i, layout in enumerate(layouts):
        d = tmp_path / f"set{i}"
        d.mkdir()
        tiles = {}
        for la, lo in layout:
            tiles[(la, lo)] = F.srtm_field(la, lo, [(11 * i + 3, 5)])
            ns, ew = ("N" if la >= 0 else "S"), ("E" if lo >= 0 else "W")
            F.write_hgt(d / f"{ns}{abs(la):02d}{ew}{abs(lo):03d}.hgt", tiles[(la, lo)])
        paths.append(sorted(str(p) for p in d.glob("*.hgt")))
        wants.append(F.mosaic(tiles)[0])
    print("srtm back-to-back loads", flush=True)
    s = torch.cuda.current_stream()
    order = [0, 1, 2, 0, 2, 1, 1, 0]
    got = [api.load_srtm_mosaic(paths[k], device=cuda, stream=s) for k in order]
    torch.cuda.synchronize()
    for k, dem in zip(order, got):
        assert np.array_equal(dem.data.cpu().numpy(), wants[k], equal_nan=True)
