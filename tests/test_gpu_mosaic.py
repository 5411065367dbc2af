"""GPU parity of the GLO-30 mosaic (DemMosaic::load_glo30_files + the
terrain correction's lookup, DemMosaic::elevation_at_linear over
Glo30Tile::elevation_at, dem.rs:351-372, 642-653, 730-747): the loader's
raster, the device lookup at points on and around the tiles' shared edges
and corners (load order, per-tile clamping, NaN taps handing over to the
next tile, empty cells) and a terrain correction over a 2 x 2 mosaic, each
against the oracle."""
import copy

import numpy as np
import pytest

import oracle as O
import slc_fixtures as F
from sardine_amd import _native, abi, api, synthetic
from sardine_amd.workload import DeviceScene

pytestmark = pytest.mark.gpu

DB_REL = 10 ** (1e-3 / 10) - 1


def _tiling(rows, cols, th, tw, olat, olon, ranks):
    t = abi.DemTiling()
    t.rows, t.cols, t.tile_height, t.tile_width = rows, cols, th, tw
    for i in range(rows):
        t.origin_lat_deg[i] = olat[i]
    for i in range(cols):
        t.origin_lon_deg[i] = olon[i]
    for i in range(64):
        t.rank[i] = -1
    for (r, c), k in ranks.items():
        t.rank[r * 8 + c] = k
    return t


def _probe(dem: api.DemRaster, lat, lon):
    import torch

    dev = dem.data.device
    la = torch.from_numpy(np.ascontiguousarray(lat, np.float64)).to(dev)
    lo = torch.from_numpy(np.ascontiguousarray(lon, np.float64)).to(dev)
    out = torch.empty(len(lat), dtype=torch.float32, device=dev)
    r = dem.as_c()
    assert _native.lib().sardine_diag_dem(r, la.data_ptr(), lo.data_ptr(), out.data_ptr(), len(lat), None) == 0
    return out.cpu().numpy()


def test_glo30_mosaic_loader_raster(cuda, tmp_path):
    rng = np.random.default_rng(4)
    cells = {"N46E008.tif": (47.0, 8.0), "N46E007.tif": (47.0, 7.0), "N45E008.tif": (46.0, 8.0)}
    tiles, paths = {}, []
    for name, (la, lo) in cells.items():
        a = rng.normal(500, 100, (40, 40)).astype(np.float32)
        tiles[name] = a
        p = tmp_path / name
        p.write_bytes(F.geotiff_dem_bytes(a, scale=(1 / 40, 1 / 40, 0), tiepoint=(0, 0, 0, lo, la, 0),
                                          compression="deflate", predictor=3))
        paths.append(str(p))
    dem = api.load_glo30_mosaic(paths, device=cuda)
    t = dem.tiling
    assert (t.rows, t.cols, t.tile_height, t.tile_width) == (2, 2, 40, 40)
    got = dem.data.cpu().numpy()
    want = np.full((80, 80), np.nan, np.float32)
    want[0:40, 0:40] = tiles["N46E007.tif"]
    want[0:40, 40:80] = tiles["N46E008.tif"]
    want[40:80, 40:80] = tiles["N45E008.tif"]
    assert np.array_equal(got, want, equal_nan=True)
    assert [t.rank[0], t.rank[1], t.rank[8], t.rank[9]] == [1, 0, -1, 2]


@pytest.fixture(scope="module")
def small(cuda):
    sc = synthetic.scene_small_plain(5, lines=512, samples=1024, out_rows=400, out_cols=400, dem_side=512)
    ds = DeviceScene.build(sc, cuda, host_raw=True)
    info, s0, nz, _ = O.prepare_merged(sc)
    return sc, ds, info, s0, nz


def _split(sc, ds, ranks, nan_rows=()):
    """The scene's 512 x 512 DEM declared a 2 x 2 mosaic of 256 x 256 tiles
    (each at its own origin, one tile apart); rows listed in nan_rows set to
    NaN in the copy."""
    import torch

    sp = sc.dem_spec
    h = np.array(sc.dem(), np.float32, copy=True)
    for r in nan_rows:
        h[r, :] = np.nan
    olat = [sp.origin_lat_deg - 256 * k * sp.pixel_deg for k in range(2)]
    olon = [sp.origin_lon_deg + 256 * k * sp.pixel_deg for k in range(2)]
    t = _tiling(2, 2, 256, 256, olat, olon, ranks)
    dev = api.DemRaster(torch.from_numpy(h).to(ds.dem.data.device), sp.origin_lat_deg, sp.origin_lon_deg,
                        sp.pixel_deg, t)
    spec = copy.copy(sp)
    spec.tiling = t
    return h, dev, spec, olat, olon


@pytest.mark.parametrize("ranks,nan_rows", [({(0, 0): 0, (0, 1): 1, (1, 0): 2, (1, 1): 3}, ()),
                                            ({(0, 0): 3, (0, 1): 2, (1, 0): 1}, (255, 256))])
def test_mosaic_lookup_matches_oracle(small, ranks, nan_rows):
    sc, ds, info, s0, nz = small
    h, dev, spec, olat, olon = _split(sc, ds, ranks, nan_rows)
    sp = sc.dem_spec
    rng = np.random.default_rng(len(ranks))
    n = 4000
    lat = sp.origin_lat_deg - rng.uniform(-1.0, 512.0, n) * sp.pixel_deg
    lon = sp.origin_lon_deg + rng.uniform(-1.0, 512.0, n) * sp.pixel_deg
    # the shared edges and the corner, and a few ulps around them
    e_lat = olat[0] - 255.5 * sp.pixel_deg
    e_lon = olon[0] + 255.5 * sp.pixel_deg
    eps = np.array([0.0, 1e-13, -1e-13, 1e-11, -1e-11, 0.3 * sp.pixel_deg, -0.3 * sp.pixel_deg])
    lon_e = sp.origin_lon_deg + rng.uniform(0.0, 511.0, len(eps) * 20) * sp.pixel_deg
    lat_e = sp.origin_lat_deg - rng.uniform(0.0, 511.0, len(eps) * 20) * sp.pixel_deg
    lat = np.concatenate([lat, np.repeat(e_lat + eps, 20), lat_e, e_lat + eps, np.full(len(eps), e_lat)])
    lon = np.concatenate([lon, lon_e, np.repeat(e_lon + eps, 20), np.full(len(eps), e_lon), e_lon + eps])
    got = _probe(dev, lat, lon)
    want = np.array([np.nan if (v := O.dem_mosaic_elevation(h, sp.pixel_deg, spec.tiling, a, b)) is None else v
                     for a, b in zip(lat, lon)], np.float32)
    assert np.array_equal(got, want, equal_nan=True), np.argwhere(~((got == want) | (np.isnan(got) & np.isnan(want))))[:5]
    # the mosaic differs from one raster near the shared edges (per-tile clamping)
    single = api.DemRaster(dev.data, sp.origin_lat_deg, sp.origin_lon_deg, sp.pixel_deg)
    one = _probe(single, lat, lon)
    assert not np.array_equal(one, got, equal_nan=True)


def test_tc_over_glo30_mosaic_matches_oracle(small):
    sc, ds, info, s0, nz = small
    ranks = {(0, 0): 2, (0, 1): 0, (1, 0): 3, (1, 1): 1}
    h, dev, spec, _, _ = _split(sc, ds, ranks, nan_rows=(256,))
    m = ds.prepare_merged()
    cfg = ds.tc_config(compute_lia=True)
    img = api.terrain_correction(m, ds.scene, dev, cfg, radar_coords=True)
    keep = []
    meta = dict(lines=info.lines, samples=info.samples, azimuth_start_time_us=info.azimuth_start_time_us,
                near_slant_range_time_s=info.near_slant_range_time_s, range_pixel_spacing_m=info.range_pixel_spacing_m)
    ref = O.terrain_correction(s0, nz, meta, sc.scene_c(keep), h, spec, sc.tc_config(compute_lia=1), coords=True)
    assert ref.status == 0, ref.error
    assert np.array_equal(img.mask.cpu().numpy(), ref.mask)
    c = ref.counters()
    assert (img.valid_pixel_count, img.dem_missing_count, img.flat_masked_count) == (
        c["valid"], c["dem_missing"], c["flat_masked"])
    assert img.newton_iter_histogram == c["hist"]
    rc = img.radar_coords.cpu().numpy()
    both = np.isfinite(ref.coords[..., 0])
    assert np.array_equal(np.isfinite(rc[..., 0]), both)
    assert np.abs(rc[..., 0][both] - ref.coords[..., 0][both]).max() <= 1e-3
    assert np.abs(rc[..., 1][both] - ref.coords[..., 1][both]).max() <= 1e-3
    d = img.data.cpu().numpy()
    assert np.array_equal(np.isnan(d), np.isnan(ref.data))
    v = np.isfinite(ref.data)
    assert (np.abs(d[v] - ref.data[v]) <= DB_REL * np.abs(ref.data[v])).all()
    # the mosaic changed the answer somewhere (per-tile clamping, the NaN row)
    one = api.terrain_correction(m, ds.scene, api.DemRaster(dev.data, dev.origin_lat_deg, dev.origin_lon_deg,
                                                            dev.pixel_deg), cfg)
    assert not np.array_equal(one.mask.cpu().numpy(), ref.mask) or not np.array_equal(one.data.cpu().numpy(), d)
    print(f"mosaic TC: valid {img.valid_pixel_count}, dem missing {img.dem_missing_count}", flush=True)
