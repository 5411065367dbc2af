"""GPU parity of the SRTM mosaic loader (DemMosaic::load_srtm_files,
dem.rs:628-640, as one device raster): the assembled raster equals the
tiles' values at their places (big-endian decode, voids and holes NaN), and
the terrain correction's DEM lookup on it gives the per-tile heights."""
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
    for la, lo in layout:
        tiles[(la, lo)] = F.srtm_field(la, lo, [(7, 9), (3600, 3600)] if (la, lo) == layout[0] else [])
        ns, ew = ("N" if la >= 0 else "S"), ("E" if lo >= 0 else "W")
        F.write_hgt(tmp_path / f"{ns}{abs(la):02d}{ew}{abs(lo):03d}.hgt", tiles[(la, lo)])
    paths = sorted(str(p) for p in tmp_path.glob("*.hgt"))
    dem = api.load_srtm_mosaic(paths, device=cuda)
    want, north, west = F.mosaic(tiles)
    got = dem.data.cpu().numpy()
    assert (dem.origin_lat_deg, dem.origin_lon_deg, dem.pixel_deg) == (north, west, 1.0 / 3600.0)
    assert got.shape == want.shape and np.array_equal(got, want, equal_nan=True)
    # the device raster under the TC's DEM rule vs the per-tile lookup
    rng = np.random.default_rng(len(layout))
    la0, lo0 = layout[0]
    lat = rng.uniform(la0 + 0.001, la0 + 0.999, 300)
    lon = rng.uniform(lo0 + 0.001, lo0 + 0.999, 300)
    st, ref, ok = O.srtm_sample(paths, lat, lon)
    spec = abi.DemRaster(0, got.shape[1], got.shape[0], north, west, 1.0 / 3600.0)
    single = np.array([np.nan if (v := O.dem_elevation(got, spec, a, b)) is None else v for a, b in zip(lat, lon)],
                      np.float32)
    assert st == 0 and np.array_equal(np.isfinite(single), ok.astype(bool))
