"""sardine_terrain_correction_noalloc (the entry point that allocates no
device memory) gives the same results as sardine_terrain_correction, bit
for bit, through the C ABI: two calls in a row on one stream, the second on
a smaller AOI (its scratch keeps the first call's contents), on a WGS84 and
a UTM grid; and a call whose work exceeds the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

from sardine_amd import _native, abi, api, synthetic
from sardine_amd.workload import DeviceScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(cuda):
    a = synthetic.scene_small_plain(5, lines=512, samples=1024, out_rows=400, out_cols=400, dem_side=512)
    b = synthetic.scene_iw_full(2, n_swaths=3, n_bursts=3, lines_per_burst=400, raw_cols=1600, overlap_lines=60,
                                swath_overlap=120, utm=True)
    out = []
    for sc in (a, b):
        ds = DeviceScene.build(sc, cuda)
        out.append((sc, ds, ds.prepare_merged()))
    return out


def _call(entry, ds, m, cfg, rows, cols):
    """One terrain-correction call through the C ABI into fresh outputs."""
    import torch

    dev = m.data.device
    data = torch.full((rows, cols), -7.0, dtype=torch.float32, device=dev)
    cl = torch.full((rows, cols), -7.0, dtype=torch.float32, device=dev)
    mk = torch.full((rows, cols), 77, dtype=torch.uint8, device=dev)
    img = abi.GeocodedImage()
    img.data, img.cos_lia, img.mask = data.data_ptr(), cl.data_ptr(), mk.data_ptr()
    img.radar_coords = None
    sc, c, rim, dr = ds.scene.as_c(), cfg.as_c(), m.as_c(), ds.dem.as_c()
    st = getattr(_native.lib(), entry)(C.byref(rim), C.byref(sc), C.byref(dr), C.byref(c), C.byref(img),
                                       C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == abi.OK, (entry, st, _native.last_error())
    torch.cuda.synchronize()
    counters = (img.rows, img.cols, tuple(img.geotransform), img.total_pixel_count, img.valid_pixel_count,
                img.dem_missing_count, img.outside_footprint_count, img.non_converged_count,
                img.degenerate_geometry_count, img.flat_masked_count, img.noise_masked_count,
                tuple(img.newton_iter_histogram))
    return data.cpu().numpy(), cl.cpu().numpy(), mk.cpu().numpy(), counters


@pytest.mark.parametrize("which", [0, 1])
def test_noalloc_equals_allocating_entry_point(scenes, which):
    sc, ds, m = scenes[which]
    rows, cols, _ = api.output_grid(ds.scene, ds.tc_config())
    big = ds.tc_config(compute_lia=True, noise_floor_margin_db=1.0)
    r0, c0, h, w = rows // 5, cols // 7, rows // 3, cols // 2
    small = ds.tc_config(compute_lia=True, flatten=True, resampling=abi.RESAMPLE_BICUBIC, aoi=(r0, c0, h, w))
    for cfg, shape in ((big, (rows, cols)), (small, (h, w))):
        print(f"noalloc scene {which} shape {shape}", flush=True)
        got = _call("sardine_terrain_correction_noalloc", ds, m, cfg, *shape)
        ref = _call("sardine_terrain_correction", ds, m, cfg, *shape)
        assert got[3] == ref[3]
        assert got[3][4] > 0  # valid pixels
        for g, r in zip(got[:3], ref[:3]):
            assert np.array_equal(g, r, equal_nan=True)


def test_noalloc_refuses_work_beyond_its_scratch(scenes):
    import torch

    sc, ds, m = scenes[0]
    bb = sc.bbox
    # a 12000 x 12000 WGS84 grid: its chain records alone need ~72 MB
    cfg = ds.tc_config(pixel_spacing_deg=min(bb[1] - bb[0], bb[3] - bb[2]) / 12000.0)
    rows, cols, _ = api.output_grid(ds.scene, cfg)
    assert rows * cols > 130_000_000
    data = torch.empty((rows, cols), dtype=torch.float32, device=m.data.device)
    img = abi.GeocodedImage()
    img.data = data.data_ptr()
    s, c, rim, dr = ds.scene.as_c(), cfg.as_c(), m.as_c(), ds.dem.as_c()
    st = _native.lib().sardine_terrain_correction_noalloc(C.byref(rim), C.byref(s), C.byref(dr), C.byref(c),
                                                          C.byref(img), None)
    assert st == abi.ERR_CAPACITY
    assert "scratch" in _native.last_error()
