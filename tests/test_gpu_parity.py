"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bars (from the path's contract): masks, counters and integer layouts
bit-exact; radar geolocation within 1e-3 line/sample; sigma0 within 1e-3 dB
(relative 2.3e-4); calibration and speckle window statistics bit-exact where
the arithmetic is the same IEEE sequence.
"""
import numpy as np
import pytest

import oracle as O
from sardine_amd import TerrainCorrectionError, abi, api, synthetic
from test_oracle_reference import _ww15mgh_text
from sardine_amd.workload import DeviceScene

pytestmark = pytest.mark.gpu

DB_REL = 10 ** (1e-3 / 10) - 1  # 1e-3 dB as a relative bound (≈2.3e-4)


def _meta(info):
    return dict(lines=info.lines, samples=info.samples, azimuth_start_time_us=info.azimuth_start_time_us,
                near_slant_range_time_s=info.near_slant_range_time_s,
                range_pixel_spacing_m=info.range_pixel_spacing_m)


@pytest.fixture(scope="module")
def small(cuda):
    sc = synthetic.scene_small_plain(5, lines=512, samples=1024, out_rows=400, out_cols=400, dem_side=512)
    ds = DeviceScene.build(sc, cuda, host_raw=True)
    info, s0, nz, _ = O.prepare_merged(sc)
    return sc, ds, info, s0, nz


@pytest.fixture(scope="module")
def iw(cuda):
    sc = synthetic.scene_iw_full(2, n_swaths=3, n_bursts=3, lines_per_burst=400, raw_cols=1600,
                                 overlap_lines=60, swath_overlap=120, utm=True, dem_max_side=None)
    ds = DeviceScene.build(sc, cuda, host_raw=True)
    info, s0, nz, layouts = O.prepare_merged(sc)
    return sc, ds, info, s0, nz, layouts


def test_calibration_small_bit_exact(small):
    sc, ds, info, s0, nz = small
    m = ds.prepare_merged()
    assert (m.lines, m.samples) == (info.lines, info.samples)
    assert np.array_equal(m.data.cpu().numpy(), s0)
    assert np.array_equal(m.nesz.cpu().numpy(), nz)
    assert m.near_slant_range_time_s == info.near_slant_range_time_s
    assert m.azimuth_start_time_us == info.azimuth_start_time_us


def test_deburst_merge_calibration_iw_bit_exact(iw):
    sc, ds, info, s0, nz, layouts = iw
    m = ds.prepare_merged()
    gi = m.info
    assert (gi.lines, gi.samples) == (info.lines, info.samples)
    assert list(gi.sw_out_offset)[:3] == list(info.sw_out_offset)[:3]
    assert list(gi.seam)[:2] == list(info.seam)[:2]
    assert gi.cal_lut_extrapolation_gap_px == info.cal_lut_extrapolation_gap_px
    assert gi.noise_lut_extrapolation_gap_px == info.noise_lut_extrapolation_gap_px
    a, b = m.data.cpu().numpy(), m.nesz.cpu().numpy()
    assert np.array_equal(np.isnan(a), np.isnan(s0))
    assert np.array_equal(a, s0, equal_nan=True)
    assert np.array_equal(b, nz, equal_nan=True)


def _tc_compare(sc, ds, info, s0, nz, m, geoid=None, **kw):
    cfg = ds.tc_config(compute_lia=True, **kw)
    if geoid is not None:
        cfg.geoid_grid = geoid
    img = ds.terrain_correction(m, cfg, radar_coords=True)
    keep = []
    ccfg = sc.tc_config(compute_lia=1, **{k: (int(v) if isinstance(v, bool) else v) for k, v in kw.items()})
    if geoid is not None:
        g = np.ascontiguousarray(geoid, np.float32)
        keep.append(g)
        ccfg.geoid = abi.Geoid(abi.GEOID_GRID_2P5, g.ctypes.data)
    ref = O.terrain_correction(s0, nz, _meta(info), sc.scene_c(keep), sc.dem(), sc.dem_spec, ccfg, coords=True)
    assert ref.status == 0, ref.error
    mk = img.mask.cpu().numpy()
    assert np.array_equal(mk, ref.mask), np.bincount((mk != ref.mask).ravel())
    rc = img.radar_coords.cpu().numpy()
    both = np.isfinite(ref.coords[..., 0])
    assert np.array_equal(np.isfinite(rc[..., 0]), both)
    dl = np.abs(rc[..., 0][both] - ref.coords[..., 0][both])
    ds_ = np.abs(rc[..., 1][both] - ref.coords[..., 1][both])
    assert dl.max() <= 1e-3 and ds_.max() <= 1e-3, (dl.max(), ds_.max())
    d = img.data.cpu().numpy()
    assert np.array_equal(np.isnan(d), np.isnan(ref.data))
    v = np.isfinite(ref.data)
    rel = np.abs(d[v] - ref.data[v]) / np.maximum(np.abs(ref.data[v]), 1e-30)
    assert rel.max() <= DB_REL, rel.max()
    c = ref.counters()
    assert img.valid_pixel_count == c["valid"]
    assert img.dem_missing_count == c["dem_missing"]
    assert img.outside_footprint_count == c["outside_footprint"]
    assert img.non_converged_count == c["non_converged"]
    assert img.degenerate_geometry_count == c["degenerate"]
    assert img.flat_masked_count == c["flat_masked"]
    assert img.noise_masked_count == c["noise_masked"]
    assert img.newton_iter_histogram == c["hist"]
    cl = img.cos_lia.cpu().numpy()
    assert np.array_equal(np.isnan(cl), np.isnan(ref.cos_lia))
    f = np.isfinite(ref.cos_lia)
    if f.any():
        assert np.abs(cl[f] - ref.cos_lia[f]).max() <= 1e-5
    return img, ref


def test_tc_wgs84_bilinear(small):
    sc, ds, info, s0, nz = small
    m = ds.prepare_merged()
    img, ref = _tc_compare(sc, ds, info, s0, nz, m)
    assert img.valid_pixel_count > 0.3 * img.total_pixel_count
    assert tuple(img.geotransform) == tuple(ref.img.geotransform)


@pytest.mark.parametrize("kernel", [abi.RESAMPLE_BICUBIC, abi.RESAMPLE_LANCZOS3])
def test_tc_wgs84_resamplers(small, kernel):
    sc, ds, info, s0, nz = small
    m = ds.prepare_merged()
    _tc_compare(sc, ds, info, s0, nz, m, resampling=kernel)


def test_tc_flatten_and_nesz(small):
    sc, ds, info, s0, nz = small
    m = ds.prepare_merged()
    img, ref = _tc_compare(sc, ds, info, s0, nz, m, flatten=True, noise_floor_margin_db=3.0)
    assert img.flat_masked_count + img.noise_masked_count > 0


def test_tc_utm_iw(iw):
    sc, ds, info, s0, nz, _ = iw
    m = ds.prepare_merged()
    img, ref = _tc_compare(sc, ds, info, s0, nz, m)
    assert img.valid_pixel_count > 0


def test_tc_aoi_window(small):
    sc, ds, info, s0, nz = small
    m = ds.prepare_merged()
    cfg = ds.tc_config(compute_lia=True, aoi=(37, 51, 130, 177))
    img = ds.terrain_correction(m, cfg)
    keep = []
    ccfg = sc.tc_config(compute_lia=1, aoi_row0=37, aoi_col0=51, aoi_rows=130, aoi_cols=177)
    ref = O.terrain_correction(s0, nz, _meta(info), sc.scene_c(keep), sc.dem(), sc.dem_spec, ccfg)
    assert np.array_equal(img.mask.cpu().numpy(), ref.mask)
    d = img.data.cpu().numpy()
    v = np.isfinite(ref.data)
    assert np.array_equal(np.isfinite(d), v)
    assert (np.abs(d[v] - ref.data[v]) / ref.data[v]).max() <= DB_REL


@pytest.mark.parametrize("aoi,resampling", [((5, 17, 97, 301), 2), ((100, 2216, 50, 1), 0), ((0, 900, 1, 700), 1),
                                            ((1800, 0, 52, 2217), 0)])
def test_tc_utm_aoi_windows(iw, aoi, resampling):
    """UTM AOI windows (the kernels store only the project_row anchors the
    window's columns blend between) against the oracle on the same window."""
    sc, ds, info, s0, nz, _ = iw
    m = ds.prepare_merged()
    img = ds.terrain_correction(m, ds.tc_config(compute_lia=True, aoi=aoi, resampling=resampling))
    keep = []
    ccfg = sc.tc_config(compute_lia=1, aoi_row0=aoi[0], aoi_col0=aoi[1], aoi_rows=aoi[2], aoi_cols=aoi[3],
                        resampling=resampling)
    ref = O.terrain_correction(s0, nz, _meta(info), sc.scene_c(keep), sc.dem(), sc.dem_spec, ccfg)
    assert ref.status == 0, ref.error
    assert np.array_equal(img.mask.cpu().numpy(), ref.mask)
    d = img.data.cpu().numpy()
    v = np.isfinite(ref.data)
    assert np.array_equal(np.isfinite(d), v)
    if v.any():
        assert (np.abs(d[v] - ref.data[v]) / ref.data[v]).max() <= DB_REL
    assert img.valid_pixel_count == ref.counters()["valid"]


SPECKLES = [
    (abi.SPECKLE_BOXCAR, 3, 1.0, 1.0), (abi.SPECKLE_BOXCAR, 7, 1.0, 1.0),
    (abi.SPECKLE_LEE, 5, 1.0, 1.0), (abi.SPECKLE_LEE, 5, 4.0, 1.0),
    (abi.SPECKLE_FROST, 5, 1.0, 1.0), (abi.SPECKLE_FROST, 7, 1.0, 2.0),
    (abi.SPECKLE_GAMMA_MAP, 5, 1.0, 1.0), (abi.SPECKLE_GAMMA_MAP, 7, 4.0, 1.0),
    (abi.SPECKLE_REFINED_LEE, 7, 1.0, 1.0), (abi.SPECKLE_REFINED_LEE, 7, 4.0, 1.0),
    (abi.SPECKLE_BOXCAR, 9, 1.0, 1.0),
]


@pytest.mark.parametrize("kind,window,enl,damping", SPECKLES)
def test_speckle_filters(cuda, kind, window, enl, damping):
    import torch

    rng = np.random.default_rng(kind * 100 + window)
    rows, cols = 137, 211
    x = rng.exponential(0.1, (rows, cols)).astype(np.float32)
    x[rng.random((rows, cols)) < 0.03] = np.nan  # nodata, as the TC output has
    x[:5, :9] = np.nan
    st, ref = O.speckle(x, kind, window, enl, damping)
    assert st == 0
    f = api.SpeckleFilter(kind, window, enl, damping)
    out = f.apply(torch.from_numpy(x).to(cuda), cols, rows).cpu().numpy()
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    v = np.isfinite(ref)
    rel = np.abs(out[v] - ref[v]) / np.maximum(np.abs(ref[v]), 1e-30)
    assert rel.max() <= DB_REL, rel.max()
    if kind in (abi.SPECKLE_BOXCAR, abi.SPECKLE_LEE, abi.SPECKLE_REFINED_LEE):
        assert np.array_equal(out[v], ref[v])  # same IEEE sequence end to end


@pytest.mark.parametrize("enl", [1.0, 4.4])
def test_refined_lee_dense_wide_range(cuda, enl):
    """No missing samples (every window takes the full-count path) and values
    over nine decades, so the per-direction sums and the constant-count
    divisions see widely spread operands."""
    import torch

    rng = np.random.default_rng(int(enl * 10))
    rows, cols = 301, 257
    x = (10.0 ** rng.uniform(-6.0, 3.0, (rows, cols))).astype(np.float32)
    st, ref = O.speckle(x, abi.SPECKLE_REFINED_LEE, 7, enl, 1.0)
    assert st == 0
    out = api.SpeckleFilter(abi.SPECKLE_REFINED_LEE, 7, enl, 1.0).apply(
        torch.from_numpy(x).to(cuda), cols, rows).cpu().numpy()
    assert np.array_equal(out, ref)


def test_refined_lee_pick_edge_cases(cuda):
    """The direction pick's fast path (reciprocal CV²) must hand every case it
    cannot decide to exact division: constant patches (all CV² = 0, ties
    resolved to the first direction), mirror-symmetric patches (equal
    non-zero CV² in pairs of directions), means whose square underflows or
    is subnormal, means whose square overflows or whose reciprocal would be
    subnormal; bit-exact against the oracle."""
    import torch

    rng = np.random.default_rng(11)
    rows, cols = 160, 192
    x = rng.exponential(1.0, (rows, cols)).astype(np.float32)
    x[0:32, 0:32] = 2.5                                   # constant
    sym = rng.exponential(1.0, (16, 32)).astype(np.float32)
    x[32:64, 0:32] = np.concatenate([sym, sym[::-1]])     # mirrored rows: N/S ties
    x[64:96, 0:32] = np.concatenate([sym.T[:, :16], sym.T[:, :16][:, ::-1]], axis=1)  # mirrored columns: E/W ties
    x[96:128, 0:32] = 1e-22 * rng.exponential(1.0, (32, 32))   # mean² underflows
    x[128:160, 0:32] = 1e-20 * rng.exponential(1.0, (32, 32))  # mean² subnormal
    x[0:32, 32:64] = 1e19 * rng.exponential(1.0, (32, 32))     # mean² above 2^126
    x[32:64, 32:64] = 3e19 * rng.exponential(1.0, (32, 32))    # mean² overflows
    x[64:96, 32:64] = 0.0                                      # all-zero windows
    x[96:128, 32:64] = np.where(rng.random((32, 32)) < 0.5, 0.0, 1.0)
    x[128:160, 32:64] = 7.0
    x[130:140, 40:50] = 7.0000005                             # near ties
    for enl in (1.0, 4.0):
        st, ref = O.speckle(x, abi.SPECKLE_REFINED_LEE, 7, enl, 1.0)
        assert st == 0
        out = api.SpeckleFilter(abi.SPECKLE_REFINED_LEE, 7, enl, 1.0).apply(
            torch.from_numpy(x).to(cuda), cols, rows).cpu().numpy()
        assert np.array_equal(out, ref, equal_nan=True), np.argwhere(~((out == ref) | (np.isnan(out) & np.isnan(ref))))[:5]


@pytest.mark.parametrize("rows,cols,nan_frac,neg", [
    (300, 200, 0.4, False),   # most windows hold a no-data sample: nearly every pixel takes the exact path
    (250, 131, 0.0, True),    # signed values (means near zero, skipped directions), partial last strip
    (1000, 5, 0.02, False),   # one strip narrower than the halo, eight row blocks
    (3, 70, 0.0, False),      # fewer rows than the window
    (1, 1, 0.0, False),
])
def test_refined_lee_walk_shapes(cuda, rows, cols, nan_frac, neg):
    """The column walk at raster edges: partial strips and row blocks, rasters
    smaller than the window, dense no-data (the exact path's pixel list fills
    up) and signed samples; bit-exact against the oracle."""
    import torch

    rng = np.random.default_rng(rows * 7 + cols)
    x = rng.exponential(1.0, (rows, cols)).astype(np.float32)
    if neg:
        x = (x - 1.0).astype(np.float32)
    x[rng.random((rows, cols)) < nan_frac] = np.nan
    st, ref = O.speckle(x, abi.SPECKLE_REFINED_LEE, 7, 2.0, 1.0)
    assert st == 0
    out = api.SpeckleFilter(abi.SPECKLE_REFINED_LEE, 7, 2.0, 1.0).apply(
        torch.from_numpy(x).to(cuda), cols, rows).cpu().numpy()
    assert np.array_equal(out, ref, equal_nan=True)


def test_refined_lee_walk_nodata_bands(cuda):
    """No-data across whole strips: full-width NaN bands of 6, 7, 8 and 20
    rows (windows wholly inside no-data, and the exact path on the rows
    around them), a band of +inf rows (a row of the walk with no finite
    centre gives NaN for infinite centres too), a no-data corner like a
    rotated footprint's and a partial band; bit-exact against the oracle."""
    import torch

    rows, cols = 420, 300
    rng = np.random.default_rng(11)
    x = rng.exponential(1.0, (rows, cols)).astype(np.float32)
    for r0, h in ((20, 6), (60, 7), (100, 8), (150, 20)):
        x[r0:r0 + h] = np.nan
    x[200:212] = np.inf
    yy, xx = np.mgrid[0:rows, 0:cols]
    x[(yy > 260) & (xx + (yy - 260) * 2 < 250)] = np.nan
    x[330:345, :140] = np.nan
    x[400:] = np.nan  # the last rows: a run that reaches the raster's end
    st, ref = O.speckle(x, abi.SPECKLE_REFINED_LEE, 7, 2.0, 1.0)
    assert st == 0
    out = api.SpeckleFilter(abi.SPECKLE_REFINED_LEE, 7, 2.0, 1.0).apply(
        torch.from_numpy(x).to(cuda), cols, rows).cpu().numpy()
    assert np.array_equal(out, ref, equal_nan=True), np.argwhere(~((out == ref) | (np.isnan(out) & np.isnan(ref))))[:5]


def test_to_db(cuda):
    import torch

    rng = np.random.default_rng(3)
    x = rng.exponential(0.1, 100_000).astype(np.float32)
    x[::17] = np.nan
    x[::23] = 1e-6
    ref = x.copy()
    st, conv, masked = O.to_db_inplace(ref, 1e-5)
    t = torch.from_numpy(x).to(cuda)
    c2, m2 = api.to_db_inplace(t, 1e-5)
    assert (c2, m2) == (conv, masked)
    out = t.cpu().numpy()
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    v = np.isfinite(ref)
    assert np.abs(out[v] - ref[v]).max() <= 1e-4


def test_synth_generator_matches_host(cuda):
    import torch

    t = torch.empty((3000, 2), dtype=torch.int16, device=cuda)
    api.synth_cint16(t, 77, 1234, 100.0)
    host = synthetic.complex_gaussian_dn(77, 1234, 3000)
    assert np.abs(t.cpu().numpy().astype(np.int32) - host.astype(np.int32)).max() <= 1
    assert (t.cpu().numpy() == host).mean() > 0.999


def test_orbit_window_error_aborts(small):
    sc, ds, info, s0, nz = small
    m = ds.prepare_merged()
    scene = api.SceneInputs.from_synthetic(sc)
    scene.sv_time_us = scene.sv_time_us[:9]  # orbit ends long before the scene → Newton leaves the window
    scene.sv_pos = scene.sv_pos[:9]
    scene.sv_vel = scene.sv_vel[:9]
    with pytest.raises(TerrainCorrectionError) as e:
        api.terrain_correction(m, scene, ds.dem, ds.tc_config())
    assert getattr(e.value, "status", None) == abi.ERR_TC_ORBIT_OUTSIDE_WINDOW
    # Orbit(OutsideOrbitWindow { requested, first, last }): the window of the
    # 9 vectors, and an iterate past its end
    assert (e.value.first_us, e.value.last_us) == (int(scene.sv_time_us[0]), int(scene.sv_time_us[-1]))
    assert e.value.requested_us > e.value.last_us


@pytest.mark.parametrize("min_seg,lead", [(1, 2), (5, 2), (64, 2), (100000, 2), (5, 0), (16, 0), (32, 1)])
def test_tc_chain_segmentation_invariant(small, min_seg, lead):
    """Segment length (lanes chaining 1 column .. whole rows) and the phase-A
    lead must not change a single output bit: the seed-chain fix-up makes
    every decomposition equal the reference's sequential chain.  lead 0
    starts every segment at the row seed, so the fix-up replays most of them."""
    from sardine_amd import _native

    sc, ds, info, s0, nz = small
    m = ds.prepare_merged()
    L = _native.lib()
    try:
        L.sardine_set_tc_segment(min_seg)
        L.sardine_set_tc_lead(lead)
        img, ref = _tc_compare(sc, ds, info, s0, nz, m)
    finally:
        L.sardine_set_tc_segment(0)
        L.sardine_set_tc_lead(2)
    assert img.valid_pixel_count == ref.counters()["valid"]
    if lead == 0 and min_seg < 64:
        assert img.chain_redo_rounds > 0  # the replay path ran


@pytest.mark.parametrize("band,lead", [(0.4, 2), (0.4, 0), (0.05, 2)])
def test_tc_tie_replays_change_no_bit(small, band, lead):
    """Steps inside the tie band are replayed with the reference's own
    arithmetic (the parallel tie-list pass for rows whose seeds all match,
    the sequential walk for the others and for list overflow).  A wide band
    sends most pixels through those replays: the output must be the same
    bits as with the default band, and match the oracle."""
    from sardine_amd import _native

    sc, ds, info, s0, nz = small
    m = ds.prepare_merged()
    L = _native.lib()
    cfg = ds.tc_config(compute_lia=True, resampling=abi.RESAMPLE_BICUBIC)
    base = ds.terrain_correction(m, cfg, radar_coords=True)
    try:
        L.sardine_set_tc_tie_band(band)
        L.sardine_set_tc_lead(lead)
        img, ref = _tc_compare(sc, ds, info, s0, nz, m, resampling=abi.RESAMPLE_BICUBIC)
    finally:
        L.sardine_set_tc_tie_band(0.0)
        L.sardine_set_tc_lead(2)
    assert img.tie_steps > 0.2 * band * img.valid_pixel_count  # ~2 band of every step is a tie
    assert img.tie_decisions_changed <= img.tie_steps
    assert img.tie_max_dev_us < 2.0 ** -18
    for a, b in ((img.data, base.data), (img.mask, base.mask), (img.cos_lia, base.cos_lia),
                 (img.radar_coords, base.radar_coords)):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy(), equal_nan=True)
    assert img.newton_iter_histogram == base.newton_iter_histogram


@pytest.mark.parametrize("n_sv,spacing_s", [(9, 20.0), (33, 5.0), (64, 2.5)])
def test_tc_orbit_lengths(cuda, n_sv, spacing_s):
    """Orbits of 9, 33 and 64 state vectors (the chain kernels' LDS copy of
    the orbit is sized to its vectors; 64 is the most the device form
    holds): the per-call path against the oracle, bilinear and bicubic."""
    import dataclasses

    print(f"orbit of {n_sv} vectors at {spacing_s} s", flush=True)
    sc = synthetic.scene_small_plain(5, lines=512, samples=1024, out_rows=400, out_cols=400, dem_side=512)
    t, pos, vel = synthetic.state_vectors(sc.orbit, n=n_sv, spacing_s=spacing_s,
                                          t0_s=80.0 - 0.5 * spacing_s * (n_sv - 1))  # centred on the scene
    sc = dataclasses.replace(sc, sv_t_us=t, sv_pos=pos, sv_vel=vel)
    ds = DeviceScene.build(sc, cuda, host_raw=True)
    info, s0, nz, _ = O.prepare_merged(sc)
    m = ds.prepare_merged()
    for rs in (abi.RESAMPLE_BILINEAR, abi.RESAMPLE_BICUBIC):
        _tc_compare(sc, ds, info, s0, nz, m, resampling=rs)


@pytest.mark.parametrize("sat_lon", [None, -5.8])
def test_tc_geoid_grid_ww15mgh(cuda, tmp_path, sat_lon):
    """GeoidModel::Egm96 from a grid file (geoid.rs:220-251; added per pixel
    at terrain_correction.rs:1545): a smooth synthetic 73 x 144 grid written
    as an NGA WW15MGH file, read by sardine_geoid_load_ww15mgh and used by
    the kernels, against the oracle.  sat_lon -5.8 puts the scene at about
    -1.1° longitude, in the cell that wraps from 357.5° to 360° = 0°."""
    print(f"geoid scene sat_lon {sat_lon}", flush=True)
    sc = synthetic.scene_small_plain(9, lines=512, samples=1024, out_rows=300, out_cols=300, dem_side=512,
                                     sat_lon_deg=sat_lon)
    if sat_lon is not None:
        assert -2.5 < sc.bbox[2] and sc.bbox[3] < 0.0
    ds = DeviceScene.build(sc, cuda, host_raw=True)
    info, s0, nz, _ = O.prepare_merged(sc)
    lat = 90.0 - 2.5 * np.arange(73)[:, None]
    lon = 2.5 * np.arange(144)[None, :]
    g = (40 * np.sin(np.radians(2 * lat)) * np.cos(np.radians(lon)) + 12 * np.cos(np.radians(3 * lon))
         + 5 * np.sin(np.radians(5 * lat + lon))).round(6)
    p = tmp_path / "WW15MGH.GRD"
    p.write_text(_ww15mgh_text(g))
    grid = api.load_geoid_grid(p, ascii=True)
    assert grid.shape == (73, 144) and np.array_equal(grid, g.astype(np.float32))
    m = ds.prepare_merged()
    img, ref = _tc_compare(sc, ds, info, s0, nz, m, geoid=grid, flatten=True)
    assert img.valid_pixel_count > 10000
    # the undulation (tens of metres here) moves the geolocation
    flat = ds.terrain_correction(m, ds.tc_config(compute_lia=True, flatten=True), radar_coords=True)
    moved = np.nanmax(np.abs(flat.radar_coords.cpu().numpy()[..., 1] - img.radar_coords.cpu().numpy()[..., 1]))
    assert moved > 1.0
