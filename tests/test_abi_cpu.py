"""CPU-side checks of the C ABI (no GPU needed).

The library must load, export every symbol include/sardine_hip.h declares,
compute the host-only layouts exactly as the oracle does, reject invalid
arguments with the reference's error variants, and refuse compute work
without a gfx950 device (there is no CPU fallback).
"""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

import oracle as O
from sardine_amd import DeburstError, _native, abi, api, synthetic

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_header_symbols_are_exported():
    hdr = (ROOT / "include" / "sardine_hip.h").read_text()
    declared = set(re.findall(r"\b(sardine_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(abi.EXPORTED_SYMBOLS), declared ^ set(abi.EXPORTED_SYMBOLS)
    lib = C.CDLL(str(_native.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    assert _native.lib().sardine_abi_version() == abi.ABI_VERSION


def test_struct_sizes_match_header_layout():
    # field order/size mismatches would show up as a size mismatch here
    assert C.sizeof(abi.StateVector) == 56
    assert C.sizeof(abi.GridPoint) == 16
    assert C.sizeof(abi.SpeckleFilterC) == 16
    assert C.sizeof(abi.DeburstLayout) == 8 * 3 + 8 + 4 * 64 * 8


def _scene(sc, keep):
    return sc.scene_c(keep)


@pytest.mark.parametrize("utm", [False, True])
def test_output_grid_matches_oracle(utm):
    sc = synthetic.scene_iw_full(3, n_swaths=2, n_bursts=2, lines_per_burst=300, raw_cols=1000, overlap_lines=60,
                                 swath_overlap=80, utm=utm)
    keep = []
    cfg = sc.tc_config()
    st, r, c, gt = O.tc_output_grid(_scene(sc, keep), cfg)
    r2, c2, gt2 = api.output_grid(api.SceneInputs.from_synthetic(sc), api.TerrainCorrectionConfig(
        pixel_spacing_deg=sc.pixel_spacing, crs=sc.crs, utm_zone=sc.utm_zone))
    assert st == 0 and (r, c) == (r2, c2)
    assert np.allclose(gt, gt2, rtol=0, atol=1e-6 if utm else 1e-12)


def test_orbit_model_matches_oracle():
    """The kernels' orbit model (per-interval Lagrange-8 polynomials, the
    reference's own values at node times), evaluated on the host through the
    C ABI, against the oracle's product-form interpolate_orbit_and_accel
    (orbit.rs:498-596) and against the exact interpolant (mpmath, 50 digits).

    Exact equality at every state-vector time (where the reference's 1 ns
    guard changes its acceleration formula); elsewhere within the product
    form's own rounding.  The reference forms L'_k as L_k (Σ_j 1/(t-t_j) -
    1/(t-t_k)), which cancels near a node: the sum is ~1/|t - t_node| and
    its rounding, ~2.2e-16/|t - t_node|, is multiplied by velocities of
    ~7.5e3 m/s over 8 terms (1 µs away: 1.4e-7 m/s² observed), so the
    acceleration tolerance against the oracle widens as 2e-11 / |t - t_node|
    there; against the exact interpolant the model stays within 1e-12 m/s²
    everywhere."""
    mp = pytest.importorskip("mpmath")
    sc = synthetic.scene_iw_full(1)
    keep = []
    orb = O.orbit_c(sc.sv_t_us, sc.sv_pos, sc.sv_vel, keep)
    L = _native.lib()
    T = [int(t) for t in sc.sv_t_us]
    t0, t1 = T[0], T[-1]
    rng = np.random.default_rng(7)
    times = T + [int(t) for t in rng.integers(t0, t1 + 1, 4000)] + [t0 + 1, t1 - 1, T[5] + 1, T[5] - 1, T[9] + 7]
    p, v, a = np.zeros(3), np.zeros(3), np.zeros(3)

    def model(t):
        assert L.sardine_orbit_interpolate(C.byref(orb), t, _np_ptr(p), _np_ptr(v), _np_ptr(a)) == abi.OK
        return p.copy(), v.copy(), a.copy()

    for t in times:
        mp_, mv, ma = model(t)
        st, rp, rv, ra = O.interpolate_orbit_and_accel(orb, t - t0)
        assert st == 0
        if t in T:
            assert np.array_equal(mp_, rp) and np.array_equal(mv, rv) and np.array_equal(ma, ra), t
            continue
        gap_s = min(abs(t - tn) for tn in T) * 1e-6
        assert np.abs(mp_ - rp).max() <= 1e-8, t
        assert np.abs(mv - rv).max() <= 1e-10, t
        assert np.abs(ma - ra).max() <= 1e-10 + 2e-11 / gap_s, t

    # exact interpolant (same window rule) at a few times, near nodes included
    mp.mp.dps = 50
    n = len(T)
    xs_all = [mp.mpf(tn - t0) / 10**6 for tn in T]
    for t in [t0 + 1, T[5] + 1, T[9] + 7, t1 - 1] + [int(x) for x in rng.integers(t0, t1, 6)]:
        c = max(i for i in range(n) if T[i] <= t)
        s = 0 if c < 4 else (n - 8 if c + 4 >= n else c - 3)
        ts = mp.mpf(t - t0) / 10**6
        xs = xs_all[s:s + 8]
        exact_p, exact_a = [mp.mpf(0)] * 3, [mp.mpf(0)] * 3
        for k in range(8):
            lk = mp.mpf(1)
            dk = mp.mpf(0)
            for j in range(8):
                if j != k:
                    lk *= (ts - xs[j]) / (xs[k] - xs[j])
            for m in range(8):
                if m == k:
                    continue
                term = 1 / (xs[k] - xs[m])
                for j in range(8):
                    if j not in (k, m):
                        term *= (ts - xs[j]) / (xs[k] - xs[j])
                dk += term
            for d in range(3):
                exact_p[d] += lk * mp.mpf(float(sc.sv_pos[s + k][d]))
                exact_a[d] += dk * mp.mpf(float(sc.sv_vel[s + k][d]))
        mp_, _, ma = model(t)
        assert max(abs(float(mp_[d] - exact_p[d])) for d in range(3)) <= 2e-9, t
        assert max(abs(float(ma[d] - exact_a[d])) for d in range(3)) <= 1e-12, t
    assert L.sardine_orbit_interpolate(C.byref(orb), t1 + 1, _np_ptr(p), _np_ptr(v), _np_ptr(a)) == \
        abi.ERR_TC_ORBIT_OUTSIDE_WINDOW
    # OutsideOrbitWindow { requested, first, last } (orbit.rs:58-66, 513-517)
    d = _native.last_error_detail()
    assert d.status == abi.ERR_TC_ORBIT_OUTSIDE_WINDOW
    assert (d.requested_us, d.first_us, d.last_us) == (t1 + 1, int(sc.sv_t_us[0]), int(sc.sv_t_us[-1]))
    assert L.sardine_orbit_interpolate(C.byref(orb), t0 - 7, _np_ptr(p), _np_ptr(v), _np_ptr(a)) == \
        abi.ERR_TC_ORBIT_OUTSIDE_WINDOW
    assert _native.last_error_detail().requested_us == t0 - 7
    short = O.orbit_c(sc.sv_t_us[:7], sc.sv_pos[:7], sc.sv_vel[:7], keep)
    assert L.sardine_orbit_interpolate(C.byref(short), t0 + 5, _np_ptr(p), _np_ptr(v), _np_ptr(a)) == \
        abi.ERR_TC_ORBIT_INSUFFICIENT
    d = _native.last_error_detail()
    assert (d.status, d.have_vectors, d.need_vectors, d.requested_us) == (abi.ERR_TC_ORBIT_INSUFFICIENT, 7, 8, 0)
    # a later failure of another kind resets the detail
    assert L.sardine_orbit_interpolate(None, 0, _np_ptr(p), _np_ptr(v), _np_ptr(a)) == abi.ERR_INVALID_ARGUMENT
    d = _native.last_error_detail()
    assert (d.status, d.have_vectors) == (abi.ERR_INVALID_ARGUMENT, 0)


def test_tc_dem_error_carries_the_dem_variant():
    """TerrainCorrectionError::Dem(DemError::…) (terrain_correction.rs:51):
    an unusable DEM raster is reported with the DemError variant (host
    validation, before the device check)."""
    from sardine_amd import TerrainCorrectionError
    sc = synthetic.scene_small_plain(0, lines=64, samples=64, out_rows=8, out_cols=8, dem_side=8)
    keep = []
    scene = sc.scene_c(keep)
    img = abi.RadarImage(1, 0, 64, 64, int(sc.sv_t_us[0]) + 10_000_000, 0.0053, float(sc.rps_m))
    cfg = sc.tc_config()
    out = abi.GeocodedImage()
    L = _native.lib()
    for dem, want in ((abi.DemRaster(0, 8, 8, 50.0, 10.0, 1e-3), abi.ERR_DEM_NO_TILES),
                      (abi.DemRaster(1, 1, 8, 50.0, 10.0, 1e-3), abi.ERR_DEM_BAD_SIZE),
                      (abi.DemRaster(1, 8, 8, 50.0, 10.0, 0.0), abi.ERR_DEM_TIFF_DECODE)):
        st = L.sardine_terrain_correction(C.byref(img), C.byref(scene), C.byref(dem), C.byref(cfg), C.byref(out), None)
        assert st == abi.ERR_TC_DEM
        d = _native.last_error_detail()
        assert (d.status, d.dem_status) == (abi.ERR_TC_DEM, want)
        if want == abi.ERR_DEM_BAD_SIZE:
            assert (d.dem_expected_bytes, d.dem_actual_bytes) == (16, 32)
        e = TerrainCorrectionError(st, _native.last_error(), d)
        assert e.dem_variant == abi.STATUS_NAMES[want]


def test_tc_rejects_a_dem_tiling_the_lookup_cannot_serve():
    """A caller-built tiling (sardine_dem_tiling) is checked on the host
    before the device: grid shape against the raster, ranks inside the grid,
    tile rows one tile apart north to south and columns west to east."""
    sc = synthetic.scene_small_plain(0, lines=64, samples=64, out_rows=8, out_cols=8, dem_side=8)
    keep = []
    scene = sc.scene_c(keep)
    img = abi.RadarImage(1, 0, 64, 64, int(sc.sv_t_us[0]) + 10_000_000, 0.0053, float(sc.rps_m))
    cfg = sc.tc_config()
    out = abi.GeocodedImage()
    L = _native.lib()

    def tiling(**kw):
        t = abi.DemTiling()
        t.rows, t.cols, t.tile_height, t.tile_width = 2, 2, 4, 4
        t.origin_lat_deg[0], t.origin_lat_deg[1] = 50.0, 50.0 - 4e-3
        t.origin_lon_deg[0], t.origin_lon_deg[1] = 10.0, 10.0 + 4e-3
        for i in range(64):
            t.rank[i] = -1
        t.rank[0], t.rank[9] = 0, 1
        for k, v in kw.items():
            if k in ("olat1", "olon1"):
                getattr(t, "origin_lat_deg" if k == "olat1" else "origin_lon_deg")[1] = v
            elif k == "rank":
                t.rank[v[0]] = v[1]
            else:
                setattr(t, k, v)
        return t

    for kw, ok in (({}, True), ({"olat1": 50.0 - 4.2e-3}, True), ({"olat1": 50.0 - 4.3e-3}, False),
                   ({"olat1": 50.0 + 4e-3}, False), ({"olon1": 10.0 + 3.7e-3}, False), ({"olon1": float("nan")}, False),
                   ({"rows": 3}, False), ({"tile_width": 5}, False), ({"rank": (2, 0)}, False),
                   ({"rank": (0, 64)}, False)):
        t = tiling(**kw)
        dem = abi.DemRaster(1, 8, 8, 50.0, 10.0, 1e-3)
        dem.tiling = C.pointer(t)
        st = L.sardine_terrain_correction(C.byref(img), C.byref(scene), C.byref(dem), C.byref(cfg), C.byref(out), None)
        # a good tiling passes and the call stops at a later check (no output buffers)
        assert (st == abi.ERR_INVALID_ARGUMENT and "DEM tiling" in _native.last_error()) != ok, \
            (kw, _native.last_error())


def _np_ptr(x):
    return x.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("full", [False, True])
def test_deburst_and_merge_layout_match_oracle(full):
    """The integer deburst/merge layout (deburst.rs:172-305,
    merge_subswaths.rs:207-311), host only: a small scene and the full
    3 x 9 x 1500 x 21000 IW scene of the benchmark (50-line burst overlaps)."""
    if full:
        sc = synthetic.scene_iw_full(1, out_rows=18000, out_cols=25000)
    else:
        sc = synthetic.scene_iw_full(4, n_swaths=3, n_bursts=4, lines_per_burst=500, raw_cols=1500,
                                     overlap_lines=77, swath_overlap=100)
    nb = sc.swaths[0].n_bursts
    keep = []
    swc = sc.swath_inputs_c([0, 0, 0], keep)
    info = api.merged_layout(swc, 3)
    for i, sw in enumerate(sc.swaths):
        b = sw.bursts_c()
        s = abi.SubSwath(sw.n_bursts, sw.lines_per_burst, sc.ati, sw.slant_range_time_s, sc.rps_m)
        st, Lo = O.deburst_layout(s, b, sw.n_bursts, sw.tiff_height)
        Ln = api.deburst_layout(s, b, sw.n_bursts, sw.tiff_height)
        assert st == 0
        for f in ("lines", "samples", "valid_sample_offset", "n_bursts"):
            assert getattr(Lo, f) == getattr(Ln, f)
        assert list(Lo.overlap)[:nb - 1] == list(Ln.overlap)[:nb - 1]
        assert list(Lo.local_first)[:nb] == list(Ln.local_first)[:nb]
        assert list(Lo.local_end)[:nb] == list(Ln.local_end)[:nb]
        assert info.swath_lines[i] == Lo.lines
    # merge offsets/seams from the oracle's merge on dummy arrays of the same shapes
    sig = [np.zeros((int(info.swath_lines[i]), int(info.swath_samples[i])), np.float32) for i in range(3)]
    st, mi, _, _ = O.merge(sig, sig, [int(info.swath_valid_offset[i]) for i in range(3)],
                           [sw.slant_range_time_s for sw in sc.swaths], [sc.rps_m] * 3,
                           [int(sw.burst_times_us[0]) for sw in sc.swaths], query_only=True)
    assert st == 0
    assert (mi.lines, mi.samples) == (info.lines, info.samples)
    assert list(mi.sw_out_offset)[:3] == list(info.sw_out_offset)[:3]
    assert list(mi.seam)[:2] == list(info.seam)[:2]
    assert mi.near_slant_range_time_s == info.near_slant_range_time_s
    assert mi.azimuth_start_time_us == info.azimuth_start_time_us
    if full:
        assert (info.lines, info.samples) == (13100, 60814)
        assert all(int(x) == 50 for x in list(Lo.overlap)[:nb - 1])


def test_deburst_errors_through_abi():
    sw = abi.SubSwath(2, 1000, 2e-3, 0.005, 2.33)
    b = (abi.Burst * 2)(abi.Burst(0, 0, 0, 100, 0), abi.Burst(100_000, 1000, 0, 100, 1))
    with pytest.raises(DeburstError) as e:
        api.deburst_layout(sw, b, 2, 2000)
    assert e.value.status == abi.ERR_DEBURST_IMPLAUSIBLE_OVERLAP
    with pytest.raises(Exception) as e:
        api.deburst_layout(sw, b, 0, 2000)
    assert e.value.status == abi.ERR_DEBURST_NO_BURSTS


def test_speckle_argument_errors_precede_device_check():
    L = _native.lib()
    d = np.zeros(25, np.float32)
    cases = [((abi.SPECKLE_BOXCAR, 4, 1.0, 1.0), 25, abi.ERR_SPECKLE_BAD_WINDOW),
             ((abi.SPECKLE_LEE, 3, 0.0, 1.0), 25, abi.ERR_SPECKLE_BAD_ENL),
             ((abi.SPECKLE_REFINED_LEE, 0, float("nan"), 1.0), 25, abi.ERR_SPECKLE_BAD_ENL),
             ((abi.SPECKLE_FROST, 3, 1.0, -1.0), 25, abi.ERR_SPECKLE_BAD_DAMPING),
             ((abi.SPECKLE_BOXCAR, 3, 1.0, 1.0), 24, abi.ERR_SPECKLE_DIMENSION)]
    for (k, w, e, dmp), n, exp in cases:
        f = abi.SpeckleFilterC(k, w, e, dmp)
        st = L.sardine_speckle_apply(d.ctypes.data, n, 5, 5, C.byref(f), d.ctypes.data, None)
        assert st == exp, (k, w, st)


def test_compute_without_device_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = _native.lib()
    f = abi.SpeckleFilterC(abi.SPECKLE_BOXCAR, 3, 1.0, 1.0)
    a = np.zeros(25, np.float32)
    b = np.zeros(25, np.float32)
    st = L.sardine_speckle_apply(a.ctypes.data, 25, 5, 5, C.byref(f), b.ctypes.data, None)
    assert st == abi.ERR_NO_DEVICE
    assert "no CPU fallback" in _native.last_error()
    counts = (C.c_uint64 * 2)()
    assert L.sardine_to_db_inplace(a.ctypes.data, 25, 0.0, counts, None) == abi.ERR_NO_DEVICE


def test_tc_config_validation_errors():
    sc = synthetic.scene_small_plain(0, lines=64, samples=64, out_rows=8, out_cols=8, dem_side=8)
    scene = api.SceneInputs.from_synthetic(sc)
    cfg = api.TerrainCorrectionConfig(pixel_spacing_deg=-1.0)
    with pytest.raises(Exception) as e:
        api.output_grid(scene, cfg)
    assert e.value.status == abi.ERR_TC_CONFIG
    L = _native.lib()
    keep = []
    s = sc.scene_c(keep)
    s.n_grid_points = 0
    c = sc.tc_config()
    img = abi.GeocodedImage()
    st = L.sardine_terrain_correction(C.byref(abi.RadarImage()), C.byref(s), C.byref(abi.DemRaster()), C.byref(c),
                                      C.byref(img), None)
    assert st == abi.ERR_TC_NO_GRID_POINTS
    s = sc.scene_c(keep)
    s.n_subswaths = 0
    st = L.sardine_terrain_correction(C.byref(abi.RadarImage()), C.byref(s), C.byref(abi.DemRaster()), C.byref(c),
                                      C.byref(img), None)
    assert st == abi.ERR_TC_NO_SUBSWATHS


# ── COG export: the host half (layout, header bytes, validation order) ──
COG_GT = [8.0, 0.0001, 0.0, 52.0, 0.0, -0.0001]


@pytest.mark.parametrize("cols,rows,ts,crs,zone", [
    (1024, 512, 512, abi.CRS_WGS84, 0), (300, 200, 64, abi.CRS_UTM_NORTH, 33), (129, 130, 64, abi.CRS_UTM_SOUTH, 1),
    (1, 1, 64, abi.CRS_WGS84, 0), (0, 0, 64, abi.CRS_WGS84, 0), (0, 500, 128, abi.CRS_WGS84, 0),
    (25000, 1, 4096, abi.CRS_UTM_NORTH, 60), (32768, 32767, 512, abi.CRS_WGS84, 0),
    (32768, 32768, 512, abi.CRS_WGS84, 0), (100000, 70001, 256, abi.CRS_UTM_SOUTH, 58)])
def test_cog_header_matches_oracle(cols, rows, ts, crs, zone):
    """Layout and header bytes (both classic TIFF and BigTIFF) are host code:
    they are compared with the oracle without a device."""
    L = api.cog_layout(cols, rows, ts, crs)
    gt = [500000.0, 10.0, 0.0, 5.2e6, 0.0, -10.0] if crs else COG_GT
    head = api.cog_header(L, gt, crs, zone)
    st, ref = O.cog_header(cols, rows, gt, crs, zone, ts)
    assert st == 0 and head == ref
    assert L.bigtiff == int(rows * cols * 4 > 0xFFFFFFFF)
    assert L.level[L.n_levels - 1].cols == cols and L.level[L.n_levels - 1].rows == rows
    n_tiles = sum(L.level[k].tiles_x * L.level[k].tiles_y for k in range(L.n_levels))
    assert L.tile_bytes == n_tiles * ts * ts * 4 and L.file_bytes == L.header_bytes + L.tile_bytes
    assert L.level[0].tile_data_offset == L.header_bytes == len(head)


def test_cog_argument_errors_precede_device_check(tmp_path):
    L = _native.lib()
    d = np.ones(4, np.float32)
    gt = (C.c_double * 6)(*COG_GT)
    bad_gt = (C.c_double * 6)(8.0, 0.0, 0.0, 52.0, 0.0, -0.0001)
    p = str(tmp_path / "x.tiff").encode()
    for ts in (100, 32, 8192, 0):
        assert L.sardine_write_cog_with_crs(p, d.ctypes.data, 2, 2, gt, 0, 0, ts, None) == \
            abi.ERR_EXPORT_INVALID_TILE_SIZE
    assert L.sardine_write_cog_with_crs(p, d.ctypes.data, 2, 2, bad_gt, 0, 0, 100, None) == \
        abi.ERR_EXPORT_INVALID_GEOTRANSFORM
    assert L.sardine_write_cog_with_crs(p, d.ctypes.data, 2, 2, gt, abi.CRS_UTM_NORTH, 61, 512, None) == \
        abi.ERR_INVALID_ARGUMENT
    assert L.sardine_write_cog_with_crs(p, d.ctypes.data, 2, 2, gt, 7, 0, 512, None) == abi.ERR_INVALID_ARGUMENT
    lay = abi.CogLayout()
    assert L.sardine_cog_layout_compute(2, 2, 96, 0, C.byref(lay)) == abi.ERR_EXPORT_INVALID_TILE_SIZE
    assert L.sardine_cog_layout_compute(-1, 2, 64, 0, C.byref(lay)) == abi.ERR_INVALID_ARGUMENT
    # a layout edited after it was computed is refused
    lay = api.cog_layout(300, 200, 64)
    lay.tile_bytes += 4
    buf = (C.c_uint8 * 4096)()
    assert L.sardine_cog_header(C.byref(lay), gt, 0, 0, buf, 4096) == abi.ERR_INVALID_ARGUMENT
    lay = api.cog_layout(300, 200, 64)
    assert L.sardine_cog_header(C.byref(lay), gt, abi.CRS_UTM_NORTH, 33, buf, 4096) == abi.ERR_INVALID_ARGUMENT
    assert L.sardine_cog_header(C.byref(lay), gt, 0, 0, buf, 10) == abi.ERR_INVALID_ARGUMENT
    import torch

    if not torch.cuda.is_available():
        assert L.sardine_write_cog_with_crs(p, d.ctypes.data, 2, 2, gt, 0, 0, 64, None) == abi.ERR_NO_DEVICE
        assert L.sardine_cog_pack(d.ctypes.data, C.byref(lay), d.ctypes.data, None) == abi.ERR_NO_DEVICE
        assert not (tmp_path / "x.tiff").exists()


@pytest.mark.parametrize("cols,rows,bpp,crs,zone", [
    (2, 2, 4, abi.CRS_WGS84, 0), (5, 7, 4, abi.CRS_UTM_NORTH, 32), (7, 65, 1, abi.CRS_UTM_SOUTH, 1),
    (33, 200, 4, None, 0), (0, 3, 4, abi.CRS_WGS84, 0), (1, 1, 1, None, 0),
    (33_000, 33_000, 4, abi.CRS_WGS84, 0), (66_000, 66_000, 1, abi.CRS_UTM_NORTH, 60), (33_000, 33_000, 1, None, 0),
    (100_000, 20_000, 4, None, 0), (1_000_000, 64, 4, abi.CRS_WGS84, 0), (16_777_216, 64, 4, abi.CRS_WGS84, 0),
    (16_777_000, 1, 4, abi.CRS_UTM_SOUTH, 7)])
def test_geotiff_header_matches_oracle(cols, rows, bpp, crs, zone):
    """The striped writers' header bytes (classic and BigTIFF, float32 and
    uint8, with and without GeoTIFF tags) are host code, compared with the
    oracle without a device."""
    gt = None if crs is None else ([500000.0, 10.0, 0.0, 5.2e6, 0.0, -10.0] if crs else COG_GT)
    head = api.geotiff_header(cols, rows, bpp, gt, crs or 0, zone)
    st, ref = O.geotiff_header(cols, rows, bpp, gt, crs or 0, zone)
    assert st == 0 and head == ref
    big = O.needs_bigtiff(rows, cols, bpp, 0 if crs is None else (16 if crs == 0 else 20))
    assert head[2] == (43 if big else 42)


def test_geotiff_argument_errors_precede_device_check(tmp_path):
    L = _native.lib()
    d = np.ones(4, np.float32)
    gt = (C.c_double * 6)(*COG_GT)
    p = str(tmp_path / "g.tif").encode()
    for bad in ((8.0, 0.0, 0.0, 51.0, 0.0, -1e-4), (8.0, 1e-4, 0.0, 51.0, 0.0, 1e-4)):
        bgt = (C.c_double * 6)(*bad)
        assert L.sardine_write_geotiff_with_crs(p, d.ctypes.data, 2, 2, bgt, 0, 0, None) == \
            abi.ERR_EXPORT_INVALID_GEOTRANSFORM
        assert L.sardine_write_geotiff_u8_with_crs(p, d.ctypes.data, 2, 2, bgt, 0, 0, None) == \
            abi.ERR_EXPORT_INVALID_GEOTRANSFORM
    assert L.sardine_write_geotiff_with_crs(p, d.ctypes.data, 2, 0, gt, 0, 0, None) == abi.ERR_INVALID_ARGUMENT
    assert L.sardine_write_geotiff_with_crs(p, d.ctypes.data, 2, 2, gt, abi.CRS_UTM_SOUTH, 0, None) == \
        abi.ERR_INVALID_ARGUMENT
    n = C.c_uint64()
    assert L.sardine_geotiff_header(2, 2, 2, gt, 0, 0, None, 0, C.byref(n)) == abi.ERR_INVALID_ARGUMENT
    assert L.sardine_geotiff_header(2, 2, 4, gt, 0, 0, None, 0, C.byref(n)) == 0 and n.value == 334
    buf = (C.c_uint8 * 16)()
    assert L.sardine_geotiff_header(2, 2, 4, gt, 0, 0, buf, 16, C.byref(n)) == abi.ERR_INVALID_ARGUMENT
    import torch

    if not torch.cuda.is_available():
        assert L.sardine_write_geotiff_with_crs(p, d.ctypes.data, 2, 2, gt, 0, 0, None) == abi.ERR_NO_DEVICE
        assert L.sardine_write_tiff_no_crs(p, d.ctypes.data, 2, 2, None) == abi.ERR_NO_DEVICE
        assert not (tmp_path / "g.tif").exists()


def _ml_info(lines, samples):
    i = abi.MergedInfo()
    i.lines, i.samples = lines, samples
    i.near_slant_range_time_s = 0.0053
    i.range_pixel_spacing_m = 2.329562
    i.range_pixel_spacing_s = 2.329562 / 299_792_458.0 * 2.0
    i.azimuth_start_time_us = 1_548_221_628_000_000
    i.n_swaths = 3
    i.seam[0] = 17
    return i


@pytest.mark.parametrize("r,a,ati", [(1, 1, 0.0), (4, 1, 0.0), (3, 2, 2.055e-3), (1, 5, 1e-3), (7, 7, 2.2e-3)])
def test_multilook_geometry_matches_oracle(r, a, ati):
    """The shape query (null output rasters) is host code: its geometry equals
    the oracle's apply_multilook."""
    L = _native.lib()
    src = _ml_info(50, 90)
    out = abi.MergedInfo()
    assert L.sardine_apply_multilook(None, None, C.byref(src), r, a, ati, None, None, C.byref(out), None) == 0
    z = np.zeros((50, 90), np.float32)
    st, _, _, ref = O.multilook(z, z, src, r, a, ati)
    assert st == 0
    assert bytes(out) == bytes(ref)
    assert out.seam[0] == 17 and out.n_swaths == 3


def test_multilook_argument_errors_precede_device_check():
    L = _native.lib()
    src = _ml_info(2, 4)
    out = abi.MergedInfo()
    q = lambda r, a, ati: L.sardine_apply_multilook(None, None, C.byref(src), r, a, ati, None, None, C.byref(out), None)
    assert q(0, 1, 0.0) == abi.ERR_INVALID_ARGUMENT and "factors must be" in _native.last_error()
    assert q(8, 1, 0.0) == abi.ERR_INVALID_ARGUMENT and "larger than" in _native.last_error()
    assert q(1, 2, 0.0) == abi.ERR_INVALID_ARGUMENT and q(1, 2, float("nan")) == abi.ERR_INVALID_ARGUMENT
    assert q(-1, 2, 1e-3) == abi.ERR_INVALID_ARGUMENT
    import torch

    if not torch.cuda.is_available():
        d = np.zeros(8, np.float32)
        assert L.sardine_apply_multilook(d.ctypes.data, d.ctypes.data, C.byref(src), 2, 1, 0.0, d.ctypes.data,
                                         d.ctypes.data, C.byref(out), None) == abi.ERR_NO_DEVICE


def _grd_info(lines, samples):
    i = abi.MergedInfo()
    i.lines, i.samples = lines, samples
    i.near_slant_range_time_s = 5.0e-3
    i.range_pixel_spacing_s = 2.33 * 2.0 / 299_792_458.0
    i.range_pixel_spacing_m = 2.33
    return i


def _grd_grids(samples, n_sw=3, lines=(0, 50, 99)):
    rps = 2.33 * 2.0 / 299_792_458.0
    per = samples // n_sw
    grids = []
    for k in range(n_sw):
        c0, c1 = k * per - (8 if k else 0), (k + 1) * per - 1 if k < n_sw - 1 else samples - 1
        grids.append([(ln, 5.0e-3 + c * rps, 30.0 + 16.0 * c / samples + 0.01 * k)
                      for ln in lines for c in np.linspace(c0, c1, 6)])
    return grids


@pytest.mark.parametrize("samples,target,az", [(200, 10.0, 14.08), (3001, 10.0, 3.4), (999, 25.0, 2.0),
                                               (5000, 12.5, 14.0)])
def test_ground_range_shape_matches_oracle(samples, target, az):
    """The shape query (out = NULL) is host code: output size, looks and
    spacings equal the oracle's."""
    L = _native.lib()
    info = _grd_info(100, samples)
    grids = _grd_grids(samples)
    pts, counts = O.grd_points(grids)
    img = abi.GrdImage()
    assert L.sardine_to_ground_range(None, C.byref(info), pts, counts.ctypes.data, len(grids), target, az, None,
                                     C.byref(img), None) == 0
    st, _, ref = O.to_ground_range(np.zeros((100, samples), np.float32), info, grids, target, az, shape_only=True)
    assert st == 0 and bytes(img) == bytes(ref)


def test_ground_range_errors_match_oracle():
    L = _native.lib()
    grids = _grd_grids(100)
    pts, counts = O.grd_points(grids)
    img = abi.GrdImage()

    def both(info, g, target, az):
        p, c = O.grd_points(g)
        st = L.sardine_to_ground_range(None, C.byref(info), p, c.ctypes.data, len(g), target, az, None,
                                       C.byref(img), None)
        ref = O.to_ground_range(np.zeros((max(info.lines, 0), max(info.samples, 0)), np.float32), info, g, target,
                                az, shape_only=True)[0]
        assert st == ref, (st, ref)
        return st

    assert both(_grd_info(0, 0), grids, 10.0, 14.08) == abi.ERR_GRD_EMPTY_IMAGE
    assert both(_grd_info(10, 100), grids, -10.0, 14.08) == abi.ERR_GRD_INVALID_SPACING
    assert both(_grd_info(10, 100), grids, float("nan"), 14.08) == abi.ERR_GRD_INVALID_SPACING
    for az in (0.0, -1.0, float("nan"), float("inf")):
        assert both(_grd_info(10, 100), grids, 10.0, az) == abi.ERR_GRD_INVALID_AZIMUTH_SPACING
    assert both(_grd_info(10, 100), [], 10.0, 14.08) == abi.ERR_GRD_INTERPOLATION_FAILED
    assert both(_grd_info(10, 100), [grids[0], []], 10.0, 14.08) == abi.ERR_GRD_NO_GRID_POINTS
    assert both(_grd_info(10, 100), [grids[0]], 10.0, 14.08) == abi.ERR_GRD_INTERPOLATION_FAILED
    assert both(_grd_info(10, 100), grids, 1e6, 14.08) == abi.ERR_GRD_INVALID_SPACING
    import torch

    if not torch.cuda.is_available():
        d = np.zeros(1000, np.float32)
        assert L.sardine_to_ground_range(d.ctypes.data, C.byref(_grd_info(10, 100)), pts, counts.ctypes.data, 3,
                                         10.0, 14.08, d.ctypes.data, C.byref(img), None) == abi.ERR_NO_DEVICE


@pytest.mark.parametrize("case", ["ok", "be", "short", "single", "gap", "bigtiff", "magic", "comp", "fmt", "bps",
                                  "rps", "count", "noncontig", "type_rational", "drop256", "drop257", "drop258",
                                  "drop259", "drop278", "drop339", "drop273", "drop277", "text", "empty", "missing",
                                  "be_short_single"])
def test_slc_open_matches_oracle(tmp_path, case):
    """SlcReader::open is host code: status and layout equal the oracle's."""
    import slc_fixtures as F
    d = np.random.default_rng(5).integers(-32768, 32768, (1 if "single" in case else 12, 9, 2)).astype(np.int16)
    kw = {"be": dict(le=False), "short": dict(offsets_type=3), "gap": dict(data_gap=40), "bigtiff": dict(magic=43),
          "magic": dict(magic=41), "comp": dict(overrides={259: (3, 5)}), "fmt": dict(overrides={339: (3, 1)}),
          "bps": dict(overrides={258: (3, 16)}), "rps": dict(overrides={278: (4, 2)}), "count": dict(strip_count=11),
          "noncontig": dict(gap_last=4), "type_rational": dict(overrides={256: (5, 9)}),
          "be_short_single": dict(le=False, offsets_type=3)}.get(case, {})
    if case.startswith("drop"):
        kw = dict(drop=(int(case[4:]),))
    p = tmp_path / "c.tiff"
    if case == "text":
        p.write_bytes(b"This is not a TIFF file at all!!!")
    elif case == "empty":
        p.write_bytes(b"")
    elif case != "missing":
        p.write_bytes(F.slc_tiff_bytes(d, **kw))
    t = abi.SlcTiff()
    st = _native.lib().sardine_slc_open(str(p).encode(), C.byref(t))
    ref_st, ref, _ = O.slc_open(p)
    assert st == ref_st, (st, ref_st, _native.last_error())
    if st == 0:
        assert bytes(t) == bytes(ref)


def test_slc_read_errors_precede_device_check(tmp_path):
    import slc_fixtures as F
    d = np.zeros((12, 9, 2), np.int16)
    F.write_slc_tiff(tmp_path / "s.tiff", d)
    L = _native.lib()
    t = abi.SlcTiff()
    assert L.sardine_slc_open(str(tmp_path / "s.tiff").encode(), C.byref(t)) == 0
    p = str(tmp_path / "s.tiff").encode()
    buf = np.zeros(12 * 9 * 2, np.int16)
    assert L.sardine_slc_read_lines(p, C.byref(t), 12, 1, buf.ctypes.data, None) == abi.ERR_SLC_LINE_OUT_OF_BOUNDS
    assert L.sardine_slc_read_lines(p, C.byref(t), 10, 10, buf.ctypes.data, None) == abi.ERR_SLC_LINE_OUT_OF_BOUNDS
    assert L.sardine_slc_read_lines(p, C.byref(t), -1, 1, buf.ctypes.data, None) == abi.ERR_SLC_LINE_OUT_OF_BOUNDS
    assert L.sardine_slc_read_lines(p, C.byref(t), 3, 0, None, None) == 0  # nothing to read
    import torch

    if not torch.cuda.is_available():
        assert L.sardine_slc_read_lines(p, C.byref(t), 0, 12, buf.ctypes.data, None) == abi.ERR_NO_DEVICE


@pytest.mark.parametrize("case", ["bin96", "bin2008", "short96", "short2008", "missing", "ascii", "ascii_crlf",
                                  "ascii_one_header", "ascii_short_lines", "ascii_bad_token", "ascii_hex",
                                  "ascii_specials", "ascii_no_header", "ascii_extra_value"])
def test_geoid_loaders_match_oracle(tmp_path, case):
    """The grid loaders are host code: status and values equal the oracle's."""
    import test_oracle_reference as T
    g = (np.random.default_rng(4).normal(0, 30, (73, 144)) * 1e5).round() / 1e5
    p = tmp_path / "g"
    model = 1 if "2008" in case else 0
    if case.startswith("bin"):
        p.write_bytes(g.astype("<f4").tobytes())
    elif case.startswith("short"):
        p.write_bytes(b"short")
    elif case == "ascii":
        p.write_text(T._ww15mgh_text(g))
    elif case == "ascii_crlf":
        p.write_text(T._ww15mgh_text(g, eol="\r\n"))
    elif case == "ascii_one_header":
        p.write_text(T._ww15mgh_text(g, header=("-90 90 0 360 2.5 2.5",)))
    elif case == "ascii_short_lines":
        p.write_text(T._ww15mgh_text(g, per_line=6))
    elif case == "ascii_no_header":
        p.write_text(T._ww15mgh_text(g, header=()))
    elif case == "ascii_extra_value":
        p.write_text(T._ww15mgh_text(g) + "1.0\n")
    elif case in ("ascii_bad_token", "ascii_hex", "ascii_specials"):
        lines = T._ww15mgh_text(g).split("\n")
        first = lines[7].split()[0]
        lines[7] = lines[7].replace(first, {"ascii_bad_token": "1.2.3", "ascii_hex": "0x1p3",
                                            "ascii_specials": "-inf"}[case], 1)
        lines[8] = lines[8].replace(lines[8].split()[0], "NaN" if case == "ascii_specials" else "1e", 1)
        p.write_text("\n".join(lines))
    ascii_ = case.startswith("ascii")
    out = np.zeros(73 * 144, np.float32)
    L = _native.lib()
    st = (L.sardine_geoid_load_ww15mgh(str(p).encode(), out.ctypes.data) if ascii_ else
          L.sardine_geoid_load_binary(str(p).encode(), model, out.ctypes.data))
    ref_st, ref = O.geoid_load(p, ascii=ascii_, model=model)
    assert st == ref_st, (st, ref_st, _native.last_error())
    if st == 0:
        assert np.array_equal(out.view(np.uint32), ref.ravel().view(np.uint32))


def test_srtm_mosaic_layout_and_errors(tmp_path):
    L = _native.lib()

    def q(names, dst=None):
        arr = (C.c_char_p * max(len(names), 1))(*[str(tmp_path / n).encode() for n in names])
        r = abi.DemRaster()
        return L.sardine_dem_srtm_mosaic(arr, len(names), dst, C.byref(r), None), r

    st, r = q(["N50E010.hgt", "N51E011.hgt"])
    assert st == 0 and (r.width, r.height) == (7201, 7201)
    assert (r.origin_lat_deg, r.origin_lon_deg, r.pixel_deg) == (52.0, 10.0, 1.0 / 3600.0)
    st, r = q(["s05w120.hgt", "S06W120.hgt", "s05W119_extra.hgt"])
    assert st == 0 and (r.width, r.height, r.origin_lat_deg, r.origin_lon_deg) == (7201, 7201, -4.0, -120.0)
    assert q([])[0] == abi.ERR_DEM_NO_TILES
    for bad in ("X50E010.hgt", "N50E01.hgt", "N5xE010.hgt", "N50Q010.hgt"):
        assert q([bad])[0] == abi.ERR_DEM_UNRECOGNISED_TILE_NAME, bad
    # device path: files are checked in order (name, then the file) before any device work
    import torch

    buf = np.zeros(4, np.float32)
    (tmp_path / "N50E010.hgt").write_bytes(b"\0" * 10)
    assert q(["N50E010.hgt", "X.hgt"], buf.ctypes.data)[0] == abi.ERR_DEM_BAD_SIZE
    assert q(["N51E010.hgt"], buf.ctypes.data)[0] == abi.ERR_DEM_IO
    if not torch.cuda.is_available():
        (tmp_path / "N52E010.hgt").write_bytes(b"\0" * (3601 * 3601 * 2))
        assert q(["N52E010.hgt"], buf.ctypes.data)[0] == abi.ERR_NO_DEVICE


def test_multi_slc_errors_precede_device_check(tmp_path):
    import slc_fixtures as F
    d = np.zeros((30, 11, 2), np.int16)
    F.write_slc_tiff(tmp_path / "a.tiff", d[:18])
    F.write_slc_tiff(tmp_path / "b.tiff", d[18:])
    F.write_slc_tiff(tmp_path / "w.tiff", np.zeros((5, 12, 2), np.int16))
    ra, rb, rw = (api.SlcReader.open(tmp_path / n) for n in ("a.tiff", "b.tiff", "w.tiff"))
    m = api.MultiSlcReader.from_readers([ra, rb])
    assert (m.width, m.height) == (11, 30)
    from sardine_amd import SlcReadError
    with pytest.raises(SlcReadError) as e:
        api.MultiSlcReader.from_readers([ra, rw])
    assert e.value.status == abi.ERR_SLC_MULTI_WIDTH_MISMATCH
    L = _native.lib()
    buf = np.zeros(30 * 11 * 2, np.int16)
    ts = (abi.SlcTiff * 2)(ra.tiff, rb.tiff)
    ps = (C.c_char_p * 2)(ra.path.encode(), rb.path.encode())
    assert L.sardine_slc_read_lines_multi(ps, ts, 0, 0, 1, buf.ctypes.data, None) == abi.ERR_SLC_MULTI_EMPTY_INPUT
    assert L.sardine_slc_read_lines_multi(ps, ts, 2, 15, 5, buf.ctypes.data, None) == \
        abi.ERR_SLC_CROSS_SLICE_BOUNDARY
    assert L.sardine_slc_read_lines_multi(ps, ts, 2, 25, 6, buf.ctypes.data, None) == abi.ERR_SLC_LINE_OUT_OF_BOUNDS
    import torch

    if not torch.cuda.is_available():
        assert L.sardine_slc_read_lines_multi(ps, ts, 2, 20, 10, buf.ctypes.data, None) == abi.ERR_NO_DEVICE


def test_glo30_tile_georeferencing_and_errors(tmp_path):
    """Glo30Tile::load's georeferencing (dem.rs:262-303) and error order, host
    side (decode errors surface before the device check)."""
    import slc_fixtures as F
    L = _native.lib()
    a = np.full((36, 36), 100.0, np.float32)

    def head(name, **kw):
        p = tmp_path / name
        p.write_bytes(F.geotiff_dem_bytes(a, **kw))
        r = abi.DemRaster()
        return L.sardine_dem_glo30_tile(str(p).encode(), None, C.byref(r), None), r

    # tiepoint at the NW corner, pixel-is-area: the first pixel centre is half a pixel inside
    st, r = head("t.tif", scale=(1 / 3600, 1 / 3600, 0), tiepoint=(0, 0, 0, 10.0, 51.0, 0))
    assert st == 0 and (r.width, r.height) == (36, 36)
    assert r.origin_lon_deg == 10.0 + 0.5 * (1 / 3600) and r.origin_lat_deg == 51.0 - 0.5 * (1 / 3600)
    assert r.pixel_deg == 1 / 3600
    # a tiepoint at a pixel centre (I = J = 0.5) and unequal scales: their mean
    st, r = head("t2.tif", scale=(2e-4, 4e-4, 0), tiepoint=(0.5, 0.5, 0, 10.0, 51.0, 0))
    assert st == 0 and (r.origin_lon_deg, r.origin_lat_deg) == (10.0, 51.0) and r.pixel_deg == (2e-4 + 4e-4) / 2
    # without GeoTIFF tags: the origin from the name, the pixel from the width
    for name in ("Copernicus_DSM_COG_10_N50_00_E010_00_DEM.tif", "N50E010.tif", "s05_00_w120_00.tif"):
        st, r = head(name)
        la, lo = (50, 10) if "N50" in name else (-5, -120)
        if name.startswith("s05"):
            assert st == abi.ERR_DEM_UNRECOGNISED_TILE_NAME  # lower-case n/s are not Copernicus tokens
            continue
        assert st == 0 and r.pixel_deg == 1 / 36
        assert r.origin_lat_deg == (la + 1) - 0.5 / 36 and r.origin_lon_deg == lo + 0.5 / 36
    assert head("Copernicus_DSM_COG_10_N50_01_E010_00_DEM.tif")[0] == abi.ERR_DEM_UNRECOGNISED_TILE_NAME
    r = abi.DemRaster()
    assert L.sardine_dem_glo30_tile(str(tmp_path / "missing.tif").encode(), None, C.byref(r), None) == abi.ERR_DEM_IO
    (tmp_path / "junk.tif").write_bytes(b"not a tiff at all")
    assert L.sardine_dem_glo30_tile(str(tmp_path / "junk.tif").encode(), None, C.byref(r), None) == \
        abi.ERR_DEM_TIFF_DECODE
    buf = np.zeros(4, np.float32)
    # decode errors and pixel types come before the device check
    p = tmp_path / "u8.tif"
    p.write_bytes(F.geotiff_dem_bytes(np.zeros((8, 8), np.uint8), scale=(1, 1, 0), tiepoint=(0, 0, 0, 0, 0, 0)))
    assert L.sardine_dem_glo30_tile(str(p).encode(), buf.ctypes.data, C.byref(r), None) == \
        abi.ERR_DEM_UNSUPPORTED_PIXEL_TYPE
    b = bytearray(F.geotiff_dem_bytes(a, compression="deflate", scale=(1, 1, 0), tiepoint=(0, 0, 0, 0, 0, 0)))
    b[-40:] = b"\xff" * 40  # a corrupt deflate stream
    (tmp_path / "bad.tif").write_bytes(bytes(b))
    assert L.sardine_dem_glo30_tile(str(tmp_path / "bad.tif").encode(), buf.ctypes.data, C.byref(r), None) == \
        abi.ERR_DEM_TIFF_DECODE


def _glo30_tiles(tmp_path, cells, px=0.25, n=4, shape=None):
    """GLO-30 tiles of n x n pixels of px degrees with pixel-corner tiepoints
    at (lat_n, lon_w); cells: {name: (lat_n, lon_w)}."""
    import slc_fixtures as F
    paths = {}
    for k, (name, (la, lo)) in enumerate(cells.items()):
        a = np.full(shape or (n, n), 10.0 * (k + 1), np.float32)
        p = tmp_path / name
        p.write_bytes(F.geotiff_dem_bytes(a, scale=(px, px, 0), tiepoint=(0, 0, 0, lo, la, 0)))
        paths[name] = str(p)
    return paths


def test_glo30_mosaic_layout_and_errors(tmp_path):
    """DemMosaic::load_glo30_files as one raster (host side, headers only):
    tile rows north first and columns west first at each tile's own origin,
    ranks = load order, empty cells; and the grid errors."""
    L = _native.lib()
    paths = _glo30_tiles(tmp_path, {"a.tif": (47.0, 8.0), "b.tif": (47.0, 7.0), "c.tif": (46.0, 8.0)})

    def load(names):
        arr = (C.c_char_p * len(names))(*[paths[x].encode() for x in names]) if names else None
        r, t = abi.DemRaster(), abi.DemTiling()
        st = L.sardine_dem_glo30_mosaic(arr, len(names), None, C.byref(r), C.byref(t), None)
        return st, r, t

    st, r, t = load(["a.tif", "b.tif", "c.tif"])
    assert st == 0
    assert (t.rows, t.cols, t.tile_height, t.tile_width) == (2, 2, 4, 4) and (r.height, r.width) == (8, 8)
    assert list(t.origin_lat_deg[:2]) == [47.0 - 0.125, 46.0 - 0.125]
    assert list(t.origin_lon_deg[:2]) == [7.0 + 0.125, 8.0 + 0.125]
    assert (r.origin_lat_deg, r.origin_lon_deg, r.pixel_deg) == (47.0 - 0.125, 7.0 + 0.125, 0.25)
    assert [t.rank[0], t.rank[1], t.rank[8], t.rank[9]] == [1, 0, -1, 2]  # b at NW, a at NE, none at SW, c at SE
    assert all(t.rank[i] == -1 for i in range(64) if i not in (0, 1, 8, 9))
    assert load([])[0] == abi.ERR_DEM_NO_TILES
    # tiles of another size, a gap between columns, two tiles in one cell, a missing file first
    paths.update(_glo30_tiles(tmp_path, {"big.tif": (46.0, 7.0)}, shape=(5, 4)))
    paths.update(_glo30_tiles(tmp_path, {"gap.tif": (47.0, 10.0)}))
    paths.update(_glo30_tiles(tmp_path, {"dup.tif": (47.0, 8.0)}))
    assert load(["a.tif", "big.tif"])[0] == abi.ERR_INVALID_ARGUMENT
    assert load(["a.tif", "gap.tif"])[0] == abi.ERR_INVALID_ARGUMENT
    assert load(["a.tif", "dup.tif"])[0] == abi.ERR_INVALID_ARGUMENT
    paths["nofile.tif"] = str(tmp_path / "nofile.tif")
    assert load(["nofile.tif", "big.tif"])[0] == abi.ERR_DEM_IO
    # the tiles load whole in path order: a bad block in the first tile is
    # reported before the second tile's missing file, and after it when second
    import slc_fixtures as F
    b = bytearray(F.geotiff_dem_bytes(np.zeros((64, 64), np.float32), scale=(1 / 64, 1 / 64, 0),
                                      tiepoint=(0, 0, 0, 5.0, 47.0, 0), tile=(16, 16), compression="deflate"))
    _patch_tile(b, 3, corrupt=True)
    paths["bad.tif"] = str(tmp_path / "bad.tif")
    (tmp_path / "bad.tif").write_bytes(bytes(b))
    assert load(["bad.tif"])[0] == 0  # the header is good: the layout query does not decode
    assert load(["bad.tif", "nofile.tif"])[0] == abi.ERR_DEM_TIFF_DECODE
    assert "corrupt compressed block" in _native.last_error()
    assert load(["nofile.tif", "bad.tif"])[0] == abi.ERR_DEM_IO
    # nine tile columns: more than the grid holds
    nine = _glo30_tiles(tmp_path, {f"w{i}.tif": (47.0, 20.0 + i) for i in range(9)})
    paths.update(nine)
    assert load(list(nine))[0] == abi.ERR_INVALID_ARGUMENT


def _patch_tile(buf: bytearray, k: int, outside: bool = False, corrupt: bool = False) -> None:
    """Block k of a little-endian classic tiled TIFF: its offset moved past
    the end of the file, or its Deflate stream overwritten with garbage."""
    import struct

    import tiffread as T
    ifd = T.parse(bytes(buf))["ifds"][0]
    offs, counts = ifd["tags"][324][1], ifd["tags"][325][1]
    if corrupt:
        buf[offs[k]:offs[k] + counts[k]] = b"\xff" * counts[k]
    if outside:
        n = struct.unpack_from("<H", buf, ifd["offset"])[0]
        for i in range(n):
            e = ifd["offset"] + 2 + 12 * i
            if struct.unpack_from("<H", buf, e)[0] == 324:
                table = struct.unpack_from("<I", buf, e + 8)[0]
                struct.pack_into("<I", buf, table + 4 * k, len(buf) + 4096)


def test_glo30_decode_errors_in_file_order(tmp_path):
    """Blocks decode in parallel, but the error reported is the first failing
    block's in file order, as a sequential decode (the `tiff` crate's)
    reports it; and decode errors come after the georeferencing checks
    (Glo30Tile::load reads the georeferencing before the image)."""
    import slc_fixtures as F
    L = _native.lib()
    a = np.arange(64 * 64, dtype=np.float32).reshape(64, 64)
    good = bytearray(F.geotiff_dem_bytes(a, tile=(16, 16), compression="deflate"))
    buf = np.zeros(4, np.float32)
    r = abi.DemRaster()
    for first, second, want in (("corrupt", "outside", "corrupt compressed block"),
                                ("outside", "corrupt", "block outside the file")):
        b = bytearray(good)
        _patch_tile(b, 2, **{first: True})
        _patch_tile(b, 11, **{second: True})
        p = tmp_path / "N50E010.tif"
        p.write_bytes(bytes(b))
        for _ in range(5):  # the same answer whatever the threads' timing
            assert L.sardine_dem_glo30_tile(str(p).encode(), buf.ctypes.data, C.byref(r), None) == \
                abi.ERR_DEM_TIFF_DECODE
            assert want in _native.last_error()
        # the layout query does not decode: the header is good
        assert L.sardine_dem_glo30_tile(str(p).encode(), None, C.byref(r), None) == 0
        # a bad name without GeoTIFF tags is reported before the bad block
        q = tmp_path / "not_a_tile.tif"
        q.write_bytes(bytes(b))
        assert L.sardine_dem_glo30_tile(str(q).encode(), buf.ctypes.data, C.byref(r), None) == \
            abi.ERR_DEM_UNRECOGNISED_TILE_NAME


def _crafted_tiff(entries, big=False, tail=b""):
    """A little-endian TIFF (classic or BigTIFF) with one IFD of raw
    (tag, type, count, field-bytes) entries, sorted by tag, then `tail`."""
    import struct

    if big:
        out = bytearray(b"II") + struct.pack("<HHHQ", 43, 8, 0, 16) + struct.pack("<Q", len(entries))
        for tag, typ, count, field in sorted(entries):
            out += struct.pack("<HHQ", tag, typ, count) + field.ljust(8, b"\0")
        out += struct.pack("<Q", 0)
    else:
        out = bytearray(b"II") + struct.pack("<HI", 42, 8) + struct.pack("<H", len(entries))
        for tag, typ, count, field in sorted(entries):
            out += struct.pack("<HHI", tag, typ, count) + field.ljust(4, b"\0")
        out += struct.pack("<I", 0)
    return bytes(out + tail)


def _u32(*v):
    import struct
    return struct.pack("<" + "I" * len(v), *v)


def _u16(v):
    import struct
    return struct.pack("<H", v)


@pytest.mark.parametrize("case", ["huge_tiles", "huge_image", "bigtiff_count_wrap", "bigtiff_ifd_wrap",
                                  "huge_strip_count"])
def test_glo30_crafted_headers_are_decode_errors(tmp_path, case):
    """Headers whose sizes overflow 64-bit products or ask for absurd
    allocations are TIFF decode errors (the reference's `tiff` crate refuses
    them with an error), never a crash: the reader checks every count x size,
    block and image product before allocating, and host exceptions do not
    cross the C ABI.  Decode errors come before the device check, so no GPU
    is needed."""
    L = _native.lib()
    base = [(258, 3, 1, _u16(32)), (259, 3, 1, _u16(1)), (277, 3, 1, _u16(1)), (339, 3, 1, _u16(3))]
    if case == "huge_tiles":  # 2^31 x 2^31 tiles of a 16 x 16 image: 2^64 bytes per tile wraps to 0
        ents = base + [(256, 4, 1, _u32(16)), (257, 4, 1, _u32(16)), (322, 4, 1, _u32(1 << 31)),
                       (323, 4, 1, _u32(1 << 31)), (324, 4, 1, _u32(200)), (325, 4, 1, _u32(16))]
        data = _crafted_tiff(ents, tail=b"\0" * 256)
    elif case == "huge_image":  # 2^30 x 2^30 samples of 8 bytes: 2^63 bytes
        ents = [(256, 4, 1, _u32(1 << 30)), (257, 4, 1, _u32(1 << 30)), (258, 3, 1, _u16(64)),
                (259, 3, 1, _u16(1)), (273, 4, 1, _u32(0)), (277, 3, 1, _u16(1)), (278, 4, 1, _u32(1 << 30)),
                (279, 4, 1, _u32(8)), (339, 3, 1, _u16(3))]
        data = _crafted_tiff(ents)
    elif case == "bigtiff_count_wrap":  # ModelPixelScale with 2^61 doubles: count x 8 wraps to 0
        import struct
        ents = base + [(256, 4, 1, _u32(4)), (257, 4, 1, _u32(4)), (273, 4, 1, _u32(0)),
                       (279, 4, 1, _u32(64)), (33550, 12, 1 << 61, struct.pack("<Q", 0))]
        data = _crafted_tiff(ents, big=True)
    elif case == "bigtiff_ifd_wrap":  # an IFD entry count whose x 20 wraps
        import struct
        data = bytes(bytearray(b"II") + struct.pack("<HHHQ", 43, 8, 0, 16) + struct.pack("<Q", (1 << 64) // 20 + 1)
                     + b"\0" * 64)
    else:  # one strip per row of a 2^30-row image, but a 4-entry strip table
        ents = base + [(256, 4, 1, _u32(1 << 30)), (257, 4, 1, _u32(1 << 30)), (273, 4, 4, _u32(200)),
                       (278, 4, 1, _u32(1)), (279, 4, 4, _u32(216))]
        data = _crafted_tiff(ents, tail=b"\0" * 64)
    p = tmp_path / "N50E010.tif"
    p.write_bytes(data)
    r = abi.DemRaster()
    if case == "bigtiff_count_wrap":
        # the unreadable scale counts as absent (no read past the buffer):
        # the georeferencing comes from the name
        assert L.sardine_dem_glo30_tile(str(p).encode(), None, C.byref(r), None) == 0
        assert r.pixel_deg == 1 / 4 and (r.width, r.height) == (4, 4)
        return
    buf = np.zeros(4, np.float32)
    st = L.sardine_dem_glo30_tile(str(p).encode(), buf.ctypes.data, C.byref(r), None)
    assert st == abi.ERR_DEM_TIFF_DECODE, (case, st, _native.last_error())


def test_glo30_mistyped_georeferencing_counts_as_absent(tmp_path):
    """A ModelPixelScale / ModelTiepoint of a non-floating type is not read
    as zeros: the tile falls back to the name's georeferencing."""
    import slc_fixtures as F
    L = _native.lib()
    a = np.full((36, 36), 5.0, np.float32)
    raw = bytearray(F.geotiff_dem_bytes(a, scale=(1.0, 1.0, 0.0), tiepoint=(0, 0, 0, 3.0, 4.0, 0)))
    import tiffread as T
    import struct
    ifd = T.parse(bytes(raw))["ifds"][0]
    n = struct.unpack_from("<H", raw, ifd["offset"])[0]
    for i in range(n):  # retype both tags as SHORT (3)
        e = ifd["offset"] + 2 + 12 * i
        if struct.unpack_from("<H", raw, e)[0] in (33550, 33922):
            struct.pack_into("<H", raw, e + 2, 3)
    p = tmp_path / "N50E010.tif"
    p.write_bytes(bytes(raw))
    r = abi.DemRaster()
    assert L.sardine_dem_glo30_tile(str(p).encode(), None, C.byref(r), None) == 0
    assert r.pixel_deg == 1 / 36 and r.origin_lat_deg == 51 - 0.5 / 36 and r.origin_lon_deg == 10 + 0.5 / 36
