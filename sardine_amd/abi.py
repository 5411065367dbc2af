"""ctypes mirror of include/sardine_hip.h (struct layouts and enums).

Both the native library wrapper (sardine_amd._native) and the test-side oracle
wrapper build their arguments from these definitions, so one layout serves
both sides of every parity comparison.
"""
from __future__ import annotations

import ctypes as C

ABI_VERSION = 3
NEWTON_HIST_SIZE = 16
MAX_BURSTS = 64
MAX_SWATHS = 8
COG_MAX_LEVELS = 32

# sardine_status
OK = 0
ERR_TC_ORBIT_OUTSIDE_WINDOW = 10
ERR_TC_ORBIT_INSUFFICIENT = 11
ERR_TC_NO_GRID_POINTS = 12
ERR_TC_NO_SUBSWATHS = 13
ERR_TC_CONFIG = 14
ERR_TC_PROJECTION = 15
ERR_TC_DEM = 16
ERR_SPECKLE_DIMENSION = 20
ERR_SPECKLE_BAD_WINDOW = 21
ERR_SPECKLE_BAD_ENL = 22
ERR_SPECKLE_BAD_DAMPING = 23
ERR_CAL_NO_CAL_VECTORS = 30
ERR_CAL_NO_NOISE_VECTORS = 31
ERR_CAL_LUT_RANGE = 32
ERR_DEBURST_NO_BURSTS = 40
ERR_DEBURST_COUNT_MISMATCH = 41
ERR_DEBURST_TIFF_TOO_SMALL = 42
ERR_DEBURST_EMPTY_VALID_WINDOW = 43
ERR_DEBURST_INVALID_OVERLAP = 44
ERR_DEBURST_IMPLAUSIBLE_OVERLAP = 45
ERR_MERGE_EMPTY_INPUTS = 50
ERR_MERGE_PIXEL_SPACING = 51
ERR_MERGE_EMPTY_SWATH = 52
ERR_EXPORT_INVALID_GEOTRANSFORM = 60
ERR_EXPORT_INVALID_TILE_SIZE = 61
ERR_EXPORT_IO = 62
ERR_GRD_NO_GRID_POINTS = 70
ERR_GRD_EMPTY_IMAGE = 71
ERR_GRD_INVALID_SPACING = 72
ERR_GRD_INVALID_AZIMUTH_SPACING = 73
ERR_GRD_INTERPOLATION_FAILED = 74
ERR_SLC_IO = 80
ERR_SLC_NOT_TIFF = 81
ERR_SLC_BIGTIFF = 82
ERR_SLC_MISSING_TAG = 83
ERR_SLC_UNSUPPORTED_FORMAT = 84
ERR_SLC_NON_CONTIGUOUS_STRIPS = 85
ERR_SLC_LINE_OUT_OF_BOUNDS = 86
ERR_SLC_CROSS_SLICE_BOUNDARY = 87
ERR_SLC_MULTI_EMPTY_INPUT = 88
ERR_SLC_MULTI_WIDTH_MISMATCH = 89
ERR_GEOID_IO = 100
ERR_GEOID_WRONG_SIZE = 101
ERR_GEOID_WRONG_SIZE_EGM2008 = 102
ERR_GEOID_MALFORMED_ASCII = 103
ERR_DEM_IO = 110
ERR_DEM_BAD_SIZE = 111
ERR_DEM_TIFF_DECODE = 112
ERR_DEM_UNSUPPORTED_PIXEL_TYPE = 113
ERR_DEM_UNRECOGNISED_TILE_NAME = 114
ERR_DEM_NO_TILES = 115
GEOID_MODEL_EGM96, GEOID_MODEL_EGM2008 = 0, 1
ERR_INVALID_ARGUMENT = 90
ERR_NO_DEVICE = 91
ERR_HIP = 92
ERR_OUT_OF_MEMORY = 93
ERR_CAPACITY = 94

STATUS_NAMES = {v: k for k, v in globals().items() if k.startswith("ERR_") or k == "OK"}

# enums
GEOID_ZERO, GEOID_GRID_2P5 = 0, 1
CRS_WGS84, CRS_UTM_NORTH, CRS_UTM_SOUTH = 0, 1, 2
RESAMPLE_BILINEAR, RESAMPLE_BICUBIC, RESAMPLE_LANCZOS3 = 0, 1, 2
PRECISION_FP64, PRECISION_FP32_CENTRED, PRECISION_MIXED = 0, 1, 2
SPECKLE_BOXCAR, SPECKLE_LEE, SPECKLE_FROST, SPECKLE_GAMMA_MAP, SPECKLE_REFINED_LEE = 0, 1, 2, 3, 4

# mask_code (terrain_correction.rs:346-376)
MASK_VALID = 0
MASK_OUTSIDE_FOOTPRINT = 1
MASK_DEM_MISSING = 2
MASK_NON_CONVERGED = 3
MASK_DEGENERATE_GEOMETRY = 4
MASK_SHADOW = 5
MASK_LAYOVER = 6
MASK_FORESHORTENING = 7
MASK_NOISE_MASKED = 8
MASK_DEM_NEIGHBOUR_MISSING = 9


class StateVector(C.Structure):
    _fields_ = [("time_us", C.c_int64), ("position_m", C.c_double * 3), ("velocity_m_s", C.c_double * 3)]


class Orbit(C.Structure):
    _fields_ = [("state_vectors", C.POINTER(StateVector)), ("count", C.c_int32)]


class GridPoint(C.Structure):
    _fields_ = [("latitude_deg", C.c_double), ("azimuth_time_us", C.c_int64)]


class BBox(C.Structure):
    _fields_ = [("min_lat_deg", C.c_double), ("max_lat_deg", C.c_double),
                ("min_lon_deg", C.c_double), ("max_lon_deg", C.c_double)]


class Scene(C.Structure):
    _fields_ = [("orbit", Orbit), ("bounding_box", BBox), ("azimuth_time_interval_s", C.c_double),
                ("grid_points", C.POINTER(GridPoint)), ("n_grid_points", C.c_int32),
                ("n_subswaths", C.c_int32)]


class RadarImage(C.Structure):
    _fields_ = [("data", C.c_void_p), ("nesz", C.c_void_p), ("lines", C.c_int64), ("samples", C.c_int64),
                ("azimuth_start_time_us", C.c_int64), ("near_slant_range_time_s", C.c_double),
                ("range_pixel_spacing_m", C.c_double)]


DEM_MAX_TILES_AXIS = 8


class DemTiling(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("tile_height", C.c_int32), ("tile_width", C.c_int32),
                ("origin_lat_deg", C.c_double * DEM_MAX_TILES_AXIS), ("origin_lon_deg", C.c_double * DEM_MAX_TILES_AXIS),
                ("rank", C.c_int32 * (DEM_MAX_TILES_AXIS * DEM_MAX_TILES_AXIS))]


class DemRaster(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_int64), ("height", C.c_int64),
                ("origin_lat_deg", C.c_double), ("origin_lon_deg", C.c_double), ("pixel_deg", C.c_double),
                ("tiling", C.POINTER(DemTiling))]


class Geoid(C.Structure):
    _fields_ = [("kind", C.c_int32), ("grid_73x144", C.c_void_p)]


class TcConfig(C.Structure):
    _fields_ = [("pixel_spacing", C.c_double), ("crs", C.c_int32), ("utm_zone", C.c_int32),
                ("max_newton_iterations", C.c_int32), ("newton_tolerance_s", C.c_double),
                ("flatten", C.c_int32), ("compute_lia", C.c_int32), ("noise_floor_margin_db", C.c_float),
                ("resampling", C.c_int32), ("projector_anchors", C.c_int32), ("precision", C.c_int32),
                ("geoid", Geoid),
                ("aoi_row0", C.c_int64), ("aoi_col0", C.c_int64), ("aoi_rows", C.c_int64),
                ("aoi_cols", C.c_int64)]


class GeocodedImage(C.Structure):
    _fields_ = [("data", C.c_void_p), ("cos_lia", C.c_void_p), ("mask", C.c_void_p),
                ("radar_coords", C.c_void_p),
                ("rows", C.c_int64), ("cols", C.c_int64), ("full_rows", C.c_int64), ("full_cols", C.c_int64),
                ("geotransform", C.c_double * 6),
                ("total_pixel_count", C.c_uint64), ("valid_pixel_count", C.c_uint64),
                ("dem_missing_count", C.c_uint64), ("outside_footprint_count", C.c_uint64),
                ("non_converged_count", C.c_uint64), ("degenerate_geometry_count", C.c_uint64),
                ("flat_masked_count", C.c_uint64), ("noise_masked_count", C.c_uint64),
                ("newton_iter_histogram", C.c_uint64 * NEWTON_HIST_SIZE),
                ("chain_redo_rounds", C.c_uint64), ("tie_steps", C.c_uint64),
                ("tie_decisions_changed", C.c_uint64), ("tie_max_dev_us", C.c_double)]


class SpeckleFilterC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("window", C.c_int32), ("enl", C.c_float), ("damping", C.c_float)]


class BatchJob(C.Structure):
    _fields_ = [("scene", C.c_int32), ("status", C.c_int32), ("config", TcConfig),
                ("out", C.POINTER(GeocodedImage)), ("speckle", SpeckleFilterC), ("speckle_out", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("segment_columns", C.c_int32), ("tc_groups_per_cu", C.c_int32),
                ("speckle_groups_per_cu", C.c_int32)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch", C.c_int32), ("slots", C.c_int32),
                ("use_graph", C.c_int32), ("pad", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_int32), ("pad", C.c_int32), ("config", TcConfig), ("speckle", SpeckleFilterC)]


class Response(C.Structure):
    _fields_ = [("status", C.c_int32), ("done", C.c_int32), ("image", GeocodedImage), ("speckled", C.c_void_p),
                ("t_submit_us", C.c_double), ("t_admit_us", C.c_double), ("t_launch_us", C.c_double),
                ("t_done_us", C.c_double), ("batch_size", C.c_int32), ("pad", C.c_int32)]


class Burst(C.Structure):
    _fields_ = [("azimuth_time_us", C.c_int64), ("first_line", C.c_int64),
                ("first_valid_sample", C.c_int64), ("last_valid_sample", C.c_int64),
                ("burst_index", C.c_int32)]


class SubSwath(C.Structure):
    _fields_ = [("burst_count", C.c_int32), ("lines_per_burst", C.c_int64),
                ("azimuth_time_interval_s", C.c_double), ("slant_range_time_s", C.c_double),
                ("range_pixel_spacing_m", C.c_double)]


class DeburstLayout(C.Structure):
    _fields_ = [("lines", C.c_int64), ("samples", C.c_int64), ("valid_sample_offset", C.c_int64),
                ("n_bursts", C.c_int32),
                ("overlap", C.c_int64 * MAX_BURSTS), ("local_first", C.c_int64 * MAX_BURSTS),
                ("local_end", C.c_int64 * MAX_BURSTS), ("out_line_start", C.c_int64 * MAX_BURSTS)]


class Lut(C.Structure):
    _fields_ = [("n_vectors", C.c_int32), ("lines", C.c_void_p), ("offsets", C.c_void_p),
                ("pixels", C.c_void_p), ("values", C.c_void_p)]


class AzLut(C.Structure):
    _fields_ = [("n", C.c_int32), ("lines", C.c_void_p), ("values", C.c_void_p)]


class SwathInput(C.Structure):
    _fields_ = [("raw", C.c_void_p), ("raw_cols", C.c_int64), ("tiff_height", C.c_int64),
                ("swath", SubSwath), ("bursts", C.POINTER(Burst)), ("n_bursts", C.c_int32),
                ("cal", Lut), ("noise_range", Lut), ("noise_azimuth", AzLut)]


class MergedInfo(C.Structure):
    _fields_ = [("lines", C.c_int64), ("samples", C.c_int64), ("near_slant_range_time_s", C.c_double),
                ("range_pixel_spacing_s", C.c_double), ("range_pixel_spacing_m", C.c_double),
                ("azimuth_start_time_us", C.c_int64),
                ("cal_lut_extrapolation_gap_px", C.c_uint32), ("noise_lut_extrapolation_gap_px", C.c_uint32),
                ("n_swaths", C.c_int32),
                ("sw_out_offset", C.c_int64 * MAX_SWATHS), ("seam", C.c_int64 * MAX_SWATHS),
                ("swath_lines", C.c_int64 * MAX_SWATHS), ("swath_samples", C.c_int64 * MAX_SWATHS),
                ("swath_valid_offset", C.c_int64 * MAX_SWATHS)]


class SlcTiff(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("little_endian", C.c_int32), ("pad", C.c_int32),
                ("base_offset", C.c_uint64), ("row_bytes", C.c_uint64)]


class GrdGridPoint(C.Structure):
    _fields_ = [("slant_range_time_s", C.c_double), ("incidence_angle_deg", C.c_double), ("line", C.c_uint32),
                ("pad", C.c_uint32)]


class GrdImage(C.Structure):
    _fields_ = [("lines", C.c_int64), ("samples", C.c_int64), ("range_pixel_spacing_m", C.c_double),
                ("azimuth_pixel_spacing_m", C.c_double), ("range_looks", C.c_int32), ("azimuth_looks", C.c_int32)]


class CogLevel(C.Structure):
    _fields_ = [("cols", C.c_int64), ("rows", C.c_int64), ("tiles_x", C.c_int64), ("tiles_y", C.c_int64),
                ("ifd_offset", C.c_uint64), ("tile_offsets_offset", C.c_uint64),
                ("tile_counts_offset", C.c_uint64), ("tile_data_offset", C.c_uint64)]


class CogLayout(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("bigtiff", C.c_int32), ("tile_size", C.c_uint32),
                ("geo_key_shorts", C.c_int32), ("header_bytes", C.c_uint64), ("tile_bytes", C.c_uint64),
                ("file_bytes", C.c_uint64), ("pixel_scale_offset", C.c_uint64), ("tiepoint_offset", C.c_uint64),
                ("geo_keys_offset", C.c_uint64), ("level", CogLevel * COG_MAX_LEVELS)]


class ErrorDetail(C.Structure):
    _fields_ = [("status", C.c_int32), ("dem_status", C.c_int32), ("requested_us", C.c_int64),
                ("first_us", C.c_int64), ("last_us", C.c_int64), ("have_vectors", C.c_int32),
                ("need_vectors", C.c_int32), ("dem_expected_bytes", C.c_uint64), ("dem_actual_bytes", C.c_uint64)]


# Exported symbols of libsardine_hip.so (every function include/sardine_hip.h declares).
EXPORTED_SYMBOLS = (
    "sardine_last_error",
    "sardine_last_error_detail",
    "sardine_abi_version",
    "sardine_device_count",
    "sardine_tc_config_default",
    "sardine_tc_output_grid",
    "sardine_orbit_interpolate",
    "sardine_terrain_correction",
    "sardine_terrain_correction_noalloc",
    "sardine_speckle_apply",
    "sardine_deburst_layout_compute",
    "sardine_prepare_merged_layout",
    "sardine_prepare_merged",
    "sardine_apply_calibration",
    "sardine_to_db_inplace",
    "sardine_apply_multilook",
    "sardine_geoid_load_binary",
    "sardine_dem_srtm_mosaic",
    "sardine_dem_glo30_tile",
    "sardine_dem_glo30_mosaic",
    "sardine_geoid_load_ww15mgh",
    "sardine_slc_open",
    "sardine_slc_read_lines",
    "sardine_slc_read_lines_multi",
    "sardine_to_ground_range",
    "sardine_cog_layout_compute",
    "sardine_cog_header",
    "sardine_cog_pack",
    "sardine_write_cog_with_crs",
    "sardine_geotiff_header",
    "sardine_write_geotiff_with_crs",
    "sardine_write_geotiff_u8_with_crs",
    "sardine_write_tiff_no_crs",
    "sardine_synth_cint16",
    "sardine_diag_math",
    "sardine_diag_dem",
    "sardine_set_profiling",
    "sardine_last_kernel_ms",
    "sardine_recent_kernel_ms",
    "sardine_set_tc_segment",
    "sardine_set_tc_lead",
    "sardine_set_tc_tie_band",
    "sardine_scene_create",
    "sardine_scene_destroy",
    "sardine_batch_create",
    "sardine_batch_destroy",
    "sardine_batch_set_tuning",
    "sardine_batch_get_tuning",
    "sardine_batch_run",
    "sardine_batch_autotune",
    "sardine_server_create",
    "sardine_server_submit",
    "sardine_server_wait",
    "sardine_server_release",
    "sardine_server_destroy",
)
KERNEL_TC, KERNEL_CALIBRATE, KERNEL_SPECKLE, KERNEL_TO_DB, KERNEL_TC_SEGMENT, KERNEL_COG_PACK = 0, 1, 2, 3, 4, 5
KERNEL_MULTILOOK, KERNEL_GRD, KERNEL_DEM = 6, 7, 8
