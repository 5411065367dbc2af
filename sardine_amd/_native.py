"""Loader of the native library sardine_amd/libsardine_hip.so.

There is no fallback: if the library is missing or cannot be loaded, every
call raises NativeLibraryError.  Build it with `python -m sardine_amd.build`
(or __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

from . import abi

# SARDINE_LIB selects another build of the same library (the host-sanitizer
# build, libsardine_hip_san.so); the default is the in-tree release build.
LIB_PATH = pathlib.Path(os.environ.get("SARDINE_LIB") or pathlib.Path(__file__).resolve().parent / "libsardine_hip.so")


class NativeLibraryError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NativeLibraryError(f"{LIB_PATH} is not built; run `python -m sardine_amd.build`")
    try:
        L = C.CDLL(str(LIB_PATH))
    except OSError as e:
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    P, i32, i64, d, f, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_float, C.c_uint64
    L.sardine_last_error.restype = C.c_char_p
    L.sardine_last_error.argtypes = []
    L.sardine_last_error_detail.argtypes = [C.POINTER(abi.ErrorDetail)]
    L.sardine_abi_version.restype = C.c_int
    L.sardine_device_count.restype = C.c_int
    L.sardine_tc_config_default.argtypes = [C.POINTER(abi.TcConfig)]
    L.sardine_tc_config_default.restype = None
    L.sardine_tc_output_grid.argtypes = [C.POINTER(abi.Scene), C.POINTER(abi.TcConfig), C.POINTER(i64),
                                         C.POINTER(i64), P]
    L.sardine_orbit_interpolate.argtypes = [C.POINTER(abi.Orbit), i64, P, P, P]
    L.sardine_terrain_correction.argtypes = [C.POINTER(abi.RadarImage), C.POINTER(abi.Scene),
                                             C.POINTER(abi.DemRaster), C.POINTER(abi.TcConfig),
                                             C.POINTER(abi.GeocodedImage), P]
    L.sardine_terrain_correction_noalloc.argtypes = L.sardine_terrain_correction.argtypes
    L.sardine_speckle_apply.argtypes = [P, i64, i64, i64, C.POINTER(abi.SpeckleFilterC), P, P]
    L.sardine_deburst_layout_compute.argtypes = [C.POINTER(abi.SubSwath), C.POINTER(abi.Burst), i32, i64,
                                                 C.POINTER(abi.DeburstLayout)]
    L.sardine_prepare_merged_layout.argtypes = [C.POINTER(abi.SwathInput), i32, C.POINTER(abi.MergedInfo)]
    L.sardine_prepare_merged.argtypes = [C.POINTER(abi.SwathInput), i32, P, P, C.POINTER(abi.MergedInfo), P]
    L.sardine_apply_calibration.argtypes = [P, i64, i64, i64, C.POINTER(abi.Lut), C.POINTER(abi.Lut),
                                            C.POINTER(abi.AzLut), i64, P, P, C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32), P]
    L.sardine_to_db_inplace.argtypes = [P, i64, f, P, P]
    L.sardine_cog_layout_compute.argtypes = [i64, i64, C.c_uint32, i32, C.POINTER(abi.CogLayout)]
    L.sardine_cog_header.argtypes = [C.POINTER(abi.CogLayout), P, i32, i32, P, u64]
    L.sardine_cog_pack.argtypes = [P, C.POINTER(abi.CogLayout), P, P]
    L.sardine_write_cog_with_crs.argtypes = [C.c_char_p, P, i64, i64, P, i32, i32, C.c_uint32, P]
    L.sardine_geotiff_header.argtypes = [i64, i64, i32, P, i32, i32, P, u64, C.POINTER(u64)]
    L.sardine_write_geotiff_with_crs.argtypes = [C.c_char_p, P, i64, i64, P, i32, i32, P]
    L.sardine_write_geotiff_u8_with_crs.argtypes = [C.c_char_p, P, i64, i64, P, i32, i32, P]
    L.sardine_write_tiff_no_crs.argtypes = [C.c_char_p, P, i64, i64, P]
    L.sardine_apply_multilook.argtypes = [P, P, C.POINTER(abi.MergedInfo), i32, i32, d, P, P,
                                          C.POINTER(abi.MergedInfo), P]
    L.sardine_to_ground_range.argtypes = [P, C.POINTER(abi.MergedInfo), C.POINTER(abi.GrdGridPoint), P, i32, d, d,
                                          P, C.POINTER(abi.GrdImage), P]
    L.sardine_dem_glo30_tile.argtypes = [C.c_char_p, P, C.POINTER(abi.DemRaster), P]
    L.sardine_dem_srtm_mosaic.argtypes = [P, i32, P, C.POINTER(abi.DemRaster), P]
    L.sardine_dem_glo30_mosaic.argtypes = [P, i32, P, C.POINTER(abi.DemRaster), C.POINTER(abi.DemTiling), P]
    L.sardine_geoid_load_binary.argtypes = [C.c_char_p, i32, P]
    L.sardine_geoid_load_ww15mgh.argtypes = [C.c_char_p, P]
    L.sardine_slc_open.argtypes = [C.c_char_p, C.POINTER(abi.SlcTiff)]
    L.sardine_slc_read_lines_multi.argtypes = [P, C.POINTER(abi.SlcTiff), i32, i64, i64, P, P]
    L.sardine_slc_read_lines.argtypes = [C.c_char_p, C.POINTER(abi.SlcTiff), i64, i64, P, P]
    L.sardine_synth_cint16.argtypes = [P, i64, u64, u64, d, P]
    L.sardine_set_profiling.argtypes = [C.c_int]
    L.sardine_set_profiling.restype = None
    L.sardine_last_kernel_ms.argtypes = [C.c_int]
    L.sardine_last_kernel_ms.restype = C.c_double
    L.sardine_recent_kernel_ms.argtypes = [C.c_int, P, C.c_int32]
    L.sardine_set_tc_segment.argtypes = [C.c_int64]
    L.sardine_set_tc_segment.restype = None
    L.sardine_set_tc_lead.argtypes = [C.c_int32]
    L.sardine_set_tc_lead.restype = None
    L.sardine_set_tc_tie_band.argtypes = [C.c_double]
    L.sardine_set_tc_tie_band.restype = None
    PP = C.POINTER(C.c_void_p)
    L.sardine_scene_create.argtypes = [C.POINTER(abi.RadarImage), C.POINTER(abi.Scene), C.POINTER(abi.DemRaster), PP]
    L.sardine_scene_destroy.argtypes = [P]
    L.sardine_scene_destroy.restype = None
    L.sardine_batch_create.argtypes = [i32, i32, PP]
    L.sardine_batch_destroy.argtypes = [P]
    L.sardine_batch_destroy.restype = None
    L.sardine_batch_set_tuning.argtypes = [P, C.POINTER(abi.BatchTuning)]
    L.sardine_batch_get_tuning.argtypes = [P, C.POINTER(abi.BatchTuning)]
    L.sardine_batch_run.argtypes = [P, PP, i32, C.POINTER(abi.BatchJob), i32, P, C.POINTER(d)]
    L.sardine_batch_autotune.argtypes = [P, PP, i32, C.POINTER(abi.BatchJob), i32, P, C.POINTER(abi.BatchTuning),
                                         C.POINTER(d)]
    L.sardine_server_create.argtypes = [PP, i32, C.POINTER(abi.ServerOptions), PP]
    L.sardine_server_submit.argtypes = [P, C.POINTER(abi.Request), C.POINTER(u64)]
    L.sardine_server_wait.argtypes = [P, u64, i32, C.POINTER(abi.Response)]
    L.sardine_server_release.argtypes = [P, u64, P]
    L.sardine_server_destroy.argtypes = [P]
    L.sardine_diag_math.argtypes = [i32, P, P, P, i64, P]
    L.sardine_diag_dem.argtypes = [C.POINTER(abi.DemRaster), P, P, P, i64, P]
    L.sardine_server_destroy.restype = None
    for name in abi.EXPORTED_SYMBOLS:
        if name not in ("sardine_last_error", "sardine_abi_version", "sardine_device_count",
                        "sardine_tc_config_default", "sardine_set_profiling", "sardine_last_kernel_ms",
                        "sardine_set_tc_segment", "sardine_set_tc_lead", "sardine_set_tc_tie_band",
                        "sardine_scene_destroy",
                        "sardine_batch_destroy", "sardine_server_destroy"):
            getattr(L, name).restype = C.c_int
    if L.sardine_abi_version() != abi.ABI_VERSION:
        raise NativeLibraryError(f"ABI version mismatch: library {L.sardine_abi_version()}, "
                                 f"python {abi.ABI_VERSION}")
    _lib = L
    return L


def last_error() -> str:
    return lib().sardine_last_error().decode(errors="replace")


def last_error_detail() -> "abi.ErrorDetail":
    d = abi.ErrorDetail()
    lib().sardine_last_error_detail(C.byref(d))
    return d
