"""sardine_amd — MI355X (gfx950) implementation of SARdine's
calibrate-and-geocode path behind a C ABI (include/sardine_hip.h).

The compute lives in libsardine_hip.so (hand-written HIP kernels); this
package mirrors the reference's Rust interface for the path on top of it.
"""
from . import abi
from .api import (DemRaster, GeocodedImage, GroundRangeImage, MergedSigma0, MultiSlcReader, ResamplingKernel, SceneInputs,
                  Sigma0Array, SlcReader, SpeckleFilter, TerrainCorrectionConfig, apply_calibration, apply_multilook,
                  apply_speckle_filter, apply_speckle_step, cog_header, cog_layout, cog_pack, deburst_layout,
                  geotiff_header, load_geoid_grid, load_glo30_tile, load_srtm_mosaic, merged_layout, output_grid, prepare_merged_scene, synth_cint16,
                  terrain_correction, to_db_inplace, to_ground_range, write_cog_with_crs, write_geotiff_u8_with_crs,
                  write_geotiff_with_crs, write_tiff_no_crs)
from .errors import (CalibrationError, DeburstError, DemError, DeviceError, ExportError, GeoidError, GroundRangeError, MergeError,
                     SardineError, SlcReadError, SpeckleError, TerrainCorrectionError)

__all__ = [
    "abi", "DemRaster", "GeocodedImage", "GroundRangeImage", "MergedSigma0", "MultiSlcReader", "ResamplingKernel", "SceneInputs",
    "Sigma0Array", "SlcReader", "SpeckleFilter", "TerrainCorrectionConfig", "apply_calibration", "apply_multilook",
    "apply_speckle_filter", "apply_speckle_step", "cog_header", "cog_layout", "cog_pack", "deburst_layout",
    "geotiff_header", "load_geoid_grid", "load_glo30_tile", "load_srtm_mosaic", "merged_layout", "output_grid", "prepare_merged_scene", "synth_cint16", "terrain_correction",
    "to_db_inplace", "to_ground_range", "write_cog_with_crs", "write_geotiff_u8_with_crs", "write_geotiff_with_crs",
    "write_tiff_no_crs", "CalibrationError", "DeburstError", "DemError", "DeviceError", "ExportError", "GroundRangeError",
    "MergeError", "SardineError", "SlcReadError", "SpeckleError", "TerrainCorrectionError",
]
