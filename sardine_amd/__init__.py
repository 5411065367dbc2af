"""sardine_amd — MI355X (gfx950) implementation of SARdine's
calibrate-and-geocode path behind a C ABI (include/sardine_hip.h).

The compute lives in libsardine_hip.so (hand-written HIP kernels); this
package mirrors the reference's Rust interface for the path on top of it.
"""
from . import abi
from .api import (DemRaster, GeocodedImage, MergedSigma0, ResamplingKernel, SceneInputs, Sigma0Array,
                  SpeckleFilter, TerrainCorrectionConfig, apply_calibration, apply_speckle_filter,
                  apply_speckle_step, deburst_layout, merged_layout, output_grid, prepare_merged_scene,
                  synth_cint16, terrain_correction, to_db_inplace)
from .errors import (CalibrationError, DeburstError, DeviceError, MergeError, SardineError, SpeckleError,
                     TerrainCorrectionError)

__all__ = [
    "abi", "DemRaster", "GeocodedImage", "MergedSigma0", "ResamplingKernel", "SceneInputs", "Sigma0Array",
    "SpeckleFilter", "TerrainCorrectionConfig", "apply_calibration", "apply_speckle_filter", "apply_speckle_step",
    "deburst_layout", "merged_layout", "output_grid", "prepare_merged_scene", "synth_cint16", "terrain_correction",
    "to_db_inplace", "CalibrationError", "DeburstError", "DeviceError", "MergeError", "SardineError",
    "SpeckleError", "TerrainCorrectionError",
]
