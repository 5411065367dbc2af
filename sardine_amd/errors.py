"""Exception types mirroring the reference's error enums on the path.

  TerrainCorrectionError  sardine/src/terrain_correction.rs:45-67
  SpeckleError            sardine/src/speckle.rs:96-108
  CalibrationError        sardine/src/apply_calibration.rs:64-84
  DeburstError            sardine/src/deburst.rs:56-121
  MergeError              sardine/src/merge_subswaths.rs:62-81
  ExportError             sardine/src/export.rs:26-39
  GroundRangeError        sardine/src/ground_range.rs:31-47
  SlcReadError            sardine/src/slc_reader.rs:66-131
  GeoidError              sardine/src/geoid.rs:87-104
  DemError                sardine/src/dem.rs:54-109
Each carries the C-ABI status code (`.status`) and its name (`.variant`).
"""
from __future__ import annotations

from . import abi


class SardineError(RuntimeError):
    def __init__(self, status: int, message: str, detail=None):
        self.status = status
        self.variant = abi.STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.variant}: {message}")


class TerrainCorrectionError(SardineError):
    """With the fields of the reference's nested variants:
    Orbit(OutsideOrbitWindow { requested, first, last }) → `requested_us`,
    `first_us`, `last_us` (µs on the state vectors' clock);
    Orbit(InsufficientCoverage) → `have_vectors`, `need_vectors`;
    Dem(DemError::…) → `dem_variant` (a DemError status name)."""

    def __init__(self, status: int, message: str, detail=None):
        super().__init__(status, message)
        self.requested_us = self.first_us = self.last_us = None
        self.have_vectors = self.need_vectors = None
        self.dem_variant = None
        if detail is not None and detail.status == status:
            if status == abi.ERR_TC_ORBIT_OUTSIDE_WINDOW:
                self.requested_us, self.first_us, self.last_us = detail.requested_us, detail.first_us, detail.last_us
            elif status == abi.ERR_TC_ORBIT_INSUFFICIENT:
                self.have_vectors, self.need_vectors = detail.have_vectors, detail.need_vectors
            elif status == abi.ERR_TC_DEM:
                self.dem_variant = abi.STATUS_NAMES.get(detail.dem_status, str(detail.dem_status))


class SpeckleError(SardineError):
    pass


class CalibrationError(SardineError):
    pass


class DeburstError(SardineError):
    pass


class MergeError(SardineError):
    pass


class ExportError(SardineError):
    pass


class GroundRangeError(SardineError):
    pass


class SlcReadError(SardineError):
    pass


class GeoidError(SardineError):
    pass


class DemError(SardineError):
    pass


class DeviceError(SardineError):
    pass


def raise_for(status: int, message: str, detail=None):
    if status == abi.OK:
        return
    if 10 <= status < 20:
        raise TerrainCorrectionError(status, message, detail)
    if 20 <= status < 30:
        raise SpeckleError(status, message)
    if 30 <= status < 40:
        raise CalibrationError(status, message)
    if 40 <= status < 50:
        raise DeburstError(status, message)
    if 50 <= status < 60:
        raise MergeError(status, message)
    if 60 <= status < 70:
        raise ExportError(status, message)
    if 70 <= status < 80:
        raise GroundRangeError(status, message)
    if 80 <= status < 90:
        raise SlcReadError(status, message)
    if 100 <= status < 110:
        raise GeoidError(status, message)
    if 110 <= status < 120:
        raise DemError(status, message)
    raise DeviceError(status, message)
