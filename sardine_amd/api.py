"""Python mirror of the reference interface of the calibrate-and-geocode path.

Names and argument meaning follow the Rust crate:
  terrain_correction(merged, scene, dem, config) -> GeocodedImage
        sardine/src/terrain_correction.rs:1245-1251 (the `grids` argument is
        carried by SceneInputs.grid_points: the IW2 grid or the first one)
  TerrainCorrectionConfig.new(geoid)         terrain_correction.rs:223-235
  SpeckleFilter.{Boxcar,Lee,Frost,GammaMap,RefinedLee}(...).apply(data, cols, rows)
        speckle.rs:61-91, 586-603;  apply_speckle_filter speckle.rs:131
  prepare_merged_scene(swaths) -> MergedSigma0
        deburst → apply_calibration → merge_subswaths, scene_prep.rs:598-699
  apply_calibration(...) -> Sigma0Array      apply_calibration.rs:374
  deburst_layout(...)                        deburst.rs:172-305
  apply_multilook(merged, range_looks, azimuth_looks, ati) -> MergedSigma0
        scene_prep.rs:1060-1161
  to_ground_range(merged, grids, target_spacing_m, azimuth_pixel_spacing_m) -> GroundRangeImage
        ground_range.rs:303-442
  SlcReader.open(path).read_burst_raw(first_line, line_count) -> device CInt16
        slc_reader.rs:504-575
  to_db_inplace(data, noise_floor_linear)    export.rs:59-94
  write_cog_with_crs(path, data, cols, rows, geotransform, crs, tile_size)
        export.rs:329-354 (cog_layout / cog_header / cog_pack expose its parts)
  write_geotiff_with_crs / write_geotiff_u8_with_crs / write_tiff_no_crs
        export.rs:139-284 (geotiff_header exposes the bytes before the raster)
Rasters are torch tensors on the GPU (torch is used only to own device
memory and streams); every computation runs in libsardine_hip.so.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi
from ._native import last_error, last_error_detail, lib
from .errors import SpeckleError, raise_for


def _torch():
    import torch

    return torch


def _stream_ptr(stream) -> Optional[int]:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _check(status: int):
    if status != abi.OK:
        raise_for(status, last_error(), last_error_detail())


# ─── inputs ───────────────────────────────────────────────────────────────
@dataclass
class SceneInputs:
    """The parts of SceneMetadata + geolocation grids terrain_correction reads."""
    sv_time_us: np.ndarray      # (n,) int64
    sv_pos: np.ndarray          # (n, 3)
    sv_vel: np.ndarray          # (n, 3)
    bbox: tuple                 # (min_lat, max_lat, min_lon, max_lon)
    azimuth_time_interval_s: float
    grid_lat: np.ndarray        # seed-LUT grid points
    grid_time_us: np.ndarray
    n_subswaths: int = 1
    _keep: list = field(default_factory=list, repr=False)

    def as_c(self) -> abi.Scene:
        n = len(self.sv_time_us)
        svs = (abi.StateVector * max(n, 1))()
        for i in range(n):
            svs[i].time_us = int(self.sv_time_us[i])
            for c in range(3):
                svs[i].position_m[c] = float(self.sv_pos[i][c])
                svs[i].velocity_m_s[c] = float(self.sv_vel[i][c])
        m = len(self.grid_time_us)
        gp = (abi.GridPoint * max(m, 1))()
        for i in range(m):
            gp[i] = abi.GridPoint(float(self.grid_lat[i]), int(self.grid_time_us[i]))
        self._keep = [svs, gp]
        return abi.Scene(abi.Orbit(svs, n), abi.BBox(*self.bbox), float(self.azimuth_time_interval_s), gp, m,
                         int(self.n_subswaths))

    @classmethod
    def from_synthetic(cls, sc) -> "SceneInputs":
        return cls(sc.sv_t_us, sc.sv_pos, sc.sv_vel, sc.bbox, sc.ati, sc.grid_points, sc.grid_t_us, len(sc.swaths))


@dataclass
class MergedSigma0:
    """Raster-backed RadarImage (merge_subswaths.rs:106-162) in HBM."""
    data: object                # torch.float32 (lines, samples) on cuda
    nesz: object                # same shape, or None
    near_slant_range_time_s: float
    range_pixel_spacing_m: float
    azimuth_start_time_us: int
    range_pixel_spacing_s: float = 0.0
    cal_lut_extrapolation_gap_px: int = 0
    noise_lut_extrapolation_gap_px: int = 0
    info: Optional[abi.MergedInfo] = None

    @property
    def lines(self):
        return int(self.data.shape[0])

    @property
    def samples(self):
        return int(self.data.shape[1])

    def as_c(self) -> abi.RadarImage:
        return abi.RadarImage(self.data.data_ptr(), None if self.nesz is None else self.nesz.data_ptr(),
                              self.lines, self.samples, int(self.azimuth_start_time_us),
                              float(self.near_slant_range_time_s), float(self.range_pixel_spacing_m))


@dataclass
class DemRaster:
    """Geographic DEM raster with the GLO-30 sampling rule (dem.rs:351-409);
    with `tiling`, a GLO-30 mosaic held as one raster (DemMosaic, dem.rs:559-747)."""
    data: object                # torch.float32 (height, width) on cuda
    origin_lat_deg: float       # centre of pixel (0, 0)
    origin_lon_deg: float
    pixel_deg: float
    tiling: object = None       # abi.DemTiling or None

    def as_c(self) -> abi.DemRaster:
        return abi.DemRaster(self.data.data_ptr(), int(self.data.shape[1]), int(self.data.shape[0]),
                             float(self.origin_lat_deg), float(self.origin_lon_deg), float(self.pixel_deg),
                             C.pointer(self.tiling) if self.tiling is not None else None)


# ─── terrain correction ───────────────────────────────────────────────────
class ResamplingKernel:
    Bilinear = abi.RESAMPLE_BILINEAR
    Bicubic = abi.RESAMPLE_BICUBIC
    Lanczos3 = abi.RESAMPLE_LANCZOS3


@dataclass
class TerrainCorrectionConfig:
    """TerrainCorrectionConfig (terrain_correction.rs:100-235)."""
    geoid_grid: Optional[np.ndarray] = None       # None → GeoidModel::Zero; else 73x144 f32 grid
    pixel_spacing_deg: float = 0.0001
    crs: int = abi.CRS_WGS84
    utm_zone: int = 0
    max_newton_iterations: int = 10
    newton_tolerance_s: float = 1e-6
    flatten: bool = False
    compute_lia: bool = False
    noise_floor_margin_db: float = float("-inf")
    resampling: int = abi.RESAMPLE_BILINEAR
    projector_anchors: int = 256
    precision: int = abi.PRECISION_FP64
    aoi: Optional[tuple] = None                    # (row0, col0, rows, cols)

    @classmethod
    def new(cls, geoid_grid=None) -> "TerrainCorrectionConfig":
        return cls(geoid_grid=geoid_grid)

    def as_c(self) -> abi.TcConfig:
        c = abi.TcConfig()
        lib().sardine_tc_config_default(C.byref(c))
        c.pixel_spacing = self.pixel_spacing_deg
        c.crs = self.crs
        c.utm_zone = self.utm_zone
        c.max_newton_iterations = self.max_newton_iterations
        c.newton_tolerance_s = self.newton_tolerance_s
        c.flatten = int(self.flatten)
        c.compute_lia = int(self.compute_lia)
        c.noise_floor_margin_db = self.noise_floor_margin_db
        c.resampling = self.resampling
        c.projector_anchors = self.projector_anchors
        c.precision = self.precision
        if self.geoid_grid is not None:
            g = np.ascontiguousarray(self.geoid_grid, dtype=np.float32).reshape(73, 144)
            self._geoid_keep = g
            c.geoid = abi.Geoid(abi.GEOID_GRID_2P5, g.ctypes.data)
        if self.aoi is not None:
            c.aoi_row0, c.aoi_col0, c.aoi_rows, c.aoi_cols = (int(v) for v in self.aoi)
        return c


@dataclass
class GeocodedImage:
    """GeocodedImage (terrain_correction.rs:245-307) with device rasters."""
    data: object
    rows: int
    cols: int
    geotransform: tuple
    total_pixel_count: int
    valid_pixel_count: int
    dem_missing_count: int
    outside_footprint_count: int
    non_converged_count: int
    degenerate_geometry_count: int
    flat_masked_count: int
    noise_masked_count: int
    newton_iter_histogram: list
    cos_lia: object = None
    mask: object = None
    full_rows: int = 0
    full_cols: int = 0
    chain_redo_rounds: int = 0
    radar_coords: object = None
    tie_steps: int = 0
    tie_decisions_changed: int = 0
    tie_max_dev_us: float = 0.0

    @property
    def origin_lat_deg(self):
        return self.geotransform[3]

    @property
    def origin_lon_deg(self):
        return self.geotransform[0]

    def valid_fraction(self) -> float:
        return self.valid_pixel_count / self.total_pixel_count


def output_grid(scene: SceneInputs, config: TerrainCorrectionConfig):
    sc = scene.as_c()
    cfg = config.as_c()
    r, c = C.c_int64(), C.c_int64()
    gt = (C.c_double * 6)()
    _check(lib().sardine_tc_output_grid(C.byref(sc), C.byref(cfg), C.byref(r), C.byref(c), gt))
    return r.value, c.value, tuple(gt)


def terrain_correction(merged: MergedSigma0, scene: SceneInputs, dem: DemRaster,
                       config: TerrainCorrectionConfig, stream=None, out=None,
                       radar_coords: bool = False) -> GeocodedImage:
    torch = _torch()
    sc = scene.as_c()
    cfg = config.as_c()
    r, c, _ = output_grid(scene, config)
    if config.aoi is not None:
        r, c = int(config.aoi[2]), int(config.aoi[3])
    dev = merged.data.device
    if out is None:
        data = torch.empty((r, c), dtype=torch.float32, device=dev)
        cl = torch.empty((r, c), dtype=torch.float32, device=dev) if config.compute_lia else None
        mk = torch.empty((r, c), dtype=torch.uint8, device=dev) if config.compute_lia else None
        rc = torch.empty((r, c, 2), dtype=torch.float64, device=dev) if radar_coords else None
    else:
        data, cl, mk = out[:3]
        rc = out[3] if len(out) > 3 else None
    img = abi.GeocodedImage()
    img.radar_coords = None if rc is None else rc.data_ptr()
    img.data = data.data_ptr()
    img.cos_lia = None if cl is None else cl.data_ptr()
    img.mask = None if mk is None else mk.data_ptr()
    rim = merged.as_c()
    dr = dem.as_c()
    _check(lib().sardine_terrain_correction(C.byref(rim), C.byref(sc), C.byref(dr), C.byref(cfg), C.byref(img),
                                            _stream_ptr(stream)))
    return GeocodedImage(data, img.rows, img.cols, tuple(img.geotransform), img.total_pixel_count,
                         img.valid_pixel_count, img.dem_missing_count, img.outside_footprint_count,
                         img.non_converged_count, img.degenerate_geometry_count, img.flat_masked_count,
                         img.noise_masked_count, list(img.newton_iter_histogram), cl, mk, img.full_rows,
                         img.full_cols, img.chain_redo_rounds, rc,
                         img.tie_steps, img.tie_decisions_changed, img.tie_max_dev_us)


# ─── speckle ──────────────────────────────────────────────────────────────
@dataclass(frozen=True)
class SpeckleFilter:
    """SpeckleFilter (speckle.rs:61-91); `.apply` is SpeckleKernel::apply."""
    kind: int
    window: int = 7
    enl: float = 1.0
    damping: float = 1.0

    @staticmethod
    def Boxcar(window: int) -> "SpeckleFilter":
        return SpeckleFilter(abi.SPECKLE_BOXCAR, window)

    @staticmethod
    def Lee(window: int, enl: float) -> "SpeckleFilter":
        return SpeckleFilter(abi.SPECKLE_LEE, window, enl=enl)

    @staticmethod
    def Frost(window: int, damping: float) -> "SpeckleFilter":
        return SpeckleFilter(abi.SPECKLE_FROST, window, damping=damping)

    @staticmethod
    def GammaMap(window: int, enl: float) -> "SpeckleFilter":
        return SpeckleFilter(abi.SPECKLE_GAMMA_MAP, window, enl=enl)

    @staticmethod
    def RefinedLee(enl: float) -> "SpeckleFilter":
        return SpeckleFilter(abi.SPECKLE_REFINED_LEE, 7, enl=enl)

    def as_c(self) -> abi.SpeckleFilterC:
        return abi.SpeckleFilterC(self.kind, self.window, self.enl, self.damping)

    def apply(self, data, cols: int, rows: int, stream=None, out=None):
        return apply_speckle_filter(data, cols, rows, self, stream=stream, out=out)


def apply_speckle_filter(data, cols: int, rows: int, filt: SpeckleFilter, stream=None, out=None):
    """Returns a fresh device tensor (speckle.rs:131-180 is out-of-place)."""
    torch = _torch()
    flat = data.reshape(-1)
    if out is None:
        out = torch.empty(max(rows * cols, 0), dtype=torch.float32, device=data.device)
    f = filt.as_c()
    _check(lib().sardine_speckle_apply(flat.data_ptr() if flat.numel() else None, flat.numel(), cols, rows,
                                       C.byref(f), out.data_ptr() if out.numel() else None, _stream_ptr(stream)))
    return out.reshape(rows, cols) if rows * cols > 0 else out


def apply_speckle_step(data, cols: int, rows: int, choice: Optional[SpeckleFilter], stream=None):
    """run.rs:108-123: replaces `data` with the filtered raster (or leaves it)."""
    if choice is None:
        return data
    return choice.apply(data, cols, rows, stream=stream)


# ─── deburst / calibration / merge ────────────────────────────────────────
def deburst_layout(swath: abi.SubSwath, bursts, n_bursts: int, tiff_height: int) -> abi.DeburstLayout:
    L = abi.DeburstLayout()
    _check(lib().sardine_deburst_layout_compute(C.byref(swath), bursts, n_bursts, tiff_height, C.byref(L)))
    return L


def merged_layout(swath_inputs, n: int) -> abi.MergedInfo:
    info = abi.MergedInfo()
    _check(lib().sardine_prepare_merged_layout(swath_inputs, n, C.byref(info)))
    return info


def prepare_merged_scene(swath_inputs, n: int, device="cuda", stream=None, out=None) -> MergedSigma0:
    """deburst → apply_calibration → merge (scene_prep.rs:598-699) in one pass."""
    torch = _torch()
    info = merged_layout(swath_inputs, n)
    if out is None:
        s0 = torch.empty((info.lines, info.samples), dtype=torch.float32, device=device)
        nz = torch.empty((info.lines, info.samples), dtype=torch.float32, device=device)
    else:
        s0, nz = out
    _check(lib().sardine_prepare_merged(swath_inputs, n, s0.data_ptr(), nz.data_ptr(), C.byref(info),
                                        _stream_ptr(stream)))
    return MergedSigma0(s0, nz, info.near_slant_range_time_s, info.range_pixel_spacing_m,
                        info.azimuth_start_time_us, info.range_pixel_spacing_s, info.cal_lut_extrapolation_gap_px,
                        info.noise_lut_extrapolation_gap_px, info)


@dataclass
class Sigma0Array:
    data: object
    nesz: object
    lines: int
    samples: int
    valid_sample_offset: int
    cal_lut_extrapolation_gap_px: int
    noise_lut_extrapolation_gap_px: int


def apply_calibration(deburst, valid_sample_offset: int, cal: abi.Lut, noise_range: abi.Lut,
                      noise_azimuth: abi.AzLut, tiff_line_origin: int, stream=None) -> Sigma0Array:
    """apply_calibration (apply_calibration.rs:374-572); `deburst` is an int16
    device tensor of shape (lines, samples, 2)."""
    torch = _torch()
    lines, samples = int(deburst.shape[0]), int(deburst.shape[1])
    out = torch.empty((lines, samples), dtype=torch.float32, device=deburst.device)
    nz = torch.empty((lines, samples), dtype=torch.float32, device=deburst.device)
    g1, g2 = C.c_uint32(), C.c_uint32()
    _check(lib().sardine_apply_calibration(deburst.data_ptr(), lines, samples, valid_sample_offset, C.byref(cal),
                                           C.byref(noise_range), C.byref(noise_azimuth), tiff_line_origin,
                                           out.data_ptr(), nz.data_ptr(), C.byref(g1), C.byref(g2),
                                           _stream_ptr(stream)))
    return Sigma0Array(out, nz, lines, samples, valid_sample_offset, g1.value, g2.value)


def apply_multilook(merged: MergedSigma0, range_looks: int, azimuth_looks: int, azimuth_time_interval_s: float,
                    stream=None, out=None) -> MergedSigma0:
    """apply_multilook (scene_prep.rs:1060-1161): NaN-aware boxcar decimation
    of σ⁰ and NESZ with the centre-aligned geometry update.  1 x 1 returns
    `merged` itself, as the reference does."""
    if range_looks == 1 and azimuth_looks == 1:
        return merged
    torch = _torch()
    if merged.nesz is None:
        raise ValueError("multilook needs the NESZ raster")
    if merged.info is not None:
        info = abi.MergedInfo.from_buffer_copy(merged.info)
    else:
        info = abi.MergedInfo()
    info.lines, info.samples = merged.lines, merged.samples
    info.near_slant_range_time_s = merged.near_slant_range_time_s
    info.range_pixel_spacing_s = merged.range_pixel_spacing_s
    info.range_pixel_spacing_m = merged.range_pixel_spacing_m
    info.azimuth_start_time_us = int(merged.azimuth_start_time_us)
    info.cal_lut_extrapolation_gap_px = merged.cal_lut_extrapolation_gap_px
    info.noise_lut_extrapolation_gap_px = merged.noise_lut_extrapolation_gap_px
    shape = abi.MergedInfo()
    _check(lib().sardine_apply_multilook(None, None, C.byref(info), range_looks, azimuth_looks,
                                         azimuth_time_interval_s, None, None, C.byref(shape), None))
    if out is None:
        od = torch.empty((shape.lines, shape.samples), dtype=torch.float32, device=merged.data.device)
        on = torch.empty_like(od)
    else:
        od, on = out
    res = abi.MergedInfo()
    _check(lib().sardine_apply_multilook(merged.data.data_ptr(), merged.nesz.data_ptr(), C.byref(info), range_looks,
                                         azimuth_looks, azimuth_time_interval_s, od.data_ptr(), on.data_ptr(),
                                         C.byref(res), _stream_ptr(stream)))
    return MergedSigma0(od, on, res.near_slant_range_time_s, res.range_pixel_spacing_m, res.azimuth_start_time_us,
                        res.range_pixel_spacing_s, res.cal_lut_extrapolation_gap_px,
                        res.noise_lut_extrapolation_gap_px, res)


def load_geoid_grid(path, model: int = abi.GEOID_MODEL_EGM96, ascii: bool = False) -> np.ndarray:
    """Egm96Grid::load_binary / Egm2008Grid::load_binary / Egm96Grid::from_ww15mgh
    (geoid.rs:132-299) → the (73, 144) float32 grid TerrainCorrectionConfig
    takes as `geoid_grid`."""
    g = np.zeros(73 * 144, np.float32)
    st = (lib().sardine_geoid_load_ww15mgh(str(path).encode(), g.ctypes.data) if ascii else
          lib().sardine_geoid_load_binary(str(path).encode(), model, g.ctypes.data))
    _check(st)
    return g.reshape(73, 144)


def load_srtm_mosaic(paths, device="cuda", stream=None) -> "DemRaster":
    """DemMosaic::load_srtm_files (dem.rs:628-640) as one device DEM raster
    for terrain_correction (NaN for voids and for cells no tile covers)."""
    torch = _torch()
    arr = (C.c_char_p * max(len(paths), 1))(*[str(p).encode() for p in paths])
    r = abi.DemRaster()
    _check(lib().sardine_dem_srtm_mosaic(arr, len(paths), None, C.byref(r), None))
    data = torch.empty((r.height, r.width), dtype=torch.float32, device=device)
    _check(lib().sardine_dem_srtm_mosaic(arr, len(paths), data.data_ptr(), C.byref(r), _stream_ptr(stream)))
    return DemRaster(data, r.origin_lat_deg, r.origin_lon_deg, r.pixel_deg)


def load_glo30_tile(path, device="cuda", stream=None) -> "DemRaster":
    """Glo30Tile::load (dem.rs:248-352): one Copernicus GLO-30 GeoTIFF tile as
    a device DEM raster (NaN for nodata and non-finite samples)."""
    torch = _torch()
    r = abi.DemRaster()
    _check(lib().sardine_dem_glo30_tile(str(path).encode(), None, C.byref(r), None))
    data = torch.empty((r.height, r.width), dtype=torch.float32, device=device)
    _check(lib().sardine_dem_glo30_tile(str(path).encode(), data.data_ptr(), C.byref(r), _stream_ptr(stream)))
    return DemRaster(data, r.origin_lat_deg, r.origin_lon_deg, r.pixel_deg)


def load_glo30_mosaic(paths, device="cuda", stream=None) -> "DemRaster":
    """DemMosaic::load_glo30_files (dem.rs:642-653): GLO-30 tiles, in load
    order, as one device raster with its tile grid; the terrain correction
    samples it as DemMosaic::elevation_at_linear (first tile in load order
    that answers, each with its own origin and edge rule)."""
    torch = _torch()
    arr = (C.c_char_p * max(len(paths), 1))(*[str(p).encode() for p in paths])
    r, t = abi.DemRaster(), abi.DemTiling()
    _check(lib().sardine_dem_glo30_mosaic(arr, len(paths), None, C.byref(r), C.byref(t), None))
    data = torch.empty((r.height, r.width), dtype=torch.float32, device=device)
    _check(lib().sardine_dem_glo30_mosaic(arr, len(paths), data.data_ptr(), C.byref(r), C.byref(t),
                                          _stream_ptr(stream)))
    return DemRaster(data, r.origin_lat_deg, r.origin_lon_deg, r.pixel_deg, t)


class SlcReader:
    """SlcReader (slc_reader.rs:466-590) reading rows straight into HBM."""

    def __init__(self, path, tiff: abi.SlcTiff):
        self.path = str(path)
        self.tiff = tiff

    @classmethod
    def open(cls, path) -> "SlcReader":
        t = abi.SlcTiff()
        _check(lib().sardine_slc_open(str(path).encode(), C.byref(t)))
        return cls(path, t)

    @property
    def width(self) -> int:
        return self.tiff.width

    @property
    def height(self) -> int:
        return self.tiff.height

    @property
    def little_endian(self) -> bool:
        return bool(self.tiff.little_endian)

    def read_burst_raw(self, first_line: int, line_count: int, out=None, device="cuda", stream=None):
        """line_count rows from first_line as an int16 device tensor
        (line_count, width, 2) of [i, q] pairs."""
        torch = _torch()
        if out is None:
            out = torch.empty((line_count, self.width, 2), dtype=torch.int16, device=device)
        _check(lib().sardine_slc_read_lines(self.path.encode(), C.byref(self.tiff), first_line, line_count,
                                            out.data_ptr() if out.numel() else None, _stream_ptr(stream)))
        return out


class MultiSlcReader:
    """MultiSlcReader (multi_slc_reader.rs:40-193): the slices of one swath
    read in their logical line space, straight into HBM."""

    def __init__(self, readers):
        self.readers = list(readers)
        paths = [r.path.encode() for r in self.readers]
        self._paths = (C.c_char_p * max(len(paths), 1))(*paths)
        self._tiffs = (abi.SlcTiff * max(len(self.readers), 1))(*[r.tiff for r in self.readers])

    @classmethod
    def from_readers(cls, readers) -> "MultiSlcReader":
        m = cls(readers)
        # construction errors are raised here, as the reference raises them
        _check(lib().sardine_slc_read_lines_multi(m._paths, m._tiffs, len(m.readers), 0, 0, None, None))
        return m

    @property
    def width(self) -> int:
        return self.readers[0].width

    @property
    def height(self) -> int:
        return sum(r.height for r in self.readers)

    def read_burst_raw(self, first_line: int, line_count: int, out=None, device="cuda", stream=None):
        torch = _torch()
        if out is None:
            out = torch.empty((line_count, self.width, 2), dtype=torch.int16, device=device)
        _check(lib().sardine_slc_read_lines_multi(self._paths, self._tiffs, len(self.readers), first_line, line_count,
                                                  out.data_ptr() if out.numel() else None, _stream_ptr(stream)))
        return out


@dataclass
class GroundRangeImage:
    """GroundRangeImage (ground_range.rs:54-76) with a device raster."""
    data: object
    lines: int
    samples: int
    range_pixel_spacing_m: float
    azimuth_pixel_spacing_m: float
    range_looks: int
    azimuth_looks: int


def _grd_grids(grids):
    flat = []
    counts = []
    for g in grids:
        pts = g[1] if isinstance(g, tuple) and len(g) == 2 and not isinstance(g[0], (int, float)) else g
        n = 0
        for p in pts:
            if isinstance(p, tuple):
                line, t, a = p
            else:
                line, t, a = p.line, p.slant_range_time_s, p.incidence_angle_deg
            flat.append(abi.GrdGridPoint(float(t), float(a), int(line), 0))
            n += 1
        counts.append(n)
    arr = (abi.GrdGridPoint * max(len(flat), 1))(*flat)
    cnt = (C.c_int32 * max(len(counts), 1))(*counts)
    return arr, cnt, len(counts)


def to_ground_range(merged: MergedSigma0, grids, target_spacing_m: float, azimuth_pixel_spacing_m: float,
                    stream=None, out=None) -> GroundRangeImage:
    """to_ground_range (ground_range.rs:303-442).  `grids`: one entry per
    subswath, either (id, points) or points, each point a (line, slant range
    time s, incidence °) tuple or an object with those attributes."""
    torch = _torch()
    info = abi.MergedInfo()
    info.lines, info.samples = merged.lines, merged.samples
    info.near_slant_range_time_s = merged.near_slant_range_time_s
    info.range_pixel_spacing_s = merged.range_pixel_spacing_s
    info.range_pixel_spacing_m = merged.range_pixel_spacing_m
    pts, cnt, n = _grd_grids(grids)
    img = abi.GrdImage()
    _check(lib().sardine_to_ground_range(None, C.byref(info), pts, cnt, n, target_spacing_m, azimuth_pixel_spacing_m,
                                         None, C.byref(img), None))
    o = out if out is not None else torch.empty((img.lines, img.samples), dtype=torch.float32,
                                                device=merged.data.device)
    _check(lib().sardine_to_ground_range(merged.data.data_ptr(), C.byref(info), pts, cnt, n, target_spacing_m,
                                         azimuth_pixel_spacing_m, o.data_ptr(), C.byref(img), _stream_ptr(stream)))
    return GroundRangeImage(o, img.lines, img.samples, img.range_pixel_spacing_m, img.azimuth_pixel_spacing_m,
                            img.range_looks, img.azimuth_looks)


def to_db_inplace(data, noise_floor_linear: float, stream=None):
    """export.rs:59-94; returns (converted, noise_masked)."""
    counts = (C.c_uint64 * 2)()
    _check(lib().sardine_to_db_inplace(data.data_ptr(), data.numel(), noise_floor_linear, counts,
                                       _stream_ptr(stream)))
    return int(counts[0]), int(counts[1])


def cog_layout(cols: int, rows: int, tile_size: int, crs: int = abi.CRS_WGS84) -> abi.CogLayout:
    """The pyramid levels, TIFF flavour and file offsets of a COG of a
    cols x rows raster (write_cog_raw_inner, export.rs:430-535)."""
    L = abi.CogLayout()
    _check(lib().sardine_cog_layout_compute(cols, rows, tile_size, crs, C.byref(L)))
    return L


def cog_header(layout: abi.CogLayout, geotransform, crs: int = abi.CRS_WGS84, utm_zone: int = 0) -> bytes:
    """The bytes of the file before its first tile (export.rs:537-670)."""
    gt = (C.c_double * 6)(*geotransform)
    buf = (C.c_uint8 * layout.header_bytes)()
    _check(lib().sardine_cog_header(C.byref(layout), gt, crs, utm_zone, buf, layout.header_bytes))
    return bytes(buf)


def cog_pack(data, layout: abi.CogLayout, tiles=None, stream=None):
    """The file's tile region (every level, coarsest first, NaN-padded
    tiles) packed on the device from a float32 device raster; returns the
    float32 device tensor of layout.tile_bytes / 4 values."""
    torch = _torch()
    full = layout.level[layout.n_levels - 1]
    if data.dtype != torch.float32 or data.numel() != full.cols * full.rows or not data.is_contiguous():
        raise ValueError("data must be a contiguous float32 raster of the layout's size")
    if tiles is None:
        tiles = torch.empty(layout.tile_bytes // 4, dtype=torch.float32, device=data.device)
    _check(lib().sardine_cog_pack(data.data_ptr(), C.byref(layout), tiles.data_ptr(), _stream_ptr(stream)))
    return tiles


def write_cog_with_crs(path, data, cols: int, rows: int, geotransform, crs: int = abi.CRS_WGS84,
                       tile_size: int = 512, utm_zone: int = 0, stream=None):
    """write_cog_with_crs (export.rs:329-354): a tiled float32 COG with its
    overview pyramid; `data` is a float32 device raster of rows x cols."""
    torch = _torch()
    if data.dtype != torch.float32 or data.numel() != cols * rows or not data.is_contiguous():
        raise ValueError("data must be a contiguous float32 raster of rows x cols")
    gt = (C.c_double * 6)(*geotransform)
    _check(lib().sardine_write_cog_with_crs(str(path).encode(), data.data_ptr(), cols, rows, gt, crs, utm_zone,
                                            tile_size, _stream_ptr(stream)))


def geotiff_header(cols: int, rows: int, bytes_per_pixel: int = 4, geotransform=None, crs: int = abi.CRS_WGS84,
                   utm_zone: int = 0) -> bytes:
    """The bytes before the raster of the striped writers (export.rs:788-1219);
    geotransform None = write_tiff_no_crs's plain TIFF."""
    gt = None if geotransform is None else (C.c_double * 6)(*geotransform)
    n = C.c_uint64()
    _check(lib().sardine_geotiff_header(cols, rows, bytes_per_pixel, gt, crs, utm_zone, None, 0, C.byref(n)))
    buf = (C.c_uint8 * n.value)()
    _check(lib().sardine_geotiff_header(cols, rows, bytes_per_pixel, gt, crs, utm_zone, buf, n.value, C.byref(n)))
    return bytes(buf)


def _raster_arg(data, cols, rows, dtype):
    if data.dtype != dtype or data.numel() != cols * rows or not data.is_contiguous():
        raise ValueError(f"data must be a contiguous {dtype} raster of rows x cols")


def write_geotiff_with_crs(path, data, cols: int, rows: int, geotransform, crs: int = abi.CRS_WGS84,
                           utm_zone: int = 0, stream=None):
    """write_geotiff_with_crs (export.rs:139-187) of a float32 device raster."""
    _raster_arg(data, cols, rows, _torch().float32)
    _check(lib().sardine_write_geotiff_with_crs(str(path).encode(), data.data_ptr(), cols, rows,
                                                (C.c_double * 6)(*geotransform), crs, utm_zone,
                                                _stream_ptr(stream)))


def write_geotiff_u8_with_crs(path, data, cols: int, rows: int, geotransform, crs: int = abi.CRS_WGS84,
                              utm_zone: int = 0, stream=None):
    """write_geotiff_u8_with_crs (export.rs:204-245) of a uint8 device raster."""
    _raster_arg(data, cols, rows, _torch().uint8)
    _check(lib().sardine_write_geotiff_u8_with_crs(str(path).encode(), data.data_ptr(), cols, rows,
                                                   (C.c_double * 6)(*geotransform), crs, utm_zone,
                                                   _stream_ptr(stream)))


def write_tiff_no_crs(path, data, cols: int, rows: int, stream=None):
    """write_tiff_no_crs (export.rs:252-284) of a float32 device raster."""
    _raster_arg(data, cols, rows, _torch().float32)
    _check(lib().sardine_write_tiff_no_crs(str(path).encode(), data.data_ptr(), cols, rows, _stream_ptr(stream)))


def synth_cint16(out, seed: int, index_base: int = 0, sigma: float = 100.0, stream=None):
    """Fill an int16 device tensor (..., 2) with the synthetic DN generator."""
    _check(lib().sardine_synth_cint16(out.data_ptr(), out.numel() // 2, seed, index_base, sigma,
                                      _stream_ptr(stream)))
    return out


# ─── resident scenes and batched geocoding (no reference counterpart) ─────
class ResidentScene:
    """A scene kept on the device for batched geocoding: orbit and seed table
    uploaded once; the merged and DEM rasters stay the caller's (keep
    references to them for the scene's lifetime)."""

    def __init__(self, merged: MergedSigma0, scene: SceneInputs, dem: DemRaster):
        self.merged, self.scene, self.dem = merged, scene, dem
        self._c = (merged.as_c(), scene.as_c(), dem.as_c())
        h = C.c_void_p()
        _check(lib().sardine_scene_create(C.byref(self._c[0]), C.byref(self._c[1]), C.byref(self._c[2]),
                                          C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().sardine_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class BatchJob:
    """One AOI geocode request: the scene index, its TerrainCorrectionConfig
    (AOI window, CRS, resampler, ...) and an optional speckle filter."""
    scene: int
    config: TerrainCorrectionConfig
    speckle: Optional[SpeckleFilter] = None
    out: Optional[tuple] = None       # (data, cos_lia, mask) device tensors; allocated when None
    speckle_out: object = None


@dataclass
class BatchResult:
    image: GeocodedImage
    speckled: object  # the speckle output tensor, or None
    status: int


class Batch:
    """Batched executor (sardine_batch_*): every job of a run shares the
    persistent chain and speckle launches (replayed from one HIP graph when
    use_graph)."""

    def __init__(self, max_jobs: int = 256, use_graph: bool = True):
        h = C.c_void_p()
        _check(lib().sardine_batch_create(int(max_jobs), 1 if use_graph else 0, C.byref(h)))
        self.handle = h
        self.max_jobs = max_jobs
        self.last_device_ms = None

    def close(self):
        if self.handle:
            lib().sardine_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tuning(self) -> abi.BatchTuning:
        t = abi.BatchTuning()
        _check(lib().sardine_batch_get_tuning(self.handle, C.byref(t)))
        return t

    @tuning.setter
    def tuning(self, t: abi.BatchTuning):
        _check(lib().sardine_batch_set_tuning(self.handle, C.byref(t)))

    def _prepare(self, scenes, jobs):
        torch = _torch()
        arr = (C.c_void_p * len(scenes))(*[s.handle for s in scenes])
        cj = (abi.BatchJob * len(jobs))()
        keep = []
        for i, j in enumerate(jobs):
            sc = scenes[j.scene]
            r, c, _ = output_grid(sc.scene, j.config)
            if j.config.aoi is not None:
                r, c = int(j.config.aoi[2]), int(j.config.aoi[3])
            dev = sc.merged.data.device
            if j.out is None:
                data = torch.empty((r, c), dtype=torch.float32, device=dev)
                cl = torch.empty((r, c), dtype=torch.float32, device=dev) if j.config.compute_lia else None
                mk = torch.empty((r, c), dtype=torch.uint8, device=dev) if j.config.compute_lia else None
                j.out = (data, cl, mk)
            if j.speckle is not None and j.speckle_out is None:
                j.speckle_out = torch.empty((r, c), dtype=torch.float32, device=dev)
            data, cl, mk = j.out
            img = abi.GeocodedImage()
            img.data = data.data_ptr()
            img.cos_lia = None if cl is None else cl.data_ptr()
            img.mask = None if mk is None else mk.data_ptr()
            keep.append(img)
            cj[i].scene = j.scene
            cj[i].config = j.config.as_c()
            cj[i].out = C.pointer(img)
            cj[i].speckle = j.speckle.as_c() if j.speckle is not None else abi.SpeckleFilterC(-1, 0, 0.0, 0.0)
            cj[i].speckle_out = None if j.speckle is None else j.speckle_out.data_ptr()
        return arr, cj, keep

    def prepare(self, scenes, jobs):
        """Builds the C job table once (outputs are allocated on first use);
        `execute` can then run it any number of times."""
        return (list(scenes), list(jobs)) + self._prepare(scenes, jobs)

    def execute(self, prepared, stream=None, check: bool = True):
        scenes, jobs, arr, cj, keep = prepared
        ms = C.c_double(0.0)
        st = lib().sardine_batch_run(self.handle, arr, len(scenes), cj, len(jobs), _stream_ptr(stream), C.byref(ms))
        if check:
            _check(st)
        self.last_device_ms = ms.value
        out = []
        for i, j in enumerate(jobs):
            img = keep[i]
            data, cl, mk = j.out
            gi = GeocodedImage(data, img.rows, img.cols, tuple(img.geotransform), img.total_pixel_count,
                               img.valid_pixel_count, img.dem_missing_count, img.outside_footprint_count,
                               img.non_converged_count, img.degenerate_geometry_count, img.flat_masked_count,
                               img.noise_masked_count, list(img.newton_iter_histogram), cl, mk, img.full_rows,
                               img.full_cols, img.chain_redo_rounds, None,
                         img.tie_steps, img.tie_decisions_changed, img.tie_max_dev_us)
            out.append(BatchResult(gi, j.speckle_out if j.speckle is not None else None, cj[i].status))
        return out

    def run(self, scenes, jobs, stream=None, check: bool = True):
        return self.execute(self.prepare(scenes, jobs), stream=stream, check=check)

    def autotune(self, scenes, jobs, stream=None):
        arr, cj, keep = self._prepare(scenes, jobs)
        best = abi.BatchTuning()
        ms = C.c_double(0.0)
        _check(lib().sardine_batch_autotune(self.handle, arr, len(scenes), cj, len(jobs), _stream_ptr(stream),
                                            C.byref(best), C.byref(ms)))
        return best, ms.value


# ─── serving (no reference counterpart) ───────────────────────────────────
class _DeviceView:
    """Zero-copy view of server-owned device memory for torch.as_tensor."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (int(ptr), False),
                                         "version": 3, "strides": None}


@dataclass
class ServedResult:
    status: int
    image: Optional[GeocodedImage]
    speckled: object
    t_submit_us: float
    t_admit_us: float
    t_launch_us: float
    t_done_us: float
    batch_size: int

    @property
    def latency_ms(self) -> float:
        return (self.t_done_us - self.t_submit_us) / 1e3


class Server:
    """The native request loop (sardine_server_*): requests on resident
    scenes are admitted by HBM budget, batched as they arrive and run with
    several batches in flight.  Result tensors view server-owned memory
    until `release`."""

    def __init__(self, scenes, hbm_budget_bytes: int, max_batch: int = 64, slots: int = 2, use_graph: bool = True):
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[s.handle for s in self.scenes])
        opt = abi.ServerOptions(int(hbm_budget_bytes), int(max_batch), int(slots), 1 if use_graph else 0, 0)
        h = C.c_void_p()
        _check(lib().sardine_server_create(arr, len(self.scenes), C.byref(opt), C.byref(h)))
        self.handle = h

    def submit(self, scene: int, config: TerrainCorrectionConfig, speckle: Optional[SpeckleFilter] = None) -> int:
        r = abi.Request()
        r.scene = scene
        r.config = config.as_c()
        r.speckle = speckle.as_c() if speckle is not None else abi.SpeckleFilterC(-1, 0, 0.0, 0.0)
        t = C.c_uint64()
        _check(lib().sardine_server_submit(self.handle, C.byref(r), C.byref(t)))
        return t.value

    def wait(self, ticket: int, block: bool = True) -> ServedResult:
        torch = _torch()
        r = abi.Response()
        _check(lib().sardine_server_wait(self.handle, ticket, 1 if block else 0, C.byref(r)))
        img = None
        sp = None
        if r.done and r.status == 0:
            g = r.image
            data = torch.as_tensor(_DeviceView(g.data, (g.rows, g.cols), "<f4"), device="cuda")
            img = GeocodedImage(data, g.rows, g.cols, tuple(g.geotransform), g.total_pixel_count,
                                g.valid_pixel_count, g.dem_missing_count, g.outside_footprint_count,
                                g.non_converged_count, g.degenerate_geometry_count, g.flat_masked_count,
                                g.noise_masked_count, list(g.newton_iter_histogram), None, None, g.full_rows,
                                g.full_cols, g.chain_redo_rounds, None,
                                g.tie_steps, g.tie_decisions_changed, g.tie_max_dev_us)
            if r.speckled:
                sp = torch.as_tensor(_DeviceView(r.speckled, (g.rows, g.cols), "<f4"), device="cuda")
        return ServedResult(r.status, img, sp, r.t_submit_us, r.t_admit_us, r.t_launch_us, r.t_done_us,
                            r.batch_size)

    def release(self, ticket: int, stream=None):
        """Returns the request's output memory once the work enqueued so far
        on `stream` (default: torch's current stream) has run; its tensors
        must not be used afterwards."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(lib().sardine_server_release(self.handle, ticket, C.c_void_p(s.cuda_stream)))

    def close(self):
        if self.handle:
            lib().sardine_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
