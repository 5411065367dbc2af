"""Device-resident inputs for a synthetic scene (tests, smoke, benchmark).

`DeviceScene` keeps every raster of one scene in HBM: the CInt16 TIFF rows of
each subswath (generated on the device by the same counter-based generator as
the host one, or uploaded), the DEM and the small metadata, and runs the
path's stages through the C ABI.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import abi, api
from .synthetic import SyntheticScene


@dataclass
class DeviceScene:
    sc: SyntheticScene
    raws: list
    dem: api.DemRaster
    scene: api.SceneInputs
    device: object
    _keep: list = field(default_factory=list, repr=False)
    _swc: object = None

    @classmethod
    def build(cls, sc: SyntheticScene, device="cuda", host_raw: bool = False) -> "DeviceScene":
        import torch

        raws = []
        for sw in sc.swaths:
            if host_raw:
                raws.append(torch.from_numpy(sw.raw_host()).to(device))
            else:
                t = torch.empty((sw.tiff_height, sw.raw_cols, 2), dtype=torch.int16, device=device)
                api.synth_cint16(t, sw.seed, 0, 100.0)
                raws.append(t)
        dem_np = sc.dem(device=device)
        dem = api.DemRaster(torch.from_numpy(np.ascontiguousarray(dem_np)).to(device), sc.dem_spec.origin_lat_deg,
                            sc.dem_spec.origin_lon_deg, sc.dem_spec.pixel_deg)
        ds = cls(sc, raws, dem, api.SceneInputs.from_synthetic(sc), device)
        ds._swc = sc.swath_inputs_c([r.data_ptr() for r in raws], ds._keep)
        return ds

    def swath_inputs(self):
        return self._swc, len(self.sc.swaths)

    def prepare_merged(self, stream=None, out=None) -> api.MergedSigma0:
        return api.prepare_merged_scene(self._swc, len(self.sc.swaths), device=self.device, stream=stream, out=out)

    def merged_layout(self) -> abi.MergedInfo:
        return api.merged_layout(self._swc, len(self.sc.swaths))

    def tc_config(self, **kw) -> api.TerrainCorrectionConfig:
        cfg = api.TerrainCorrectionConfig.new()
        cfg.pixel_spacing_deg = self.sc.pixel_spacing
        cfg.crs = self.sc.crs
        cfg.utm_zone = self.sc.utm_zone
        for k, v in kw.items():
            setattr(cfg, k, v)
        return cfg

    def terrain_correction(self, merged: api.MergedSigma0, cfg: api.TerrainCorrectionConfig, stream=None, out=None,
                           radar_coords: bool = False):
        return api.terrain_correction(merged, self.scene, self.dem, cfg, stream=stream, out=out,
                                      radar_coords=radar_coords)
