"""Seeded synthetic Sentinel-1 IW scenes for the calibrate-and-geocode path.

Nothing here is downloaded: every input (orbit, burst timing, LUTs, CInt16
DN, DEM, geolocation grid, bounding box) is generated from a seed.

Workloads (sizes as named in the scene constructors below):
  * ``scene_small_plain`` — one merged 2048 x 4096 intensity image from CInt16
    DN, LUT calibration, a 1024 x 1024 1-arcsec DEM (flat 200 m plain + three
    Gaussian hills), output 2048 x 2048 EPSG:4326 at 10/111320 deg.
  * ``scene_iw_full`` — 3 subswaths x 9 bursts x 1500 lines x 21000 samples,
    midpoint deburst + IW merge, fBm DEM (Hurst 0.8, 0-3000 m relief) at
    1 arcsec, output a 10 m UTM grid.
  * ``scaled(...)`` variants keep the geometry and shrink the rasters so the
    CPU oracle finishes in seconds.

Orbit: a circular orbit of radius a = 6378137 + 693000 m and inclination
98.18 deg, propagated exactly in inertial space and rotated into ECEF with
the Earth rate; 17 state vectors at 10 s spacing.

Geometry: right-looking zero-Doppler forward geocoding (the ellipsoid /
range / Doppler system that sardine/src/insar/coreg.rs:280-362 solves), used
for the geolocation grid, the footprint and the bounding box.

The CInt16 generator is counter based (splitmix64 + Box-Muller), so any
sub-block of a large raster can be regenerated independently on the host;
the device kernel sardine_synth_cint16 implements the same function.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

WGS84_A = 6378137.0
WGS84_F = 1.0 / 298.257223563
WGS84_E2 = 2.0 * WGS84_F - WGS84_F * WGS84_F
WGS84_B = WGS84_A * (1.0 - WGS84_F)
GM = 3.986004418e14
OMEGA_E = 7.2921150e-5
C_LIGHT = 299792458.0
METRES_PER_DEG = 111320.0


# ─── geodesy (numpy) ──────────────────────────────────────────────────────
def geodetic_to_ecef(lat_deg, lon_deg, h):
    lat = np.radians(lat_deg)
    lon = np.radians(lon_deg)
    sl, cl = np.sin(lat), np.cos(lat)
    n = WGS84_A / np.sqrt(1.0 - WGS84_E2 * sl * sl)
    return np.stack([(n + h) * cl * np.cos(lon), (n + h) * cl * np.sin(lon),
                     (n * (1.0 - WGS84_E2) + h) * sl], axis=-1)


def ecef_to_geodetic(p):
    p = np.asarray(p, dtype=np.float64)
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    lon = np.arctan2(y, x)
    r = np.hypot(x, y)
    lat = np.arctan2(z, r * (1.0 - WGS84_E2))
    for _ in range(6):
        sl = np.sin(lat)
        n = WGS84_A / np.sqrt(1.0 - WGS84_E2 * sl * sl)
        h = r / np.cos(lat) - n
        lat = np.arctan2(z, r * (1.0 - WGS84_E2 * n / (n + h)))
    sl = np.sin(lat)
    n = WGS84_A / np.sqrt(1.0 - WGS84_E2 * sl * sl)
    h = r / np.cos(lat) - n
    return np.degrees(lat), np.degrees(lon), h


# ─── orbit ────────────────────────────────────────────────────────────────
@dataclass
class CircularOrbit:
    """Exact circular orbit in inertial space, expressed in ECEF.

    ECI and ECEF coincide at t = t_mid_s (Greenwich angle 0 there)."""
    a: float
    inc_rad: float
    raan_rad: float
    u_mid_rad: float
    t_mid_s: float

    @property
    def mean_motion(self) -> float:
        return math.sqrt(GM / self.a ** 3)

    def state(self, t_s):
        t = np.asarray(t_s, dtype=np.float64)
        u = self.u_mid_rad + self.mean_motion * (t - self.t_mid_s)
        ci, si = math.cos(self.inc_rad), math.sin(self.inc_rad)
        cO, sO = math.cos(self.raan_rad), math.sin(self.raan_rad)
        P = np.array([cO, sO, 0.0])
        Q = np.array([-ci * sO, ci * cO, si])
        cu, su = np.cos(u)[..., None], np.sin(u)[..., None]
        r_eci = self.a * (cu * P + su * Q)
        v_eci = self.a * self.mean_motion * (-su * P + cu * Q)
        th = OMEGA_E * (t - self.t_mid_s)
        ct, st = np.cos(th)[..., None], np.sin(th)[..., None]

        def rot(v):  # Rz(-theta)
            return np.stack([ct[..., 0] * v[..., 0] + st[..., 0] * v[..., 1],
                             -st[..., 0] * v[..., 0] + ct[..., 0] * v[..., 1], v[..., 2]], axis=-1)

        r = rot(r_eci)
        v = rot(v_eci)
        # v_ecef = R v_eci − ω × r_ecef
        v = v - np.stack([-OMEGA_E * r[..., 1], OMEGA_E * r[..., 0], np.zeros_like(r[..., 0])], axis=-1)
        return r, v


def make_orbit(sat_lat_deg: float, sat_lon_deg: float, t_mid_s: float,
               altitude_m: float = 693000.0, inc_deg: float = 98.18) -> CircularOrbit:
    """Circular orbit whose (geocentric) sub-satellite point at t_mid_s is
    (sat_lat_deg, sat_lon_deg), on an ascending pass."""
    inc = math.radians(inc_deg)
    a = WGS84_A + altitude_m
    u = math.asin(math.sin(math.radians(sat_lat_deg)) / math.sin(inc))
    raan = math.radians(sat_lon_deg) - math.atan2(math.cos(inc) * math.sin(u), math.cos(u))
    return CircularOrbit(a, inc, raan, u, t_mid_s)


def state_vectors(orbit: CircularOrbit, n: int = 17, spacing_s: float = 10.0, t0_s: float = 0.0):
    """n ECEF state vectors at t0 + k*spacing (times in µs on the scene epoch)."""
    t = t0_s + spacing_s * np.arange(n, dtype=np.float64)
    r, v = orbit.state(t)
    t_us = np.round(t * 1e6).astype(np.int64)
    return t_us, r, v


# ─── forward zero-Doppler geocoding ───────────────────────────────────────
def forward_geocode(orbit: CircularOrbit, t_s, slant_m, h_m=0.0):
    """Right-looking zero-Doppler ground point (lat, lon) for azimuth time t_s
    and slant range slant_m at ellipsoidal height h_m (vectorised)."""
    t_s, slant_m = np.broadcast_arrays(np.asarray(t_s, float), np.asarray(slant_m, float))
    h_m = np.broadcast_to(np.asarray(h_m, float), t_s.shape)
    S, V = orbit.state(t_s)
    vhat = V / np.linalg.norm(V, axis=-1, keepdims=True)
    shat = S / np.linalg.norm(S, axis=-1, keepdims=True)
    s_perp = shat - np.sum(shat * vhat, axis=-1, keepdims=True) * vhat
    s_perp /= np.linalg.norm(s_perp, axis=-1, keepdims=True)
    right = np.cross(vhat, s_perp)
    Rs = np.linalg.norm(S, axis=-1)
    Re = 6367000.0 + h_m
    cos_th = (Rs ** 2 + slant_m ** 2 - Re ** 2) / (2 * Rs * slant_m)
    th = np.arccos(np.clip(cos_th, -1, 1))
    P = S + slant_m[..., None] * (-np.cos(th)[..., None] * s_perp + np.sin(th)[..., None] * right)
    lat, lon, _ = ecef_to_geodetic(P)
    for _ in range(8):
        def resid(la, lo):
            Pp = geodetic_to_ecef(la, lo, h_m)
            d = Pp - S
            return np.linalg.norm(d, axis=-1) - slant_m, np.sum(d * V, axis=-1) / np.linalg.norm(V, axis=-1)
        f1, f2 = resid(lat, lon)
        e = 1e-6
        a1, a2 = resid(lat + e, lon)
        b1, b2 = resid(lat, lon + e)
        j11, j21 = (a1 - f1) / e, (a2 - f2) / e
        j12, j22 = (b1 - f1) / e, (b2 - f2) / e
        det = j11 * j22 - j12 * j21
        dlat = (-f1 * j22 + f2 * j12) / det
        dlon = (-j11 * f2 + j21 * f1) / det
        lat = lat + dlat
        lon = lon + dlon
    return lat, lon


# ─── counter-based CInt16 generator ───────────────────────────────────────
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def complex_gaussian_dn(seed: int, index_base: int, n: int, sigma: float = 100.0) -> np.ndarray:
    """n CInt16 samples (shape (n, 2), int16): round(sigma * N(0,1)) per part.

    Element i uses key k = seed * 0x9E3779B97F4A7C15 + index_base + i and
    u1 = top53(splitmix64(2k)), u2 = top53(splitmix64(2k+1));
    r = sqrt(-2 ln(1-u1)), re = r cos(2 pi u2), im = r sin(2 pi u2);
    DN = clamp(floor(sigma*x + 0.5), -32767, 32767)."""
    idx = np.arange(n, dtype=np.uint64) + np.uint64(index_base)
    with np.errstate(over="ignore"):
        k = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + idx
        h1 = _splitmix64(k * np.uint64(2))
        h2 = _splitmix64(k * np.uint64(2) + np.uint64(1))
    u1 = (h1 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    ang = 2.0 * np.pi * u2
    out = np.empty((n, 2), dtype=np.int16)
    out[:, 0] = np.clip(np.floor(sigma * r * np.cos(ang) + 0.5), -32767, 32767)
    out[:, 1] = np.clip(np.floor(sigma * r * np.sin(ang) + 0.5), -32767, 32767)
    return out


# ─── LUTs ─────────────────────────────────────────────────────────────────
@dataclass
class LutVectors:
    """Flattened calibration/noise range LUT (calibration.rs:38-82)."""
    lines: np.ndarray    # int32 (n_vectors,)
    offsets: np.ndarray  # int64 (n_vectors+1,)
    pixels: np.ndarray   # uint32
    values: np.ndarray   # float32

    def vector(self, i):
        a, b = int(self.offsets[i]), int(self.offsets[i + 1])
        return self.pixels[a:b], self.values[a:b]

    def as_c(self) -> abi.Lut:
        return abi.Lut(len(self.lines), self.lines.ctypes.data, self.offsets.ctypes.data,
                       self.pixels.ctypes.data, self.values.ctypes.data)


@dataclass
class AzLutVector:
    lines: np.ndarray   # int32
    values: np.ndarray  # float32

    def as_c(self) -> abi.AzLut:
        return abi.AzLut(len(self.lines), self.lines.ctypes.data if len(self.lines) else None,
                         self.values.ctypes.data if len(self.values) else None)


def make_range_lut(rng, vec_lines, raw_cols, step, lo, hi, phase=0.0) -> LutVectors:
    """Values in [lo, hi] varying smoothly in range on a `step`-pixel grid."""
    pix = np.arange(0, raw_cols + step, step, dtype=np.int64)
    pix = pix[pix <= raw_cols - 1 + step - 1]
    pix = np.unique(np.minimum(pix, raw_cols - 1)).astype(np.uint32)
    lines, offs, allp, allv = [], [0], [], []
    for j, ln in enumerate(vec_lines):
        x = pix.astype(np.float64) / max(raw_cols - 1, 1)
        wob = 0.02 * math.sin(0.7 * j + phase + rng.uniform(0, 6.28))
        v = lo + (hi - lo) * (0.5 - 0.5 * np.cos(np.pi * np.clip(x + wob, 0, 1)))
        v = np.clip(v, lo, hi).astype(np.float32)
        lines.append(int(ln))
        allp.append(pix)
        allv.append(v)
        offs.append(offs[-1] + len(pix))
    return LutVectors(np.array(lines, np.int32), np.array(offs, np.int64),
                      np.concatenate(allp).astype(np.uint32), np.concatenate(allv).astype(np.float32))


# ─── DEM ──────────────────────────────────────────────────────────────────
@dataclass
class DemSpec:
    width: int
    height: int
    origin_lat_deg: float   # centre of pixel (0, 0)
    origin_lon_deg: float
    pixel_deg: float

    def lat_lon_grids(self):
        lat = self.origin_lat_deg - np.arange(self.height) * self.pixel_deg
        lon = self.origin_lon_deg + np.arange(self.width) * self.pixel_deg
        return lat, lon


def dem_hills(spec: DemSpec, rng, centre_lat, centre_lon, extent_deg) -> np.ndarray:
    """200 m plain + 3 Gaussian hills (peaks 500-1500 m, sigma 1-3 km)."""
    lat, lon = spec.lat_lon_grids()
    LAT, LON = np.meshgrid(lat, lon, indexing="ij")
    h = np.full(LAT.shape, 200.0)
    coslat = math.cos(math.radians(centre_lat))
    for _ in range(3):
        clat = centre_lat + rng.uniform(-0.35, 0.35) * extent_deg
        clon = centre_lon + rng.uniform(-0.35, 0.35) * extent_deg
        peak = rng.uniform(500.0, 1500.0)
        sig = rng.uniform(1000.0, 3000.0)
        dy = (LAT - clat) * METRES_PER_DEG
        dx = (LON - clon) * METRES_PER_DEG * coslat
        h += (peak - 200.0) * np.exp(-(dx * dx + dy * dy) / (2 * sig * sig))
    return h.astype(np.float32)


def dem_fbm(spec: DemSpec, seed: int, hurst: float = 0.8, relief_m: float = 3000.0, device=None) -> np.ndarray:
    """fBm-like terrain by octave value noise (amplitude 2^(-H*octave)),
    scaled to [0, relief_m].  Uses torch for the upsampling (CPU or GPU)."""
    import torch
    import torch.nn.functional as F

    g = torch.Generator(device="cpu").manual_seed(seed)
    H, W = spec.height, spec.width
    dev = device or "cpu"
    acc = torch.zeros((1, 1, H, W), dtype=torch.float32, device=dev)
    base = 4
    amp = 1.0
    octave = 0
    while base <= max(H, W):
        noise = torch.rand((1, 1, base + 1, base + 1), generator=g, dtype=torch.float32).to(dev)
        acc += amp * F.interpolate(noise, size=(H, W), mode="bilinear", align_corners=True)
        base *= 2
        amp *= 2.0 ** (-hurst)
        octave += 1
    a = acc[0, 0]
    a = (a - a.min()) / (a.max() - a.min()) * relief_m
    return a.cpu().numpy().astype(np.float32)


# ─── scene description ────────────────────────────────────────────────────
@dataclass
class SwathSpec:
    index: int
    n_bursts: int
    lines_per_burst: int
    raw_cols: int
    slant_range_time_s: float
    burst_times_us: np.ndarray
    first_valid: np.ndarray
    last_valid: np.ndarray
    cal: LutVectors
    noise: LutVectors
    noise_az: AzLutVector
    seed: int

    @property
    def tiff_height(self):
        return self.n_bursts * self.lines_per_burst

    def bursts_c(self):
        arr = (abi.Burst * self.n_bursts)()
        for k in range(self.n_bursts):
            arr[k] = abi.Burst(int(self.burst_times_us[k]), k * self.lines_per_burst,
                               int(self.first_valid[k]), int(self.last_valid[k]), k)
        return arr

    def raw_host(self) -> np.ndarray:
        """The whole swath TIFF raster as (tiff_height, raw_cols, 2) int16."""
        n = self.tiff_height * self.raw_cols
        return complex_gaussian_dn(self.seed, 0, n).reshape(self.tiff_height, self.raw_cols, 2)


@dataclass
class SyntheticScene:
    name: str
    seed: int
    orbit: CircularOrbit
    sv_t_us: np.ndarray
    sv_pos: np.ndarray
    sv_vel: np.ndarray
    ati: float
    rps_m: float
    swaths: list
    bbox: tuple            # (min_lat, max_lat, min_lon, max_lon)
    grid_points: np.ndarray  # (n, 2): lat, t_us (float64 / int64 in two arrays)
    grid_t_us: np.ndarray
    dem_spec: DemSpec
    dem_kind: str
    crs: int = abi.CRS_WGS84
    utm_zone: int = 0
    pixel_spacing: float = 10.0 / METRES_PER_DEG
    _dem: np.ndarray = field(default=None, repr=False)

    # ── ctypes views ──
    def orbit_c(self):
        n = len(self.sv_t_us)
        arr = (abi.StateVector * n)()
        for i in range(n):
            arr[i].time_us = int(self.sv_t_us[i])
            for c in range(3):
                arr[i].position_m[c] = float(self.sv_pos[i, c])
                arr[i].velocity_m_s[c] = float(self.sv_vel[i, c])
        return arr

    def scene_c(self, keep: list) -> abi.Scene:
        svs = self.orbit_c()
        gp = (abi.GridPoint * len(self.grid_t_us))()
        for i in range(len(self.grid_t_us)):
            gp[i] = abi.GridPoint(float(self.grid_points[i]), int(self.grid_t_us[i]))
        keep += [svs, gp]
        bb = abi.BBox(*self.bbox)
        return abi.Scene(abi.Orbit(svs, len(self.sv_t_us)), bb, self.ati, gp, len(self.grid_t_us),
                         len(self.swaths))

    def tc_config(self, **kw) -> abi.TcConfig:
        cfg = abi.TcConfig()
        cfg.pixel_spacing = self.pixel_spacing
        cfg.crs = self.crs
        cfg.utm_zone = self.utm_zone
        cfg.max_newton_iterations = 10
        cfg.newton_tolerance_s = 1e-6
        cfg.flatten = 0
        cfg.compute_lia = 0
        cfg.noise_floor_margin_db = float("-inf")
        cfg.resampling = abi.RESAMPLE_BILINEAR
        cfg.projector_anchors = 256
        cfg.precision = abi.PRECISION_FP64
        cfg.geoid = abi.Geoid(abi.GEOID_ZERO, None)
        for k, v in kw.items():
            setattr(cfg, k, v)
        return cfg

    def dem(self, device=None) -> np.ndarray:
        if self._dem is None:
            if self.dem_kind == "hills":
                rng = np.random.default_rng(self.seed + 7)
                lat0 = 0.5 * (self.bbox[0] + self.bbox[1])
                lon0 = 0.5 * (self.bbox[2] + self.bbox[3])
                self._dem = dem_hills(self.dem_spec, rng, lat0, lon0, self.bbox[1] - self.bbox[0])
            else:
                self._dem = dem_fbm(self.dem_spec, self.seed + 7, device=device)
        return self._dem

    def swath_inputs_c(self, raw_ptrs, keep: list):
        """sardine_swath_input array; raw_ptrs[i] is the device (or host, for
        the oracle) pointer of swath i's TIFF raster."""
        arr = (abi.SwathInput * len(self.swaths))()
        for i, sw in enumerate(self.swaths):
            b = sw.bursts_c()
            keep.append(b)
            arr[i].raw = raw_ptrs[i]
            arr[i].raw_cols = sw.raw_cols
            arr[i].tiff_height = sw.tiff_height
            arr[i].swath = abi.SubSwath(sw.n_bursts, sw.lines_per_burst, self.ati, sw.slant_range_time_s,
                                        self.rps_m)
            arr[i].bursts = b
            arr[i].n_bursts = sw.n_bursts
            arr[i].cal = sw.cal.as_c()
            arr[i].noise_range = sw.noise.as_c()
            arr[i].noise_azimuth = sw.noise_az.as_c()
        return arr


def _dem_spec_covering(bbox, margin_deg, pixel_deg=1.0 / 3600.0, max_side=None):
    lat_top = bbox[1] + margin_deg
    lat_bot = bbox[0] - margin_deg
    lon_w = bbox[2] - margin_deg
    lon_e = bbox[3] + margin_deg
    h = int(math.ceil((lat_top - lat_bot) / pixel_deg)) + 1
    w = int(math.ceil((lon_e - lon_w) / pixel_deg)) + 1
    if max_side:
        h, w = min(h, max_side), min(w, max_side)
    return DemSpec(w, h, lat_top, lon_w, pixel_deg)


def _aligned_bbox(centre_lat, centre_lon, rows, cols, sp):
    """A bbox whose snapped output grid (terrain_correction.rs:1174-1216) is
    exactly rows x cols at spacing sp."""
    k_lat = math.floor(centre_lat / sp) - rows // 2
    k_lon = math.floor(centre_lon / sp) - cols // 2
    return ((k_lat + 0.25) * sp, (k_lat + rows - 0.25) * sp, (k_lon + 0.25) * sp, (k_lon + cols - 0.25) * sp)


def _grid_points(orbit, t_start_s, ati, lines, near_range, rps, samples, n_az=10, n_rg=21, h=0.0):
    ls = np.linspace(0, lines - 1, n_az)
    ss = np.linspace(0, samples - 1, n_rg)
    L, S = np.meshgrid(ls, ss, indexing="ij")
    t = t_start_s + L * ati
    lat, lon = forward_geocode(orbit, t.ravel(), (near_range + S * rps).ravel(), h)
    t_us = np.round(t.ravel() * 1e6).astype(np.int64)
    return lat, t_us


def scene_small_plain(seed: int = 0, lines: int = 2048, samples: int = 4096, out_rows: int = 2048,
                      out_cols: int = 2048, dem_side: int = 1024, sat_lon_deg: float = None) -> SyntheticScene:
    """Plain method, small: merged 2048 x 4096 intensity from CInt16 DN, LUT
    calibration, 1-arcsec 1024 x 1024 DEM of hills, 2048 x 2048 EPSG:4326
    output at 10 m (10/111320 deg).  Smaller arguments give scaled copies of
    the same geometry."""
    rng = np.random.default_rng(seed)
    ati = 2.055556e-3
    rps = 2.329562
    near_range = 800000.0
    t_mid = 80.0
    sat_lat = 45.3 + 0.2 * rng.uniform(-1, 1)
    sat_lon = 4.4 + 0.3 * rng.uniform(-1, 1)  # drawn in any case: the seed's stream stays the same
    # (the scene centre lies ~4.7° east of the sub-satellite longitude)
    orbit = make_orbit(sat_lat, sat_lon if sat_lon_deg is None else sat_lon_deg, t_mid)
    t_us, pos, vel = state_vectors(orbit)
    t_start_s = t_mid - 0.5 * lines * ati
    t_start_us = int(round(t_start_s * 1e6))
    t_start_s = t_start_us / 1e6
    # scene centre
    clat, clon = forward_geocode(orbit, t_start_s + 0.5 * lines * ati, near_range + 0.5 * samples * rps, 200.0)
    sp = 10.0 / METRES_PER_DEG
    bbox = _aligned_bbox(float(clat), float(clon), out_rows, out_cols, sp)
    gl, gt = _grid_points(orbit, t_start_s, ati, lines, near_range, rps, samples)
    dem_spec = DemSpec(dem_side, dem_side, 0.5 * (bbox[0] + bbox[1]) + 0.5 * dem_side / 3600.0,
                       0.5 * (bbox[2] + bbox[3]) - 0.5 * dem_side / 3600.0, 1.0 / 3600.0)
    # one swath, one burst of `lines` lines; LUT vectors at the burst bounds
    vec_lines = [0, lines // 2, lines]
    cal = make_range_lut(rng, vec_lines, samples, 40, 400.0, 700.0)
    noise = make_range_lut(rng, vec_lines, samples, 40, 50.0, 150.0, phase=1.0)
    az_lines = np.linspace(0, lines, 5).astype(np.int32)
    noise_az = AzLutVector(az_lines, (1.0 + 0.05 * np.sin(np.arange(5))).astype(np.float32))
    sw = SwathSpec(0, 1, lines, samples, 2.0 * near_range / C_LIGHT, np.array([t_start_us], np.int64),
                   np.array([0]), np.array([samples]), cal, noise, noise_az, seed * 1000 + 1)
    return SyntheticScene("small_plain", seed, orbit, t_us, pos, vel, ati, rps, [sw], bbox, gl, gt,
                          dem_spec, "hills")


def scene_iw_full(seed: int = 1, n_swaths: int = 3, n_bursts: int = 9, lines_per_burst: int = 1500,
                  raw_cols: int = 21000, overlap_lines: int = 50, swath_overlap: int = 900,
                  out_rows: int = None, out_cols: int = None, utm: bool = True,
                  dem_max_side: int = None) -> SyntheticScene:
    """Full single-IW-scene: 3 subswaths x 9 bursts of 1500 x 21000 CInt16,
    integer midpoint deburst (burst overlap `overlap_lines`) and IW merge,
    fBm DEM at 1 arcsec, 10 m UTM output over the footprint's bounding box.  Smaller arguments give scaled
    copies with the same structure."""
    rng = np.random.default_rng(seed)
    ati = 2.055556e-3
    rps = 2.329562
    near_range = 800000.0
    dt_s = 2.0 * rps / C_LIGHT
    t_mid = 80.0
    orbit = make_orbit(45.0 + 0.2 * rng.uniform(-1, 1), 4.6 + 0.3 * rng.uniform(-1, 1), t_mid)
    t_us, pos, vel = state_vectors(orbit)
    step_lines = lines_per_burst - overlap_lines
    deb_lines = n_bursts * lines_per_burst - (n_bursts - 1) * overlap_lines
    t_start_s = t_mid - 0.5 * deb_lines * ati
    t_start_us = int(round(t_start_s * 1e6))
    swaths = []
    valid_w = []
    srt = 2.0 * near_range / C_LIGHT
    for i in range(n_swaths):
        bt = np.array([t_start_us + i * 300 + int(round(k * step_lines * ati * 1e6)) for k in range(n_bursts)],
                      np.int64)
        fv = 60 + rng.integers(0, 40, n_bursts)
        lv = raw_cols - 60 - rng.integers(0, 40, n_bursts)
        vec_lines = [k * lines_per_burst for k in range(n_bursts + 1)]
        cal = make_range_lut(rng, vec_lines, raw_cols, 40, 400.0, 700.0, phase=i)
        noise = make_range_lut(rng, vec_lines, raw_cols, 40, 50.0, 150.0, phase=1.0 + i)
        az_l = np.linspace(0, n_bursts * lines_per_burst, 2 * n_bursts + 1).astype(np.int32)
        noise_az = AzLutVector(az_l, (1.0 + 0.05 * np.sin(np.arange(len(az_l)) + i)).astype(np.float32))
        swaths.append(SwathSpec(i, n_bursts, lines_per_burst, raw_cols, srt, bt, fv, lv, cal, noise, noise_az,
                                seed * 1000 + 11 + i))
        valid_w.append(int(lv.min() - fv.max()))
        # next swath's TIFF column 0 sits `swath_overlap` samples before this
        # swath's valid far edge
        srt = srt + (int(lv.min()) - swath_overlap) * dt_s
    total_samples = int(round((srt - 2.0 * near_range / C_LIGHT) / dt_s)) + swath_overlap
    far_range = near_range + total_samples * rps
    clat, clon = forward_geocode(orbit, t_start_s + 0.5 * deb_lines * ati, 0.5 * (near_range + far_range), 300.0)
    clat, clon = float(clat), float(clon)
    crs = abi.CRS_UTM_NORTH if utm else abi.CRS_WGS84
    zone = int((clon + 180.0) // 6.0) + 1
    # footprint corners → an inner lat/lon bbox
    corners_t = np.array([0, 0, deb_lines - 1, deb_lines - 1]) * ati + t_start_s
    corners_r = np.array([near_range, far_range, near_range, far_range])
    flat, flon = forward_geocode(orbit, corners_t, corners_r, 300.0)
    if out_rows is None:
        # the footprint's bounding box, as an annotation's geolocation grid gives it
        bbox = (float(flat.min()), float(flat.max()), float(flon.min()), float(flon.max()))
    elif utm:
        # a lat/lon box spanning about out_rows x out_cols 10 m cells around
        # the scene centre (the grid is the UTM extent of its corners)
        half_lat = 0.5 * out_rows * 10.0 / METRES_PER_DEG
        half_lon = 0.5 * out_cols * 10.0 / (METRES_PER_DEG * math.cos(math.radians(clat)))
        bbox = (clat - half_lat, clat + half_lat, clon - half_lon, clon + half_lon)
    else:
        sp_deg = 10.0 / METRES_PER_DEG
        bbox = _aligned_bbox(clat, clon, out_rows, out_cols, sp_deg)
    gl, gt = _grid_points(orbit, t_start_s, ati, deb_lines, near_range, rps, total_samples)
    dem_spec = _dem_spec_covering(bbox, 0.03, max_side=dem_max_side)
    sc = SyntheticScene("iw_full", seed, orbit, t_us, pos, vel, ati, rps, swaths, bbox, gl, gt, dem_spec, "fbm",
                        crs=crs, utm_zone=zone if utm else 0, pixel_spacing=10.0 if utm else 10.0 / METRES_PER_DEG)
    return sc
