// sardine_oracle.cpp — CPU restatement of SARdine's calibrate-and-geocode
// path, used ONLY as test infrastructure: the checker in tests/, the
// reference side of __graft_entry__.smoke(), and the CPU baseline leg of
// bench.py.  Nothing in the product (sardine_amd/, include/) calls it.
//
// It follows the reference Rust sources function by function (file:line cited
// on each), in f64/f32 exactly where the reference uses them, compiled with
// -ffp-contract=off so that no fused multiply-add changes a rounding the Rust
// code performs separately.  Row chunks of 64 are run on a thread pool, the
// reference's rayon structure (terrain_correction.rs:1414-1417).
//
// Parity pins (what checks each function; see tests/test_oracle_*.py):
//   geodetic_to_ecef           reference tests geodesy.rs:141-221 + closed-form
//                              WGS84 values + Bowring round trip
//   interpolate_orbit_and_accel exactness on degree-≤7 polynomials (linear
//                              orbit, orbit.rs:984-1030), zero accel on a linear
//                              orbit (orbit.rs:1238-1275), accel vs central FD
//   solve_zero_doppler         nadir (terrain_correction.rs:2217-2247), forced
//                              non-convergence (:2250-2275), residual invariant
//                              and brute-force bisection on the root
//   resamplers                 terrain_correction.rs:2322-2454, 2751-2809
//   terrain geometry           terrain_correction.rs:1982-2183, 2558-2600
//   calibration                apply_calibration.rs:583-747 known values
//   deburst / merge layout     deburst.rs:467-503, merge_subswaths.rs:500-725
//   speckle                    speckle.rs:621-1000 known values
//   to_db                      export.rs:1227-1275 known values
//   downsample_2x / COG writer export.rs:1583-1733 (tile-size and geotransform
//                              errors, header and 17-entry IFDs, tile data
//                              round trip, 2x2 means and NaN rule)
//   apply_multilook            run.rs:1449-1588 (pass-through, uniform and
//                              distinct averages, NaN exclusion, all-NaN
//                              block, geometry, azimuth shift, errors)
//   to_ground_range            ground_range.rs:506-724 (ground spacing, column
//                              map monotonic / covering, range interpolator,
//                              uniform synthetic scene, the three rejections)
//   geoid grid files           geoid.rs:596-604 (wrong size) + round trips,
//                              the WW15MGH header rules, malformed tokens
//   SRTM tiles + mosaic        dem.rs:905-1026 style known values on synthetic
//                              tiles, the name rules, void → missing, edges
//   SLC TIFF reader            slc_reader.rs:618-820 (non-TIFF, empty, missing
//                              file, SamplesPerPixel 2, line bounds, row count)
//   striped GeoTIFF writers    export.rs:1276-1580 (geotransform errors, key
//                              offsets, BigTIFF threshold and layout, u8 tags,
//                              the no-CRS IFD, UTM keys)
//   UTM                        published Krüger-series (Karney 2011) values,
//                              forward/inverse round trip (proj4rs is not
//                              vendored: parity unpinned below ±1 m)
#include "../include/sardine_hip.h"
// The libm of this restatement: the reference's results depend on the
// platform's sin/cos/atan/... (Rust's f64 methods call the system libm), so
// the oracle fixes one, a portable libm built from correctly rounded IEEE
// operations (pinned to ~1 ulp of 50-digit values by the tests), which the
// GPU kernels evaluate with the same bits.  The geodesy and the projection
// use it; everything else keeps <cmath>.
#include "../sardine_amd/csrc/pmath.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <limits>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;
// geodesy.rs:13-25
constexpr double WGS84_A = 6378137.0;
constexpr double WGS84_F = 1.0 / 298.257223563;
constexpr double WGS84_B = WGS84_A * (1.0 - WGS84_F);
constexpr double WGS84_E2 = 2.0 * WGS84_F - WGS84_F * WGS84_F;
constexpr double WGS84_EP2 = WGS84_E2 / (1.0 - WGS84_E2);
constexpr double SPEED_OF_LIGHT = 299792458.0;  // types.rs:11

struct V3 {
    double x, y, z;
};
inline V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm3(V3 v) { return std::sqrt(dot3(v, v)); }
inline V3 cross3(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 normalize3(V3 v) {  // geodesy.rs:125-132
    double n = norm3(v);
    if (n == 0.0) return {0.0, 0.0, 0.0};
    return {v.x / n, v.y / n, v.z / n};
}

// geodesy.rs:38-55
V3 geodetic_to_ecef(double lat_deg, double lon_deg, double h) {
    double lat = lat_deg * kPi / 180.0;
    double lon = lon_deg * kPi / 180.0;
    double sl, cl, so, co;
    pm::sincos(lat, &sl, &cl);
    pm::sincos(lon, &so, &co);
    double n = WGS84_A / std::sqrt(1.0 - WGS84_E2 * sl * sl);
    return {(n + h) * cl * co, (n + h) * cl * so, (n * (1.0 - WGS84_E2) + h) * sl};
}

// geodesy.rs:63-94
void ecef_to_geodetic(V3 e, double* lat_o, double* lon_o, double* h_o) {
    double lon = std::atan2(e.y, e.x);
    double p = std::sqrt(e.x * e.x + e.y * e.y);
    double theta = std::atan2(e.z * WGS84_A, p * WGS84_B);
    double st = std::sin(theta), ct = std::cos(theta);
    double lat = std::atan2(e.z + WGS84_EP2 * WGS84_B * st * st * st,
                            p - WGS84_E2 * WGS84_A * ct * ct * ct);
    for (int i = 0; i < 5; ++i) {
        double sl = std::sin(lat);
        double n = WGS84_A / std::sqrt(1.0 - WGS84_E2 * sl * sl);
        double nl = std::atan2(e.z + WGS84_E2 * n * sl, p);
        if (std::fabs(nl - lat) < 1e-12) break;
        lat = nl;
    }
    double sl = std::sin(lat);
    double n = WGS84_A / std::sqrt(1.0 - WGS84_E2 * sl * sl);
    double h = std::fabs(std::cos(lat)) > 1e-10 ? p / std::cos(lat) - n
                                                : e.z / std::sin(lat) - n * (1.0 - WGS84_E2);
    *lat_o = lat * 180.0 / kPi;
    *lon_o = lon * 180.0 / kPi;
    *h_o = h;
}

// ── Orbit (orbit.rs:339-596) ─────────────────────────────────────────────
constexpr int LAGRANGE_DEGREE = 8;

struct Orbit {
    std::vector<int64_t> t;   // µs relative to the first state vector
    std::vector<V3> pos, vel;
};

Orbit make_orbit(const sardine_orbit* o) {
    Orbit r;
    int64_t t0 = o->count > 0 ? o->state_vectors[0].time_us : 0;
    for (int i = 0; i < o->count; ++i) {
        const sardine_state_vector& s = o->state_vectors[i];
        r.t.push_back(s.time_us - t0);
        r.pos.push_back({s.position_m[0], s.position_m[1], s.position_m[2]});
        r.vel.push_back({s.velocity_m_s[0], s.velocity_m_s[1], s.velocity_m_s[2]});
    }
    return r;
}

// interpolate_orbit_and_accel, orbit.rs:498-596.  `t` is µs relative to the
// first state vector.  Returns 0, or an OrbitError code.
int interp_orbit_accel(const Orbit& o, int64_t t, V3* pos, V3* vel, V3* acc) {
    const int n = (int)o.t.size();
    if (n < LAGRANGE_DEGREE) return SARDINE_ERR_TC_ORBIT_INSUFFICIENT;
    if (t < o.t[0] || t > o.t[n - 1]) return SARDINE_ERR_TC_ORBIT_OUTSIDE_WINDOW;
    double t_sec = (double)(t - o.t[0]) / 1e6;
    // centre: largest i with t_i <= t (orbit.rs:522-528; the hint never
    // changes the result)
    int centre = 0;
    while (centre + 1 < n && o.t[centre + 1] <= t) ++centre;
    const int half = LAGRANGE_DEGREE / 2;
    int start;
    if (centre < half) start = 0;
    else if (centre + half >= n) start = n - LAGRANGE_DEGREE;
    else start = centre - half + 1;
    double times[LAGRANGE_DEGREE];
    for (int i = 0; i < LAGRANGE_DEGREE; ++i) times[i] = (double)(o.t[start + i] - o.t[0]) / 1e6;
    double rdiff[LAGRANGE_DEGREE] = {0};
    double total_rdiff = 0.0;
    for (int j = 0; j < LAGRANGE_DEGREE; ++j) {
        double diff = t_sec - times[j];
        if (std::fabs(diff) > 1e-9) {
            rdiff[j] = 1.0 / diff;
            total_rdiff += rdiff[j];
        }
    }
    V3 p = {0, 0, 0}, v = {0, 0, 0}, a = {0, 0, 0};
    for (int k = 0; k < LAGRANGE_DEGREE; ++k) {
        double basis = 1.0;
        for (int j = 0; j < LAGRANGE_DEGREE; ++j)
            if (j != k) basis *= (t_sec - times[j]) / (times[k] - times[j]);
        double d_basis = basis * (total_rdiff - rdiff[k]);
        const V3& P = o.pos[start + k];
        const V3& W = o.vel[start + k];
        p.x += basis * P.x; p.y += basis * P.y; p.z += basis * P.z;
        v.x += basis * W.x; v.y += basis * W.y; v.z += basis * W.z;
        a.x += d_basis * W.x; a.y += d_basis * W.y; a.z += d_basis * W.z;
    }
    *pos = p; *vel = v; *acc = a;
    return 0;
}

enum { ZD_CONVERGED = 0, ZD_NOT_CONVERGED = 1, ZD_DEGENERATE = 2 };

// solve_zero_doppler, terrain_correction.rs:1010-1073.  Returns an outcome
// (>=0) or a negative OrbitError code.
int solve_zero_doppler(V3 target, const Orbit& o, int64_t seed, int max_iter, double tol,
                       int64_t* t_out, int* iters, V3* sat_pos) {
    const double DERIV_FLOOR = 1.0;
    int64_t t = seed;
    for (int i = 0; i < max_iter; ++i) {
        V3 p, v, a;
        int e = interp_orbit_accel(o, t, &p, &v, &a);
        if (e) return -e;
        V3 r = sub3(target, p);
        double f = dot3(r, v);
        double v_sq = dot3(v, v);
        double fp = -v_sq + dot3(r, a);
        if (std::fabs(fp) < DERIV_FLOOR) return ZD_DEGENERATE;
        double delta = -f / fp;
        if (std::fabs(delta) < tol) {
            *t_out = t;
            *iters = std::min(i + 1, 255);
            *sat_pos = p;
            return ZD_CONVERGED;
        }
        t = t + (int64_t)(delta * 1e6);  // Duration::microseconds((δ·1e6) as i64)
    }
    return ZD_NOT_CONVERGED;
}

// ── Resampling (terrain_correction.rs:770-983) ────────────────────────────
float bilinear_sample_slice(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    int64_t max_line = nl > 0 ? nl - 1 : 0, max_sample = ns > 0 ? ns - 1 : 0;
    int64_t l0 = std::clamp<int64_t>((int64_t)std::floor(line), 0, max_line);
    int64_t l1 = std::min(l0 + 1, max_line);
    int64_t s0 = std::clamp<int64_t>((int64_t)std::floor(sample), 0, max_sample);
    int64_t s1 = std::min(s0 + 1, max_sample);
    double dl = line - (double)l0, ds = sample - (double)s0;
    float v00 = d[l0 * ns + s0], v01 = d[l0 * ns + s1], v10 = d[l1 * ns + s0], v11 = d[l1 * ns + s1];
    if (std::isnan(v00) || std::isnan(v01) || std::isnan(v10) || std::isnan(v11))
        return std::numeric_limits<float>::quiet_NaN();
    double w00 = (1.0 - dl) * (1.0 - ds), w01 = (1.0 - dl) * ds, w10 = dl * (1.0 - ds), w11 = dl * ds;
    return (float)((double)v00 * w00 + (double)v01 * w01 + (double)v10 * w10 + (double)v11 * w11);
}

double keys_weight(double t) {  // :820-829
    t = std::fabs(t);
    if (t < 1.0) return 1.5 * t * t * t - 2.5 * t * t + 1.0;
    if (t < 2.0) return -0.5 * t * t * t + 2.5 * t * t - 4.0 * t + 2.0;
    return 0.0;
}

float bicubic_sample_slice(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    int64_t max_l = nl - 1, max_s = ns - 1;
    int64_t lf = (int64_t)std::floor(line), sf = (int64_t)std::floor(sample);
    double dl = line - (double)lf, ds = sample - (double)sf;
    double acc = 0.0;
    for (int64_t li = -1; li <= 2; ++li) {
        int64_t row = std::clamp<int64_t>(lf + li, 0, max_l);
        double wl = keys_weight((double)li - dl);
        for (int64_t si = -1; si <= 2; ++si) {
            int64_t col = std::clamp<int64_t>(sf + si, 0, max_s);
            float v = d[row * ns + col];
            if (std::isnan(v)) return std::numeric_limits<float>::quiet_NaN();
            acc += (double)v * wl * keys_weight((double)si - ds);
        }
    }
    return (float)acc;
}

double lanczos_weight(double x) {  // :938-950
    const double A = 3.0;
    double ax = std::fabs(x);
    if (ax < 1e-10) return 1.0;
    if (ax < A) {
        double px = kPi * x;
        double pxa = kPi * x / A;
        return std::sin(px) * std::sin(pxa) / (px * pxa);
    }
    return 0.0;
}

float lanczos_sample_slice(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    int64_t max_l = nl - 1, max_s = ns - 1;
    int64_t lf = (int64_t)std::floor(line), sf = (int64_t)std::floor(sample);
    double acc = 0.0;
    for (int64_t li = -2; li <= 3; ++li) {
        int64_t row = std::clamp<int64_t>(lf + li, 0, max_l);
        double wl = lanczos_weight(line - (double)(lf + li));
        for (int64_t si = -2; si <= 3; ++si) {
            int64_t col = std::clamp<int64_t>(sf + si, 0, max_s);
            float v = d[row * ns + col];
            if (std::isnan(v)) return std::numeric_limits<float>::quiet_NaN();
            acc += (double)v * wl * lanczos_weight(sample - (double)(sf + si));
        }
    }
    return (float)acc;
}

float sample_at(int kind, const float* d, int64_t nl, int64_t ns, double line, double sample) {
    switch (kind) {  // :888-901
        case SARDINE_RESAMPLE_BICUBIC: return bicubic_sample_slice(d, nl, ns, line, sample);
        case SARDINE_RESAMPLE_LANCZOS3: return lanczos_sample_slice(d, nl, ns, line, sample);
        default: return bilinear_sample_slice(d, nl, ns, line, sample);
    }
}

// ── DEM (dem.rs:351-409, 694-723) ─────────────────────────────────────────
struct HostDem {
    const float* data;
    int64_t width, height;
    double olat, olon, pix;
    const sardine_dem_tiling* tiling = nullptr;  // a GLO-30 mosaic held as one raster
    int64_t stride = 0;                          // row pitch of data in floats (0: width)
};

bool dem_elevation_one(const HostDem& d, double lat, double lon, float* h);

// Returns false for "no value" (outside coverage or a NaN tap): both become
// DEM_MISSING / a missing neighbour in terrain_correction.
bool dem_elevation(const HostDem& d, double lat, double lon, float* h) {
    if (!d.tiling) return dem_elevation_one(d, lat, lon, h);
    // DemMosaic::elevation_at_linear (dem.rs:730-747): the tiles in load
    // order, each a Glo30Tile (its block of the raster, its own origin);
    // the first that answers Some wins.  The mosaic normalises the
    // longitude before the tile sees it (dem.rs:735), as the tile would.
    const sardine_dem_tiling& t = *d.tiling;
    const int M = SARDINE_DEM_MAX_TILES_AXIS;
    std::vector<std::pair<int, int>> order;  // (rank, cell)
    for (int i = 0; i < M * M; ++i)
        if (t.rank[i] >= 0) order.push_back({t.rank[i], i});
    std::sort(order.begin(), order.end());
    for (const auto& rc : order) {
        const int r = rc.second / M, c = rc.second % M;
        // the tile's block: rows r*th.., cols c*tw.., row pitch the raster width
        HostDem tile{d.data + (int64_t)r * t.tile_height * d.width + (int64_t)c * t.tile_width, t.tile_width,
                     t.tile_height, t.origin_lat_deg[r], t.origin_lon_deg[c], d.pix, nullptr, d.width};
        if (dem_elevation_one(tile, lat, lon, h)) return true;
    }
    return false;
}

// Glo30Tile::elevation_at (dem.rs:351-372) on one raster.
bool dem_elevation_one(const HostDem& d, double lat, double lon, float* h) {
    double lon_n = lon >= 180.0 ? lon - 360.0 : lon;  // dem.rs:697
    double row_f = (d.olat - lat) / d.pix;
    double col_f = (lon_n - d.olon) / d.pix;
    double max_row = (double)(d.height - 1), max_col = (double)(d.width - 1);
    const double margin = 0.5;
    if (row_f < -margin || row_f > max_row + margin || col_f < -margin || col_f > max_col + margin)
        return false;
    row_f = std::clamp(row_f, 0.0, max_row);
    col_f = std::clamp(col_f, 0.0, max_col);
    // bilinear_on_grid(data, width, row_f, col_f, height-1)
    int64_t dim = d.width, max_rc = d.height - 1, pitch = d.stride ? d.stride : d.width;
    int64_t r0 = std::min<int64_t>((int64_t)row_f, max_rc > 0 ? max_rc - 1 : 0);
    int64_t c0 = std::min<int64_t>((int64_t)col_f, dim >= 2 ? dim - 2 : 0);
    int64_t r1 = r0 + 1, c1 = c0 + 1;
    double dr = row_f - (double)r0, dc = col_f - (double)c0;
    float v00 = d.data[r0 * pitch + c0], v01 = d.data[r0 * pitch + c1];
    float v10 = d.data[r1 * pitch + c0], v11 = d.data[r1 * pitch + c1];
    if (std::isnan(v00) || std::isnan(v01) || std::isnan(v10) || std::isnan(v11)) return false;
    double v = (double)v00 * (1.0 - dr) * (1.0 - dc) + (double)v01 * (1.0 - dr) * dc +
               (double)v10 * dr * (1.0 - dc) + (double)v11 * dr * dc;
    *h = (float)v;
    return true;
}

// ── Geoid (geoid.rs:220-251, 395-401) ────────────────────────────────────
double geoid_undulation(const sardine_geoid& g, double lat, double lon) {
    if (g.kind != SARDINE_GEOID_GRID_2P5 || !g.grid_73x144) return 0.0;
    const int NLAT = 73, NLON = 144;
    const double STEP = 2.5;
    double l = std::fmod(lon, 360.0);
    if (l < 0.0) l += 360.0;  // rem_euclid
    double row_f = (90.0 - std::clamp(lat, -90.0, 90.0)) / STEP;
    double col_f = l / STEP;
    int r0 = std::min((int)row_f, NLAT - 2), r1 = r0 + 1;
    int c0 = ((int)col_f) % NLON, c1 = (c0 + 1) % NLON;
    double dr = row_f - r0;
    double dc = std::clamp(col_f - c0, 0.0, 1.0);
    const float* u = g.grid_73x144;
    double v00 = u[r0 * NLON + c0], v01 = u[r0 * NLON + c1], v10 = u[r1 * NLON + c0], v11 = u[r1 * NLON + c1];
    return v00 * (1.0 - dr) * (1.0 - dc) + v01 * (1.0 - dr) * dc + v10 * dr * (1.0 - dc) + v11 * dr * dc;
}

// ── Terrain geometry (terrain_correction.rs:452-651) ──────────────────────
enum { TG_VALID = 0, TG_FORESHORTENING = 1, TG_LAYOVER = 2, TG_SHADOW = 3 };

double flat_cell_area(double lat, double lon, double s) {  // :576-583
    V3 fe = geodetic_to_ecef(lat, lon + s, 0.0), fw = geodetic_to_ecef(lat, lon - s, 0.0);
    V3 fn = geodetic_to_ecef(lat + s, lon, 0.0), fs = geodetic_to_ecef(lat - s, lon, 0.0);
    return norm3(cross3(sub3(fe, fw), sub3(fn, fs)));
}

int terrain_geometry(double lat, double lon, double he, double hw, double hn, double hs,
                     V3 sat, V3 target, double s, double flat_area, double* w_o, double* cos_lia_o) {
    V3 pe = geodetic_to_ecef(lat, lon + s, he), pw = geodetic_to_ecef(lat, lon - s, hw);
    V3 pn = geodetic_to_ecef(lat + s, lon, hn), ps = geodetic_to_ecef(lat - s, lon, hs);
    V3 tn = cross3(sub3(pe, pw), sub3(pn, ps));
    V3 look = normalize3(sub3(sat, target));
    double local_proj = dot3(tn, look);
    if (local_proj <= 0.0) return TG_SHADOW;
    double terrain_area = norm3(tn);
    if (terrain_area == 0.0) return TG_SHADOW;
    double cos_lia = std::clamp(local_proj / terrain_area, 0.0, 1.0);
    double w = local_proj / flat_area;
    *cos_lia_o = cos_lia;
    *w_o = w;
    if (w < 0.05) return TG_FORESHORTENING;
    if (w > 2.0) return TG_LAYOVER;
    return TG_VALID;
}

// ── UTM (output_crs.rs, via proj4rs 0.1.10 — not vendored) ───────────────
// Transverse Mercator by the Krüger series to order n^6 (C. F. F. Karney,
// "Transverse Mercator with an accuracy of a few nanometers", J. Geodesy 85
// (2011) 475-485, eqs. 7-11, 35-36).  Sub-millimetre agreement with any exact
// TM implementation; the reference pins proj4rs only to ±1 m.
struct TmCoef {
    double n, A, alpha[7], beta[7], e;
};
TmCoef tm_coef() {
    TmCoef c;
    double f = WGS84_F;
    double n = f / (2.0 - f), n2 = n * n, n3 = n2 * n, n4 = n3 * n, n5 = n4 * n, n6 = n5 * n;
    c.n = n;
    c.e = std::sqrt(WGS84_E2);
    c.A = WGS84_A / (1.0 + n) * (1.0 + n2 / 4.0 + n4 / 64.0 + n6 / 256.0);
    c.alpha[1] = n / 2 - 2 * n2 / 3 + 5 * n3 / 16 + 41 * n4 / 180 - 127 * n5 / 288 + 7891 * n6 / 37800;
    c.alpha[2] = 13 * n2 / 48 - 3 * n3 / 5 + 557 * n4 / 1440 + 281 * n5 / 630 - 1983433 * n6 / 1935360;
    c.alpha[3] = 61 * n3 / 240 - 103 * n4 / 140 + 15061 * n5 / 26880 + 167603 * n6 / 181440;
    c.alpha[4] = 49561 * n4 / 161280 - 179 * n5 / 168 + 6601661 * n6 / 7257600;
    c.alpha[5] = 34729 * n5 / 80640 - 3418889 * n6 / 1995840;
    c.alpha[6] = 212378941 * n6 / 319334400;
    c.beta[1] = n / 2 - 2 * n2 / 3 + 37 * n3 / 96 - n4 / 360 - 81 * n5 / 512 + 96199 * n6 / 604800;
    c.beta[2] = n2 / 48 + n3 / 15 - 437 * n4 / 1440 + 46 * n5 / 105 - 1118711 * n6 / 3870720;
    c.beta[3] = 17 * n3 / 480 - 37 * n4 / 840 - 209 * n5 / 4480 + 5569 * n6 / 90720;
    c.beta[4] = 4397 * n4 / 161280 - 11 * n5 / 504 - 830251 * n6 / 7257600;
    c.beta[5] = 4583 * n5 / 161280 - 108847 * n6 / 3991680;
    c.beta[6] = 20648693 * n6 / 638668800;
    return c;
}
const TmCoef& TM() {
    static const TmCoef c = tm_coef();
    return c;
}
constexpr double UTM_K0 = 0.9996, UTM_E0 = 500000.0, UTM_N0_SOUTH = 10000000.0;

// Σ_j c_j sin(2jξ) cosh(2jη) and Σ_j c_j cos(2jξ) sinh(2jη) (Karney 2011,
// eqs. 11 and 35) summed by Clenshaw's recurrence on z = ξ + iη:
// sum = Im/Re parts of sin(2z)·y0 with y_k = 2cos(2z)·y_{k+1} - y_{k+2} + c_k.
void tm_clenshaw(const double* c /* c[1..6] */, double xi, double eta, double* sx, double* sy) {
    double s2, c2;
    pm::sincos(2 * xi, &s2, &c2);
    const double sh2 = pm::sinh(2 * eta), ch2 = pm::cosh(2 * eta);
    double ar = 2 * c2 * ch2, ai = -2 * s2 * sh2;
    double y0r = 0, y0i = 0, y1r = 0, y1i = 0;
    for (int k = 6; k >= 1; --k) {
        double tr = ar * y0r - ai * y0i - y1r + c[k];
        double ti = ar * y0i + ai * y0r - y1i;
        y1r = y0r; y1i = y0i;
        y0r = tr; y0i = ti;
    }
    double zr = s2 * ch2, zi = c2 * sh2;
    *sx = zr * y0r - zi * y0i;
    *sy = zr * y0i + zi * y0r;
}

void utm_forward(int zone, bool south, double lon_deg, double lat_deg, double* x, double* y) {
    const TmCoef& c = TM();
    double lon0 = (double)(zone * 6 - 183);
    double phi = lat_deg * kPi / 180.0, lam = (lon_deg - lon0) * kPi / 180.0;
    double sphi = pm::sin(phi);
    double sl, cl;
    pm::sincos(lam, &sl, &cl);
    // conformal latitude: τ' = sinh(atanh(sin φ) - e atanh(e sin φ)) (eq. 7-9)
    double t = pm::sinh(pm::atanh(sphi) - c.e * pm::atanh(c.e * sphi));
    double xip = pm::atan2(t, cl);
    double etap = pm::atanh(sl / std::sqrt(1.0 + t * t));
    double sx, sy;
    tm_clenshaw(c.alpha, xip, etap, &sx, &sy);
    *x = UTM_E0 + UTM_K0 * c.A * (etap + sy);
    *y = (south ? UTM_N0_SOUTH : 0.0) + UTM_K0 * c.A * (xip + sx);
}

void utm_inverse(int zone, bool south, double x, double y, double* lon_deg, double* lat_deg) {
    const TmCoef& c = TM();
    double xi = (y - (south ? UTM_N0_SOUTH : 0.0)) / (UTM_K0 * c.A);
    double eta = (x - UTM_E0) / (UTM_K0 * c.A);
    double sx, sy;
    tm_clenshaw(c.beta, xi, eta, &sx, &sy);
    double xip = xi - sx, etap = eta - sy;
    double sxp, cx;
    pm::sincos(xip, &sxp, &cx);
    double sh = pm::sinh(etap);
    double taup = sxp / std::sqrt(sh * sh + cx * cx);
    double lam = pm::atan2(sh, cx);
    // Solve tau from tau' (Karney 2011 eqs. 19-21), Newton iteration.
    double e2 = WGS84_E2, e = c.e;
    double tau = taup;
    for (int i = 0; i < 8; ++i) {
        double s1 = std::sqrt(1.0 + tau * tau);
        double sig = pm::sinh(e * pm::atanh(e * tau / s1));
        double taupi = tau * std::sqrt(1.0 + sig * sig) - sig * s1;
        double dtau = (taup - taupi) / std::sqrt(1.0 + taupi * taupi) * (1.0 + (1.0 - e2) * tau * tau) /
                      ((1.0 - e2) * s1);
        tau += dtau;
        if (std::fabs(dtau) < 1e-14) break;
    }
    *lat_deg = pm::atan(tau) * 180.0 / kPi;
    *lon_deg = (double)(zone * 6 - 183) + lam * 180.0 / kPi;
}

// Projector::project_row, output_crs.rs:364-417
void project_row(int zone, bool south, double y, double x_start, double x_step, int64_t n_cols,
                 int64_t n_anchors, double* lon_out, double* lat_out) {
    if (n_cols == 0) return;
    n_anchors = std::min(n_anchors, n_cols);
    double x_end = x_start + ((double)n_cols - 1.0) * x_step;
    int64_t n_pts = n_anchors + 1;
    double anchor_dx = (x_end - x_start) / (double)n_anchors;
    std::vector<double> ax(n_pts), alon(n_pts), alat(n_pts);
    for (int64_t i = 0; i < n_pts; ++i) {
        ax[i] = x_start + (double)i * anchor_dx;
        utm_inverse(zone, south, ax[i], y, &alon[i], &alat[i]);
    }
    int64_t seg = 0;
    for (int64_t col = 0; col < n_cols; ++col) {
        double x = x_start + (double)col * x_step;
        while (seg + 1 < n_anchors && x > ax[seg + 1]) ++seg;
        double x0 = ax[seg], x1 = ax[seg + 1];
        double t = (x1 == x0) ? 0.0 : (x - x0) / (x1 - x0);
        lon_out[col] = alon[seg] + t * (alon[seg + 1] - alon[seg]);
        lat_out[col] = alat[seg] + t * (alat[seg + 1] - alat[seg]);
    }
}

// ── Seed LUT (terrain_correction.rs:1080-1137) ───────────────────────────
struct SeedLut {
    std::vector<double> lat;
    std::vector<int64_t> t;
};
SeedLut build_seed_lut(const sardine_grid_point* pts, int n) {
    std::vector<std::pair<double, int64_t>> v;
    for (int i = 0; i < n; ++i) v.push_back({pts[i].latitude_deg, pts[i].azimuth_time_us});
    std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    SeedLut l;
    for (auto& p : v) {  // dedup_by(|a,b| |a-b| < 1e-6): keep the first of a run
        if (!l.lat.empty() && std::fabs(p.first - l.lat.back()) < 1e-6) continue;
        l.lat.push_back(p.first);
        l.t.push_back(p.second);
    }
    return l;
}
int64_t guess_azimuth_time(const SeedLut& l, double lat) {
    size_t n = l.lat.size();
    if (n == 1) return l.t[0];
    // binary_search_by: exact match returns that entry
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (l.lat[mid] < lat) lo = mid + 1;
        else if (l.lat[mid] > lat) hi = mid;
        else return l.t[mid];
    }
    size_t idx = lo;
    if (idx == 0) return l.t[0];
    if (idx >= n) return l.t[n - 1];
    double frac = (lat - l.lat[idx - 1]) / (l.lat[idx] - l.lat[idx - 1]);
    double dt_us = (double)(l.t[idx] - l.t[idx - 1]);
    return l.t[idx - 1] + (int64_t)(frac * dt_us);
}

// RadarGeometry::to_image_coords (terrain_correction.rs:745-762); times in
// µs on any common epoch.  Returns false outside [-0.5, n-0.5].
bool to_image_coords(int64_t az, int64_t az_start, double ati, double slant, double near_range, double rps,
                     int64_t n_lines, int64_t n_samples, double* line_o, double* sample_o) {
    double line = (double)(az - az_start) / 1e6 / ati;
    double sample = (slant - near_range) / rps;
    *line_o = line;
    *sample_o = sample;
    return !(line < -0.5 || line > (double)n_lines - 0.5 || sample < -0.5 || sample > (double)n_samples - 0.5);
}

void set_err(char* buf, const char* msg) {
    if (buf) std::snprintf(buf, 256, "%s", msg);
}

}  // namespace

// ═════════════════════════════════════════════════════════════════════════
// extern "C" entry points (ctypes from tests/ and bench.py)
// ═════════════════════════════════════════════════════════════════════════
extern "C" {

void oracle_geodetic_to_ecef(double lat, double lon, double h, double out[3]) {
    V3 e = geodetic_to_ecef(lat, lon, h);
    out[0] = e.x; out[1] = e.y; out[2] = e.z;
}

void oracle_ecef_to_geodetic(const double e[3], double out[3]) {
    ecef_to_geodetic({e[0], e[1], e[2]}, &out[0], &out[1], &out[2]);
}

int oracle_interpolate_orbit_and_accel(const sardine_orbit* orbit, int64_t t_rel_us, double pos[3],
                                       double vel[3], double acc[3]) {
    Orbit o = make_orbit(orbit);
    V3 p, v, a;
    int e = interp_orbit_accel(o, t_rel_us, &p, &v, &a);
    if (e) return e;
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
    vel[0] = v.x; vel[1] = v.y; vel[2] = v.z;
    acc[0] = a.x; acc[1] = a.y; acc[2] = a.z;
    return 0;
}

// Returns outcome (0 converged, 1 not converged, 2 degenerate) or -status.
int oracle_solve_zero_doppler(const double target[3], const sardine_orbit* orbit, int64_t seed_rel_us,
                              int max_iter, double tol, int64_t* t_rel_out, int* iterations,
                              double sat_pos[3]) {
    Orbit o = make_orbit(orbit);
    V3 p = {0, 0, 0};
    int it = 0;
    int64_t t = 0;
    int r = solve_zero_doppler({target[0], target[1], target[2]}, o, seed_rel_us, max_iter, tol, &t, &it, &p);
    *t_rel_out = t;
    *iterations = it;
    sat_pos[0] = p.x; sat_pos[1] = p.y; sat_pos[2] = p.z;
    return r;
}

int oracle_to_image_coords(int64_t az, int64_t az_start, double ati, double slant, double near_range, double rps,
                           int64_t n_lines, int64_t n_samples, double* line, double* sample) {
    return to_image_coords(az, az_start, ati, slant, near_range, rps, n_lines, n_samples, line, sample) ? 1 : 0;
}

float oracle_sample(int kind, const float* data, int64_t lines, int64_t samples, double line, double sample) {
    return sample_at(kind, data, lines, samples, line, sample);
}
double oracle_keys_weight(double t) { return keys_weight(t); }
double oracle_lanczos_weight(double x) { return lanczos_weight(x); }

double oracle_flat_cell_area(double lat, double lon, double s) { return flat_cell_area(lat, lon, s); }

// compute_terrain_geometry (terrain_correction.rs:545-562): flat area from
// the pixel's own latitude.  Returns the outcome code; w/cos_lia as available.
int oracle_terrain_geometry(double lat, double lon, double he, double hw, double hn, double hs,
                            const double sat[3], const double target[3], double spacing, double* w,
                            double* cos_lia) {
    double fa = flat_cell_area(lat, lon, spacing);
    *w = std::numeric_limits<double>::quiet_NaN();
    *cos_lia = std::numeric_limits<double>::quiet_NaN();
    return terrain_geometry(lat, lon, he, hw, hn, hs, {sat[0], sat[1], sat[2]},
                            {target[0], target[1], target[2]}, spacing, fa, w, cos_lia);
}

int oracle_dem_elevation(const float* data, int64_t width, int64_t height, double olat, double olon,
                         double pix, double lat, double lon, float* h) {
    HostDem d{data, width, height, olat, olon, pix};
    return dem_elevation(d, lat, lon, h) ? 1 : 0;
}

// The same over a GLO-30 mosaic held as one raster (sardine_dem_tiling).
int oracle_dem_mosaic_elevation(const float* data, int64_t width, int64_t height, double pix,
                                const sardine_dem_tiling* tiling, double lat, double lon, float* h) {
    HostDem d{data, width, height, tiling->origin_lat_deg[0], tiling->origin_lon_deg[0], pix, tiling};
    return dem_elevation(d, lat, lon, h) ? 1 : 0;
}

double oracle_geoid_undulation(const float* grid, double lat, double lon) {
    sardine_geoid g{grid ? SARDINE_GEOID_GRID_2P5 : SARDINE_GEOID_ZERO, grid};
    return geoid_undulation(g, lat, lon);
}

void oracle_utm_forward(int zone, int south, double lon, double lat, double* x, double* y) {
    utm_forward(zone, south != 0, lon, lat, x, y);
}
void oracle_utm_inverse(int zone, int south, double x, double y, double* lon, double* lat) {
    utm_inverse(zone, south != 0, x, y, lon, lat);
}
void oracle_project_row(int zone, int south, double y, double x_start, double x_step, int64_t n_cols,
                        int64_t n_anchors, double* lon, double* lat) {
    project_row(zone, south != 0, y, x_start, x_step, n_cols, n_anchors, lon, lat);
}

int64_t oracle_guess_azimuth_time(const sardine_grid_point* pts, int n, double lat) {
    SeedLut l = build_seed_lut(pts, n);
    if (l.lat.empty()) return 0;
    return guess_azimuth_time(l, lat);
}

// compute_output_grid_extent + shape (terrain_correction.rs:1174-1216, 1304-1317)
int oracle_tc_output_grid(const sardine_scene* scene, const sardine_tc_config* cfg, int64_t* rows,
                          int64_t* cols, double gt[6]) {
    double sp = cfg->pixel_spacing;
    if (!(sp > 0.0)) return SARDINE_ERR_TC_CONFIG;
    const sardine_bbox& bb = scene->bounding_box;
    double yt, yb, xl, xr;
    if (cfg->crs == SARDINE_CRS_WGS84) {
        yt = std::ceil(bb.max_lat_deg / sp) * sp;
        yb = std::floor(bb.min_lat_deg / sp) * sp;
        xl = std::floor(bb.min_lon_deg / sp) * sp;
        xr = std::ceil(bb.max_lon_deg / sp) * sp;
    } else {
        bool south = cfg->crs == SARDINE_CRS_UTM_SOUTH;
        double corners[4][2] = {{bb.min_lon_deg, bb.min_lat_deg}, {bb.min_lon_deg, bb.max_lat_deg},
                                {bb.max_lon_deg, bb.min_lat_deg}, {bb.max_lon_deg, bb.max_lat_deg}};
        double x_min = INFINITY, x_max = -INFINITY, y_min = INFINITY, y_max = -INFINITY;
        for (auto& c : corners) {
            double x, y;
            utm_forward(cfg->utm_zone, south, c[0], c[1], &x, &y);
            x_min = std::min(x_min, x); x_max = std::max(x_max, x);
            y_min = std::min(y_min, y); y_max = std::max(y_max, y);
        }
        yt = std::ceil(y_max / sp) * sp;
        yb = std::floor(y_min / sp) * sp;
        xl = std::floor(x_min / sp) * sp;
        xr = std::ceil(x_max / sp) * sp;
    }
    double nr = std::round((yt - yb) / sp), nc = std::round((xr - xl) / sp);
    *rows = nr > 0 ? (int64_t)nr : 0;
    *cols = nc > 0 ? (int64_t)nc : 0;
    gt[0] = xl; gt[1] = sp; gt[2] = 0.0; gt[3] = yt; gt[4] = 0.0; gt[5] = -sp;
    if (*rows == 0 || *cols == 0) return SARDINE_ERR_TC_CONFIG;
    return 0;
}

// terrain_correction, terrain_correction.rs:1245-1836.  All raster pointers
// here are HOST memory.  The AOI window restarts the per-row seed chain at
// the window's first column (the reference has no AOI: aoi = full grid is the
// reference call).  out->data etc. hold aoi_rows*aoi_cols elements.
int oracle_terrain_correction(const sardine_radar_image* merged, const sardine_scene* scene,
                              const sardine_dem_raster* dem_r, const sardine_tc_config* cfg,
                              sardine_geocoded_image* out, int n_threads, char* errbuf) {
    if (scene->n_grid_points <= 0) { set_err(errbuf, "no geolocation grid points"); return SARDINE_ERR_TC_NO_GRID_POINTS; }
    if (!(cfg->pixel_spacing > 0.0)) { set_err(errbuf, "pixel spacing must be positive"); return SARDINE_ERR_TC_CONFIG; }
    if (scene->n_subswaths <= 0) { set_err(errbuf, "no subswaths"); return SARDINE_ERR_TC_NO_SUBSWATHS; }
    double ati = scene->azimuth_time_interval_s;
    if (!std::isfinite(ati) || ati <= 0.0) { set_err(errbuf, "invalid azimuth_time_interval_s"); return SARDINE_ERR_TC_CONFIG; }
    Orbit orbit = make_orbit(&scene->orbit);
    const int64_t epoch = scene->orbit.count > 0 ? scene->orbit.state_vectors[0].time_us : 0;
    SeedLut lut = build_seed_lut(scene->grid_points, scene->n_grid_points);
    if (lut.lat.empty()) return SARDINE_ERR_TC_NO_GRID_POINTS;

    int64_t n_rows, n_cols;
    double gt[6];
    int st = oracle_tc_output_grid(scene, cfg, &n_rows, &n_cols, gt);
    if (st) { set_err(errbuf, "output grid is empty"); return st; }
    const double spacing = cfg->pixel_spacing;
    const double lat_north = gt[3], lon_west = gt[0];
    const bool utm = cfg->crs != SARDINE_CRS_WGS84;
    const bool south = cfg->crs == SARDINE_CRS_UTM_SOUTH;
    const int64_t anchors = cfg->projector_anchors > 0 ? cfg->projector_anchors : 256;

    int64_t r0 = cfg->aoi_row0, c0 = cfg->aoi_col0, nr = cfg->aoi_rows, nc = cfg->aoi_cols;
    if (nr == 0) { r0 = 0; c0 = 0; nr = n_rows; nc = n_cols; }
    if (r0 < 0 || c0 < 0 || nr <= 0 || nc <= 0 || r0 + nr > n_rows || c0 + nc > n_cols) {
        set_err(errbuf, "AOI window outside the output grid");
        return SARDINE_ERR_TC_CONFIG;
    }

    // RadarGeometry (terrain_correction.rs:701-741)
    const double near_range_m = merged->near_slant_range_time_s * SPEED_OF_LIGHT / 2.0;
    const double rps = merged->range_pixel_spacing_m;
    const int64_t n_lines = merged->lines, n_samples = merged->samples;
    const int64_t az_start = merged->azimuth_start_time_us - epoch;
    const double step_deg = utm ? spacing / 111320.0 : spacing;  // :452-465
    const bool need_geom = cfg->flatten || cfg->compute_lia;
    const bool nesz_on = std::isfinite(cfg->noise_floor_margin_db) && merged->nesz != nullptr;
    const float nesz_factor = std::pow(10.0f, cfg->noise_floor_margin_db / 10.0f);
    HostDem dem{dem_r->data, dem_r->width, dem_r->height, dem_r->origin_lat_deg, dem_r->origin_lon_deg,
                dem_r->pixel_deg, dem_r->tiling};

    const float NaN = std::numeric_limits<float>::quiet_NaN();
    struct RowAcc {
        uint64_t valid = 0, dem_missing = 0, outside = 0, nonconv = 0, degen = 0, flat = 0, noise = 0;
        uint64_t hist[SARDINE_NEWTON_HIST_SIZE] = {0};
        int err = 0;
    };
    const int64_t CHUNK = 64;
    const int64_t n_chunks = (nr + CHUNK - 1) / CHUNK;
    std::vector<RowAcc> acc((size_t)n_chunks);
    std::atomic<int64_t> next{0};

    auto worker = [&]() {
        std::vector<double> row_lon(utm ? (size_t)n_cols : 0), row_lat(utm ? (size_t)n_cols : 0);
        for (;;) {
            int64_t ci = next.fetch_add(1);
            if (ci >= n_chunks) break;
            RowAcc& A = acc[(size_t)ci];
            for (int64_t rr = ci * CHUNK; rr < std::min(nr, (ci + 1) * CHUNK) && !A.err; ++rr) {
                const int64_t row = r0 + rr;
                const double row_y = lat_north - ((double)row + 0.5) * spacing;
                double row_lat_seed = row_y;
                if (utm) {
                    double mid_x = lon_west + ((double)n_cols / 2.0) * spacing, lon_m;
                    utm_inverse(cfg->utm_zone, south, mid_x, row_y, &lon_m, &row_lat_seed);
                    project_row(cfg->utm_zone, south, row_y, lon_west + 0.5 * spacing, spacing, n_cols, anchors,
                                row_lon.data(), row_lat.data());
                }
                int64_t prev = guess_azimuth_time(lut, row_lat_seed) - epoch;
                double row_flat_area = 0.0;
                if (need_geom) {
                    double mid_lon = lon_west + ((double)n_cols / 2.0) * spacing;
                    double ref_lon = utm ? row_lon[(size_t)(n_cols / 2)] : mid_lon;
                    row_flat_area = flat_cell_area(row_lat_seed, ref_lon, step_deg);
                }
                float* od = out->data + rr * nc;
                float* ol = cfg->compute_lia ? out->cos_lia + rr * nc : nullptr;
                uint8_t* om = cfg->compute_lia ? out->mask + rr * nc : nullptr;
                double* orc = out->radar_coords ? out->radar_coords + rr * nc * 2 : nullptr;
                for (int64_t cc = 0; cc < nc; ++cc) {
                    od[cc] = NaN;
                    if (orc) { orc[2 * cc] = NAN; orc[2 * cc + 1] = NAN; }
                    if (ol) { ol[cc] = NaN; om[cc] = SARDINE_MASK_OUTSIDE_FOOTPRINT; }
                }
                for (int64_t cc = 0; cc < nc; ++cc) {
                    const int64_t col = c0 + cc;
                    double lon, lat;
                    if (utm) { lon = row_lon[(size_t)col]; lat = row_lat[(size_t)col]; }
                    else { lon = lon_west + ((double)col + 0.5) * spacing; lat = row_y; }
                    float hf;
                    if (!dem_elevation(dem, lat, lon, &hf) || !std::isfinite(hf)) {
                        A.dem_missing++;
                        if (om) om[cc] = SARDINE_MASK_DEM_MISSING;
                        continue;
                    }
                    double h = (double)hf;
                    double h_ell = h + geoid_undulation(cfg->geoid, lat, lon);
                    V3 target = geodetic_to_ecef(lat, lon, h_ell);
                    int64_t az;
                    int iters;
                    V3 sat;
                    int oc = solve_zero_doppler(target, orbit, prev, cfg->max_newton_iterations,
                                                cfg->newton_tolerance_s, &az, &iters, &sat);
                    if (oc < 0) { A.err = -oc; break; }
                    if (oc == ZD_NOT_CONVERGED) { A.nonconv++; if (om) om[cc] = SARDINE_MASK_NON_CONVERGED; continue; }
                    if (oc == ZD_DEGENERATE) { A.degen++; if (om) om[cc] = SARDINE_MASK_DEGENERATE_GEOMETRY; continue; }
                    A.hist[std::min(iters, SARDINE_NEWTON_HIST_SIZE - 1)]++;
                    prev = az;
                    double slant = norm3(sub3(target, sat));
                    double line, sample;
                    bool inside = to_image_coords(az, az_start, ati, slant, near_range_m, rps, n_lines, n_samples,
                                                  &line, &sample);
                    if (orc) { orc[2 * cc] = line; orc[2 * cc + 1] = sample; }
                    if (!inside) {
                        A.outside++;
                        if (om) om[cc] = SARDINE_MASK_OUTSIDE_FOOTPRINT;
                        continue;
                    }
                    float sigma0 = sample_at(cfg->resampling, merged->data, n_lines, n_samples, line, sample);
                    if (std::isnan(sigma0)) {
                        A.outside++;
                        if (om) om[cc] = SARDINE_MASK_OUTSIDE_FOOTPRINT;
                        continue;
                    }
                    if (need_geom) {
                        float he, hw, hn, hs;
                        bool ok = dem_elevation(dem, lat, lon + step_deg, &he) &&
                                  dem_elevation(dem, lat, lon - step_deg, &hw) &&
                                  dem_elevation(dem, lat + step_deg, lon, &hn) &&
                                  dem_elevation(dem, lat - step_deg, lon, &hs);
                        // evaluation order of the reference tuple is E,W,N,S;
                        // all four are evaluated there, but only "all finite"
                        // matters for the outcome.
                        if (ok) ok = std::isfinite(he) && std::isfinite(hw) && std::isfinite(hn) && std::isfinite(hs);
                        if (!ok) {
                            if (om) om[cc] = SARDINE_MASK_DEM_NEIGHBOUR_MISSING;
                            if (cfg->flatten) { A.flat++; continue; }
                        } else {
                            double w = 0, cl = 0;
                            int g = terrain_geometry(lat, lon, he, hw, hn, hs, sat, target, step_deg,
                                                     row_flat_area, &w, &cl);
                            if (g == TG_VALID) {
                                if (ol) ol[cc] = (float)cl;
                                if (cfg->flatten) sigma0 /= (float)w;
                            } else if (g == TG_FORESHORTENING || g == TG_LAYOVER) {
                                if (ol) { ol[cc] = (float)cl; om[cc] = g == TG_LAYOVER ? SARDINE_MASK_LAYOVER : SARDINE_MASK_FORESHORTENING; }
                                if (cfg->flatten) { A.flat++; continue; }
                            } else {
                                if (om) om[cc] = SARDINE_MASK_SHADOW;
                                if (cfg->flatten) { A.flat++; continue; }
                            }
                        }
                    }
                    if (nesz_on) {
                        float nesz = bilinear_sample_slice(merged->nesz, n_lines, n_samples, line, sample);
                        if (!std::isnan(nesz)) {
                            float thr = nesz * nesz_factor;
                            if (sigma0 <= thr) {
                                A.noise++;
                                if (om) om[cc] = SARDINE_MASK_NOISE_MASKED;
                                continue;
                            }
                        }
                    }
                    od[cc] = sigma0;
                    if (om) om[cc] = SARDINE_MASK_VALID;
                    A.valid++;
                }
            }
        }
    };
    int nt = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; ++i) pool.emplace_back(worker);
    for (auto& t : pool) t.join();

    std::memset(out->newton_iter_histogram, 0, sizeof(out->newton_iter_histogram));
    out->valid_pixel_count = out->dem_missing_count = out->outside_footprint_count = 0;
    out->non_converged_count = out->degenerate_geometry_count = out->flat_masked_count = 0;
    out->noise_masked_count = 0;
    for (auto& A : acc) {
        if (A.err) {
            set_err(errbuf, A.err == SARDINE_ERR_TC_ORBIT_OUTSIDE_WINDOW
                                ? "orbit interpolation failed: outside orbit window"
                                : "orbit interpolation failed: insufficient state vectors");
            return A.err;
        }
        out->valid_pixel_count += A.valid;
        out->dem_missing_count += A.dem_missing;
        out->outside_footprint_count += A.outside;
        out->non_converged_count += A.nonconv;
        out->degenerate_geometry_count += A.degen;
        out->flat_masked_count += A.flat;
        out->noise_masked_count += A.noise;
        for (int i = 0; i < SARDINE_NEWTON_HIST_SIZE; ++i) out->newton_iter_histogram[i] += A.hist[i];
    }
    out->rows = nr;
    out->cols = nc;
    out->full_rows = n_rows;
    out->full_cols = n_cols;
    out->total_pixel_count = (uint64_t)(nr * nc);
    out->geotransform[0] = lon_west + (double)c0 * spacing;
    out->geotransform[1] = spacing;
    out->geotransform[2] = 0.0;
    out->geotransform[3] = lat_north - (double)r0 * spacing;
    out->geotransform[4] = 0.0;
    out->geotransform[5] = -spacing;
    out->chain_redo_rounds = 0;
    return 0;
}

// ── Deburst (deburst.rs:172-344) ─────────────────────────────────────────
int oracle_deburst_layout(const sardine_subswath* sw, const sardine_burst* b, int n, int64_t tiff_height,
                          sardine_deburst_layout* out) {
    if (n <= 0) return SARDINE_ERR_DEBURST_NO_BURSTS;
    if (n != sw->burst_count) return SARDINE_ERR_DEBURST_COUNT_MISMATCH;
    if (n > SARDINE_MAX_BURSTS) return SARDINE_ERR_INVALID_ARGUMENT;
    int64_t lps = sw->lines_per_burst;
    if (tiff_height < (int64_t)n * lps) return SARDINE_ERR_DEBURST_TIFF_TOO_SMALL;
    int64_t fv = b[0].first_valid_sample, lv = b[0].last_valid_sample;
    for (int k = 1; k < n; ++k) { fv = std::max(fv, b[k].first_valid_sample); lv = std::min(lv, b[k].last_valid_sample); }
    if (fv >= lv) return SARDINE_ERR_DEBURST_EMPTY_VALID_WINDOW;
    for (int k = 0; k + 1 < n; ++k) {
        int64_t dt_us = b[k + 1].azimuth_time_us - b[k].azimuth_time_us;
        double dt_s = (double)dt_us * 1e-6;
        int64_t step = (int64_t)std::round(dt_s / sw->azimuth_time_interval_s);
        int64_t ov = lps - step;
        if (ov <= 0) return SARDINE_ERR_DEBURST_INVALID_OVERLAP;
        if (ov < 50 || ov > 400) return SARDINE_ERR_DEBURST_IMPLAUSIBLE_OVERLAP;
        out->overlap[k] = ov;
    }
    int64_t total = 0;
    for (int k = 0; k < n; ++k) {
        int64_t lf = k == 0 ? 0 : (out->overlap[k - 1] + 1) / 2;
        int64_t le = k == n - 1 ? lps : lps - out->overlap[k] / 2;
        out->local_first[k] = lf;
        out->local_end[k] = le;
        out->out_line_start[k] = total;
        total += le - lf;
    }
    if (n > 0) out->overlap[n - 1] = 0;
    out->n_bursts = n;
    out->lines = total;
    out->samples = lv - fv;
    out->valid_sample_offset = fv;
    return 0;
}

// deburst_subswath's row copy; raw is HOST tiff_height x raw_cols CInt16.
void oracle_deburst_copy(const int16_t* raw, int64_t raw_cols, const sardine_burst* b,
                         const sardine_deburst_layout* L, int16_t* out) {
    int64_t vc = L->samples, fv = L->valid_sample_offset, ol = 0;
    for (int k = 0; k < L->n_bursts; ++k) {
        int64_t nl = L->local_end[k] - L->local_first[k];
        int64_t tf = b[k].first_line + L->local_first[k];
        for (int64_t l = 0; l < nl; ++l)
            std::memcpy(out + (ol + l) * vc * 2, raw + ((tf + l) * raw_cols + fv) * 2, (size_t)vc * 4);
        ol += nl;
    }
}

}  // extern "C"

// ── Calibration (apply_calibration.rs:121-572) ───────────────────────────
namespace {
inline float lerpf(float a, float b, float t) {  // :128-131
    t = std::clamp(t, 0.0f, 1.0f);
    return a + (b - a) * t;
}
void bracket_azimuth(const std::vector<int64_t>& lines, int64_t tl, size_t* lo, size_t* hi) {
    auto it = std::lower_bound(lines.begin(), lines.end(), tl);
    size_t i = (size_t)(it - lines.begin());
    if (it != lines.end() && *it == tl) { *lo = *hi = i; return; }
    if (i == 0) { *lo = *hi = 0; return; }
    if (i >= lines.size()) { *lo = *hi = lines.size() - 1; return; }
    *lo = i - 1; *hi = i;
}
float lut_cursor_eval(const uint32_t* px, const float* val, size_t n, size_t cur, uint32_t col) {
    if (px[cur] > col) return val[0];
    if (cur + 1 < n) {
        float t = (float)(col - px[cur]) / (float)(px[cur + 1] - px[cur]);
        return lerpf(val[cur], val[cur + 1], t);
    }
    return val[cur];
}
void bilinear_row_into(const sardine_lut& L, const std::vector<int64_t>& lines, int64_t tl, int64_t col_off,
                       float* buf, int64_t n_cols) {
    size_t lo, hi;
    bracket_azimuth(lines, tl, &lo, &hi);
    const uint32_t* plo = L.pixels + L.offsets[lo];
    const float* vlo = L.values + L.offsets[lo];
    size_t nlo = (size_t)(L.offsets[lo + 1] - L.offsets[lo]);
    size_t clo = 0;
    if (lo == hi) {
        for (int64_t c = 0; c < n_cols; ++c) {
            uint32_t tc = (uint32_t)(c + col_off);
            while (clo + 1 < nlo && plo[clo + 1] <= tc) ++clo;
            buf[c] = lut_cursor_eval(plo, vlo, nlo, clo, tc);
        }
    } else {
        float t_line = (float)(tl - lines[lo]) / (float)(lines[hi] - lines[lo]);
        const uint32_t* phi = L.pixels + L.offsets[hi];
        const float* vhi = L.values + L.offsets[hi];
        size_t nhi = (size_t)(L.offsets[hi + 1] - L.offsets[hi]);
        size_t chi = 0;
        for (int64_t c = 0; c < n_cols; ++c) {
            uint32_t tc = (uint32_t)(c + col_off);
            while (clo + 1 < nlo && plo[clo + 1] <= tc) ++clo;
            while (chi + 1 < nhi && phi[chi + 1] <= tc) ++chi;
            buf[c] = lerpf(lut_cursor_eval(plo, vlo, nlo, clo, tc), lut_cursor_eval(phi, vhi, nhi, chi, tc), t_line);
        }
    }
}
}  // namespace

extern "C" {

int oracle_apply_calibration(const int16_t* deb, int64_t lines, int64_t samples, int64_t col_off,
                             const sardine_lut* cal, const sardine_lut* nr, const sardine_az_lut* az,
                             int64_t origin, float* out, float* nesz, uint32_t* cal_gap, uint32_t* nr_gap,
                             int n_threads) {
    if (cal->n_vectors <= 0) return SARDINE_ERR_CAL_NO_CAL_VECTORS;
    if (nr->n_vectors <= 0) return SARDINE_ERR_CAL_NO_NOISE_VECTORS;
    const uint32_t SLACK = 600;
    uint32_t req = (uint32_t)std::max<int64_t>(col_off + samples - 1, 0);
    uint32_t gaps[2] = {0, 0};
    const sardine_lut* Ls[2] = {cal, nr};
    for (int q = 0; q < 2; ++q) {
        for (int i = 0; i < Ls[q]->n_vectors; ++i) {
            int64_t n = Ls[q]->offsets[i + 1] - Ls[q]->offsets[i];
            if (n <= 0) return q == 0 ? SARDINE_ERR_CAL_NO_CAL_VECTORS : SARDINE_ERR_CAL_NO_NOISE_VECTORS;
            uint32_t mx = Ls[q]->pixels[Ls[q]->offsets[i + 1] - 1];
            if (mx < req) {
                uint32_t gap = req - mx;
                if (gap > SLACK) return SARDINE_ERR_CAL_LUT_RANGE;
                gaps[q] = std::max(gaps[q], gap);
            }
        }
    }
    *cal_gap = gaps[0];
    *nr_gap = gaps[1];
    std::vector<int64_t> cl(cal->lines, cal->lines + cal->n_vectors), nl(nr->lines, nr->lines + nr->n_vectors);
    std::vector<int64_t> al;
    for (int i = 0; i < az->n; ++i) al.push_back(az->lines[i]);
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        std::vector<float> k(samples), nrow(samples);
        for (;;) {
            int64_t line = next.fetch_add(1);
            if (line >= lines) break;
            int64_t tl = origin + line;
            bilinear_row_into(*cal, cl, tl, col_off, k.data(), samples);
            bilinear_row_into(*nr, nl, tl, col_off, nrow.data(), samples);
            float azf = 1.0f;
            if (!al.empty()) {
                size_t lo, hi;
                bracket_azimuth(al, tl, &lo, &hi);
                if (al[hi] == al[lo]) azf = az->values[lo];
                else {
                    float t = (float)(tl - al[lo]) / (float)(al[hi] - al[lo]);
                    azf = lerpf(az->values[lo], az->values[hi], t);
                }
            }
            const int16_t* in = deb + line * samples * 2;
            float* o = out + line * samples;
            float* ns = nesz + line * samples;
            for (int64_t c = 0; c < samples; ++c) {
                int32_t i = in[2 * c], q = in[2 * c + 1];
                float p = (float)(i * i + q * q);
                float k2 = k[c] * k[c];
                float valid = (k2 > 0.0f) ? 1.0f : 0.0f;
                float inv_k2 = k2 > 0.0f ? 1.0f / k2 : 0.0f;
                float nv = nrow[c] * azf;
                o[c] = std::max(p - nv, 0.0f) * inv_k2 * valid;
                ns[c] = nv * inv_k2 * valid;
            }
        }
    };
    int nt = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; ++i) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    return 0;
}

// merge_subswaths (merge_subswaths.rs:207-450), single-swath passthrough
// included (:213-233).  sigma0[i]/nesz[i] are HOST arrays of
// lines[i]*samples[i]; slant_t0/rps from each SubSwathMetadata.  Returns the
// layout in info and fills out_sigma0/out_nesz (info->lines*info->samples).
int oracle_merge(int n, const float* const* sigma0, const float* const* nesz, const int64_t* lines,
                 const int64_t* samples, const int64_t* valid_off, const double* slant_t0, const double* rps,
                 const int64_t* az_start_us, float* out_sigma0, float* out_nesz, sardine_merged_info* info,
                 int query_only) {
    if (n <= 0) return SARDINE_ERR_MERGE_EMPTY_INPUTS;
    if (n > SARDINE_MAX_SWATHS) return SARDINE_ERR_INVALID_ARGUMENT;
    double rps_m = rps[0];
    if (n > 1)
        for (int i = 1; i < n; ++i)
            if (std::fabs(rps[i] - rps_m) > 1e-3) return SARDINE_ERR_MERGE_PIXEL_SPACING;
    for (int i = 0; i < n; ++i)
        if (samples[i] == 0) return SARDINE_ERR_MERGE_EMPTY_SWATH;
    double dt = rps_m * 2.0 / SPEED_OF_LIGHT;
    std::vector<double> st(n);
    for (int i = 0; i < n; ++i) st[i] = slant_t0[i] + (double)valid_off[i] * dt;
    std::vector<int64_t> off(n), end(n);
    for (int i = 0; i < n; ++i) {
        off[i] = n == 1 ? 0 : (int64_t)std::round((st[i] - st[0]) / dt);
        end[i] = off[i] + samples[i];
    }
    int64_t out_cols = end[n - 1];
    std::vector<int64_t> seam(std::max(n - 1, 0));
    for (int i = 0; i + 1 < n; ++i) seam[i] = (end[i] + off[i + 1]) / 2;
    int64_t out_lines = lines[0];
    for (int i = 1; i < n; ++i) out_lines = std::min(out_lines, lines[i]);
    info->lines = out_lines;
    info->samples = out_cols;
    info->near_slant_range_time_s = st[0];
    info->range_pixel_spacing_s = dt;
    info->range_pixel_spacing_m = rps_m;
    int64_t az0 = az_start_us[0];
    for (int i = 1; i < n; ++i) az0 = std::min(az0, az_start_us[i]);
    info->azimuth_start_time_us = n == 1 ? az_start_us[0] : az0;
    info->n_swaths = n;
    for (int i = 0; i < n; ++i) {
        info->sw_out_offset[i] = off[i];
        info->seam[i] = i + 1 < n ? seam[i] : out_cols;
        info->swath_lines[i] = lines[i];
        info->swath_samples[i] = samples[i];
        info->swath_valid_offset[i] = valid_off[i];
    }
    if (query_only) return 0;
    const float NaN = std::numeric_limits<float>::quiet_NaN();
    for (int64_t l = 0; l < out_lines; ++l) {
        float* orow = out_sigma0 + l * out_cols;
        float* nrow = out_nesz + l * out_cols;
        int64_t fill_end = 0;
        for (int s = 0; s < n; ++s) {
            int64_t cs = s == 0 ? 0 : seam[s - 1];
            int64_t ce = s == n - 1 ? out_cols : seam[s];
            int64_t vs = std::max(cs, off[s]);
            int64_t ve = std::min(ce, off[s] + samples[s]);
            int64_t in_start = vs > off[s] ? vs - off[s] : 0;
            for (int64_t c = fill_end; c < vs; ++c) { orow[c] = NaN; nrow[c] = NaN; }
            if (vs < ve) {
                std::memcpy(orow + vs, sigma0[s] + l * samples[s] + in_start, (size_t)(ve - vs) * 4);
                std::memcpy(nrow + vs, nesz[s] + l * samples[s] + in_start, (size_t)(ve - vs) * 4);
            }
            for (int64_t c = std::max(ve, vs); c < ce; ++c) { orow[c] = NaN; nrow[c] = NaN; }
            fill_end = ce;
        }
        for (int64_t c = fill_end; c < out_cols; ++c) { orow[c] = NaN; nrow[c] = NaN; }
    }
    return 0;
}

// ── Speckle (speckle.rs:131-603) ─────────────────────────────────────────
}  // extern "C"

namespace {
bool window_mean_var(const float* d, int64_t cols, int64_t rows, int64_t r, int64_t c, int half, float* mean,
                     float* var) {
    double sum = 0.0, sq = 0.0;
    int64_t n = 0;
    for (int dy = -half; dy <= half; ++dy)
        for (int dx = -half; dx <= half; ++dx) {
            int64_t rr = r + dy, cc = c + dx;
            if (rr < 0 || cc < 0 || rr >= rows || cc >= cols) continue;
            float v = d[rr * cols + cc];
            if (!std::isfinite(v)) continue;
            double x = (double)v;
            sum += x;
            sq += x * x;
            ++n;
        }
    if (n == 0) return false;
    double m = sum / (double)n;
    double va = std::max(sq / (double)n - m * m, 0.0);
    *mean = (float)m;
    *var = (float)va;
    return true;
}
bool rl_in_mask(int dir, int dy, int dx) {  // speckle.rs:432-457
    if (dy == 0 && dx == 0) return false;
    switch (dir) {
        case 0: return dy >= 1 && dy <= 3;
        case 1: return dy >= -3 && dy <= -1;
        case 2: return dx >= 1 && dx <= 3;
        case 3: return dx >= -3 && dx <= -1;
        case 4: return std::abs(dy - dx) <= 1 && (dy + dx) >= 1;
        case 5: return std::abs(dy - dx) <= 1 && (dy + dx) <= -1;
        case 6: return std::abs(dy + dx) <= 1 && (dx - dy) >= 1;
        case 7: return std::abs(dy + dx) <= 1 && (dx - dy) <= -1;
    }
    return false;
}
bool rl_dir_mean_var(const float* d, int64_t cols, int64_t rows, int64_t r, int64_t c, int dir, float* mean,
                     float* var) {
    double sum = 0.0, sq = 0.0;
    int64_t n = 0;
    for (int dy = -3; dy <= 3; ++dy)
        for (int dx = -3; dx <= 3; ++dx) {
            if (!rl_in_mask(dir, dy, dx)) continue;
            int64_t rr = r + dy, cc = c + dx;
            if (rr < 0 || cc < 0 || rr >= rows || cc >= cols) continue;
            float v = d[rr * cols + cc];
            if (!std::isfinite(v)) continue;
            double x = (double)v;
            sum += x;
            sq += x * x;
            ++n;
        }
    if (n == 0) return false;
    double m = sum / (double)n;
    double va = std::max(sq / (double)n - m * m, 0.0);
    *mean = (float)m;
    *var = (float)va;
    return true;
}

float speckle_pixel(const float* d, int64_t cols, int64_t rows, int64_t r, int64_t c,
                    const sardine_speckle_filter& f) {
    const float NaN = std::numeric_limits<float>::quiet_NaN();
    float centre = d[r * cols + c];
    if (!std::isfinite(centre)) return NaN;
    int half = f.window / 2;
    float mean, var;
    switch (f.kind) {
        case SARDINE_SPECKLE_BOXCAR:  // :247-263
            return window_mean_var(d, cols, rows, r, c, half, &mean, &var) ? mean : NaN;
        case SARDINE_SPECKLE_LEE: {  // :269-298
            float cv_n_sq = 1.0f / f.enl;
            if (!window_mean_var(d, cols, rows, r, c, half, &mean, &var)) return NaN;
            float noise_var = mean * mean * cv_n_sq;
            float signal_var = std::max(var - noise_var, 0.0f);
            float denom = signal_var + noise_var;
            float k = denom > 0.0f ? signal_var / denom : 0.0f;
            return mean + k * (centre - mean);
        }
        case SARDINE_SPECKLE_FROST: {  // :304-342
            if (!window_mean_var(d, cols, rows, r, c, half, &mean, &var)) return NaN;
            float cv_sq = mean > 0.0f ? var / (mean * mean) : 0.0f;
            float alpha = f.damping * cv_sq;
            double wsum = 0.0, vwsum = 0.0;
            for (int dy = -half; dy <= half; ++dy)
                for (int dx = -half; dx <= half; ++dx) {
                    int64_t rr = r + dy, cc = c + dx;
                    if (rr < 0 || cc < 0 || rr >= rows || cc >= cols) continue;
                    float v = d[rr * cols + cc];
                    if (!std::isfinite(v)) continue;
                    float dist = (float)(std::abs(dy) + std::abs(dx));
                    double w = (double)std::exp(-alpha * dist);
                    wsum += w;
                    vwsum += w * (double)v;
                }
            return wsum > 0.0 ? (float)(vwsum / wsum) : NaN;
        }
        case SARDINE_SPECKLE_GAMMA_MAP: {  // :348-411
            float enl = f.enl;
            float cv_n_sq = 1.0f / enl;
            float cv_max_sq = cv_n_sq < 1.0f ? 2.0f * cv_n_sq / (1.0f - cv_n_sq) : INFINITY;
            if (!window_mean_var(d, cols, rows, r, c, half, &mean, &var)) return NaN;
            float cv_sq = mean > 0.0f ? var / (mean * mean) : 0.0f;
            if (cv_sq <= cv_n_sq) return mean;
            if (cv_sq >= cv_max_sq) return centre;
            float alpha = (1.0f + cv_n_sq) / (cv_sq - cv_n_sq);
            float b = (alpha - enl - 1.0f) * mean;
            float disc = b * b + 4.0f * alpha * enl * mean * centre;
            return disc >= 0.0f ? ((enl - alpha + 1.0f) * mean + std::sqrt(disc)) / (2.0f * enl) : mean;
        }
        case SARDINE_SPECKLE_REFINED_LEE: {  // :522-569
            float cv_n_sq = 1.0f / f.enl;
            float best_cv = INFINITY, best_mean = centre, best_var = 0.0f;
            for (int dir = 0; dir < 8; ++dir) {
                float m, v;
                if (!rl_dir_mean_var(d, cols, rows, r, c, dir, &m, &v)) continue;
                if (!(m > 0.0f)) continue;
                float cv = v / (m * m);
                if (cv < best_cv) { best_cv = cv; best_mean = m; best_var = v; }
            }
            if (!std::isfinite(best_cv)) return centre;
            float noise_var = best_mean * best_mean * cv_n_sq;
            float signal_var = std::max(best_var - noise_var, 0.0f);
            float denom = signal_var + noise_var;
            float k = denom > 0.0f ? signal_var / denom : 0.0f;
            return best_mean + k * (centre - best_mean);
        }
    }
    return NaN;
}
}  // namespace

extern "C" {

int oracle_speckle(const float* data, int64_t data_len, int64_t cols, int64_t rows,
                   const sardine_speckle_filter* f, float* out, int n_threads) {
    if (cols < 0 || rows < 0 || (rows > 0 && cols > INT64_MAX / rows)) return SARDINE_ERR_SPECKLE_DIMENSION;
    if (data_len != cols * rows) return SARDINE_ERR_SPECKLE_DIMENSION;
    int win = f->kind == SARDINE_SPECKLE_REFINED_LEE ? 7 : f->window;
    if (win < 3 || win % 2 == 0) return SARDINE_ERR_SPECKLE_BAD_WINDOW;
    if (f->kind == SARDINE_SPECKLE_LEE || f->kind == SARDINE_SPECKLE_GAMMA_MAP || f->kind == SARDINE_SPECKLE_REFINED_LEE)
        if (!(f->enl > 0.0f) || !std::isfinite(f->enl)) return SARDINE_ERR_SPECKLE_BAD_ENL;
    if (f->kind == SARDINE_SPECKLE_FROST)
        if (!std::isfinite(f->damping) || f->damping < 0.0f) return SARDINE_ERR_SPECKLE_BAD_DAMPING;
    if (f->kind < 0 || f->kind > 4) return SARDINE_ERR_INVALID_ARGUMENT;
    sardine_speckle_filter ff = *f;
    ff.window = win;
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        for (;;) {
            int64_t r = next.fetch_add(1);
            if (r >= rows) break;
            for (int64_t c = 0; c < cols; ++c) out[r * cols + c] = speckle_pixel(data, cols, rows, r, c, ff);
        }
    };
    int nt = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; ++i) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    return 0;
}

// to_db_inplace, export.rs:59-94
int oracle_to_db_inplace(float* data, int64_t n, float floor_lin, uint64_t counts[2]) {
    if (floor_lin < 0.0f) return SARDINE_ERR_INVALID_ARGUMENT;
    uint64_t conv = 0, masked = 0;
    for (int64_t i = 0; i < n; ++i) {
        float v = data[i];
        if (std::isnan(v)) continue;
        if (v <= floor_lin) { data[i] = std::numeric_limits<float>::quiet_NaN(); ++masked; }
        else { data[i] = 10.0f * std::log10(v); ++conv; }
    }
    counts[0] = conv;
    counts[1] = masked;
    return 0;
}

}  // extern "C"

// ── COG export (export.rs:329-760) ─────────────────────────────────────────
namespace {

// downsample_2x (export.rs:362-392): 2x2 means of the non-NaN pixels in f64,
// NaN where all four (inside the raster) are NaN.
void downsample_2x(const float* src, int64_t sc, int64_t sr, std::vector<float>& dst, int64_t* dc_out,
                   int64_t* dr_out) {
    const int64_t dc = (sc + 1) / 2, dr = (sr + 1) / 2;
    dst.assign((size_t)(dc * dr), std::numeric_limits<float>::quiet_NaN());
    for (int64_t r = 0; r < dr; ++r)
        for (int64_t c = 0; c < dc; ++c) {
            double sum = 0.0;
            uint32_t count = 0;
            for (int64_t a = 0; a < 2; ++a)
                for (int64_t b = 0; b < 2; ++b) {
                    const int64_t y = 2 * r + a, x = 2 * c + b;
                    if (y < sr && x < sc) {
                        const float v = src[y * sc + x];
                        if (!std::isnan(v)) {
                            sum += (double)v;
                            ++count;
                        }
                    }
                }
            if (count > 0) dst[(size_t)(r * dc + c)] = (float)(sum / (double)count);
        }
    *dc_out = dc;
    *dr_out = dr;
}

// OutputCrs::geo_key_directory (output_crs.rs:265-287) for WGS84 / UTM.
std::vector<uint16_t> geo_key_directory(int crs, int zone) {
    if (crs == SARDINE_CRS_WGS84) return {1, 1, 0, 3, 1024, 0, 1, 2, 1025, 0, 1, 1, 2048, 0, 1, 4326};
    const uint16_t epsg = (uint16_t)((crs == SARDINE_CRS_UTM_SOUTH ? 32700 : 32600) + zone);
    return {1, 1, 0, 4, 1024, 0, 1, 1, 1025, 0, 1, 1, 3072, 0, 1, epsg, 3076, 0, 1, 9001};
}

template <typename T>
void put(std::vector<uint8_t>& b, T v) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
}

// write_cog_raw_inner's pyramid sizes, format choice, offsets and the bytes
// before the first tile (export.rs:430-670).  levels: coarsest overview
// first, full raster last.
struct CogPlan {
    std::vector<std::pair<int64_t, int64_t>> levels;
    std::vector<uint64_t> tile_data;  // first tile's offset per level
    std::vector<uint8_t> head;
};

int cog_plan(int64_t cols, int64_t rows, const double gt[6], int crs, int zone, int32_t tile_size, CogPlan* P) {
    const double pw = gt[1], ph = -gt[5];
    if (!std::isfinite(pw) || pw <= 0.0 || !std::isfinite(ph) || ph <= 0.0) return SARDINE_ERR_EXPORT_INVALID_GEOTRANSFORM;
    if (tile_size < 64 || tile_size > 4096 || (tile_size & (tile_size - 1)) != 0) return SARDINE_ERR_EXPORT_INVALID_TILE_SIZE;
    const std::vector<uint16_t> keys = geo_key_directory(crs, zone);
    const int64_t ts = tile_size;
    std::vector<std::pair<int64_t, int64_t>> ov;
    int64_t c = (cols + 1) / 2, r = (rows + 1) / 2;
    while (c > ts || r > ts) {
        ov.push_back({c, r});
        c = (c + 1) / 2;
        r = (r + 1) / 2;
    }
    ov.push_back({c, r});
    std::reverse(ov.begin(), ov.end());
    P->levels = ov;
    P->levels.push_back({cols, rows});
    const bool big = (unsigned __int128)rows * (unsigned __int128)cols * 4 > 0xffffffffull;
    const uint64_t n_entries = 17, geo_bytes = keys.size() * 2;
    const uint64_t ifd_fixed = big ? 8 + n_entries * 20 + 8 : 2 + n_entries * 12 + 4;
    const uint64_t arr_entry = big ? 8 : 4, header = big ? 16 : 8;
    auto ntiles = [&](int64_t cc, int64_t rr) { return ((cc + ts - 1) / ts) * ((rr + ts - 1) / ts); };
    uint64_t off = header;
    std::vector<uint64_t> ifd_off, toff_arr, tbc_arr;
    for (auto& l : P->levels) {
        ifd_off.push_back(off);
        off += ifd_fixed;
        const uint64_t n = (uint64_t)ntiles(l.first, l.second);
        toff_arr.push_back(off);
        off += n * arr_entry;
        tbc_arr.push_back(off);
        off += n * arr_entry;
    }
    const uint64_t scale_off = off;
    off += 24;
    const uint64_t tie_off = off;
    off += 48;
    const uint64_t key_off = off;
    off += geo_bytes;
    const uint64_t tile_bytes = (uint64_t)ts * (uint64_t)ts * 4;
    std::vector<std::vector<uint64_t>> tile_off;
    for (auto& l : P->levels) {
        std::vector<uint64_t> t;
        P->tile_data.push_back(off);
        for (int64_t i = 0; i < ntiles(l.first, l.second); ++i) {
            t.push_back(off);
            off += tile_bytes;
        }
        tile_off.push_back(t);
    }
    std::vector<uint8_t>& b = P->head;
    b.clear();
    b.push_back('I');
    b.push_back('I');
    if (big) {
        put<uint16_t>(b, 43);
        put<uint16_t>(b, 8);
        put<uint16_t>(b, 0);
        put<uint64_t>(b, ifd_off[0]);
    } else {
        put<uint16_t>(b, 42);
        put<uint32_t>(b, (uint32_t)ifd_off[0]);
    }
    for (size_t k = 0; k < P->levels.size(); ++k) {
        const int64_t lc = P->levels[k].first, lr = P->levels[k].second;
        const uint64_t n = (uint64_t)ntiles(lc, lr);
        const uint64_t next = k + 1 < P->levels.size() ? ifd_off[k + 1] : 0;
        if (big) {
            put<uint64_t>(b, n_entries);
            auto e = [&](uint16_t tag, uint16_t typ, uint64_t cnt, uint64_t val) {
                put<uint16_t>(b, tag);
                put<uint16_t>(b, typ);
                put<uint64_t>(b, cnt);
                put<uint64_t>(b, val);
            };
            const uint64_t rat = (1ull << 32) | 72ull;
            e(256, 4, 1, (uint64_t)lc);
            e(257, 4, 1, (uint64_t)lr);
            e(258, 3, 1, 32);
            e(259, 3, 1, 1);
            e(262, 3, 1, 1);
            e(277, 3, 1, 1);
            e(282, 5, 1, rat);
            e(283, 5, 1, rat);
            e(296, 3, 1, 1);
            e(322, 4, 1, (uint64_t)ts);
            e(323, 4, 1, (uint64_t)ts);
            e(324, 16, n, toff_arr[k]);
            e(325, 16, n, tbc_arr[k]);
            e(339, 3, 1, 3);
            e(33550, 12, 3, scale_off);
            e(33922, 12, 6, tie_off);
            e(34735, 3, keys.size(), key_off);
            put<uint64_t>(b, next);
        } else {
            put<uint16_t>(b, (uint16_t)n_entries);
            auto e = [&](uint16_t tag, uint16_t typ, uint32_t cnt, uint32_t val) {
                put<uint16_t>(b, tag);
                put<uint16_t>(b, typ);
                put<uint32_t>(b, cnt);
                put<uint32_t>(b, val);
            };
            e(256, 4, 1, (uint32_t)lc);
            e(257, 4, 1, (uint32_t)lr);
            e(258, 3, 1, 32);
            e(259, 3, 1, 1);
            e(262, 3, 1, 1);
            e(277, 3, 1, 1);
            e(282, 4, 1, 72);
            e(283, 4, 1, 72);
            e(296, 3, 1, 1);
            e(322, 4, 1, (uint32_t)ts);
            e(323, 4, 1, (uint32_t)ts);
            e(324, 4, (uint32_t)n, (uint32_t)toff_arr[k]);
            e(325, 4, (uint32_t)n, (uint32_t)tbc_arr[k]);
            e(339, 3, 1, 3);
            e(33550, 12, 3, (uint32_t)scale_off);
            e(33922, 12, 6, (uint32_t)tie_off);
            e(34735, 3, (uint32_t)keys.size(), (uint32_t)key_off);
            put<uint32_t>(b, (uint32_t)next);
        }
        for (uint64_t o : tile_off[k]) {
            if (big) put<uint64_t>(b, o);
            else put<uint32_t>(b, (uint32_t)o);
        }
        for (uint64_t i = 0; i < n; ++i) {
            if (big) put<uint64_t>(b, tile_bytes);
            else put<uint32_t>(b, (uint32_t)tile_bytes);
        }
    }
    put<double>(b, gt[1]);
    put<double>(b, -gt[5]);
    put<double>(b, 0.0);
    put<double>(b, 0.0);
    put<double>(b, 0.0);
    put<double>(b, 0.0);
    put<double>(b, gt[0]);
    put<double>(b, gt[3]);
    put<double>(b, 0.0);
    for (uint16_t k : keys) put<uint16_t>(b, k);
    return 0;
}

}  // namespace

extern "C" {

int oracle_downsample_2x(const float* src, int64_t sc, int64_t sr, float* dst) {
    std::vector<float> d;
    int64_t dc, dr;
    downsample_2x(src, sc, sr, d, &dc, &dr);
    std::memcpy(dst, d.data(), d.size() * sizeof(float));
    return 0;
}

// The bytes before the first tile (header, IFDs, tile arrays, GeoTIFF
// blocks) of write_cog_with_crs's file; *size = their count (written to
// `out` when it holds them).
int oracle_cog_header(int64_t cols, int64_t rows, const double gt[6], int crs, int zone, int32_t tile_size,
                      uint8_t* out, uint64_t cap, uint64_t* size) {
    CogPlan P;
    const int st = cog_plan(cols, rows, gt, crs, zone, tile_size, &P);
    if (st) return st;
    *size = P.head.size();
    if (out && cap >= P.head.size()) std::memcpy(out, P.head.data(), P.head.size());
    return 0;
}

// write_cog_with_crs + write_cog_raw_inner (export.rs:329-760): validation,
// the overview pyramid (coarsest first), the header and every level's
// NaN-padded tiles, written to `path`.
int oracle_write_cog(const char* path, const float* data, int64_t cols, int64_t rows, const double gt[6], int crs,
                     int zone, int32_t tile_size) {
    CogPlan P;
    const int st = cog_plan(cols, rows, gt, crs, zone, tile_size, &P);
    if (st) return st;
    // overview rasters, finest first as the reference builds them
    std::vector<std::vector<float>> ov;
    {
        std::vector<float> d;
        int64_t c, r;
        downsample_2x(data, cols, rows, d, &c, &r);
        for (size_t k = 0; k + 2 < P.levels.size(); ++k) {
            std::vector<float> n;
            int64_t nc, nr;
            downsample_2x(d.data(), c, r, n, &nc, &nr);
            ov.push_back(std::move(d));
            d.swap(n);
            c = nc;
            r = nr;
        }
        ov.push_back(std::move(d));
    }
    std::reverse(ov.begin(), ov.end());
    std::vector<const float*> srcs;
    for (auto& v : ov) srcs.push_back(v.data());
    srcs.push_back(data);
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return SARDINE_ERR_EXPORT_IO;
    bool ok = std::fwrite(P.head.data(), 1, P.head.size(), f) == P.head.size();
    const int64_t ts = tile_size;
    std::vector<float> tile((size_t)(ts * ts));
    for (size_t k = 0; k < P.levels.size() && ok; ++k) {
        const int64_t c = P.levels[k].first, r = P.levels[k].second;
        const float* src = srcs[k];
        const int64_t nx = (c + ts - 1) / ts, ny = (r + ts - 1) / ts;
        for (int64_t ty = 0; ty < ny && ok; ++ty)
            for (int64_t tx = 0; tx < nx && ok; ++tx) {
                std::fill(tile.begin(), tile.end(), std::numeric_limits<float>::quiet_NaN());
                const int64_t r0 = ty * ts, c0 = tx * ts, r1 = std::min(r0 + ts, r), c1 = std::min(c0 + ts, c);
                for (int64_t y = r0; y < r1; ++y)
                    std::memcpy(&tile[(size_t)((y - r0) * ts)], src + y * c + c0, (size_t)(c1 - c0) * sizeof(float));
                ok = std::fwrite(tile.data(), sizeof(float), tile.size(), f) == tile.size();
            }
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : SARDINE_ERR_EXPORT_IO;
}

}  // extern "C"

// ── striped GeoTIFF / BigTIFF (export.rs:139-284, 762-1219) ─────────────────
namespace {

// needs_bigtiff (export.rs:762-786), with the reference's saturating sums.
bool needs_bigtiff(uint64_t rows, uint64_t cols, uint64_t bpp, uint64_t geo_shorts) {
    const uint64_t prelude = geo_shorts > 0 ? 8 + 198 + 16 + 24 + 48 + geo_shorts * 2 : 8 + 162 + 16;
    auto sat_mul = [](uint64_t a, uint64_t b) { return (b != 0 && a > UINT64_MAX / b) ? UINT64_MAX : a * b; };
    auto sat_add = [](uint64_t a, uint64_t b) { return a > UINT64_MAX - b ? UINT64_MAX : a + b; };
    const uint64_t image = sat_mul(sat_mul(rows, cols), bpp);
    const uint64_t strips = (rows + 63) / 64;
    return sat_add(sat_add(prelude, sat_mul(strips, 8)), image) > 0xffffffffull;
}

// The bytes before the raster of write_geotiff_raw_inner (export.rs:788-1010,
// classic) or write_bigtiff_raw_inner (export.rs:1023-1219); keys empty = no
// GeoTIFF tags.  rows >= 1.
void striped_header(uint64_t cols, uint64_t rows, uint32_t bpp, const double* gt, const std::vector<uint16_t>& keys,
                    bool big, std::vector<uint8_t>& b) {
    const bool geo = !keys.empty();
    const uint16_t bits = (uint16_t)(bpp * 8), sf = bpp == 1 ? 1 : 3;
    b.clear();
    if (!big) {
        const uint16_t n_entries = geo ? 16 : 13;
        const uint32_t geo_bytes = (uint32_t)keys.size() * 2;
        const uint32_t n_strips = ((uint32_t)rows + 63u) / 64u;
        const uint32_t row_bytes = (uint32_t)cols * bpp;
        const uint32_t full = 64u * row_bytes;
        const uint32_t last = ((uint32_t)rows - (n_strips - 1u) * 64u) * row_bytes;
        const uint32_t ifd_size = 2u + n_entries * 12u + 4u;
        const uint32_t so = 8u + ifd_size, sb = so + n_strips * 4u, xr = sb + n_strips * 4u, yr = xr + 8u;
        const uint32_t ps = yr + 8u, tp = ps + 24u, gk = tp + 48u, im = geo ? gk + geo_bytes : yr + 8u;
        b.push_back('I');
        b.push_back('I');
        put<uint16_t>(b, 42);
        put<uint32_t>(b, 8);
        put<uint16_t>(b, n_entries);
        auto e = [&](uint16_t tag, uint16_t typ, uint32_t cnt, uint32_t val) {
            put<uint16_t>(b, tag);
            put<uint16_t>(b, typ);
            put<uint32_t>(b, cnt);
            put<uint32_t>(b, val);
        };
        e(256, 4, 1, (uint32_t)cols);
        e(257, 4, 1, (uint32_t)rows);
        e(258, 3, 1, bits);
        e(259, 3, 1, 1);
        e(262, 3, 1, 1);
        e(273, 4, n_strips, so);
        e(277, 3, 1, 1);
        e(278, 4, 1, 64);
        e(279, 4, n_strips, sb);
        e(282, 5, 1, xr);
        e(283, 5, 1, yr);
        e(296, 3, 1, 1);
        e(339, 3, 1, sf);
        if (geo) {
            e(33550, 12, 3, ps);
            e(33922, 12, 6, tp);
            e(34735, 3, (uint32_t)keys.size(), gk);
        }
        put<uint32_t>(b, 0);
        for (uint32_t i = 0; i < n_strips; ++i) put<uint32_t>(b, (uint32_t)((uint64_t)im + (uint64_t)i * full));
        for (uint32_t i = 0; i < n_strips; ++i) put<uint32_t>(b, i == n_strips - 1 ? last : full);
        put<uint32_t>(b, 72);
        put<uint32_t>(b, 1);
        put<uint32_t>(b, 72);
        put<uint32_t>(b, 1);
    } else {
        const uint64_t n_entries = geo ? 16 : 13;
        const uint64_t geo_bytes = keys.size() * 2;
        const uint64_t n_strips = (rows + 63) / 64;
        const uint64_t row_bytes = cols * bpp, full = 64 * row_bytes;
        const uint64_t last = (rows - (n_strips - 1) * 64) * row_bytes;
        const uint64_t ifd_end = 16 + 8 + n_entries * 20 + 8;
        const bool inl = n_strips == 1;
        const uint64_t so = inl ? 0 : ifd_end, sb = inl ? 0 : ifd_end + n_strips * 8;
        const uint64_t after = inl ? ifd_end : ifd_end + 2 * n_strips * 8;
        const uint64_t ps = geo ? after : 0, tp = geo ? after + 24 : 0, gk = geo ? after + 72 : 0;
        const uint64_t im = geo ? gk + geo_bytes : after;
        b.push_back('I');
        b.push_back('I');
        put<uint16_t>(b, 43);
        put<uint16_t>(b, 8);
        put<uint16_t>(b, 0);
        put<uint64_t>(b, 16);
        put<uint64_t>(b, n_entries);
        auto e = [&](uint16_t tag, uint16_t typ, uint64_t cnt, uint64_t val) {
            put<uint16_t>(b, tag);
            put<uint16_t>(b, typ);
            put<uint64_t>(b, cnt);
            put<uint64_t>(b, val);
        };
        const uint64_t rat = (1ull << 32) | 72ull;
        e(256, 4, 1, cols);
        e(257, 4, 1, rows);
        e(258, 3, 1, bits);
        e(259, 3, 1, 1);
        e(262, 3, 1, 1);
        e(273, 16, n_strips, inl ? im : so);
        e(277, 3, 1, 1);
        e(278, 4, 1, 64);
        e(279, 16, n_strips, inl ? last : sb);
        e(282, 5, 1, rat);
        e(283, 5, 1, rat);
        e(296, 3, 1, 1);
        e(339, 3, 1, sf);
        if (geo) {
            e(33550, 12, 3, ps);
            e(33922, 12, 6, tp);
            e(34735, 3, keys.size(), gk);
        }
        put<uint64_t>(b, 0);
        if (!inl) {
            for (uint64_t i = 0; i < n_strips; ++i) put<uint64_t>(b, im + i * full);
            for (uint64_t i = 0; i < n_strips; ++i) put<uint64_t>(b, i == n_strips - 1 ? last : full);
        }
    }
    if (geo) {
        put<double>(b, gt[1]);
        put<double>(b, -gt[5]);
        put<double>(b, 0.0);
        put<double>(b, 0.0);
        put<double>(b, 0.0);
        put<double>(b, 0.0);
        put<double>(b, gt[0]);
        put<double>(b, gt[3]);
        put<double>(b, 0.0);
        for (uint16_t k : keys) put<uint16_t>(b, k);
    }
}

// gt == null: write_tiff_no_crs; else write_geotiff(_u8)_with_crs.  force:
// 0 = the reference's choice, 1 = classic, 2 = BigTIFF (the inner writers,
// as the reference's tests call them).
int striped_plan(int64_t cols, int64_t rows, int32_t bpp, const double* gt, int crs, int zone, int force,
                 std::vector<uint8_t>* head) {
    if (gt) {
        const double pw = gt[1], ph = -gt[5];
        if (!std::isfinite(pw) || pw <= 0.0 || !std::isfinite(ph) || ph <= 0.0)
            return SARDINE_ERR_EXPORT_INVALID_GEOTRANSFORM;
    }
    if (rows <= 0 || cols < 0 || (bpp != 1 && bpp != 4)) return SARDINE_ERR_INVALID_ARGUMENT;
    const std::vector<uint16_t> keys = gt ? geo_key_directory(crs, zone) : std::vector<uint16_t>();
    const bool big = force == 2 || (force == 0 && needs_bigtiff((uint64_t)rows, (uint64_t)cols, (uint64_t)bpp, keys.size()));
    striped_header((uint64_t)cols, (uint64_t)rows, (uint32_t)bpp, gt, keys, big, *head);
    return 0;
}

}  // namespace

extern "C" {

int oracle_needs_bigtiff(uint64_t rows, uint64_t cols, uint64_t bpp, uint64_t geo_shorts) {
    return needs_bigtiff(rows, cols, bpp, geo_shorts) ? 1 : 0;
}

// The bytes before the raster; *size = their count (copied when cap allows).
int oracle_geotiff_header(int64_t cols, int64_t rows, int32_t bpp, const double* gt, int crs, int zone, int force,
                          uint8_t* out, uint64_t cap, uint64_t* size) {
    std::vector<uint8_t> h;
    const int st = striped_plan(cols, rows, bpp, gt, crs, zone, force, &h);
    if (st) return st;
    *size = h.size();
    if (out && cap >= h.size()) std::memcpy(out, h.data(), h.size());
    return 0;
}

// write_geotiff_with_crs / write_geotiff_u8_with_crs / write_tiff_no_crs:
// the header, then the raster bytes as they lie in memory.
int oracle_write_geotiff(const char* path, const void* data, int64_t cols, int64_t rows, int32_t bpp, const double* gt,
                         int crs, int zone, int force) {
    std::vector<uint8_t> h;
    const int st = striped_plan(cols, rows, bpp, gt, crs, zone, force, &h);
    if (st) return st;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return SARDINE_ERR_EXPORT_IO;
    const size_t n = (size_t)cols * (size_t)rows * (size_t)bpp;
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    if (ok && n) ok = std::fwrite(data, 1, n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : SARDINE_ERR_EXPORT_IO;
}

}  // extern "C"

// ── pre-TC multilook (scene_prep.rs:1060-1161) ─────────────────────────────
extern "C" {

// apply_multilook: NaN-aware boxcar decimation of σ⁰ (and of NESZ over the
// same non-NaN σ⁰ samples), output lines/samples = floor(in / looks), with
// the centre-aligned geometry update.  range_looks = azimuth_looks = 1
// returns the input unchanged.  Errors (anyhow bails in the reference):
// zero factors, factors larger than the image, a non-positive ATI with
// azimuth_looks > 1 — all SARDINE_ERR_INVALID_ARGUMENT.
int oracle_multilook(const float* data, const float* nesz, const sardine_merged_info* in, int32_t range_looks,
                     int32_t azimuth_looks, double ati_s, float* out_data, float* out_nesz,
                     sardine_merged_info* out) {
    const int64_t R = range_looks, A = azimuth_looks;
    if (R == 1 && A == 1) {
        *out = *in;
        const size_t n = (size_t)(in->lines * in->samples);
        if (n) {
            std::memcpy(out_data, data, n * sizeof(float));
            std::memcpy(out_nesz, nesz, n * sizeof(float));
        }
        return 0;
    }
    if (R <= 0 || A <= 0) return SARDINE_ERR_INVALID_ARGUMENT;
    const int64_t il = in->lines, is = in->samples, ol = il / A, os = is / R;
    if (ol == 0 || os == 0) return SARDINE_ERR_INVALID_ARGUMENT;
    *out = *in;
    out->lines = ol;
    out->samples = os;
    out->range_pixel_spacing_s = in->range_pixel_spacing_s * (double)R;
    out->range_pixel_spacing_m = in->range_pixel_spacing_m * (double)R;
    out->near_slant_range_time_s = in->near_slant_range_time_s + (double)(R - 1) / 2.0 * in->range_pixel_spacing_s;
    if (A > 1) {
        if (!std::isfinite(ati_s) || ati_s <= 0.0) return SARDINE_ERR_INVALID_ARGUMENT;
        out->azimuth_start_time_us = in->azimuth_start_time_us + (int64_t)((double)(A - 1) / 2.0 * ati_s * 1e6);
    }
    for (int64_t l = 0; l < ol; ++l)
        for (int64_t c = 0; c < os; ++c) {
            double sum = 0.0, sum_n = 0.0;
            uint32_t count = 0;
            for (int64_t da = 0; da < A; ++da)
                for (int64_t dr = 0; dr < R; ++dr) {
                    const int64_t k = (l * A + da) * is + c * R + dr;
                    const float v = data[k];
                    if (!std::isnan(v)) {
                        sum += (double)v;
                        sum_n += (double)nesz[k];
                        ++count;
                    }
                }
            const float qnan = std::numeric_limits<float>::quiet_NaN();
            out_data[l * os + c] = count ? (float)(sum / (double)count) : qnan;
            out_nesz[l * os + c] = count ? (float)(sum_n / (double)count) : qnan;
        }
    return 0;
}

}  // extern "C"

// ── ground-range projection (ground_range.rs:82-490) ───────────────────────
namespace {

// RangeInterpolator (ground_range.rs:157-236): the mid-azimuth grid row of
// one swath as (slant time, incidence °), sorted by slant time.
struct GrdInterp {
    std::vector<double> t, a;
    double tmin = 0, tmax = 0;
    // interpolate (ground_range.rs:203-235)
    bool at(double x, double* out) const {
        if (x < tmin || x > tmax) return false;
        const size_t i = (size_t)(std::lower_bound(t.begin(), t.end(), x) - t.begin());
        if (i < t.size() && t[i] == x) {
            *out = a[i];
            return true;
        }
        if (i == 0) {
            *out = a[0];
            return true;
        }
        if (i >= t.size()) {
            *out = a.back();
            return true;
        }
        const double f = (x - t[i - 1]) / (t[i] - t[i - 1]);
        *out = a[i - 1] + f * (a[i] - a[i - 1]);
        return true;
    }
};

GrdInterp grd_interp(const sardine_grd_grid_point* p, int32_t n) {
    std::vector<uint32_t> lines;
    for (int32_t i = 0; i < n; ++i) lines.push_back(p[i].line);
    std::sort(lines.begin(), lines.end());
    lines.erase(std::unique(lines.begin(), lines.end()), lines.end());
    const uint32_t mid = lines[lines.size() / 2];
    std::vector<std::pair<double, double>> pts;
    for (int32_t i = 0; i < n; ++i)
        if (p[i].line == mid) pts.push_back({p[i].slant_range_time_s, p[i].incidence_angle_deg});
    std::stable_sort(pts.begin(), pts.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    GrdInterp g;
    for (auto& q : pts) {
        g.t.push_back(q.first);
        g.a.push_back(q.second);
    }
    g.tmin = g.t.front();
    g.tmax = g.t.back();
    return g;
}

// interpolate_at_slant_time (ground_range.rs:238-290)
bool grd_angle(const std::vector<GrdInterp>& gs, double x, double* out) {
    bool have = false;
    double best_d = 0, best_a = 0;
    for (const GrdInterp& g : gs) {
        double a;
        if (g.at(x, &a)) {
            const double d = std::fabs(x - (g.tmin + g.tmax) / 2.0);
            if (!have || d < best_d) {
                have = true;
                best_d = d;
                best_a = a;
            }
        }
    }
    if (have) {
        *out = best_a;
        return true;
    }
    const GrdInterp* near = nullptr;
    double near_d = 0;
    for (const GrdInterp& g : gs) {
        const double d = x < g.tmin ? g.tmin - x : x - g.tmax;
        if (!near || d < near_d) {
            near = &g;
            near_d = d;
        }
    }
    if (near && near_d < (near->tmax - near->tmin) * 0.01)
        return near->at(std::min(std::max(x, near->tmin), near->tmax), out);
    return false;
}

}  // namespace

extern "C" {

// build_incidence_profile (ground_range.rs:103-155): radians per merged column;
// returns the first failing column + 1 as a negative number, or a status.
int oracle_grd_incidence_profile(const sardine_grd_grid_point* points, const int32_t* counts, int32_t n_swaths,
                                 double near_t, double rps_s, int64_t samples, double* out_rad) {
    std::vector<GrdInterp> gs;
    int64_t off = 0;
    for (int32_t k = 0; k < n_swaths; ++k) {
        if (counts[k] <= 0) return SARDINE_ERR_GRD_NO_GRID_POINTS;
        gs.push_back(grd_interp(points + off, counts[k]));
        off += counts[k];
    }
    std::stable_sort(gs.begin(), gs.end(), [](const GrdInterp& x, const GrdInterp& y) { return x.tmin < y.tmin; });
    for (int64_t c = 0; c < samples; ++c) {
        double deg;
        if (!grd_angle(gs, near_t + (double)c * rps_s, &deg)) return SARDINE_ERR_GRD_INTERPOLATION_FAILED;
        out_rad[c] = deg * (3.14159265358979323846 / 180.0);
    }
    return 0;
}

// build_ground_to_slant_mapping (ground_range.rs:446-490): kind 0 = exact
// (left = the column), 1 = interpolate between left and left + 1.
void oracle_grd_column_map(const double* cum, int64_t n, double target, int64_t out_samples, int32_t* kind,
                           int64_t* left, double* frac) {
    int64_t start = 0;
    for (int64_t j = 0; j < out_samples; ++j) {
        const double g = (double)j * target;
        const double* b = cum + start;
        const double* e = cum + n;
        const double* lb = std::lower_bound(b, e, g);
        if (lb != e && *lb == g) {
            kind[j] = 0;
            left[j] = start + (lb - b);
            frac[j] = 0.0;
            start = left[j];
            continue;
        }
        const int64_t idx = start + (lb - b);
        if (idx == 0) {
            kind[j] = 0;
            left[j] = 0;
            frac[j] = 0.0;
        } else if (idx >= n) {
            kind[j] = 0;
            left[j] = n - 1;
            frac[j] = 0.0;
        } else {
            kind[j] = 1;
            left[j] = idx - 1;
            frac[j] = (g - cum[idx - 1]) / (cum[idx] - cum[idx - 1]);
            start = idx - 1;
        }
    }
}

// to_ground_range (ground_range.rs:303-442).  out == null: only *img.
int oracle_to_ground_range(const float* data, const sardine_merged_info* m, const sardine_grd_grid_point* points,
                           const int32_t* counts, int32_t n_swaths, double target, double az_spacing, float* out,
                           sardine_grd_image* img) {
    if (m->lines == 0 || m->samples == 0) return SARDINE_ERR_GRD_EMPTY_IMAGE;
    if (target <= 0.0) return SARDINE_ERR_GRD_INVALID_SPACING;
    if (!std::isfinite(az_spacing) || az_spacing <= 0.0) return SARDINE_ERR_GRD_INVALID_AZIMUTH_SPACING;
    const int64_t S = m->samples, Lines = m->lines;
    std::vector<double> rad((size_t)S);
    const int st = oracle_grd_incidence_profile(points, counts, n_swaths, m->near_slant_range_time_s,
                                                m->range_pixel_spacing_s, S, rad.data());
    if (st) return st;
    std::vector<double> gsp((size_t)S), cum((size_t)S);
    for (int64_t c = 0; c < S; ++c) gsp[c] = m->range_pixel_spacing_m / std::sin(rad[c]);
    double acc = 0.0;
    cum[0] = 0.0;
    for (int64_t c = 1; c < S; ++c) {
        acc += (gsp[c - 1] + gsp[c]) / 2.0;
        cum[c] = acc;
    }
    const double q = std::floor(cum[S - 1] / target);
    const int64_t os = (q >= 1.0 && q < 9.2e18) ? (int64_t)q : 0;  // `as usize`: NaN and negatives → 0
    if (os == 0) return SARDINE_ERR_GRD_INVALID_SPACING;
    const double aq = std::max(std::round(target / az_spacing), 1.0);
    const int64_t A = (int64_t)aq;
    img->lines = Lines / A;
    img->samples = os;
    img->range_pixel_spacing_m = target;
    img->azimuth_pixel_spacing_m = az_spacing * (double)A;
    img->range_looks = 1;
    img->azimuth_looks = (int32_t)A;
    if (!out) return 0;
    std::vector<int32_t> kind((size_t)os);
    std::vector<int64_t> left((size_t)os);
    std::vector<double> frac((size_t)os);
    oracle_grd_column_map(cum.data(), S, target, os, kind.data(), left.data(), frac.data());
    std::vector<float> proj((size_t)(Lines * os));
    for (int64_t l = 0; l < Lines; ++l)
        for (int64_t j = 0; j < os; ++j) {
            const float* row = data + l * S;
            float v;
            if (kind[j] == 0) {
                v = row[left[j]];
            } else {
                const float v0 = row[left[j]], v1 = row[left[j] + 1];
                v = (std::isnan(v0) || std::isnan(v1)) ? std::numeric_limits<float>::quiet_NaN()
                                                       : v0 + (v1 - v0) * (float)frac[j];
            }
            proj[(size_t)(l * os + j)] = v;
        }
    for (int64_t ol = 0; ol < img->lines; ++ol)
        for (int64_t j = 0; j < os; ++j) {
            double sum = 0.0;
            uint32_t count = 0;
            for (int64_t l = ol * A; l < ol * A + A; ++l) {
                const float v = proj[(size_t)(l * os + j)];
                if (!std::isnan(v)) {
                    sum += (double)v;
                    ++count;
                }
            }
            out[ol * os + j] = count ? (float)(sum / (double)count) : std::numeric_limits<float>::quiet_NaN();
        }
    return 0;
}

}  // extern "C"

// ── SLC measurement TIFF reader (slc_reader.rs:200-590) ────────────────────
namespace {

struct SlcFile {
    std::FILE* f = nullptr;
    ~SlcFile() {
        if (f) std::fclose(f);
    }
};

bool read_at(std::FILE* f, uint64_t off, void* dst, size_t n) {
    return std::fseek(f, (long)off, SEEK_SET) == 0 && std::fread(dst, 1, n, f) == n;
}

uint16_t get16(const uint8_t* p, bool le) { return le ? (uint16_t)(p[0] | p[1] << 8) : (uint16_t)(p[0] << 8 | p[1]); }
uint32_t get32(const uint8_t* p, bool le) {
    return le ? (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24
              : (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}

}  // namespace

extern "C" {

// parse_ifd (slc_reader.rs:268-461).  *missing_tag receives the tag of a
// MissingTag error.
int oracle_slc_open(const char* path, sardine_slc_tiff* out, int32_t* missing_tag) {
    SlcFile sf;
    sf.f = std::fopen(path, "rb");
    if (!sf.f) return SARDINE_ERR_SLC_IO;
    uint8_t h[8];
    if (!read_at(sf.f, 0, h, 8)) return SARDINE_ERR_SLC_IO;
    bool le;
    if (h[0] == 'I' && h[1] == 'I') le = true;
    else if (h[0] == 'M' && h[1] == 'M') le = false;
    else return SARDINE_ERR_SLC_NOT_TIFF;
    const uint16_t magic = get16(h + 2, le);
    if (magic == 43) return SARDINE_ERR_SLC_BIGTIFF;
    if (magic != 42) return SARDINE_ERR_SLC_NOT_TIFF;
    const uint64_t ifd = get32(h + 4, le);
    uint8_t nb[2];
    if (!read_at(sf.f, ifd, nb, 2)) return SARDINE_ERR_SLC_IO;
    const size_t n = get16(nb, le);
    std::vector<uint8_t> e(n * 12);
    if (n && std::fread(e.data(), 1, e.size(), sf.f) != e.size()) return SARDINE_ERR_SLC_IO;
    // tag → (have, value); StripOffsets keeps (type, count, raw field)
    bool hw = false, hh = false, hb = false, hc = false, hr = false, hs = false, hf = false, ho = false;
    uint32_t w = 0, ht = 0, bps = 0, comp = 0, rps = 0, spp = 1, fmt = 0;
    uint16_t so_type = 0;
    uint32_t so_count = 0, so_value = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t* p = &e[i * 12];
        const uint16_t tag = get16(p, le), type = get16(p + 2, le);
        const uint32_t count = get32(p + 4, le);
        // entry_scalar: SHORT → first 2 bytes, LONG → 4 bytes, else none
        const bool ok = type == 3 || type == 4;
        const uint32_t v = type == 3 ? get16(p + 8, le) : get32(p + 8, le);
        switch (tag) {
            case 256: hw = ok; w = v; break;
            case 257: hh = ok; ht = v; break;
            case 258: hb = ok; bps = (uint16_t)v; break;
            case 259: hc = ok; comp = (uint16_t)v; break;
            case 278: hr = ok; rps = v; break;
            case 277: hs = ok; spp = (uint16_t)v; break;
            case 339: hf = ok; fmt = (uint16_t)v; break;
            case 273: ho = true; so_type = type; so_count = count; so_value = get32(p + 8, le); break;
            default: break;
        }
    }
    auto missing = [&](int32_t tag) {
        if (missing_tag) *missing_tag = tag;
        return SARDINE_ERR_SLC_MISSING_TAG;
    };
    if (!hw) return missing(256);
    if (!hh) return missing(257);
    if (!hb) return missing(258);
    if (!hc) return missing(259);
    if (!hr) return missing(278);
    if (!hs) spp = 1;  // SamplesPerPixel defaults to 1
    if (!hf) return missing(339);
    if (!ho) return missing(273);
    if (comp != 1 || fmt != 5 || bps != 32 || rps != 1 || spp != 1 || so_count != ht)
        return SARDINE_ERR_SLC_UNSUPPORTED_FORMAT;
    const uint64_t row = (uint64_t)w * 4;
    uint64_t base, last;
    if (so_count == 1) {
        base = last = so_value;
    } else {
        if (so_type != 3 && so_type != 4) return SARDINE_ERR_SLC_UNSUPPORTED_FORMAT;
        const uint64_t each = so_type == 3 ? 2 : 4;
        uint8_t b[4];
        if (!read_at(sf.f, so_value, b, each)) return SARDINE_ERR_SLC_IO;
        base = each == 2 ? get16(b, le) : get32(b, le);
        if (!read_at(sf.f, so_value + ((uint64_t)so_count - 1) * each, b, each)) return SARDINE_ERR_SLC_IO;
        last = each == 2 ? get16(b, le) : get32(b, le);
    }
    if (last != base + ((uint64_t)ht - 1) * row) return SARDINE_ERR_SLC_NON_CONTIGUOUS_STRIPS;
    out->width = w;
    out->height = ht;
    out->little_endian = le ? 1 : 0;
    out->pad = 0;
    out->base_offset = base;
    out->row_bytes = row;
    return 0;
}

// read_burst_raw (slc_reader.rs:532-575) into host memory.
int oracle_slc_read(const char* path, const sardine_slc_tiff* t, int64_t first, int64_t count, int16_t* dst) {
    if (first < 0 || count < 0 || first > (int64_t)t->height || count > (int64_t)t->height - first)
        return SARDINE_ERR_SLC_LINE_OUT_OF_BOUNDS;
    SlcFile sf;
    sf.f = std::fopen(path, "rb");
    if (!sf.f) return SARDINE_ERR_SLC_IO;
    const size_t n = (size_t)count * t->width * 4;
    if (n && !read_at(sf.f, t->base_offset + (uint64_t)first * t->row_bytes, dst, n)) return SARDINE_ERR_SLC_IO;
    if (!t->little_endian)
        for (size_t i = 0; i < n / 2; ++i) dst[i] = (int16_t)(((uint16_t)dst[i] >> 8) | ((uint16_t)dst[i] << 8));
    return 0;
}

}  // extern "C"

// ── geoid grid files (geoid.rs:132-299) ────────────────────────────────────
namespace {

// Rust's float grammar for `str::parse` (sign; inf / infinity / nan in any
// case; or digits[.digits] / .digits with an optional exponent).
bool rust_float_token(const char* s) {
    if (*s == '+' || *s == '-') ++s;
    std::string low;
    for (const char* p = s; *p; ++p) low += (char)std::tolower((unsigned char)*p);
    if (low == "inf" || low == "infinity" || low == "nan") return true;
    int digits = 0;
    while (std::isdigit((unsigned char)*s)) ++s, ++digits;
    if (*s == '.') {
        ++s;
        while (std::isdigit((unsigned char)*s)) ++s, ++digits;
    }
    if (!digits) return false;
    if (*s == 'e' || *s == 'E') {
        ++s;
        if (*s == '+' || *s == '-') ++s;
        int e = 0;
        while (std::isdigit((unsigned char)*s)) ++s, ++e;
        if (!e) return false;
    }
    return *s == 0;
}

}  // namespace

extern "C" {

int oracle_geoid_load_binary(const char* path, int model, float* grid) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return SARDINE_ERR_GEOID_IO;
    std::vector<unsigned char> raw;
    int c;
    while ((c = std::fgetc(f)) != EOF) raw.push_back((unsigned char)c);
    std::fclose(f);
    if (raw.size() != 73 * 144 * 4) return model == 0 ? SARDINE_ERR_GEOID_WRONG_SIZE : SARDINE_ERR_GEOID_WRONG_SIZE_EGM2008;
    for (size_t i = 0; i < 73 * 144; ++i) {
        uint32_t u = (uint32_t)raw[4 * i] | (uint32_t)raw[4 * i + 1] << 8 | (uint32_t)raw[4 * i + 2] << 16 |
                     (uint32_t)raw[4 * i + 3] << 24;
        std::memcpy(&grid[i], &u, 4);
    }
    return 0;
}

int oracle_geoid_load_ww15mgh(const char* path, float* grid) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return SARDINE_ERR_GEOID_IO;
    std::vector<float> vals;
    int headers = 0;
    std::string line;
    auto flush_line = [&](const std::string& ln) -> int {
        std::vector<std::string> tok;
        std::string cur;
        for (char ch : ln) {
            if (std::isspace((unsigned char)ch)) {
                if (!cur.empty()) tok.push_back(cur);
                cur.clear();
            } else {
                cur += ch;
            }
        }
        if (!cur.empty()) tok.push_back(cur);
        if (tok.empty()) return 0;
        if (headers < 2) {
            bool all = true;
            for (auto& t : tok) all = all && rust_float_token(t.c_str());
            if (!all || tok.size() <= 6) {
                ++headers;
                return 0;
            }
        }
        for (auto& t : tok) {
            if (!rust_float_token(t.c_str())) return SARDINE_ERR_GEOID_MALFORMED_ASCII;
            vals.push_back(std::strtof(t.c_str(), nullptr));
        }
        return 0;
    };
    int c, st = 0;
    while ((c = std::fgetc(f)) != EOF && !st) {
        if (c == '\n') {
            st = flush_line(line);
            line.clear();
        } else {
            line += (char)c;
        }
    }
    if (!st) st = flush_line(line);
    std::fclose(f);
    if (st) return st;
    if (vals.size() != 73 * 144) return SARDINE_ERR_GEOID_MALFORMED_ASCII;
    std::memcpy(grid, vals.data(), vals.size() * sizeof(float));
    return 0;
}

}  // extern "C"

// ── SRTM 1" tiles and the mosaic lookup (dem.rs:129-178, 385-437, 694-739) ──
namespace {

struct OracleSrtm {
    int lat_sw, lon_sw;
    std::vector<float> data;  // 3601 x 3601, north row first, NaN voids
};

// bilinear_on_grid (dem.rs:385-409); false = a NaN tap (None)
bool grid_bilinear(const std::vector<float>& d, size_t dim, double row_f, double col_f, size_t max_rc, float* out) {
    const size_t r0 = std::min((size_t)row_f, max_rc >= 1 ? max_rc - 1 : 0);
    const size_t c0 = std::min((size_t)col_f, dim >= 2 ? dim - 2 : 0);
    const size_t r1 = r0 + 1, c1 = c0 + 1;
    const double dr = row_f - (double)r0, dc = col_f - (double)c0;
    const float v00 = d[r0 * dim + c0], v01 = d[r0 * dim + c1], v10 = d[r1 * dim + c0], v11 = d[r1 * dim + c1];
    if (std::isnan(v00) || std::isnan(v01) || std::isnan(v10) || std::isnan(v11)) return false;
    const double v = (double)v00 * (1.0 - dr) * (1.0 - dc) + (double)v01 * (1.0 - dr) * dc +
                     (double)v10 * dr * (1.0 - dc) + (double)v11 * dr * dc;
    *out = (float)v;
    return true;
}

}  // namespace

extern "C" {

// SrtmTile::load for each path then DemMosaic::elevation_at_linear per
// point: ok[i] = 1 with out[i] the height (NaN for a void), 0 = OutOfBounds.
int oracle_srtm_sample(const char* const* paths, int32_t n, const double* lat, const double* lon, int64_t m,
                       float* out, int32_t* ok) {
    std::vector<OracleSrtm> tiles;
    for (int32_t i = 0; i < n; ++i) {
        std::string name = paths[i];
        const size_t sl = name.find_last_of('/');
        if (sl != std::string::npos) name = name.substr(sl + 1);
        const size_t dot = name.find_last_of('.');
        if (dot != std::string::npos && dot > 0) name = name.substr(0, dot);
        if (name.size() < 7) return SARDINE_ERR_DEM_UNRECOGNISED_TILE_NAME;
        auto field = [](const std::string& f, int* v) {
            char* end = nullptr;
            if (f.empty() || std::isspace((unsigned char)f[0])) return false;
            const long x = std::strtol(f.c_str(), &end, 10);
            if (*end || (f.size() == 1 && (f[0] == '+' || f[0] == '-'))) return false;
            *v = (int)x;
            return true;
        };
        int la, lo, ls, ms;
        if (name[0] == 'N' || name[0] == 'n') ls = 1;
        else if (name[0] == 'S' || name[0] == 's') ls = -1;
        else return SARDINE_ERR_DEM_UNRECOGNISED_TILE_NAME;
        if (!field(name.substr(1, 2), &la)) return SARDINE_ERR_DEM_UNRECOGNISED_TILE_NAME;
        if (name[3] == 'E' || name[3] == 'e') ms = 1;
        else if (name[3] == 'W' || name[3] == 'w') ms = -1;
        else return SARDINE_ERR_DEM_UNRECOGNISED_TILE_NAME;
        if (!field(name.substr(4, 3), &lo)) return SARDINE_ERR_DEM_UNRECOGNISED_TILE_NAME;
        std::FILE* f = std::fopen(paths[i], "rb");
        if (!f) return SARDINE_ERR_DEM_IO;
        // the whole file (std::fs::read), at most one byte past the expected size
        std::vector<unsigned char> raw(3601u * 3601u * 2u + 1u);
        raw.resize(std::fread(raw.data(), 1, raw.size(), f));
        std::fclose(f);
        if (raw.size() != 3601u * 3601u * 2u) return SARDINE_ERR_DEM_BAD_SIZE;
        OracleSrtm t{la * ls, lo * ms, std::vector<float>(3601u * 3601u)};
        for (size_t k = 0; k < t.data.size(); ++k) {
            const int16_t v = (int16_t)(uint16_t)(raw[2 * k] << 8 | raw[2 * k + 1]);
            t.data[k] = v == -32768 ? std::numeric_limits<float>::quiet_NaN() : (float)v;
        }
        tiles.push_back(std::move(t));
    }
    if (tiles.empty()) return SARDINE_ERR_DEM_NO_TILES;
    for (int64_t i = 0; i < m; ++i) {
        const double lo = lon[i] >= 180.0 ? lon[i] - 360.0 : lon[i];
        ok[i] = 0;
        for (const OracleSrtm& t : tiles) {
            const double fl = lat[i] - (double)t.lat_sw, fo = lo - (double)t.lon_sw;
            if (fl < 0.0 || fl > 1.0 || fo < 0.0 || fo > 1.0) continue;
            float v;
            // a NaN tap makes the tile answer None and the scan go on
            if (grid_bilinear(t.data, 3601, (1.0 - fl) * 3600.0, fo * 3600.0, 3600, &v)) {
                out[i] = v;
                ok[i] = 1;
                break;
            }
        }
        if (!ok[i]) out[i] = std::numeric_limits<float>::quiet_NaN();
    }
    return 0;
}

}  // extern "C"

// ── MultiSlcReader (multi_slc_reader.rs:88-193) ─────────────────────────────
extern "C" int oracle_slc_multi_read(const char* const* paths, const sardine_slc_tiff* t, int32_t n, int64_t first,
                                     int64_t count, int16_t* dst) {
    if (n <= 0) return SARDINE_ERR_SLC_MULTI_EMPTY_INPUT;
    for (int32_t i = 1; i < n; ++i)
        if (t[i].width != t[0].width) return SARDINE_ERR_SLC_MULTI_WIDTH_MISMATCH;
    std::vector<int64_t> offsets;
    int64_t cum = 0;
    for (int32_t i = 0; i < n; ++i) {
        offsets.push_back(cum);
        cum += t[i].height;
    }
    if (first < 0 || count < 0 || first + count > cum) return SARDINE_ERR_SLC_LINE_OUT_OF_BOUNDS;
    // partition_point(off <= first) - 1, saturating
    int64_t k = 0;
    while (k < n && offsets[(size_t)k] <= first) ++k;
    k = k > 0 ? k - 1 : 0;
    const int64_t end = k + 1 < n ? offsets[(size_t)k + 1] : cum;
    if (first + count > end) return SARDINE_ERR_SLC_CROSS_SLICE_BOUNDARY;
    return oracle_slc_read(paths[k], &t[k], first - offsets[(size_t)k], count, dst);
}

// The portable libm and the IEEE division / square root on the host, element
// by element (fn as sardine_diag_math): the host side of the bit comparison.
extern "C" void oracle_pmath(int fn, const double* x, const double* y, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const double a = x[i], b = y[i];
        double r;
        switch (fn) {
            case 0: r = pm::sin(a); break;
            case 1: r = pm::cos(a); break;
            case 2: r = pm::exp(a); break;
            case 3: r = pm::log(a); break;
            case 4: r = pm::sinh(a); break;
            case 5: r = pm::cosh(a); break;
            case 6: r = pm::atanh(a); break;
            case 7: r = pm::atan(a); break;
            case 8: r = pm::atan2(b, a); break;
            case 9: r = a / b; break;
            default: r = std::sqrt(a); break;
        }
        out[i] = r;
    }
}
