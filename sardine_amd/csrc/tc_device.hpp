// tc_device.hpp — device pieces of terrain correction shared by the
// per-call kernels (tc.hip) and the batched kernels (tc_batch.hip): per-pixel
// preparation, Newton solve, resampling and geometry, output staging, the
// per-row seed and anchor computations, and the chain bodies of the segment
// and fix-up passes.  Reference: sardine/src/terrain_correction.rs:1245-1836.
#pragma once

#include "geom.hpp"
#include "tc_params.hpp"
#include "tc_pixel.hpp"

namespace sd {

// Outcome categories of one pixel (the counters of GeocodedImage).
enum : int { CAT_VALID = 0, CAT_DEM = 1, CAT_OUT = 2, CAT_NC = 3, CAT_DEG = 4, CAT_FLAT = 5, CAT_NOISE = 6,
             CAT_ORBIT = 7 };

struct LaneCounts {
    int c[7];
    int h1, h2, h3;  // histogram bins 1..3 (the rest go straight to global counters)
    int err_outside, err_short;
};

__device__ __forceinline__ void count(LaneCounts& L, int cat, int it, int sign, unsigned long long* g) {
    switch (cat) {
        case CAT_VALID: L.c[0] += sign; break;
        case CAT_DEM: L.c[1] += sign; break;
        case CAT_OUT: L.c[2] += sign; break;
        case CAT_NC: L.c[3] += sign; break;
        case CAT_DEG: L.c[4] += sign; break;
        case CAT_FLAT: L.c[5] += sign; break;
        case CAT_NOISE: L.c[6] += sign; break;
        default: break;
    }
    if (cat == CAT_VALID || cat == CAT_OUT || cat == CAT_FLAT || cat == CAT_NOISE) {
        if (it == 1) L.h1 += sign;
        else if (it == 2) L.h2 += sign;
        else if (it == 3) L.h3 += sign;
        else atomicAdd(&g[7 + (it < kHistSize - 1 ? it : kHistSize - 1)], (unsigned long long)(long long)sign);
    }
}

struct PixelPos {
    double lat, lon;
    double sl, cl, so, co;  // sines/cosines of lat/lon (radians of the degrees)
    double n;               // prime-vertical radius at lat
    d3 target;
    bool dem_ok;
    NbHeights nb;  // the terrain geometry's neighbour heights (when p.need_geom)
};

// The project_row anchors of one anchor interval.  The interval index
// depends on the column only: in the chain pass, where the lanes of a wave
// walk the same columns of 64 rows, a walk reloads the anchors once per
// interval (anchor_dx / spacing columns) instead of per pixel.  (Passes whose
// lanes sit at different columns do not use the cache.)
struct AnchorCache {
    int a;
    double lon0, lon1, lat0, lat1;
};
__device__ __forceinline__ AnchorCache anchor_cache_empty() { return AnchorCache{-1, 0.0, 0.0, 0.0, 0.0}; }

// The project_row blend of grid column `col` (output_crs.rs:395-414): the
// anchor interval holding x (relative to the first stored anchor) and the
// weight of its right anchor.
__device__ __forceinline__ void col_blend(const TcParams& p, int64_t col, int* a_out, double* t_out) {
    const double x = p.x_start + (double)col * p.spacing;
    const int na = p.n_anchor_pts - 1;
    int seg = (int)floor((x - p.x_start) * p.inv_anchor_dx);  // a guess: the walk below decides
    seg = seg < 0 ? 0 : (seg > na - 1 ? na - 1 : seg);
    while (seg > 0 && !(x > p.x_start + (double)seg * p.anchor_dx)) --seg;
    while (seg + 1 < na && x > p.x_start + (double)(seg + 1) * p.anchor_dx) ++seg;
    const double x0 = p.x_start + (double)seg * p.anchor_dx, x1 = p.x_start + (double)(seg + 1) * p.anchor_dx;
    *a_out = seg - p.anchor0;  // the row stores anchors anchor0.. only
    *t_out = (x1 == x0) ? 0.0 : (x - x0) / (x1 - x0);
}

// Output-grid position of a column and its DEM/ECEF target.  The centre's
// and (for the terrain geometry) the neighbours' DEM taps are issued
// together: one memory round trip instead of two; only the four neighbour
// heights stay live across the Newton solve.
template <bool UTM, bool NB = true>
__device__ __forceinline__ PixelPos prep_pixel(const TcParams& p, int64_t col, double row_y, const double* alon,
                                               const double* alat, AnchorCache* ac = nullptr) {
    PixelPos q;
    if constexpr (!UTM) {
        q.lon = p.lon_west + ((double)col + 0.5) * p.spacing;
        q.lat = row_y;
    } else {  // Projector::project_row blend (output_crs.rs:395-414)
        int a;
        double t;
        if (p.col_blend) {
            const ColBlend cb = p.col_blend[col - p.col0];
            a = cb.a;
            t = cb.t;
        } else {
            col_blend(p, col, &a, &t);
        }
        double lon0, lon1, lat0, lat1;
        if (ac && ac->a == a) {
            lon0 = ac->lon0;
            lon1 = ac->lon1;
            lat0 = ac->lat0;
            lat1 = ac->lat1;
        } else {
            lon0 = alon[a];  // the four loads in flight together
            lon1 = alon[a + 1];
            lat0 = alat[a];
            lat1 = alat[a + 1];
            if (ac) *ac = AnchorCache{a, lon0, lon1, lat0, lat1};
        }
        q.lon = lon0 + t * (lon1 - lon0);
        q.lat = lat0 + t * (lat1 - lat0);
    }
    float hf;
    if (p.dem_tiled) {  // GLO-30 mosaic: one tile lookup per point (dem.rs:730-747)
        const MosArgs m = mos_args(p);
        hf = dem_height_mosaic(m, q.lat, q.lon);
        if (NB && p.need_geom) {
            const double st = p.step_deg;
            q.nb.e = dem_height_mosaic(m, q.lat, q.lon + st);
            q.nb.w = dem_height_mosaic(m, q.lat, q.lon - st);
            q.nb.n = dem_height_mosaic(m, q.lat + st, q.lon);
            q.nb.s = dem_height_mosaic(m, q.lat - st, q.lon);
            q.nb.ok = isfinite(q.nb.e) && isfinite(q.nb.w) && isfinite(q.nb.n) && isfinite(q.nb.s);
        }
    } else {
        const DemAxis rc = dem_row_axis(p, q.lat), cc = dem_col_axis(p, q.lon);
        const DemTaps ct = dem_load(p, rc, cc);
        if (NB && p.need_geom) q.nb = nb_heights(p, q.lat, q.lon, rc, cc);
        hf = dem_interp(rc, cc, ct);
    }
    q.dem_ok = isfinite(hf);
    q.target = {0.0, 0.0, 0.0};
    q.sl = q.cl = q.so = q.co = q.n = 0.0;
    if (q.dem_ok) {
        sincos_deg(q.lat, &q.sl, &q.cl);
        sincos_deg(q.lon, &q.so, &q.co);
        q.n = prime_vertical(q.sl);
        q.target = ecef_from_n(q.n, q.sl, q.cl, q.so, q.co, (double)hf + geoid_n(p, q.lat, q.lon));
    }
    return q;
}

// ── EPSG:4326 grids: row and column terms computed once ─────────────────
// On a geographic grid a pixel's latitude is its row's and its longitude
// its column's, so the sine/cosine and prime-vertical radius of the
// latitude and the DEM row axes (centre, north and south neighbours) are
// the same along a lane's walk (RowGeo, per lane), and the longitude's
// sine/cosine and DEM column axes the same down a column (ColGeo, a table
// filled by wgs_col_kernel).  The values are those prep_pixel computes per
// pixel, by the same arithmetic.
struct RowGeo {
    double lat, sl, cl, n;
    DemAxis rc, rn, rs;
};

__device__ __forceinline__ RowGeo row_geo(const TcParams& p, double lat) {
    RowGeo g;
    g.lat = lat;
    sincos_deg(lat, &g.sl, &g.cl);
    g.n = prime_vertical(g.sl);
    g.rc = dem_row_axis(p, lat);
    g.rn = g.rs = g.rc;
    if (p.need_geom) {
        g.rn = dem_row_axis(p, lat + p.step_deg);
        g.rs = dem_row_axis(p, lat - p.step_deg);
    }
    return g;
}

__device__ __forceinline__ ColGeo col_geo(const TcParams& p, int64_t col) {
    const double lon = p.lon_west + ((double)col + 0.5) * p.spacing;
    ColGeo c;
    sincos_deg(lon, &c.so, &c.co);
    const DemAxis cc = dem_col_axis(p, lon);
    DemAxis ce = cc, cw = cc;
    if (p.need_geom) {
        ce = dem_col_axis(p, lon + p.step_deg);
        cw = dem_col_axis(p, lon - p.step_deg);
    }
    c.ic = cc.i0;
    c.fc = cc.f;
    c.ie = ce.i0;
    c.fe = ce.f;
    c.iw = cw.i0;
    c.fw = cw.f;
    c.pad = 0;
    return c;
}

// prep_pixel for a raster DEM on an EPSG:4326 grid from the row's and the
// column's terms.
template <bool NB = true>
__device__ __forceinline__ PixelPos prep_pixel_wgs(const TcParams& p, const RowGeo& g, const ColGeo& cg,
                                                   int64_t col) {
    PixelPos q;
    q.lat = g.lat;
    q.lon = p.lon_west + ((double)col + 0.5) * p.spacing;
    float hf;
    if (p.dem_tiled) {  // GLO-30 mosaic: one tile lookup per point (dem.rs:730-747)
        const MosArgs m = mos_args(p);
        hf = dem_height_mosaic(m, q.lat, q.lon);
        if (NB && p.need_geom) {
            const double st = p.step_deg;
            q.nb.e = dem_height_mosaic(m, q.lat, q.lon + st);
            q.nb.w = dem_height_mosaic(m, q.lat, q.lon - st);
            q.nb.n = dem_height_mosaic(m, q.lat + st, q.lon);
            q.nb.s = dem_height_mosaic(m, q.lat - st, q.lon);
            q.nb.ok = isfinite(q.nb.e) && isfinite(q.nb.w) && isfinite(q.nb.n) && isfinite(q.nb.s);
        }
    } else {
        const DemAxis cc{cg.ic, cg.fc};
        const DemTaps ct = dem_load(p, g.rc, cc);
        if (NB && p.need_geom) {
            const DemAxis ce{cg.ie, cg.fe}, cw{cg.iw, cg.fw};
            const DemTaps te = dem_load(p, g.rc, ce), tw = dem_load(p, g.rc, cw);
            const DemTaps tn = dem_load(p, g.rn, cc), ts = dem_load(p, g.rs, cc);
            q.nb.e = dem_interp(g.rc, ce, te);
            q.nb.w = dem_interp(g.rc, cw, tw);
            q.nb.n = dem_interp(g.rn, cc, tn);
            q.nb.s = dem_interp(g.rs, cc, ts);
            q.nb.ok = isfinite(q.nb.e) && isfinite(q.nb.w) && isfinite(q.nb.n) && isfinite(q.nb.s);
        }
        hf = dem_interp(g.rc, cc, ct);
    }
    q.dem_ok = isfinite(hf);
    q.sl = g.sl;
    q.cl = g.cl;
    q.so = cg.so;
    q.co = cg.co;
    q.n = g.n;
    q.target = {0.0, 0.0, 0.0};
    if (q.dem_ok) q.target = ecef_from_n(g.n, g.sl, g.cl, cg.so, cg.co, (double)hf + geoid_n(p, q.lat, q.lon));
    return q;
}

struct PixelOut {
    float sigma0, cos_lia;
    double line, sample;
    int cat;
    unsigned char mask;
};

// Everything after the Newton solve (terrain_correction.rs:1558-1739).
template <int RS>
__device__ __forceinline__ PixelOut finish_pixel(const TcParams& p, const PixelPos& q, int oc, int64_t t, d3 sat,
                                                 double inv_flat_area) {
    PixelOut o;
    o.sigma0 = __builtin_nanf("");
    o.cos_lia = __builtin_nanf("");
    o.line = __builtin_nan("");
    o.sample = __builtin_nan("");
    o.mask = 1;  // OUTSIDE_FOOTPRINT until decided
    if (oc == ZD_ORBIT_OUTSIDE || oc == ZD_ORBIT_INSUFFICIENT) {
        o.cat = CAT_ORBIT;
        return o;
    }
    if (oc == ZD_NOT_CONVERGED) {
        o.cat = CAT_NC;
        o.mask = 3;
        return o;
    }
    if (oc == ZD_DEGENERATE) {
        o.cat = CAT_DEG;
        o.mask = 4;
        return o;
    }
    d3 rv = q.target - sat;
    const double slant = sqrt(dot(rv, rv));
    // to_image_coords (terrain_correction.rs:742-744): µs / 1e6 / ati and
    // (R - R0) / rps, each division correctly rounded as in the reference
    const double line = div_rn(div_rn((double)(t - p.az_start_us), 1e6, 1e-6), p.ati, p.inv_ati);
    const double sample = div_rn(slant - p.near_range_m, p.rps, p.inv_rps);
    o.line = line;
    o.sample = sample;
    if (line < -0.5 || line > (double)p.lines - 0.5 || sample < -0.5 || sample > (double)p.samples - 0.5) {
        o.cat = CAT_OUT;
        return o;
    }
    // the NESZ lookup's taps (same position) are loaded with the image's;
    // RS >= 3 is the mixed form (FP32 resampling and terrain geometry)
    constexpr bool MIX = RS >= 3;
    const Bilin nb_idx = bilin_index(p.lines, p.samples, line, sample);
    float sigma0;
    DemTaps nzt = {0.0f, 0.0f, 0.0f, 0.0f};
    // The terrain geometry does not depend on the image: it is computed
    // while the bilinear taps are in flight.
    DemTaps it = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (RS == 0 || RS == 3) {
        it = bilin_load(p.img, p.samples, nb_idx);
    }
    if (p.nesz) nzt = bilin_load(p.nesz, p.samples, nb_idx);
    double w = 0.0, cl = 0.0;
    int g = TG_SHADOW;
    if (p.need_geom && q.nb.ok) {
        if constexpr (MIX) {
            float wf = 0.0f, clf = 0.0f;
            g = terrain_geometry_mixed(q.sl, q.cl, q.so, q.co, q.n, p.step_sin, p.step_cos, q.nb.e, q.nb.w,
                                       q.nb.n, q.nb.s, sat, q.target, inv_flat_area, &wf, &clf);
            w = wf;
            cl = clf;
        } else {
            g = terrain_geometry_sc(q.sl, q.cl, q.so, q.co, q.n, p.step_sin, p.step_cos, q.nb.e, q.nb.w, q.nb.n,
                                    q.nb.s, sat, q.target, inv_flat_area, &w, &cl);
        }
    }
    if constexpr (RS == 0 || RS == 3) {
        sigma0 = MIX ? bilin_interp_mx(nb_idx, it) : bilin_interp(nb_idx, it);
    } else {
        sigma0 = resample<RS>(p, line, sample);
    }
    if (isnan(sigma0)) {
        o.cat = CAT_OUT;
        return o;
    }
    if (p.need_geom) {
        if (!q.nb.ok) {
            o.mask = 9;
            if (p.flatten) {
                o.cat = CAT_FLAT;
                return o;
            }
        } else {
            if (g == TG_VALID) {
                o.cos_lia = (float)cl;
                if (p.flatten) sigma0 /= (float)w;
            } else if (g == TG_FORESHORTENING || g == TG_LAYOVER) {
                o.cos_lia = (float)cl;
                o.mask = g == TG_LAYOVER ? 6 : 7;
                if (p.flatten) {
                    o.cat = CAT_FLAT;
                    return o;
                }
            } else {
                o.mask = 5;
                if (p.flatten) {
                    o.cat = CAT_FLAT;
                    return o;
                }
            }
        }
    }
    if (p.nesz) {
        float nz = MIX ? bilin_interp_mx(nb_idx, nzt) : bilin_interp(nb_idx, nzt);
        if (!isnan(nz) && sigma0 <= nz * p.nesz_factor) {
            o.cat = CAT_NOISE;
            o.mask = 8;
            return o;
        }
    }
    o.sigma0 = sigma0;
    o.mask = 0;
    o.cat = CAT_VALID;
    return o;
}

__device__ __forceinline__ void write_pixel(const TcParams& p, int64_t o, const PixelOut& v) {
    p.out[o] = v.sigma0;
    if (p.compute_lia) {
        p.cos_lia[o] = v.cos_lia;
        p.mask[o] = v.mask;
    }
    if (p.radar_coords) {
        p.radar_coords[2 * o] = v.line;
        p.radar_coords[2 * o + 1] = v.sample;
    }
}

__device__ __forceinline__ void write_dem_missing(const TcParams& p, int64_t o) {
    p.out[o] = __builtin_nanf("");
    if (p.compute_lia) {
        p.cos_lia[o] = __builtin_nanf("");
        p.mask[o] = 2;
    }
    if (p.radar_coords) {
        p.radar_coords[2 * o] = __builtin_nan("");
        p.radar_coords[2 * o + 1] = __builtin_nan("");
    }
}

// Newton from `seed` for one pixel (dem_ok pixels only); `tie` is set when
// a step was a tie (decided by the reference's arithmetic with EXACT, by the
// kernel's own otherwise).
struct Solve {
    int oc, it;
    int64_t t;
    d3 sat;
    bool tie;
};

template <bool EXACT = false>
__device__ __forceinline__ Solve solve(const TcParams& p, const OrbitDev& orb, const PixelPos& q, int64_t seed,
                                       int max_iter, OrbitCache& cache) {
    Solve s;
    s.t = 0;
    s.it = 0;
    s.tie = false;
    d3 sat_unused;
    s.oc = zero_doppler<EXACT>(orb, q.target, seed, max_iter, p.tol, p.tie_band, &s.t, &s.it, &sat_unused, &cache,
                               &s.tie, EXACT ? p.counters + kCntTie : nullptr);
    // a converged solve's last evaluation is at its time: the satellite
    // position is the evaluation cache's (the registers it already holds,
    // not a copy per exit of the Newton loop); unused otherwise
    s.sat = cache.P;
    return s;
}

__device__ __forceinline__ void zero_counts(LaneCounts& C) {
#pragma unroll
    for (int i = 0; i < 7; ++i) C.c[i] = 0;
    C.h1 = C.h2 = C.h3 = 0;
    C.err_outside = C.err_short = 0;
}

// Records how far an iterate outside the orbit window reached beyond it
// (OutsideOrbitWindow's `requested`, orbit.rs:513-517).
__device__ __forceinline__ void note_orbit_outside(const TcParams& p, const OrbitDev& o, int64_t t) {
    if (t < o.t_us[0])
        atomicMax(&p.counters[kCntOrbitBelow], (unsigned long long)(o.t_us[0] - t));
    else
        atomicMax(&p.counters[kCntOrbitAbove], (unsigned long long)(t - o.t_us[o.n - 1]));
}

// Wave-reduce each lane's counts and add them to the global counters.
__device__ __forceinline__ void flush_counts(const TcParams& p, const LaneCounts& C, bool set_errors) {
    unsigned long long* G = p.counters;
    const int lane = threadIdx.x & 63;
    int v[10] = {C.c[0], C.c[1], C.c[2], C.c[3], C.c[4], C.c[5], C.c[6], C.h1, C.h2, C.h3};
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        long long x = v[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        if (lane == 0 && x) atomicAdd(&G[i < 7 ? i : 7 + (i - 6)], (unsigned long long)x);
    }
    if (set_errors && C.err_outside > 0) atomicOr(p.error_flag, 1);
    if (set_errors && C.err_short > 0) atomicOr(p.error_flag, 2);
}

// The chain pass's per-lane counts, one byte each (a lane counts at most
// the pixels of one segment, <= 64): bytes 0..6 the categories CAT_VALID ..
// CAT_NOISE, 7..9 histogram bins 1..3, 10 and 11 the orbit errors
// (outside the window, too few vectors).  Three registers instead of twelve
// live across the walk.
struct PackedCounts {
    uint32_t w[3];
};
constexpr int kPcErrOutside = 10, kPcErrShort = 11;

__device__ __forceinline__ void pc_add(PackedCounts& P, int byte) {
    const uint32_t one = 1u << (8 * (byte & 3));
    P.w[0] += byte < 4 ? one : 0u;
    P.w[1] += (byte >> 2) == 1 ? one : 0u;
    P.w[2] += byte >= 8 ? one : 0u;
}

// count() for a pixel of the chain pass (sign +1); histogram bins past 3 go
// straight to the global counters as there.
__device__ __forceinline__ void pc_count(PackedCounts& P, int cat, int it, unsigned long long* g) {
    if (cat <= CAT_NOISE) pc_add(P, cat);
    if (cat == CAT_VALID || cat == CAT_OUT || cat == CAT_FLAT || cat == CAT_NOISE) {
        if (it >= 1 && it <= 3) pc_add(P, 6 + it);
        else atomicAdd(&g[7 + (it < kHistSize - 1 ? it : kHistSize - 1)], 1ull);
    }
}

__device__ __forceinline__ int pc_get(const PackedCounts& P, int byte) {
    return (int)((P.w[byte >> 2] >> (8 * (byte & 3))) & 0xffu);
}

__device__ __forceinline__ LaneCounts pc_unpack(const PackedCounts& P) {
    LaneCounts C;
#pragma unroll
    for (int i = 0; i < 7; ++i) C.c[i] = pc_get(P, i);
    C.h1 = pc_get(P, 7);
    C.h2 = pc_get(P, 8);
    C.h3 = pc_get(P, 9);
    C.err_outside = pc_get(P, kPcErrOutside);
    C.err_short = pc_get(P, kPcErrShort);
    return C;
}

// Copies the orbit into LDS up to its last interval polynomial.
__device__ __forceinline__ void load_orbit(const TcParams& p, OrbitDev& orb) {
    const uint4* src = reinterpret_cast<const uint4*>(p.orbit);
    uint4* dst = reinterpret_cast<uint4*>(&orb);
    const int n16 = (int)(orbit_lds_bytes(p.orbit->n) / sizeof(uint4));
    for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

// guess_azimuth_time (terrain_correction.rs:1107-1136): linear
// interpolation in the latitude-sorted grid-point table, clamped at both
// ends; an exact latitude hit returns that entry.
__device__ __forceinline__ int64_t guess_azimuth_time(const RowPrepParams& q, double lat) {
    const int n = q.n_lut;
    if (n == 1) return q.lut_t[0];
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        const double v = q.lut_lat[mid];
        if (v < lat) lo = mid + 1;
        else if (v > lat) hi = mid;
        else return q.lut_t[mid];
    }
    if (lo == 0) return q.lut_t[0];
    if (lo >= n) return q.lut_t[n - 1];
    const double frac = (lat - q.lut_lat[lo - 1]) / (q.lut_lat[lo] - q.lut_lat[lo - 1]);
    const double dt_us = (double)(q.lut_t[lo] - q.lut_t[lo - 1]);
    return q.lut_t[lo - 1] + (int64_t)(frac * dt_us);
}

// One lane per AOI row: the row's Newton seed (the grid-point guess at the
// row's seed latitude) and, for the terrain geometry, the flat-cell area at
// (seed latitude, centre longitude) (terrain_correction.rs:1424-1511).
__device__ __forceinline__ void row_prep(const RowPrepParams& q, int64_t rr) {
    const int64_t row = q.row0 + rr;
    const double row_y = q.y_top - ((double)row + 0.5) * q.spacing;
    double lat_seed = row_y;
    double ref_lon = q.x_left + ((double)q.full_cols / 2.0) * q.spacing;
    if (q.utm) {
        const TmSeries ts = tm_series();
        double lon_mid;
        utm_inverse(ts, q.zone, q.south != 0, q.x_left + ((double)q.full_cols / 2.0) * q.spacing, row_y, &lon_mid,
                    &lat_seed);
        if (q.need_geom) {  // row_lonlat[n_cols/2].lon from the anchor blend (output_crs.rs:395-414)
            const int64_t col = q.full_cols / 2;
            const double x = q.x_start + (double)col * q.spacing;
            int seg = 0;
            while (seg + 1 < q.n_anch && x > q.x_start + (double)(seg + 1) * q.anchor_dx) ++seg;
            const double xa = q.x_start + (double)seg * q.anchor_dx, xb = q.x_start + (double)(seg + 1) * q.anchor_dx;
            double la0, lo0, la1, lo1;
            utm_inverse(ts, q.zone, q.south != 0, xa, row_y, &lo0, &la0);
            utm_inverse(ts, q.zone, q.south != 0, xb, row_y, &lo1, &la1);
            const double t = xb == xa ? 0.0 : (x - xa) / (xb - xa);
            ref_lon = lo0 + t * (lo1 - lo0);
        }
    }
    q.row_seed[rr] = guess_azimuth_time(q, lat_seed) - q.epoch;
    if (q.need_geom) q.row_flat_area[rr] = flat_cell_area(lat_seed, ref_lon, q.step_deg);
}

// Project_row anchors of every AOI row (output_crs.rs:364-393):
// anchors[row][0..stride) = lon, anchors[row][stride..2 stride) = lat of
// anchor points anchor0 .. anchor0 + stride - 1 (the ones the AOI's columns
// blend between); i indexes (row, stored point).
__device__ __forceinline__ void anchor_point(const TcParams& p, int64_t i) {
    const int npts = p.anchor_stride;
    const int64_t rr = i / npts;
    const int k = (int)(i - rr * npts);
    const double y = p.lat_north - ((double)(p.row0 + rr) + 0.5) * p.spacing;
    const TmSeries ts = tm_series();
    double lo, la;
    utm_inverse(ts, p.utm_zone, p.utm_south != 0, p.x_start + (double)(p.anchor0 + k) * p.anchor_dx, y, &lo, &la);
    p.anchors[rr * 2 * npts + k] = lo;
    p.anchors[rr * 2 * npts + npts + k] = la;
}

// Output staging.  Lane = row means a wave's per-pixel stores land on 64
// different rows; stored one pixel at a time they leave 64 partly written
// cache lines per array in L2 and each is written back many times.  The
// segment kernel instead stages each wave's pixels in LDS and writes them out
// every kStageCols columns (f32 arrays) or kMaskCols columns (the byte mask)
// as contiguous row pieces.
constexpr int kStageCols = 8;
constexpr int kMaskCols = 32;
struct WaveStage {
    float d[64][kStageCols + 1];  // +1 word: row-to-row bank skew
    float c[64][kStageCols + 1];
    unsigned char m[64][kMaskCols + 4];
    // each lane's tie bookkeeping (rarely written, so kept out of registers):
    // the seed entering its first tie pixel, the first and last tie columns,
    // and flags (bit 0: a tie was met; bit 1: a pixel converged before it)
    long long tie_seed[64];
    int tie_first[64], tie_last[64], tie_flags[64];
};

// Orders the LDS accesses of one wave's lanes (lanes of a wave execute in
// lockstep; this keeps the compiler from moving accesses across it).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Writes the n staged f32 columns starting at output column cs of rows
// rr0..rr0+63 (rows past p.rows are skipped).
__device__ __forceinline__ void flush_f32(const TcParams& p, const WaveStage& st, int64_t rr0, int64_t cs, int n) {
    const int lane = threadIdx.x & 63;
    wave_sync();
#pragma unroll 1
    for (int i = 0; i < kStageCols; ++i) {
        const int e = lane + 64 * i, r = e / kStageCols, c = e % kStageCols;
        if (c < n && rr0 + r < p.rows) {
            const int64_t o = (rr0 + r) * p.cols + cs + c;
            p.out[o] = st.d[r][c];
            if (p.compute_lia) p.cos_lia[o] = st.c[r][c];
        }
    }
    wave_sync();
}

__device__ __forceinline__ void flush_mask(const TcParams& p, const WaveStage& st, int64_t rr0, int64_t cs, int n) {
    const int lane = threadIdx.x & 63;
    wave_sync();
#pragma unroll 1
    for (int i = 0; i < kMaskCols; ++i) {
        const int e = lane + 64 * i, r = e / kMaskCols, c = e % kMaskCols;
        if (c < n && rr0 + r < p.rows) p.mask[(rr0 + r) * p.cols + cs + c] = st.m[r][c];
    }
    wave_sync();
}

// Occupancy of the chain kernels (waves per SIMD the register allocator
// must allow), measured per form on the full scene.  At three: every
// per-call chain but the FP64 Lanczos-3 ones (155-168 VGPRs with the
// parameters reloaded in the pixel loop; UTM bilinear 18.1 against 20.9 ms
// at two, mixed Lanczos-3 27.5 against 31.4; EPSG:4326 bilinear 24.7 ->
// 20.0 ms, mixed 24.3 -> 21.2, mixed Lanczos-3 38.0 -> 35.7 on an
// 18001 x 36067 grid).  The orbit's LDS copy is sized to the orbit
// (sd_launch_tc_fp64), which leaves room for three workgroups per CU.  At
// two: the FP64 Lanczos-3 chains (214-219 VGPRs; they spill at three: UTM
// 36.0 -> 44.1 ms).  The batched chain kernels (one per job class) follow
// the same table.
template <int RS, bool UTM>
constexpr int chain_waves() {
    return RS == 2 ? 2 : 3;
}

// Lane (row, segment): wave = 64 adjacent rows of one segment; the 4 waves of
// a workgroup take 4 consecutive segments.  Lanes past the last row compute
// the last row again and discard it, so every lane reaches the wave-wide
// flushes.
// COLS: EPSG:4326 chains take their row and column terms from RowGeo and
// the column table (the batched kernel, which holds every job class in one
// kernel, keeps the per-pixel form: the row terms would raise its spills).
// The chain's parameters read again where the pixel loop uses them: the
// empty asm hides their address, so the loads cannot be hoisted out of the
// loop.  Each use becomes a scalar load (scalar memory, not VALU) instead of
// a uniform value held across the loop, which at the register limit spills
// to a VGPR lane and costs a VALU readlane per use.  RELOAD 0: p is used as
// it is; 1: p is the kernel's only explicit argument, which the
// kernel-argument segment holds at offset 0; 2: p lies in device memory
// that no thread writes during the kernel, at a wave-uniform address (a
// batch's job table).
template <int RELOAD>
__device__ __forceinline__ const TcParams& params_reload(const TcParams& p) {
    using KPtr = const __attribute__((address_space(4))) TcParams*;
    if constexpr (RELOAD == 0) {
        return p;
    } else {
        KPtr k = RELOAD == 1 ? (KPtr)__builtin_amdgcn_kernarg_segment_ptr() : (KPtr)&p;
        asm volatile("" : "+s"(k));
        return *(const TcParams*)k;
    }
}

template <int RS, bool UTM, bool COLS = true, int RELOAD = 0>
__device__ __forceinline__ void segment_wave(const TcParams& p0, const OrbitDev& orb, WaveStage& st, int64_t rr0,
                                             int64_t seg) {
    const TcParams& p = p0;
    const int lane = threadIdx.x & 63;
    const bool row_ok = rr0 + lane < p.rows;
    const int64_t rr = row_ok ? rr0 + lane : p.rows - 1;
    const int64_t row = p.row0 + rr;
    const double row_y = p.lat_north - ((double)row + 0.5) * p.spacing;
    const int npts = p.anchor_stride;
    const double* alon = UTM ? p.anchors + rr * 2 * npts : nullptr;
    const double* alat = UTM ? alon + npts : nullptr;
    const int64_t c_begin = seg * p.seg_len;
    const int64_t c_end = min(c_begin + p.seg_len, p.cols);
    const long long row_seed = p.row_seed[rr];
    const double inv_flat_area = p.need_geom ? 1.0 / p.row_flat_area[rr] : 0.0;
    const int64_t obase = rr * p.cols;

    // phase A: the chain's value at the end of the previous segment, from a
    // short chain over the p.lead columns before this segment started at the
    // row seed.  Its first solve (from the far row seed) may settle on either
    // side of its root; the following ones approach their roots from the
    // previous pixel's side, as the reference's chain does, and so usually
    // reproduce its value exactly.
    long long S = row_seed;
    OrbitCache cache = orbit_cache_empty();
    if (seg > 0) {
        bool have = false;
        for (int64_t cc = c_begin > p.lead ? c_begin - p.lead : 0; cc < c_begin; ++cc) {
            const TcParams& p = params_reload<RELOAD>(p0);
            PixelPos q = prep_pixel<UTM, false>(p, p.col0 + cc, row_y, alon, alat);
            if (!q.dem_ok) continue;
            Solve s = solve(p, orb, q, S, have ? p.max_iter : 40, cache);
            if (s.oc == ZD_CONVERGED) {
                S = s.t;
                have = true;
            }
        }
    }

    // phase B: the reference's sequential chain over the segment
    PackedCounts C = {{0u, 0u, 0u}};
    long long seed = S, last_t = 0;
    bool last_has = false, any_newton = false;
    st.tie_flags[lane] = 0;
    AnchorCache ac = anchor_cache_empty();
    // EPSG:4326 bilinear: the row's terms once, the column's from the table
    // (every EPSG:4326 job has one).  Not for bicubic and Lanczos-3: with the
    // row terms held across the walk their chains spill at the occupancy
    // they run at and gain nothing (FP64 bicubic 25.5 -> 25.3 ms, Lanczos-3
    // 43.6 -> 44.3, mixed Lanczos-3 35.7 -> 37.1 on an 18001 x 36067 grid).
    constexpr bool WGS = COLS && !UTM && RS % 3 == 0;
    RowGeo rg;
    if constexpr (WGS) rg = row_geo(p, row_y);
    for (int64_t cc = c_begin; cc < c_end; ++cc) {
        const TcParams& p = params_reload<RELOAD>(p0);
        const int j = (int)(cc - c_begin);
        PixelPos q;
        if constexpr (WGS) q = prep_pixel_wgs(p, rg, p.col_geo[cc], p.col0 + cc);
        else q = prep_pixel<UTM>(p, p.col0 + cc, row_y, alon, alat, &ac);
        float v_s0, v_cos;
        unsigned char v_mask;
        if (!q.dem_ok) {  // DEM_MISSING (terrain_correction.rs:1529-1536)
            v_s0 = v_cos = __builtin_nanf("");
            v_mask = 2;
            if (row_ok) {
                pc_add(C, CAT_DEM);
                if (p.radar_coords) {
                    p.radar_coords[2 * (obase + cc)] = __builtin_nan("");
                    p.radar_coords[2 * (obase + cc) + 1] = __builtin_nan("");
                }
            }
        } else {
            any_newton = true;
            Solve s = solve(p, orb, q, seed, p.max_iter, cache);
            if (s.tie) {  // the fix-up passes replay this pixel with the reference's arithmetic
                if (!st.tie_flags[lane]) {
                    st.tie_seed[lane] = seed;
                    st.tie_first[lane] = (int)cc;
                    st.tie_flags[lane] = 1 | (last_has ? 2 : 0);
                }
                st.tie_last[lane] = (int)cc;
            }
            if (s.oc == ZD_CONVERGED) {
                seed = s.t;
                last_t = s.t;
                last_has = true;
            }
            PixelOut v = finish_pixel<RS>(p, q, s.oc, s.t, s.sat, inv_flat_area);
            if (row_ok) {
                if (v.cat == CAT_ORBIT) {
                    if (s.oc == ZD_ORBIT_OUTSIDE) {
                        pc_add(C, kPcErrOutside);
                        note_orbit_outside(p, orb, s.t);
                    } else {
                        pc_add(C, kPcErrShort);
                    }
                } else {
                    pc_count(C, v.cat, s.it, p.counters);
                }
                if (p.radar_coords) {
                    p.radar_coords[2 * (obase + cc)] = v.line;
                    p.radar_coords[2 * (obase + cc) + 1] = v.sample;
                }
            }
            v_s0 = v.sigma0;
            v_cos = v.cos_lia;
            v_mask = v.mask;
        }
        st.d[lane][j % kStageCols] = v_s0;
        st.c[lane][j % kStageCols] = v_cos;
        st.m[lane][j % kMaskCols] = v_mask;
        const bool last = cc + 1 == c_end;
        if ((j + 1) % kStageCols == 0 || last) {
            const int n = j % kStageCols + 1;
            flush_f32(p, st, rr0, cc + 1 - n, n);
        }
        if (p.compute_lia && ((j + 1) % kMaskCols == 0 || last)) {
            const int n = j % kMaskCols + 1;
            flush_mask(p, st, rr0, cc + 1 - n, n);
        }
    }
    if (row_ok) {
        const int tf = st.tie_flags[lane];
        SegRecord rec;
        rec.seed = S;
        rec.last = last_t;
        rec.flags = (last_has ? 1 : 0) | (any_newton ? 2 : 0) | ((tf & 1) ? 4 : 0) | ((tf & 2) ? 8 : 0);
        rec.err_outside = pc_get(C, kPcErrOutside);
        rec.err_short = pc_get(C, kPcErrShort);
        rec.tie_first = tf ? st.tie_first[lane] : 0;
        rec.tie_seed = tf ? st.tie_seed[lane] : 0;
        rec.tie_last = tf ? st.tie_last[lane] : 0;
        rec.pad = 0;
        p.records[rr * p.nseg + seg] = rec;
    }
    flush_counts(p, pc_unpack(C), false);  // orbit errors are decided after the fix-up
}

// One wave per AOI row, before the fix-up pass: a segment that ran a Newton
// solve must have started from the chained value to its left (the last
// converged time of the nearest segment to its left that has one, or the
// row seed), and must have met no tie step.  The row's records are read 64
// at a time (coalesced); the chained value entering each lane is the `last`
// of the nearest lane below it whose segment has one (ballot + shuffle),
// else the value carried in from the previous 64.  row_flags:
//   0  nothing to correct: the row's orbit-error counts are added here;
//   2  every seed matches and only tie segments need a replay: they are
//      appended to the tie list (p.tie_list), whose replays run in parallel
//      (tie_item) unless one of them finds the chains do not meet;
//   1  the sequential walk (fixup_row) corrects the row, and adds its
//      orbit-error counts (also for tie rows with orbit errors, tie rows
//      without a list, and list overflow).
__device__ __forceinline__ void check_row(const TcParams& p, int64_t rr) {
    const int lane = threadIdx.x & 63;
    const SegRecord* recs = p.records + rr * p.nseg;
    long long carry = p.row_seed[rr];
    bool bad = false, tie = false;
    int eo = 0, es = 0;
    for (int64_t k0 = 0; k0 < p.nseg; k0 += 64) {
        const int64_t k = k0 + lane;
        SegRecord r;
        r.seed = r.last = r.tie_seed = 0;
        r.flags = r.err_outside = r.err_short = r.tie_first = r.tie_last = r.pad = 0;
        if (k < p.nseg) r = recs[k];
        const unsigned long long has = __ballot((r.flags & 1) != 0);
        const unsigned long long below = has & ((1ull << lane) - 1ull);
        const int src = below ? 63 - __clzll((long long)below) : lane;
        const long long left = __shfl(r.last, src);
        const long long E = below ? left : carry;
        if ((r.flags & 2) && r.seed != E) bad = true;
        if ((r.flags & 6) == 6) tie = true;
        eo += r.err_outside;
        es += r.err_short;
        if (has) carry = __shfl(r.last, 63 - __clzll((long long)has));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        eo += __shfl_xor(eo, o);
        es += __shfl_xor(es, o);
    }
    const bool row_tie = __any(tie);
    bool row_bad = __any(bad) || (row_tie && (eo > 0 || es > 0 || !p.tie_list));
    if (row_tie && !row_bad) {  // append the tie segments to the list
        bool overflow = false;
        for (int64_t k0 = 0; k0 < p.nseg; k0 += 64) {
            const int64_t k = k0 + lane;
            const bool mine = k < p.nseg && (recs[k].flags & 6) == 6;
            const unsigned long long m = __ballot(mine);
            if (!m) continue;
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(&p.counters[kCntTieList], (unsigned long long)__popcll(m));
            base = __shfl(base, 0);
            const unsigned long long idx = base + __popcll(m & ((1ull << lane) - 1ull));
            if (mine) {
                if (idx < (unsigned long long)p.tie_cap) p.tie_list[idx] = make_int2((int)rr, (int)k);
                else overflow = true;
            }
        }
        row_bad = __any(overflow);
    }
    if (lane == 0) {
        p.row_flags[rr] = row_bad ? 1 : (row_tie ? 2 : 0);
        if (!row_bad && eo > 0) atomicOr(p.error_flag, 1);
        if (!row_bad && es > 0) atomicOr(p.error_flag, 2);
    }
}

// Replays one segment's old chain (the chain pass's arithmetic, from `so`)
// and its corrected chain (ties decided by the reference's arithmetic, from
// `sn`) in lockstep over AOI columns [c_begin, c_end), until both converge
// to the same µs at a column >= sync_from; from there on the two are the
// same chain.  With WRITE every pixel before that point is rewritten with
// the corrected chain's outcome and C receives the counter corrections;
// without, nothing is written (a dry run that finds out whether the chains
// meet).  Returns whether they met; *new_last / *new_has are the corrected
// chain's last converged time over the replayed columns, and *changed
// whether any pixel's outcome differs between the chains.  A pixel both
// chains converge on at the same µs in the same number of iterations has
// the same outcome bits (the satellite position is the evaluation of that
// µs), so it is neither finished nor written again.
template <int RS, bool UTM, bool WRITE>
__device__ __forceinline__ bool replay_segment(const TcParams& p, const OrbitDev& orb, int64_t rr, int64_t c_begin,
                                               int64_t c_end, int64_t sync_from, long long so, long long sn,
                                               LaneCounts& C, long long* new_last, bool* new_has,
                                               bool* changed = nullptr) {
    const int64_t row = p.row0 + rr;
    const double row_y = p.lat_north - ((double)row + 0.5) * p.spacing;
    const int npts = p.anchor_stride;
    const double* alon = UTM ? p.anchors + rr * 2 * npts : nullptr;
    const double* alat = UTM ? alon + npts : nullptr;
    const double inv_flat_area = p.need_geom ? 1.0 / p.row_flat_area[rr] : 0.0;
    const int64_t obase = rr * p.cols;
    OrbitCache co = orbit_cache_empty(), cn = orbit_cache_empty();
    *new_has = false;
    *new_last = 0;
    for (int64_t cc = c_begin; cc < c_end; ++cc) {
        PixelPos q = prep_pixel<UTM>(p, p.col0 + cc, row_y, alon, alat);
        if (!q.dem_ok) continue;  // outcome independent of the seed
        Solve o = solve(p, orb, q, so, p.max_iter, co);
        // From the same seed the corrected chain takes the old chain's steps
        // exactly unless one of them is a tie step (the reference's
        // arithmetic replaces only those, and the tie test is the same in
        // both), so its own solve is needed only then.
        const Solve n = (so == sn && !o.tie) ? o : solve<true>(p, orb, q, sn, p.max_iter, cn);
        const bool same_t = o.oc == ZD_CONVERGED && n.oc == ZD_CONVERGED && o.t == n.t;
        if (changed && !(same_t && o.it == n.it)) *changed = true;
        if (same_t && (cc >= sync_from || o.it == n.it)) {
            if (WRITE && o.it != n.it) {  // same result, different iteration count
                PixelOut v = finish_pixel<RS>(p, q, o.oc, o.t, o.sat, inv_flat_area);
                count(C, v.cat, o.it, -1, p.counters);
                count(C, v.cat, n.it, 1, p.counters);
            }
            if (cc >= sync_from) return true;
            so = o.t;  // the same outcome: nothing to rewrite
            sn = n.t;
            *new_last = n.t;
            *new_has = true;
            continue;
        }
        if (WRITE) {
            PixelOut vo = finish_pixel<RS>(p, q, o.oc, o.t, o.sat, inv_flat_area);
            PixelOut vn = finish_pixel<RS>(p, q, n.oc, n.t, n.sat, inv_flat_area);
            if (vo.cat == CAT_ORBIT) {
                if (o.oc == ZD_ORBIT_OUTSIDE) C.err_outside -= 1;
                else C.err_short -= 1;
            } else {
                count(C, vo.cat, o.it, -1, p.counters);
            }
            if (vn.cat == CAT_ORBIT) {
                if (n.oc == ZD_ORBIT_OUTSIDE) {
                    C.err_outside += 1;
                    note_orbit_outside(p, orb, n.t);
                } else {
                    C.err_short += 1;
                }
            } else {
                count(C, vn.cat, n.it, 1, p.counters);
            }
            write_pixel(p, obase + cc, vn);
        }
        if (o.oc == ZD_CONVERGED) so = o.t;
        if (n.oc == ZD_CONVERGED) {
            sn = n.t;
            *new_last = n.t;
            *new_has = true;
        }
    }
    return false;
}

// Tie list entry flag (in the segment index): the dry pass found the
// corrected chain equal to the old one on every pixel.
constexpr int kTieUnchanged = 1 << 30;

// One entry of the tie list: the replay of one tie-only segment, from its
// first tie pixel.  The dry pass (WRITE false) sends the row to the
// sequential walk when the chains do not meet inside the segment (the
// segment's last converged time then changes, and the segments after it
// must be checked in order); the write pass replays the others.
template <int RS, bool UTM, bool WRITE>
__device__ __forceinline__ void tie_item(const TcParams& p, const OrbitDev& orb, int64_t i, int64_t n) {
    LaneCounts C;
    zero_counts(C);
    if (i < n) {
        const int2 e = p.tie_list[i];
        const int64_t rr = e.x, sg = e.y & ~kTieUnchanged;
        // the write pass skips the entries whose dry run found every outcome unchanged
        if (!(p.row_flags[rr] & 1) && !(WRITE && (e.y & kTieUnchanged))) {
            const SegRecord rec = p.records[rr * p.nseg + sg];
            const int64_t c_end = min((int64_t)(sg + 1) * p.seg_len, p.cols);
            long long nl;
            bool nh, changed = false;
            const bool met = replay_segment<RS, UTM, WRITE>(p, orb, rr, rec.tie_first, c_end, rec.tie_last,
                                                            rec.tie_seed, rec.tie_seed, C, &nl, &nh, &changed);
            if (!WRITE && !met) atomicOr(&p.row_flags[rr], 1);
            if (!WRITE && met && !changed) p.tie_list[i] = make_int2(e.x, e.y | kTieUnchanged);
        }
    }
    if (WRITE) flush_counts(p, C, true);
}

// One lane per row: walk the row's segments in order and correct every
// segment whose starting seed differs from the chained value to its left or
// whose chain met tie steps (rows check_row or the tie pass sent here,
// row_flags bit 0).  The old and corrected chains are replayed from the
// segment's start, or from its first tie pixel when only ties are at issue,
// until they meet past the last tie pixel.  The other lanes (and lanes whose
// row is past the end) only join the wave-wide counter reduction.
template <int RS, bool UTM>
__device__ __forceinline__ void fixup_row(const TcParams& p, const OrbitDev& orb, int64_t rr, bool valid) {
    LaneCounts C;
    zero_counts(C);
    unsigned redone = 0;
    if (valid && (p.row_flags[rr] & 1)) {
        SegRecord* recs = p.records + rr * p.nseg;
        long long E = p.row_seed[rr];
        SegRecord r1 = recs[0], r2 = p.nseg > 1 ? recs[1] : r1;  // two records in flight ahead
        for (int64_t sg = 0; sg < p.nseg; ++sg) {
            SegRecord rec = r1;
            r1 = r2;
            if (sg + 2 < p.nseg) r2 = recs[sg + 2];
            C.err_outside += rec.err_outside;
            C.err_short += rec.err_short;
            const bool ties = (rec.flags & 4) != 0;
            if ((rec.flags & 2) && (rec.seed != E || ties)) {
                ++redone;
                const int64_t c_seg = sg * p.seg_len, c_end = min(c_seg + p.seg_len, p.cols);
                const bool from_tie = rec.seed == E;  // the chain is right up to its first tie pixel
                long long new_last;
                bool new_has;
                const bool met = replay_segment<RS, UTM, true>(
                    p, orb, rr, from_tie ? (int64_t)rec.tie_first : c_seg, c_end,
                    ties ? (int64_t)rec.tie_last : c_seg, from_tie ? rec.tie_seed : rec.seed,
                    from_tie ? rec.tie_seed : E, C, &new_last, &new_has);
                if (!met) {
                    // the corrected chain's last converged time; when the
                    // replay began at the first tie pixel and nothing
                    // converged from there on, the one before that pixel
                    const bool has = new_has || (from_tie && (rec.flags & 8));
                    rec.last = new_has ? new_last : (from_tie ? rec.tie_seed : 0);
                    rec.flags = (rec.flags & ~1) | (has ? 1 : 0);
                }
            }
            if (rec.flags & 1) E = rec.last;
        }
    }
    // C holds the segment pass's orbit errors plus the replay's corrections
    flush_counts(p, C, true);
    if (redone) atomicAdd(&p.counters[kCntRedo], (unsigned long long)redone);
}

}  // namespace sd
