// tc_pixel.hpp — per-pixel pieces of terrain correction shared by the
// terrain-correction kernels: DEM and geoid sampling, the three resamplers.
// Reference: sardine/src/dem.rs:351-409, 694-723; geoid.rs:220-251;
// terrain_correction.rs:770-983.
#pragma once

#include "geom.hpp"
#include "tc_params.hpp"

namespace sd {

// One axis of a DEM bilinear lookup: the lower grid index and the fraction.
// A coordinate outside the raster (beyond its half-pixel margin) or NaN
// gets a NaN fraction, which makes the interpolated height NaN: the
// reference's None.  The row axis depends on latitude only and the column
// axis on longitude only, so the terrain geometry's east/west samples share
// the centre's row axis and its north/south samples share the centre's
// column axis.  Indices are 32-bit (the host rejects larger rasters).
struct DemAxis {
    int i0;
    double f;
};

// dem.rs:351-372 / 385-409 (GLO-30 grid rule, bilinear_on_grid), per axis:
// x is the fractional pixel coordinate, max_x = n - 1, hi_x = n - 0.5 and
// lim = n - 2, the largest lower index.
__device__ __forceinline__ DemAxis dem_axis(double x, double max_x, double hi_x, int lim) {
    DemAxis a;
    const bool ok = x >= -0.5 && x <= hi_x;
    x = fmin(fmax(x, 0.0), max_x);
    const int i0 = (int)x;
    a.i0 = i0 < lim ? i0 : lim;
    a.f = ok ? x - (double)a.i0 : __builtin_nan("");
    return a;
}

// (origin - lat) / pixel and (lon - origin) / pixel, divided exactly as the
// reference does (div_rn).
__device__ __forceinline__ DemAxis dem_row_axis(const TcParams& p, double lat) {
    return dem_axis(div_rn(p.dem_olat - lat, p.dem_pix, p.dem_inv_pix), p.dem_max_row, p.dem_hi_row, p.dem_lim_row);
}

__device__ __forceinline__ DemAxis dem_col_axis(const TcParams& p, double lon) {
    // dem.rs:694-723 longitude wrap
    const double lon_n = lon >= 180.0 ? lon - 360.0 : lon;
    return dem_axis(div_rn(lon_n - p.dem_olon, p.dem_pix, p.dem_inv_pix), p.dem_max_col, p.dem_hi_col,
                    p.dem_lim_col);
}

// The four taps of one bilinear lookup.  dem_axis clamps the indices into
// the raster whatever the coordinate (width and height are at least 2), so
// the loads are always in bounds and need no branch.
struct DemTaps {
    float v00, v01, v10, v11;
};

__device__ __forceinline__ DemTaps dem_load(const TcParams& p, DemAxis r, DemAxis c) {
    const float* d = p.dem + (size_t)(unsigned)r.i0 * (size_t)p.dem_w + (unsigned)c.i0;
    return {d[0], d[1], d[p.dem_w], d[p.dem_w + 1]};
}

// The reference's sum in its order; NaN when a tap is NaN (the reference
// returns None for any NaN tap) or the point is outside (NaN fraction).
// Callers test the result with isfinite.
__device__ __forceinline__ float dem_interp(DemAxis r, DemAxis c, const DemTaps& t) {
    const double dr = r.f, dc = c.f;
    const double v = (double)t.v00 * (1.0 - dr) * (1.0 - dc) + (double)t.v01 * (1.0 - dr) * dc +
                     (double)t.v10 * dr * (1.0 - dc) + (double)t.v11 * dr * dc;
    return (float)v;
}

__device__ __forceinline__ float dem_sample(const TcParams& p, double lat, double lon) {
    const DemAxis r = dem_row_axis(p, lat), c = dem_col_axis(p, lon);
    return dem_interp(r, c, dem_load(p, r, c));
}

// Heights of the terrain geometry's four cardinal neighbours (east, west,
// north, south at ±step_deg; terrain_correction.rs:599-606).  They depend on
// the pixel's position only, so the kernels load their taps together with
// the centre's, before the Newton solve; ok is false when any is outside
// the DEM or touches a void (DEM_NEIGHBOUR_MISSING).
struct NbHeights {
    float e, w, n, s;
    bool ok;
};

__device__ __forceinline__ NbHeights nb_heights(const TcParams& p, double lat, double lon, DemAxis rc, DemAxis cc) {
    const double st = p.step_deg;
    const DemAxis ce = dem_col_axis(p, lon + st), cw = dem_col_axis(p, lon - st);
    const DemAxis rn = dem_row_axis(p, lat + st), rs = dem_row_axis(p, lat - st);
    const DemTaps te = dem_load(p, rc, ce), tw = dem_load(p, rc, cw);
    const DemTaps tn = dem_load(p, rn, cc), ts = dem_load(p, rs, cc);
    NbHeights h;
    h.e = dem_interp(rc, ce, te);
    h.w = dem_interp(rc, cw, tw);
    h.n = dem_interp(rn, cc, tn);
    h.s = dem_interp(rs, cc, ts);
    h.ok = isfinite(h.e) && isfinite(h.w) && isfinite(h.n) && isfinite(h.s);
    return h;
}

// ── GLO-30 mosaic (p.dem_tiled): DemMosaic::elevation_at_linear ──────────
// (dem.rs:730-747) over Glo30Tile::elevation_at (dem.rs:351-372): the first
// tile in load order whose extent [-0.5, n - 0.5] (in its own pixel
// coordinates, from its own origin) holds the point and whose four taps are
// finite answers, clamped to its own edges.  Tiles meet on half-pixel
// strips, so at most two tile rows and two tile columns can hold a point.
// What the mosaic lookup reads (passed by value: the lookup is out of line).
struct MosArgs {
    const float* dem;
    const double* tab;  // TcParams::dem_tab
    int64_t dem_w;
    double pix, ipx;
    int trows, tcols, th, tw;
};
__device__ __forceinline__ MosArgs mos_args(const TcParams& p) {
    return MosArgs{p.dem, p.dem_tab, p.dem_w, p.dem_pix, p.dem_inv_pix, p.dem_trows, p.dem_tcols, p.dem_th, p.dem_tw};
}

// The mosaic tables: origins and load orders.
__device__ __forceinline__ double mos_olat(const MosArgs& m, int r) { return m.tab[r]; }
__device__ __forceinline__ double mos_olon(const MosArgs& m, int c) { return m.tab[8 + c]; }
__device__ __forceinline__ int mos_rank(const MosArgs& m, int r, int c) {
    const unsigned long long w = (unsigned long long)__double_as_longlong(m.tab[16 + r]);
    return (int)((w >> (8 * c)) & 0xffu);  // 0xff: no tile
}

// The tile rows (or columns) whose extent holds x: t0 (and t1 when two
// meet there, else -1) with their pixel coordinates.  `f(t)` is the exact
// per-tile coordinate; the first guess comes from tile 0's origin.
template <typename F>
__device__ __forceinline__ void mos_axis(int n, int tile, double guess, F f, int* t0, double* x0, int* t1,
                                         double* x1) {
    int t = (int)floor((guess + 0.5) / (double)tile);
    t = t < 0 ? 0 : (t > n - 1 ? n - 1 : t);
    double x = f(t);
    if (x < -0.5 && t > 0) x = f(--t);
    else if (x > (double)tile - 0.5 && t < n - 1) x = f(++t);
    *t0 = t;
    *x0 = x;
    *t1 = -1;
    *x1 = 0.0;
    if (x < 0.0 && t > 0) {  // the half pixel before its first centre: the previous tile may reach here
        const double y = f(t - 1);
        if (y <= (double)tile - 0.5) *t1 = t - 1, *x1 = y;
    } else if (x > (double)tile - 1.0 && t < n - 1) {
        const double y = f(t + 1);
        if (y >= -0.5) *t1 = t + 1, *x1 = y;
    }
}

// Glo30Tile::elevation_at of tile (tr, tc) at its pixel coordinates.
__device__ __forceinline__ float mos_tile_height(const MosArgs& m, int tr, double rf, int tc, double cf) {
    const int th = m.th, tw = m.tw;
    if (!(rf >= -0.5 && rf <= (double)th - 0.5 && cf >= -0.5 && cf <= (double)tw - 0.5)) return __builtin_nanf("");
    rf = fmin(fmax(rf, 0.0), (double)(th - 1));
    cf = fmin(fmax(cf, 0.0), (double)(tw - 1));
    DemAxis r, c;
    r.i0 = min((int)rf, th - 2);
    c.i0 = min((int)cf, tw - 2);
    r.f = rf - (double)r.i0;
    c.f = cf - (double)c.i0;
    const float* d = m.dem + (size_t)(unsigned)(r.i0 + tr * th) * (size_t)m.dem_w + (unsigned)(c.i0 + tc * tw);
    const DemTaps t{d[0], d[1], d[m.dem_w], d[m.dem_w + 1]};
    return dem_interp(r, c, t);  // NaN when a tap is NaN
}

// Out of line: the chain kernels run at their register limit and only
// mosaic DEMs take this path.
__device__ __noinline__ float dem_height_mosaic(MosArgs m, double lat, double lon) {
    const double lon_n = lon >= 180.0 ? lon - 360.0 : lon;  // dem.rs:697
    int r0, r1, c0, c1;
    double y0, y1, x0, x1;
    mos_axis(m.trows, m.th, (mos_olat(m, 0) - lat) * m.ipx,
             [&](int t) { return div_rn(mos_olat(m, t) - lat, m.pix, m.ipx); }, &r0, &y0, &r1, &y1);
    mos_axis(m.tcols, m.tw, (lon_n - mos_olon(m, 0)) * m.ipx,
             [&](int t) { return div_rn(lon_n - mos_olon(m, t), m.pix, m.ipx); }, &c0, &x0, &c1, &x1);
    // the (at most four) holding tiles in load order; the first with finite
    // taps answers
    int k0 = mos_rank(m, r0, c0), k1 = r1 >= 0 ? mos_rank(m, r1, c0) : 0xff;
    int k2 = c1 >= 0 ? mos_rank(m, r0, c1) : 0xff, k3 = (r1 >= 0 && c1 >= 0) ? mos_rank(m, r1, c1) : 0xff;
    float h = __builtin_nanf("");
    for (int pass = 0; pass < 4; ++pass) {
        const int m01 = k0 < k1 ? k0 : k1, m23 = k2 < k3 ? k2 : k3, mm = m01 < m23 ? m01 : m23;
        if (mm == 0xff) break;
        const int best = mm == k0 ? 0 : (mm == k1 ? 1 : (mm == k2 ? 2 : 3));
        k0 = best == 0 ? 0xff : k0;
        k1 = best == 1 ? 0xff : k1;
        k2 = best == 2 ? 0xff : k2;
        k3 = best == 3 ? 0xff : k3;
        const int tr = (best & 1) ? r1 : r0, tc = (best & 2) ? c1 : c0;
        const double rf = (best & 1) ? y1 : y0, cf = (best & 2) ? x1 : x0;
        h = mos_tile_height(m, tr, rf, tc, cf);
        if (!isnan(h)) break;
    }
    return h;
}

__device__ __forceinline__ double geoid_n(const TcParams& p, double lat, double lon) {
    // geoid.rs:220-251 (2.5° grid, lon wrap, lat clamp)
    if (!p.geoid) return 0.0;
    double l = (lon < 360.0 && lon > -360.0) ? lon : fmod(lon, 360.0);  // fmod is exact: identity below 360
    if (l < 0.0) l += 360.0;
    double la = fmin(fmax(lat, -90.0), 90.0);
    double row_f = (90.0 - la) / 2.5, col_f = l / 2.5;
    int r0 = (int)row_f;
    r0 = r0 < 71 ? r0 : 71;
    int c0 = ((int)col_f) % 144, c1 = (c0 + 1) % 144;
    double dr = row_f - r0;
    double dc = fmin(fmax(col_f - c0, 0.0), 1.0);
    const float* g = p.geoid;
    double v00 = g[r0 * 144 + c0], v01 = g[r0 * 144 + c1], v10 = g[(r0 + 1) * 144 + c0], v11 = g[(r0 + 1) * 144 + c1];
    return v00 * (1.0 - dr) * (1.0 - dc) + v01 * (1.0 - dr) * dc + v10 * dr * (1.0 - dc) + v11 * dr * dc;
}

__device__ __forceinline__ double keys_w(double t) {  // terrain_correction.rs:820-829
    t = fabs(t);
    if (t < 1.0) return 1.5 * t * t * t - 2.5 * t * t + 1.0;
    if (t < 2.0) return -0.5 * t * t * t + 2.5 * t * t - 4.0 * t + 2.0;
    return 0.0;
}


__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// terrain_correction.rs:770-797, split into the tap indices, the loads and
// the interpolation so that the NESZ lookup at the same (line, sample) can
// issue its loads with the image's.  The indices are clamped into the image
// (32-bit: the host rejects larger images); (line, sample) are inside
// [-0.5, n - 0.5] here.
struct Bilin {
    int l0, l1, s0, s1;
    double dl, ds;
};

__device__ __forceinline__ Bilin bilin_index(int64_t nl, int64_t ns, double line, double sample) {
    const int ml = nl > 0 ? (int)nl - 1 : 0, ms = ns > 0 ? (int)ns - 1 : 0;
    Bilin b;
    b.l0 = clampi((int)floor(line), 0, ml);
    b.l1 = b.l0 + 1 < ml ? b.l0 + 1 : ml;
    b.s0 = clampi((int)floor(sample), 0, ms);
    b.s1 = b.s0 + 1 < ms ? b.s0 + 1 : ms;
    b.dl = line - (double)b.l0;
    b.ds = sample - (double)b.s0;
    return b;
}

__device__ __forceinline__ const float* image_row(const float* d, int64_t ns, int l) {
    return d + (size_t)(unsigned)l * (size_t)ns;
}

__device__ __forceinline__ DemTaps bilin_load(const float* d, int64_t ns, const Bilin& b) {
    const float* r0 = image_row(d, ns, b.l0);
    const float* r1 = image_row(d, ns, b.l1);
    return {r0[b.s0], r0[b.s1], r1[b.s0], r1[b.s1]};
}

// NaN when any tap is NaN, as the reference's early return.
__device__ __forceinline__ float bilin_interp(const Bilin& b, const DemTaps& t) {
    const double dl = b.dl, ds = b.ds;
    double w00 = (1.0 - dl) * (1.0 - ds), w01 = (1.0 - dl) * ds, w10 = dl * (1.0 - ds), w11 = dl * ds;
    return (float)((double)t.v00 * w00 + (double)t.v01 * w01 + (double)t.v10 * w10 + (double)t.v11 * w11);
}

__device__ __forceinline__ float bilinear(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    const Bilin b = bilin_index(nl, ns, line, sample);
    return bilin_interp(b, bilin_load(d, ns, b));
}

// terrain_correction.rs:842-872
// A NaN tap makes the reference return NaN at once; here every tap is
// gathered first and a NaN tap turns the f64 sum into NaN instead.  The
// taps are summed row by row (see lanczos3).
__device__ __forceinline__ float bicubic(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    const int lf = (int)floor(line), sf = (int)floor(sample);
    double dl = line - (double)lf, ds = sample - (double)sf;
    double ws[4];
    int cs[4];
#pragma unroll
    for (int si = 0; si < 4; ++si) {
        ws[si] = keys_w((double)(si - 1) - ds);
        cs[si] = clampi(sf + si - 1, 0, (int)ns - 1);
    }
    float v[16];  // all gathers in flight before any is used
#pragma unroll
    for (int li = 0; li < 4; ++li) {
        const float* row = image_row(d, ns, clampi(lf + li - 1, 0, (int)nl - 1));
#pragma unroll
        for (int si = 0; si < 4; ++si) v[li * 4 + si] = row[cs[si]];
    }
    __builtin_amdgcn_sched_barrier(0);
    double acc = 0.0;  // row by row, as lanczos3
#pragma unroll
    for (int li = -1; li <= 2; ++li) {
        const double wl = keys_w((double)li - dl);
        double rs = 0.0;
#pragma unroll
        for (int si = 0; si < 4; ++si) rs += (double)v[(li + 1) * 4 + si] * ws[si];
        acc += rs * wl;
    }
    return (float)acc;
}

// sin and cos of |φ| <= π/6 by their Taylor series in Horner form on φ²
// (first omitted terms below 2e-19 there); no range reduction is needed.
__device__ __forceinline__ void sincos_small(double f, double* s, double* c) {
    const double z = f * f;
    double ps = -7.6471637318198164759e-13;  // -1/15!
    ps = fma(ps, z, 1.6059043836821614599e-10);
    ps = fma(ps, z, -2.5052108385441718775e-08);
    ps = fma(ps, z, 2.7557319223985890653e-06);
    ps = fma(ps, z, -1.9841269841269841270e-04);
    ps = fma(ps, z, 8.3333333333333333333e-03);
    ps = fma(ps, z, -1.6666666666666666667e-01);
    double pc = 4.7794773323873852974e-14;  // 1/16!
    pc = fma(pc, z, -1.1470745597729724714e-11);
    pc = fma(pc, z, 2.0876756987868098979e-09);
    pc = fma(pc, z, -2.7557319223985890653e-07);
    pc = fma(pc, z, 2.4801587301587301587e-05);
    pc = fma(pc, z, -1.3888888888888888889e-03);
    pc = fma(pc, z, 4.1666666666666666667e-02);
    pc = fma(pc, z, -0.5);
    *s = fma(f * z, ps, f);
    *c = fma(z, pc, 1.0);
}

// terrain_correction.rs:938-950: the six Lanczos-3 weights
// L(x) = sin(πx)·sin(πx/3) / (πx·πx/3) (1 for |x| < 1e-10, 0 for |x| >= 3)
// at the tap offsets x = u + 2 - k, k = 0..5, of one axis (u = the
// fractional coordinate in [0, 1)).  Every offset is r + m for the offset r
// nearest zero (r = u or u - 1, |r| <= 1/2, exact) and an integer m, so
// sin(πx) = (-1)^m sin(πr), sin(πx/3) = sin(πr/3) cos(mπ/3) + cos(πr/3)
// sin(mπ/3), and sin(πr) = 3s - 4s³ with s = sin(πr/3) (no cancellation for
// |s| <= 1/2).  Expanding about r keeps every factor relatively accurate
// where its tap's weight is near 1 (x → 0); the six 1/x² come from one
// division (prefix and suffix products).  The weights agree with the
// reference's twelve sin calls to ~1e-15 absolute.
__device__ __forceinline__ void lanczos_axis(double u, double w[6]) {
    const bool up = u >= 0.5;
    const double r = up ? u - 1.0 : u;
    double s3, c3;
    sincos_small(kPi * r * (1.0 / 3.0), &s3, &c3);
    const double s1 = s3 * (3.0 - 4.0 * (s3 * s3));
    constexpr double h = 0.86602540378443864676;  // sin(π/3)
    constexpr double K = 3.0 / (kPi * kPi);
    // cos and sin of mπ/3 for m = -3..3
    const double CM[7] = {-1.0, -0.5, 0.5, 1.0, 0.5, -0.5, -1.0};
    const double SM[7] = {0.0, -h, -h, 0.0, h, h, 0.0};
    double a[7], d[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const int m = i - 3;
        const double x = r + (double)m;
        const double ax = fabs(x);
        const double sp = (m & 1) ? -s1 : s1;
        if (ax < 1e-10) {
            a[i] = 1.0;
            d[i] = 1.0;
        } else if (ax < 3.0) {
            a[i] = sp * (s3 * CM[i] + c3 * SM[i]) * K;
            d[i] = x * x;
        } else {
            a[i] = 0.0;
            d[i] = 1.0;
        }
    }
    double pre[8], suf[8];
    pre[0] = 1.0;
#pragma unroll
    for (int i = 0; i < 7; ++i) pre[i + 1] = pre[i] * d[i];
    suf[7] = 1.0;
#pragma unroll
    for (int i = 6; i >= 0; --i) suf[i] = suf[i + 1] * d[i];
    const double inv = 1.0 / pre[7];
    double W[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) W[i] = a[i] * (inv * (pre[i] * suf[i + 1]));
    // tap k has m = 2 - k (r = u) or m = 3 - k (r = u - 1)
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = up ? W[6 - k] : W[5 - k];
}

// terrain_correction.rs:958-983.  The 36 taps are summed row by row (the
// row's taps times their column weights, then times the row weight); the
// reference adds v·wl·ws tap by tap, so the f64 sums differ in their last
// bits only (~1e-15 relative, below the f32 result's resolution in all but
// ~1e-8 of the samples).  A NaN tap makes the sum NaN: the reference's NaN.
__device__ __forceinline__ float lanczos3(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    const int lf = (int)floor(line), sf = (int)floor(sample);
    double ws[6], wl[6];
    lanczos_axis(sample - (double)sf, ws);
    lanczos_axis(line - (double)lf, wl);
    int cs[6];
#pragma unroll
    for (int si = 0; si < 6; ++si) cs[si] = clampi(sf + si - 2, 0, (int)ns - 1);
    float v[36];  // all gathers in flight before any is used
#pragma unroll
    for (int li = 0; li < 6; ++li) {
        const float* row = image_row(d, ns, clampi(lf + li - 2, 0, (int)nl - 1));
#pragma unroll
        for (int si = 0; si < 6; ++si) v[li * 6 + si] = row[cs[si]];
    }
    __builtin_amdgcn_sched_barrier(0);
    double acc = 0.0;
#pragma unroll
    for (int li = 0; li < 6; ++li) {
        double rs = 0.0;
#pragma unroll
        for (int si = 0; si < 6; ++si) rs += (double)v[li * 6 + si] * ws[si];
        acc += rs * wl[li];
    }
    return (float)acc;
}

// ── the mixed form's FP32 resampling ────────────────────────────────────
// The same taps and rules as above; the weights (from the fractional
// coordinates rounded to f32) and the weighted sums are f32.

__device__ __forceinline__ float bilin_interp_mx(const Bilin& b, const DemTaps& t) {
    const float dl = (float)b.dl, ds = (float)b.ds;
    const float w00 = (1.0f - dl) * (1.0f - ds), w01 = (1.0f - dl) * ds, w10 = dl * (1.0f - ds), w11 = dl * ds;
    return t.v00 * w00 + t.v01 * w01 + t.v10 * w10 + t.v11 * w11;
}

__device__ __forceinline__ float keys_w_mx(float t) {
    t = fabsf(t);
    if (t < 1.0f) return 1.5f * t * t * t - 2.5f * t * t + 1.0f;
    if (t < 2.0f) return -0.5f * t * t * t + 2.5f * t * t - 4.0f * t + 2.0f;
    return 0.0f;
}

__device__ __forceinline__ float bicubic_mx(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    const int lf = (int)floor(line), sf = (int)floor(sample);
    const float dl = (float)(line - (double)lf), ds = (float)(sample - (double)sf);
    float ws[4];
    int cs[4];
#pragma unroll
    for (int si = 0; si < 4; ++si) {
        ws[si] = keys_w_mx((float)(si - 1) - ds);
        cs[si] = clampi(sf + si - 1, 0, (int)ns - 1);
    }
    float v[16];  // all gathers in flight before any is used
#pragma unroll
    for (int li = 0; li < 4; ++li) {
        const float* row = image_row(d, ns, clampi(lf + li - 1, 0, (int)nl - 1));
#pragma unroll
        for (int si = 0; si < 4; ++si) v[li * 4 + si] = row[cs[si]];
    }
    __builtin_amdgcn_sched_barrier(0);
    float acc = 0.0f;
#pragma unroll
    for (int li = -1; li <= 2; ++li) {
        const float wl = keys_w_mx((float)li - dl);
        float rs = 0.0f;
#pragma unroll
        for (int si = 0; si < 4; ++si) rs += v[(li + 1) * 4 + si] * ws[si];
        acc += rs * wl;
    }
    return acc;
}

// sin and cos of |φ| <= π/6 by Taylor series to φ^9 / φ^10 (first omitted
// terms below 3e-11, beneath f32 resolution).
__device__ __forceinline__ void sincos_small_mx(float f, float* s, float* c) {
    const float z = f * f;
    float ps = 2.7557319e-06f;  // 1/9!
    ps = ps * z - 1.9841270e-04f;
    ps = ps * z + 8.3333333e-03f;
    ps = ps * z - 1.6666667e-01f;
    float pc = 2.4801587e-05f;  // 1/8! (the 1/10! term is below 3e-11 here)
    pc = pc * z - 1.3888889e-03f;
    pc = pc * z + 4.1666668e-02f;
    pc = pc * z - 0.5f;
    *s = f + (f * z) * ps;
    *c = 1.0f + z * pc;
}

// lanczos_axis in f32: the same expansion about the nearest offset r.
__device__ __forceinline__ void lanczos_axis_mx(double u64, float w[6]) {
    const bool up = u64 >= 0.5;
    const float r = (float)(up ? u64 - 1.0 : u64);
    float s3, c3;
    sincos_small_mx(3.14159265f * r * (1.0f / 3.0f), &s3, &c3);
    const float s1 = s3 * (3.0f - 4.0f * (s3 * s3));
    constexpr float h = 0.866025404f;  // sin(π/3)
    constexpr float K = 0.303963551f;  // 3/π²
    const float CM[7] = {-1.0f, -0.5f, 0.5f, 1.0f, 0.5f, -0.5f, -1.0f};
    const float SM[7] = {0.0f, -h, -h, 0.0f, h, h, 0.0f};
    float a[7], d[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const int m = i - 3;
        const float x = r + (float)m;
        const float ax = fabsf(x);
        const float sp = (m & 1) ? -s1 : s1;
        // selects, not branches: the chain kernel keeps one basic block here
        const bool tiny = ax < 1e-10f, in = ax < 3.0f;
        const float av = sp * (s3 * CM[i] + c3 * SM[i]) * K;
        a[i] = tiny ? 1.0f : (in ? av : 0.0f);
        d[i] = (tiny || !in) ? 1.0f : x * x;
    }
    float pre[8], suf[8];
    pre[0] = 1.0f;
#pragma unroll
    for (int i = 0; i < 7; ++i) pre[i + 1] = pre[i] * d[i];
    suf[7] = 1.0f;
#pragma unroll
    for (int i = 6; i >= 0; --i) suf[i] = suf[i + 1] * d[i];
    const float inv = __builtin_amdgcn_rcpf(pre[7]);  // within 1 ulp: the weights are f32 anyway
    float W[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) W[i] = a[i] * (inv * (pre[i] * suf[i + 1]));
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = up ? W[6 - k] : W[5 - k];
}

__device__ __forceinline__ float lanczos3_mx(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    const int lf = (int)floor(line), sf = (int)floor(sample);
    int cs[6];
#pragma unroll
    for (int si = 0; si < 6; ++si) cs[si] = clampi(sf + si - 2, 0, (int)ns - 1);
    // all 36 gathers in flight before any is used (left to itself the
    // scheduler issues them one at a time in this kernel)
    float v[36];
#pragma unroll
    for (int li = 0; li < 6; ++li) {
        const float* row = image_row(d, ns, clampi(lf + li - 2, 0, (int)nl - 1));
#pragma unroll
        for (int si = 0; si < 6; ++si) v[li * 6 + si] = row[cs[si]];
    }
    __builtin_amdgcn_sched_barrier(0);
    float ws[6], wl[6];
    lanczos_axis_mx(sample - (double)sf, ws);
    lanczos_axis_mx(line - (double)lf, wl);
    float acc = 0.0f;
#pragma unroll
    for (int li = 0; li < 6; ++li) {
        float rs = 0.0f;
#pragma unroll
        for (int si = 0; si < 6; ++si) rs += v[li * 6 + si] * ws[si];
        acc += rs * wl[li];
    }
    return acc;
}

// RS 0..2: bilinear, bicubic, Lanczos-3 in FP64 (the reference's
// arithmetic); RS 3..5: the same in the mixed form's FP32.
template <int RS>
__device__ __forceinline__ float resample(const TcParams& p, double line, double sample) {
    if constexpr (RS == 1) return bicubic(p.img, p.lines, p.samples, line, sample);
    else if constexpr (RS == 2) return lanczos3(p.img, p.lines, p.samples, line, sample);
    else if constexpr (RS == 4) return bicubic_mx(p.img, p.lines, p.samples, line, sample);
    else if constexpr (RS == 5) return lanczos3_mx(p.img, p.lines, p.samples, line, sample);
    else return bilinear(p.img, p.lines, p.samples, line, sample);
}

}  // namespace sd
