// tc_pixel.hpp — per-pixel pieces of terrain correction shared by the
// terrain-correction kernels: DEM and geoid sampling, the three resamplers.
// Reference: sardine/src/dem.rs:351-409, 694-723; geoid.rs:220-251;
// terrain_correction.rs:770-983.
#pragma once

#include "geom.hpp"
#include "tc_params.hpp"

namespace sd {

// One axis of a DEM bilinear lookup: the lower grid index, the fraction and
// whether the coordinate is inside the raster (half a pixel of margin).  The
// row axis depends on latitude only and the column axis on longitude only,
// so the terrain geometry's east/west samples share the centre's row axis and
// its north/south samples share the centre's column axis.
struct DemAxis {
    int64_t i0;
    double f;
    bool ok;
};

// dem.rs:351-372 / 385-409 (GLO-30 grid rule, bilinear_on_grid), per axis.
__device__ __forceinline__ DemAxis dem_axis(double x, double max_x, int64_t lim) {
    DemAxis a;
    a.ok = !(x < -0.5 || x > max_x + 0.5);
    x = fmin(fmax(x, 0.0), max_x);
    int64_t i0 = (int64_t)x;
    a.i0 = i0 < lim ? i0 : lim;
    a.f = x - (double)a.i0;
    return a;
}

__device__ __forceinline__ DemAxis dem_row_axis(const TcParams& p, double lat) {
    return dem_axis((p.dem_olat - lat) * p.dem_inv_pix, p.dem_max_row, p.dem_lim_row);
}

__device__ __forceinline__ DemAxis dem_col_axis(const TcParams& p, double lon) {
    // dem.rs:694-723 longitude wrap
    const double lon_n = lon >= 180.0 ? lon - 360.0 : lon;
    return dem_axis((lon_n - p.dem_olon) * p.dem_inv_pix, p.dem_max_col, p.dem_lim_col);
}

__device__ __forceinline__ bool dem_bilinear(const TcParams& p, DemAxis r, DemAxis c, float* h) {
    if (!(r.ok && c.ok)) return false;
    const float* d = p.dem + r.i0 * p.dem_w + c.i0;
    float v00 = d[0], v01 = d[1], v10 = d[p.dem_w], v11 = d[p.dem_w + 1];
    if (isnan(v00) || isnan(v01) || isnan(v10) || isnan(v11)) return false;
    const double dr = r.f, dc = c.f;
    double v = (double)v00 * (1.0 - dr) * (1.0 - dc) + (double)v01 * (1.0 - dr) * dc +
               (double)v10 * dr * (1.0 - dc) + (double)v11 * dr * dc;
    *h = (float)v;
    return true;
}

__device__ __forceinline__ bool dem_sample(const TcParams& p, double lat, double lon, float* h) {
    return dem_bilinear(p, dem_row_axis(p, lat), dem_col_axis(p, lon), h);
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


__device__ __forceinline__ int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// terrain_correction.rs:770-797
__device__ __forceinline__ float bilinear(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    int64_t ml = nl > 0 ? nl - 1 : 0, ms = ns > 0 ? ns - 1 : 0;
    int64_t l0 = clampi((int64_t)floor(line), 0, ml), l1 = l0 + 1 < ml ? l0 + 1 : ml;
    int64_t s0 = clampi((int64_t)floor(sample), 0, ms), s1 = s0 + 1 < ms ? s0 + 1 : ms;
    double dl = line - (double)l0, ds = sample - (double)s0;
    float v00 = d[l0 * ns + s0], v01 = d[l0 * ns + s1], v10 = d[l1 * ns + s0], v11 = d[l1 * ns + s1];
    if (isnan(v00) || isnan(v01) || isnan(v10) || isnan(v11)) return __builtin_nanf("");
    double w00 = (1.0 - dl) * (1.0 - ds), w01 = (1.0 - dl) * ds, w10 = dl * (1.0 - ds), w11 = dl * ds;
    return (float)((double)v00 * w00 + (double)v01 * w01 + (double)v10 * w10 + (double)v11 * w11);
}

// terrain_correction.rs:842-872
// A NaN tap makes the reference return NaN at once; here every tap is
// gathered first and a NaN tap turns the f64 sum into NaN instead, which
// yields the same NaN result and the same bits otherwise (the sum runs in
// the reference's order).
__device__ __forceinline__ float bicubic(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    int64_t lf = (int64_t)floor(line), sf = (int64_t)floor(sample);
    double dl = line - (double)lf, ds = sample - (double)sf;
    double ws[4];
    int64_t cs[4];
#pragma unroll
    for (int si = 0; si < 4; ++si) {
        ws[si] = keys_w((double)(si - 1) - ds);
        cs[si] = clampi(sf + si - 1, 0, ns - 1);
    }
    double acc = 0.0;
#pragma unroll
    for (int li = -1; li <= 2; ++li) {
        const float* row = d + clampi(lf + li, 0, nl - 1) * ns;
        double wl = keys_w((double)li - dl);
#pragma unroll
        for (int si = 0; si < 4; ++si) acc += (double)row[cs[si]] * wl * ws[si];
    }
    return (float)acc;
}

// terrain_correction.rs:958-983
// The six Lanczos-3 weights L(x) = sin(πx)·sin(πx/3) / (πx·πx/3) (1 for
// |x| < 1e-10, 0 for |x| >= 3; terrain_correction.rs:938-950) at
// x = u + 2 - k, k = 0..5, of one axis (u = the fractional coordinate in
// [0, 1)).  sin(π(u+m)) = (-1)^m sin(πu)
// and sin(π(u+m)/3) follows from sin/cos(πu/3) by angle addition, so one
// sin and one sincos replace the reference's twelve sin calls; the weights
// agree with the direct form to ~1e-15.
__device__ __forceinline__ void lanczos_axis(double u, double w[6]) {
    const double s1 = sin(kPi * u);
    double s3, c3;
    sincos(kPi * u / 3.0, &s3, &c3);
    constexpr double h = 0.86602540378443864676;  // sin(π/3)
    // cos and sin of π(2 - k)/3 for k = 0..5
    const double C[6] = {-0.5, 0.5, 1.0, 0.5, -0.5, -1.0};
    const double S[6] = {h, h, 0.0, -h, -h, 0.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double x = u + (double)(2 - k);
        const double ax = fabs(x);
        double v = 0.0;
        if (ax < 1e-10) {
            v = 1.0;
        } else if (ax < 3.0) {
            const double px = kPi * x, pxa = kPi * x / 3.0;
            const double sp = (k & 1) ? -s1 : s1;
            v = sp * (s3 * C[k] + c3 * S[k]) / (px * pxa);
        }
        w[k] = v;
    }
}

__device__ __forceinline__ float lanczos3(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    int64_t lf = (int64_t)floor(line), sf = (int64_t)floor(sample);
    double ws[6], wl[6];
    lanczos_axis(sample - (double)sf, ws);
    lanczos_axis(line - (double)lf, wl);
    int64_t cs[6];
#pragma unroll
    for (int si = 0; si < 6; ++si) cs[si] = clampi(sf + si - 2, 0, ns - 1);
    double acc = 0.0;  // NaN taps: see bicubic
#pragma unroll
    for (int li = 0; li < 6; ++li) {
        const float* row = d + clampi(lf + li - 2, 0, nl - 1) * ns;
#pragma unroll
        for (int si = 0; si < 6; ++si) acc += (double)row[cs[si]] * wl[li] * ws[si];
    }
    return (float)acc;
}

template <int RS>
__device__ __forceinline__ float resample(const TcParams& p, double line, double sample) {
    if constexpr (RS == 1) return bicubic(p.img, p.lines, p.samples, line, sample);
    else if constexpr (RS == 2) return lanczos3(p.img, p.lines, p.samples, line, sample);
    else return bilinear(p.img, p.lines, p.samples, line, sample);
}

}  // namespace sd
