// geom.hpp — FP64 geodesy, orbit interpolation and transverse Mercator used
// by the terrain-correction kernels (device) and its per-row set-up (host).
//
// Reference semantics:
//   geodetic_to_ecef          sardine/src/geodesy.rs:38-55
//   Lagrange-8 pos/vel/accel  sardine/src/orbit.rs:498-596
//   zero-Doppler Newton       sardine/src/terrain_correction.rs:1010-1073
//   flat_cell_area / geometry sardine/src/terrain_correction.rs:576-651
//   inverse UTM               proj4rs 0.1.10 behind output_crs.rs:313-340
//                             (here: Krüger series to n^6, Karney 2011)
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pmath.hpp"

#define SD_HD __host__ __device__ __forceinline__

namespace sd {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg2Rad = kPi / 180.0;
constexpr double kA = 6378137.0;
constexpr double kF = 1.0 / 298.257223563;
constexpr double kE2 = 2.0 * kF - kF * kF;
constexpr double kC = 299792458.0;
constexpr int kLagrange = 8;
constexpr int kMaxSv = 64;

struct d3 {
    double x, y, z;
};
SD_HD d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
SD_HD double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SD_HD d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// x / d correctly rounded, for a divisor d known with its reciprocal
// y = RN(1/d): q0 = x·y is within an ulp of the quotient, r = x - q0·d is
// exact with an FMA, and q0 + r·y rounds to the correctly rounded quotient
// (Markstein, IBM J. Res. Dev. 34(1), 1990).  Three FMA-pipe operations
// instead of the ~10 of a general division; the reference's divisions by
// constants (1e6, the DEM pixel size, the azimuth interval, the range pixel
// spacing) are reproduced bit for bit.
SD_HD double div_rn(double x, double d, double y) {
    const double q0 = x * y;
    const double r = fma(-q0, d, x);
    return fma(r, y, q0);
}

// geodesy.rs:38-55 (same operation order: lat_deg*PI/180, n = a/sqrt(1-e2 s^2)).
// The sines and cosines are the portable ones (pmath.hpp), the same bits
// on the host and on the device.
SD_HD d3 geodetic_to_ecef(double lat_deg, double lon_deg, double h) {
    double lat = lat_deg * kPi / 180.0;
    double lon = lon_deg * kPi / 180.0;
    double sl, cl, so, co;
    pm::sincos(lat, &sl, &cl);
    pm::sincos(lon, &so, &co);
    double n = kA / sqrt(1.0 - kE2 * sl * sl);
    return {(n + h) * cl * co, (n + h) * cl * so, (n * (1.0 - kE2) + h) * sl};
}

// sin and cos of deg·π/180, the radians formed as the reference does
// ((deg·π) / 180, the division exact by div_rn).
SD_HD void sincos_deg(double deg, double* s, double* c) {
    const double r = div_rn(deg * kPi, 180.0, 1.0 / 180.0);
    pm::sincos(r, s, c);
}

// geodetic_to_ecef from the sines/cosines of its (degree) arguments: the
// same operations after the trigonometry, so sharing one angle's sin/cos
// between several calls changes no bit.
SD_HD double prime_vertical(double sl) { return kA / sqrt(1.0 - kE2 * sl * sl); }

SD_HD d3 ecef_from_n(double n, double sl, double cl, double so, double co, double h) {
    return {(n + h) * cl * co, (n + h) * cl * so, (n * (1.0 - kE2) + h) * sl};
}

SD_HD d3 ecef_from_sc(double sl, double cl, double so, double co, double h) {
    return ecef_from_n(prime_vertical(sl), sl, cl, so, co, h);
}

// The Lagrange-8 interpolants of one central interval [t_us[c], t_us[c+1])
// (the interval whose times share one window under the reference's window
// rule) as polynomials in τ = ts - mid, mid = the interval's midpoint in
// seconds: position and velocity of degree 7, acceleration (the derivative
// of the velocity interpolant, which is the reference's Σ v_k L'_k) of
// degree 6, ascending powers.  A degree-7 interpolant is exactly a degree-7
// polynomial, so this is the same function as the reference's product form;
// the host computes the coefficients in binary128 from the double node times
// and state vectors, and |τ| <= half a node spacing keeps Horner's rounding
// at ~1 ulp of the result (~1e-9 m, below the reference's own product-form
// rounding of ~3e-9 m against the exact interpolant).
struct OrbitPoly {
    double mid;
    double pad0;
    double cp[3][8];
    double cv[3][8];
    double ca[3][7];
    double pad1;
};

// Orbit in device-friendly form.  Times are µs relative to the first state
// vector; tsec is the same in seconds ((t_us) / 1e6, orbit.rs:545-547).  The
// kernels copy the part up to poly[n-2] into LDS (orbit_lds_bytes).
struct alignas(16) OrbitDev {
    int32_t n;
    int32_t pad;
    double guess_scale;  // (n - 1) / (t_us[n-1] - t_us[0]): first guess of the bracketing node
    int64_t t_us[kMaxSv];
    double tsec[kMaxSv];
    double pos[kMaxSv][3];
    double vel[kMaxSv][3];
    OrbitPoly poly[kMaxSv - 1];
};

SD_HD size_t orbit_lds_bytes(int n) {
    return offsetof(OrbitDev, poly) + (size_t)(n > 1 ? n - 1 : 0) * sizeof(OrbitPoly);
}

enum : int { ORB_OK = 0, ORB_OUTSIDE = 1, ORB_INSUFFICIENT = 2 };

// Start of the 8-point window of the node c bracketing t (orbit.rs:530-533).
SD_HD int orbit_window_start(int c, int n) {
    if (c < 4) return 0;
    if (c + 4 >= n) return n - kLagrange;
    return c - 3;
}

// interpolate_orbit_and_accel (orbit.rs:498-596): the reference's window
// rule and its refusal to extrapolate.  Between nodes it evaluates the
// interval's polynomials (OrbitPoly); at a node time (t = t_us[c] exactly;
// times are whole µs, so the reference's 1 ns reciprocal guard fires there
// and only there) it takes the reference's own values: every basis but the
// node's is exactly 0 and the node's is exactly 1, so P and V are the node's
// state vector and A = v_c · Σ_{j≠c} 1/(t - t_j) over the window, summed in
// window order.
// `hint` (optional): the interval of the caller's previous evaluation; a
// time strictly inside it (no node, inside the window) skips the search,
// and the interval found is stored back.
SD_HD int orbit_interp(const OrbitDev& o, int64_t t, d3* P, d3* V, d3* A, int* hint = nullptr) {
    const int n = o.n;
    if (n < kLagrange) return ORB_INSUFFICIENT;
    const double tr = (double)(t - o.t_us[0]);
    const double ts = div_rn(tr, 1e6, 1e-6);
    int c = hint ? *hint : -1;
    if (!(c >= 0 && c + 1 < n && o.t_us[c] < t && t < o.t_us[c + 1])) {
        if (t < o.t_us[0] || t > o.t_us[n - 1]) return ORB_OUTSIDE;
        c = (int)(tr * o.guess_scale);  // a guess: the walks below decide
        c = c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
        while (c + 1 < n && o.t_us[c + 1] <= t) ++c;
        while (c > 0 && o.t_us[c] > t) --c;
        if (hint) *hint = c;
    }
    if (t == o.t_us[c]) {
        const int start = orbit_window_start(c, n);
        double tot = 0.0;
        for (int j = 0; j < kLagrange; ++j)
            if (start + j != c) tot += 1.0 / (ts - o.tsec[start + j]);
        *P = {o.pos[c][0], o.pos[c][1], o.pos[c][2]};
        *V = {o.vel[c][0], o.vel[c][1], o.vel[c][2]};
        *A = {tot * o.vel[c][0], tot * o.vel[c][1], tot * o.vel[c][2]};
        return ORB_OK;
    }
    const OrbitPoly& q = o.poly[c];
    const double tau = ts - q.mid;
    double r[9];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double pp = q.cp[d][7], vv = q.cv[d][7], aa = q.ca[d][6];
#pragma unroll
        for (int i = 6; i >= 0; --i) {
            pp = fma(pp, tau, q.cp[d][i]);
            vv = fma(vv, tau, q.cv[d][i]);
        }
#pragma unroll
        for (int i = 5; i >= 0; --i) aa = fma(aa, tau, q.ca[d][i]);
        r[d] = pp;
        r[3 + d] = vv;
        r[6 + d] = aa;
    }
    *P = {r[0], r[1], r[2]};
    *V = {r[3], r[4], r[5]};
    *A = {r[6], r[7], r[8]};
    return ORB_OK;
}

// interpolate_orbit_and_accel (orbit.rs:536-596) evaluated as the reference
// evaluates it, for a time t inside the window of an orbit of at least 8
// vectors: the product-form Lagrange basis with its 56 divisions, the 1 ns
// reciprocal guard, and every sum over the window in the reference's order,
// each operation separately rounded.  Given the same inputs it returns the
// reference's bits; the Newton solve calls it only for the steps the
// interval polynomials above cannot decide (zero_doppler).
SD_HD void orbit_interp_reference(const OrbitDev& o, int64_t t, d3* P, d3* V, d3* A) {
    const int n = o.n;
    int c = 0;
    while (c + 1 < n && o.t_us[c + 1] <= t) ++c;
    const int start = orbit_window_start(c, n);
    const double ts = (double)(t - o.t_us[0]) / 1e6;
    double rd[kLagrange];
    double tot = 0.0;
    for (int j = 0; j < kLagrange; ++j) {
        const double diff = ts - o.tsec[start + j];
        rd[j] = 0.0;
        if (fabs(diff) > 1e-9) {
            rd[j] = 1.0 / diff;
            tot += rd[j];
        }
    }
    d3 p = {0.0, 0.0, 0.0}, v = {0.0, 0.0, 0.0}, a = {0.0, 0.0, 0.0};
    for (int k = 0; k < kLagrange; ++k) {
        double basis = 1.0;
        for (int j = 0; j < kLagrange; ++j)
            if (j != k) basis *= (ts - o.tsec[start + j]) / (o.tsec[start + k] - o.tsec[start + j]);
        const double db = basis * (tot - rd[k]);
        const double* pp = o.pos[start + k];
        const double* vv = o.vel[start + k];
        p.x += basis * pp[0];
        p.y += basis * pp[1];
        p.z += basis * pp[2];
        v.x += basis * vv[0];
        v.y += basis * vv[1];
        v.z += basis * vv[2];
        a.x += db * vv[0];
        a.y += db * vv[1];
        a.z += db * vv[2];
    }
    *P = p;
    *V = v;
    *A = a;
}

// One Newton step's δ and f' at time t exactly as solve_zero_doppler
// computes them (terrain_correction.rs:1035-1044): the reference's orbit
// evaluation, then f = (T - P)·V, f' = -|V|² + (T - P)·A, δ = -f / f'.
SD_HD double zero_doppler_delta_reference(const OrbitDev& o, d3 target, int64_t t, double* fp_out) {
    d3 P, V, A;
    orbit_interp_reference(o, t, &P, &V, &A);
    const double rx = target.x - P.x, ry = target.y - P.y, rz = target.z - P.z;
    const double f = rx * V.x + ry * V.y + rz * V.z;
    const double vsq = V.x * V.x + V.y * V.y + V.z * V.z;
    const double fp = -vsq + (rx * A.x + ry * A.y + rz * A.z);
    *fp_out = fp;
    return -f / fp;
}

// Default half-width (µs) of the band around a decision point inside which
// a Newton step is a tie: the kernels' arithmetic cannot decide it, and it
// takes the reference's.  The polynomial orbit (within 1e-9 m of the exact
// interpolant, as the reference's product form is within 3e-9 m of it) and
// the contracted f, f' move δ·1e6 by at most ~6e-7 µs from the reference's
// value (measured: 3.1e-7 over thousands of such steps); the band is 2^-18 µs
// = 3.8e-6 µs.
constexpr double kTieBandUs = 0x1p-18;

// Counters of the re-evaluated steps: how many, how many changed a decision
// (the µs truncation or the convergence test), and the largest |Δ(δ·1e6)|
// between the two evaluations (the bits of a non-negative double).
enum : int { TIE_STEPS = 0, TIE_FLIPS = 1, TIE_MAX_DEV = 2 };

enum : int { ZD_CONVERGED = 0, ZD_NOT_CONVERGED = 1, ZD_DEGENERATE = 2, ZD_ORBIT_OUTSIDE = 3, ZD_ORBIT_INSUFFICIENT = 4 };

// The last orbit evaluation of a chain.  A pixel's Newton solve starts at
// the previous pixel's converged time, which is exactly the time of that
// pixel's last evaluation, so the first iteration of most pixels finds its
// (P, V, A) here.  orbit_interp is a pure function of t: a hit returns the
// same bits as a fresh evaluation.
struct OrbitCache {
    int64_t t;
    d3 P, V, A;
    int c;  // interval of the last evaluation (orbit_interp's hint)
    bool valid;
};

SD_HD OrbitCache orbit_cache_empty() {
    OrbitCache c;
    c.c = -1;
    c.t = 0;
    c.P = c.V = c.A = d3{0.0, 0.0, 0.0};
    c.valid = false;
    return c;
}

// solve_zero_doppler (terrain_correction.rs:1010-1073): convergence is tested
// before the step, steps are truncated to whole µs.  A tie step (δ·1e6
// within `band` µs of an integer, where the truncation changes, or |δ| that
// close to the tolerance) sets *tie; with EXACT it takes the reference's own
// δ (zero_doppler_delta_reference), so every decision is the reference's
// given the same target, and tie_stats (device, optional) counts such steps.
// The chain kernels solve without EXACT and record where ties fell; the
// fix-up pass replays those pixels with EXACT.
template <bool EXACT>
SD_HD int zero_doppler(const OrbitDev& o, d3 target, int64_t seed, int max_iter, double tol, double band,
                       int64_t* t_out, int* iters, d3* sat, OrbitCache* cache, bool* tie,
                       unsigned long long* tie_stats) {
#pragma clang fp contract(fast)
    int64_t t = seed;
    for (int i = 0; i < max_iter; ++i) {
        // the evaluation writes the cache itself: one copy of (P, V, A)
        // lives across the iterations, not a cache and a working set
        if (!(cache->valid && cache->t == t)) {
            int e = orbit_interp(o, t, &cache->P, &cache->V, &cache->A, &cache->c);
            if (e != ORB_OK) {
                cache->valid = false;
                if (e == ORB_OUTSIDE) {
                    *t_out = t;  // the requested time, for the error report
                    return ZD_ORBIT_OUTSIDE;
                }
                return ZD_ORBIT_INSUFFICIENT;
            }
            cache->t = t;
            cache->valid = true;
        }
        const d3 P = cache->P, V = cache->V, A = cache->A;
        d3 r = target - P;
        double f = dot(r, V);
        double vsq = dot(V, V);
        double fp = -vsq + dot(r, A);
        if (fabs(fp) < 1.0) return ZD_DEGENERATE;
#ifdef __HIP_DEVICE_COMPILE__
        // -f / fp from the hardware reciprocal and two Newton-Raphson steps:
        // within 2 ulp of the quotient (|fp| >= 1 here), far inside the tie
        // band the decisions below are checked against
        double delta;
        {
            double r = __builtin_amdgcn_rcp(fp);
            r = fma(fma(-fp, r, 1.0), r, r);
            r = fma(fma(-fp, r, 1.0), r, r);
            delta = -f * r;
        }
#else
        double delta = -f / fp;
#endif
        const double x = delta * 1e6;
        if (fabs(x - rint(x)) < band || fabs(fabs(delta) - tol) < band * 1e-6) {
            *tie = true;
            if constexpr (EXACT) {
                double fp_ref;
                const double d_ref = zero_doppler_delta_reference(o, target, t, &fp_ref);
#ifdef __HIP_DEVICE_COMPILE__
                if (tie_stats) {
                    const double x_ref = d_ref * 1e6;
                    const bool flip =
                        (int64_t)x != (int64_t)x_ref || (fabs(delta) < tol) != (fabs(d_ref) < tol);
                    const double dev = fabs(x - x_ref);
                    atomicAdd(&tie_stats[TIE_STEPS], 1ull);
                    if (flip) atomicAdd(&tie_stats[TIE_FLIPS], 1ull);
                    atomicMax(&tie_stats[TIE_MAX_DEV], (unsigned long long)__double_as_longlong(dev));
                }
#endif
                if (fabs(fp_ref) < 1.0) return ZD_DEGENERATE;
                delta = d_ref;
            }
        }
        if (fabs(delta) < tol) {
            *t_out = t;
            *iters = i + 1 < 255 ? i + 1 : 255;
            *sat = P;
            return ZD_CONVERGED;
        }
#ifdef __HIP_DEVICE_COMPILE__
        // the µs step truncated toward zero as the 64-bit conversion does;
        // steps below 2^31 µs (all but divergent ones) by the 32-bit one
        const double xs = delta * 1e6;
        int64_t step;
        if (__builtin_expect(fabs(xs) < 0x1p31, 1)) step = (int64_t)(int32_t)xs;
        else step = (int64_t)xs;
        t = t + step;
#else
        t = t + (int64_t)(delta * 1e6);
#endif
    }
    return ZD_NOT_CONVERGED;
}

// flat_cell_area (terrain_correction.rs:576-583)
SD_HD double flat_cell_area(double lat, double lon, double s) {
    d3 fe = geodetic_to_ecef(lat, lon + s, 0.0), fw = geodetic_to_ecef(lat, lon - s, 0.0);
    d3 fn = geodetic_to_ecef(lat + s, lon, 0.0), fs = geodetic_to_ecef(lat - s, lon, 0.0);
    d3 c = cross(fe - fw, fn - fs);
    return sqrt(dot(c, c));
}

enum : int { TG_VALID = 0, TG_FORESHORTENING = 1, TG_LAYOVER = 2, TG_SHADOW = 3 };



// compute_terrain_geometry_with_flat_area (terrain_correction.rs:593-651).
// The five points share trigonometry: (sl, cl, so, co) are the sines/cosines
// of the centre latitude/longitude and (ss, cs) those of the cardinal step
// s (all in radians of the degree arguments); the neighbours' sines/cosines
// follow by angle addition, sin(a±b) = sin a cos b ± cos a sin b.  That
// moves each neighbour by ~1e-10 m against evaluating sin((lon+s)π/180)
// directly, far below anything the facet normal or the 0.05 / 2.0 mask
// thresholds resolve.  The east/west points have the centre's latitude and
// so its prime-vertical radius n_c; the north/south radii follow from n_c by
// the binomial series of n = a (x_c (1 + ε))^-1/2 in ε = (x - x_c) / x_c,
// x = 1 - e² sin²φ (|ε| < 1e-5 for steps below 1°; the series to ε³ is exact
// to ~1e-20).  The look vector is not normalised: lp = n·(S - T) / |S - T|
// with one reciprocal square root, cos_lia = lp / |n| likewise, and
// w = lp · (1 / flat_area).  Each quantity is within a few ulp of the
// reference's three divisions and two square roots, which moves w and
// cos_lia by ~1e-15 relative: the masks (thresholds 0, 0.05, 2) do not
// resolve it.
SD_HD double prime_vertical_near(double n_c, double sl, double sl2) {
    const double inv_xc = (n_c * (1.0 / kA)) * (n_c * (1.0 / kA));
    const double eps = kE2 * ((sl - sl2) * (sl + sl2)) * inv_xc;
    return n_c * (1.0 + eps * (-0.5 + eps * (0.375 - 0.3125 * eps)));
}

SD_HD double rsqrt_d(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    return rsqrt(x);
#else
    return 1.0 / sqrt(x);
#endif
}

SD_HD int terrain_geometry_sc(double sl, double cl, double so, double co, double n_c, double ss, double cs,
                              double he, double hw, double hn, double hs, d3 sat, d3 target, double inv_flat_area,
                              double* w_out, double* cos_lia_out) {
    const double soe = so * cs + co * ss, coe = co * cs - so * ss;  // lon + s
    const double sow = so * cs - co * ss, cow = co * cs + so * ss;  // lon - s
    const double sln = sl * cs + cl * ss, cln = cl * cs - sl * ss;  // lat + s
    const double sls = sl * cs - cl * ss, cls = cl * cs + sl * ss;  // lat - s
    d3 pe = ecef_from_n(n_c, sl, cl, soe, coe, he), pw = ecef_from_n(n_c, sl, cl, sow, cow, hw);
    d3 pn = ecef_from_n(prime_vertical_near(n_c, sl, sln), sln, cln, so, co, hn);
    d3 ps = ecef_from_n(prime_vertical_near(n_c, sl, sls), sls, cls, so, co, hs);
    d3 tn = cross(pe - pw, pn - ps);
    d3 lk = sat - target;
    const double l2 = dot(lk, lk);
    const double lr = dot(tn, lk);
    if (!(l2 > 0.0) || lr <= 0.0) return TG_SHADOW;  // |S - T| = 0: zero look vector, lp = 0
    const double lp = lr * rsqrt_d(l2);
    const double t2 = dot(tn, tn);
    if (t2 == 0.0) return TG_SHADOW;
    double cosl = lp * rsqrt_d(t2);
    cosl = cosl < 0.0 ? 0.0 : (cosl > 1.0 ? 1.0 : cosl);
    double w = lp * inv_flat_area;
    *w_out = w;
    *cos_lia_out = cosl;
    if (w < 0.05) return TG_FORESHORTENING;
    if (w > 2.0) return TG_LAYOVER;
    return TG_VALID;
}

// terrain_geometry_sc in the mixed form: the neighbour points and their
// differences in FP64 (the differences of ECEF coordinates ~6.4e6 m need
// it), the normal, the look vector's products, the two inverse square roots
// and the thresholds in FP32.
SD_HD int terrain_geometry_mixed(double sl, double cl, double so, double co, double n_c, double ss, double cs,
                                 double he, double hw, double hn, double hs, d3 sat, d3 target, double inv_flat_area,
                                 float* w_out, float* cos_lia_out) {
    const double soe = so * cs + co * ss, coe = co * cs - so * ss;  // lon + s
    const double sow = so * cs - co * ss, cow = co * cs + so * ss;  // lon - s
    const double sln = sl * cs + cl * ss, cln = cl * cs - sl * ss;  // lat + s
    const double sls = sl * cs - cl * ss, cls = cl * cs + sl * ss;  // lat - s
    const d3 pe = ecef_from_n(n_c, sl, cl, soe, coe, he), pw = ecef_from_n(n_c, sl, cl, sow, cow, hw);
    const d3 pn = ecef_from_n(prime_vertical_near(n_c, sl, sln), sln, cln, so, co, hn);
    const d3 ps = ecef_from_n(prime_vertical_near(n_c, sl, sls), sls, cls, so, co, hs);
    const d3 dew = pe - pw, dns = pn - ps, lk64 = sat - target;
    const float ax = (float)dew.x, ay = (float)dew.y, az = (float)dew.z;
    const float bx = (float)dns.x, by = (float)dns.y, bz = (float)dns.z;
    const float lx = (float)lk64.x, ly = (float)lk64.y, lz = (float)lk64.z;
    const float tx = ay * bz - az * by, ty = az * bx - ax * bz, tz = ax * by - ay * bx;
    const float l2 = lx * lx + ly * ly + lz * lz;
    const float lr = tx * lx + ty * ly + tz * lz;
    if (!(l2 > 0.0f) || lr <= 0.0f) return TG_SHADOW;
    const float lp = lr * (1.0f / sqrtf(l2));
    const float t2 = tx * tx + ty * ty + tz * tz;
    if (t2 == 0.0f) return TG_SHADOW;
    float cosl = lp * (1.0f / sqrtf(t2));
    cosl = cosl < 0.0f ? 0.0f : (cosl > 1.0f ? 1.0f : cosl);
    const float w = lp * (float)inv_flat_area;
    *w_out = w;
    *cos_lia_out = cosl;
    if (w < 0.05f) return TG_FORESHORTENING;
    if (w > 2.0f) return TG_LAYOVER;
    return TG_VALID;
}

// ── inverse transverse Mercator (UTM) ────────────────────────────────────
// Krüger series of order n^6 in the third flattening n (Karney 2011,
// "Transverse Mercator with an accuracy of a few nanometers", eqs. 7, 11, 35,
// 36), evaluated with Clenshaw summation; τ from τ' by Newton (eqs. 19-21).
struct TmSeries {
    double A;       // rectifying radius
    double beta[6]; // inverse series
    double alpha[6];
    double e;
};

SD_HD TmSeries tm_series() {
    const double n = kF / (2.0 - kF);
    const double n2 = n * n, n3 = n2 * n, n4 = n3 * n, n5 = n4 * n, n6 = n5 * n;
    TmSeries s;
    s.e = sqrt(kE2);
    s.A = kA / (1.0 + n) * (1.0 + n2 / 4.0 + n4 / 64.0 + n6 / 256.0);
    s.alpha[0] = n / 2 - 2 * n2 / 3 + 5 * n3 / 16 + 41 * n4 / 180 - 127 * n5 / 288 + 7891 * n6 / 37800;
    s.alpha[1] = 13 * n2 / 48 - 3 * n3 / 5 + 557 * n4 / 1440 + 281 * n5 / 630 - 1983433 * n6 / 1935360;
    s.alpha[2] = 61 * n3 / 240 - 103 * n4 / 140 + 15061 * n5 / 26880 + 167603 * n6 / 181440;
    s.alpha[3] = 49561 * n4 / 161280 - 179 * n5 / 168 + 6601661 * n6 / 7257600;
    s.alpha[4] = 34729 * n5 / 80640 - 3418889 * n6 / 1995840;
    s.alpha[5] = 212378941 * n6 / 319334400;
    s.beta[0] = n / 2 - 2 * n2 / 3 + 37 * n3 / 96 - n4 / 360 - 81 * n5 / 512 + 96199 * n6 / 604800;
    s.beta[1] = n2 / 48 + n3 / 15 - 437 * n4 / 1440 + 46 * n5 / 105 - 1118711 * n6 / 3870720;
    s.beta[2] = 17 * n3 / 480 - 37 * n4 / 840 - 209 * n5 / 4480 + 5569 * n6 / 90720;
    s.beta[3] = 4397 * n4 / 161280 - 11 * n5 / 504 - 830251 * n6 / 7257600;
    s.beta[4] = 4583 * n5 / 161280 - 108847 * n6 / 3991680;
    s.beta[5] = 20648693 * n6 / 638668800;
    return s;
}

constexpr double kUtmK0 = 0.9996;
constexpr double kUtmE0 = 500000.0;
constexpr double kUtmN0South = 10000000.0;

// Σ_j c_j sin(2j ξ) cosh(2j η) and Σ_j c_j cos(2j ξ) sinh(2j η) by the complex
// Clenshaw recurrence on z = ξ + iη.
SD_HD void clenshaw_series(const double* c, double xi, double eta, double* sx, double* sy) {
    double s2, c2;
    pm::sincos(2 * xi, &s2, &c2);
    const double sh2 = pm::sinh(2 * eta), ch2 = pm::cosh(2 * eta);
    // a = 2 cos(2z) = 2(c2 ch2 - i s2 sh2)
    double ar = 2 * c2 * ch2, ai = -2 * s2 * sh2;
    double y0r = 0, y0i = 0, y1r = 0, y1i = 0;
    for (int k = 5; k >= 0; --k) {
        double tr = ar * y0r - ai * y0i - y1r + c[k];
        double ti = ar * y0i + ai * y0r - y1i;
        y1r = y0r; y1i = y0i;
        y0r = tr; y0i = ti;
    }
    // sum = sin(2z) * y0 ; sin(2z) = s2 ch2 + i c2 sh2
    double zr = s2 * ch2, zi = c2 * sh2;
    *sx = zr * y0r - zi * y0i;
    *sy = zr * y0i + zi * y0r;
}

SD_HD void utm_inverse(const TmSeries& s, int zone, bool south, double x, double y, double* lon_deg,
                       double* lat_deg) {
    double xi = (y - (south ? kUtmN0South : 0.0)) / (kUtmK0 * s.A);
    double eta = (x - kUtmE0) / (kUtmK0 * s.A);
    double sx, sy;
    clenshaw_series(s.beta, xi, eta, &sx, &sy);
    double xip = xi - sx, etap = eta - sy;
    double sxp, cx;
    pm::sincos(xip, &sxp, &cx);
    double sh = pm::sinh(etap);
    double taup = sxp / sqrt(sh * sh + cx * cx);
    double lam = pm::atan2(sh, cx);
    const double e = s.e, e2 = kE2;
    double tau = taup;
    for (int i = 0; i < 8; ++i) {
        double s1 = sqrt(1.0 + tau * tau);
        double sig = pm::sinh(e * pm::atanh(e * tau / s1));
        double taui = tau * sqrt(1.0 + sig * sig) - sig * s1;
        double dt = (taup - taui) / sqrt(1.0 + taui * taui) * (1.0 + (1.0 - e2) * tau * tau) / ((1.0 - e2) * s1);
        tau += dt;
        if (fabs(dt) < 1e-14) break;
    }
    *lat_deg = pm::atan(tau) * 180.0 / kPi;
    *lon_deg = (double)(zone * 6 - 183) + lam * 180.0 / kPi;
}

SD_HD void utm_forward(const TmSeries& s, int zone, bool south, double lon_deg, double lat_deg, double* x,
                       double* y) {
    double phi = lat_deg * kPi / 180.0, lam = (lon_deg - (double)(zone * 6 - 183)) * kPi / 180.0;
    double sp = pm::sin(phi);
    double sl, cl;
    pm::sincos(lam, &sl, &cl);
    double t = pm::sinh(pm::atanh(sp) - s.e * pm::atanh(s.e * sp));
    double xip = pm::atan2(t, cl);
    double etap = pm::atanh(sl / sqrt(1.0 + t * t));
    double sx, sy;
    clenshaw_series(s.alpha, xip, etap, &sx, &sy);
    *x = kUtmE0 + kUtmK0 * s.A * (etap + sy);
    *y = (south ? kUtmN0South : 0.0) + kUtmK0 * s.A * (xip + sx);
}

}  // namespace sd
