// tc32.hip — the FP32 scene-centred variant of terrain correction.
//
// The grid position, DEM sample and geodesy are the FP64 path's (tc_device.hpp);
// from the target onwards the pixel runs in single precision on well-scaled
// quantities: ECEF offsets from a scene centre (|x| ≲ 1e6 m) and times in
// seconds from a reference time near the scene (|t| ≲ 1e2 s).  The 8-point
// Lagrange orbit (barycentric, the FP64 path's window rule), the zero-Doppler
// Newton iteration, the slant range, the resampling weights and sums, and
// the terrain geometry (facet normal from FP64 neighbour offsets rounded to
// FP32) are FP32.  Newton steps are not quantised to µs and each pixel is
// solved on its own from its row's seed, so there is no seed chain and a
// lane owns one column (coalesced stores).  Single precision cannot resolve
// the zero-Doppler condition below ~1e-5 s (f = (T−P)·V carries rounding
// noise ~|T−P||V|·2⁻²⁴), which is why the variant has its own tolerance;
// every result of it is reported with its error against the FP64 path.
#include "tc_device.hpp"

namespace sd {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// orbit_interp in FP32 (same window rule and derivative form).
__device__ __forceinline__ int orbit_interp32(const OrbitF32& o, float t, f3* P, f3* V, f3* A) {
    const int n = o.n;
    if (n < kLagrange) return ORB_INSUFFICIENT;
    if (t < o.t[0] || t > o.t[n - 1]) return ORB_OUTSIDE;
    int c = (int)((float)(n - 1) * (t - o.t[0]) / (o.t[n - 1] - o.t[0]));
    c = c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
    while (c + 1 < n && o.t[c + 1] <= t) ++c;
    while (c > 0 && o.t[c] > t) --c;
    int start;
    if (c < 4) start = 0;
    else if (c + 4 >= n) start = n - kLagrange;
    else start = c - 3;
    float d[kLagrange], pre[kLagrange + 1], suf[kLagrange + 1];
#pragma unroll
    for (int j = 0; j < kLagrange; ++j) d[j] = t - o.t[start + j];
    pre[0] = 1.0f;
#pragma unroll
    for (int j = 0; j < kLagrange; ++j) pre[j + 1] = pre[j] * d[j];
    suf[kLagrange] = 1.0f;
#pragma unroll
    for (int j = kLagrange - 1; j >= 0; --j) suf[j] = suf[j + 1] * d[j];
    // reciprocals 1/(t - t_j) from one division of the full product, as in
    // the FP64 path; a node closer than ~1e-6 s takes the per-node form
    float rd[kLagrange], tot = 0.0f;
    bool guard = false;
#pragma unroll
    for (int j = 0; j < kLagrange; ++j) guard |= !(fabsf(d[j]) > 1e-6f);
    if (!guard) {
        const float inv_all = 1.0f / pre[kLagrange];
#pragma unroll
        for (int j = 0; j < kLagrange; ++j) {
            rd[j] = inv_all * (pre[j] * suf[j + 1]);
            tot += rd[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < kLagrange; ++j) {
            rd[j] = fabsf(d[j]) > 1e-9f ? 1.0f / d[j] : 0.0f;
            tot += rd[j];
        }
    }
    const float* w = o.bary[start];
    f3 p = {0, 0, 0}, v = {0, 0, 0}, a = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < kLagrange; ++k) {
        const float basis = w[k] * (pre[k] * suf[k + 1]);
        const float db = basis * (tot - rd[k]);
        const float* pp = o.pos[start + k];
        const float* vv = o.vel[start + k];
        p.x += basis * pp[0]; p.y += basis * pp[1]; p.z += basis * pp[2];
        v.x += basis * vv[0]; v.y += basis * vv[1]; v.z += basis * vv[2];
        a.x += db * vv[0]; a.y += db * vv[1]; a.z += db * vv[2];
    }
    *P = p; *V = v; *A = a;
    return ORB_OK;
}

// solve_zero_doppler in FP32 with unquantised steps; on convergence the
// last step is applied (to first order in the position) rather than
// discarded, since no µs chain is emulated here.
__device__ __forceinline__ int zero_doppler32(const OrbitF32& o, f3 T, float t0, int max_iter, float tol, float* t_out,
                                              int* iters, f3* sat) {
    float t = t0;
    for (int i = 0; i < max_iter; ++i) {
        f3 P, V, A;
        const int e = orbit_interp32(o, t, &P, &V, &A);
        if (e == ORB_OUTSIDE) {
            *t_out = t;  // the requested time, for the error report
            return ZD_ORBIT_OUTSIDE;
        }
        if (e == ORB_INSUFFICIENT) return ZD_ORBIT_INSUFFICIENT;
        const f3 r = T - P;
        const float f = dot3(r, V);
        const float fp = -dot3(V, V) + dot3(r, A);
        if (fabsf(fp) < 1.0f) return ZD_DEGENERATE;
        const float delta = -f / fp;
        if (fabsf(delta) < tol) {  // take the last step too: t is then within the rounding noise
            *t_out = t + delta;
            *iters = i + 1 < 255 ? i + 1 : 255;
            *sat = {P.x + V.x * delta, P.y + V.y * delta, P.z + V.z * delta};
            return ZD_CONVERGED;
        }
        t += delta;
    }
    return ZD_NOT_CONVERGED;
}

__device__ __forceinline__ float keys_w32(float t) {
    t = fabsf(t);
    if (t < 1.0f) return 1.5f * t * t * t - 2.5f * t * t + 1.0f;
    if (t < 2.0f) return -0.5f * t * t * t + 2.5f * t * t - 4.0f * t + 2.0f;
    return 0.0f;
}

// Lanczos-3 weights of one axis in FP32, by the FP64 path's expansion about
// the tap offset r nearest zero (lanczos_axis): one polynomial sincos of
// πr/3 (|πr/3| <= π/6; Taylor to φ⁹ / φ⁸), sin(πr) = 3s - 4s³, angle
// addition for the other taps and one division for the six 1/x².
__device__ __forceinline__ void lanczos_axis32(float u, float w[6]) {
    const bool up = u >= 0.5f;
    const float r = up ? u - 1.0f : u;
    const float f = 1.04719755f * r;  // π/3 · r
    const float z = f * f;
    const float s3 = f + f * z * (-1.66666667e-1f + z * (8.33333333e-3f + z * (-1.98412698e-4f + z * 2.75573192e-6f)));
    const float c3 = 1.0f + z * (-0.5f + z * (4.16666667e-2f + z * (-1.38888889e-3f + z * 2.48015873e-5f)));
    const float s1 = s3 * (3.0f - 4.0f * (s3 * s3));
    const float h = 0.866025404f;
    const float K = 0.303963551f;  // 3 / π²
    const float CM[7] = {-1.0f, -0.5f, 0.5f, 1.0f, 0.5f, -0.5f, -1.0f};
    const float SM[7] = {0.0f, -h, -h, 0.0f, h, h, 0.0f};
    float a[7], d[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const int m = i - 3;
        const float x = r + (float)m, ax = fabsf(x);
        const float sp = (m & 1) ? -s1 : s1;
        if (ax < 1e-6f) {
            a[i] = 1.0f;
            d[i] = 1.0f;
        } else if (ax < 3.0f) {
            a[i] = sp * (s3 * CM[i] + c3 * SM[i]) * K;
            d[i] = x * x;
        } else {
            a[i] = 0.0f;
            d[i] = 1.0f;
        }
    }
    float pre[8], suf[8];
    pre[0] = 1.0f;
#pragma unroll
    for (int i = 0; i < 7; ++i) pre[i + 1] = pre[i] * d[i];
    suf[7] = 1.0f;
#pragma unroll
    for (int i = 6; i >= 0; --i) suf[i] = suf[i + 1] * d[i];
    const float inv = 1.0f / pre[7];
    float W[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) W[i] = a[i] * (inv * (pre[i] * suf[i + 1]));
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = up ? W[6 - k] : W[5 - k];
}

// The three resamplers with FP32 weights and sums (NaN taps propagate).
template <int RS>
__device__ __forceinline__ float resample32(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    const int lf = (int)floor(line), sf = (int)floor(sample);
    const float dl = (float)(line - (double)lf), ds = (float)(sample - (double)sf);
    if constexpr (RS == 0) {
        const int ml = nl > 0 ? (int)nl - 1 : 0, ms = ns > 0 ? (int)ns - 1 : 0;
        const int l0 = clampi(lf, 0, ml), l1 = l0 + 1 < ml ? l0 + 1 : ml;
        const int s0 = clampi(sf, 0, ms), s1 = s0 + 1 < ms ? s0 + 1 : ms;
        const float el = (float)(line - (double)l0), es = (float)(sample - (double)s0);
        const float* r0 = image_row(d, ns, l0);
        const float* r1 = image_row(d, ns, l1);
        const float v00 = r0[s0], v01 = r0[s1], v10 = r1[s0], v11 = r1[s1];
        return v00 * (1.0f - el) * (1.0f - es) + v01 * (1.0f - el) * es + v10 * el * (1.0f - es) + v11 * el * es;
    } else if constexpr (RS == 1) {
        float ws[4], acc = 0.0f;
        int cs[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ws[k] = keys_w32((float)(k - 1) - ds);
            cs[k] = clampi(sf + k - 1, 0, (int)ns - 1);
        }
#pragma unroll
        for (int li = -1; li <= 2; ++li) {
            const float* row = image_row(d, ns, clampi(lf + li, 0, (int)nl - 1));
            const float wl = keys_w32((float)li - dl);
            float rs = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) rs += row[cs[k]] * ws[k];
            acc += rs * wl;
        }
        return acc;
    } else {
        float ws[6], wl[6], acc = 0.0f;
        lanczos_axis32(ds, ws);
        lanczos_axis32(dl, wl);
        int cs[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) cs[k] = clampi(sf + k - 2, 0, (int)ns - 1);
#pragma unroll
        for (int li = 0; li < 6; ++li) {
            const float* row = image_row(d, ns, clampi(lf + li - 2, 0, (int)nl - 1));
            float rs = 0.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) rs += row[cs[k]] * ws[k];
            acc += rs * wl[li];
        }
        return acc;
    }
}

__device__ __forceinline__ f3 to_f3(d3 v) { return {(float)v.x, (float)v.y, (float)v.z}; }

// Everything after the Newton solve, mirroring finish_pixel with FP32
// resampling and geometry.
template <int RS>
__device__ __forceinline__ PixelOut finish32(const TcParams& p, const PixelPos& q, f3 T, int oc, float t, f3 sat,
                                             double flat_area) {
    PixelOut o;
    o.sigma0 = __builtin_nanf("");
    o.cos_lia = __builtin_nanf("");
    o.line = __builtin_nan("");
    o.sample = __builtin_nan("");
    o.mask = 1;
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
    const f3 rv = T - sat;
    const float slant = sqrtf(dot3(rv, rv));
    const double line = ((double)(p.t_ref_us - p.az_start_us) + (double)t * 1e6) * p.line_per_us;
    const double sample = ((double)slant - p.near_range_m) * p.inv_rps;  // |T−P| is translation-free
    o.line = line;
    o.sample = sample;
    if (line < -0.5 || line > (double)p.lines - 0.5 || sample < -0.5 || sample > (double)p.samples - 0.5) {
        o.cat = CAT_OUT;
        return o;
    }
    float sigma0 = resample32<RS>(p.img, p.lines, p.samples, line, sample);
    if (isnan(sigma0)) {
        o.cat = CAT_OUT;
        return o;
    }
    if (p.need_geom) {
        const float he = q.nb.e, hw = q.nb.w, hn = q.nb.n, hs = q.nb.s;
        if (!q.nb.ok) {
            o.mask = 9;
            if (p.flatten) {
                o.cat = CAT_FLAT;
                return o;
            }
        } else {
            // neighbour offsets from the target: FP64 geodesy, FP32 from here on
            const double ss = p.step_sin, cs = p.step_cos;
            const double soe = q.so * cs + q.co * ss, coe = q.co * cs - q.so * ss;
            const double sow = q.so * cs - q.co * ss, cow = q.co * cs + q.so * ss;
            const double sln = q.sl * cs + q.cl * ss, cln = q.cl * cs - q.sl * ss;
            const double sls = q.sl * cs - q.cl * ss, cls = q.cl * cs + q.sl * ss;
            const f3 de = to_f3(ecef_from_n(q.n, q.sl, q.cl, soe, coe, he) - q.target);
            const f3 dw = to_f3(ecef_from_n(q.n, q.sl, q.cl, sow, cow, hw) - q.target);
            const f3 dn = to_f3(ecef_from_sc(sln, cln, q.so, q.co, hn) - q.target);
            const f3 ds = to_f3(ecef_from_sc(sls, cls, q.so, q.co, hs) - q.target);
            const f3 tn = cross3(de - dw, dn - ds);
            const f3 lk = sat - T;
            const float ln = sqrtf(dot3(lk, lk));
            const float lp = ln > 0.0f ? dot3(tn, lk) / ln : 0.0f;
            const float ta = sqrtf(dot3(tn, tn));
            int g;
            float w = 0.0f, cl = 0.0f;
            if (lp <= 0.0f || ta == 0.0f) {
                g = TG_SHADOW;
            } else {
                cl = fminf(fmaxf(lp / ta, 0.0f), 1.0f);
                w = lp / (float)flat_area;
                g = w < 0.05f ? TG_FORESHORTENING : (w > 2.0f ? TG_LAYOVER : TG_VALID);
            }
            if (g == TG_VALID) {
                o.cos_lia = cl;
                if (p.flatten) sigma0 /= w;
            } else if (g == TG_FORESHORTENING || g == TG_LAYOVER) {
                o.cos_lia = cl;
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
        const float nz = resample32<0>(p.nesz, p.lines, p.samples, line, sample);
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

// Per AOI row: FP64 zero-Doppler roots of the row's first and last columns
// (from the row seed), stored as seconds from the reference time; the FP32
// pass seeds each column by linear interpolation between them, which puts
// it within about a millisecond of its root.  Rows whose end pixels have no
// DEM sample keep the row seed at both ends.
template <bool UTM>
__global__ __launch_bounds__(256) void tc32_seed_kernel(TcParams p, float* ends) {
    const int64_t rr = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (rr >= p.rows) return;
    const int64_t row = p.row0 + rr;
    const double row_y = p.lat_north - ((double)row + 0.5) * p.spacing;
    const int npts = p.anchor_stride;
    const double* alon = UTM ? p.anchors + rr * 2 * npts : nullptr;
    const double* alat = UTM ? alon + npts : nullptr;
    const double t_row = (double)(p.row_seed[rr] - p.t_ref_us) * 1e-6;
    double e[2] = {t_row, t_row};
    for (int k = 0; k < 2; ++k) {
        const PixelPos q = prep_pixel<UTM, false>(p, p.col0 + (k ? p.cols - 1 : 0), row_y, alon, alat);
        if (!q.dem_ok) continue;
        OrbitCache cache = orbit_cache_empty();
        int64_t t = 0;
        int it = 0;
        d3 sat;
        bool tie = false;  // a seed only: the kernel's own decision is good enough
        if (zero_doppler<false>(*p.orbit, q.target, p.row_seed[rr], 40, p.tol, kTieBandUs, &t, &it, &sat, &cache,
                                &tie, nullptr) == ZD_CONVERGED)
            e[k] = (double)(t - p.t_ref_us) * 1e-6;
    }
    ends[2 * rr] = (float)e[0];
    ends[2 * rr + 1] = (float)e[1];
}

// One lane per output column; a workgroup takes kRows32 rows of 256
// columns (the orbit copy into LDS and the argument loads are per workgroup).
constexpr int kRows32 = 16;
template <int RS, bool UTM>
__global__ __launch_bounds__(256) void tc32_kernel(TcParams p, const float* ends) {
    __shared__ OrbitF32 orb;
    {
        const uint4* src = reinterpret_cast<const uint4*>(p.orbit32);
        uint4* dst = reinterpret_cast<uint4*>(&orb);
        for (int i = threadIdx.x; i < (int)(sizeof(OrbitF32) / sizeof(uint4)); i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const int64_t cc0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool ok_lane = cc0 < p.cols;
    const int64_t cc = ok_lane ? cc0 : p.cols - 1;  // every lane reaches the wave-wide steps
    LaneCounts C;
    zero_counts(C);
    const int64_t r_end = min((int64_t)(blockIdx.y + 1) * kRows32, p.rows);
    const float frac = p.cols > 1 ? (float)cc / (float)(p.cols - 1) : 0.0f;
    float prev_t = 0.0f, pe0 = 0.0f, pe1 = 0.0f;
    bool have_prev = false;
    for (int64_t rr = (int64_t)blockIdx.y * kRows32; rr < r_end; ++rr) {
        const int64_t row = p.row0 + rr;
        const double row_y = p.lat_north - ((double)row + 0.5) * p.spacing;
        const int npts = p.anchor_stride;
        const double* alon = UTM ? p.anchors + rr * 2 * npts : nullptr;
        const double* alat = UTM ? alon + npts : nullptr;
        const PixelPos q = prep_pixel<UTM>(p, p.col0 + cc, row_y, alon, alat);
        const f3 T = to_f3(q.target - d3{p.centre[0], p.centre[1], p.centre[2]});
        // seed: this column's root in the previous row moved by the change of
        // the row's end roots, else linear between this row's end roots
        const float e0 = ends[2 * rr], e1 = ends[2 * rr + 1];
        const float seed = have_prev ? prev_t + (e0 - pe0) + ((e1 - pe1) - (e0 - pe0)) * frac
                                     : e0 + (e1 - e0) * frac;
        pe0 = e0;
        pe1 = e1;
        have_prev = false;
        if (ok_lane) {
            const int64_t o = rr * p.cols + cc;
            PixelOut v;
            if (!q.dem_ok) {
                v.sigma0 = v.cos_lia = __builtin_nanf("");
                v.line = v.sample = __builtin_nan("");
                v.mask = 2;
                C.c[1] += 1;
            } else {
                float t = 0.0f;
                int it = 0;
                f3 sat = {0, 0, 0};
                const int oc = zero_doppler32(orb, T, seed, p.max_iter, p.tol32, &t, &it, &sat);
                if (oc == ZD_CONVERGED) {
                    prev_t = t;
                    have_prev = true;
                }
                const double flat_area = p.need_geom ? p.row_flat_area[rr] : 0.0;
                v = finish32<RS>(p, q, T, oc, t, sat, flat_area);
                if (v.cat == CAT_ORBIT) {
                    if (oc == ZD_ORBIT_OUTSIDE) {
                        C.err_outside += 1;
                        note_orbit_outside(p, *p.orbit, p.t_ref_us + (int64_t)llrint((double)t * 1e6));
                    } else {
                        C.err_short += 1;
                    }
                } else {
                    count(C, v.cat, it, 1, p.counters);
                }
            }
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
    }
    flush_counts(p, C, true);
}

}  // namespace sd

// The FP32 passes (the caller launches the row and anchor passes first);
// `ends` holds 2 floats per AOI row.
extern "C" hipError_t sd_launch_tc_fp32(const sd::TcParams* p, float* ends, hipStream_t stream) {
    const bool utm = p->crs != 0;
    const dim3 rg((unsigned)((p->rows + 255) / 256));
    if (utm) hipLaunchKernelGGL((sd::tc32_seed_kernel<true>), rg, dim3(256), 0, stream, *p, ends);
    else hipLaunchKernelGGL((sd::tc32_seed_kernel<false>), rg, dim3(256), 0, stream, *p, ends);
    const dim3 grid((unsigned)((p->cols + 255) / 256), (unsigned)((p->rows + sd::kRows32 - 1) / sd::kRows32));
    switch (p->resampling * 2 + (utm ? 1 : 0)) {
        case 0: hipLaunchKernelGGL((sd::tc32_kernel<0, false>), grid, dim3(256), 0, stream, *p, ends); break;
        case 1: hipLaunchKernelGGL((sd::tc32_kernel<0, true>), grid, dim3(256), 0, stream, *p, ends); break;
        case 2: hipLaunchKernelGGL((sd::tc32_kernel<1, false>), grid, dim3(256), 0, stream, *p, ends); break;
        case 3: hipLaunchKernelGGL((sd::tc32_kernel<1, true>), grid, dim3(256), 0, stream, *p, ends); break;
        case 4: hipLaunchKernelGGL((sd::tc32_kernel<2, false>), grid, dim3(256), 0, stream, *p, ends); break;
        default: hipLaunchKernelGGL((sd::tc32_kernel<2, true>), grid, dim3(256), 0, stream, *p, ends); break;
    }
    return hipGetLastError();
}
