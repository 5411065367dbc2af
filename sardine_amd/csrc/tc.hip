// tc.hip — backward Range-Doppler terrain correction on gfx950 (FP64 path).
//
// One workgroup of 256 threads (4 wave64) geocodes one output row, walking
// the row in chunks of 256 columns.  Per pixel: DEM bilinear → geoid → ECEF →
// zero-Doppler Newton on the 8-point Lagrange orbit → slant range → radar
// (line, sample) → resample → optional Small-2011 geometry / flattening →
// NESZ mask → σ⁰ / mask / cos_lia writes; counters and the Newton histogram
// are reduced per wave and added to global counters.
// Reference: sardine/src/terrain_correction.rs:1245-1836.
//
// The per-row seed chain.  The reference seeds each pixel's Newton solve
// with the previous converged pixel's time in the same row
// (terrain_correction.rs:1443,1580) and truncates every step to whole µs, so
// a converged time is floor or ceil of the root depending on the side the
// seed comes from.  Lanes cannot run that chain sequentially, so each chunk:
//   1. solves every pixel speculatively from an extrapolated seed;
//   2. re-solves every pixel with the reference's exact arithmetic, seeded by
//      the speculative time of its nearest converged left neighbour (or the
//      row's carried seed);
//   3. verifies that every seed equals the neighbour's re-solved time and
//      re-solves the lanes whose seed differed, until none does.
// Step 3 makes the result equal to the sequential chain by induction along
// the row (the leftmost mismatching lane always becomes exact), so the
// output does not depend on the speculation, only its cost does.
#include "geom.hpp"
#include "tc_params.hpp"

namespace sd {

__device__ __forceinline__ bool dem_sample(const TcParams& p, double lat, double lon, float* h) {
    // dem.rs:694-723 lon wrap, GLO-30 tile rule dem.rs:351-372, bilinear_on_grid dem.rs:385-409
    double lon_n = lon >= 180.0 ? lon - 360.0 : lon;
    double row_f = (p.dem_olat - lat) / p.dem_pix;
    double col_f = (lon_n - p.dem_olon) / p.dem_pix;
    double max_row = (double)(p.dem_h - 1), max_col = (double)(p.dem_w - 1);
    if (row_f < -0.5 || row_f > max_row + 0.5 || col_f < -0.5 || col_f > max_col + 0.5) return false;
    row_f = fmin(fmax(row_f, 0.0), max_row);
    col_f = fmin(fmax(col_f, 0.0), max_col);
    int64_t max_rc = p.dem_h - 1;
    int64_t r0 = (int64_t)row_f;
    int64_t lim_r = max_rc > 0 ? max_rc - 1 : 0;
    r0 = r0 < lim_r ? r0 : lim_r;
    int64_t c0 = (int64_t)col_f;
    int64_t lim_c = p.dem_w >= 2 ? p.dem_w - 2 : 0;
    c0 = c0 < lim_c ? c0 : lim_c;
    double dr = row_f - (double)r0, dc = col_f - (double)c0;
    const float* d = p.dem + r0 * p.dem_w + c0;
    float v00 = d[0], v01 = d[1], v10 = d[p.dem_w], v11 = d[p.dem_w + 1];
    if (isnan(v00) || isnan(v01) || isnan(v10) || isnan(v11)) return false;
    double v = (double)v00 * (1.0 - dr) * (1.0 - dc) + (double)v01 * (1.0 - dr) * dc +
               (double)v10 * dr * (1.0 - dc) + (double)v11 * dr * dc;
    *h = (float)v;
    return true;
}

__device__ __forceinline__ double geoid_n(const TcParams& p, double lat, double lon) {
    // geoid.rs:220-251 (2.5° grid, lon wrap, lat clamp)
    if (!p.geoid) return 0.0;
    double l = fmod(lon, 360.0);
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

__device__ __forceinline__ double lanczos_w(double x) {  // terrain_correction.rs:938-950
    double ax = fabs(x);
    if (ax < 1e-10) return 1.0;
    if (ax < 3.0) {
        double px = kPi * x;
        double pxa = kPi * x / 3.0;
        return sin(px) * sin(pxa) / (px * pxa);
    }
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
__device__ __forceinline__ float bicubic(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    int64_t lf = (int64_t)floor(line), sf = (int64_t)floor(sample);
    double dl = line - (double)lf, ds = sample - (double)sf;
    double ws[4];
#pragma unroll
    for (int si = 0; si < 4; ++si) ws[si] = keys_w((double)(si - 1) - ds);
    double acc = 0.0;
#pragma unroll
    for (int li = -1; li <= 2; ++li) {
        const float* row = d + clampi(lf + li, 0, nl - 1) * ns;
        double wl = keys_w((double)li - dl);
#pragma unroll
        for (int si = -1; si <= 2; ++si) {
            float v = row[clampi(sf + si, 0, ns - 1)];
            if (isnan(v)) return __builtin_nanf("");
            acc += (double)v * wl * ws[si + 1];
        }
    }
    return (float)acc;
}

// terrain_correction.rs:958-983
__device__ __forceinline__ float lanczos3(const float* d, int64_t nl, int64_t ns, double line, double sample) {
    int64_t lf = (int64_t)floor(line), sf = (int64_t)floor(sample);
    double ws[6];
#pragma unroll
    for (int si = 0; si < 6; ++si) ws[si] = lanczos_w(sample - (double)(sf + si - 2));
    double acc = 0.0;
    for (int li = -2; li <= 3; ++li) {
        const float* row = d + clampi(lf + li, 0, nl - 1) * ns;
        double wl = lanczos_w(line - (double)(lf + li));
#pragma unroll
        for (int si = -2; si <= 3; ++si) {
            float v = row[clampi(sf + si, 0, ns - 1)];
            if (isnan(v)) return __builtin_nanf("");
            acc += (double)v * wl * ws[si + 2];
        }
    }
    return (float)acc;
}

__device__ __forceinline__ float resample(const TcParams& p, double line, double sample) {
    if (p.resampling == 1) return bicubic(p.img, p.lines, p.samples, line, sample);
    if (p.resampling == 2) return lanczos3(p.img, p.lines, p.samples, line, sample);
    return bilinear(p.img, p.lines, p.samples, line, sample);
}

// Value of the nearest lane before this thread (in thread order, across the
// 256-thread workgroup) whose flag is set, else `fallback`.  Every thread of
// the workgroup must call it (it contains barriers).
__device__ __forceinline__ long long excl_last(bool flag, long long val, long long fallback, long long* sh_val,
                                                int* sh_has) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long m = __ballot(flag);
    unsigned long long below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
    int src = below ? 63 - __clzll(below) : lane;
    long long v = __shfl(val, src);
    int last = m ? 63 - __clzll(m) : 0;
    long long wl = __shfl(val, last);
    __syncthreads();  // previous users of sh_* are done
    if (lane == 0) {
        sh_has[w] = m != 0ull;
        sh_val[w] = wl;
    }
    __syncthreads();
    if (below) return v;
    for (int q = w - 1; q >= 0; --q)
        if (sh_has[q]) return sh_val[q];
    return fallback;
}

// (value, column) of the last flagged lane of the whole workgroup, if any.
__device__ __forceinline__ bool wg_last(bool flag, long long val, long long* out_val, int* out_idx,
                                        long long* sh_val, int* sh_has) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long m = __ballot(flag);
    int last = m ? 63 - __clzll(m) : 0;
    long long wl = __shfl(val, last);
    __syncthreads();
    if (lane == 0) {
        sh_has[w] = m != 0ull ? (w * 64 + last + 1) : 0;
        sh_val[w] = wl;
    }
    __syncthreads();
    for (int q = 3; q >= 0; --q)
        if (sh_has[q]) {
            *out_val = sh_val[q];
            *out_idx = sh_has[q] - 1;
            return true;
        }
    return false;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(256) void tc_fp64_kernel(TcParams p) {
    __shared__ long long sh_val[4];
    __shared__ int sh_has[4];
    __shared__ double sh_alon[kMaxAnchorPts];
    __shared__ double sh_alat[kMaxAnchorPts];

    const int tid = threadIdx.x;
    const int64_t rr = blockIdx.x;       // AOI row
    const int64_t row = p.row0 + rr;     // full-grid row
    const double row_y = p.lat_north - ((double)row + 0.5) * p.spacing;
    const OrbitDev& orb = *p.orbit;

    // UTM: the row's n_anchors+1 exact inverse projections (output_crs.rs:364-417)
    if (p.crs != 0) {
        TmSeries ts = tm_series();
        for (int i = tid; i < p.n_anchor_pts; i += 256) {
            double xi = p.x_start + (double)i * p.anchor_dx;
            utm_inverse(ts, p.utm_zone, p.utm_south != 0, xi, row_y, &sh_alon[i], &sh_alat[i]);
        }
        __syncthreads();
    }

    long long carry = p.row_seed[rr];  // prev_az_time (terrain_correction.rs:1443)
    long long carry_col = -1;
    double slope = 0.0;  // µs per column from the previous chunk
    const double flat_area = p.need_geom ? p.row_flat_area[rr] : 0.0;

    unsigned cnt_valid = 0, cnt_dem = 0, cnt_out = 0, cnt_nc = 0, cnt_deg = 0, cnt_flat = 0, cnt_noise = 0;
    unsigned hist[kHistSize];
#pragma unroll
    for (int i = 0; i < kHistSize; ++i) hist[i] = 0;
    unsigned redo_rounds = 0;
    int err = 0;

    for (int64_t c0 = 0; c0 < p.cols; c0 += 256) {
        const int64_t cc = c0 + tid;
        const bool active = cc < p.cols;
        const int64_t col = p.col0 + cc;

        // ── pixel position, DEM, target ──
        double lat = 0.0, lon = 0.0;
        if (p.crs == 0) {
            lon = p.lon_west + ((double)col + 0.5) * p.spacing;
            lat = row_y;
        } else if (active) {
            double x = p.x_start + (double)col * p.spacing;
            int na = p.n_anchor_pts - 1;
            int seg = (int)floor((x - p.x_start) / p.anchor_dx);
            seg = seg < 0 ? 0 : (seg > na - 1 ? na - 1 : seg);
            // the reference's walk: smallest seg with x <= ax[seg+1], capped
            while (seg > 0 && !(x > p.x_start + (double)seg * p.anchor_dx)) --seg;
            while (seg + 1 < na && x > p.x_start + (double)(seg + 1) * p.anchor_dx) ++seg;
            double x0 = p.x_start + (double)seg * p.anchor_dx, x1 = p.x_start + (double)(seg + 1) * p.anchor_dx;
            double t = (x1 == x0) ? 0.0 : (x - x0) / (x1 - x0);
            lon = sh_alon[seg] + t * (sh_alon[seg + 1] - sh_alon[seg]);
            lat = sh_alat[seg] + t * (sh_alat[seg + 1] - sh_alat[seg]);
        }
        float hf = 0.0f;
        bool dem_ok = active && dem_sample(p, lat, lon, &hf) && isfinite(hf);
        d3 target = {0, 0, 0};
        if (dem_ok) {
            double h_ell = (double)hf + geoid_n(p, lat, lon);
            target = geodetic_to_ecef(lat, lon, h_ell);
        }

        // ── 1. speculative solve from an extrapolated seed ──
        long long t1 = 0;
        bool c1 = false;
        if (dem_ok) {
            double ext = slope * (double)(cc - carry_col);
            ext = fmin(fmax(ext, -1.0e8), 1.0e8);
            long long seed = carry + (long long)ext;
            int it;
            d3 sat;
            int64_t tt;
            int oc = zero_doppler(orb, target, seed, 24, p.tol, &tt, &it, &sat);
            c1 = oc == ZD_CONVERGED;
            t1 = tt;
        }
        // ── 2. exact re-solve seeded by the left neighbour's speculation ──
        long long s = excl_last(c1, t1, carry, sh_val, sh_has);
        int oc2 = ZD_NOT_CONVERGED, it2 = 0;
        int64_t t2 = 0;
        d3 sat2 = {0, 0, 0};
        if (dem_ok) oc2 = zero_doppler(orb, target, s, p.max_iter, p.tol, &t2, &it2, &sat2);
        // ── 3. verify the chain, re-solve lanes whose seed was wrong ──
        for (int round = 0; round < 257; ++round) {
            bool conv2 = dem_ok && oc2 == ZD_CONVERGED;
            long long e = excl_last(conv2, (long long)t2, carry, sh_val, sh_has);
            bool mism = dem_ok && e != s;
            if (!__syncthreads_or(mism)) break;
            if (mism) {
                s = e;
                oc2 = zero_doppler(orb, target, s, p.max_iter, p.tol, &t2, &it2, &sat2);
            }
            ++redo_rounds;
        }
        {
            bool conv2 = dem_ok && oc2 == ZD_CONVERGED;
            long long lv;
            int li;
            if (wg_last(conv2, (long long)t2, &lv, &li, sh_val, sh_has)) {
                long long new_col = c0 + li;
                if (carry_col >= 0 && new_col > carry_col) slope = (double)(lv - carry) / (double)(new_col - carry_col);
                carry = lv;
                carry_col = new_col;
            }
        }

        // ── finish the pixel (terrain_correction.rs:1529-1739) ──
        if (!active) continue;
        float out = __builtin_nanf("");
        float cl_out = __builtin_nanf("");
        double rc_line = __builtin_nan(""), rc_sample = __builtin_nan("");
        unsigned char mk = 1;  // OUTSIDE_FOOTPRINT default
        if (!dem_ok) {
            ++cnt_dem;
            mk = 2;
        } else if (oc2 == ZD_ORBIT_OUTSIDE || oc2 == ZD_ORBIT_INSUFFICIENT) {
            err = oc2 == ZD_ORBIT_OUTSIDE ? 1 : 2;
        } else if (oc2 == ZD_NOT_CONVERGED) {
            ++cnt_nc;
            mk = 3;
        } else if (oc2 == ZD_DEGENERATE) {
            ++cnt_deg;
            mk = 4;
        } else {
            hist[it2 < kHistSize - 1 ? it2 : kHistSize - 1]++;
            d3 rv = target - sat2;
            double slant = sqrt(dot(rv, rv));
            double line = (double)(t2 - p.az_start_us) / 1e6 / p.ati;
            double sample = (slant - p.near_range_m) / p.rps;
            rc_line = line;
            rc_sample = sample;
            bool done = false;
            float sigma0 = 0.0f;
            if (line < -0.5 || line > (double)p.lines - 0.5 || sample < -0.5 || sample > (double)p.samples - 0.5) {
                ++cnt_out;
                done = true;
            } else {
                sigma0 = resample(p, line, sample);
                if (isnan(sigma0)) {
                    ++cnt_out;
                    done = true;
                }
            }
            if (!done && p.need_geom) {
                const double st = p.step_deg;
                float he, hw, hn, hs;
                bool ok = dem_sample(p, lat, lon + st, &he) && dem_sample(p, lat, lon - st, &hw) &&
                          dem_sample(p, lat + st, lon, &hn) && dem_sample(p, lat - st, lon, &hs);
                if (ok) ok = isfinite(he) && isfinite(hw) && isfinite(hn) && isfinite(hs);
                if (!ok) {
                    mk = 9;
                    if (p.flatten) { ++cnt_flat; done = true; }
                } else {
                    double w = 0.0, cl = 0.0;
                    int g = terrain_geometry(lat, lon, he, hw, hn, hs, sat2, target, st, flat_area, &w, &cl);
                    if (g == TG_VALID) {
                        cl_out = (float)cl;
                        if (p.flatten) sigma0 /= (float)w;
                    } else if (g == TG_FORESHORTENING || g == TG_LAYOVER) {
                        cl_out = (float)cl;
                        mk = g == TG_LAYOVER ? 6 : 7;
                        if (p.flatten) { ++cnt_flat; done = true; }
                    } else {
                        mk = 5;
                        if (p.flatten) { ++cnt_flat; done = true; }
                    }
                }
            }
            if (!done && p.nesz) {
                float nz = bilinear(p.nesz, p.lines, p.samples, line, sample);
                if (!isnan(nz) && sigma0 <= nz * p.nesz_factor) {
                    ++cnt_noise;
                    mk = 8;
                    done = true;
                }
            }
            if (!done) {
                out = sigma0;
                mk = 0;
                ++cnt_valid;
            }
        }
        const int64_t o = rr * p.cols + cc;
        p.out[o] = out;
        if (p.compute_lia) {
            p.cos_lia[o] = cl_out;
            p.mask[o] = mk;
        }
        if (p.radar_coords) {
            p.radar_coords[2 * o] = rc_line;
            p.radar_coords[2 * o + 1] = rc_sample;
        }
    }

    // ── counters (wave reduction, one atomic per counter per wave) ──
    unsigned long long* C = p.counters;
    unsigned v;
    const int lane = tid & 63;
#define SD_FLUSH(idx, x)                                   \
    v = wave_sum(x);                                       \
    if (lane == 0 && v) atomicAdd(&C[idx], (unsigned long long)v);
    SD_FLUSH(0, cnt_valid)
    SD_FLUSH(1, cnt_dem)
    SD_FLUSH(2, cnt_out)
    SD_FLUSH(3, cnt_nc)
    SD_FLUSH(4, cnt_deg)
    SD_FLUSH(5, cnt_flat)
    SD_FLUSH(6, cnt_noise)
#pragma unroll
    for (int i = 0; i < kHistSize; ++i) { SD_FLUSH(7 + i, hist[i]) }
#undef SD_FLUSH
    if (tid == 0 && redo_rounds) atomicAdd(&C[kCntRedo], (unsigned long long)redo_rounds);
    if (err) atomicOr(p.error_flag, err);
}

}  // namespace sd

extern "C" hipError_t sd_launch_tc_fp64(const sd::TcParams* p, hipStream_t stream) {
    hipLaunchKernelGGL(sd::tc_fp64_kernel, dim3((unsigned)p->rows), dim3(256), 0, stream, *p);
    return hipGetLastError();
}
