// speckle.hip — the five SpeckleKernel filters on gfx950.
//
// Reference: sardine/src/speckle.rs:131-603 (apply_speckle_filter,
// window_mean_var, boxcar, lee, frost, gamma_map, refined_lee).
//
// Each 256-thread workgroup filters a 64 x 16 output tile.  The input tile
// plus a `half`-pixel halo is staged once in LDS; samples outside the raster
// are stored as NaN, which the reference's "finite samples only" rule skips
// exactly like out-of-bounds samples.  Window sums are accumulated in f64 in
// the reference's raster order (dy outer, dx inner), so mean/variance are
// bit-identical to the CPU path; the filter formulas run in f32 as there.
// Refined Lee has two forms: the tiled one (batched filtering) accumulates
// all eight directional sub-windows in one pass over the 7 x 7
// neighbourhood with compile-time membership masks, each direction seeing
// its samples in raster order; the column walk (one raster per call) shares
// partial sums between neighbouring pixels and hands the pixels it cannot
// decide as the reference would to an exact path.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace sd {

constexpr int kTileW = 64;
constexpr int kTileH = 16;
constexpr int kRowsPerThread = kTileH / 4;

enum { SK_BOXCAR = 0, SK_LEE = 1, SK_FROST = 2, SK_GAMMA = 3, SK_RLEE = 4 };

struct SpeckleParams {
    const float* in;
    float* out;
    int64_t cols, rows;
    int half;
    float enl;
    float damping;
};

// Stage the (kTileH + 2h) x (kTileW + 2h) input tile; pitch = kTileW + 2h.
__device__ __forceinline__ void stage_tile(const SpeckleParams& p, float* tile, int half, int64_t r0, int64_t c0) {
    const int tw = kTileW + 2 * half, th = kTileH + 2 * half;
    const float qnan = __builtin_nanf("");
    for (int i = threadIdx.x; i < tw * th; i += 256) {
        int ty = i / tw, tx = i - ty * tw;
        int64_t r = r0 - half + ty, c = c0 - half + tx;
        float v = qnan;
        if (r >= 0 && r < p.rows && c >= 0 && c < p.cols) v = p.in[r * p.cols + c];
        tile[i] = v;
    }
}

// Pair tiles: each staged sample as {x, x²} in f64 (NaN outside the
// raster), so a window tap is one 16-byte LDS read and no pixel repeats a
// neighbour's conversion or square.  Used by Refined Lee and by the filters
// that need the window variance (Lee, Frost, Gamma-MAP) with compile-time
// windows up to 7 x 7: 5-10 % faster there.  Boxcar (the mean only) and
// larger windows keep the f32 tile; for Boxcar the pairs' fourfold LDS
// bytes cost 40 %.
constexpr int kMaxPairHalf = 3;
constexpr bool pair_tile(int kind, int half) {
    return kind == SK_RLEE || (kind != SK_BOXCAR && half >= 1 && half <= kMaxPairHalf);
}
constexpr size_t pair_tile_bytes(int half) {
    return (size_t)(kTileW + 2 * half) * (kTileH + 2 * half) * sizeof(double2);
}
constexpr size_t kRlTileBytes = pair_tile_bytes(3);

__device__ __forceinline__ void stage_tile_pairs(const SpeckleParams& p, double2* tile, int half, int64_t r0,
                                                 int64_t c0) {
    const int tw = kTileW + 2 * half, th = kTileH + 2 * half;
    const double qnan = __builtin_nan("");
    for (int i = threadIdx.x; i < tw * th; i += 256) {
        int ty = i / tw, tx = i - ty * tw;
        int64_t r = r0 - half + ty, c = c0 - half + tx;
        double x = qnan;
        if (r >= 0 && r < p.rows && c >= 0 && c < p.cols) x = (double)p.in[r * p.cols + c];
        tile[i] = double2{x, x * x};
    }
}

// A zero the compiler cannot see through.  Taps read again at t + this
// (the skipping pass, Frost's weighting pass) are reloaded from LDS instead
// of the first pass's loads being kept live in registers across the branch.
__device__ __forceinline__ int opaque_zero() {
    int z = 0;
    asm volatile("" : "+s"(z));
    return z;
}

// A window tap as {x, x²} in f64, from either tile.
__device__ __forceinline__ double2 tap(const float* t, int i) {
    const double x = (double)t[i];
    return double2{x, x * x};
}
__device__ __forceinline__ double2 tap(const double2* t, int i) { return t[i]; }

template <int D>
__device__ __forceinline__ double div_const(double x);

// Mean and population variance of the finite window samples (speckle.rs:218-241).
// With a compile-time window the first pass assumes every sample is finite
// (then the count is the full window and the division is by a constant);
// a non-finite sample makes the sum non-finite and the window is redone
// with the skipping rule.  Both passes add in the reference's raster order.
template <int HALF, bool NEED_VAR, typename T>
__device__ __forceinline__ bool win_stats(const T* t, int pitch, int half_rt, float* mean, float* var) {
    const int half = HALF > 0 ? HALF : half_rt;
    double sum = 0.0, sq = 0.0;
    if constexpr (HALF > 0) {
#pragma unroll
        for (int dy = -HALF; dy <= HALF; ++dy)
#pragma unroll
            for (int dx = -HALF; dx <= HALF; ++dx) {
                const double2 e = tap(t, dy * pitch + dx);
                sum += e.x;
                sq += e.y;
                if (dx == HALF) __builtin_amdgcn_sched_barrier(0);  // one window row at a time
            }
        // finite exactly when every sample is (f64 sums of f32 values and
        // their squares cannot overflow); where the variance is used, testing
        // sq too keeps its sum in this pass instead of sunk past the branch
        // with every square live
        if (isfinite(NEED_VAR ? sum + sq : sum)) {
            constexpr int N = (2 * HALF + 1) * (2 * HALF + 1);
            const double m = div_const<N>(sum);
            double va = div_const<N>(sq) - m * m;
            va = va > 0.0 ? va : 0.0;
            *mean = (float)m;
            *var = (float)va;
            return true;
        }
        sum = 0.0;
        sq = 0.0;
    }
    int n = 0;
#pragma unroll
    for (int dy = -half; dy <= half; ++dy)
#pragma unroll
        for (int dx = -half; dx <= half; ++dx) {
            if constexpr (sizeof(T) == sizeof(float)) {
                const float v = t[dy * pitch + dx];
                if (isfinite(v)) {
                    const double x = (double)v;
                    sum += x;
                    sq += x * x;
                    ++n;
                }
            } else {
                const double2 e = t[dy * pitch + dx];
                if (isfinite(e.x)) {
                    sum += e.x;
                    sq += e.y;
                    ++n;
                }
            }
        }
    if (n == 0) return false;
    double m = sum / (double)n;
    double va = sq / (double)n - m * m;
    va = va > 0.0 ? va : 0.0;
    *mean = (float)m;
    *var = (float)va;
    return true;
}

template <int KIND, int HALF, typename T>
__device__ __forceinline__ float filter_pixel(const SpeckleParams& p, const T* t, int pitch, float centre) {
    const int half = HALF > 0 ? HALF : p.half;
    const float qnan = __builtin_nanf("");
    float mean, var;
    if (!win_stats<HALF, KIND != SK_BOXCAR>(t, pitch, half, &mean, &var)) return qnan;
    if (KIND == SK_BOXCAR) return mean;  // speckle.rs:247-263
    if (KIND == SK_LEE) {                // speckle.rs:269-298
        float cv_n_sq = 1.0f / p.enl;
        float noise_var = mean * mean * cv_n_sq;
        float signal_var = fmaxf(var - noise_var, 0.0f);
        float denom = signal_var + noise_var;
        float k = denom > 0.0f ? signal_var / denom : 0.0f;
        return mean + k * (centre - mean);
    }
    if (KIND == SK_FROST) {  // speckle.rs:304-342
        float cv_sq = mean > 0.0f ? var / (mean * mean) : 0.0f;
        float alpha = p.damping * cv_sq;
        double wd[2 * 31 + 1];
        for (int d = 0; d <= 2 * half; ++d) wd[d] = (double)expf(-alpha * (float)d);
        double wsum = 0.0, vw = 0.0;
        const T* tw = t + opaque_zero();
#pragma unroll
        for (int dy = -half; dy <= half; ++dy)
#pragma unroll
            for (int dx = -half; dx <= half; ++dx) {
                const double x = tap(tw, dy * pitch + dx).x;
                if (isfinite(x)) {
                    double w = wd[abs(dy) + abs(dx)];
                    wsum += w;
                    vw += w * x;
                }
            }
        return wsum > 0.0 ? (float)(vw / wsum) : qnan;
    }
    // Gamma-MAP, speckle.rs:348-411
    float enl = p.enl;
    float cv_n_sq = 1.0f / enl;
    float cv_max_sq = cv_n_sq < 1.0f ? 2.0f * cv_n_sq / (1.0f - cv_n_sq) : __builtin_inff();
    float cv_sq = mean > 0.0f ? var / (mean * mean) : 0.0f;
    if (cv_sq <= cv_n_sq) return mean;
    if (cv_sq >= cv_max_sq) return centre;
    float alpha = (1.0f + cv_n_sq) / (cv_sq - cv_n_sq);
    float b = (alpha - enl - 1.0f) * mean;
    float disc = b * b + 4.0f * alpha * enl * mean * centre;
    return disc >= 0.0f ? ((enl - alpha + 1.0f) * mean + sqrtf(disc)) / (2.0f * enl) : mean;
}

// refined_lee_in_mask (speckle.rs:432-457) as an 8-bit direction set per offset.
__host__ __device__ constexpr unsigned rl_dirs(int dy, int dx) {
    if (dy == 0 && dx == 0) return 0u;
    unsigned m = 0;
    if (dy >= 1 && dy <= 3) m |= 1u << 0;                                           // S
    if (dy >= -3 && dy <= -1) m |= 1u << 1;                                         // N
    if (dx >= 1 && dx <= 3) m |= 1u << 2;                                           // E
    if (dx >= -3 && dx <= -1) m |= 1u << 3;                                         // W
    if ((dy - dx <= 1 && dx - dy <= 1) && dy + dx >= 1) m |= 1u << 4;               // SE
    if ((dy - dx <= 1 && dx - dy <= 1) && dy + dx <= -1) m |= 1u << 5;              // NW
    if ((dy + dx <= 1 && -(dy + dx) <= 1) && dx - dy >= 1) m |= 1u << 6;            // NE
    if ((dy + dx <= 1 && -(dy + dx) <= 1) && dx - dy <= -1) m |= 1u << 7;           // SW
    return m;
}

// The directions whose first member in raster order (dy outer, dx inner)
// is (dy, dx): there a direction's sums start, so its accumulators take the
// member's values instead of adding them to 0.0.  (0.0 + x differs from x
// only for x = -0.0, which changes nothing observable: a direction whose
// sum is a zero has mean 0 and is never picked, speckle.rs:546.)
struct RlFirstTable {
    unsigned m[7][7];
};
constexpr RlFirstTable rl_first_table() {
    RlFirstTable t{};
    unsigned seen = 0;
    for (int y = -3; y <= 3; ++y)
        for (int x = -3; x <= 3; ++x) {
            t.m[y + 3][x + 3] = rl_dirs(y, x) & ~seen;
            seen |= rl_dirs(y, x);
        }
    return t;
}
constexpr RlFirstTable kRlFirst = rl_first_table();

// x / D for a constant window count D, correctly rounded: q0 = x·RN(1/D)
// is within an ulp of x/D and one FMA residual step (Markstein 1990, the
// closing step of the IEEE division sequence) rounds it correctly.  Checked
// bit-exact against x / D over 2.4e9 operands for D = 9 and 21.
template <int D>
__device__ __forceinline__ double div_const(double x) {
    constexpr double r = 1.0 / (double)D;
    const double q0 = x * r;
    const double e = __builtin_fma(-q0, (double)D, x);
    return __builtin_fma(e, r, q0);
}

// Lee update with the most homogeneous direction (speckle.rs:536-569).
__device__ __forceinline__ void rl_pick(double m, double va, float* best_cv, float* best_mean, float* best_var) {
    va = va > 0.0 ? va : 0.0;
    const float mf = (float)m, vf = (float)va;
    if (!(mf > 0.0f)) return;
    const float cv = vf / (mf * mf);
    if (cv < *best_cv) {
        *best_cv = cv;
        *best_mean = mf;
        *best_var = vf;
    }
}

__device__ __forceinline__ float rl_update(float best_cv, float best_mean, float best_var, float centre,
                                           float cv_n_sq) {
    if (!isfinite(best_cv)) return centre;
    float noise_var = best_mean * best_mean * cv_n_sq;
    float signal_var = fmaxf(best_var - noise_var, 0.0f);
    float denom = signal_var + noise_var;
    float k = denom > 0.0f ? signal_var / denom : 0.0f;
    return best_mean + k * (centre - best_mean);
}

// Refined Lee with the reference's skipping rule, exactly as
// speckle.rs:465-569 forms it: each direction sums its finite members in
// raster order (dy outer, dx inner), divides by its own count, and the pick
// divides var by mean² in f32; the first direction with the smallest CV²
// wins.  tap(dy, dx) returns the sample as f64, NaN outside the raster.
template <typename Tap>
__device__ __forceinline__ float rl_skipping(float centre, float cv_n_sq, Tap tap) {
    double sum[8], sq[8];
    int n[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) { sum[d] = 0.0; sq[d] = 0.0; n[d] = 0; }
#pragma unroll
    for (int dy = -3; dy <= 3; ++dy)
#pragma unroll
        for (int dx = -3; dx <= 3; ++dx) {
            const unsigned m = rl_dirs(dy, dx);
            if (m == 0u) continue;
            const double x = tap(dy, dx);
            if (!isfinite(x)) continue;
            const double xx = x * x;
#pragma unroll
            for (int d = 0; d < 8; ++d)
                if (m & (1u << d)) {
                    sum[d] += x;
                    sq[d] += xx;
                    ++n[d];
                }
        }
    float best_cv = __builtin_inff(), best_mean = centre, best_var = 0.0f;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        if (n[d] == 0) continue;
        const double m = sum[d] / (double)n[d];
        rl_pick(m, sq[d] / (double)n[d] - m * m, &best_cv, &best_mean, &best_var);
    }
    return rl_update(best_cv, best_mean, best_var, centre, cv_n_sq);
}

// Refined Lee, speckle.rs:522-569.  Each direction's sums run over its
// members in raster order.  The first pass assumes all 48 neighbours are
// finite: then no sample is skipped, every count is the full 21 (strips) or
// 9 (diagonal bands), and the sums are the reference's.  A non-finite
// neighbour makes some direction's sum non-finite (every neighbour belongs
// to at least one direction) and the pixel is redone with the skipping rule.
__device__ __forceinline__ float refined_lee_pixel(const double2* t, int pitch, float centre, float cv_n_sq) {
    double sum[8], sq[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) { sum[d] = 0.0; sq[d] = 0.0; }
#pragma unroll
    for (int dy = -3; dy <= 3; ++dy)
#pragma unroll
        for (int dx = -3; dx <= 3; ++dx) {
            const unsigned m = rl_dirs(dy, dx);
            if (m == 0u) continue;
            const unsigned first = kRlFirst.m[dy + 3][dx + 3];
            const double2 e = t[dy * pitch + dx];
            const double x = e.x, xx = e.y;
#pragma unroll
            for (int d = 0; d < 8; ++d)
                if (m & (1u << d)) {
                    if (first & (1u << d)) {
                        sum[d] = x;
                        sq[d] = xx;
                    } else {
                        sum[d] += x;
                        sq[d] += xx;
                    }
                }
            // one window row at a time: keeps the scheduler from hoisting
            // all 48 LDS reads (and their registers) to the top
            if (dx == 3) __builtin_amdgcn_sched_barrier(0);
        }
    // finite exactly when every neighbour is (f64 sums of f32 values and
    // their squares cannot overflow); the squares' sums are part of the test
    // so that they are formed in this pass, not kept as 48 live squares
    double all = 0.0;
#pragma unroll
    for (int d = 0; d < 8; ++d) all += sum[d] + sq[d];
    float best_cv = __builtin_inff(), best_mean = centre, best_var = 0.0f;
    if (isfinite(all)) {
        // Every direction's mean and variance, and its CV² = var / mean² from
        // a reciprocal (within 2.4e-7 relative of the correctly rounded
        // quotient).  These decide the pick unless another candidate is
        // within 2^-19 relative of the smallest (more than twice the error
        // of both), or a mean² or its reciprocal is not a normal number;
        // then every CV² is recomputed by exact division, which resolves
        // ties to the first direction as the reference does.  (A quotient
        // below 2^-126 needs var < mean²·2^-126, and var is either 0 or at
        // least the ~1e-16·mean² rounding residue of sq/n - mean².)
        float mfv[8], vfv[8], cva[8];
        bool okv[8], special = false;
#pragma unroll
        for (int d = 0; d < 8; ++d) {  // S, N, E, W strips: 21 members; diagonal bands: 9
            const double m = d < 4 ? div_const<21>(sum[d]) : div_const<9>(sum[d]);
            double va = (d < 4 ? div_const<21>(sq[d]) : div_const<9>(sq[d])) - m * m;
            va = va > 0.0 ? va : 0.0;
            const float mf = (float)m, vf = (float)va, m2 = mf * mf;
            mfv[d] = mf;
            vfv[d] = vf;
            okv[d] = mf > 0.0f;
            cva[d] = vf * __builtin_amdgcn_rcpf(m2);
            // mean² and its reciprocal normal: biased exponent 1..252 (an
            // infinite or NaN CV² is never picked on either path, and a
            // NaN one fails the near-tie test below)
            special |= okv[d] && (__float_as_uint(m2) >> 23) - 1u >= 252u;
        }
        int bi = -1;
#pragma unroll
        for (int d = 0; d < 8; ++d)
            if (okv[d] && cva[d] < best_cv) {
                best_cv = cva[d];
                best_mean = mfv[d];
                best_var = vfv[d];
                bi = d;
            }
        const float lim = best_cv * (1.0f + 0x1p-19f);
        bool close = special;
#pragma unroll
        for (int d = 0; d < 8; ++d) close |= okv[d] && d != bi && !(cva[d] > lim);
        if (close) {
            best_cv = __builtin_inff();
            best_mean = centre;
            best_var = 0.0f;
#pragma unroll
            for (int d = 0; d < 8; ++d)
                if (okv[d]) {
                    const float cv = vfv[d] / (mfv[d] * mfv[d]);
                    if (cv < best_cv) {
                        best_cv = cv;
                        best_mean = mfv[d];
                        best_var = vfv[d];
                    }
                }
        }
        return rl_update(best_cv, best_mean, best_var, centre, cv_n_sq);
    }
    // some neighbour is NaN/inf or outside the raster: skip it per direction
    return rl_skipping(centre, cv_n_sq, [&](int dy, int dx) { return t[dy * pitch + dx].x; });
}

template <int KIND, int HALF>
__device__ __forceinline__ void speckle_tile(const SpeckleParams& p, float* tile, int64_t r0, int64_t c0) {
    const int half = KIND == SK_RLEE ? 3 : (HALF > 0 ? HALF : p.half);
    const int pitch = kTileW + 2 * half;
    constexpr bool kPairs = pair_tile(KIND, HALF);
    double2* pairs = reinterpret_cast<double2*>(tile);
    if constexpr (kPairs) stage_tile_pairs(p, pairs, half, r0, c0);
    else stage_tile(p, tile, half, r0, c0);
    __syncthreads();
    const int tx = threadIdx.x & (kTileW - 1);
    const int ty0 = threadIdx.x >> 6;  // 0..3
    const float cv_n_sq = 1.0f / p.enl;
#pragma unroll 1
    for (int k = 0; k < kRowsPerThread; ++k) {
        const int ty = ty0 + 4 * k;
        const int64_t r = r0 + ty, c = c0 + tx;
        if (r >= p.rows || c >= p.cols) continue;
        float o = __builtin_nanf("");
        if constexpr (kPairs) {
            const double2* t = pairs + (ty + half) * pitch + (tx + half);
            const float centre = (float)t[0].x;  // exact: the sample was an f32
            if (isfinite(centre)) {
                if constexpr (KIND == SK_RLEE) o = refined_lee_pixel(t, pitch, centre, cv_n_sq);
                else o = filter_pixel<KIND, HALF>(p, t, pitch, centre);
            }
        } else {
            const float* t = tile + (ty + half) * pitch + (tx + half);
            const float centre = t[0];
            if (isfinite(centre)) o = filter_pixel<KIND, HALF>(p, t, pitch, centre);
        }
        p.out[r * p.cols + c] = o;
    }
}

template <int KIND, int HALF>
__global__ __launch_bounds__(256) void speckle_kernel(SpeckleParams p) {
    extern __shared__ __attribute__((aligned(16))) float tile[];
    speckle_tile<KIND, HALF>(p, tile, (int64_t)blockIdx.y * kTileH, (int64_t)blockIdx.x * kTileW);
}

// Tiled Refined Lee (batched filtering) keeps 16 f64 accumulators per
// pixel; 4 waves per SIMD fit without spilling.
#ifndef SD_RL_WAVES
#define SD_RL_WAVES 4
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SD_RL_WAVES))) void refined_lee_kernel(
    SpeckleParams p) {
    extern __shared__ __attribute__((aligned(16))) float tile[];
    speckle_tile<SK_RLEE, 3>(p, tile, (int64_t)blockIdx.y * kTileH, (int64_t)blockIdx.x * kTileW);
}

// ── Refined Lee as a column walk ────────────────────────────────────────
//
// One 64-lane wave filters a strip of 64 columns (one per lane) over
// kRlRows output rows, taking one input row per step.  Each step stages the
// next row in LDS, reads the lane's seven horizontal neighbours once and
// forms the row's partial sums {x, x²} over the column groups the eight
// directions use (x-3..x-1, x+1..x+3, x-2..x, x..x+2, x-3..x-2, x+2..x+3,
// all seven); the directional window sums are then short vertical sums of
// those partials kept in registers:
//   S  = rows y+1..y+3 of the 7-sums          N = S of row y-4
//   E  = rows y-3..y+3 of x+1..x+3            W = the same of x-3..x-1
//   SE = (y, x+1) + (y+1, x..x+2) + (y+2, x+1..x+3) + (y+3, x+2..x+3)
//   NE = (y, x+1) + (y-1, x..x+2) + (y-2, x+1..x+3) + (y-3, x+2..x+3)
//   SW, NW the mirror images of SE, NE (speckle.rs:432-457).
// That is about 30 f64 pair additions per pixel instead of 120.
//
// The sums are the reference's up to the order of the additions.  Each
// direction's mean, variance and CV² are then formed as the reference forms
// them (mean and variance rounded to f32; CV² = var / mean², here through a
// reciprocal within two ulps), and the smallest CV² wins.  A pixel goes to
// the exact path, which repeats speckle.rs:465-569 operation for operation
// from the staged rows, when the fast path cannot decide it as the reference
// would: a non-finite sample in its window (the reference skips it), a
// candidate CV² within 16 ulps of the best seen before it (ties keep the
// first direction), or a mean outside [2^-63, 2^63) (mean² or its reciprocal
// not a normal f32).  Elsewhere the result equals the reference's except
// where the reordered f64 sums round a mean or variance to the other
// neighbouring f32 (a difference of the last f64 bits, rare).
constexpr int kRlRows = 122;            // output rows per wave
constexpr int kRlSteps = kRlRows + 6;   // input rows walked: 3 above, 3 below
constexpr int kRlBlock = 8;             // steps unrolled (register ring period)
constexpr int kRlPitch = 72;            // staged columns x0-3 .. x0+66, padded
constexpr int kRlRing = 16;             // f32 rows kept for the exact path
static_assert(kRlSteps % kRlBlock == 0, "walk length is whole blocks");
static_assert(kRlRing >= kRlBlock + 7, "exact path needs the block's windows");

struct RlWalkLds {
    double2 pair[2][kRlPitch];          // {x, x²} of the row being summed
    float ring[kRlRing][kRlPitch];      // the last rows as f32
    float4 held[3][4][64];              // S (read as N four rows on), NE, NW
    float spare_f[64];                  // halo writes of lanes 6..63 land here
    double2 spare_p[64];
};

struct P2 {
    double s, q;
};
__device__ __forceinline__ P2 operator+(P2 a, P2 b) { return P2{a.s + b.s, a.q + b.q}; }

// One direction's candidate: mean and variance as the reference rounds them
// to f32 and a key ordered like its CV²: the bits of vf / mf² when mf lies
// in [2^-63, 2^63) (mf² and its reciprocal normal), else a skip key above
// +inf (distinct per direction D, so that two skipped directions are never
// "close"); `bad` marks mf > 0 outside that range (the reference would not
// skip it: the exact path decides).  Held candidates carry `bad` as the key
// kRlBad.
constexpr uint32_t kRlBad = 0xFFFFFFFFu;
constexpr uint32_t kRlInf = 0x7F800000u;
constexpr uint32_t rl_skip_key(int d) { return 0xFFFFFF00u + 32u * (uint32_t)d; }
struct RlCand {
    float mf, vf;
    uint32_t key;
    bool bad;
};

template <int N, int D>
__device__ __forceinline__ RlCand rl_cand(P2 a) {
    constexpr double rn = 1.0 / N;
    const double m = a.s * rn;
    // var >= +0: the difference is never -0 (qn - m² = 0 rounds to +0)
    const double va = __builtin_fmax(__builtin_fma(-m, m, a.q * rn), 0.0);
    const float mf = (float)m, vf = (float)va;
    const float cv = vf * __builtin_amdgcn_rcpf(mf * mf);
    // [2^-63, 2^63) as bit patterns 0x20000000 .. 0x5F000000
    const bool inr = __float_as_uint(mf) - 0x20000000u < 0x3F000000u;
    return RlCand{mf, vf, inr ? __float_as_uint(cv) : rl_skip_key(D), !inr && mf > 0.0f};
}

__device__ __forceinline__ float4 rl_put(const RlCand& c) {
    return float4{c.mf, c.vf, __uint_as_float(c.bad ? kRlBad : c.key), 0.0f};
}
__device__ __forceinline__ RlCand rl_get(const float4& v) {
    const uint32_t k = __float_as_uint(v.z);
    return RlCand{v.x, v.y, k, k == kRlBad};   // a bad key is never "close" to another
}

// The smaller-key candidate of two, `a` being the earlier direction in the
// reference's order (it keeps a tie).  Keys within 16 ulps of each other
// leave the pixel undecided (skip keys are 32 apart and never are): by the reference's strict <
// the earlier of two nearly equal CV² can win where the reciprocal form
// cannot tell them apart.  Played as a tree over the eight directions this
// finds every candidate within 16 ulps of the winner: such a candidate lost
// to one between it and the winner.
__device__ __forceinline__ RlCand rl_match(const RlCand& a, const RlCand& b, bool& undecided) {
    // |a - b| <= 16 in wrapping unsigned arithmetic
    undecided |= a.bad | b.bad | (a.key - b.key + 16u <= 32u);
    const bool bw = b.key < a.key;
    return RlCand{bw ? b.mf : a.mf, bw ? b.vf : a.vf, bw ? b.key : a.key, false};
}

struct RlWalkArgs {
    const float* in;
    float* out;
    int64_t cols, rows;
    int64_t x0, y0;                     // strip origin
    int64_t y_end;                      // first row past the strip's rows
    int lane, halo;                     // halo: lane % 6, the column x0 + 61 + halo
    int cc0, cc1;                       // the lane's two columns clamped into the raster
    float cv_n_sq;
    float in0, in1;                     // -0 where the lane's two columns are inside, else NaN
};

struct RlWalk {
    float l0[4], l1[4];                 // prefetched rows (slot: step & 3)
    P2 hs[2];                           // 7-sums of the two previous rows
    P2 he_prev, hw_prev;                // x+1..x+3 / x-3..x-1 of the previous row
    P2 eh2[4], eh4[4], wh2[4], wh4[4];  // pair and quad row sums for E / W
    P2 e_prev, c_prev, se_b, se_a, sw_b, sw_a;
    P2 r23_prev, l23_prev, ne_c2, ne_c3, nw_c2, nw_c3;
    unsigned long long bad_row[8];      // lanes with a non-finite sample per row
    float o[kRlBlock];                  // the block's outputs, stored after it
    uint32_t pend;                      // phases whose pixel takes the exact path
};

// Row r clamped into the raster (a wave-uniform address).  Samples outside
// the raster are turned into NaN when the row is staged (adding -0 or NaN),
// so that a prefetched load is not waited for until then.
__device__ __forceinline__ const float* rl_row(const RlWalkArgs& a, int64_t r) {
    return a.in + (r < 0 ? 0 : (r >= a.rows ? a.rows - 1 : r)) * a.cols;
}

// Stage the row of step `t` (v0: column x0-3+lane; v1: column x0+61+lane%6,
// kept by lanes 0..5, written to a spare slot by the others).  SLOT = t & 1.
template <int SLOT>
__device__ __forceinline__ void rl_stage(RlWalkLds& L, int t, int lane, float v0, float v1) {
    const double x0 = (double)v0, x1 = (double)v1;
    float* rr = L.ring[t & (kRlRing - 1)];
    double2* pr = L.pair[SLOT];
    rr[lane] = v0;
    pr[lane] = double2{x0, x0 * x0};
    (lane < 6 ? rr + 64 + lane : L.spare_f + lane)[0] = v1;
    (lane < 6 ? pr + 64 + lane : L.spare_p + lane)[0] = double2{x1, x1 * x1};
}

// Lanes of one wave exchange rows through LDS: keep every LDS access on
// its side of this point (the wave executes LDS operations in order, so no
// wait is needed, only that the compiler does not move them).
__device__ __forceinline__ void rl_wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Step t = blk * 8 + PH: input row y0 - 3 + t, output row y0 - 6 + t.
// No lane-dependent branch: outputs are kept in registers and stored after
// the block.
template <int PH>
__device__ __forceinline__ void rl_step(RlWalk& w, RlWalkLds& L, const RlWalkArgs& a, int tb) {
    const int t = tb + PH;
    const int lane = a.lane;
    const int64_t r = a.y0 - 3 + t;
    // prefetch the row three steps on; stage the next row
    {
        const float* rp = rl_row(a, r + 3);
        w.l0[(PH + 3) & 3] = rp[a.cc0];
        w.l1[(PH + 3) & 3] = rp[a.cc1];
    }
    {
        const int64_t rs = r + 1;
        const bool row_in = rs >= 0 && rs < a.rows;
        const float qnan = __builtin_nanf("");
        rl_stage<(PH + 1) & 1>(L, t + 1, lane, w.l0[(PH + 1) & 3] + (row_in ? a.in0 : qnan),
                 w.l1[(PH + 1) & 3] + (row_in ? a.in1 : qnan));
    }
    rl_wave_fence();

    const double2* row = L.pair[PH & 1] + lane;
    const P2 pa{row[0].x, row[0].y}, pb{row[1].x, row[1].y}, pc{row[2].x, row[2].y}, pd{row[3].x, row[3].y},
        pe{row[4].x, row[4].y}, pf{row[5].x, row[5].y}, pg{row[6].x, row[6].y};
    // row partials
    const P2 l23 = pa + pb, hw3 = l23 + pc, r23 = pf + pg, he3 = pe + r23;
    const P2 q2 = (pb + pc) + pd, q1 = (pd + pe) + pf;
    const P2 hs7 = (hw3 + he3) + pd;
    w.bad_row[PH] = __builtin_amdgcn_ballot_w64(!isfinite(hs7.q));

    // S of the output row, rows +1..+3
    const P2 S = (w.hs[PH & 1] + w.hs[(PH + 1) & 1]) + hs7;
    w.hs[PH & 1] = hs7;
    // E / W: quad (rows -3..0) + pair (+1, +2) + row +3
    const P2 E = (w.eh4[(PH + 2) & 3] + w.eh2[(PH + 2) & 3]) + he3;
    const P2 W = (w.wh4[(PH + 2) & 3] + w.wh2[(PH + 2) & 3]) + hw3;
    w.eh2[(PH + 3) & 3] = w.he_prev + he3;
    w.eh4[(PH + 1) & 3] = w.eh2[(PH + 1) & 3] + w.eh2[(PH + 3) & 3];
    w.wh2[(PH + 3) & 3] = w.hw_prev + hw3;
    w.wh4[(PH + 1) & 3] = w.wh2[(PH + 1) & 3] + w.wh2[(PH + 3) & 3];
    w.he_prev = he3;
    w.hw_prev = hw3;
    // SE / SW complete on the row three below the output row
    const P2 SE = w.se_a + r23, SW = w.sw_a + l23;
    w.se_a = w.se_b + he3;
    w.sw_a = w.sw_b + hw3;
    w.se_b = w.e_prev + q1;
    w.sw_b = w.c_prev + q2;
    w.e_prev = pe;
    w.c_prev = pc;
    // NE / NW complete on their own row (this input row) and wait three steps
    const P2 NE = w.ne_c3 + pe, NW = w.nw_c3 + pc;
    w.ne_c3 = w.ne_c2 + q1;
    w.nw_c3 = w.nw_c2 + q2;
    w.ne_c2 = w.r23_prev + he3;
    w.nw_c2 = w.l23_prev + hw3;
    w.r23_prev = r23;
    w.l23_prev = l23;

    // held candidates: N = S of four rows up, NE / NW of three rows up
    const RlCand cn = rl_get(L.held[0][PH & 3][lane]), cne = rl_get(L.held[1][(PH + 1) & 3][lane]),
                 cnw = rl_get(L.held[2][(PH + 1) & 3][lane]);
    const RlCand cs = rl_cand<21, 0>(S);
    L.held[0][PH & 3][lane] = rl_put(cs);
    L.held[1][PH & 3][lane] = rl_put(rl_cand<9, 6>(NE));
    L.held[2][PH & 3][lane] = rl_put(rl_cand<9, 5>(NW));

    const float centre = L.ring[(t - 3) & (kRlRing - 1)][lane + 3];
    // a row with no finite centre in the strip (no-data) skips the pick;
    // the branch also bounds the scheduler's region to one step
    if (__builtin_amdgcn_ballot_w64(isfinite(centre)) == 0) {
        w.o[PH] = __builtin_nanf("");  // a non-finite centre gives NaN (an infinite one too)
        return;
    }
    bool undecided = false;
    // N is S of four rows up: move its skip key to direction 1's
    const RlCand cn1{cn.mf, cn.vf, cn.key == rl_skip_key(0) ? rl_skip_key(1) : cn.key, cn.bad};
    const RlCand c_sn = rl_match(cs, cn1, undecided);
    const RlCand c_ew = rl_match(rl_cand<21, 2>(E), rl_cand<21, 3>(W), undecided);
    const RlCand c_d1 = rl_match(rl_cand<9, 4>(SE), cnw, undecided);
    const RlCand c_d2 = rl_match(cne, rl_cand<9, 7>(SW), undecided);
    const RlCand c_a = rl_match(c_sn, c_ew, undecided);
    const RlCand c_b = rl_match(c_d1, c_d2, undecided);
    const RlCand best = rl_match(c_a, c_b, undecided);
    unsigned long long win_bad = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) win_bad |= w.bad_row[(PH + 8 - k) & 7];
    const bool finite = isfinite(centre);
    const bool exact = finite && (undecided || ((win_bad >> lane) & 1ull));
    w.pend |= exact ? 1u << PH : 0u;
    // Lee update with the chosen direction (speckle.rs:555-567)
    const float noise_var = best.mf * best.mf * a.cv_n_sq;
    const float signal_var = fmaxf(best.vf - noise_var, 0.0f);
    const float denom = signal_var + noise_var;
    const float k = denom > 0.0f ? signal_var / denom : 0.0f;
    const float lee = best.mf + k * (centre - best.mf);
    // no direction with a finite CV²: the centre passes (speckle.rs:550-553)
    w.o[PH] = !finite ? __builtin_nanf("") : (best.key >= kRlInf ? centre : lee);
}

#ifndef SD_RLW_WAVES
#define SD_RLW_WAVES 2
#endif
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SD_RLW_WAVES))) void refined_lee_walk_kernel(
    SpeckleParams p) {
    __shared__ RlWalkLds L;
    RlWalkArgs a;
    a.in = p.in;
    a.out = p.out;
    a.cols = p.cols;
    a.rows = p.rows;
    a.x0 = (int64_t)blockIdx.x * 64;
    a.y0 = (int64_t)blockIdx.y * kRlRows;
    a.y_end = a.y0 + kRlRows < p.rows ? a.y0 + kRlRows : p.rows;
    a.lane = threadIdx.x;
    a.halo = a.lane % 6;
    a.cv_n_sq = 1.0f / p.enl;
    const int lane = a.lane;
    const int64_t x = a.x0 + lane;
    {
        const int64_t c0 = a.x0 - 3 + lane, c1 = a.x0 + 61 + a.halo;
        a.in0 = (c0 >= 0 && c0 < a.cols) ? -0.0f : __builtin_nanf("");
        a.in1 = (c1 >= 0 && c1 < a.cols) ? -0.0f : __builtin_nanf("");
        a.cc0 = (int)(c0 < 0 ? 0 : (c0 >= a.cols ? a.cols - 1 : c0));
        a.cc1 = (int)(c1 < 0 ? 0 : (c1 >= a.cols ? a.cols - 1 : c1));
    }
    RlWalk w;
    {
        const P2 z{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            w.eh2[i] = z;
            w.eh4[i] = z;
            w.wh2[i] = z;
            w.wh4[i] = z;
        }
        w.hs[0] = w.hs[1] = z;
        w.he_prev = w.hw_prev = w.e_prev = w.c_prev = z;
        w.se_b = w.se_a = w.sw_b = w.sw_a = z;
        w.r23_prev = w.l23_prev = w.ne_c2 = w.ne_c3 = w.nw_c2 = w.nw_c3 = z;
#pragma unroll
        for (int i = 0; i < 8; ++i) w.bad_row[i] = 0;
        w.pend = 0;
    }
    // rows of steps 0..2 in flight, step 0's row staged
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float* rp = rl_row(a, a.y0 - 3 + i);
        w.l0[i] = rp[a.cc0];
        w.l1[i] = rp[a.cc1];
    }
    w.l0[3] = w.l1[3] = 0.0f;
    {
        const bool row_in = a.y0 - 3 >= 0;   // y0 - 3 < rows: the strip has rows
        const float qnan = __builtin_nanf("");
        rl_stage<0>(L, 0, lane, w.l0[0] + (row_in ? a.in0 : qnan), w.l1[0] + (row_in ? a.in1 : qnan));
    }
    rl_wave_fence();
#pragma unroll 1
    for (int tb = 0; tb < kRlSteps; tb += kRlBlock) {
        rl_step<0>(w, L, a, tb);
        rl_step<1>(w, L, a, tb);
        rl_step<2>(w, L, a, tb);
        rl_step<3>(w, L, a, tb);
        rl_step<4>(w, L, a, tb);
        rl_step<5>(w, L, a, tb);
        rl_step<6>(w, L, a, tb);
        rl_step<7>(w, L, a, tb);
        // the block's rows y0-6+tb .. y0+1+tb; those before y0 belong to the
        // strip above (their steps only fill the sums)
#pragma unroll
        for (int ph = 0; ph < kRlBlock; ++ph) {
            const int64_t y = a.y0 - 6 + tb + ph;
            if (y < a.y0 || y >= a.y_end || x >= a.cols) w.pend &= ~(1u << ph);
            else if (!((w.pend >> ph) & 1u)) (a.out + y * a.cols)[x] = w.o[ph];
        }
        // Pixels the fast path could not decide, listed in LDS and taken 64
        // at a time (any lane takes any listed pixel: the block's rows,
        // steps tb-6 .. tb+7, are all still in the f32 ring).
        if (__builtin_amdgcn_ballot_w64(w.pend != 0) != 0) {
            uint16_t* list = reinterpret_cast<uint16_t*>(L.spare_p);   // 1 KiB: up to 8 x 64 entries
            int n = 0;
#pragma unroll
            for (int ph = 0; ph < kRlBlock; ++ph) {
                const bool f = (w.pend >> ph) & 1u;
                const unsigned long long m = __builtin_amdgcn_ballot_w64(f);
                const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                if (f) list[n + below] = (uint16_t)((ph << 8) | lane);
                n += __builtin_popcountll(m);
            }
            rl_wave_fence();
#pragma unroll 1
            for (int base = 0; base < n; base += 64) {
                if (base + lane < n) {
                    const int e = list[base + lane];
                    const int t = tb + (e >> 8), l = e & 63;
                    const float* col = &L.ring[0][l + 3];
                    const float centre = col[((t - 3) & (kRlRing - 1)) * kRlPitch];
                    const float o = rl_skipping(centre, a.cv_n_sq, [&](int dy, int dx) {
                        return (double)col[((t - 3 + dy) & (kRlRing - 1)) * kRlPitch + dx];
                    });
                    a.out[(a.y0 - 6 + t) * a.cols + a.x0 + l] = o;
                }
            }
            rl_wave_fence();   // the list area is a staging spare again
            w.pend = 0;
        }
    }
}

// ── batched filtering (speckle steps of a batch of geocoded tiles) ──────

// One job: a raster and its filter.  Tiles of all jobs form one index
// space; tile_begin is the job's first tile, tiles_x its tiles per row.
struct SpeckleJob {
    SpeckleParams p;
    int32_t kind;
    int32_t pad;
    int64_t tile_begin, tiles_x;
};

struct SpeckleBatchHeader {
    int32_t n_jobs;
    int32_t pad;
    int64_t n_tiles;
};

template <int KIND>
__device__ __forceinline__ void batch_tile_kind(const SpeckleParams& p, float* tile, int64_t r0, int64_t c0) {
    switch (p.half) {
        case 1: speckle_tile<KIND, 1>(p, tile, r0, c0); break;
        case 2: speckle_tile<KIND, 2>(p, tile, r0, c0); break;
        case 3: speckle_tile<KIND, 3>(p, tile, r0, c0); break;
        default: speckle_tile<KIND, 0>(p, tile, r0, c0); break;
    }
}

// Persistent: workgroups pull tiles of any job from `queue` (zero at start);
// the LDS tile is sized for the largest window of the batch.
__global__ __launch_bounds__(256) void batch_speckle_kernel(const SpeckleBatchHeader* __restrict__ H,
                                                            const SpeckleJob* __restrict__ J, unsigned int* queue) {
    extern __shared__ __attribute__((aligned(16))) float tile[];
    __shared__ unsigned int s_item;
    const int n_jobs = H->n_jobs;
    const int64_t n = H->n_tiles;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
        __syncthreads();
        const int64_t item = (int64_t)s_item;
        if (item >= n) break;
        int lo = 0, hi = n_jobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (J[mid].tile_begin <= item) lo = mid;
            else hi = mid - 1;
        }
        const int j = __builtin_amdgcn_readfirstlane(lo);
        const SpeckleJob& job = J[j];
        const int64_t t = item - job.tile_begin;
        const int64_t r0 = (t / job.tiles_x) * kTileH, c0 = (t % job.tiles_x) * kTileW;
        switch (job.kind) {
            case SK_BOXCAR: batch_tile_kind<SK_BOXCAR>(job.p, tile, r0, c0); break;
            case SK_LEE: batch_tile_kind<SK_LEE>(job.p, tile, r0, c0); break;
            case SK_FROST: batch_tile_kind<SK_FROST>(job.p, tile, r0, c0); break;
            case SK_GAMMA: batch_tile_kind<SK_GAMMA>(job.p, tile, r0, c0); break;
            default: speckle_tile<SK_RLEE, 3>(job.p, tile, r0, c0); break;
        }
        __syncthreads();  // the tile and the item slot are rewritten next
    }
}

template <int KIND, int HALF>
static hipError_t launch_one(const SpeckleParams& p, hipStream_t s) {
    const int half = KIND == SK_RLEE ? 3 : p.half;
    size_t lds = pair_tile(KIND, HALF) ? pair_tile_bytes(half) : (size_t)(kTileW + 2 * half) * (kTileH + 2 * half) * sizeof(float);
    dim3 grid((unsigned)((p.cols + kTileW - 1) / kTileW), (unsigned)((p.rows + kTileH - 1) / kTileH));
    if constexpr (KIND == SK_RLEE) {
        const dim3 walk((unsigned)((p.cols + 63) / 64), (unsigned)((p.rows + kRlRows - 1) / kRlRows));
        hipLaunchKernelGGL(refined_lee_walk_kernel, walk, dim3(64), 0, s, p);
        (void)lds;
        (void)grid;
    }
    else hipLaunchKernelGGL((speckle_kernel<KIND, HALF>), grid, dim3(256), lds, s, p);
    return hipGetLastError();
}

template <int KIND>
static hipError_t launch_kind(const SpeckleParams& p, hipStream_t s) {
    switch (p.half) {
        case 1: return launch_one<KIND, 1>(p, s);
        case 2: return launch_one<KIND, 2>(p, s);
        case 3: return launch_one<KIND, 3>(p, s);
        default: return launch_one<KIND, 0>(p, s);
    }
}

}  // namespace sd

// kind: sardine_speckle_kind; half = window/2 (≤ 31, checked by the caller).
extern "C" hipError_t sd_launch_speckle(int kind, const float* in, float* out, int64_t cols, int64_t rows, int half,
                                        float enl, float damping, hipStream_t stream) {
    if (cols <= 0 || rows <= 0) return hipSuccess;
    sd::SpeckleParams p{in, out, cols, rows, half, enl, damping};
    switch (kind) {
        case sd::SK_BOXCAR: return sd::launch_kind<sd::SK_BOXCAR>(p, stream);
        case sd::SK_LEE: return sd::launch_kind<sd::SK_LEE>(p, stream);
        case sd::SK_FROST: return sd::launch_kind<sd::SK_FROST>(p, stream);
        case sd::SK_GAMMA: return sd::launch_kind<sd::SK_GAMMA>(p, stream);
        default: p.half = 3; return sd::launch_one<sd::SK_RLEE, 3>(p, stream);
    }
}

// Tiles per job for the batched pass (host helper for the job table).
extern "C" int64_t sd_speckle_tiles(int64_t cols, int64_t rows, int64_t* tiles_x) {
    *tiles_x = (cols + sd::kTileW - 1) / sd::kTileW;
    return *tiles_x * ((rows + sd::kTileH - 1) / sd::kTileH);
}

extern "C" size_t sd_speckle_job_size(void) { return sizeof(sd::SpeckleJob); }

// Fills one job-table entry (kept here so the host needs no kernel types).
extern "C" void sd_speckle_job_fill(void* entry, int kind, const float* in, float* out, int64_t cols, int64_t rows,
                                    int half, float enl, float damping, int64_t tile_begin, int64_t tiles_x) {
    sd::SpeckleJob& j = *static_cast<sd::SpeckleJob*>(entry);
    j.p = sd::SpeckleParams{in, out, cols, rows, kind == sd::SK_RLEE ? 3 : half, enl, damping};
    j.kind = kind;
    j.pad = 0;
    j.tile_begin = tile_begin;
    j.tiles_x = tiles_x;
}

extern "C" hipError_t sd_launch_batch_speckle(const void* H, const void* J, unsigned int* queue, int n_groups,
                                              int max_half, hipStream_t stream) {
    // jobs with windows up to 7 x 7 (and Refined Lee) stage {x, x²} pairs:
    // the larger of the two tiles
    size_t lds = (size_t)(sd::kTileW + 2 * max_half) * (sd::kTileH + 2 * max_half) * sizeof(float);
    const size_t pairs = sd::pair_tile_bytes(max_half < sd::kMaxPairHalf ? max_half : sd::kMaxPairHalf);
    if (max_half >= 1 && lds < pairs) lds = pairs;
    hipLaunchKernelGGL(sd::batch_speckle_kernel, dim3((unsigned)n_groups), dim3(256), lds, stream,
                       static_cast<const sd::SpeckleBatchHeader*>(H), static_cast<const sd::SpeckleJob*>(J), queue);
    return hipGetLastError();
}
