// multilook.hip — the decimating radar-geometry stages: the pre-TC
// multilook of the merged σ⁰ and NESZ, and the ground-range (GRD)
// projection's resampling + azimuth multilook.
//
// Multilook
// (apply_multilook, scene_prep.rs:1060-1161): NaN-aware boxcar decimation by
// range_looks x azimuth_looks, output pixel (l, c) = mean of the non-NaN σ⁰
// samples of input block [l*A, l*A+A) x [c*R, c*R+R), NESZ averaged over the
// same samples.  Sums run in f64 in the reference's order (azimuth outer,
// range inner) and are divided by the count in f64, so results are the
// reference's bit for bit.
//
// A thread owns one output pixel of one output line and walks its block; a
// wave's loads of one block row cover 64*R consecutive input samples, so
// every input byte is fetched once from HBM (the R loads of one lane hit the
// lines its neighbours bring in).  HBM-bound: 8 B read per input pixel and
// 8 B written per output pixel.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sd {

#ifndef SD_ML_NT
#define SD_ML_NT 0
#endif

template <typename T>
__device__ __forceinline__ T ld(const T* p) {
    if (SD_ML_NT) return __builtin_nontemporal_load(p);
    return *p;
}

typedef float f2v __attribute__((ext_vector_type(2)));

// With one azimuth look and an unrolled range factor read by scalar loads,
// a thread takes kMlRows output lines (gridDim.y apart) and issues all their
// loads before it sums any: 5 x 1 looks 1.54 → 1.48 ms on the 13100 x 60814
// image.  The 8-byte-pair reads (V2) stay one line per thread: there two
// lines took 4 x 1 from 1.55 to 1.62 ms (and four lines were slower still).
#ifndef SD_ML_ROWS
#define SD_ML_ROWS 2
#endif
constexpr int kMlRows = SD_ML_ROWS;

// One block row of RT samples of σ⁰ and NESZ.
template <int RT, bool V2>
__device__ __forceinline__ void ml_load_row(const float* pd, const float* pn, float* v, float* n) {
    if (V2) {
#pragma unroll
        for (int j = 0; j < RT / 2; ++j) {
            const f2v a = ld(reinterpret_cast<const f2v*>(pd) + j);
            const f2v b = ld(reinterpret_cast<const f2v*>(pn) + j);
            v[2 * j] = a.x;
            v[2 * j + 1] = a.y;
            n[2 * j] = b.x;
            n[2 * j + 1] = b.y;
        }
    } else {
#pragma unroll
        for (int dr = 0; dr < RT; ++dr) {
            v[dr] = ld(pd + dr);
            n[dr] = ld(pn + dr);
        }
    }
}

// RT > 0: range_looks known at compile time (the block row is unrolled).
// V2: rows and blocks start at even samples and the rasters are 8-byte
// aligned, so a block row is read as RT/2 8-byte pairs (RT even).
template <int RT, bool V2>
__global__ __launch_bounds__(256) void multilook_kernel(const float* __restrict__ data,
                                                        const float* __restrict__ nesz, int64_t in_samples,
                                                        int64_t out_lines, int64_t out_samples, int r_looks,
                                                        int a_looks, float* __restrict__ out_data,
                                                        float* __restrict__ out_nesz) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= out_samples) return;
    const int R = RT > 0 ? RT : r_looks;
    const float qnan = __builtin_nanf("");
    if constexpr (RT > 0 && !V2 && kMlRows > 1) {
        if (a_looks == 1) {
            const int64_t step = gridDim.y;
            for (int64_t l0 = blockIdx.y; l0 < out_lines; l0 += kMlRows * step) {
                float v[kMlRows][RT], n[kMlRows][RT];
#pragma unroll
                for (int u = 0; u < kMlRows; ++u) {
                    const int64_t l = l0 + u * step < out_lines ? l0 + u * step : l0;  // in bounds; unused
                    ml_load_row<RT, V2>(data + l * in_samples + c * RT, nesz + l * in_samples + c * RT, v[u], n[u]);
                }
#pragma unroll
                for (int u = 0; u < kMlRows; ++u) {
                    const int64_t l = l0 + u * step;
                    if (l >= out_lines) break;
                    double sum = 0.0, sum_n = 0.0;
                    unsigned count = 0;
#pragma unroll
                    for (int dr = 0; dr < RT; ++dr) {
                        if (!__builtin_isnan(v[u][dr])) {
                            sum += (double)v[u][dr];
                            sum_n += (double)n[u][dr];
                            ++count;
                        }
                    }
                    out_data[l * out_samples + c] = count ? (float)(sum / (double)count) : qnan;
                    out_nesz[l * out_samples + c] = count ? (float)(sum_n / (double)count) : qnan;
                }
            }
            return;
        }
    }
    for (int64_t l = blockIdx.y; l < out_lines; l += gridDim.y) {
        double sum = 0.0, sum_n = 0.0;
        unsigned count = 0;
        const int64_t base = l * a_looks * in_samples + c * R;
        for (int da = 0; da < a_looks; ++da) {
            const float* pd = data + base + da * in_samples;
            const float* pn = nesz + base + da * in_samples;
            if (RT > 0) {
                float v[RT > 0 ? RT : 1], n[RT > 0 ? RT : 1];
                if (V2) {
#pragma unroll
                    for (int j = 0; j < RT / 2; ++j) {
                        const f2v a = ld(reinterpret_cast<const f2v*>(pd) + j);
                        const f2v b = ld(reinterpret_cast<const f2v*>(pn) + j);
                        v[2 * j] = a.x;
                        v[2 * j + 1] = a.y;
                        n[2 * j] = b.x;
                        n[2 * j + 1] = b.y;
                    }
                } else {
#pragma unroll
                    for (int dr = 0; dr < RT; ++dr) {
                        v[dr] = ld(pd + dr);
                        n[dr] = ld(pn + dr);
                    }
                }
#pragma unroll
                for (int dr = 0; dr < RT; ++dr) {
                    if (!__builtin_isnan(v[dr])) {
                        sum += (double)v[dr];
                        sum_n += (double)n[dr];
                        ++count;
                    }
                }
            } else {
                for (int dr = 0; dr < R; ++dr) {
                    const float v = ld(pd + dr), n = ld(pn + dr);
                    if (!__builtin_isnan(v)) {
                        sum += (double)v;
                        sum_n += (double)n;
                        ++count;
                    }
                }
            }
        }
        out_data[l * out_samples + c] = count ? (float)(sum / (double)count) : qnan;
        out_nesz[l * out_samples + c] = count ? (float)(sum_n / (double)count) : qnan;
    }
}

#ifndef SD_ML_VEC
#define SD_ML_VEC 1
#endif

template <int RT>
hipError_t launch_multilook(const float* data, const float* nesz, int64_t in_samples, int64_t out_lines,
                            int64_t out_samples, int r, int a, float* od, float* on, hipStream_t s) {
    const unsigned gx = (unsigned)((out_samples + 255) / 256);
    const bool v2 = SD_ML_VEC && RT > 0 && RT % 2 == 0 && in_samples % 2 == 0 &&
                    reinterpret_cast<uintptr_t>(data) % 8 == 0 && reinterpret_cast<uintptr_t>(nesz) % 8 == 0;
    // the multi-line form (scalar loads, one azimuth look) takes kMlRows lines per thread
    const int64_t rows = RT > 0 && !v2 && a == 1 ? (out_lines + kMlRows - 1) / kMlRows : out_lines;
    const unsigned gy = (unsigned)(rows < 65535 ? rows : 65535);
    if (v2)
        hipLaunchKernelGGL((multilook_kernel<RT, true>), dim3(gx, gy), dim3(256), 0, s, data, nesz, in_samples,
                           out_lines, out_samples, r, a, od, on);
    else
        hipLaunchKernelGGL((multilook_kernel<RT, false>), dim3(gx, gy), dim3(256), 0, s, data, nesz, in_samples,
                           out_lines, out_samples, r, a, od, on);
    return hipGetLastError();
}

// Ground-range projection (to_ground_range, ground_range.rs:364-428): output
// column j reads merged column `left` (exact) or interpolates between `left`
// and `left + 1` with the f32 weight `frac` (NaN if either tap is NaN); the
// azimuth multilook then averages the non-NaN projected values of A
// consecutive lines in f64.  Both steps run per thread, so the projected
// raster is never stored.  The column map is the host's (exact: frac < 0).
struct GrdColumn {
    int32_t left;
    float frac;  // < 0: exact column
};

// One output column per thread.  With one azimuth look (the IW case) a
// thread takes kGrdRows output lines, gridDim.y apart, and issues all their
// tap loads before it uses any.
#ifndef SD_GRD_ROWS
#define SD_GRD_ROWS 4
#endif
constexpr int kGrdRows = SD_GRD_ROWS;

__device__ __forceinline__ float grd_tap(const float* row, const GrdColumn& m, bool exact) {
    if (exact) return row[0];
    const float v0 = row[0], v1 = row[1];
    return (__builtin_isnan(v0) || __builtin_isnan(v1)) ? __builtin_nanf("") : v0 + (v1 - v0) * m.frac;
}

__global__ __launch_bounds__(256) void grd_kernel(const float* __restrict__ data, int64_t in_samples,
                                                  const GrdColumn* __restrict__ map, int64_t out_lines,
                                                  int64_t out_samples, int a_looks, float* __restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= out_samples) return;
    const GrdColumn m = map[j];
    const bool exact = m.frac < 0.0f;
    const float qnan = __builtin_nanf("");
    const int64_t step = gridDim.y;
    if (a_looks == 1) {
        for (int64_t l0 = blockIdx.y; l0 < out_lines; l0 += kGrdRows * step) {
            float v0[kGrdRows], v1[kGrdRows];
#pragma unroll
            for (int u = 0; u < kGrdRows; ++u) {
                const int64_t l = l0 + u * step;
                const float* row = data + (l < out_lines ? l : l0) * in_samples + m.left;
                v0[u] = row[0];
                v1[u] = exact ? 0.0f : row[1];
            }
#pragma unroll
            for (int u = 0; u < kGrdRows; ++u) {
                const int64_t l = l0 + u * step;
                if (l >= out_lines) break;
                float v = v0[u];
                if (!exact) v = (__builtin_isnan(v0[u]) || __builtin_isnan(v1[u])) ? qnan : v0[u] + (v1[u] - v0[u]) * m.frac;
                // one look: the reference's f64 accumulation of a single value
                // from +0.0 (so −0 becomes +0), divided by a count of 1
                double sum = 0.0;
                sum += (double)v;
                out[l * out_samples + j] = __builtin_isnan(v) ? qnan : (float)(sum / 1.0);
            }
        }
        return;
    }
    for (int64_t l = blockIdx.y; l < out_lines; l += step) {
        double sum = 0.0;
        unsigned count = 0;
        const float* row = data + l * a_looks * in_samples + m.left;
        for (int da = 0; da < a_looks; ++da, row += in_samples) {
            const float v = grd_tap(row, m, exact);
            if (!__builtin_isnan(v)) {
                sum += (double)v;
                ++count;
            }
        }
        out[l * out_samples + j] = count ? (float)(sum / (double)count) : qnan;
    }
}

}  // namespace sd

extern "C" hipError_t sd_launch_grd(const float* data, int64_t in_samples, const void* map, int64_t out_lines,
                                    int64_t out_samples, int a, float* out, hipStream_t s) {
    if (out_lines <= 0 || out_samples <= 0) return hipSuccess;
    const unsigned gx = (unsigned)((out_samples + 255) / 256);
    const int64_t rows = a == 1 ? (out_lines + sd::kGrdRows - 1) / sd::kGrdRows : out_lines;
    const unsigned gy = (unsigned)(rows < 65535 ? rows : 65535);
    hipLaunchKernelGGL(sd::grd_kernel, dim3(gx, gy), dim3(256), 0, s, data, in_samples,
                       static_cast<const sd::GrdColumn*>(map), out_lines, out_samples, a, out);
    return hipGetLastError();
}

// Sizes as the host validated them: out_lines = lines / a, out_samples =
// samples / r, both >= 1.
extern "C" hipError_t sd_launch_multilook(const float* data, const float* nesz, int64_t in_samples,
                                          int64_t out_lines, int64_t out_samples, int r, int a, float* od,
                                          float* on, hipStream_t s) {
    if (out_lines <= 0 || out_samples <= 0) return hipSuccess;
    switch (r) {
        case 1: return sd::launch_multilook<1>(data, nesz, in_samples, out_lines, out_samples, r, a, od, on, s);
        case 2: return sd::launch_multilook<2>(data, nesz, in_samples, out_lines, out_samples, r, a, od, on, s);
        case 3: return sd::launch_multilook<3>(data, nesz, in_samples, out_lines, out_samples, r, a, od, on, s);
        case 4: return sd::launch_multilook<4>(data, nesz, in_samples, out_lines, out_samples, r, a, od, on, s);
        case 5: return sd::launch_multilook<5>(data, nesz, in_samples, out_lines, out_samples, r, a, od, on, s);
        case 8: return sd::launch_multilook<8>(data, nesz, in_samples, out_lines, out_samples, r, a, od, on, s);
        case 10: return sd::launch_multilook<10>(data, nesz, in_samples, out_lines, out_samples, r, a, od, on, s);
        default: return sd::launch_multilook<0>(data, nesz, in_samples, out_lines, out_samples, r, a, od, on, s);
    }
}
