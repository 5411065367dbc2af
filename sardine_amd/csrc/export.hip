// export.hip — the device half of the Cloud-Optimised GeoTIFF writer
// (export.rs:329-760): the overview pyramid (downsample_2x,
// export.rs:362-392) and the packing of every level into the file's tile
// region, so that the host only prepends the header and streams bytes out.
//
// The tile region holds the levels coarsest first; level k's tiles are
// row-major over its tile grid, each tile_size^2 float32 values row-major,
// NaN outside the level's raster.  Every launch here produces one level of
// the pyramid from the level below it:
//   - the first reads the full raster (row-major) and writes both the full
//     raster's tiles (the values it read) and the first overview's tiles;
//   - each later one reads the previous overview from its tiles and writes
//     the next overview's tiles.
// A thread owns two adjacent output pixels of the level it writes, so it
// reads a 4-column x 2-row patch of the source (one 16-byte run per row).
// Both grids cover the padded tile area of the output level, which always
// contains the padded area of the full raster (2*ceil(ceil(c/2)/ts)*ts >=
// ceil(c/ts)*ts), so every padding value of every tile is written.
//
// Bandwidth: HBM-bound streaming (the first launch reads 4 B and writes
// 5 B per full-resolution pixel; the rest add about 1/4 of that).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace sd {

struct CogLevelDev {
    float* tiles;   // the level's first tile
    int64_t cols, rows;
    int64_t tiles_x;
    int64_t pad_cols, pad_rows;  // tiles_x * ts, tiles_y * ts
};

// Offset (in floats) of pixel (x, y) of a level in its tiles; x, y inside the
// padded area.
__device__ __forceinline__ int64_t tile_index(int64_t x, int64_t y, int64_t tiles_x, int lg) {
    const int64_t m = (1ll << lg) - 1;
    return ((((y >> lg) * tiles_x + (x >> lg)) << (2 * lg)) + ((y & m) << lg) + (x & m));
}

// 2x2 mean of the non-NaN values in the reference's order (row by row, left
// to right), accumulated in f64 from +0.0; NaN when none is a number.
__device__ __forceinline__ float patch_mean(float a, float b, float c, float d) {
    double sum = 0.0;
    int count = 0;
    if (!__builtin_isnan(a)) { sum += (double)a; ++count; }
    if (!__builtin_isnan(b)) { sum += (double)b; ++count; }
    if (!__builtin_isnan(c)) { sum += (double)c; ++count; }
    if (!__builtin_isnan(d)) { sum += (double)d; ++count; }
    return count > 0 ? (float)(sum / (double)count) : __builtin_nanf("");
}

// The 4-column x 2-row source patch of output pair (xo, yo) into v.
// TILED: the source is a level in its tiles (else the row-major full
// raster).  VEC: row-major source rows are 16-byte aligned at every
// multiple of 4 columns.
template <bool TILED, bool VEC>
__device__ __forceinline__ void load_patch(const float* __restrict__ src, const CogLevelDev& S, int64_t xs,
                                           int64_t yo, int lg, float (&v)[2][4]) {
    const float qnan = __builtin_nanf("");
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int64_t ys = 2 * yo + r;
        if (TILED) {
            // inside the source's padded area the tiles hold NaN beyond its
            // raster, so only the padded bound is tested
            if (xs < S.pad_cols && ys < S.pad_rows) {
                const float4 q = *reinterpret_cast<const float4*>(src + tile_index(xs, ys, S.tiles_x, lg));
                v[r][0] = q.x;
                v[r][1] = q.y;
                v[r][2] = q.z;
                v[r][3] = q.w;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) v[r][j] = qnan;
            }
        } else {
            if (ys < S.rows && VEC && xs + 3 < S.cols) {
                const float4 q = *reinterpret_cast<const float4*>(src + ys * S.cols + xs);
                v[r][0] = q.x;
                v[r][1] = q.y;
                v[r][2] = q.z;
                v[r][3] = q.w;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    v[r][j] = (ys < S.rows && xs + j < S.cols) ? src[ys * S.cols + xs + j] : qnan;
            }
        }
    }
}

// Rows of output pairs a thread handles per loop trip (their loads are all
// issued before any store), the workgroups per launch the grid is capped at
// (0: no cap) and whether tiles are written with non-temporal stores (they
// are not read again on the device except as the next level's source).
// Build options, so that they can be tuned per device; on MI355X at
// 25479 x 18590 (tile 512) 2 rows, no cap and non-temporal stores measured
// best (capping the grid at 1-8k workgroups cost 8-20%).
#ifndef SD_COG_ROWS
#define SD_COG_ROWS 2
#endif
#ifndef SD_COG_WGS
#define SD_COG_WGS 0
#endif
#ifndef SD_COG_NT
#define SD_COG_NT 1
#endif
constexpr int kCogRows = SD_COG_ROWS;
constexpr int64_t kCogWgs = SD_COG_WGS;

typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
template <typename V>
__device__ __forceinline__ void put(V* p, V v) {
    if (SD_COG_NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// COPY: also write the source into `copy`'s tiles (the full raster's).
template <bool TILED, bool VEC, bool COPY>
__global__ __launch_bounds__(256) void cog_level_kernel(const float* __restrict__ src, CogLevelDev S,
                                                        CogLevelDev D, CogLevelDev copy, int lg) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // output column pair
    const int64_t xo = 2 * g, xs = 4 * g;
    if (xo >= D.pad_cols) return;
    const int64_t step = gridDim.y;
    for (int64_t y0 = blockIdx.y; y0 < D.pad_rows; y0 += kCogRows * step) {
        float v[kCogRows][2][4];
#pragma unroll
        for (int u = 0; u < kCogRows; ++u) {
            const int64_t yo = y0 + u * step;
            if (u == 0 || yo < D.pad_rows) load_patch<TILED, VEC>(src, S, xs, yo, lg, v[u]);
        }
#pragma unroll
        for (int u = 0; u < kCogRows; ++u) {
            const int64_t yo = y0 + u * step;
            if (u > 0 && yo >= D.pad_rows) break;
            if (COPY) {
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int64_t ys = 2 * yo + r;
                    if (xs < copy.pad_cols && ys < copy.pad_rows)
                        put(reinterpret_cast<f4v*>(copy.tiles + tile_index(xs, ys, copy.tiles_x, lg)),
                            f4v{v[u][r][0], v[u][r][1], v[u][r][2], v[u][r][3]});
                }
            }
            const float m0 = patch_mean(v[u][0][0], v[u][0][1], v[u][1][0], v[u][1][1]);
            const float m1 = patch_mean(v[u][0][2], v[u][0][3], v[u][1][2], v[u][1][3]);
            put(reinterpret_cast<f2v*>(D.tiles + tile_index(xo, yo, D.tiles_x, lg)), f2v{m0, m1});
        }
    }
}

template <bool TILED, bool VEC, bool COPY>
hipError_t launch_level(const float* src, const CogLevelDev& S, const CogLevelDev& D, const CogLevelDev& copy,
                        int lg, hipStream_t s) {
    const int64_t pairs = D.pad_cols / 2;
    const int64_t gx = (pairs + 255) / 256;
    int64_t gy = D.pad_rows < 65535 ? D.pad_rows : 65535;
    if (kCogWgs > 0) gy = std::min<int64_t>(gy, std::max<int64_t>(1, (kCogWgs + gx - 1) / gx));
    hipLaunchKernelGGL((cog_level_kernel<TILED, VEC, COPY>), dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, src,
                       S, D, copy, lg);
    return hipGetLastError();
}

}  // namespace sd

// One pyramid step.  `full` (row-major, level S) → level D's tiles, also
// copying S into `copy_tiles` when that is non-null.  Otherwise `src` is
// level S's tiles.  Sizes must satisfy D = ceil(S / 2), pad = tiles * 2^lg.
extern "C" hipError_t sd_launch_cog_level(const float* src, int src_tiled, const sd::CogLevelDev* S,
                                          const sd::CogLevelDev* D, const sd::CogLevelDev* copy, int lg,
                                          hipStream_t s) {
    if (D->pad_cols <= 0 || D->pad_rows <= 0) return hipSuccess;
    if (src_tiled) return sd::launch_level<true, false, false>(src, *S, *D, *copy, lg, s);
    const bool vec = (S->cols % 4) == 0 && (reinterpret_cast<uintptr_t>(src) % 16) == 0;
    if (copy->tiles) {
        return vec ? sd::launch_level<false, true, true>(src, *S, *D, *copy, lg, s)
                   : sd::launch_level<false, false, true>(src, *S, *D, *copy, lg, s);
    }
    return vec ? sd::launch_level<false, true, false>(src, *S, *D, *copy, lg, s)
               : sd::launch_level<false, false, false>(src, *S, *D, *copy, lg, s);
}
