// cal.hip — LUT σ⁰ calibration with NESZ, with the TOPS deburst and the IW
// sub-swath merge fused into its addressing (one HBM pass: 4 B CInt16 read,
// 4 B σ⁰ + 4 B NESZ written per merged pixel).
//
// Reference chain (scene_prep.rs:598-699):
//   deburst_subswath   sardine/src/deburst.rs:230-344   (row copy → tiff_row table)
//   apply_calibration  sardine/src/apply_calibration.rs:374-572
//   merge_subswaths    sardine/src/merge_subswaths.rs:207-450 (column copy → owner test)
//
// The range LUT evaluation (lut_cursor_eval, apply_calibration.rs:212-225)
// depends only on (LUT vector, TIFF column), so it is tabulated once per
// call per vector by lut_dense_kernel; the azimuth bracket and line
// fraction depend only on the line and come from a host table.  The
// per-pixel arithmetic keeps the reference's f32 order:
//   k2 = K*K; inv = k2 > 0 ? 1/k2 : 0; nv = N*az;
//   σ⁰ = max(p - nv, 0) * inv * valid; NESZ = nv * inv * valid.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cal_params.hpp"

namespace sd {

__device__ __forceinline__ float lerp_clamped(float a, float b, float t) {  // apply_calibration.rs:128-131
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return a + (b - a) * t;
}

// dense[v][c] = lut_cursor_eval(pixels_v, values_v, cursor(c), c) for
// c in [0, dense_cols).  cursor(c) = the position the reference's monotone
// walk reaches: the last index i >= 1 with pixels[i] <= c, else 0.
__global__ void lut_dense_kernel(const uint32_t* __restrict__ pixels, const float* __restrict__ values,
                                 const int64_t* __restrict__ offsets, int n_vec, int64_t dense_cols,
                                 float* __restrict__ dense) {
    const int v = blockIdx.y;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vec || c >= dense_cols) return;
    const int64_t a = offsets[v], n = offsets[v + 1] - a;
    const uint32_t* px = pixels + a;
    const float* val = values + a;
    const uint32_t col = (uint32_t)c;
    // upper_bound over px[1..n) of col
    int64_t lo = 1, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (px[mid] <= col) lo = mid + 1;
        else hi = mid;
    }
    int64_t cur = lo - 1;
    float out;
    if (px[cur] > col) out = val[0];
    else if (cur + 1 < n) {
        float t = (float)(col - px[cur]) / (float)(px[cur + 1] - px[cur]);
        out = lerp_clamped(val[cur], val[cur + 1], t);
    } else out = val[cur];
    dense[(int64_t)v * dense_cols + c] = out;
}

// One merged pixel (apply_calibration.rs:520-572 per-pixel arithmetic in the
// reference's f32 order) from its CInt16 word w (I in the low half) and the
// range-LUT taps of its column in the line's two bracketing vectors.
__device__ __forceinline__ void cal_pixel(uint32_t w, float klo, float khi, float nlo, float nhi, const CalLine& L,
                                          float* o_s, float* o_n) {
    const int32_t i = (int16_t)(w & 0xffffu), q = (int16_t)(w >> 16);
    // i32 power; 2 * 32768^2 wraps as the reference's release build does
    const float p = (float)(int32_t)((uint32_t)(i * i) + (uint32_t)(q * q));
    const float k = L.cal_lo == L.cal_hi ? klo : lerp_clamped(klo, khi, L.cal_t);
    const float nr = L.nr_lo == L.nr_hi ? nlo : lerp_clamped(nlo, nhi, L.nr_t);
    const float k2 = k * k;
    const float valid = k2 > 0.0f ? 1.0f : 0.0f;
    const float inv = k2 > 0.0f ? 1.0f / k2 : 0.0f;
    const float nv = nr * L.az;
    *o_s = fmaxf(p - nv, 0.0f) * inv * valid;
    *o_n = nv * inv * valid;
}

// A workgroup takes kCalLines merged lines x kCalCols merged columns (8
// columns per thread, 256 apart, so every access of a wave is one contiguous
// run).  The sub-swaths own consecutive column ranges (swath s up to its
// seam, the last one to the end of the line), so the block is walked swath
// by swath and every swath constant is uniform.  Lines of one burst share
// their LUT brackets: the range-LUT taps of a column are loaded once and
// reloaded only when a line's bracket differs, and the raw words of all the
// block's lines are loaded before any is used.  Column arithmetic is 32-bit
// (the host rejects wider rasters).
// Block shape (build options, so that it can be tuned per device): on the
// full IW scene 4 lines x 8 columns per thread 1.85 ms, 8 x 4 1.92, 4 x 4
// 1.91, 8 x 8 1.96, 2 x 8 1.96, 16 x 2 2.05, 4 x 16 2.47, 16 x 4 4.54.
#ifndef SD_CAL_LINES
#define SD_CAL_LINES 4
#endif
#ifndef SD_CAL_PER
#define SD_CAL_PER 8
#endif
constexpr int kCalLines = SD_CAL_LINES;
constexpr int kCalPer = SD_CAL_PER;
constexpr int kCalCols = 256 * kCalPer;

__global__ __launch_bounds__(256) void cal_merge_kernel(CalParams P) {
    __shared__ CalLine ln[kMaxSwaths][kCalLines];
    const int c_begin = (int)blockIdx.x * kCalCols;
    const int c_end = min(c_begin + kCalCols, (int)P.out_cols);
    const int64_t l_begin = (int64_t)blockIdx.y * kCalLines;
    const int nl = (int)min((int64_t)kCalLines, P.out_lines - l_begin);
    for (int i = threadIdx.x; i < P.n_sw * kCalLines; i += 256) {
        const int s = i / kCalLines, l = i % kCalLines;
        if (l < nl) ln[s][l] = P.sw[s].lines[l_begin + l];
    }
    __syncthreads();
    const float qnan = __builtin_nanf("");
    int own_start = 0;
    for (int s = 0; s < P.n_sw; ++s) {
        const CalSwathDev& S = P.sw[s];
        const int own_end = s + 1 < P.n_sw ? max((int)S.own_end, own_start) : (int)P.out_cols;
        const int lo = max(own_start, c_begin), hi = min(own_end, c_end);
        own_start = own_end;
        if (lo >= hi) continue;
        const int vs = (int)S.valid_start, ve = (int)S.valid_end;
        const int shift = (int)(S.first_valid - S.out_off);  // TIFF column = merged column + shift
        const uint32_t* raw = reinterpret_cast<const uint32_t*>(S.raw);
        int oc[kCalPer];
        bool own[kCalPer], val[kCalPer];
#pragma unroll
        for (int j = 0; j < kCalPer; ++j) {
            oc[j] = lo + (int)threadIdx.x + 256 * j;
            own[j] = oc[j] < hi;
            val[j] = own[j] && oc[j] >= vs && oc[j] < ve;
        }
        uint32_t w[kCalLines][kCalPer];
#pragma unroll
        for (int l = 0; l < kCalLines; ++l) {
            const uint32_t* row = raw + (l < nl ? ln[s][l].tiff_row : 0) * S.raw_cols + shift;
#pragma unroll
            for (int j = 0; j < kCalPer; ++j) w[l][j] = (l < nl && val[j]) ? row[oc[j]] : 0u;
        }
        float klo[kCalPer], khi[kCalPer], nlo[kCalPer], nhi[kCalPer];
        int bk_lo = -1, bk_hi = -1, bn_lo = -1, bn_hi = -1;  // brackets the taps hold
#pragma unroll
        for (int l = 0; l < kCalLines; ++l) {
            if (l >= nl) break;
            const CalLine L = ln[s][l];
            if (L.cal_lo != bk_lo || L.cal_hi != bk_hi) {
                const float* a = S.cal_dense + (int64_t)L.cal_lo * S.dense_cols + shift;
                const float* b = S.cal_dense + (int64_t)L.cal_hi * S.dense_cols + shift;
#pragma unroll
                for (int j = 0; j < kCalPer; ++j) {
                    klo[j] = val[j] ? a[oc[j]] : 0.0f;
                    khi[j] = val[j] ? b[oc[j]] : 0.0f;
                }
                bk_lo = L.cal_lo;
                bk_hi = L.cal_hi;
            }
            if (L.nr_lo != bn_lo || L.nr_hi != bn_hi) {
                const float* a = S.nr_dense + (int64_t)L.nr_lo * S.dense_cols + shift;
                const float* b = S.nr_dense + (int64_t)L.nr_hi * S.dense_cols + shift;
#pragma unroll
                for (int j = 0; j < kCalPer; ++j) {
                    nlo[j] = val[j] ? a[oc[j]] : 0.0f;
                    nhi[j] = val[j] ? b[oc[j]] : 0.0f;
                }
                bn_lo = L.nr_lo;
                bn_hi = L.nr_hi;
            }
            float* const out_s = P.sigma0 + (l_begin + l) * P.out_cols;
            float* const out_n = P.nesz + (l_begin + l) * P.out_cols;
#pragma unroll
            for (int j = 0; j < kCalPer; ++j) {
                float vs_ = qnan, vn_ = qnan;
                if (val[j]) cal_pixel(w[l][j], klo[j], khi[j], nlo[j], nhi[j], L, &vs_, &vn_);
                if (own[j]) {
                    out_s[oc[j]] = vs_;
                    out_n[oc[j]] = vn_;
                }
            }
        }
    }
}

}  // namespace sd

extern "C" hipError_t sd_launch_lut_dense(const uint32_t* pixels, const float* values, const int64_t* offsets,
                                          int n_vec, int64_t dense_cols, float* dense, hipStream_t stream) {
    if (n_vec <= 0 || dense_cols <= 0) return hipSuccess;
    dim3 grid((unsigned)((dense_cols + 255) / 256), (unsigned)n_vec);
    hipLaunchKernelGGL(sd::lut_dense_kernel, grid, dim3(256), 0, stream, pixels, values, offsets, n_vec, dense_cols,
                       dense);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_cal_merge(const sd::CalParams* p, hipStream_t stream) {
    if (p->out_lines <= 0 || p->out_cols <= 0) return hipSuccess;
    dim3 grid((unsigned)((p->out_cols + sd::kCalCols - 1) / sd::kCalCols),
              (unsigned)((p->out_lines + sd::kCalLines - 1) / sd::kCalLines));
    hipLaunchKernelGGL(sd::cal_merge_kernel, grid, dim3(256), 0, stream, *p);
    return hipGetLastError();
}
