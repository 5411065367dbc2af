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

__device__ __forceinline__ void cal_pixel(const CalSwathDev& s, const CalLine& L, int64_t in_col, float* o_s,
                                          float* o_n) {
    const int64_t tc = in_col + s.first_valid;  // TIFF column
    const int16_t* px = s.raw + (L.tiff_row * s.raw_cols + tc) * 2;
    const int32_t i = px[0], q = px[1];
    const float p = (float)(i * i + q * q);
    const float* cd = s.cal_dense;
    const float* nd = s.nr_dense;
    float k = L.cal_lo == L.cal_hi ? cd[L.cal_lo * s.dense_cols + tc]
                                   : lerp_clamped(cd[L.cal_lo * s.dense_cols + tc], cd[L.cal_hi * s.dense_cols + tc], L.cal_t);
    float nr = L.nr_lo == L.nr_hi ? nd[L.nr_lo * s.dense_cols + tc]
                                  : lerp_clamped(nd[L.nr_lo * s.dense_cols + tc], nd[L.nr_hi * s.dense_cols + tc], L.nr_t);
    float k2 = k * k;
    float valid = k2 > 0.0f ? 1.0f : 0.0f;
    float inv = k2 > 0.0f ? 1.0f / k2 : 0.0f;
    float nv = nr * L.az;
    *o_s = fmaxf(p - nv, 0.0f) * inv * valid;
    *o_n = nv * inv * valid;
}

// One workgroup per merged line; consecutive lanes take consecutive merged
// columns, so every raw, LUT and output access of a wave is one contiguous
// run.  The line's per-swath constants are loaded once into LDS.
constexpr int kCalUnroll = 4;

__global__ __launch_bounds__(256) void cal_merge_kernel(CalParams P) {
    __shared__ CalSwathDev sw[kMaxSwaths];
    __shared__ CalLine ln[kMaxSwaths];
    const int64_t line = blockIdx.x;
    if (threadIdx.x < (unsigned)P.n_sw) {
        sw[threadIdx.x] = P.sw[threadIdx.x];
        ln[threadIdx.x] = P.sw[threadIdx.x].lines[line];
    }
    __syncthreads();
    const float qnan = __builtin_nanf("");
    const int n_sw = P.n_sw;
    float* const out_s = P.sigma0 + line * P.out_cols;
    float* const out_n = P.nesz + line * P.out_cols;
    for (int64_t c0 = threadIdx.x; c0 < P.out_cols; c0 += 256 * kCalUnroll) {
        float vs[kCalUnroll], vn[kCalUnroll];
#pragma unroll
        for (int j = 0; j < kCalUnroll; ++j) {
            const int64_t oc = c0 + 256 * j;
            vs[j] = qnan;
            vn[j] = qnan;
            if (oc >= P.out_cols) continue;
            int s = 0;
            while (s + 1 < n_sw && oc >= sw[s].own_end) ++s;
            if (oc >= sw[s].valid_start && oc < sw[s].valid_end) cal_pixel(sw[s], ln[s], oc - sw[s].out_off, &vs[j], &vn[j]);
        }
#pragma unroll
        for (int j = 0; j < kCalUnroll; ++j) {
            const int64_t oc = c0 + 256 * j;
            if (oc < P.out_cols) {
                out_s[oc] = vs[j];
                out_n[oc] = vn[j];
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
    hipLaunchKernelGGL(sd::cal_merge_kernel, dim3((unsigned)p->out_lines), dim3(256), 0, stream, *p);
    return hipGetLastError();
}
