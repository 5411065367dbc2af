// misc.hip — small streaming kernels: the dB epilogue (export.rs:59-94), the
// byte swap of big-endian CInt16 rows (slc_reader.rs:565-571),
// the synthetic CInt16 generator used for benchmark inputs, and a zero fill
// (used instead of hipMemsetAsync so the work can be captured in graphs).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pmath.hpp"

namespace sd {

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// to_db_inplace: NaN stays NaN; v <= floor → NaN (masked); else 10*log10(v).
__global__ __launch_bounds__(256) void to_db_kernel(float* d, int64_t n, float floor_lin,
                                                    unsigned long long* counts) {
    unsigned conv = 0, masked = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v = d[i];
        if (isnan(v)) continue;
        if (v <= floor_lin) {
            d[i] = __builtin_nanf("");
            ++masked;
        } else {
            d[i] = 10.0f * log10f(v);
            ++conv;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        conv += __shfl_xor(conv, o);
        masked += __shfl_xor(masked, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (conv) atomicAdd(&counts[0], (unsigned long long)conv);
        if (masked) atomicAdd(&counts[1], (unsigned long long)masked);
    }
}

// Swaps the two bytes of every 16-bit value (big-endian TIFF rows → host
// order), four values per 8-byte word; `n` counts 16-bit values.
__global__ __launch_bounds__(256) void swap16_kernel(uint16_t* d, int64_t n) {
    const int64_t words = n / 4;
    uint64_t* w = reinterpret_cast<uint64_t*>(d);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = w[i];
        w[i] = ((x & 0x00ff00ff00ff00ffull) << 8) | ((x >> 8) & 0x00ff00ff00ff00ffull);
    }
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n - words * 4) {
        const uint16_t x = d[words * 4 + t];
        d[words * 4 + t] = (uint16_t)((x >> 8) | (x << 8));
    }
}

// The same, one 16-bit value per thread, for buffers not 8-byte aligned.
__global__ __launch_bounds__(256) void swap16_scalar_kernel(uint16_t* d, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint16_t x = d[i];
        d[i] = (uint16_t)((x >> 8) | (x << 8));
    }
}

// One SRTM 1" tile (big-endian int16, -32768 = void) into its place in a
// float DEM raster: row r, column c of the tile → (row0 + r, col0 + c).
__global__ __launch_bounds__(256) void srtm_tile_kernel(const uint16_t* __restrict__ be, int dim, float* dst,
                                                        int64_t dst_w, int64_t row0, int64_t col0) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= dim) return;
    const uint16_t u = be[(size_t)r * dim + c];
    const int16_t v = (int16_t)(uint16_t)((u >> 8) | (u << 8));
    dst[(row0 + r) * dst_w + col0 + c] = v == -32768 ? __builtin_nanf("") : (float)v;
}

// Glo30Tile::load's sample rule (dem.rs:317-342) on decoded samples in host
// byte order: kind 0 = f32 (non-finite or equal to the nodata value → NaN),
// 1 = i16 and 2 = u16 (equal to the nodata value, already cast to the sample
// type, → NaN; otherwise widened to f32).
__global__ __launch_bounds__(256) void dem_values_kernel(const void* __restrict__ raw, int kind, int have_nd,
                                                         float nd_f, int32_t nd_i, float* __restrict__ dst,
                                                         int64_t n) {
    const float qnan = __builtin_nanf("");
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (kind == 0) {
            const float x = static_cast<const float*>(raw)[i];
            dst[i] = (!__builtin_isfinite(x) || (have_nd && x == nd_f)) ? qnan : x;
        } else {
            const int32_t x = kind == 1 ? (int32_t) static_cast<const int16_t*>(raw)[i]
                                        : (int32_t) static_cast<const uint16_t*>(raw)[i];
            dst[i] = (have_nd && x == nd_i) ? qnan : (float)x;
        }
    }
}

// TIFF's floating-point predictor (3) undone on the device, fused with
// Glo30Tile::load's f32 sample rule.  One workgroup per stored block row
// (blocks in file order, block_h rows of bw * 4 bytes each): the row's bytes
// are an inclusive prefix sum mod 256 of the stored ones; the result holds 4
// byte planes, most significant first, from which sample s is rebuilt and
// written to the image (cropped at the right and bottom edges).
__global__ __launch_bounds__(256) void fp_predictor_kernel(const uint8_t* __restrict__ raw, int64_t bw, int64_t bh,
                                                           int64_t nx, int64_t w, int64_t h, int have_nd, float nd,
                                                           float* __restrict__ dst) {
    extern __shared__ uint8_t row[];  // bw * 4 bytes
    __shared__ uint32_t part[256];
    const int64_t k = blockIdx.x / bh, r = blockIdx.x % bh;
    const int64_t by = k / nx, bx = k % nx;
    const int64_t y = by * bh + r;
    if (y >= h) return;  // the same for the whole workgroup
    const int64_t rb = bw * 4;
    const uint8_t* src = raw + (size_t)blockIdx.x * rb;
    // each thread a run of consecutive bytes; the runs' sums scanned in LDS
    const int64_t chunk = (rb + 255) / 256;
    const int64_t b0 = min((int64_t)threadIdx.x * chunk, rb), b1 = min(b0 + chunk, rb);
    uint32_t acc = 0;
    for (int64_t i = b0; i < b1; ++i) acc += src[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - acc;  // exclusive: the bytes before this run
    for (int64_t i = b0; i < b1; ++i) {
        run += src[i];
        row[i] = (uint8_t)run;
    }
    __syncthreads();
    const int64_t x0 = bx * bw, cols = min(bw, w - x0);
    const float qnan = __builtin_nanf("");
    for (int64_t s = threadIdx.x; s < cols; s += 256) {
        const uint32_t u = (uint32_t)row[s] << 24 | (uint32_t)row[bw + s] << 16 | (uint32_t)row[2 * bw + s] << 8 |
                           (uint32_t)row[3 * bw + s];
        const float x = __uint_as_float(u);
        dst[y * w + x0 + s] = (!__builtin_isfinite(x) || (have_nd && x == nd)) ? qnan : x;
    }
}

__global__ __launch_bounds__(256) void fill_nan_kernel(float* d, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = __builtin_nanf("");
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Same function as sardine_amd.synthetic.complex_gaussian_dn.
__global__ __launch_bounds__(256) void synth_cint16_kernel(int16_t* out, int64_t n, uint64_t seed, uint64_t base,
                                                           double sigma) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t k = seed * 0x9E3779B97F4A7C15ull + base + (uint64_t)i;
        uint64_t h1 = splitmix64(k * 2ull), h2 = splitmix64(k * 2ull + 1ull);
        double u1 = (double)(h1 >> 11) * (1.0 / 9007199254740992.0);
        double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log1p(-u1));
        double ang = 2.0 * 3.14159265358979323846 * u2;
        double re = floor(sigma * r * cos(ang) + 0.5), im = floor(sigma * r * sin(ang) + 0.5);
        re = fmin(fmax(re, -32767.0), 32767.0);
        im = fmin(fmax(im, -32767.0), 32767.0);
        reinterpret_cast<short2*>(out)[i] = make_short2((short)re, (short)im);
    }
}

}  // namespace sd

extern "C" hipError_t sd_launch_zero_u64(unsigned long long* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sd::zero_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_swap16(uint16_t* d, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const bool aligned = reinterpret_cast<uintptr_t>(d) % 8 == 0;
    int64_t blocks = ((aligned ? n / 4 : n) + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 65536) blocks = 65536;
    if (aligned) hipLaunchKernelGGL(sd::swap16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, n);
    else hipLaunchKernelGGL(sd::swap16_scalar_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, n);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_srtm_tile(const uint16_t* be, int dim, float* dst, int64_t dst_w, int64_t row0,
                                          int64_t col0, hipStream_t s) {
    hipLaunchKernelGGL(sd::srtm_tile_kernel, dim3((unsigned)((dim + 255) / 256), (unsigned)dim), dim3(256), 0, s, be,
                       dim, dst, dst_w, row0, col0);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_dem_values(const void* raw, int kind, int have_nd, float nd_f, int32_t nd_i,
                                           float* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(sd::dem_values_kernel, dim3((unsigned)blocks), dim3(256), 0, s, raw, kind, have_nd, nd_f, nd_i,
                       dst, n);
    return hipGetLastError();
}

// raw: blocks_x * blocks_y blocks of bh rows of bw * 4 bytes (bw * 4 <=
// 32768, the dynamic LDS row); blocks_x * blocks_y * bh < 2^31 rows.
extern "C" hipError_t sd_launch_fp_predictor(const uint8_t* raw, int64_t bw, int64_t bh, int64_t blocks_x,
                                             int64_t blocks_y, int64_t w, int64_t h, int have_nd, float nd,
                                             float* dst, hipStream_t s) {
    const int64_t rows = blocks_x * blocks_y * bh;
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(sd::fp_predictor_kernel, dim3((unsigned)rows), dim3(256), (size_t)(bw * 4), s, raw, bw, bh,
                       blocks_x, w, h, have_nd, nd, dst);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_fill_nan(float* d, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(sd::fill_nan_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, n);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_to_db(float* d, int64_t n, float floor_lin, unsigned long long* counts,
                                      hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(sd::to_db_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, n, floor_lin, counts);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_synth_cint16(int16_t* out, int64_t n, uint64_t seed, uint64_t base, double sigma,
                                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 262144) blocks = 262144;
    hipLaunchKernelGGL(sd::synth_cint16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, out, n, seed, base, sigma);
    return hipGetLastError();
}

// The portable libm (pmath.hpp) and the IEEE division and square root on the
// device, element by element: a diagnostic that lets the host compare the
// device's bits with its own evaluation of the same functions.
namespace sd {
__global__ __launch_bounds__(256) void pmath_kernel(int fn, const double* x, const double* y, double* out,
                                                    int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double a = x[i], b = y[i];
    double r;
    switch (fn) {
        case 0: r = pm::sin(a); break;
        case 1: r = pm::cos(a); break;
        case 2: r = pm::exp(a); break;
        case 3: r = pm::log(a); break;
        case 4: r = pm::sinh(a); break;
        case 5: r = pm::cosh(a); break;
        case 6: r = pm::atanh(a); break;
        case 7: r = pm::atan(a); break;
        case 8: r = pm::atan2(b, a); break;
        case 9: r = a / b; break;
        default: r = sqrt(a); break;
    }
    out[i] = r;
}
}  // namespace sd

extern "C" hipError_t sd_launch_pmath(int fn, const double* x, const double* y, double* out, int64_t n,
                                      hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sd::pmath_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fn, x, y, out, n);
    return hipGetLastError();
}
