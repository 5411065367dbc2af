// tc.hip — backward Range-Doppler terrain correction on gfx950 (FP64 path).
// Reference: sardine/src/terrain_correction.rs:1245-1836.
//
// Per pixel: DEM bilinear → geoid → ECEF → zero-Doppler Newton on the
// 8-point Lagrange orbit → slant range → radar (line, sample) → resample →
// optional Small-2011 geometry / flattening → NESZ mask → σ⁰ / mask / cos_lia
// writes.  Counters and the Newton histogram are accumulated per lane and
// added to global counters once per lane.
//
// Work decomposition and the per-row seed chain.  The reference seeds each
// pixel's Newton solve with the previous converged pixel's time in the same
// row (terrain_correction.rs:1443,1580) and truncates every step to whole µs,
// so a converged time is floor or ceil of the root depending on the side the
// seed comes from.  Here every lane runs that chain sequentially over one
// segment of L consecutive columns of a row, which costs the reference's ~2
// orbit evaluations per pixel.  A wave holds 64 adjacent rows of the same
// segment, so its DEM and radar gathers fall on neighbouring cache lines.
// Only a segment's starting seed is unknown in parallel.  A converged time
// is floor or ceil of the root by the side the seed approaches from, and the
// chain approaches each root from the previous pixel's; phase A therefore
// runs a short chain over the last few columns of the previous segment,
// started at the row seed, and takes its end as the speculative seed.  Each
// segment records its seed and its last converged
// time; tc_fixup_kernel then walks every row's segments in order, and where
// a seed differs from the chained value to its left it replays the old and
// the corrected chains in lockstep, rewriting pixels until both converge to
// the same µs (after which the rest of the segment is unchanged).  The output
// is therefore exactly the sequential chain of the reference.
#include "tc_device.hpp"

namespace sd {

// One lane per AOI row (row_prep).
__global__ __launch_bounds__(256) void tc_rows_kernel(RowPrepParams q) {
    const int64_t rr = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (rr < q.rows) row_prep(q, rr);
}

// One lane per (AOI row, anchor point).
__global__ __launch_bounds__(256) void utm_anchor_kernel(TcParams p) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < p.rows * p.anchor_stride) anchor_point(p, i);
}

// One lane per AOI column (the column's project_row blend).
__global__ __launch_bounds__(256) void utm_col_kernel(TcParams p) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.cols) return;
    ColBlend b;
    col_blend(p, p.col0 + i, &b.a, &b.t);
    b.pad = 0;
    p.col_blend[i] = b;
}

// One lane per AOI column of an EPSG:4326 grid (the column's longitude terms).
__global__ __launch_bounds__(256) void wgs_col_kernel(TcParams p) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < p.cols) p.col_geo[i] = col_geo(p, p.col0 + i);
}

// The orbit's LDS copy is dynamic, sized to the orbit (p.orbit_lds bytes:
// 13 KiB for 17 state vectors, at most 40 KiB), so that three workgroups fit
// a CU where the chain runs three waves per SIMD.
template <int RS, bool UTM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(chain_waves<RS, UTM>()))) void tc_segment_kernel(
    TcParams p) {
    extern __shared__ __attribute__((aligned(16))) unsigned char tc_orbit_lds[];
    OrbitDev& orb = *reinterpret_cast<OrbitDev*>(tc_orbit_lds);
    __shared__ WaveStage stage[4];
    load_orbit(p, orb);
    const int w = threadIdx.x >> 6;
    const int64_t seg = (int64_t)blockIdx.y * 4 + w;
    if (seg >= p.nseg) return;  // whole wave; no block barriers below
    segment_wave<RS, UTM, true, 1>(p, orb, stage[w], (int64_t)blockIdx.x * 64, seg);
}

// One wave per AOI row (check_row).
__global__ __launch_bounds__(256) void tc_check_kernel(TcParams p) {
    const int64_t rr = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (rr < p.rows) check_row(p, rr);
}

// The tie list's replays, one lane per entry (the list's length is on the
// device; the grid covers its capacity).  The counter keeps counting past
// the capacity (those entries were not stored, and their rows went to the
// sequential walk), so the entries read are the first min(count, capacity).
template <int RS, bool UTM, bool WRITE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void tc_tie_kernel(TcParams p) {
    __shared__ OrbitDev orb;
    const int64_t n = min((int64_t)p.counters[kCntTieList], (int64_t)p.tie_cap);
    if ((int64_t)blockIdx.x * 256 >= n) return;  // whole workgroup
    load_orbit(p, orb);
    tie_item<RS, UTM, WRITE>(p, orb, (int64_t)blockIdx.x * 256 + threadIdx.x, n);
}

template <int RS, bool UTM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void tc_fixup_kernel(TcParams p) {
    __shared__ OrbitDev orb;
    load_orbit(p, orb);
    const int64_t rr = (int64_t)blockIdx.x * 256 + threadIdx.x;
    fixup_row<RS, UTM>(p, orb, rr, rr < p.rows);
}

// The terrain correction's DEM lookup at given points (diagnostics).
__global__ __launch_bounds__(256) void dem_probe_kernel(TcParams p, const double* lat, const double* lon, float* out,
                                                        int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[i] = p.dem_tiled ? dem_height_mosaic(mos_args(p), lat[i], lon[i]) : dem_sample(p, lat[i], lon[i]);
}

}  // namespace sd

namespace {
template <int RS, bool UTM>
void launch_chain(const sd::TcParams* p, hipStream_t stream, hipEvent_t seg_start, hipEvent_t seg_end) {
    dim3 grid((unsigned)((p->rows + 63) / 64), (unsigned)((p->nseg + 3) / 4));
    if (seg_start) (void)hipEventRecord(seg_start, stream);
    const size_t orbit_lds = ((size_t)p->orbit_lds + 15) & ~(size_t)15;
    hipLaunchKernelGGL((sd::tc_segment_kernel<RS, UTM>), grid, dim3(256), orbit_lds, stream, *p);
    if (seg_end) (void)hipEventRecord(seg_end, stream);
    hipLaunchKernelGGL(sd::tc_check_kernel, dim3((unsigned)((p->rows + 3) / 4)), dim3(256), 0, stream, *p);
    if (p->tie_list && p->tie_cap > 0) {
        const dim3 g((unsigned)((p->tie_cap + 255) / 256));
        hipLaunchKernelGGL((sd::tc_tie_kernel<RS, UTM, false>), g, dim3(256), 0, stream, *p);
        hipLaunchKernelGGL((sd::tc_tie_kernel<RS, UTM, true>), g, dim3(256), 0, stream, *p);
    }
    hipLaunchKernelGGL((sd::tc_fixup_kernel<RS, UTM>), dim3((unsigned)((p->rows + 255) / 256)), dim3(256), 0, stream,
                       *p);
}
}  // namespace

extern "C" hipError_t sd_launch_dem_probe(const sd::TcParams* p, const double* lat, const double* lon, float* out,
                                          int64_t n, hipStream_t stream) {
    if (n > 0)
        hipLaunchKernelGGL(sd::dem_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *p, lat, lon,
                           out, n);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_tc_rows(const sd::RowPrepParams* q, hipStream_t stream) {
    if (q->rows > 0)
        hipLaunchKernelGGL(sd::tc_rows_kernel, dim3((unsigned)((q->rows + 255) / 256)), dim3(256), 0, stream, *q);
    return hipGetLastError();
}

// seg_start / seg_end: optional events recorded around the segment kernel.
// project_row anchors of the AOI rows (UTM grids only).
extern "C" hipError_t sd_launch_tc_anchors(const sd::TcParams* p, hipStream_t stream) {
    if (p->crs == 0) {
        if (p->col_geo)
            hipLaunchKernelGGL(sd::wgs_col_kernel, dim3((unsigned)((p->cols + 255) / 256)), dim3(256), 0, stream, *p);
        return hipGetLastError();
    }
    const int64_t n = p->rows * p->anchor_stride;
    hipLaunchKernelGGL(sd::utm_anchor_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *p);
    if (p->col_blend)
        hipLaunchKernelGGL(sd::utm_col_kernel, dim3((unsigned)((p->cols + 255) / 256)), dim3(256), 0, stream, *p);
    return hipGetLastError();
}

extern "C" hipError_t sd_launch_tc_fp64(const sd::TcParams* p, hipStream_t stream, hipEvent_t seg_start,
                                        hipEvent_t seg_end) {
    const bool utm = p->crs != 0;
    // EPSG:4326 chains read their grid's column table (UTM ones fall back to
    // the per-pixel blend without one); the orbit's LDS copy must be sized
    if (!utm && p->col_geo == nullptr) return hipErrorInvalidValue;
    if (p->orbit_lds <= 0 || (size_t)p->orbit_lds > sizeof(sd::OrbitDev)) return hipErrorInvalidValue;
    hipError_t e = sd_launch_tc_anchors(p, stream);
    if (e != hipSuccess) return e;
    // RS 0..2: the resampler in FP64; 3..5: the same in the mixed form
    switch ((p->resampling + (p->mixed ? 3 : 0)) * 2 + (utm ? 1 : 0)) {
        case 0: launch_chain<0, false>(p, stream, seg_start, seg_end); break;
        case 1: launch_chain<0, true>(p, stream, seg_start, seg_end); break;
        case 2: launch_chain<1, false>(p, stream, seg_start, seg_end); break;
        case 3: launch_chain<1, true>(p, stream, seg_start, seg_end); break;
        case 4: launch_chain<2, false>(p, stream, seg_start, seg_end); break;
        case 5: launch_chain<2, true>(p, stream, seg_start, seg_end); break;
        case 6: launch_chain<3, false>(p, stream, seg_start, seg_end); break;
        case 7: launch_chain<3, true>(p, stream, seg_start, seg_end); break;
        case 8: launch_chain<4, false>(p, stream, seg_start, seg_end); break;
        case 9: launch_chain<4, true>(p, stream, seg_start, seg_end); break;
        case 10: launch_chain<5, false>(p, stream, seg_start, seg_end); break;
        default: launch_chain<5, true>(p, stream, seg_start, seg_end); break;
    }
    return hipGetLastError();
}
