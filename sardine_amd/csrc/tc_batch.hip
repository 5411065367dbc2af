// tc_batch.hip — many terrain-correction jobs (AOI tiles of uneven size, on
// one or several resident scenes, with mixed resamplers and CRSs) in one set
// of launches.  Each pass indexes the concatenation of the jobs' work; the
// chain pass is persistent: a fixed grid of workgroups pulls (job, 64-row
// block, 4-segment group) items from a device work queue until it is empty,
// so small and large tiles share the GPU without per-job launches.  Grid
// sizes depend only on the tuning, and every count is read from the device
// job table, so the whole sequence can be captured once as a HIP graph and
// replayed with new tables.  Per-pixel work is exactly tc.hip's.
#include "tc_device.hpp"

namespace sd {

// Largest j with begin(j) <= x (begin is non-decreasing in j, begin(0) = 0).
template <typename F>
__device__ __forceinline__ int find_job(int n_jobs, int64_t x, F begin) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin(mid) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void batch_rows_kernel(const BatchHeader* __restrict__ H,
                                                         const BatchJob* __restrict__ J) {
    const int n_jobs = H->n_jobs;
    const int64_t n = H->n_rows;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int j = find_job(n_jobs, i, [&](int k) { return J[k].row_begin; });
        row_prep(J[j].q, i - J[j].row_begin);
    }
}

__global__ __launch_bounds__(256) void batch_anchor_kernel(const BatchHeader* __restrict__ H,
                                                           const BatchJob* __restrict__ J) {
    const int n_jobs = H->n_jobs;
    const int64_t n = H->n_anchor;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int j = find_job(n_jobs, i, [&](int k) { return J[k].anchor_begin; });
        const TcParams& p = J[j].p;
        const int64_t k = i - J[j].anchor_begin;
        if (p.crs == 0) {  // EPSG:4326: the column terms only
            p.col_geo[k] = col_geo(p, p.col0 + k);
            continue;
        }
        const int64_t na = p.rows * p.anchor_stride;
        if (k < na) {
            anchor_point(p, k);
        } else {  // then the job's column blends
            ColBlend b;
            col_blend(p, p.col0 + (k - na), &b.a, &b.t);
            b.pad = 0;
            p.col_blend[k - na] = b;
        }
    }
}

template <int RS, bool UTM>
__device__ __forceinline__ void batch_item(const TcParams& p, const OrbitDev& orb, WaveStage& st, int64_t rr0,
                                           int64_t seg) {
    segment_wave<RS, UTM, false, 2>(p, orb, st, rr0, seg);   // the job table is read-only here
}

// Persistent chain pass: `queue` is zero when the launch starts.
__global__ __launch_bounds__(256) SD_SEGMENT_OCCUPANCY void batch_tc_kernel(const BatchHeader* __restrict__ H,
                                                                            const BatchJob* __restrict__ J,
                                                                            unsigned int* queue) {
    __shared__ OrbitDev orb;
    __shared__ WaveStage stage[4];
    __shared__ unsigned int s_item;
    const int w = threadIdx.x >> 6;
    const int n_jobs = H->n_jobs;
    const int64_t n_items = H->n_tc;
    const OrbitDev* loaded = nullptr;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
        __syncthreads();
        const int64_t item = (int64_t)s_item;
        if (item >= n_items) break;  // the same value for the whole workgroup
        const int j = __builtin_amdgcn_readfirstlane(find_job(n_jobs, item, [&](int k) { return J[k].tc_begin; }));
        const BatchJob& job = J[j];
        if (job.p.orbit != loaded) {  // workgroup-uniform
            load_orbit(job.p, orb);   // ends with a barrier
            loaded = job.p.orbit;
        }
        const int64_t local = item - job.tc_begin;
        const int64_t rb = local % job.tc_row_blocks, sg = local / job.tc_row_blocks;
        const int64_t seg = sg * 4 + w;
        if (seg < job.p.nseg) {
            const int64_t rr0 = rb * 64;
            switch (job.cls) {
                case 0: batch_item<0, false>(job.p, orb, stage[w], rr0, seg); break;
                case 1: batch_item<0, true>(job.p, orb, stage[w], rr0, seg); break;
                case 2: batch_item<1, false>(job.p, orb, stage[w], rr0, seg); break;
                case 3: batch_item<1, true>(job.p, orb, stage[w], rr0, seg); break;
                case 4: batch_item<2, false>(job.p, orb, stage[w], rr0, seg); break;
                default: batch_item<2, true>(job.p, orb, stage[w], rr0, seg); break;
            }
        }
        __syncthreads();  // the item slot and the orbit are rewritten next
    }
}

// Row check before the fix-up pass: one wave per row of every job.
__global__ __launch_bounds__(256) void batch_check_kernel(const BatchHeader* __restrict__ H,
                                                          const BatchJob* __restrict__ J) {
    const int n_jobs = H->n_jobs;
    const int64_t n = H->n_rows;
    const int64_t wave0 = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6), waves = (int64_t)gridDim.x * 4;
    for (int64_t i = wave0; i < n; i += waves) {
        const int j = __builtin_amdgcn_readfirstlane(find_job(n_jobs, i, [&](int k) { return J[k].row_begin; }));
        check_row(J[j].p, i - J[j].row_begin);
    }
}

// Fix-up pass: blocks of 256 rows of one job each.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void batch_fixup_kernel(
    const BatchHeader* __restrict__ H, const BatchJob* __restrict__ J) {
    __shared__ OrbitDev orb;
    const int n_jobs = H->n_jobs;
    const int64_t n = H->n_fix;
    const OrbitDev* loaded = nullptr;
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        const int j = __builtin_amdgcn_readfirstlane(find_job(n_jobs, b, [&](int k) { return J[k].fix_begin; }));
        const BatchJob& job = J[j];
        if (job.p.orbit != loaded) {
            __syncthreads();  // everyone is done with the previous orbit
            load_orbit(job.p, orb);
            loaded = job.p.orbit;
        }
        const int64_t rr = (b - job.fix_begin) * 256 + threadIdx.x;
        const bool valid = rr < job.p.rows;
        switch (job.cls) {
            case 0: fixup_row<0, false>(job.p, orb, rr, valid); break;
            case 1: fixup_row<0, true>(job.p, orb, rr, valid); break;
            case 2: fixup_row<1, false>(job.p, orb, rr, valid); break;
            case 3: fixup_row<1, true>(job.p, orb, rr, valid); break;
            case 4: fixup_row<2, false>(job.p, orb, rr, valid); break;
            default: fixup_row<2, true>(job.p, orb, rr, valid); break;
        }
    }
}

}  // namespace sd

// The batch's TC passes: rows, anchors, persistent chain, fix-up.  `queue`
// must be zero on entry (stream order).
extern "C" hipError_t sd_launch_batch_tc(const sd::BatchHeader* H, const sd::BatchJob* J, unsigned int* queue,
                                         int n_cu, int tc_groups_per_cu, hipStream_t stream) {
    hipLaunchKernelGGL(sd::batch_rows_kernel, dim3(256), dim3(256), 0, stream, H, J);
    hipLaunchKernelGGL(sd::batch_anchor_kernel, dim3(1024), dim3(256), 0, stream, H, J);
    hipLaunchKernelGGL(sd::batch_tc_kernel, dim3((unsigned)(n_cu * tc_groups_per_cu)), dim3(256), 0, stream, H, J,
                       queue);
    hipLaunchKernelGGL(sd::batch_check_kernel, dim3((unsigned)(n_cu * 8)), dim3(256), 0, stream, H, J);
    hipLaunchKernelGGL(sd::batch_fixup_kernel, dim3((unsigned)(n_cu * 2)), dim3(256), 0, stream, H, J);
    return hipGetLastError();
}
