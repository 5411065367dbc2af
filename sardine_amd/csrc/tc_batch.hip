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

// Persistent chain pass of one job class (resampler RS, UTM or EPSG:4326),
// at that class's occupancy: each class has its own kernel, so the bilinear
// and bicubic jobs run at three waves per SIMD while the Lanczos-3 ones run
// at two.  The job table is ordered by class, so a class's items are the
// contiguous range [tc_cls[c], tc_cls[c + 1]) of the combined index space;
// queue[c] counts the items taken (zero when the launch starts).  The
// orbit's LDS copy is dynamic, sized by the host to the largest orbit of
// the batch's scenes.
template <int RS, bool UTM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(chain_waves<RS, UTM>()))) void batch_tc_kernel(
    const BatchHeader* __restrict__ H, const BatchJob* __restrict__ J, unsigned int* queue) {
    extern __shared__ __attribute__((aligned(16))) unsigned char batch_orbit_lds[];
    OrbitDev& orb = *reinterpret_cast<OrbitDev*>(batch_orbit_lds);
    __shared__ WaveStage stage[4];
    __shared__ unsigned int s_item;
    constexpr int CLS = RS * 2 + (UTM ? 1 : 0);
    const int w = threadIdx.x >> 6;
    const int n_jobs = H->n_jobs;
    const int64_t i0 = H->tc_cls[CLS], n_items = H->tc_cls[CLS + 1] - i0;
    const OrbitDev* loaded = nullptr;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue + CLS, 1u);
        __syncthreads();
        const int64_t taken = (int64_t)s_item;
        if (taken >= n_items) break;  // the same value for the whole workgroup
        const int64_t item = i0 + taken;
        const int j = __builtin_amdgcn_readfirstlane(find_job(n_jobs, item, [&](int k) { return J[k].tc_begin; }));
        const BatchJob& job = J[j];
        if (job.p.orbit != loaded) {  // workgroup-uniform
            load_orbit(job.p, orb);   // ends with a barrier
            loaded = job.p.orbit;
        }
        const int64_t local = item - job.tc_begin;
        const int64_t rb = local % job.tc_row_blocks, sg = local / job.tc_row_blocks;
        const int64_t seg = sg * 4 + w;
        if (seg < job.p.nseg) segment_wave<RS, UTM, false, 2>(job.p, orb, stage[w], rb * 64, seg);
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

namespace {
// One class's persistent chain kernel: tc_groups_per_cu 1 puts one
// workgroup on each CU, more fills the class's occupancy.
template <int RS, bool UTM>
void launch_batch_chain(const sd::BatchHeader* H, const sd::BatchJob* J, unsigned int* queue, int n_cu,
                        int tc_groups_per_cu, size_t orbit_lds, hipStream_t stream) {
    const int groups = tc_groups_per_cu <= 1 ? 1 : sd::chain_waves<RS, UTM>();
    hipLaunchKernelGGL((sd::batch_tc_kernel<RS, UTM>), dim3((unsigned)(n_cu * groups)), dim3(256), orbit_lds, stream,
                       H, J, queue);
}
}  // namespace

// The batch's TC passes: rows, anchors, the persistent chain kernels (one per
// job class), fix-up.  queue[0..5] (the classes' item counters) must be zero
// on entry (stream order); orbit_lds: bytes of the largest orbit's LDS copy.
extern "C" hipError_t sd_launch_batch_tc(const sd::BatchHeader* H, const sd::BatchJob* J, unsigned int* queue,
                                         int n_cu, int tc_groups_per_cu, int orbit_lds, hipStream_t stream) {
    if (orbit_lds <= 0 || (size_t)orbit_lds > sizeof(sd::OrbitDev)) return hipErrorInvalidValue;
    const size_t lds = ((size_t)orbit_lds + 15) & ~(size_t)15;
    hipLaunchKernelGGL(sd::batch_rows_kernel, dim3(256), dim3(256), 0, stream, H, J);
    hipLaunchKernelGGL(sd::batch_anchor_kernel, dim3(1024), dim3(256), 0, stream, H, J);
    launch_batch_chain<0, false>(H, J, queue, n_cu, tc_groups_per_cu, lds, stream);
    launch_batch_chain<0, true>(H, J, queue, n_cu, tc_groups_per_cu, lds, stream);
    launch_batch_chain<1, false>(H, J, queue, n_cu, tc_groups_per_cu, lds, stream);
    launch_batch_chain<1, true>(H, J, queue, n_cu, tc_groups_per_cu, lds, stream);
    launch_batch_chain<2, false>(H, J, queue, n_cu, tc_groups_per_cu, lds, stream);
    launch_batch_chain<2, true>(H, J, queue, n_cu, tc_groups_per_cu, lds, stream);
    hipLaunchKernelGGL(sd::batch_check_kernel, dim3((unsigned)(n_cu * 8)), dim3(256), 0, stream, H, J);
    hipLaunchKernelGGL(sd::batch_fixup_kernel, dim3((unsigned)(n_cu * 2)), dim3(256), 0, stream, H, J);
    return hipGetLastError();
}
