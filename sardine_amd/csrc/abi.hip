// abi.hip — host side of the C ABI declared in include/sardine_hip.h.
//
// This is the MI355X replacement of the Rust host code of the path: it
// performs the reference's host-level set-up exactly (grid extent, seed LUT,
// deburst/merge integer layout, LUT coverage checks, per-row constants),
// stages the small metadata, launches the kernels and maps device-side
// outcomes back onto the reference's error taxonomy.  There is no CPU
// compute path: without a device every compute entry point fails with
// SARDINE_ERR_NO_DEVICE.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sardine_hip.h"
#include "cal_params.hpp"
#include "dem_io.hpp"
#include "geom.hpp"
#include "tc_params.hpp"

extern "C" hipError_t sd_launch_tc_fp64(const sd::TcParams* p, hipStream_t stream, hipEvent_t seg_start,
                                        hipEvent_t seg_end);
extern "C" hipError_t sd_launch_tc_rows(const sd::RowPrepParams* q, hipStream_t stream);
extern "C" hipError_t sd_launch_tc_anchors(const sd::TcParams* p, hipStream_t stream);
extern "C" hipError_t sd_launch_tc_fp32(const sd::TcParams* p, float* ends, hipStream_t stream);
extern "C" hipError_t sd_launch_lut_dense(const uint32_t* pixels, const float* values, const int64_t* offsets,
                                          int n_vec, int64_t dense_cols, float* dense, hipStream_t stream);
extern "C" hipError_t sd_launch_cal_merge(const sd::CalParams* p, hipStream_t stream);
extern "C" hipError_t sd_launch_dem_probe(const sd::TcParams* p, const double* lat, const double* lon, float* out,
                                          int64_t n, hipStream_t stream);
extern "C" hipError_t sd_launch_speckle(int kind, const float* in, float* out, int64_t cols, int64_t rows, int half,
                                        float enl, float damping, hipStream_t stream);
extern "C" hipError_t sd_launch_zero_u64(unsigned long long* p, int64_t n, hipStream_t s);
extern "C" hipError_t sd_launch_to_db(float* d, int64_t n, float floor_lin, unsigned long long* counts,
                                      hipStream_t s);
namespace sd {
struct CogLevelDev {  // layout of the struct in export.hip
    float* tiles;
    int64_t cols, rows;
    int64_t tiles_x;
    int64_t pad_cols, pad_rows;
};
}  // namespace sd
extern "C" hipError_t sd_launch_cog_level(const float* src, int src_tiled, const sd::CogLevelDev* S,
                                          const sd::CogLevelDev* D, const sd::CogLevelDev* copy, int lg,
                                          hipStream_t s);
extern "C" hipError_t sd_launch_multilook(const float* data, const float* nesz, int64_t in_samples,
                                          int64_t out_lines, int64_t out_samples, int r, int a, float* od,
                                          float* on, hipStream_t s);
extern "C" hipError_t sd_launch_grd(const float* data, int64_t in_samples, const void* map, int64_t out_lines,
                                    int64_t out_samples, int a, float* out, hipStream_t s);
extern "C" hipError_t sd_launch_swap16(uint16_t* d, int64_t n, hipStream_t s);
extern "C" hipError_t sd_launch_srtm_tile(const uint16_t* be, int dim, float* dst, int64_t dst_w, int64_t row0,
                                          int64_t col0, hipStream_t s);
extern "C" hipError_t sd_launch_fill_nan(float* d, int64_t n, hipStream_t s);
extern "C" hipError_t sd_launch_pmath(int fn, const double* x, const double* y, double* out, int64_t n,
                                      hipStream_t s);
extern "C" hipError_t sd_launch_dem_values(const void* raw, int kind, int have_nd, float nd_f, int32_t nd_i,
                                           float* dst, int64_t n, hipStream_t s);
extern "C" hipError_t sd_launch_fp_predictor(const uint8_t* raw, int64_t bw, int64_t bh, int64_t blocks_x,
                                             int64_t blocks_y, int64_t w, int64_t h, int have_nd, float nd,
                                             float* dst, hipStream_t s);
extern "C" hipError_t sd_launch_synth_cint16(int16_t* out, int64_t n, uint64_t seed, uint64_t base, double sigma,
                                             hipStream_t s);

namespace {

thread_local std::string g_last_error;
thread_local sardine_error_detail g_last_detail;
thread_local bool g_profile = false;
std::atomic<int64_t> g_tc_seg{0};  // columns per lane segment, 0 = automatic (sardine_set_tc_segment)
std::atomic<int32_t> g_tc_lead{2};  // phase-A lead columns (sardine_set_tc_lead)
std::atomic<double> g_tc_tie_band{sd::kTieBandUs};  // µs (sardine_set_tc_tie_band)
constexpr int kFamilies = 9;

// Brackets one kernel launch with events when profiling is on.  Nothing
// waits here: sardine_last_kernel_ms waits for the end event when read.
// Kernel timers: per family, a ring of event pairs recorded around the
// family's main kernel on its launch stream.  Recording never blocks; a
// reader synchronises on the events it reads, so a caller can time a run
// without waiting between its steps and read the durations afterwards.
constexpr int kTimerRing = 64;
struct TimerEntry {
    hipEvent_t a = nullptr, b = nullptr;
    bool pending = false;
    double ms = -1.0;
};
struct TimerSlot {
    TimerEntry e[kTimerRing];
    int64_t next = 0;  // entries next-kTimerRing .. next-1 are the most recent
};
thread_local TimerSlot g_slots[kFamilies];

// The next ring entry of a family, with its events created; null when
// profiling is off or event creation fails.
TimerEntry* timer_next(int family) {
    if (!g_profile) return nullptr;
    TimerSlot& t = g_slots[family];
    TimerEntry& e = t.e[t.next % kTimerRing];
    if (!e.a && hipEventCreate(&e.a) != hipSuccess) return nullptr;
    if (!e.b && hipEventCreate(&e.b) != hipSuccess) return nullptr;
    e.pending = false;
    e.ms = -1.0;
    ++t.next;
    return &e;
}

double timer_read(TimerEntry& e) {
    if (e.pending) {
        e.pending = false;
        float ms = 0.0f;
        if (hipEventSynchronize(e.b) == hipSuccess && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) e.ms = ms;
    }
    return e.ms;
}

struct KernelTimer {
    TimerEntry* entry = nullptr;
    hipStream_t s;
    KernelTimer(int family, hipStream_t stream) : s(stream) {
        TimerEntry* e = timer_next(family);
        if (e && hipEventRecord(e->a, s) == hipSuccess) entry = e;
    }
    void stop() {
        if (entry && hipEventRecord(entry->b, s) == hipSuccess) entry->pending = true;
        entry = nullptr;
    }
};

sardine_status fail(sardine_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    std::memset(&g_last_detail, 0, sizeof g_last_detail);
    g_last_detail.status = st;
    return st;
}

#define SD_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t _e = (call);                                                                       \
        if (_e != hipSuccess)                                                                         \
            return fail(_e == hipErrorOutOfMemory ? SARDINE_ERR_OUT_OF_MEMORY : SARDINE_ERR_HIP,      \
                        "%s failed: %s", #call, hipGetErrorString(_e));                               \
    } while (0)

// Host exceptions never cross the C ABI: an allocation failure or any other
// std::exception raised by host code (file parsing, planning, staging)
// becomes a status and a message.
template <typename F>
sardine_status guarded(sardine_status on_error, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(on_error == SARDINE_ERR_INVALID_ARGUMENT ? SARDINE_ERR_OUT_OF_MEMORY : on_error,
                    "host memory allocation failed");
    } catch (const std::exception& e) {
        return fail(on_error, "%s", e.what());
    }
}

// Scratch of the no-allocation terrain-correction entry point: every device
// buffer of such a call (counters, orbit, seed table, chain records, row
// seeds, anchors, flat areas, geoid) is a slice of this array.  Each GPU the
// library runs on holds its own copy until the process ends.  Host code must
// not use the array's name as a device address: its address on a device
// comes from hipGetSymbolAddress (scratch_on_current_device).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
sardine_status scratch_on_current_device(char** out) {
    int dev = 0;
    const hipError_t ge = hipGetDevice(&dev);
    if (ge != hipSuccess) return fail(SARDINE_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(ge));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return fail(SARDINE_ERR_CAPACITY, "device %d is beyond the %d devices the scratch table holds", dev,
                    kMaxScratchDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SARDINE_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<char*>(g_scratch_ptr[dev]);
    return SARDINE_OK;
}

// Device presence + gfx950 check, once per process.
sardine_status require_device() {
    static int state = -1;  // -1 unknown, 0 ok, 1 none
    static std::string why;
    if (state < 0) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            state = 1;
            why = "no HIP device visible";
        } else {
            int dev = 0;
            hipDeviceProp_t prop;
            if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
                state = 1;
                why = "cannot query the HIP device";
            } else if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
                state = 1;
                why = std::string("device is ") + prop.gcnArchName + ", the library is built for gfx950";
            } else {
                state = 0;
                // Keep memory freed by hipFreeAsync in the device's default pool
                // (the per-call work buffers are reallocated on every call).
                hipMemPool_t pool;
                if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                    uint64_t keep = UINT64_MAX;
                    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
                }
            }
        }
    }
    if (state) return fail(SARDINE_ERR_NO_DEVICE, "%s (no CPU fallback exists)", why.c_str());
    return SARDINE_OK;
}

// Stream-ordered device scratch.
struct DevBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
    hipError_t alloc(size_t bytes, hipStream_t stream) {
        s = stream;
        return hipMallocAsync(&p, bytes ? bytes : 1, stream);
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// A piece of a gathered host→device copy: n bytes from src to offset off.
struct UploadPart {
    const void* src;
    size_t off, n;
};

// Host→device copies go through pinned staging blocks: from pageable memory
// hipMemcpyAsync copies synchronously, which would make every call wait for
// all earlier work on its stream.  A block is reused once the event recorded
// after its copy has completed.
class StagingPool {
   public:
    hipError_t copy(void* dst, const void* src, size_t n, hipStream_t s) {
        const UploadPart part{src, 0, n};
        return copy_parts(dst, &part, 1, n, s);
    }
    // One copy of `total` bytes to dst whose pieces are gathered from
    // several host arrays (the bytes between pieces are undefined).
    hipError_t copy_parts(void* dst, const UploadPart* parts, size_t np, size_t total, hipStream_t s) {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        std::lock_guard<std::mutex> g(mu_);
        Block* b = nullptr;
        size_t busy = 0;
        for (Block& x : blocks_) {
            if (x.busy && hipEventQuery(x.ev) == hipSuccess) x.busy = false;
            busy += x.busy ? 1 : 0;
            if (!x.busy && x.dev == dev && x.cap >= total && (!b || x.cap < b->cap)) b = &x;
        }
        if (!b && busy >= kMaxBusy) {  // bound the pinned footprint: wait for the oldest copy
            for (Block& x : blocks_)
                if (x.busy && x.dev == dev && x.cap >= total) {
                    if ((e = hipEventSynchronize(x.ev)) != hipSuccess) return e;
                    x.busy = false;
                    b = &x;
                    break;
                }
        }
        if (!b) {
            Block nb;
            nb.dev = dev;
            nb.cap = 64 * 1024;
            while (nb.cap < total) nb.cap *= 2;
            if ((e = hipHostMalloc(&nb.p, nb.cap, hipHostMallocDefault)) != hipSuccess) return e;
            if ((e = hipEventCreateWithFlags(&nb.ev, hipEventDisableTiming)) != hipSuccess) {
                (void)hipHostFree(nb.p);
                return e;
            }
            blocks_.push_back(nb);
            b = &blocks_.back();
        }
        for (size_t i = 0; i < np; ++i)
            if (parts[i].n) std::memcpy(static_cast<char*>(b->p) + parts[i].off, parts[i].src, parts[i].n);
        if ((e = hipMemcpyAsync(dst, b->p, total, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
        if ((e = hipEventRecord(b->ev, s)) != hipSuccess) return e;
        b->busy = true;
        return hipSuccess;
    }

   private:
    struct Block {
        void* p = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
        int dev = 0;
        bool busy = false;
    };
    static constexpr size_t kMaxBusy = 64;
    std::mutex mu_;
    std::deque<Block> blocks_;  // stable addresses; blocks live until the process ends
};

StagingPool& staging() {
    static StagingPool* pool = new StagingPool();  // never destroyed: pinned memory outlives static teardown
    return *pool;
}

template <typename T>
hipError_t upload(DevBuf& b, const T* host, size_t n, hipStream_t s) {
    hipError_t e = b.alloc(n * sizeof(T), s);
    if (e != hipSuccess) return e;
    if (n) return staging().copy(b.p, host, n * sizeof(T), s);
    return hipSuccess;
}

// A call's host inputs gathered into one host→device copy.  Every copy from
// the host is a transfer the stream waits for, which leaves the GPU idle for
// tens of microseconds each whatever its size, so the small arrays a call
// uploads (LUTs, line tables, the orbit, the seed table) travel together:
// add() places each at a 256-byte aligned offset of one device region.
class UploadPack {
   public:
    // offset of `n` bytes from src; `reserve` (>= n) keeps more room
    size_t add(const void* src, size_t n, size_t reserve = 0) {
        const size_t off = round256(size_);
        parts_.push_back(UploadPart{src, off, n});
        size_ = off + std::max(n, reserve);
        return off;
    }
    size_t size() const { return size_ ? size_ : 1; }
    // through a pinned staging block: the host arrays may go once this returns
    hipError_t copy(void* dst, hipStream_t s) const {
        return staging().copy_parts(dst, parts_.data(), parts_.size(), size(), s);
    }
    // straight from host memory (the caller keeps `host` and the arrays
    // alive until the stream has passed the copy)
    hipError_t copy_direct(void* dst, std::vector<char>& host, hipStream_t s) const {
        host.assign(size(), 0);
        for (const UploadPart& q : parts_)
            if (q.n) std::memcpy(host.data() + q.off, q.src, q.n);
        return hipMemcpyAsync(dst, host.data(), size(), hipMemcpyHostToDevice, s);
    }

   private:
    std::vector<UploadPart> parts_;
    size_t size_ = 0;
};

// ── seed LUT (terrain_correction.rs:1080-1137) ──────────────────────────
struct SeedLut {
    std::vector<double> lat;
    std::vector<int64_t> t;
};

SeedLut build_seed_lut(const sardine_grid_point* pts, int n) {
    std::vector<std::pair<double, int64_t>> v;
    v.reserve(n);
    for (int i = 0; i < n; ++i) v.emplace_back(pts[i].latitude_deg, pts[i].azimuth_time_us);
    std::stable_sort(v.begin(), v.end(), [](auto& a, auto& b) { return a.first < b.first; });
    SeedLut l;
    for (auto& e : v) {
        if (!l.lat.empty() && std::fabs(e.first - l.lat.back()) < 1e-6) continue;  // dedup_by keeps the first
        l.lat.push_back(e.first);
        l.t.push_back(e.second);
    }
    return l;
}

int64_t guess_azimuth_time(const SeedLut& l, double lat) {
    const size_t n = l.lat.size();
    if (n == 1) return l.t[0];
    size_t lo = 0, hi = n;
    while (lo < hi) {  // binary_search_by: an exact hit returns that entry
        size_t mid = lo + (hi - lo) / 2;
        if (l.lat[mid] < lat) lo = mid + 1;
        else if (l.lat[mid] > lat) hi = mid;
        else return l.t[mid];
    }
    if (lo == 0) return l.t[0];
    if (lo >= n) return l.t[n - 1];
    double frac = (lat - l.lat[lo - 1]) / (l.lat[lo] - l.lat[lo - 1]);
    double dt_us = (double)(l.t[lo] - l.t[lo - 1]);
    return l.t[lo - 1] + (int64_t)(frac * dt_us);
}

struct Grid {
    int64_t rows = 0, cols = 0;
    double y_top = 0, y_bot = 0, x_left = 0, x_right = 0;
};

// compute_output_grid_extent + n_rows/n_cols (terrain_correction.rs:1174-1216, 1304-1317)
sardine_status output_grid(const sardine_scene* sc, const sardine_tc_config* cfg, Grid* g) {
    const double sp = cfg->pixel_spacing;
    if (!(sp > 0.0)) return fail(SARDINE_ERR_TC_CONFIG, "pixel spacing must be positive");
    const sardine_bbox& bb = sc->bounding_box;
    if (cfg->crs == SARDINE_CRS_WGS84) {
        g->y_top = std::ceil(bb.max_lat_deg / sp) * sp;
        g->y_bot = std::floor(bb.min_lat_deg / sp) * sp;
        g->x_left = std::floor(bb.min_lon_deg / sp) * sp;
        g->x_right = std::ceil(bb.max_lon_deg / sp) * sp;
    } else if (cfg->crs == SARDINE_CRS_UTM_NORTH || cfg->crs == SARDINE_CRS_UTM_SOUTH) {
        if (cfg->utm_zone < 1 || cfg->utm_zone > 60)
            return fail(SARDINE_ERR_TC_PROJECTION, "UTM zone %d outside 1..60", cfg->utm_zone);
        const sd::TmSeries ts = sd::tm_series();
        const bool south = cfg->crs == SARDINE_CRS_UTM_SOUTH;
        const double c[4][2] = {{bb.min_lon_deg, bb.min_lat_deg}, {bb.min_lon_deg, bb.max_lat_deg},
                                {bb.max_lon_deg, bb.min_lat_deg}, {bb.max_lon_deg, bb.max_lat_deg}};
        double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
        for (auto& q : c) {
            double x, y;
            sd::utm_forward(ts, cfg->utm_zone, south, q[0], q[1], &x, &y);
            if (!std::isfinite(x) || !std::isfinite(y))
                return fail(SARDINE_ERR_TC_PROJECTION, "scene corner (%.4f, %.4f) does not project", q[0], q[1]);
            x0 = std::min(x0, x); x1 = std::max(x1, x);
            y0 = std::min(y0, y); y1 = std::max(y1, y);
        }
        g->y_top = std::ceil(y1 / sp) * sp;
        g->y_bot = std::floor(y0 / sp) * sp;
        g->x_left = std::floor(x0 / sp) * sp;
        g->x_right = std::ceil(x1 / sp) * sp;
    } else {
        return fail(SARDINE_ERR_TC_PROJECTION, "unsupported output CRS kind %d", cfg->crs);
    }
    const double nr = std::round((g->y_top - g->y_bot) / sp), nc = std::round((g->x_right - g->x_left) / sp);
    g->rows = nr > 0 ? (int64_t)nr : 0;
    g->cols = nc > 0 ? (int64_t)nc : 0;
    if (g->rows == 0 || g->cols == 0)
        return fail(SARDINE_ERR_TC_CONFIG, "output grid is empty: %lld rows x %lld cols", (long long)g->rows,
                    (long long)g->cols);
    return SARDINE_OK;
}

// ── terrain correction: planning shared by the per-call and batched paths ──

// apply_speckle_filter's parameter checks (speckle.rs:131-180), in order.
sardine_status validate_filter(const sardine_speckle_filter* f, int* win_out) {
    if (f->kind < SARDINE_SPECKLE_BOXCAR || f->kind > SARDINE_SPECKLE_REFINED_LEE)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "unknown speckle filter kind %d", f->kind);
    const int win = f->kind == SARDINE_SPECKLE_REFINED_LEE ? 7 : f->window;
    if (win < 3 || win % 2 == 0) return fail(SARDINE_ERR_SPECKLE_BAD_WINDOW, "window size %d must be an odd integer >= 3", win);
    if (f->kind == SARDINE_SPECKLE_LEE || f->kind == SARDINE_SPECKLE_GAMMA_MAP || f->kind == SARDINE_SPECKLE_REFINED_LEE)
        if (!(f->enl > 0.0f) || !std::isfinite(f->enl))
            return fail(SARDINE_ERR_SPECKLE_BAD_ENL, "ENL must be > 0; got %g", (double)f->enl);
    if (f->kind == SARDINE_SPECKLE_FROST && (!std::isfinite(f->damping) || f->damping < 0.0f))
        return fail(SARDINE_ERR_SPECKLE_BAD_DAMPING, "Frost damping factor must be >= 0 and finite; got %g",
                    (double)f->damping);
    if (win > 63) return fail(SARDINE_ERR_CAPACITY, "window %d above the supported 63", win);
    *win_out = win;
    return SARDINE_OK;
}

// Coefficients of each central interval's Lagrange-8 interpolants as
// polynomials in τ = ts - mid (sd::OrbitPoly), in binary128: each basis
// polynomial Π_{j≠k}(τ + mid - x_j) / Π_{j≠k}(x_k - x_j) is expanded in
// powers of τ and weighted by the node's position / velocity; the
// acceleration polynomial is the derivative of the velocity one.  The node
// times x_j are the double seconds the reference interpolates with.
void orbit_fill_poly(sd::OrbitDev& o) {
    using q128 = __float128;
    for (int c = 0; c + 1 < o.n; ++c) {
        sd::OrbitPoly& P = o.poly[c];
        const int s = sd::orbit_window_start(c, o.n);
        if (s < 0 || s + sd::kLagrange > o.n) continue;  // fewer than 8 vectors: never evaluated
        P.mid = (double)(((q128)o.tsec[c] + (q128)o.tsec[c + 1]) / 2);
        q128 cp[3][8] = {}, cv[3][8] = {};
        for (int k = 0; k < sd::kLagrange; ++k) {
            q128 poly[8] = {1};
            int deg = 0;
            q128 den = 1;
            for (int j = 0; j < sd::kLagrange; ++j) {
                if (j == k) continue;
                const q128 a = (q128)P.mid - (q128)o.tsec[s + j];
                for (int i = deg + 1; i >= 1; --i) poly[i] = poly[i] * a + poly[i - 1];
                poly[0] = poly[0] * a;
                ++deg;
                den *= (q128)o.tsec[s + k] - (q128)o.tsec[s + j];
            }
            for (int d = 0; d < 3; ++d) {
                const q128 yp = (q128)o.pos[s + k][d] / den, yv = (q128)o.vel[s + k][d] / den;
                for (int i = 0; i < 8; ++i) {
                    cp[d][i] += poly[i] * yp;
                    cv[d][i] += poly[i] * yv;
                }
            }
        }
        for (int d = 0; d < 3; ++d) {
            for (int i = 0; i < 8; ++i) {
                P.cp[d][i] = (double)cp[d][i];
                P.cv[d][i] = (double)cv[d][i];
            }
            for (int i = 1; i < 8; ++i) P.ca[d][i - 1] = (double)(cv[d][i] * i);
        }
    }
}

// Orbit in device form, times relative to the first state vector.
sd::OrbitDev make_orbit_dev(const sardine_orbit& o) {
    sd::OrbitDev orb;
    std::memset(&orb, 0, sizeof orb);
    orb.n = o.count;
    const int64_t epoch = orb.n > 0 ? o.state_vectors[0].time_us : 0;
    for (int i = 0; i < orb.n; ++i) {
        const sardine_state_vector& s = o.state_vectors[i];
        orb.t_us[i] = s.time_us - epoch;
        orb.tsec[i] = (double)(s.time_us - epoch) / 1e6;
        for (int c = 0; c < 3; ++c) {
            orb.pos[i][c] = s.position_m[c];
            orb.vel[i][c] = s.velocity_m_s[c];
        }
    }
    orbit_fill_poly(orb);
    if (orb.n >= 2 && orb.t_us[orb.n - 1] > 0) orb.guess_scale = (double)(orb.n - 1) / (double)orb.t_us[orb.n - 1];
    return orb;
}

// The binary128 interval polynomials cost ~0.2 ms of host time per orbit;
// calls on the same state vectors (compared bit for bit) reuse the last few
// results.
class OrbitDevCache {
   public:
    std::shared_ptr<const sd::OrbitDev> get(const sardine_orbit& o) {
        const size_t bytes = (size_t)std::max(o.count, 0) * sizeof(sardine_state_vector);
        std::lock_guard<std::mutex> g(mu_);
        ++tick_;
        for (Entry& e : entries_)
            if (e.key.size() == bytes && (bytes == 0 || std::memcmp(e.key.data(), o.state_vectors, bytes) == 0)) {
                e.used = tick_;
                return e.orb;
            }
        auto orb = std::make_shared<const sd::OrbitDev>(make_orbit_dev(o));
        Entry ne;
        ne.key.assign(reinterpret_cast<const char*>(o.state_vectors), reinterpret_cast<const char*>(o.state_vectors) + bytes);
        ne.orb = orb;
        ne.used = tick_;
        if (entries_.size() < kEntries) {
            entries_.push_back(std::move(ne));
        } else {
            auto lru = std::min_element(entries_.begin(), entries_.end(),
                                        [](const Entry& a, const Entry& b) { return a.used < b.used; });
            *lru = std::move(ne);
        }
        return orb;
    }

   private:
    struct Entry {
        std::vector<char> key;
        std::shared_ptr<const sd::OrbitDev> orb;
        uint64_t used = 0;
    };
    static constexpr size_t kEntries = 8;
    std::mutex mu_;
    std::vector<Entry> entries_;
    uint64_t tick_ = 0;
};

OrbitDevCache& orbit_cache() {
    static OrbitDevCache* c = new OrbitDevCache();
    return *c;
}

// Uploads an orbit: the whole struct is allocated, the part the kernels
// read (up to its last interval polynomial) is copied.
hipError_t upload_orbit(DevBuf& b, const sd::OrbitDev& o, hipStream_t s) {
    hipError_t e = b.alloc(sizeof(sd::OrbitDev), s);
    if (e != hipSuccess) return e;
    return staging().copy(b.p, &o, sd::orbit_lds_bytes(o.n), s);
}

// Device copies of a scene's orbit and seed table.
// The tiling of a mosaic raster: grid within bounds, covering the raster,
// ranks in range and only in cells of the grid.
sardine_status validate_dem_tiling(const sardine_dem_raster* dem) {
    const sardine_dem_tiling* tl = dem->tiling;
    if (!tl) return SARDINE_OK;
    const int M = SARDINE_DEM_MAX_TILES_AXIS;
    if (tl->rows < 1 || tl->rows > M || tl->cols < 1 || tl->cols > M || tl->tile_height < 2 || tl->tile_width < 2 ||
        (int64_t)tl->rows * tl->tile_height != dem->height || (int64_t)tl->cols * tl->tile_width != dem->width)
        return fail(SARDINE_ERR_INVALID_ARGUMENT,
                    "DEM tiling: %d x %d tiles of %d x %d pixels (at most %d x %d) must cover the %lld x %lld raster",
                    tl->rows, tl->cols, tl->tile_height, tl->tile_width, M, M, (long long)dem->height,
                    (long long)dem->width);
    for (int i = 0; i < M * M; ++i)
        if (tl->rank[i] >= M * M || (tl->rank[i] >= 0 && (i / M >= tl->rows || i % M >= tl->cols)))
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "DEM tiling: bad rank at cell %d", i);
    // The lookup finds a point's tile from tile 0's origin, one step either
    // way, and looks for a second holding tile only within half a pixel of
    // an edge: adjacent tile rows (columns) must sit one tile apart, north
    // to south (west to east), to a quarter pixel.
    const double px = dem->pixel_deg;
    for (int i = 0; i < tl->rows; ++i) {
        const double d = i ? (tl->origin_lat_deg[i - 1] - tl->origin_lat_deg[i]) / px - tl->tile_height : 0.0;
        if (!std::isfinite(tl->origin_lat_deg[i]) || !(std::fabs(d) <= 0.25))
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "DEM tiling: tile row %d is not one tile south of row %d", i,
                        i - 1);
    }
    for (int i = 0; i < tl->cols; ++i) {
        const double d = i ? (tl->origin_lon_deg[i] - tl->origin_lon_deg[i - 1]) / px - tl->tile_width : 0.0;
        if (!std::isfinite(tl->origin_lon_deg[i]) || !(std::fabs(d) <= 0.25))
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "DEM tiling: tile column %d is not one tile east of column %d",
                        i, i - 1);
    }
    return SARDINE_OK;
}

// The DEM fields of the kernel arguments (raster validated: 2 <= width,
// height < 2^31, tiling checked); tab: the device copy of the tiling table.
void fill_dem_params(sd::TcParams& p, const sardine_dem_raster* dem, const double* tab) {
    p.dem = dem->data;
    p.dem_w = dem->width;
    p.dem_h = dem->height;
    p.dem_olat = dem->origin_lat_deg;
    p.dem_olon = dem->origin_lon_deg;
    p.dem_pix = dem->pixel_deg;
    p.dem_inv_pix = 1.0 / dem->pixel_deg;
    p.dem_max_row = (double)(dem->height - 1);
    p.dem_max_col = (double)(dem->width - 1);
    p.dem_hi_row = p.dem_max_row + 0.5;
    p.dem_hi_col = p.dem_max_col + 0.5;
    p.dem_lim_row = (int)(dem->height - 2);
    p.dem_lim_col = (int)(dem->width - 2);
    if (const sardine_dem_tiling* tl = dem->tiling) {
        p.dem_tiled = 1;
        p.dem_trows = tl->rows;
        p.dem_tcols = tl->cols;
        p.dem_th = tl->tile_height;
        p.dem_tw = tl->tile_width;
        p.dem_tab = tab;
    }
}

// The device table of a DEM tiling (TcParams::dem_tab's layout).
void dem_tile_table(const sardine_dem_tiling* tl, double tab[sd::kDemTabWords]) {
    const int M = SARDINE_DEM_MAX_TILES_AXIS;
    for (int i = 0; i < sd::kDemTabWords; ++i) tab[i] = 0.0;
    for (int i = 0; i < M; ++i) {
        tab[i] = tl->origin_lat_deg[i];
        tab[M + i] = tl->origin_lon_deg[i];
    }
    for (int r = 0; r < M; ++r) {
        uint64_t w = 0;
        for (int c = 0; c < M; ++c) {
            const int32_t k = tl->rank[r * M + c];
            w |= (uint64_t)(k < 0 ? 0xff : k) << (8 * c);
        }
        std::memcpy(&tab[2 * M + r], &w, sizeof w);
    }
}

struct SceneDev {
    const sd::OrbitDev* orb = nullptr;
    const double* dem_tab = nullptr;  // GLO-30 mosaic table (sd::kDemTabWords), mosaics only
    const double* lut_lat = nullptr;
    const int64_t* lut_t = nullptr;
    int n_lut = 0;
    int64_t epoch = 0;
};

// One terrain-correction job: validated, with its grid and kernel arguments
// filled except for the per-call work buffers.
struct TcPlan {
    Grid g;
    int64_t r0 = 0, c0 = 0, nr = 0, nc = 0;
    bool utm = false, need_geom = false, compute_lia = false;
    int n_anch = 0;
    sd::TcParams p;
    sd::RowPrepParams q;
};

// Validation in the reference's order (terrain_correction.rs:1262-1317).
sardine_status validate_tc(const sardine_radar_image* merged, const sardine_scene* scene,
                           const sardine_dem_raster* dem, const sardine_tc_config* cfg, Grid* g) {
    if (scene->n_grid_points <= 0 || !scene->grid_points)
        return fail(SARDINE_ERR_TC_NO_GRID_POINTS, "no geolocation grid points provided");
    if (!(cfg->pixel_spacing > 0.0)) return fail(SARDINE_ERR_TC_CONFIG, "pixel spacing must be positive");
    if (scene->n_subswaths <= 0) return fail(SARDINE_ERR_TC_NO_SUBSWATHS, "scene has no subswaths");
    const double ati = scene->azimuth_time_interval_s;
    if (!std::isfinite(ati) || ati <= 0.0)
        return fail(SARDINE_ERR_TC_CONFIG, "invalid azimuth_time_interval_s in subswath metadata: %g", ati);
    sardine_status st = output_grid(scene, cfg, g);
    if (st) return st;
    const int anchors = cfg->projector_anchors == 0 ? 256 : cfg->projector_anchors;
    if (anchors < 1) return fail(SARDINE_ERR_TC_CONFIG, "SARDINE_TC_PROJECTOR_ANCHORS must be >= 1 (default 256)");
    if (anchors > sd::kMaxAnchorPts - 1)
        return fail(SARDINE_ERR_TC_CONFIG, "projector anchors %d above the supported %d", anchors,
                    sd::kMaxAnchorPts - 1);
    if (cfg->max_newton_iterations < 0 || cfg->max_newton_iterations > 255)
        return fail(SARDINE_ERR_TC_CONFIG, "max_newton_iterations %d outside 0..255", cfg->max_newton_iterations);
    if (cfg->resampling < 0 || cfg->resampling > 2) return fail(SARDINE_ERR_TC_CONFIG, "bad resampling kernel");
    if (cfg->precision != SARDINE_PRECISION_FP64 && cfg->precision != SARDINE_PRECISION_FP32_CENTRED &&
        cfg->precision != SARDINE_PRECISION_MIXED)
        return fail(SARDINE_ERR_TC_CONFIG, "bad precision selector");
    if (scene->orbit.count > sd::kMaxSv || scene->orbit.count < 0)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "orbit has %d state vectors (max %d)", scene->orbit.count, sd::kMaxSv);
    for (int i = 1; i < scene->orbit.count; ++i)
        if (!(scene->orbit.state_vectors[i].time_us > scene->orbit.state_vectors[i - 1].time_us))
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "state vector times must be strictly increasing (vector %d)", i);
    if (!merged->data || merged->lines <= 0 || merged->samples <= 0)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "radar image has no raster");
    if (merged->lines >= INT32_MAX || merged->samples >= INT32_MAX)
        return fail(SARDINE_ERR_CAPACITY, "radar image dimensions must be below 2^31");
    if (!dem->data || dem->width < 2 || dem->height < 2 || !(dem->pixel_deg > 0.0)) {
        // TerrainCorrectionError::Dem(DemError::…): no raster is NoTiles, a
        // raster below 2 x 2 samples BadSize, a bad spacing TiffDecode
        const sardine_status st = fail(SARDINE_ERR_TC_DEM, "DEM error: %s",
                                       !dem->data ? "no tiles loaded"
                                       : (dem->width < 2 || dem->height < 2)
                                           ? "DEM raster must be at least 2x2 samples"
                                           : "DEM raster spacing must be positive");
        g_last_detail.dem_status = !dem->data ? SARDINE_ERR_DEM_NO_TILES
                                   : (dem->width < 2 || dem->height < 2) ? SARDINE_ERR_DEM_BAD_SIZE
                                                                         : SARDINE_ERR_DEM_TIFF_DECODE;
        if (g_last_detail.dem_status == SARDINE_ERR_DEM_BAD_SIZE) {
            g_last_detail.dem_expected_bytes = 2 * 2 * sizeof(float);
            g_last_detail.dem_actual_bytes =
                (uint64_t)std::max<int64_t>(dem->width, 0) * (uint64_t)std::max<int64_t>(dem->height, 0) * sizeof(float);
        }
        return st;
    }
    if (dem->width >= INT32_MAX || dem->height >= INT32_MAX)
        return fail(SARDINE_ERR_CAPACITY, "DEM dimensions must be below 2^31");
    if (sardine_status st = validate_dem_tiling(dem)) return st;
    if (cfg->geoid.kind == SARDINE_GEOID_GRID_2P5 && !cfg->geoid.grid_73x144)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "geoid grid pointer is null");
    return SARDINE_OK;
}

// The project_row anchor segment a grid column blends in: the same
// arithmetic as the kernels' walk (output_crs.rs:395-414).
int anchor_segment(const sd::TcParams& p, int64_t col) {
    const double x = p.x_start + (double)col * p.spacing;
    const int na = p.n_anchor_pts - 1;
    int seg = (int)std::floor((x - p.x_start) * p.inv_anchor_dx);
    seg = seg < 0 ? 0 : (seg > na - 1 ? na - 1 : seg);
    while (seg > 0 && !(x > p.x_start + (double)seg * p.anchor_dx)) --seg;
    while (seg + 1 < na && x > p.x_start + (double)(seg + 1) * p.anchor_dx) ++seg;
    return seg;
}

// Everything of one job that needs no device work: grid, AOI, kernel
// arguments (terrain_correction.rs:1304-1511).  Work buffers (records,
// anchors, row seeds, flat areas, counters, geoid) are set by the caller.
sardine_status plan_tc(const sardine_radar_image* merged, const sardine_scene* scene, const sardine_dem_raster* dem,
                       const sardine_tc_config* cfg, const sardine_geocoded_image* out, const SceneDev& sdv,
                       TcPlan* pl) {
    sardine_status st = validate_tc(merged, scene, dem, cfg, &pl->g);
    if (st) return st;
    const Grid& g = pl->g;
    int64_t r0 = cfg->aoi_row0, c0 = cfg->aoi_col0, nr = cfg->aoi_rows, nc = cfg->aoi_cols;
    if (nr == 0) { r0 = 0; c0 = 0; nr = g.rows; nc = g.cols; }
    if (r0 < 0 || c0 < 0 || nr <= 0 || nc <= 0 || r0 + nr > g.rows || c0 + nc > g.cols)
        return fail(SARDINE_ERR_TC_CONFIG, "AOI window outside the %lld x %lld output grid", (long long)g.rows,
                    (long long)g.cols);
    if (nr > 0x7fffffffLL) return fail(SARDINE_ERR_CAPACITY, "AOI has too many rows");
    const bool compute_lia = cfg->compute_lia != 0;
    if (!out->data || (compute_lia && (!out->cos_lia || !out->mask)))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "output buffers missing");
    pl->r0 = r0; pl->c0 = c0; pl->nr = nr; pl->nc = nc;
    const int anchors = cfg->projector_anchors == 0 ? 256 : cfg->projector_anchors;
    const double sp = cfg->pixel_spacing;
    const bool utm = cfg->crs != SARDINE_CRS_WGS84;
    const bool south = cfg->crs == SARDINE_CRS_UTM_SOUTH;
    const bool need_geom = cfg->flatten || cfg->compute_lia;
    const double step_deg = utm ? sp / 111320.0 : sp;
    const int n_anch = (int)std::min<int64_t>(anchors, g.cols);
    const double x_start = g.x_left + 0.5 * sp;
    const double x_end = x_start + ((double)g.cols - 1.0) * sp;
    const double anchor_dx = (x_end - x_start) / (double)n_anch;
    pl->utm = utm; pl->need_geom = need_geom; pl->compute_lia = compute_lia; pl->n_anch = n_anch;

    sd::RowPrepParams& q = pl->q;
    std::memset(&q, 0, sizeof q);
    q.lut_lat = sdv.lut_lat;
    q.lut_t = sdv.lut_t;
    q.n_lut = sdv.n_lut;
    q.epoch = sdv.epoch;
    q.y_top = g.y_top;
    q.x_left = g.x_left;
    q.spacing = sp;
    q.x_start = x_start;
    q.anchor_dx = anchor_dx;
    q.step_deg = step_deg;
    q.full_cols = g.cols;
    q.row0 = r0;
    q.rows = nr;
    q.n_anch = n_anch;
    q.utm = utm ? 1 : 0;
    q.south = south ? 1 : 0;
    q.zone = cfg->utm_zone;
    q.need_geom = need_geom ? 1 : 0;

    const double ati = scene->azimuth_time_interval_s;
    sd::TcParams& p = pl->p;
    std::memset(&p, 0, sizeof p);
    p.orbit = sdv.orb;
    p.orbit_lds = (int32_t)sd::orbit_lds_bytes(scene->orbit.count);
    p.img = merged->data;
    p.nesz = (std::isfinite(cfg->noise_floor_margin_db) && merged->nesz) ? merged->nesz : nullptr;
    p.lines = merged->lines;
    p.samples = merged->samples;
    p.az_start_us = merged->azimuth_start_time_us - sdv.epoch;
    p.near_range_m = merged->near_slant_range_time_s * sd::kC / 2.0;
    p.rps = merged->range_pixel_spacing_m;
    p.ati = ati;
    p.line_per_us = 1.0 / (1e6 * ati);
    p.inv_ati = 1.0 / ati;
    p.inv_rps = 1.0 / p.rps;
    fill_dem_params(p, dem, sdv.dem_tab);
    p.crs = utm ? 1 : 0;
    p.utm_zone = cfg->utm_zone;
    p.utm_south = south ? 1 : 0;
    p.n_anchor_pts = n_anch + 1;
    p.spacing = sp;
    p.lat_north = g.y_top;
    p.lon_west = g.x_left;
    p.x_start = x_start;
    p.anchor_dx = anchor_dx;
    p.inv_anchor_dx = anchor_dx > 0.0 ? 1.0 / anchor_dx : 0.0;
    p.full_cols = g.cols;
    p.row0 = r0;
    p.col0 = c0;
    p.rows = nr;
    p.cols = nc;
    p.max_iter = cfg->max_newton_iterations;
    p.tol = cfg->newton_tolerance_s;
    p.tie_band = g_tc_tie_band.load();
    p.flatten = cfg->flatten ? 1 : 0;
    p.compute_lia = compute_lia ? 1 : 0;
    p.need_geom = need_geom ? 1 : 0;
    p.resampling = cfg->resampling;
    p.mixed = cfg->precision == SARDINE_PRECISION_MIXED ? 1 : 0;
    p.step_deg = step_deg;
    p.step_sin = std::sin(step_deg * sd::kPi / 180.0);
    p.step_cos = std::cos(step_deg * sd::kPi / 180.0);
    // 10f32.powf(margin / 10.0) (terrain_correction.rs:1727), evaluated once
    p.nesz_factor = std::pow(10.0f, cfg->noise_floor_margin_db / 10.0f);
    p.out = out->data;
    p.cos_lia = compute_lia ? out->cos_lia : nullptr;
    p.mask = compute_lia ? out->mask : nullptr;
    p.radar_coords = out->radar_coords;
    p.lead = g_tc_lead.load();
    if (utm) {  // the anchors the AOI's columns blend between
        p.anchor0 = anchor_segment(p, c0);
        p.anchor_stride = anchor_segment(p, c0 + nc - 1) + 2 - p.anchor0;
    }
    return SARDINE_OK;
}

// The FP32 variant's orbit: node times in seconds from t_ref, positions in
// metres from `centre`, barycentric weights of the rounded node times.
sd::OrbitF32 make_orbit_f32(const sd::OrbitDev& o, const double centre[3], int64_t t_ref_us) {
    sd::OrbitF32 f;
    std::memset(&f, 0, sizeof f);
    f.n = o.n;
    for (int i = 0; i < o.n; ++i) {
        f.t[i] = (float)((double)(o.t_us[i] - t_ref_us) * 1e-6);
        for (int c = 0; c < 3; ++c) {
            f.pos[i][c] = (float)(o.pos[i][c] - centre[c]);
            f.vel[i][c] = (float)o.vel[i][c];
        }
    }
    for (int s = 0; s + sd::kLagrange <= o.n; ++s)
        for (int k = 0; k < sd::kLagrange; ++k) {
            double prod = 1.0;
            for (int j = 0; j < sd::kLagrange; ++j)
                if (j != k) prod *= (double)f.t[s + k] - (double)f.t[s + j];
            f.bary[s][k] = (float)(1.0 / prod);
        }
    return f;
}

// Segment length of the chain decomposition: `requested` > 0 forces it,
// otherwise aim for ~2 waves of 64 lanes per SIMD over the whole grid.
void set_decomposition(sd::TcParams& p, int64_t requested) {
    int64_t L = requested;
    if (L <= 0) {
        const int64_t target_lanes = (int64_t)256 * 4 * 2 * 64;
        // the 2-column lead of every segment is overhead: as long a segment
        // as still fills the GPU (measured on the 473.7 Mpx UTM grid:
        // L = 32 25.5 ms, 64 25.1 ms, 128 25.1 ms)
        L = std::min<int64_t>(64, std::max<int64_t>(16, (p.rows * p.cols + target_lanes - 1) / target_lanes));
    }
    // at most 255: the chain pass counts a segment's pixels in bytes
    L = std::min<int64_t>(std::min<int64_t>(L, 255), p.cols);
    p.seg_len = (int)L;
    p.nseg = (int)((p.cols + L - 1) / L);
    p.rows_per_wg = 64;
}

// Bytes of the per-job work buffers, each rounded to 256.
struct TcWork {
    size_t records = 0, flags = 0, ties = 0, anchors = 0, blend = 0, seed = 0, flat = 0, geoid = 0;
    int64_t tie_cap = 0;
    size_t total() const { return records + flags + ties + anchors + blend + seed + flat + geoid; }
};
TcWork tc_work_bytes(const TcPlan& pl, const sardine_tc_config* cfg) {
    TcWork w;
    // the FP32 variant solves every pixel on its own: no chain records
    w.records = cfg->precision == SARDINE_PRECISION_FP32_CENTRED
                    ? 0
                    : round256((size_t)pl.nr * pl.p.nseg * sizeof(sd::SegRecord));
    w.flags = w.records ? round256((size_t)pl.nr * sizeof(int32_t)) : 0;
    // tie-only segments replayed in parallel: room for 8 per row (more send
    // their rows to the sequential walk)
    w.tie_cap = w.records ? std::min<int64_t>(pl.nr * (int64_t)pl.p.nseg, std::max<int64_t>(64, pl.nr * 8)) : 0;
    w.ties = round256((size_t)w.tie_cap * sizeof(int2));
    w.anchors = pl.utm ? round256((size_t)pl.nr * 2 * pl.p.anchor_stride * sizeof(double)) : 0;
    w.blend = round256((size_t)pl.nc * (pl.utm ? sizeof(sd::ColBlend) : sizeof(sd::ColGeo)));
    w.seed = round256((size_t)pl.nr * sizeof(int64_t));
    w.flat = pl.need_geom ? round256((size_t)pl.nr * sizeof(double)) : 0;
    w.geoid = cfg->geoid.kind == SARDINE_GEOID_GRID_2P5 ? round256(73 * 144 * sizeof(float)) : 0;
    return w;
}
// Points the job's kernel arguments at its slice of a work arena.
void assign_work(TcPlan& pl, const TcWork& w, char* base) {
    pl.p.records = reinterpret_cast<sd::SegRecord*>(base);
    base += w.records;
    pl.p.row_flags = w.flags ? reinterpret_cast<int32_t*>(base) : nullptr;
    base += w.flags;
    pl.p.tie_list = w.tie_cap ? reinterpret_cast<int2*>(base) : nullptr;
    pl.p.tie_cap = (int)w.tie_cap;
    base += w.ties;
    pl.p.anchors = w.anchors ? reinterpret_cast<double*>(base) : nullptr;
    base += w.anchors;
    pl.p.col_blend = pl.utm ? reinterpret_cast<sd::ColBlend*>(base) : nullptr;
    pl.p.col_geo = pl.utm ? nullptr : reinterpret_cast<sd::ColGeo*>(base);
    base += w.blend;
    pl.p.row_seed = reinterpret_cast<const int64_t*>(base);
    pl.q.row_seed = reinterpret_cast<int64_t*>(base);
    base += w.seed;
    pl.p.row_flat_area = w.flat ? reinterpret_cast<const double*>(base) : nullptr;
    pl.q.row_flat_area = w.flat ? reinterpret_cast<double*>(base) : nullptr;
    base += w.flat;
    pl.p.geoid = w.geoid ? reinterpret_cast<const float*>(base) : nullptr;
}

// Host result of one job from its counters (GeocodedImage fields).
sardine_status finish_tc(const TcPlan& pl, const unsigned long long* cnt, const sardine_orbit& orbit,
                         sardine_geocoded_image* out) {
    const int err = (int)(cnt[sd::kNumCounters] & 0xffffffffull);
    const int n_sv = orbit.count;
    if (err & 1) {
        // OutsideOrbitWindow { requested, first, last } (orbit.rs:513-517);
        // requested is the iterate furthest outside the window
        const int64_t first = orbit.state_vectors[0].time_us, last = orbit.state_vectors[n_sv - 1].time_us;
        const int64_t req = cnt[sd::kCntOrbitBelow] ? first - (int64_t)cnt[sd::kCntOrbitBelow]
                                                    : last + (int64_t)cnt[sd::kCntOrbitAbove];
        const sardine_status st =
            fail(SARDINE_ERR_TC_ORBIT_OUTSIDE_WINDOW,
                 "orbit interpolation failed: interpolation time %lld outside orbit window [%lld, %lld]; refusing "
                 "to extrapolate", (long long)req, (long long)first, (long long)last);
        g_last_detail.requested_us = req;
        g_last_detail.first_us = first;
        g_last_detail.last_us = last;
        return st;
    }
    if (err & 2) {
        const sardine_status st =
            fail(SARDINE_ERR_TC_ORBIT_INSUFFICIENT,
                 "orbit interpolation failed: EOF orbit does not cover scene: need at least 8 vectors for "
                 "interpolation, have %d", n_sv);
        g_last_detail.have_vectors = n_sv;
        g_last_detail.need_vectors = sd::kLagrange;
        return st;
    }
    const Grid& g = pl.g;
    const double sp = pl.p.spacing;
    out->rows = pl.nr;
    out->cols = pl.nc;
    out->full_rows = g.rows;
    out->full_cols = g.cols;
    out->geotransform[0] = g.x_left + (double)pl.c0 * sp;
    out->geotransform[1] = sp;
    out->geotransform[2] = 0.0;
    out->geotransform[3] = g.y_top - (double)pl.r0 * sp;
    out->geotransform[4] = 0.0;
    out->geotransform[5] = -sp;
    out->total_pixel_count = (uint64_t)(pl.nr * pl.nc);
    out->valid_pixel_count = cnt[0];
    out->dem_missing_count = cnt[1];
    out->outside_footprint_count = cnt[2];
    out->non_converged_count = cnt[3];
    out->degenerate_geometry_count = cnt[4];
    out->flat_masked_count = cnt[5];
    out->noise_masked_count = cnt[6];
    for (int i = 0; i < SARDINE_NEWTON_HIST_SIZE; ++i) out->newton_iter_histogram[i] = cnt[7 + i];
    out->chain_redo_rounds = cnt[sd::kCntRedo];
    out->tie_steps = cnt[sd::kCntTie + sd::TIE_STEPS];
    out->tie_decisions_changed = cnt[sd::kCntTie + sd::TIE_FLIPS];
    std::memcpy(&out->tie_max_dev_us, &cnt[sd::kCntTie + sd::TIE_MAX_DEV], sizeof(double));
    return SARDINE_OK;
}

}  // namespace

// ═════════════════════════════════════════════════════════════════════════
extern "C" {

const char* sardine_last_error(void) { return g_last_error.c_str(); }
sardine_status sardine_last_error_detail(sardine_error_detail* out) {
    if (!out) return SARDINE_ERR_INVALID_ARGUMENT;
    *out = g_last_detail;
    return SARDINE_OK;
}
void sardine_set_profiling(int on) {
    g_profile = on != 0;
    if (!g_profile) return;
    // create this thread's timer events now, outside whatever is timed next
    for (TimerSlot& t : g_slots)
        for (TimerEntry& e : t.e) {
            if (!e.a) (void)hipEventCreate(&e.a);
            if (!e.b) (void)hipEventCreate(&e.b);
        }
}
void sardine_set_tc_segment(int64_t columns) { g_tc_seg = columns < 0 ? 0 : (columns > 255 ? 255 : columns); }
void sardine_set_tc_lead(int32_t columns) { g_tc_lead = columns < 0 ? 0 : (columns > 64 ? 64 : columns); }
void sardine_set_tc_tie_band(double us) { g_tc_tie_band = us > 0.0 && us <= 0.5 ? us : sd::kTieBandUs; }
double sardine_last_kernel_ms(int family) {
    if (family < 0 || family >= kFamilies) return -1.0;
    TimerSlot& t = g_slots[family];
    if (t.next == 0) return -1.0;
    return timer_read(t.e[(t.next - 1) % kTimerRing]);
}
int32_t sardine_recent_kernel_ms(int family, double* ms, int32_t max) {
    if (family < 0 || family >= kFamilies || !ms || max <= 0) return 0;
    TimerSlot& t = g_slots[family];
    const int64_t n = std::min<int64_t>({(int64_t)max, t.next, (int64_t)kTimerRing});
    for (int64_t i = 0; i < n; ++i) ms[i] = timer_read(t.e[(t.next - n + i) % kTimerRing]);
    return (int32_t)n;
}
int sardine_abi_version(void) { return SARDINE_ABI_VERSION; }

int sardine_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void sardine_tc_config_default(sardine_tc_config* c) {
    // TerrainCorrectionConfig::new (terrain_correction.rs:223-235), GeoidModel::Zero
    std::memset(c, 0, sizeof *c);
    c->pixel_spacing = 0.0001;
    c->crs = SARDINE_CRS_WGS84;
    c->max_newton_iterations = 10;
    c->newton_tolerance_s = 1e-6;
    c->noise_floor_margin_db = -INFINITY;
    c->resampling = SARDINE_RESAMPLE_BILINEAR;
    c->projector_anchors = 256;
    c->precision = SARDINE_PRECISION_FP64;
    c->geoid.kind = SARDINE_GEOID_ZERO;
}

sardine_status sardine_tc_output_grid(const sardine_scene* scene, const sardine_tc_config* cfg, int64_t* rows,
                                      int64_t* cols, double gt[6]) {
    if (!scene || !cfg || !rows || !cols) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    Grid g;
    sardine_status st = output_grid(scene, cfg, &g);
    if (st) return st;
    *rows = g.rows;
    *cols = g.cols;
    if (gt) {  // terrain_correction.rs:1805-1812
        gt[0] = g.x_left; gt[1] = cfg->pixel_spacing; gt[2] = 0.0;
        gt[3] = g.y_top; gt[4] = 0.0; gt[5] = -cfg->pixel_spacing;
    }
    return SARDINE_OK;
}

sardine_status sardine_orbit_interpolate(const sardine_orbit* orbit, int64_t time_us, double pos[3], double vel[3],
                                         double acc[3]) {
    if (!orbit || !pos || !vel || !acc || (orbit->count > 0 && !orbit->state_vectors))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    if (orbit->count > sd::kMaxSv || orbit->count < 0)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "orbit has %d state vectors (max %d)", orbit->count, sd::kMaxSv);
    for (int i = 1; i < orbit->count; ++i)
        if (!(orbit->state_vectors[i].time_us > orbit->state_vectors[i - 1].time_us))
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "state vector times must be strictly increasing (vector %d)", i);
    const std::shared_ptr<const sd::OrbitDev> o = orbit_cache().get(*orbit);
    const int64_t epoch = orbit->count > 0 ? orbit->state_vectors[0].time_us : 0;
    sd::d3 P, V, A;
    const int e = sd::orbit_interp(*o, time_us - epoch, &P, &V, &A);
    if (e == sd::ORB_INSUFFICIENT) {
        const sardine_status st = fail(SARDINE_ERR_TC_ORBIT_INSUFFICIENT,
                                       "EOF orbit does not cover scene: need at least 8 vectors for interpolation, "
                                       "have %d", orbit->count);
        g_last_detail.have_vectors = orbit->count;
        g_last_detail.need_vectors = sd::kLagrange;
        return st;
    }
    if (e == sd::ORB_OUTSIDE) {
        const int64_t first = orbit->state_vectors[0].time_us, last = orbit->state_vectors[orbit->count - 1].time_us;
        const sardine_status st =
            fail(SARDINE_ERR_TC_ORBIT_OUTSIDE_WINDOW,
                 "interpolation time %lld outside orbit window [%lld, %lld]; refusing to extrapolate",
                 (long long)time_us, (long long)first, (long long)last);
        g_last_detail.requested_us = time_us;
        g_last_detail.first_us = first;
        g_last_detail.last_us = last;
        return st;
    }
    pos[0] = P.x; pos[1] = P.y; pos[2] = P.z;
    vel[0] = V.x; vel[1] = V.y; vel[2] = V.z;
    acc[0] = A.x; acc[1] = A.y; acc[2] = A.z;
    return SARDINE_OK;
}

// A pinned block per host thread for a call's counters (allocated on the
// thread's first call, kept until the process ends).
hipError_t pinned_counters(unsigned long long** out) {
    thread_local unsigned long long* block = nullptr;
    if (!block) {
        void* p = nullptr;
        const hipError_t e = hipHostMalloc(&p, sd::kCntSlots * sizeof(unsigned long long), hipHostMallocDefault);
        if (e != hipSuccess) return e;
        block = static_cast<unsigned long long*>(p);
    }
    *out = block;
    return hipSuccess;
}

// Device memory of one terrain-correction call: stream-ordered allocations,
// or consecutive 256-byte aligned slices of a fixed region (the
// no-allocation entry point's scratch array).
class TcMemory {
   public:
    TcMemory(hipStream_t s, char* fixed, size_t cap) : s_(s), fixed_(fixed), cap_(cap) {}
    bool fixed() const { return fixed_ != nullptr; }
    size_t used() const { return used_; }
    // n bytes, or nullptr with *e set (hipErrorOutOfMemory when the fixed
    // region is full)
    void* get(size_t n, hipError_t* e) {
        *e = hipSuccess;
        if (fixed_) {
            const size_t at = round256(used_);
            if (at + n > cap_) {
                *e = hipErrorOutOfMemory;
                used_ = at + n;  // reported in the capacity error
                return nullptr;
            }
            used_ = at + n;
            return fixed_ + at;
        }
        bufs_.emplace_back(new DevBuf());
        if ((*e = bufs_.back()->alloc(n, s_)) != hipSuccess) return nullptr;
        return bufs_.back()->p;
    }
    // host → device: pinned staging for allocated buffers; for the fixed
    // region a direct copy, which needs no pinned block (the call waits for
    // its stream before returning, so the host source outlives the copy)
    hipError_t copy_in(void* dst, const void* src, size_t n) {
        if (!n) return hipSuccess;
        if (fixed_) return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s_);
        return staging().copy(dst, src, n, s_);
    }
    // free the allocated buffers (stream-ordered, after the work enqueued
    // so far); the fixed region needs nothing
    void release() { bufs_.clear(); }
    // a gathered upload (size() bytes at dst), the same way
    hipError_t copy_in(void* dst, const UploadPack& up) {
        if (fixed_) return up.copy_direct(dst, host_, s_);
        return up.copy(dst, s_);
    }

   private:
    hipStream_t s_;
    char* fixed_;
    size_t cap_;
    size_t used_ = 0;
    std::vector<std::unique_ptr<DevBuf>> bufs_;
    std::vector<char> host_;  // a direct gathered upload's source
};

// terrain_correction over `mem`: planning, uploads, the kernels, counters.
sardine_status tc_call(const sardine_radar_image* merged, const sardine_scene* scene, const sardine_dem_raster* dem,
                       const sardine_tc_config* cfg, sardine_geocoded_image* out, hipStream_t stream, TcMemory& mem) {
    if (!merged || !scene || !dem || !cfg || !out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    const SeedLut lut = scene->grid_points && scene->n_grid_points > 0
                            ? build_seed_lut(scene->grid_points, scene->n_grid_points)
                            : SeedLut();
    SceneDev sdv;  // device pointers are set once the scene is uploaded
    sdv.n_lut = (int)lut.lat.size();
    sdv.epoch = scene->orbit.count > 0 ? scene->orbit.state_vectors[0].time_us : 0;
    TcPlan pl;
    sardine_status st = plan_tc(merged, scene, dem, cfg, out, sdv, &pl);
    if (st) return st;
    if ((st = require_device())) return st;

    // scene context of this call: orbit and seed table on the device
    const std::shared_ptr<const sd::OrbitDev> orb_p = orbit_cache().get(scene->orbit);
    const sd::OrbitDev& orb = *orb_p;
    set_decomposition(pl.p, g_tc_seg.load());
    const TcWork w = tc_work_bytes(pl, cfg);
    const bool fp32 = cfg->precision == SARDINE_PRECISION_FP32_CENTRED;
    hipError_t e = hipSuccess;
    auto get = [&](size_t n) { return e == hipSuccess ? static_cast<char*>(mem.get(n, &e)) : nullptr; };
    // the orbit, the seed table, the DEM tile table and the geoid grid: one upload
    double tab[sd::kDemTabWords];
    if (dem->tiling) dem_tile_table(dem->tiling, tab);
    UploadPack up;
    const size_t o_orb = up.add(&orb, sd::orbit_lds_bytes(orb.n), sizeof(sd::OrbitDev));
    const size_t o_lat = up.add(lut.lat.data(), lut.lat.size() * sizeof(double));
    const size_t o_t = up.add(lut.t.data(), lut.t.size() * sizeof(int64_t));
    const size_t o_tab = dem->tiling ? up.add(tab, sizeof tab) : 0;
    const size_t o_geoid = w.geoid ? up.add(cfg->geoid.grid_73x144, 73 * 144 * sizeof(float)) : 0;
    char* d_cnt = get(sd::kCntSlots * sizeof(unsigned long long));
    char* d_up = get(up.size());
    char* d_work = get(w.total());
    char* d_orb32 = fp32 ? get(sizeof(sd::OrbitF32)) : nullptr;
    char* d_ends = fp32 ? get((size_t)pl.nr * 2 * sizeof(float)) : nullptr;
    if (e != hipSuccess) {
        if (mem.fixed() && e == hipErrorOutOfMemory)
            return fail(SARDINE_ERR_CAPACITY,
                        "terrain correction of %lld x %lld pixels needs %zu bytes of device scratch; the "
                        "no-allocation entry point has %zu",
                        (long long)pl.nr, (long long)pl.nc, round256(mem.used()), kNoAllocScratchBytes);
        SD_HIP(e);
    }
    SD_HIP(mem.copy_in(d_up, up));
    pl.p.orbit = reinterpret_cast<const sd::OrbitDev*>(d_up + o_orb);
    pl.p.orbit_lds = (int32_t)sd::orbit_lds_bytes(orb.n);
    if (dem->tiling) pl.p.dem_tab = reinterpret_cast<const double*>(d_up + o_tab);
    pl.q.lut_lat = reinterpret_cast<const double*>(d_up + o_lat);
    pl.q.lut_t = reinterpret_cast<const int64_t*>(d_up + o_t);
    assign_work(pl, w, d_work);
    if (w.geoid) pl.p.geoid = reinterpret_cast<const float*>(d_up + o_geoid);
    // every counter starts at zero (a kernel, not a memset: capturable)
    SD_HIP(sd_launch_zero_u64(reinterpret_cast<unsigned long long*>(d_cnt), sd::kCntSlots, stream));
    pl.p.counters = reinterpret_cast<unsigned long long*>(d_cnt);
    pl.p.error_flag = reinterpret_cast<int*>(pl.p.counters + sd::kNumCounters);
    SD_HIP(sd_launch_tc_rows(&pl.q, stream));
    if (fp32) {
        // scene centre: the bounding box centre on the ellipsoid; reference
        // time: the middle of the orbit
        const sardine_bbox& bb = scene->bounding_box;
        const sd::d3 c = sd::geodetic_to_ecef(0.5 * (bb.min_lat_deg + bb.max_lat_deg),
                                              0.5 * (bb.min_lon_deg + bb.max_lon_deg), 0.0);
        pl.p.centre[0] = c.x;
        pl.p.centre[1] = c.y;
        pl.p.centre[2] = c.z;
        pl.p.t_ref_us = orb.n > 0 ? (orb.t_us[0] + orb.t_us[orb.n - 1]) / 2 : 0;
        pl.p.tol32 = 3e-5f;  // above the ~1e-5 s rounding noise of the FP32 Doppler residual
        const sd::OrbitF32 o32 = make_orbit_f32(orb, pl.p.centre, pl.p.t_ref_us);
        SD_HIP(mem.copy_in(d_orb32, &o32, sizeof o32));
        pl.p.orbit32 = reinterpret_cast<const sd::OrbitF32*>(d_orb32);
        SD_HIP(sd_launch_tc_anchors(&pl.p, stream));
        KernelTimer kt(SARDINE_KERNEL_TC, stream);
        KernelTimer kt2(SARDINE_KERNEL_TC_SEGMENT, stream);
        SD_HIP(sd_launch_tc_fp32(&pl.p, reinterpret_cast<float*>(d_ends), stream));
        kt2.stop();
        kt.stop();
    } else {
        KernelTimer kt(SARDINE_KERNEL_TC, stream);
        // the per-pixel chain kernel alone is timed as its own family
        TimerEntry* seg = timer_next(SARDINE_KERNEL_TC_SEGMENT);
        SD_HIP(sd_launch_tc_fp64(&pl.p, stream, seg ? seg->a : nullptr, seg ? seg->b : nullptr));
        if (seg) seg->pending = true;
        kt.stop();
    }
    // The counters come back through a pinned block (a copy into pageable
    // memory would wait for the kernels by itself), and the call's buffers
    // are released before the wait: the frees are stream-ordered, and their
    // host cost then overlaps the kernels instead of delaying the return.
    unsigned long long* hc = nullptr;
    SD_HIP(pinned_counters(&hc));
    SD_HIP(hipMemcpyAsync(hc, d_cnt, sd::kCntSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    mem.release();
    SD_HIP(hipStreamSynchronize(stream));
    unsigned long long cnt[sd::kCntSlots];
    std::memcpy(cnt, hc, sizeof cnt);
    return finish_tc(pl, cnt, scene->orbit, out);
}

static sardine_status sardine_terrain_correction_impl(const sardine_radar_image* merged, const sardine_scene* scene,
                                                      const sardine_dem_raster* dem, const sardine_tc_config* cfg,
                                                      sardine_geocoded_image* out, void* stream_v) {
    TcMemory mem((hipStream_t)stream_v, nullptr, 0);
    return tc_call(merged, scene, dem, cfg, out, (hipStream_t)stream_v, mem);
}
sardine_status sardine_terrain_correction(const sardine_radar_image* merged, const sardine_scene* scene,
                                          const sardine_dem_raster* dem, const sardine_tc_config* cfg,
                                          sardine_geocoded_image* out, void* stream_v) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] {
        return sardine_terrain_correction_impl(merged, scene, dem, cfg, out, stream_v);
    });
}

sardine_status sardine_terrain_correction_noalloc(const sardine_radar_image* merged, const sardine_scene* scene,
                                                  const sardine_dem_raster* dem, const sardine_tc_config* cfg,
                                                  sardine_geocoded_image* out, void* stream_v) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&]() -> sardine_status {
        sardine_status st;
        if ((st = require_device())) return st;
        char* base = nullptr;
        if ((st = scratch_on_current_device(&base))) return st;
        TcMemory mem((hipStream_t)stream_v, base, kNoAllocScratchBytes);
        return tc_call(merged, scene, dem, cfg, out, (hipStream_t)stream_v, mem);
    });
}

// ── speckle (speckle.rs:131-180) ─────────────────────────────────────────
sardine_status sardine_speckle_apply(const float* data, int64_t data_len, int64_t cols, int64_t rows,
                                     const sardine_speckle_filter* f, float* out, void* stream) {
    if (!f) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null filter");
    if (cols < 0 || rows < 0 || (rows > 0 && cols > INT64_MAX / rows))
        return fail(SARDINE_ERR_SPECKLE_DIMENSION, "data length %lld does not match cols x rows", (long long)data_len);
    const int64_t expected = cols * rows;
    if (data_len != expected)
        return fail(SARDINE_ERR_SPECKLE_DIMENSION, "data length %lld does not match cols x rows = %lld",
                    (long long)data_len, (long long)expected);
    int win = 0;
    sardine_status st = validate_filter(f, &win);
    if (st) return st;
    if (expected == 0) return SARDINE_OK;
    if (!data || !out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null raster");
    if (data == out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "speckle filtering is out-of-place: out aliases data");
    if ((st = require_device())) return st;
    KernelTimer kt(SARDINE_KERNEL_SPECKLE, (hipStream_t)stream);
    SD_HIP(sd_launch_speckle(f->kind, data, out, cols, rows, win / 2, f->enl, f->damping, (hipStream_t)stream));
    kt.stop();
    return SARDINE_OK;
}

// ── deburst layout (deburst.rs:172-305) ──────────────────────────────────
static sardine_status sardine_deburst_layout_compute_impl(const sardine_subswath* sw, const sardine_burst* b, int32_t n,
                                              int64_t tiff_height, sardine_deburst_layout* out) {
    if (!sw || !out || (n > 0 && !b)) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof *out);
    if (n <= 0) return fail(SARDINE_ERR_DEBURST_NO_BURSTS, "burst list is empty");
    if (n != sw->burst_count)
        return fail(SARDINE_ERR_DEBURST_COUNT_MISMATCH, "subswath declares %d bursts but %d were given",
                    sw->burst_count, n);
    if (n > SARDINE_MAX_BURSTS) return fail(SARDINE_ERR_CAPACITY, "%d bursts above the supported %d", n, SARDINE_MAX_BURSTS);
    const int64_t lps = sw->lines_per_burst;
    if (tiff_height < (int64_t)n * lps)
        return fail(SARDINE_ERR_DEBURST_TIFF_TOO_SMALL, "TIFF height %lld is too small for %d bursts x %lld lines",
                    (long long)tiff_height, n, (long long)lps);
    int64_t fv = b[0].first_valid_sample, lv = b[0].last_valid_sample;
    for (int k = 1; k < n; ++k) {
        fv = std::max(fv, b[k].first_valid_sample);
        lv = std::min(lv, b[k].last_valid_sample);
    }
    if (fv >= lv)
        return fail(SARDINE_ERR_DEBURST_EMPTY_VALID_WINDOW, "no valid range window: first_valid=%lld, last_valid=%lld",
                    (long long)fv, (long long)lv);
    for (int k = 0; k + 1 < n; ++k) {
        const int64_t dt_us = b[k + 1].azimuth_time_us - b[k].azimuth_time_us;
        const double dt_s = (double)dt_us * 1e-6;
        const int64_t step = (int64_t)std::round(dt_s / sw->azimuth_time_interval_s);
        const int64_t ov = lps - step;
        if (ov <= 0)
            return fail(SARDINE_ERR_DEBURST_INVALID_OVERLAP, "non-positive overlap between burst %d and burst %d: %lld",
                        b[k].burst_index, b[k + 1].burst_index, (long long)ov);
        if (ov < 50 || ov > 400)
            return fail(SARDINE_ERR_DEBURST_IMPLAUSIBLE_OVERLAP,
                        "implausible overlap between burst %d and burst %d: %lld lines is outside [50, 400]",
                        b[k].burst_index, b[k + 1].burst_index, (long long)ov);
        out->overlap[k] = ov;
    }
    int64_t total = 0;
    for (int k = 0; k < n; ++k) {
        const int64_t lf = k == 0 ? 0 : (out->overlap[k - 1] + 1) / 2;  // ⌈ov/2⌉
        const int64_t le = k == n - 1 ? lps : lps - out->overlap[k] / 2;  // lps − ⌊ov/2⌋
        out->local_first[k] = lf;
        out->local_end[k] = le;
        out->out_line_start[k] = total;
        total += le - lf;
    }
    out->n_bursts = n;
    out->lines = total;
    out->samples = lv - fv;
    out->valid_sample_offset = fv;
    return SARDINE_OK;
}
sardine_status sardine_deburst_layout_compute(const sardine_subswath* sw, const sardine_burst* b, int32_t n,
                                              int64_t tiff_height, sardine_deburst_layout* out) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] {
        return sardine_deburst_layout_compute_impl(sw, b, n, tiff_height, out);
    });
}

}  // extern "C"

// ── calibration + merge host set-up ───────────────────────────────────────
namespace {

void bracket_azimuth(const int32_t* lines, int n, int64_t tl, int* lo, int* hi) {  // apply_calibration.rs:190-202
    const int32_t* it = std::lower_bound(lines, lines + n, tl, [](int32_t a, int64_t b) { return (int64_t)a < b; });
    int i = (int)(it - lines);
    if (it != lines + n && (int64_t)*it == tl) { *lo = *hi = i; return; }
    if (i == 0) { *lo = *hi = 0; return; }
    if (i >= n) { *lo = *hi = n - 1; return; }
    *lo = i - 1;
    *hi = i;
}

float lerp_f(float a, float b, float t) {
    t = std::min(std::max(t, 0.0f), 1.0f);
    return a + (b - a) * t;
}

// LUT coverage check (apply_calibration.rs:421-462); returns the max gap.
sardine_status lut_gap(const sardine_lut* L, int64_t col_off, int64_t samples, bool is_cal, uint32_t* gap) {
    if (!L || L->n_vectors <= 0)
        return is_cal ? fail(SARDINE_ERR_CAL_NO_CAL_VECTORS, "calibration LUT has no vectors")
                      : fail(SARDINE_ERR_CAL_NO_NOISE_VECTORS, "noise range LUT has no vectors");
    const uint32_t req = (uint32_t)std::max<int64_t>(col_off + samples - 1, 0);
    uint32_t mg = 0;
    for (int i = 0; i < L->n_vectors; ++i) {
        const int64_t a = L->offsets[i], e = L->offsets[i + 1];
        if (e <= a)
            return is_cal ? fail(SARDINE_ERR_CAL_NO_CAL_VECTORS, "calibration LUT has no vectors")
                          : fail(SARDINE_ERR_CAL_NO_NOISE_VECTORS, "noise range LUT has no vectors");
        const uint32_t mx = L->pixels[e - 1];
        if (mx < req) {
            const uint32_t g = req - mx;
            if (g > 600)
                return fail(SARDINE_ERR_CAL_LUT_RANGE,
                            "calibration vector at line %d covers pixels up to %u, but output requires up to pixel %u",
                            L->lines[i], mx, req);
            mg = std::max(mg, g);
        }
    }
    *gap = mg;
    return SARDINE_OK;
}

// Per-line constants of one swath for `n_lines` merged lines.
std::vector<sd::CalLine> line_table(const sardine_lut* cal, const sardine_lut* nr, const sardine_az_lut* az,
                                    int64_t origin, int64_t n_lines, const sardine_deburst_layout* L,
                                    const sardine_burst* bursts) {
    std::vector<sd::CalLine> t((size_t)n_lines);
    int k = 0;
    for (int64_t l = 0; l < n_lines; ++l) {
        sd::CalLine& c = t[(size_t)l];
        std::memset(&c, 0, sizeof c);
        if (L) {
            while (k + 1 < L->n_bursts && l >= L->out_line_start[k + 1]) ++k;
            c.tiff_row = bursts[k].first_line + L->local_first[k] + (l - L->out_line_start[k]);
        } else {
            c.tiff_row = l;
        }
        const int64_t tl = origin + l;
        int lo, hi;
        bracket_azimuth(cal->lines, cal->n_vectors, tl, &lo, &hi);
        c.cal_lo = lo;
        c.cal_hi = hi;
        c.cal_t = lo == hi ? 0.0f
                           : (float)(tl - (int64_t)cal->lines[lo]) / (float)((int64_t)cal->lines[hi] - (int64_t)cal->lines[lo]);
        bracket_azimuth(nr->lines, nr->n_vectors, tl, &lo, &hi);
        c.nr_lo = lo;
        c.nr_hi = hi;
        c.nr_t = lo == hi ? 0.0f
                          : (float)(tl - (int64_t)nr->lines[lo]) / (float)((int64_t)nr->lines[hi] - (int64_t)nr->lines[lo]);
        c.az = 1.0f;
        if (az && az->n > 0) {  // interp_azimuth (apply_calibration.rs:168-181)
            bracket_azimuth(az->lines, az->n, tl, &lo, &hi);
            if (az->lines[hi] == az->lines[lo]) c.az = az->values[lo];
            else {
                float tt = (float)(tl - (int64_t)az->lines[lo]) / (float)((int64_t)az->lines[hi] - (int64_t)az->lines[lo]);
                c.az = lerp_f(az->values[lo], az->values[hi], tt);
            }
        }
    }
    return t;
}

// Device inputs of the calibration kernel for one or more swaths: every
// swath's LUT vectors and line table in one gathered upload, the dense LUTs
// in one allocation.
struct SwathDev {
    const float* cal_dense;
    const float* nr_dense;
    const sd::CalLine* lines;
};
struct SwathsDev {
    DevBuf packed, dense;
    std::vector<SwathDev> sw;
};

sardine_status stage_swaths(int n, const sardine_lut* const* cal, const sardine_lut* const* nr,
                            const int64_t* dense_cols, const std::vector<sd::CalLine>* lt, SwathsDev& d,
                            hipStream_t s) {
    UploadPack up;
    std::vector<size_t> o((size_t)n * 7);
    std::vector<size_t> od((size_t)n * 2);
    size_t dense = 0;
    for (int i = 0; i < n; ++i) {
        const int64_t nc = cal[i]->offsets[cal[i]->n_vectors], nn = nr[i]->offsets[nr[i]->n_vectors];
        size_t* q = &o[(size_t)i * 7];
        q[0] = up.add(cal[i]->pixels, (size_t)nc * sizeof(uint32_t));
        q[1] = up.add(cal[i]->values, (size_t)nc * sizeof(float));
        q[2] = up.add(cal[i]->offsets, ((size_t)cal[i]->n_vectors + 1) * sizeof(int64_t));
        q[3] = up.add(nr[i]->pixels, (size_t)nn * sizeof(uint32_t));
        q[4] = up.add(nr[i]->values, (size_t)nn * sizeof(float));
        q[5] = up.add(nr[i]->offsets, ((size_t)nr[i]->n_vectors + 1) * sizeof(int64_t));
        q[6] = up.add(lt[i].data(), lt[i].size() * sizeof(sd::CalLine));
        od[(size_t)i * 2] = dense;
        dense = round256(dense + (size_t)cal[i]->n_vectors * dense_cols[i] * sizeof(float));
        od[(size_t)i * 2 + 1] = dense;
        dense = round256(dense + (size_t)nr[i]->n_vectors * dense_cols[i] * sizeof(float));
    }
    SD_HIP(d.packed.alloc(up.size(), s));
    SD_HIP(up.copy(d.packed.p, s));
    SD_HIP(d.dense.alloc(dense, s));
    char* base = d.packed.as<char>();
    char* dbase = d.dense.as<char>();
    d.sw.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        const size_t* q = &o[(size_t)i * 7];
        float* cd = reinterpret_cast<float*>(dbase + od[(size_t)i * 2]);
        float* nd = reinterpret_cast<float*>(dbase + od[(size_t)i * 2 + 1]);
        SD_HIP(sd_launch_lut_dense(reinterpret_cast<const uint32_t*>(base + q[0]), reinterpret_cast<const float*>(base + q[1]),
                                   reinterpret_cast<const int64_t*>(base + q[2]), cal[i]->n_vectors, dense_cols[i], cd,
                                   s));
        SD_HIP(sd_launch_lut_dense(reinterpret_cast<const uint32_t*>(base + q[3]), reinterpret_cast<const float*>(base + q[4]),
                                   reinterpret_cast<const int64_t*>(base + q[5]), nr[i]->n_vectors, dense_cols[i], nd,
                                   s));
        d.sw[(size_t)i] = SwathDev{cd, nd, reinterpret_cast<const sd::CalLine*>(base + q[6])};
    }
    return SARDINE_OK;
}

}  // namespace

extern "C" {

static sardine_status sardine_prepare_merged_layout_impl(const sardine_swath_input* sw, int32_t n, sardine_merged_info* info) {
    if (!info || (n > 0 && !sw)) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(info, 0, sizeof *info);
    if (n <= 0) return fail(SARDINE_ERR_MERGE_EMPTY_INPUTS, "no subswaths to merge");
    if (n > SARDINE_MAX_SWATHS) return fail(SARDINE_ERR_CAPACITY, "%d subswaths above the supported %d", n, SARDINE_MAX_SWATHS);
    std::vector<sardine_deburst_layout> L((size_t)n);
    uint32_t gc = 0, gn = 0;
    for (int i = 0; i < n; ++i) {
        sardine_status st = sardine_deburst_layout_compute(&sw[i].swath, sw[i].bursts, sw[i].n_bursts, sw[i].tiff_height,
                                                           &L[(size_t)i]);
        if (st) return st;
        uint32_t a, b;
        if ((st = lut_gap(&sw[i].cal, L[(size_t)i].valid_sample_offset, L[(size_t)i].samples, true, &a))) return st;
        if ((st = lut_gap(&sw[i].noise_range, L[(size_t)i].valid_sample_offset, L[(size_t)i].samples, false, &b))) return st;
        gc = std::max(gc, a);
        gn = std::max(gn, b);
    }
    // merge_subswaths (merge_subswaths.rs:207-330)
    const double rps = sw[0].swath.range_pixel_spacing_m;
    for (int i = 1; i < n; ++i)
        if (std::fabs(sw[i].swath.range_pixel_spacing_m - rps) > 1e-3)
            return fail(SARDINE_ERR_MERGE_PIXEL_SPACING, "inconsistent range pixel spacing: first %g, swath %d has %g", rps, i,
                        sw[i].swath.range_pixel_spacing_m);
    for (int i = 0; i < n; ++i)
        if (L[(size_t)i].samples == 0) return fail(SARDINE_ERR_MERGE_EMPTY_SWATH, "subswath %d is empty", i);
    const double dt = rps * 2.0 / sd::kC;
    std::vector<double> t0((size_t)n);
    for (int i = 0; i < n; ++i) t0[(size_t)i] = sw[i].swath.slant_range_time_s + (double)L[(size_t)i].valid_sample_offset * dt;
    int64_t out_lines = L[0].lines;
    int64_t az0 = sw[0].bursts[0].azimuth_time_us;
    for (int i = 0; i < n; ++i) {
        const int64_t off = n == 1 ? 0 : (int64_t)std::round((t0[(size_t)i] - t0[0]) / dt);
        info->sw_out_offset[i] = off;
        info->swath_lines[i] = L[(size_t)i].lines;
        info->swath_samples[i] = L[(size_t)i].samples;
        info->swath_valid_offset[i] = L[(size_t)i].valid_sample_offset;
        out_lines = std::min(out_lines, L[(size_t)i].lines);
        az0 = std::min(az0, sw[i].bursts[0].azimuth_time_us);
    }
    const int64_t out_cols = info->sw_out_offset[n - 1] + L[(size_t)n - 1].samples;
    for (int i = 0; i + 1 < n; ++i)
        info->seam[i] = (info->sw_out_offset[i] + L[(size_t)i].samples + info->sw_out_offset[i + 1]) / 2;
    info->seam[n - 1] = out_cols;
    info->lines = out_lines;
    info->samples = out_cols;
    info->near_slant_range_time_s = t0[0];
    info->range_pixel_spacing_s = dt;
    info->range_pixel_spacing_m = rps;
    info->azimuth_start_time_us = az0;
    info->cal_lut_extrapolation_gap_px = gc;
    info->noise_lut_extrapolation_gap_px = gn;
    info->n_swaths = n;
    return SARDINE_OK;
}
sardine_status sardine_prepare_merged_layout(const sardine_swath_input* sw, int32_t n, sardine_merged_info* info) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] { return sardine_prepare_merged_layout_impl(sw, n, info); });
}

static sardine_status sardine_prepare_merged_impl(const sardine_swath_input* sw, int32_t n, float* sigma0, float* nesz,
                                      sardine_merged_info* info, void* stream_v) {
    sardine_status st = sardine_prepare_merged_layout(sw, n, info);
    if (st) return st;
    if (!sigma0 || !nesz) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null output raster");
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream_v;
    std::vector<std::vector<sd::CalLine>> tables((size_t)n);
    std::vector<const sardine_lut*> cals((size_t)n), nrs((size_t)n);
    std::vector<int64_t> dcols((size_t)n);
    sd::CalParams P;
    std::memset(&P, 0, sizeof P);
    if (info->samples >= INT32_MAX) return fail(SARDINE_ERR_CAPACITY, "merged raster wider than 2^31");
    P.n_sw = n;
    P.out_lines = info->lines;
    P.out_cols = info->samples;
    P.sigma0 = sigma0;
    P.nesz = nesz;
    for (int i = 0; i < n; ++i) {
        sardine_deburst_layout L;
        if ((st = sardine_deburst_layout_compute(&sw[i].swath, sw[i].bursts, sw[i].n_bursts, sw[i].tiff_height, &L)))
            return st;
        if (!sw[i].raw) return fail(SARDINE_ERR_INVALID_ARGUMENT, "swath %d raster is null", i);
        if (reinterpret_cast<uintptr_t>(sw[i].raw) % 4 != 0)
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "swath %d CInt16 raster must be 4-byte aligned", i);
        if (sw[i].raw_cols >= INT32_MAX) return fail(SARDINE_ERR_CAPACITY, "swath %d raster wider than 2^31", i);
        if (L.valid_sample_offset + L.samples > sw[i].raw_cols)
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "swath %d valid window exceeds the raster width", i);
        const int64_t dense_cols = L.valid_sample_offset + L.samples;
        tables[(size_t)i] = line_table(&sw[i].cal, &sw[i].noise_range, &sw[i].noise_azimuth, sw[i].bursts[0].first_line,
                                       info->lines, &L, sw[i].bursts);
        cals[(size_t)i] = &sw[i].cal;
        nrs[(size_t)i] = &sw[i].noise_range;
        dcols[(size_t)i] = dense_cols;
        sd::CalSwathDev& D = P.sw[i];
        D.raw = sw[i].raw;
        D.raw_cols = sw[i].raw_cols;
        D.first_valid = L.valid_sample_offset;
        D.dense_cols = dense_cols;
        D.out_off = info->sw_out_offset[i];
        const int64_t col_start = i == 0 ? 0 : info->seam[i - 1];
        const int64_t col_end = i == n - 1 ? info->samples : info->seam[i];
        D.own_end = col_end;
        D.valid_start = std::max(col_start, D.out_off);
        D.valid_end = std::min(col_end, D.out_off + L.samples);
    }
    SwathsDev dev;
    if ((st = stage_swaths(n, cals.data(), nrs.data(), dcols.data(), tables.data(), dev, s))) return st;
    for (int i = 0; i < n; ++i) {
        P.sw[i].cal_dense = dev.sw[(size_t)i].cal_dense;
        P.sw[i].nr_dense = dev.sw[(size_t)i].nr_dense;
        P.sw[i].lines = dev.sw[(size_t)i].lines;
    }
    {
        KernelTimer kt(SARDINE_KERNEL_CALIBRATE, s);
        SD_HIP(sd_launch_cal_merge(&P, s));
        kt.stop();
    }
    return SARDINE_OK;  // stream-ordered: the work buffers are released after the kernel
}
sardine_status sardine_prepare_merged(const sardine_swath_input* sw, int32_t n, float* sigma0, float* nesz,
                                      sardine_merged_info* info, void* stream_v) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] {
        return sardine_prepare_merged_impl(sw, n, sigma0, nesz, info, stream_v);
    });
}

static sardine_status sardine_apply_calibration_impl(const int16_t* deburst, int64_t lines, int64_t samples, int64_t col_off,
                                         const sardine_lut* cal, const sardine_lut* nr, const sardine_az_lut* az,
                                         int64_t origin, float* out, float* nesz, uint32_t* cal_gap, uint32_t* nr_gap,
                                         void* stream_v) {
    if (!cal || !nr) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null LUT");
    uint32_t a = 0, b = 0;
    sardine_status st;
    if ((st = lut_gap(cal, col_off, samples, true, &a))) return st;
    if ((st = lut_gap(nr, col_off, samples, false, &b))) return st;
    if (cal_gap) *cal_gap = a;
    if (nr_gap) *nr_gap = b;
    if (lines <= 0 || samples <= 0) return SARDINE_OK;
    if (!deburst || !out || !nesz) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null raster");
    if (reinterpret_cast<uintptr_t>(deburst) % 4 != 0)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "CInt16 raster must be 4-byte aligned");
    if (col_off + samples >= INT32_MAX) return fail(SARDINE_ERR_CAPACITY, "raster wider than 2^31");
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream_v;
    auto lt = line_table(cal, nr, az, origin, lines, nullptr, nullptr);
    SwathsDev dev;
    const int64_t dcols = col_off + samples;
    if ((st = stage_swaths(1, &cal, &nr, &dcols, &lt, dev, s))) return st;
    const SwathDev& d = dev.sw[0];
    sd::CalParams P;
    std::memset(&P, 0, sizeof P);
    P.n_sw = 1;
    P.out_lines = lines;
    P.out_cols = samples;
    P.sigma0 = out;
    P.nesz = nesz;
    sd::CalSwathDev& D = P.sw[0];
    // the debursted array is its own "TIFF": row = line, column 0 = TIFF col_off
    D.raw = deburst - 2 * col_off;  // address arithmetic only; indices stay in [col_off, col_off+samples)
    D.raw_cols = samples;
    D.first_valid = col_off;
    D.cal_dense = d.cal_dense;
    D.nr_dense = d.nr_dense;
    D.dense_cols = dcols;
    D.lines = d.lines;
    D.out_off = 0;
    D.own_end = samples;
    D.valid_start = 0;
    D.valid_end = samples;
    SD_HIP(sd_launch_cal_merge(&P, s));
    return SARDINE_OK;
}
sardine_status sardine_apply_calibration(const int16_t* deburst, int64_t lines, int64_t samples, int64_t col_off,
                                         const sardine_lut* cal, const sardine_lut* nr, const sardine_az_lut* az,
                                         int64_t origin, float* out, float* nesz, uint32_t* cal_gap, uint32_t* nr_gap,
                                         void* stream_v) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] {
        return sardine_apply_calibration_impl(deburst, lines, samples, col_off, cal, nr, az, origin, out, nesz, cal_gap, nr_gap, stream_v);
    });
}

sardine_status sardine_to_db_inplace(float* data, int64_t n, float floor_lin, uint64_t counts[2], void* stream_v) {
    if (floor_lin < 0.0f) return fail(SARDINE_ERR_INVALID_ARGUMENT, "invalid noise floor %g", (double)floor_lin);
    if (n < 0 || (n > 0 && !data)) return fail(SARDINE_ERR_INVALID_ARGUMENT, "bad raster");
    sardine_status st;
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream_v;
    DevBuf c;
    SD_HIP(c.alloc(2 * sizeof(unsigned long long), s));
    SD_HIP(sd_launch_zero_u64(c.as<unsigned long long>(), 2, s));
    {
        KernelTimer kt(SARDINE_KERNEL_TO_DB, s);
        SD_HIP(sd_launch_to_db(data, n, floor_lin, c.as<unsigned long long>(), s));
        kt.stop();
    }
    unsigned long long h[2];
    SD_HIP(hipMemcpyAsync(h, c.p, sizeof h, hipMemcpyDeviceToHost, s));
    SD_HIP(hipStreamSynchronize(s));
    if (counts) {
        counts[0] = h[0];
        counts[1] = h[1];
    }
    return SARDINE_OK;
}

sardine_status sardine_diag_math(int32_t fn, const double* x, const double* y, double* out, int64_t n,
                                 void* stream) {
    if (fn < 0 || fn > 10 || n < 0 || (n > 0 && (!x || !y || !out)))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "bad diagnostic math request");
    sardine_status st;
    if ((st = require_device())) return st;
    SD_HIP(sd_launch_pmath(fn, x, y, out, n, (hipStream_t)stream));
    return SARDINE_OK;
}

sardine_status sardine_diag_dem(const sardine_dem_raster* dem, const double* lat, const double* lon, float* out,
                                int64_t n, void* stream) {
    if (!dem || n < 0 || (n > 0 && (!lat || !lon || !out)))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "bad diagnostic DEM request");
    if (!dem->data || dem->width < 2 || dem->height < 2 || dem->width >= INT32_MAX || dem->height >= INT32_MAX ||
        !(dem->pixel_deg > 0.0))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "bad DEM raster");
    sardine_status st;
    if ((st = validate_dem_tiling(dem))) return st;
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream;
    sd::TcParams p;
    std::memset(&p, 0, sizeof p);
    DevBuf tab;
    double host_tab[sd::kDemTabWords];
    if (dem->tiling) {
        dem_tile_table(dem->tiling, host_tab);
        SD_HIP(tab.alloc(sizeof host_tab, s));
        SD_HIP(hipMemcpyAsync(tab.p, host_tab, sizeof host_tab, hipMemcpyHostToDevice, s));
    }
    fill_dem_params(p, dem, tab.as<double>());
    SD_HIP(sd_launch_dem_probe(&p, lat, lon, out, n, s));
    if (dem->tiling) SD_HIP(hipStreamSynchronize(s));  // host_tab and the table buffer end here
    return SARDINE_OK;
}

sardine_status sardine_synth_cint16(int16_t* out, int64_t n, uint64_t seed, uint64_t base, double sigma, void* stream) {
    if (n < 0 || (n > 0 && !out)) return fail(SARDINE_ERR_INVALID_ARGUMENT, "bad raster");
    sardine_status st;
    if ((st = require_device())) return st;
    SD_HIP(sd_launch_synth_cint16(out, n, seed, base, sigma, (hipStream_t)stream));
    return SARDINE_OK;
}

}  // extern "C"

// ── geoid grid files (geoid.rs:132-299) ──────────────────────────────────
namespace {

constexpr size_t kGeoidValues = 73 * 144;

bool read_file(const char* path, std::string* out) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    out->clear();
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out->append(buf, n);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

// A token Rust's `str::parse::<f64>()` / `::<f32>()` accepts: an optional
// sign, then "inf", "infinity" or "nan" (any case), or digits with an
// optional fraction (at least one digit in all) and an optional exponent
// with at least one digit.  No hex, no partial parses.
bool is_rust_float(const std::string& t) {
    size_t i = 0;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
    std::string rest = t.substr(i);
    for (char& c : rest) c = (char)std::tolower((unsigned char)c);
    if (rest == "inf" || rest == "infinity" || rest == "nan") return true;
    size_t digits = 0;
    while (i < t.size() && std::isdigit((unsigned char)t[i])) ++i, ++digits;
    if (i < t.size() && t[i] == '.') {
        ++i;
        while (i < t.size() && std::isdigit((unsigned char)t[i])) ++i, ++digits;
    }
    if (digits == 0) return false;
    if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
        ++i;
        if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
        size_t ed = 0;
        while (i < t.size() && std::isdigit((unsigned char)t[i])) ++i, ++ed;
        if (ed == 0) return false;
    }
    return i == t.size();
}

std::vector<std::string> split_ws(const std::string& line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
        const size_t a = i;
        while (i < line.size() && !std::isspace((unsigned char)line[i])) ++i;
        if (i > a) out.push_back(line.substr(a, i - a));
    }
    return out;
}

}  // namespace

extern "C" {

static sardine_status sardine_geoid_load_binary_impl(const char* path, int32_t model, float grid[10512]) {
    if (!path || !grid) return fail(SARDINE_ERR_INVALID_ARGUMENT, "path or grid is null");
    if (model != SARDINE_GEOID_MODEL_EGM96 && model != SARDINE_GEOID_MODEL_EGM2008)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "unknown geoid model %d", model);
    std::string raw;
    if (!read_file(path, &raw)) return fail(SARDINE_ERR_GEOID_IO, "IO error reading %s: %s", path, std::strerror(errno));
    if (raw.size() != kGeoidValues * 4)
        return fail(model == SARDINE_GEOID_MODEL_EGM96 ? SARDINE_ERR_GEOID_WRONG_SIZE : SARDINE_ERR_GEOID_WRONG_SIZE_EGM2008,
                    "%s binary grid file %s has wrong size: expected %zu bytes, got %zu",
                    model == SARDINE_GEOID_MODEL_EGM96 ? "EGM96" : "EGM2008", path, kGeoidValues * 4, raw.size());
    std::memcpy(grid, raw.data(), raw.size());  // little-endian host, as the file
    return SARDINE_OK;
}
sardine_status sardine_geoid_load_binary(const char* path, int32_t model, float grid[10512]) {
    return guarded(SARDINE_ERR_GEOID_IO, [&] { return sardine_geoid_load_binary_impl(path, model, grid); });
}

static sardine_status sardine_geoid_load_ww15mgh_impl(const char* path, float grid[10512]) {
    if (!path || !grid) return fail(SARDINE_ERR_INVALID_ARGUMENT, "path or grid is null");
    std::string text;
    if (!read_file(path, &text)) return fail(SARDINE_ERR_GEOID_IO, "IO error reading %s: %s", path, std::strerror(errno));
    std::vector<float> values;
    values.reserve(kGeoidValues);
    int headers = 0;
    size_t pos = 0;
    while (pos <= text.size()) {
        size_t nl = text.find('\n', pos);
        if (nl == std::string::npos) nl = text.size();
        const std::vector<std::string> tok = split_ws(text.substr(pos, nl - pos));
        pos = nl + 1;
        if (tok.empty()) continue;
        // up to two header lines: text, or the numeric bounds line (<= 6 numbers)
        if (headers < 2) {
            bool numeric = true;
            for (const std::string& t : tok) numeric = numeric && is_rust_float(t);
            if (!numeric || tok.size() <= 6) {
                ++headers;
                continue;
            }
        }
        for (const std::string& t : tok) {
            if (!is_rust_float(t))
                return fail(SARDINE_ERR_GEOID_MALFORMED_ASCII,
                            "EGM96 ASCII grid file %s is malformed: cannot parse token '%s' as f32", path, t.c_str());
            values.push_back(std::strtof(t.c_str(), nullptr));
        }
    }
    if (values.size() != kGeoidValues)
        return fail(SARDINE_ERR_GEOID_MALFORMED_ASCII,
                    "EGM96 ASCII grid file %s is malformed: expected %zu undulation values, found %zu", path,
                    kGeoidValues, values.size());
    std::memcpy(grid, values.data(), kGeoidValues * sizeof(float));
    return SARDINE_OK;
}
sardine_status sardine_geoid_load_ww15mgh(const char* path, float grid[10512]) {
    return guarded(SARDINE_ERR_GEOID_IO, [&] { return sardine_geoid_load_ww15mgh_impl(path, grid); });
}

}  // extern "C"

// ── SLC measurement TIFF → HBM (slc_reader.rs:200-590) ───────────────────
namespace {

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) ::close(fd);
    }
};

// pread of exactly n bytes (retrying short reads); false on error or EOF.
bool pread_all(int fd, void* dst, size_t n, uint64_t off) {
    uint8_t* p = static_cast<uint8_t*>(dst);
    while (n > 0) {
        const ssize_t r = ::pread(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += (uint64_t)r;
    }
    return true;
}

uint16_t rd16(const uint8_t* p, bool le) { return le ? (uint16_t)(p[0] | (p[1] << 8)) : (uint16_t)((p[0] << 8) | p[1]); }
uint32_t rd32(const uint8_t* p, bool le) {
    return le ? ((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24))
              : (((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3]);
}

// parse_ifd (slc_reader.rs:268-461).
sardine_status slc_parse(int fd, const char* path, sardine_slc_tiff* out) {
    uint8_t h[8];
    if (!pread_all(fd, h, 8, 0)) return fail(SARDINE_ERR_SLC_IO, "I/O error reading the TIFF header of %s", path);
    bool le;
    if (h[0] == 'I' && h[1] == 'I') le = true;
    else if (h[0] == 'M' && h[1] == 'M') le = false;
    else return fail(SARDINE_ERR_SLC_NOT_TIFF, "not a classic TIFF (bad byte-order marker or magic=0)");
    const uint16_t magic = rd16(h + 2, le);
    if (magic == 43)
        return fail(SARDINE_ERR_SLC_BIGTIFF, "BigTIFF (magic=43) is not supported; only classic TIFF (magic=42) is accepted");
    if (magic != 42) return fail(SARDINE_ERR_SLC_NOT_TIFF, "not a classic TIFF (bad byte-order marker or magic=%u)", magic);
    const uint64_t ifd = rd32(h + 4, le);
    uint8_t nb[2];
    if (!pread_all(fd, nb, 2, ifd)) return fail(SARDINE_ERR_SLC_IO, "I/O error reading the IFD of %s", path);
    const size_t n = rd16(nb, le);
    std::vector<uint8_t> ent(n * 12);
    if (n && !pread_all(fd, ent.data(), ent.size(), ifd + 2))
        return fail(SARDINE_ERR_SLC_IO, "I/O error reading the IFD entries of %s", path);
    // the tags the reader needs: an entry of another type than SHORT/LONG
    // counts as absent (entry_scalar), a later entry replaces an earlier one
    struct Tag {
        bool have = false;
        uint32_t v = 0;
    } width, height, bps, comp, rps, spp, fmt;
    bool have_so = false;
    uint16_t so_type = 0;
    uint32_t so_count = 0, so_field = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t* e = &ent[i * 12];
        const uint16_t tag = rd16(e, le), type = rd16(e + 2, le);
        Tag t;
        t.have = type == 3 || type == 4;
        t.v = type == 3 ? rd16(e + 8, le) : rd32(e + 8, le);
        switch (tag) {
            case 256: width = t; break;
            case 257: height = t; break;
            case 258: bps = {t.have, t.v & 0xffffu}; break;
            case 259: comp = {t.have, t.v & 0xffffu}; break;
            case 277: spp = {t.have, t.v & 0xffffu}; break;
            case 278: rps = t; break;
            case 339: fmt = {t.have, t.v & 0xffffu}; break;
            case 273:
                have_so = true;
                so_type = type;
                so_count = rd32(e + 4, le);
                so_field = rd32(e + 8, le);
                break;
            default: break;
        }
    }
    const std::pair<const Tag*, int> required[] = {{&width, 256}, {&height, 257}, {&bps, 258}, {&comp, 259},
                                                   {&rps, 278}};
    for (const auto& r : required)
        if (!r.first->have) return fail(SARDINE_ERR_SLC_MISSING_TAG, "required TIFF tag %d is absent from the IFD", r.second);
    const uint32_t samples_pp = spp.have ? spp.v : 1;  // TIFF default
    if (!fmt.have) return fail(SARDINE_ERR_SLC_MISSING_TAG, "required TIFF tag 339 is absent from the IFD");
    if (!have_so) return fail(SARDINE_ERR_SLC_MISSING_TAG, "required TIFF tag 273 is absent from the IFD");
    if (comp.v != 1) return fail(SARDINE_ERR_SLC_UNSUPPORTED_FORMAT, "unsupported TIFF format: Compression=%u (expected 1=None)", comp.v);
    if (fmt.v != 5) return fail(SARDINE_ERR_SLC_UNSUPPORTED_FORMAT, "unsupported TIFF format: SampleFormat=%u (expected 5=CInt)", fmt.v);
    if (bps.v != 32) return fail(SARDINE_ERR_SLC_UNSUPPORTED_FORMAT, "unsupported TIFF format: BitsPerSample=%u (expected 32 for CInt16)", bps.v);
    if (rps.v != 1) return fail(SARDINE_ERR_SLC_UNSUPPORTED_FORMAT, "unsupported TIFF format: RowsPerStrip=%u (expected 1)", rps.v);
    if (samples_pp != 1)
        return fail(SARDINE_ERR_SLC_UNSUPPORTED_FORMAT, "unsupported TIFF format: SamplesPerPixel=%u (expected 1 for CInt16)", samples_pp);
    if (so_count != height.v)
        return fail(SARDINE_ERR_SLC_UNSUPPORTED_FORMAT, "unsupported TIFF format: strip count (%u) != image height (%u)", so_count, height.v);
    const uint64_t row = (uint64_t)width.v * 4;
    uint64_t first, last;
    if (so_count == 1) {  // the single offset sits in the entry's value field
        first = last = so_field;
    } else {
        if (so_type != 3 && so_type != 4)
            return fail(SARDINE_ERR_SLC_UNSUPPORTED_FORMAT, "unsupported TIFF format: StripOffsets field type %u is not SHORT(3) or LONG(4)", so_type);
        const uint64_t each = so_type == 3 ? 2 : 4;
        uint8_t b[4];
        if (!pread_all(fd, b, each, so_field)) return fail(SARDINE_ERR_SLC_IO, "I/O error reading the strip offsets of %s", path);
        first = each == 2 ? rd16(b, le) : rd32(b, le);
        if (!pread_all(fd, b, each, so_field + ((uint64_t)so_count - 1) * each))
            return fail(SARDINE_ERR_SLC_IO, "I/O error reading the strip offsets of %s", path);
        last = each == 2 ? rd16(b, le) : rd32(b, le);
    }
    const uint64_t expected = first + ((uint64_t)height.v - 1) * row;
    if (last != expected)
        return fail(SARDINE_ERR_SLC_NON_CONTIGUOUS_STRIPS,
                    "strip offsets not contiguous (strip[0]=%llu, strip[%u-1]=%llu, expected=%llu)",
                    (unsigned long long)first, so_count, (unsigned long long)last, (unsigned long long)expected);
    out->width = width.v;
    out->height = height.v;
    out->little_endian = le ? 1 : 0;
    out->pad = 0;
    out->base_offset = first;
    out->row_bytes = row;
    return SARDINE_OK;
}

// Pinned chunks for file → HBM reads, kept for the process: two sets of
// kSlcSet chunks; a set is filled by parallel preads while the other one's
// copies run.  One reader at a time.
constexpr int kSlcSet = 8;
constexpr size_t kSlcChunk = 8u << 20;
struct SlcPinned {
    std::mutex mu;
    int dev = -1;
    void* buf[2][kSlcSet] = {};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    int next = 0;  // the set the next read fills first (alternates across calls)
};
SlcPinned& slc_pinned() {
    static SlcPinned* p = new SlcPinned();  // never destroyed: pinned memory outlives static teardown
    return *p;
}

// Bytes [base, base + bytes) of an open file into device memory at `out`,
// in stream order on s.  A short read fails with io_err and io_msg.  The
// chunks are reused by the next call only after these copies: a call waits
// on a set's event before refilling it.
sardine_status file_to_device(int fd, uint64_t base, uint64_t bytes, uint8_t* out, hipStream_t s,
                              sardine_status io_err, const std::string& io_msg) {
    if (bytes == 0) return SARDINE_OK;
    SlcPinned& P = slc_pinned();
    std::lock_guard<std::mutex> g(P.mu);
    int dev = 0;
    SD_HIP(hipGetDevice(&dev));
    if (P.dev != dev) {  // (re)create the chunks on this device's context
        for (int k = 0; k < 2; ++k) {
            if (P.done[k]) {
                SD_HIP(hipEventSynchronize(P.done[k]));
                (void)hipEventDestroy(P.done[k]);
                P.done[k] = nullptr;
            }
            for (int i = 0; i < kSlcSet; ++i)
                if (P.buf[k][i]) {
                    (void)hipHostFree(P.buf[k][i]);
                    P.buf[k][i] = nullptr;
                }
            P.used[k] = false;
        }
        P.next = 0;
        for (int k = 0; k < 2; ++k) {
            for (int i = 0; i < kSlcSet; ++i) SD_HIP(hipHostMalloc(&P.buf[k][i], kSlcChunk, hipHostMallocDefault));
            SD_HIP(hipEventCreateWithFlags(&P.done[k], hipEventDisableTiming));
        }
        P.dev = dev;
    }
    const uint64_t per_set = (uint64_t)kSlcSet * kSlcChunk;
    // start on the set the previous read did not end on, so back-to-back
    // reads (tiles, bursts) overlap one read's copies with the next's preads
    int set = P.next;
    for (uint64_t done = 0; done < bytes; done += per_set, set ^= 1) {
        if (P.used[set]) SD_HIP(hipEventSynchronize(P.done[set]));  // its previous copies have finished
        const uint64_t n_set = std::min(per_set, bytes - done);
        // the set's bytes spread over all its chunks (4 KiB multiples), so a
        // short read still runs kSlcSet preads in parallel
        const uint64_t chunk = std::min<uint64_t>(kSlcChunk, ((n_set + kSlcSet - 1) / kSlcSet + 4095) & ~(uint64_t)4095);
        const int n_chunks = (int)((n_set + chunk - 1) / chunk);
        bool ok[kSlcSet] = {};
        std::vector<std::thread> workers;
        for (int i = 1; i < n_chunks; ++i) {
            const uint64_t off = (uint64_t)i * chunk, n = std::min<uint64_t>(chunk, n_set - off);
            workers.emplace_back([&, i, off, n] { ok[i] = pread_all(fd, P.buf[set][i], (size_t)n, base + done + off); });
        }
        ok[0] = pread_all(fd, P.buf[set][0], (size_t)std::min<uint64_t>(chunk, n_set), base + done);
        for (auto& w : workers) w.join();
        for (int i = 0; i < n_chunks; ++i)
            if (!ok[i]) return fail(io_err, "%s", io_msg.c_str());
        for (int i = 0; i < n_chunks; ++i) {
            const uint64_t off = (uint64_t)i * chunk, n = std::min<uint64_t>(chunk, n_set - off);
            SD_HIP(hipMemcpyAsync(out + done + off, P.buf[set][i], (size_t)n, hipMemcpyHostToDevice, s));
        }
        SD_HIP(hipEventRecord(P.done[set], s));
        P.used[set] = true;
        P.next = set ^ 1;
    }
    return SARDINE_OK;
}

}  // namespace

extern "C" {

static sardine_status sardine_slc_open_impl(const char* path, sardine_slc_tiff* tiff) {
    if (!path || !tiff) return fail(SARDINE_ERR_INVALID_ARGUMENT, "path or tiff is null");
    Fd f;
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) return fail(SARDINE_ERR_SLC_IO, "I/O error: %s: %s", path, std::strerror(errno));
    return slc_parse(f.fd, path, tiff);
}
sardine_status sardine_slc_open(const char* path, sardine_slc_tiff* tiff) {
    return guarded(SARDINE_ERR_SLC_IO, [&] { return sardine_slc_open_impl(path, tiff); });
}

static sardine_status sardine_slc_read_lines_impl(const char* path, const sardine_slc_tiff* t, int64_t first_line,
                                      int64_t line_count, int16_t* dst, void* stream_v) {
    if (!path || !t) return fail(SARDINE_ERR_INVALID_ARGUMENT, "path or tiff is null");
    if (first_line < 0 || line_count < 0 || first_line > (int64_t)t->height ||
        line_count > (int64_t)t->height - first_line)
        return fail(SARDINE_ERR_SLC_LINE_OUT_OF_BOUNDS, "line range [%lld, %lld) is out of image bounds [0, %u)",
                    (long long)first_line, (long long)(first_line + line_count), t->height);
    if (t->row_bytes != (uint64_t)t->width * 4) return fail(SARDINE_ERR_INVALID_ARGUMENT, "inconsistent TIFF layout");
    const uint64_t bytes = (uint64_t)line_count * t->row_bytes;
    if (bytes == 0) return SARDINE_OK;
    if (!dst) return fail(SARDINE_ERR_INVALID_ARGUMENT, "dst is null");
    sardine_status st;
    if ((st = require_device())) return st;
    Fd f;
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) return fail(SARDINE_ERR_SLC_IO, "I/O error: %s: %s", path, std::strerror(errno));
    hipStream_t s = (hipStream_t)stream_v;
    const uint64_t base = t->base_offset + (uint64_t)first_line * t->row_bytes;
    if ((st = file_to_device(f.fd, base, bytes, reinterpret_cast<uint8_t*>(dst), s, SARDINE_ERR_SLC_IO,
                             std::string("I/O error reading ") + path + " (file shorter than its IFD says?)")))
        return st;
    if (!t->little_endian) SD_HIP(sd_launch_swap16(reinterpret_cast<uint16_t*>(dst), (int64_t)(bytes / 2), s));
    return SARDINE_OK;
}
sardine_status sardine_slc_read_lines(const char* path, const sardine_slc_tiff* t, int64_t first_line,
                                      int64_t line_count, int16_t* dst, void* stream_v) {
    return guarded(SARDINE_ERR_SLC_IO, [&] {
        return sardine_slc_read_lines_impl(path, t, first_line, line_count, dst, stream_v);
    });
}

static sardine_status sardine_slc_read_lines_multi_impl(const char* const* paths, const sardine_slc_tiff* tiffs,
                                            int32_t n_slices, int64_t first_line, int64_t line_count, int16_t* dst,
                                            void* stream) {
    if (n_slices <= 0) return fail(SARDINE_ERR_SLC_MULTI_EMPTY_INPUT, "MultiSlcReader requires at least one slice reader");
    if (!paths || !tiffs) return fail(SARDINE_ERR_INVALID_ARGUMENT, "paths or tiffs is null");
    for (int32_t k = 1; k < n_slices; ++k)
        if (tiffs[k].width != tiffs[0].width)
            return fail(SARDINE_ERR_SLC_MULTI_WIDTH_MISMATCH, "slice %d width %u != slice 0 width %u", k, tiffs[k].width,
                        tiffs[0].width);
    // offset[k] = the heights of the slices before k
    std::vector<int64_t> off((size_t)n_slices);
    int64_t total = 0;
    for (int32_t k = 0; k < n_slices; ++k) {
        off[(size_t)k] = total;
        total += tiffs[k].height;
    }
    if (first_line < 0 || line_count < 0 || first_line > total || line_count > total - first_line)
        return fail(SARDINE_ERR_SLC_LINE_OUT_OF_BOUNDS, "line range [%lld, %lld) is out of image bounds [0, %lld)",
                    (long long)first_line, (long long)(first_line + line_count), (long long)total);
    // the owning slice: the last k with offset[k] <= first_line
    const int32_t k = (int32_t)(std::upper_bound(off.begin(), off.end(), first_line) - off.begin()) - 1;
    const int64_t slice_end = k + 1 < n_slices ? off[(size_t)k + 1] : total;
    if (first_line + line_count > slice_end)
        return fail(SARDINE_ERR_SLC_CROSS_SLICE_BOUNDARY,
                    "burst read [%lld, %lld) crosses a slice boundary at logical line %lld",
                    (long long)first_line, (long long)(first_line + line_count), (long long)slice_end);
    return sardine_slc_read_lines(paths[k], &tiffs[k], first_line - off[(size_t)k], line_count, dst, stream);
}
sardine_status sardine_slc_read_lines_multi(const char* const* paths, const sardine_slc_tiff* tiffs,
                                            int32_t n_slices, int64_t first_line, int64_t line_count, int16_t* dst,
                                            void* stream) {
    return guarded(SARDINE_ERR_SLC_IO, [&] {
        return sardine_slc_read_lines_multi_impl(paths, tiffs, n_slices, first_line, line_count, dst, stream);
    });
}

}  // extern "C"

// ── SRTM 1" tiles → one device DEM raster (dem.rs:129-178, 414-437) ──────
namespace {

constexpr int kSrtmDim = 3601;

// Path::file_stem: the file name without its last extension (a leading dot
// is part of the name).
std::string file_stem(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = name.find_last_of('.');
    if (dot != std::string::npos && dot > 0) name = name.substr(0, dot);
    return name;
}

// Rust's `i32::from_str` on a short field: an optional sign then digits.
bool parse_i32_field(const std::string& s, int* out) {
    size_t i = 0;
    bool neg = false;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (i == s.size()) return false;
    long v = 0;
    for (; i < s.size(); ++i) {
        if (!std::isdigit((unsigned char)s[i])) return false;
        v = v * 10 + (s[i] - '0');
    }
    *out = (int)(neg ? -v : v);
    return true;
}

// parse_srtm_tile_name (dem.rs:414-437): "N50E010" → (50, 10); characters
// after the seventh are ignored.
bool srtm_tile_name(const std::string& stem, int* lat, int* lon) {
    if (stem.size() < 7) return false;
    const char a = stem[0], b = stem[3];
    int ls, ms, la, lo;
    if (a == 'N' || a == 'n') ls = 1;
    else if (a == 'S' || a == 's') ls = -1;
    else return false;
    if (!parse_i32_field(stem.substr(1, 2), &la)) return false;
    if (b == 'E' || b == 'e') ms = 1;
    else if (b == 'W' || b == 'w') ms = -1;
    else return false;
    if (!parse_i32_field(stem.substr(4, 3), &lo)) return false;
    *lat = la * ls;
    *lon = lo * ms;
    return true;
}

}  // namespace

extern "C" {

static sardine_status sardine_dem_srtm_mosaic_impl(const char* const* paths, int32_t n_paths, float* dst,
                                       sardine_dem_raster* out, void* stream) {
    if (!out || (n_paths > 0 && !paths)) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    if (n_paths <= 0) return fail(SARDINE_ERR_DEM_NO_TILES, "no DEM tiles");
    // tile by tile in the given order, as load_srtm_files loads them: the
    // name, then (device path) the file and its size
    const size_t bytes = (size_t)kSrtmDim * kSrtmDim * 2;
    std::vector<std::pair<int, int>> sw((size_t)n_paths);
    for (int32_t i = 0; i < n_paths; ++i) {
        const std::string stem = file_stem(paths[i] ? paths[i] : "");
        if (!srtm_tile_name(stem, &sw[i].first, &sw[i].second))
            return fail(SARDINE_ERR_DEM_UNRECOGNISED_TILE_NAME,
                        "cannot determine tile origin for '%s': filename does not match SRTM (N50E010) naming",
                        stem.c_str());
        if (!dst) continue;
        Fd f;
        f.fd = ::open(paths[i], O_RDONLY | O_CLOEXEC);
        struct stat stt;
        if (f.fd < 0 || ::fstat(f.fd, &stt) != 0)
            return fail(SARDINE_ERR_DEM_IO, "IO error reading %s: %s", paths[i], std::strerror(errno));
        if ((size_t)stt.st_size != bytes)
            return fail(SARDINE_ERR_DEM_BAD_SIZE, "tile %s has wrong size: expected %zu bytes, got %lld", paths[i],
                        bytes, (long long)stt.st_size);
    }
    int lat_lo = INT32_MAX, lat_hi = INT32_MIN, lon_lo = INT32_MAX, lon_hi = INT32_MIN;
    for (const auto& t : sw) {
        lat_lo = std::min(lat_lo, t.first);
        lat_hi = std::max(lat_hi, t.first + 1);
        lon_lo = std::min(lon_lo, t.second);
        lon_hi = std::max(lon_hi, t.second + 1);
    }
    const int64_t w = (int64_t)(lon_hi - lon_lo) * (kSrtmDim - 1) + 1;
    const int64_t h = (int64_t)(lat_hi - lat_lo) * (kSrtmDim - 1) + 1;
    if (w > INT32_MAX || h > INT32_MAX || w * h > (int64_t)1 << 40)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "tiles span %lld x %lld samples", (long long)w, (long long)h);
    sardine_dem_raster r;
    r.tiling = nullptr;
    r.data = dst;
    r.width = w;
    r.height = h;
    r.origin_lat_deg = (double)lat_hi;
    r.origin_lon_deg = (double)lon_lo;
    r.pixel_deg = 1.0 / 3600.0;
    if (!dst) {  // layout query: names only (the files are read on the device path)
        *out = r;
        return SARDINE_OK;
    }
    sardine_status st;
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream;
    SD_HIP(sd_launch_fill_nan(dst, w * h, s));
    // each tile's bytes go file → pinned chunks → HBM (the next tile's reads
    // overlap this tile's copies and decode), then the big-endian decode
    // places it in the mosaic; the tile buffer is reused in stream order
    // Tiles are placed last to first, so where tiles overlap (the shared
    // edge rows and columns of adjacent SRTM tiles) the first tile in the
    // list supplies the value, as DemMosaic's lookup answers from the first
    // covering tile.
    DevBuf tile;
    SD_HIP(tile.alloc(bytes, s));
    for (int32_t i = n_paths - 1; i >= 0; --i) {
        Fd f;
        f.fd = ::open(paths[i], O_RDONLY | O_CLOEXEC);
        if (f.fd < 0) return fail(SARDINE_ERR_DEM_IO, "IO error reading %s", paths[i]);
        if ((st = file_to_device(f.fd, 0, bytes, tile.as<uint8_t>(), s, SARDINE_ERR_DEM_IO,
                                 std::string("IO error reading ") + paths[i])))
            return st;
        const int64_t row0 = (int64_t)(lat_hi - (sw[i].first + 1)) * (kSrtmDim - 1);
        const int64_t col0 = (int64_t)(sw[i].second - lon_lo) * (kSrtmDim - 1);
        KernelTimer kt(SARDINE_KERNEL_DEM, s);
        SD_HIP(sd_launch_srtm_tile(tile.as<uint16_t>(), kSrtmDim, dst, w, row0, col0, s));
        kt.stop();
    }
    *out = r;
    return SARDINE_OK;
}
sardine_status sardine_dem_srtm_mosaic(const char* const* paths, int32_t n_paths, float* dst,
                                       sardine_dem_raster* out, void* stream) {
    return guarded(SARDINE_ERR_DEM_IO, [&] {
        return sardine_dem_srtm_mosaic_impl(paths, n_paths, dst, out, stream);
    });
}

}  // extern "C"

// ── Copernicus GLO-30 GeoTIFF tile → device DEM raster (dem.rs:248-352) ──
namespace {

// parse_glo30_tile_name (dem.rs:445-500): an SRTM-style 7-character stem, or
// the Copernicus name's N../S.. and E../W.. tokens each followed by "00".
bool glo30_tile_name(const std::string& stem, int* lat, int* lon) {
    if (stem.size() == 7 && srtm_tile_name(stem, lat, lon)) return true;
    std::vector<std::string> parts;
    size_t a = 0;
    while (true) {
        const size_t u = stem.find('_', a);
        parts.push_back(stem.substr(a, u == std::string::npos ? std::string::npos : u - a));
        if (u == std::string::npos) break;
        a = u + 1;
    }
    bool have_lat = false, have_lon = false;
    for (size_t i = 0; i < parts.size(); ++i) {
        const std::string& p = parts[i];
        const bool next00 = i + 1 < parts.size() && parts[i + 1] == "00";
        int v;
        if (!have_lat) {
            if (p.size() >= 3 && (p[0] == 'N' || p[0] == 'S') && parse_i32_field(p.substr(1), &v) && next00) {
                *lat = p[0] == 'N' ? v : -v;
                have_lat = true;
            }
        } else if (!have_lon) {
            if (p.size() >= 3 && (p[0] == 'E' || p[0] == 'W') && parse_i32_field(p.substr(1), &v) && next00) {
                *lon = p[0] == 'E' ? v : -v;
                have_lon = true;
            }
        }
        if (have_lat && have_lon) break;
    }
    return have_lat && have_lon;
}

// Rust's saturating float → integer cast (NaN → 0).
template <typename T>
T sat_cast(float x) {
    if (std::isnan(x)) return 0;
    const double lo = (double)std::numeric_limits<T>::min(), hi = (double)std::numeric_limits<T>::max();
    const double t = std::trunc((double)x);
    return (T)(t < lo ? lo : t > hi ? hi : t);
}

}  // namespace

extern "C" {

// dst null: the header and georeferencing only, unless full_check, which
// also decodes the blocks on the host (every error Glo30Tile::load can
// report) and uploads nothing.
static sardine_status sardine_dem_glo30_tile_impl(const char* path, float* dst, sardine_dem_raster* out, void* stream,
                                                  bool full_check = false) {
    if (!path || !out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "path or out is null");
    sd::GeoTiffRaster g;
    std::string why;
    // one read of the file; decode errors (rc 3) are reported after the
    // georeferencing, as the reference decodes the image after it
    const int rc = sd::read_geotiff(path, &g, &why, /*decode=*/dst != nullptr || full_check,
                                    /*defer_fp_predictor=*/true);
    if (rc == 1) return fail(SARDINE_ERR_DEM_IO, "IO error reading %s: %s", path, why.c_str());
    if (rc == 2) return fail(SARDINE_ERR_DEM_TIFF_DECODE, "TIFF decode error reading %s: %s", path, why.c_str());
    double lat0, lon0, px;
    if (g.pixel_scale.size() >= 2 && g.tiepoint.size() >= 6) {
        const double sx = g.pixel_scale[0], sy = g.pixel_scale[1];
        lon0 = g.tiepoint[3] + (0.5 - g.tiepoint[0]) * sx;
        lat0 = g.tiepoint[4] - (0.5 - g.tiepoint[1]) * sy;
        px = (sx + sy) / 2.0;
    } else {
        int la, lo;
        const std::string stem = file_stem(path);
        if (!glo30_tile_name(stem, &la, &lo))
            return fail(SARDINE_ERR_DEM_UNRECOGNISED_TILE_NAME,
                        "cannot determine tile origin for '%s': filename does not match SRTM (N50E010) or GLO-30 "
                        "(Copernicus_DSM_COG_10_N50_00_E010_00_DEM) naming conventions", stem.c_str());
        px = 1.0 / (double)g.width;
        lat0 = (double)(la + 1) - 0.5 * px;
        lon0 = (double)lo + 0.5 * px;
    }
    sardine_dem_raster r;
    r.tiling = nullptr;
    r.data = dst;
    r.width = g.width;
    r.height = g.height;
    r.origin_lat_deg = lat0;
    r.origin_lon_deg = lon0;
    r.pixel_deg = px;
    if (!dst && !full_check) {
        *out = r;
        return SARDINE_OK;
    }
    if (rc == 3) return fail(SARDINE_ERR_DEM_TIFF_DECODE, "TIFF decode error reading %s: %s", path, why.c_str());
    if (g.sample == sd::TiffSample::Other)
        return fail(SARDINE_ERR_DEM_UNSUPPORTED_PIXEL_TYPE,
                    "GLO-30 tile %s has unsupported pixel type (expected f32 or i16)", path);
    if (!dst) {
        *out = r;
        return SARDINE_OK;
    }
    // GDAL_NODATA parsed as Rust parses an f32 (after trimming)
    bool have_nd = false;
    float nd = 0.0f;
    if (g.has_nodata) {
        size_t a = 0, b = g.nodata.size();
        while (a < b && std::isspace((unsigned char)g.nodata[a])) ++a;
        while (b > a && std::isspace((unsigned char)g.nodata[b - 1])) --b;
        const std::string t = g.nodata.substr(a, b - a);
        if (is_rust_float(t)) {
            have_nd = true;
            nd = std::strtof(t.c_str(), nullptr);
        }
    }
    const int64_t n = g.width * g.height;
    const int kind = g.sample == sd::TiffSample::F32 ? 0 : g.sample == sd::TiffSample::I16 ? 1 : 2;
    // the nodata value cast to the sample type as Rust's `as` casts (saturating, NaN → 0)
    const int32_t nd_i = kind == 1 ? (int32_t)sat_cast<int16_t>(nd) : (int32_t)sat_cast<uint16_t>(nd);
    sardine_status st;
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream;
    // the decoded samples go to HBM as they are; the sample rule runs there
    DevBuf raw;
    SD_HIP(raw.alloc(g.pixels.size(), s));
    SD_HIP(staging().copy(raw.p, g.pixels.data(), g.pixels.size(), s));
    KernelTimer kt(SARDINE_KERNEL_DEM, s);
    if (g.fp_deferred)  // the floating-point predictor and the sample rule in one pass
        SD_HIP(sd_launch_fp_predictor(raw.as<uint8_t>(), g.block_w, g.block_h, g.blocks_x, g.blocks_y, g.width,
                                      g.height, have_nd ? 1 : 0, nd, dst, s));
    else
        SD_HIP(sd_launch_dem_values(raw.p, kind, have_nd ? 1 : 0, nd, nd_i, dst, n, s));
    kt.stop();
    *out = r;
    return SARDINE_OK;
}
sardine_status sardine_dem_glo30_tile(const char* path, float* dst, sardine_dem_raster* out, void* stream) {
    return guarded(SARDINE_ERR_DEM_TIFF_DECODE, [&] { return sardine_dem_glo30_tile_impl(path, dst, out, stream); });
}

// DemMosaic::load_glo30_files (dem.rs:642-653): every tile loaded as one
// GLO-30 tile (errors in path order), then placed into one raster on a grid
// of tile rows (origin latitude, north first) and columns (origin
// longitude, west first); the lookup keeps each tile's own origin and edge
// rule and the load order (TcParams::dem_tab, tc_pixel.hpp).
static sardine_status sardine_dem_glo30_mosaic_impl(const char* const* paths, int32_t n, float* dst,
                                                    sardine_dem_raster* out, sardine_dem_tiling* tiling,
                                                    void* stream) {
    if (!out || !tiling || (n > 0 && !paths)) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    if (n <= 0) return fail(SARDINE_ERR_DEM_NO_TILES, "no DEM tiles");
    const int M = SARDINE_DEM_MAX_TILES_AXIS;
    std::vector<sardine_dem_raster> hd((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        const sardine_status st = sardine_dem_glo30_tile_impl(paths[i], nullptr, &hd[(size_t)i], nullptr);
        if (st) {
            // the reference loads tiles whole, in path order: an earlier
            // tile's decode or pixel-type error comes first
            const std::string msg = g_last_error;
            const sardine_error_detail det = g_last_detail;
            for (int32_t j = 0; j < i; ++j) {
                sardine_dem_raster hj;
                const sardine_status sj = sardine_dem_glo30_tile_impl(paths[j], nullptr, &hj, nullptr, true);
                if (sj) return sj;
            }
            g_last_error = msg;
            g_last_detail = det;
            return st;
        }
    }
    const sardine_dem_raster& h0 = hd[0];
    std::vector<double> lats, lons;
    for (int32_t i = 0; i < n; ++i) {
        const sardine_dem_raster& h = hd[(size_t)i];
        if (h.width != h0.width || h.height != h0.height || h.pixel_deg != h0.pixel_deg)
            return fail(SARDINE_ERR_INVALID_ARGUMENT,
                        "GLO-30 mosaic: tile %d (%s) is %lld x %lld at %.17g deg, tile 0 %lld x %lld at %.17g deg; "
                        "the tiles must share size and pixel size", i, paths[i], (long long)h.height,
                        (long long)h.width, h.pixel_deg, (long long)h0.height, (long long)h0.width, h0.pixel_deg);
        if (std::find(lats.begin(), lats.end(), h.origin_lat_deg) == lats.end()) lats.push_back(h.origin_lat_deg);
        if (std::find(lons.begin(), lons.end(), h.origin_lon_deg) == lons.end()) lons.push_back(h.origin_lon_deg);
    }
    if ((int)lats.size() > M || (int)lons.size() > M)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "GLO-30 mosaic: %zu x %zu tile rows x columns (at most %d x %d)",
                    lats.size(), lons.size(), M, M);
    std::sort(lats.begin(), lats.end(), [](double a, double b) { return a > b; });
    std::sort(lons.begin(), lons.end());
    // adjacent rows and columns abut: origins one tile apart (to 1e-6 pixel)
    const double px = h0.pixel_deg, tol = 1e-6 * px;
    for (size_t i = 1; i < lats.size(); ++i)
        if (std::fabs((lats[i - 1] - lats[i]) - (double)h0.height * px) > tol)
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "GLO-30 mosaic: tile rows at %.12g and %.12g do not abut",
                        lats[i - 1], lats[i]);
    for (size_t i = 1; i < lons.size(); ++i)
        if (std::fabs((lons[i] - lons[i - 1]) - (double)h0.width * px) > tol)
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "GLO-30 mosaic: tile columns at %.12g and %.12g do not abut",
                        lons[i - 1], lons[i]);
    sardine_dem_tiling t;
    std::memset(&t, 0, sizeof t);
    t.rows = (int32_t)lats.size();
    t.cols = (int32_t)lons.size();
    t.tile_height = (int32_t)h0.height;
    t.tile_width = (int32_t)h0.width;
    for (int i = 0; i < M * M; ++i) t.rank[i] = -1;
    for (size_t i = 0; i < lats.size(); ++i) t.origin_lat_deg[i] = lats[i];
    for (size_t i = 0; i < lons.size(); ++i) t.origin_lon_deg[i] = lons[i];
    std::vector<int> cell((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        const int r = (int)(std::find(lats.begin(), lats.end(), hd[(size_t)i].origin_lat_deg) - lats.begin());
        const int c = (int)(std::find(lons.begin(), lons.end(), hd[(size_t)i].origin_lon_deg) - lons.begin());
        if (t.rank[r * M + c] >= 0)
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "GLO-30 mosaic: tiles %d and %d (%s) share a grid cell",
                        t.rank[r * M + c], i, paths[i]);
        t.rank[r * M + c] = i;
        cell[(size_t)i] = r * M + c;
    }
    const int64_t W = (int64_t)t.cols * h0.width, H = (int64_t)t.rows * h0.height;
    if (W >= INT32_MAX || H >= INT32_MAX) return fail(SARDINE_ERR_CAPACITY, "GLO-30 mosaic raster too large");
    sardine_dem_raster r = h0;
    r.data = dst;
    r.width = W;
    r.height = H;
    r.origin_lat_deg = lats[0];
    r.origin_lon_deg = lons[0];
    r.tiling = tiling;
    if (!dst) {
        *tiling = t;
        *out = r;
        return SARDINE_OK;
    }
    sardine_status st;
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream;
    // cells without a tile stay NaN (no tile answers there)
    SD_HIP(hipMemsetAsync(dst, 0xff, (size_t)W * (size_t)H * sizeof(float), s));
    DevBuf tmp;
    SD_HIP(tmp.alloc((size_t)h0.width * (size_t)h0.height * sizeof(float), s));
    for (int32_t i = 0; i < n; ++i) {
        sardine_dem_raster ri;
        if ((st = sardine_dem_glo30_tile_impl(paths[i], tmp.as<float>(), &ri, stream))) return st;
        const int rr = cell[(size_t)i] / M, cc = cell[(size_t)i] % M;
        SD_HIP(hipMemcpy2DAsync(dst + (size_t)rr * h0.height * W + (size_t)cc * h0.width, (size_t)W * sizeof(float),
                                tmp.p, (size_t)h0.width * sizeof(float), (size_t)h0.width * sizeof(float),
                                (size_t)h0.height, hipMemcpyDeviceToDevice, s));
    }
    *tiling = t;
    *out = r;
    return SARDINE_OK;
}
sardine_status sardine_dem_glo30_mosaic(const char* const* paths, int32_t n_paths, float* dst, sardine_dem_raster* out,
                                        sardine_dem_tiling* tiling, void* stream) {
    return guarded(SARDINE_ERR_DEM_TIFF_DECODE,
                   [&] { return sardine_dem_glo30_mosaic_impl(paths, n_paths, dst, out, tiling, stream); });
}

}  // extern "C"

// ── pre-TC multilook (scene_prep.rs:1060-1161) ──────────────────────────
extern "C" {

sardine_status sardine_apply_multilook(const float* data, const float* nesz, const sardine_merged_info* in,
                                       int32_t range_looks, int32_t azimuth_looks, double azimuth_time_interval_s,
                                       float* out_data, float* out_nesz, sardine_merged_info* out, void* stream) {
    if (!in || !out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "info is null");
    if (in->lines < 0 || in->samples < 0) return fail(SARDINE_ERR_INVALID_ARGUMENT, "negative image size");
    const sardine_merged_info src = *in;  // `out` may alias `in`
    sardine_merged_info o = src;
    const bool pass = range_looks == 1 && azimuth_looks == 1;
    if (!pass) {
        if (range_looks <= 0 || azimuth_looks <= 0)
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "multilook factors must be >= 1 (range=%d, azimuth=%d)",
                        range_looks, azimuth_looks);
        o.lines = src.lines / azimuth_looks;
        o.samples = src.samples / range_looks;
        if (o.lines == 0 || o.samples == 0)
            return fail(SARDINE_ERR_INVALID_ARGUMENT,
                        "multilook (%dx%d) is larger than image dimensions (%lld samples x %lld lines)", range_looks,
                        azimuth_looks, (long long)src.samples, (long long)src.lines);
        // centre-aligned geometry (scene_prep.rs:1094-1117)
        o.range_pixel_spacing_s = src.range_pixel_spacing_s * (double)range_looks;
        o.range_pixel_spacing_m = src.range_pixel_spacing_m * (double)range_looks;
        o.near_slant_range_time_s =
            src.near_slant_range_time_s + (double)(range_looks - 1) / 2.0 * src.range_pixel_spacing_s;
        if (azimuth_looks > 1) {
            if (!std::isfinite(azimuth_time_interval_s) || azimuth_time_interval_s <= 0.0)
                return fail(SARDINE_ERR_INVALID_ARGUMENT,
                            "azimuth_time_interval_s must be > 0 when azimuth_looks > 1 (got %g)",
                            azimuth_time_interval_s);
            o.azimuth_start_time_us =
                src.azimuth_start_time_us +
                (int64_t)((double)(azimuth_looks - 1) / 2.0 * azimuth_time_interval_s * 1e6);
        }
    }
    if (!out_data && !out_nesz) {  // shape query: no device work
        *out = o;
        return SARDINE_OK;
    }
    const int64_t n_in = src.lines * src.samples;
    if (n_in > 0 && (!data || !nesz || !out_data || !out_nesz))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "null raster");
    sardine_status st;
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream;
    KernelTimer kt(SARDINE_KERNEL_MULTILOOK, s);
    if (pass) {
        // range_looks = azimuth_looks = 1 returns the input unchanged
        if (n_in > 0 && out_data != data)
            SD_HIP(hipMemcpyAsync(out_data, data, (size_t)n_in * sizeof(float), hipMemcpyDeviceToDevice, s));
        if (n_in > 0 && out_nesz != nesz)
            SD_HIP(hipMemcpyAsync(out_nesz, nesz, (size_t)n_in * sizeof(float), hipMemcpyDeviceToDevice, s));
    } else {
        SD_HIP(sd_launch_multilook(data, nesz, src.samples, o.lines, o.samples, range_looks, azimuth_looks, out_data,
                                   out_nesz, s));
    }
    kt.stop();
    *out = o;
    return SARDINE_OK;
}

}  // extern "C"

// ── ground-range (GRD) projection (ground_range.rs:82-490) ──────────────
namespace {

// One swath's incidence profile along slant range: the grid's mid-azimuth
// row, sorted by slant time (RangeInterpolator, ground_range.rs:157-236).
struct IncidenceRow {
    std::vector<double> t, deg;
    double lo = 0.0, hi = 0.0;
    double centre() const { return (lo + hi) / 2.0; }
    // Linear in slant time inside [lo, hi]; an exact node returns its value.
    bool eval(double x, double* out) const {
        if (x < lo || x > hi) return false;
        const size_t k = (size_t)(std::lower_bound(t.begin(), t.end(), x) - t.begin());
        if (k < t.size() && t[k] == x) *out = deg[k];
        else if (k == 0) *out = deg[0];
        else if (k >= t.size()) *out = deg.back();
        else *out = deg[k - 1] + (x - t[k - 1]) / (t[k] - t[k - 1]) * (deg[k] - deg[k - 1]);
        return true;
    }
};

IncidenceRow incidence_row(const sardine_grd_grid_point* p, int32_t n) {
    std::vector<uint32_t> lines(n);
    for (int32_t i = 0; i < n; ++i) lines[i] = p[i].line;
    std::sort(lines.begin(), lines.end());
    lines.erase(std::unique(lines.begin(), lines.end()), lines.end());
    const uint32_t mid = lines[lines.size() / 2];
    std::vector<std::pair<double, double>> row;
    for (int32_t i = 0; i < n; ++i)
        if (p[i].line == mid) row.emplace_back(p[i].slant_range_time_s, p[i].incidence_angle_deg);
    std::stable_sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    IncidenceRow r;
    for (const auto& q : row) {
        r.t.push_back(q.first);
        r.deg.push_back(q.second);
    }
    r.lo = r.t.front();
    r.hi = r.t.back();
    return r;
}

// interpolate_at_slant_time (ground_range.rs:238-290): the covering swath
// whose centre is nearest (the first on a tie), else the nearest swath's
// edge value when within 1 % of its width.
bool incidence_at(const std::vector<IncidenceRow>& rows, double x, double* deg) {
    const IncidenceRow* best = nullptr;
    double best_d = 0.0, best_v = 0.0;
    for (const IncidenceRow& r : rows) {
        double v;
        if (!r.eval(x, &v)) continue;
        const double d = std::fabs(x - r.centre());
        if (!best || d < best_d) {
            best = &r;
            best_d = d;
            best_v = v;
        }
    }
    if (best) {
        *deg = best_v;
        return true;
    }
    const IncidenceRow* near = nullptr;
    double near_d = 0.0;
    for (const IncidenceRow& r : rows) {
        const double d = x < r.lo ? r.lo - x : x - r.hi;
        if (!near || d < near_d) {
            near = &r;
            near_d = d;
        }
    }
    if (!near || !(near_d < (near->hi - near->lo) * 0.01)) return false;
    return near->eval(std::min(std::max(x, near->lo), near->hi), deg);
}

struct GrdColumnHost {  // layout of sd::GrdColumn (multilook.hip)
    int32_t left;
    float frac;
};

// Everything of to_ground_range except the per-line work: validation, the
// per-column incidence, the cumulative ground range, the output shape and
// (when map != null) the column map.
sardine_status grd_plan(const sardine_merged_info* m, const sardine_grd_grid_point* points, const int32_t* counts,
                        int32_t n_swaths, double target, double az_spacing, sardine_grd_image* img,
                        std::vector<GrdColumnHost>* map) {
    if (m->lines == 0 || m->samples == 0)
        return fail(SARDINE_ERR_GRD_EMPTY_IMAGE, "merged image is empty (%lld lines x %lld samples)",
                    (long long)m->lines, (long long)m->samples);
    if (m->lines < 0 || m->samples < 0) return fail(SARDINE_ERR_INVALID_ARGUMENT, "negative image size");
    if (target <= 0.0) return fail(SARDINE_ERR_GRD_INVALID_SPACING, "target pixel spacing must be positive, got %g", target);
    if (!std::isfinite(az_spacing) || az_spacing <= 0.0)
        return fail(SARDINE_ERR_GRD_INVALID_AZIMUTH_SPACING,
                    "input azimuth pixel spacing must be positive and finite, got %g", az_spacing);
    if (n_swaths < 0 || (n_swaths > 0 && (!points || !counts))) return fail(SARDINE_ERR_INVALID_ARGUMENT, "bad grids");
    std::vector<IncidenceRow> rows;
    int64_t off = 0;
    for (int32_t k = 0; k < n_swaths; ++k) {
        if (counts[k] <= 0) return fail(SARDINE_ERR_GRD_NO_GRID_POINTS, "no geolocation grid points for swath %d", k);
        rows.push_back(incidence_row(points + off, counts[k]));
        off += counts[k];
    }
    std::stable_sort(rows.begin(), rows.end(), [](const IncidenceRow& a, const IncidenceRow& b) { return a.lo < b.lo; });
    const int64_t S = m->samples;
    if (S > INT32_MAX) return fail(SARDINE_ERR_INVALID_ARGUMENT, "merged image wider than 2^31-1 samples");
    // ground spacing per column: slant spacing / sin(incidence), incidence
    // in radians as deg * (π / 180) (f64::to_radians)
    std::vector<double> cum((size_t)S);
    double prev_gs = 0.0, acc = 0.0;
    for (int64_t c = 0; c < S; ++c) {
        double deg;
        const double t = m->near_slant_range_time_s + (double)c * m->range_pixel_spacing_s;
        if (!incidence_at(rows, t, &deg))
            return fail(SARDINE_ERR_GRD_INTERPOLATION_FAILED,
                        "incidence angle interpolation failed at column %lld: slant time %.6e s outside all grid "
                        "ranges", (long long)c, t);
        const double gs = m->range_pixel_spacing_m / std::sin(deg * (3.14159265358979323846 / 180.0));
        if (c > 0) acc += (prev_gs + gs) / 2.0;
        cum[(size_t)c] = acc;
        prev_gs = gs;
    }
    const double q = std::floor(cum[(size_t)S - 1] / target);
    if (!(q >= 1.0)) return fail(SARDINE_ERR_GRD_INVALID_SPACING, "target pixel spacing must be positive, got %g", target);
    if (q > (double)INT32_MAX)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "%g ground-range columns exceed 2^31-1", q);
    const int64_t os = (int64_t)q;
    const int64_t A = (int64_t)std::max(std::round(target / az_spacing), 1.0);
    if (A > INT32_MAX) return fail(SARDINE_ERR_INVALID_ARGUMENT, "azimuth looks exceed 2^31-1");
    img->lines = m->lines / A;
    img->samples = os;
    img->range_pixel_spacing_m = target;
    img->azimuth_pixel_spacing_m = az_spacing * (double)A;
    img->range_looks = 1;
    img->azimuth_looks = (int32_t)A;
    if (!map) return SARDINE_OK;
    // build_ground_to_slant_mapping (ground_range.rs:446-490): a forward
    // search from the last bracket (the table is increasing)
    map->resize((size_t)os);
    int64_t from = 0;
    for (int64_t j = 0; j < os; ++j) {
        const double g = (double)j * target;
        const int64_t k = (int64_t)(std::lower_bound(cum.begin() + from, cum.end(), g) - cum.begin());
        GrdColumnHost& e = (*map)[(size_t)j];
        if (k < S && cum[(size_t)k] == g) {
            e = {(int32_t)k, -1.0f};
            from = k;
        } else if (k == 0) {
            e = {0, -1.0f};
        } else if (k >= S) {
            e = {(int32_t)(S - 1), -1.0f};
        } else {
            const double f = (g - cum[(size_t)k - 1]) / (cum[(size_t)k] - cum[(size_t)k - 1]);
            e = {(int32_t)(k - 1), (float)f};  // the kernel weighs with frac as f32
            from = k - 1;
        }
    }
    return SARDINE_OK;
}

struct GrdPlanned {
    sardine_grd_image img;
    std::vector<GrdColumnHost> map;
};

// The plan depends only on the image geometry, the grids and the two
// spacings, and a scene is usually projected more than once (bands,
// re-runs, services), so plans are kept per input bytes (8, least recently
// used first out).  Failing inputs are not kept.
class GrdPlanCache {
   public:
    sardine_status get(const sardine_merged_info* m, const sardine_grd_grid_point* points, const int32_t* counts,
                       int32_t n_swaths, double target, double az, std::shared_ptr<const GrdPlanned>* out) {
        std::vector<char> key;
        auto add = [&](const void* p, size_t n) {
            key.insert(key.end(), static_cast<const char*>(p), static_cast<const char*>(p) + n);
        };
        add(&m->lines, sizeof m->lines);
        add(&m->samples, sizeof m->samples);
        add(&m->near_slant_range_time_s, sizeof(double));
        add(&m->range_pixel_spacing_s, sizeof(double));
        add(&m->range_pixel_spacing_m, sizeof(double));
        add(&target, sizeof target);
        add(&az, sizeof az);
        add(&n_swaths, sizeof n_swaths);
        int64_t n_pts = 0;
        if (n_swaths > 0 && counts) {
            add(counts, sizeof(int32_t) * (size_t)n_swaths);
            for (int32_t k = 0; k < n_swaths; ++k) n_pts += std::max(counts[k], 0);
        }
        if (n_pts > 0 && points) add(points, sizeof(sardine_grd_grid_point) * (size_t)n_pts);
        {
            std::lock_guard<std::mutex> g(mu_);
            ++tick_;
            for (Entry& e : entries_)
                if (e.key == key) {
                    e.used = tick_;
                    *out = e.plan;
                    return SARDINE_OK;
                }
        }
        auto plan = std::make_shared<GrdPlanned>();
        sardine_status st = grd_plan(m, points, counts, n_swaths, target, az, &plan->img, &plan->map);
        if (st) return st;
        std::lock_guard<std::mutex> g(mu_);
        Entry ne{std::move(key), plan, ++tick_};
        if (entries_.size() < kEntries) {
            entries_.push_back(std::move(ne));
        } else {
            auto lru = std::min_element(entries_.begin(), entries_.end(),
                                        [](const Entry& a, const Entry& b) { return a.used < b.used; });
            *lru = std::move(ne);
        }
        *out = plan;
        return SARDINE_OK;
    }

   private:
    struct Entry {
        std::vector<char> key;
        std::shared_ptr<const GrdPlanned> plan;
        uint64_t used = 0;
    };
    static constexpr size_t kEntries = 8;
    std::mutex mu_;
    std::vector<Entry> entries_;
    uint64_t tick_ = 0;
};

GrdPlanCache& grd_plan_cache() {
    static GrdPlanCache* c = new GrdPlanCache();
    return *c;
}

}  // namespace

extern "C" {

static sardine_status sardine_to_ground_range_impl(const float* data, const sardine_merged_info* merged,
                                       const sardine_grd_grid_point* points, const int32_t* counts,
                                       int32_t n_swaths, double target_spacing_m, double azimuth_pixel_spacing_m,
                                       float* out, sardine_grd_image* image, void* stream) {
    if (!merged || !image) return fail(SARDINE_ERR_INVALID_ARGUMENT, "merged info or image is null");
    sardine_status st;
    if (!out) {
        sardine_grd_image img;
        if ((st = grd_plan(merged, points, counts, n_swaths, target_spacing_m, azimuth_pixel_spacing_m, &img,
                           nullptr)))
            return st;
        *image = img;
        return SARDINE_OK;
    }
    std::shared_ptr<const GrdPlanned> plan;
    if ((st = grd_plan_cache().get(merged, points, counts, n_swaths, target_spacing_m, azimuth_pixel_spacing_m,
                                   &plan)))
        return st;
    if (!data) return fail(SARDINE_ERR_INVALID_ARGUMENT, "data is null");
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream;
    const sardine_grd_image& img = plan->img;
    DevBuf dm;
    SD_HIP(upload(dm, plan->map.data(), plan->map.size(), s));
    KernelTimer kt(SARDINE_KERNEL_GRD, s);
    SD_HIP(sd_launch_grd(data, merged->samples, dm.p, img.lines, img.samples, img.azimuth_looks, out, s));
    kt.stop();
    *image = img;
    return SARDINE_OK;
}
sardine_status sardine_to_ground_range(const float* data, const sardine_merged_info* merged,
                                       const sardine_grd_grid_point* points, const int32_t* counts,
                                       int32_t n_swaths, double target_spacing_m, double azimuth_pixel_spacing_m,
                                       float* out, sardine_grd_image* image, void* stream) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] {
        return sardine_to_ground_range_impl(data, merged, points, counts, n_swaths, target_spacing_m, azimuth_pixel_spacing_m, out, image, stream);
    });
}

}  // extern "C"

// ── Cloud-Optimised GeoTIFF export (export.rs:329-760) ───────────────────
namespace {

bool cog_tile_size_ok(uint32_t ts) { return ts >= 64 && ts <= 4096 && (ts & (ts - 1)) == 0; }

// export.rs:338-342 (InvalidGeotransform)
sardine_status cog_check_geotransform(const double* gt) {
    if (!gt) return fail(SARDINE_ERR_INVALID_ARGUMENT, "geotransform is null");
    const double pw = gt[1], ph = -gt[5];
    if (!std::isfinite(pw) || pw <= 0.0 || !std::isfinite(ph) || ph <= 0.0)
        return fail(SARDINE_ERR_EXPORT_INVALID_GEOTRANSFORM,
                    "pixel width/height must be finite and positive: geotransform [%g, %g, %g, %g, %g, %g]", gt[0],
                    gt[1], gt[2], gt[3], gt[4], gt[5]);
    return SARDINE_OK;
}

// OutputCrs::geo_key_directory (output_crs.rs:265-287) for the CRSs the path
// produces; epsg() (output_crs.rs:167-175).
sardine_status cog_geo_keys(int32_t crs, int32_t zone, std::vector<uint16_t>* keys) {
    if (crs == SARDINE_CRS_WGS84) {
        *keys = {1, 1, 0, 3, 1024, 0, 1, 2, 1025, 0, 1, 1, 2048, 0, 1, 4326};
        return SARDINE_OK;
    }
    if (crs != SARDINE_CRS_UTM_NORTH && crs != SARDINE_CRS_UTM_SOUTH)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "unknown CRS %d", crs);
    if (zone < 1 || zone > 60) return fail(SARDINE_ERR_INVALID_ARGUMENT, "UTM zone %d outside 1..60", zone);
    const uint16_t epsg = (uint16_t)((crs == SARDINE_CRS_UTM_NORTH ? 32600 : 32700) + zone);
    *keys = {1, 1, 0, 4, 1024, 0, 1, 1, 1025, 0, 1, 1, 3072, 0, 1, epsg, 3076, 0, 1, 9001};
    return SARDINE_OK;
}

// write_cog_raw_inner's pyramid, format choice and offset pass
// (export.rs:430-535).
sardine_status cog_layout(int64_t cols, int64_t rows, uint32_t ts, int32_t geo_shorts, sardine_cog_layout* L) {
    if (!cog_tile_size_ok(ts))
        return fail(SARDINE_ERR_EXPORT_INVALID_TILE_SIZE,
                    "COG tile size must be a power of two >= 64 and <= 4096; got %u", ts);
    if (cols < 0 || rows < 0) return fail(SARDINE_ERR_INVALID_ARGUMENT, "negative raster size");
    std::memset(L, 0, sizeof *L);
    // overview sizes, finest first, then reversed (coarsest first)
    std::vector<std::pair<int64_t, int64_t>> lv;
    int64_t c = (cols + 1) / 2, r = (rows + 1) / 2;
    while (c > (int64_t)ts || r > (int64_t)ts) {
        lv.push_back({c, r});
        c = (c + 1) / 2;
        r = (r + 1) / 2;
    }
    lv.push_back({c, r});
    std::reverse(lv.begin(), lv.end());
    lv.push_back({cols, rows});
    if ((int)lv.size() > SARDINE_COG_MAX_LEVELS)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "%zu pyramid levels exceed %d", lv.size(), SARDINE_COG_MAX_LEVELS);
    // rows*cols*4 > u32::MAX with the reference's saturating products
    const uint64_t uc = (uint64_t)cols, ur = (uint64_t)rows;
    const bool big = uc != 0 && ur != 0 && (ur > UINT64_MAX / uc || ur * uc > 0xffffffffull / 4);
    const uint64_t n_entries = 17, geo_bytes = (uint64_t)geo_shorts * 2;
    const uint64_t ifd_fixed = big ? 8 + n_entries * 20 + 8 : 2 + n_entries * 12 + 4;
    const uint64_t entry = big ? 8 : 4;
    L->n_levels = (int32_t)lv.size();
    L->bigtiff = big ? 1 : 0;
    L->tile_size = ts;
    L->geo_key_shorts = geo_shorts;
    uint64_t off = big ? 16 : 8;
    for (size_t k = 0; k < lv.size(); ++k) {
        sardine_cog_level& l = L->level[k];
        l.cols = lv[k].first;
        l.rows = lv[k].second;
        l.tiles_x = (l.cols + ts - 1) / ts;
        l.tiles_y = (l.rows + ts - 1) / ts;
        const uint64_t n = (uint64_t)(l.tiles_x * l.tiles_y);
        l.ifd_offset = off;
        off += ifd_fixed;
        l.tile_offsets_offset = off;
        off += n * entry;
        l.tile_counts_offset = off;
        off += n * entry;
    }
    L->pixel_scale_offset = off;
    off += 24;
    L->tiepoint_offset = off;
    off += 48;
    L->geo_keys_offset = off;
    off += geo_bytes;
    L->header_bytes = off;
    const uint64_t tile_bytes = (uint64_t)ts * ts * 4;
    for (int k = 0; k < L->n_levels; ++k) {
        sardine_cog_level& l = L->level[k];
        l.tile_data_offset = off;
        off += (uint64_t)(l.tiles_x * l.tiles_y) * tile_bytes;
    }
    L->file_bytes = off;
    L->tile_bytes = off - L->header_bytes;
    return SARDINE_OK;
}

template <typename T>
void put_le(std::vector<uint8_t>& b, T v) {  // the host is little-endian (x86-64)
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
}

// Header, IFDs, tile arrays and GeoTIFF blocks (export.rs:537-670).  Classic
// TIFF offsets are the low 32 bits of the 64-bit ones, as the reference
// writes them.
void cog_header_bytes(const sardine_cog_layout& L, const double* gt, const std::vector<uint16_t>& keys,
                      std::vector<uint8_t>& b) {
    const bool big = L.bigtiff != 0;
    const uint64_t ts = L.tile_size, tile_bytes = ts * ts * 4;
    b.clear();
    b.reserve(L.header_bytes);
    b.push_back('I');
    b.push_back('I');
    if (big) {
        put_le<uint16_t>(b, 43);
        put_le<uint16_t>(b, 8);
        put_le<uint16_t>(b, 0);
        put_le<uint64_t>(b, L.level[0].ifd_offset);
    } else {
        put_le<uint16_t>(b, 42);
        put_le<uint32_t>(b, (uint32_t)L.level[0].ifd_offset);
    }
    auto entry = [&](uint16_t tag, uint16_t type, uint64_t count, uint64_t value) {
        put_le<uint16_t>(b, tag);
        put_le<uint16_t>(b, type);
        if (big) {
            put_le<uint64_t>(b, count);
            put_le<uint64_t>(b, value);
        } else {
            put_le<uint32_t>(b, (uint32_t)count);
            put_le<uint32_t>(b, (uint32_t)value);
        }
    };
    for (int k = 0; k < L.n_levels; ++k) {
        const sardine_cog_level& l = L.level[k];
        const uint64_t n = (uint64_t)(l.tiles_x * l.tiles_y);
        if (big) put_le<uint64_t>(b, 17);
        else put_le<uint16_t>(b, 17);
        entry(256, 4, 1, (uint64_t)l.cols);  // ImageWidth
        entry(257, 4, 1, (uint64_t)l.rows);  // ImageLength
        entry(258, 3, 1, 32);                // BitsPerSample
        entry(259, 3, 1, 1);                 // Compression: none
        entry(262, 3, 1, 1);                 // PhotometricInterpretation
        entry(277, 3, 1, 1);                 // SamplesPerPixel
        // X/YResolution: an inline RATIONAL 72/1 in BigTIFF, a LONG 72 in
        // classic TIFF (export.rs:638-646)
        entry(282, big ? 5 : 4, 1, big ? ((1ull << 32) | 72ull) : 72ull);
        entry(283, big ? 5 : 4, 1, big ? ((1ull << 32) | 72ull) : 72ull);
        entry(296, 3, 1, 1);                       // ResolutionUnit
        entry(322, 4, 1, ts);                      // TileWidth
        entry(323, 4, 1, ts);                      // TileLength
        entry(324, big ? 16 : 4, n, l.tile_offsets_offset);
        entry(325, big ? 16 : 4, n, l.tile_counts_offset);
        entry(339, 3, 1, 3);                       // SampleFormat: IEEE float
        entry(33550, 12, 3, L.pixel_scale_offset);  // ModelPixelScale
        entry(33922, 12, 6, L.tiepoint_offset);     // ModelTiepoint
        entry(34735, 3, keys.size(), L.geo_keys_offset);
        const uint64_t next = k + 1 < L.n_levels ? L.level[k + 1].ifd_offset : 0;
        if (big) put_le<uint64_t>(b, next);
        else put_le<uint32_t>(b, (uint32_t)next);
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t o = l.tile_data_offset + i * tile_bytes;
            if (big) put_le<uint64_t>(b, o);
            else put_le<uint32_t>(b, (uint32_t)o);
        }
        for (uint64_t i = 0; i < n; ++i) {
            if (big) put_le<uint64_t>(b, tile_bytes);
            else put_le<uint32_t>(b, (uint32_t)tile_bytes);
        }
    }
    const double blocks[9] = {gt[1], -gt[5], 0.0, 0.0, 0.0, 0.0, gt[0], gt[3], 0.0};
    for (double v : blocks) put_le<double>(b, v);
    for (uint16_t k : keys) put_le<uint16_t>(b, k);
}

// The layout a caller passes back, re-derived from its full size, tile size
// and key count so that a stale or edited struct cannot steer the kernels.
sardine_status cog_relayout(const sardine_cog_layout* in, sardine_cog_layout* L) {
    if (!in || in->n_levels < 2 || in->n_levels > SARDINE_COG_MAX_LEVELS)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "invalid COG layout");
    if (in->geo_key_shorts != 16 && in->geo_key_shorts != 20)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "invalid COG layout (geo key count %d)", in->geo_key_shorts);
    const sardine_cog_level& f = in->level[in->n_levels - 1];
    sardine_status st;
    if ((st = cog_layout(f.cols, f.rows, in->tile_size, in->geo_key_shorts, L))) return st;
    if (L->n_levels != in->n_levels || L->tile_bytes != in->tile_bytes || L->header_bytes != in->header_bytes)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "COG layout does not match its full raster size");
    return SARDINE_OK;
}

// The device pyramid + pack of every level into `tiles` (export.rs:672-712).
sardine_status cog_pack(const float* data, const sardine_cog_layout& L, float* tiles, hipStream_t s) {
    if (L.tile_bytes == 0) return SARDINE_OK;
    int lg = 0;
    while ((1u << lg) < L.tile_size) ++lg;
    auto dev_level = [&](int k) {
        const sardine_cog_level& l = L.level[k];
        sd::CogLevelDev d;
        d.tiles = tiles + (l.tile_data_offset - L.header_bytes) / sizeof(float);
        d.cols = l.cols;
        d.rows = l.rows;
        d.tiles_x = l.tiles_x;
        d.pad_cols = l.tiles_x * (int64_t)L.tile_size;
        d.pad_rows = l.tiles_y * (int64_t)L.tile_size;
        return d;
    };
    const int n = L.n_levels;
    KernelTimer kt(SARDINE_KERNEL_COG_PACK, s);
    const sd::CogLevelDev full = dev_level(n - 1);
    sd::CogLevelDev d = dev_level(n - 2);
    SD_HIP(sd_launch_cog_level(data, 0, &full, &d, &full, lg, s));
    for (int k = n - 3; k >= 0; --k) {
        const sd::CogLevelDev src = d, none{};
        d = dev_level(k);
        SD_HIP(sd_launch_cog_level(src.tiles, 1, &src, &d, &none, lg, s));
    }
    kt.stop();
    return SARDINE_OK;
}

struct PinnedPair {  // the two host buffers of the file writer's copy/write overlap
    void* p[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~PinnedPair() {
        for (int i = 0; i < 2; ++i) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (p[i]) (void)hipHostFree(p[i]);
        }
    }
};

struct FileCloser {
    std::FILE* f = nullptr;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

// Device bytes → file: two pinned chunks in turn, each written while the
// next one is copied.
sardine_status stream_to_file(std::FILE* f, const void* dev, uint64_t bytes, hipStream_t s, const char* path) {
    const uint64_t chunk = std::min<uint64_t>(bytes, 32ull << 20);
    if (chunk == 0) return SARDINE_OK;
    PinnedPair pp;
    for (int i = 0; i < 2; ++i) {
        SD_HIP(hipHostMalloc(&pp.p[i], chunk, hipHostMallocDefault));
        SD_HIP(hipEventCreateWithFlags(&pp.ev[i], hipEventDisableTiming));
    }
    const uint8_t* src = static_cast<const uint8_t*>(dev);
    const uint64_t n_chunks = (bytes + chunk - 1) / chunk;
    auto issue = [&](uint64_t i) -> hipError_t {
        const uint64_t off = i * chunk, n = std::min(chunk, bytes - off);
        hipError_t e = hipMemcpyAsync(pp.p[i % 2], src + off, n, hipMemcpyDeviceToHost, s);
        return e != hipSuccess ? e : hipEventRecord(pp.ev[i % 2], s);
    };
    SD_HIP(issue(0));
    for (uint64_t i = 0; i < n_chunks; ++i) {
        SD_HIP(hipEventSynchronize(pp.ev[i % 2]));
        if (i + 1 < n_chunks) SD_HIP(issue(i + 1));
        const uint64_t n = std::min(chunk, bytes - i * chunk);
        if (std::fwrite(pp.p[i % 2], 1, n, f) != n)
            return fail(SARDINE_ERR_EXPORT_IO, "writing %s: %s", path, std::strerror(errno));
    }
    return SARDINE_OK;
}

// needs_bigtiff (export.rs:762-786): the classic layout's size estimate
// (header, 16- or 13-entry IFD, resolution rationals, GeoTIFF blocks, an
// 8-byte strip table entry per 64 rows, the raster) above 2^32 - 1, with the
// reference's saturating arithmetic.
bool striped_needs_bigtiff(uint64_t rows, uint64_t cols, uint64_t bpp, uint64_t geo_shorts) {
    auto mul = [](uint64_t a, uint64_t b) { return (b != 0 && a > UINT64_MAX / b) ? UINT64_MAX : a * b; };
    auto add = [](uint64_t a, uint64_t b) { return a > UINT64_MAX - b ? UINT64_MAX : a + b; };
    const uint64_t fixed = geo_shorts > 0 ? 8 + 198 + 16 + 24 + 48 + geo_shorts * 2 : 8 + 162 + 16;
    const uint64_t total = add(add(fixed, mul((rows + 63) / 64, 8)), mul(mul(rows, cols), bpp));
    return total > 0xffffffffull;
}

// The bytes before the raster of a single-IFD striped TIFF, 64 rows per
// strip: write_geotiff_raw_inner (classic, export.rs:788-1010) or
// write_bigtiff_raw_inner (export.rs:1023-1219).  keys empty: no GeoTIFF
// tags (write_tiff_no_crs).  rows >= 1.  Classic offsets are 32-bit
// arithmetic, as the reference computes them.
void striped_header_bytes(uint64_t cols, uint64_t rows, uint32_t bpp, const double* gt,
                          const std::vector<uint16_t>& keys, bool big, std::vector<uint8_t>& b) {
    const bool geo = !keys.empty();
    const uint64_t n_entries = geo ? 16 : 13;
    const uint64_t bits = bpp * 8ull, fmt = bpp == 1 ? 1 : 3;  // SampleFormat: unsigned int / IEEE float
    b.clear();
    b.push_back('I');
    b.push_back('I');
    uint64_t strips = 0, full = 0, last = 0, arr_off = 0, cnt_off = 0, data_off = 0;
    uint64_t res_off = 0, scale_off = 0, tie_off = 0, key_off = 0;
    if (big) {
        strips = (rows + 63) / 64;
        full = 64 * cols * bpp;
        last = (rows - (strips - 1) * 64) * cols * bpp;
        const uint64_t ifd_end = 16 + 8 + n_entries * 20 + 8;
        uint64_t after = ifd_end;  // one strip: its offset and count sit inline in the IFD
        if (strips > 1) {
            arr_off = ifd_end;
            cnt_off = ifd_end + strips * 8;
            after = ifd_end + 2 * strips * 8;
        }
        if (geo) {
            scale_off = after;
            tie_off = after + 24;
            key_off = after + 72;
            data_off = key_off + keys.size() * 2;
        } else {
            data_off = after;
        }
        put_le<uint16_t>(b, 43);
        put_le<uint16_t>(b, 8);
        put_le<uint16_t>(b, 0);
        put_le<uint64_t>(b, 16);
        put_le<uint64_t>(b, n_entries);
    } else {
        const uint32_t r32 = (uint32_t)rows, c32 = (uint32_t)cols;
        const uint32_t st32 = (r32 + 63u) / 64u, row32 = c32 * bpp;
        strips = st32;
        full = 64u * row32;
        last = (uint32_t)((r32 - (st32 - 1u) * 64u) * row32);
        arr_off = 8u + (2u + (uint32_t)n_entries * 12u + 4u);
        cnt_off = (uint32_t)arr_off + st32 * 4u;
        res_off = (uint32_t)cnt_off + st32 * 4u;
        scale_off = (uint32_t)res_off + 16u;
        tie_off = scale_off + 24;
        key_off = tie_off + 48;
        data_off = geo ? (uint32_t)(key_off + keys.size() * 2) : (uint32_t)res_off + 16u;
        put_le<uint16_t>(b, 42);
        put_le<uint32_t>(b, 8);
        put_le<uint16_t>(b, (uint16_t)n_entries);
    }
    auto entry = [&](uint16_t tag, uint16_t type, uint64_t count, uint64_t value) {
        put_le<uint16_t>(b, tag);
        put_le<uint16_t>(b, type);
        if (big) {
            put_le<uint64_t>(b, count);
            put_le<uint64_t>(b, value);
        } else {
            put_le<uint32_t>(b, (uint32_t)count);
            put_le<uint32_t>(b, (uint32_t)value);
        }
    };
    const uint64_t rat = (1ull << 32) | 72ull;  // RATIONAL 72/1 inline (BigTIFF)
    entry(256, 4, 1, cols);
    entry(257, 4, 1, rows);
    entry(258, 3, 1, bits);
    entry(259, 3, 1, 1);
    entry(262, 3, 1, 1);
    entry(273, big ? 16 : 4, strips, big ? (strips > 1 ? arr_off : data_off) : arr_off);  // StripOffsets
    entry(277, 3, 1, 1);
    entry(278, 4, 1, 64);                                                            // RowsPerStrip
    entry(279, big ? 16 : 4, strips, big ? (strips > 1 ? cnt_off : last) : cnt_off);  // StripByteCounts
    entry(282, 5, 1, big ? rat : res_off);
    entry(283, 5, 1, big ? rat : res_off + 8);
    entry(296, 3, 1, 1);
    entry(339, 3, 1, fmt);
    if (geo) {
        entry(33550, 12, 3, scale_off);
        entry(33922, 12, 6, tie_off);
        entry(34735, 3, keys.size(), key_off);
    }
    if (big) put_le<uint64_t>(b, 0);
    else put_le<uint32_t>(b, 0);
    if (big) {
        if (strips > 1) {
            for (uint64_t i = 0; i < strips; ++i) put_le<uint64_t>(b, data_off + i * full);
            for (uint64_t i = 0; i < strips; ++i) put_le<uint64_t>(b, i + 1 == strips ? last : full);
        }
    } else {
        for (uint64_t i = 0; i < strips; ++i) put_le<uint32_t>(b, (uint32_t)(data_off + i * (uint64_t)(uint32_t)full));
        for (uint64_t i = 0; i < strips; ++i) put_le<uint32_t>(b, (uint32_t)(i + 1 == strips ? last : full));
        for (int i = 0; i < 2; ++i) {
            put_le<uint32_t>(b, 72);
            put_le<uint32_t>(b, 1);
        }
    }
    if (geo) {
        const double blocks[9] = {gt[1], -gt[5], 0.0, 0.0, 0.0, 0.0, gt[0], gt[3], 0.0};
        for (double v : blocks) put_le<double>(b, v);
        for (uint16_t k : keys) put_le<uint16_t>(b, k);
    }
}

// Validation and header of the striped writers: gt null = no CRS.
sardine_status striped_plan(int64_t cols, int64_t rows, int32_t bpp, const double* gt, int32_t crs, int32_t zone,
                            std::vector<uint8_t>* head) {
    sardine_status st;
    std::vector<uint16_t> keys;
    if (gt) {
        if ((st = cog_check_geotransform(gt))) return st;
        if ((st = cog_geo_keys(crs, zone, &keys))) return st;
    }
    if (bpp != 1 && bpp != 4) return fail(SARDINE_ERR_INVALID_ARGUMENT, "bytes per pixel must be 1 or 4");
    // rows == 0 underflows the reference's strip arithmetic (export.rs:861, 1061)
    if (rows <= 0 || cols < 0) return fail(SARDINE_ERR_INVALID_ARGUMENT, "raster needs at least one row");
    const bool big = striped_needs_bigtiff((uint64_t)rows, (uint64_t)cols, (uint64_t)bpp, keys.size());
    striped_header_bytes((uint64_t)cols, (uint64_t)rows, (uint32_t)bpp, gt, keys, big, *head);
    return SARDINE_OK;
}

sardine_status striped_write(const char* path, const void* data, int64_t cols, int64_t rows, int32_t bpp,
                             const double* gt, int32_t crs, int32_t zone, void* stream_v) {
    sardine_status st;
    std::vector<uint8_t> head;
    if ((st = striped_plan(cols, rows, bpp, gt, crs, zone, &head))) return st;
    if (!path) return fail(SARDINE_ERR_INVALID_ARGUMENT, "path is null");
    const uint64_t n = (uint64_t)cols * (uint64_t)rows * (uint64_t)bpp;
    if (n > 0 && !data) return fail(SARDINE_ERR_INVALID_ARGUMENT, "data is null");
    if ((st = require_device())) return st;
    FileCloser fc;
    fc.f = std::fopen(path, "wb");
    if (!fc.f) return fail(SARDINE_ERR_EXPORT_IO, "cannot create %s: %s", path, std::strerror(errno));
    if (std::fwrite(head.data(), 1, head.size(), fc.f) != head.size())
        return fail(SARDINE_ERR_EXPORT_IO, "writing %s: %s", path, std::strerror(errno));
    if ((st = stream_to_file(fc.f, data, n, (hipStream_t)stream_v, path))) return st;
    const int rc = std::fclose(fc.f);
    fc.f = nullptr;
    if (rc != 0) return fail(SARDINE_ERR_EXPORT_IO, "closing %s: %s", path, std::strerror(errno));
    return SARDINE_OK;
}

}  // namespace

extern "C" {

sardine_status sardine_cog_layout_compute(int64_t cols, int64_t rows, uint32_t tile_size, int32_t crs,
                                          sardine_cog_layout* layout) {
    if (!layout) return fail(SARDINE_ERR_INVALID_ARGUMENT, "layout is null");
    if (crs != SARDINE_CRS_WGS84 && crs != SARDINE_CRS_UTM_NORTH && crs != SARDINE_CRS_UTM_SOUTH)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "unknown CRS %d", crs);
    return cog_layout(cols, rows, tile_size, crs == SARDINE_CRS_WGS84 ? 16 : 20, layout);
}

sardine_status sardine_cog_header(const sardine_cog_layout* layout, const double geotransform[6], int32_t crs,
                                  int32_t utm_zone, uint8_t* out, uint64_t capacity) {
    sardine_status st;
    if ((st = cog_check_geotransform(geotransform))) return st;
    std::vector<uint16_t> keys;
    if ((st = cog_geo_keys(crs, utm_zone, &keys))) return st;
    sardine_cog_layout L;
    if ((st = cog_relayout(layout, &L))) return st;
    if ((int32_t)keys.size() != L.geo_key_shorts)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "layout was computed for another CRS");
    if (!out || capacity < L.header_bytes)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "header needs %llu bytes", (unsigned long long)L.header_bytes);
    std::vector<uint8_t> b;
    cog_header_bytes(L, geotransform, keys, b);
    std::memcpy(out, b.data(), b.size());
    return SARDINE_OK;
}

sardine_status sardine_cog_pack(const float* data, const sardine_cog_layout* layout, float* tiles, void* stream) {
    sardine_status st;
    sardine_cog_layout L;
    if ((st = cog_relayout(layout, &L))) return st;
    const sardine_cog_level& f = L.level[L.n_levels - 1];
    if (f.cols > 0 && f.rows > 0 && !data) return fail(SARDINE_ERR_INVALID_ARGUMENT, "data is null");
    if (L.tile_bytes > 0 && (!tiles || reinterpret_cast<uintptr_t>(tiles) % 16 != 0))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "tiles must be a 16-byte aligned device buffer");
    if ((st = require_device())) return st;
    return cog_pack(data, L, tiles, (hipStream_t)stream);
}

sardine_status sardine_geotiff_header(int64_t cols, int64_t rows, int32_t bytes_per_pixel, const double geotransform[6],
                                     int32_t crs, int32_t utm_zone, uint8_t* out, uint64_t capacity, uint64_t* size) {
    sardine_status st;
    std::vector<uint8_t> head;
    if ((st = striped_plan(cols, rows, bytes_per_pixel, geotransform, crs, utm_zone, &head))) return st;
    if (size) *size = head.size();
    if (out) {
        if (capacity < head.size())
            return fail(SARDINE_ERR_INVALID_ARGUMENT, "header needs %zu bytes", head.size());
        std::memcpy(out, head.data(), head.size());
    }
    return SARDINE_OK;
}

static sardine_status sardine_write_geotiff_with_crs_impl(const char* path, const float* data, int64_t cols, int64_t rows,
                                              const double geotransform[6], int32_t crs, int32_t utm_zone,
                                              void* stream) {
    if (!geotransform) return fail(SARDINE_ERR_INVALID_ARGUMENT, "geotransform is null");
    return striped_write(path, data, cols, rows, 4, geotransform, crs, utm_zone, stream);
}
sardine_status sardine_write_geotiff_with_crs(const char* path, const float* data, int64_t cols, int64_t rows,
                                              const double geotransform[6], int32_t crs, int32_t utm_zone,
                                              void* stream) {
    return guarded(SARDINE_ERR_EXPORT_IO, [&] {
        return sardine_write_geotiff_with_crs_impl(path, data, cols, rows, geotransform, crs, utm_zone, stream);
    });
}

static sardine_status sardine_write_geotiff_u8_with_crs_impl(const char* path, const uint8_t* data, int64_t cols, int64_t rows,
                                                 const double geotransform[6], int32_t crs, int32_t utm_zone,
                                                 void* stream) {
    if (!geotransform) return fail(SARDINE_ERR_INVALID_ARGUMENT, "geotransform is null");
    return striped_write(path, data, cols, rows, 1, geotransform, crs, utm_zone, stream);
}
sardine_status sardine_write_geotiff_u8_with_crs(const char* path, const uint8_t* data, int64_t cols, int64_t rows,
                                                 const double geotransform[6], int32_t crs, int32_t utm_zone,
                                                 void* stream) {
    return guarded(SARDINE_ERR_EXPORT_IO, [&] {
        return sardine_write_geotiff_u8_with_crs_impl(path, data, cols, rows, geotransform, crs, utm_zone, stream);
    });
}

static sardine_status sardine_write_tiff_no_crs_impl(const char* path, const float* data, int64_t cols, int64_t rows,
                                         void* stream) {
    return striped_write(path, data, cols, rows, 4, nullptr, 0, 0, stream);
}
sardine_status sardine_write_tiff_no_crs(const char* path, const float* data, int64_t cols, int64_t rows,
                                         void* stream) {
    return guarded(SARDINE_ERR_EXPORT_IO, [&] {
        return sardine_write_tiff_no_crs_impl(path, data, cols, rows, stream);
    });
}

static sardine_status sardine_write_cog_with_crs_impl(const char* path, const float* data, int64_t cols, int64_t rows,
                                          const double geotransform[6], int32_t crs, int32_t utm_zone,
                                          uint32_t tile_size, void* stream_v) {
    sardine_status st;
    if ((st = cog_check_geotransform(geotransform))) return st;
    if (!cog_tile_size_ok(tile_size))
        return fail(SARDINE_ERR_EXPORT_INVALID_TILE_SIZE,
                    "COG tile size must be a power of two >= 64 and <= 4096; got %u", tile_size);
    std::vector<uint16_t> keys;
    if ((st = cog_geo_keys(crs, utm_zone, &keys))) return st;
    if (!path) return fail(SARDINE_ERR_INVALID_ARGUMENT, "path is null");
    if (cols < 0 || rows < 0 || (cols > 0 && rows > 0 && !data))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "bad raster");
    sardine_cog_layout L;
    if ((st = cog_layout(cols, rows, tile_size, (int32_t)keys.size(), &L))) return st;
    if ((st = require_device())) return st;
    hipStream_t s = (hipStream_t)stream_v;
    DevBuf tiles;
    SD_HIP(tiles.alloc(L.tile_bytes, s));
    if ((st = cog_pack(data, L, tiles.as<float>(), s))) return st;
    std::vector<uint8_t> head;
    cog_header_bytes(L, geotransform, keys, head);
    FileCloser fc;
    fc.f = std::fopen(path, "wb");
    if (!fc.f) return fail(SARDINE_ERR_EXPORT_IO, "cannot create %s: %s", path, std::strerror(errno));
    if (std::fwrite(head.data(), 1, head.size(), fc.f) != head.size())
        return fail(SARDINE_ERR_EXPORT_IO, "writing %s: %s", path, std::strerror(errno));
    if ((st = stream_to_file(fc.f, tiles.p, L.tile_bytes, s, path))) return st;
    const int rc = std::fclose(fc.f);
    fc.f = nullptr;
    if (rc != 0) return fail(SARDINE_ERR_EXPORT_IO, "closing %s: %s", path, std::strerror(errno));
    return SARDINE_OK;
}
sardine_status sardine_write_cog_with_crs(const char* path, const float* data, int64_t cols, int64_t rows,
                                          const double geotransform[6], int32_t crs, int32_t utm_zone,
                                          uint32_t tile_size, void* stream_v) {
    return guarded(SARDINE_ERR_EXPORT_IO, [&] {
        return sardine_write_cog_with_crs_impl(path, data, cols, rows, geotransform, crs, utm_zone, tile_size, stream_v);
    });
}

}  // extern "C"

// ── resident scenes and batched geocoding ────────────────────────────────
namespace {

extern "C" hipError_t sd_launch_batch_tc(const sd::BatchHeader* H, const sd::BatchJob* J, unsigned int* queue,
                                         int n_cu, int tc_groups_per_cu, int orbit_lds, hipStream_t stream);
extern "C" int64_t sd_speckle_tiles(int64_t cols, int64_t rows, int64_t* tiles_x);
extern "C" size_t sd_speckle_job_size(void);
extern "C" void sd_speckle_job_fill(void* entry, int kind, const float* in, float* out, int64_t cols, int64_t rows,
                                    int half, float enl, float damping, int64_t tile_begin, int64_t tiles_x);
extern "C" hipError_t sd_launch_batch_speckle(const void* H, const void* J, unsigned int* queue, int n_groups,
                                              int max_half, hipStream_t stream);

struct SpeckleBatchHeaderHost {  // layout of sd::SpeckleBatchHeader (speckle.hip)
    int32_t n_jobs;
    int32_t pad;
    int64_t n_tiles;
};

}  // namespace

struct sardine_resident_scene_s {
    sardine_radar_image merged;
    sardine_dem_raster dem;
    sardine_dem_tiling tiling;  // dem.tiling points here for a mosaic
    sardine_scene scene;  // points into the vectors below
    std::vector<sardine_state_vector> svs;
    std::vector<sardine_grid_point> gps;
    int device = 0;
    void* d_mem = nullptr;  // orbit + seed table
    SceneDev dev;
};

struct sardine_batch_s {
    int max_jobs = 0;
    bool use_graph = false;
    int n_cu = 0;
    int device = 0;
    sardine_batch_tuning tuning{0, 2, 4};
    size_t sp_job_size = 0;
    // device tables (one allocation) and their pinned host images
    void* d_mem = nullptr;
    sd::BatchHeader* d_H = nullptr;
    sd::BatchJob* d_J = nullptr;
    void* d_SH = nullptr;
    void* d_SJ = nullptr;
    unsigned long long* d_cnt = nullptr;  // max_jobs x kCntSlots
    unsigned long long* d_queue = nullptr;  // as 32-bit counters: [0..5] chain items per job class, [6] speckle tiles
    void* h_mem = nullptr;
    sd::BatchHeader* h_H = nullptr;
    sd::BatchJob* h_J = nullptr;
    SpeckleBatchHeaderHost* h_SH = nullptr;
    char* h_SJ = nullptr;
    unsigned long long* h_cnt = nullptr;
    hipStream_t work = nullptr;  // graph replays run here, ordered against the caller's stream by events
    hipEvent_t ev_in = nullptr, ev_out = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraphExec_t exec = nullptr;
    sardine_batch_tuning graph_tuning{-1, -1, -1};
    int graph_half = -1;
    int orbit_lds = 0;        // the chain kernels' orbit LDS (bytes) of the next launch
    int graph_orbit_lds = -1; // and of the captured graph
};

namespace {

size_t cnt_stride() { return sd::kCntSlots; }

// The launch sequence of one run (eager or under capture).
hipError_t batch_launches(sardine_batch_s* b, int max_half, hipStream_t s) {
    hipError_t e;
    if ((e = sd_launch_zero_u64(b->d_cnt, (int64_t)b->max_jobs * cnt_stride(), s)) != hipSuccess) return e;
    if ((e = sd_launch_zero_u64(b->d_queue, 4, s)) != hipSuccess) return e;
    unsigned int* q = reinterpret_cast<unsigned int*>(b->d_queue);
    if ((e = sd_launch_batch_tc(b->d_H, b->d_J, q, b->n_cu, b->tuning.tc_groups_per_cu, b->orbit_lds, s)) !=
        hipSuccess)
        return e;
    return sd_launch_batch_speckle(b->d_SH, b->d_SJ, q + 6, b->n_cu * b->tuning.speckle_groups_per_cu, max_half, s);
}

}  // namespace

extern "C" {

sardine_status sardine_scene_create(const sardine_radar_image* merged, const sardine_scene* scene,
                                    const sardine_dem_raster* dem, sardine_resident_scene* out) {
    if (!merged || !scene || !dem || !out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (scene->n_grid_points <= 0 || !scene->grid_points)
        return fail(SARDINE_ERR_TC_NO_GRID_POINTS, "no geolocation grid points provided");
    if (scene->orbit.count > sd::kMaxSv || scene->orbit.count < 0 || (scene->orbit.count > 0 && !scene->orbit.state_vectors))
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "orbit has %d state vectors (max %d)", scene->orbit.count, sd::kMaxSv);
    sardine_status st;
    if ((st = require_device())) return st;
    std::unique_ptr<sardine_resident_scene_s> r(new sardine_resident_scene_s());
    r->merged = *merged;
    r->dem = *dem;
    if (dem->tiling) {
        r->tiling = *dem->tiling;
        r->dem.tiling = &r->tiling;
    }
    r->scene = *scene;
    r->svs.assign(scene->orbit.state_vectors, scene->orbit.state_vectors + scene->orbit.count);
    r->gps.assign(scene->grid_points, scene->grid_points + scene->n_grid_points);
    r->scene.orbit.state_vectors = r->svs.data();
    r->scene.grid_points = r->gps.data();
    SD_HIP(hipGetDevice(&r->device));
    const sd::OrbitDev orb = make_orbit_dev(r->scene.orbit);
    const SeedLut lut = build_seed_lut(r->scene.grid_points, r->scene.n_grid_points);
    const size_t n = lut.lat.size();
    const size_t tab_off = round256(sizeof orb) + round256(n * sizeof(double)) + round256(n * sizeof(int64_t));
    const size_t bytes = tab_off + (dem->tiling ? sd::kDemTabWords * sizeof(double) : 0);
    SD_HIP(hipMalloc(&r->d_mem, bytes));
    char* base = static_cast<char*>(r->d_mem);
    SD_HIP(hipMemcpy(base, &orb, sizeof orb, hipMemcpyHostToDevice));
    SD_HIP(hipMemcpy(base + round256(sizeof orb), lut.lat.data(), n * sizeof(double), hipMemcpyHostToDevice));
    SD_HIP(hipMemcpy(base + round256(sizeof orb) + round256(n * sizeof(double)), lut.t.data(), n * sizeof(int64_t),
                     hipMemcpyHostToDevice));
    r->dev.orb = reinterpret_cast<const sd::OrbitDev*>(base);
    r->dev.lut_lat = reinterpret_cast<const double*>(base + round256(sizeof orb));
    r->dev.lut_t = reinterpret_cast<const int64_t*>(base + round256(sizeof orb) + round256(n * sizeof(double)));
    r->dev.n_lut = (int)n;
    r->dev.epoch = orb.n > 0 ? r->svs[0].time_us : 0;
    if (dem->tiling) {
        double tab[sd::kDemTabWords];
        dem_tile_table(dem->tiling, tab);
        SD_HIP(hipMemcpy(base + tab_off, tab, sizeof tab, hipMemcpyHostToDevice));
        r->dev.dem_tab = reinterpret_cast<const double*>(base + tab_off);
    }
    *out = r.release();
    return SARDINE_OK;
}

void sardine_scene_destroy(sardine_resident_scene r) {
    if (!r) return;
    if (r->d_mem) (void)hipFree(r->d_mem);
    delete r;
}

sardine_status sardine_batch_create(int32_t max_jobs, int32_t use_graph, sardine_batch* out) {
    if (!out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (max_jobs < 1 || max_jobs > 65536) return fail(SARDINE_ERR_CAPACITY, "max_jobs %d outside 1..65536", max_jobs);
    sardine_status st;
    if ((st = require_device())) return st;
    std::unique_ptr<sardine_batch_s> b(new sardine_batch_s());
    b->max_jobs = max_jobs;
    b->use_graph = use_graph != 0;
    SD_HIP(hipGetDevice(&b->device));
    hipDeviceProp_t prop;
    SD_HIP(hipGetDeviceProperties(&prop, b->device));
    b->n_cu = prop.multiProcessorCount;
    b->sp_job_size = sd_speckle_job_size();
    const size_t szH = round256(sizeof(sd::BatchHeader)), szJ = round256((size_t)max_jobs * sizeof(sd::BatchJob));
    const size_t szSH = round256(sizeof(SpeckleBatchHeaderHost)), szSJ = round256((size_t)max_jobs * b->sp_job_size);
    const size_t szC = round256((size_t)max_jobs * cnt_stride() * sizeof(unsigned long long));
    const size_t szQ = 256;
    SD_HIP(hipMalloc(&b->d_mem, szH + szJ + szSH + szSJ + szC + szQ));
    char* d = static_cast<char*>(b->d_mem);
    b->d_H = reinterpret_cast<sd::BatchHeader*>(d);
    b->d_J = reinterpret_cast<sd::BatchJob*>(d + szH);
    b->d_SH = d + szH + szJ;
    b->d_SJ = d + szH + szJ + szSH;
    b->d_cnt = reinterpret_cast<unsigned long long*>(d + szH + szJ + szSH + szSJ);
    b->d_queue = reinterpret_cast<unsigned long long*>(d + szH + szJ + szSH + szSJ + szC);
    SD_HIP(hipHostMalloc(&b->h_mem, szH + szJ + szSH + szSJ + szC, hipHostMallocDefault));
    char* h = static_cast<char*>(b->h_mem);
    b->h_H = reinterpret_cast<sd::BatchHeader*>(h);
    b->h_J = reinterpret_cast<sd::BatchJob*>(h + szH);
    b->h_SH = reinterpret_cast<SpeckleBatchHeaderHost*>(h + szH + szJ);
    b->h_SJ = h + szH + szJ + szSH;
    b->h_cnt = reinterpret_cast<unsigned long long*>(h + szH + szJ + szSH + szSJ);
    SD_HIP(hipStreamCreateWithFlags(&b->work, hipStreamNonBlocking));
    SD_HIP(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
    SD_HIP(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
    SD_HIP(hipEventCreate(&b->t0));
    SD_HIP(hipEventCreate(&b->t1));
    *out = b.release();
    return SARDINE_OK;
}

void sardine_batch_destroy(sardine_batch b) {
    if (!b) return;
    if (b->work) (void)hipStreamSynchronize(b->work);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    for (hipEvent_t e : {b->ev_in, b->ev_out, b->t0, b->t1})
        if (e) (void)hipEventDestroy(e);
    if (b->work) (void)hipStreamDestroy(b->work);
    if (b->d_mem) (void)hipFree(b->d_mem);
    if (b->h_mem) (void)hipHostFree(b->h_mem);
    delete b;
}

sardine_status sardine_batch_set_tuning(sardine_batch b, const sardine_batch_tuning* t) {
    if (!b || !t) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    if (t->segment_columns < 0 || t->segment_columns > 4096 || t->tc_groups_per_cu < 1 || t->tc_groups_per_cu > 8 ||
        t->speckle_groups_per_cu < 1 || t->speckle_groups_per_cu > 16)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "tuning out of range");
    b->tuning = *t;
    return SARDINE_OK;
}

sardine_status sardine_batch_get_tuning(sardine_batch b, sardine_batch_tuning* t) {
    if (!b || !t) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    *t = b->tuning;
    return SARDINE_OK;
}

}  // extern "C"

namespace {

// One batch run in flight: the plans of its jobs and its work arena.
struct BatchRun {
    std::vector<TcPlan> plans;
    std::vector<int> run_idx;
    DevBuf work;
    hipStream_t s = nullptr;  // where the launches went
    sardine_status first = SARDINE_OK;
    std::string first_msg;
    bool launched = false;
};

void note_failure(BatchRun& run, sardine_status st) {
    if (st && !run.first) {
        run.first = st;
        run.first_msg = g_last_error;
    }
}

// Plans every job, fills and uploads the job tables and enqueues the
// launches (graph replay or eager) and the counter read-back on the
// batch's stream; b->ev_out marks the end.  Per-job validation failures go
// to jobs[i].status.
sardine_status batch_begin(sardine_batch_s* b, const sardine_resident_scene* scenes, int32_t n_scenes,
                           sardine_batch_job* jobs, int32_t n_jobs, hipStream_t stream, BatchRun& run) {
    if (!b || (n_jobs > 0 && (!jobs || !scenes))) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    if (n_jobs < 0 || n_jobs > b->max_jobs)
        return fail(SARDINE_ERR_CAPACITY, "%d jobs for a batch of at most %d", n_jobs, b->max_jobs);
    sardine_status st;
    run.plans.assign((size_t)n_jobs, TcPlan());
    std::vector<TcWork> works((size_t)n_jobs);
    std::vector<int> halves((size_t)n_jobs, 0);
    size_t arena = 0;
    for (int i = 0; i < n_jobs; ++i) {
        sardine_batch_job& jb = jobs[i];
        st = SARDINE_OK;
        if (jb.scene < 0 || jb.scene >= n_scenes || !scenes[jb.scene] || !jb.out)
            st = fail(SARDINE_ERR_INVALID_ARGUMENT, "job %d: bad scene index or output", i);
        const sardine_resident_scene_s* r = st ? nullptr : scenes[jb.scene];
        if (!st && r->device != b->device) st = fail(SARDINE_ERR_INVALID_ARGUMENT, "job %d: scene on another device", i);
        if (!st) st = plan_tc(&r->merged, &r->scene, &r->dem, &jb.config, jb.out, r->dev, &run.plans[(size_t)i]);
        if (!st && jb.config.precision != SARDINE_PRECISION_FP64)
            st = fail(SARDINE_ERR_TC_CONFIG, "job %d: the FP32 and mixed variants run through sardine_terrain_correction",
                      i);
        if (!st && jb.speckle.kind >= 0) {
            int win = 0;
            st = validate_filter(&jb.speckle, &win);
            if (!st && (!jb.speckle_out || jb.speckle_out == jb.out->data))
                st = fail(SARDINE_ERR_INVALID_ARGUMENT, "job %d: speckle output missing or aliasing the data", i);
            halves[(size_t)i] = win / 2;
        }
        jb.status = st;
        if (st) {
            note_failure(run, st);
            continue;
        }
        set_decomposition(run.plans[(size_t)i].p, b->tuning.segment_columns);
        works[(size_t)i] = tc_work_bytes(run.plans[(size_t)i], &jb.config);
        arena += works[(size_t)i].total();
        run.run_idx.push_back(i);
    }
    if (run.run_idx.empty()) return SARDINE_OK;
    // the job table ordered by class (resampler * 2 + UTM): each class's chain
    // kernel takes a contiguous range of the chain items
    auto job_class = [&](int i) { return run.plans[(size_t)i].p.resampling * 2 + (run.plans[(size_t)i].utm ? 1 : 0); };
    std::stable_sort(run.run_idx.begin(), run.run_idx.end(), [&](int a, int c) { return job_class(a) < job_class(c); });
    if ((st = require_device())) return st;
    hipStream_t s = b->use_graph ? b->work : stream;
    run.s = s;
    if (b->use_graph) {  // the batch's stream starts after the caller's earlier work
        SD_HIP(hipEventRecord(b->ev_in, stream));
        SD_HIP(hipStreamWaitEvent(s, b->ev_in, 0));
    }
    SD_HIP(run.work.alloc(arena, s));
    char* base = run.work.as<char>();
    sd::BatchHeader& H = *b->h_H;
    std::memset(&H, 0, sizeof H);
    SpeckleBatchHeaderHost& SH = *b->h_SH;
    std::memset(&SH, 0, sizeof SH);
    int max_half = 1, orbit_lds = 0;
    int k = 0;
    for (int i : run.run_idx) {
        TcPlan& pl = run.plans[(size_t)i];
        assign_work(pl, works[(size_t)i], base);
        // the batch's fix-up pass walks tie rows itself (no parallel tie pass)
        pl.p.tie_list = nullptr;
        pl.p.tie_cap = 0;
        if (works[(size_t)i].geoid)
            SD_HIP(staging().copy(const_cast<float*>(pl.p.geoid), jobs[i].config.geoid.grid_73x144,
                                  73 * 144 * sizeof(float), s));
        base += works[(size_t)i].total();
        pl.p.counters = b->d_cnt + (size_t)k * cnt_stride();
        pl.p.error_flag = reinterpret_cast<int*>(pl.p.counters + sd::kNumCounters);
        sd::BatchJob& J = b->h_J[k];
        std::memset(&J, 0, sizeof J);
        J.p = pl.p;
        J.q = pl.q;
        J.cls = job_class(i);
        J.tc_row_blocks = (pl.nr + 63) / 64;
        J.tc_begin = H.n_tc;
        H.n_tc += J.tc_row_blocks * ((pl.p.nseg + 3) / 4);
        for (int c = J.cls + 1; c < 7; ++c) H.tc_cls[c] = H.n_tc;  // classes ascend along the table
        orbit_lds = std::max(orbit_lds, (int)pl.p.orbit_lds);
        J.fix_begin = H.n_fix;
        H.n_fix += (pl.nr + 255) / 256;
        J.row_begin = H.n_rows;
        H.n_rows += pl.nr;
        J.anchor_begin = H.n_anchor;
        // UTM: anchors, then column blends; EPSG:4326: the column terms
        H.n_anchor += pl.utm ? pl.nr * pl.p.anchor_stride + pl.nc : pl.nc;
        const sardine_batch_job& jb = jobs[i];
        int64_t tiles_x = 1, tiles = 0;
        if (jb.speckle.kind >= 0) {
            tiles = sd_speckle_tiles(pl.nc, pl.nr, &tiles_x);
            max_half = std::max(max_half, halves[(size_t)i]);
        }
        sd_speckle_job_fill(b->h_SJ + (size_t)k * b->sp_job_size, jb.speckle.kind >= 0 ? jb.speckle.kind : 0,
                            jb.out->data, jb.speckle_out, pl.nc, pl.nr, halves[(size_t)i], jb.speckle.enl,
                            jb.speckle.damping, SH.n_tiles, tiles_x);
        SH.n_tiles += tiles;
        ++k;
    }
    H.n_jobs = k;
    SH.n_jobs = k;
    b->orbit_lds = orbit_lds;
    if (H.n_tc >= 0xffffffffLL || SH.n_tiles >= 0xffffffffLL)
        return fail(SARDINE_ERR_CAPACITY, "batch work above 2^32 items");
    // the four tables lie at the same offsets in the pinned and the device
    // block: one transfer from the header to the last used speckle job (each
    // transfer leaves the stream idle for tens of microseconds)
    {
        const size_t n = (size_t)(b->h_SJ - reinterpret_cast<char*>(b->h_H)) + (size_t)k * b->sp_job_size;
        SD_HIP(hipMemcpyAsync(b->d_H, b->h_H, n, hipMemcpyHostToDevice, s));
    }
    SD_HIP(hipEventRecord(b->t0, s));
    if (b->use_graph) {
        const sardine_batch_tuning& t = b->tuning;
        const sardine_batch_tuning& g = b->graph_tuning;
        if (!b->exec || g.tc_groups_per_cu != t.tc_groups_per_cu || g.speckle_groups_per_cu != t.speckle_groups_per_cu ||
            b->graph_half != max_half || b->graph_orbit_lds != b->orbit_lds) {
            if (b->exec) {
                SD_HIP(hipGraphExecDestroy(b->exec));
                b->exec = nullptr;
            }
            hipGraph_t graph = nullptr;
            SD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const hipError_t le = batch_launches(b, max_half, s);
            const hipError_t ce = hipStreamEndCapture(s, &graph);
            if (le != hipSuccess || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(SARDINE_ERR_HIP, "graph capture failed: %s", hipGetErrorString(le != hipSuccess ? le : ce));
            }
            const hipError_t ie = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return fail(SARDINE_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
            b->graph_tuning = t;
            b->graph_half = max_half;
            b->graph_orbit_lds = b->orbit_lds;
        }
        SD_HIP(hipGraphLaunch(b->exec, s));
    } else {
        SD_HIP(batch_launches(b, max_half, s));
    }
    SD_HIP(hipEventRecord(b->t1, s));
    SD_HIP(hipMemcpyAsync(b->h_cnt, b->d_cnt, (size_t)k * cnt_stride() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    SD_HIP(hipEventRecord(b->ev_out, s));
    if (b->use_graph) SD_HIP(hipStreamWaitEvent(stream, b->ev_out, 0));  // later caller work sees the outputs
    run.launched = true;
    return SARDINE_OK;
}

// Waits for a launched run and fills every job's result.
sardine_status batch_end(sardine_batch_s* b, const sardine_resident_scene* scenes, sardine_batch_job* jobs,
                         BatchRun& run, double* device_ms) {
    if (run.launched) {
        SD_HIP(hipEventSynchronize(b->ev_out));
        if (device_ms) {
            float ms = 0.0f;
            SD_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
            *device_ms = ms;
        }
        int k = 0;
        for (int i : run.run_idx) {
            const sardine_status st = finish_tc(run.plans[(size_t)i], b->h_cnt + (size_t)k * cnt_stride(),
                                                scenes[jobs[i].scene]->scene.orbit, jobs[i].out);
            jobs[i].status = st;
            note_failure(run, st);
            ++k;
        }
        run.launched = false;
    }
    if (run.first) g_last_error = run.first_msg;
    return run.first;
}

}  // namespace

extern "C" {

static sardine_status sardine_batch_run_impl(sardine_batch b, const sardine_resident_scene* scenes, int32_t n_scenes,
                                 sardine_batch_job* jobs, int32_t n_jobs, void* stream_v, double* device_ms) {
    BatchRun run;
    sardine_status st = batch_begin(b, scenes, n_scenes, jobs, n_jobs, (hipStream_t)stream_v, run);
    if (st) {
        if (run.launched) (void)hipEventSynchronize(b->ev_out);  // never leave work reading freed tables
        return st;
    }
    return batch_end(b, scenes, jobs, run, device_ms);
}
sardine_status sardine_batch_run(sardine_batch b, const sardine_resident_scene* scenes, int32_t n_scenes,
                                 sardine_batch_job* jobs, int32_t n_jobs, void* stream_v, double* device_ms) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] {
        return sardine_batch_run_impl(b, scenes, n_scenes, jobs, n_jobs, stream_v, device_ms);
    });
}

static sardine_status sardine_batch_autotune_impl(sardine_batch b, const sardine_resident_scene* scenes, int32_t n_scenes,
                                      sardine_batch_job* jobs, int32_t n_jobs, void* stream,
                                      sardine_batch_tuning* best, double* best_ms) {
    if (!b) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    static const int segs[] = {16, 24, 32, 48};
    static const int tcg[] = {1, 2};
    static const int spg[] = {2, 4, 8};
    const sardine_batch_tuning keep = b->tuning;
    sardine_batch_tuning chosen = keep;
    double chosen_ms = 1e30;
    sardine_status st;
    auto measure = [&](const sardine_batch_tuning& t, double* ms) -> sardine_status {
        b->tuning = t;
        double a = 0.0, c = 0.0;
        sardine_status s1 = sardine_batch_run(b, scenes, n_scenes, jobs, n_jobs, stream, &a);
        if (s1) return s1;
        s1 = sardine_batch_run(b, scenes, n_scenes, jobs, n_jobs, stream, &c);
        if (s1) return s1;
        *ms = std::min(a, c);
        return SARDINE_OK;
    };
    // chain pass shape first (segment length x workgroups per CU), then the
    // speckle pass's workgroups per CU with the chosen chain shape
    for (int sg : segs)
        for (int g : tcg) {
            sardine_batch_tuning t{sg, g, keep.speckle_groups_per_cu};
            double ms = 0.0;
            if ((st = measure(t, &ms))) {
                b->tuning = keep;
                return st;
            }
            if (ms < chosen_ms) {
                chosen_ms = ms;
                chosen = t;
            }
        }
    for (int g : spg) {
        sardine_batch_tuning t{chosen.segment_columns, chosen.tc_groups_per_cu, g};
        if (g == chosen.speckle_groups_per_cu) continue;
        double ms = 0.0;
        if ((st = measure(t, &ms))) {
            b->tuning = keep;
            return st;
        }
        if (ms < chosen_ms) {
            chosen_ms = ms;
            chosen = t;
        }
    }
    b->tuning = chosen;
    if (best) *best = chosen;
    if (best_ms) *best_ms = chosen_ms;
    return SARDINE_OK;
}
sardine_status sardine_batch_autotune(sardine_batch b, const sardine_resident_scene* scenes, int32_t n_scenes,
                                      sardine_batch_job* jobs, int32_t n_jobs, void* stream,
                                      sardine_batch_tuning* best, double* best_ms) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] {
        return sardine_batch_autotune_impl(b, scenes, n_scenes, jobs, n_jobs, stream, best, best_ms);
    });
}

}  // extern "C"

// ── serving ──────────────────────────────────────────────────────────────
namespace {

double now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ServerReq {
    uint64_t ticket = 0;
    sardine_request req{};
    int64_t rows = 0, cols = 0;
    size_t out_bytes = 0, work_bytes = 0;
    float* d_out = nullptr;
    float* d_speckled = nullptr;
    sardine_geocoded_image img{};
    sardine_response resp{};
};

struct ServerSlot {
    sardine_batch_s* b = nullptr;
    hipStream_t stream = nullptr;
    std::vector<ServerReq*> reqs;
    std::vector<sardine_batch_job> jobs;
    std::unique_ptr<BatchRun> run{new BatchRun()};
    size_t work_bytes = 0;
    bool busy = false;
    ServerSlot() = default;
    ServerSlot(const ServerSlot&) = delete;
    ServerSlot& operator=(const ServerSlot&) = delete;
    ~ServerSlot() {
        run.reset();
        if (b) sardine_batch_destroy(b);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace

struct sardine_server_s {
    std::vector<sardine_resident_scene> scenes;
    sardine_server_options opt{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerReq*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<ServerReq>> reqs;
    std::vector<std::unique_ptr<ServerSlot>> slots;
    hipStream_t release_stream = nullptr;
    uint64_t next_ticket = 1;
    uint64_t used = 0;  // bytes held: admitted outputs + work of batches in flight
    bool stop = false;
    std::thread loop;
    std::string loop_error;
};

namespace {

// Output grid shape and memory estimate of a request (host only).
sardine_status request_shape(const sardine_server_s* sv, const sardine_request* r, ServerReq* q) {
    if (r->scene < 0 || r->scene >= (int)sv->scenes.size())
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "scene index %d outside 0..%d", r->scene, (int)sv->scenes.size() - 1);
    Grid g;
    sardine_status st = output_grid(&sv->scenes[(size_t)r->scene]->scene, &r->config, &g);
    if (st) return st;
    int64_t nr = r->config.aoi_rows, nc = r->config.aoi_cols;
    if (nr == 0) {
        nr = g.rows;
        nc = g.cols;
    }
    if (nr <= 0 || nc <= 0) return fail(SARDINE_ERR_TC_CONFIG, "empty AOI window");
    if (r->speckle.kind >= 0) {
        int win = 0;
        if ((st = validate_filter(&r->speckle, &win))) return st;
    }
    q->rows = nr;
    q->cols = nc;
    const size_t px = (size_t)nr * (size_t)nc;
    q->out_bytes = round256(px * sizeof(float)) + (r->speckle.kind >= 0 ? round256(px * sizeof(float)) : 0);
    // chain records (segments >= 16 columns), row seeds, flat areas, anchors
    q->work_bytes = px * 2 + (size_t)nr * (16 + 2 * 8 * 260) + 4096;
    return SARDINE_OK;
}

// Finishes a completed slot: results, timestamps, memory accounting.
void complete_slot(sardine_server_s* sv, ServerSlot& sl) {
    (void)batch_end(sl.b, sv->scenes.data(), sl.jobs.data(), *sl.run, nullptr);
    const double t = now_us();
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        ServerReq* q = sl.reqs[i];
        q->resp.status = sl.jobs[i].status;
        q->resp.image = q->img;
        q->resp.speckled = q->d_speckled;
        q->resp.t_done_us = t;
        q->resp.done = 1;
    }
    sl.run.reset(new BatchRun());  // releases the work arena (stream-ordered)
    sv->used -= sl.work_bytes;
    sl.work_bytes = 0;
    sl.reqs.clear();
    sl.jobs.clear();
    sl.busy = false;
}

// Admits pending requests in order into a free slot while memory allows,
// then launches the batch.  Called with the lock held.
bool admit_and_launch(sardine_server_s* sv, ServerSlot& sl) {
    const double t = now_us();
    while (!sv->pending.empty() && (int)sl.reqs.size() < sv->opt.max_batch) {
        ServerReq* q = sv->pending.front();
        const size_t need = q->out_bytes + q->work_bytes;
        if (sv->used + need > sv->opt.hbm_budget_bytes) break;  // in order: wait for memory
        const size_t px = (size_t)q->rows * (size_t)q->cols;
        void* a = nullptr;
        if (hipMallocAsync(&a, q->out_bytes, sl.stream) != hipSuccess) break;
        q->d_out = static_cast<float*>(a);
        q->d_speckled = q->req.speckle.kind >= 0 ? q->d_out + round256(px * sizeof(float)) / sizeof(float) : nullptr;
        sv->used += need;  // outputs until release, work until the batch completes
        sl.work_bytes += q->work_bytes;
        q->resp.t_admit_us = t;
        std::memset(&q->img, 0, sizeof q->img);
        q->img.data = q->d_out;
        sardine_batch_job j;
        std::memset(&j, 0, sizeof j);
        j.scene = q->req.scene;
        j.config = q->req.config;
        j.out = &q->img;
        j.speckle = q->req.speckle;
        j.speckle_out = q->d_speckled;
        sl.reqs.push_back(q);
        sl.jobs.push_back(j);
        sv->pending.pop_front();
    }
    if (sl.reqs.empty()) return false;
    const double tl = now_us();
    for (ServerReq* q : sl.reqs) {
        q->resp.t_launch_us = tl;
        q->resp.batch_size = (int)sl.reqs.size();
    }
    sl.busy = true;
    const sardine_status st = batch_begin(sl.b, sv->scenes.data(), (int)sv->scenes.size(), sl.jobs.data(),
                                          (int)sl.jobs.size(), sl.stream, *sl.run);
    if (st) {  // the batch could not be launched: every request fails with this status
        if (sl.run->launched) (void)hipEventSynchronize(sl.b->ev_out);
        for (size_t i = 0; i < sl.jobs.size(); ++i)
            if (!sl.jobs[i].status) sl.jobs[i].status = st;
        sl.run->launched = false;
        complete_slot(sv, sl);
    }
    return true;
}

void server_loop(sardine_server_s* sv) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        bool progressed = false;
        for (auto& sp : sv->slots) {
            ServerSlot& sl = *sp;
            if (sl.busy && (!sl.run->launched || hipEventQuery(sl.b->ev_out) == hipSuccess)) {
                complete_slot(sv, sl);
                progressed = true;
            }
        }
        if (progressed) sv->cv_done.notify_all();
        for (auto& sp : sv->slots)
            if (!sp->busy && !sv->pending.empty() && admit_and_launch(sv, *sp)) progressed = true;
        bool busy = false;
        for (auto& sp : sv->slots) busy |= sp->busy;
        if (sv->stop && !busy && (sv->pending.empty() || !progressed)) break;
        if (!progressed) sv->cv_work.wait_for(lk, std::chrono::microseconds(busy ? 50 : 2000));
    }
}

}  // namespace

extern "C" {

sardine_status sardine_server_create(const sardine_resident_scene* scenes, int32_t n_scenes,
                                     const sardine_server_options* opt, sardine_server* out) {
    if (!scenes || n_scenes <= 0 || !opt || !out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (opt->max_batch < 1 || opt->max_batch > 4096 || opt->slots < 1 || opt->slots > 8 || opt->hbm_budget_bytes == 0)
        return fail(SARDINE_ERR_INVALID_ARGUMENT, "server options out of range");
    sardine_status st;
    if ((st = require_device())) return st;
    std::unique_ptr<sardine_server_s> sv(new sardine_server_s());
    sv->scenes.assign(scenes, scenes + n_scenes);
    for (auto s : sv->scenes)
        if (!s) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null scene");
    sv->opt = *opt;
    SD_HIP(hipGetDevice(&sv->device));
    SD_HIP(hipStreamCreateWithFlags(&sv->release_stream, hipStreamNonBlocking));
    for (int i = 0; i < opt->slots; ++i) {
        sv->slots.emplace_back(new ServerSlot());
        ServerSlot& sl = *sv->slots.back();
        if ((st = sardine_batch_create(opt->max_batch, opt->use_graph, &sl.b))) return st;
        SD_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    }
    sardine_server_s* raw = sv.get();
    sv->loop = std::thread(server_loop, raw);
    *out = sv.release();
    return SARDINE_OK;
}

static sardine_status sardine_server_submit_impl(sardine_server sv, const sardine_request* r, uint64_t* ticket) {
    if (!sv || !r || !ticket) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    std::unique_ptr<ServerReq> q(new ServerReq());
    q->req = *r;
    sardine_status st = request_shape(sv, r, q.get());
    if (st) return st;
    if (q->out_bytes + q->work_bytes > sv->opt.hbm_budget_bytes)
        return fail(SARDINE_ERR_CAPACITY, "request needs %zu bytes, above the server budget", q->out_bytes + q->work_bytes);
    q->resp.t_submit_us = now_us();
    std::lock_guard<std::mutex> g(sv->mu);
    if (sv->stop) return fail(SARDINE_ERR_INVALID_ARGUMENT, "server is stopping");
    q->ticket = sv->next_ticket++;
    *ticket = q->ticket;
    sv->pending.push_back(q.get());
    sv->reqs.emplace(q->ticket, std::move(q));
    sv->cv_work.notify_one();
    return SARDINE_OK;
}
sardine_status sardine_server_submit(sardine_server sv, const sardine_request* r, uint64_t* ticket) {
    return guarded(SARDINE_ERR_INVALID_ARGUMENT, [&] { return sardine_server_submit_impl(sv, r, ticket); });
}

sardine_status sardine_server_wait(sardine_server sv, uint64_t ticket, int32_t block, sardine_response* out) {
    if (!sv || !out) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return fail(SARDINE_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
    ServerReq* q = it->second.get();
    if (block) sv->cv_done.wait(lk, [&] { return q->resp.done != 0; });
    *out = q->resp;
    return SARDINE_OK;
}

sardine_status sardine_server_release(sardine_server sv, uint64_t ticket, void* stream_v) {
    if (!sv) return fail(SARDINE_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> g(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return fail(SARDINE_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
    ServerReq* q = it->second.get();
    if (!q->resp.done) return fail(SARDINE_ERR_INVALID_ARGUMENT, "request %llu has not finished", (unsigned long long)ticket);
    if (q->d_out) {
        // the free waits for the caller's work on `stream` enqueued so far
        // (its reads of the outputs), then the memory returns to the pool
        hipEvent_t ev = nullptr;
        SD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        hipError_t e = hipEventRecord(ev, (hipStream_t)stream_v);
        if (e == hipSuccess) e = hipStreamWaitEvent(sv->release_stream, ev, 0);
        (void)hipEventDestroy(ev);
        SD_HIP(e);
        (void)hipFreeAsync(q->d_out, sv->release_stream);
        sv->used -= q->out_bytes;
    }
    sv->reqs.erase(it);
    sv->cv_work.notify_one();
    return SARDINE_OK;
}

void sardine_server_destroy(sardine_server sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
        // requests not yet admitted are dropped
        for (ServerReq* q : sv->pending) {
            q->resp.status = SARDINE_ERR_INVALID_ARGUMENT;
            q->resp.done = 1;
        }
        sv->pending.clear();
        sv->cv_work.notify_one();
        sv->cv_done.notify_all();
    }
    if (sv->loop.joinable()) sv->loop.join();
    // outputs never released may still be read by the caller's queued work
    (void)hipSetDevice(sv->device);
    (void)hipDeviceSynchronize();
    for (auto& kv : sv->reqs)
        if (kv.second->d_out) (void)hipFreeAsync(kv.second->d_out, sv->release_stream);
    (void)hipStreamSynchronize(sv->release_stream);
    sv->slots.clear();
    (void)hipStreamDestroy(sv->release_stream);
    delete sv;
}

}  // extern "C"
