// tc_params.hpp — kernel-argument block of the terrain-correction kernels.
#pragma once

#include <stdint.h>

#include "geom.hpp"

namespace sd {

constexpr int kHistSize = 16;                // NEWTON_ITER_HIST_SIZE (terrain_correction.rs:95)
constexpr int kCntRedo = 7 + kHistSize;      // counters[kCntRedo] = chain redo rounds
constexpr int kNumCounters = kCntRedo + 1;
// After the counters: the orbit-error flags word, then how far the Newton
// iterates that fell outside the orbit window reached below its first and
// above its last state vector (µs, 0 = none).
constexpr int kCntOrbitBelow = kNumCounters + 1;
constexpr int kCntOrbitAbove = kNumCounters + 2;
// Then the Newton steps re-evaluated with the reference's arithmetic
// (zero_doppler's tie_stats: steps, decisions changed, largest deviation).
constexpr int kCntTie = kNumCounters + 3;
// Then the length of the tie list (check_row).
constexpr int kCntTieList = kNumCounters + 6;
constexpr int kCntSlots = kNumCounters + 7;
constexpr int kMaxAnchorPts = 1025;          // SARDINE_TC_PROJECTOR_ANCHORS ≤ 1024
constexpr size_t kMaxTcLds = 64 * 1024;      // dynamic LDS budget of the UTM anchors

// Chain bookkeeping of one (row, segment): the seed the segment started
// from, its last converged time, its orbit-error counts, and where its chain
// met tie steps (Newton steps its arithmetic could not decide): the seed
// entering the first such pixel and the first and last such columns.
struct SegRecord {
    int64_t seed;
    int64_t last;
    int32_t flags;      // bit 0: last is valid; bit 1: the segment ran a Newton solve; bit 2: ties;
                        // bit 3: a pixel converged before the first tie pixel
    int32_t err_outside;
    int32_t err_short;
    int32_t tie_first;  // AOI column of the first tie pixel (flags bit 2)
    int64_t tie_seed;   // the chain's seed entering that pixel
    int32_t tie_last;   // AOI column of the last tie pixel
    int32_t pad;
};

// Single-precision orbit of the FP32 scene-centred variant: times in
// seconds from a reference time, positions in metres from a scene centre.
struct alignas(16) OrbitF32 {
    int32_t n;
    int32_t pad[3];
    float t[kMaxSv];
    float pos[kMaxSv][3];
    float vel[kMaxSv][3];
    float bary[kMaxSv][kLagrange];
};

constexpr int kDemTabWords = 24;

// Projector::project_row's blend of one output column (output_crs.rs:395-414):
// the anchor interval (relative to the first stored anchor) and the weight
// t = (x - x0) / (x1 - x0).  Both depend on the column only, so they are
// computed once per AOI column (utm_col_kernel) instead of per pixel.
struct ColBlend {
    double t;
    int32_t a;
    int32_t pad;
};

// The longitude-only terms of one output column of an EPSG:4326 grid (the
// longitude is the same on every row): the sine and cosine of the target's
// longitude and the DEM column axes of the centre and of the east / west
// terrain-geometry neighbours (index and fraction as dem_col_axis returns
// them).  Computed once per AOI column (wgs_col_kernel).
struct ColGeo {
    double so, co;
    double fc, fe, fw;
    int32_t ic, ie, iw;
    int32_t pad;
};

struct TcParams {
    const OrbitDev* orbit;
    int32_t orbit_lds;      // bytes of its LDS copy, orbit_lds_bytes(n) (the per-call chain sizes its LDS by it)
    int32_t orbit_lds_pad;
    // FP32 scene-centred variant (precision SARDINE_PRECISION_FP32_CENTRED)
    const OrbitF32* orbit32;
    double centre[3];       // scene centre, ECEF metres
    int64_t t_ref_us;       // reference time, µs relative to the first state vector
    float tol32;            // Newton tolerance of the variant, seconds
    // radar image (MergedSigma0 buffers)
    const float* img;
    const float* nesz;      // NULL → no NESZ masking
    int64_t lines, samples;
    int64_t az_start_us;    // relative to the first state vector
    double near_range_m, rps, ati;
    double line_per_us;     // 1 / (1e6 · ati): azimuth µs → line (FP32 variant)
    double inv_ati;         // RN(1 / ati), for the exact division by ati
    double inv_rps;         // RN(1 / range pixel spacing)
    // DEM raster
    const float* dem;
    int64_t dem_w, dem_h;
    double dem_olat, dem_olon, dem_pix, dem_inv_pix;
    double dem_max_row, dem_max_col;   // height-1, width-1
    double dem_hi_row, dem_hi_col;     // height-0.5, width-0.5 (the outer margin)
    int dem_lim_row, dem_lim_col;      // largest lower bilinear index
    // GLO-30 mosaic (dem_tiled): a grid of tiles of dem_th x dem_tw pixels;
    // dem_tab (device, kDemTabWords): [0, 8) the origin latitude of each
    // tile row, [8, 16) the origin longitude of each tile column, [16, 24)
    // per tile row the load order of its tiles as bytes (column c in byte c,
    // 0xff: no tile), stored as the bits of a double
    int dem_tiled, dem_trows, dem_tcols, dem_th, dem_tw;
    const double* dem_tab;
    const float* geoid;     // device 73x144 or NULL
    // output grid
    int crs, utm_zone, utm_south, n_anchor_pts;
    double spacing, lat_north, lon_west;
    double x_start, anchor_dx, inv_anchor_dx;
    int64_t full_cols;
    int64_t row0, col0, rows, cols;
    int seg_len, nseg, rows_per_wg;  // work decomposition (see tc.hip)
    int lead;                        // phase-A lead columns
    double* anchors;                  // UTM: per AOI row lon[stride], lat[stride] of anchors anchor0..
    ColBlend* col_blend;              // UTM: per AOI column, or NULL (then computed per pixel)
    ColGeo* col_geo;                  // EPSG:4326: per AOI column, or NULL (then computed per pixel)
    int anchor0, anchor_stride;       // first stored anchor index, stored anchors per row
    SegRecord* records;               // per (AOI row, segment)
    int32_t* row_flags;               // per AOI row: 1 = the fix-up pass walks the row, 2 = tie list only
    int2* tie_list;                   // (AOI row, segment) of tie-only segments, or NULL
    int tie_cap;                      // its capacity
    const int64_t* row_seed;      // per AOI row, µs relative to the first state vector
    const double* row_flat_area;  // per AOI row (need_geom)
    // config
    int max_iter;
    double tol;
    double tie_band;  // µs (kTieBandUs unless set with sardine_set_tc_tie_band)
    int flatten, compute_lia, need_geom, resampling;
    int mixed;  // FP32 resampling and terrain geometry (the chain kernels' RS + 3)
    double step_deg;
    double step_sin, step_cos;  // sin/cos of step_deg in radians
    float nesz_factor;
    // outputs
    float* out;
    float* cos_lia;
    uint8_t* mask;
    double* radar_coords;
    unsigned long long* counters;  // kNumCounters
    int* error_flag;
};

// Per-row seeds and flat-cell areas (terrain_correction.rs:1424-1511),
// computed on the device by tc_rows_kernel.
struct RowPrepParams {
    const double* lut_lat;  // geolocation-grid seed table, sorted by latitude
    const int64_t* lut_t;   // azimuth time (absolute µs) of each entry
    int n_lut;
    int64_t epoch;          // time of the first state vector (µs)
    double y_top, x_left, spacing, x_start, anchor_dx, step_deg;
    int64_t full_cols, row0, rows;
    int n_anch, utm, south, zone, need_geom;
    int64_t* row_seed;      // out: µs relative to epoch
    double* row_flat_area;  // out (need_geom)
};

// One terrain-correction job of a batch (tc_batch.hip).  The *_begin
// fields are the job's first index in each pass's combined index space.
struct BatchJob {
    TcParams p;
    RowPrepParams q;
    int32_t cls;            // resampler * 2 + (UTM ? 1 : 0)
    int32_t pad;
    int64_t tc_begin;       // chain items: 64 rows x 4 segments each
    int64_t tc_row_blocks;  // ceil(rows / 64)
    int64_t fix_begin;      // fix-up blocks of 256 rows
    int64_t row_begin;      // rows pass (one lane per AOI row)
    int64_t anchor_begin;   // anchor pass (one lane per row x anchor point; UTM jobs)
};

struct BatchHeader {
    int32_t n_jobs;
    int32_t pad;
    int64_t n_tc, n_fix, n_rows, n_anchor;  // totals of the passes
    // chain items of job class c (resampler * 2 + UTM): [tc_cls[c], tc_cls[c + 1])
    // (the job table is ordered by class)
    int64_t tc_cls[7];
};

}  // namespace sd
