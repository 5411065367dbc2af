/*
 * sardine_hip.h — C ABI of the MI355X (gfx950) calibrate-and-geocode path.
 *
 * Every entry point replaces one CPU stage of SARdine's Rust pipeline.  The
 * comment above each declaration names the reference interface it replaces
 * (paths relative to the SARdine repository root).
 *
 * Conventions
 *  - Times are int64 microseconds on one caller-chosen epoch.  The reference
 *    does all time arithmetic in whole microseconds (chrono `Duration`
 *    quantisation, sardine/src/terrain_correction.rs:746-749,1067,1135 and
 *    sardine/src/orbit.rs:520,546), so this representation is exact.
 *  - Raster pointers marked [device] are HIP device pointers (inputs already
 *    resident in HBM).  Small metadata (orbit, LUT vectors, grid points) is
 *    passed as host memory and staged by the library.
 *  - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *    Calls are stream-ordered.  Entry points whose results include host
 *    values computed on the device (sardine_terrain_correction's counters,
 *    sardine_to_db_inplace's counts) return after their work has finished;
 *    the others (calibration, speckle, synthetic fill) return once the work
 *    is enqueued, and their outputs are valid for later work on `stream`.
 *    Host metadata is copied before the call returns.
 *  - Every function returns a sardine_status; on failure a message is
 *    available from sardine_last_error() (thread-local).
 *  - There is no CPU fallback: without a usable gfx950 device every compute
 *    entry point returns SARDINE_ERR_NO_DEVICE.
 */
#ifndef SARDINE_HIP_H
#define SARDINE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SARDINE_ABI_VERSION 3

/* ── Status codes ─────────────────────────────────────────────────────────
 * One code per reference error variant on the path:
 *   TerrainCorrectionError  sardine/src/terrain_correction.rs:45-67
 *   OrbitError              sardine/src/orbit.rs:40-75
 *   SpeckleError            sardine/src/speckle.rs:96-108
 *   CalibrationError        sardine/src/apply_calibration.rs:64-84
 *   DeburstError            sardine/src/deburst.rs:56-121
 *   MergeError              sardine/src/merge_subswaths.rs:62-81          */
typedef enum sardine_status {
    SARDINE_OK = 0,
    /* TerrainCorrectionError */
    SARDINE_ERR_TC_ORBIT_OUTSIDE_WINDOW = 10, /* Orbit(OutsideOrbitWindow): aborts the whole call */
    SARDINE_ERR_TC_ORBIT_INSUFFICIENT = 11,   /* Orbit(InsufficientCoverage): < 8 state vectors */
    SARDINE_ERR_TC_NO_GRID_POINTS = 12,
    SARDINE_ERR_TC_NO_SUBSWATHS = 13,
    SARDINE_ERR_TC_CONFIG = 14,
    SARDINE_ERR_TC_PROJECTION = 15,
    SARDINE_ERR_TC_DEM = 16,
    /* SpeckleError */
    SARDINE_ERR_SPECKLE_DIMENSION = 20,
    SARDINE_ERR_SPECKLE_BAD_WINDOW = 21,
    SARDINE_ERR_SPECKLE_BAD_ENL = 22,
    SARDINE_ERR_SPECKLE_BAD_DAMPING = 23,
    /* CalibrationError */
    SARDINE_ERR_CAL_NO_CAL_VECTORS = 30,
    SARDINE_ERR_CAL_NO_NOISE_VECTORS = 31,
    SARDINE_ERR_CAL_LUT_RANGE = 32,
    /* DeburstError */
    SARDINE_ERR_DEBURST_NO_BURSTS = 40,
    SARDINE_ERR_DEBURST_COUNT_MISMATCH = 41,
    SARDINE_ERR_DEBURST_TIFF_TOO_SMALL = 42,
    SARDINE_ERR_DEBURST_EMPTY_VALID_WINDOW = 43,
    SARDINE_ERR_DEBURST_INVALID_OVERLAP = 44,
    SARDINE_ERR_DEBURST_IMPLAUSIBLE_OVERLAP = 45,
    /* MergeError */
    SARDINE_ERR_MERGE_EMPTY_INPUTS = 50,
    SARDINE_ERR_MERGE_PIXEL_SPACING = 51,
    SARDINE_ERR_MERGE_EMPTY_SWATH = 52,
    /* ExportError (export.rs) */
    SARDINE_ERR_EXPORT_INVALID_GEOTRANSFORM = 60,
    SARDINE_ERR_EXPORT_INVALID_TILE_SIZE = 61,
    SARDINE_ERR_EXPORT_IO = 62,
    /* GroundRangeError (ground_range.rs:31-47) */
    SARDINE_ERR_GRD_NO_GRID_POINTS = 70,
    SARDINE_ERR_GRD_EMPTY_IMAGE = 71,
    SARDINE_ERR_GRD_INVALID_SPACING = 72,
    SARDINE_ERR_GRD_INVALID_AZIMUTH_SPACING = 73,
    SARDINE_ERR_GRD_INTERPOLATION_FAILED = 74,
    /* SlcReadError (slc_reader.rs:66-131; the multi-slice variants have no
     * counterpart here) */
    SARDINE_ERR_SLC_IO = 80,
    SARDINE_ERR_SLC_NOT_TIFF = 81,
    SARDINE_ERR_SLC_BIGTIFF = 82,
    SARDINE_ERR_SLC_MISSING_TAG = 83,
    SARDINE_ERR_SLC_UNSUPPORTED_FORMAT = 84,
    SARDINE_ERR_SLC_NON_CONTIGUOUS_STRIPS = 85,
    SARDINE_ERR_SLC_LINE_OUT_OF_BOUNDS = 86,
    SARDINE_ERR_SLC_CROSS_SLICE_BOUNDARY = 87,
    /* MultiSlcOpenError (multi_slc_reader.rs:54-71) */
    SARDINE_ERR_SLC_MULTI_EMPTY_INPUT = 88,
    SARDINE_ERR_SLC_MULTI_WIDTH_MISMATCH = 89,
    /* GeoidError (geoid.rs:87-104) */
    SARDINE_ERR_GEOID_IO = 100,
    SARDINE_ERR_GEOID_WRONG_SIZE = 101,
    SARDINE_ERR_GEOID_WRONG_SIZE_EGM2008 = 102,
    SARDINE_ERR_GEOID_MALFORMED_ASCII = 103,
    /* DemError (dem.rs:54-109), the variants the SRTM loader raises */
    SARDINE_ERR_DEM_IO = 110,
    SARDINE_ERR_DEM_BAD_SIZE = 111,
    SARDINE_ERR_DEM_TIFF_DECODE = 112,
    SARDINE_ERR_DEM_UNSUPPORTED_PIXEL_TYPE = 113,
    SARDINE_ERR_DEM_UNRECOGNISED_TILE_NAME = 114,
    SARDINE_ERR_DEM_NO_TILES = 115,
    /* library / device */
    SARDINE_ERR_INVALID_ARGUMENT = 90,
    SARDINE_ERR_NO_DEVICE = 91,
    SARDINE_ERR_HIP = 92,
    SARDINE_ERR_OUT_OF_MEMORY = 93,
    SARDINE_ERR_CAPACITY = 94
} sardine_status;

/* Thread-local message for the last failing call on this thread. */
const char* sardine_last_error(void);

/* The fields of the reference's error variants that a status alone does not
 * carry, for the last failing call on this thread (all zero except `status`
 * where the variant has no such field):
 *   SARDINE_ERR_TC_ORBIT_OUTSIDE_WINDOW  OrbitError::OutsideOrbitWindow
 *       { requested, first, last } (orbit.rs:58-66, raised at orbit.rs:513-517)
 *       as µs on the state vectors' clock; for terrain correction `requested`
 *       is the Newton iterate that fell furthest outside the window (the
 *       reference reports whichever pixel its threads reach first);
 *   SARDINE_ERR_TC_ORBIT_INSUFFICIENT    OrbitError::InsufficientCoverage
 *       (orbit.rs:54-55, raised at orbit.rs:503-511): vectors present/needed;
 *   SARDINE_ERR_TC_DEM                   TerrainCorrectionError::Dem(DemError)
 *       (terrain_correction.rs:51, dem.rs:54-109): dem_status names the
 *       DemError variant (SARDINE_ERR_DEM_*), with BadSize's byte counts. */
typedef struct sardine_error_detail {
    int32_t status;
    int32_t dem_status;
    int64_t requested_us, first_us, last_us;
    int32_t have_vectors, need_vectors;
    uint64_t dem_expected_bytes, dem_actual_bytes;
} sardine_error_detail;
sardine_status sardine_last_error_detail(sardine_error_detail* out);
int sardine_abi_version(void);
/* Number of visible gfx9xx devices (0 when none); never initialises kernels. */
int sardine_device_count(void);

/* ── Geometry / orbit types ────────────────────────────────────────────── */

/* StateVector, sardine/src/types.rs:142 */
typedef struct sardine_state_vector {
    int64_t time_us;        /* on the caller's epoch */
    double position_m[3];   /* ECEF */
    double velocity_m_s[3]; /* ECEF */
} sardine_state_vector;

/* OrbitData, sardine/src/types.rs:157 (state vectors sorted by time) */
typedef struct sardine_orbit {
    const sardine_state_vector* state_vectors; /* host */
    int32_t count;
} sardine_orbit;

/* GeolocationGridPoint, sardine/src/types.rs:171 — the two fields the
 * azimuth-time seed LUT reads (terrain_correction.rs:1080-1104). */
typedef struct sardine_grid_point {
    double latitude_deg;
    int64_t azimuth_time_us;
} sardine_grid_point;

/* BoundingBox, sardine/src/types.rs:267 */
typedef struct sardine_bbox {
    double min_lat_deg, max_lat_deg, min_lon_deg, max_lon_deg;
} sardine_bbox;

/* The parts of SceneMetadata (types.rs:406-451) terrain_correction reads:
 * the orbit, the bounding box, the first subswath's azimuth time interval
 * (terrain_correction.rs:707-711) and the seed-LUT grid (the caller passes the
 * IW2 grid when present, else the first one — the selection rule of
 * terrain_correction.rs:1087-1091). */
typedef struct sardine_scene {
    sardine_orbit orbit;
    sardine_bbox bounding_box;
    double azimuth_time_interval_s;
    const sardine_grid_point* grid_points; /* host */
    int32_t n_grid_points;
    int32_t n_subswaths; /* 0 → SARDINE_ERR_TC_NO_SUBSWATHS */
} sardine_scene;

/* A raster-backed RadarImage (pipeline_options.rs:100-135), i.e. the
 * MergedSigma0 buffers (merge_subswaths.rs:106-162).  Opaque per-sample
 * implementors have no GPU form and are rejected by construction. */
typedef struct sardine_radar_image {
    const float* data;  /* [device] lines*samples, row-major */
    const float* nesz;  /* [device] same shape, or NULL (nesz_at → None) */
    int64_t lines;
    int64_t samples;
    int64_t azimuth_start_time_us;
    double near_slant_range_time_s;
    double range_pixel_spacing_m;
} sardine_radar_image;

/* A geographic DEM raster with the GLO-30 tile sampling rule
 * (dem.rs:351-372 + bilinear_on_grid dem.rs:385-409): origin is the centre of
 * pixel (0,0), rows run south.  NaN cells are voids. */
/* A DemMosaic of GLO-30 tiles (dem.rs:559-747) held as one raster: the
 * tiles, all of one size and pixel size, sit in a grid of `rows` x `cols`
 * blocks of the raster (north row first, west column first), each block at
 * its tile's own origin.  The lookup is DemMosaic::elevation_at_linear's:
 * the first tile in load order (`rank`) whose half-pixel-margined extent
 * holds the point and whose four taps are all finite answers, with its own
 * edge clamping (Glo30Tile::elevation_at, dem.rs:351-372); no tile → no
 * height.  Adjacent tile rows (columns) sit one tile apart, north to south
 * (west to east), to a quarter pixel; the call rejects other tilings.
 * (elevation_at, the indexed form, tries the thread's last-hit tile first
 * and so may answer from the other tile on the line where two tiles'
 * extents meet, both bounds being inclusive; the linear form is
 * deterministic.) */
#define SARDINE_DEM_MAX_TILES_AXIS 8
typedef struct sardine_dem_tiling {
    int32_t rows, cols;               /* tile grid, each 1..SARDINE_DEM_MAX_TILES_AXIS */
    int32_t tile_height, tile_width;  /* pixels per tile (>= 2) */
    double origin_lat_deg[SARDINE_DEM_MAX_TILES_AXIS]; /* NW pixel centre latitude of each tile row */
    double origin_lon_deg[SARDINE_DEM_MAX_TILES_AXIS]; /* NW pixel centre longitude of each tile column */
    int32_t rank[SARDINE_DEM_MAX_TILES_AXIS * SARDINE_DEM_MAX_TILES_AXIS]; /* load order of tile
                                      [row * 8 + col]; -1: no tile there (NaN block) */
} sardine_dem_tiling;

typedef struct sardine_dem_raster {
    const float* data; /* [device] height*width */
    int64_t width;
    int64_t height;
    double origin_lat_deg;
    double origin_lon_deg;
    double pixel_deg;
    /* NULL: one raster sampled by the tile rule (dem.rs:351-372) over its
     * whole extent; else the mosaic described above (height = rows x
     * tile_height, width = cols x tile_width). */
    const sardine_dem_tiling* tiling;
} sardine_dem_raster;

/* GeoidModel, geoid.rs:342-380 */
typedef enum sardine_geoid_kind {
    SARDINE_GEOID_ZERO = 0,
    SARDINE_GEOID_GRID_2P5 = 1 /* Egm96 / Egm2008 73x144 2.5° grid, geoid.rs:220-251 */
} sardine_geoid_kind;

typedef struct sardine_geoid {
    int32_t kind;
    const float* grid_73x144; /* host, row 0 = 90°N, col 0 = 0°E; NULL for ZERO */
} sardine_geoid;

/* OutputCrs, output_crs.rs:70-96 (WGS84 and UTM are supported) */
typedef enum sardine_crs_kind {
    SARDINE_CRS_WGS84 = 0,
    SARDINE_CRS_UTM_NORTH = 1,
    SARDINE_CRS_UTM_SOUTH = 2
} sardine_crs_kind;

/* ResamplingKernel, pipeline_options.rs:74-90 */
typedef enum sardine_resampling {
    SARDINE_RESAMPLE_BILINEAR = 0,
    SARDINE_RESAMPLE_BICUBIC = 1,
    SARDINE_RESAMPLE_LANCZOS3 = 2
} sardine_resampling;

/* Arithmetic of the zero-Doppler / orbit / geometry part. */
typedef enum sardine_precision {
    SARDINE_PRECISION_FP64 = 0,        /* the reference's arithmetic */
    SARDINE_PRECISION_FP32_CENTRED = 1, /* scene-centred FP32 variant */
    /* FP64 orbit, Newton, slant range and image coordinates; FP32 resampling
     * weights and sums, NESZ lookup, and terrain geometry (normal, look
     * vector, cos LIA, flattening weight from FP64 ECEF differences).
     * sardine_terrain_correction only (not batched or served jobs). */
    SARDINE_PRECISION_MIXED = 2
} sardine_precision;

/* TerrainCorrectionConfig, terrain_correction.rs:100-235, plus the two env
 * knobs of terrain_correction.rs:1328-1378 and an optional AOI window. */
typedef struct sardine_tc_config {
    double pixel_spacing;          /* degrees for WGS84, metres for UTM */
    int32_t crs;                   /* sardine_crs_kind */
    int32_t utm_zone;              /* 1..60 for UTM */
    int32_t max_newton_iterations; /* reference default 10 */
    double newton_tolerance_s;     /* reference default 1e-6 */
    int32_t flatten;
    int32_t compute_lia;
    float noise_floor_margin_db;   /* -INFINITY disables NESZ masking */
    int32_t resampling;            /* sardine_resampling */
    int32_t projector_anchors;     /* SARDINE_TC_PROJECTOR_ANCHORS; 0 → 256 */
    int32_t precision;             /* sardine_precision */
    sardine_geoid geoid;
    /* AOI window on the full output grid: rows [aoi_row0, aoi_row0+aoi_rows),
     * cols [aoi_col0, aoi_col0+aoi_cols).  aoi_rows == 0 → the full grid.
     * The per-row seed chain starts at aoi_col0 (the reference has no AOI;
     * its full-grid result equals aoi = full grid). */
    int64_t aoi_row0, aoi_col0, aoi_rows, aoi_cols;
} sardine_tc_config;

/* Fills cfg with TerrainCorrectionConfig::new(GeoidModel::Zero) defaults. */
void sardine_tc_config_default(sardine_tc_config* cfg);

#define SARDINE_NEWTON_HIST_SIZE 16 /* NEWTON_ITER_HIST_SIZE, terrain_correction.rs:95 */

/* mask_code, terrain_correction.rs:346-376 */
enum {
    SARDINE_MASK_VALID = 0,
    SARDINE_MASK_OUTSIDE_FOOTPRINT = 1,
    SARDINE_MASK_DEM_MISSING = 2,
    SARDINE_MASK_NON_CONVERGED = 3,
    SARDINE_MASK_DEGENERATE_GEOMETRY = 4,
    SARDINE_MASK_SHADOW = 5,
    SARDINE_MASK_LAYOVER = 6,
    SARDINE_MASK_FORESHORTENING = 7,
    SARDINE_MASK_NOISE_MASKED = 8,
    SARDINE_MASK_DEM_NEIGHBOUR_MISSING = 9
};

/* GeocodedImage, terrain_correction.rs:245-307.  Buffers are caller-owned
 * device memory of rows*cols elements (AOI rows*cols when an AOI is set);
 * cos_lia / mask are written only when compute_lia (may be NULL otherwise). */
typedef struct sardine_geocoded_image {
    float* data;     /* [device] */
    float* cos_lia;  /* [device] or NULL */
    uint8_t* mask;   /* [device] or NULL */
    /* [device] or NULL: per pixel (line, sample) radar coordinates of every
     * converged pixel, NaN elsewhere (a geolocation diagnostic with no
     * reference counterpart; 2*rows*cols doubles). */
    double* radar_coords;
    int64_t rows, cols;                 /* filled by the call */
    int64_t full_rows, full_cols;       /* full-grid shape */
    double geotransform[6];             /* of the written window */
    uint64_t total_pixel_count;
    uint64_t valid_pixel_count;
    uint64_t dem_missing_count;
    uint64_t outside_footprint_count;
    uint64_t non_converged_count;
    uint64_t degenerate_geometry_count;
    uint64_t flat_masked_count;
    uint64_t noise_masked_count;
    uint64_t newton_iter_histogram[SARDINE_NEWTON_HIST_SIZE];
    /* diagnostics of the GPU seed-chain emulation (no reference counterpart) */
    uint64_t chain_redo_rounds;
    /* diagnostics of the Newton steps whose µs truncation or convergence test
     * lay within the band where the kernels' polynomial orbit cannot decide
     * it and were re-evaluated with the reference's own arithmetic (counted
     * over every chain the kernels run, speculative and replayed ones
     * included): how many, how many decisions that changed, and the largest
     * difference in δ·1e6 (µs) between the two evaluations */
    uint64_t tie_steps;
    uint64_t tie_decisions_changed;
    double tie_max_dev_us;
} sardine_geocoded_image;

/* Output grid of terrain_correction: compute_output_grid_extent
 * (terrain_correction.rs:1174-1216) + n_rows/n_cols (:1310-1317) + the
 * geotransform (:1805-1812).  Host-only, no device needed. */
sardine_status sardine_tc_output_grid(const sardine_scene* scene, const sardine_tc_config* cfg,
                                      int64_t* rows, int64_t* cols, double geotransform[6]);

/* interpolate_orbit_and_accel (orbit.rs:498-596): position, velocity and
 * acceleration at time_us (absolute µs, the state vectors' clock), evaluated
 * on the host by the same code the terrain-correction kernels run (the
 * per-interval polynomial form, the reference's values at node times).
 * Errors as the reference: SARDINE_ERR_TC_ORBIT_INSUFFICIENT below 8
 * vectors, SARDINE_ERR_TC_ORBIT_OUTSIDE_WINDOW outside the orbit.
 * Host-only, no device needed. */
sardine_status sardine_orbit_interpolate(const sardine_orbit* orbit, int64_t time_us, double pos[3],
                                         double vel[3], double acc[3]);

/* terrain_correction(merged, scene, dem, grids, config) -> GeocodedImage,
 * terrain_correction.rs:1245-1836.  Synchronous: returns after the counters
 * are on the host.  Output buffers must hold out->rows*out->cols elements of
 * the (AOI) window; query the shape with sardine_tc_output_grid. */
sardine_status sardine_terrain_correction(const sardine_radar_image* merged,
                                          const sardine_scene* scene,
                                          const sardine_dem_raster* dem,
                                          const sardine_tc_config* cfg,
                                          sardine_geocoded_image* out,
                                          void* stream);

/* sardine_terrain_correction without device memory allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and results as sardine_terrain_correction;
 * every device buffer of the call is a slice of a fixed 64 MiB scratch array
 * that each GPU holds until the process ends (the caller still passes the
 * inputs and outputs).  A call whose work needs more scratch (chain records,
 * row seeds, the tie list and anchors grow with the AOI: about 0.75-3 bytes
 * per pixel plus 0.1 KiB per row, 4 KiB per UTM row and 16 bytes per UTM
 * column) returns
 * SARDINE_ERR_CAPACITY.  All calls on one GPU share its
 * array: calls that overlap on one GPU (from several threads or streams)
 * race; calls in sequence do not, since each returns after its work. */
sardine_status sardine_terrain_correction_noalloc(const sardine_radar_image* merged,
                                                  const sardine_scene* scene,
                                                  const sardine_dem_raster* dem,
                                                  const sardine_tc_config* cfg,
                                                  sardine_geocoded_image* out,
                                                  void* stream);

/* ── Speckle ────────────────────────────────────────────────────────────── */

/* SpeckleFilter, speckle.rs:61-91 */
typedef enum sardine_speckle_kind {
    SARDINE_SPECKLE_BOXCAR = 0,
    SARDINE_SPECKLE_LEE = 1,
    SARDINE_SPECKLE_FROST = 2,
    SARDINE_SPECKLE_GAMMA_MAP = 3,
    SARDINE_SPECKLE_REFINED_LEE = 4
} sardine_speckle_kind;

typedef struct sardine_speckle_filter {
    int32_t kind;
    int32_t window;  /* odd ≥ 3; ignored for REFINED_LEE (fixed 7) */
    float enl;       /* LEE, GAMMA_MAP, REFINED_LEE */
    float damping;   /* FROST */
} sardine_speckle_filter;

/* SpeckleKernel::apply(data, cols, rows) -> Vec<f32> (speckle.rs:586-603,
 * apply_speckle_filter speckle.rs:131-180).  Out-of-place: `out` must not
 * alias `data`.  `data_len` is the element count of `data` (the reference's
 * DimensionMismatch check). */
sardine_status sardine_speckle_apply(const float* data /*[device]*/, int64_t data_len,
                                     int64_t cols, int64_t rows,
                                     const sardine_speckle_filter* filter,
                                     float* out /*[device]*/, void* stream);

/* ── Deburst / calibration / merge ─────────────────────────────────────── */

/* BurstEntry, types.rs:351-395 (fields the deburst reads) */
typedef struct sardine_burst {
    int64_t azimuth_time_us;
    int64_t first_line;          /* TIFF row of the burst's line 0 */
    int64_t first_valid_sample;
    int64_t last_valid_sample;   /* exclusive */
    int32_t burst_index;
} sardine_burst;

/* SubSwathMetadata, types.rs:281-337 (fields the path reads) */
typedef struct sardine_subswath {
    int32_t burst_count;
    int64_t lines_per_burst;
    double azimuth_time_interval_s;
    double slant_range_time_s;
    double range_pixel_spacing_m;
} sardine_subswath;

/* Result of deburst_subswath's integer layout (deburst.rs:230-344). */
#define SARDINE_MAX_BURSTS 64
typedef struct sardine_deburst_layout {
    int64_t lines;                /* total debursted lines */
    int64_t samples;              /* valid_cols */
    int64_t valid_sample_offset;  /* first_valid */
    int32_t n_bursts;
    int64_t overlap[SARDINE_MAX_BURSTS];      /* overlap[k] between burst k and k+1 */
    int64_t local_first[SARDINE_MAX_BURSTS];  /* ⌈ov/2⌉ rule */
    int64_t local_end[SARDINE_MAX_BURSTS];    /* lps − ⌊ov/2⌋ rule */
    int64_t out_line_start[SARDINE_MAX_BURSTS];
} sardine_deburst_layout;

/* compute_overlap_lines + the contribution ranges of deburst_subswath
 * (deburst.rs:172-305).  Host-only; integer results are bit-exact. */
sardine_status sardine_deburst_layout_compute(const sardine_subswath* sw,
                                              const sardine_burst* bursts, int32_t n_bursts,
                                              int64_t tiff_height,
                                              sardine_deburst_layout* out);

/* One LUT vector list: CalibrationVector.sigma_nought or
 * NoiseRangeVector.noise_range_lut (calibration.rs:38-82), flattened. */
typedef struct sardine_lut {
    int32_t n_vectors;
    const int32_t* lines;        /* host, n_vectors (sorted ascending) */
    const int64_t* offsets;      /* host, n_vectors+1 into pixels/values */
    const uint32_t* pixels;      /* host */
    const float* values;         /* host */
} sardine_lut;

/* NoiseAzimuthVector (calibration.rs:86-96); n == 0 → factor 1.0 */
typedef struct sardine_az_lut {
    int32_t n;
    const int32_t* lines;   /* host */
    const float* values;    /* host */
} sardine_az_lut;

/* One subswath input to the fused deburst+calibrate+merge pass. */
typedef struct sardine_swath_input {
    const int16_t* raw;          /* [device] CInt16 TIFF rows: tiff_height x raw_cols x 2 */
    int64_t raw_cols;
    int64_t tiff_height;
    sardine_subswath swath;
    const sardine_burst* bursts; /* host, sorted by burst_index */
    int32_t n_bursts;
    sardine_lut cal;             /* sigma_nought LUT */
    sardine_lut noise_range;
    sardine_az_lut noise_azimuth;
} sardine_swath_input;

/* MergedSigma0 metadata (merge_subswaths.rs:106-162) + the merge layout. */
#define SARDINE_MAX_SWATHS 8
typedef struct sardine_merged_info {
    int64_t lines, samples;
    double near_slant_range_time_s;
    double range_pixel_spacing_s;
    double range_pixel_spacing_m;
    int64_t azimuth_start_time_us;
    uint32_t cal_lut_extrapolation_gap_px;
    uint32_t noise_lut_extrapolation_gap_px;
    int32_t n_swaths;
    int64_t sw_out_offset[SARDINE_MAX_SWATHS];
    int64_t seam[SARDINE_MAX_SWATHS];          /* seam[i] between swath i and i+1 */
    int64_t swath_lines[SARDINE_MAX_SWATHS];   /* debursted lines per swath */
    int64_t swath_samples[SARDINE_MAX_SWATHS];
    int64_t swath_valid_offset[SARDINE_MAX_SWATHS];
} sardine_merged_info;

/* Shape + layout only (no device work): deburst_subswath layout per swath,
 * apply_calibration's LUT coverage checks (apply_calibration.rs:421-462) and
 * merge_subswaths' offsets/seams (merge_subswaths.rs:207-330). */
sardine_status sardine_prepare_merged_layout(const sardine_swath_input* swaths, int32_t n_swaths,
                                             sardine_merged_info* info);

/* prepare_merged_scene's deburst → apply_calibration → merge chain
 * (scene_prep.rs:598-699; deburst.rs:230, apply_calibration.rs:374,
 * merge_subswaths.rs:207) fused into one HBM pass: the deburst row copy and
 * the merge column copy become output addressing of the calibration kernel.
 * sigma0/nesz [device] hold info.lines*info.samples floats. */
sardine_status sardine_prepare_merged(const sardine_swath_input* swaths, int32_t n_swaths,
                                      float* sigma0, float* nesz,
                                      sardine_merged_info* info, void* stream);

/* apply_calibration(deburst, cal, noise, tiff_line_origin) -> Sigma0Array
 * (apply_calibration.rs:374-572) on an already debursted CInt16 array.
 * out/nesz [device] hold lines*samples floats. */
sardine_status sardine_apply_calibration(const int16_t* deburst /*[device] lines*samples*2*/,
                                         int64_t lines, int64_t samples,
                                         int64_t valid_sample_offset,
                                         const sardine_lut* cal, const sardine_lut* noise_range,
                                         const sardine_az_lut* noise_azimuth,
                                         int64_t tiff_line_origin,
                                         float* out, float* nesz,
                                         uint32_t* cal_gap_px, uint32_t* noise_gap_px,
                                         void* stream);

/* ── Geoid grid files ───────────────────────────────────────────────────── */
/* The 73 x 144 2.5-degree undulation grid the terrain correction samples
 * (sardine_geoid kind SARDINE_GEOID_GRID_2P5): row 0 = 90N, col 0 = 0E,
 * metres.  EGM96 and EGM2008 share layout and interpolation; the model only
 * selects the size error's variant.  Host only. */
typedef enum sardine_geoid_model {
    SARDINE_GEOID_MODEL_EGM96 = 0,
    SARDINE_GEOID_MODEL_EGM2008 = 1
} sardine_geoid_model;

/* Egm96Grid::load_binary / Egm2008Grid::load_binary (geoid.rs:132-160,
 * 276-299): exactly 10512 little-endian float32 values. */
sardine_status sardine_geoid_load_binary(const char* path, int32_t model, float grid[10512]);

/* Egm96Grid::from_ww15mgh (geoid.rs:164-217): the NGA ASCII grid (up to two
 * header lines, then 10512 values). */
sardine_status sardine_geoid_load_ww15mgh(const char* path, float grid[10512]);

/* ── SRTM 1" tiles → one device DEM raster ─────────────────────────────── */

/* DemMosaic::load_srtm_files (dem.rs:628-640; SrtmTile::load dem.rs:129-178)
 * as one raster for the terrain correction: each `.hgt` file (named by its
 * SW corner, N50E010.hgt; 3601 x 3601 big-endian int16, -32768 = void) lands
 * at its place in the bounding rectangle of all tiles, neighbours sharing
 * their edge row/column; voids and cells no tile covers are NaN (the
 * reference counts both as dem_missing).  The raster: origin = the NW
 * corner (the first sample), pixel_deg = 1/3600.  With dst NULL only `out`
 * is filled (host only); else dst [device] holds out->width * out->height
 * floats and out->data = dst.  Errors: SARDINE_ERR_DEM_NO_TILES,
 * _UNRECOGNISED_TILE_NAME, _IO, _BAD_SIZE in the reference's order. */
sardine_status sardine_dem_srtm_mosaic(const char* const* paths, int32_t n_paths, float* dst /*[device]*/,
                                       sardine_dem_raster* out, void* stream);

/* Glo30Tile::load (dem.rs:248-352): one Copernicus GLO-30 GeoTIFF tile
 * (float32, int16 or uint16; any of the usual compressions and predictors)
 * as a device DEM raster.  Georeferencing from ModelPixelScale +
 * ModelTiepoint (the NW pixel centre; pixel size the mean of the two
 * scales), else from the file name (Copernicus_DSM_COG_10_N50_00_E010_00_DEM
 * or N50E010); GDAL_NODATA and non-finite floats become NaN.  The tile's
 * sampling rule is the terrain correction's DEM rule, so the raster is the
 * reference's tile exactly.  With dst NULL only `out` is filled from the
 * header (host only).  Errors: SARDINE_ERR_DEM_IO, _TIFF_DECODE,
 * _UNRECOGNISED_TILE_NAME, _UNSUPPORTED_PIXEL_TYPE in the reference's order. */
sardine_status sardine_dem_glo30_tile(const char* path, float* dst /*[device]*/, sardine_dem_raster* out,
                                      void* stream);

/* DemMosaic::load_glo30_files (dem.rs:642-653) + elevation_at_linear: the
 * tiles of `paths`, in that order (the load order decides which tile answers
 * where two meet; load_directory sorts the paths), decoded as
 * sardine_dem_glo30_tile does and placed into one raster dst of
 * out->height x out->width with the grid in *tiling (out->tiling = tiling).
 * The tiles must share size and pixel size, tiles of one grid row one origin
 * latitude and of one grid column one origin longitude, adjacent rows and
 * columns must abut (origins one tile apart), at most one tile per cell,
 * at most 8 x 8 cells; cells without a tile are NaN.  With dst NULL only
 * `out` and `tiling` are filled from the headers (host only).  Errors: those
 * of sardine_dem_glo30_tile per file in order, SARDINE_ERR_DEM_NO_TILES for
 * an empty list, SARDINE_ERR_INVALID_ARGUMENT for tiles that do not form
 * such a grid. */
sardine_status sardine_dem_glo30_mosaic(const char* const* paths, int32_t n_paths, float* dst /*[device]*/,
                                        sardine_dem_raster* out, sardine_dem_tiling* tiling, void* stream);

/* ── SLC measurement TIFF → HBM ──────────────────────────────────────────── */

/* The layout SlcReader::open validates (slc_reader.rs:268-461): a classic
 * TIFF, uncompressed CInt16 (SampleFormat 5, 32 bits, 1 sample per pixel),
 * one row per strip, contiguous strips (first and last offsets checked). */
typedef struct sardine_slc_tiff {
    uint32_t width, height;
    int32_t little_endian;
    int32_t pad;
    uint64_t base_offset;  /* file offset of row 0 */
    uint64_t row_bytes;    /* width * 4 */
} sardine_slc_tiff;

/* SlcReader::open (slc_reader.rs:504-515): parses and validates the IFD.
 * Host only.  Errors: SARDINE_ERR_SLC_* as parse_ifd raises them. */
sardine_status sardine_slc_open(const char* path, sardine_slc_tiff* tiff);

/* SlcReader::read_burst_raw (slc_reader.rs:532-575) into device memory:
 * line_count rows from first_line, as CInt16 [i, q] pairs in host order, to
 * dst [device] (line_count * width * 2 int16).  The rows are read with
 * parallel preads into pinned chunks, each set copied to HBM while the next
 * is read; big-endian files are byte-swapped on the device.  Returns after
 * the file reads; the copies are ordered on `stream`.  The line range is
 * checked first (SARDINE_ERR_SLC_LINE_OUT_OF_BOUNDS). */
sardine_status sardine_slc_read_lines(const char* path, const sardine_slc_tiff* tiff, int64_t first_line,
                                      int64_t line_count, int16_t* dst /*[device]*/, void* stream);

/* MultiSlcReader (multi_slc_reader.rs:75-193): one swath stored as
 * n_slices consecutive measurement TIFFs (slice order), read in its logical
 * line space (the slices' heights summed).  The read must lie inside one
 * slice.  Errors: SARDINE_ERR_SLC_MULTI_EMPTY_INPUT, _MULTI_WIDTH_MISMATCH
 * (construction), then _LINE_OUT_OF_BOUNDS, _CROSS_SLICE_BOUNDARY and the
 * single-slice read's errors. */
sardine_status sardine_slc_read_lines_multi(const char* const* paths, const sardine_slc_tiff* tiffs,
                                            int32_t n_slices, int64_t first_line, int64_t line_count,
                                            int16_t* dst /*[device]*/, void* stream);

/* ── Pre-TC multilook ─────────────────────────────────────────────────────── */

/* apply_multilook(merged, range_looks, azimuth_looks, ati) (scene_prep.rs:
 * 1060-1161): NaN-aware boxcar decimation of σ⁰ and NESZ (NESZ averaged over
 * the non-NaN σ⁰ samples) to floor(lines / azimuth_looks) x floor(samples /
 * range_looks), with the centre-aligned geometry update written to `out`
 * (the per-swath fields are copied unchanged; they describe the input
 * columns).  1 x 1 returns the input unchanged (copied when the output
 * buffers differ).  With out_data and out_nesz both NULL only `out` is
 * computed (host only: the caller's shape query).  `out` may alias `in`.
 * Errors (anyhow bails in the reference): SARDINE_ERR_INVALID_ARGUMENT for a
 * zero factor, factors larger than the image, and a non-finite or
 * non-positive ATI with azimuth_looks > 1. */
sardine_status sardine_apply_multilook(const float* data /*[device]*/, const float* nesz /*[device]*/,
                                       const sardine_merged_info* in, int32_t range_looks, int32_t azimuth_looks,
                                       double azimuth_time_interval_s, float* out_data /*[device]*/,
                                       float* out_nesz /*[device]*/, sardine_merged_info* out, void* stream);

/* ── Ground-range (GRD) projection ──────────────────────────────────────── */

/* One geolocation grid point as the ground-range projection reads it
 * (GeolocationGridPoint, types.rs: line, slant_range_time_s,
 * incidence_angle_deg). */
typedef struct sardine_grd_grid_point {
    double slant_range_time_s;
    double incidence_angle_deg;
    uint32_t line;
    uint32_t pad;
} sardine_grd_grid_point;

/* GroundRangeImage (ground_range.rs:54-76) without its raster. */
typedef struct sardine_grd_image {
    int64_t lines, samples;
    double range_pixel_spacing_m;
    double azimuth_pixel_spacing_m;
    int32_t range_looks;   /* always 1 */
    int32_t azimuth_looks;
} sardine_grd_image;

/* to_ground_range(merged, grids, target_spacing_m, azimuth_pixel_spacing_m)
 * (ground_range.rs:303-442).  The merged image is `data` [device] with the
 * size and range timing of `merged` (lines, samples, near_slant_range_time_s,
 * range_pixel_spacing_s, range_pixel_spacing_m).  The grids are n_swaths
 * consecutive runs of `points`, counts[i] points each.  The incidence profile
 * (mid-azimuth grid row per swath, ground_range.rs:103-259), the cumulative
 * ground range and the column map (ground_range.rs:446-490) are host work;
 * the device resamples every line onto the ground-range columns and averages
 * azimuth_looks lines in one pass.  `out` [device] holds image->lines x
 * image->samples floats; with out NULL only `image` is computed (host only).
 * Errors: SARDINE_ERR_GRD_* as the reference raises them, in its order. */
sardine_status sardine_to_ground_range(const float* data /*[device]*/, const sardine_merged_info* merged,
                                       const sardine_grd_grid_point* points, const int32_t* counts,
                                       int32_t n_swaths, double target_spacing_m, double azimuth_pixel_spacing_m,
                                       float* out /*[device]*/, sardine_grd_image* image, void* stream);

/* ── dB epilogue ────────────────────────────────────────────────────────── */

/* to_db_inplace(data, noise_floor_linear) (export.rs:59-94).  counts[0] =
 * converted, counts[1] = noise-masked. */
sardine_status sardine_to_db_inplace(float* data /*[device]*/, int64_t n, float noise_floor_linear,
                                     uint64_t counts[2], void* stream);

/* ── Cloud-Optimised GeoTIFF export (export.rs:329-760) ───────────────────
 * The file is the reference's byte for byte: a classic TIFF (BigTIFF when
 * rows*cols*4 exceeds 2^32-1), one 17-entry IFD per level with the coarsest
 * overview first and the full resolution last, then the pixel-scale,
 * tiepoint and GeoKey blocks, then every level's tiles (row-major tile
 * order, each tile tile_size^2 float32 values, padded with NaN).  Overviews
 * are 2x2 means of the non-NaN pixels (downsample_2x, export.rs:362-392),
 * halved until both sides are <= tile_size; there is always at least one.
 * The pyramid and the tile packing run on the device; the host writes the
 * header it computes and streams the packed tiles to the file. */
#define SARDINE_COG_MAX_LEVELS 32

typedef struct sardine_cog_level {
    int64_t cols, rows;            /* raster size of the level */
    int64_t tiles_x, tiles_y;      /* ceil(cols / tile_size), ceil(rows / tile_size) */
    uint64_t ifd_offset;           /* file offset of the level's IFD */
    uint64_t tile_offsets_offset;  /* file offset of its TileOffsets array */
    uint64_t tile_counts_offset;   /* file offset of its TileByteCounts array */
    uint64_t tile_data_offset;     /* file offset of its first tile */
} sardine_cog_level;

typedef struct sardine_cog_layout {
    int32_t n_levels;              /* overviews + 1; level[n_levels-1] is the full raster */
    int32_t bigtiff;               /* 1: BigTIFF, 0: classic TIFF */
    uint32_t tile_size;
    int32_t geo_key_shorts;        /* 16 (WGS84) or 20 (UTM) */
    uint64_t header_bytes;         /* bytes before the first tile (= level[0].tile_data_offset) */
    uint64_t tile_bytes;           /* bytes of all tiles of all levels */
    uint64_t file_bytes;           /* header_bytes + tile_bytes */
    uint64_t pixel_scale_offset, tiepoint_offset, geo_keys_offset;
    sardine_cog_level level[SARDINE_COG_MAX_LEVELS];
} sardine_cog_layout;

/* The pyramid, the TIFF flavour and every offset of write_cog_raw_inner
 * (export.rs:430-535) for a cols x rows raster.  Host only.  Errors:
 * SARDINE_ERR_EXPORT_INVALID_TILE_SIZE (not a power of two in [64, 4096]),
 * SARDINE_ERR_INVALID_ARGUMENT (negative size, unknown CRS). */
sardine_status sardine_cog_layout_compute(int64_t cols, int64_t rows, uint32_t tile_size, int32_t crs,
                                          sardine_cog_layout* layout);

/* The header region of the file (header, IFDs, tile arrays, GeoTIFF blocks;
 * export.rs:537-670): layout->header_bytes bytes into `out`.  Host only.
 * crs is a sardine_crs_kind, utm_zone 1..60 for UTM (OutputCrs::epsg and
 * geo_key_directory, output_crs.rs:167-175,265-287).  Errors:
 * SARDINE_ERR_EXPORT_INVALID_GEOTRANSFORM (pixel width or -gt[5] not finite
 * and > 0), SARDINE_ERR_INVALID_ARGUMENT (capacity, CRS, zone). */
sardine_status sardine_cog_header(const sardine_cog_layout* layout, const double geotransform[6], int32_t crs,
                                  int32_t utm_zone, uint8_t* out, uint64_t capacity);

/* The tile region of the file (export.rs:672-712, with downsample_2x for
 * every overview): layout->tile_bytes bytes at `tiles` [device, 16-byte
 * aligned], from the full raster `data` [device, layout's cols x rows,
 * row-major].  Enqueued on `stream`; returns without waiting. */
sardine_status sardine_cog_pack(const float* data, const sardine_cog_layout* layout, float* tiles, void* stream);

/* write_cog_with_crs (export.rs:329-354): validation in the reference's order
 * (geotransform, then tile size), the layout, the device pack, and the file
 * at `path`.  Returns after the file is closed.  I/O failures return
 * SARDINE_ERR_EXPORT_IO. */
sardine_status sardine_write_cog_with_crs(const char* path, const float* data /*[device]*/, int64_t cols,
                                          int64_t rows, const double geotransform[6], int32_t crs,
                                          int32_t utm_zone, uint32_t tile_size, void* stream);

/* ── Striped GeoTIFF / BigTIFF export (export.rs:139-284, 762-1219) ──────
 * One IFD, 64 rows per strip, the raster bytes as they lie in memory after
 * the header; BigTIFF when needs_bigtiff's estimate (export.rs:762-786)
 * exceeds 2^32-1.  The device raster is streamed to the file through two
 * pinned chunks (no kernel: the file holds the raster unchanged).  Files are
 * the reference's byte for byte.  A raster needs at least one row (the
 * reference's strip arithmetic underflows on zero rows). */

/* The bytes before the raster (*size of them; copied to `out` when it is
 * non-null and `capacity` holds them).  bytes_per_pixel 4 (float32) or 1
 * (uint8); geotransform NULL = write_tiff_no_crs's plain TIFF.  Host only. */
sardine_status sardine_geotiff_header(int64_t cols, int64_t rows, int32_t bytes_per_pixel,
                                      const double geotransform[6], int32_t crs, int32_t utm_zone, uint8_t* out,
                                      uint64_t capacity, uint64_t* size);

/* write_geotiff_with_crs (export.rs:139-187): float32 raster [device]. */
sardine_status sardine_write_geotiff_with_crs(const char* path, const float* data /*[device]*/, int64_t cols,
                                              int64_t rows, const double geotransform[6], int32_t crs,
                                              int32_t utm_zone, void* stream);

/* write_geotiff_u8_with_crs (export.rs:204-245): uint8 raster [device], e.g.
 * the terrain-correction quality mask. */
sardine_status sardine_write_geotiff_u8_with_crs(const char* path, const uint8_t* data /*[device]*/, int64_t cols,
                                                 int64_t rows, const double geotransform[6], int32_t crs,
                                                 int32_t utm_zone, void* stream);

/* write_tiff_no_crs (export.rs:252-284): float32 raster [device], no
 * GeoTIFF tags (radar-geometry rasters). */
sardine_status sardine_write_tiff_no_crs(const char* path, const float* data /*[device]*/, int64_t cols,
                                         int64_t rows, void* stream);

/* ── Kernel timing (diagnostic, no reference counterpart) ─────────────────
 * With profiling on, every compute entry point brackets its main kernel with
 * HIP events on the stream it launches on (without waiting); the duration of
 * the last such kernel on this thread is then available by kernel family. */
typedef enum sardine_kernel_family {
    SARDINE_KERNEL_TC = 0,        /* terrain-correction kernel */
    SARDINE_KERNEL_CALIBRATE = 1, /* fused deburst + calibration + merge kernel */
    SARDINE_KERNEL_SPECKLE = 2,
    SARDINE_KERNEL_TO_DB = 3,
    SARDINE_KERNEL_TC_SEGMENT = 4, /* the per-pixel chain kernel of terrain correction alone */
    SARDINE_KERNEL_COG_PACK = 5,   /* the COG export's pyramid + tile pack (all its launches) */
    SARDINE_KERNEL_MULTILOOK = 6,
    SARDINE_KERNEL_GRD = 7,
    SARDINE_KERNEL_DEM = 8 /* a DEM loader's device decode (one entry per SRTM tile; the GLO-30 sample rule) */
} sardine_kernel_family;
/* Switches the kernel timers of the calling thread on (creating their
 * events) or off. */
void sardine_set_profiling(int on);
/* Number of output columns one lane chains sequentially in the
 * terrain-correction kernel (0 = automatic, the default; at most 255).  A
 * tuning knob with no effect on results: the seed-chain fix-up makes every
 * value exact. */
void sardine_set_tc_segment(int64_t columns);
/* Columns before each segment over which the terrain-correction kernel runs
 * a short chain to predict the segment's starting seed (default 2; 0 starts
 * every segment at the row seed).  A tuning knob with no effect on results:
 * a wrong prediction only costs the fix-up pass a replay. */
void sardine_set_tc_lead(int32_t columns);
/* Half-width (µs) of the band around a Newton step's decision points (the
 * µs truncation, the convergence test) inside which the step is decided by
 * the reference's own arithmetic instead of the kernels' (default 2^-18;
 * values outside (0, 0.5] restore it).  A tuning knob with no effect on
 * results as long as the band holds the kernels' deviation (about 3e-7 µs);
 * a wide band sends many more pixels through the exact replays. */
void sardine_set_tc_tie_band(double us);
/* Milliseconds of the last profiled kernel of `family` on this thread, -1 if
 * none; waits for that kernel to finish. */
double sardine_last_kernel_ms(int family);
/* The durations (ms) of the family's most recent launches, oldest first,
 * at most min(max, 64); returns how many were written.  The timers record
 * events without waiting, so a caller can time many calls back to back
 * and read them afterwards; reading waits for the events it reads.
 * No reference counterpart (diagnostics). */
int32_t sardine_recent_kernel_ms(int family, double* ms, int32_t max);

/* ── Resident scenes and batched geocoding (no reference counterpart: the
 * reference geocodes one scene per call; these entry points run many
 * terrain_correction + SpeckleKernel::apply jobs — AOI tiles of uneven size,
 * mixed resamplers, CRSs and filters, on one or more device-resident scenes
 * — in one set of persistent launches with the results each job would get
 * from sardine_terrain_correction and sardine_speckle_apply alone). ──── */

typedef struct sardine_resident_scene_s* sardine_resident_scene;

/* Keeps a scene's orbit and azimuth-time seed table on the device.  The
 * merged and DEM rasters stay caller-owned device memory and must outlive
 * the handle; the host metadata is copied. */
sardine_status sardine_scene_create(const sardine_radar_image* merged, const sardine_scene* scene,
                                    const sardine_dem_raster* dem, sardine_resident_scene* out);
void sardine_scene_destroy(sardine_resident_scene scene);

typedef struct sardine_batch_job {
    int32_t scene;                  /* index into the scenes passed with the batch */
    int32_t status;                 /* out: this job's sardine_status */
    sardine_tc_config config;       /* output grid, AOI window, resampler, ... */
    sardine_geocoded_image* out;    /* device buffers in; shape and counters out */
    sardine_speckle_filter speckle; /* kind < 0: no speckle step */
    float* speckle_out;             /* [device] rows*cols floats when speckle.kind >= 0 */
} sardine_batch_job;

/* Launch shape of a batch: chain segment length (0 = automatic), persistent
 * workgroups per CU of the chain pass (1: one per CU; 2 or more: as many as
 * each job class's chain kernel holds, three for bilinear and bicubic and
 * two for Lanczos-3) and of the speckle pass. */
typedef struct sardine_batch_tuning {
    int32_t segment_columns;
    int32_t tc_groups_per_cu;
    int32_t speckle_groups_per_cu;
} sardine_batch_tuning;

typedef struct sardine_batch_s* sardine_batch;

/* An executor for up to max_jobs jobs per run.  use_graph != 0 captures the
 * launch sequence once as a HIP graph (recaptured when the tuning, the
 * largest speckle window or the largest orbit changes) and replays it every
 * run. */
sardine_status sardine_batch_create(int32_t max_jobs, int32_t use_graph, sardine_batch* out);
void sardine_batch_destroy(sardine_batch batch);
sardine_status sardine_batch_set_tuning(sardine_batch batch, const sardine_batch_tuning* tuning);
sardine_status sardine_batch_get_tuning(sardine_batch batch, sardine_batch_tuning* tuning);

/* Runs every job and returns when all have finished.  A job that fails
 * validation is skipped with its status set; the others still run.  Returns
 * SARDINE_OK when every job succeeded, else the first failing job's status.
 * device_ms (optional) receives the device time of the run's launches. */
sardine_status sardine_batch_run(sardine_batch batch, const sardine_resident_scene* scenes, int32_t n_scenes,
                                 sardine_batch_job* jobs, int32_t n_jobs, void* stream, double* device_ms);

/* Project autotuner: times candidate tunings on this batch (two runs each,
 * the faster counts), keeps the fastest in the batch and returns it. */
sardine_status sardine_batch_autotune(sardine_batch batch, const sardine_resident_scene* scenes, int32_t n_scenes,
                                      sardine_batch_job* jobs, int32_t n_jobs, void* stream,
                                      sardine_batch_tuning* best, double* best_ms);

/* ── Serving (no reference counterpart) ────────────────────────────────────
 * A request loop over resident scenes.  Requests are admitted in arrival
 * order while their output and work memory fits the HBM budget, grouped
 * into a batch as they arrive (up to max_batch), and run with up to `slots`
 * batches in flight at once on separate streams, so work admitted while
 * another batch is on the GPU overlaps it.  Each result equals what the
 * per-call path gives for that request alone. */

typedef struct sardine_server_s* sardine_server;

typedef struct sardine_server_options {
    uint64_t hbm_budget_bytes; /* outputs held until release + work of batches in flight */
    int32_t max_batch;         /* requests per batch */
    int32_t slots;             /* batches in flight */
    int32_t use_graph;         /* replay each slot's launches from a HIP graph */
    int32_t pad;
} sardine_server_options;

typedef struct sardine_request {
    int32_t scene;                  /* index into the server's scenes */
    int32_t pad;
    sardine_tc_config config;       /* output grid, AOI window, resampler, ... */
    sardine_speckle_filter speckle; /* kind < 0: none */
} sardine_request;

typedef struct sardine_response {
    int32_t status;                 /* the request's sardine_status */
    int32_t done;                   /* 0 while queued or running */
    sardine_geocoded_image image;   /* data: [device] server-owned rows*cols floats until release */
    float* speckled;                /* [device] server-owned filtered data, or NULL */
    double t_submit_us, t_admit_us, t_launch_us, t_done_us; /* host steady clock */
    int32_t batch_size;             /* requests in the batch it ran in */
    int32_t pad;
} sardine_response;

sardine_status sardine_server_create(const sardine_resident_scene* scenes, int32_t n_scenes,
                                     const sardine_server_options* options, sardine_server* out);
/* Queues a request (thread-safe, never waits for the GPU). */
sardine_status sardine_server_submit(sardine_server server, const sardine_request* request, uint64_t* ticket);
/* Copies the request's response; with block != 0 waits until it is done. */
sardine_status sardine_server_wait(sardine_server server, uint64_t ticket, int32_t block, sardine_response* out);
/* Frees a finished request's outputs and forgets the ticket.  The memory is
 * returned only after all work enqueued on `stream` before this call has
 * run (the caller's reads of image.data / speckled belong there; NULL is the
 * null stream); work on other streams that reads the outputs must be
 * ordered before that point by the caller.  The pointers are invalid once
 * this returns: later requests reuse the memory. */
sardine_status sardine_server_release(sardine_server server, uint64_t ticket, void* stream);
/* Stops the loop after the requests already admitted have finished; waits
 * for all work on the device before it frees outputs still unreleased. */
void sardine_server_destroy(sardine_server server);

/* ── Diagnostics (no reference counterpart) ──────────────────────────────
 * Evaluates, on the device, element by element over n doubles (device
 * buffers): fn 0-8 the portable libm of the geodesy and projection (sin,
 * cos, exp, log, sinh, cosh, atanh, atan, atan2(y, x)), 9 x / y, 10 sqrt x.
 * For comparing the device's bits with a host evaluation of the same
 * functions.  Returns once the work is enqueued. */
sardine_status sardine_diag_math(int32_t fn, const double* x /*[device]*/, const double* y /*[device]*/,
                                 double* out /*[device]*/, int64_t n, void* stream);

/* Diagnostics, no reference counterpart: the terrain correction's DEM
 * lookup (the tile rule on one raster, dem.rs:351-409, or a GLO-30 mosaic,
 * DemMosaic::elevation_at_linear) at n points; NaN where the DEM gives no
 * height.  For comparing the device lookup with a host one. */
sardine_status sardine_diag_dem(const sardine_dem_raster* dem, const double* lat /*[device]*/,
                                const double* lon /*[device]*/, float* out /*[device]*/, int64_t n, void* stream);

/* ── Synthetic inputs (test / benchmark data, no reference counterpart) ── */

/* Fill a CInt16 raster with round(sigma * N(0,1)) real/imag parts from a
 * counter-based generator keyed on (seed, element index); identical to the
 * host generator sardine_amd.synthetic.complex_gaussian_dn. */
sardine_status sardine_synth_cint16(int16_t* out /*[device]*/, int64_t n_complex,
                                    uint64_t seed, uint64_t index_base, double sigma,
                                    void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SARDINE_HIP_H */
