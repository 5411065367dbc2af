// cal_params.hpp — kernel-argument block of the fused calibration pass.
#pragma once

#include <stdint.h>

namespace sd {

constexpr int kMaxSwaths = 8;

// Per (swath, merged line) constants, computed on the host.
struct CalLine {
    int64_t tiff_row;      // deburst: TIFF row of this debursted line
    int32_t cal_lo, cal_hi;
    int32_t nr_lo, nr_hi;
    float cal_t, nr_t;     // azimuth line fractions (apply_calibration.rs:343-344)
    float az;              // noise azimuth factor (apply_calibration.rs:535-545)
    float pad;
};

struct CalSwathDev {
    const int16_t* raw;    // CInt16 TIFF rows
    int64_t raw_cols;
    int64_t first_valid;   // TIFF column of debursted column 0
    const float* cal_dense;
    const float* nr_dense;
    int64_t dense_cols;
    const CalLine* lines;  // per merged line
    int64_t out_off;       // merged column of this swath's column 0
    int64_t own_end;       // exclusive end of the owned merged columns (seam)
    int64_t valid_start, valid_end;
};

struct CalParams {
    CalSwathDev sw[kMaxSwaths];
    int32_t n_sw;
    int64_t out_lines, out_cols;
    float* sigma0;
    float* nesz;
};

}  // namespace sd
