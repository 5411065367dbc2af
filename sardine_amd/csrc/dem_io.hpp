// dem_io.hpp — host-side reader of single-band GeoTIFF DEM tiles (the
// Copernicus GLO-30 case of Glo30Tile::load, dem.rs:248-352).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace sd {

enum class TiffSample { F32, I16, U16, Other };

struct GeoTiffRaster {
    int64_t width = 0, height = 0;
    TiffSample sample = TiffSample::Other;
    std::vector<uint8_t> pixels;  // width * height samples, host byte order, row-major
    std::vector<double> pixel_scale, tiepoint;  // ModelPixelScale (33550), ModelTiepoint (33922); empty if absent
    bool has_nodata = false;
    std::string nodata;  // GDAL_NODATA (42113), as stored
    // Set when the floating-point predictor was left for the device (see
    // read_geotiff): `pixels` then holds the inflated blocks as stored, block
    // after block in file order, each block_h rows of block_w * 4 bytes.
    bool fp_deferred = false;
    int64_t block_w = 0, block_h = 0, blocks_x = 0, blocks_y = 0;
};

// Decodes the first image of a TIFF file (with decode false: its size,
// sample type and GeoTIFF tags only).  Returns 0, or with *err set: 1 when
// the file cannot be read, 2 when its header is not one this reader takes,
// 3 when the header was read (the size, sample type and tags are set) but
// the image data cannot be decoded (the message says why).
// With defer_fp_predictor, an f32 image with the floating-point predictor
// (3) and block rows of at most kMaxDeferredRowBytes is only inflated: the
// predictor is undone on the device (fp_deferred is set).
constexpr int64_t kMaxDeferredRowBytes = 32768;
int read_geotiff(const char* path, GeoTiffRaster* out, std::string* err, bool decode = true,
                 bool defer_fp_predictor = false);

}  // namespace sd
