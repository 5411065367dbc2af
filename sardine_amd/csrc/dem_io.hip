// dem_io.hip — host-side reader of single-band GeoTIFF DEM tiles, the file
// side of Glo30Tile::load (dem.rs:248-352), which the reference decodes with
// the `tiff` crate (0.11).  Classic TIFF and BigTIFF, either byte order,
// strips or tiles, no compression / LZW / Deflate / PackBits, horizontal
// (2) and floating-point (3) predictors; the first image only.
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <thread>

#include "dem_io.hpp"

namespace sd {
namespace {

// worker threads for block decoding (a share of a host's cores)
constexpr uint64_t kDecodeThreads = 16;
// Bounds on what a header may ask the decoder to allocate: the decoded image
// (a GLO-30 tile is 3600 x 3600 x 4 B = 52 MB), one block, and a block side.
constexpr uint64_t kMaxImageBytes = 1ull << 32;
constexpr uint64_t kMaxBlockBytes = 1ull << 28;
constexpr uint64_t kMaxBlockSide = 1ull << 16;

// a * b, false when it does not fit in 64 bits
bool mul_ok(uint64_t a, uint64_t b, uint64_t* r) { return !__builtin_mul_overflow(a, b, r); }

struct File {
    std::vector<uint8_t> b;
    bool le = true, big = false;
    bool has(uint64_t off, uint64_t n) const { return off <= b.size() && n <= b.size() - off; }
    uint64_t get(uint64_t off, int n) const {  // n bytes in file order
        uint64_t v = 0;
        for (int i = 0; i < n; ++i) {
            const uint64_t byte = b[off + (le ? i : n - 1 - i)];
            v |= byte << (8 * i);
        }
        return v;
    }
};

struct Entry {
    uint16_t tag = 0, type = 0;
    uint64_t count = 0;
    uint64_t field = 0;  // file offset of the entry's value field
};

int type_size(uint16_t t) {
    switch (t) {
        case 1: case 2: case 6: case 7: return 1;
        case 3: case 8: return 2;
        case 4: case 9: case 11: case 13: return 4;
        case 5: case 10: case 12: case 16: case 17: case 18: return 8;
        default: return 0;
    }
}

// Where an entry's values start: inline in the value field when they fit.
bool values_at(const File& f, const Entry& e, uint64_t* off) {
    const int sz = type_size(e.type);
    uint64_t n;
    if (!sz || !mul_ok(e.count, (uint64_t)sz, &n)) return false;
    const int field = f.big ? 8 : 4;
    *off = n <= (uint64_t)field ? e.field : f.get(e.field, field);
    return f.has(*off, n);
}

bool ints(const File& f, const Entry& e, std::vector<uint64_t>* out) {
    uint64_t off;
    if (!values_at(f, e, &off)) return false;
    const int sz = type_size(e.type);
    if (e.type != 1 && e.type != 3 && e.type != 4 && e.type != 16) return false;
    out->resize(e.count);
    for (uint64_t i = 0; i < e.count; ++i) (*out)[i] = f.get(off + i * sz, sz);
    return true;
}

// The entry's values as doubles; on failure `out` is left empty (an entry of
// another type counts as absent, never as zeros).
bool doubles(const File& f, const Entry& e, std::vector<double>* out) {
    out->clear();
    uint64_t off;
    if (!values_at(f, e, &off) || (e.type != 11 && e.type != 12)) return false;
    out->resize(e.count);
    for (uint64_t i = 0; i < e.count; ++i) {
        if (e.type == 12) {
            const uint64_t u = f.get(off + 8 * i, 8);
            std::memcpy(&(*out)[i], &u, 8);
        } else if (e.type == 11) {
            const uint32_t u = (uint32_t)f.get(off + 4 * i, 4);
            float x;
            std::memcpy(&x, &u, 4);
            (*out)[i] = x;
        }
    }
    return true;
}

// TIFF LZW (MSB-first codes, 9-12 bits, the code width grows one code early).
bool lzw(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t want) {
    uint16_t prefix[4096];
    uint8_t suffix[4096], first[4096];
    uint16_t len[4096];
    for (int i = 0; i < 256; ++i) {
        prefix[i] = 0xffff;
        suffix[i] = first[i] = (uint8_t)i;
        len[i] = 1;
    }
    out.clear();
    size_t bitpos = 0;
    int width = 9, next = 258, old = -1;
    std::vector<uint8_t> tmp(4096);
    auto read = [&](int w) -> int {
        if (bitpos + (size_t)w > n * 8) return 257;  // out of data: as end of information
        int v = 0;
        for (int k = 0; k < w; ++k, ++bitpos) v = (v << 1) | ((in[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
        return v;
    };
    auto emit = [&](int code) {
        const int l = len[code];
        for (int k = l - 1, c = code; k >= 0; --k, c = prefix[c]) tmp[(size_t)k] = suffix[c];
        out.insert(out.end(), tmp.begin(), tmp.begin() + l);
    };
    while (out.size() < want) {
        int code = read(width);
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            code = read(width);
            if (code == 257) break;
            if (code > 255) return false;
            emit(code);
            old = code;
            continue;
        }
        if (old < 0) return false;
        if (code < next) {
            emit(code);
            if (next < 4096) {
                prefix[next] = (uint16_t)old;
                suffix[next] = first[code];
                first[next] = first[old];
                len[next] = (uint16_t)(len[old] + 1);
                ++next;
            }
        } else if (code == next && next < 4096) {
            prefix[next] = (uint16_t)old;
            suffix[next] = first[old];
            first[next] = first[old];
            len[next] = (uint16_t)(len[old] + 1);
            ++next;
            emit(code);
        } else {
            return false;
        }
        old = code;
        if (next == 511) width = 10;
        else if (next == 1023) width = 11;
        else if (next == 2047) width = 12;
    }
    return out.size() >= want;
}

bool packbits(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t want) {
    out.clear();
    size_t i = 0;
    while (i < n && out.size() < want) {
        const int8_t h = (int8_t)in[i++];
        if (h >= 0) {
            if (i + (size_t)h + 1 > n) return false;
            out.insert(out.end(), in + i, in + i + h + 1);
            i += (size_t)h + 1;
        } else if (h != -128) {
            if (i >= n) return false;
            out.insert(out.end(), (size_t)(1 - h), in[i++]);
        }
    }
    return out.size() >= want;
}

bool inflate_all(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t want) {
    out.assign(want, 0);
    z_stream z{};
    if (inflateInit(&z) != Z_OK) return false;
    z.next_in = const_cast<Bytef*>(in);
    z.avail_in = (uInt)n;
    z.next_out = out.data();
    z.avail_out = (uInt)want;
    const int r = inflate(&z, Z_FINISH);
    const size_t got = want - z.avail_out;
    inflateEnd(&z);
    return (r == Z_STREAM_END || r == Z_OK || r == Z_BUF_ERROR) && got == want;
}

}  // namespace

namespace {

int read_geotiff_impl(const char* path, GeoTiffRaster* out, std::string* err, bool decode, bool defer_fp_predictor) {
    File f;
    {
        std::FILE* fp = std::fopen(path, "rb");
        if (!fp) {
            *err = std::string("cannot open ") + path;
            return 1;
        }
        // a regular file, read whole in one call
        struct stat st;
        bool bad = fstat(fileno(fp), &st) != 0 || !S_ISREG(st.st_mode);
        const long long size = bad ? 0 : (long long)st.st_size;
        if (!bad) {
            f.b.resize((size_t)size);
            bad = std::fread(f.b.data(), 1, f.b.size(), fp) != f.b.size();
        }
        std::fclose(fp);
        if (bad) {
            *err = std::string("cannot read ") + path;
            return 1;
        }
    }
    auto bad = [&](const std::string& why) {
        *err = why;
        return 2;
    };
    if (!f.has(0, 8)) return bad("file too short for a TIFF header");
    if (f.b[0] == 'I' && f.b[1] == 'I') f.le = true;
    else if (f.b[0] == 'M' && f.b[1] == 'M') f.le = false;
    else return bad("not a TIFF file (byte order)");
    const uint64_t magic = f.get(2, 2);
    uint64_t ifd;
    if (magic == 42) {
        ifd = f.get(4, 4);
    } else if (magic == 43) {
        if (!f.has(0, 16) || f.get(4, 2) != 8) return bad("bad BigTIFF header");
        f.big = true;
        ifd = f.get(8, 8);
    } else {
        return bad("not a TIFF file (magic)");
    }
    const int cnt_size = f.big ? 8 : 2, ent_size = f.big ? 20 : 12;
    if (!f.has(ifd, cnt_size)) return bad("IFD outside the file");
    const uint64_t n = f.get(ifd, cnt_size);
    if (n > f.b.size() / ent_size || !f.has(ifd + cnt_size, n * ent_size)) return bad("IFD entries outside the file");
    std::vector<Entry> ents;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t o = ifd + cnt_size + i * ent_size;
        Entry e;
        e.tag = (uint16_t)f.get(o, 2);
        e.type = (uint16_t)f.get(o + 2, 2);
        e.count = f.get(o + 4, f.big ? 8 : 4);
        e.field = o + (f.big ? 12 : 8);
        ents.push_back(e);
    }
    auto find = [&](uint16_t tag) -> const Entry* {
        for (const Entry& e : ents)
            if (e.tag == tag) return &e;
        return nullptr;
    };
    auto scalar = [&](uint16_t tag, uint64_t dflt, uint64_t* v) -> bool {
        const Entry* e = find(tag);
        if (!e) {
            *v = dflt;
            return true;
        }
        std::vector<uint64_t> x;
        if (!ints(f, *e, &x) || x.empty()) return false;
        *v = x[0];
        return true;
    };
    uint64_t w, h, bits, comp, spp, fmt, pred, planar;
    if (!find(256) || !find(257) || !scalar(256, 0, &w) || !scalar(257, 0, &h)) return bad("missing image size");
    if (!scalar(258, 1, &bits) || !scalar(259, 1, &comp) || !scalar(277, 1, &spp) || !scalar(339, 1, &fmt) ||
        !scalar(317, 1, &pred) || !scalar(284, 1, &planar))
        return bad("unreadable image tags");
    if (w == 0 || h == 0 || w > (1u << 30) || h > (1u << 30)) return bad("image size out of range");
    if (spp != 1) return bad("more than one sample per pixel");
    if (bits % 8 != 0 || bits == 0 || bits > 64) return bad("bits per sample not a whole number of bytes");
    const int bps = (int)(bits / 8);
    out->width = (int64_t)w;
    out->height = (int64_t)h;
    out->sample = bits == 32 && fmt == 3 ? TiffSample::F32
                  : bits == 16 && fmt == 2 ? TiffSample::I16
                  : bits == 16 && fmt == 1 ? TiffSample::U16
                                           : TiffSample::Other;
    if (const Entry* e = find(33550)) doubles(f, *e, &out->pixel_scale);
    if (const Entry* e = find(33922)) doubles(f, *e, &out->tiepoint);
    if (const Entry* e = find(42113)) {
        uint64_t off;
        if (e->type == 2 && values_at(f, *e, &off)) {
            std::string s(reinterpret_cast<const char*>(&f.b[off]), (size_t)e->count);
            const size_t z = s.find('\0');
            out->nodata = z == std::string::npos ? s : s.substr(0, z);
            out->has_nodata = true;
        }
    }
    if (!decode) return 0;
    // from here on the header fields above are set: errors are decode errors
    auto bad_block = [&](const std::string& why) {
        *err = why;
        return 3;
    };
    // blocks: tiles or strips
    const bool tiled = find(322) != nullptr;
    uint64_t bw, bh;
    std::vector<uint64_t> offs, counts;
    if (tiled) {
        if (!scalar(322, 0, &bw) || !scalar(323, 0, &bh) || !find(324) || !find(325) ||
            !ints(f, *find(324), &offs) || !ints(f, *find(325), &counts))
            return bad_block("unreadable tile tags");
    } else {
        uint64_t rps;
        if (!scalar(278, h, &rps) || !find(273) || !find(279) || !ints(f, *find(273), &offs) ||
            !ints(f, *find(279), &counts))
            return bad_block("unreadable strip tags");
        bw = w;
        bh = std::min<uint64_t>(std::max<uint64_t>(rps, 1), h);
    }
    if (bw == 0 || bh == 0) return bad_block("zero block size");
    // every size below is checked before anything is allocated from it
    uint64_t img_bytes, block_bytes;
    if (!mul_ok(w, h, &img_bytes) || !mul_ok(img_bytes, (uint64_t)bps, &img_bytes) || img_bytes > kMaxImageBytes)
        return bad_block("image of " + std::to_string(w) + " x " + std::to_string(h) + " samples is too large");
    // (a strip is never larger than the image; a tile may be, so its sides
    // and bytes have their own bounds)
    if ((tiled && (bw > kMaxBlockSide || bh > kMaxBlockSide)) || !mul_ok(bw, bh, &block_bytes) ||
        !mul_ok(block_bytes, (uint64_t)bps, &block_bytes) || (tiled && block_bytes > kMaxBlockBytes))
        return bad_block("block of " + std::to_string(bw) + " x " + std::to_string(bh) + " samples is too large");
    const uint64_t nx = (w + bw - 1) / bw, ny = (h + bh - 1) / bh;
    if (offs.size() != nx * ny || counts.size() != nx * ny) return bad_block("block count does not match the image");
    if (comp != 1 && comp != 5 && comp != 8 && comp != 32946 && comp != 32773)
        return bad_block("unsupported compression " + std::to_string(comp));
    if (pred != 1 && pred != 2 && pred != 3) return bad_block("unsupported predictor " + std::to_string(pred));
    const bool defer = defer_fp_predictor && pred == 3 && bps == 4 && out->sample == TiffSample::F32 &&
                       (int64_t)(bw * bps) <= kMaxDeferredRowBytes;
    uint64_t deferred_bytes = 0;
    if (defer && (!mul_ok(nx * ny, block_bytes, &deferred_bytes) || deferred_bytes > kMaxImageBytes))
        return bad_block("tiled image is too large");
    if (defer) {
        out->fp_deferred = true;
        out->block_w = (int64_t)bw;
        out->block_h = (int64_t)bh;
        out->blocks_x = (int64_t)nx;
        out->blocks_y = (int64_t)ny;
        out->pixels.assign((size_t)deferred_bytes, 0);
    } else {
        out->pixels.assign((size_t)img_bytes, 0);
    }
    // Blocks decode independently into disjoint parts of the image, so they
    // are spread over worker threads; the error reported is the first block's
    // in file order, as a sequential decode would report it.
    const uint64_t n_blocks = nx * ny;
    std::vector<uint8_t> fault((size_t)n_blocks, 0);  // 1: outside the file, 2: corrupt, 3: out of memory
    auto decode_block = [&](uint64_t k, std::vector<uint8_t>& blk, std::vector<uint8_t>& tmp) -> uint8_t {
        const uint64_t by = k / nx, bx = k % nx;
        // a strip holds its own rows only; a tile is always full size
        const uint64_t rows = tiled ? bh : std::min(bh, h - by * bh);
        const size_t want = (size_t)(rows * bw * bps);  // <= block_bytes, checked above
        if (!f.has(offs[k], counts[k])) return 1;
        const uint8_t* src = &f.b[offs[k]];
        const size_t sn = (size_t)counts[k];
        bool ok;
        if (comp == 1) {
            ok = sn >= want;
            if (ok) blk.assign(src, src + want);
        } else if (comp == 5) {
            ok = lzw(src, sn, blk, want);
        } else if (comp == 32773) {
            ok = packbits(src, sn, blk, want);
        } else {
            ok = inflate_all(src, sn, blk, want);
        }
        if (!ok) return 2;
        if (defer) {  // the block as inflated, into its slot
            std::memcpy(&out->pixels[(size_t)(k * bh * bw * bps)], blk.data(), want);
            return 0;
        }
        blk.resize(want);
        const size_t row_bytes = (size_t)(bw * bps);
        for (uint64_t r = 0; r < rows; ++r) {
            uint8_t* p = &blk[(size_t)r * row_bytes];
            if (pred == 3) {
                // floating-point predictor: byte differences along the row,
                // then byte planes (most significant first) to samples
                for (size_t i = 1; i < row_bytes; ++i) p[i] = (uint8_t)(p[i] + p[i - 1]);
                tmp.assign(p, p + row_bytes);
                for (uint64_t s = 0; s < bw; ++s)
                    for (int byte = 0; byte < bps; ++byte) p[s * bps + byte] = tmp[(size_t)(bps - byte - 1) * bw + s];
            } else {
                if (!f.le && bps > 1)
                    for (uint64_t s = 0; s < bw; ++s) std::reverse(p + s * bps, p + (s + 1) * bps);
                if (pred == 2) {
                    for (uint64_t s = 1; s < bw; ++s) {
                        uint64_t a = 0, b = 0;
                        std::memcpy(&a, p + (s - 1) * bps, (size_t)bps);
                        std::memcpy(&b, p + s * bps, (size_t)bps);
                        const uint64_t sum = a + b;
                        std::memcpy(p + s * bps, &sum, (size_t)bps);
                    }
                }
            }
            // the block's row into the image (tiles are cropped at the edges)
            const uint64_t y = by * bh + r;
            if (y >= h) break;
            const uint64_t x0 = bx * bw, cols = std::min(bw, w - x0);
            std::memcpy(&out->pixels[(size_t)((y * w + x0) * bps)], p, (size_t)(cols * bps));
        }
        return 0;
    };
    std::atomic<uint64_t> next{0};
    auto worker = [&]() {
        std::vector<uint8_t> blk, tmp;
        for (uint64_t k; (k = next.fetch_add(1)) < n_blocks;) {
            try {  // nothing may escape a worker thread (std::terminate)
                fault[(size_t)k] = decode_block(k, blk, tmp);
            } catch (const std::exception&) {
                fault[(size_t)k] = 3;
            }
        }
    };
    const uint64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint64_t n_threads = std::min<uint64_t>(std::min<uint64_t>(hw, kDecodeThreads), n_blocks);
    std::vector<std::thread> pool;
    for (uint64_t t = 1; t < n_threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    for (uint64_t k = 0; k < n_blocks; ++k) {
        if (fault[(size_t)k] == 1) return bad_block("block outside the file");
        if (fault[(size_t)k] == 2) return bad_block("corrupt compressed block");
        if (fault[(size_t)k] == 3) return bad_block("out of memory decoding a block");
    }
    return 0;
}

}  // namespace

int read_geotiff(const char* path, GeoTiffRaster* out, std::string* err, bool decode, bool defer_fp_predictor) {
    try {
        return read_geotiff_impl(path, out, err, decode, defer_fp_predictor);
    } catch (const std::bad_alloc&) {
        *err = "out of memory reading the file";
    } catch (const std::exception& e) {
        *err = e.what();
    }
    out->pixels.clear();
    return 3;
}

}  // namespace sd
